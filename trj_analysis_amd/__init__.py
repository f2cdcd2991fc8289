"""MI355X-native structural hot path of trj_analysis: species-resolved pair
histograms g_ab(r) and the static structure factor S(Q) as hand-written HIP
kernels for gfx950 behind a C ABI (include/trj_hip.h).

Modules
  trjhip     ctypes binding of libtrjhip.so (the C ABI)
  mod_rdf    the reference's mod_rdf interface (RDF_init, RDFcomp, printrdf, rdf_clear)
  mod_sq     the reference's mod_sq interface (sq_init, SQcalc, sq_transfer_gpu_cpu,
             printSQ, sq_clear)
  namelist   /INPUT/, /INPUT_RDF/, /INPUT_SQ/ namelist reader
  workloads  seeded synthetic frames of the benchmark configurations
  build      hipcc build of libtrjhip.so
"""
__all__ = ["trjhip", "mod_rdf", "mod_sq", "namelist", "workloads", "build"]
