"""Build the trj_hip shared library for gfx950 with hipcc (in-tree).

    python -m trj_analysis_amd.build        # or trj_analysis_amd.build.build()

Produces trj_analysis_amd/libtrjhip.so next to this file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtrjhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TRJ_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "trj_ctx.hip", "batch.hip", "serve.hip", "plan.cpp", "ncread.cpp"]
# -ffp-contract=off: the RDF pair rule and the FP32 q-set bookkeeping must
# round once per operation, as written.  Packed-FP32 instructions are disabled.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result", "-Wno-unused-value"]
HIPFLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "c++"] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = ["trj_hip.f90", "trj_hip_modules.f90"]
FPROGRAMS = ["example_driver.f90", "abi_check.f90"]
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
REF_MODULES = ["mod_precision.f90", "mod_input.f90", "mod_common.f90", "mod_rdf.f90", "mod_sq.f90"]


def build_fortran(verbose: bool = False) -> list:
    """ISO_C_BINDING modules and the Fortran programs (amdflang): the two over
    the trj_hip modules and ref_driver over the reference-signature modules,
    linked against libtrjhip.so with an rpath relative to the executable."""
    fdir = os.path.join(HERE, "fortran")
    objdir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in FORTRAN:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [FLANG, "-O2", "-fPIC", "-J", objdir, "-c", os.path.join(fdir, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    out = []
    for prog in FPROGRAMS:
        exe = os.path.join(fdir, os.path.splitext(prog)[0])
        cmd = [FLANG, "-O2", "-J", objdir, os.path.join(fdir, prog)] + objs + \
            ["-L" + HERE, "-ltrjhip", "-Wl,-rpath,$ORIGIN/..", "-o", exe + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    # The reference-signature modules (mod_precision, mod_input, mod_common,
    # mod_rdf, mod_sq) in a module directory of their own, over the trj_hip
    # interface module, and the driver that runs the reference's frame loop.
    rdir = os.path.join(fdir, "reference_api")
    robj = os.path.join(objdir, "reference_api")
    os.makedirs(robj, exist_ok=True)
    robjs = [objs[0]]  # trj_hip.o: the C-ABI interfaces
    for src in REF_MODULES:
        obj = os.path.join(robj, os.path.splitext(src)[0] + ".o")
        cmd = [FLANG, "-O2", "-fPIC", "-I", objdir, "-J", robj, "-c", os.path.join(rdir, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        robjs.append(obj)
    exe = os.path.join(rdir, "ref_driver")
    cmd = [FLANG, "-O2", "-I", objdir, "-J", robj, os.path.join(rdir, "ref_driver.f90")] + robjs + \
        ["-L" + HERE, "-ltrjhip", "-Wl,-rpath,$ORIGIN/../..", "-o", exe + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(exe + ".tmp", exe)
    out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
