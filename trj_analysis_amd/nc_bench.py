"""Throughput of the NetCDF path: trajectory read alone (host) and the whole
driver (read + device frame + output files) on a synthetic LAMMPS-style
CDF-2 file of the lattice_262k workload (262144 atoms, 3 species, float32
coordinates as LAMMPS writes them).

    python -m trj_analysis_amd.nc_bench [--frames 6] [--precision fp64] [--read-only]

Prints one JSON line.  The file is written to a scratch folder first (not
timed)."""
from __future__ import annotations

import argparse
import json
import os
import tempfile
import time

import numpy as np

from . import trjhip as T
from . import workloads as W


def write_lattice_nc(path: str, nframes: int) -> W.Workload:
    from scipy.io import netcdf_file

    wl = W.LATTICE_262K
    frames = list(W.lattice_262k(nframes=nframes))
    n = len(frames[0].itype)
    with netcdf_file(path, "w", version=2) as nc:
        nc.createDimension("frame", None)
        nc.createDimension("atom", n)
        nc.createDimension("spatial", 3)
        nc.createDimension("cell_spatial", 3)
        t = nc.createVariable("time", "f", ("frame",))
        t.units = "lj"
        co = nc.createVariable("cell_origin", "f", ("frame", "cell_spatial"))
        cl = nc.createVariable("cell_lengths", "f", ("frame", "cell_spatial"))
        ty = nc.createVariable("type", "i", ("frame", "atom"))
        xyz = nc.createVariable("coordinates", "f", ("frame", "atom", "spatial"))
        for f, fr in enumerate(frames):
            t[f] = 100.0 * (f + 1)
            co[f] = np.zeros(3, dtype=np.float32)
            cl[f] = fr.box.astype(np.float32)
            ty[f] = fr.itype
            xyz[f] = fr.pos.astype(np.float32)
    return wl


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=6)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--pipeline", choices=["serial", "overlap"], default="serial")
    ap.add_argument("--read-only", action="store_true")
    ap.add_argument("--dir", default=None)
    a = ap.parse_args(argv)
    work = a.dir or tempfile.mkdtemp(prefix="trj_nc_bench_")
    os.makedirs(work, exist_ok=True)
    path = os.path.join(work, "run.nc")
    wl = write_lattice_nc(path, a.frames)
    size = os.path.getsize(path)
    out = {"workload": "lattice_262k NetCDF (CDF-2, float32 coordinates)", "frames": a.frames,
           "file_bytes": size}
    with T.NcTrajectory(path) as tr:
        t0 = time.perf_counter()
        for f in range(tr.nframes):
            tr.frame_ref(f, [1, 2, 3])
        dt = time.perf_counter() - t0
        out["read_ms_per_frame"] = 1e3 * dt / tr.nframes
        out["read_GBps"] = size / dt / 1e9
    if not a.read_only:
        from . import driver as D

        with open(os.path.join(work, "input.nml"), "w") as fh:
            fh.write(f"""&INPUT
  trj_input_file='run.nc'
  ndim=3
  nsp=3
  ncfs_from_to=0 1 0
  rdf_sq_cl_dyn_sqw_conf_ord=T T F F F F F
/
&INPUT_SP
  sp_types_selected=1 2 3
/
&INPUT_RDF
  deltar={wl.deltar}
  rcrdf={wl.rcrdf}
/
&INPUT_SQ
  qmax={wl.qmax}
  qmin={wl.qmax}
  bsc={' '.join(str(b) for b in wl.bsc)}
/
""")
        pipe = T.TRJ_PIPE_OVERLAP if a.pipeline == "overlap" else T.TRJ_PIPE_SERIAL
        res = D.run(os.path.join(work, "input.nml"), outdir=os.path.join(work, "out"),
                    precision=a.precision, pipeline=pipe)
        out.update(precision=a.precision, pipeline=a.pipeline, frames_processed=res.frames_processed,
                   driver_frames_per_s=res.frames_processed / res.t_total,
                   driver_ms_per_frame=1e3 * res.t_total / res.frames_processed,
                   driver_read_ms_per_frame=1e3 * res.t_read / res.frames_processed,
                   driver_init_ms=1e3 * res.t_init, driver_files_ms=1e3 * res.t_files,
                   driver_steady_ms_per_frame=1e3 * (res.t_total - res.t_init - res.t_files)
                   / res.frames_processed)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
