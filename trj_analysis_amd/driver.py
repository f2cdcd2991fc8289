"""The reference's main program for the structural path: namelist in, NetCDF
trajectory through the frame loop, gmixsim.dat / sq.dat / sqmix.dat out.

Follows src/main/trj_analysis.f90:135-253 with the modules other than RDF and
S(Q) left out:

  read_input_file                    input.f90:78-217 (namelist.read_input)
  open_netcdf_input + first read     trj_analysis.f90:139-149, netcdf.f90:253-264
  select_species                     util.f90:83-126
  nstep == 0 on the first frame      trj_analysis.f90:153-157 (nstepi0)
  reset_confs                        util.f90:128-145
  frame loop                         trj_analysis.f90:173-242: ncstart stride,
                                     EOF exit, init on the first frame read,
                                     frames with nstep == 0 skipped
  print_results                      util.f90:280-300 -> printrdf / printSQ

Every frame goes to the device in one trj_frame call that asks for both the
pair histogram and S(Q) (the reference transfers once and then runs RDFcomp
and SQcalc on the resident copy, trj_analysis.f90:206-216).

Normalisation uses Nconf as reset_confs leaves it (ncfs_from_to(1), capped at
the frames in the file), as printrdf / printSQ do (log.f90:363-380, 504-509),
also when frames were skipped or the loop met the end of the file first.

Species (netcdf.f90:456-531, select_species): the selected LAMMPS types are
numbered in order of first appearance in the first frame read, and bsc follows
them (bsc(k) belongs to sp_types_selected(k)).  When sp_types_selected leaves
entries at 0 and the file holds exactly nsp types, the reference keeps every
atom (Nsites = natoms, util.f90:89-90); here the zero entries take the file's
unselected types in order of first appearance, which gives that atom set a
defined species map.
"""
from __future__ import annotations

import os
import sys
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Union

import numpy as np

from . import mod_rdf, mod_sq
from . import trjhip as T
from .namelist import StructuralInput, read_input_file


@dataclass
class RunResult:
    nconf: int                      # the normaliser printrdf / printSQ use
    frames_processed: int           # Nccount of trj_analysis.f90:241
    frames_skipped: List[int] = field(default_factory=list)  # 1-based, nstep == 0
    frames_read: List[int] = field(default_factory=list)     # 1-based ncstart values
    nsites: int = 0
    species_types: List[int] = field(default_factory=list)   # LAMMPS type of species 1..nsp
    lsmax: Optional[int] = None
    rmingr: Optional[float] = None
    sq_info: Optional[object] = None
    rdf_counts: Optional[np.ndarray] = None   # (nit, lsmax) int64, as trj_rdf_counts (slices: per slice)
    sq_shells: Optional[tuple] = None         # (sqf, sqfp, nq), as trj_sq_shells
    charges: bool = False                     # the file holds q (ex_qc): g_qq and S_qq
    rdf_qq: Optional[np.ndarray] = None       # charge-weighted accumulator, as trj_rdf_qq
    sq_qq: Optional[np.ndarray] = None        # charge-weighted shell sums, as trj_sq_qq
    s2n_sums: Optional[tuple] = None          # (gi, g2i) of the number fluctuations (nrandom > 0)
    files: List[str] = field(default_factory=list)
    t_read: float = 0.0             # seconds in trj_nc_frame_ref (host read + conversion)
    t_total: float = 0.0            # seconds from the first read to the written files
    t_init: float = 0.0             # seconds in RDF_init / sq_init and the first-frame rcl
    t_files: float = 0.0            # seconds writing the output tables (after the final sync)


def _fdiv(a: int, b: int) -> int:
    """Fortran integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _nstep(tr: T.NcTrajectory, frame: int) -> int:
    """time(frame) read into an integer, as nf90_get_var converts (truncation).
    Without a time variable every frame counts as a nonzero step."""
    return int(tr.time(frame)) if tr.has_time else 1


def resolve_species(nsp: int, selected, types_first_frame: np.ndarray):
    """select_species (util.f90:83-126) and the type discovery of read_nc_cfg
    (netcdf.f90:456-513): returns (sp_types for trj_nc_frame_ref, LAMMPS type
    of each species 1..nsp, bsc index of each species)."""
    sel = [int(x) for x in selected][:nsp]
    sel += [0] * (nsp - len(sel))
    if sel[0] == 0:
        raise T.TrjError(T.TRJ_EINVAL, "atom IDs must be specified in sp_types_selected (input.f90:118-121)")
    order: List[int] = []
    seen = set()
    for t in types_first_frame.tolist():
        if t not in seen:
            seen.add(t)
            order.append(int(t))
    ntypes_tr = len(order)
    if nsp > ntypes_tr:
        raise T.TrjError(T.TRJ_EINVAL, f"nsp={nsp} is larger than the {ntypes_tr} types in the trajectory "
                                       "(util.f90:88-90)")
    if 0 in sel:
        if nsp < ntypes_tr:
            raise T.TrjError(T.TRJ_EINVAL, "select the species to analyze: sp_types_selected has zeros "
                                           "(util.f90:96-101)")
        rest = [t for t in order if t not in sel]
        it = iter(rest)
        sel = [s if s != 0 else next(it) for s in sel]
    if len(set(sel)) != nsp:
        raise T.TrjError(T.TRJ_EINVAL, "a type appears more than once in sp_types_selected (netcdf.f90:563-566)")
    missing = [s for s in sel if s not in seen]
    if missing:
        raise T.TrjError(T.TRJ_EINVAL, f"species type(s) {missing} not in the trajectory (util.f90:102-110)")
    species_types = [t for t in order if t in sel]
    bsc_index = [sel.index(t) for t in species_types]
    return sel, species_types, bsc_index


def frame_schedule(ncfs_from_to, nmconf: int, nstep_of: Callable[[int], int],
                   say: Callable[[str], None] = lambda s: None):
    """Frames the reference's loop visits: (Nconf, [(ncstart, skipped)]) with
    1-based ncstart.  Follows the first read (netcdf.f90:253-264), the
    nstepi0 test (trj_analysis.f90:153-157), reset_confs (util.f90:128-145)
    and the loop's stride and end-of-file exit (trj_analysis.f90:173-190);
    nstep_of(frame) is the frame's integer time step, 0 marks a skipped frame
    (trj_analysis.f90:196-198)."""
    n1, n2, n3 = (list(ncfs_from_to) + [0, 0, 0])[:3]
    if n2 == 0:  # input.f90:137-140
        n2 = 1
    if n1 == 0:  # first read_nc_cfg call
        n1, n3 = nmconf, nmconf
    if n2 > nmconf or n2 < 1:
        raise T.TrjError(T.TRJ_ERANGE, f"read NetCDF trajectory file beyond end of file: frame {n2} "
                                       f"of {nmconf} (trj_analysis.f90:145-149)")
    nstepi0 = nstep_of(n2) == 0
    if nstepi0:
        say(" !!*** Warning: initial configuration for step 0, skipping ...")
    nconf = n1  # reset_confs; n1 is nonzero here, so its whole-file branch is already taken
    n3 += 1
    if nconf > nmconf:
        say(f" !!*** Warning: resetting Nconf to {nmconf}, last conf. in trajectory !")
        nconf = nmconf
    out = []
    for i in range(1, n1 + 1):
        ncstart = n2 + _fdiv((i - 1) * (n3 - n2), n1) + (1 if (nstepi0 and n2 == 1) else 0)
        if ncstart > nmconf:  # netcdf.f90:260-264 -> reformat_input_conf, exit
            say(f" ** Warning : trying to read past last configuration no. {nmconf}")
            break
        if ncstart < 1:
            raise T.TrjError(T.TRJ_ERANGE, f"frame index {ncstart} below 1 (check ncfs_from_to)")
        out.append((ncstart, nstep_of(ncstart) == 0))
    return nconf, out


def _file_z(tr: T.NcTrajectory, frame: int, per_in):
    """z of all atoms as read_nc_cfg has it (netcdf.f90:317-325), FP32: minus
    cell_origin when the /INPUT/ z axis is periodic; and the z cell length."""
    raw, _, org, cell = tr.read(frame)
    z = raw[:, 2].astype(np.float32)
    if per_in[2]:
        z = z - np.float32(org[2])
    return z, cell[2]


def _setup_slices(inp: StructuralInput, tr: T.NcTrajectory, frame: int, per_in, ctx: T.Context, say):
    """The z-slices on the first frame read (netcdf.f90:326-346, input.f90:
    176-191): auto slices (nslice < 1) from the z range of all atoms, in FP32;
    with a zero z cell length the walls pwall = min z - 6, pwallp = pwall +
    span + 8 bound the slice centres (and, per frame, the atoms).  Returns the
    walls or None."""
    z, lz = _file_z(tr, frame, per_in)
    zmin, zmax = z.min(), z.max()
    walls = None
    if lz == 0.0:
        pwall = np.float32(zmin - np.float32(6.0))
        pwallp = np.float32(pwall + np.float32(abs(zmax - zmin) + np.float32(8.0)))
        walls = (pwall, pwallp)
    if inp.auto_zslice:
        if lz != 0.0:
            raise T.TrjError(T.TRJ_EINVAL, "automatic z-slices (nslice < 1) need a zero z cell length "
                                           "(netcdf.f90:333-339)")
        zs = [np.float32((zmax + zmin) / np.float32(2.0))]
        zg = np.float32((zmax - zmin) / np.float32(10.0))
        say(f" ** Auto-defined z-slices: zslice(1) = {zs[0]:8.2f}, zgrid = {zg:8.2f}")
    else:
        zs = np.asarray(inp.zslice, dtype=np.float32)
        if inp.zgrid is None:
            raise T.TrjError(T.TRJ_EINVAL, "zgrid is not set in /INPUT_CONF/")
        zg = np.float32(inp.zgrid)
        if walls is not None and (zs.min() < walls[0] or zs.max() > walls[1]):
            raise T.TrjError(T.TRJ_EINVAL, "z-slices out of box limits, check your input parameters "
                                           "(netcdf.f90:341-344)")
    ctx.set_slices(zs, zg)
    return walls


def _slice_z(tr: T.NcTrajectory, frame: int, pos: np.ndarray, box, per, per_in, walls) -> np.ndarray:
    """z as select_ncdfinput leaves it (netcdf.f90:679-682): wrapped by
    r - L*int(r/L) in FP32 on a periodic z axis, as read otherwise (checked
    against the walls, netcdf.f90:348-355)."""
    if per[2]:
        z = pos[2].astype(np.float32)
        L = np.float32(box[2])
        pos[2] = (z - L * np.trunc(z / L).astype(np.float32)).astype(np.float64)
    elif walls is not None:
        z, _ = _file_z(tr, frame, per_in)
        if z.min() < walls[0] or z.max() > walls[1]:
            raise T.TrjError(T.TRJ_EINVAL, f"frame {frame + 1}: atoms outside the walls "
                                           f"[{walls[0]:.2f}, {walls[1]:.2f}] (netcdf.f90:348-355)")
    return pos


def run(inp: Union[str, StructuralInput], *, outdir: str = ".", trj_file: Optional[str] = None,
        precision: str = "fp64", pipeline: Optional[int] = None, device: Optional[int] = None,
        log: Optional[Callable[[str], None]] = None) -> RunResult:
    """The frame loop of trj_analysis.f90:135-253 for RDF and S(Q).  `inp` is a
    namelist file or a parsed StructuralInput; trj_file overrides its
    trj_input_file (relative names resolve against the namelist's folder)."""
    say = log or (lambda s: None)
    base = "."
    if isinstance(inp, str):
        base = os.path.dirname(os.path.abspath(inp))
        inp = read_input_file(inp)
    path = trj_file or inp.trj_input_file
    if not path:
        raise T.TrjError(T.TRJ_EINVAL, "no trj_input_file")
    if not os.path.isabs(path):
        path = os.path.join(base, path)
    if not os.path.exists(path):
        raise T.TrjError(T.TRJ_EINVAL, f"trajectory file {path} does not exist (input.f90:113-117)")
    if not (inp.run_rdf or inp.run_sq):
        raise T.TrjError(T.TRJ_EINVAL, "neither RDF nor S(Q) is selected in rdf_sq_cl_dyn_sqw_conf_ord")
    nsp, ndim = inp.nsp, inp.ndim
    if ndim not in (2, 3):
        raise T.TrjError(T.TRJ_EINVAL, "ndim must be 2 or 3")
    os.makedirs(outdir, exist_ok=True)

    if inp.slices and ndim != 3:
        raise T.TrjError(T.TRJ_EINVAL, "system must be 3D to compute confinement properties (input.f90:167-170)")
    with T.NcTrajectory(path) as tr:
        nmconf = tr.nframes
        if inp.slices:  # slice centres are absolute z: no centring of non-periodic axes
            tr.keep_nonperiodic(True)
        if not tr.has_type:
            raise T.TrjError(T.TRJ_EINVAL, "the trajectory has no type variable (netcdf.f90:546-548)")
        nconf, schedule = frame_schedule(inp.ncfs_from_to, nmconf, lambda f: _nstep(tr, f - 1), say)
        _, types0, _, _ = tr.read(max(inp.ncfs_from_to[1], 1) - 1)
        sel, species_types, bsc_index = resolve_species(nsp, inp.sp_types_selected, types0)
        bsc = [inp.bsc[k] if k < len(inp.bsc) else 1.0 for k in bsc_index]
        ctx = T.Context(int(os.environ.get("TRJ_DEVICE", "0")) if device is None else device)
        res = RunResult(nconf=nconf, frames_processed=0, species_types=species_types,
                        charges=tr.has_charge)
        qch = None
        per_in = tuple(bool(p) for p in inp.periodic)
        first = True
        what = (T.TRJ_WANT_RDF if inp.run_rdf else 0) | (T.TRJ_WANT_SQ if inp.run_sq else 0)
        walls = None
        try:
            if pipeline is not None:
                ctx.set_pipeline(pipeline)
            t_start = time.perf_counter()
            for i, (ncstart, skip) in enumerate(schedule, start=1):
                t0 = time.perf_counter()
                if res.charges:  # netcdf.f90:404-408: a q variable switches ex_qc on
                    pos, itype, box, per, qch = tr.frame_ref(ncstart - 1, sel, per_in, charges=True)
                else:
                    pos, itype, box, per = tr.frame_ref(ncstart - 1, sel, per_in)
                res.t_read += time.perf_counter() - t0
                res.frames_read.append(ncstart)
                if inp.slices:
                    if first:
                        walls = _setup_slices(inp, tr, ncstart - 1, per_in, ctx, say)
                    pos = _slice_z(tr, ncstart - 1, pos, box, per, per_in, walls)
                if first:  # init_modules (util.f90:180-260) on the first frame read
                    t0 = time.perf_counter()
                    res.nsites = len(itype)
                    if inp.run_rdf:
                        res.lsmax = mod_rdf.RDF_init(nsp, ndim=ndim, deltar=inp.deltar, rcrdf=inp.rcrdf,
                                                     sidel=box, periodic=tuple(bool(p) for p in per),
                                                     charges=res.charges,
                                                     nrandom=0 if inp.slices else inp.nrandom, ctx=ctx)
                    if inp.run_sq:
                        res.sq_info = mod_sq.sq_init(len(itype), nsp, ndim=ndim, qmax=inp.qmax, qmin=inp.qmin,
                                                     bsc=bsc, sidel=box, precision=precision,
                                                     flags=T.TRJ_SQ_CHARGE if res.charges else 0, ctx=ctx)
                    first = False
                    res.t_init += time.perf_counter() - t0
                if skip:
                    res.frames_skipped.append(ncstart)
                    continue
                if len(itype) != res.nsites:
                    raise T.TrjError(T.TRJ_EINVAL, f"frame {ncstart} holds {len(itype)} selected atoms, "
                                                   f"the first frame read {res.nsites}")
                ctx.frame(pos, itype, box, what=what, layout=T.TRJ_SOA, charge=qch)
                res.frames_processed += 1
                if inp.run_rdf and i == 1 and not inp.slices:  # rdf.cuf:343-352 (Iconf == 1)
                    t0 = time.perf_counter()
                    res.rmingr = ctx.rdf_first_minimum()
                    res.t_init += time.perf_counter() - t0
            if res.frames_processed == 0:
                raise T.TrjError(T.TRJ_EINVAL, "no configuration was analysed")
            ctx.sync()
            t_files = time.perf_counter()
            # print_results (util.f90:280-300)
            if inp.run_sq:
                if inp.slices:  # (sqf, sqfp, sqfq) per slice
                    res.sq_shells = ctx.sq_xy(charges=res.charges)
                else:
                    res.sq_shells = ctx.sq_shells()
                    if res.charges:
                        res.sq_qq = ctx.sq_qq()
                res.files += mod_sq.printSQ(res.nsites, nconf=nconf, outdir=outdir)
            if inp.run_rdf:
                if inp.slices:  # raw slice pair counts (nslice, nit, lsmax)
                    res.rdf_counts = ctx.rdf_xy()[1]
                    if res.charges:
                        res.rdf_qq = ctx.rdf_xy_qq()
                else:
                    res.rdf_counts = ctx.rdf_counts()
                    if res.charges:
                        res.rdf_qq = ctx.rdf_qq()
                    if inp.nrandom > 0:
                        res.s2n_sums = ctx.rdf_s2n_sums()
                res.files += mod_rdf.printrdf(0.0, res.lsmax, nconf=nconf, outdir=outdir)
            res.t_total = time.perf_counter() - t_start
            res.t_files = time.perf_counter() - t_files
        finally:
            mod_rdf._state.clear()
            mod_sq._state.clear()
            ctx.close()
    return res


def main(argv=None) -> int:
    """`python -m trj_analysis_amd.driver input.nml [device]`, the reference's
    command line (trj_analysis.f90:114-130): outputs go to the current folder."""
    argv = sys.argv[1:] if argv is None else argv
    if not argv or len(argv) > 2:
        print("usage: python -m trj_analysis_amd.driver <input.nml> [device]", file=sys.stderr)
        return 2
    dev = int(argv[1]) if len(argv) == 2 else 0
    res = run(argv[0], outdir=".", device=dev, log=print)
    print(f" ** frames analysed {res.frames_processed} (Nconf {res.nconf}), atoms {res.nsites}, "
          f"species types {res.species_types}")
    if res.frames_processed:
        print(f" ** total {res.t_total:.3f} s, {res.frames_processed / res.t_total:.2f} frames/s; "
              f"trajectory read {res.t_read:.3f} s")
    for f in res.files:
        print(" ** wrote", f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
