"""Frame selection and species resolution of the end-to-end driver (host
only).  The expected frame lists are worked by hand from the reference's
loop (trj_analysis.f90:144-190, netcdf.f90:253-264, util.f90:128-145):
ncstart = from + (i-1)*(to+1-from)/n with Fortran integer division, +1 when
the first frame read has step 0 and from == 1; Nconf = n capped at the file
length."""
import numpy as np
import pytest

from trj_analysis_amd import driver as D
from trj_analysis_amd import namelist as N
from trj_analysis_amd import trjhip as T


def _steps(zero=()):
    return lambda f: 0 if f in zero else 10 * f


def test_whole_file_step0_first():
    # ncfs_from_to = 0 -> whole file; frame 1 has step 0: frames 2..5, Nconf 5
    nconf, sch = D.frame_schedule((0, 1, 0), 5, lambda f: (f - 1) * 100)
    assert nconf == 5
    assert sch == [(2, False), (3, False), (4, False), (5, False)]


def test_whole_file_nonzero_steps():
    nconf, sch = D.frame_schedule((0, 0, 0), 4, _steps())
    assert nconf == 4 and sch == [(1, False), (2, False), (3, False), (4, False)]


def test_stride():
    # 3 frames from 2 to 10: to becomes 11, stride (11-2)/3 = 3 -> 2, 5, 8
    nconf, sch = D.frame_schedule((3, 2, 10), 10, _steps())
    assert nconf == 3 and [f for f, _ in sch] == [2, 5, 8]
    # 4 frames from 1 to 10: (i-1)*10/4 = 0, 2, 5, 7 -> 1, 3, 6, 8
    nconf, sch = D.frame_schedule((4, 1, 10), 10, _steps())
    assert [f for f, _ in sch] == [1, 3, 6, 8]


def test_more_frames_than_file():
    # 20 frames over 10: (i-1)*10/20 repeats frames; Nconf capped at 10
    nconf, sch = D.frame_schedule((20, 1, 10), 10, _steps())
    assert nconf == 10
    assert [f for f, _ in sch] == [1 + (i * 10) // 20 for i in range(20)]


def test_eof_exit_and_midfile_zero_step():
    # from 3, step 0 at frame 3 does not shift (from != 1); frame 5 with step 0 skipped
    nconf, sch = D.frame_schedule((4, 3, 9), 6, _steps(zero=(3, 5)))
    # to -> 10, stride (10-3)/4: 0, 1, 3, 5 -> 3, 4, 6, 8; 8 > 6 ends the loop
    assert nconf == 4
    assert sch == [(3, True), (4, False), (6, False)]


def test_first_frame_beyond_file():
    with pytest.raises(T.TrjError):
        D.frame_schedule((2, 7, 8), 6, _steps())


def test_resolve_species():
    types = np.array([3, 1, 3, 2, 1, 4], dtype=np.int32)  # first appearance 3, 1, 2, 4
    sel, sp, bi = D.resolve_species(2, [1, 3], types)
    assert sel == [1, 3] and sp == [3, 1] and bi == [1, 0]  # species 1 is type 3 -> bsc(2)
    # zeros with exactly nsp types in the file: filled from the unselected types
    sel, sp, bi = D.resolve_species(4, [2], types)
    assert sel == [2, 3, 1, 4] and sp == [3, 1, 2, 4] and bi == [1, 2, 0, 3]
    with pytest.raises(T.TrjError):
        D.resolve_species(2, [0, 1], types)       # sp_types_selected(1) == 0
    with pytest.raises(T.TrjError):
        D.resolve_species(3, [1, 0, 0], types)    # zeros while the file has more types
    with pytest.raises(T.TrjError):
        D.resolve_species(5, [1, 2, 3, 4, 5], types)  # nsp above the types in the file
    with pytest.raises(T.TrjError):
        D.resolve_species(2, [1, 7], types)       # type not in the file
    with pytest.raises(T.TrjError):
        D.resolve_species(2, [1, 1], types)       # repeated type


def test_namelist_trajectory_and_species():
    txt = """&INPUT
  trj_input_file='run.nc'
  ndim=3
  nsp=2
  ncfs_from_to=10 2 20
  rdf_sq_cl_dyn_sqw_conf_ord=T T F F F F F
/
&INPUT_SP
  sp_types_selected=1
  mat=1.0 2.0
/
&INPUT_RDF
  deltar=0.25
  rcrdf=15.0
/
&INPUT_SQ
  qmax=25.0
  qmin=6.4
/
"""
    inp = N.read_input(txt)
    assert inp.trj_input_file == "run.nc"
    assert inp.sp_types_selected == [1, 0]
    assert inp.ncfs_from_to == (10, 2, 20)
