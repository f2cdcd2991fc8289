"""The namelist reader on the reference's example inputs (copied as data into
tests/golden/inputs) and on the defaults of src/io/input.f90."""
import os

import pytest

from trj_analysis_amd import namelist as N

GOLD = os.path.join(os.path.dirname(__file__), "golden", "inputs")


def test_ljmix_input():
    inp = N.read_input_file(os.path.join(GOLD, "3D_LJmix.nml"))
    assert inp.ndim == 3 and inp.nsp == 2 and inp.nthread == 64
    assert inp.ncfs_from_to == (1000, 1, 1000)
    assert inp.flags == (True, True, False, False, False, False, True)
    assert inp.deltar == 0.25 and inp.rcrdf == 15.0
    assert inp.qmax == 25.0 and inp.qmin == 6.4
    assert inp.bsc == [1.0, 1.0]  # default (input.f90:106-107)
    assert inp.periodic == (True, True, True)


def test_meam_input_periodic_and_comments():
    inp = N.read_input_file(os.path.join(GOLD, "3D_meam-dyn.nml"))
    assert inp.periodic == (True, True, True)
    assert inp.rcrdf == 25.0 and inp.deltar == 0.25  # '! Cutoff in rdf' comment stripped
    assert inp.qmax == 5.5 and inp.qmin == 5.5


def test_2d_input_nonperiodic():
    inp = N.read_input_file(os.path.join(GOLD, "2D_salr.nml"))
    assert inp.ndim == 2 and inp.periodic == (False, False, False)
    assert inp.qmin == 0.2 and inp.rcrdf == 300.0


def test_groups_skipped_when_module_off():
    text = """&INPUT
  ndim=3
  nsp=2
  rdf_sq_cl_dyn_sqw_conf_ord=F T F F F F F
/
&INPUT_RDF
  deltar=0.1
  rcrdf=3.0
/
&INPUT_SQ
  qmax=12.0
  bsc=0.5 2*1.5
/
"""
    inp = N.read_input(text)
    assert inp.deltar == -1.0 and inp.rcrdf == -1.0  # /INPUT_RDF/ not read (input.f90:143-145)
    assert inp.qmax == 12.0 and inp.qmin == 12.0     # qmin defaults to qmax here
    assert inp.bsc == [0.5, 1.5]                     # repeat count, truncated to nsp
    assert inp.ncfs_from_to[1] == 1                  # input.f90:137-140


def test_repeat_and_logicals():
    g = N.parse_groups("&A\n x=3*0\n y=.t. .f. T\n s='run.nc'\n/\n")
    assert g["a"]["x"] == [0, 0, 0]
    assert g["a"]["y"] == [True, False, True]
    assert g["a"]["s"] == ["run.nc"]


def test_confinement_groups_of_the_conf_example():
    """examples/3D/Conf/input.nml: conf flag with RDF and S(Q) -> z-slices
    (input.f90:166-191); nslice < 1 switches to one automatic slice."""
    from trj_analysis_amd.namelist import read_input
    text = """&INPUT
  trj_input_file='run.nc'
  ndim=3
  nsp=7
  nslice=1
  ncfs_from_to=1000 1 1000
  rdf_sq_cl_dyn_sqw_conf_ord=T T F F F T F
/
&INPUT_SP
  sp_types_selected= 1 2 3 4 5 6 7
/
&INPUT_RDF
  deltar=0.25
  rcrdf=150.0
/
&INPUT_SQ
  qmax=8.0
  qmin=1.4
/
&INPUT_CONF
zslice=0.0
zgrid=4.0
/
"""
    inp = read_input(text)
    assert inp.slices and inp.nslice == 1 and inp.zslice == [0.0] and inp.zgrid == 4.0
    assert not inp.auto_zslice
    auto = read_input(text.replace("  nslice=1\n", ""))
    assert auto.auto_zslice and auto.nslice == 1
    plain = read_input(text.replace("T T F F F T F", "T T F F F F F"))
    assert not plain.slices
