"""The C-ABI library loads here (no GPU) and exports every symbol that
include/trj_hip.h declares; host-only entry points work without a device."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "trj_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|int64_t|const char\*)\s+\**(trj_\w+)\s*\(",
                                 text, re.M)))


def test_header_symbols_exported():
    from trj_analysis_amd import trjhip as T
    L = T.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(T.EXPORTS) == names


def test_abi_version():
    from trj_analysis_amd import trjhip as T
    assert T.lib().trj_abi_version() == 1


def test_plan_rejects_bad_params():
    from trj_analysis_amd import trjhip as T
    import pytest
    with pytest.raises(T.TrjError):
        T.sq_plan(T.sq_params(4, 1, 5.0), [10, 10, 10])
    with pytest.raises(T.TrjError):
        T.sq_plan(T.sq_params(3, 1, -5.0), [10, 10, 10])
    with pytest.raises(T.TrjError):
        T.rdf_plan(T.rdf_params(3, 1, 0.0, 5.0), [10, 10, 10])


def test_halfspace_plan_counts():
    from trj_analysis_amd import trjhip as T
    p = T.sq_params(3, 3, 1e9, qset=T.TRJ_QSET_HALFSPACE, grid=(8, 8, 8))
    info, nq = T.sq_plan(p, [10.0, 10.0, 10.0])
    assert info.nvec_grid == 8 * 8 * 4 - 1
    assert nq.sum() == info.nvec_grid
    assert list(info.box_ext) == [8, 8, 4]


def test_no_gpu_context_fails_loudly():
    """Creating a context without a HIP device returns an error, never a CPU path."""
    import torch
    from trj_analysis_amd import trjhip as T
    if torch.cuda.is_available():
        return
    import pytest
    with pytest.raises(T.TrjError):
        T.Context(0)


def test_plan_large_qmax_is_clamped():
    """qmax/dq above the int range: half-space clamps to the grid's largest shell,
    the reference q-set reports a limit error instead of overflowing."""
    from trj_analysis_amd import trjhip as T
    import pytest
    L = (1e6 / 0.85) ** (1 / 3)
    info, nq = T.sq_plan(T.sq_params(3, 3, 1e9, qset=T.TRJ_QSET_HALFSPACE, grid=(64, 64, 64)), [L] * 3)
    assert info.nvec_grid == 64 * 64 * 32 - 1 and nq.sum() == info.nvec_grid and info.nqmax == 55
    with pytest.raises(T.TrjError) as e:
        T.sq_plan(T.sq_params(3, 1, 1e9), [L] * 3)
    assert e.value.code == T.TRJ_ELIMIT


def test_batch_layout_host_only():
    from trj_analysis_amd import trjhip as T
    fa = T.batch_frames([(1000, 2, [20.0] * 3, (1, 1, 1)), (5000, 4, [30.0] * 3, (1, 1, 0))])
    ls, nit, off, tot = T.batch_layout(T.batch_params(3, 0.01, 4.0), fa)
    assert list(ls) == [401, 401] and list(nit) == [3, 10]
    assert list(off) == [0, 3 * 401] and tot == 13 * 401
