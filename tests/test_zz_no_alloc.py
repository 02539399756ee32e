"""glide_enth_column_sweep_noalloc (scratch in a static device array, no
allocation) against glide_enth_column_sweep (library-allocated scratch): two
calls in a row through the C ABI, the second on a smaller grid, each compared
bitwise (state and report) with the allocating entry point on the same input.
"""
import ctypes

import numpy as np
import pytest
import torch

from tests.util import enth_case

pytestmark = pytest.mark.gpu

DT = 10.0 * 365.25 * 86400.0


def _sweep(entry, ny, nx, nz, seed, n_iter, rtol):
    from glide_amd import _lib
    from tests.test_parity_enthalpy_gpu import make_pair
    th, ops = make_pair(enth_case(ny, nx, nz, seed), nz)
    C = ops._column()
    cfg = ops.smoother_config
    cfg.relative_tolerance = rtol
    rep = _lib.EnthReport()
    _lib.call(entry, ctypes.byref(C), float(DT), int(n_iter), ctypes.byref(cfg.to_c(True)), ctypes.byref(rep),
              _lib.stream_handle())
    torch.cuda.synchronize()
    return ops.enthalpy_state.E.detach().cpu().numpy(), rep


@pytest.mark.parametrize("calls", [[(24, 20, 41, 3, 8, 1e-5), (9, 14, 21, 4, 5, 1e-3)]])
def test_noalloc_sweep_matches_allocating_sweep(calls):
    from glide_amd import _lib
    for ny, nx, nz, seed, n_iter, rtol in calls:
        print(f"noalloc sweep {ny}x{nx}x{nz} seed={seed} n_iter={n_iter}", flush=True)
        E_ref, rep_ref = _sweep("glide_enth_column_sweep", ny, nx, nz, seed, n_iter, rtol)
        E_na, rep_na = _sweep("glide_enth_column_sweep_noalloc", ny, nx, nz, seed, n_iter, rtol)
        assert np.array_equal(E_na, E_ref)
        assert rep_na.sweeps == rep_ref.sweeps
        assert rep_na.r0 == rep_ref.r0 and rep_na.r == rep_ref.r
        assert list(rep_na.norms[:rep_na.sweeps]) == list(rep_ref.norms[:rep_ref.sweeps])
        assert rep_na.sweeps >= 1
    # sizes beyond the scratch's sweep record are refused before any work
    lib = _lib.load()
    C = _lib.EnthColumn(ny=4, nx=4, nz=11, term_flags=7, dx=100.0)
    cfg = _lib.EnthSweepConfigC(1.0, 1.0, 1e-5, 1e-2, 3, 1)
    assert lib.glide_enth_column_sweep_noalloc(ctypes.byref(C), 1.0, _lib.ENTH_MAX_SWEEPS + 1, ctypes.byref(cfg),
                                               None, None) == -1
    assert "GLIDE_ENTH_MAX_SWEEPS" in _lib.last_error()
