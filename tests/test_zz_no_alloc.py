"""pnx_forward_noalloc, the forward that allocates nothing (its intensity partial
sums live in a fixed device scratch array), against pnx_forward on the same
problems: the largest batch its scratch admits, then two smaller problems, the
no-allocation calls made back to back after the references, so every call
finds the previous call's contents in the scratch.  A problem beyond the
scratch is refused before anything is launched."""
import numpy as np
import pytest

from pnepma_amd import workloads
from pnepma_amd._lib import PnxError, check, dptr, lib
from pnepma_amd.problem import solve_inputs
from pnepma_amd.solver import DeviceProblem

from .helpers import assert_physical, rel_err, small_1d, small_2d

pytestmark = pytest.mark.gpu


def _noalloc(dp):
    out = np.zeros((dp.problem.n_beams, len(dp.problem.x_rays)))
    check(lib().pnx_forward_noalloc(dp.handle, dptr(out)))
    return out


def test_noalloc_forward_equals_forward():
    cases = [workloads.raster2d(n_steps=400, n_beams=32),     # 59 MB of the 64 MiB scratch
             small_2d(N=5, M=(16, 12), n_beams=3),
             small_1d(N=3, cells=30)]
    dps, refs = [], []
    for p, mat, _, det in cases:
        dp = DeviceProblem(p, solve_inputs(p, mat, det), adjoint=False)
        dps.append(dp)
        refs.append(dp.forward())
    outs = [_noalloc(dp) for dp in dps]          # back to back, each smaller than the one before
    outs.append(_noalloc(dps[0]))                # and the largest again after the smallest
    refs.append(refs[0])
    for k, (I, I_ref) in enumerate(zip(outs, refs)):
        e = rel_err(I, I_ref)
        print(f"call {k}: shape {I.shape}, rel err vs pnx_forward {e:.3e}", flush=True)
        assert_physical(I)
        assert e < 1e-10
    for dp in dps:
        dp.close()


def test_noalloc_refuses_a_problem_beyond_its_scratch():
    p, mat, _, det = workloads.raster2d(n_steps=400, n_beams=40)  # 74 MB of partial sums
    dp = DeviceProblem(p, solve_inputs(p, mat, det), adjoint=False)
    with pytest.raises(PnxError, match="scratch"):
        _noalloc(dp)
    I = dp.forward()                                              # the problem is untouched
    assert np.all(np.isfinite(I))
    dp.close()
