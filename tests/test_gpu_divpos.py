"""The positive-denominator division of the device code (div_pos: reciprocal
estimate, two Newton-Raphson steps, residual correction) against the IEEE
division, over the denominator range the code feeds it (WENO weight sums
~1e-24..1e60, densities, heat capacities, concentration totals, transport
sums >= 1e-30): never more than one ulp apart, and almost always equal."""
import ctypes as C

import numpy as np
import pytest

from mesoflow_amd import cases, runtime

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lo,hi", [(-100, -60), (-60, 60), (60, 200)])
def test_div_pos_within_one_ulp(lo, hi):
    c = cases.porous_reactor(n_cell=(8, 8, 8), nspec=5, nreact=2, geometry="none", order_hyp=5)
    ctx = runtime.Context(c)
    counts = np.zeros(3, dtype=np.int64)
    n = 1 << 24
    runtime.check(runtime.lib().mf_util_div_check(ctx.h, n, lo, hi, counts.ctypes.data_as(C.c_void_p)), "mf_util_div_check")
    ctx.close()
    print("exponents [%d, %d]: equal %d, one ulp %d, more %d" % (lo, hi, counts[0], counts[1], counts[2]), flush=True)
    assert counts.sum() == n
    assert counts[2] == 0
    assert counts[0] >= 0.99 * n
