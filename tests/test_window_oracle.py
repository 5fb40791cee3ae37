"""The halo radius behind window parity, pinned on the CPU: a window of an
oracle run equals, bit for bit, the oracle run on the window grown by
HALO_PER_STEP cells per step with periodic (or the true x) conditions at the
block's edge. The full-size GPU checks (test_gpu_fullsize.py) rely on it."""
import numpy as np
import pytest

from mesoflow_amd import cases
from window import HALO_PER_STEP, perturb_block, sub_case, window_in_block, window_of, windows_for


@pytest.mark.parametrize("nsteps,mgs", [(1, None), (2, None), (1, 16), (2, 16)])
def test_window_of_oracle_run_equals_oracle_on_grown_window(oracle, nsteps, mgs):
    n = 24 + 2 * HALO_PER_STEP * nsteps + 1
    c = cases.porous_reactor(n_cell=(n, n, n), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=3.0, rmax=5.0,
                             order_hyp=5, seed=41, max_grid_size=mgs)
    s0 = c.make_state()
    perturb_block(s0, (0, 0, 0), (n, n, n), seed=7)
    L = oracle.OracleLevel(c)
    L.set_state(s0)
    c.params.fixed_dt = 1
    c.params.dt_fixed = 0.5 * L.est_dt()
    L = oracle.OracleLevel(c)
    L.set_state(s0)
    L.step(nsteps)
    full = L.get_state()
    size = 6
    for wlo, blo, bhi in windows_for(c.n_cell, nsteps, size, seed=3):
        sc = sub_case(c, s0, blo, bhi)
        Ls = oracle.OracleLevel(sc)
        Ls.set_state(sc.make_state())
        Ls.step(nsteps)
        got = window_in_block(Ls.get_state(), wlo, blo, size)
        ref = window_of(full, wlo, size)
        assert np.array_equal(got, ref), (wlo, np.max(np.abs(got - ref)))
    if mgs:
        # the box layout matters (cut cells of other boxes are read before
        # their update): the same block as one box differs
        wlo, blo, bhi = windows_for(c.n_cell, nsteps, size, seed=3)[2]
        sc = sub_case(c, s0, blo, bhi)
        sc.boxes = []
        Ls = oracle.OracleLevel(sc)
        Ls.set_state(sc.make_state())
        Ls.step(nsteps)
        assert not np.array_equal(window_in_block(Ls.get_state(), wlo, blo, size), window_of(full, wlo, size))
    # a halo shorter than the WENO5 reach alone (3 cells per stage) must show the seam
    h = 9 * nsteps - 1
    wlo = (n // 2 - size // 2,) * 3
    blo = tuple(w - h for w in wlo)
    bhi = tuple(w + size + h for w in wlo)
    sc = sub_case(c, s0, blo, bhi)
    Ls = oracle.OracleLevel(sc)
    Ls.set_state(sc.make_state())
    Ls.step(nsteps)
    assert not np.array_equal(window_in_block(Ls.get_state(), wlo, blo, size), window_of(full, wlo, size))
