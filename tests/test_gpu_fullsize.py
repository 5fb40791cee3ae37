"""Full-size porous-reactor runs on the GPU, checked through properties that do
not depend on the domain size:

* window parity (tests/window.py): windows of the full-size GPU result match
  the oracle run on the windows grown by the dependence halo -- the 256^3
  level of 512 boxes with 5 species (2 steps) and the 512^3 level with 10
  species and 8 reactions (1 step), the 256^3 level of about 1200 uneven
  boxes on a GRF solid (1 step); windows at the inflow, at the outflow and
  inside, each block given a seeded perturbation;
* launch-mode invariance: the 256^3 state after two adaptive-dt steps is the
  same for the persistent batched launches, the per-box launches (the
  reference's launch shape) and the captured-graph replay.
"""
import numpy as np
import pytest

from mesoflow_amd import abi, cases, runtime
from helpers import rel_err
from window import perturb_block, sub_case, window_in_block, window_of, windows_for

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def _window_parity(case, nsteps, size, seed):
    from oracle.pyoracle import OracleLevel
    say("%s %s: making the state" % (case.name, case.n_cell))
    s0 = case.make_state()
    wins = windows_for(case.n_cell, nsteps, size, seed)
    for q, (_, blo, bhi) in enumerate(wins):
        perturb_block(s0, blo, bhi, seed=100 + q)
    # fixed dt: half the oracle's estimate on the first block
    sc0 = sub_case(case, s0, wins[0][1], wins[0][2])
    o = OracleLevel(sc0)
    o.set_state(sc0.make_state())
    case.params.fixed_dt = 1
    case.params.dt_fixed = 0.5 * o.est_dt()
    say("dt %.6e; GPU run of %d step(s) on %d boxes" % (case.params.dt_fixed, nsteps, len(case.boxes) or 1))
    ctx = runtime.Context(case)
    L = runtime.Level(ctx)
    L.set_state(s0)
    L.step(nsteps, abi.MODE_BATCHED)
    got = L.get_state()
    L.close()
    ctx.close()
    assert np.all(np.isfinite(got))
    say("GPU done; oracle on %d blocks" % len(wins))
    worst = 0.0
    for wlo, blo, bhi in wins:
        sc = sub_case(case, s0, blo, bhi)
        o = OracleLevel(sc)
        o.set_state(sc.make_state())
        o.step(nsteps)
        ref = window_in_block(o.get_state(), wlo, blo, size)
        err = rel_err(ref, window_of(got, wlo, size))
        say("window at %s (block %s..%s): max rel err %.2e" % (wlo, blo, bhi, err.max()))
        assert err.max() <= TOL, err
        worst = max(worst, err.max())
    return worst


@pytest.mark.timeout(600)
def test_porous256_ns5_window_parity():
    c = cases.porous_reactor(n_cell=(256, 256, 256), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=6.0,
                             rmax=14.0, order_hyp=5, cfl=0.3, seed=20261203, max_grid_size=32)
    _window_parity(c, 2, 24, seed=11)


@pytest.mark.timeout(900)
def test_porous512_ns10_window_parity():
    c = cases.porous_reactor(n_cell=(512, 512, 512), nspec=10, nreact=8, geometry="grf", porosity=0.45, corr_len=8.0,
                             order_hyp=5, cfl=0.3, seed=20261207, max_grid_size=64)
    _window_parity(c, 1, 16, seed=12)


@pytest.mark.timeout(600)
def test_uneven256_grf_window_parity():
    """the single-level stand-in of the many-box configuration: about 1200
    uneven boxes (edges 8..64), GRF solid at porosity 0.5"""
    boxes = cases.random_boxes((256, 256, 256), 8, 64, seed=7, p_stop=0.5)
    c = cases.porous_reactor(n_cell=(256, 256, 256), nspec=5, nreact=2, geometry="grf", porosity=0.5, corr_len=8.0,
                             order_hyp=5, cfl=0.3, seed=20261205, boxes=boxes)
    _window_parity(c, 1, 16, seed=13)


@pytest.mark.timeout(600)
def test_porous256_launch_modes_agree():
    n = (256, 256, 256)
    c = cases.porous_reactor(n_cell=n, nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=6.0, rmax=14.0,
                             order_hyp=5, cfl=0.3, seed=20261203, max_grid_size=32)
    say("launch modes: making the state")
    s0 = c.make_state()
    perturb_block(s0, (0, 0, 0), n, seed=5)
    ref = None
    for name, mode in (("batched", abi.MODE_BATCHED), ("per box", abi.MODE_PER_BOX), ("graph", abi.MODE_GRAPH)):
        ctx = runtime.Context(c)
        L = runtime.Level(ctx)
        L.set_state(s0)
        L.step(2, mode)
        got = L.get_state()
        t, dt, ist = L.info()[:3]
        L.close()
        ctx.close()
        say("%s: t %.9e dt %.9e" % (name, t, dt))
        assert np.all(np.isfinite(got))
        if ref is None:
            ref = (got, t, dt)
        else:
            err = rel_err(ref[0], got)
            say("   max rel diff to batched %.2e" % err.max())
            assert err.max() <= 1e-13, err
            assert abs(t - ref[1]) <= 1e-14 * ref[1] and abs(dt - ref[2]) <= 1e-14 * ref[2]
