"""The named workloads at their full size and step count on the GPU.

* Sod 64^3, NS=1, WENO5, CFL 0.5, 100 steps: the GPU against the oracle on
  the whole domain (<= 1e-11 per component), one 64^3 box launched per box,
  eight 32^3 boxes as persistent launches and as a captured graph (without
  cut cells the result does not depend on the box layout).
* Porous reactor 256^3, NS=5, 2 reactions, sphere pack: all 50 steps in the
  batched, per-box and graph launch modes -- finite, positive pressure where
  the fluid is, volume fraction untouched, species non-negative, and the
  three modes equal (the 2-step window parity of test_gpu_fullsize.py ties
  this run to the oracle).
* Porous reactor 512^3, NS=10, 8 reactions, GRF solid: all 20 steps, batched
  against graph, same properties.
* Discrete adjoint at 128^3, NS=5, 3 reactions, sphere pack, 200 fixed-dt
  steps: the three log-A gradients against GPU central differences
  (h = 1e-6 relative) with every state stored and with 20 binomial
  checkpoints; the inflow-velocity gradient at 128^3 over enough steps for the
  inflow's acoustic front to reach the outlet.
"""
import time

import numpy as np
import pytest

from mesoflow_amd import abi, cases, runtime
from helpers import rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _run(case, nsteps, mode, s0, boxes=None):
    ctx = runtime.Context(case)
    L = runtime.Level(ctx, boxes)
    L.set_state(s0)
    t0 = time.time()
    L.step(nsteps, mode)
    out = L.get_state()
    info = L.info()
    say("   mode %d: %d steps in %.2f s, t %.9e dt %.6e" % (mode, nsteps, time.time() - t0, info[0], info[1]))
    L.close()
    ctx.close()
    return out, info


@pytest.mark.timeout(900)
def test_sod64_100_steps_against_oracle():
    from oracle.pyoracle import OracleLevel
    c = cases.shocktube_n2(n_cell=(64, 64, 64), prob_hi=(1.0, 1.0, 1.0), order_hyp=5, cfl=0.5)
    s0 = c.make_state()
    say("Sod 64^3, 100 steps: oracle ...")
    t0 = time.time()
    o = OracleLevel(c)
    o.set_state(s0)
    o.step(100)
    ref = o.get_state()
    rinfo = o.info()
    say("   oracle %.1f s, t %.9e" % (time.time() - t0, rinfo[0]))
    eight = cases.chop_boxes(c.n_cell, 32)
    for mode, boxes in ((abi.MODE_PER_BOX, None), (abi.MODE_BATCHED, eight), (abi.MODE_GRAPH, eight)):
        got, ginfo = _run(c, 100, mode, s0, boxes)
        err = rel_err(ref, got)
        say("   max rel err %.3e" % err.max())
        assert np.all(np.isfinite(got))
        assert err.max() <= 1e-11, err
        assert ginfo[2] == rinfo[2] == 100
        assert abs(ginfo[0] - rinfo[0]) <= 1e-12 * rinfo[0]


def _physical(case, s, s0):
    nsp = case.nspec
    fluid = s0[abi.VFRAC] == 1.0
    assert np.all(np.isfinite(s))
    assert np.all(s[abi.PRES][fluid] > 0.0)
    assert np.array_equal(s[abi.VFRAC], s0[abi.VFRAC])
    assert np.all(s[abi.FLO_NVARS:abi.FLO_NVARS + nsp] >= 0.0)


@pytest.mark.timeout(900)
def test_porous256_all_50_steps():
    c = cases.porous_reactor(n_cell=(256, 256, 256), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=6.0,
                             rmax=14.0, order_hyp=5, cfl=0.3, seed=20261203, max_grid_size=32)
    say("porous 256^3, 50 steps: making the state")
    s0 = c.make_state()
    ref = None
    for mode in (abi.MODE_BATCHED, abi.MODE_GRAPH, abi.MODE_PER_BOX):
        got, info = _run(c, 50, mode, s0)
        _physical(c, got, s0)
        assert info[2] == 50
        if ref is None:
            ref = (got, info)
            # the flow has moved: the solution is not the initial state
            assert rel_err(s0, got)[abi.RHOU] > 1e-6
        else:
            err = rel_err(ref[0], got)
            say("   max rel diff to batched %.2e" % err.max())
            assert err.max() <= 1e-13, err
            assert info[0] == ref[1][0] and info[1] == ref[1][1]
        del got


@pytest.mark.timeout(1100)
def test_porous512_ns10_all_20_steps():
    c = cases.porous_reactor(n_cell=(512, 512, 512), nspec=10, nreact=8, geometry="grf", porosity=0.45, corr_len=8.0,
                             order_hyp=5, cfl=0.3, seed=20261207, max_grid_size=64)
    say("porous 512^3 NS=10, 20 steps: making the state")
    s0 = c.make_state()
    a, ia = _run(c, 20, abi.MODE_BATCHED, s0)
    _physical(c, a, s0)
    b, ib = _run(c, 20, abi.MODE_GRAPH, s0)
    worst = 0.0
    for q in range(c.nv):
        d = np.max(np.abs(a[q] - b[q]))
        sc = np.max(np.abs(a[q]))
        worst = max(worst, d / sc if sc > 0 else d)
    say("   max rel diff graph vs batched %.2e" % worst)
    assert worst <= 1e-13
    assert ia[:3] == ib[:3] and ia[2] == 20


# ---------------------------------------------------------------------------
# discrete adjoint at 128^3
# ---------------------------------------------------------------------------
def _adj_case():
    return cases.porous_adjoint()


def _perturbed(case, seed=1, amp=0.02):
    return cases.perturbed_state(case, seed, amp)


def _fixed_dt(case, s0):
    ctx = runtime.Context(case)
    L = runtime.Level(ctx)
    L.set_state(s0)
    dt = L.est_dt()
    L.close()
    ctx.close()
    case.params.fixed_dt = 1
    case.params.dt_fixed = dt
    return dt


@pytest.mark.timeout(1100)
def test_adjoint128_200_steps_log_A_gradient():
    from mesoflow_amd.autograd import run_objective, theta_of
    c = _adj_case()
    say("adjoint 128^3, 200 steps: making the state")
    s0 = _perturbed(c)
    dt = _fixed_dt(c, s0)
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    theta = theta_of(c.plugin)
    nsteps, spec = 200, 1
    grads = {}
    for ncheck in (200, 20):
        t0 = time.time()
        J, g = run_objective(L, theta, s0, nsteps, ncheck, spec)
        say("   ncheck %d: J %.14e grad %s (%.1f s)" % (ncheck, J, np.array2string(g, precision=8), time.time() - t0))
        grads[ncheck] = (J, g)
    J, g = grads[200]
    assert grads[20][0] == J
    np.testing.assert_allclose(grads[20][1], g, rtol=1e-9, atol=1e-12 * float(np.max(np.abs(g))))
    assert abs(J) > 0.0
    # the objective responds to the kinetics: the log-A gradients are not a numerical tail
    assert np.max(np.abs(g[:3])) > 1e-3 * abs(J), g
    def fd_at(p, h):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        Jp, _ = run_objective(L, tp, s0, nsteps, 200, spec, with_grad=False)
        Jm, _ = run_objective(L, tm, s0, nsteps, 200, spec, with_grad=False)
        return (Jp - Jm) / (2 * h)

    for p in range(3):
        # central differences at h = 1e-6 relative carry an O(h^2) truncation
        # error: the rate constants enter through exp(log A) with k t ~ 10, so
        # the third derivative times h^2 / 6 is about 1e-5 of the gradient;
        # Richardson extrapolation of h and h/2 removes that term
        h = 1e-6 * abs(theta[p])
        fd = fd_at(p, h)
        fd2 = fd_at(p, 0.5 * h)
        fdr = (4.0 * fd2 - fd) / 3.0
        say("   log_A[%d]: adjoint %.10e fd(h) %.10e fd(h/2) %.10e Richardson %.10e" % (p, g[p], fd, fd2, fdr))
        assert abs(g[p] - fd) <= 5e-5 * abs(fd), (p, g[p], fd)
        assert abs(g[p] - fdr) <= 2e-6 * abs(fdr), (p, g[p], fdr)
    # 200 steps at CFL 0.3 carry the inflow's acoustic front about 60 cells: the
    # outlet does not see the inflow velocity, in both the adjoint and the differences
    p = 3
    h = 1e-6 * abs(theta[p])
    tp, tm = theta.copy(), theta.copy()
    tp[p] += h
    tm[p] -= h
    Jp, _ = run_objective(L, tp, s0, nsteps, 200, spec, with_grad=False)
    Jm, _ = run_objective(L, tm, s0, nsteps, 200, spec, with_grad=False)
    fd = (Jp - Jm) / (2 * h)
    say("   in_vel: adjoint %.6e fd %.6e (dt %.3e)" % (g[p], fd, dt))
    assert abs(g[p]) <= 1e-12 * abs(J) / abs(theta[p])
    assert abs(fd) <= 1e-6 * abs(J) / abs(theta[p])
    L.close()
    ctx.close()


@pytest.mark.timeout(1100)
def test_adjoint128_inflow_velocity_gradient():
    """600 fixed-dt steps at CFL 0.3 (the inflow front passes x = 127 after
    about 430 steps), 30 binomial checkpoints."""
    from mesoflow_amd.autograd import run_objective, theta_of
    c = _adj_case()
    say("adjoint 128^3, 600 steps, d/d(in_vel): making the state")
    s0 = _perturbed(c, seed=2)
    _fixed_dt(c, s0)
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    theta = theta_of(c.plugin)
    nsteps, spec, ncheck = 600, 0, 30
    t0 = time.time()
    J, g = run_objective(L, theta, s0, nsteps, ncheck, spec)
    say("   J %.14e grad %s (%.1f s)" % (J, np.array2string(g, precision=8), time.time() - t0))
    p = 3
    h = 1e-6 * abs(theta[p])
    tp, tm = theta.copy(), theta.copy()
    tp[p] += h
    tm[p] -= h
    Jp, _ = run_objective(L, tp, s0, nsteps, ncheck, spec, with_grad=False)
    Jm, _ = run_objective(L, tm, s0, nsteps, ncheck, spec, with_grad=False)
    fd = (Jp - Jm) / (2 * h)
    say("   in_vel: adjoint %.10e fd %.10e" % (g[p], fd))
    assert abs(fd) > 1e-6 * abs(J)
    assert abs(g[p] - fd) <= 1e-5 * abs(fd), (g[p], fd)
    L.close()
    ctx.close()
