"""Discrete adjoint on the GPU (mf_adjoint_forward / mf_adjoint_backward).

The objective J is the time-integrated outlet species mass flux of a fixed-dt
run of the porous reactor. The adjoint gradients are checked against
  * central finite differences of J computed by the same GPU forward,
  * central finite differences of J computed by the oracle (the CPU
    restatement of the reference solver), i.e. the reference gradient,
  * directional derivatives of J with respect to the initial state
    (dot-product test of the whole reverse sweep: faces, cut cells, ghost
    fill, chemistry, RK3 arithmetic), with one box and with uneven boxes,
and the binomial checkpointing must not change the result.

The runs start from the seeded initial state with a 2 % seeded perturbation of
the conserved components and species (primitives made consistent): in an
exactly uniform flow the LLF eigenvalue max(|v_L|, |v_R|), max(a_L, a_R) ties at
every face, where the discrete map is not differentiable (central differences
average both branches, the adjoint takes one).
"""
import numpy as np
import pytest
import torch

from mesoflow_amd import cases, runtime
from mesoflow_amd.autograd import OutletFluxObjective, run_objective, theta_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def adj_case(n_cell, geometry="spheres", nreact=3, seed=5, boxes=None, strong=True, nspec=5):
    c = cases.porous_reactor(n_cell=n_cell, nspec=nspec, nreact=nreact, geometry=geometry, porosity=0.6, rmin=2.0, rmax=3.5,
                             order_hyp=5, cfl=0.3, seed=seed, boxes=boxes)
    if strong:
        # fast chemistry so that the rate constants visibly change J within a few steps
        for r in range(nreact):
            c.plugin.log_A[r] = np.log(1.0e12) - 0.5 * r
            c.plugin.Ea[r] = 60.0e3 + 2.0e3 * r
    from oracle.pyoracle import OracleLevel
    o = OracleLevel(c)
    o.set_state(c.make_state())
    dt = o.est_dt()
    c.params.fixed_dt = 1
    c.params.dt_fixed = dt
    return c


def perturbed_state(case, seed=1, amp=0.02):
    s = case.make_state()
    rng = np.random.default_rng(seed)
    # conserved components and the gas species (the last species is the site)
    for q in list(range(5)) + list(range(12, 12 + case.nspec - 1)):
        s[q] *= 1.0 + amp * rng.uniform(-1.0, 1.0, s[q].shape)
    s[1:4] += amp * rng.uniform(-1.0, 1.0, s[1:4].shape) * s[0]
    s[6:9] = s[1:4] / s[0]
    s[1:4] *= s[11]
    s[6:9] *= s[11]
    return s


def outlet_J_oracle(case, nsteps, spec, state0=None):
    """J of the oracle: sum over steps of dt dy dz sum_{j,k} M_s c_s (rho u)/rho at i = nx-1."""
    from oracle.pyoracle import OracleLevel
    o = OracleLevel(case)
    o.set_state(case.make_state() if state0 is None else state0)
    dt = case.params.dt_fixed
    dx = case.dx
    M = case.plugin.molwt[spec]
    J = 0.0
    for _ in range(nsteps):
        o.step(1)
        s = o.get_state()
        J += dt * dx[1] * dx[2] * float(np.sum(M * s[12 + spec, :, :, -1] * s[1, :, :, -1] / s[0, :, :, -1]))
    return J


def gpu_level(case):
    ctx = runtime.Context(case)
    L = runtime.Level(ctx)
    return ctx, L


@pytest.mark.parametrize("variant", ["no_ib", "spheres", "spheres_uneven_boxes"])
def test_adjoint_initial_state_directional_derivative(variant):
    say("adjoint dot-product test", variant)
    n = (16, 8, 8)
    boxes = cases.random_boxes(n, 4, 8, seed=4) if variant == "spheres_uneven_boxes" else None
    c = adj_case(n, geometry="none" if variant == "no_ib" else "spheres", boxes=boxes, seed=13)
    nsteps, spec = 5, 1
    ctx, L = gpu_level(c)
    s0 = perturbed_state(c)
    theta = theta_of(c.plugin)
    J0, _ = run_objective(L, theta, s0, nsteps, 2, spec)
    lam = L.adjoint_state_grad()
    rng = np.random.default_rng(3)
    # relative step 1e-7: the WENO weights keep the map close to linear over it
    # (at 1e-5 the central difference itself is off by ~2e-5)
    d = rng.uniform(-1.0, 1.0, s0.shape) * np.abs(s0) * 1e-7
    d[11] = 0.0  # vfrac is geometry
    Jp, _ = run_objective(L, theta, s0 + d, nsteps, 2, spec, with_grad=False)
    Jm, _ = run_objective(L, theta, s0 - d, nsteps, 2, spec, with_grad=False)
    fd = (Jp - Jm) / 2.0
    ad = float(np.sum(lam * d))
    say("J %.12e  <dJ/dU0, d> %.12e  fd %.12e" % (J0, ad, fd))
    assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-18 * abs(J0), (ad, fd)
    L.close()
    ctx.close()


def test_adjoint_ns10_eight_reactions_directional_derivative():
    """the 10-species family (8 reactions): initial-state dot-product test and
    one rate-constant gradient against GPU central differences"""
    say("adjoint NS=10")
    c = adj_case((12, 8, 8), geometry="spheres", seed=19, nspec=10, nreact=8)
    nsteps, spec = 4, 3
    ctx, L = gpu_level(c)
    s0 = perturbed_state(c)
    theta = theta_of(c.plugin)
    J0, g = run_objective(L, theta, s0, nsteps, 2, spec)
    lam = L.adjoint_state_grad()
    rng = np.random.default_rng(8)
    d = rng.uniform(-1.0, 1.0, s0.shape) * np.abs(s0) * 1e-7
    d[11] = 0.0
    Jp, _ = run_objective(L, theta, s0 + d, nsteps, 2, spec, with_grad=False)
    Jm, _ = run_objective(L, theta, s0 - d, nsteps, 2, spec, with_grad=False)
    fd = (Jp - Jm) / 2.0
    ad = float(np.sum(lam * d))
    say("J %.12e  <dJ/dU0, d> %.12e  fd %.12e" % (J0, ad, fd))
    assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-18 * abs(J0), (ad, fd)
    p = int(np.argmax(np.abs(g[:8])))
    h = 1e-6 * abs(theta[p])
    tp, tm = theta.copy(), theta.copy()
    tp[p] += h
    tm[p] -= h
    Jp, _ = run_objective(L, tp, s0, nsteps, 2, spec, with_grad=False)
    Jm, _ = run_objective(L, tm, s0, nsteps, 2, spec, with_grad=False)
    fdp = (Jp - Jm) / (2 * h)
    say("log_A[%d]: adjoint %.10e fd %.10e" % (p, g[p], fdp))
    assert abs(g[p] - fdp) <= 1e-5 * abs(fdp), (p, g[p], fdp)
    L.close()
    ctx.close()


def test_adjoint_parameter_gradient_vs_gpu_finite_differences():
    say("adjoint parameter gradient vs GPU central differences")
    c = adj_case((16, 8, 8), geometry="spheres", seed=21)
    nsteps, spec = 8, 2
    ctx, L = gpu_level(c)
    s0 = perturbed_state(c)
    theta = theta_of(c.plugin)
    J, g = run_objective(L, theta, s0, nsteps, 3, spec)
    nr = int(c.plugin.nreact)
    for p in range(nr):
        h = 1e-6 * abs(theta[p])
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        Jp, _ = run_objective(L, tp, s0, nsteps, 3, spec, with_grad=False)
        Jm, _ = run_objective(L, tm, s0, nsteps, 3, spec, with_grad=False)
        fd = (Jp - Jm) / (2 * h)
        say("log_A[%d]: adjoint %.10e fd %.10e" % (p, g[p], fd))
        assert abs(g[p] - fd) <= 1e-5 * abs(fd), (p, g[p], fd)
    L.close()
    ctx.close()


def test_adjoint_inflow_velocity_gradient_vs_gpu_finite_differences():
    say("adjoint d/d(in_vel) vs GPU central differences")
    # short channel and enough steps for the inflow to reach the outlet
    c = adj_case((12, 4, 4), geometry="none", seed=2)
    nsteps, spec = 48, 0
    ctx, L = gpu_level(c)
    s0 = perturbed_state(c)
    theta = theta_of(c.plugin)
    J, g = run_objective(L, theta, s0, nsteps, 6, spec)
    p = int(c.plugin.nreact)
    h = 1e-6 * abs(theta[p])
    tp, tm = theta.copy(), theta.copy()
    tp[p] += h
    tm[p] -= h
    Jp, _ = run_objective(L, tp, s0, nsteps, 6, spec, with_grad=False)
    Jm, _ = run_objective(L, tm, s0, nsteps, 6, spec, with_grad=False)
    fd = (Jp - Jm) / (2 * h)
    say("in_vel: J %.10e adjoint %.10e fd %.10e" % (J, g[p], fd))
    assert abs(fd) > 1e-6 * abs(J)
    assert abs(g[p] - fd) <= 1e-5 * abs(fd), (g[p], fd)
    L.close()
    ctx.close()


def test_adjoint_matches_oracle_finite_differences():
    """Reference gradient: central differences of J computed by the CPU oracle."""
    say("adjoint vs oracle central differences")
    c = adj_case((12, 6, 6), geometry="spheres", seed=8)
    nsteps, spec = 6, 2
    ctx, L = gpu_level(c)
    s0 = perturbed_state(c)
    theta = theta_of(c.plugin)
    J, g = run_objective(L, theta, s0, nsteps, 2, spec)
    J_orc = outlet_J_oracle(c, nsteps, spec, s0)
    say("J gpu %.14e oracle %.14e" % (J, J_orc))
    assert abs(J - J_orc) <= 1e-11 * abs(J_orc)
    nr = int(c.plugin.nreact)
    base = c.plugin
    for p in range(nr + 1):
        h = 1e-6 * abs(theta[p])
        vals = []
        for sgn in (1, -1):
            t = theta.copy()
            t[p] += sgn * h
            from mesoflow_amd.autograd import plugin_with
            c.plugin = plugin_with(base, t)
            vals.append(outlet_J_oracle(c, nsteps, spec, s0))
        c.plugin = base
        fd = (vals[0] - vals[1]) / (2 * h)
        say("theta[%d]: adjoint %.10e oracle fd %.10e" % (p, g[p], fd))
        if p < nr:
            assert abs(g[p] - fd) <= 1e-5 * abs(fd), (p, g[p], fd)
        else:
            # six steps: the inflow barely reaches the outlet; compare absolutely
            assert abs(g[p] - fd) <= 1e-5 * abs(J) + 1e-5 * abs(fd), (p, g[p], fd)
    L.close()
    ctx.close()


def test_adjoint_checkpoint_count_does_not_change_the_gradient():
    say("adjoint checkpointing invariance")
    c = adj_case((16, 8, 8), geometry="spheres", seed=17)
    nsteps, spec = 9, 1
    ctx, L = gpu_level(c)
    s0 = perturbed_state(c)
    theta = theta_of(c.plugin)
    ref = None
    for ncheck in (9, 3, 1):
        J, g = run_objective(L, theta, s0, nsteps, ncheck, spec)
        say("ncheck %d: J %.14e grad %s" % (ncheck, J, np.array2string(g, precision=10)))
        if ref is None:
            ref = (J, g)
        else:
            # the forward is deterministic; the reverse sweep accumulates with
            # atomics, so its sums may round differently
            assert J == ref[0]
            np.testing.assert_allclose(g, ref[1], rtol=1e-9, atol=1e-12 * float(np.max(np.abs(ref[1]))))
    L.close()
    ctx.close()


def test_autograd_function_gradcheck():
    say("torch.autograd.Function gradcheck")
    c = adj_case((12, 6, 6), geometry="spheres", seed=31)
    nsteps, spec = 4, 2
    ctx, L = gpu_level(c)
    s0 = perturbed_state(c)
    theta = torch.tensor(theta_of(c.plugin), dtype=torch.float64, requires_grad=True)
    J = OutletFluxObjective.apply(theta, L, s0, nsteps, 2, spec)
    (g,) = torch.autograd.grad(J, theta)
    assert torch.all(torch.isfinite(g))
    f = lambda th: OutletFluxObjective.apply(th, L, s0, nsteps, 2, spec)
    assert torch.autograd.gradcheck(f, (theta,), eps=1e-6, atol=1e-12 * abs(float(J)) + 1e-20, rtol=1e-4)
    L.close()
    ctx.close()
