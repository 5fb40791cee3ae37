"""Pins of the oracle's split-chemistry mode (Source/mflo_splitchem.cpp).

* Chemistry phase with one first-order gas reaction 1 -> 2 at constant
  temperature: the oracle must reproduce, bit for bit, the recurrence of the
  reference's Advance_chemistry (:172-195) written out here in the same
  operation order -- an explicit midpoint rule whose first stage is
  evaluated on the state before the previous substep (phi_new after the
  swap), not on the current one. The lag is visible: the result differs from
  the plain midpoint rule. Species 2 gains what species 1 loses.
* update_vars_after_chemsolve (:74-140): rho = sum over the gas species of
  c_s M_s, momentum = rho * velocity, rhoE unchanged, T = p / (R sum c_gas).
* EvolveAMR(final_time, only_flow) (mflo_evolve.cpp:18-89): steps while
  t < final_time; the residual norms of a flow-only step are
  ||(U^{n+1} - U^n)/dt * vfrac|| (no chemistry swap).
"""
import math

import numpy as np

from mesoflow_amd import abi, cases


def _one_reaction_case(n=6):
    c = cases.porous_reactor(n_cell=(n, n, n), nspec=5, nreact=1, geometry="none", order_hyp=5, seed=3)
    t = c.plugin
    for s in range(5):
        t.stoich[0][s] = 0.0
    t.stoich[0][1] = -1.0
    t.stoich[0][2] = 1.0
    t.log_A[0] = math.log(3.0e10)
    t.Ea[0] = 90.0e3
    return c


def _lagged_midpoint(y, ylag, k, dt, nsub):
    for _ in range(nsub):
        d1 = 0.0 + (0.0 + (-1.0) * (k * ylag))
        s1 = 1.0 * y + (0.5 * dt) * d1
        d2 = 0.0 + (0.0 + (-1.0) * (k * s1))
        ylag, y = y, 1.0 * y + dt * d2
    return y


def _substeps(dt_react, dt_rk):
    n, t = 0, 0.0
    while t < dt_react:
        n += 1
        t += dt_rk
    return n


def test_chemistry_phase_is_the_lagged_midpoint_rule(oracle):
    c = _one_reaction_case()
    s0 = c.make_state()
    L = oracle.OracleLevel(c)
    L.set_state(s0)
    dt_rk = 2.0e-5
    dt_react = 37 * dt_rk
    nsub = _substeps(dt_react, dt_rk)
    L.advance_chemistry(0.0, dt_react, dt_rk)
    s1 = L.get_state()
    q = (0, 2, 3, 4)
    T = s0[abi.TEMP][q[1:]]
    k = math.exp(c.plugin.log_A[0] - c.plugin.Ea[0] / (c.plugin.R_univ * T))
    y0 = s0[abi.FLO_NVARS + 1][q[1:]]
    assert 0.05 < k * dt_rk < 0.5
    want = _lagged_midpoint(y0, y0, k, dt_rk, nsub)
    got = s1[abi.FLO_NVARS + 1][q[1:]]
    assert got == want, (got, want)
    # the lag matters: the plain midpoint rule gives another value
    plain = y0
    for _ in range(nsub):
        plain = plain + dt_rk * (-k * (plain + 0.5 * dt_rk * (-k * plain)))
    assert abs(plain - got) > 1e-6 * abs(got)
    # species 2 gains what species 1 loses; the others are untouched
    tot0 = s0[abi.FLO_NVARS + 1] + s0[abi.FLO_NVARS + 2]
    tot1 = s1[abi.FLO_NVARS + 1] + s1[abi.FLO_NVARS + 2]
    np.testing.assert_allclose(tot1, tot0, rtol=1e-14)
    for s in (0, 3, 4):
        assert np.array_equal(s1[abi.FLO_NVARS + s], s0[abi.FLO_NVARS + s])


def test_update_vars_after_chemsolve(oracle):
    c = _one_reaction_case()
    s0 = c.make_state()
    rng = np.random.default_rng(2)
    s0[abi.VELY] = 0.1 * rng.standard_normal(s0[abi.VELY].shape)
    L = oracle.OracleLevel(c)
    L.set_state(s0)
    L.advance_chemistry(0.0, 10 * 2.0e-5, 2.0e-5)
    s1 = L.get_state()
    ng = c.plugin.ngas
    rho = np.zeros_like(s1[0])
    for s in range(ng):
        rho = rho + s1[abi.FLO_NVARS + s] * c.plugin.molwt[s]
    assert np.array_equal(s1[abi.RHO], rho) and np.array_equal(s1[abi.DENS], rho)
    for a in range(3):
        assert np.array_equal(s1[abi.RHOU + a], rho * s0[abi.VELX + a])
        np.testing.assert_allclose(s1[abi.VELX + a], s0[abi.VELX + a], rtol=1e-14, atol=1e-300)
    assert np.array_equal(s1[abi.RHOE], s0[abi.RHOE])
    csum = np.zeros_like(rho)
    for s in range(ng):
        csum = csum + s1[abi.FLO_NVARS + s]
    np.testing.assert_allclose(s1[abi.TEMP], s1[abi.PRES] / (csum * c.plugin.R_univ), rtol=1e-14)
    # the reaction changed the mixture: pressure follows rhoE at the new composition
    assert not np.allclose(s1[abi.PRES], s0[abi.PRES], rtol=1e-12, atol=0.0)


def test_evolve_only_flow_steps_and_residuals(oracle):
    c = cases.porous_reactor(n_cell=(12, 8, 8), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=2.0, rmax=3.0,
                             order_hyp=5, seed=9)
    c.params.fixed_dt = 1
    c.params.dt_fixed = 1.0e-9
    L = oracle.OracleLevel(c)
    L.set_state(c.make_state())
    assert L.evolve(3.5e-9, only_flow=True) == 4
    t, dt, ist, res = L.info()
    assert ist == 4 and abs(t - 4.0e-9) < 1e-23
    before = L.get_state()
    assert L.evolve(5.0e-9, only_flow=True) == 1
    after = L.get_state()
    _, dt, _, res = L.info()
    vf = before[abi.VFRAC]
    for q in range(after.shape[0]):
        d = (1.0 / dt) * after[q] + (-1.0 / dt) * before[q]
        if abi.RHO <= q < abi.VFRAC:
            d = d * vf
        np.testing.assert_allclose(res[q], np.sqrt(np.sum(d * d)), rtol=1e-12, atol=1e-300)
    # max_step bounds the loop
    assert L.evolve(1.0, max_step=7, only_flow=True) == 2
