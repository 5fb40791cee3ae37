"""GPU parity of the split-chemistry mode (mf_level_evolve,
mf_level_advance_chemistry, mf_level_evolve_split) against the oracle on the
same seeded inputs, to the 1e-11 bar of the coupled path."""
import numpy as np
import pytest

from mesoflow_amd import abi, cases, runtime
from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def _case(n=(16, 12, 12), strong=True, seed=21):
    c = cases.porous_reactor(n_cell=n, nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=2.0, rmax=3.5,
                             order_hyp=5, cfl=0.3, seed=seed)
    if strong:
        for r in range(2):
            c.plugin.log_A[r] = np.log(3.0e10) + 0.5 * r
            c.plugin.Ea[r] = 90.0e3
    return c


def _both(c, s0):
    from oracle.pyoracle import OracleLevel
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    L.set_state(s0)
    O = OracleLevel(c)
    O.set_state(s0)
    return ctx, L, O


def _compare(tag, L, O):
    got, ref = L.get_state(), O.get_state()
    err = rel_err(ref, got)
    gi, oi = L.info(), O.info()
    say("%s: max rel err %.2e  t %.9e/%.9e  steps %d/%d" % (tag, err.max(), gi[0], oi[0], gi[2], oi[2]))
    assert np.all(np.isfinite(got))
    assert err.max() <= TOL, err
    assert gi[2] == oi[2]
    assert abs(gi[0] - oi[0]) <= 1e-12 * abs(oi[0])


def test_chemistry_phase_after_set_state():
    """lag = the state itself; 2600 substeps run as three launches of <= 1024"""
    c = _case()
    ctx, L, O = _both(c, c.make_state())
    args = (0.0, 2600 * 1.0e-6, 1.0e-6)
    L.advance_chemistry(*args)
    O.advance_chemistry(*args)
    _compare("chemistry phase", L, O)
    L.close()
    ctx.close()


def test_evolve_only_flow_then_chemistry():
    c = _case()
    ctx, L, O = _both(c, c.make_state())
    dt = L.est_dt()
    n1 = L.evolve(3.5 * dt, only_flow=True)
    n2 = O.evolve(3.5 * dt, only_flow=True)
    assert n1 == n2 >= 3
    _compare("flow-only steps", L, O)
    # residual norms of a flow-only step: ||(U^{n+1} - U^n)/dt * vfrac|| (exact
    # zeros for the components the step leaves unchanged: vfrac, site species)
    r, g = O.info()[3], L.info()[3]
    assert r[abi.VFRAC] == 0.0 and g[abi.VFRAC] == 0.0
    np.testing.assert_allclose(g, r, rtol=1e-9, atol=0.0)
    # the chemistry phase's first stage reads the state before the last flow step
    L.advance_chemistry(0.0, 300 * 1.0e-6, 1.0e-6)
    O.advance_chemistry(0.0, 300 * 1.0e-6, 1.0e-6)
    _compare("chemistry after flow", L, O)
    L.close()
    ctx.close()


@pytest.mark.parametrize("mode", [abi.MODE_BATCHED, abi.MODE_GRAPH])
def test_evolve_split(mode):
    c = _case(seed=23)
    ctx, L, O = _both(c, c.make_state())
    dt = L.est_dt()
    args = (3.0 * dt, 2.0 * 400e-6, 400e-6, 2.0e-6, 2.0 * dt)
    L.evolve_split(*args, mode=mode)
    O.evolve_split(*args)
    _compare("evolve_split mode %d" % mode, L, O)
    L.close()
    ctx.close()


def test_chemistry_phase_refused_after_coupled_steps():
    c = _case()
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    L.set_state(c.make_state())
    L.step(1)
    with pytest.raises(abi.MFError) as e:
        L.advance_chemistry(0.0, 1e-5, 1e-6)
    assert e.value.status == abi.MF_ERR_UNSUPPORTED
    L.evolve(L.info()[0] * 3.0, only_flow=True)
    L.advance_chemistry(0.0, 1e-5, 1e-6)
    L.close()
    ctx.close()
