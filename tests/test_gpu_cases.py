"""GPU parity of the cases written against the plugin API (plugin families
compiled from mesoflow_amd/plugins/<case>/mflo_case.H) against the oracle,
which compiles the same case headers: SpeciesEqTests along x, y, z (species
only), Channel (AUSM+-up with the limiter, and WENO5), the background-inert-gas
reactor (coupled steps, and the split mode with mflo.using_bg_inertgas = 1).
Bar: norm-wise relative error per component <= 1e-11 after N steps."""
import numpy as np
import pytest

from mesoflow_amd import abi, cases, runtime
from helpers import rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def _check(case, nsteps, mode=abi.MODE_BATCHED, state=None, boxes=None, tol=TOL):
    s0 = case.make_state() if state is None else state
    ref, rinfo = run_oracle(case, nsteps, s0)
    got, ginfo = run_gpu(case, nsteps, mode=mode, state=s0, boxes=boxes)
    err = rel_err(ref, got)
    say("%s: %d steps, max rel err %.2e" % (case.name, nsteps, err.max()))
    assert np.all(np.isfinite(got))
    assert err.max() <= tol, err
    assert ginfo[2] == rinfo[2]
    assert abs(ginfo[0] - rinfo[0]) <= 1e-12 * abs(rinfo[0])
    return ref, got


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_species_eq_tests(direction):
    n = [4, 4, 4]
    n[direction] = 32
    c = cases.species_eq_tests(n_cell=tuple(n), direction=direction)
    _check(c, 60)
    mg = [4, 4, 4]
    mg[direction] = 8
    c.boxes = cases.chop_boxes(c.n_cell, mg)
    _check(c, 30, mode=abi.MODE_PER_BOX)


@pytest.mark.parametrize("order", [2, 5])
def test_channel(order):
    c = cases.channel(n_cell=(24, 16, 16), order_hyp=order, cfl=0.3)
    _check(c, 20)


@pytest.mark.parametrize("do_ns", [1, 0])
def test_bg_inert_gas_coupled_steps(do_ns):
    """the background-gas thermo, transport and inflow in the coupled step; the
    example's kinetics (desorption 1.7e10 /s) cannot be integrated explicitly
    at the flow's dt (~2e-9 s), so chemistry_on = 0 here and the split mode
    below runs them"""
    c = cases.bg_inert_gas(n_cell=(24, 16, 12), do_ns=do_ns, porosity=0.6, rmin=2.0, rmax=3.5, h2_initial=2.0)
    c.plugin.user[0] = 0.0
    c.boxes = cases.random_boxes(c.n_cell, 4, 12, seed=6)
    _check(c, 8, mode=abi.MODE_GRAPH)


def test_bg_inert_gas_split_mode():
    """Evolve_split with the background gas: flow-only steps, chemistry phases
    (store_inertgas_conc, midpoint substeps, update_vars_after_chemsolve with
    the background concentration), flow-only steps"""
    from oracle.pyoracle import OracleLevel
    c = cases.bg_inert_gas(n_cell=(20, 12, 12), porosity=0.6, rmin=2.0, rmax=3.0, h2_initial=5.0)
    s0 = c.make_state()
    o = OracleLevel(c)
    o.set_state(s0)
    dt = o.est_dt()
    t_ss, dt_ss = 3.0 * dt, 2.0 * dt
    args = (t_ss, 2.0e-12, 1.0e-12, 2.0e-13, dt_ss)
    o.evolve_split(*args)
    ref = o.get_state()
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    L.set_state(s0)
    L.evolve_split(*args)
    got = L.get_state()
    L.close()
    ctx.close()
    err = rel_err(ref, got)
    say("bg inert gas split mode: max rel err %.2e" % err.max())
    assert np.all(np.isfinite(got)) and err.max() <= TOL, err
    # the chemistry ran: adsorbed hydrogen appeared
    assert np.max(ref[abi.FLO_NVARS + 1]) > 0.0


def test_family_without_background_gas_refuses_the_flag():
    c = cases.species_eq_tests(n_cell=(16, 4, 4))
    c.params.using_bg_inertgas = 1
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    L.set_state(c.make_state())
    with pytest.raises(abi.MFError) as e:
        L.advance_chemistry(0.0, 1e-3, 1e-4)
    assert e.value.status == abi.MF_ERR_UNSUPPORTED
    L.close()
    ctx.close()
