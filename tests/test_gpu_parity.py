"""GPU parity of the HIP level path against the CPU oracle (same seeded inputs).

Bar: norm-wise relative error per component <= 1e-11 after N steps (FP64).
"""
import numpy as np
import pytest

from mesoflow_amd import abi, cases
from helpers import rel_err, run_gpu, run_oracle

TOL = 1e-11
pytestmark = pytest.mark.gpu


def _check(case, nsteps, mode=abi.MODE_BATCHED, tile=None, boxes=None, tol=TOL):
    s0 = case.make_state()
    ref, rinfo = run_oracle(case, nsteps, s0)
    got, ginfo = run_gpu(case, nsteps, mode=mode, state=s0, tile=tile, boxes=boxes)
    err = rel_err(ref, got)
    print(case.name, "max rel err", err.max(), "per comp", np.array2string(err, precision=2))
    assert np.all(np.isfinite(got))
    assert err.max() <= tol, err
    # time, dt and step count follow the same ComputeDt sequence
    assert ginfo[2] == rinfo[2]
    assert abs(ginfo[0] - rinfo[0]) <= 1e-12 * abs(rinfo[0])
    assert abs(ginfo[1] - rinfo[1]) <= 1e-12 * abs(rinfo[1])
    return ref, got, rinfo, ginfo


@pytest.mark.parametrize("mode", [abi.MODE_PER_BOX, abi.MODE_BATCHED, abi.MODE_GRAPH])
def test_sod_weno5_modes(mode):
    c = cases.shocktube_n2(n_cell=(64, 8, 8), prob_hi=(1.0, 0.125, 0.125), order_hyp=5, cfl=0.5)
    _check(c, 12, mode=mode)


@pytest.mark.parametrize("order", [1, 2, 4])
def test_sod_other_schemes(order):
    c = cases.shocktube_n2(n_cell=(48, 4, 4), prob_hi=(1.0, 1.0 / 12, 1.0 / 12), order_hyp=order, cfl=0.5)
    _check(c, 10)


@pytest.mark.parametrize("direction", [1, 2])
def test_sod_other_directions(direction):
    n = [4, 4, 4]
    n[direction] = 48
    hi = [1.0 / 12] * 3
    hi[direction] = 1.0
    c = cases.shocktube_n2(n_cell=tuple(n), prob_hi=tuple(hi), order_hyp=5, cfl=0.5, direction=direction)
    _check(c, 10)


def test_sod_many_boxes():
    c = cases.shocktube_n2(n_cell=(64, 8, 8), prob_hi=(1.0, 0.125, 0.125), order_hyp=5, cfl=0.5, max_grid_size=(16, 4, 8))
    _check(c, 8, mode=abi.MODE_PER_BOX)
    _check(c, 8, mode=abi.MODE_BATCHED)


def test_porous_ib_chemistry_single_box():
    c = cases.porous_reactor(n_cell=(24, 20, 16), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=3, rmax=5,
                             order_hyp=5, seed=7)
    _check(c, 6)


def test_porous_ib_uneven_boxes():
    c = cases.porous_reactor(n_cell=(24, 20, 16), nspec=5, nreact=3, geometry="spheres", porosity=0.6, rmin=3, rmax=5,
                             order_hyp=5, seed=11)
    c.boxes = cases.random_boxes(c.n_cell, 4, 12, seed=3)
    _check(c, 5, mode=abi.MODE_PER_BOX)
    _check(c, 5, mode=abi.MODE_GRAPH)


@pytest.mark.parametrize("order", [1, 2, 4])
def test_porous_other_schemes(order):
    c = cases.porous_reactor(n_cell=(20, 16, 12), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=3, rmax=4,
                             order_hyp=order, seed=5)
    _check(c, 4)


def test_porous_ns10_grf():
    c = cases.porous_reactor(n_cell=(20, 16, 16), nspec=10, nreact=8, geometry="grf", porosity=0.45, corr_len=3.0,
                             order_hyp=5, seed=9)
    _check(c, 4)


def test_catalyst_cylinder():
    c = cases.catalyst_cylinder(n_cell=(48, 48, 4), order_hyp=5, cylrad=2.5e-3)
    _check(c, 5)


def test_species_only_mode():
    c = cases.porous_reactor(n_cell=(20, 16, 12), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=3, rmax=4,
                             order_hyp=5, seed=5)
    c.params.do_ns = 0
    _check(c, 4)


def test_residual_norms_match():
    """compute_residual_norms: ||(S_new - S_old)/dt * vfrac||_2. The species
    norms are differences of nearly equal states (the chemistry change of one
    step), so their error is bounded by the state error times the operand
    scale ||S_c||_2/dt, not by the norm itself."""
    c = cases.porous_reactor(n_cell=(20, 16, 12), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=3, rmax=4,
                             order_hyp=5, seed=5)
    ref, _, rinfo, ginfo = _check(c, 3)
    r, g = rinfo[3], ginfo[3]
    dt = rinfo[1]
    operand = np.sqrt((ref.reshape(ref.shape[0], -1) ** 2).sum(axis=1)) / dt
    assert np.all(np.abs(r - g) <= 1e-11 * operand + 1e-12 * np.abs(r)), (r, g)
    assert np.all(r[:5] == 0.0) and np.all(g[:5] == 0.0)   # conserved: chemistry leaves them unchanged


def test_porous_tiny_boxes():
    """boxes with edges of 1 to 3 cells (435 boxes): tiles cut to the box, face
    sweeps whose stencil reaches three boxes away, cut cells of other boxes
    read raw"""
    c = cases.porous_reactor(n_cell=(14, 9, 7), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=2, rmax=3,
                             order_hyp=5, seed=29)
    c.boxes = cases.random_boxes(c.n_cell, 1, 3, seed=5, p_stop=0.3)
    assert len(c.boxes) > 100
    _check(c, 3, mode=abi.MODE_BATCHED)
    _check(c, 2, mode=abi.MODE_PER_BOX)
