"""Pins of the oracle's two-level driver (oracle/mf_hier_oracle.h), CPU only.

The driver restates mflo::EvolveAMR over levels with subcycling
(Source/mflo_evolve.cpp:18-89, timeStep :175-249), FillPatchTwoLevels per fine
box, the flux registers, Reflux and AverageDownTo. AMReX itself is absent, so
its orders are unpinned; these tests tie the driver to what is fixed:
  * one step equals, bit for bit, the same step composed from the oracle's
    per-box RHS and its coupling operators in the order of timeStep;
  * with periodic boundaries the conserved totals are unchanged to round-off,
    and without the reflux they are not (several fine boxes, one across the
    periodic boundary, and several coarse boxes);
  * a fine level over the whole periodic domain evolves exactly as the fine
    level alone with half the time step, and the coarse level is its average;
  * a regrid keeps the data: RemakeLevel copies the old fine data and
    interpolates the coarse level elsewhere, conservation holds across it, and
    ClearLevel returns to the single-level run.
"""
import numpy as np
import pytest

from compose_two_level import FHI, FLO, NC, R, compose_oracle_step, initial_state, periodic_case
from hier_cases import all_periodic, cons_comps, fine_init, shear_state
from mesoflow_amd import abi, cases
from oracle import pyoracle as po
from oracle.pyoracle import OracleHier, OracleLevel

def test_step_equals_the_per_box_composition():
    cc = periodic_case(NC)
    s0 = initial_state(cc)
    O = OracleLevel(cc)
    O.set_state(s0)
    dt = 0.5 * O.est_dt()
    want_c, want_f = compose_oracle_step(dt)
    c2 = periodic_case(NC)
    c2.params.do_chemistry = 0          # the composition's old coarse state is U^n
    c2.params.fixed_dt = 1
    c2.params.dt_fixed = dt
    flo, fhi = (FLO * R,) * 3, (FHI * R + R - 1,) * 3
    cs, fs = fine_init(s0, [(flo, fhi)])
    L0 = OracleLevel(c2)
    L0.set_state(cs)
    H = OracleHier(L0, R)
    H.add_level([(flo, fhi)], fs)
    H.step(1)
    got_c = H.get_state(0)
    got_f = H.get_state(1)[:, flo[2]:fhi[2] + 1, flo[1]:fhi[1] + 1, flo[0]:fhi[0] + 1]
    k = cons_comps(s0.shape[0])
    assert np.array_equal(got_c[k], want_c[k])
    assert np.array_equal(got_f[k], want_f[k])


def _conservation_run(reflux, nsteps=3):
    n = 16
    cboxes = cases.chop_boxes((n, n, n), 8)
    cc = all_periodic(n, cboxes)
    s0 = shear_state(cc)
    # three fine boxes: two sharing a face, one across the periodic x boundary
    fboxes = [((8, 8, 8), (15, 23, 19)), ((16, 8, 8), (21, 15, 19)), ((0, 20, 4), (5, 27, 11))]
    cs, fs = fine_init(s0, fboxes)
    L0 = OracleLevel(cc)
    L0.set_state(cs)
    H = OracleHier(L0, R)
    H.set_reflux(reflux)
    H.add_level(fboxes, fs)
    dv = (1.0 / n) ** 3
    k = cons_comps(s0.shape[0])
    m0 = cs[k].sum(axis=(1, 2, 3)) * dv
    H.step(nsteps)
    c1 = H.get_state(0)
    m1 = c1[k].sum(axis=(1, 2, 3)) * dv
    mag = np.abs(c1[k]).sum(axis=(1, 2, 3)) * dv
    return np.abs(m1 - m0) / np.where(mag > 0, mag, 1.0), H


def test_reflux_conserves_with_periodic_boundaries():
    rel, H = _conservation_run(True)
    rel0, _ = _conservation_run(False)
    assert H.nlevels() == 2
    assert np.all(rel <= 2e-14), rel
    assert rel0.max() > 1e3 * max(rel.max(), 1e-16), (rel, rel0)


def test_full_cover_equals_the_fine_level_alone():
    n = 8
    cc = all_periodic(n, cases.chop_boxes((n, n, n), 4))
    s0 = shear_state(cc)
    fboxes = [((0, 0, 0), (15, 15, 7)), ((0, 0, 8), (15, 15, 15))]
    cs, fs = fine_init(s0, fboxes)
    O = OracleLevel(cc)
    O.set_state(cs)
    dt = 0.5 * O.est_dt()
    cc.params.fixed_dt = 1
    cc.params.dt_fixed = dt
    L0 = OracleLevel(cc)
    L0.set_state(cs)
    H = OracleHier(L0, R)
    H.add_level(fboxes, fs)
    H.step(2)
    cf = all_periodic(2 * n, [((lo), (hi)) for lo, hi in fboxes])
    cf.params.fixed_dt = 1
    cf.params.dt_fixed = dt / 2
    F = OracleLevel(cf)
    F.set_state(fs)
    F.step(4)
    want = F.get_state()
    got = H.get_state(1)
    assert np.array_equal(got, want)
    # the coarse level is the average of the fine level
    avg = got.reshape(got.shape[0], n, 2, n, 2, n, 2).sum(axis=(2, 4, 6))
    c = H.get_state(0)
    k = cons_comps(c.shape[0])
    assert np.allclose(c[k], avg[k] / 8.0, rtol=1e-14, atol=1e-14)


def test_regrid_keeps_data_and_conservation():
    n = 16
    cc = all_periodic(n, cases.chop_boxes((n, n, n), 8))
    s0 = shear_state(cc)
    fb0 = [((8, 8, 8), (15, 23, 19))]
    cs, fs = fine_init(s0, fb0)
    L0 = OracleLevel(cc)
    L0.set_state(cs)
    H = OracleHier(L0, R)
    H.add_level(fb0, fs)
    H.step(1)
    old_f = H.get_state(1)
    crse = H.get_state(0)
    t = H.info()[0]
    fb1 = [((4, 8, 8), (15, 19, 19)), ((16, 8, 8), (23, 15, 15))]
    H.regrid(fb1, t)
    new_f = H.get_state(1)
    # old fine data where the old level covered, interpolation elsewhere
    pad = np.ascontiguousarray(np.pad(crse, ((0, 0), (1, 1), (1, 1), (1, 1)), mode="wrap"))
    for lo, hi in fb1:
        m = [hi[d] - lo[d] + 1 for d in range(3)]
        want = np.zeros((crse.shape[0], m[2], m[1], m[0]))
        po.interp_cons((lo, hi), (pad, (-1, -1, -1)), (want, lo), R)
        for olo, ohi in fb0:
            xl = [max(lo[d], olo[d]) for d in range(3)]
            xh = [min(hi[d], ohi[d]) for d in range(3)]
            if all(xh[d] >= xl[d] for d in range(3)):
                want[:, xl[2] - lo[2]:xh[2] - lo[2] + 1, xl[1] - lo[1]:xh[1] - lo[1] + 1,
                     xl[0] - lo[0]:xh[0] - lo[0] + 1] = old_f[:, xl[2]:xh[2] + 1, xl[1]:xh[1] + 1, xl[0]:xh[0] + 1]
        got = new_f[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        assert np.array_equal(got, want), (lo, hi)
    # conservation across steps after the regrid
    k = cons_comps(crse.shape[0])
    dv = (1.0 / n) ** 3
    m0 = crse[k].sum(axis=(1, 2, 3)) * dv
    H.step(2)
    c2 = H.get_state(0)
    mag = np.abs(c2[k]).sum(axis=(1, 2, 3)) * dv
    rel = np.abs(c2[k].sum(axis=(1, 2, 3)) * dv - m0) / np.where(mag > 0, mag, 1.0)
    assert np.all(rel <= 2e-14), rel


def test_regrid_without_boxes_clears_the_level():
    n = 8
    cc = all_periodic(n)
    s0 = shear_state(cc)
    fb0 = [((4, 4, 4), (11, 11, 11))]
    cs, fs = fine_init(s0, fb0)
    L0 = OracleLevel(cc)
    L0.set_state(cs)
    H = OracleHier(L0, R)
    H.add_level(fb0, fs)
    H.step(1, regrid_boxes=[])
    assert H.nlevels() == 1
    ref = OracleLevel(cc)
    ref.set_state(cs)
    ref.step(1)
    assert np.array_equal(H.get_state(0), ref.get_state())


def test_fine_boxes_must_be_aligned_and_disjoint():
    cc = all_periodic(8)
    L0 = OracleLevel(cc)
    L0.set_state(shear_state(cc))
    H = OracleHier(L0, R)
    f = np.zeros(H.fine_shape)
    with pytest.raises(abi.MFError):
        H.add_level([((1, 0, 0), (4, 3, 3))], f)
    with pytest.raises(abi.MFError):
        H.add_level([((0, 0, 0), (7, 7, 7)), ((6, 0, 0), (9, 7, 7))], f)
