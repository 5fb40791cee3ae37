"""The two-level driver on the GPU (mf_hier_*) against the oracle's driver
(oracle/mf_hier_oracle.h): EvolveAMR over levels 0 and 1 with subcycling,
FillPatchTwoLevels, flux registers, Reflux, AverageDownTo and regrid
(Source/mflo_evolve.cpp:18-89, 175-318, 655-681; Source/mflo_gridwork.cpp:
21-193). Bar: norm-wise relative error per component <= 1e-11 on both
levels (the fine level over the cells of its boxes), the same time and step
count; conservation to round-off with periodic boundaries."""
import numpy as np
import pytest

from helpers import rel_err
from hier_cases import all_periodic, cons_comps, fine_init, fine_mask, shear_state
from mesoflow_amd import abi, amr, cases, runtime
from oracle.pyoracle import OracleHier, OracleLevel

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def _pair(case, fboxes, cstate, fstate):
    """(oracle hierarchy, GPU context, level 0, hierarchy) on the same data"""
    O0 = OracleLevel(case)
    O0.set_state(cstate)
    OH = OracleHier(O0, 2)
    OH.add_level(fboxes, fstate)
    ctx = runtime.Context(case)
    L0 = runtime.Level(ctx)
    L0.set_state(cstate)
    H = runtime.Hierarchy(L0, 2)
    H.add_level(fboxes, fstate)
    return (O0, OH), (ctx, L0, H)


def _compare(OH, H, fboxes, where):
    oc, gc = OH.get_state(0), H.get_state(0)
    of, gf = OH.get_state(1), H.get_state(1)
    m = fine_mask(of.shape, fboxes)
    ec = rel_err(oc, gc)
    ef = rel_err(of[:, m], gf[:, m])
    ot, odt, ois, ores = OH.info()
    gt, gdt, gis, gres = H.info()
    say("%s: coarse %.2e, fine %.2e, t %.6e / %.6e, steps %d / %d" % (where, ec.max(), ef.max(), ot, gt, ois, gis))
    assert np.all(np.isfinite(gc)) and np.all(np.isfinite(gf[:, m]))
    assert ec.max() <= TOL, ec
    assert ef.max() <= TOL, ef
    assert gis == ois
    assert abs(gt - ot) <= 1e-12 * abs(ot)
    # residual norms of level 0: sums of squares of (S_new - S_old) / dt; a
    # component that changes only by rounding (vfrac under average_down) has a
    # norm of noise, so the bar is relative to the norm or to that noise floor
    floor = 1e-12 * np.abs(oc).reshape(oc.shape[0], -1).max(axis=1) * np.sqrt(oc[0].size) / odt
    assert np.all(np.abs(gres - ores) <= 1e-9 * np.abs(ores) + floor), (ores, gres)


def _close(ctx, L0, H):
    H.close()
    L0.close()
    ctx.close()


@pytest.mark.parametrize("mode", [abi.MODE_BATCHED, abi.MODE_GRAPH, abi.MODE_PER_BOX])
def test_periodic_hierarchy_matches_the_oracle(mode):
    n = 16
    cc = all_periodic(n, cases.chop_boxes((n, n, n), 8))
    s0 = shear_state(cc)
    # two fine boxes sharing a face, one across the periodic x boundary
    fboxes = [((8, 8, 8), (15, 23, 19)), ((16, 8, 8), (21, 15, 19)), ((0, 20, 4), (5, 27, 11))]
    cs, fs = fine_init(s0, fboxes)
    (O0, OH), (ctx, L0, H) = _pair(cc, fboxes, cs, fs)
    OH.step(4)
    H.step(4, mode)
    _compare(OH, H, fboxes, "periodic Sod + shear, 4 steps, mode %d" % mode)
    _close(ctx, L0, H)


def test_conservation_with_periodic_boundaries():
    n = 16
    cc = all_periodic(n, cases.chop_boxes((n, n, n), 8))
    s0 = shear_state(cc)
    fboxes = [((8, 8, 8), (15, 23, 19)), ((16, 8, 8), (21, 15, 19)), ((0, 20, 4), (5, 27, 11))]
    cs, fs = fine_init(s0, fboxes)
    k = cons_comps(cs.shape[0])
    dv = (1.0 / n) ** 3
    out = []
    for reflux in (True, False):
        ctx = runtime.Context(cc)
        L0 = runtime.Level(ctx)
        L0.set_state(cs)
        H = runtime.Hierarchy(L0, 2)
        H.set_reflux(reflux)
        H.add_level(fboxes, fs)
        H.step(5)
        c1 = H.get_state(0)
        mag = np.abs(c1[k]).sum(axis=(1, 2, 3)) * dv
        out.append(np.abs(c1[k].sum(axis=(1, 2, 3)) * dv - cs[k].sum(axis=(1, 2, 3)) * dv) / np.where(mag > 0, mag, 1))
        _close(ctx, L0, H)
    say("relative change of the totals: with reflux %.2e, without %.2e" % (out[0].max(), out[1].max()))
    assert np.all(out[0] <= 1e-13)
    assert out[1].max() > 1e3 * max(out[0].max(), 1e-16)


def test_walls_inflow_and_chemistry_flag():
    """Sod along x (ext_dir inflow at x-lo, wall at x-hi, periodic y/z) with
    fine boxes against both x boundaries: the coarse patch and the fine
    ghost cells take the physical boundary rules; do_chemistry = 1 makes the
    coarse level's old data of the fine fills the post-flow state"""
    cc = cases.shocktube_n2(n_cell=(24, 8, 8), order_hyp=5, cfl=0.4, max_grid_size=8)
    s0 = shear_state(cc)
    fboxes = [((0, 0, 0), (11, 15, 7)), ((36, 4, 4), (47, 11, 15)), ((16, 0, 8), (27, 7, 15))]
    cs, fs = fine_init(s0, fboxes, periodic=(0, 1, 1))
    (O0, OH), (ctx, L0, H) = _pair(cc, fboxes, cs, fs)
    OH.step(4)
    H.step(4)
    _compare(OH, H, fboxes, "Sod with walls, 4 steps")
    _close(ctx, L0, H)


def test_porous_reactor_cut_cells_and_chemistry():
    """the 5-species Arrhenius reactor with a sphere pack: cut cells on both
    levels (the fine level's geometry from its interpolated ghost vfrac),
    inflow / outflow, chemistry on the coarse step and every fine substep"""
    cc = cases.porous_reactor(n_cell=(24, 16, 16), nspec=5, porosity=0.6, rmin=2.0, rmax=3.5, max_grid_size=8)
    s0 = cc.make_state()
    fboxes = [((0, 8, 8), (15, 23, 23)), ((16, 0, 8), (31, 15, 23)), ((32, 16, 0), (47, 31, 15))]
    cs, fs = fine_init(s0, fboxes, periodic=(0, 1, 1))
    (O0, OH), (ctx, L0, H) = _pair(cc, fboxes, cs, fs)
    OH.step(3)
    H.step(3)
    _compare(OH, H, fboxes, "porous reactor, 3 steps")
    _close(ctx, L0, H)


def _oracle_regrid_boxes(case, O0, t, crit, bf, mgs, eff):
    """the regrid's boxes from the oracle's level 0: ErrorEst on the filled
    state (orc_box_tag) and the tags-to-boxes step"""
    from oracle import pyoracle as po
    phi, grad, comps = crit
    nx, ny, nz = case.n_cell
    sb = np.ascontiguousarray(O0.fill_ghosts(t))
    tags = np.zeros((nz, ny, nx), dtype=np.uint8)
    po.box_tag(((0, 0, 0), (nx - 1, ny - 1, nz - 1)), (sb, (-3, -3, -3)), tags, phi, grad, comps)
    per = tuple(int(case.geom.is_periodic[d]) for d in range(3))
    return amr.tags_to_boxes(tags, 1, bf, mgs, eff, 2, periodic=per)


def test_regrid_cadence_matches_the_oracle():
    """regrid_int = 2: at coarse steps 2 and 4 the tags of level 0 remake
    level 1 (old fine data where the boxes overlap, interpolation elsewhere)"""
    n = 16
    cc = all_periodic(n, cases.chop_boxes((n, n, n), 8))
    s0 = shear_state(cc, amp=0.3)
    fb0 = [((8, 8, 8), (23, 23, 23))]
    cs, fs = fine_init(s0, fb0)
    crit = ([1e30], [0.05], [abi.DENS])
    bf, mgs, eff = 4, 16, 0.7
    (O0, OH), (ctx, L0, H) = _pair(cc, fb0, cs, fs)
    H.set_regrid(2, crit[0], crit[1], crit[2], n_error_buf=1, blocking_factor=bf, max_grid_size=mgs, grid_eff=eff)
    H.step(5)
    boxes = None
    for s in range(5):
        t, _, ist, _ = OH.info()
        if ist > 0 and ist % 2 == 0:
            boxes = _oracle_regrid_boxes(cc, O0, t, crit, bf, mgs, eff)
            OH.step(1, regrid_boxes=boxes)
        else:
            OH.step(1)
    got_boxes = H.boxes(1)
    say("regrid: %d fine boxes" % len(got_boxes))
    assert sorted(got_boxes) == sorted(tuple(map(tuple, b)) for b in boxes)
    _compare(OH, H, got_boxes, "regrid every 2 steps, 5 steps")
    _close(ctx, L0, H)


def test_level0_is_stepped_through_the_hierarchy():
    cc = all_periodic(8)
    ctx = runtime.Context(cc)
    L0 = runtime.Level(ctx)
    L0.set_state(shear_state(cc))
    H = runtime.Hierarchy(L0, 2)
    with pytest.raises(abi.MFError) as e:
        L0.step(1)
    assert e.value.status == abi.MF_ERR_UNSUPPORTED
    assert runtime.lib().mf_level_destroy(L0.h) == abi.MF_ERR_BAD_ARG
    f = np.zeros(H.fine_shape)
    with pytest.raises(abi.MFError):
        H.add_level([((1, 0, 0), (4, 3, 3))], f)          # not aligned to the ratio
    with pytest.raises(abi.MFError):
        H.add_level([((0, 0, 0), (7, 7, 7)), ((6, 0, 0), (9, 7, 7))], f)   # overlapping
    H.close()
    L0.step(1)                                           # a plain level again
    L0.close()
    ctx.close()


def _tagged_boxes_gpu(L0, time=0.0, **grid):
    tg = cases.AMR_TAGS
    tags, _ = amr.level_tag(L0, time, tg["refine_phi"], tg["refine_phigrad"], tg["comps"])
    return amr.tags_to_boxes(tags, grid["n_error_buf"], grid["blocking_factor"], grid["max_grid_size"],
                             grid["grid_eff"], 2, periodic=(0, 1, 1))


def test_grf_two_level_reactor_matches_the_oracle():
    """the two-level porous reactor (GRF solid, uneven coarse boxes, the fine
    level made from the coarse one where vfrac jumps, MakeNewLevelFromCoarse)
    at 32^3, with a regrid every 2 steps, against the oracle"""
    cc = cases.porous_amr(n_cell=(32, 32, 32), edge=(8, 16), corr_len=4.0)
    s0 = cc.make_state()
    grid = dict(cases.AMR_GRID, max_grid_size=16)
    tg = cases.AMR_TAGS
    crit = (tg["refine_phi"], tg["refine_phigrad"], tg["comps"])
    O0 = OracleLevel(cc)
    O0.set_state(s0)
    OH = OracleHier(O0, 2)
    ctx = runtime.Context(cc)
    L0 = runtime.Level(ctx)
    L0.set_state(s0)
    fb = _tagged_boxes_gpu(L0, **grid)
    ob = _oracle_regrid_boxes(cc, O0, 0.0, crit, grid["blocking_factor"], grid["max_grid_size"], grid["grid_eff"])
    assert sorted(fb) == sorted(ob) and fb
    H = runtime.Hierarchy(L0, 2)
    H.regrid(fb)
    OH.regrid(ob, 0.0)
    H.set_regrid(2, *crit, **grid)
    H.step(4)
    for s in range(4):
        t, _, ist, _ = OH.info()
        if ist > 0 and ist % 2 == 0:
            OH.step(1, regrid_boxes=_oracle_regrid_boxes(cc, O0, t, crit, grid["blocking_factor"],
                                                        grid["max_grid_size"], grid["grid_eff"]))
        else:
            OH.step(1)
    got_boxes = H.boxes(1)
    say("GRF reactor 32^3: %d coarse boxes, %d fine boxes" % (len(cc.boxes), len(got_boxes)))
    _compare(OH, H, got_boxes, "GRF two-level reactor, 4 steps")
    _close(ctx, L0, H)


def test_grf_two_level_reactor_full_size():
    """the two-level reactor at its full size: 128^3 coarse level in uneven
    boxes, the fine level from the tags (about 10.7M fine cells in about 1300
    boxes): 6 coarse steps as persistent launches with the default tiles and
    replayed from a graph with each level's autotuned tiles give the same
    state bit for bit, finite, with positive pressure in the fluid and
    non-negative species"""
    cc = cases.porous_amr()
    s0 = cc.make_state()
    out = []
    for mode in (abi.MODE_BATCHED, abi.MODE_GRAPH):
        ctx = runtime.Context(cc)
        L0 = runtime.Level(ctx)
        L0.set_state(s0)
        fb = _tagged_boxes_gpu(L0, **cases.AMR_GRID)
        H = runtime.Hierarchy(L0, 2)
        H.regrid(fb)
        if mode == abi.MODE_GRAPH:
            H.autotune()
        ncf = sum(np.prod([h[d] - l[d] + 1 for d in range(3)]) for l, h in fb)
        H.step(1, mode)
        ctx.sync()
        import time
        t0 = time.time()
        H.step(5, mode)
        ctx.sync()
        el = time.time() - t0
        cu = 5 * (128 ** 3 + 2 * ncf) / el
        say("two-level reactor 128^3: %d coarse + %d fine boxes, %.2fM fine cells, mode %d: %.1f ms per coarse step, "
            "%.3g cell-updates/s" % (len(cc.boxes), len(fb), ncf / 1e6, mode, 1e3 * el / 5, cu))
        c = H.get_state(0)
        f = H.get_state(1)
        m = fine_mask(f.shape, fb)
        out.append((c, f[:, m]))
        _close(ctx, L0, H)
    (c0, f0), (c1, f1) = out
    assert np.array_equal(c0, c1) and np.array_equal(f0, f1)
    for a in (c0, f0):
        assert np.all(np.isfinite(a))
        fl = a[abi.VFRAC] == 1.0
        assert np.all(a[abi.PRES][fl] > 0)
        assert np.all(a[abi.FLO_NVARS:] >= 0)


def _hier_run(cc, fboxes, cs, fs, crit):
    ctx = runtime.Context(cc)
    L0 = runtime.Level(ctx)
    L0.set_state(cs)
    H = runtime.Hierarchy(L0, 2)
    H.add_level(fboxes, fs)
    H.set_regrid(2, *crit, n_error_buf=1, blocking_factor=4, max_grid_size=16, grid_eff=0.7)
    return ctx, L0, H


def test_hierarchy_checkpoint_restart_is_bitwise(tmp_path):
    """WriteCheckpointFile / ReadCheckpointFile of both levels: a run that
    writes a checkpoint after 3 coarse steps and takes 3 more equals, bit for
    bit, a fresh hierarchy that reads the checkpoint and takes the same 3
    steps (a regrid every 2 steps on both sides)"""
    from mesoflow_amd import amrexio
    n = 16
    cc = all_periodic(n, cases.chop_boxes((n, n, n), 8))
    s0 = shear_state(cc, amp=0.3)
    fb0 = [((8, 8, 8), (23, 23, 23))]
    cs, fs = fine_init(s0, fb0)
    crit = ([1e30], [0.05], [abi.DENS])
    chk = str(tmp_path / "chk00003")
    ctx, L0, H = _hier_run(cc, fb0, cs, fs, crit)
    H.step(3)
    H.write_checkpoint(chk)
    ck = amrexio.read_checkpoint_levels(chk, ctx.nv, [cc.n_cell, tuple(2 * q for q in cc.n_cell)])
    say("checkpoint: finest %d, istep %s, t %s" % (ck["finest_level"], ck["istep"], ck["time"]))
    assert ck["finest_level"] == 1 and ck["istep"] == [3, 6]
    H.step(3)
    a_c, a_f, a_info, a_boxes = H.get_state(0), H.get_state(1), H.info(), H.boxes(1)
    _close(ctx, L0, H)
    ctx2 = runtime.Context(cc)
    B0 = runtime.Level(ctx2)
    B0.set_state(cs)
    HB = runtime.Hierarchy(B0, 2)
    HB.set_regrid(2, *crit, n_error_buf=1, blocking_factor=4, max_grid_size=16, grid_eff=0.7)
    HB.read_checkpoint(chk)
    HB.step(3)
    b_c, b_f, b_info, b_boxes = HB.get_state(0), HB.get_state(1), HB.info(), HB.boxes(1)
    m = fine_mask(a_f.shape, a_boxes)
    say("restart: boxes %d / %d, t %.17g / %.17g" % (len(a_boxes), len(b_boxes), a_info[0], b_info[0]))
    assert sorted(a_boxes) == sorted(b_boxes)
    assert np.array_equal(a_c, b_c) and np.array_equal(a_f[:, m], b_f[:, m])
    assert a_info[0] == b_info[0] and a_info[1] == b_info[1] and a_info[2] == b_info[2]
    assert np.array_equal(a_info[3], b_info[3])
    _close(ctx2, B0, HB)


def test_hierarchy_plotfile(tmp_path):
    """WriteMultiLevelPlotfile of the two levels: the Header's levels,
    refinement ratio and steps, each level's data equal to its state"""
    from mesoflow_amd import amrexio
    n = 16
    cc = all_periodic(n, cases.chop_boxes((n, n, n), 8))
    s0 = shear_state(cc)
    fboxes = [((8, 8, 8), (15, 23, 19)), ((16, 8, 8), (21, 15, 19)), ((0, 20, 4), (5, 27, 11))]
    cs, fs = fine_init(s0, fboxes)
    ctx = runtime.Context(cc)
    L0 = runtime.Level(ctx)
    L0.set_state(cs)
    H = runtime.Hierarchy(L0, 2)
    H.add_level(fboxes, fs)
    H.step(2)
    p = str(tmp_path / "plt00002")
    H.write_plotfile(p)
    pf = amrexio.read_plotfile(p)
    assert pf["finest_level"] == 1 and pf["ref_ratio"] == [2]
    assert [lv["level_steps"] for lv in pf["levels"]] == [2, 4]
    c, f = H.get_state(0), H.get_state(1)
    assert np.array_equal(pf["levels"][0]["data"], c)
    m = fine_mask(f.shape, fboxes)
    got = pf["levels"][1]["data"]
    assert np.array_equal(got[:, m], f[:, m]) and np.all(np.isnan(got[:, ~m]))
    assert sorted(pf["levels"][1]["boxes"]) == sorted(fboxes)
    _close(ctx, L0, H)
