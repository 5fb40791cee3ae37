"""The two-level coupling operators on the GPU (mf_box_average_down,
mf_box_time_interp, mf_box_interp_cons, mf_fluxreg_*) against the oracle's
(oracle/mf_amr_oracle.cpp) on the same random fabs: bit for bit, since both
round every product on its own and sum in the same order."""
import numpy as np
import pytest
import torch  # loaded before the library opens the device, as in the other tests that use torch

from mesoflow_amd import amr, cases, runtime

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = cases.porous_reactor(n_cell=(8, 8, 8), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=2.0,
                             rmax=3.0, order_hyp=5, seed=1)
    x = runtime.Context(c)
    yield x
    x.close()


def _rand(rng, lo, hi, nv, scale=1.0):
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    return (rng.standard_normal((nv, n[2], n[1], n[0])) * scale, tuple(lo))


def _dev(fab):
    a, lo = fab
    return (torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0"), lo)


def _host(dfab, ctx):
    ctx.sync()
    return dfab[0].cpu().numpy()


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


@pytest.mark.parametrize("r", [2, 3])
def test_average_down(ctx, r):
    from oracle import pyoracle as po
    rng = np.random.default_rng(r)
    nv = ctx.nv
    cbx = ((-2, 1, 0), (3, 4, 2))
    fine = _rand(rng, tuple(cbx[0][d] * r - 2 for d in range(3)), tuple(cbx[1][d] * r + r + 1 for d in range(3)), nv)
    crse = _rand(rng, (-3, 0, -1), (4, 5, 3), nv)
    want = (crse[0].copy(), crse[1])
    po.average_down(cbx, fine, want, r)
    dc = _dev(crse)
    amr.average_down(ctx, cbx, _dev(fine), dc, r)
    got = _host(dc, ctx)
    say("average_down r=%d: bitwise %s" % (r, _same(got, want[0])))
    assert _same(got, want[0])


def test_time_interp(ctx):
    from oracle import pyoracle as po
    rng = np.random.default_rng(5)
    nv = ctx.nv
    bx = ((-3, -3, -3), (10, 6, 4))
    so, sn = _rand(rng, *bx, nv), _rand(rng, *bx, nv)
    for t in (0.1, 0.1 + 2e-5, 0.3, 0.42, 0.5 - 1e-5, 0.5):
        want = _rand(rng, *bx, nv)
        dd = _dev(want)
        po.time_interp(bx, want, so, sn, 0.1, 0.5, t)
        amr.time_interp(ctx, bx, dd, _dev(so), _dev(sn), 0.1, 0.5, t)
        got = _host(dd, ctx)
        assert _same(got, want[0]), t
    say("time_interp: bitwise at 6 times")


@pytest.mark.parametrize("r", [2, 3])
def test_interp_cons(ctx, r):
    from oracle import pyoracle as po
    rng = np.random.default_rng(11 + r)
    nv = ctx.nv
    crse = _rand(rng, (-4, -3, -2), (6, 5, 4), nv)
    crse[0][3] = 2.0   # a constant component: flat slopes
    fbx = ((-3 * r + 1, -2 * r, -1 * r), (5 * r + r - 2, 4 * r, 3 * r + r - 1))
    want = _rand(rng, *fbx, nv)
    df = _dev(want)
    po.interp_cons(fbx, crse, want, r)
    amr.interp_cons(ctx, fbx, _dev(crse), df, r)
    got = _host(df, ctx)
    say("interp_cons r=%d: bitwise %s" % (r, _same(got, want[0])))
    assert _same(got, want[0])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_flux_register_sequence(ctx, d):
    """CrseInit, two fine substeps of FineAdd with the RK3 weights, then Reflux
    on both sides of a coarsened fine box"""
    from oracle import pyoracle as po
    rng = np.random.default_rng(20 + d)
    nv, r = ctx.nv, 2
    cl, ch = (2, 1, 3), (6, 4, 5)
    state = _rand(rng, (0, 0, 0), (9, 7, 8), nv)
    dstate = _dev(state)
    want_state = (state[0].copy(), state[1])
    for side in (0, 1):
        face = [list(cl), list(ch)]
        fidx = cl[d] if side == 0 else ch[d] + 1
        face[0][d] = face[1][d] = fidx
        face = (tuple(face[0]), tuple(face[1]))
        reg = _rand(rng, *face, nv, 0.0)
        dreg = _dev(reg)
        cflux = _rand(rng, *face, nv)
        flo = [face[0][q] * r for q in range(3)]
        fhi = [face[1][q] * r + r - 1 for q in range(3)]
        fhi[d] = flo[d]
        dt_c = 1.0e-3
        for w in (1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0):
            po.fluxreg_crse_add(d, face, reg, cflux, -dt_c * w * 0.25)
            amr.fluxreg_crse_add(ctx, d, face, dreg, _dev(cflux), -dt_c * w * 0.25)
        for sub in range(2):
            for w in (1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0):
                ff = _rand(rng, tuple(flo), tuple(fhi), nv)
                po.fluxreg_fine_add(d, face, reg, ff, r, 0.5 * dt_c * w * 0.0625)
                amr.fluxreg_fine_add(ctx, d, face, dreg, _dev(ff), r, 0.5 * dt_c * w * 0.0625)
        assert _same(_host(dreg, ctx), reg[0])
        po.fluxreg_reflux(d, side, face, want_state, reg, 1.0, 0.125)
        amr.fluxreg_reflux(ctx, d, side, face, dstate, dreg, 1.0, 0.125)
    got = _host(dstate, ctx)
    moved = np.any(got != state[0], axis=0).sum()
    say("flux register dir %d: bitwise %s, %d coarse cells refluxed" % (d, _same(got, want_state[0]), moved))
    assert _same(got, want_state[0])
    assert moved == 2 * np.prod([ch[q] - cl[q] + 1 for q in range(3) if q != d])


def test_box_tag(ctx):
    from oracle import pyoracle as po
    rng = np.random.default_rng(31)
    nv = ctx.nv
    bx = ((-2, 1, 0), (13, 9, 6))
    st = _rand(rng, tuple(bx[0][d] - 1 for d in range(3)), tuple(bx[1][d] + 1 for d in range(3)), nv)
    # NaN neighbours: the maxima keep std::max's NaN semantics (the oracle's)
    a = st[0]
    for q in range(12):
        a[7, rng.integers(a.shape[1]), rng.integers(a.shape[2]), rng.integers(a.shape[3])] = np.nan
    phi, grad, comps = [2.5, 1e30, 3.0], [3.9, 4.2, 1e30], [0, 7, nv - 1]
    n = [bx[1][d] - bx[0][d] + 1 for d in range(3)]
    init = np.zeros((n[2], n[1], n[0]), np.uint8)
    init[1, 2, 3] = 9
    want = init.copy()
    po.box_tag(bx, st, want, phi, grad, comps, 2)
    dt = torch.from_numpy(init).to("cuda:0")
    amr.box_tag(ctx, bx, _dev(st), dt, phi, grad, comps, 2)
    ctx.sync()
    got = dt.cpu().numpy()
    say("box tag: %d of %d cells tagged, equal %s" % ((got == 2).sum(), got.size, np.array_equal(got, want)))
    assert np.array_equal(got, want)
    assert 0 < (got == 2).sum() < got.size


def test_level_tag():
    from oracle.pyoracle import OracleLevel, box_tag
    c = cases.porous_reactor(n_cell=(20, 16, 12), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=2.0,
                             rmax=3.0, order_hyp=5, seed=6)
    s = c.make_state()
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    L.set_state(s)
    O = OracleLevel(c)
    O.set_state(s)
    from mesoflow_amd import abi
    phi, grad, comps = [0.5, 1e30], [1e30, 50.0], [abi.VFRAC, abi.TEMP]
    got, cnt = amr.level_tag(L, 0.0, phi, grad, comps)
    pad = O.fill_ghosts(0.0)
    nx, ny, nz = c.n_cell
    want = np.zeros((nz, ny, nx), np.uint8)
    box_tag(((0, 0, 0), (nx - 1, ny - 1, nz - 1)), (pad, (-3, -3, -3)), want, phi, grad, comps, 1)
    say("level tag: %d tagged (oracle %d)" % (cnt, int(want.sum())))
    assert np.array_equal(got, want) and cnt == int(want.sum()) > 0
    # the regrid of these tags: fine boxes (ratio 2) over every tagged cell
    boxes = amr.tags_to_boxes(got, 1, 4, 16, 0.8, 2, periodic=tuple(c.geom.is_periodic))
    cov = np.zeros((nz * 2, ny * 2, nx * 2), np.int32)
    for lo, hi in boxes:
        cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    fine = np.repeat(np.repeat(np.repeat(got.astype(bool), 2, 0), 2, 1), 2, 2)
    say("regrid: %d fine boxes, %.1f %% of the fine domain" % (len(boxes), 100.0 * (cov > 0).mean()))
    assert cov.max() == 1 and np.all(cov[fine] == 1)
    L.close()
    ctx.close()


def test_fill_patch_two_levels():
    """four fine boxes of a periodic 32^3 fine level (ratio 2), two of them
    neighbours across the periodic seam: each box's Sborder against the oracle's time interpolation +
    cell_cons_interp followed by copies of the fine data and its periodic images"""
    from oracle import pyoracle as po
    c = cases.shocktube_n2(n_cell=(32, 32, 32))
    c.geom.is_periodic[0] = 1
    c.params.lo_bc[0] = c.params.hi_bc[0] = 0
    fctx = runtime.Context(c)
    nv = fctx.nv
    rng = np.random.default_rng(41)
    co = _rand(rng, (-3, -3, -3), (18, 18, 18), nv)
    cn = _rand(rng, (-3, -3, -3), (18, 18, 18), nv)
    boxes = [((8, 8, 8), (15, 23, 23)), ((16, 8, 8), (23, 23, 23)), ((0, 0, 26), (5, 7, 31)), ((26, 0, 0), (31, 7, 5))]
    fabs = [_rand(rng, *b, nv) for b in boxes]
    dfabs = [(b, _dev(f)) for b, f in zip(boxes, fabs)]
    dco, dcn = _dev(co), _dev(cn)
    N = 32
    for q, (lo, hi) in enumerate(boxes):
        glo, ghi = tuple(v - 3 for v in lo), tuple(v + 3 for v in hi)
        clo, chi = tuple(v // 2 - 1 for v in glo), tuple(v // 2 + 1 for v in ghi)
        ct = _rand(rng, clo, chi, nv)
        po.time_interp((clo, chi), ct, co, cn, 0.0, 1.0, 0.25)
        want = _rand(rng, glo, ghi, nv)
        po.interp_cons((glo, ghi), ct, want, 2)
        w = want[0]
        for (blo, bhi), (fa, _) in zip(boxes, fabs):
            for sh in [(a, b_, c_) for a in (-N, 0, N) for b_ in (-N, 0, N) for c_ in (-N, 0, N)]:
                xl = [max(glo[d], blo[d] + sh[d]) for d in range(3)]
                xh = [min(ghi[d], bhi[d] + sh[d]) for d in range(3)]
                if any(xh[d] < xl[d] for d in range(3)):
                    continue
                w[:, xl[2] - glo[2]:xh[2] - glo[2] + 1, xl[1] - glo[1]:xh[1] - glo[1] + 1,
                  xl[0] - glo[0]:xh[0] - glo[0] + 1] = \
                    fa[:, xl[2] - sh[2] - blo[2]:xh[2] - sh[2] - blo[2] + 1,
                       xl[1] - sh[1] - blo[1]:xh[1] - sh[1] - blo[1] + 1,
                       xl[0] - sh[0] - blo[0]:xh[0] - sh[0] - blo[0] + 1]
        dsb = _dev(_rand(rng, glo, ghi, nv))
        amr.fill_patch_two_levels(fctx, (lo, hi), dsb, dfabs, dco, dcn, 0.0, 1.0, 0.25, 2)
        got = _host(dsb, fctx)
        say("fill patch two levels, box %d: bitwise %s" % (q, _same(got, w)))
        assert _same(got, w)
    fctx.close()


def test_fill_patch_two_levels_at_physical_boundaries():
    """fine boxes whose ghost regions leave the domain: x ext_dir (low) / wall
    (high), y foextrap (low) / wall (high), z periodic. Each box's Sborder
    against the oracle composition: time interpolation + cell_cons_interp, the
    fine data and its periodic images, then the physical-boundary rules of the
    fine level (orc_fill_physbc) on the cells outside the domain, corners
    included"""
    from oracle import pyoracle as po
    c = cases.shocktube_n2(n_cell=(32, 32, 32))
    c.geom.is_periodic[1] = 0
    c.params.lo_bc[1], c.params.hi_bc[1] = 2, 6
    fctx = runtime.Context(c)
    nv = fctx.nv
    rng = np.random.default_rng(43)
    co = _rand(rng, (-3, -3, -3), (18, 18, 18), nv)
    cn = _rand(rng, (-3, -3, -3), (18, 18, 18), nv)
    boxes = [((0, 0, 8), (7, 7, 15)), ((24, 24, 28), (31, 31, 31)), ((8, 0, 0), (15, 7, 3)), ((0, 10, 12), (1, 20, 19))]
    fabs = [_rand(rng, *b, nv) for b in boxes]
    dfabs = [(b, _dev(f)) for b, f in zip(boxes, fabs)]
    dco, dcn = _dev(co), _dev(cn)
    N = 32
    for q, (lo, hi) in enumerate(boxes):
        glo, ghi = tuple(v - 3 for v in lo), tuple(v + 3 for v in hi)
        clo, chi = tuple(v // 2 - 1 for v in glo), tuple(v // 2 + 1 for v in ghi)
        ct = _rand(rng, clo, chi, nv)
        po.time_interp((clo, chi), ct, co, cn, 0.0, 1.0, 0.25)
        want = _rand(rng, glo, ghi, nv)
        po.interp_cons((glo, ghi), ct, want, 2)
        w = want[0]
        for (blo, bhi), (fa, _) in zip(boxes, fabs):
            for sh in [(0, 0, c_) for c_ in (-N, 0, N)]:             # periodic in z only
                xl = [max(glo[d], blo[d] + sh[d]) for d in range(3)]
                xh = [min(ghi[d], bhi[d] + sh[d]) for d in range(3)]
                if any(xh[d] < xl[d] for d in range(3)):
                    continue
                w[:, xl[2] - glo[2]:xh[2] - glo[2] + 1, xl[1] - glo[1]:xh[1] - glo[1] + 1,
                  xl[0] - glo[0]:xh[0] - glo[0] + 1] = \
                    fa[:, xl[2] - sh[2] - blo[2]:xh[2] - sh[2] - blo[2] + 1,
                       xl[1] - sh[1] - blo[1]:xh[1] - sh[1] - blo[1] + 1,
                       xl[0] - sh[0] - blo[0]:xh[0] - sh[0] - blo[0] + 1]
        po.fill_physbc(c, (glo, ghi), want, 0.25)
        dsb = _dev(_rand(rng, glo, ghi, nv))
        amr.fill_patch_two_levels(fctx, (lo, hi), dsb, dfabs, dco, dcn, 0.0, 1.0, 0.25, 2)
        got = _host(dsb, fctx)
        nout = sum(1 for d in range(2) if glo[d] < 0 or ghi[d] >= N)
        say("fill patch at the boundary, box %d (%d non-periodic sides crossed): bitwise %s" % (q, nout, _same(got, w)))
        assert _same(got, w)
    fctx.close()
