"""The plain per-box C-ABI (mf_box_cutcells, mf_box_rhs with face fluxes,
mf_box_rk3_combine) on device views, against the oracle's per-box functions on
the same Sborder arrays: the drop-in form of the MFIter bodies of
Source/mflo_evolve.cpp:509-517 (update_cutcell_data), :579-652
(compute_dsdt_flow) and the stage arithmetic of :283, :296-298, :309-311."""
import ctypes as C

import numpy as np
import pytest

from mesoflow_amd import abi, cases, runtime
from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _view(t, begin, end, nv):
    return abi.make_array4(t.data_ptr(), begin, end, nv)


def _case(order=5, seed=13):
    return cases.porous_reactor(n_cell=(18, 14, 12), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=2.0,
                                rmax=3.5, order_hyp=order, cfl=0.3, seed=seed)


BOXES = [((0, 0, 0), (17, 13, 11)), ((3, 2, 1), (10, 9, 7)), ((11, 0, 5), (17, 13, 11)), ((6, 6, 6), (6, 9, 6)),
         ((0, 13, 0), (17, 13, 0))]


def _sborder(O, lo, hi):
    """the level's ghost-filled state cut to the box grown by 3"""
    sb = O.fill_ghosts(0.0)
    return np.ascontiguousarray(sb[:, lo[2]:hi[2] + 7, lo[1]:hi[1] + 7, lo[0]:hi[0] + 7])


@pytest.mark.parametrize("order", [5, 2])
def test_box_cutcells_and_rhs(order):
    import torch
    from oracle.pyoracle import OracleLevel, box_cutcells, box_rhs
    c = _case(order)
    s = c.make_state()
    r = np.random.default_rng(4)
    fl = s[abi.VFRAC] == 1.0
    ke0 = sum(s[abi.VELX + q] ** 2 for q in range(3))
    for q in range(3):   # a non-uniform velocity field, so every flux term is exercised
        v = r.uniform(-0.3, 0.3, size=fl.shape) * fl
        s[abi.VELX + q] += v
        s[abi.RHOU + q] += s[abi.DENS] * v
    s[abi.RHOE] += 0.5 * s[abi.DENS] * (sum(s[abi.VELX + q] ** 2 for q in range(3)) - ke0)
    O = OracleLevel(c)
    O.set_state(s)
    ctx = runtime.Context(c)
    nv = ctx.nv
    L = runtime.lib()
    for lo, hi in BOXES:
        glo = [lo[d] - 3 for d in range(3)]
        ghi = [hi[d] + 4 for d in range(3)]
        sb = _sborder(O, lo, hi)
        ref_cut = box_cutcells(c, lo, hi, sb)
        d_sb = _dev(sb)
        torch.cuda.synchronize()
        runtime.check(L.mf_box_cutcells(ctx.h, abi.make_box(lo, hi), _view(d_sb, glo, ghi, nv), 0.0),
                      "mf_box_cutcells")
        ctx.sync()
        got_cut = d_sb.cpu().numpy()
        e_cut = rel_err(ref_cut, got_cut).max()
        # the right-hand side of the oracle's cut-cell result, with fluxes
        ref_d, ref_f = box_rhs(c, lo, hi, ref_cut, with_flux=True)
        d_sb = _dev(ref_cut)
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        d_ds = torch.full((nv, n[2], n[1], n[0]), np.nan, dtype=torch.float64, device="cuda:0")
        fv = (abi.Array4 * 3)()
        d_fl = []
        for d in range(3):
            fe = [hi[q] + 1 for q in range(3)]
            fe[d] += 1
            t = torch.zeros((nv, fe[2] - lo[2], fe[1] - lo[1], fe[0] - lo[0]), dtype=torch.float64, device="cuda:0")
            d_fl.append(t)
            fv[d] = _view(t, lo, fe, nv)
        torch.cuda.synchronize()
        runtime.check(L.mf_box_rhs(ctx.h, abi.make_box(lo, hi), _view(d_sb, glo, ghi, nv),
                                   _view(d_ds, lo, [hi[d] + 1 for d in range(3)], nv), fv, 0.0), "mf_box_rhs")
        ctx.sync()
        e_d = rel_err(ref_d, d_ds.cpu().numpy()).max()
        e_f = max(rel_err(ref_f[d], d_fl[d].cpu().numpy()).max() for d in range(3))
        say("order %d box %s-%s: cutcells %.2e  dsdt %.2e  flux %.2e" % (order, lo, hi, e_cut, e_d, e_f))
        assert np.all(np.isfinite(d_ds.cpu().numpy()))
        assert e_cut <= TOL and e_d <= TOL and e_f <= TOL
    ctx.close()


def test_box_rk3_combine():
    import torch
    c = _case()
    ctx = runtime.Context(c)
    nv = ctx.nv
    L = runtime.lib()
    r = np.random.default_rng(8)
    lo, hi = (2, 1, 3), (9, 6, 8)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    shp = (nv, n[2] + 2, n[1] + 2, n[0] + 2)   # views one cell larger than the box: only the box is written
    begin = [lo[d] - 1 for d in range(3)]
    end = [hi[d] + 2 for d in range(3)]
    src, sn, ds = (r.standard_normal(shp) for _ in range(3))
    dt = 3.7e-3
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    for stage in (1, 2, 3):
        want = sn.copy()
        a, b, x = sn[inner], src[inner], ds[inner]
        if stage == 1:
            want[inner] = 1.0 * b + dt * x
        elif stage == 2:
            want[inner] = ((a + dt * x) + 3.0 * b) * 0.25
        else:
            want[inner] = (b + 2.0 * (a + dt * x)) * (1.0 / 3.0)
        d_sn, d_src, d_ds = _dev(sn), _dev(src), _dev(ds)
        torch.cuda.synchronize()
        runtime.check(L.mf_box_rk3_combine(ctx.h, stage, abi.make_box(lo, hi), _view(d_sn, begin, end, nv),
                                           _view(d_src, begin, end, nv), _view(d_ds, begin, end, nv), dt),
                      "mf_box_rk3_combine")
        ctx.sync()
        got = d_sn.cpu().numpy()
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        say("rk3 combine stage %d: max rel err %.2e" % (stage, err))
        assert err <= 1e-15
        outside = np.ones(shp, bool)
        outside[inner] = False
        assert np.array_equal(got[outside], sn[outside])
    ctx.close()


def test_box_views_must_cover_the_box():
    """a view that does not hold the region a call writes or reads is rejected on
    the host (MF_ERR_BAD_ARG) before any launch"""
    import torch
    c = _case()
    ctx = runtime.Context(c)
    nv = ctx.nv
    L = runtime.lib()
    lo, hi = (2, 2, 2), (7, 6, 5)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    glo = [lo[d] - 3 for d in range(3)]
    ghi = [hi[d] + 4 for d in range(3)]
    d_sb = torch.zeros((nv, n[2] + 6, n[1] + 6, n[0] + 6), dtype=torch.float64, device="cuda:0")
    d_ds = torch.zeros((nv, n[2], n[1], n[0]), dtype=torch.float64, device="cuda:0")
    box = abi.make_box(lo, hi)
    fv = (abi.Array4 * 3)()
    keep = []
    for d in range(3):
        t = torch.zeros((nv, n[2], n[1], n[0]), dtype=torch.float64, device="cuda:0")   # no extra face
        keep.append(t)
        fv[d] = _view(t, lo, [hi[q] + 1 for q in range(3)], nv)
    st = L.mf_box_rhs(ctx.h, box, _view(d_sb, glo, ghi, nv), _view(d_ds, lo, [hi[q] + 1 for q in range(3)], nv), fv, 0.0)
    assert st == abi.MF_ERR_BAD_ARG
    small = _view(d_ds, lo, [hi[q] + 1 for q in range(3)], nv)
    big = abi.make_box(lo, (hi[0] + 1, hi[1], hi[2]))
    st = L.mf_box_rk3_combine(ctx.h, 1, big, small, small, small, 1.0)
    assert st == abi.MF_ERR_BAD_ARG
    st = L.mf_box_cutcells(ctx.h, big, _view(d_sb, glo, ghi, nv), 0.0)
    assert st == abi.MF_ERR_BAD_ARG
    ctx.sync()
    ctx.close()


@pytest.mark.parametrize("flags", [1, 2, 4, 7])
def test_box_post(flags):
    """mf_box_post (chemistry, primitive update, clip) against the oracle's
    per-box form on a porous state with strong reactions, negative species and
    perturbed conserved variables"""
    import torch
    from oracle.pyoracle import box_post
    c = _case()
    for r in range(2):
        c.plugin.log_A[r] = np.log(3.0e10) + 0.5 * r
        c.plugin.Ea[r] = 90.0e3
    s = c.make_state()
    r = np.random.default_rng(flags)
    fl = s[abi.VFRAC] == 1.0
    s[abi.RHO] *= 1.0 + 0.01 * r.standard_normal(fl.shape)
    s[abi.RHOU] += 0.05 * s[abi.RHO] * r.standard_normal(fl.shape) * fl
    s[abi.FLO_NVARS + 1] -= 0.02 * np.abs(s[abi.FLO_NVARS + 1]).max() * r.random(fl.shape)
    lo, hi = (2, 1, 3), (15, 12, 9)
    fab = np.ascontiguousarray(s[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])
    want = fab.copy()
    st_o = box_post(c, (lo, hi), (want, lo), 2.0e-7, flags)
    ctx = runtime.Context(c)
    d = _dev(fab)
    torch.cuda.synchronize()
    st = runtime.lib().mf_box_post(ctx.h, abi.make_box(lo, hi), _view(d, lo, [hi[q] + 1 for q in range(3)], ctx.nv),
                                   2.0e-7, flags)
    assert st == abi.MF_OK
    st_g = abi.MF_OK
    try:
        ctx.sync()
    except abi.MFError as e:
        st_g = e.status
    got = d.cpu().numpy()
    err = rel_err(want, got).max()
    changed = np.any(want != fab)
    say("box post flags %d: max rel err %.2e, oracle status %d, gpu %d, state changed %s" % (flags, err, st_o, st_g,
                                                                                            changed))
    assert st_o == st_g == abi.MF_OK and changed
    assert err <= TOL
    ctx.close()


def test_box_post_negative_pressure():
    from oracle.pyoracle import box_post
    import torch
    c = _case()
    s = c.make_state()
    solid = np.argwhere(s[abi.VFRAC] == 0.0)
    s[abi.PRES][tuple(solid[len(solid) // 2])] = -1.0
    fab = np.ascontiguousarray(s)
    nx, ny, nz = c.n_cell
    bx = ((0, 0, 0), (nx - 1, ny - 1, nz - 1))
    want = fab.copy()
    st_o = box_post(c, bx, (want, (0, 0, 0)), 1e-9, abi.MF_POST_PRIMITIVES)
    ctx = runtime.Context(c)
    d = _dev(fab)
    torch.cuda.synchronize()
    assert runtime.lib().mf_box_post(ctx.h, abi.make_box(*bx), _view(d, (0, 0, 0), (nx, ny, nz), ctx.nv), 1e-9,
                                     abi.MF_POST_PRIMITIVES) == abi.MF_OK
    with pytest.raises(abi.MFError) as e:
        ctx.sync()
    say("box post negative pressure: oracle %d gpu %d" % (st_o, e.value.status))
    assert st_o == e.value.status == abi.MF_ERR_NEG_PRESSURE
    ctx.close()
