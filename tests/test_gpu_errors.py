"""Error paths of the C-ABI on the GPU, each against the oracle where it has the
same check. The reference aborts at these sites (amrex::Abort): NaN in dsdt
(div_3d.H:72-75), negative pressure after update_primitive_vars
(mflo_evolve.cpp:162-167), an unknown hyperbolic scheme
(compute_flux_3d.H:273-276) and a boundary code it does not map
(mflo.cpp:113,141). The conditions come from the inputs: every kernel still
runs to completion and the device error word carries the status."""
import numpy as np
import pytest

from mesoflow_amd import abi, cases, runtime

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _status_of(fn):
    try:
        fn()
    except abi.MFError as e:
        return e.status
    return abi.MF_OK


def _porous(n=(12, 8, 8)):
    return cases.porous_reactor(n_cell=n, nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=2.0, rmax=3.0,
                                order_hyp=5, cfl=0.3, seed=4)


def _step_status(case, state):
    from oracle.pyoracle import OracleLevel
    ctx = runtime.Context(case)
    L = runtime.Level(ctx)
    L.set_state(state)
    g = _status_of(lambda: L.step(2))
    L.close()
    ctx.close()
    O = OracleLevel(case)
    O.set_state(state)
    o = _status_of(lambda: O.step(2))
    return g, o


def test_nan_in_state_is_reported_as_nan_dsdt():
    c = _porous()
    c.params.fixed_dt = 1
    c.params.dt_fixed = 1.0e-9
    s = c.make_state()
    fluid = np.argwhere(s[abi.VFRAC] == 1.0)[len(np.argwhere(s[abi.VFRAC] == 1.0)) // 2]
    s[abi.RHOU][tuple(fluid)] = np.nan
    g, o = _step_status(c, s)
    say("NaN: gpu %s oracle %s" % (abi.status_name(g), abi.status_name(o)))
    assert g == o == abi.MF_ERR_NAN_DSDT


def test_negative_pressure_is_reported():
    """a negative pressure component in a solid cell: no flux reads it (the face
    fluxes form p from the conserved state, the cut cells read fluid cells
    only), update_primitive_vars leaves it (vfrac < 1), and the minimum over
    the box is negative"""
    c = _porous()
    c.params.fixed_dt = 1
    c.params.dt_fixed = 1.0e-9
    s = c.make_state()
    solid = np.argwhere(s[abi.VFRAC] == 0.0)
    assert len(solid)
    s[abi.PRES][tuple(solid[len(solid) // 2])] = -1.0
    g, o = _step_status(c, s)
    say("negative pressure: gpu %s oracle %s" % (abi.status_name(g), abi.status_name(o)))
    assert g == o == abi.MF_ERR_NEG_PRESSURE


def test_negative_internal_energy_fails_like_the_oracle():
    """total energy below the kinetic energy in a fluid cell: the first stage's
    eigenvalue takes the root of a negative pressure and dsdt is NaN"""
    c = _porous()
    c.params.fixed_dt = 1
    c.params.dt_fixed = 1.0e-9
    s = c.make_state()
    fl = np.argwhere(s[abi.VFRAC] == 1.0)
    q = tuple(fl[len(fl) // 3])
    s[abi.RHOE][q] = 0.25 * s[abi.RHO][q] * s[abi.VELX][q] ** 2 - 1.0e4
    g, o = _step_status(c, s)
    say("negative rhoe: gpu %s oracle %s" % (abi.status_name(g), abi.status_name(o)))
    assert g == o == abi.MF_ERR_NAN_DSDT


@pytest.mark.parametrize("what", ["scheme", "bc", "plugin"])
def test_context_checks(what):
    from oracle.pyoracle import OracleLevel
    c = _porous()
    if what == "scheme":
        c.params.order_hyp = 3
        want = abi.MF_ERR_UNKNOWN_SCHEME
    elif what == "bc":
        c.params.hi_bc[1] = 5
        c.params.lo_bc[1] = 5
        c.geom.is_periodic[1] = 0
        want = abi.MF_ERR_BAD_BC
    else:
        c.plugin.nspec = 7
        c.plugin.ngas = 6
        want = abi.MF_ERR_UNSUPPORTED
    g = _status_of(lambda: runtime.Context(c))
    say("%s: gpu %s" % (what, abi.status_name(g)))
    assert g == want
    if what != "bc":
        assert _status_of(lambda: OracleLevel(c)) == want


def test_adjoint_needs_fixed_dt():
    c = _porous()
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    L.set_state(c.make_state())
    assert _status_of(lambda: L.adjoint_forward(2, 2, 1)) == abi.MF_ERR_UNSUPPORTED
    L.close()
    ctx.close()


def test_failed_scratch_allocation_leaves_a_usable_context():
    """mf_box_rhs without flux views keeps its face fluxes in a context-owned
    scratch buffer. A request that cannot be allocated (HBM held by another
    allocation) returns MF_ERR_HIP and leaves the context without a buffer;
    the next smaller call allocates again and computes the same dsdt as the
    oracle (a stale size with a freed buffer would make it write through a
    null pointer)."""
    import torch
    from oracle.pyoracle import box_rhs
    from helpers import rel_err
    c = cases.porous_reactor(n_cell=(128, 128, 128), nspec=5, nreact=2, geometry="none", order_hyp=5, cfl=0.3, seed=4)
    ctx = runtime.Context(c)
    nv = ctx.nv
    L = runtime.lib()
    s = c.make_state()
    rng = np.random.default_rng(2)
    s[abi.RHOE] *= 1.0 + 1e-3 * rng.uniform(-1, 1, s[abi.RHOE].shape)
    sb = np.pad(s, ((0, 0), (3, 3), (3, 3), (3, 3)), mode="edge")
    d_sb = torch.from_numpy(sb).to("cuda:0")
    big = ((0, 0, 0), (127, 127, 127))
    d_ds = torch.zeros((nv, 128, 128, 128), dtype=torch.float64, device="cuda:0")
    v_sb = abi.make_array4(d_sb.data_ptr(), (-3, -3, -3), (131, 131, 131), nv)
    v_ds = abi.make_array4(d_ds.data_ptr(), (0, 0, 0), (128, 128, 128), nv)
    small = ((10, 20, 30), (17, 25, 33))

    def rhs(bx):
        return L.mf_box_rhs(ctx.h, abi.make_box(*bx), v_sb, v_ds, None, 0.0)

    assert rhs(small) == abi.MF_OK                 # a small scratch buffer exists
    ctx.sync()
    torch.cuda.synchronize()
    # hold the device memory in chunks until allocation fails (1 GiB, then 32 MiB)
    hog = []
    for chunk in (1 << 30, 32 << 20):
        while True:
            try:
                hog.append(torch.empty(chunk, dtype=torch.uint8, device="cuda:0"))
            except torch.cuda.OutOfMemoryError:
                break
    say("held %.1f GB of device memory" % (sum(h.numel() for h in hog) / 1e9))
    st = rhs(big)                                  # about 510 MB of face fluxes: does not fit
    say("oversized scratch request: status %d" % st)
    del hog
    torch.cuda.empty_cache()
    assert st == abi.MF_ERR_HIP
    d_ds.zero_()
    assert rhs(small) == abi.MF_OK
    ctx.sync()
    lo, hi = small
    sub = np.ascontiguousarray(sb[:, lo[2]:hi[2] + 7, lo[1]:hi[1] + 7, lo[0]:hi[0] + 7])
    want = box_rhs(c, lo, hi, sub)
    got = d_ds.cpu().numpy()[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
    # the same call on a fresh context: bit for bit
    ctx2 = runtime.Context(c)
    d_ds2 = torch.zeros_like(d_ds)
    v_ds2 = abi.make_array4(d_ds2.data_ptr(), (0, 0, 0), (128, 128, 128), nv)
    assert L.mf_box_rhs(ctx2.h, abi.make_box(*small), v_sb, v_ds2, None, 0.0) == abi.MF_OK
    ctx2.sync()
    fresh = d_ds2.cpu().numpy()[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
    assert np.array_equal(got.view(np.int64), fresh.view(np.int64))
    ctx2.close()
    # and the oracle, on the components the perturbed energy drives (density and
    # species stay uniform: their dsdt is rounding noise on both sides)
    err = rel_err(want, got)
    say("small box after the failure: rel err per component %s" % np.array2string(err, precision=2))
    assert err[1:5].max() <= 1e-11
    ctx.close()


def test_level_boxes_must_not_overlap():
    """boxes whose volumes add up to the domain but overlap (and so leave a gap)
    are rejected: tiles of both boxes would write the overlap and no tile the gap"""
    c = cases.shocktube_n2(n_cell=(16, 8, 8), prob_hi=(1.0, 0.5, 0.5), order_hyp=5)
    ctx = runtime.Context(c)
    bad = [((0, 0, 0), (8, 7, 7)), ((8, 0, 0), (14, 7, 7))]      # x = 8 twice, x = 15 never
    assert _status_of(lambda: runtime.Level(ctx, bad)) == abi.MF_ERR_BAD_ARG
    good = [((0, 0, 0), (7, 7, 7)), ((8, 0, 0), (15, 7, 7))]
    L = runtime.Level(ctx, good)
    L.close()
    ctx.close()
