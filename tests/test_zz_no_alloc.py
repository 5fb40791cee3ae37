"""mf_box_rhs_noalloc (the per-box RHS that allocates nothing) through the
C-ABI, against mf_box_rhs without flux views on the same device views: two
calls in a row, the second on a smaller box, each bit for bit; and a box too
large for its scratch array is refused before any launch."""
import numpy as np
import pytest
import torch  # loaded before the library opens the device

from mesoflow_amd import abi, cases, runtime

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_noalloc_rhs_matches_the_allocating_entry_point():
    c = cases.porous_reactor(n_cell=(64, 48, 40), nspec=5, nreact=2, geometry="spheres", porosity=0.6, rmin=3.0,
                             rmax=6.0, order_hyp=5, cfl=0.3, seed=17)
    from oracle.pyoracle import OracleLevel
    s = c.make_state()
    rng = np.random.default_rng(5)
    fl = s[abi.VFRAC] == 1.0
    for q in range(3):
        v = rng.uniform(-0.2, 0.2, size=fl.shape) * fl
        s[abi.VELX + q] += v
        s[abi.RHOU + q] += s[abi.DENS] * v
    O = OracleLevel(c)
    O.set_state(s)
    sb = O.fill_ghosts(0.0)                          # [nv][nz+6][ny+6][nx+6], index -3 first
    ctx = runtime.Context(c)
    nv = ctx.nv
    L = runtime.lib()
    nx, ny, nz = c.n_cell
    d_sb = torch.from_numpy(sb).to("cuda:0")
    v_sb = abi.make_array4(d_sb.data_ptr(), (-3, -3, -3), (nx + 3, ny + 3, nz + 3), nv)
    for bx in (((0, 0, 0), (63, 47, 39)), ((5, 9, 2), (20, 30, 11))):
        lo, hi = bx
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        outs = []
        for name in ("mf_box_rhs", "mf_box_rhs_noalloc"):
            d_ds = torch.full((nv, n[2], n[1], n[0]), np.nan, dtype=torch.float64, device="cuda:0")
            v_ds = abi.make_array4(d_ds.data_ptr(), lo, (hi[0] + 1, hi[1] + 1, hi[2] + 1), nv)
            if name == "mf_box_rhs":
                st = L.mf_box_rhs(ctx.h, abi.make_box(lo, hi), v_sb, v_ds, None, 0.0)
            else:
                st = L.mf_box_rhs_noalloc(ctx.h, abi.make_box(lo, hi), v_sb, v_ds, 0.0)
            assert st == abi.MF_OK, (name, st)
            ctx.sync()
            outs.append(d_ds.cpu().numpy())
        a, b = outs
        same = np.array_equal(a.view(np.int64), b.view(np.int64))
        say("box %s..%s: no-alloc equals mf_box_rhs bit for bit: %s" % (lo, hi, same))
        assert np.all(np.isfinite(a)) and same
    # 3 * 65 * 64 * 64 * 10 doubles fit, 3 * 73 * 72 * 72 * 10 do not
    big = ((0, 0, 0), (71, 71, 71))
    d_big = torch.zeros((nv, 72, 72, 72), dtype=torch.float64, device="cuda:0")
    d_sbig = torch.ones((nv, 78, 78, 78), dtype=torch.float64, device="cuda:0")
    v_big = abi.make_array4(d_big.data_ptr(), (0, 0, 0), (72, 72, 72), nv)
    v_sbig = abi.make_array4(d_sbig.data_ptr(), (-3, -3, -3), (75, 75, 75), nv)
    assert L.mf_box_rhs_noalloc(ctx.h, abi.make_box(*big), v_sbig, v_big, 0.0) == abi.MF_ERR_TOO_LARGE
    ctx.sync()
    assert float(d_big.abs().max()) == 0.0                                   # nothing was launched
    ctx.close()
