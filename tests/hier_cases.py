"""Cases and runners of the two-level tests (CPU oracle and GPU): periodic Sod
with a shear layer, the fine data of a level made from the coarse one
(cell_cons_interp of each fine box, the coarse level averaged down over it)."""
import numpy as np

from mesoflow_amd import abi, cases
from oracle import pyoracle as po


def cons_comps(nv):
    return [0, 1, 2, 3, 4] + list(range(abi.FLO_NVARS, nv))


def all_periodic(n, boxes=None, cfl=0.4):
    c = cases.shocktube_n2(n_cell=(n, n, n), order_hyp=5, cfl=cfl)
    for d in range(3):
        c.geom.is_periodic[d] = 1
        c.params.lo_bc[d] = abi.BC_PERIODIC
        c.params.hi_bc[d] = abi.BC_PERIODIC
    if boxes is not None:
        c.boxes = boxes
    return c


def shear_state(cc, amp=0.2):
    s0 = cc.make_state()
    n = cc.n_cell[1]
    y = (np.arange(n) + 0.5) / n
    v = amp * np.sin(2 * np.pi * y)[None, :, None] * np.ones(s0.shape[1:])
    s0[abi.VELY] = v
    s0[abi.RHOV] = s0[abi.DENS] * v
    s0[abi.RHOE] += 0.5 * s0[abi.DENS] * v * v
    return s0


def fine_init(cstate, fboxes, ratio=2, periodic=(1, 1, 1)):
    """(coarse averaged down, fine dense over the fine domain)"""
    nv, nz, ny, nx = cstate.shape
    c = np.ascontiguousarray(cstate.copy())
    f = np.zeros((nv, nz * ratio, ny * ratio, nx * ratio))
    pad = c
    for ax, d in ((3, 0), (2, 1), (1, 2)):
        w = [(0, 0)] * 4
        w[ax] = (1, 1)
        pad = np.pad(pad, w, mode="wrap" if periodic[d] else "edge")
    pad = np.ascontiguousarray(pad)
    for lo, hi in fboxes:
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        fab = np.zeros((nv, n[2], n[1], n[0]))
        po.interp_cons((lo, hi), (pad, (-1, -1, -1)), (fab, lo), ratio)
        f[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = fab
    for lo, hi in fboxes:
        clo = tuple(q // ratio for q in lo)
        chi = tuple(q // ratio for q in hi)
        po.average_down((clo, chi), (np.ascontiguousarray(f), (0, 0, 0)), (c, (0, 0, 0)), ratio)
    return c, f


def fine_mask(shape, fboxes):
    m = np.zeros(shape[1:], dtype=bool)
    for lo, hi in fboxes:
        m[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return m
