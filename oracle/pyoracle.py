"""TEST INFRASTRUCTURE ONLY: ctypes loader of the CPU oracle (liboracle.so).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from mesoflow_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force=False):
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_level_create.argtypes = [P(abi.PluginParams), P(abi.Params), P(abi.Geom), P(abi.Box), C.c_int,
                                       P(C.c_void_p)]
        L.orc_level_create.restype = C.c_int
        L.orc_level_destroy.argtypes = [C.c_void_p]
        L.orc_level_set_state.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_level_get_state.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_level_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_level_step.restype = C.c_int
        L.orc_level_evolve.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, P(C.c_int)]
        L.orc_level_evolve.restype = C.c_int
        L.orc_level_advance_chemistry.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_level_advance_chemistry.restype = C.c_int
        L.orc_level_evolve_split.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.c_int]
        L.orc_level_evolve_split.restype = C.c_int
        L.orc_level_est_dt.argtypes = [C.c_void_p, P(C.c_double)]
        L.orc_level_est_dt.restype = C.c_int
        L.orc_level_fill_ghosts.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
        L.orc_level_fill_ghosts.restype = C.c_int
        L.orc_level_info.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_int), C.c_void_p]
        L.orc_level_set_time.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.orc_box_rhs.argtypes = [P(abi.PluginParams), P(abi.Params), P(abi.Geom), abi.Box, abi.Array4,
                                  abi.Array4, P(abi.Array4), C.c_double]
        L.orc_box_rhs.restype = C.c_int
        L.orc_box_cutcells.argtypes = [P(abi.PluginParams), P(abi.Params), P(abi.Geom), abi.Box, abi.Array4]
        L.orc_box_cutcells.restype = C.c_int
        L.orc_box_post.argtypes = [P(abi.PluginParams), P(abi.Params), abi.Box, abi.Array4, C.c_double, C.c_int]
        L.orc_box_post.restype = C.c_int
        L.orc_fill_physbc.argtypes = [P(abi.PluginParams), P(abi.Params), P(abi.Geom), abi.Box, abi.Array4, C.c_double]
        L.orc_fill_physbc.restype = C.c_int
        L.orc_kat.argtypes = [P(abi.PluginParams), C.c_int, C.c_void_p, C.c_void_p]
        L.orc_kat.restype = C.c_int
        L.orc_hier_create.argtypes = [P(abi.PluginParams), C.c_void_p, C.c_int, P(C.c_void_p)]
        L.orc_hier_create.restype = C.c_int
        L.orc_hier_destroy.argtypes = [C.c_void_p]
        L.orc_hier_add_level.argtypes = [C.c_void_p, P(abi.Box), C.c_int, C.c_void_p]
        L.orc_hier_add_level.restype = C.c_int
        L.orc_hier_regrid.argtypes = [C.c_void_p, P(abi.Box), C.c_int, C.c_double]
        L.orc_hier_regrid.restype = C.c_int
        L.orc_hier_step.argtypes = [C.c_void_p, C.c_int, C.c_int, P(abi.Box), C.c_int]
        L.orc_hier_step.restype = C.c_int
        L.orc_hier_level.argtypes = [C.c_void_p, C.c_int]
        L.orc_hier_level.restype = C.c_void_p
        L.orc_hier_nlevels.argtypes = [C.c_void_p]
        L.orc_hier_nlevels.restype = C.c_int
        L.orc_hier_info.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_int), C.c_void_p]
        L.orc_hier_set_only_flow.argtypes = [C.c_void_p, C.c_int]
        L.orc_hier_set_reflux.argtypes = [C.c_void_p, C.c_int]
        A4, B = abi.Array4, abi.Box
        for name, args in (("orc_average_down", [B, A4, A4, C.c_int, C.c_int]),
                           ("orc_time_interp", [B, A4, A4, A4, C.c_double, C.c_double, C.c_double, C.c_int]),
                           ("orc_interp_cons", [B, A4, A4, C.c_int, C.c_int]),
                           ("orc_fluxreg_crse_add", [C.c_int, B, A4, A4, C.c_double, C.c_int]),
                           ("orc_fluxreg_fine_add", [C.c_int, B, A4, A4, C.c_int, C.c_double, C.c_int]),
                           ("orc_fluxreg_reflux", [C.c_int, C.c_int, B, A4, A4, C.c_double, C.c_double, C.c_int]),
                           ("orc_box_tag", [B, A4, C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_int), C.c_int,
                                            C.c_int])):
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleLevel:
    """One single-level AMR level of the oracle, mirroring mf_level_* of the C-ABI."""

    def __init__(self, case):
        self.case = case
        boxes, nb = case.box_structs()
        h = C.c_void_p()
        st = lib().orc_level_create(C.byref(case.plugin), C.byref(case.params), C.byref(case.geom), boxes, nb,
                                    C.byref(h))
        if st != 0:
            raise abi.MFError("orc_level_create", st)
        self.h = h
        self.nv = case.nv
        self.shape = (self.nv,) + tuple(case.n_cell[::-1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_level_destroy(self.h)
            self.h = None

    def set_state(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        assert s.shape == self.shape
        lib().orc_level_set_state(self.h, _ptr(s))

    def get_state(self):
        out = np.empty(self.shape)
        lib().orc_level_get_state(self.h, _ptr(out))
        return out

    def step(self, n=1):
        st = lib().orc_level_step(self.h, int(n))
        if st != 0:
            raise abi.MFError("orc_level_step", st)

    def evolve(self, final_time, max_step=2 ** 30, only_flow=True):
        """EvolveAMR(final_time, only_flow); returns the number of steps taken."""
        n = C.c_int()
        st = lib().orc_level_evolve(self.h, float(final_time), int(max_step), int(bool(only_flow)), C.byref(n))
        if st != 0:
            raise abi.MFError("orc_level_evolve", st)
        return n.value

    def advance_chemistry(self, react_time, dt_react, dt_react_rk):
        st = lib().orc_level_advance_chemistry(self.h, float(react_time), float(dt_react), float(dt_react_rk))
        if st != 0:
            raise abi.MFError("orc_level_advance_chemistry", st)

    def evolve_split(self, t_ss, t_react, dt_react, dt_react_rk, dt_ss, max_step=2 ** 30):
        st = lib().orc_level_evolve_split(self.h, float(t_ss), float(t_react), float(dt_react), float(dt_react_rk),
                                          float(dt_ss), int(max_step))
        if st != 0:
            raise abi.MFError("orc_level_evolve_split", st)

    def est_dt(self):
        d = C.c_double()
        lib().orc_level_est_dt(self.h, C.byref(d))
        return d.value

    def fill_ghosts(self, time=0.0):
        nx, ny, nz = self.case.n_cell
        out = np.zeros((self.nv, nz + 6, ny + 6, nx + 6))
        lib().orc_level_fill_ghosts(self.h, _ptr(out), float(time))
        return out

    def info(self):
        t, dt, ist = C.c_double(), C.c_double(), C.c_int()
        res = np.zeros(self.nv)
        lib().orc_level_info(self.h, C.byref(t), C.byref(dt), C.byref(ist), _ptr(res))
        return t.value, dt.value, ist.value, res

    def set_time(self, t, dt_prev, istep):
        lib().orc_level_set_time(self.h, float(t), float(dt_prev), int(istep))


def _box_array(boxes):
    boxes = list(boxes)
    arr = (abi.Box * max(len(boxes), 1))(*[abi.make_box(lo, hi) for lo, hi in boxes])
    return arr, len(boxes)


class OracleHier:
    """The oracle's two-level hierarchy (mf_hier_oracle.h) over an OracleLevel
    as level 0: EvolveAMR's step with subcycling, flux registers, reflux,
    average_down and regrid."""

    def __init__(self, level0, ratio=2):
        self.L0 = level0
        self.ratio = ratio
        self.nv = level0.nv
        h = C.c_void_p()
        _ok(lib().orc_hier_create(C.byref(level0.case.plugin), level0.h, int(ratio), C.byref(h)), "orc_hier_create")
        self.h = h
        n = level0.case.n_cell
        self.fine_shape = (self.nv,) + tuple(q * ratio for q in n[::-1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_hier_destroy(self.h)
            self.h = None

    def add_level(self, boxes, state):
        """MakeNewLevelFromScratch: level 1 with the boxes ((lo, hi) in fine
        cells) and its initial data, dense over the fine domain"""
        s = np.ascontiguousarray(state, dtype=np.float64)
        assert s.shape == self.fine_shape
        arr, nb = _box_array(boxes)
        _ok(lib().orc_hier_add_level(self.h, arr, nb, _ptr(s)), "orc_hier_add_level")

    def regrid(self, boxes, time):
        arr, nb = _box_array(boxes)
        _ok(lib().orc_hier_regrid(self.h, arr, nb, float(time)), "orc_hier_regrid")

    def step(self, n=1, regrid_boxes=None):
        """n coarse steps; regrid_boxes: a regrid of level 1 at the start of the first"""
        if regrid_boxes is None:
            arr, nb = _box_array([])
            _ok(lib().orc_hier_step(self.h, int(n), 0, arr, 0), "orc_hier_step")
        else:
            arr, nb = _box_array(regrid_boxes)
            _ok(lib().orc_hier_step(self.h, int(n), 1, arr, nb), "orc_hier_step")

    def nlevels(self):
        return lib().orc_hier_nlevels(self.h)

    def get_state(self, lev):
        p = lib().orc_hier_level(self.h, int(lev))
        if not p:
            raise abi.MFError("orc_hier_level", abi.MF_ERR_BAD_ARG)
        out = np.empty(self.L0.shape if lev == 0 else self.fine_shape)
        lib().orc_level_get_state(p, _ptr(out))
        return out

    def info(self):
        t, dt, ist = C.c_double(), C.c_double(), C.c_int()
        res = np.zeros(self.nv)
        lib().orc_hier_info(self.h, C.byref(t), C.byref(dt), C.byref(ist), _ptr(res))
        return t.value, dt.value, ist.value, res

    def set_only_flow(self, flag):
        lib().orc_hier_set_only_flow(self.h, int(bool(flag)))

    def set_reflux(self, flag):
        lib().orc_hier_set_reflux(self.h, int(bool(flag)))


def kat(plugin, func, inp, nout):
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.zeros(nout)
    st = lib().orc_kat(C.byref(plugin), int(func), _ptr(inp), _ptr(out))
    if st != 0:
        raise abi.MFError("orc_kat", st)
    return out


def box_rhs(case, lo, hi, sborder, time=0.0, with_flux=False):
    """Per-box RHS of the oracle on a (NV, nz+6, ny+6, nx+6) Sborder array of the box."""
    nv = case.nv
    sb = np.ascontiguousarray(sborder, dtype=np.float64)
    glo = [lo[d] - 3 for d in range(3)]
    ghi = [hi[d] + 4 for d in range(3)]
    a_sb = abi.make_array4(sb.ctypes.data, glo, ghi, nv)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    dsdt = np.zeros((nv, n[2], n[1], n[0]))
    a_ds = abi.make_array4(dsdt.ctypes.data, lo, [hi[d] + 1 for d in range(3)], nv)
    fl = None
    fluxes = []
    if with_flux:
        fl = (abi.Array4 * 3)()
        for d in range(3):
            fe = [hi[q] + 1 for q in range(3)]
            fe[d] += 1
            f = np.zeros((nv, fe[2] - lo[2], fe[1] - lo[1], fe[0] - lo[0]))
            fluxes.append(f)
            fl[d] = abi.make_array4(f.ctypes.data, lo, fe, nv)
    st = lib().orc_box_rhs(C.byref(case.plugin), C.byref(case.params), C.byref(case.geom),
                           abi.make_box(lo, hi), a_sb, a_ds, fl, float(time))
    if st != 0:
        raise abi.MFError("orc_box_rhs", st)
    return (dsdt, fluxes) if with_flux else dsdt


def box_cutcells(case, lo, hi, sborder):
    sb = np.ascontiguousarray(sborder, dtype=np.float64).copy()
    glo = [lo[d] - 3 for d in range(3)]
    ghi = [hi[d] + 4 for d in range(3)]
    a_sb = abi.make_array4(sb.ctypes.data, glo, ghi, case.nv)
    st = lib().orc_box_cutcells(C.byref(case.plugin), C.byref(case.params), C.byref(case.geom),
                                abi.make_box(lo, hi), a_sb)
    if st != 0:
        raise abi.MFError("orc_box_cutcells", st)
    return sb


# --------------------------------------------------------------------------
# two-level coupling operators (mf_amr_oracle.cpp) on numpy fabs: a fab is
# (array [nc][nz][ny][nx], lo (x, y, z) of its first cell or face)
# --------------------------------------------------------------------------
def _a4(fab):
    a, lo = fab
    assert a.flags.c_contiguous and a.dtype == np.float64
    nc, nz, ny, nx = a.shape
    return abi.make_array4(a.ctypes.data, lo, (lo[0] + nx, lo[1] + ny, lo[2] + nz), nc)


def _ok(st, where):
    if st != 0:
        raise abi.MFError(where, st)


def average_down(cbx, fine, crse, ratio):
    _ok(lib().orc_average_down(abi.make_box(*cbx), _a4(fine), _a4(crse), ratio, crse[0].shape[0]), "orc_average_down")


def time_interp(bx, dst, s_old, s_new, t_old, t_new, time):
    _ok(lib().orc_time_interp(abi.make_box(*bx), _a4(dst), _a4(s_old), _a4(s_new), t_old, t_new, time,
                              dst[0].shape[0]), "orc_time_interp")


def interp_cons(fbx, crse, fine, ratio):
    _ok(lib().orc_interp_cons(abi.make_box(*fbx), _a4(crse), _a4(fine), ratio, fine[0].shape[0]), "orc_interp_cons")


def fluxreg_crse_add(d, face, reg, cflux, scale):
    _ok(lib().orc_fluxreg_crse_add(d, abi.make_box(*face), _a4(reg), _a4(cflux), scale, reg[0].shape[0]),
        "orc_fluxreg_crse_add")


def fluxreg_fine_add(d, face, reg, fflux, ratio, scale):
    _ok(lib().orc_fluxreg_fine_add(d, abi.make_box(*face), _a4(reg), _a4(fflux), ratio, scale, reg[0].shape[0]),
        "orc_fluxreg_fine_add")


def fluxreg_reflux(d, side, face, state, reg, scale, vol):
    _ok(lib().orc_fluxreg_reflux(d, side, abi.make_box(*face), _a4(state), _a4(reg), scale, vol, reg[0].shape[0]),
        "orc_fluxreg_reflux")


def box_tag(bx, state, tags, refine_phi, refine_phigrad, comps, tagval=1):
    """ErrorEst on one box; tags: uint8 array [nz][ny][nx] over bx, updated in place"""
    assert tags.dtype == np.uint8 and tags.flags.c_contiguous
    n = len(comps)
    ph = (C.c_double * max(n, 1))(*refine_phi)
    gr = (C.c_double * max(n, 1))(*refine_phigrad)
    cp = (C.c_int * max(n, 1))(*comps)
    _ok(lib().orc_box_tag(abi.make_box(*bx), _a4(state), tags.ctypes.data, ph, gr, cp, n, tagval), "orc_box_tag")


def fill_physbc(case, region, fab, time):
    """physical-boundary rules on the cells of region of a numpy fab outside the domain (in place)"""
    _ok(lib().orc_fill_physbc(C.byref(case.plugin), C.byref(case.params), C.byref(case.geom), abi.make_box(*region),
                              _a4(fab), float(time)), "orc_fill_physbc")


def box_post(case, bx, fab, dt, flags):
    """chemistry / primitive update / clip on a numpy fab (in place); returns the status"""
    return lib().orc_box_post(C.byref(case.plugin), C.byref(case.params), abi.make_box(*bx), _a4(fab), float(dt),
                              int(flags))
