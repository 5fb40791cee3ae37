"""Python host side of the C-ABI (libmesoflow_amd.so): context and level objects.

Mirrors the reference's per-level driver (mflo::EvolveAMR, Source/mflo_evolve.cpp:18-89)
and its per-box advance (mflo::compute_dsdt_flow body, :579-652) on top of the
C entry points declared in include/mesoflow_amd.h. There is no fallback: if the
HIP library or a GPU is missing every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
# MESOFLOW_AMD_LIB selects another build of the same library (e.g. one compiled
# with different tuning macros); default: the in-tree build
LIB_PATH = os.environ.get("MESOFLOW_AMD_LIB", os.path.join(_HERE, "libmesoflow_amd.so"))
_LIB = None

# every function of include/mesoflow_amd.h with its argument types
_P = C.POINTER
SIGNATURES = {
    "mf_status_string": ([C.c_int], C.c_char_p),
    "mf_device_count": ([_P(C.c_int)], C.c_int),
    "mf_ctx_create": ([_P(abi.PluginParams), _P(abi.Params), _P(abi.Geom), _P(C.c_void_p)], C.c_int),
    "mf_ctx_destroy": ([C.c_void_p], C.c_int),
    "mf_plugin_defaults": ([C.c_int, _P(abi.PluginParams)], C.c_int),
    "mf_plugin_name": ([C.c_int, _P(C.c_char_p)], C.c_int),
    "mf_ctx_set_plugin_params": ([C.c_void_p, C.c_void_p], C.c_int),
    "mf_ctx_set_stream": ([C.c_void_p, C.c_void_p], C.c_int),
    "mf_ctx_nvars": ([C.c_void_p, _P(C.c_int)], C.c_int),
    "mf_ctx_sync": ([C.c_void_p], C.c_int),
    "mf_box_cutcells": ([C.c_void_p, abi.Box, abi.Array4, C.c_double], C.c_int),
    "mf_box_rhs": ([C.c_void_p, abi.Box, abi.Array4, abi.Array4, _P(abi.Array4), C.c_double], C.c_int),
    "mf_box_rhs_noalloc": ([C.c_void_p, abi.Box, abi.Array4, abi.Array4, C.c_double], C.c_int),
    "mf_box_rk3_combine": ([C.c_void_p, C.c_int, abi.Box, abi.Array4, abi.Array4, abi.Array4, C.c_double], C.c_int),
    "mf_box_post": ([C.c_void_p, abi.Box, abi.Array4, C.c_double, C.c_int], C.c_int),
    "mf_box_average_down": ([C.c_void_p, abi.Box, abi.Array4, abi.Array4, C.c_int], C.c_int),
    "mf_box_time_interp": ([C.c_void_p, abi.Box, abi.Array4, abi.Array4, abi.Array4, C.c_double, C.c_double,
                            C.c_double], C.c_int),
    "mf_box_interp_cons": ([C.c_void_p, abi.Box, abi.Array4, abi.Array4, C.c_int], C.c_int),
    "mf_fill_patch_two_levels": ([C.c_void_p, abi.Box, abi.Array4, _P(abi.Box), _P(abi.Array4), C.c_int, abi.Array4,
                                  abi.Array4, C.c_double, C.c_double, C.c_double, C.c_int], C.c_int),
    "mf_fluxreg_crse_add": ([C.c_void_p, C.c_int, abi.Box, abi.Array4, abi.Array4, C.c_double], C.c_int),
    "mf_fluxreg_fine_add": ([C.c_void_p, C.c_int, abi.Box, abi.Array4, abi.Array4, C.c_int, C.c_double], C.c_int),
    "mf_fluxreg_reflux": ([C.c_void_p, C.c_int, C.c_int, abi.Box, abi.Array4, abi.Array4, C.c_double, C.c_double],
                          C.c_int),
    "mf_box_tag": ([C.c_void_p, abi.Box, abi.Array4, C.c_void_p, _P(C.c_double), _P(C.c_double), _P(C.c_int), C.c_int,
                    C.c_int], C.c_int),
    "mf_level_tag": ([C.c_void_p, C.c_double, _P(C.c_double), _P(C.c_double), _P(C.c_int), C.c_int, C.c_void_p,
                      _P(C.c_int64)], C.c_int),
    "mf_tags_to_boxes": ([C.c_void_p, _P(C.c_int32), _P(C.c_int32), C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _P(abi.Box),
                          C.c_int, _P(C.c_int)], C.c_int),
    "mf_level_create": ([C.c_void_p, _P(abi.Box), C.c_int, _P(C.c_void_p)], C.c_int),
    "mf_level_destroy": ([C.c_void_p], C.c_int),
    "mf_level_set_state": ([C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "mf_level_get_state": ([C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "mf_level_state_view": ([C.c_void_p, _P(abi.Array4)], C.c_int),
    "mf_fill_ghosts_single_level": ([C.c_void_p, C.c_double], C.c_int),
    "mf_est_dt": ([C.c_void_p, _P(C.c_double)], C.c_int),
    "mf_level_step": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "mf_level_step_timed": ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
    "mf_level_evolve": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, _P(C.c_int)], C.c_int),
    "mf_level_advance_chemistry": ([C.c_void_p, C.c_double, C.c_double, C.c_double], C.c_int),
    "mf_level_evolve_split": ([C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int],
                              C.c_int),
    "mf_level_info": ([C.c_void_p, _P(C.c_double), _P(C.c_double), _P(C.c_int), C.c_void_p], C.c_int),
    "mf_level_set_time": ([C.c_void_p, C.c_double, C.c_double, C.c_int], C.c_int),
    "mf_level_boxes": ([C.c_void_p, _P(abi.Box), C.c_int, _P(C.c_int), _P(abi.Geom), _P(C.c_int)], C.c_int),
    "mf_level_write_plotfile": ([C.c_void_p, C.c_char_p, _P(C.c_char_p)], C.c_int),
    "mf_level_write_checkpoint": ([C.c_void_p, C.c_char_p, _P(C.c_char_p)], C.c_int),
    "mf_level_read_checkpoint": ([C.c_void_p, C.c_char_p, _P(C.c_char_p)], C.c_int),
    "mf_write_plotfile": ([C.c_char_p, _P(abi.Geom), _P(abi.Box), C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_int,
                           _P(C.c_char_p)], C.c_int),
    "mf_write_checkpoint": ([C.c_char_p, _P(abi.Geom), _P(abi.Box), C.c_int, C.c_int, C.c_void_p, C.c_double,
                             C.c_double, C.c_int, _P(C.c_char_p)], C.c_int),
    "mf_read_checkpoint": ([C.c_char_p, _P(abi.Geom), _P(abi.Box), C.c_int, C.c_int, C.c_void_p, _P(C.c_double),
                            _P(C.c_double), _P(C.c_int), _P(C.c_char_p)], C.c_int),
    "mf_level_set_tile": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "mf_level_get_tile": ([C.c_void_p, C.c_int, _P(C.c_int), _P(C.c_int), _P(C.c_int)], C.c_int),
    "mf_level_autotune": ([C.c_void_p], C.c_int),
    "mf_adjoint_forward": ([C.c_void_p, C.c_int, C.c_int, C.c_int, _P(C.c_double)], C.c_int),
    "mf_adjoint_backward": ([C.c_void_p, C.c_void_p, _P(C.c_double)], C.c_int),
    "mf_adjoint_state_grad": ([C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "mf_hier_create": ([C.c_void_p, C.c_int, _P(C.c_void_p)], C.c_int),
    "mf_hier_destroy": ([C.c_void_p], C.c_int),
    "mf_hier_add_level": ([C.c_void_p, _P(abi.Box), C.c_int, C.c_void_p, C.c_int, _P(C.c_void_p)], C.c_int),
    "mf_hier_level": ([C.c_void_p, C.c_int, _P(C.c_void_p)], C.c_int),
    "mf_hier_nlevels": ([C.c_void_p, _P(C.c_int)], C.c_int),
    "mf_hier_set_reflux": ([C.c_void_p, C.c_int], C.c_int),
    "mf_hier_autotune": ([C.c_void_p], C.c_int),
    "mf_hier_set_regrid": ([C.c_void_p, C.c_int, _P(C.c_double), _P(C.c_double), _P(C.c_int), C.c_int, C.c_int, C.c_int,
                            C.c_int, C.c_double], C.c_int),
    "mf_hier_regrid": ([C.c_void_p, _P(abi.Box), C.c_int], C.c_int),
    "mf_hier_step": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "mf_hier_evolve": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, _P(C.c_int)], C.c_int),
    "mf_hier_info": ([C.c_void_p, _P(C.c_double), _P(C.c_double), _P(C.c_int), C.c_void_p], C.c_int),
    "mf_hier_write_plotfile": ([C.c_void_p, C.c_char_p, _P(C.c_char_p)], C.c_int),
    "mf_hier_write_checkpoint": ([C.c_void_p, C.c_char_p, _P(C.c_char_p)], C.c_int),
    "mf_hier_read_checkpoint": ([C.c_void_p, C.c_char_p, _P(C.c_char_p)], C.c_int),
    "mf_util_copy": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64], C.c_int),
    "mf_util_div_check": ([C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p], C.c_int),
}


def lib():
    """Loads the HIP library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmesoflow_amd.so is not built; run __graft_entry__.build()")
        # torch ships its own HIP runtime under the same sonames
        # (libamdhip64.so.7, libhsa-runtime64.so.1) as the ROCm one this
        # library links: whichever loads first serves the whole process. Loading
        # torch first keeps torch on the runtime it was built with (this library
        # runs on it as well); loading this library first would bind torch to the
        # system runtime, on which torch then finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def check(st, where):
    if st != abi.MF_OK:
        raise abi.MFError(where, st)


def plugin_defaults(kind):
    """The starting table of a case plugin family (mf_plugin_defaults)."""
    t = abi.PluginParams()
    check(lib().mf_plugin_defaults(int(kind), C.byref(t)), "mf_plugin_defaults")
    return t


def plugin_name(kind):
    n = C.c_char_p()
    check(lib().mf_plugin_name(int(kind), C.byref(n)), "mf_plugin_name")
    return n.value.decode()


def device_count():
    n = C.c_int(0)
    lib().mf_device_count(C.byref(n))
    return n.value


class Context:
    """mf_ctx: plugin tables, mflo.* parameters and geometry on the current HIP device."""

    def __init__(self, case):
        self.case = case
        h = C.c_void_p()
        check(lib().mf_ctx_create(C.byref(case.plugin), C.byref(case.params), C.byref(case.geom), C.byref(h)),
              "mf_ctx_create")
        self.h = h
        nv = C.c_int()
        lib().mf_ctx_nvars(self.h, C.byref(nv))
        self.nv = nv.value

    def close(self):
        if getattr(self, "h", None):
            lib().mf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        st = lib().mf_ctx_sync(self.h)
        check(st, "device work")

    def set_stream(self, stream_ptr):
        check(lib().mf_ctx_set_stream(self.h, C.c_void_p(stream_ptr)), "mf_ctx_set_stream")

    def set_plugin_params(self, plugin):
        """Replace the plugin table (same family and species counts)."""
        check(lib().mf_ctx_set_plugin_params(self.h, C.byref(plugin)), "mf_ctx_set_plugin_params")


class Level:
    """One single-level AMR level (mf_level_*): boxes tiling the domain, state in HBM."""

    def __init__(self, ctx, boxes=None):
        self.ctx = ctx
        case = ctx.case
        if boxes is None:
            arr, nb = case.box_structs()
        else:
            arr = (abi.Box * len(boxes))()
            for i, (lo, hi) in enumerate(boxes):
                arr[i] = abi.make_box(lo, hi)
            nb = len(boxes)
        h = C.c_void_p()
        check(lib().mf_level_create(ctx.h, arr, nb, C.byref(h)), "mf_level_create")
        self.h = h
        self.nboxes = nb
        self.shape = (ctx.nv,) + tuple(case.n_cell[::-1])

    def close(self):
        if getattr(self, "h", None):
            lib().mf_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, s):
        if hasattr(s, "data_ptr"):          # a torch tensor on the device
            assert tuple(s.shape) == self.shape and s.is_contiguous()
            check(lib().mf_level_set_state(self.h, C.c_void_p(s.data_ptr()), 1), "mf_level_set_state")
            return
        s = np.ascontiguousarray(s, dtype=np.float64)
        assert s.shape == self.shape, (s.shape, self.shape)
        check(lib().mf_level_set_state(self.h, s.ctypes.data_as(C.c_void_p), 0), "mf_level_set_state")

    def get_state(self, out=None):
        if out is not None and hasattr(out, "data_ptr"):
            check(lib().mf_level_get_state(self.h, C.c_void_p(out.data_ptr()), 1), "mf_level_get_state")
            return out
        o = np.empty(self.shape)
        check(lib().mf_level_get_state(self.h, o.ctypes.data_as(C.c_void_p), 0), "mf_level_get_state")
        return o

    def step(self, n=1, mode=abi.MODE_BATCHED):
        check(lib().mf_level_step(self.h, int(n), int(mode)), "mf_level_step")

    KERNEL_CLASSES = ("compute_dt", "fill_ghosts", "cutcells", "face_flux", "residual_combine", "finalize", "post_step")

    def evolve(self, final_time, max_step=2 ** 30, only_flow=True, mode=abi.MODE_BATCHED):
        """EvolveAMR(final_time, only_flow); returns the number of steps taken."""
        n = C.c_int()
        check(lib().mf_level_evolve(self.h, float(final_time), int(max_step), int(bool(only_flow)), int(mode), C.byref(n)),
              "mf_level_evolve")
        return n.value

    def advance_chemistry(self, react_time, dt_react, dt_react_rk):
        """Chemistry phase of the split mode (substeps + update_vars_after_chemsolve)."""
        check(lib().mf_level_advance_chemistry(self.h, float(react_time), float(dt_react), float(dt_react_rk)),
              "mf_level_advance_chemistry")

    def evolve_split(self, t_ss, t_react, dt_react, dt_react_rk, dt_ss, max_step=2 ** 30, mode=abi.MODE_BATCHED):
        """Evolve_split: flow to t_ss, then chemistry phases and flow-coupling intervals."""
        check(lib().mf_level_evolve_split(self.h, float(t_ss), float(t_react), float(dt_react), float(dt_react_rk),
                                          float(dt_ss), int(max_step), int(mode)), "mf_level_evolve_split")

    def step_timed(self, n=1, mode=abi.MODE_BATCHED):
        """Steps with HIP events around each kernel class; returns {class: (ms, launches)}."""
        ms = np.zeros(7)
        cnt = np.zeros(7, dtype=np.int32)
        check(lib().mf_level_step_timed(self.h, int(n), int(mode), ms.ctypes.data_as(C.c_void_p),
                                        cnt.ctypes.data_as(C.c_void_p)), "mf_level_step_timed")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.KERNEL_CLASSES)}

    def est_dt(self):
        d = C.c_double()
        check(lib().mf_est_dt(self.h, C.byref(d)), "mf_est_dt")
        return d.value

    def fill_ghosts(self, time=0.0):
        check(lib().mf_fill_ghosts_single_level(self.h, float(time)), "mf_fill_ghosts_single_level")

    def padded_state(self):
        """Copy of the padded level array (ghost layers included) as numpy."""
        import torch
        v = abi.Array4()
        check(lib().mf_level_state_view(self.h, C.byref(v)), "mf_level_state_view")
        self.ctx.sync()
        n = [v.end[d] - v.begin[d] for d in range(3)]
        count = v.nstride * v.ncomp
        # device copy into a torch buffer on the level's stream, then to the host
        t = torch.empty(count, dtype=torch.float64, device="cuda")
        check(lib().mf_util_copy(self.ctx.h, C.c_void_p(t.data_ptr()), C.c_void_p(v.p), C.c_int64(count)), "mf_util_copy")
        self.ctx.sync()
        return t.cpu().numpy().reshape(v.ncomp, n[2], n[1], n[0])

    def info(self):
        t, dt, ist = C.c_double(), C.c_double(), C.c_int()
        res = np.zeros(self.ctx.nv)
        check(lib().mf_level_info(self.h, C.byref(t), C.byref(dt), C.byref(ist), res.ctypes.data_as(C.c_void_p)),
              "mf_level_info")
        return t.value, dt.value, ist.value, res

    def set_time(self, t, dt_prev, istep):
        check(lib().mf_level_set_time(self.h, float(t), float(dt_prev), int(istep)), "mf_level_set_time")

    def boxes(self):
        nb, nv = C.c_int(), C.c_int()
        g = abi.Geom()
        check(lib().mf_level_boxes(self.h, None, 0, C.byref(nb), C.byref(g), C.byref(nv)), "mf_level_boxes")
        arr = (abi.Box * nb.value)()
        check(lib().mf_level_boxes(self.h, arr, nb.value, C.byref(nb), None, None), "mf_level_boxes")
        return [(tuple(b.lo), tuple(b.hi)) for b in arr]

    @staticmethod
    def _names(varnames):
        if varnames is None:
            return None
        arr = (C.c_char_p * len(varnames))(*[v.encode() for v in varnames])
        return arr

    def write_plotfile(self, path, varnames=None):
        """WritePlotFile (Source/mflo_plot.cpp:45-56)"""
        check(lib().mf_level_write_plotfile(self.h, os.fsencode(path), self._names(varnames)),
              "mf_level_write_plotfile")

    def write_checkpoint(self, path, varnames=None):
        """WriteCheckpointFile (Source/mflo_plot.cpp:57-145)"""
        check(lib().mf_level_write_checkpoint(self.h, os.fsencode(path), self._names(varnames)),
              "mf_level_write_checkpoint")

    def read_checkpoint(self, path, varnames=None):
        """ReadCheckpointFile (Source/mflo_plot.cpp:147-239) into this level"""
        check(lib().mf_level_read_checkpoint(self.h, os.fsencode(path), self._names(varnames)),
              "mf_level_read_checkpoint")

    TILE_CLASSES = ("faces_x", "faces_y", "faces_z", "cells")

    def set_tile(self, cls, tx, ty, tz):
        check(lib().mf_level_set_tile(self.h, int(cls), int(tx), int(ty), int(tz)), "mf_level_set_tile")

    def get_tile(self, cls):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().mf_level_get_tile(self.h, int(cls), C.byref(a), C.byref(b), C.byref(c)), "mf_level_get_tile")
        return a.value, b.value, c.value

    def tiles(self):
        return [self.get_tile(q) for q in range(4)]

    def set_tiles(self, tiles):
        for q, t in enumerate(tiles):
            self.set_tile(q, *t)

    def autotune(self):
        check(lib().mf_level_autotune(self.h), "mf_level_autotune")
        return self.tiles()


    # ----- discrete adjoint ----------------------------------------------------
    def adjoint_forward(self, nsteps, ncheck, spec):
        """Run nsteps fixed-dt steps keeping binomial checkpoints; returns the
        objective J (time-integrated outlet mass flux of species `spec`)."""
        J = C.c_double()
        check(lib().mf_adjoint_forward(self.h, int(nsteps), int(ncheck), int(spec), C.byref(J)), "mf_adjoint_forward")
        return J.value

    def adjoint_backward(self):
        """Reverse sweep of the last adjoint_forward: (dJ/dlog_A[nreact], dJ/din_vel)."""
        nr = int(self.ctx.case.plugin.nreact)
        g = np.zeros(max(nr, 1), dtype=np.float64)
        gv = C.c_double()
        check(lib().mf_adjoint_backward(self.h, g.ctypes.data_as(C.c_void_p), C.byref(gv)), "mf_adjoint_backward")
        return g[:nr].copy(), gv.value

    def adjoint_state_grad(self):
        """dJ/d(initial state) of the last backward sweep, (nv, nz, ny, nx)."""
        nx, ny, nz = self.ctx.case.n_cell
        out = np.empty((self.ctx.nv, nz, ny, nx), dtype=np.float64)
        check(lib().mf_adjoint_state_grad(self.h, out.ctypes.data_as(C.c_void_p), 0), "mf_adjoint_state_grad")
        return out




def _box_array(boxes):
    boxes = list(boxes)
    arr = (abi.Box * max(len(boxes), 1))()
    for i, (lo, hi) in enumerate(boxes):
        arr[i] = abi.make_box(lo, hi)
    return arr, len(boxes)


class _LevelView(Level):
    """A level owned by a hierarchy (level 1): the Level calls on its handle,
    released with the hierarchy."""

    def __init__(self, h, ctx, shape, nboxes):
        self.h = h
        self.ctx = ctx
        self.shape = shape
        self.nboxes = nboxes

    def close(self):
        self.h = None


class Hierarchy:
    """Two-level hierarchy (mf_hier_*) over a Level as level 0: mflo's
    EvolveAMR with subcycling, flux registers, reflux, average_down and
    regrid, stepped on the device."""

    def __init__(self, level0, ratio=2):
        self.L0 = level0
        self.ratio = ratio
        h = C.c_void_p()
        check(lib().mf_hier_create(level0.h, int(ratio), C.byref(h)), "mf_hier_create")
        self.h = h
        nv = level0.shape[0]
        self.fine_shape = (nv,) + tuple(q * ratio for q in level0.shape[1:])

    def close(self):
        if getattr(self, "h", None):
            lib().mf_hier_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_level(self, boxes, state):
        """MakeNewLevelFromScratch: level 1 over boxes ((lo, hi) in fine
        indices) with its initial data, dense over the fine domain (numpy or a
        device tensor)."""
        arr, nb = _box_array(boxes)
        h = C.c_void_p()
        if hasattr(state, "data_ptr"):
            assert tuple(state.shape) == self.fine_shape and state.is_contiguous()
            check(lib().mf_hier_add_level(self.h, arr, nb, C.c_void_p(state.data_ptr()), 1, C.byref(h)),
                  "mf_hier_add_level")
        else:
            s = np.ascontiguousarray(state, dtype=np.float64)
            assert s.shape == self.fine_shape
            check(lib().mf_hier_add_level(self.h, arr, nb, s.ctypes.data_as(C.c_void_p), 0, C.byref(h)),
                  "mf_hier_add_level")

    def nlevels(self):
        n = C.c_int()
        check(lib().mf_hier_nlevels(self.h, C.byref(n)), "mf_hier_nlevels")
        return n.value

    def level(self, lev):
        if lev == 0:
            return self.L0
        h = C.c_void_p()
        check(lib().mf_hier_level(self.h, int(lev), C.byref(h)), "mf_hier_level")
        nb = C.c_int()
        check(lib().mf_level_boxes(h, None, 0, C.byref(nb), None, None), "mf_level_boxes")
        return _LevelView(h, self.L0.ctx, self.fine_shape, nb.value)

    def get_state(self, lev):
        return self.level(lev).get_state()

    def set_reflux(self, flag):
        check(lib().mf_hier_set_reflux(self.h, int(bool(flag))), "mf_hier_set_reflux")

    def autotune(self):
        """tile shapes of each level tuned on its own boxes"""
        check(lib().mf_hier_autotune(self.h), "mf_hier_autotune")

    def set_regrid(self, regrid_int, refine_phi, refine_phigrad, comps, n_error_buf=1, blocking_factor=8,
                   max_grid_size=32, grid_eff=0.7):
        n = len(comps)
        ph = (C.c_double * max(n, 1))(*refine_phi)
        gr = (C.c_double * max(n, 1))(*refine_phigrad)
        cp = (C.c_int * max(n, 1))(*comps)
        check(lib().mf_hier_set_regrid(self.h, int(regrid_int), ph, gr, cp, n, int(n_error_buf), int(blocking_factor),
                                       int(max_grid_size), float(grid_eff)), "mf_hier_set_regrid")

    def regrid(self, boxes):
        arr, nb = _box_array(boxes)
        check(lib().mf_hier_regrid(self.h, arr, nb), "mf_hier_regrid")

    def step(self, n=1, mode=abi.MODE_BATCHED):
        check(lib().mf_hier_step(self.h, int(n), int(mode)), "mf_hier_step")

    def evolve(self, final_time, max_step=2 ** 30, only_flow=False, mode=abi.MODE_BATCHED):
        n = C.c_int()
        check(lib().mf_hier_evolve(self.h, float(final_time), int(max_step), int(bool(only_flow)), int(mode),
                                   C.byref(n)), "mf_hier_evolve")
        return n.value

    def info(self):
        t, dt, ist = C.c_double(), C.c_double(), C.c_int()
        res = np.zeros(self.L0.shape[0])
        check(lib().mf_hier_info(self.h, C.byref(t), C.byref(dt), C.byref(ist), res.ctypes.data_as(C.c_void_p)),
              "mf_hier_info")
        return t.value, dt.value, ist.value, res

    def boxes(self, lev=1):
        return self.level(lev).boxes()

    def write_plotfile(self, path, varnames=None):
        check(lib().mf_hier_write_plotfile(self.h, os.fsencode(str(path)), Level._names(varnames)),
              "mf_hier_write_plotfile")

    def write_checkpoint(self, path, varnames=None):
        check(lib().mf_hier_write_checkpoint(self.h, os.fsencode(str(path)), Level._names(varnames)),
              "mf_hier_write_checkpoint")

    def read_checkpoint(self, path, varnames=None):
        check(lib().mf_hier_read_checkpoint(self.h, os.fsencode(str(path)), Level._names(varnames)),
              "mf_hier_read_checkpoint")


