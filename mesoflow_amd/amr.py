"""Two-level coupling operators on device fabs (mf_box_average_down,
mf_box_time_interp, mf_box_interp_cons, mf_fluxreg_*) and the tagging of
ErrorEst (mf_box_tag, mf_level_tag): the AMReX calls of
Source/mflo_gridwork.cpp:125-193, Source/mflo.cpp:310-330 and
Source/mflo_evolve.cpp:232-248, 655-681. A fab is (tensor [nc][nz][ny][nx]
on the device, lo (x, y, z) of its first cell or face). Work is enqueued on
the context's stream; Context.sync() waits for it."""
import ctypes as C

import numpy as np

from . import abi, runtime


def _a4(fab):
    t, lo = fab
    assert t.is_contiguous() and t.is_cuda
    nc, nz, ny, nx = t.shape
    return abi.make_array4(t.data_ptr(), lo, (lo[0] + nx, lo[1] + ny, lo[2] + nz), nc)


def average_down(ctx, cbx, fine, crse, ratio):
    runtime.check(runtime.lib().mf_box_average_down(ctx.h, abi.make_box(*cbx), _a4(fine), _a4(crse), ratio),
                  "mf_box_average_down")


def time_interp(ctx, bx, dst, s_old, s_new, t_old, t_new, time):
    runtime.check(runtime.lib().mf_box_time_interp(ctx.h, abi.make_box(*bx), _a4(dst), _a4(s_old), _a4(s_new),
                                                   float(t_old), float(t_new), float(time)), "mf_box_time_interp")


def interp_cons(ctx, fbx, crse, fine, ratio):
    runtime.check(runtime.lib().mf_box_interp_cons(ctx.h, abi.make_box(*fbx), _a4(crse), _a4(fine), ratio),
                  "mf_box_interp_cons")


def fill_patch_two_levels(ctx, fbx, sborder, fine_fabs, crse_old, crse_new, t_old, t_new, time, ratio):
    """FillPatchTwoLevels for one fine box: sborder (fbx grown by 3) from the
    coarse states at time and the fine fabs [(box, fab)] (ctx: the fine level's)"""
    nb = len(fine_fabs)
    boxes = (abi.Box * max(nb, 1))()
    views = (abi.Array4 * max(nb, 1))()
    for q, (bx, fab) in enumerate(fine_fabs):
        boxes[q] = abi.make_box(*bx)
        views[q] = _a4(fab)
    runtime.check(runtime.lib().mf_fill_patch_two_levels(ctx.h, abi.make_box(*fbx), _a4(sborder), boxes, views, nb,
                                                         _a4(crse_old), _a4(crse_new), float(t_old), float(t_new),
                                                         float(time), ratio), "mf_fill_patch_two_levels")


def fluxreg_crse_add(ctx, d, face, reg, cflux, scale):
    runtime.check(runtime.lib().mf_fluxreg_crse_add(ctx.h, d, abi.make_box(*face), _a4(reg), _a4(cflux),
                                                    float(scale)), "mf_fluxreg_crse_add")


def fluxreg_fine_add(ctx, d, face, reg, fflux, ratio, scale):
    runtime.check(runtime.lib().mf_fluxreg_fine_add(ctx.h, d, abi.make_box(*face), _a4(reg), _a4(fflux), ratio,
                                                    float(scale)), "mf_fluxreg_fine_add")


def fluxreg_reflux(ctx, d, side, face, state, reg, scale, vol):
    runtime.check(runtime.lib().mf_fluxreg_reflux(ctx.h, d, side, abi.make_box(*face), _a4(state), _a4(reg),
                                                  float(scale), float(vol)), "mf_fluxreg_reflux")


def _tag_params(refine_phi, refine_phigrad, comps):
    n = len(comps)
    return ((C.c_double * max(n, 1))(*refine_phi), (C.c_double * max(n, 1))(*refine_phigrad),
            (C.c_int * max(n, 1))(*comps), n)


def box_tag(ctx, bx, state, tags, refine_phi, refine_phigrad, comps, tagval=1):
    """ErrorEst on one box (Source/mflo.cpp:183-260); tags: uint8 device tensor [nz][ny][nx] over bx"""
    ph, gr, cp, n = _tag_params(refine_phi, refine_phigrad, comps)
    assert tags.is_contiguous() and tags.is_cuda
    runtime.check(runtime.lib().mf_box_tag(ctx.h, abi.make_box(*bx), _a4(state), C.c_void_p(tags.data_ptr()), ph, gr,
                                           cp, n, tagval), "mf_box_tag")


def level_tag(level, time, refine_phi, refine_phigrad, comps):
    """ErrorEst over a level: (tags uint8 [nz][ny][nx] on the host, number tagged)"""
    ph, gr, cp, n = _tag_params(refine_phi, refine_phigrad, comps)
    nx, ny, nz = level.ctx.case.n_cell
    tags = np.zeros((nz, ny, nx), dtype=np.uint8)
    cnt = C.c_int64()
    runtime.check(runtime.lib().mf_level_tag(level.h, float(time), ph, gr, cp, n, tags.ctypes.data_as(C.c_void_p),
                                             C.byref(cnt)), "mf_level_tag")
    return tags, cnt.value


def tags_to_boxes(tags, n_error_buf=1, blocking_factor=8, max_grid_size=32, grid_eff=0.7, ratio=2, periodic=None):
    """new fine-level boxes [(lo, hi)] from host tags [nz][ny][nx] (mf_tags_to_boxes);
    periodic: (px, py, pz) flags of the directions where buffered tags wrap"""
    t = np.ascontiguousarray(tags, dtype=np.uint8)
    nz, ny, nx = t.shape
    n = (C.c_int32 * 3)(nx, ny, nz)
    per = None if periodic is None else (C.c_int32 * 3)(*[int(bool(v)) for v in periodic])
    nb = C.c_int()
    args = (n, per, int(n_error_buf), int(blocking_factor), int(max_grid_size), float(grid_eff), int(ratio))
    lib = runtime.lib()
    runtime.check(lib.mf_tags_to_boxes(t.ctypes.data_as(C.c_void_p), *args, None, 0, C.byref(nb)), "mf_tags_to_boxes")
    out = (abi.Box * max(nb.value, 1))()
    runtime.check(lib.mf_tags_to_boxes(t.ctypes.data_as(C.c_void_p), *args, out, nb.value, C.byref(nb)),
                  "mf_tags_to_boxes")
    return [(tuple(out[q].lo), tuple(out[q].hi)) for q in range(nb.value)]
