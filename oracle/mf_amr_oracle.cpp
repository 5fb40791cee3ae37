// mf_amr_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the two-level coupling operators that mesoflow calls from
// AMReX (Source/mflo_gridwork.cpp:125-193, Source/mflo.cpp:310-330,
// Source/mflo_evolve.cpp:232-248, 655-681): average_down, the time
// interpolation of FillPatch, the conservative linear interpolation
// (cell_cons_interp) and the FluxRegister operations; and the tagging of
// ErrorEst (Source/mflo.cpp:183-260, Source/mflo_tagging.H). AMReX is not in the
// tree: the algorithms are restated from AMReX's published descriptions, so
// parity with AMReX itself is unpinned; tests/test_oracle_amr.py pins these
// functions with closed forms and invariants (exactness on linear data,
// conservation of the interpolation and of a refluxed two-level update, no
// new extrema, restriction of the interpolation = identity).
//
// Views: mf_array4 (component-major, x fastest). Boxes: inclusive cell
// indices. A face box (FluxRegister) holds face indices: along dir its lo ==
// hi is the face index, the other directions are cell ranges.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include "mesoflow_amd.h"

namespace {

struct V {
    double *p;
    int64_t js, ks, ns;
    int b[3];
    double &operator()(int i, int j, int k, int n) const {
        return p[(int64_t)(i - b[0]) + (int64_t)(j - b[1]) * js + (int64_t)(k - b[2]) * ks + (int64_t)n * ns];
    }
};

V mk(const mf_array4 &a) {
    V v;
    v.p = a.p;
    v.js = a.jstride;
    v.ks = a.kstride;
    v.ns = a.nstride;
    for (int d = 0; d < 3; ++d) v.b[d] = a.begin[d];
    return v;
}

bool holds(const mf_array4 &a, const int lo[3], const int hi[3], int nc) {
    if (!a.p || a.ncomp < nc) return false;
    for (int d = 0; d < 3; ++d)
        if (lo[d] < a.begin[d] || hi[d] >= a.end[d] || hi[d] < lo[d]) return false;
    return true;
}

int floor_div(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

// MC-limited slope of one coarse cell along one direction (centred difference
// limited by twice the one-sided differences)
double mc_slope(double um, double u0, double up) {
    const double dc = 0.5 * (up - um);
    const double df = 2.0 * (up - u0);
    const double db = 2.0 * (u0 - um);
    const double lim = (df * db >= 0.0) ? std::fmin(std::fabs(df), std::fabs(db)) : 0.0;
    return std::copysign(1.0, dc) * std::fmin(lim, std::fabs(dc));
}

}  // namespace

extern "C" {

// amrex::average_down, Cartesian (AverageDownTo, Source/mflo_gridwork.cpp:125-130):
// crse(i,j,k,n) = (sum over the r^3 fine cells, k outer, i inner) * (1/r^3)
int orc_average_down(mf_box cbx, mf_array4 fine, mf_array4 crse, int ratio, int ncomp) {
    if (ratio < 1) return MF_ERR_BAD_ARG;
    int flo[3], fhi[3];
    for (int d = 0; d < 3; ++d) { flo[d] = cbx.lo[d] * ratio; fhi[d] = cbx.hi[d] * ratio + ratio - 1; }
    if (!holds(crse, cbx.lo, cbx.hi, ncomp) || !holds(fine, flo, fhi, ncomp)) return MF_ERR_BAD_ARG;
    const V f = mk(fine), c = mk(crse);
    const double volfrac = 1.0 / (double)(ratio * ratio * ratio);
    for (int n = 0; n < ncomp; ++n)
        for (int k = cbx.lo[2]; k <= cbx.hi[2]; ++k)
            for (int j = cbx.lo[1]; j <= cbx.hi[1]; ++j)
                for (int i = cbx.lo[0]; i <= cbx.hi[0]; ++i) {
                    double s = 0.0;
                    for (int kr = 0; kr < ratio; ++kr)
                        for (int jr = 0; jr < ratio; ++jr)
                            for (int ir = 0; ir < ratio; ++ir) s += f(i * ratio + ir, j * ratio + jr, k * ratio + kr, n);
                    c(i, j, k, n) = s * volfrac;
                }
    return MF_OK;
}

// GetData (Source/mflo.cpp:310-330) + FillPatchSingleLevel's time
// interpolation: within teps = 1e-3 (t_new - t_old) of t_new the new data,
// of t_old the old data, else alpha*old + beta*new with
// alpha = (t_new - time)/(t_new - t_old), beta = (time - t_old)/(t_new - t_old)
int orc_time_interp(mf_box bx, mf_array4 dst, mf_array4 s_old, mf_array4 s_new, double t_old, double t_new, double time,
                    int ncomp) {
    if (!holds(dst, bx.lo, bx.hi, ncomp) || !holds(s_old, bx.lo, bx.hi, ncomp) || !holds(s_new, bx.lo, bx.hi, ncomp))
        return MF_ERR_BAD_ARG;
    const V d = mk(dst), o = mk(s_old), w = mk(s_new);
    const double teps = (t_new - t_old) * 1.e-3;
    int mode = 2;
    if (time > t_new - teps && time < t_new + teps) mode = 1;
    else if (time > t_old - teps && time < t_old + teps) mode = 0;
    const double alpha = (mode == 2) ? (t_new - time) / (t_new - t_old) : 0.0;
    const double beta = (mode == 2) ? (time - t_old) / (t_new - t_old) : 0.0;
    for (int n = 0; n < ncomp; ++n)
        for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                    if (mode == 1) d(i, j, k, n) = w(i, j, k, n);
                    else if (mode == 0) d(i, j, k, n) = o(i, j, k, n);
                    else {
                        const double a = alpha * o(i, j, k, n);
                        const double b = beta * w(i, j, k, n);
                        d(i, j, k, n) = a + b;
                    }
                }
    return MF_OK;
}

// cell_cons_interp (the mapper of FillPatchTwoLevels / InterpFromCoarseLevel,
// Source/mflo_gridwork.cpp:151-165, 170-193): per coarse cell and component
// MC-limited slopes in x, y, z, scaled by one factor alpha <= 1 so that no
// fine value leaves [min, max] of the 27 coarse neighbours; a fine cell is
// crse + xoff*sx + yoff*sy + zoff*sz with offsets (i - ic r + 1/2)/r - 1/2.
// crse must hold coarsen(fbx) grown by one cell.
int orc_interp_cons(mf_box fbx, mf_array4 crse, mf_array4 fine, int ratio, int ncomp) {
    if (ratio < 1) return MF_ERR_BAD_ARG;
    int clo[3], chi[3];
    for (int d = 0; d < 3; ++d) { clo[d] = floor_div(fbx.lo[d], ratio) - 1; chi[d] = floor_div(fbx.hi[d], ratio) + 1; }
    if (!holds(fine, fbx.lo, fbx.hi, ncomp) || !holds(crse, clo, chi, ncomp)) return MF_ERR_BAD_ARG;
    const V c = mk(crse), f = mk(fine);
    const double rr = (double)ratio;
    const double dfac = (double)(ratio - 1) / (double)(2 * ratio);
    for (int n = 0; n < ncomp; ++n)
        for (int k = fbx.lo[2]; k <= fbx.hi[2]; ++k)
            for (int j = fbx.lo[1]; j <= fbx.hi[1]; ++j)
                for (int i = fbx.lo[0]; i <= fbx.hi[0]; ++i) {
                    const int ic = floor_div(i, ratio), jc = floor_div(j, ratio), kc = floor_div(k, ratio);
                    const double u = c(ic, jc, kc, n);
                    double sx = mc_slope(c(ic - 1, jc, kc, n), u, c(ic + 1, jc, kc, n));
                    double sy = mc_slope(c(ic, jc - 1, kc, n), u, c(ic, jc + 1, kc, n));
                    double sz = mc_slope(c(ic, jc, kc - 1, n), u, c(ic, jc, kc + 1, n));
                    double alpha = 1.0;
                    if (sx != 0.0 || sy != 0.0 || sz != 0.0) {
                        const double dumax = std::fabs(sx) * dfac + std::fabs(sy) * dfac + std::fabs(sz) * dfac;
                        double umax = u, umin = u;
                        for (int ko = -1; ko <= 1; ++ko)
                            for (int jo = -1; jo <= 1; ++jo)
                                for (int io = -1; io <= 1; ++io) {
                                    const double v = c(ic + io, jc + jo, kc + ko, n);
                                    umax = std::fmax(umax, v);
                                    umin = std::fmin(umin, v);
                                }
                        if (dumax * alpha > (umax - u)) alpha = (umax - u) / dumax;
                        if (dumax * alpha > (u - umin)) alpha = (u - umin) / dumax;
                    }
                    sx *= alpha;
                    sy *= alpha;
                    sz *= alpha;
                    const double xo = ((double)(i - ic * ratio) + 0.5) / rr - 0.5;
                    const double yo = ((double)(j - jc * ratio) + 0.5) / rr - 0.5;
                    const double zo = ((double)(k - kc * ratio) + 0.5) / rr - 0.5;
                    const double a = xo * sx, b = yo * sy, e = zo * sz;
                    f(i, j, k, n) = ((u + a) + b) + e;
                }
    return MF_OK;
}

// FluxRegister::CrseInit(flux, dir, 0, 0, ncomp, scale, ADD)
// (Source/mflo_evolve.cpp:655-667): reg(f) += scale * cflux(f) on the coarse
// faces of the face box
int orc_fluxreg_crse_add(int dir, mf_box face, mf_array4 reg, mf_array4 cflux, double scale, int ncomp) {
    if (dir < 0 || dir > 2 || face.lo[dir] != face.hi[dir]) return MF_ERR_BAD_ARG;
    if (!holds(reg, face.lo, face.hi, ncomp) || !holds(cflux, face.lo, face.hi, ncomp)) return MF_ERR_BAD_ARG;
    const V r = mk(reg), fl = mk(cflux);
    for (int n = 0; n < ncomp; ++n)
        for (int k = face.lo[2]; k <= face.hi[2]; ++k)
            for (int j = face.lo[1]; j <= face.hi[1]; ++j)
                for (int i = face.lo[0]; i <= face.hi[0]; ++i) {
                    const double t = scale * fl(i, j, k, n);
                    r(i, j, k, n) = r(i, j, k, n) + t;
                }
    return MF_OK;
}

// FluxRegister::FineAdd(flux, dir, 0, 0, ncomp, scale) (:668-679): for every
// coarse face of the face box, the ratio^2 fine faces on it, each scaled, are
// added (second transverse index outer, first inner)
int orc_fluxreg_fine_add(int dir, mf_box face, mf_array4 reg, mf_array4 fflux, int ratio, double scale, int ncomp) {
    if (dir < 0 || dir > 2 || face.lo[dir] != face.hi[dir] || ratio < 1) return MF_ERR_BAD_ARG;
    const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
    int flo[3], fhi[3];
    flo[dir] = fhi[dir] = face.lo[dir] * ratio;
    flo[t1] = face.lo[t1] * ratio;
    fhi[t1] = face.hi[t1] * ratio + ratio - 1;
    flo[t2] = face.lo[t2] * ratio;
    fhi[t2] = face.hi[t2] * ratio + ratio - 1;
    if (!holds(reg, face.lo, face.hi, ncomp) || !holds(fflux, flo, fhi, ncomp)) return MF_ERR_BAD_ARG;
    const V r = mk(reg), fl = mk(fflux);
    for (int n = 0; n < ncomp; ++n)
        for (int k = face.lo[2]; k <= face.hi[2]; ++k)
            for (int j = face.lo[1]; j <= face.hi[1]; ++j)
                for (int i = face.lo[0]; i <= face.hi[0]; ++i) {
                    const int c[3] = {i, j, k};
                    double acc = r(i, j, k, n);
                    for (int b = 0; b < ratio; ++b)
                        for (int a = 0; a < ratio; ++a) {
                            int q[3];
                            q[dir] = c[dir] * ratio;
                            q[t1] = c[t1] * ratio + a;
                            q[t2] = c[t2] * ratio + b;
                            const double t = scale * fl(q[0], q[1], q[2], n);
                            acc = acc + t;
                        }
                    r(i, j, k, n) = acc;
                }
    return MF_OK;
}

// FluxRegister::Reflux(S, scale, 0, 0, ncomp, geom) (:243-246): the coarse
// cells that share the register's faces from outside the fine region get
// S += mult * reg / vol, mult = -scale where the fine region lies above the
// face (side 0, the cell below), +scale where it lies below (side 1)
int orc_fluxreg_reflux(int dir, int side, mf_box face, mf_array4 state, mf_array4 reg, double scale, double vol,
                       int ncomp) {
    if (dir < 0 || dir > 2 || face.lo[dir] != face.hi[dir] || (side != 0 && side != 1) || !(vol > 0.0))
        return MF_ERR_BAD_ARG;
    mf_box cells = face;
    if (side == 0) { cells.lo[dir] -= 1; cells.hi[dir] -= 1; }
    if (!holds(reg, face.lo, face.hi, ncomp) || !holds(state, cells.lo, cells.hi, ncomp)) return MF_ERR_BAD_ARG;
    const V r = mk(reg), s = mk(state);
    const double mult = side == 0 ? -scale : scale;
    for (int n = 0; n < ncomp; ++n)
        for (int k = face.lo[2]; k <= face.hi[2]; ++k)
            for (int j = face.lo[1]; j <= face.hi[1]; ++j)
                for (int i = face.lo[0]; i <= face.hi[0]; ++i) {
                    const int ci = i - (dir == 0 && side == 0 ? 1 : 0);
                    const int cj = j - (dir == 1 && side == 0 ? 1 : 0);
                    const int ck = k - (dir == 2 && side == 0 ? 1 : 0);
                    const double t = (mult * r(i, j, k, n)) / vol;
                    s(ci, cj, ck, n) = s(ci, cj, ck, n) + t;
                }
    return MF_OK;
}

}  // extern "C"

extern "C" {

// ErrorEst (Source/mflo.cpp:183-260) on one box: state holds bx grown by one
// (FillPatch); tags is a dense byte array over bx (x fastest). For each tagged
// variable c, state_based_refinement (Source/mflo_tagging.H:6-27) sets the
// tag where |state(comp_c)| > refine_phi[c], then stategrad_based_refinement
// (:29-70) where the largest of the one-sided absolute differences in x, y
// and z exceeds refine_phigrad[c]. Tags are only ever set.
int orc_box_tag(mf_box bx, mf_array4 state, unsigned char *tags, const double *refine_phi,
                const double *refine_phigrad, const int *comps, int ntag, int tagval) {
    if (!tags || ntag < 0 || (ntag > 0 && (!refine_phi || !refine_phigrad || !comps))) return MF_ERR_BAD_ARG;
    int glo[3], ghi[3];
    for (int d = 0; d < 3; ++d) { glo[d] = bx.lo[d] - 1; ghi[d] = bx.hi[d] + 1; }
    int maxc = 0;
    for (int c = 0; c < ntag; ++c) {
        if (comps[c] < 0) return MF_ERR_BAD_ARG;
        maxc = std::max(maxc, comps[c] + 1);
    }
    if (!holds(state, glo, ghi, maxc)) return MF_ERR_BAD_ARG;
    const V s = mk(state);
    const int ex = bx.hi[0] - bx.lo[0] + 1, ey = bx.hi[1] - bx.lo[1] + 1;
    const unsigned char tv = (unsigned char)tagval;
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
        for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
            for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                unsigned char &t = tags[(int64_t)(i - bx.lo[0]) + (int64_t)ex * ((j - bx.lo[1]) + (int64_t)ey * (k - bx.lo[2]))];
                for (int c = 0; c < ntag; ++c)
                    if (std::fabs(s(i, j, k, comps[c])) > refine_phi[c]) t = tv;
                for (int c = 0; c < ntag; ++c) {
                    const int n = comps[c];
                    const double u = s(i, j, k, n);
                    const double gx = std::max(std::fabs(s(i + 1, j, k, n) - u), std::fabs(u - s(i - 1, j, k, n)));
                    const double gy = std::max(std::fabs(s(i, j + 1, k, n) - u), std::fabs(u - s(i, j - 1, k, n)));
                    const double gz = std::max(std::fabs(s(i, j, k + 1, n) - u), std::fabs(u - s(i, j, k - 1, n)));
                    // amrex::max(gx, grady, gradz) = max(max(gx, gy), gz) (mflo_tagging.H:66)
                    if (std::max(std::max(gx, gy), gz) > refine_phigrad[c]) t = tv;
                }
            }
    return MF_OK;
}

}  // extern "C"
