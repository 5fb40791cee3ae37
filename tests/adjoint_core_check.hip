// Host-side check of the discrete-adjoint numerics (mesoflow_amd/csrc/mf_adjoint.h)
// against central finite differences of the forward functions they reverse,
// compiled for the host from the same sources the kernels use. For a random
// output weighting w and input direction v, each reverse function gives
// <rev(w), v>, which must match (phi(x + h v) - phi(x - h v)) / (2 h) with
// phi(x) = <w, f(x)>. Prints one line per check and exits non-zero on failure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "mf_adjoint.h"

using namespace mfd;
using P = PPorous<5>;
using R = Rev<P>;
constexpr int NS = P::NS, NG = P::NG, NC = NCV + NS, NV = FLO + NS;

static std::mt19937_64 rng(12345);
static double urand(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

static int failures = 0;
static void report(const char *name, double ad, double fd, double tol) {
    double err = fabs(ad - fd) / fmax(fabs(fd), 1e-300);
    bool ok = err < tol || fabs(ad - fd) < 1e-14;
    printf("%-34s adjoint % .12e  fd % .12e  rel %.2e %s\n", name, ad, fd, err, ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

static Tab make_tab() {
    Tab t;
    memset(&t, 0, sizeof(t));
    t.kind = MF_PLUGIN_POROUS_ARRHENIUS;
    t.nspec = NS;
    t.ngas = NG;
    t.nreact = 2;
    t.R_univ = 8.31446261815324;
    for (int i = 0; i < NG; ++i) {
        t.molwt[i] = i == 0 ? 0.0289 : 0.030 + 0.004 * i;
        t.gamma[i] = i == 0 ? 1.4 : 1.3 - 0.01 * i;
        t.advect_flag[i] = 1.0;
        t.dcoef_gas[i] = 2.0e-5 + 2.0e-6 * i;
        t.dcoef_solid[i] = 2.0e-7;
    }
    t.molwt[NG] = 0.001;
    t.gamma[NG] = 1.0;
    t.visc = 3.0e-5;
    t.thcond = 0.05;
    // reaction 0: 1 + site -> 2 ; reaction 1: 2 -> 3 (gas phase)
    t.stoich[0][1] = -1.0; t.stoich[0][2] = 1.0; t.stoich[0][NG] = -1.0;
    t.stoich[1][2] = -1.0; t.stoich[1][3] = 1.0;
    t.log_A[0] = log(3.0e6); t.Ea[0] = 60.0e3;
    t.log_A[1] = log(2.0e5); t.Ea[1] = 55.0e3;
    t.in_vel = 1.0; t.in_temp = 700.0; t.in_pres = 1.0e5;
    t.in_spec[0] = 0.8; t.in_spec[1] = 0.2;
    t.out_pres = 1.0e5;
    for (int q = 0; q < MF_MAX_SPECIES; ++q) {
        t.gm1[q] = t.gamma[q] - 1.0;
        t.igm1[q] = 1.0 / (t.gamma[q] - 1.0);
        t.cv[q] = t.R_univ / (t.gamma[q] - 1.0);
        t.cvm[q] = t.R_univ / (t.gamma[q] - 1.0) / t.molwt[q];
    }
    return t;
}

// a physical random cell: species in mol/m^3, T ~ 700 K, |u| ~ 1 m/s
static void random_cell(const Tab &T, double *S, double vfrac) {
    double sp[NS];
    for (int s = 0; s < NG; ++s) sp[s] = urand(1.0, 10.0);
    sp[NG] = urand(2.0, 10.0);
    const double Tm = urand(650.0, 750.0);
    double S_ = 0.0, rho = 0.0;
    for (int s = 0; s < NG; ++s) { S_ += sp[s]; rho += sp[s] * T.molwt[s]; }
    const double p = S_ * T.R_univ * Tm;
    const double v[3] = {urand(-1.5, 1.5), urand(-1.0, 1.0), urand(-1.0, 1.0)};
    const double e = P::get_e_from_rpc(T, rho, p, sp);
    S[RHO] = rho;
    S[RHOU] = rho * v[0]; S[RHOV] = rho * v[1]; S[RHOW] = rho * v[2];
    S[RHOE] = rho * e + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    S[DENS] = rho; S[VELX] = v[0]; S[VELY] = v[1]; S[VELZ] = v[2]; S[PRES] = p; S[TEMP] = Tm;
    S[VFRAC] = vfrac;
    for (int s = 0; s < NS; ++s) S[FLO + s] = sp[s];
}

// padded host level array
struct HostLevel {
    int n[3];
    std::vector<double> a;
    View v;
    explicit HostLevel(int nx, int ny, int nz) {
        n[0] = nx; n[1] = ny; n[2] = nz;
        const int G = MF_NGROW;
        v.b0 = v.b1 = v.b2 = -G;
        v.js = nx + 2 * G;
        v.ks = v.js * (ny + 2 * G);
        v.ns = v.ks * (nz + 2 * G);
        a.assign((size_t)v.ns * NV, 0.0);
        v.p = a.data();
    }
    void fill_random(const Tab &T) {
        const int G = MF_NGROW;
        for (int k = -G; k < n[2] + G; ++k)
            for (int j = -G; j < n[1] + G; ++j)
                for (int i = -G; i < n[0] + G; ++i) {
                    double S[NV];
                    random_cell(T, S, 1.0);
                    for (int c = 0; c < NV; ++c) v(i, j, k, c) = S[c];
                }
    }
};

static double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    for (size_t q = 0; q < a.size(); ++q) s += a[q] * b[q];
    return s;
}

// ---------------------------------------------------------------------------
static void check_weno() {
    for (int trial = 0; trial < 4; ++trial) {
        double s[5], v[5], w = urand(-2, 2);
        for (int q = 0; q < 5; ++q) { s[q] = urand(-1, 1) * (trial + 1); v[q] = urand(-1, 1); }
        if (trial == 3) for (int q = 0; q < 5; ++q) s[q] = 0.3 + 0.01 * q + 1e-4 * urand(-1, 1);  // smooth data
        double g[5] = {0, 0, 0, 0, 0}, gm[5] = {0, 0, 0, 0, 0};
        const double vp = weno_plus_rev(s, w, g);
        const double vm = weno_minus_rev(s, w, gm);
        // the reverse passes return the forward reconstructions (to rounding)
        {
            const double fp = Num<P>::weno_plus(s[0], s[1], s[2], s[3], s[4]);
            const double fm = Num<P>::weno_minus(s[0], s[1], s[2], s[3], s[4]);
            char nm[64];
            snprintf(nm, sizeof nm, "weno_plus value trial %d", trial);
            report(nm, vp, fp, 1e-14);
            snprintf(nm, sizeof nm, "weno_minus value trial %d", trial);
            report(nm, vm, fm, 1e-14);
        }
        double ad = 0, adm = 0;
        for (int q = 0; q < 5; ++q) { ad += g[q] * v[q]; adm += gm[q] * v[q]; }
        const double h = 1e-6;
        double sp[5], sm[5];
        for (int q = 0; q < 5; ++q) { sp[q] = s[q] + h * v[q]; sm[q] = s[q] - h * v[q]; }
        const double fd = w * (Num<P>::weno_plus(sp[0], sp[1], sp[2], sp[3], sp[4]) -
                               Num<P>::weno_plus(sm[0], sm[1], sm[2], sm[3], sm[4])) / (2 * h);
        const double fdm = w * (Num<P>::weno_minus(sp[0], sp[1], sp[2], sp[3], sp[4]) -
                                Num<P>::weno_minus(sm[0], sm[1], sm[2], sm[3], sm[4])) / (2 * h);
        char nm[64];
        snprintf(nm, sizeof nm, "weno_plus trial %d", trial);
        report(nm, ad, fd, 1e-6);
        snprintf(nm, sizeof nm, "weno_minus trial %d", trial);
        report(nm, adm, fdm, 1e-6);
    }
}

// ---------------------------------------------------------------------------
// one face of direction D: forward face_flux<P,5,D> vs the reverse functions
struct HostStencil {
    const Tab *T;
    const View *s;
    int cell[6][3];
    double pr[6][4];                 // vx, vy, vz, p
    std::vector<double> *lam;        // adjoint of the array
    const double *Fb;                // adjoint of the flux components
    double vb[6][3], pb[6];
    double fb(int c) const { return Fb[c]; }
    double u(int q, int c) const { return (*s)(cell[q][0], cell[q][1], cell[q][2], c < NCV ? c : FLO + (c - NCV)); }
    double vel(int q, int a) const { return pr[q][a]; }
    double pres(int q) const { return pr[q][3]; }
    void add_u(int q, int c, double x) {
        (*lam)[s->off(cell[q][0], cell[q][1], cell[q][2]) + (size_t)s->ns * (c < NCV ? c : FLO + (c - NCV))] += x;
    }
    void add_v(int q, int a, double x) { vb[q][a] += x; }
    void add_p(int q, double x) { pb[q] += x; }
};

template <int D>
static void face_value(const Tab &T, const View &s, int i, int j, int k, double *F) {
    View Fv;
    static std::vector<double> fbuf;
    fbuf.assign(NC, 0.0);
    Fv.p = fbuf.data();
    Fv.b0 = i; Fv.b1 = j; Fv.b2 = k;
    Fv.js = 1; Fv.ks = 1; Fv.ns = 1;
    Acc acc;
    memset(&acc, 0, sizeof(acc));
    acc.s = s;
    acc.cutid = nullptr;
    StageArgs a;
    memset(&a, 0, sizeof(a));
    a.dx[0] = 4e-6; a.dx[1] = 5e-6; a.dx[2] = 3e-6;
    a.nsflag = 1;
    a.order = 5;
    face_flux<P, 5, D>(T, acc, a, i, j, k, Fv, 0.0);
    for (int c = 0; c < NC; ++c) F[c] = fbuf[c];
}

template <int D>
static void check_face(const char *tag, HostLevel &L, const Tab &T, int i, int j, int k) {
    const View &s = L.v;
    constexpr int e0 = D == 0, e1 = D == 1, e2 = D == 2;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    const double dx[3] = {4e-6, 5e-6, 3e-6};
    double Fb[NC];
    for (int c = 0; c < NC; ++c) Fb[c] = urand(-1, 1) * (c == RHOE ? 1e-5 : 1.0);
    std::vector<double> lam(L.a.size(), 0.0);
    // stencil and first-order mapping
    double vf[6];
    for (int q = 0; q < 6; ++q) vf[q] = s(i + (q - 3) * e0, j + (q - 3) * e1, k + (q - 3) * e2, VFRAC);
    const double facevfrac = 0.5 * (vf[2] + vf[3]);
    const bool hyp = facevfrac >= 0.5;
    bool first = false;
    for (int q = 0; q < 6; ++q) first = first || vf[q] < 1.0;
    const int ms[6] = {first ? 2 : 0, first ? 2 : 1, 2, 3, first ? 3 : 4, first ? 3 : 5};
    HostStencil st;
    st.T = &T;
    st.s = &s;
    st.lam = &lam;
    st.Fb = Fb;
    for (int q = 0; q < 6; ++q) {
        const int m = ms[q] - 3;
        st.cell[q][0] = i + m * e0; st.cell[q][1] = j + m * e1; st.cell[q][2] = k + m * e2;
        double u[NC];
        for (int c = 0; c < NC; ++c) u[c] = st.u(q, c);
        Num<P>::euler_prims(T, u, st.pr[q][0], st.pr[q][1], st.pr[q][2], st.pr[q][3]);
        st.vb[q][0] = st.vb[q][1] = st.vb[q][2] = 0.0;
        st.pb[q] = 0.0;
    }
    if (hyp) R::face_hyp_rev<D>(T, st);
    for (int q = 0; q < 6; ++q) {
        double u[NC], ub[NC];
        for (int c = 0; c < NC; ++c) { u[c] = st.u(q, c); ub[c] = 0.0; }
        R::prims_rev(T, u, st.vb[q], st.pb[q], ub);
        for (int c = 0; c < NC; ++c) st.add_u(q, c, ub[c]);
    }
    // pass 2
    const int iL = i - e0, jL = j - e1, kL = k - e2;
    constexpr int p1[3] = {T1 == 0, T1 == 1, T1 == 2}, p2[3] = {T2 == 0, T2 == 1, T2 == 2};
    const int ti[8][3] = {{iL + p1[0], jL + p1[1], kL + p1[2]}, {i + p1[0], j + p1[1], k + p1[2]},
                          {iL - p1[0], jL - p1[1], kL - p1[2]}, {i - p1[0], j - p1[1], k - p1[2]},
                          {iL + p2[0], jL + p2[1], kL + p2[2]}, {i + p2[0], j + p2[1], k + p2[2]},
                          {iL - p2[0], jL - p2[1], kL - p2[2]}, {i - p2[0], j - p2[1], k - p2[2]}};
    double vL[3], vR[3], tv[8][3], tvb[8][3], vLb[3] = {0, 0, 0}, vRb[3] = {0, 0, 0}, TLb = 0, TRb = 0, cLb[NS], cRb[NS];
    for (int q = 0; q < 3; ++q) { vL[q] = s(iL, jL, kL, VELX + q); vR[q] = s(i, j, k, VELX + q); }
    for (int m = 0; m < 8; ++m)
        for (int q = 0; q < 3; ++q) { tv[m][q] = s(ti[m][0], ti[m][1], ti[m][2], VELX + q); tvb[m][q] = 0.0; }
    for (int q = 0; q < NS; ++q) { cLb[q] = 0.0; cRb[q] = 0.0; }
    double Ds[NS];
    {
        double dl[NS], dr[NS];
        // spec_dcoeff reads VFRAC only: gather the cells into compact arrays
        double Sl[NV], Sr[NV];
        for (int c = 0; c < NV; ++c) { Sl[c] = s(iL, jL, kL, c); Sr[c] = s(i, j, k, c); }
        P::spec_dcoeff(T, CellArr<NV>{Sl}, dl);
        P::spec_dcoeff(T, CellArr<NV>{Sr}, dr);
        for (int q = 0; q < NS; ++q) Ds[q] = Num<P>::harmonic_mean(dl[q], dr[q]);
    }
    const double visc = Num<P>::harmonic_mean(T.visc, T.visc), thcond = Num<P>::harmonic_mean(T.thcond, T.thcond);
    double G[3][3], h1[3], h2[3];
    R::grad_tensor<D>(dx, vL, vR, tv, G);
    R::pass2_rev<D>(Fb, true, hyp, visc, thcond, Ds, dx, vL, vR, G, vLb, vRb, h1, h2, TLb, TRb, cLb, cRb);
    for (int m = 0; m < 8; ++m)
        for (int q = 0; q < 3; ++q) tvb[m][q] = (m < 4 ? h1[q] : h2[q]) * ((m & 2) ? -0.5 : 0.5);
    auto addc = [&](int a, int b, int c, int n, double x) { lam[s.off(a, b, c) + (size_t)s.ns * n] += x; };
    for (int q = 0; q < 3; ++q) { addc(iL, jL, kL, VELX + q, vLb[q]); addc(i, j, k, VELX + q, vRb[q]); }
    addc(iL, jL, kL, TEMP, TLb);
    addc(i, j, k, TEMP, TRb);
    for (int m = 0; m < 8; ++m)
        for (int q = 0; q < 3; ++q) addc(ti[m][0], ti[m][1], ti[m][2], VELX + q, tvb[m][q]);
    for (int q = 0; q < NS; ++q) { addc(iL, jL, kL, FLO + q, cLb[q]); addc(i, j, k, FLO + q, cRb[q]); }
    // finite differences along a random direction scaled per component
    std::vector<double> dir(L.a.size(), 0.0);
    for (size_t q = 0; q < dir.size(); ++q) {
        const int c = (int)(q / s.ns);
        if (c == VFRAC) continue;
        dir[q] = urand(-1, 1) * fabs(L.a[q]) * 1e-7;
    }
    const double ad = dot(lam, dir);
    double Fp[NC], Fm[NC];
    std::vector<double> save = L.a;
    for (size_t q = 0; q < dir.size(); ++q) L.a[q] = save[q] + dir[q];
    face_value<D>(T, s, i, j, k, Fp);
    for (size_t q = 0; q < dir.size(); ++q) L.a[q] = save[q] - dir[q];
    face_value<D>(T, s, i, j, k, Fm);
    L.a = save;
    double fd = 0.0;
    for (int c = 0; c < NC; ++c) fd += Fb[c] * (Fp[c] - Fm[c]) / 2.0;
    char nm[80];
    snprintf(nm, sizeof nm, "face D=%d %s", D, tag);
    report(nm, ad, fd, 1e-5);
}

// ---------------------------------------------------------------------------
static void check_post(const Tab &T0) {
    for (int trial = 0; trial < 3; ++trial) {
        Tab T = T0;
        double S3[NV];
        random_cell(T, S3, trial == 2 ? 0.625 : 1.0);
        if (trial == 1) S3[FLO + 3] = -1e-3;   // clipped species
        // stored primitives not consistent with the conserved state (RK-combined)
        S3[VELX] += 0.1; S3[TEMP] += 5.0; S3[PRES] *= 1.01;
        const double dt = 2e-9 * 1e3;
        double lb[NV], w[NV], dir[NV];
        for (int n = 0; n < NV; ++n) { w[n] = urand(-1, 1); lb[n] = w[n]; dir[n] = urand(-1, 1) * fabs(S3[n]) * 1e-7; }
        dir[VFRAC] = 0.0;
        double gA[MF_MAX_REACTIONS] = {0};
        R::post_rev(T, S3, dt, 1, lb, gA);
        double ad = 0.0;
        for (int n = 0; n < NV; ++n) ad += lb[n] * dir[n];
        // <w, f(x1) - f(x2)> with the difference taken per component
        auto dphi = [&](const double *x1, const Tab &T1, const double *x2, const Tab &T2) {
            double S1[NV], S2[NV];
            for (int n = 0; n < NV; ++n) { S1[n] = x1[n]; S2[n] = x2[n]; }
            R::post_cell(T1, S1, dt, 1);
            R::post_cell(T2, S2, dt, 1);
            double v = 0.0;
            for (int n = 0; n < NV; ++n) v += w[n] * (S1[n] - S2[n]);
            return v;
        };
        double xp[NV], xm[NV];
        for (int n = 0; n < NV; ++n) { xp[n] = S3[n] + dir[n]; xm[n] = S3[n] - dir[n]; }
        const double fd = dphi(xp, T, xm, T) / 2.0;
        char nm[80];
        snprintf(nm, sizeof nm, "post state trial %d", trial);
        report(nm, ad, fd, 1e-6);
        for (int r = 0; r < T.nreact; ++r) {
            const double h = 1e-4;
            Tab Tp = T, Tm = T;
            Tp.log_A[r] += h;
            Tm.log_A[r] -= h;
            const double fdA = dphi(S3, Tp, S3, Tm) / (2 * h);
            snprintf(nm, sizeof nm, "post log_A[%d] trial %d", r, trial);
            report(nm, gA[r], fdA, 1e-5);
        }
    }
}

// ---------------------------------------------------------------------------
static void check_externalbc(const Tab &T0) {
    for (int inflow = 0; inflow < 2; ++inflow) {
        Tab T = T0;
        double si[NV];
        random_cell(T, si, 1.0);
        double w[NV], dir[NV];
        for (int n = 0; n < NV; ++n) { w[n] = urand(-1, 1); dir[n] = urand(-1, 1) * fabs(si[n]) * 1e-7; }
        const int idir = 0, sgn = inflow ? 1 : -1;
        double lint[NV] = {0}, invb = 0.0;
        R::externalbc_rev(T, si, w, idir, sgn, lint, invb);
        double ad = 0.0;
        for (int n = 0; n < NV; ++n) ad += lint[n] * dir[n];
        auto dphi = [&](const double *x1, const Tab &T1, const double *x2, const Tab &T2) {
            double s1[NV], s2[NV];
            for (int n = 0; n < NV; ++n) { s1[n] = 0.0; s2[n] = 0.0; }
            const double xx[3] = {0, 0, 0};
            P::externalbc(T1, xx, x1, s1, idir, sgn, 0.0);
            P::externalbc(T2, xx, x2, s2, idir, sgn, 0.0);
            double v = 0.0;
            for (int n = 0; n < NV; ++n) v += w[n] * (s1[n] - s2[n]);
            return v;
        };
        double xp[NV], xm[NV];
        for (int n = 0; n < NV; ++n) { xp[n] = si[n] + dir[n]; xm[n] = si[n] - dir[n]; }
        char nm[80];
        snprintf(nm, sizeof nm, "externalbc %s state", inflow ? "inflow" : "outflow");
        report(nm, ad, dphi(xp, T, xm, T) / 2.0, 1e-6);
        if (inflow) {
            Tab Tp = T, Tm = T;
            const double h = 1e-4;
            Tp.in_vel += h;
            Tm.in_vel -= h;
            report("externalbc inflow in_vel", invb, dphi(si, Tp, si, Tm) / (2 * h), 1e-5);
        }
    }
}

// ---------------------------------------------------------------------------
static void check_cut(const Tab &T) {
    HostLevel L(9, 9, 9);
    L.fill_random(T);
    const View &s = L.v;
    // planar solid x + 0.3 y + 0.15 z < 5.3: vfrac from 2x2x2 corner sampling (k/8)
    for (int k = -3; k < 12; ++k)
        for (int j = -3; j < 12; ++j)
            for (int i = -3; i < 12; ++i) {
                int fluid = 0;
                for (int c = 0; c < 8; ++c) {
                    const double x = i + 0.25 + 0.5 * (c & 1), y = j + 0.25 + 0.5 * ((c >> 1) & 1);
                    const double z = k + 0.25 + 0.5 * ((c >> 2) & 1);
                    if (!(x + 0.3 * y + 0.15 * z < 5.3)) ++fluid;
                }
                s(i, j, k, VFRAC) = fluid / 8.0;
            }
    CutArgs a;
    memset(&a, 0, sizeof(a));
    a.s = s;
    a.dx[0] = 4e-6; a.dx[1] = 5e-6; a.dx[2] = 3e-6;
    int tested = 0;
    for (int q = 0; q < 125 && tested < 4; ++q) {
        const int i = 2 + q % 5, j = 2 + (q / 5) % 5, k = 2 + q / 25;
        const double vf = s(i, j, k, VFRAC);
        if (!(vf > 0.0 && vf < 1.0)) continue;
        double w[NV];
        for (int n = 0; n < NV; ++n) w[n] = urand(-1, 1);
        std::vector<double> lam(L.a.size(), 0.0);
        auto add = [&](int a_, int b_, int c_, int n, double x) { lam[s.off(a_, b_, c_) + (size_t)s.ns * n] += x; };
        R::cut_rev(T, a, i, j, k, w, add);
        std::vector<double> dir(L.a.size(), 0.0);
        for (size_t q = 0; q < dir.size(); ++q) {
            if ((int)(q / s.ns) == VFRAC) continue;
            dir[q] = urand(-1, 1) * fabs(L.a[q]) * 1e-7;
        }
        const double ad = dot(lam, dir);
        std::vector<double> save = L.a;
        double op[NV], om[NV];
        for (size_t q = 0; q < dir.size(); ++q) L.a[q] = save[q] + dir[q];
        cutcell_values<P>(T, a, i, j, k, op);
        for (size_t q = 0; q < dir.size(); ++q) L.a[q] = save[q] - dir[q];
        cutcell_values<P>(T, a, i, j, k, om);
        L.a = save;
        double fd = 0.0;
        for (int n = 0; n < NV; ++n) fd += w[n] * (op[n] - om[n]) / 2.0;
        char nm[80];
        snprintf(nm, sizeof nm, "cut cell (%d,%d,%d) vfrac %.3f", i, j, k, vf);
        report(nm, ad, fd, 1e-6);
        ++tested;
    }
    if (!tested) { printf("cut: no cut cell found\n"); ++failures; }
}

// the gather form used by the GPU (cut_geom_cell + cut_coef per cut cell,
// cut_gather per receiving fluid cell) against cut_rev over every cut cell of
// a planar wall: the same adjoint up to rounding
static void check_cut_gather(const Tab &T) {
    HostLevel L(9, 9, 9);
    L.fill_random(T);
    const View &s = L.v;
    for (int k = -3; k < 12; ++k)
        for (int j = -3; j < 12; ++j)
            for (int i = -3; i < 12; ++i) {
                int fluid = 0;
                for (int c = 0; c < 8; ++c) {
                    const double x = i + 0.25 + 0.5 * (c & 1), y = j + 0.25 + 0.5 * ((c >> 1) & 1);
                    const double z = k + 0.25 + 0.5 * ((c >> 2) & 1);
                    if (!(0.8 * x + 0.5 * y - 0.35 * z < 3.9)) ++fluid;
                }
                s(i, j, k, VFRAC) = fluid / 8.0;
            }
    CutArgs a;
    memset(&a, 0, sizeof(a));
    a.s = s;
    a.dx[0] = 4e-6; a.dx[1] = 5e-6; a.dx[2] = 3e-6;
    std::vector<int> cid(9 * 9 * 9, -1), ci, cj, ck;
    for (int k = 0; k < 9; ++k)
        for (int j = 0; j < 9; ++j)
            for (int i = 0; i < 9; ++i) {
                const double vf = s(i, j, k, VFRAC);
                if (vf > 0.0 && vf < 1.0) {
                    cid[i + 9 * (j + 9 * k)] = (int)ci.size();
                    ci.push_back(i); cj.push_back(j); ck.push_back(k);
                }
            }
    const int ncut = (int)ci.size();
    std::vector<double> lref(L.a.size(), 0.0), lnew(L.a.size(), 0.0), recs((size_t)ncut * R::CUTREC);
    auto at = [&](std::vector<double> &v, int i, int j, int k, int n) -> double & { return v[s.off(i, j, k) + (size_t)s.ns * n]; };
    for (int q = 0; q < ncut; ++q) {
        double ob[NV], out[NV], g[6];
        unsigned mb[4];
        for (int n = 0; n < NV; ++n) ob[n] = urand(-1, 1);
        cutcell_values<P>(T, a, ci[q], cj[q], ck[q], out);
        auto add = [&](int i_, int j_, int k_, int n, double x) { at(lref, i_, j_, k_, n) += x; };
        R::cut_rev(T, a, ci[q], cj[q], ck[q], ob, add, out, 1);
        cut_geom_cell(a, ci[q], cj[q], ck[q], g, mb);
        auto self = [&](int n, double x) { at(lnew, ci[q], cj[q], ck[q], n) += x; };
        R::cut_coef(T, a, g, ob, out, &recs[(size_t)q * R::CUTREC], self);
    }
    int receivers = 0;
    for (int k = -2; k < 11; ++k)
        for (int j = -2; j < 11; ++j)
            for (int i = -2; i < 11; ++i) {
                if (s(i, j, k, VFRAC) != 1.0) continue;
                unsigned mb[4] = {0u, 0u, 0u, 0u};
                for (int nb = 0; nb < 125; ++nb) {
                    const int l = nb % 5 - 2, m = (nb / 5) % 5 - 2, n = nb / 25 - 2;
                    const int x = i - l, y = j - m, z = k - n;
                    if (x < 0 || x >= 9 || y < 0 || y >= 9 || z < 0 || z >= 9) continue;
                    if (cid[x + 9 * (y + 9 * z)] >= 0) mb[nb >> 5] |= 1u << (nb & 31);
                }
                if (!(mb[0] | mb[1] | mb[2] | mb[3])) continue;
                ++receivers;
                auto rec_of = [&](int l, int m, int n) -> const double * {
                    return &recs[(size_t)cid[(i - l) + 9 * ((j - m) + 9 * (k - n))] * R::CUTREC];
                };
                double res[6 + NS];
                R::cut_gather(a.dx, mb, rec_of, res);
                for (int d = 0; d < 3; ++d) at(lnew, i, j, k, VELX + d) += res[d];
                at(lnew, i, j, k, PRES) += res[3];
                at(lnew, i, j, k, TEMP) += res[4];
                at(lnew, i, j, k, DENS) += res[5];
                for (int q = 0; q < NS; ++q) at(lnew, i, j, k, FLO + q) += res[6 + q];
            }
    double dmax = 0.0, rmax = 0.0;
    for (size_t q = 0; q < lref.size(); ++q) {
        dmax = fmax(dmax, fabs(lnew[q] - lref[q]));
        rmax = fmax(rmax, fabs(lref[q]));
    }
    const bool ok = ncut > 20 && receivers > 20 && dmax <= 1e-12 * rmax;
    printf("%-34s %d cut cells %d receivers  max|diff| %.3e  max|adjoint| %.3e %s\n", "cut cell gather form", ncut, receivers,
           dmax, rmax, ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

int main() {
    const Tab T = make_tab();
    check_weno();
    {
        HostLevel L(8, 8, 8);
        L.fill_random(T);
        check_face<0>("smooth", L, T, 4, 4, 4);
        check_face<1>("smooth", L, T, 4, 4, 4);
        check_face<2>("smooth", L, T, 4, 4, 4);
        // first-order fallback: a solid-touched cell in the stencil
        L.v(1, 4, 4, VFRAC) = 0.75;
        L.v(4, 2, 4, VFRAC) = 0.75;
        L.v(4, 4, 6, VFRAC) = 0.75;
        check_face<0>("IB first order", L, T, 4, 4, 4);
        check_face<1>("IB first order", L, T, 4, 4, 4);
        check_face<2>("IB first order", L, T, 4, 4, 4);
        // face below half open: conduction and species diffusion only
        L.v(3, 4, 4, VFRAC) = 0.25;
        L.v(4, 4, 4, VFRAC) = 0.5;
        check_face<0>("no hyperbolic part", L, T, 4, 4, 4);
    }
    check_post(T);
    check_externalbc(T);
    check_cut(T);
    check_cut_gather(T);
    printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
