// mf_adjoint.h -- reverse-mode (discrete adjoint) numerics of the RHS + RK3 path.
//
// Each function here is the exact transpose of the derivative of one forward
// function of mf_device.h / mf_kernels.inc.h at a given state: given the
// adjoint of the forward outputs it adds (the derivative)^T * adjoint to the
// adjoints of the inputs. Forward values are recomputed from the inputs.
// Branches (fmax/fmin/fabs, clipping, IB masks) take the branch the forward took.
// Covered plugin family: the Arrhenius porous reactor (PPorous<NS>: mixture
// thermo of gas species 0..NG-1, constant transport, state-independent
// species diffusion coefficients given vfrac).
#pragma once
#include "mf_kernels.inc.h"

namespace mfd {

// ---------------------------------------------------------------------------
// WENO5-JS (single-division form of Num::weno_plus): gradient w.r.t. s0..s4
// times gbar, added to sb; returns the reconstructed value (Num::weno_plus up
// to rounding: one reciprocal of the weight sum instead of a division).
// weno_minus(s1..s5) == weno_plus(s5,s4,s3,s2,s1).
// ---------------------------------------------------------------------------
MFD double weno_plus_rev(const double *s, double gbar, double *sb) {
    const double eps = 1e-6;
    const double t2 = s[0] - 2.0 * s[1] + s[2], q2 = s[0] - 4.0 * s[1] + 3.0 * s[2];
    const double b2 = (13.0 / 12.0) * (t2 * t2) + 0.25 * (q2 * q2);
    const double t1 = s[1] - 2.0 * s[2] + s[3], r1 = s[1] - s[3];
    const double b1 = (13.0 / 12.0) * (t1 * t1) + 0.25 * (r1 * r1);
    const double t0 = s[2] - 2.0 * s[3] + s[4], q0 = 3.0 * s[2] - 4.0 * s[3] + s[4];
    const double b0 = (13.0 / 12.0) * (t0 * t0) + 0.25 * (q0 * q0);
    const double v2 = 2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2];
    const double v1 = -s[1] + 5.0 * s[2] + 2.0 * s[3];
    const double v0 = 2.0 * s[2] + 5.0 * s[3] - s[4];
    const double e2 = eps + b2, e1 = eps + b1, e0 = eps + b0;
    const double Q2 = e2 * e2, Q1 = e1 * e1, Q0 = e0 * e0;
    const double w2 = Q1 * Q0, w1 = 6.0 * (Q2 * Q0), w0 = 3.0 * (Q2 * Q1);
    const double den = 6.0 * (w2 + w1 + w0);
    const double iden = div_pos(1.0, den);  // one reciprocal of the positive weight sum
    const double val = (w2 * v2 + w1 * v1 + w0 * v0) * iden;
    const double nb = gbar * iden;
    const double db = -gbar * val * iden;
    const double w2b = nb * v2 + 6.0 * db, w1b = nb * v1 + 6.0 * db, w0b = nb * v0 + 6.0 * db;
    const double v2b = nb * w2, v1b = nb * w1, v0b = nb * w0;
    const double Q2b = w1b * 6.0 * Q0 + w0b * 3.0 * Q1;
    const double Q1b = w2b * Q0 + w0b * 3.0 * Q2;
    const double Q0b = w2b * Q1 + w1b * 6.0 * Q2;
    const double b2b = Q2b * 2.0 * e2, b1b = Q1b * 2.0 * e1, b0b = Q0b * 2.0 * e0;
    const double t2b = b2b * (13.0 / 6.0) * t2, q2b = b2b * 0.5 * q2;
    const double t1b = b1b * (13.0 / 6.0) * t1, r1b = b1b * 0.5 * r1;
    const double t0b = b0b * (13.0 / 6.0) * t0, q0b = b0b * 0.5 * q0;
    sb[0] += t2b + q2b + 2.0 * v2b;
    sb[1] += -2.0 * t2b - 4.0 * q2b + t1b + r1b - 7.0 * v2b - v1b;
    sb[2] += t2b + 3.0 * q2b - 2.0 * t1b + t0b + 3.0 * q0b + 11.0 * v2b + 5.0 * v1b + 2.0 * v0b;
    sb[3] += t1b - r1b - 2.0 * t0b - 4.0 * q0b + 2.0 * v1b + 5.0 * v0b;
    sb[4] += t0b + q0b - v0b;
    return val;
}
// gradient of weno_minus(m[0..4]) (m = s1..s5) times gbar, added to mb;
// returns the reconstructed value
MFD double weno_minus_rev(const double *m, double gbar, double *mb) {
    const double r[5] = {m[4], m[3], m[2], m[1], m[0]};
    double rb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double val = weno_plus_rev(r, gbar, rb);
    mb[0] += rb[4]; mb[1] += rb[3]; mb[2] += rb[2]; mb[3] += rb[1]; mb[4] += rb[0];
    return val;
}

template <class P>
struct Rev {
    static constexpr int NS = P::NS, NG = P::NG, NC = NCV + NS, NV = FLO + NS;
    using N = Num<P>;

    // get_p_from_rec (mixture): p = R * (rhoe / sum cv_s c_s) * sum c_s
    MFD static void p_rev(const Tab &T, double rhoe, const double *sp, double pb, double &rhoeb, double *spb) {
        double cvs = kZero, S = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) { cvs += T.cv[s] * sp[s]; S += sp[s]; }
        const double Tm = rhoe / cvs;
        rhoeb += pb * T.R_univ * S / cvs;
#pragma unroll
        for (int s = 0; s < NG; ++s) spb[s] += pb * T.R_univ * Tm * (kOne - S * T.cv[s] / cvs);
    }
    // get_gama_from_rpc (mixture): gamma = 1 + 1 / sum_s (c_s/S)/(gamma_s-1)
    MFD static void gama_rev(const Tab &T, const double *sp, double gb, double *spb) {
        double S = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) S += sp[s];
        double term2 = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) term2 += (sp[s] / S) / T.gm1[s];
        const double f = -gb / (term2 * term2) / S;
#pragma unroll
        for (int s = 0; s < NG; ++s) spb[s] += f * (kOne / T.gm1[s] - term2);
    }
    // get_t_from_rpc: T = p / (S R)
    MFD static void t_rev(const Tab &T, double pres, const double *sp, double tb, double &presb, double *spb) {
        double S = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) S += sp[s];
        const double Tm = pres / (S * T.R_univ);
        presb += tb / (S * T.R_univ);
#pragma unroll
        for (int s = 0; s < NG; ++s) spb[s] += -tb * Tm / S;
    }
    // get_e_from_rpc: e = (sum cv_s c_s) * T / rho, T = p / (S R)
    MFD static void e_rev(const Tab &T, double rho, double pres, const double *sp, double eb, double &rhob, double &presb,
                          double *spb) {
        double S = kZero, cvc = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) { S += sp[s]; cvc += T.cv[s] * sp[s]; }
        const double Tm = pres / (S * T.R_univ);
        const double e = cvc * Tm / rho;
        rhob += -eb * e / rho;
        presb += eb * cvc / (S * T.R_univ * rho);
#pragma unroll
        for (int s = 0; s < NG; ++s) spb[s] += eb * (T.cv[s] * Tm - cvc * Tm / S) / rho;
    }
    // Num::euler_prims / cons_to_prim velocities and pressure:
    // vbar (3) and pbar are the adjoints of vx, vy, vz, p; adds to ub[NC]
    MFD static void prims_rev(const Tab &T, const double *u, const double *vbar, double pbar, double *ub) {
        const double rho = u[RHO];
        const double vx = u[RHOU] / rho, vy = u[RHOV] / rho, vz = u[RHOW] / rho;
        const double rhoe = u[RHOE] - kHalf * u[RHO] * (vx * vx + vy * vy + vz * vz);
        double rhoeb = 0.0;
        p_rev(T, rhoe, u + NCV, pbar, rhoeb, ub + NCV);
        ub[RHOE] += rhoeb;
        double rhob = -kHalf * (vx * vx + vy * vy + vz * vz) * rhoeb;
        const double vxb = vbar[0] - rho * vx * rhoeb;
        const double vyb = vbar[1] - rho * vy * rhoeb;
        const double vzb = vbar[2] - rho * vz * rhoeb;
        ub[RHOU] += vxb / rho;
        ub[RHOV] += vyb / rho;
        ub[RHOW] += vzb / rho;
        rhob += -(vx * vxb + vy * vyb + vz * vzb) / rho;
        ub[RHO] += rhob;
    }
    // rhoe = u4 - 0.5 (u1^2 + u2^2 + u3^2) / u0 (max-eigenvalue form)
    MFD static void rhoe_rev(const double *u, double rhoeb, double *ub) {
        ub[RHOE] += rhoeb;
        ub[RHOU] += -rhoeb * u[RHOU] / u[RHO];
        ub[RHOV] += -rhoeb * u[RHOV] / u[RHO];
        ub[RHOW] += -rhoeb * u[RHOW] / u[RHO];
        ub[RHO] += rhoeb * kHalf * (u[RHOU] * u[RHOU] + u[RHOV] * u[RHOV] + u[RHOW] * u[RHOW]) / (u[RHO] * u[RHO]);
    }
    // LLF eigenvalue of the WENO split (flux_schemes.H:10-46, presR with the left species)
    MFD static double eigen(const Tab &T, int D, const double *ul, const double *ur) {
        double rhoeL = ul[RHOE] - kHalf * (ul[RHOU] * ul[RHOU] + ul[RHOV] * ul[RHOV] + ul[RHOW] * ul[RHOW]) / ul[RHO];
        double rhoeR = ur[RHOE] - kHalf * (ur[RHOU] * ur[RHOU] + ur[RHOV] * ur[RHOV] + ur[RHOW] * ur[RHOW]) / ur[RHO];
        double presL = P::get_p_from_rec(T, rhoeL, ul[RHO], ul + NCV);
        double presR = P::get_p_from_rec(T, rhoeR, ur[RHO], ul + NCV);
        double veln = fmax(fabs(ul[RHOU + D] / ul[RHO]), fabs(ur[RHOU + D] / ur[RHO]));
        double gamaL = P::get_gama_from_rpc(T, ul[RHO], presL, ul + NCV);
        double gamaR = P::get_gama_from_rpc(T, ur[RHO], presR, ur + NCV);
        return fmax(sqrt(gamaL * presL / ul[RHO]), sqrt(gamaR * presR / ur[RHO])) + veln;
    }
    MFD static void eigen_rev(const Tab &T, int D, const double *ul, const double *ur, double lamb, double *ubl, double *ubr) {
        const double rhoeL = ul[RHOE] - kHalf * (ul[RHOU] * ul[RHOU] + ul[RHOV] * ul[RHOV] + ul[RHOW] * ul[RHOW]) / ul[RHO];
        const double rhoeR = ur[RHOE] - kHalf * (ur[RHOU] * ur[RHOU] + ur[RHOV] * ur[RHOV] + ur[RHOW] * ur[RHOW]) / ur[RHO];
        const double presL = P::get_p_from_rec(T, rhoeL, ul[RHO], ul + NCV);
        const double presR = P::get_p_from_rec(T, rhoeR, ur[RHO], ul + NCV);
        const double vL = ul[RHOU + D] / ul[RHO], vR = ur[RHOU + D] / ur[RHO];
        const double gamaL = P::get_gama_from_rpc(T, ul[RHO], presL, ul + NCV);
        const double gamaR = P::get_gama_from_rpc(T, ur[RHO], presR, ur + NCV);
        const double aL = sqrt(gamaL * presL / ul[RHO]), aR = sqrt(gamaR * presR / ur[RHO]);
        // veln = fmax(|vL|, |vR|); the branch the forward took is selected by
        // value (a select between the two register arrays would put them in
        // private memory)
        {
            const bool left = fabs(vL) >= fabs(vR);
            const double v = left ? vL : vR, r = left ? ul[RHO] : ur[RHO];
            const double vb = lamb * (double)((v > 0.0) - (v < 0.0));
            const double du = vb / r, dr = -vb * v / r;
            ubl[RHOU + D] += left ? du : 0.0;
            ubl[RHO] += left ? dr : 0.0;
            ubr[RHOU + D] += left ? 0.0 : du;
            ubr[RHO] += left ? 0.0 : dr;
        }
        double presLb, presRb;
        {
            const bool left = aL >= aR;
            const double a = left ? aL : aR, r = left ? ul[RHO] : ur[RHO];
            const double g = left ? gamaL : gamaR, pr = left ? presL : presR;
            double sp[NG], spb[NG];
#pragma unroll
            for (int q = 0; q < NG; ++q) {
                sp[q] = left ? ul[NCV + q] : ur[NCV + q];
                spb[q] = 0.0;
            }
            const double Xb = lamb / (kTwo * a);
            gama_rev(T, sp, Xb * pr / r, spb);
#pragma unroll
            for (int q = 0; q < NG; ++q) {
                ubl[NCV + q] += left ? spb[q] : 0.0;
                ubr[NCV + q] += left ? 0.0 : spb[q];
            }
            presLb = left ? Xb * g / r : 0.0;
            presRb = left ? 0.0 : Xb * g / r;
            const double rb = -Xb * (g * pr / r) / r;
            ubl[RHO] += left ? rb : 0.0;
            ubr[RHO] += left ? 0.0 : rb;
        }
        double rhoeLb = 0.0, rhoeRb = 0.0;
        if (presLb != 0.0) p_rev(T, rhoeL, ul + NCV, presLb, rhoeLb, ubl + NCV);
        if (presRb != 0.0) p_rev(T, rhoeR, ul + NCV, presRb, rhoeRb, ubl + NCV);
        rhoe_rev(ul, rhoeLb, ubl);
        rhoe_rev(ur, rhoeRb, ubr);
    }

    // -----------------------------------------------------------------------
    // Hyperbolic (WENO5 LLF) part of a face flux, pass 1 of face_flux<P,5,D>
    // for a face with hyp = true. St provides the stencil along D at positions
    // q = 0..5 (offsets -3..+2, with the IB first-order mapping applied):
    //   u(q, c) conserved component c (0..NC-1), vel(q, a), pres(q) the
    //   geteulerflux primitives; fb(c) the adjoint of flux component c;
    //   add_u(q, c, x), add_v(q, a, x), add_p(q, x) accumulate adjoints of u
    //   and of the primitives (the caller turns the primitive adjoints into
    //   conserved ones with prims_rev per cell).
    // Positions 2 and 3 are always the true left and right cells. Loop-indexed
    // values come through St (no register array is indexed by the component).
    // -----------------------------------------------------------------------
    template <int D, class St>
    MFD static void face_hyp_rev(const Tab &T, St &st) {
        double lam;
        {
            double ul[NC], ur[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) { ul[c] = st.u(2, c); ur[c] = st.u(3, c); }
            lam = eigen(T, D, ul, ur);
        }
        double lamb = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double hb = (c < NCV ? kOne : T.advect_flag[c - NCV]) * st.fb(c);
            if (hb == 0.0) continue;
            double uc[6], fc[6], g[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                uc[q] = st.u(q, c);
                const double vn = st.vel(q, D);
                const double p = st.pres(q);
                if (c == RHO) fc[q] = uc[q] * vn;
                else if (c >= RHOU && c <= RHOW) {
                    const double va = st.vel(q, c - RHOU);
                    const double mm = st.u(q, RHO) * va;
                    fc[q] = (c - RHOU == D) ? (mm * va + p) : (mm * vn);
                } else if (c == RHOE) fc[q] = (uc[q] + p) * vn;
                else fc[q] = uc[q] * vn;
            }
            // reconstructions of u_c: their adjoints go to u_c directly; the
            // reverse passes return the reconstructed values lam multiplies
#pragma unroll
            for (int q = 0; q < 6; ++q) g[q] = 0.0;
            const double up = weno_plus_rev(uc, kHalf * lam * hb, g);
            const double um = weno_minus_rev(uc + 1, -kHalf * lam * hb, g + 1);
            lamb += kHalf * hb * (up - um);
#pragma unroll
            for (int q = 0; q < 6; ++q) st.add_u(q, c, g[q]);
            // reconstructions of the flux component: through the cell's flux
            // into u_c, rho, the velocities and the pressure
#pragma unroll
            for (int q = 0; q < 6; ++q) g[q] = 0.0;
            weno_plus_rev(fc, kHalf * hb, g);
            weno_minus_rev(fc + 1, kHalf * hb, g + 1);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double f = g[q];
                const double vn = st.vel(q, D);
                if (c == RHO) {
                    st.add_u(q, c, f * vn);
                    st.add_v(q, D, f * uc[q]);
                } else if (c >= RHOU && c <= RHOW) {
                    const int a = c - RHOU;
                    const double va = st.vel(q, a), r0 = st.u(q, RHO);
                    if (a == D) {
                        st.add_u(q, RHO, f * va * va);
                        st.add_v(q, D, f * kTwo * r0 * va);
                        st.add_p(q, f);
                    } else {
                        st.add_u(q, RHO, f * va * vn);
                        st.add_v(q, a, f * r0 * vn);
                        st.add_v(q, D, f * r0 * va);
                    }
                } else if (c == RHOE) {
                    const double p = st.pres(q);
                    st.add_u(q, c, f * vn);
                    st.add_p(q, f * vn);
                    st.add_v(q, D, f * (uc[q] + p));
                } else {
                    st.add_u(q, c, f * vn);
                    st.add_v(q, D, f * uc[q]);
                }
            }
        }
        double ul[NC], ur[NC], ubl[NC], ubr[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) { ul[c] = st.u(2, c); ur[c] = st.u(3, c); ubl[c] = 0.0; ubr[c] = 0.0; }
        eigen_rev(T, D, ul, ur, lamb, ubl, ubr);
#pragma unroll
        for (int c = 0; c < NC; ++c) { st.add_u(2, c, ubl[c]); st.add_u(3, c, ubr[c]); }
    }

    // -----------------------------------------------------------------------
    // Pass 2 of a face (face_pass2<P,D>): viscous stress and work, conduction,
    // species diffusion. Inputs: vL, vR (stored VELX..VELZ of the left/right
    // cells), G[q][axis] = d(vel q)/d(axis) as the forward forms it (normal
    // slot from vR - vL, transverse slots from the averaged neighbours),
    // visc/thcond/Ds the harmonic means. Outputs add to vLb, vRb (through the
    // normal derivative and the work term), TLb, TRb, cLb[NS], cRb[NS]; h1[q],
    // h2[q] are the adjoints of d(vel q) along T1 and T2 divided by 2 dx: the
    // transverse neighbours get +-h/2 (L+t1, R+t1: +h1/2; L-t1, R-t1: -h1/2;
    // the same with h2 for t2).
    // -----------------------------------------------------------------------
    template <int D>
    MFD static void pass2_rev(const double *Fb, bool ns, bool hyp, double visc, double thcond, const double *Ds, const double *dx,
                              const double *vL, const double *vR, const double (*G)[3], double *vLb, double *vRb, double *h1,
                              double *h2, double &TLb, double &TRb, double *cLb, double *cRb) {
        constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
        const double dxs = dx[D];
#pragma unroll
        for (int q = 0; q < 3; ++q) { h1[q] = 0.0; h2[q] = 0.0; }
        if (ns) {
            if (hyp) {
                const double delu = G[0][0] + G[1][1] + G[2][2];
                double t[3];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    t[a] = (a == D) ? visc * (kTwo * G[D][D] - kTwo3rd * delu) : visc * (G[D][a] + G[a][D]);
                double tb[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double avg = kHalf * (vL[a] + vR[a]);
                    tb[a] = -Fb[RHOU + a] - Fb[RHOE] * avg;
                    const double avgb = -Fb[RHOE] * t[a];
                    vLb[a] += kHalf * avgb;
                    vRb[a] += kHalf * avgb;
                }
                double Gb[3][3];
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int a = 0; a < 3; ++a) Gb[q][a] = 0.0;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (a == D) {
                        Gb[D][D] += kTwo * visc * tb[a];
#pragma unroll
                        for (int b = 0; b < 3; ++b) Gb[b][b] += -kTwo3rd * visc * tb[a];
                    } else {
                        Gb[D][a] += visc * tb[a];
                        Gb[a][D] += visc * tb[a];
                    }
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double gn = Gb[q][D] / dxs;
                    vRb[q] += gn;
                    vLb[q] -= gn;
                    h1[q] = Gb[q][T1] / (kTwo * dx[T1]);
                    h2[q] = Gb[q][T2] / (kTwo * dx[T2]);
                }
            }
            TRb += -Fb[RHOE] * thcond / dxs;
            TLb += Fb[RHOE] * thcond / dxs;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            cRb[s] += -Fb[NCV + s] * Ds[s] / dxs;
            cLb[s] += Fb[NCV + s] * Ds[s] / dxs;
        }
    }
    // the forward's velocity-gradient tensor of a face from the stored
    // velocities of L, R and the eight transverse neighbours (order L+t1, R+t1,
    // L-t1, R-t1, L+t2, R+t2, L-t2, R-t2)
    template <int D>
    MFD static void grad_tensor(const double *dx, const double *vL, const double *vR, const double (*tv)[3], double (*G)[3]) {
        constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            G[q][D] = (vR[q] - vL[q]) / dx[D];
            const double top = kHalf * (tv[0][q] + tv[1][q]), bot = kHalf * (tv[2][q] + tv[3][q]);
            G[q][T1] = (top - bot) / (kTwo * dx[T1]);
            const double frt = kHalf * (tv[4][q] + tv[5][q]), bck = kHalf * (tv[6][q] + tv[7][q]);
            G[q][T2] = (frt - bck) / (kTwo * dx[T2]);
        }
    }

    // -----------------------------------------------------------------------
    // Chemistry: spec_source of PPorous and the SSP-RK3 of chemistry_rk3
    // -----------------------------------------------------------------------
    // adjoint of src = spec_source(c, Tm): adds to cb, Tb, logAb
    MFD static void spec_source_rev(const Tab &T, const double *c, double Tm, const double *srcb, double *cb, double &Tb,
                                    double *logAb) {
        // unrolled over the table size so that logAb[r] stays in registers
#pragma unroll
        for (int r = 0; r < MF_MAX_REACTIONS; ++r) {
            if (r >= T.nreact) break;
            double rateb = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double nu = T.stoich[r][s];
                if (fabs(nu) > 0.0) rateb += nu * srcb[s];
            }
            if (rateb == 0.0) continue;
            double concmult = 1.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double nu = T.stoich[r][s];
                if (nu < 0.0) {
                    const double a = fabs(nu);
                    concmult *= (a == 1.0) ? c[s] : ((a == 2.0) ? c[s] * c[s] : pow(c[s], a));
                }
            }
            const double krate = exp(T.log_A[r] - T.Ea[r] / (T.R_univ * Tm));
            const double kb = rateb * concmult;
            logAb[r] += kb * krate;
            Tb += kb * krate * T.Ea[r] / (T.R_univ * Tm * Tm);
            const double cmb = rateb * krate;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double nu = T.stoich[r][s];
                if (!(nu < 0.0)) continue;
                // d concmult / d c_s: product of the other factors times d(c_s^a)
                double other = 1.0;
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) {
                    const double nu2 = T.stoich[r][s2];
                    if (s2 == s || !(nu2 < 0.0)) continue;
                    const double a2 = fabs(nu2);
                    other *= (a2 == 1.0) ? c[s2] : ((a2 == 2.0) ? c[s2] * c[s2] : pow(c[s2], a2));
                }
                const double a = fabs(nu);
                const double dfac = (a == 1.0) ? 1.0 : ((a == 2.0) ? 2.0 * c[s] : a * pow(c[s], a - 1.0));
                cb[s] += cmb * other * dfac;
            }
        }
    }
    // forward source at a stage state (same arithmetic as the plugin call in chemistry_rk3)
    MFD static void source_at(const Tab &T, const double *S, const double *sp, double *F) {
        double src[NS];
        P::spec_source(T, ChemCellFwd{S, sp}, src);
#pragma unroll
        for (int s = 0; s < NS; ++s) F[s] = 0.0 + src[s];
    }
    struct ChemCellFwd {
        const double *S;
        const double *sp;
        MFD double operator()(int n) const { return n >= FLO ? sp[n - FLO] : S[n]; }
    };
    // adjoint of chemistry_rk3 on the species of S (TEMP read from S); lb holds the
    // adjoint of the output species on entry and of the input species on exit
    MFD static void chem_rev(const Tab &T, const double *S, double dt, double *lb, double &Tb, double *logAb) {
        double S0[NS], F0[NS], F1[NS], st1[NS], st2[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) S0[s] = S[FLO + s];
        const double a10 = 1.0, a20 = 0.25, a21 = 0.25;
        const double w0 = 1. / 6., w1 = 1. / 6., w2 = 2. / 3.;
        source_at(T, S, S0, F0);
#pragma unroll
        for (int s = 0; s < NS; ++s) st1[s] = S0[s] + (dt * a10) * F0[s];
        source_at(T, S, st1, F1);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double v = S0[s];
            v += (dt * a20) * F0[s];
            v += (dt * a21) * F1[s];
            st2[s] = v;
        }
        const double Tm = S[TEMP];
        double S0b[NS], F0b[NS], F1b[NS], F2b[NS], stb[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            S0b[s] = lb[s];
            F0b[s] = (dt * w0) * lb[s];
            F1b[s] = (dt * w1) * lb[s];
            F2b[s] = (dt * w2) * lb[s];
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) stb[s] = 0.0;
        spec_source_rev(T, st2, Tm, F2b, stb, Tb, logAb);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            S0b[s] += stb[s];
            F0b[s] += (dt * a20) * stb[s];
            F1b[s] += (dt * a21) * stb[s];
            stb[s] = 0.0;
        }
        spec_source_rev(T, st1, Tm, F1b, stb, Tb, logAb);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            S0b[s] += stb[s];
            F0b[s] += (dt * a10) * stb[s];
            stb[s] = 0.0;
        }
        spec_source_rev(T, S0, Tm, F0b, stb, Tb, logAb);
#pragma unroll
        for (int s = 0; s < NS; ++s) lb[s] = S0b[s] + stb[s];
    }

    // -----------------------------------------------------------------------
    // Post-step pass of one cell: chemistry, update_primitive_vars, clip.
    // post_cell is the forward (the arithmetic of k_post); post_rev turns the
    // adjoint of the output (lb, NV) into the adjoint of the input S3 in place.
    // -----------------------------------------------------------------------
    MFD static void post_cell(const Tab &T, double *S, double dt, int clip) {
        if (P::kHasChemistry) chemistry_rk3<P>(T, S, dt);
        if (S[VFRAC] == kOne) {
            double u[NC], p[NCV];
#pragma unroll
            for (int c = 0; c < NCV; ++c) u[c] = S[c];
#pragma unroll
            for (int s = 0; s < NS; ++s) u[NCV + s] = S[FLO + s];
            N::cons_to_prim(T, u, p);
#pragma unroll
            for (int c = 0; c < NCV; ++c) S[DENS + c] = p[c];
            S[TEMP] = P::get_t_from_rpc(T, S[DENS], S[PRES], u + NCV);
        }
        if (clip) {
#pragma unroll
            for (int s = 0; s < NS; ++s) if (S[FLO + s] < 0.0) S[FLO + s] = 0.0;
        }
    }
    MFD static void post_rev(const Tab &T, const double *S3, double dt, int clip, double *lb, double *logAb) {
        double S[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) S[n] = S3[n];
        if (P::kHasChemistry) chemistry_rk3<P>(T, S, dt);
        // S: after chemistry (input of the primitive update)
        const bool prim = S[VFRAC] == kOne;
        double u[NC], p[NCV];
#pragma unroll
        for (int c = 0; c < NCV; ++c) u[c] = S[c];
#pragma unroll
        for (int s = 0; s < NS; ++s) u[NCV + s] = S[FLO + s];
        if (prim) N::cons_to_prim(T, u, p);
        if (clip) {
#pragma unroll
            for (int s = 0; s < NS; ++s) if (S[FLO + s] < 0.0) lb[FLO + s] = 0.0;
        }
        if (prim) {
            double ub[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) ub[c] = 0.0;
            double presb = lb[PRES];
            t_rev(T, p[4], u + NCV, lb[TEMP], presb, ub + NCV);
            const double vbar[3] = {lb[VELX], lb[VELY], lb[VELZ]};
            prims_rev(T, u, vbar, presb, ub);
            ub[RHO] += lb[DENS];
#pragma unroll
            for (int n = DENS; n <= TEMP; ++n) lb[n] = 0.0;
#pragma unroll
            for (int c = 0; c < NCV; ++c) lb[c] += ub[c];
#pragma unroll
            for (int s = 0; s < NS; ++s) lb[FLO + s] += ub[NCV + s];
        }
        if (P::kHasChemistry) {
            double Tb = 0.0;
            chem_rev(T, S3, dt, lb + FLO, Tb, logAb);
            lb[TEMP] += Tb;
        }
    }

    // -----------------------------------------------------------------------
    // update_cutcells of one cut cell (cutcell_values): ob[NV] is the adjoint
    // of the cell's updated values; add(i, j, k, n, x) receives the adjoints of
    // the raw neighbourhood values it was computed from. The IDW weights depend
    // on vfrac only.
    // adjoint of the cut cell's prim_to_cons and of its primitive outputs: ob
    // is the adjoint of the updated cell (NV), out its forward values; the
    // results are the adjoints of the interpolated density, pressure,
    // temperature, velocities and species
    MFD static void cut_prim_bars(const Tab &T, const double *out, const double *ob, double &bD, double &bP, double &bT,
                                  double *bV, double *bS) {
        const double pv[5] = {out[DENS], out[VELX], out[VELY], out[VELZ], out[PRES]};
        const double *sp = out + FLO;
        const double e = P::get_e_from_rpc(T, pv[0], pv[4], sp);
        double pvb[5] = {ob[RHO], ob[RHOU] * pv[0], ob[RHOV] * pv[0], ob[RHOW] * pv[0], 0.0};
        double spb[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) spb[q] = 0.0;
        pvb[0] += ob[RHOU] * pv[1] + ob[RHOV] * pv[2] + ob[RHOW] * pv[3];
        pvb[0] += ob[RHOE] * (e + kHalf * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]));
        pvb[1] += ob[RHOE] * pv[0] * pv[1];
        pvb[2] += ob[RHOE] * pv[0] * pv[2];
        pvb[3] += ob[RHOE] * pv[0] * pv[3];
        e_rev(T, pv[0], pv[4], sp, ob[RHOE] * pv[0], pvb[0], pvb[4], spb);
        bD = ob[DENS] + pvb[0];
        bP = ob[PRES] + pvb[4];
        bT = ob[TEMP];
        bV[0] = ob[VELX] + pvb[1];
        bV[1] = ob[VELY] + pvb[2];
        bV[2] = ob[VELZ] + pvb[3];
#pragma unroll
        for (int q = 0; q < NS; ++q) bS[q] = ob[FLO + q] + spb[q];
    }

    // -----------------------------------------------------------------------
    // fwd (may be null): the cell's forward result, read with stride fstride
    // instead of recomputing it
    template <class Add>
    MFD static void cut_rev(const Tab &T, const CutArgs &a, int i, int j, int k, const double *ob, Add &add,
                            const double *fwd = nullptr, int64_t fstride = 1) {
        const View &s = a.s;
        const double vf = s(i, j, k, VFRAC);
        if (!(vf > kZero && vf < kOne)) {
            for (int n = 0; n < NV; ++n) add(i, j, k, n, ob[n]);
            return;
        }
        double out[NV];
        if (fwd) {
#pragma unroll
            for (int n = 0; n < NV; ++n) out[n] = fwd[(int64_t)n * fstride];
        } else {
            cutcell_values<P>(T, a, i, j, k, out);
        }
        const double *dx = a.dx;
        double lsg[3] = {0.0, 0.0, 0.0};
        const double zero = 0.0;
        for (int n = -1; n <= 1; n++)
            for (int m = -1; m <= 1; m++)
                for (int l = -1; l <= 1; l++) {
                    double v = s(i + l, j + m, k + n, VFRAC);
                    double bx = ib_basisder1d(l, zero / dx[0]) / dx[0] * ib_basis1d(m, zero / dx[1]) * ib_basis1d(n, zero / dx[2]);
                    double by = ib_basis1d(l, zero / dx[0]) * (ib_basisder1d(m, zero / dx[1]) / dx[1]) * ib_basis1d(n, zero / dx[2]);
                    double bz = ib_basis1d(l, zero / dx[0]) * ib_basis1d(m, zero / dx[1]) * (ib_basisder1d(n, zero / dx[2]) / dx[2]);
                    lsg[0] += v * bx;
                    lsg[1] += v * by;
                    lsg[2] += v * bz;
                }
        const double mag = sqrt(lsg[0] * lsg[0] + lsg[1] * lsg[1] + lsg[2] * lsg[2]);
        double dist = kZero, xp[3] = {kZero, kZero, kZero};
        if (mag > kTiny) {
            dist = (vf - kHalf) / mag;
#pragma unroll
            for (int d = 0; d < 3; d++) xp[d] = -dist * (lsg[d] / mag);
        }
        const double dxmin = fmin(dx[0], fmin(dx[1], dx[2]));
        const double dnd[3] = {dx[0] / dxmin, dx[1] / dxmin, dx[2] / dxmin};
        const double xnd[3] = {xp[0] / dxmin, xp[1] / dxmin, xp[2] / dxmin};
        const double cutoff = dxmin * 1e-5;
        const bool doT = !a.conj_ht, doS = !a.spec_in_solid;
        double bD, bP, bT, bV[3], bS[NS];
        cut_prim_bars(T, out, ob, bD, bP, bT, bV, bS);
        add(i, j, k, VFRAC, ob[VFRAC]);
        if (!doS) {
#pragma unroll
            for (int q = 0; q < NS; ++q) add(i, j, k, FLO + q, bS[q]);
        }
        if (!(fabs(dist) < cutoff)) {
            double sw = 0.0;
            for (int n = -2; n <= 2; n++)
                for (int m = -2; m <= 2; m++)
                    for (int l = -2; l <= 2; l++) {
                        if (l == 0 && m == 0 && n == 0) continue;
                        if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
                            double a0 = l * dnd[0], a1 = m * dnd[1], a2 = n * dnd[2];
                            sw += 1.0 / (a0 * a0 + a1 * a1 + a2 * a2 + kTiny2);
                        }
                    }
            const double dn = fabs(dist) / dxmin;
            const double den = sw + 1.0 / (dn * dn);
            for (int n = -2; n <= 2; n++)
                for (int m = -2; m <= 2; m++)
                    for (int l = -2; l <= 2; l++) {
                        if (l == 0 && m == 0 && n == 0) continue;
                        if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
                            double a0 = l * dnd[0], a1 = m * dnd[1], a2 = n * dnd[2];
                            double w2 = a0 * a0 + a1 * a1 + a2 * a2 + kTiny2;
#pragma unroll
                            for (int q = 0; q < 3; ++q) add(i + l, j + m, k + n, VELX + q, bV[q] / w2 / den);
                        }
                    }
        }
        {
            double sw = 0.0;
            for (int n = -2; n <= 2; n++)
                for (int m = -2; m <= 2; m++)
                    for (int l = -2; l <= 2; l++)
                        if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
                            double a0 = l * dnd[0] - xnd[0], a1 = m * dnd[1] - xnd[1], a2 = n * dnd[2] - xnd[2];
                            sw += 1.0 / (a0 * a0 + a1 * a1 + a2 * a2 + kTiny2);
                        }
            if (sw > 0.0) {
                for (int n = -2; n <= 2; n++)
                    for (int m = -2; m <= 2; m++)
                        for (int l = -2; l <= 2; l++)
                            if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
                                double a0 = l * dnd[0] - xnd[0], a1 = m * dnd[1] - xnd[1], a2 = n * dnd[2] - xnd[2];
                                const double f = 1.0 / (a0 * a0 + a1 * a1 + a2 * a2 + kTiny2) / sw;
                                add(i + l, j + m, k + n, PRES, bP * f);
                                if (doT) add(i + l, j + m, k + n, TEMP, bT * f);
                                add(i + l, j + m, k + n, DENS, bD * f);
                                if (doS) {
#pragma unroll
                                    for (int q = 0; q < NS; ++q) add(i + l, j + m, k + n, FLO + q, bS[q] * f);
                                }
                            }
            } else {
                add(i, j, k, PRES, bP);
                if (doT) add(i, j, k, TEMP, bT);
                add(i, j, k, DENS, bD);
                if (doS) {
#pragma unroll
                    for (int q = 0; q < NS; ++q) add(i, j, k, FLO + q, bS[q]);
                }
            }
        }
        if (a.conj_ht) {
            for (int n = -1; n <= 1; n++)
                for (int m = -1; m <= 1; m++)
                    for (int l = -1; l <= 1; l++)
                        if (!(l == 0 && m == 0 && n == 0)) add(i + l, j + m, k + n, TEMP, bT / 26.0);
        }
    }

    // -----------------------------------------------------------------------
    // cut_rev in gather form (conj_ht = 0). cut_coef, per cut cell with static
    // geometry g (wall point xnd[3], distance, velocity IDW denominator,
    // scalar weight sum): the terms the cell's fluid neighbours receive,
    // divided by the weight sums, as a record of CUTREC doubles
    //   rec[0..2] velocity adjoints / den (0 when the wall is within the
    //   cutoff), rec[3] pressure, rec[4] temperature, rec[5] density,
    //   rec[6..6+NS) species (scalars / sw, 0 without fluid neighbours),
    //   rec[6+NS..9+NS) the wall point xnd;
    // the terms of the cell itself go to self(n, v). cut_gather, per fluid
    // cell: the sum over the cut cells c = p - o of its mask bits o of the
    // records times the IDW weights of offset o, in the same record order.
    // -----------------------------------------------------------------------
    static constexpr int CUTREC = (9 + NS + 1) & ~1;
    template <class Self>
    MFD static void cut_coef(const Tab &T, const CutArgs &a, const double *g, const double *ob, const double *out, double *rec,
                             Self &self) {
#pragma unroll
        for (int q = 0; q < CUTREC; ++q) rec[q] = 0.0;
        const double vf = out[VFRAC];
        if (!(vf > kZero && vf < kOne)) {
            for (int n = 0; n < NV; ++n) self(n, ob[n]);
            return;
        }
        const double *dx = a.dx;
        const double dxmin = fmin(dx[0], fmin(dx[1], dx[2]));
        const double cutoff = dxmin * 1e-5;
        const bool doT = !a.conj_ht, doS = !a.spec_in_solid;
        double bD, bP, bT, bV[3], bS[NS];
        cut_prim_bars(T, out, ob, bD, bP, bT, bV, bS);
        self(VFRAC, ob[VFRAC]);
        if (!doS) {
#pragma unroll
            for (int q = 0; q < NS; ++q) self(FLO + q, bS[q]);
        }
        if (!(fabs(g[3]) < cutoff)) {
#pragma unroll
            for (int q = 0; q < 3; ++q) rec[q] = bV[q] / g[4];
        }
        const double sw = g[5];
        if (sw > 0.0) {
            rec[3] = bP / sw;
            rec[4] = doT ? bT / sw : 0.0;
            rec[5] = bD / sw;
#pragma unroll
            for (int q = 0; q < NS; ++q) rec[6 + q] = doS ? bS[q] / sw : 0.0;
        } else {
            self(PRES, bP);
            if (doT) self(TEMP, bT);
            self(DENS, bD);
            if (doS) {
#pragma unroll
                for (int q = 0; q < NS; ++q) self(FLO + q, bS[q]);
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) rec[6 + NS + d] = g[d];
    }
    // rec_of(o0, o1, o2) returns the record of the cut cell at offset -o from
    // the fluid cell; res[0..6+NS) as the record layout
    template <class RecOf>
    MFD static void cut_gather(const double *dx, const unsigned *mb, RecOf &rec_of, double *res) {
        const double dxmin = fmin(dx[0], fmin(dx[1], dx[2]));
        const double dnd[3] = {dx[0] / dxmin, dx[1] / dxmin, dx[2] / dxmin};
#pragma unroll
        for (int q = 0; q < 6 + NS; ++q) res[q] = 0.0;
        for (int w = 0; w < 4; ++w) {
            unsigned bits = mb[w];
            while (bits) {
                const int b = __builtin_ctz(bits);
                bits &= bits - 1u;
                const int nb = w * 32 + b;
                const int l = nb % 5 - 2, m = (nb / 5) % 5 - 2, n = nb / 25 - 2;
                const double *r = rec_of(l, m, n);
                const double a0 = l * dnd[0], a1 = m * dnd[1], a2 = n * dnd[2];
                const double wv = 1.0 / (a0 * a0 + a1 * a1 + a2 * a2 + kTiny2);
                const double b0 = a0 - r[6 + NS], b1 = a1 - r[7 + NS], b2 = a2 - r[8 + NS];
                const double ws = 1.0 / (b0 * b0 + b1 * b1 + b2 * b2 + kTiny2);
#pragma unroll
                for (int q = 0; q < 3; ++q) res[q] += r[q] * wv;
#pragma unroll
                for (int q = 3; q < 6 + NS; ++q) res[q] += r[q] * ws;
            }
        }
    }

    // -----------------------------------------------------------------------
    // externalbc of PPorous: adjoint of s_ext w.r.t. s_int and in_vel
    // -----------------------------------------------------------------------
    MFD static void externalbc_rev(const Tab &T, const double *s_int, const double *lext, int idir, int sgn, double *lint,
                                   double &invelb) {
        double spec[NS], specb[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) { spec[s] = kZero; specb[s] = lext[FLO + s]; }
        double dens, pres, vel[3];
        const bool inflow = idir == 0 && sgn > 0;
        if (inflow) {
            pres = s_int[PRES];
            const double ctot = pres / (T.R_univ * T.in_temp);
#pragma unroll
            for (int s = 0; s < NG; ++s) spec[s] = T.in_spec[s] * ctot;
            dens = P::get_r_from_c(T, spec);
            vel[0] = T.in_vel; vel[1] = kZero; vel[2] = kZero;
        } else {
            dens = s_int[DENS];
            pres = T.out_pres;
            vel[0] = s_int[VELX]; vel[1] = s_int[VELY]; vel[2] = s_int[VELZ];
#pragma unroll
            for (int s = 0; s < NS; ++s) spec[s] = s_int[FLO + s];
        }
        const double e = P::get_e_from_rpc(T, dens, pres, spec);
        const double v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
        double densb = lext[DENS] + lext[RHO] + lext[RHOE] * (e + kHalf * v2);
        double velb[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            densb += lext[RHOU + a] * vel[a];
            velb[a] = lext[VELX + a] + lext[RHOU + a] * dens + lext[RHOE] * dens * vel[a];
        }
        double presb = lext[PRES];
        e_rev(T, dens, pres, spec, lext[RHOE] * dens, densb, presb, specb);
        if (inflow) {
            invelb += velb[0];
#pragma unroll
            for (int s = 0; s < NG; ++s) specb[s] += densb * T.molwt[s];
            double ctotb = 0.0;
#pragma unroll
            for (int s = 0; s < NG; ++s) ctotb += specb[s] * T.in_spec[s];
            presb += ctotb / (T.R_univ * T.in_temp);
            lint[PRES] += presb;
        } else {
            lint[DENS] += densb;
            lint[VELX] += velb[0];
            lint[VELY] += velb[1];
            lint[VELZ] += velb[2];
            lint[TEMP] += lext[TEMP];
            lint[VFRAC] += lext[VFRAC];
#pragma unroll
            for (int s = 0; s < NS; ++s) lint[FLO + s] += specb[s];
        }
    }
};

}  // namespace mfd
