// mf_hier_oracle.h -- TEST INFRASTRUCTURE ONLY (part of mf_oracle.cpp).
//
// CPU restatement of mesoflow's multi-level advance with subcycling:
//   EvolveAMR's step body over levels (Source/mflo_evolve.cpp:18-89),
//   timeStep's recursion (:175-249), Advance_coupled_strang per level
//   (:251-318) with the flux-register updates of compute_dsdt_flow (:655-681),
//   FillPatch at levels > 0 = FillPatchTwoLevels (Source/mflo_gridwork.cpp:
//   147-165) with GetData's time selection (Source/mflo.cpp:310-330),
//   ComputeDt over levels (Source/mflo_timestep.cpp:18-50), Reflux and
//   AverageDownTo (Source/mflo_evolve.cpp:243-248, Source/mflo_gridwork.cpp:
//   125-130), RemakeLevel / MakeNewLevelFromCoarse / ClearLevel of a regrid
//   (Source/mflo_gridwork.cpp:21-89).
// Two levels (max_level = 1), the configuration of the two-level runs.
//
// Every fine box gets its own Sborder, filled as AMReX does per fab: the
// coarse patch coarsen(grow(box, 3)) grown by one cell is the time
// interpolation of the coarse valid data (periodic images included) followed
// by the coarse physical boundary conditions at the fill time
// (FillPatchSingleLevel), cell_cons_interp of it over the whole region, the
// fine level's valid data (and its periodic images) copied over it, then the
// fine physical boundary conditions.
//
// Flux registers. The register of level 1 holds the coarse faces of the
// coarsened fine boxes whose outside coarse cell (periodic image included) is
// a valid coarse cell not covered by the fine level; faces whose outside cell
// is covered are left out (their reflux lands in cells that average_down
// overwrites). CrseInit takes the coarse face flux of the coarse box that owns
// the outside cell (where two coarse boxes share the face AMReX's copy order
// picks one of them; the owner of the refluxed cell is the choice under which
// the reflux restores exact conservation). A refluxed coarse cell takes the
// corrections of its faces in the order x-lo, x-hi, y-lo, y-hi, z-lo, z-hi.
// These orders are AMReX's (absent from the tree): parity unpinned; the
// tests pin the driver by conservation with periodic boundaries, by the
// single-level run it reduces to when the fine level covers the domain, and by
// the per-box composition of the coupling operators.

extern "C" int orc_average_down(mf_box cbx, mf_array4 fine, mf_array4 crse, int ratio, int ncomp);
extern "C" int orc_interp_cons(mf_box fbx, mf_array4 crse, mf_array4 fine, int ratio, int ncomp);

static mf_array4 a4_to_mf(const A4 &a) {
    mf_array4 v;
    std::memset(&v, 0, sizeof(v));
    v.p = a.p;
    v.jstride = a.js;
    v.kstride = a.ks;
    v.nstride = a.ns;
    v.ncomp = a.nc;
    for (int d = 0; d < 3; ++d) { v.begin[d] = a.b[d]; v.end[d] = a.e[d]; }
    return v;
}
static inline int fdiv(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

struct HierBase {
    virtual ~HierBase() {}
    virtual int add_level(const mf_box *boxes, int nb, const double *state) = 0;
    virtual int regrid_level(const mf_box *boxes, int nb, double time) = 0;
    virtual int step(int nsteps, int regrid_first, const mf_box *boxes, int nb) = 0;
    virtual int nlevels() const = 0;
    virtual LevelBase *level(int l) = 0;
    virtual void info(double *t, double *dt0, int *istep, double *res) = 0;
    int only_flow = 0;
    int do_reflux = 1;  // mflo.do_reflux
};

template <class P>
struct Hier : HierBase {
    static const int NV = Ops<P>::NV;
    std::vector<Level<P> *> L;   // L[0] belongs to the caller
    int ratio = 2;
    std::vector<int> nsub;       // nsubsteps (Source/mflo.cpp:56-61)
    std::vector<double> dt;      // dt[lev]
    std::vector<int> istep;
    struct RegFace {
        int dir, side;           // side 0: the fine region lies above the face
        int fbox, cbox;          // fine box of the inside, coarse box owning the outside cell
        int face[3];             // coarse face index (fine box's coarsened frame)
        int cface[3];            // the same face in the owner box's frame (periodic shift)
        int out[3];              // the outside coarse cell (inside the domain)
    };
    std::vector<RegFace> reg;
    std::vector<double> regv;    // [face][NV]
    struct Target { int c[3]; int e[6]; };
    std::vector<Target> targets;
    std::vector<unsigned char> ccov;  // coarse cells covered by the coarsened fine boxes

    ~Hier() override {
        for (size_t l = 1; l < L.size(); ++l) delete L[l];
        if (!L.empty()) { L[0]->fill_box = nullptr; L[0]->after_stage = nullptr; L[0]->keep_flux = false; }
    }
    int nlevels() const override { return (int)L.size(); }
    LevelBase *level(int l) override { return (l >= 0 && l < (int)L.size()) ? L[l] : nullptr; }
    void info(double *t, double *dt0, int *is, double *res) override {
        if (t) *t = L[0]->t_new;
        if (dt0) *dt0 = dt[0];
        if (is) *is = istep[0];
        if (res) for (int c = 0; c < NV; ++c) res[c] = L[0]->resnorms[c];
    }

    static long cidx(const Level<P> &lv, const int *c) {
        return (long)c[0] + (long)lv.n[0] * ((long)c[1] + (long)lv.n[1] * c[2]);
    }
    // periodic image inside the domain; false when outside a non-periodic side
    static bool wrap(const Level<P> &lv, const int *iv, int *w) {
        for (int d = 0; d < 3; ++d) {
            w[d] = iv[d];
            if (lv.geom.is_periodic[d]) {
                w[d] = ((iv[d] % lv.n[d]) + lv.n[d]) % lv.n[d];
            } else if (iv[d] < 0 || iv[d] >= lv.n[d]) {
                return false;
            }
        }
        return true;
    }
    int owner_box(const Level<P> &lv, const int *c) const {
        for (size_t b = 0; b < lv.boxes.size(); ++b) {
            const Box &B = lv.boxes[b];
            if (c[0] >= B.lo[0] && c[0] <= B.hi[0] && c[1] >= B.lo[1] && c[1] <= B.hi[1] && c[2] >= B.lo[2] &&
                c[2] <= B.hi[2])
                return (int)b;
        }
        return -1;
    }

    // fine boxes: inside the fine domain, aligned to the refinement ratio, disjoint
    int check_boxes(const mf_box *boxes, int nb, const int *nf) const {
        if (nb < 0 || (nb > 0 && !boxes)) return MF_ERR_BAD_ARG;
        std::vector<unsigned char> m((size_t)nf[0] * nf[1] * nf[2], 0);
        for (int b = 0; b < nb; ++b) {
            for (int d = 0; d < 3; ++d) {
                const int lo = boxes[b].lo[d], hi = boxes[b].hi[d];
                if (lo < 0 || hi >= nf[d] || hi < lo || lo % ratio || (hi + 1) % ratio) return MF_ERR_BAD_ARG;
            }
            for (int k = boxes[b].lo[2]; k <= boxes[b].hi[2]; ++k)
                for (int j = boxes[b].lo[1]; j <= boxes[b].hi[1]; ++j)
                    for (int i = boxes[b].lo[0]; i <= boxes[b].hi[0]; ++i) {
                        unsigned char &x = m[(size_t)i + (size_t)nf[0] * (j + (size_t)nf[1] * k)];
                        if (x) return MF_ERR_BAD_ARG;
                        x = 1;
                    }
        }
        return MF_OK;
    }

    Level<P> *new_fine(const mf_box *boxes, int nb) {
        const Level<P> &c = *L[0];
        mf_geom g = c.geom;
        for (int d = 0; d < 3; ++d) g.n_cell[d] = c.n[d] * ratio;
        Level<P> *F = static_cast<Level<P> *>(make_level<P>(&c.tab, &c.prm, &g, boxes, nb));
        F->vmask.assign((size_t)F->ncell(), 0);
        for (const Box &B : F->boxes)
            for (int k = B.lo[2]; k <= B.hi[2]; ++k)
                for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                    for (int i = B.lo[0]; i <= B.hi[0]; ++i)
                        F->vmask[(size_t)i + (size_t)F->n[0] * (j + (size_t)F->n[1] * k)] = 1;
        F->keep_flux = true;
        F->fill_box = [this](const Box &b, const std::vector<double> &src, double time, A4 &sb) {
            fill_two_levels(b, src, time, sb);
        };
        F->after_stage = [this](int stage, double dt_lev) { fine_add(stage, dt_lev); };
        F->resnorms.assign(NV, 0.0);
        return F;
    }

    void attach_coarse() {
        Level<P> &c = *L[0];
        c.keep_flux = true;
        c.after_stage = [this](int stage, double dt_lev) {
            if (L.size() > 1) crse_init(stage, dt_lev);
        };
    }

    // register faces and reflux targets of the current fine boxes
    void build_registers() {
        reg.clear();
        targets.clear();
        if (L.size() < 2) return;
        const Level<P> &c = *L[0];
        const Level<P> &f = *L[1];
        ccov.assign((size_t)c.ncell(), 0);
        for (const Box &B : f.boxes)
            for (int k = B.lo[2] / ratio; k <= B.hi[2] / ratio; ++k)
                for (int j = B.lo[1] / ratio; j <= B.hi[1] / ratio; ++j)
                    for (int i = B.lo[0] / ratio; i <= B.hi[0] / ratio; ++i) {
                        const int cc[3] = {i, j, k};
                        ccov[(size_t)cidx(c, cc)] = 1;
                    }
        std::map<long, size_t> tmap;
        for (size_t fb = 0; fb < f.boxes.size(); ++fb) {
            int clo[3], chi[3];
            for (int d = 0; d < 3; ++d) { clo[d] = f.boxes[fb].lo[d] / ratio; chi[d] = f.boxes[fb].hi[d] / ratio; }
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                    const int fidx = side == 0 ? clo[d] : chi[d] + 1;
                    for (int b2 = clo[t2]; b2 <= chi[t2]; ++b2)
                        for (int a1 = clo[t1]; a1 <= chi[t1]; ++a1) {
                            RegFace e;
                            e.dir = d;
                            e.side = side;
                            e.fbox = (int)fb;
                            e.face[d] = fidx;
                            e.face[t1] = a1;
                            e.face[t2] = b2;
                            int o[3] = {e.face[0], e.face[1], e.face[2]};
                            if (side == 0) o[d] -= 1;
                            if (!wrap(c, o, e.out)) continue;
                            if (ccov[(size_t)cidx(c, e.out)]) continue;
                            for (int q = 0; q < 3; ++q) e.cface[q] = e.face[q];
                            e.cface[d] += e.out[d] - o[d];
                            e.cbox = owner_box(c, e.out);
                            const long key = cidx(c, e.out);
                            auto it = tmap.find(key);
                            if (it == tmap.end()) {
                                Target t;
                                for (int q = 0; q < 3; ++q) t.c[q] = e.out[q];
                                for (int q = 0; q < 6; ++q) t.e[q] = -1;
                                targets.push_back(t);
                                it = tmap.emplace(key, targets.size() - 1).first;
                            }
                            // side 1: the fine region lies below the face, the face is the cell's low face
                            targets[it->second].e[2 * d + (side == 1 ? 0 : 1)] = (int)reg.size();
                            reg.push_back(e);
                        }
                }
        }
        regv.assign(reg.size() * NV, 0.0);
    }

    // FluxRegister::CrseInit(flux[d], d, 0, 0, ncomp, -dt dA w, ADD) after a coarse stage
    void crse_init(int stage, double dt_lev) {
        Level<P> &c = *L[0];
        const double w = stage == 3 ? two3rd : two3rd * half * half;
        for (size_t q = 0; q < reg.size(); ++q) {
            const RegFace &e = reg[q];
            const int d = e.dir;
            const double dA = (d == 0) ? c.dx[1] * c.dx[2] : ((d == 1) ? c.dx[0] * c.dx[2] : c.dx[0] * c.dx[1]);
            const double scale = -dt_lev * dA * w;
            const A4 F = c.box_flux_view((size_t)e.cbox, d);
            for (int n = 0; n < NV; ++n) regv[q * NV + n] += scale * F(e.cface[0], e.cface[1], e.cface[2], n);
        }
    }
    // FluxRegister::FineAdd(flux[d], d, 0, 0, ncomp, dt dA w) after a fine stage
    void fine_add(int stage, double dt_lev) {
        Level<P> &f = *L[1];
        const double w = stage == 3 ? two3rd : two3rd * half * half;
        for (size_t q = 0; q < reg.size(); ++q) {
            const RegFace &e = reg[q];
            const int d = e.dir, t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            const double dA = (d == 0) ? f.dx[1] * f.dx[2] : ((d == 1) ? f.dx[0] * f.dx[2] : f.dx[0] * f.dx[1]);
            const double scale = dt_lev * dA * w;
            const A4 F = f.box_flux_view((size_t)e.fbox, d);
            for (int n = 0; n < NV; ++n) {
                double acc = regv[q * NV + n];
                for (int b = 0; b < ratio; ++b)
                    for (int a = 0; a < ratio; ++a) {
                        int p[3];
                        p[d] = e.face[d] * ratio;
                        p[t1] = e.face[t1] * ratio + a;
                        p[t2] = e.face[t2] * ratio + b;
                        acc = acc + scale * F(p[0], p[1], p[2], n);
                    }
                regv[q * NV + n] = acc;
            }
        }
    }
    // FluxRegister::Reflux(phi_new[0], 1.0, 0, 0, ncomp, geom[0])
    void reflux() {
        Level<P> &c = *L[0];
        const double vol = c.dx[0] * c.dx[1] * c.dx[2];
        A4 S = c.valid(c.phi_new);
        for (const Target &t : targets)
            for (int s = 0; s < 6; ++s) {
                const int q = t.e[s];
                if (q < 0) continue;
                const double mult = reg[(size_t)q].side == 0 ? -1.0 : 1.0;
                for (int n = 0; n < NV; ++n)
                    S(t.c[0], t.c[1], t.c[2], n) = S(t.c[0], t.c[1], t.c[2], n) + (mult * regv[(size_t)q * NV + n]) / vol;
            }
    }
    void average_down() {
        Level<P> &c = *L[0];
        Level<P> &f = *L[1];
        A4 Fv = f.valid(f.phi_new), Cv = c.valid(c.phi_new);
        for (const Box &B : f.boxes) {
            mf_box cb;
            for (int d = 0; d < 3; ++d) { cb.lo[d] = B.lo[d] / ratio; cb.hi[d] = B.hi[d] / ratio; }
            orc_average_down(cb, a4_to_mf(Fv), a4_to_mf(Cv), ratio, NV);
        }
    }

    // FillPatchTwoLevels of the region `sb` (a fine box grown by 3, or a box of
    // a regrid) at `time`; src: the fine level's data at that time (the data
    // of GetData), over the fine level's current boxes
    void fill_region(const std::vector<double> &src, const std::vector<Box> &fboxes, double time, A4 &sb) {
        Level<P> &c = *L[0];
        Level<P> &f = *L[1];
        int glo[3], ghi[3], clo[3], chi[3];
        for (int d = 0; d < 3; ++d) {
            glo[d] = sb.b[d];
            ghi[d] = sb.e[d] - 1;
            clo[d] = fdiv(glo[d], ratio) - 1;
            chi[d] = fdiv(ghi[d], ratio) + 1;
        }
        // coarse patch: time interpolation of the valid coarse data (GetData of level 0)
        std::vector<double> ctv((size_t)(chi[0] - clo[0] + 1) * (chi[1] - clo[1] + 1) * (chi[2] - clo[2] + 1) * NV, 0.0);
        A4 ct = make_a4(ctv.data(), clo, chi, NV);
        const double t_old = c.t_old, t_new = c.t_new;
        const double teps = (t_new - t_old) * 1.e-3;
        int mode = 2;
        if (time > t_new - teps && time < t_new + teps) mode = 1;
        else if (time > t_old - teps && time < t_old + teps) mode = 0;
        const double alpha = (mode == 2) ? (t_new - time) / (t_new - t_old) : 0.0;
        const double beta = (mode == 2) ? (time - t_old) / (t_new - t_old) : 0.0;
        A4 So = c.valid(c.phi_old), Sn = c.valid(c.phi_new);
        for (int k = clo[2]; k <= chi[2]; ++k)
            for (int j = clo[1]; j <= chi[1]; ++j)
                for (int i = clo[0]; i <= chi[0]; ++i) {
                    const int iv[3] = {i, j, k};
                    int w[3];
                    if (!wrap(c, iv, w)) continue;
                    for (int n = 0; n < NV; ++n) {
                        if (mode == 1) ct(i, j, k, n) = Sn(w[0], w[1], w[2], n);
                        else if (mode == 0) ct(i, j, k, n) = So(w[0], w[1], w[2], n);
                        else {
                            const double a = alpha * So(w[0], w[1], w[2], n);
                            const double b = beta * Sn(w[0], w[1], w[2], n);
                            ct(i, j, k, n) = a + b;
                        }
                    }
                }
        // the coarse physical boundary conditions at the fill time
        for (int phase = 1; phase <= 3; ++phase)
            for (int k = clo[2]; k <= chi[2]; ++k)
                for (int j = clo[1]; j <= chi[1]; ++j)
                    for (int i = clo[0]; i <= chi[0]; ++i) {
                        int iv[3] = {i, j, k}, out_np = 0;
                        for (int d = 0; d < 3; ++d)
                            if (!c.geom.is_periodic[d] && (iv[d] < 0 || iv[d] >= c.n[d])) out_np++;
                        if (out_np == phase) c.fill_cell(ct, iv, time);
                    }
        // cell_cons_interp over the region
        mf_box gb;
        for (int d = 0; d < 3; ++d) { gb.lo[d] = glo[d]; gb.hi[d] = ghi[d]; }
        orc_interp_cons(gb, a4_to_mf(ct), a4_to_mf(sb), ratio, NV);
        // the fine valid data and its periodic images
        A4 Fs = f.valid(const_cast<std::vector<double> &>(src));
        int span[3][2];
        for (int d = 0; d < 3; ++d) {
            span[d][0] = f.geom.is_periodic[d] ? -1 : 0;
            span[d][1] = f.geom.is_periodic[d] ? 1 : 0;
        }
        for (const Box &B : fboxes)
            for (int pk = span[2][0]; pk <= span[2][1]; ++pk)
                for (int pj = span[1][0]; pj <= span[1][1]; ++pj)
                    for (int pi = span[0][0]; pi <= span[0][1]; ++pi) {
                        const int sh[3] = {pi * f.n[0], pj * f.n[1], pk * f.n[2]};
                        int lo[3], hi[3];
                        bool empty = false;
                        for (int d = 0; d < 3; ++d) {
                            lo[d] = std::max(glo[d], B.lo[d] + sh[d]);
                            hi[d] = std::min(ghi[d], B.hi[d] + sh[d]);
                            if (hi[d] < lo[d]) empty = true;
                        }
                        if (empty) continue;
                        for (int n = 0; n < NV; ++n)
                            for (int k = lo[2]; k <= hi[2]; ++k)
                                for (int j = lo[1]; j <= hi[1]; ++j)
                                    for (int i = lo[0]; i <= hi[0]; ++i)
                                        sb(i, j, k, n) = Fs(i - sh[0], j - sh[1], k - sh[2], n);
                    }
        // the fine physical boundary conditions
        for (int phase = 1; phase <= 3; ++phase)
            for (int k = glo[2]; k <= ghi[2]; ++k)
                for (int j = glo[1]; j <= ghi[1]; ++j)
                    for (int i = glo[0]; i <= ghi[0]; ++i) {
                        int iv[3] = {i, j, k}, out_np = 0;
                        for (int d = 0; d < 3; ++d)
                            if (!f.geom.is_periodic[d] && (iv[d] < 0 || iv[d] >= f.n[d])) out_np++;
                        if (out_np == phase) f.fill_cell(sb, iv, time);
                    }
    }
    void fill_two_levels(const Box &b, const std::vector<double> &src, double time, A4 &sb) {
        (void)b;
        fill_region(src, L[1]->boxes, time, sb);
    }

    // MakeNewLevelFromScratch with the caller's initial data
    int add_level(const mf_box *boxes, int nb, const double *state) override {
        if (L.size() != 1 || nb < 1 || !state) return MF_ERR_UNSUPPORTED;
        int nf[3];
        for (int d = 0; d < 3; ++d) nf[d] = L[0]->n[d] * ratio;
        int st = check_boxes(boxes, nb, nf);
        if (st) return st;
        Level<P> *F = new_fine(boxes, nb);
        F->set_state(state);
        F->t_new = L[0]->t_new;
        F->t_old = -1.e100;
        L.push_back(F);
        nsub.push_back(ratio);
        dt.push_back(1.e100);
        istep.push_back(0);
        attach_coarse();
        build_registers();
        return MF_OK;
    }

    // regrid of level 1 at `time` with the new boxes: RemakeLevel (level 1
    // exists), MakeNewLevelFromCoarse (it does not), ClearLevel (no boxes)
    int regrid_level(const mf_box *boxes, int nb, double time) override {
        int nf[3];
        for (int d = 0; d < 3; ++d) nf[d] = L[0]->n[d] * ratio;
        int st = check_boxes(boxes, nb, nf);
        if (st) return st;
        if (nb == 0) {
            if (L.size() > 1) {
                delete L[1];
                L.resize(1);
                nsub.resize(1);
                dt.resize(1);
                istep.resize(1);
            }
            build_registers();
            return MF_OK;
        }
        const bool had = L.size() > 1;
        std::vector<Box> oldb;
        std::vector<double> olds;
        if (had) { oldb = L[1]->boxes; olds = L[1]->phi_new; }
        Level<P> *F = new_fine(boxes, nb);
        if (!had) {
            // InterpFromCoarseLevel: no fine data
            L.push_back(F);
            nsub.push_back(ratio);
            dt.push_back(dt[0] / ratio);
            istep.push_back(istep[0]);
        } else {
            F->istep = L[1]->istep;
            delete L[1];
            L[1] = F;
        }
        // FillPatch(1, time) over each new box from the old fine data and level 0
        std::vector<double> ns((size_t)NV * F->ncell(), 0.0);
        A4 NSv = F->valid(ns);
        for (const Box &B : F->boxes) {
            std::vector<double> bv((size_t)(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1) * (B.hi[2] - B.lo[2] + 1) * NV);
            A4 sb = make_a4(bv.data(), B.lo, B.hi, NV);
            fill_region(had ? olds : ns, oldb, time, sb);
            for (int n = 0; n < NV; ++n)
                for (int k = B.lo[2]; k <= B.hi[2]; ++k)
                    for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                        for (int i = B.lo[0]; i <= B.hi[0]; ++i) NSv(i, j, k, n) = sb(i, j, k, n);
        }
        F->phi_new = ns;
        F->phi_old = ns;  // never read: t_old is 1e200 before t_new
        F->t_new = time;
        F->t_old = time - 1.e200;
        attach_coarse();
        build_registers();
        return MF_OK;
    }

    // ComputeDt (Source/mflo_timestep.cpp:18-50) over the levels
    void compute_dt() {
        const mf_params &prm = L[0]->prm;
        const int nl = (int)L.size();
        std::vector<double> est(nl);
        for (int l = 0; l < nl; ++l) {
            if (prm.fixed_dt) est[l] = prm.dt_fixed;
            else L[l]->est_time_step(&est[l]);
        }
        const double change_max = 1.1;
        double dt_0 = est[0];
        int n_factor = 1;
        for (int l = 0; l < nl; ++l) {
            est[l] = std::min(est[l], change_max * dt[l]);
            n_factor *= nsub[l];
            dt_0 = std::min(dt_0, n_factor * est[l]);
        }
        if (prm.fixed_dt) dt_0 = prm.dt_fixed;
        const double eps = 1.e-3 * dt_0;
        if (L[0]->t_new + dt_0 > prm.stop_time - eps) dt_0 = prm.stop_time - L[0]->t_new;
        dt[0] = dt_0;
        for (int l = 1; l < nl; ++l) dt[l] = dt[l - 1] / nsub[l];
        for (int l = 0; l < nl; ++l) L[l]->dt = dt[l];
    }

    // timeStep (Source/mflo_evolve.cpp:175-249) without the regrid check
    int time_step(int lev, double time) {
        if (lev + 1 < (int)L.size()) std::fill(regv.begin(), regv.end(), 0.0);
        Level<P> &lv = *L[lev];
        lv.only_flow = lev == 0 ? only_flow : 0;  // the recursion passes no only_flow (:237)
        int st = lv.advance(time, dt[lev]);
        lv.only_flow = 0;
        if (st) return st;
        ++istep[lev];
        lv.istep = istep[lev];
        if (lev + 1 < (int)L.size()) {
            for (int i = 1; i <= nsub[lev + 1]; ++i) {
                st = time_step(lev + 1, time + (i - 1) * dt[lev + 1]);
                if (st) return st;
            }
            if (do_reflux) reflux();
            average_down();
        }
        return MF_OK;
    }

    // EvolveAMR's step body (Source/mflo_evolve.cpp:20-82), nsteps coarse
    // steps; regrid_first: a regrid of level 1 with `boxes` at the start of
    // the first step (after ComputeDt, as timeStep does it)
    int step(int nsteps, int regrid_first, const mf_box *boxes, int nb) override {
        const mf_params &prm = L[0]->prm;
        double cur_time = L[0]->t_new;
        for (int s = 0; s < nsteps && cur_time < prm.stop_time; ++s) {
            compute_dt();
            if (s == 0 && regrid_first) {
                int st = regrid_level(boxes, nb, cur_time);
                if (st) return st;
            }
            int st = time_step(0, cur_time);
            if (st) return st;
            for (size_t l = 0; l < L.size(); ++l) {
                st = L[l]->update_primitive_vars();
                if (st) return st;
            }
            if (prm.clip_species)
                for (size_t l = 0; l < L.size(); ++l) L[l]->clip_neg_speciesconc();
            if (prm.track_residual_norms) L[0]->compute_residual_norms();
            cur_time += dt[0];
            for (size_t l = 0; l < L.size(); ++l) L[l]->t_new = cur_time;
            if (cur_time >= prm.stop_time - 1.e-6 * dt[0]) break;
        }
        return MF_OK;
    }
};
