// mf_io.hip -- plotfile and checkpoint files of a level (host code).
//
// Layouts of the reference's output (Source/mflo_plot.cpp):
//   plotfile (WritePlotFile -> amrex::WriteMultiLevelPlotfile, :45-56):
//     <dir>/Header            "HyperCLaw-V1.1", variable names, time, domain,
//                             level steps, cell sizes, one line of physical
//                             bounds per grid and dimension, "Level_0/Cell"
//     <dir>/Level_0/Cell_H    VisMF header: version, how, ncomp, nghost, the
//                             BoxArray, one "FabOnDisk: <file> <offset>" per
//                             box, per-box minima and maxima
//     <dir>/Level_0/Cell_D_00000  one FAB per box: a text header
//                             "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))"
//                             "((lo) (hi) (0,0,0)) ncomp" and the box's doubles,
//                             component-major, x fastest (little-endian IEEE)
//   checkpoint (WriteCheckpointFile / ReadCheckpointFile, :58-239):
//     <dir>/Header            "Checkpoint file for mflo", finest_level, istep,
//                             dt, t_new (precision 17), the BoxArray
//     <dir>/Level_0/<var0>_H, <var0>_D_00000   the state as a VisMF file
// AMReX itself is not available here: these files follow the formats AMReX
// documents; reading them back (and restarting bit for bit) is what the tests
// check.
#include <sys/stat.h>

#include <array>
#include <cerrno>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "mesoflow_amd.h"
#include "mf_io.h"

int mf_level_dt_prev(void *level, double *dt_prev);  // mf_capi.hip

namespace {

struct LevelInfo {
    int n[3];
    int nv;
    std::vector<mf_box> boxes;
    mf_geom geom;
};

bool make_dir(const std::string &d) {
    if (mkdir(d.c_str(), 0755) == 0) return true;
    return errno == EEXIST;
}

std::string box_str(const mf_box &b) {
    std::ostringstream os;
    os << "((" << b.lo[0] << ',' << b.lo[1] << ',' << b.lo[2] << ") (" << b.hi[0] << ',' << b.hi[1] << ',' << b.hi[2]
       << ") (0,0,0))";
    return os.str();
}

// BoxArray::writeOn
void write_boxarray(std::ostream &os, const std::vector<mf_box> &bx) {
    os << '(' << bx.size() << ' ' << 0 << '\n';
    for (const mf_box &b : bx) os << box_str(b) << '\n';
    os << ')';
}

bool parse_box(const std::string &s, mf_box &b) {
    int t0, t1, t2;
    return std::sscanf(s.c_str(), " ((%d,%d,%d) (%d,%d,%d) (%d,%d,%d))", &b.lo[0], &b.lo[1], &b.lo[2], &b.hi[0], &b.hi[1],
                       &b.hi[2], &t0, &t1, &t2) == 9;
}

bool read_boxarray(std::istream &is, std::vector<mf_box> &bx) {
    std::string line;
    if (!std::getline(is, line)) return false;
    size_t n = 0;
    int hash = 0;
    if (std::sscanf(line.c_str(), " (%zu %d", &n, &hash) != 2) return false;
    bx.resize(n);
    for (size_t q = 0; q < n; ++q) {
        if (!std::getline(is, line) || !parse_box(line, bx[q])) return false;
    }
    return static_cast<bool>(std::getline(is, line));  // ")"
}

int64_t cells_of(const mf_box &b) {
    return (int64_t)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
}

const char *kFabHeader = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))";

// VisMF::Write of a level state [nv][nz][ny][nx]: <prefix>_H and <prefix>_D_00000
bool write_vismf(const std::string &prefix, const std::string &dname, const LevelInfo &L, const double *state) {
    const int nx = L.n[0], ny = L.n[1];
    const int64_t ncell = (int64_t)L.n[0] * L.n[1] * L.n[2];
    std::ofstream df(prefix + "_D_00000", std::ios::binary | std::ios::trunc);
    if (!df) return false;
    std::vector<int64_t> offs;
    std::vector<std::vector<double>> mn, mx;
    std::vector<double> fab;
    for (const mf_box &b : L.boxes) {
        offs.push_back((int64_t)df.tellp());
        df << kFabHeader << box_str(b) << ' ' << L.nv << '\n';
        const int64_t nb = cells_of(b);
        fab.resize((size_t)nb * L.nv);
        std::vector<double> lo(L.nv, std::numeric_limits<double>::max()), hi(L.nv, -std::numeric_limits<double>::max());
        size_t w = 0;
        for (int c = 0; c < L.nv; ++c)
            for (int k = b.lo[2]; k <= b.hi[2]; ++k)
                for (int j = b.lo[1]; j <= b.hi[1]; ++j)
                    for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
                        const double v = state[(size_t)c * ncell + i + (size_t)nx * (j + (size_t)ny * k)];
                        fab[w++] = v;
                        if (v < lo[c]) lo[c] = v;
                        if (v > hi[c]) hi[c] = v;
                    }
        df.write(reinterpret_cast<const char *>(fab.data()), (std::streamsize)(fab.size() * sizeof(double)));
        mn.push_back(lo);
        mx.push_back(hi);
    }
    if (!df) return false;
    df.close();
    std::ofstream hf(prefix + "_H", std::ios::trunc);
    if (!hf) return false;
    hf.setf(std::ios::scientific, std::ios::floatfield);
    hf << std::setprecision(16);  // 17 significant digits: the extrema read back exactly
    hf << 1 << '\n' << 0 << '\n' << L.nv << '\n' << 0 << '\n';
    write_boxarray(hf, L.boxes);
    hf << '\n';
    hf << L.boxes.size() << '\n';
    for (size_t q = 0; q < offs.size(); ++q) hf << "FabOnDisk: " << dname << "_D_00000 " << offs[q] << '\n';
    hf << '\n';
    for (auto *mm : {&mn, &mx}) {
        hf << mm->size() << ',' << L.nv << '\n';
        for (const auto &row : *mm) {
            for (double v : row) hf << v << ',';
            hf << '\n';
        }
        hf << '\n';
    }
    return static_cast<bool>(hf);
}

// VisMF::Read into a state [nv][nz][ny][nx] for the boxes of L
bool read_vismf(const std::string &dir, const std::string &prefix, const LevelInfo &L, double *state) {
    std::ifstream hf(prefix + "_H");
    if (!hf) return false;
    int vers, how, ncomp, ngrow;
    hf >> vers >> how >> ncomp >> ngrow;
    std::string line;
    std::getline(hf, line);
    if (ncomp != L.nv) return false;
    std::vector<mf_box> bx;
    if (!read_boxarray(hf, bx) || bx.size() != L.boxes.size()) return false;
    size_t nfab = 0;
    hf >> nfab;
    if (nfab != bx.size()) return false;
    const int nx = L.n[0], ny = L.n[1];
    const int64_t ncell = (int64_t)L.n[0] * L.n[1] * L.n[2];
    std::vector<double> fab;
    for (size_t q = 0; q < nfab; ++q) {
        std::string tag, fname;
        int64_t off;
        hf >> tag >> fname >> off;
        if (tag != "FabOnDisk:") return false;
        std::ifstream df(dir + "/" + fname, std::ios::binary);
        if (!df) return false;
        df.seekg(off);
        std::string head;
        std::getline(df, head);
        if (head.compare(0, std::strlen(kFabHeader), kFabHeader) != 0) return false;
        mf_box b;
        int nc = 0;
        const std::string rest = head.substr(std::strlen(kFabHeader));
        const size_t cut = rest.rfind(' ');
        if (cut == std::string::npos || !parse_box(rest.substr(0, cut), b)) return false;
        nc = std::atoi(rest.substr(cut + 1).c_str());
        if (nc != L.nv || std::memcmp(&b, &bx[q], sizeof(mf_box)) != 0) return false;
        const int64_t nb = cells_of(b);
        fab.resize((size_t)nb * nc);
        df.read(reinterpret_cast<char *>(fab.data()), (std::streamsize)(fab.size() * sizeof(double)));
        if (!df) return false;
        size_t r = 0;
        for (int c = 0; c < nc; ++c)
            for (int k = b.lo[2]; k <= b.hi[2]; ++k)
                for (int j = b.lo[1]; j <= b.hi[1]; ++j)
                    for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                        state[(size_t)c * ncell + i + (size_t)nx * (j + (size_t)ny * k)] = fab[r++];
    }
    return true;
}

const char *kBaseNames[MF_FLO_NVARS] = {"rho",   "rhou", "rhov", "rhow",     "rhoe",        "density",
                                        "velx",  "vely", "velz", "pressure", "temperature", "volfrac"};

std::string var_name(const char *const *names, int c) {
    if (names && names[c]) return names[c];
    if (c < MF_FLO_NVARS) return kBaseNames[c];
    return "S" + std::to_string(c - MF_FLO_NVARS);
}

int level_info(void *level, LevelInfo &L) {
    int nb = 0;
    int st = mf_level_boxes(level, nullptr, 0, &nb, &L.geom, &L.nv);
    if (st) return st;
    L.boxes.resize(nb);
    st = mf_level_boxes(level, L.boxes.data(), nb, &nb, &L.geom, &L.nv);
    for (int d = 0; d < 3; ++d) L.n[d] = L.geom.n_cell[d];
    return st;
}

// a host description of a level: boxes inside the domain, 1..(12 + species) components
int host_info(const mf_geom *geom, const mf_box *boxes, int nboxes, int nv, LevelInfo &L) {
    if (!geom || !boxes || nboxes < 1 || nv < 1 || nv > MF_FLO_NVARS + MF_MAX_SPECIES) return MF_ERR_BAD_ARG;
    L.geom = *geom;
    L.nv = nv;
    for (int d = 0; d < 3; ++d) {
        L.n[d] = geom->n_cell[d];
        if (L.n[d] < 1) return MF_ERR_BAD_ARG;
    }
    L.boxes.assign(boxes, boxes + nboxes);
    for (const mf_box &b : L.boxes)
        for (int d = 0; d < 3; ++d)
            if (b.lo[d] < 0 || b.hi[d] < b.lo[d] || b.hi[d] >= L.n[d]) return MF_ERR_BAD_ARG;
    return MF_OK;
}

}  // namespace

namespace mfio {

// amrex::WriteMultiLevelPlotfile (Source/mflo_plot.cpp:45-56): the Header of
// WriteGenericPlotfileHeader and one Level_<l>/Cell MultiFab per level; time
// is t_new of level 0 for every level, isteps the levels' step counts
namespace {
LevelInfo info_of(const LevelData &D) {
    LevelInfo L;
    L.geom = D.geom;
    L.nv = D.nv;
    L.boxes = D.boxes;
    for (int d = 0; d < 3; ++d) L.n[d] = D.geom.n_cell[d];
    return L;
}
}  // namespace

int write_plotfile_levels(const std::vector<LevelData> &levels, double t, const std::vector<int> &isteps, int ratio,
                          const char *dir, const char *const *varnames) {
    const std::string d(dir);
    const int nl = (int)levels.size();
    if (nl < 1) return MF_ERR_BAD_ARG;
    std::vector<LevelInfo> Ls;
    std::vector<const double *> states;
    for (const LevelData &D : levels) { Ls.push_back(info_of(D)); states.push_back(D.state); }
    if (!make_dir(d)) return MF_ERR_BAD_ARG;
    for (int l = 0; l < nl; ++l) {
        const std::string ld = d + "/Level_" + std::to_string(l);
        if (!make_dir(ld) || !write_vismf(ld + "/Cell", "Cell", Ls[l], states[l])) return MF_ERR_BAD_ARG;
    }
    std::ofstream h(d + "/Header", std::ios::trunc);
    if (!h) return MF_ERR_BAD_ARG;
    h << std::setprecision(17);
    const LevelInfo &L0 = Ls[0];
    const mf_geom &g = L0.geom;
    h << "HyperCLaw-V1.1\n" << L0.nv << '\n';
    for (int c = 0; c < L0.nv; ++c) h << var_name(varnames, c) << '\n';
    h << 3 << '\n' << t << '\n' << nl - 1 << '\n';
    for (int q = 0; q < 3; ++q) h << g.prob_lo[q] << ' ';
    h << '\n';
    for (int q = 0; q < 3; ++q) h << g.prob_hi[q] << ' ';
    h << '\n';
    for (int l = 0; l < nl - 1; ++l) h << ratio << ' ';  // refinement ratio below each level
    h << '\n';
    for (int l = 0; l < nl; ++l) {
        mf_box dom;
        for (int q = 0; q < 3; ++q) { dom.lo[q] = 0; dom.hi[q] = Ls[l].geom.n_cell[q] - 1; }
        h << box_str(dom) << ' ';
    }
    h << '\n';
    for (int l = 0; l < nl; ++l) h << isteps[l] << ' ';
    h << '\n';
    std::vector<std::array<double, 3>> dx(nl);
    for (int l = 0; l < nl; ++l) {
        const mf_geom &gl = Ls[l].geom;
        for (int q = 0; q < 3; ++q) {
            dx[l][q] = (gl.prob_hi[q] - gl.prob_lo[q]) / gl.n_cell[q];
            h << dx[l][q] << ' ';
        }
        h << '\n';
    }
    h << 0 << '\n' << 0 << '\n';  // coordinate system (Cartesian), boundary width
    for (int l = 0; l < nl; ++l) {
        h << l << ' ' << Ls[l].boxes.size() << ' ' << t << '\n' << isteps[l] << '\n';
        for (const mf_box &b : Ls[l].boxes)
            for (int q = 0; q < 3; ++q)
                h << g.prob_lo[q] + b.lo[q] * dx[l][q] << ' ' << g.prob_lo[q] + (b.hi[q] + 1) * dx[l][q] << '\n';
        h << "Level_" << l << "/Cell\n";
    }
    return h ? MF_OK : MF_ERR_BAD_ARG;
}

// WriteCheckpointFile (Source/mflo_plot.cpp:57-145): finest_level, istep, dt
// and t_new of every possible level (nlevs_max entries each), the BoxArray and
// the state (Level_<l>/<var0>) of every existing level
int write_checkpoint_levels(const std::vector<LevelData> &levels, const std::vector<double> &t_new,
                            const std::vector<double> &dt, const std::vector<int> &isteps, const char *dir,
                            const char *const *varnames) {
    const std::string d(dir);
    const int nl = (int)levels.size();
    if (nl < 1) return MF_ERR_BAD_ARG;
    std::vector<LevelInfo> Ls;
    std::vector<const double *> states;
    for (const LevelData &D : levels) { Ls.push_back(info_of(D)); states.push_back(D.state); }
    if (!make_dir(d)) return MF_ERR_BAD_ARG;
    const std::string v0 = var_name(varnames, 0);
    for (int l = 0; l < nl; ++l) {
        const std::string ld = d + "/Level_" + std::to_string(l);
        if (!make_dir(ld) || !write_vismf(ld + "/" + v0, v0, Ls[l], states[l])) return MF_ERR_BAD_ARG;
    }
    std::ofstream h(d + "/Header", std::ios::trunc | std::ios::binary);
    if (!h) return MF_ERR_BAD_ARG;
    h << std::setprecision(17);
    h << "Checkpoint file for mflo\n" << nl - 1 << '\n';
    for (int v : isteps) h << v << ' ';
    h << '\n';
    for (double v : dt) h << v << ' ';
    h << '\n';
    for (double v : t_new) h << v << ' ';
    h << '\n';
    for (int l = 0; l < nl; ++l) {
        write_boxarray(h, Ls[l].boxes);
        h << '\n';
    }
    return h ? MF_OK : MF_ERR_BAD_ARG;
}

// the Header of ReadCheckpointFile (:147-239): finest_level, the arrays and
// the BoxArray of each level
int read_checkpoint_header(const char *dir, int max_finest, int *finest, std::vector<int> &isteps,
                           std::vector<double> &dt, std::vector<double> &t_new, std::vector<std::vector<mf_box>> &boxes) {
    const std::string d(dir);
    std::ifstream h(d + "/Header");
    if (!h) return MF_ERR_BAD_ARG;
    std::string line, word;
    std::getline(h, line);  // title
    if (line != "Checkpoint file for mflo") return MF_ERR_BAD_ARG;
    *finest = -1;
    h >> *finest;
    std::getline(h, line);
    if (*finest < 0) return MF_ERR_BAD_ARG;
    if (*finest > max_finest) return MF_ERR_UNSUPPORTED;
    isteps.clear();
    dt.clear();
    t_new.clear();
    std::getline(h, line);
    {
        std::istringstream ls(line);
        while (ls >> word) isteps.push_back(std::atoi(word.c_str()));
    }
    std::getline(h, line);
    {
        std::istringstream ls(line);
        while (ls >> word) dt.push_back(std::strtod(word.c_str(), nullptr));
    }
    std::getline(h, line);
    {
        std::istringstream ls(line);
        while (ls >> word) t_new.push_back(std::strtod(word.c_str(), nullptr));
    }
    if ((int)isteps.size() <= *finest || (int)dt.size() <= *finest || (int)t_new.size() <= *finest) return MF_ERR_BAD_ARG;
    boxes.assign((size_t)*finest + 1, {});
    for (int l = 0; l <= *finest; ++l)
        if (!read_boxarray(h, boxes[(size_t)l])) return MF_ERR_BAD_ARG;
    return MF_OK;
}

// the state of level l of a checkpoint into a dense array over L's domain
int read_checkpoint_level(const char *dir, int l, const mf_geom &geom, const std::vector<mf_box> &boxes, int nv,
                          double *state, const char *const *varnames) {
    LevelData D{geom, boxes, nv, nullptr};
    const LevelInfo L = info_of(D);
    const std::string ld = std::string(dir) + "/Level_" + std::to_string(l);
    return read_vismf(ld, ld + "/" + var_name(varnames, 0), L, state) ? MF_OK : MF_ERR_BAD_ARG;
}

}  // namespace mfio

namespace {

mfio::LevelData data_of(const LevelInfo &L, const double *state) { return mfio::LevelData{L.geom, L.boxes, L.nv, state}; }

int write_plotfile(const LevelInfo &L, const double *state, double t, int istep, const char *dir,
                   const char *const *varnames) {
    return mfio::write_plotfile_levels({data_of(L, state)}, t, {istep}, 1, dir, varnames);
}

int write_checkpoint(const LevelInfo &L, const double *state, double t, double dt, int istep, const char *dir,
                     const char *const *varnames) {
    return mfio::write_checkpoint_levels({data_of(L, state)}, {t}, {dt}, {istep}, dir, varnames);
}

int read_checkpoint(const LevelInfo &L, double *state, double *t, double *dt, int *istep, const char *dir,
                    const char *const *varnames) {
    int finest = -1;
    std::vector<int> is;
    std::vector<double> dts, tn;
    std::vector<std::vector<mf_box>> bx;
    int st = mfio::read_checkpoint_header(dir, 0, &finest, is, dts, tn, bx);
    if (st) return st;
    if (finest != 0) return MF_ERR_UNSUPPORTED;
    *istep = is[0];
    *dt = dts[0];
    *t = tn[0];
    if (bx[0].size() != L.boxes.size() || std::memcmp(bx[0].data(), L.boxes.data(), bx[0].size() * sizeof(mf_box)) != 0)
        return MF_ERR_BAD_ARG;
    return mfio::read_checkpoint_level(dir, 0, L.geom, L.boxes, L.nv, state, varnames);
}

}  // namespace

extern "C" {

int mf_write_plotfile(const char *dir, const mf_geom *geom, const mf_box *boxes, int nboxes, int nv,
                      const double *state, double time, int istep, const char *const *varnames) {
    LevelInfo L;
    if (!dir || !state) return MF_ERR_BAD_ARG;
    int st = host_info(geom, boxes, nboxes, nv, L);
    if (st) return st;
    return write_plotfile(L, state, time, istep, dir, varnames);
}

int mf_write_checkpoint(const char *dir, const mf_geom *geom, const mf_box *boxes, int nboxes, int nv,
                        const double *state, double time, double dt, int istep, const char *const *varnames) {
    LevelInfo L;
    if (!dir || !state) return MF_ERR_BAD_ARG;
    int st = host_info(geom, boxes, nboxes, nv, L);
    if (st) return st;
    return write_checkpoint(L, state, time, dt, istep, dir, varnames);
}

int mf_read_checkpoint(const char *dir, const mf_geom *geom, const mf_box *boxes, int nboxes, int nv, double *state,
                       double *time, double *dt, int *istep, const char *const *varnames) {
    LevelInfo L;
    if (!dir || !state || !time || !dt || !istep) return MF_ERR_BAD_ARG;
    int st = host_info(geom, boxes, nboxes, nv, L);
    if (st) return st;
    return read_checkpoint(L, state, time, dt, istep, dir, varnames);
}

int mf_level_write_plotfile(void *level, const char *dir, const char *const *varnames) {
    if (!level || !dir) return MF_ERR_BAD_ARG;
    LevelInfo L;
    int st = level_info(level, L);
    if (st) return st;
    double t = 0.0;
    int istep = 0;
    st = mf_level_info(level, &t, nullptr, &istep, nullptr);
    if (st) return st;
    std::vector<double> state((size_t)L.nv * L.n[0] * L.n[1] * L.n[2]);
    st = mf_level_get_state(level, state.data(), 0);
    if (st) return st;
    return write_plotfile(L, state.data(), t, istep, dir, varnames);
}

int mf_level_write_checkpoint(void *level, const char *dir, const char *const *varnames) {
    if (!level || !dir) return MF_ERR_BAD_ARG;
    LevelInfo L;
    int st = level_info(level, L);
    if (st) return st;
    double t = 0.0, dt = 0.0;
    int istep = 0;
    st = mf_level_info(level, &t, nullptr, &istep, nullptr);
    if (st) return st;
    st = mf_level_dt_prev(level, &dt);
    if (st) return st;
    std::vector<double> state((size_t)L.nv * L.n[0] * L.n[1] * L.n[2]);
    st = mf_level_get_state(level, state.data(), 0);
    if (st) return st;
    return write_checkpoint(L, state.data(), t, dt, istep, dir, varnames);
}

int mf_level_read_checkpoint(void *level, const char *dir, const char *const *varnames) {
    if (!level || !dir) return MF_ERR_BAD_ARG;
    LevelInfo L;
    int st = level_info(level, L);
    if (st) return st;
    // cells outside the boxes keep the level's current values
    std::vector<double> state((size_t)L.nv * L.n[0] * L.n[1] * L.n[2]);
    st = mf_level_get_state(level, state.data(), 0);
    if (st) return st;
    double t = 0.0, dt = 0.0;
    int istep = 0;
    st = read_checkpoint(L, state.data(), &t, &dt, &istep, dir, varnames);
    if (st) return st;
    st = mf_level_set_state(level, state.data(), 0);
    if (st) return st;
    return mf_level_set_time(level, t, dt, istep);
}

}  // extern "C"
