// mf_regrid.hip -- new fine-level boxes from refinement tags (host code).
//
// The regrid step of mesoflow's hierarchy (Source/mflo_evolve.cpp:177-209 ->
// AmrMesh::regrid -> MakeNewGrids, driven by the tags of ErrorEst,
// Source/mflo.cpp:183-260) turns tagged coarse cells into the BoxArray of the
// next finer level with the amr.* parameters n_error_buf, blocking_factor,
// max_grid_size and grid_eff. AMReX is not in the tree; the steps restate its
// documented algorithm:
//   1. every tag grows by n_error_buf cells, wrapping around the domain in
//      periodic directions (AMReX maps buffered tags across periodic
//      boundaries) and clipped to it in the others;
//   2. tags are coarsened by bf = blocking_factor / ratio (a block is tagged if
//      any of its cells is), so every box edge lands on the blocking factor;
//   3. Berger-Rigoutsos clustering of the coarsened tags: a cluster is shrunk
//      to the bounding box of its tags and accepted when tags/volume >=
//      grid_eff; otherwise it is cut at the hole (an empty plane of the tag
//      signature) nearest its centre, else at the strongest inflection of the
//      signature's second difference, else in half along its longest side;
//   4. the accepted boxes are refined by bf and by ratio and chopped to at most
//      max_grid_size cells per edge in pieces that are multiples of
//      blocking_factor.
// The result covers every buffered tag, its boxes are disjoint and aligned.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "mesoflow_amd.h"

namespace {

struct IBox {
    int lo[3], hi[3];
    int64_t vol() const { return (int64_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1); }
};

struct Grid {  // tags on the coarsened index space
    int n[3];
    std::vector<unsigned char> t;
    bool at(int i, int j, int k) const { return t[(size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k)] != 0; }
};

// tags of b; the bounding box of its tags (false if none)
bool shrink(const Grid &g, const IBox &b, IBox &out, int64_t &ntag) {
    ntag = 0;
    IBox m{{b.hi[0] + 1, b.hi[1] + 1, b.hi[2] + 1}, {b.lo[0] - 1, b.lo[1] - 1, b.lo[2] - 1}};
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                if (g.at(i, j, k)) {
                    ++ntag;
                    const int c[3] = {i, j, k};
                    for (int d = 0; d < 3; ++d) {
                        m.lo[d] = std::min(m.lo[d], c[d]);
                        m.hi[d] = std::max(m.hi[d], c[d]);
                    }
                }
    out = m;
    return ntag > 0;
}

void cluster(const Grid &g, const IBox &b0, double eff, std::vector<IBox> &acc) {
    IBox b;
    int64_t ntag;
    if (!shrink(g, b0, b, ntag)) return;
    if ((double)ntag >= eff * (double)b.vol() || b.vol() == 1) {
        acc.push_back(b);
        return;
    }
    // signatures
    std::vector<int64_t> sig[3];
    for (int d = 0; d < 3; ++d) sig[d].assign(b.hi[d] - b.lo[d] + 1, 0);
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                if (g.at(i, j, k)) {
                    sig[0][i - b.lo[0]]++;
                    sig[1][j - b.lo[1]]++;
                    sig[2][k - b.lo[2]]++;
                }
    int cut_dir = -1, cut_at = 0;  // cut: left = [lo, cut_at - 1], right = [cut_at, hi]
    // 1. the hole nearest the centre
    int best = 1 << 30;
    for (int d = 0; d < 3; ++d) {
        const int len = b.hi[d] - b.lo[d] + 1;
        for (int x = 1; x + 1 < len; ++x)
            if (sig[d][x] == 0) {
                const int dist = std::abs(2 * x - (len - 1));
                if (dist < best) { best = dist; cut_dir = d; cut_at = b.lo[d] + x; }
            }
    }
    // 2. the strongest sign change of the second difference
    if (cut_dir < 0) {
        int64_t jump = 0;
        for (int d = 0; d < 3; ++d) {
            const int len = b.hi[d] - b.lo[d] + 1;
            if (len < 4) continue;
            std::vector<int64_t> lap(len, 0);
            for (int x = 1; x + 1 < len; ++x) lap[x] = sig[d][x - 1] - 2 * sig[d][x] + sig[d][x + 1];
            for (int x = 2; x + 1 < len; ++x)
                if ((lap[x - 1] < 0 && lap[x] > 0) || (lap[x - 1] > 0 && lap[x] < 0)) {
                    const int64_t j = std::llabs(lap[x - 1] - lap[x]);
                    if (j > jump) { jump = j; cut_dir = d; cut_at = b.lo[d] + x; }
                }
        }
    }
    // 3. halve the longest side
    if (cut_dir < 0) {
        int len = 0;
        for (int d = 0; d < 3; ++d)
            if (b.hi[d] - b.lo[d] + 1 > len) { len = b.hi[d] - b.lo[d] + 1; cut_dir = d; }
        cut_at = b.lo[cut_dir] + len / 2;
    }
    IBox l = b, r = b;
    l.hi[cut_dir] = cut_at - 1;
    r.lo[cut_dir] = cut_at;
    if (l.hi[cut_dir] >= l.lo[cut_dir]) cluster(g, l, eff, acc);
    if (r.hi[cut_dir] >= r.lo[cut_dir]) cluster(g, r, eff, acc);
}

}  // namespace

extern "C" {

int mf_tags_to_boxes(const unsigned char *tags, const int32_t n_cell[3], const int32_t *is_periodic, int n_error_buf,
                     int blocking_factor, int max_grid_size, double grid_eff, int ratio, mf_box *boxes, int cap,
                     int *nboxes) {
    if (!tags || !n_cell || !nboxes || n_error_buf < 0 || ratio < 1 || blocking_factor < ratio ||
        blocking_factor % ratio != 0 || max_grid_size < blocking_factor || max_grid_size % blocking_factor != 0 ||
        !(grid_eff > 0.0 && grid_eff <= 1.0) || cap < 0 || (cap > 0 && !boxes))
        return MF_ERR_BAD_ARG;
    const int bf = blocking_factor / ratio;  // blocking factor on the tagged level
    int n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = n_cell[d];
        if (n[d] < 1 || n[d] % bf != 0) return MF_ERR_BAD_ARG;
    }
    // 1 + 2: buffered tags, coarsened by bf
    Grid g;
    for (int d = 0; d < 3; ++d) g.n[d] = n[d] / bf;
    g.t.assign((size_t)g.n[0] * g.n[1] * g.n[2], 0);
    const int e = n_error_buf;
    int per[3];
    for (int d = 0; d < 3; ++d) per[d] = is_periodic ? (is_periodic[d] != 0) : 0;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                if (!tags[(size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k)]) continue;
                const int c[3] = {i, j, k};
                int lo[3], hi[3];
                for (int d = 0; d < 3; ++d) {
                    lo[d] = c[d] - e;
                    hi[d] = c[d] + e;
                    if (!per[d] || hi[d] - lo[d] + 1 >= n[d]) {  // clip (a buffer wider than a periodic domain covers it)
                        lo[d] = std::max(lo[d], 0);
                        hi[d] = std::min(hi[d], n[d] - 1);
                    }
                }
                // buffered cells, each mapped to its periodic image in the domain
                for (int kk = lo[2]; kk <= hi[2]; ++kk) {
                    const int kw = ((kk % n[2]) + n[2]) % n[2];
                    for (int jj = lo[1]; jj <= hi[1]; ++jj) {
                        const int jw = ((jj % n[1]) + n[1]) % n[1];
                        for (int ii = lo[0]; ii <= hi[0]; ++ii) {
                            const int iw = ((ii % n[0]) + n[0]) % n[0];
                            g.t[(size_t)(iw / bf) + (size_t)g.n[0] * (jw / bf + (size_t)g.n[1] * (kw / bf))] = 1;
                        }
                    }
                }
            }
    // 3: clusters
    std::vector<IBox> cl;
    cluster(g, IBox{{0, 0, 0}, {g.n[0] - 1, g.n[1] - 1, g.n[2] - 1}}, grid_eff, cl);
    // 4: refine to the fine level, chop to max_grid_size in multiples of blocking_factor
    const int rf = bf * ratio;
    std::vector<mf_box> out;
    for (const IBox &c : cl) {
        int flo[3], len[3], npc[3];
        for (int d = 0; d < 3; ++d) {
            flo[d] = c.lo[d] * rf;
            len[d] = (c.hi[d] - c.lo[d] + 1) * rf;
            npc[d] = (len[d] + max_grid_size - 1) / max_grid_size;
        }
        for (int pk = 0; pk < npc[2]; ++pk)
            for (int pj = 0; pj < npc[1]; ++pj)
                for (int pi = 0; pi < npc[0]; ++pi) {
                    const int p[3] = {pi, pj, pk};
                    mf_box b;
                    for (int d = 0; d < 3; ++d) {
                        // piece q of np: block range [q*nb/np, (q+1)*nb/np) of nb blocking-factor blocks
                        const int nb = len[d] / blocking_factor;
                        const int b0 = (int)((int64_t)p[d] * nb / npc[d]);
                        const int b1 = (int)((int64_t)(p[d] + 1) * nb / npc[d]);
                        b.lo[d] = flo[d] + b0 * blocking_factor;
                        b.hi[d] = flo[d] + b1 * blocking_factor - 1;
                    }
                    out.push_back(b);
                }
    }
    *nboxes = (int)out.size();
    if ((int)out.size() > cap) return boxes ? MF_ERR_BAD_ARG : MF_OK;
    std::copy(out.begin(), out.end(), boxes);
    return MF_OK;
}

}  // extern "C"
