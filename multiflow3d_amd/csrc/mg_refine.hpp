// mg_refine.hpp -- ghost tables of refined block grids (host side).
//
// On a grid whose blocks differ in rdiv (localparameters.for:26-56), a
// level's ghosts are no longer plain copies: across a face whose two blocks
// differ in rdv at that level, exchangepp sends the coarse side 0.125 x the
// 8 fine cells its ghost covers (exchangepp.for:163-172, 1110-1120) and the
// fine side 2 x the coarse cell over each of its ghosts, which the fine side
// stores as (received + 3 phi_1 - phi_2) / 4 (:173-188, 4046-4053,
// 4257-4262).  The coarse side keeps the interior rows of that ghost face;
// the reference's fine side also rewrites the face's edge rows from receive
// buffer entries no sender filled (left over from earlier exchanges): here
// those rows keep what they held, and a level whose ghosts would end up
// holding anything but a function of the level's interior values is
// refused.
//
// build_ghost_exprs runs the reference's whole sequence symbolically --
// face passes x, y, z (each: every block packs, then every block unpacks),
// mgbound (faces W, E, S, N, B, T), exchbc_mgpp -- on value names: an
// interior cell names itself, an 8-cell average or an extrapolation names a
// new expression, a copy copies the name.  Every ghost ends as one
// GhostExpr over interior cells, so one gather launch (k_gather_expr)
// applies the whole exchange.
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include "mf3d_mg.h"
#include "mg_common.hpp"

namespace mf3d {

// Per-block level data of the symbolic run.
struct SymBlock {
    int n[3];
    long long off;     // arena offset
    int rdv;           // rdv(dom, glevel) at this level
    const mf3d_block_desc *desc;
};

class GhostSym {
  public:
    explicit GhostSym(const std::vector<SymBlock> &b) : blk_(b), val_(b.size()), buf_(2 * b.size()) {
        for (size_t i = 0; i < b.size(); ++i) {
            const SymBlock &s = b[i];
            val_[i].assign((size_t)(s.n[0] + 2) * (s.n[1] + 2) * (s.n[2] + 2), kPre);
            for (int k = 1; k <= s.n[2]; ++k)
                for (int j = 1; j <= s.n[1]; ++j)
                    for (int x = 1; x <= s.n[0]; ++x) val_[i][gi(i, x, j, k)] = leaf(s.off + (long long)gi(i, x, j, k));
            const size_t face = std::max({(size_t)(s.n[0] + 2) * (s.n[1] + 2), (size_t)(s.n[0] + 2) * (s.n[2] + 2),
                                          (size_t)(s.n[1] + 2) * (s.n[2] + 2)});
            buf_[2 * i].assign(face, kPre);
            buf_[2 * i + 1].assign(face, kPre);
        }
    }

    // exchangepp (face passes), mgbound, exchbc_mgpp
    void run() {
        for (int d = 0; d < 3; ++d) {
            for (int b = 0; b < (int)blk_.size(); ++b) {
                const mf3d_block_desc &B = *blk_[b].desc;
                if (B.nbr[2 * d] >= 0) pack(b, B.nbr[2 * d], d, 0, buf_[2 * B.nbr[2 * d] + 1]);
                if (B.nbr[2 * d + 1] >= 0) pack(b, B.nbr[2 * d + 1], d, 1, buf_[2 * B.nbr[2 * d + 1]]);
            }
            for (int b = 0; b < (int)blk_.size(); ++b) {
                const mf3d_block_desc &B = *blk_[b].desc;
                if (B.nbr[2 * d] >= 0) unpack(b, B.nbr[2 * d], d, 0, buf_[2 * b]);
                if (B.nbr[2 * d + 1] >= 0) unpack(b, B.nbr[2 * d + 1], d, 1, buf_[2 * b + 1]);
            }
        }
        for (int b = 0; b < (int)blk_.size(); ++b) {  // mgbound (mgsolver.for:907-992)
            const mf3d_block_desc &B = *blk_[b].desc;
            for (int d = 0; d < 3; ++d) {
                if (B.nbr[2 * d] < 0 && B.bc[2 * d] != 5) plane_copy(b, d, 0, 1);
                if (B.nbr[2 * d + 1] < 0 && B.bc[2 * d + 1] != 5) plane_copy(b, d, blk_[b].n[d] + 1, blk_[b].n[d]);
            }
        }
        bool periodic = false;
        for (const SymBlock &s : blk_)
            for (int f = 0; f < 6; ++f) periodic = periodic || s.desc->bc[f] == 5;
        if (!periodic) return;
        for (int d = 0; d < 3; ++d) {  // exchbc_mgpp (exchangepp.for:4768-5383)
            for (int b = 0; b < (int)blk_.size(); ++b) {
                const mf3d_block_desc &B = *blk_[b].desc;
                if (!(B.bc[2 * d] == 5 || B.bc[2 * d + 1] == 5)) continue;
                const int prv = B.nbr[2 * d], nxt = B.nbr[2 * d + 1], n = blk_[b].n[d];
                if (prv < 0) {
                    if (nxt < 0) plane_copy(b, d, 0, n);
                    else if (B.per[2 * d] >= 0) plane_get(b, d, 1, buf_[2 * B.per[2 * d] + 1]);
                }
                if (nxt < 0) {
                    if (prv < 0) plane_copy(b, d, n + 1, 1);
                    else if (B.per[2 * d + 1] >= 0) plane_get(b, d, n, buf_[2 * B.per[2 * d + 1]]);
                }
            }
            for (int b = 0; b < (int)blk_.size(); ++b) {
                const mf3d_block_desc &B = *blk_[b].desc;
                if (!(B.bc[2 * d] == 5 || B.bc[2 * d + 1] == 5)) continue;
                const int prv = B.nbr[2 * d], nxt = B.nbr[2 * d + 1], n = blk_[b].n[d];
                if (prv < 0 && nxt >= 0 && B.per[2 * d] >= 0) plane_set(b, d, 0, buf_[2 * b]);
                if (nxt < 0 && prv >= 0 && B.per[2 * d + 1] >= 0) plane_set(b, d, n + 1, buf_[2 * b + 1]);
            }
        }
    }

    // Every ghost as a GhostExpr; "" or the reason the level is refused.
    std::string emit(std::vector<GhostExpr> *out) const {
        out->clear();
        if (!bad_.empty()) return bad_;
        for (int b = 0; b < (int)blk_.size(); ++b) {
            const SymBlock &s = blk_[b];
            for (int k = 0; k <= s.n[2] + 1; ++k)
                for (int j = 0; j <= s.n[1] + 1; ++j)
                    for (int x = 0; x <= s.n[0] + 1; ++x) {
                        const bool ghost = x == 0 || x == s.n[0] + 1 || j == 0 || j == s.n[1] + 1 || k == 0 ||
                                           k == s.n[2] + 1;
                        if (!ghost) continue;
                        const long long v = val_[b][gi(b, x, j, k)];
                        if (v == kPre)
                            return "block " + std::to_string(b) + ": ghost (" + std::to_string(x) + "," +
                                   std::to_string(j) + "," + std::to_string(k) +
                                   ") is set by no exchange step (a 2:1 face's edge row the reference fills "
                                   "from stale receive buffers)";
                        GhostExpr e{};
                        e.dst = (int)(s.off + (long long)gi(b, x, j, k));
                        if (v <= kLeaf0) {
                            e.kind = 0;
                            e.src[0] = (int)(kLeaf0 - v);
                        } else {
                            const Node &nd = nodes_[(size_t)v];
                            e.kind = nd.kind;
                            for (int q = 0; q < 8; ++q) e.src[q] = nd.src[q];
                        }
                        out->push_back(e);
                    }
        }
        return "";
    }

  private:
    static constexpr long long kPre = -1;     // value from before the exchange
    static constexpr long long kLeaf0 = -2;   // leaf of arena offset o: kLeaf0 - o
    struct Node {
        int kind;    // 1: 8-cell average, 2: extrapolation
        int src[8];  // arena offsets of interior cells
    };

    static long long leaf(long long o) { return kLeaf0 - o; }
    bool is_leaf(long long v) const { return v <= kLeaf0; }
    // an 8-cell average or an extrapolation reads interior cells only
    int leaf_off(long long v) {
        if (!is_leaf(v)) {
            bad_ = "a 2:1 exchange reads a ghost cell";
            return 0;
        }
        return (int)(kLeaf0 - v);
    }

    size_t gi(int b, int x, int j, int k) const {
        const SymBlock &s = blk_[b];
        return (size_t)x + (size_t)(s.n[0] + 2) * ((size_t)j + (size_t)(s.n[1] + 2) * (size_t)k);
    }
    // cell with normal coordinate c on axis d and tangential (t1, t2):
    // (j, k) for x, (i, k) for y, (i, j) for z
    size_t fgi(int b, int d, int c, int t1, int t2) const {
        return d == 0 ? gi(b, c, t1, t2) : (d == 1 ? gi(b, t1, c, t2) : gi(b, t1, t2, c));
    }
    size_t fm(int b, int d, int t1, int t2) const {
        return (size_t)t1 + (size_t)(blk_[b].n[d == 0 ? 1 : 0] + 2) * (size_t)t2;
    }
    int t1n(int b, int d) const { return blk_[b].n[d == 0 ? 1 : 0]; }
    int t2n(int b, int d) const { return blk_[b].n[d == 2 ? 1 : 2]; }

    void plane_get(int b, int d, int c, std::vector<long long> &buf) const {
        for (int t2 = 0; t2 <= t2n(b, d) + 1; ++t2)
            for (int t1 = 0; t1 <= t1n(b, d) + 1; ++t1) buf[fm(b, d, t1, t2)] = val_[b][fgi(b, d, c, t1, t2)];
    }
    void plane_set(int b, int d, int c, const std::vector<long long> &buf) {
        for (int t2 = 0; t2 <= t2n(b, d) + 1; ++t2)
            for (int t1 = 0; t1 <= t1n(b, d) + 1; ++t1) val_[b][fgi(b, d, c, t1, t2)] = buf[fm(b, d, t1, t2)];
    }
    void plane_copy(int b, int d, int dst, int src) {
        for (int t2 = 0; t2 <= t2n(b, d) + 1; ++t2)
            for (int t1 = 0; t1 <= t1n(b, d) + 1; ++t1)
                val_[b][fgi(b, d, dst, t1, t2)] = val_[b][fgi(b, d, src, t1, t2)];
    }

    // the 8 fine cells of a fine-to-coarse ghost, in the summation order of
    // the x, y and z passes: offsets along (normal, t1, t2)
    static const int (*avg8(int d))[3] {
        static const int o[3][8][3] = {
            {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}},
            {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}},
            {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}}};
        return o[d];
    }

    // block b sends to its face neighbour o (hi: o is b's next neighbour);
    // buf is laid out on o's face
    void pack(int b, int o, int d, int hi, std::vector<long long> &buf) {
        const int rb = blk_[b].rdv, ro = blk_[o].rdv, n = blk_[b].n[d];
        if (rb == ro) {
            plane_get(b, d, hi ? n : 1, buf);
            return;
        }
        if (rb > ro) {  // fine to coarse: o's interior rows
            const int c0 = hi ? n - 1 : 1;
            const int(*w)[3] = avg8(d);
            for (int t2 = 1; t2 <= t2n(o, d); ++t2)
                for (int t1 = 1; t1 <= t1n(o, d); ++t1) {
                    Node nd{1, {}};
                    for (int q = 0; q < 8; ++q)
                        nd.src[q] = leaf_off(val_[b][fgi(b, d, c0 + w[q][0], 2 * t1 - 1 + w[q][1], 2 * t2 - 1 + w[q][2])]);
                    nodes_.push_back(nd);
                    buf[fm(o, d, t1, t2)] = (long long)nodes_.size() - 1;
                }
            return;
        }
        // coarse to fine: the coarse face cell (doubled by the receiver's
        // expression) to the 4 fine positions over it
        const int c = hi ? n : 1;
        for (int t2 = 1; t2 <= t2n(b, d); ++t2)
            for (int t1 = 1; t1 <= t1n(b, d); ++t1) {
                const long long v = val_[b][fgi(b, d, c, t1, t2)];
                for (int e2 = 0; e2 < 2; ++e2)
                    for (int e1 = 0; e1 < 2; ++e1) buf[fm(o, d, 2 * t1 - 1 + e1, 2 * t2 - 1 + e2)] = v;
            }
    }

    void unpack(int b, int o, int d, int hi, const std::vector<long long> &buf) {
        const int rb = blk_[b].rdv, ro = blk_[o].rdv, n = blk_[b].n[d];
        const int g = hi ? n + 1 : 0;
        if (rb == ro) {
            plane_set(b, d, g, buf);
            return;
        }
        const int c1 = hi ? n : 1, c2 = hi ? n - 1 : 2;
        for (int t2 = 1; t2 <= t2n(b, d); ++t2)
            for (int t1 = 1; t1 <= t1n(b, d); ++t1) {
                const long long r = buf[fm(b, d, t1, t2)];
                if (rb < ro) {  // coarse side: the average
                    val_[b][fgi(b, d, g, t1, t2)] = r;
                } else {  // fine side: ((2 phi_c + 3 phi_1) - phi_2) / 4
                    Node nd{2, {}};
                    nd.src[0] = leaf_off(r);
                    nd.src[1] = leaf_off(val_[b][fgi(b, d, c1, t1, t2)]);
                    nd.src[2] = leaf_off(val_[b][fgi(b, d, c2, t1, t2)]);
                    nodes_.push_back(nd);
                    val_[b][fgi(b, d, g, t1, t2)] = (long long)nodes_.size() - 1;
                }
            }
    }

    std::vector<SymBlock> blk_;
    std::vector<std::vector<long long>> val_;
    std::vector<std::vector<long long>> buf_;  // [2b]: minus-side receive, [2b+1]: plus side
    std::vector<Node> nodes_;
    std::string bad_;
};

}  // namespace mf3d
