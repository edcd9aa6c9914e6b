// mg_ghostmap.hpp -- host side: the copy sequence of exchange(op) and
// pbound on the caller's ghost shell, run on origins to give a gather map.
//
// exchange(op) and pbound only copy values, so one call of them is a fixed
// map: every entry it writes ends up holding the value some entry
// (interior or ghost) held before the call.  The map is found once, by
// running the reference's copy sequence on origins instead of values, and
// then applied on the device as a gather.  The sequence, on one rank with
// equal resolution and a shell of G = pl ghost layers:
//   exchange_bc (exchange_bc.for:14-841), when any face is periodic, for
//     ly = 0..nly: for every periodic face without a neighbour, in block
//     order and face order W, E, S, N, B, T, a block alone along that axis
//     copies its own opposite plane (ghost plane 0-ly <- plane ie-ly, plane
//     ie+1+ly <- plane 1+ly), otherwise it hands the partner its plane 1+ly
//     (ie-ly); whole ghosted planes; the hand-over loop runs twice
//     (:16-330, :332-640) and the partners unpack last (:642-838);
//   exchange_smlvl (exchange.for:342-2499), for ly = 0..nly: every block
//     packs before any block unpacks; each block then unpacks, at ly = 0
//     only, its 8 corners and 12 edges as (pl+1)-deep cubes (an entry is
//     written where it lies outside the block's interior), then its 6 faces
//     (ghost plane 0-ly or ie+1+ly over the interior rows, from the
//     sender's plane ie-ly or 1+ly);
//   pbound (press.for:77-137, pressure only): each physical non-periodic
//     face, W, E, S, N, B, T, copies the adjacent interior plane over the
//     one-cell box (isp-1..iep+1 on the other axes).
// nly is pl_ex for u, v, w and 0 for p (exchange.for:14-23).  Coordinates:
// interior 1..n, ghosts 0, -1, ... and n+1, n+2, ...; a block's shell
// spans 1-G .. n+G on every axis.
#pragma once
#include <cstdint>
#include <map>
#include <vector>

#include "../../include/mf3d_mg.h"
#include "mg_common.hpp"

namespace mf3d {

// Origins of every entry of one field over all blocks' shells.  hb holds
// the shell geometry: off is the arena offset of coordinate (0,0,0), so
// gidx(off, i, j, k) addresses coordinates 1-G .. n+G.
struct ShellOrigins {
    const std::vector<BlockLevel> &hb;
    int G;
    std::vector<uint32_t> o;

    ShellOrigins(const std::vector<BlockLevel> &h, int g, size_t arena) : hb(h), G(g), o(arena) {
        for (size_t e = 0; e < arena; ++e) o[e] = (uint32_t)e;
    }
    uint32_t at(int b, int i, int j, int k) const {
        const BlockLevel &g = hb[b];
        return (uint32_t)(g.off + i + g.sy * j + g.sz * k);
    }
    uint32_t get(int b, int i, int j, int k) const { return o[at(b, i, j, k)]; }
    void set(int b, int i, int j, int k, uint32_t v) { o[at(b, i, j, k)] = v; }
};

inline int shell_ext(const BlockLevel &g, int a) { return a == 0 ? g.n0 : (a == 1 ? g.n1 : g.n2); }

// (i, j, k) with coordinate `idx` on axis a and (p, q) on the two other
// axes in increasing order.
inline void shell_ijk(int a, int idx, int p, int q, int *ijk) {
    ijk[a] = idx;
    ijk[a == 0 ? 1 : 0] = p;
    ijk[a == 2 ? 1 : 2] = q;
}

// Last index of component comp on axis a: n, or n-1 before a physical
// non-periodic wall without a neighbour (localparameters.for:185-190);
// cell fields (comp < 0): n.
inline int shell_end(const mf3d_block_desc &B, int a, int comp) {
    const int n = B.n[a];
    return (comp == a && B.nbr[2 * a + 1] < 0 && B.bc[2 * a + 1] != 5) ? n - 1 : n;
}

// Visits every (p, q) of a whole ghosted plane normal to axis a.
template <typename Fn>
inline void shell_plane(const BlockLevel &g, int G, int a, Fn fn) {
    const int e1 = shell_ext(g, a == 0 ? 1 : 0), e2 = shell_ext(g, a == 2 ? 1 : 2);
    for (int q = 1 - G; q <= e2 + G; ++q)
        for (int p = 1 - G; p <= e1 + G; ++p) fn(p, q);
}

inline void shell_exchange_bc(const std::vector<mf3d_block_desc> &blk, ShellOrigins &S, int comp, int nly) {
    const int nb = (int)blk.size(), G = S.G;
    for (int ly = 0; ly <= nly; ++ly) {
        std::map<std::pair<int, int>, std::vector<uint32_t>> bufs;
        for (int pass = 0; pass < 2; ++pass)
            for (int b = 0; b < nb; ++b)
                for (int a = 0; a < 3; ++a)
                    for (int side = 0; side < 2; ++side) {
                        const int f = 2 * a + side;
                        if (!(blk[b].bc[f] == 5 && blk[b].nbr[f] < 0)) continue;
                        const int ie = shell_end(blk[b], a, comp);
                        if (blk[b].nbr[2 * a + 1 - side] < 0) {  // alone along a: own opposite plane
                            const int dst = side == 0 ? -ly : ie + 1 + ly, src = side == 0 ? ie - ly : 1 + ly;
                            shell_plane(S.hb[b], G, a, [&](int p, int q) {
                                int d[3], s[3];
                                shell_ijk(a, dst, p, q, d);
                                shell_ijk(a, src, p, q, s);
                                S.set(b, d[0], d[1], d[2], S.get(b, s[0], s[1], s[2]));
                            });
                        } else {
                            std::vector<uint32_t> &buf = bufs[{blk[b].per[f], 2 * a + 1 - side}];
                            buf.clear();
                            const int src = side == 0 ? 1 + ly : ie - ly;
                            shell_plane(S.hb[b], G, a, [&](int p, int q) {
                                int s[3];
                                shell_ijk(a, src, p, q, s);
                                buf.push_back(S.get(b, s[0], s[1], s[2]));
                            });
                        }
                    }
        for (int b = 0; b < nb; ++b)
            for (int a = 0; a < 3; ++a)
                for (int side = 0; side < 2; ++side) {
                    const int f = 2 * a + side;
                    if (!(blk[b].bc[f] == 5 && blk[b].nbr[f] < 0 && blk[b].nbr[2 * a + 1 - side] >= 0)) continue;
                    const int dst = side == 0 ? -ly : shell_end(blk[b], a, comp) + 1 + ly;
                    const auto it = bufs.find({b, f});
                    if (it == bufs.end()) continue;  // no partner handed a plane over
                    const std::vector<uint32_t> &buf = it->second;
                    size_t t = 0;
                    shell_plane(S.hb[b], G, a, [&](int p, int q) {
                        int d[3];
                        shell_ijk(a, dst, p, q, d);
                        S.set(b, d[0], d[1], d[2], buf[t++]);
                    });
                }
    }
}

// Block in direction t (components -1, 0, +1) through chained face
// neighbours (no periodic wrap), or -1: the edge and corner neighbours that
// alloc_dom.for:617-1163 finds by matching block coordinates.  Every order
// of the chain must agree (a conforming block grid).
inline int shell_diag(const std::vector<mf3d_block_desc> &blk, int b, const int *t, bool *ambiguous) {
    int res = -2;
    const int orders[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (const auto &ord : orders) {
        int c = b;
        for (int m = 0; m < 3 && c >= 0; ++m) {
            const int a = ord[m];
            if (t[a]) c = blk[c].nbr[2 * a + (t[a] > 0 ? 1 : 0)];
        }
        if (res == -2) res = c;
        else if (res != c) *ambiguous = true;
    }
    return res;
}

// corprev1..4, cornext1..4, edgprev1..6, edgnext1..6 (exchange.for:2118-2389)
constexpr int kShellDiag[20][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1}, {1, 1, 1},
                                   {1, -1, 1},   {-1, -1, 1}, {-1, 1, 1}, {-1, 0, -1}, {0, 1, -1},
                                   {1, 0, -1},   {0, -1, -1}, {-1, -1, 0}, {-1, 1, 0}, {1, 0, 1},
                                   {0, -1, 1},   {-1, 0, 1},  {0, 1, 1},  {1, 1, 0},   {1, -1, 0}};

inline bool shell_exchange_smlvl(const std::vector<mf3d_block_desc> &blk, ShellOrigins &S, int comp, int nly) {
    bool ambiguous = false;
    const int nb = (int)blk.size(), G = S.G;
    for (int ly = 0; ly <= nly; ++ly) {
        const std::vector<uint32_t> pre = S.o;  // every block packs before any unpacks
        auto pre_get = [&](int b, int i, int j, int k) { return pre[S.at(b, i, j, k)]; };
        for (int d = 0; d < nb; ++d) {
            const BlockLevel &gd = S.hb[d];
            const int nd[3] = {gd.n0, gd.n1, gd.n2};
            // corners and edges (ly = 0): along an axis where the sender lies
            // below, the receiver's 1-q takes the sender's n+1-q; above, the
            // receiver's n+q the sender's q, q = 0..pl; written where some
            // q >= 1; edges run over the interior of their free axis
            for (int ti = 0; ly == 0 && ti < 20; ++ti) {
                const int *t = kShellDiag[ti];
                const int s = shell_diag(blk, d, t, &ambiguous);
                if (s < 0) continue;
                const BlockLevel &gs = S.hb[s];
                const int ns[3] = {gs.n0, gs.n1, gs.n2};
                bool conform = true;
                for (int a = 0; a < 3; ++a) conform = conform && (t[a] != 0 || ns[a] == nd[a]);
                if (!conform) {
                    ambiguous = true;
                    continue;
                }
                int lo[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = t[a] ? 0 : 1;
                    hi[a] = t[a] ? G : nd[a];
                }
                for (int v2 = lo[2]; v2 <= hi[2]; ++v2)
                    for (int v1 = lo[1]; v1 <= hi[1]; ++v1)
                        for (int v0 = lo[0]; v0 <= hi[0]; ++v0) {
                            const int v[3] = {v0, v1, v2};
                            int di[3], si[3];
                            bool any = false;
                            for (int a = 0; a < 3; ++a) {
                                if (!t[a]) {
                                    di[a] = si[a] = v[a];
                                    continue;
                                }
                                const int q = v[a];
                                any = any || q >= 1;
                                if (t[a] < 0) {
                                    di[a] = 1 - q;
                                    si[a] = ns[a] + 1 - q;
                                } else {
                                    di[a] = nd[a] + q;
                                    si[a] = q;
                                }
                            }
                            if (any) S.set(d, di[0], di[1], di[2], pre_get(s, si[0], si[1], si[2]));
                        }
            }
            // faces: the interior rows of ghost plane 0-ly (ie+1+ly) from the
            // sender's plane ie-ly (1+ly)
            for (int a = 0; a < 3; ++a)
                for (int side = 0; side < 2; ++side) {
                    const int s = blk[d].nbr[2 * a + side];
                    if (s < 0) continue;
                    const int dst = side == 0 ? -ly : shell_end(blk[d], a, comp) + 1 + ly;
                    const int src = side == 0 ? shell_end(blk[s], a, comp) - ly : 1 + ly;
                    const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
                    for (int q = 1; q <= nd[o2]; ++q)
                        for (int p = 1; p <= nd[o1]; ++p) {
                            int di[3], si[3];
                            shell_ijk(a, dst, p, q, di);
                            shell_ijk(a, src, p, q, si);
                            S.set(d, di[0], di[1], di[2], pre_get(s, si[0], si[1], si[2]));
                        }
                }
        }
    }
    return !ambiguous;
}

inline void shell_pbound(const std::vector<mf3d_block_desc> &blk, ShellOrigins &S) {
    for (int b = 0; b < (int)blk.size(); ++b) {
        const BlockLevel &g = S.hb[b];
        for (int a = 0; a < 3; ++a)
            for (int side = 0; side < 2; ++side) {
                const int f = 2 * a + side;
                if (!(blk[b].nbr[f] < 0 && blk[b].bc[f] != 5)) continue;
                const int n = shell_ext(g, a);
                const int dst = side == 0 ? 0 : n + 1, src = side == 0 ? 1 : n;
                const int e1 = shell_ext(g, a == 0 ? 1 : 0), e2 = shell_ext(g, a == 2 ? 1 : 2);
                for (int q = 0; q <= e2 + 1; ++q)
                    for (int p = 0; p <= e1 + 1; ++p) {
                        int d[3], s[3];
                        shell_ijk(a, dst, p, q, d);
                        shell_ijk(a, src, p, q, s);
                        S.set(b, d[0], d[1], d[2], S.get(b, s[0], s[1], s[2]));
                    }
            }
    }
}

// {destination, source} of every entry the sequence changed.
inline void shell_entries(const ShellOrigins &S, std::vector<uint2> &out) {
    for (size_t e = 0; e < S.o.size(); ++e)
        if (S.o[e] != (uint32_t)e) out.push_back(make_uint2((uint32_t)e, S.o[e]));
}

}  // namespace mf3d
