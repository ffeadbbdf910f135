// tree.cpp -- geometry of one point set on the uniform, full-depth octree.
//
// The reference allocates all 8^l nodes of every level (BuildFMMHierarchy,
// H2_3D_Tree.cpp:917-977) and distributes points by recursive octant tests
// against node centres (DistributeFieldPoints, compute.hpp:1126-1203).  Here
// only non-empty boxes are kept, as compact per-level arrays in Morton order
// (the reference's leafIndex order: x bit most significant, then y, then z,
// top level first), with a dense lookup table from grid coordinates to the
// compact id.  Interaction lists are not stored: on a uniform grid the far
// field of a box is fixed by its octant inside its parent (class_directions).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "host.hpp"

namespace pbb {

static inline uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}
static inline uint64_t morton(uint32_t ix, uint32_t iy, uint32_t iz)
{
    return (spread3(ix) << 2) | (spread3(iy) << 1) | spread3(iz);
}

// Stable LSD radix sort of v = key << 32 | index by its key bits (the low
// 3*levels bits of the upper word), 11-bit digits, parallel over contiguous
// blocks: equal keys keep their index order, as a sort of the full 64-bit
// values would give.
static void radix_sort_keys(std::vector<uint64_t> &v, int key_bits)
{
    const int64_t n = (int64_t)v.size();
    constexpr int D = 11, R = 1 << D;
    const int64_t nblk = std::min<int64_t>(64, std::max<int64_t>(1, n / 65536));
    const int64_t per = (n + nblk - 1) / nblk;
    std::vector<uint64_t> tmp(n);
    std::vector<int64_t> cnt((size_t)nblk * R);
    for (int shift = 32; shift < 32 + key_bits; shift += D) {
        std::fill(cnt.begin(), cnt.end(), 0);
        parallel_for(nblk, [&](int64_t b) {
            int64_t *c = &cnt[(size_t)b * R];
            for (int64_t i = b * per, e = std::min(n, i + per); i < e; i++) c[(v[i] >> shift) & (R - 1)]++;
        });
        int64_t run = 0;  // digit-major, block-minor offsets: stable
        for (int d = 0; d < R; d++)
            for (int64_t b = 0; b < nblk; b++) {
                int64_t &c = cnt[(size_t)b * R + d];
                const int64_t k = c;
                c = run;
                run += k;
            }
        parallel_for(nblk, [&](int64_t b) {
            int64_t *c = &cnt[(size_t)b * R];
            for (int64_t i = b * per, e = std::min(n, i + per); i < e; i++) tmp[c[(v[i] >> shift) & (R - 1)]++] = v[i];
        });
        v.swap(tmp);
    }
}

uint32_t leaf_coord(const std::vector<Level> &lv, int levels, const double p[3])
{
    uint32_t c[3];
    for (int d = 0; d < 3; d++) {
        uint32_t ix = 0;
        for (int l = 0; l < levels; l++) ix = 2 * ix + (p[d] < lv[l].ctab[ix] ? 0u : 1u);
        c[d] = ix;
    }
    return c[0] | (c[1] << 10) | (c[2] << 20);
}

void build_tree(HostTree &T, const double *xyz, int64_t N, double L, int levels, int64_t nbox)
{
    if (levels < 1 || levels > 9) fail(PBB_ERR_INVALID_ARG, "tree_level must be in [1, 9]");
    if (N <= 0) fail(PBB_ERR_INVALID_ARG, "npoints must be positive");
    if (N >= (1LL << 31)) fail(PBB_ERR_INVALID_ARG, "npoints must be < 2^31");
    if (nbox < 1 || nbox > N) fail(PBB_ERR_INTERNAL, "bbox point count out of range");
    T.Lv = levels;
    T.N = N;
    // bbox of the targets, centre, shift (compute.hpp:92-127); every point
    // is checked to be finite
    const int64_t BLK = 65536, nblk = (N + BLK - 1) / BLK;
    std::vector<double> bmn((size_t)nblk * 3, 1e32), bmx((size_t)nblk * 3, -1e32);
    std::vector<char> bad(nblk, 0);
    parallel_for(nblk, [&](int64_t b) {
        for (int64_t i = b * BLK, e = std::min(N, i + BLK); i < e; i++)
            for (int d = 0; d < 3; d++) {
                double v = xyz[3 * i + d];
                if (!std::isfinite(v)) bad[b] = 1;
                if (i >= nbox) continue;
                bmn[3 * b + d] = std::min(bmn[3 * b + d], v);
                bmx[3 * b + d] = std::max(bmx[3 * b + d], v);
            }
    });
    double mn[3] = {1e32, 1e32, 1e32}, mx[3] = {-1e32, -1e32, -1e32};
    for (int64_t b = 0; b < nblk; b++) {
        if (bad[b]) fail(PBB_ERR_INVALID_ARG, "point coordinates must be finite");
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], bmn[3 * b + d]);
            mx[d] = std::max(mx[d], bmx[3 * b + d]);
        }
    }
    for (int d = 0; d < 3; d++) T.ctr[d] = 0.5 * (mn[d] + mx[d]);

    // per-axis node centres of every level, with the arithmetic of
    // BuildFMMHierarchy: child = parent -+ 0.25*len, len_child = 0.5*len
    T.lv.assign(levels + 1, Level());
    T.lv[0].ctab = {0.0};
    T.lv[0].len = L;
    for (int l = 0; l < levels; l++) {
        const Level &P = T.lv[l];
        Level &C = T.lv[l + 1];
        C.ctab.resize((size_t)1 << (l + 1));
        double q = 0.25 * P.len;
        for (size_t i = 0; i < P.ctab.size(); i++) {
            C.ctab[2 * i] = P.ctab[i] - q;
            C.ctab[2 * i + 1] = P.ctab[i] + q;
        }
        C.len = 0.5 * P.len;
    }

    // leaf coordinates by the reference's comparison chain (strict '<')
    std::vector<uint64_t> keyidx(N);
    const double half = 0.5 * L;
    std::vector<int64_t> bout(nblk, 0);
    parallel_for(nblk, [&](int64_t blk) {
        int64_t i0 = blk * BLK, i1 = std::min(N, i0 + BLK);
        for (int64_t i = i0; i < i1; i++) {
            uint32_t c[3];
            bool out = false;
            for (int d = 0; d < 3; d++) {
                double x = xyz[3 * i + d] - T.ctr[d];
                out |= x < -half || x > half;
                uint32_t ix = 0;
                for (int l = 0; l < levels; l++) {
                    double cen = T.lv[l].ctab[ix];
                    ix = 2 * ix + (x < cen ? 0u : 1u);
                }
                c[d] = ix;
            }
            bout[blk] += out ? 1 : 0;
            keyidx[i] = (morton(c[0], c[1], c[2]) << 32) | (uint64_t)i;
        }
    });
    T.outside = 0;
    for (int64_t o : bout) T.outside += o;
    radix_sort_keys(keyidx, 3 * levels);  // key, then input index: stable

    T.xs.resize(N); T.ys.resize(N); T.zs.resize(N); T.perm.resize(N);
    parallel_for(nblk, [&](int64_t blk) {
        for (int64_t k = blk * BLK, e = std::min(N, k + BLK); k < e; k++) {
            int64_t i = (int64_t)(keyidx[k] & 0xffffffffULL);
            T.perm[k] = (int32_t)i;
            T.xs[k] = xyz[3 * i] - T.ctr[0];
            T.ys[k] = xyz[3 * i + 1] - T.ctr[1];
            T.zs[k] = xyz[3 * i + 2] - T.ctr[2];
        }
    });

    // non-empty boxes per level (Morton order), leaf point ranges
    std::vector<uint64_t> keys;  // leaf keys of the non-empty leaves
    T.leaf_start.clear();
    for (int64_t k = 0; k < N; k++) {
        uint64_t key = keyidx[k] >> 32;
        if (keys.empty() || keys.back() != key) {
            keys.push_back(key);
            T.leaf_start.push_back((int32_t)k);
        }
    }
    T.leaf_start.push_back((int32_t)N);
    std::vector<std::vector<uint64_t>> lkeys(levels + 1);
    lkeys[levels] = keys;
    for (int l = levels - 1; l >= 0; l--) {
        auto &up = lkeys[l];
        for (uint64_t k : lkeys[l + 1]) {
            uint64_t pk = k >> 3;
            if (up.empty() || up.back() != pk) up.push_back(pk);
        }
    }
    T.max_leaf = 0;
    for (size_t b = 0; b + 1 < T.leaf_start.size(); b++)
        T.max_leaf = std::max<int64_t>(T.max_leaf, T.leaf_start[b + 1] - T.leaf_start[b]);

    for (int l = 0; l <= levels; l++) {
        Level &V = T.lv[l];
        const auto &lk = lkeys[l];
        V.nb = (int64_t)lk.size();
        V.coord.resize(V.nb);
        const int64_t S = 1LL << l;
        V.map.assign((size_t)(S * S * S), -1);
        for (int64_t b = 0; b < V.nb; b++) {
            uint64_t k = lk[b];
            uint32_t ix = 0, iy = 0, iz = 0;
            for (int bit = 0; bit < l; bit++) {
                ix |= (uint32_t)((k >> (3 * bit + 2)) & 1) << bit;
                iy |= (uint32_t)((k >> (3 * bit + 1)) & 1) << bit;
                iz |= (uint32_t)((k >> (3 * bit)) & 1) << bit;
            }
            V.coord[b] = (int32_t)(ix | (iy << 10) | (iz << 20));
            V.map[((size_t)ix * S + iy) * S + iz] = (int32_t)b;
        }
    }
    for (int l = 0; l < levels; l++) {
        Level &P = T.lv[l];
        const Level &C = T.lv[l + 1];
        P.child.assign((size_t)P.nb * 8, -1);
        const int64_t S = 1LL << (l + 1);
        for (int64_t b = 0; b < P.nb; b++) {
            uint32_t c = (uint32_t)P.coord[b];
            uint32_t ix = c & 1023, iy = (c >> 10) & 1023, iz = (c >> 20) & 1023;
            for (int o = 0; o < 8; o++) {
                uint32_t cx = 2 * ix + ((o >> 2) & 1), cy = 2 * iy + ((o >> 1) & 1), cz = 2 * iz + (o & 1);
                P.child[b * 8 + o] = C.map[((size_t)cx * S + cy) * S + cz];
            }
        }
    }
    // M2L target order: grouped by octant class, Morton order inside
    T.m2l_pairs = 0;
    std::vector<int> Ktable(343, 0);
    int cdirs[8][189][4];
    for (int o = 0; o < 8; o++) class_directions(o, Ktable, cdirs[o]);
    for (int l = 2; l <= levels; l++) {
        Level &V = T.lv[l];
        V.order.resize(V.nb);
        std::fill(std::begin(V.class_count), std::end(V.class_count), 0);
        for (int64_t b = 0; b < V.nb; b++) {
            uint32_t c = (uint32_t)V.coord[b];
            int o = (int)(((c & 1) << 2) | (((c >> 10) & 1) << 1) | ((c >> 20) & 1));
            V.class_count[o]++;
        }
        int64_t start[9] = {0};
        for (int o = 0; o < 8; o++) start[o + 1] = start[o] + V.class_count[o];
        int64_t fillp[8];
        for (int o = 0; o < 8; o++) fillp[o] = start[o];
        const int64_t S = 1LL << l;
        for (int64_t b = 0; b < V.nb; b++) {
            uint32_t c = (uint32_t)V.coord[b];
            int o = (int)(((c & 1) << 2) | (((c >> 10) & 1) << 1) | ((c >> 20) & 1));
            V.order[fillp[o]++] = (int32_t)b;
        }
        const int64_t nbb = (V.nb + 4095) / 4096;
        std::vector<int64_t> part(nbb, 0);
        parallel_for(nbb, [&](int64_t blk) {
            int64_t q = 0;
            for (int64_t b = blk * 4096, e = std::min(V.nb, b + 4096); b < e; b++) {
                uint32_t c = (uint32_t)V.coord[b];
                int ix = c & 1023, iy = (c >> 10) & 1023, iz = (c >> 20) & 1023;
                const int(*dirs)[4] = cdirs[((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1)];
                for (int k = 0; k < 189; k++) {
                    int64_t sx = ix + dirs[k][0], sy = iy + dirs[k][1], sz = iz + dirs[k][2];
                    if (sx < 0 || sy < 0 || sz < 0 || sx >= S || sy >= S || sz >= S) continue;
                    if (V.map[((size_t)sx * S + sy) * S + sz] >= 0) q++;
                }
            }
            part[blk] = q;
        });
        for (int64_t q : part) T.m2l_pairs_level[l] += q;
        T.m2l_pairs += T.m2l_pairs_level[l];
    }
    // ordered near-field point pairs (27 neighbours incl. self)
    {
        const Level &V = T.lv[levels];
        const int64_t S = 1LL << levels;
        int64_t q = 0;
        for (int64_t b = 0; b < V.nb; b++) {
            uint32_t c = (uint32_t)V.coord[b];
            int ix = c & 1023, iy = (c >> 10) & 1023, iz = (c >> 20) & 1023;
            int64_t nt = T.leaf_start[b + 1] - T.leaf_start[b];
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        int64_t sx = ix + dx, sy = iy + dy, sz = iz + dz;
                        if (sx < 0 || sy < 0 || sz < 0 || sx >= S || sy >= S || sz >= S) continue;
                        int32_t s = V.map[((size_t)sx * S + sy) * S + sz];
                        if (s >= 0) q += nt * (int64_t)(T.leaf_start[s + 1] - T.leaf_start[s]);
                    }
        }
        T.near_pairs = q;
    }
}

}  // namespace pbb
