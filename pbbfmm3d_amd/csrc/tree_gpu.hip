// tree_gpu.hip -- the tree of one point set, built on the GPU.
//
// Same result as build_tree (tree.cpp), array for array: the bbox centre
// (compute.hpp:92-127), leaf coordinates by the comparison chain against the
// node centres of BuildFMMHierarchy (H2_3D_Tree.cpp:917-977) with the strict
// '<' of DistributeFieldPoints (compute.hpp:1149-1173), points sorted by
// (Morton leaf key, input index), and per level the non-empty boxes in
// Morton order with their coordinates, dense map, children, octant-class
// order, point ranges, L2P segments, upward-pass work items, and the
// near-field leaf classes.  Radix sorts and scans are hipCUB's; every other
// step is one small kernel over points, leaves or boxes.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "tree_gpu.hpp"

namespace pbb {
namespace dev {
namespace {

constexpr int TPB = 256;

inline int nblocks(int64_t n, int cap = 65536)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + TPB - 1) / TPB, cap));
}

__device__ __forceinline__ uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

// packed coordinates ix | iy << 10 | iz << 20 of a level-l Morton key
__device__ __forceinline__ int32_t key_coord(uint32_t k, int l)
{
    uint32_t ix = 0, iy = 0, iz = 0;
    for (int bit = 0; bit < l; bit++) {
        ix |= ((k >> (3 * bit + 2)) & 1u) << bit;
        iy |= ((k >> (3 * bit + 1)) & 1u) << bit;
        iz |= ((k >> (3 * bit)) & 1u) << bit;
    }
    return (int32_t)(ix | (iy << 10) | (iz << 20));
}

// per-workgroup min and max of each axis over the first nbox points
// (part[6 g + d] = min, [6 g + 3 + d] = max); all N points are checked to be
// finite
__global__ __launch_bounds__(TPB) void bbox_part(const double *__restrict__ xyz, int64_t N, int64_t nbox,
                                                 double *__restrict__ part, int *__restrict__ bad)
{
    __shared__ double s[6][TPB];
    double mn[3] = {1e32, 1e32, 1e32}, mx[3] = {-1e32, -1e32, -1e32};
    int nf = 0;
    for (int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x; i < N; i += (int64_t)gridDim.x * TPB)
        for (int d = 0; d < 3; d++) {
            const double v = xyz[3 * i + d];
            nf |= !isfinite(v);
            if (i < nbox) {
                mn[d] = fmin(mn[d], v);
                mx[d] = fmax(mx[d], v);
            }
        }
    for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x] = mn[d];
        s[3 + d][threadIdx.x] = mx[d];
    }
    if (nf) atomicOr(bad, 1);
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int d = 0; d < 3; d++) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + h]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + h]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

struct Centre {
    double c[3];
};

// key << 32 | input index of every point, and the count of points outside
// the root cube; ctab: node centres of levels 0..L-1, level l at 2^l - 1
__global__ __launch_bounds__(TPB) void point_keys(const double *__restrict__ xyz, int64_t N, Centre ctr, double half,
                                                  const double *__restrict__ ctab, int levels,
                                                  uint64_t *__restrict__ keys, unsigned long long *__restrict__ outside)
{
    unsigned long long cnt = 0;
    for (int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x; i < N; i += (int64_t)gridDim.x * TPB) {
        uint32_t c[3];
        bool out = false;
        for (int d = 0; d < 3; d++) {
            const double x = xyz[3 * i + d] - ctr.c[d];
            out |= x < -half || x > half;
            uint32_t ix = 0;
            for (int l = 0; l < levels; l++) ix = 2 * ix + (x < ctab[(1 << l) - 1 + ix] ? 0u : 1u);
            c[d] = ix;
        }
        cnt += out ? 1 : 0;
        const uint64_t key = (spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2]);
        keys[i] = (key << 32) | (uint64_t)i;
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(outside, cnt);
}

// sorted, re-centred SoA coordinates, the permutation, leaf keys and the
// heads of equal-key runs
__global__ __launch_bounds__(TPB) void gather_sorted(const uint64_t *__restrict__ keys, const double *__restrict__ xyz,
                                                     int64_t N, Centre ctr, int levels, double *__restrict__ xs,
                                                     double *__restrict__ ys, double *__restrict__ zs,
                                                     int32_t *__restrict__ perm, uint32_t *__restrict__ lk,
                                                     int32_t *__restrict__ pt_leaf, int *__restrict__ head)
{
    for (int64_t k = blockIdx.x * (int64_t)TPB + threadIdx.x; k < N; k += (int64_t)gridDim.x * TPB) {
        const uint64_t v = keys[k];
        const int64_t i = (int64_t)(v & 0xffffffffULL);
        const uint32_t key = (uint32_t)(v >> 32);
        perm[k] = (int32_t)i;
        xs[k] = xyz[3 * i] - ctr.c[0];
        ys[k] = xyz[3 * i + 1] - ctr.c[1];
        zs[k] = xyz[3 * i + 2] - ctr.c[2];
        lk[k] = key;
        pt_leaf[k] = key_coord(key, levels);
        head[k] = k == 0 || key != (uint32_t)(keys[k - 1] >> 32);
    }
}

// first sorted point and key of every leaf
__global__ __launch_bounds__(TPB) void leaf_runs(const int *__restrict__ head, const int *__restrict__ pos,
                                                 const uint32_t *__restrict__ lk, int64_t N,
                                                 int32_t *__restrict__ leaf_start, uint32_t *__restrict__ leafkey)
{
    for (int64_t k = blockIdx.x * (int64_t)TPB + threadIdx.x; k < N; k += (int64_t)gridDim.x * TPB)
        if (head[k]) {
            leaf_start[pos[k]] = (int32_t)k;
            leafkey[pos[k]] = lk[k];
        }
}

__global__ void set_i32(int32_t *p, int32_t v) { *p = v; }

// heads of the runs of equal parent keys among the nb child keys
__global__ __launch_bounds__(TPB) void parent_heads(const uint32_t *__restrict__ ck, int64_t nb, int *__restrict__ head)
{
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB)
        head[b] = b == 0 || (ck[b] >> 3) != (ck[b - 1] >> 3);
}

// parent keys and first leaves, and each child's parent index
__global__ __launch_bounds__(TPB) void parent_boxes(const uint32_t *__restrict__ ck, const int32_t *__restrict__ cfirst,
                                                    const int *__restrict__ head, const int *__restrict__ pos, int64_t nb,
                                                    uint32_t *__restrict__ pk, int32_t *__restrict__ pfirst,
                                                    int32_t *__restrict__ parent)
{
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB) {
        const int p = pos[b] + head[b] - 1;
        parent[b] = p;
        if (head[b]) {
            pk[p] = ck[b] >> 3;
            pfirst[p] = cfirst[b];
        }
    }
}

// coordinates and dense map of one level's boxes
__global__ __launch_bounds__(TPB) void level_boxes(const uint32_t *__restrict__ key, int64_t nb, int l,
                                                   int32_t *__restrict__ coord, int32_t *__restrict__ map)
{
    const int64_t S = 1LL << l;
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB) {
        const int32_t c = key_coord(key[b], l);
        coord[b] = c;
        const int64_t ix = c & 1023, iy = (c >> 10) & 1023, iz = (c >> 20) & 1023;
        map[(ix * S + iy) * S + iz] = (int32_t)b;
    }
}

__global__ __launch_bounds__(TPB) void fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    for (int64_t q = blockIdx.x * (int64_t)TPB + threadIdx.x; q < n; q += (int64_t)gridDim.x * TPB) p[q] = v;
}

// child[p * 8 + octant] of every parent p, from each child's parent index
__global__ __launch_bounds__(TPB) void child_table(const uint32_t *__restrict__ ck, const int32_t *__restrict__ parent,
                                                   int64_t nb, int32_t *__restrict__ child)
{
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB)
        child[(int64_t)parent[b] * 8 + (ck[b] & 7u)] = (int32_t)b;
}

// point range start of every box from its first leaf, the octant class of
// every box and the class histogram
__global__ __launch_bounds__(TPB) void box_ranges(const int32_t *__restrict__ first, const int32_t *__restrict__ leaf_start,
                                                  const uint32_t *__restrict__ key, int64_t nb,
                                                  int32_t *__restrict__ pstart, uint32_t *__restrict__ cls,
                                                  int32_t *__restrict__ ids, unsigned long long *__restrict__ hist)
{
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB) {
        pstart[b] = leaf_start[first[b]];
        cls[b] = key[b] & 7u;
        ids[b] = (int32_t)b;
        atomicAdd(&hist[key[b] & 7u], 1ull);
    }
}

// per box: L2P segments, upward work items, scratch slots, reductions
__global__ __launch_bounds__(TPB) void box_counts(const int32_t *__restrict__ pstart, int64_t nb, int *__restrict__ nseg,
                                                  int *__restrict__ nitem, int *__restrict__ nslot,
                                                  int *__restrict__ nred)
{
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB) {
        const int np = pstart[b + 1] - pstart[b];
        nseg[b] = (np + L2P_SEG - 1) / L2P_SEG;
        const int m = (np + P2M_SEG - 1) / P2M_SEG;
        const bool split = np > P2M_SEG;
        nitem[b] = split ? m : 1;
        nslot[b] = split ? m : 0;
        nred[b] = split ? 1 : 0;
    }
}

struct LevelBase {
    int li;           // level index l - 2
    int64_t off;      // global box offset of the level
    int item0, slot0, red0;
};

__global__ __launch_bounds__(TPB) void box_emit(const int32_t *__restrict__ pstart, int64_t nb,
                                                const int *__restrict__ sseg, const int *__restrict__ sitem,
                                                const int *__restrict__ sslot, const int *__restrict__ sred,
                                                LevelBase lb, int32_t *__restrict__ seg_box, int32_t *__restrict__ seg_p0,
                                                int4 *__restrict__ items, int4 *__restrict__ reds)
{
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB) {
        const int p0 = pstart[b], p1 = pstart[b + 1], np = p1 - p0;
        int s = sseg[b];
        for (int q = p0; q < p1; q += L2P_SEG, s++) {
            seg_box[s] = (int32_t)b;
            seg_p0[s] = q;
        }
        const int it = lb.item0 + sitem[b];
        if (np <= P2M_SEG) {
            items[it] = int4{lb.li, (int)b, p0, -1};
            continue;
        }
        const int first = lb.slot0 + sslot[b];
        int k = 0;
        for (int q = p0; q < p1; q += P2M_SEG, k++) items[it + k] = int4{lb.li, (int)b, q, first + k};
        reds[lb.red0 + sred[b]] = int4{(int)(lb.off + b), first, k, 0};
    }
}

// near-field class of every leaf (class c: at most 2 << c points; the last
// class takes the rest), its histogram, the largest leaf and the ordered
// near-field point pairs
__global__ __launch_bounds__(TPB) void leaf_stats(const int32_t *__restrict__ leaf_start, const int32_t *__restrict__ coord,
                                                  const int32_t *__restrict__ map, int64_t nb, int levels,
                                                  uint32_t *__restrict__ cls, int32_t *__restrict__ ids,
                                                  unsigned long long *__restrict__ hist, int *__restrict__ max_leaf,
                                                  unsigned long long *__restrict__ near_pairs)
{
    const int64_t S = 1LL << levels;
    unsigned long long q = 0;
    int mx = 0;
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB) {
        const int nt = leaf_start[b + 1] - leaf_start[b];
        int c = 0;
        while (c < NEAR_CLASSES - 1 && nt > (2 << c)) c++;
        cls[b] = (uint32_t)c;
        ids[b] = (int32_t)b;
        atomicAdd(&hist[c], 1ull);
        mx = max(mx, nt);
        const int32_t cc = coord[b];
        const int64_t ix = cc & 1023, iy = (cc >> 10) & 1023, iz = (cc >> 20) & 1023;
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    const int64_t sx = ix + dx, sy = iy + dy, sz = iz + dz;
                    if (sx < 0 || sy < 0 || sz < 0 || sx >= S || sy >= S || sz >= S) continue;
                    const int32_t s = map[(sx * S + sy) * S + sz];
                    if (s >= 0) q += (unsigned long long)nt * (unsigned long long)(leaf_start[s + 1] - leaf_start[s]);
                }
    }
    for (int o = 32; o > 0; o >>= 1) {
        q += __shfl_xor(q, o);
        mx = max(mx, __shfl_xor(mx, o));
    }
    if ((threadIdx.x & 63) == 0) {
        if (q) atomicAdd(near_pairs, q);
        atomicMax(max_leaf, mx);
    }
}

// far-field (target, source) box pairs of one level: the 189 offsets of
// each box's octant class that land on a non-empty box
__global__ __launch_bounds__(TPB) void m2l_pair_count(const int32_t *__restrict__ coord, const int32_t *__restrict__ map,
                                                      int64_t nb, int l, const int *__restrict__ dirs,
                                                      unsigned long long *__restrict__ pairs)
{
    const int64_t S = 1LL << l;
    unsigned long long q = 0;
    for (int64_t b = blockIdx.x * (int64_t)TPB + threadIdx.x; b < nb; b += (int64_t)gridDim.x * TPB) {
        const int32_t c = coord[b];
        const int64_t ix = c & 1023, iy = (c >> 10) & 1023, iz = (c >> 20) & 1023;
        const int *d = dirs + (((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1)) * 189 * 4;
        for (int k = 0; k < 189; k++) {
            const int64_t sx = ix + d[4 * k], sy = iy + d[4 * k + 1], sz = iz + d[4 * k + 2];
            if (sx < 0 || sy < 0 || sz < 0 || sx >= S || sy >= S || sz >= S) continue;
            if (map[(sx * S + sy) * S + sz] >= 0) q++;
        }
    }
    for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o);
    if ((threadIdx.x & 63) == 0 && q) atomicAdd(pairs, q);
}

// device allocations of one build: freed unless handed over; a buffer is
// freed only after the stream's work on it is done
struct Pool {
    hipStream_t st;
    std::vector<void *> owned;
    template <typename T>
    T *get(size_t n, const char *what)
    {
        void *p = nullptr;
        const hipError_t e = hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(e == hipErrorOutOfMemory ? PBB_ERR_OUT_OF_MEMORY : PBB_ERR_HIP,
                 std::string("hipMalloc(") + what + "): " + hipGetErrorString(e));
        }
        owned.push_back(p);
        return (T *)p;
    }
    void drop(void *p)
    {
        auto it = std::find(owned.begin(), owned.end(), p);
        if (it != owned.end()) {
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            (void)hipFree(p);
            owned.erase(it);
        }
    }
    void keep(void *p) { owned.erase(std::remove(owned.begin(), owned.end(), p), owned.end()); }
    ~Pool()
    {
        if (!owned.empty()) (void)hipStreamSynchronize(st);
        for (void *p : owned) (void)hipFree(p);
    }
};

// hipCUB scratch, grown as needed
struct Temp {
    Pool &pool;
    void *p = nullptr;
    size_t bytes = 0;
    void *need(size_t b)
    {
        if (b > bytes) {
            if (p) pool.drop(p);
            p = pool.get<char>(b, "sort / scan scratch");
            bytes = b;
        }
        return p;
    }
};

template <typename T>
T fetch(const T *d, hipStream_t st)
{
    T h{};
    check(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, st), "copy a count to the host");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return h;
}

int exclusive_count(Temp &tmp, const int *in, int *out, int64_t n, hipStream_t st)
{
    size_t b = 0;
    check(hipcub::DeviceScan::ExclusiveSum(nullptr, b, in, out, (int)n, st), "scan size");
    check(hipcub::DeviceScan::ExclusiveSum(tmp.need(b), b, in, out, (int)n, st), "scan");
    int last_in = 0, last_out = 0;
    check(hipMemcpyAsync(&last_in, in + n - 1, sizeof(int), hipMemcpyDeviceToHost, st), "copy");
    check(hipMemcpyAsync(&last_out, out + n - 1, sizeof(int), hipMemcpyDeviceToHost, st), "copy");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return last_in + last_out;
}

// stable sort of ids by a small class key (bits [0, bits))
void sort_by_class(Temp &tmp, Pool &pool, const uint32_t *cls, const int32_t *ids, int32_t *out, int64_t n, int bits,
                   hipStream_t st)
{
    uint32_t *cls_out = pool.get<uint32_t>(n, "class sort keys");
    size_t b = 0;
    check(hipcub::DeviceRadixSort::SortPairs(nullptr, b, cls, cls_out, ids, out, (int)n, 0, bits, st), "sort size");
    check(hipcub::DeviceRadixSort::SortPairs(tmp.need(b), b, cls, cls_out, ids, out, (int)n, 0, bits, st), "sort");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    pool.drop(cls_out);
}

}  // namespace

void build_tree_gpu(GpuTree &T, const double *xyz_host, int64_t N, int64_t nbox, double L, int levels,
                    const std::vector<std::vector<double>> &ctab, const int dirs[8][189][4], hipStream_t st)
{
    if (levels < 1 || levels > 9) fail(PBB_ERR_INVALID_ARG, "tree_level must be in [1, 9]");
    if (N <= 0) fail(PBB_ERR_INVALID_ARG, "npoints must be positive");
    if (N >= (1LL << 31)) fail(PBB_ERR_INVALID_ARG, "npoints must be < 2^31");
    if (nbox < 1 || nbox > N) fail(PBB_ERR_INTERNAL, "bbox point count out of range");
    Pool pool{st, {}};
    Temp tmp{pool};
    T = GpuTree{};
    T.Lv = levels;
    T.N = N;
    double *xyz = pool.get<double>((size_t)3 * N, "points");
    check(hipMemcpyAsync(xyz, xyz_host, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st), "upload points");
    // bbox and centre (compute.hpp:92-127)
    const int nbb = nblocks(N, 1024);
    double *part = pool.get<double>((size_t)6 * nbb, "bbox partials");
    int *bad = pool.get<int>(1, "flag");
    unsigned long long *cnt = pool.get<unsigned long long>(16, "counters");
    check(hipMemsetAsync(bad, 0, sizeof(int), st), "clear flag");
    check(hipMemsetAsync(cnt, 0, 16 * sizeof(unsigned long long), st), "clear counters");
    bbox_part<<<nbb, TPB, 0, st>>>(xyz, N, nbox, part, bad);
    check(hipGetLastError(), "bbox_part");
    std::vector<double> hp((size_t)6 * nbb);
    check(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st), "copy bbox");
    int hbad = 0;
    check(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st), "copy flag");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (hbad) fail(PBB_ERR_INVALID_ARG, "point coordinates must be finite");
    double mn[3] = {1e32, 1e32, 1e32}, mx[3] = {-1e32, -1e32, -1e32};
    for (int g = 0; g < nbb; g++)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], hp[6 * g + d]);
            mx[d] = std::max(mx[d], hp[6 * g + 3 + d]);
        }
    Centre ctr;
    for (int d = 0; d < 3; d++) T.ctr[d] = ctr.c[d] = 0.5 * (mn[d] + mx[d]);

    // leaf keys, stable sort by (key, input index)
    std::vector<double> cflat;
    for (int l = 0; l < levels; l++) cflat.insert(cflat.end(), ctab[l].begin(), ctab[l].end());
    double *dct = pool.get<double>(cflat.size(), "node centres");
    check(hipMemcpyAsync(dct, cflat.data(), sizeof(double) * cflat.size(), hipMemcpyHostToDevice, st), "upload centres");
    uint64_t *keys = pool.get<uint64_t>(N, "keys"), *skeys = pool.get<uint64_t>(N, "sorted keys");
    point_keys<<<nblocks(N), TPB, 0, st>>>(xyz, N, ctr, 0.5 * L, dct, levels, keys, cnt);
    check(hipGetLastError(), "point_keys");
    {
        size_t b = 0;
        check(hipcub::DeviceRadixSort::SortKeys(nullptr, b, keys, skeys, (int)N, 0, 32 + 3 * levels, st), "sort size");
        check(hipcub::DeviceRadixSort::SortKeys(tmp.need(b), b, keys, skeys, (int)N, 0, 32 + 3 * levels, st), "sort");
    }
    pool.drop(keys);
    T.x = pool.get<double>(N, "x");
    T.y = pool.get<double>(N, "y");
    T.z = pool.get<double>(N, "z");
    T.perm = pool.get<int32_t>(N, "perm");
    T.pt_leaf = pool.get<int32_t>(N, "leaf of each point");
    uint32_t *lk = pool.get<uint32_t>(N, "point leaf keys");
    int *head = pool.get<int>(N, "run heads"), *pos = pool.get<int>(N, "run positions");
    gather_sorted<<<nblocks(N), TPB, 0, st>>>(skeys, xyz, N, ctr, levels, T.x, T.y, T.z, T.perm, lk, T.pt_leaf, head);
    check(hipGetLastError(), "gather_sorted");
    pool.drop(skeys);
    pool.drop(xyz);
    const int nleaf = exclusive_count(tmp, head, pos, N, st);
    T.leaf_start = pool.get<int32_t>((size_t)nleaf + 1, "leaf_start");
    std::vector<uint32_t *> lkeys(levels + 1, nullptr);
    std::vector<int32_t *> lfirst(levels + 1, nullptr);
    lkeys[levels] = pool.get<uint32_t>(nleaf, "leaf keys");
    leaf_runs<<<nblocks(N), TPB, 0, st>>>(head, pos, lk, N, T.leaf_start, lkeys[levels]);
    check(hipGetLastError(), "leaf_runs");
    set_i32<<<1, 1, 0, st>>>(T.leaf_start + nleaf, (int32_t)N);
    check(hipGetLastError(), "set_i32");
    pool.drop(lk);
    pool.drop(head);
    pool.drop(pos);

    // levels L-1 .. 0: parents are the runs of equal key >> 3
    T.lv.assign(levels + 1, GpuLevel());
    T.lv[levels].nb = nleaf;
    lfirst[levels] = pool.get<int32_t>(nleaf, "first leaves");
    {
        std::vector<int32_t> iota(nleaf);
        for (int b = 0; b < nleaf; b++) iota[b] = b;
        check(hipMemcpyAsync(lfirst[levels], iota.data(), sizeof(int32_t) * nleaf, hipMemcpyHostToDevice, st),
              "upload leaf ids");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    std::vector<int32_t *> parent(levels + 1, nullptr);
    for (int l = levels - 1; l >= 0; l--) {
        const int64_t nbc = T.lv[l + 1].nb;
        int *h = pool.get<int>(nbc, "heads"), *p = pool.get<int>(nbc, "positions");
        parent_heads<<<nblocks(nbc), TPB, 0, st>>>(lkeys[l + 1], nbc, h);
        check(hipGetLastError(), "parent_heads");
        const int nbp = exclusive_count(tmp, h, p, nbc, st);
        T.lv[l].nb = nbp;
        lkeys[l] = pool.get<uint32_t>(nbp, "box keys");
        lfirst[l] = pool.get<int32_t>(nbp, "first leaves");
        parent[l + 1] = pool.get<int32_t>(nbc, "parent ids");
        parent_boxes<<<nblocks(nbc), TPB, 0, st>>>(lkeys[l + 1], lfirst[l + 1], h, p, nbc, lkeys[l], lfirst[l],
                                                    parent[l + 1]);
        check(hipGetLastError(), "parent_boxes");
        pool.drop(h);
        pool.drop(p);
    }
    for (int l = 0; l <= levels; l++) {
        GpuLevel &V = T.lv[l];
        const int64_t S3 = 1LL << (3 * l);
        V.coord = pool.get<int32_t>(V.nb, "coord");
        V.map = pool.get<int32_t>(S3, "map");
        fill_i32<<<nblocks(S3), TPB, 0, st>>>(V.map, S3, -1);
        level_boxes<<<nblocks(V.nb), TPB, 0, st>>>(lkeys[l], V.nb, l, V.coord, V.map);
        check(hipGetLastError(), "level_boxes");
        if (l < levels) {
            V.child = pool.get<int32_t>((size_t)V.nb * 8, "child");
            fill_i32<<<nblocks(V.nb * 8), TPB, 0, st>>>(V.child, V.nb * 8, -1);
            child_table<<<nblocks(T.lv[l + 1].nb), TPB, 0, st>>>(lkeys[l + 1], parent[l + 1], T.lv[l + 1].nb, V.child);
            check(hipGetLastError(), "child_table");
        }
    }

    // levels >= 2: point ranges, octant-class order, L2P segments, upward items
    struct Counts {
        int *seg, *item, *slot, *red, *sseg, *sitem, *sslot, *sred;
        int nseg, nitem, nslot, nred;
    };
    std::vector<Counts> cnts(levels + 1);
    int items_total = 0, slots_total = 0, reds_total = 0;
    std::vector<LevelBase> base(levels + 1);
    int64_t off = 0;
    for (int l = 2; l <= levels; l++) {
        GpuLevel &V = T.lv[l];
        V.off = off;
        off += V.nb;
        V.pstart = pool.get<int32_t>((size_t)V.nb + 1, "box point ranges");
        uint32_t *cls = pool.get<uint32_t>(V.nb, "classes");
        int32_t *ids = pool.get<int32_t>(V.nb, "ids");
        unsigned long long *hist = cnt + 8;
        check(hipMemsetAsync(hist, 0, 8 * sizeof(unsigned long long), st), "clear histogram");
        box_ranges<<<nblocks(V.nb), TPB, 0, st>>>(lfirst[l], T.leaf_start, lkeys[l], V.nb, V.pstart, cls, ids, hist);
        check(hipGetLastError(), "box_ranges");
        set_i32<<<1, 1, 0, st>>>(V.pstart + V.nb, (int32_t)N);
        unsigned long long hh[8];
        check(hipMemcpyAsync(hh, hist, sizeof(hh), hipMemcpyDeviceToHost, st), "copy histogram");
        V.order = pool.get<int32_t>(V.nb, "class order");
        sort_by_class(tmp, pool, cls, ids, V.order, V.nb, 3, st);
        for (int c = 0; c < 8; c++) V.class_count[c] = (int64_t)hh[c];
        pool.drop(cls);
        pool.drop(ids);
        Counts &C = cnts[l];
        C.seg = pool.get<int>(V.nb, "counts");
        C.item = pool.get<int>(V.nb, "counts");
        C.slot = pool.get<int>(V.nb, "counts");
        C.red = pool.get<int>(V.nb, "counts");
        C.sseg = pool.get<int>(V.nb, "scans");
        C.sitem = pool.get<int>(V.nb, "scans");
        C.sslot = pool.get<int>(V.nb, "scans");
        C.sred = pool.get<int>(V.nb, "scans");
        box_counts<<<nblocks(V.nb), TPB, 0, st>>>(V.pstart, V.nb, C.seg, C.item, C.slot, C.red);
        check(hipGetLastError(), "box_counts");
        C.nseg = exclusive_count(tmp, C.seg, C.sseg, V.nb, st);
        C.nitem = exclusive_count(tmp, C.item, C.sitem, V.nb, st);
        C.nslot = exclusive_count(tmp, C.slot, C.sslot, V.nb, st);
        C.nred = exclusive_count(tmp, C.red, C.sred, V.nb, st);
        base[l] = LevelBase{l - 2, V.off, items_total, slots_total, reds_total};
        V.up_item0 = items_total;
        V.up_red0 = reds_total;
        items_total += C.nitem;
        slots_total += C.nslot;
        reds_total += C.nred;
    }
    T.Btot = off;
    if (off >= (1LL << 31)) fail(PBB_ERR_UNSUPPORTED, "too many boxes");
    T.up_items = pool.get<int4>(items_total, "upward work items");
    T.up_reds = pool.get<int4>(reds_total, "upward reductions");
    T.n_up_items = items_total;
    T.n_up_reds = reds_total;
    T.n_up_slots = slots_total;
    for (int l = 2; l <= levels; l++) {
        GpuLevel &V = T.lv[l];
        Counts &C = cnts[l];
        V.nseg = C.nseg;
        V.seg_box = pool.get<int32_t>(C.nseg, "L2P segments");
        V.seg_p0 = pool.get<int32_t>(C.nseg, "L2P segments");
        box_emit<<<nblocks(V.nb), TPB, 0, st>>>(V.pstart, V.nb, C.sseg, C.sitem, C.sslot, C.sred, base[l], V.seg_box,
                                                 V.seg_p0, T.up_items, T.up_reds);
        check(hipGetLastError(), "box_emit");
    }

    // near-field leaf classes, largest leaf, pair counts
    {
        const GpuLevel &V = T.lv[levels];
        uint32_t *cls = pool.get<uint32_t>(V.nb, "leaf classes");
        int32_t *ids = pool.get<int32_t>(V.nb, "ids");
        int *mxl = pool.get<int>(1, "max leaf");
        check(hipMemsetAsync(mxl, 0, sizeof(int), st), "clear");
        check(hipMemsetAsync(cnt + 8, 0, 8 * sizeof(unsigned long long), st), "clear histogram");
        leaf_stats<<<nblocks(V.nb), TPB, 0, st>>>(T.leaf_start, V.coord, V.map, V.nb, levels, cls, ids, cnt + 8, mxl,
                                                   cnt + 1);
        check(hipGetLastError(), "leaf_stats");
        T.leaf_class = pool.get<int32_t>(V.nb, "leaf classes");
        sort_by_class(tmp, pool, cls, ids, T.leaf_class, V.nb, 3, st);
        unsigned long long hh[8];
        check(hipMemcpyAsync(hh, cnt + 8, sizeof(hh), hipMemcpyDeviceToHost, st), "copy histogram");
        T.max_leaf = fetch(mxl, st);
        int run = 0;
        for (int c = 0; c < NEAR_CLASSES; c++) {
            T.leaf_class_off[c] = run;
            run += (int)hh[c];
        }
        T.leaf_class_off[NEAR_CLASSES] = run;
        pool.drop(cls);
        pool.drop(ids);
    }
    int *ddirs = pool.get<int>(8 * 189 * 4, "directions");
    check(hipMemcpyAsync(ddirs, dirs, sizeof(int) * 8 * 189 * 4, hipMemcpyHostToDevice, st), "upload directions");
    // far-field box pairs per level (the shared counters' slots 8.. hold
    // the leaf-class histogram)
    unsigned long long *lpairs = pool.get<unsigned long long>(16, "pair counters");
    check(hipMemsetAsync(lpairs, 0, 16 * sizeof(unsigned long long), st), "clear pair counters");
    for (int l = 2; l <= levels; l++) {
        m2l_pair_count<<<nblocks(T.lv[l].nb, 4096), TPB, 0, st>>>(T.lv[l].coord, T.lv[l].map, T.lv[l].nb, l, ddirs,
                                                                  lpairs + l);
        check(hipGetLastError(), "m2l_pair_count");
    }
    unsigned long long hc[2], hl[16];
    check(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, st), "copy counters");
    check(hipMemcpyAsync(hl, lpairs, sizeof(hl), hipMemcpyDeviceToHost, st), "copy pair counters");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    T.outside = (int64_t)hc[0];
    T.near_pairs = (int64_t)hc[1];
    T.m2l_pairs = 0;
    for (int l = 2; l <= levels; l++) {
        T.m2l_pairs_level[l] = (int64_t)hl[l];
        T.m2l_pairs += T.m2l_pairs_level[l];
    }

    // hand the tree's arrays over; everything else is freed with the pool
    for (void *p : {(void *)T.x, (void *)T.y, (void *)T.z, (void *)T.perm, (void *)T.pt_leaf, (void *)T.leaf_start,
                    (void *)T.leaf_class, (void *)T.up_items, (void *)T.up_reds})
        pool.keep(p);
    for (auto &V : T.lv)
        for (void *p : {(void *)V.coord, (void *)V.map, (void *)V.child, (void *)V.order, (void *)V.pstart,
                        (void *)V.seg_box, (void *)V.seg_p0})
            if (p) pool.keep(p);
}

}  // namespace dev
}  // namespace pbb
