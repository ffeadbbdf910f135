// kernels_mixed.hip -- mixed-precision M2L on the 16-bit matrix cores.
//
// The M2L contraction Pf[:, t] = sum_d E_d P[:, t + d] (compute.hpp:464-493)
// runs on v_mfma_f32_16x16x32_{bf16,f16} with FP32 accumulation.  Each FP32
// operand x is split into NSPLIT 16-bit parts x = x0 + x1 (+ x2) and the
// product keeps the terms x_i y_j with i + j < NSPLIT:
//   bf16 x3: 6 MFMAs per block, ~24 significand bits per operand
//   f16  x2: 3 MFMAs per block, ~22 bits; operands are pre-scaled by powers
//            of two (E per operator set, P per level) into FP16's range and
//            the scale is removed exactly from the FP32 result
//   bf16 x2: 3 MFMAs per block, ~16 bits (the low-accuracy end)
// Layouts:
//   Es  [set][d][chunk][plane][R][32]  split E_d, rows padded to R = 16*RB,
//                                     k padded to chunks of 32, zeros
//   P32 [box][col][kpad]               FP32 proxies, k padded with zeros (the
//                                      proxy-pass levels keep rows of 16 or 32
//                                      inside their own region)
//   Pf32[box][col][rf]                 FP32 result (proxy-pass levels: rows of 32)
// Fragment layout of v_mfma_f32_16x16x32_{bf16,f16} (lane l of 64, e < 8):
//   A[i][k]: i = l % 16, k = 8*(l/16) + e     B[k][j]: j = l % 16, k = 8*(l/16) + e
//   D[i][j]: j = l % 16, i = 4*(l/16) + v     (v = 0..3, FP32)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

namespace pbb {
namespace dev {

__constant__ int c_dirs[8][189][4];  // far directions of each octant class (this module's copy)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Vec8;
template <>
struct Vec8<__bf16> {
    typedef bf16x8 type;
};
template <>
struct Vec8<_Float16> {
    typedef f16x8 type;
};

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(f16x8 a, f16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}


// ---------------------------------------------------------------------------
// Es from the FP64 operators: x = E_d[row + k*rf] * 2^escale, split in FP64.
template <typename T, int NSPLIT>
__global__ void split_e(const double *__restrict__ E, int rf, int rs, int R, int nch, int escale,
                        T *__restrict__ Es)
{
    const int64_t per_d = (int64_t)nch * NSPLIT * R * 32;
    const int64_t total = 316 * per_d;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t d = q / per_d, rem = q - d * per_d;
        int ch = (int)(rem / ((int64_t)NSPLIT * R * 32));
        rem -= (int64_t)ch * NSPLIT * R * 32;
        int p = (int)(rem / (R * 32));
        rem -= (int64_t)p * R * 32;
        int row = (int)(rem / 32), kk = (int)(rem % 32);
        int k = ch * 32 + kk;
        double x = (row < rf && k < rs) ? ldexp(E[d * (int64_t)rf * rs + row + (int64_t)k * rf], escale) : 0.0;
        T part = (T)0.0f;
        for (int i = 0; i <= p; i++) {
            part = (T)(float)x;
            x -= (double)(float)part;
        }
        Es[q] = part;
    }
}

// max |x| over the first `used` floats of rows rows of stride ld into *out
// (as ordered bits; out cleared by caller)
__global__ void absmax_f32(const float *__restrict__ x, int64_t rows, int ld, int used, unsigned *__restrict__ out)
{
    float m = 0.0f;
    const int64_t n = rows * used;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = q / used;
        m = fmaxf(m, fabsf(x[r * ld + (q - r * used)]));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// FP16 operands are scaled into range: P by 2^xs with xs from the level's
// max |P| (its FP32 bits in *xmax), E by 2^escale per operator set
__device__ __forceinline__ int p_scale_exp(const unsigned *xmax)
{
    int xs = 0;
    if (xmax) {
        const float mx = __uint_as_float(*xmax);
        if (mx > 0.0f) {
            int e;
            frexpf(mx, &e);
            xs = 12 - e;
        }
    }
    return xs;
}

// Two-part proxies, in place: each of the first cpr 32-float k-chunks of a
// row of P (128 bytes) is rewritten as 32 high then 32 low 16-bit parts of
// x * 2^xs, so that the M2L loads its B fragments ready-made.  One thread
// per chunk.  KD = 16: one 16-deep chunk per row (sets whose ks is at most
// 16), rewritten as 16 high then 16 low parts in its first 64 bytes.
template <typename T, int KD>
__global__ void split_p2(float *__restrict__ P, int64_t rows, int ld, int cpr, const unsigned *__restrict__ xmax)
{
    const float xscale = ldexpf(1.0f, p_scale_exp(xmax));
    const int64_t nchunks = rows * cpr;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nchunks;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = q / cpr;
        float4 *c = (float4 *)(P + r * ld + (q - r * cpr) * KD);
        float x[KD];
#pragma unroll
        for (int i = 0; i < KD / 4; i++) {
            const float4 v = c[i];
            x[4 * i] = v.x * xscale;
            x[4 * i + 1] = v.y * xscale;
            x[4 * i + 2] = v.z * xscale;
            x[4 * i + 3] = v.w * xscale;
        }
        T h[KD], l[KD];
#pragma unroll
        for (int e = 0; e < KD; e++) {
            h[e] = (T)x[e];
            l[e] = (T)(x[e] - (float)h[e]);
        }
        uint4 *o = (uint4 *)c;
#pragma unroll
        for (int i = 0; i < KD / 8; i++) {
            o[i] = *(const uint4 *)&h[8 * i];
            o[KD / 8 + i] = *(const uint4 *)&l[8 * i];
        }
    }
}

void launch_split_p2(int variant, float *P, int64_t rows, int ld, int cpr, const unsigned *xmax, hipStream_t st,
                     int kd)
{
    const int64_t nchunks = rows * cpr;
    if (nchunks <= 0) return;
    if (kd != 32 && !(kd == 16 && cpr == 1 && variant == PBB_PREC_MIXED_F16X2))
        fail(PBB_ERR_INTERNAL, "split_p2: 16-deep chunks are for one f16 chunk per row");
    if (ld % 4 || cpr * kd > ld) fail(PBB_ERR_INTERNAL, "split_p2: chunks past the row stride");
    const int blocks = (int)std::min<int64_t>((nchunks + 255) / 256, 65536);
    if (variant == PBB_PREC_MIXED_F16X2 && kd == 16)
        split_p2<_Float16, 16><<<blocks, 256, 0, st>>>(P, rows, ld, cpr, xmax);
    else if (variant == PBB_PREC_MIXED_F16X2) split_p2<_Float16, 32><<<blocks, 256, 0, st>>>(P, rows, ld, cpr, xmax);
    else if (variant == PBB_PREC_MIXED_BF16X2) split_p2<__bf16, 32><<<blocks, 256, 0, st>>>(P, rows, ld, cpr, xmax);
    else fail(PBB_ERR_INVALID_ARG, "proxies are pre-split for the two-part modes only");
    check(hipGetLastError(), "split_p2");
}

// ---------------------------------------------------------------------------
// One workgroup of 4 waves owns TB (<= 16) target boxes of one octant class
// x cpt columns (W = 64*NCB columns); each wave holds NCB 16-column blocks x
// RB 16-row blocks of FP32 accumulators.  The (direction, 32-deep k chunk)
// steps of the tile's non-empty directions form one software pipeline: the
// next step's E chunk (pre-split) and proxy chunk are loaded into registers
// while the current step's MFMAs run from LDS.  Steps run in the order of
// a.chunk_major (see m2l_step; direction-major by default, so that the proxy
// gathers of consecutive steps continue along the same columns).  Two
// workgroups share a
// CU and cover each other's stalls.  Two-part modes read proxies pre-split by
// split_p2 ([32 high][32 low] 16-bit parts per k chunk, the same 128 bytes
// as 32 floats); the three-part mode splits FP32 proxies in registers.
// With the row split (gridDim.z = 2) a workgroup holds RB of the row blocks
// of r_f, starting at row r0: the two halves of a tile run as two
// workgroups, adjacent in launch order so that they share the proxy gathers
// through L2, and wide tiles keep two workgroups per CU.
// Direction groups: a launch covers only the directions whose operator index
// lies in [a.dlo, a.dhi), so every workgroup of the launch reads the same few
// E_d and they stay in L2; later groups add to the partial sums in Pf
// (a.accumulate), and a tile with no direction in the group is skipped.
constexpr int TB_MAX = 16;
__device__ __attribute__((aligned(16))) float g_zero32[32];  // a zero k chunk (empty slots), never written

// KD = 16 (f16 x2, sets whose ks is at most 16): each step pairs two
// directions' 16-deep chunks into one 32-deep product, A = [E_d1 | E_d2],
// B = [P(t+d1); P(t+d2)], so one MFMA adds both directions (half the steps,
// barriers and MFMAs of one direction per step).  The proxies are pre-split
// as 16 high + 16 low parts in the first 64 bytes of each row (half the
// proxy bytes per (box, direction) pair); E_d is read as the first 16 k of
// its 32-deep chunk.
template <typename T, int NSPLIT, int RB, int NCB, int KD = 32>
__global__ __launch_bounds__(256, NCB * RB <= 24 ? 2 : 1) void m2l_mfma_split(M2LMixArgs a)
{
    static_assert(KD == 32 || (KD == 16 && NSPLIT == 2), "paired 16-deep steps are for pre-split two-part proxies");
    constexpr bool PAIR = KD == 16;
    typedef typename Vec8<T>::type V8;
    typedef V8 VK;
    constexpr int NT = 256, W = 64 * NCB, R = 16 * RB;  // R: rows of this workgroup
    constexpr int KU = 4;               // 16-byte units of one E row's 32-deep k chunk
    constexpr int KS = 40;              // LDS row stride of As (16-bit elements)
    constexpr int PU = 8;               // 16-byte units of one column's proxy chunk
    constexpr int BS = 36;              // LDS row stride of Bs (floats)
    constexpr int NA = NSPLIT * R * KU;  // 16-byte units of one E chunk
    constexpr int AR = (NA + NT - 1) / NT;
    constexpr int BR = W * PU / NT;     // 16-byte units of one proxy chunk per thread
    constexpr bool PRESPLIT = NSPLIT == 2;
    __shared__ __attribute__((aligned(16))) T As[NSPLIT * R * KS];
    __shared__ __attribute__((aligned(16))) float Bs[W * BS];
    __shared__ int s_sid[189 * TB_MAX];
    __shared__ short s_dir[189];
    __shared__ int s_nd, s_wcnt[3];
    __shared__ int s_tgt[TB_MAX];
    __shared__ int s_xyz[TB_MAX][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rf = a.rf, TB = a.TB, nch = a.nch, RE = a.R;
    const int lin0 = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                              gridDim.x * gridDim.y * gridDim.z);
    // the row halves and column chunks of one tile are adjacent in launch order
    const int r0 = (lin0 % gridDim.z) * R, lin = lin0 / gridDim.z;
    const int tile = lin / gridDim.y;
    int o = 0;
    while (o < 7 && tile >= a.tile_prefix[o + 1]) o++;
    const int64_t first = a.class_start[o] + (int64_t)(tile - a.tile_prefix[o]) * TB;
    const int64_t cend = a.class_start[o + 1];
    const int c0 = (lin - tile * (int)gridDim.y) * a.cpt;
    const int cpt = min(a.cpt, a.ncols - c0);
    const int S = a.side;
    const int xs = p_scale_exp(a.xmax);
    const float xscale = ldexpf(1.0f, xs);
    const float unscale = ldexpf(1.0f, -(xs + a.escale));
    if (tid < TB) {
        int64_t q = first + tid;
        int t = q < cend ? a.order[q] : -1;
        s_tgt[tid] = t;
        if (t >= 0) {
            uint32_t c = (uint32_t)a.coord[t];
            s_xyz[tid][0] = c & 1023;
            s_xyz[tid][1] = (c >> 10) & 1023;
            s_xyz[tid][2] = (c >> 20) & 1023;
        }
    }
    __syncthreads();
    // source box of every (direction, target) pair of the tile
    for (int q = tid; q < 189 * TB; q += NT) {
        const int k = q / TB, tb = q - k * TB;
        int sid = -1;
        const int e = c_dirs[o][k][3];
        if (s_tgt[tb] >= 0 && e >= a.dlo && e < a.dhi) {
            int sx = s_xyz[tb][0] + c_dirs[o][k][0], sy = s_xyz[tb][1] + c_dirs[o][k][1],
                sz = s_xyz[tb][2] + c_dirs[o][k][2];
            if (sx >= 0 && sy >= 0 && sz >= 0 && sx < S && sy < S && sz < S)
                sid = a.map[((int64_t)sx * S + sy) * S + sz];
        }
        s_sid[k * TB_MAX + tb] = sid;
    }
    __syncthreads();
    // compact the directions that reach a source box, in class order
    // (waves 0..2 cover k = 0..191: one ballot each, then the prefix)
    bool live = false;
    int pre = 0;
    if (tid < 192) {
        if (tid < 189)
            for (int tb = 0; tb < TB; tb++) live |= s_sid[tid * TB_MAX + tb] >= 0;
        const unsigned long long m = __ballot(live);
        pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_wcnt[wave] = __popcll(m);
    }
    __syncthreads();
    if (live) s_dir[pre + (wave > 0 ? s_wcnt[0] : 0) + (wave > 1 ? s_wcnt[1] : 0)] = (short)tid;
    if (tid == 0) s_nd = s_wcnt[0] + s_wcnt[1] + s_wcnt[2];
    __syncthreads();
    const int nsteps = PAIR ? (s_nd + 1) / 2 : s_nd * nch;
    if (a.accumulate && nsteps == 0) return;  // nothing to add for this tile

    // Loads are unconditional, so that the compiler's wait counts stay exact:
    // rows past RE repeat the last row (their results are never stored) and
    // empty (box, direction) slots read a line of zeros.
    u32x4 ra[AR];
    f32x4 pb0[BR], pb1[BR];  // proxy chunks of two consecutive steps
    const int last = nsteps - 1;
    // step -> first direction (index into s_dir) and k chunk
    auto step_dir = [&](int step, int &di, int &ch) __attribute__((always_inline)) {
        if constexpr (PAIR) {
            di = 2 * step;
            ch = 0;
        } else {
            m2l_step(step, s_nd, nch, a.chunk_major, di, ch);
        }
    };
    auto loadE = [&](int step) __attribute__((always_inline)) {
        int di, ch;
        step_dir(min(step, last), di, ch);
        const int k0 = s_dir[di], k1 = s_dir[min(di + 1, s_nd - 1)];  // (a padding direction's proxies are zero)
        const u32x4 *src0 = (const u32x4 *)((const T *)a.Es + ((size_t)c_dirs[o][k0][3] * nch + ch) * NSPLIT * RE * 32);
        const u32x4 *src1 = (const u32x4 *)((const T *)a.Es + ((size_t)c_dirs[o][k1][3] * nch + ch) * NSPLIT * RE * 32);
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int u = min(tid + r * NT, NA - 1), p = u / (R * KU), row = min(r0 + (u / KU) % R, RE - 1),
                      q = u % KU;
            if constexpr (PAIR) ra[r] = (q < 2 ? src0 : src1)[(p * RE + row) * 4 + (q & 1)];  // 16 k of each
            else ra[r] = src0[(p * RE + row) * 4 + q];  // rows of 32-deep chunks: 4 units each
        }
    };
    auto loadP = [&](int step, f32x4 (&pb)[BR]) __attribute__((always_inline)) {
        int di, ch;
        step_dir(min(step, last), di, ch);
#pragma unroll
        for (int r = 0; r < BR; r++) {
            const int u = tid + r * NT, col = u / PU, k4 = u % PU;
            const int tb = col / cpt, cl = col - tb * cpt;
            const int dd = PAIR ? di + (k4 >> 2) : di, kq = PAIR ? (k4 & 3) : k4;
            const int k = s_dir[min(dd, s_nd - 1)];
            // zeros past the last step and for the padding direction of an odd count
            const int sid = tb < TB && step <= last && dd < s_nd ? s_sid[k * TB_MAX + tb] : -1;
            const float *src = sid >= 0 ? a.P + ((size_t)(a.off + sid) * a.ncols + c0 + cl) * a.kpad + ch * 32 : g_zero32;
            pb[r] = *(const f32x4 *)(src + 4 * kq);
        }
    };
    f32x4 acc[RB][NCB];
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < NCB; j++) acc[i][j] = f32x4{0, 0, 0, 0};

    const int fr = lane % 16, fq = lane / 16;
    if (nsteps > 0) {
        loadE(0);
        loadP(0, pb0);
        loadP(1, pb1);
    }
    if (a.accumulate) {  // continue from the earlier groups' sums (exact power-of-two rescale)
        // unconditional loads (clamped to the tile's first box and row r_f - 1),
        // so that they overlap instead of waiting one by one
        const float rescale = ldexpf(1.0f, xs + a.escale);
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int col = (wave * NCB + cb) * 16 + fr;
            const int tb = col / cpt, cl = min(col - tb * cpt, cpt - 1);
            const int t = tb < TB ? s_tgt[tb] : -1;
            const float *src = a.Pf + ((size_t)(a.off + (t >= 0 ? t : s_tgt[0])) * a.ncols + c0 + cl) * a.ldf;
#pragma unroll
            for (int rb_ = 0; rb_ < RB; rb_++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = r0 + rb_ * 16 + 4 * fq + v;
                    const float x = src[min(row, rf - 1)];
                    acc[rb_][cb][v] = t >= 0 && row < rf ? x * rescale : 0.0f;
                }
        }
    }
    // one step: E chunk loaded one step ahead, proxy chunk two steps ahead
    auto run = [&](int step, f32x4 (&pb)[BR]) __attribute__((always_inline)) {
        __syncthreads();  // every wave is done with the previous step's LDS
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int u = tid + r * NT;
            if (AR * NT == NA || u < NA) {
                const int p = u / (R * KU), row = (u / KU) % R, q = u % KU;
                *(u32x4 *)&As[(p * R + row) * KS + 8 * q] = ra[r];
            }
        }
#pragma unroll
        for (int r = 0; r < BR; r++) {
            const int u = tid + r * NT;
            *(f32x4 *)&Bs[(u / PU) * BS + 4 * (u % PU)] = pb[r];
        }
        __syncthreads();
        loadE(step + 1);  // in flight during the MFMAs
        loadP(step + 2, pb);
        VK bfr[NCB][NSPLIT];
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const float *bp = &Bs[((wave * NCB + cb) * 16 + fr) * BS];
            if constexpr (PAIR) {  // [d1: 8 floats high, 8 low][d2: the same]; k 8fq.. of plane p
                bfr[cb][0] = *(const VK *)(bp + 16 * (fq >> 1) + 4 * (fq & 1));
                bfr[cb][1] = *(const VK *)(bp + 16 * (fq >> 1) + 8 + 4 * (fq & 1));
            } else if constexpr (PRESPLIT) {  // plane p of k 8fq..8fq+7 at floats 16p + 4fq
                bfr[cb][0] = *(const VK *)(bp + 4 * fq);
                bfr[cb][1] = *(const VK *)(bp + 16 + 4 * fq);
            } else {
                const float4 v0 = *(const float4 *)(bp + 8 * fq), v1 = *(const float4 *)(bp + 8 * fq + 4);
                float x[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (int e = 0; e < 8; e++) x[e] *= xscale;
#pragma unroll
                for (int p = 0; p < NSPLIT; p++)
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const T h = (T)x[e];
                        bfr[cb][p][e] = h;
                        x[e] -= (float)h;
                    }
            }
        }
#pragma unroll
        for (int rb_ = 0; rb_ < RB; rb_++) {
            VK afr[NSPLIT];
#pragma unroll
            for (int p = 0; p < NSPLIT; p++) afr[p] = *(const VK *)&As[(p * R + rb_ * 16 + fr) * KS + 8 * fq];
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) {
                // smallest terms first
                if constexpr (NSPLIT == 3) {
                    acc[rb_][cb] = mfma16(afr[2], bfr[cb][0], acc[rb_][cb]);
                    acc[rb_][cb] = mfma16(afr[1], bfr[cb][1], acc[rb_][cb]);
                    acc[rb_][cb] = mfma16(afr[0], bfr[cb][2], acc[rb_][cb]);
                }
                acc[rb_][cb] = mfma16(afr[1], bfr[cb][0], acc[rb_][cb]);
                acc[rb_][cb] = mfma16(afr[0], bfr[cb][1], acc[rb_][cb]);
                acc[rb_][cb] = mfma16(afr[0], bfr[cb][0], acc[rb_][cb]);
            }
        }
    };
    // an even number of steps: a padding step (zero proxies) adds nothing
    for (int step = 0; step < nsteps; step += 2) {
        run(step, pb0);
        run(step + 1, pb1);
    }
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
        const int col = (wave * NCB + cb) * 16 + fr;
        const int tb = col / cpt, cl = col - tb * cpt;
        if (tb >= TB) continue;
        const int t = s_tgt[tb];
        if (t < 0) continue;
        float *dst = a.Pf + ((size_t)(a.off + t) * a.ncols + c0 + cl) * a.ldf;
#pragma unroll
        for (int rb_ = 0; rb_ < RB; rb_++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = r0 + rb_ * 16 + 4 * fq + v;
                if (row < rf) dst[row] = acc[rb_][cb][v] * unscale;
            }
    }
}

// ---------------------------------------------------------------------------
// Row-wave variant: 4 waves split the rows of r_f (RBW 16-row blocks each)
// and share 16*NCB columns; WCOL = 2 puts two such groups of 4 waves side by
// side in one workgroup (W = 32*NCB columns), so that each E_d chunk serves
// twice the columns (the second wave's load of the same rows hits L1/L2
// right after the first).  A wave loads
// its rows of the E_d chunk straight into MFMA operand registers (one step
// ahead, no LDS); the proxy chunk goes through LDS once for all waves
// (loaded three steps ahead, written to the other half of a double buffer
// one step ahead, one barrier per step).  Per step a wave reads NCB B
// fragments from LDS, half the LDS traffic of the column-wave layout, each
// one block ahead of its MFMAs.
// A column block whose boxes have no source in the step's direction is
// skipped: its proxies are zero, so the sums are unchanged bit for bit.
// Step bookkeeping (direction, k chunk) runs in scalar registers; the
// compacted per-direction tables (E offset, box mask, source boxes) are in
// LDS, read before the step's barrier.  A step's loads and proxy stores are
// issued between its first MFMA blocks, so that they overlap the matrix
// cores instead of delaying the first MFMA after the barrier.
// LO (two-part modes): the cross products of the split go to accumulators of
// their own, so the sum of the leading products takes one FP32 rounding per
// k chunk instead of three; the two sums are added once at the end.
template <typename T, int NSPLIT, int RBW, int NCB, int WCOL, bool LO = false>
__global__ __launch_bounds__(256 * WCOL, WCOL == 1 ? 2 : 1) void m2l_mfma_rows(M2LMixArgs a)
{
    static_assert(!LO || NSPLIT == 2, "separate cross-product sums are for the two-part modes");
    typedef typename Vec8<T>::type V8;
    constexpr int NT = 256 * WCOL, W = 16 * NCB * WCOL, BS = 36;
    constexpr int BR = W * 8 / NT;  // 16-byte proxy units per thread and step
    constexpr bool PRESPLIT = NSPLIT == 2;
    static_assert(BR >= 1 && BR * NT == W * 8, "tile width");
    __shared__ __attribute__((aligned(16))) float Bs[2][W * BS];
    __shared__ int s_sid[189 * TB_MAX];   // by class direction
    __shared__ int s_sidc[189 * TB_MAX];  // by compacted direction
    __shared__ int s_eoff[189];           // operator index * nch, by compacted direction
    __shared__ unsigned short s_mask[189];
    __shared__ int s_nd, s_wcnt[3];
    __shared__ int s_tgt[TB_MAX];
    __shared__ int s_xyz[TB_MAX][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rf = a.rf, TB = a.TB, nch = a.nch, RE = a.R;
    // gridDim.z row groups of 4 * RBW row blocks (ranks above one
    // workgroup's rows); the row groups and column chunks of one tile are
    // adjacent in launch order
    const int gyz = (int)(gridDim.y * gridDim.z);
    const int lin = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x * gyz);
    const int tile = lin / gyz;
    const int rem = lin - tile * gyz, chunk = rem / (int)gridDim.z;
    const int rgroup = rem - chunk * (int)gridDim.z;
    int o = 0;
    while (o < 7 && tile >= a.tile_prefix[o + 1]) o++;
    const int64_t first = a.class_start[o] + (int64_t)(tile - a.tile_prefix[o]) * TB;
    const int64_t cend = a.class_start[o + 1];
    const int c0 = chunk * a.cpt;
    const int cpt = min(a.cpt, a.ncols - c0);
    const int S = a.side;
    const int xs = p_scale_exp(a.xmax);
    const float xscale = ldexpf(1.0f, xs);
    const float unscale = ldexpf(1.0f, -(xs + a.escale));
    if (tid < TB) {
        int64_t q = first + tid;
        int t = q < cend ? a.order[q] : -1;
        s_tgt[tid] = t;
        if (t >= 0) {
            uint32_t c = (uint32_t)a.coord[t];
            s_xyz[tid][0] = c & 1023;
            s_xyz[tid][1] = (c >> 10) & 1023;
            s_xyz[tid][2] = (c >> 20) & 1023;
        }
    }
    __syncthreads();
    for (int q = tid; q < 189 * TB; q += NT) {  // source box of every (direction, target) pair
        const int k = q / TB, tb = q - k * TB;
        int sid = -1;
        const int e = c_dirs[o][k][3];
        if (s_tgt[tb] >= 0 && e >= a.dlo && e < a.dhi) {
            int sx = s_xyz[tb][0] + c_dirs[o][k][0], sy = s_xyz[tb][1] + c_dirs[o][k][1],
                sz = s_xyz[tb][2] + c_dirs[o][k][2];
            if (sx >= 0 && sy >= 0 && sz >= 0 && sx < S && sy < S && sz < S)
                sid = a.map[((int64_t)sx * S + sy) * S + sz];
        }
        s_sid[k * TB_MAX + tb] = sid;
    }
    __syncthreads();
    // compact the directions that reach a source box, in class order, with
    // the mask of the boxes they reach
    unsigned bm = 0;
    int pre = 0;
    if (tid < 192) {
        if (tid < 189)
            for (int tb = 0; tb < TB; tb++) bm |= (s_sid[tid * TB_MAX + tb] >= 0 ? 1u : 0u) << tb;
        const unsigned long long m = __ballot(bm != 0);
        pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_wcnt[wave] = __popcll(m);
    }
    __syncthreads();
    if (bm) {
        const int pos = pre + (wave > 0 ? s_wcnt[0] : 0) + (wave > 1 ? s_wcnt[1] : 0);
        s_eoff[pos] = c_dirs[o][tid][3] * nch;
        s_mask[pos] = (unsigned short)bm;
        for (int tb = 0; tb < TB; tb++) s_sidc[pos * TB_MAX + tb] = s_sid[tid * TB_MAX + tb];
    }
    if (tid == 0) s_nd = s_wcnt[0] + s_wcnt[1] + s_wcnt[2];
    __syncthreads();
    const int nd = __builtin_amdgcn_readfirstlane(s_nd);
    const int nsteps = nd * nch;
    if (a.accumulate && nsteps == 0) return;  // nothing to add for this tile

    const int fr = lane % 16, fq = lane / 16, rb0 = (rgroup * 4 + (wave & 3)) * RBW;
    const int cbase = (wave >> 2) * 16 * NCB;  // first column of this wave's group
    // boxes covered by each column block (wave-uniform)
    unsigned cbm[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
        const int lo = (cbase + cb * 16) / cpt, hi = min(TB - 1, (cbase + cb * 16 + 15) / cpt);
        cbm[cb] = lo < TB ? ((2u << hi) - (1u << lo)) : 0u;
    }
    const int last = nsteps - 1;
    const bool cmaj = a.chunk_major != 0;
    // (direction, chunk) of a step, advanced one step at a time
    struct Cur {
        int step, di, ch;
    };
    auto adv = [&](Cur &c) __attribute__((always_inline)) {
        c.step++;
        if (cmaj) {
            if (++c.di == nd) {
                c.di = 0;
                c.ch++;
            }
        } else if (++c.ch == nch) {
            c.ch = 0;
            c.di++;
        }
    };
    // unconditional loads, so that the compiler's wait counts stay exact:
    // rows past RE repeat row RE - 1 (never stored), empty slots and the
    // padding step read a chunk of zeros
    // table reads for a step (issued one step before the loads that use them)
    auto readM = [&](const Cur &c) __attribute__((always_inline)) {
        return c.step <= last ? (unsigned)s_mask[c.di] : 0u;
    };
    auto readE = [&](const Cur &c) __attribute__((always_inline)) {
        return c.step <= last ? s_eoff[c.di] + c.ch : 0;
    };
    auto readP = [&](const Cur &c, int (&sid)[BR]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < BR; r++) {
            const int tb = ((tid + r * NT) / 8) / cpt;
            sid[r] = tb < TB && c.step <= last ? s_sidc[c.di * TB_MAX + tb] : -1;
        }
    };
    auto loadE = [&](int eo, V8 (&ea)[RBW][NSPLIT]) __attribute__((always_inline)) {
        const T *base = (const T *)a.Es + (size_t)eo * NSPLIT * RE * 32;
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int row = min((rb0 + i) * 16 + fr, RE - 1);
#pragma unroll
            for (int p = 0; p < NSPLIT; p++) ea[i][p] = *(const V8 *)(base + ((size_t)p * RE + row) * 32 + 8 * fq);
        }
    };
    auto loadP = [&](const int (&sid)[BR], int ch, f32x4 (&pb)[BR]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < BR; r++) {
            const int u = tid + r * NT, col = u / 8, k4 = u % 8;
            const int cl = col % cpt;
            const float *src =
                sid[r] >= 0 ? a.P + ((size_t)(a.off + sid[r]) * a.ncols + c0 + cl) * a.kpad + ch * 32 : g_zero32;
            pb[r] = *(const f32x4 *)(src + 4 * k4);
        }
    };
    f32x4 acc[RBW][NCB];
    constexpr int NLO = LO ? NCB : 1;
    f32x4 accl[RBW][NLO];  // cross-product sums (LO)
#pragma unroll
    for (int i = 0; i < RBW; i++) {
#pragma unroll
        for (int j = 0; j < NCB; j++) acc[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NLO; j++) accl[i][j] = f32x4{0, 0, 0, 0};
    }
    V8 eA[RBW][NSPLIT], eB[RBW][NSPLIT];
    f32x4 pb0[BR], pb1[BR];
    Cur cm{0, 0, 0};  // next table reads: mask, E offset, proxy sources
    Cur ce = cm, cp = cm;
    auto storeP = [&](int buf, const f32x4 (&pb)[BR]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < BR; r++) {
            const int u = tid + r * NT;
            *(f32x4 *)&Bs[buf][(u / 8) * BS + 4 * (u % 8)] = pb[r];
        }
    };
    if (nsteps > 0) {  // Bs[0] = P(0); registers: E(0), P(1), P(2)
        int sid[BR];
        loadE(readE(ce), eA);
        adv(ce);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            readP(cp, sid);
            loadP(sid, cp.ch, q == 1 ? pb1 : pb0);
            adv(cp);
            if (q == 1) storeP(0, pb0);
        }
    }
    if (a.accumulate) {  // continue from the earlier groups' sums (exact power-of-two rescale)
        const float rescale = ldexpf(1.0f, xs + a.escale);
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int col = cbase + cb * 16 + fr;
            const int tb = col / cpt, cl = min(col - tb * cpt, cpt - 1);
            const int t = tb < TB ? s_tgt[tb] : -1;
            const float *src = a.Pf + ((size_t)(a.off + (t >= 0 ? t : s_tgt[0])) * a.ncols + c0 + cl) * a.ldf;
#pragma unroll
            for (int i = 0; i < RBW; i++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = (rb0 + i) * 16 + 4 * fq + v;
                    const float x = src[min(row, rf - 1)];
                    acc[i][cb][v] = t >= 0 && row < rf ? x * rescale : 0.0f;
                }
        }
    }
    // B fragment of column block cb: 32 bytes per lane from LDS
    auto readB = [&](const float *Bb, int cb, u32x4 (&r)[2]) __attribute__((always_inline)) {
        const float *bp = &Bb[(cbase + cb * 16 + fr) * BS];
        if constexpr (PRESPLIT) {  // plane p of k 8fq..8fq+7 at floats 16p + 4fq
            r[0] = *(const u32x4 *)(bp + 4 * fq);
            r[1] = *(const u32x4 *)(bp + 16 + 4 * fq);
        } else {
            r[0] = *(const u32x4 *)(bp + 8 * fq);
            r[1] = *(const u32x4 *)(bp + 8 * fq + 4);
        }
    };
    // step s (buf = s & 1): Bs[buf] holds P(s), pb holds P(s + 1)
    auto run = [&](int buf, f32x4 (&pb)[BR], V8 (&ea)[RBW][NSPLIT], V8 (&en)[RBW][NSPLIT])
        __attribute__((always_inline)) {
        // this step's table entries, read before the barrier (the tables do
        // not change) so that they arrive while the waves synchronise
        const unsigned live = readM(cm);
        adv(cm);
        const int eo = readE(ce);
        adv(ce);
        int sid[BR];
        readP(cp, sid);
        __syncthreads();  // Bs[buf] is complete and every wave is done with Bs[buf ^ 1]
        const unsigned lv = __builtin_amdgcn_readfirstlane(live);
        const float *Bb = Bs[buf];
        u32x4 rq[2][2];
        readB(Bb, 0, rq[0]);
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            if (cb + 1 < NCB) readB(Bb, cb + 1, rq[(cb + 1) & 1]);
            // the loads and stores for later steps go between the first MFMAs
            if (cb == 1) loadE(eo, en);  // the next step's E, in flight during this step
            if (cb == NCB / 2) {
                storeP(buf ^ 1, pb);
                loadP(sid, cp.ch, pb);  // three steps ahead
                adv(cp);
            }
            if (!(lv & cbm[cb])) continue;
            V8 bf[NSPLIT];
            if constexpr (PRESPLIT) {
                bf[0] = __builtin_bit_cast(V8, rq[cb & 1][0]);
                bf[1] = __builtin_bit_cast(V8, rq[cb & 1][1]);
            } else {
                const f32x4 v0 = __builtin_bit_cast(f32x4, rq[cb & 1][0]), v1 = __builtin_bit_cast(f32x4, rq[cb & 1][1]);
                float x[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (int e = 0; e < 8; e++) x[e] *= xscale;
#pragma unroll
                for (int p = 0; p < NSPLIT; p++)
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const T h = (T)x[e];
                        bf[p][e] = h;
                        x[e] -= (float)h;
                    }
            }
#pragma unroll
            for (int i = 0; i < RBW; i++) {  // smallest terms first
                if constexpr (LO) {
                    accl[i][cb % NLO] = mfma16(ea[i][1], bf[0], accl[i][cb % NLO]);
                    accl[i][cb % NLO] = mfma16(ea[i][0], bf[1], accl[i][cb % NLO]);
                } else {
                    if constexpr (NSPLIT == 3) {
                        acc[i][cb] = mfma16(ea[i][2], bf[0], acc[i][cb]);
                        acc[i][cb] = mfma16(ea[i][1], bf[1], acc[i][cb]);
                        acc[i][cb] = mfma16(ea[i][0], bf[2], acc[i][cb]);
                    }
                    acc[i][cb] = mfma16(ea[i][1], bf[0], acc[i][cb]);
                    acc[i][cb] = mfma16(ea[i][0], bf[1], acc[i][cb]);
                }
                acc[i][cb] = mfma16(ea[i][0], bf[0], acc[i][cb]);
            }
        }
    };
    // an even number of steps: the padding step (zero proxies, empty mask) adds nothing
    for (int step = 0; step < nsteps; step += 2) {
        run(0, pb1, eA, eB);
        run(1, pb0, eB, eA);
    }
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
        const int col = cbase + cb * 16 + fr;
        const int tb = col / cpt, cl = col - tb * cpt;
        if (tb >= TB) continue;
        const int t = s_tgt[tb];
        if (t < 0) continue;
        float *dst = a.Pf + ((size_t)(a.off + t) * a.ncols + c0 + cl) * a.ldf;
#pragma unroll
        for (int i = 0; i < RBW; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = (rb0 + i) * 16 + 4 * fq + v;
                const float x = LO ? acc[i][cb][v] + accl[i][cb % NLO][v] : acc[i][cb][v];
                if (row < rf) dst[row] = x * unscale;
            }
    }
}

// ---------------------------------------------------------------------------
// The proxy (VT) and expansion (U) products of the mixed modes on the BF16
// matrix cores with three-part operands (6 products, about FP32 accuracy;
// P and Pf are FP32 in these modes):  C[M x ncol] = alpha * A B.
//   A: M x K operator, pre-split once per operator set (split_a3) into
//      [3 planes][R rows][Kp] BF16 (row-major, zero-padded to R = 16*RB rows
//      and Kp = 32*nch columns);
//   B: K x ncol, column-major (ld ldb), FP64 or FP32, split in registers.
// A workgroup of 8 waves owns 128*NCBW columns (NCBW 16-column blocks per
// wave; one for more than 12 row blocks, to stay within the registers) and
// all RB row blocks.  A goes through LDS one 32-deep k chunk at a time
// (double-buffered); B is gathered coalesced (32 consecutive k of two
// columns per load instruction), converted to FP32 into LDS, and each wave
// splits its fragments in registers.  The next chunks of A and B are loaded
// one chunk ahead.
__global__ void split_a3(const double *__restrict__ A, int M, int K, int R, int Kp, __bf16 *__restrict__ As)
{
    const int64_t per = (int64_t)R * Kp, total = 3 * per;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(q / per), rem = (int)(q - p * per), row = rem / Kp, k = rem - row * Kp;
        double x = (row < M && k < K) ? A[row + (int64_t)k * M] : 0.0;
        __bf16 part = (__bf16)0.0f;
        for (int i = 0; i <= p; i++) {
            part = (__bf16)(float)x;
            x -= (double)(float)part;
        }
        As[q] = part;
    }
}

void launch_split_a3(const double *A, int M, int K, int R, int Kp, void *As, hipStream_t st)
{
    const int64_t total = 3LL * R * Kp;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    split_a3<<<blocks, 256, 0, st>>>(A, M, K, R, Kp, (__bf16 *)As);
    check(hipGetLastError(), "split_a3");
}

template <typename TBt, typename TC, int RB, int NCBW>
__global__ __launch_bounds__(512) void gemm_split3(GemmSplitArgs g)
{
    constexpr int NT = 512, W = 128 * NCBW, R = 16 * RB, KS = 40, BST = 36;
    constexpr int NA = 3 * R * 4;           // 16-byte units of one A chunk
    constexpr int AR = (NA + NT - 1) / NT;
    constexpr int BRW = W * 32 / NT;        // B elements per thread and chunk
    __shared__ __attribute__((aligned(16))) __bf16 As[2][3 * R * KS];
    __shared__ __attribute__((aligned(16))) float Bs[W * BST];  // [column][32 k + pad], FP32
    const TBt *B = (const TBt *)g.B;
    TC *C = (TC *)g.C;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t n0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * W;
    const int nch = g.nch, fr = lane % 16, fq = lane / 16;
    u32x4 ra[AR];
    auto loadA = [&](int ch) __attribute__((always_inline)) {
        const u32x4 *src = (const u32x4 *)g.As;  // [3][R][Kp] bf16: 16-byte unit (p, row, q) at (p*R + row)*nch*4 + ch*4 + q
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int u = min(tid + r * NT, NA - 1), pr = u / 4, q = u % 4;
            ra[r] = src[(int64_t)pr * nch * 4 + ch * 4 + q];
        }
    };
    auto storeA = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int u = tid + r * NT;
            if (AR * NT == NA || u < NA) *(u32x4 *)&As[buf][(u / 4) * KS + 8 * (u % 4)] = ra[r];
        }
    };
    // B chunk (32 k x W columns), coalesced: element e = tid + r*NT is k = e % 32
    // of tile column e / 32, so one load instruction covers two columns' 32
    // consecutive k values.  Addresses are clamped, values outside zeroed.
    TBt rb[BRW];
    auto loadB = [&](int ch) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < BRW; r++) {
            const int e = tid + r * NT, k = ch * 32 + (e & 31);
            const int64_t col = n0 + (e >> 5);
            rb[r] = B[min(col, g.ncol - 1) * g.ldb + min(k, g.K - 1)];
        }
    };
    auto storeB = [&](int ch) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < BRW; r++) {
            const int e = tid + r * NT, k = ch * 32 + (e & 31);
            const bool ok = n0 + (e >> 5) < g.ncol && k < g.K;
            Bs[(e >> 5) * BST + (e & 31)] = ok ? (float)rb[r] : 0.0f;
        }
    };
    f32x4 acc[RB][NCBW];
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < NCBW; j++) acc[i][j] = f32x4{0, 0, 0, 0};
    loadA(0);
    storeA(0);
    loadA(min(1, nch - 1));
    loadB(0);
    storeB(0);
    loadB(min(1, nch - 1));
    __syncthreads();
    for (int ch = 0; ch < nch; ch++) {
        const int buf = ch & 1;
        // this wave's B fragments (NCBW 16-column blocks), split into 3 BF16 parts
        bf16x8 bfr[NCBW][3];
#pragma unroll
        for (int cb = 0; cb < NCBW; cb++) {
            const float *bp = &Bs[((wave * NCBW + cb) * 16 + fr) * BST + 8 * fq];
            const f32x4 v0 = *(const f32x4 *)bp, v1 = *(const f32x4 *)(bp + 4);
            float x[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const __bf16 h = (__bf16)x[e];
                    bfr[cb][p][e] = h;
                    x[e] -= (float)h;
                }
        }
        storeA(buf ^ 1);  // chunk ch + 1 (buffer last read in chunk ch - 1)
        loadA(min(ch + 2, nch - 1));
        __syncthreads();  // every wave holds its B fragments of chunk ch
        storeB(min(ch + 1, nch - 1));
        loadB(min(ch + 2, nch - 1));
#pragma unroll
        for (int rb_ = 0; rb_ < RB; rb_++) {
            bf16x8 afr[3];
#pragma unroll
            for (int p = 0; p < 3; p++) afr[p] = *(const bf16x8 *)&As[buf][((p * R) + rb_ * 16 + fr) * KS + 8 * fq];
#pragma unroll
            for (int cb = 0; cb < NCBW; cb++) {  // smallest terms first
                acc[rb_][cb] = mfma16(afr[2], bfr[cb][0], acc[rb_][cb]);
                acc[rb_][cb] = mfma16(afr[1], bfr[cb][1], acc[rb_][cb]);
                acc[rb_][cb] = mfma16(afr[0], bfr[cb][2], acc[rb_][cb]);
                acc[rb_][cb] = mfma16(afr[1], bfr[cb][0], acc[rb_][cb]);
                acc[rb_][cb] = mfma16(afr[0], bfr[cb][1], acc[rb_][cb]);
                acc[rb_][cb] = mfma16(afr[0], bfr[cb][0], acc[rb_][cb]);
            }
        }
        __syncthreads();  // B of chunk ch + 1 and A buffer buf ^ 1 complete; As[buf] free
    }
#pragma unroll
    for (int cb = 0; cb < NCBW; cb++) {
        const int64_t gc = n0 + (wave * NCBW + cb) * 16 + fr;
        if (gc >= g.ncol) continue;
        TC *dst = C + gc * g.ldc;
#pragma unroll
        for (int rb_ = 0; rb_ < RB; rb_++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rb_ * 16 + 4 * fq + v;
                if (row < g.M) dst[row] = (TC)(g.alpha * (double)acc[rb_][cb][v]);
                else if (row < g.ldc) dst[row] = (TC)0;
            }
        for (int row = R; row < g.ldc; row++)  // padding rows beyond the MFMA tile
            if (fq == 0) dst[row] = (TC)0;
    }
}

// ---------------------------------------------------------------------------
// The same product for thin shapes (few rows, or a short k): one wave per
// 32 columns, no LDS and no barriers.  Each wave loads its columns' B
// fragments straight from memory one k chunk at a time, splits them into
// three BF16 parts in registers and runs the RB row blocks against A
// fragments read from the pre-split operator (L2-resident).  The proxy
// product of the re-compressed deep levels (M = 32) and their expansion
// product (K = 32) move far more B or C than A, which the workgroup tile
// above would stage through LDS for nothing.
template <typename TBt, typename TC, int RB>
__global__ __launch_bounds__(256) void gemm_thin3(GemmSplitArgs g)
{
    constexpr int NCB = 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, fr = lane % 16, fq = lane / 16;
    const int64_t n0 = ((int64_t)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave) * 16 * NCB;
    if (n0 >= g.ncol) return;
    const TBt *B = (const TBt *)g.B;
    TC *C = (TC *)g.C;
    const bf16x8 *As = (const bf16x8 *)g.As;  // [3][16*RB][Kp] bf16, 8 at a time
    const int Kp = 32 * g.nch, R = 16 * RB;
    f32x4 acc[RB][NCB];
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < NCB; j++) acc[i][j] = f32x4{0, 0, 0, 0};
    for (int ch = 0; ch < g.nch; ch++) {
        bf16x8 b[NCB][3];
#pragma unroll
        for (int j = 0; j < NCB; j++) {
            const int64_t col = min(n0 + j * 16 + fr, g.ncol - 1);
            const TBt *src = B + col * g.ldb;
            const int k0 = ch * 32 + 8 * fq;
            float x[8];
            if (k0 + 8 <= g.K && (g.ldb % 4) == 0) {  // whole 16-byte groups (the columns start 16-byte aligned)
                if constexpr (sizeof(TBt) == 4) {
                    const float4 v0 = *(const float4 *)(src + k0), v1 = *(const float4 *)(src + k0 + 4);
                    x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
                    x[4] = v1.x; x[5] = v1.y; x[6] = v1.z; x[7] = v1.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; e += 2) {
                        const double2 v = *(const double2 *)(src + k0 + e);
                        x[e] = (float)v.x;
                        x[e + 1] = (float)v.y;
                    }
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) x[e] = k0 + e < g.K ? (float)src[k0 + e] : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const __bf16 h0 = (__bf16)x[e];
                const float r1 = x[e] - (float)h0;
                const __bf16 h1 = (__bf16)r1;
                b[j][0][e] = h0;
                b[j][1][e] = h1;
                b[j][2][e] = (__bf16)(r1 - (float)h1);
            }
        }
#pragma unroll
        for (int i = 0; i < RB; i++) {
            const size_t at = ((size_t)(i * 16 + fr) * Kp + ch * 32 + 8 * fq) / 8;
            const bf16x8 a0 = As[at], a1 = As[at + (size_t)R * Kp / 8], a2 = As[at + (size_t)2 * R * Kp / 8];
#pragma unroll
            for (int j = 0; j < NCB; j++) {  // smallest terms first
                acc[i][j] = mfma16(a2, b[j][0], acc[i][j]);
                acc[i][j] = mfma16(a1, b[j][1], acc[i][j]);
                acc[i][j] = mfma16(a0, b[j][2], acc[i][j]);
                acc[i][j] = mfma16(a1, b[j][0], acc[i][j]);
                acc[i][j] = mfma16(a0, b[j][1], acc[i][j]);
                acc[i][j] = mfma16(a0, b[j][0], acc[i][j]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NCB; j++) {
        const int64_t col = n0 + j * 16 + fr;
        if (col >= g.ncol) continue;
        TC *dst = C + col * g.ldc;
        const bool vec = (g.ldc % 4) == 0;  // FP32: 16-byte stores of 4 rows
#pragma unroll
        for (int i = 0; i < RB; i++) {
            const int row = i * 16 + 4 * fq;
            if constexpr (sizeof(TC) == 4) {
                if (vec && row + 4 <= g.M) {
                    float4 o;
                    o.x = (float)(g.alpha * (double)acc[i][j][0]);
                    o.y = (float)(g.alpha * (double)acc[i][j][1]);
                    o.z = (float)(g.alpha * (double)acc[i][j][2]);
                    o.w = (float)(g.alpha * (double)acc[i][j][3]);
                    *(float4 *)(dst + row) = o;
                    continue;
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (row + v < g.M) dst[row + v] = (TC)(g.alpha * (double)acc[i][j][v]);
        }
    }
}

// thin shapes (at most 64 rows, or one or two k chunks): the wave-per-32-columns
// kernel; RB row blocks of 16
template <typename TBt, typename TC>
static void gemm_thin_t(const GemmSplitArgs &g, int rb, hipStream_t st)
{
    dim3 grid((unsigned)((g.ncol + 127) / 128));
#define PBB_T3(RB)                                                        \
    case RB: gemm_thin3<TBt, TC, RB><<<grid, 256, 0, st>>>(g); break;
    switch (rb) {
        PBB_T3(1) PBB_T3(2) PBB_T3(3) PBB_T3(4) PBB_T3(5) PBB_T3(6) PBB_T3(7) PBB_T3(8) PBB_T3(9) PBB_T3(10)
        PBB_T3(11) PBB_T3(12) PBB_T3(13) PBB_T3(14)
    default: fail(PBB_ERR_UNSUPPORTED, "split GEMM supports at most 224 rows");
    }
#undef PBB_T3
}

void launch_gemm_split3(const GemmSplitArgs &g, hipStream_t st)
{
    if (g.ncol <= 0) return;
    if (g.ldb < g.K || g.ldc < g.M) fail(PBB_ERR_INTERNAL, "gemm: leading dimension too small");
    if (g.M <= 64 || g.nch <= 2) {
        const int rb = (g.M + 15) / 16;
        if (!g.b_f32 && g.c_f32) gemm_thin_t<double, float>(g, rb, st);
        else if (g.b_f32 && !g.c_f32) gemm_thin_t<float, double>(g, rb, st);
        else if (g.b_f32 && g.c_f32) gemm_thin_t<float, float>(g, rb, st);
        else fail(PBB_ERR_INTERNAL, "gemm_thin3: unsupported operand types");
        check(hipGetLastError(), "gemm_thin3");
        return;
    }
    const int rb = (g.M + 15) / 16, w = rb > 12 ? 128 : 256;
    dim3 grid((unsigned)((g.ncol + w - 1) / w));
#define PBB_G3(RB)                                                                                   \
    case RB:                                                                                         \
        if (!g.b_f32 && g.c_f32) gemm_split3<double, float, RB, RB <= 12 ? 2 : 1><<<grid, 512, 0, st>>>(g); \
        else if (g.b_f32 && !g.c_f32) gemm_split3<float, double, RB, RB <= 12 ? 2 : 1><<<grid, 512, 0, st>>>(g); \
        else if (g.b_f32 && g.c_f32) gemm_split3<float, float, RB, RB <= 12 ? 2 : 1><<<grid, 512, 0, st>>>(g); \
        else fail(PBB_ERR_INTERNAL, "gemm_split3: unsupported operand types");                       \
        break;
    switch (rb) {
        PBB_G3(1) PBB_G3(2) PBB_G3(3) PBB_G3(4) PBB_G3(5) PBB_G3(6) PBB_G3(7) PBB_G3(8) PBB_G3(9) PBB_G3(10)
        PBB_G3(11) PBB_G3(12) PBB_G3(13) PBB_G3(14)
    default: fail(PBB_ERR_UNSUPPORTED, "split GEMM supports at most 224 rows");
    }
#undef PBB_G3
    check(hipGetLastError(), "gemm_split3");
}

// ---------------------------------------------------------------------------
template <typename T, int NSPLIT>
static void split_e_t(const double *E, int rf, int rs, int R, int nch, int escale, void *Es, hipStream_t st)
{
    int64_t total = 316LL * nch * NSPLIT * R * 32;
    int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    split_e<T, NSPLIT><<<blocks, 256, 0, st>>>(E, rf, rs, R, nch, escale, (T *)Es);
    check(hipGetLastError(), "split_e");
}

void launch_split_e(int variant, const double *E, int rf, int rs, int R, int nch, int escale, void *Es,
                    hipStream_t st)
{
    switch (variant) {
    case PBB_PREC_MIXED: split_e_t<__bf16, 3>(E, rf, rs, R, nch, escale, Es, st); break;
    case PBB_PREC_MIXED_F16X2: split_e_t<_Float16, 2>(E, rf, rs, R, nch, escale, Es, st); break;
    case PBB_PREC_MIXED_BF16X2: split_e_t<__bf16, 2>(E, rf, rs, R, nch, escale, Es, st); break;
    default: fail(PBB_ERR_INVALID_ARG, "unknown mixed-precision variant");
    }
}

void upload_dirs_mixed(const int dirs[8][189][4])
{
    check(hipMemcpyToSymbol(HIP_SYMBOL(c_dirs), dirs, sizeof(int) * 8 * 189 * 4), "hipMemcpyToSymbol(c_dirs)");
}

__global__ void zero_u32(unsigned *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

void launch_zero_u32(unsigned *p, int n, hipStream_t st)
{
    if (n <= 0) return;
    zero_u32<<<(n + 255) / 256, 256, 0, st>>>(p, n);
    check(hipGetLastError(), "zero_u32");
}

void launch_absmax_rows(const float *x, int64_t rows, int ld, int used, unsigned *out, hipStream_t st)
{
    const int64_t n = rows * used;
    if (n <= 0) return;
    if (used > ld) fail(PBB_ERR_INTERNAL, "absmax: rows wider than their stride");
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    absmax_f32<<<blocks, 256, 0, st>>>(x, rows, ld, used, out);
    check(hipGetLastError(), "absmax_f32");
}

// 16-block row-wave groups spill registers for the widest row slices
template <int NSPLIT, int RBW>
constexpr bool rows_wide_ok()
{
    return !(RBW == 4 || (NSPLIT == 3 && RBW >= 3));
}

// row blocks per wave of the row-wave M2L: one workgroup takes up to 16
// row blocks (256 rows); larger ranks run as row groups (gridDim.z) of at
// most 12 row blocks each
static int rows_rbw(int R)
{
    const int rball = R / 16;
    if (rball <= 16) return (rball + 3) / 4;
    const int ngroups = (rball + 11) / 12;
    return ((rball + ngroups - 1) / ngroups + 3) / 4;
}

int m2l_rows_max_width(int variant, int R)
{
    const int rbw = rows_rbw(R);
    const bool three = variant == PBB_PREC_MIXED;
    return rbw >= 4 || (three && rbw >= 3) ? 128 : 256;
}

template <typename T, int NSPLIT, int RBW>
static void m2l_rows_launch(const M2LMixArgs &a, dim3 grid, int w, hipStream_t st)
{
    if (a.split_acc) {  // 64-column tiles: the extra sums fit the two-waves-per-SIMD budget
        if constexpr (NSPLIT == 2) {
            if (w > 64) fail(PBB_ERR_INTERNAL, "separate cross-product sums take tiles of at most 64 columns");
            m2l_mfma_rows<T, NSPLIT, RBW, 4, 1, true><<<grid, 256, 0, st>>>(a);
            return;
        } else {
            fail(PBB_ERR_INTERNAL, "separate cross-product sums are for the two-part modes");
        }
    }
    if constexpr (rows_wide_ok<NSPLIT, RBW>()) {
        if (w > 128) m2l_mfma_rows<T, NSPLIT, RBW, 8, 2><<<grid, 512, 0, st>>>(a);
        else if (w > 64) m2l_mfma_rows<T, NSPLIT, RBW, 8, 1><<<grid, 256, 0, st>>>(a);
        else m2l_mfma_rows<T, NSPLIT, RBW, 4, 1><<<grid, 256, 0, st>>>(a);
    } else {
        if (w > 128) fail(PBB_ERR_INTERNAL, "row-wave M2L tile too wide for this rank");
        if (w > 64) m2l_mfma_rows<T, NSPLIT, RBW, 4, 2><<<grid, 512, 0, st>>>(a);
        else m2l_mfma_rows<T, NSPLIT, RBW, 4, 1><<<grid, 256, 0, st>>>(a);
    }
}

template <typename T, int NSPLIT>
static void m2l_rows_t(const M2LMixArgs &a, int ntiles, int ncolchunks, hipStream_t st)
{
    const int rbw = rows_rbw(a.R);
    const int ngroups = (a.R / 16 + 4 * rbw - 1) / (4 * rbw);
    const int w = a.TB * a.cpt;
    if (w > 256) fail(PBB_ERR_INTERNAL, "row-wave M2L tiles are at most 256 columns wide");
    dim3 grid((unsigned)ntiles, (unsigned)ncolchunks, (unsigned)ngroups);
    switch (rbw) {
    case 1: m2l_rows_launch<T, NSPLIT, 1>(a, grid, w, st); break;
    case 2: m2l_rows_launch<T, NSPLIT, 2>(a, grid, w, st); break;
    case 3: m2l_rows_launch<T, NSPLIT, 3>(a, grid, w, st); break;
    case 4: m2l_rows_launch<T, NSPLIT, 4>(a, grid, w, st); break;
    default: fail(PBB_ERR_INTERNAL, "row-wave M2L row group too large");
    }
}

template <typename T, int NSPLIT>
static void m2l_split_t(const M2LMixArgs &a, int ntiles, int ncolchunks, hipStream_t st)
{
    if (a.k16) {  // 16-deep steps, column-wave tiles, all rows (at most 32) in one workgroup
        if constexpr (NSPLIT == 2 && sizeof(T) == 2 && !std::is_same<T, __bf16>::value) {
            if (a.nch != 1 || a.rf > 32) fail(PBB_ERR_INTERNAL, "16-deep M2L steps need ks <= 16 and kf <= 32");
            const int w = a.TB * a.cpt;
            if (w > 256) fail(PBB_ERR_INTERNAL, "mixed-precision M2L tile wider than 256 columns");
            dim3 grid((unsigned)ntiles, (unsigned)ncolchunks, 1);
#define PBB_K16(RB)                                                                        \
    if (w > 128) m2l_mfma_split<T, 2, RB, 4, 16><<<grid, 256, 0, st>>>(a);                 \
    else if (w > 64) m2l_mfma_split<T, 2, RB, 2, 16><<<grid, 256, 0, st>>>(a);             \
    else m2l_mfma_split<T, 2, RB, 1, 16><<<grid, 256, 0, st>>>(a);
            if (a.rf <= 16) {
                PBB_K16(1)
            } else {
                PBB_K16(2)
            }
#undef PBB_K16
            return;
        } else {
            fail(PBB_ERR_INTERNAL, "16-deep M2L steps are for the f16 x2 mode");
        }
    }
    if (a.wave_rows) {
        m2l_rows_t<T, NSPLIT>(a, ntiles, ncolchunks, st);
        return;
    }
    const int rball = (a.rf + 15) / 16, halves = a.row_split ? 2 : 1, rb = (rball + halves - 1) / halves;
    const int w = a.TB * a.cpt;  // tile width: 64, 128 or 256 columns
    if (w > 256) fail(PBB_ERR_INTERNAL, "mixed-precision M2L tile wider than 256 columns");
    dim3 grid((unsigned)ntiles, (unsigned)ncolchunks, (unsigned)((rball + rb - 1) / rb));
#define PBB_CASE(RB)                                                                   \
    case RB:                                                                           \
        if (w > 128) m2l_mfma_split<T, NSPLIT, RB, 4><<<grid, 256, 0, st>>>(a);        \
        else if (w > 64) m2l_mfma_split<T, NSPLIT, RB, 2><<<grid, 256, 0, st>>>(a);    \
        else m2l_mfma_split<T, NSPLIT, RB, 1><<<grid, 256, 0, st>>>(a);                \
        break;
    switch (rb) {
        PBB_CASE(1) PBB_CASE(2) PBB_CASE(3) PBB_CASE(4) PBB_CASE(5) PBB_CASE(6) PBB_CASE(7) PBB_CASE(8)
        PBB_CASE(9) PBB_CASE(10) PBB_CASE(11) PBB_CASE(12) PBB_CASE(13) PBB_CASE(14)
    default: fail(PBB_ERR_UNSUPPORTED, "mixed-precision M2L supports ranks up to 224");
    }
#undef PBB_CASE
}

void launch_m2l_mixed(int variant, const M2LMixArgs &a, int ntiles, int ncolchunks, hipStream_t st)
{
    if (ntiles <= 0) return;
    switch (variant) {
    case PBB_PREC_MIXED: m2l_split_t<__bf16, 3>(a, ntiles, ncolchunks, st); break;
    case PBB_PREC_MIXED_F16X2: m2l_split_t<_Float16, 2>(a, ntiles, ncolchunks, st); break;
    case PBB_PREC_MIXED_BF16X2: m2l_split_t<__bf16, 2>(a, ntiles, ncolchunks, st); break;
    default: fail(PBB_ERR_INVALID_ARG, "unknown mixed-precision variant");
    }
    check(hipGetLastError(), "m2l_mfma_split");
}

// ---------------------------------------------------------------------------
// Proxy passes of the levels with operator ranks of at most 32 (mixed modes):
//   p2m_proxy  P[b][c][r] = sum_{p in b} w_p[c] y_p[r],   y_p = VT' S_n(x_p)
//   l2p_proxy  lp[p][c] (+)= alpha sum_r z_p[r] Pf[b][c][r],   z_p = U'^T S_n(x_p)
// By associativity these equal the GEMM route (S_b = sum_p S_n(x_p) w_p^T,
// P_b = VT' S_b; F_b = U' Pf_b, value = S_n(x_p)^T F_b), without the n^3 x
// ncols expansions S and F of the level's boxes, which at a few points per
// box are several times the points' own data.
// Range k holds the boxes whose first point lies in [k*PX_PT, (k+1)*PX_PT);
// persistent workgroups walk ranges k = blockIdx.x, + gridDim.x, ..., each
// in passes of PX_PT points.
// The rank vectors of a pass (PX_PT points x 32 ranks, K = n^3) run on the
// BF16 matrix cores in three parts per operand (6 products, FP32
// accumulation), as the expansion route's proxy and expansion products do:
// A = S_n of 16 points (products of the per-axis weights, split in
// registers), B = the operator, split once per operator set on the host into
// [part][rank][k] (L2-resident, read straight into the fragments).
// The box sums (P2M) and the Pf products (L2P) run in FP64 on the VALU.
constexpr int PX_VS = PX_K + 4;  // LDS stride of a point's rank vector (16-byte aligned rows)
constexpr int PX_HALF = 64;      // points whose S_n tables are staged at once (4 waves x 16)

template <int N>
struct PxDims {
    static constexpr int N3 = N * N * N, KC = (N3 + 31) / 32, KP = 32 * KC;
    static constexpr int KS = KP + 8;  // row stride of the split operator (bf16): conflict-free 16-byte reads
    static constexpr int ZS = N + 1;                  // stride of one axis's S_n in the pass's table
    static constexpr int TAB = PX_PT * 3 * ZS;       // floats of the table
};

// The boxes [bA, bB) of range k: bA, the first box with pstart >= k*PX_PT,
// by a 256-way search (every thread probes one position per round, a
// handful of rounds), then bB - bA, the boxes that start before
// (k+1)*PX_PT (at most PX_PT), counted in one more round.  Block-uniform.
__device__ __forceinline__ void px_range(const ProxyArgs &a, int64_t k, int &bA, int &bB)
{
    const int64_t t0 = k * PX_PT, t1 = t0 + PX_PT;
    int lo = 0, hi = a.nb;  // bA lies in [lo, hi]; pstart[nb] = N
    while (lo < hi) {
        const int step = (hi - lo + 255) / 256;
        const int64_t q = (int64_t)lo + (int64_t)threadIdx.x * step;
        const int c = __syncthreads_count(q < hi && (int64_t)a.pstart[q] < t0);
        const int nlo = c > 0 ? lo + (c - 1) * step + 1 : lo;
        const int64_t qc = (int64_t)lo + (int64_t)c * step;  // the first probe at or past t0, if any
        hi = c > 0 ? (qc < hi ? (int)qc : hi) : lo;
        lo = nlo;
    }
    bA = lo;
    const int b = bA + (int)threadIdx.x;
    bB = bA + __syncthreads_count(threadIdx.x < PX_PT && b < a.nb && (int64_t)a.pstart[b] < t1);
}

// The boxes of a pass: s_ps[j] = first point of box jb + j (j <= PX_PT, the
// entry past the range's last box is its end); returns how many start
// before sub_end.  Includes the barrier after which s_ps may be read.
__device__ __forceinline__ int px_pass_boxes(const ProxyArgs &a, int *s_ps, int jb, int bB, int sub_end)
{
    const int tid = threadIdx.x;
    int pred = 0;
    if (tid <= PX_PT) {
        const int v = a.pstart[min(jb + tid, bB)];
        s_ps[tid] = v;
        pred = tid < PX_PT && jb + tid < bB && v < sub_end;
    }
    return __syncthreads_count(pred);
}

// Points [base, base + np) of a pass, whose boxes are jb + 0..nbx-1 with
// first points s_ps[0..nbx-1]: vs[pp][r] = sum_i Ob[i][r] S_n,i(x_{base+pp}).
// All threads first find each point's box and evaluate S_n along each axis
// (FP64 Clenshaw, one (point, axis) per thread) into tab[pp][axis][m]; then,
// per half of PX_HALF points, wave w forms the 16 x 32 block of points
// 16w..16w+15 on the matrix cores, with A = S_x S_y S_z (FP32 products)
// split into three BF16 parts.  Ends with a barrier after which vs is
// complete and tab is free for the caller.
template <int N>
__device__ __forceinline__ void px_point_vec(const ProxyArgs &a, const __bf16 *Ob, const int *s_ps, int nbx, int jb,
                                             int base, int np, float *vs, float *tab, int *s_bx)
{
    typedef PxDims<N> D;
    typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pl = lane & 15, g = lane >> 4;
    if (tid < np) {  // the last box of the pass that starts at or before the point
        const int q = base + tid;
        int lo = 0, hi = nbx - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_ps[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        s_bx[tid] = lo;
    }
    __syncthreads();
    const double ihalfL = 2.0 / a.len;
    for (int t = tid; t < 3 * PX_PT; t += blockDim.x) {
        const int pp = t % PX_PT, ax = t / PX_PT;
        float *dst = tab + (pp * 3 + ax) * D::ZS;
        if (pp < np) {
            const uint32_t cc = (uint32_t)a.coord[jb + s_bx[pp]];
            const double *xa = ax == 0 ? a.x : ax == 1 ? a.y : a.z;
            double S[N];
            sn_eval_n<N>(a.Tkz, ihalfL * (xa[base + pp] - a.ctab[(cc >> (10 * ax)) & 1023]), S);
#pragma unroll
            for (int m = 0; m < N; m++) dst[m] = (float)S[m];
        } else {  // points past the pass: zeros (their results are not read)
#pragma unroll
            for (int m = 0; m < N; m++) dst[m] = 0.0f;
        }
    }
    __syncthreads();
    for (int h0 = 0; h0 < PX_PT; h0 += PX_HALF) {
        if (h0 >= np) break;  // block-uniform
        const int pt = h0 + 16 * wave + pl;  // this lane's point (row of A)
        const float *sx = tab + (pt * 3) * D::ZS, *sy = sx + D::ZS, *sz = sy + D::ZS;
        f32x4 d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};
#pragma unroll 1
        for (int kc = 0; kc < D::KC; kc++) {
            bf16x8_t a0, a1, a2;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int k = 32 * kc + 8 * g + e;
                float v = 0.0f;
                if (k < D::N3) {
                    const int l12 = k / N, l3 = k - l12 * N, l1 = l12 / N, l2 = l12 - l1 * N;
                    v = sx[l1] * sy[l2] * sz[l3];
                }
                const __bf16 p0 = (__bf16)v;
                v -= (float)p0;
                const __bf16 p1 = (__bf16)v;
                v -= (float)p1;
                a0[e] = p0;
                a1[e] = p1;
                a2[e] = (__bf16)v;
            }
            const __bf16 *ob = Ob + (size_t)pl * D::KS + 32 * kc + 8 * g;
#pragma unroll
            for (int rb = 0; rb < 2; rb++) {
                const bf16x8_t b0 = *(const bf16x8_t *)(ob + (0 * PX_K + 16 * rb) * D::KS);
                const bf16x8_t b1 = *(const bf16x8_t *)(ob + (1 * PX_K + 16 * rb) * D::KS);
                const bf16x8_t b2 = *(const bf16x8_t *)(ob + (2 * PX_K + 16 * rb) * D::KS);
                f32x4 d = rb ? d1 : d0;
                d = mfma16(a2, b0, d);  // smallest terms first
                d = mfma16(a1, b1, d);
                d = mfma16(a0, b2, d);
                d = mfma16(a1, b0, d);
                d = mfma16(a0, b1, d);
                d = mfma16(a0, b0, d);
                if (rb) d1 = d;
                else d0 = d;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float *row = vs + (size_t)(h0 + 16 * wave + 4 * g + v) * PX_VS;
            row[pl] = d0[v];
            row[16 + pl] = d1[v];
        }
    }
    __syncthreads();  // vs complete, tab free
}

template <int N>
__global__ __launch_bounds__(256) void p2m_proxy(ProxyArgs a)
{
    typedef PxDims<N> D;
    constexpr int SWS = PX_K + 1;
    constexpr int UNI = D::TAB > PX_PT * SWS ? D::TAB : PX_PT * SWS;
    __shared__ __attribute__((aligned(16))) float vs[PX_PT * PX_VS];
    __shared__ float uni[UNI];  // S_n tables of the rank-vector phase, then the pass's weights
    __shared__ int s_ps[PX_PT + 1];
    __shared__ int s_bx[PX_PT];
    float *sw = uni;
    const int tid = threadIdx.x;
    const __bf16 *Ob = (const __bf16 *)a.ob;
    const int64_t nranges = (a.npts + PX_PT - 1) / PX_PT;
    const int c = tid >> 3, rq = tid & 7;  // output phase: column c, ranks 4rq..4rq+3
    float pmax = 0.0f;                     // max |P| stored by this thread
    for (int64_t k = blockIdx.x; k < nranges; k += gridDim.x) {
        int bA, bB;
        px_range(a, k, bA, bB);
        if (bA >= bB) continue;  // no box starts in this range
        const int p_begin = a.pstart[bA], p_end = a.pstart[bB];
        for (int c0 = 0; c0 < a.ncols; c0 += PX_K) {
            const int cn = min(PX_K, a.ncols - c0);
            int jb = bA;
            double acc[4] = {0.0, 0.0, 0.0, 0.0};  // the current box's sums (FP64), carried across passes
            for (int base = p_begin; base < p_end; base += PX_PT) {
                const int sub_end = min(base + PX_PT, p_end), np = sub_end - base;
                __syncthreads();  // the previous pass's readers are done
                const int nbx = px_pass_boxes(a, s_ps, jb, bB, sub_end);
                px_point_vec<N>(a, Ob, s_ps, nbx, jb, base, np, vs, uni, s_bx);
                for (int q = tid; q < PX_PT * PX_K; q += 256) {
                    const int pp = q / PX_K, cl = q - pp * PX_K;
                    sw[pp * SWS + cl] =
                        pp < np && cl < cn ? (float)a.wsT[(int64_t)(base + pp) * a.ncols + c0 + cl] : 0.0f;
                }
                __syncthreads();
                int done = 0;
                for (int j = 0; j < nbx; j++) {
                    const int lo = max(s_ps[j], base) - base, hi = min(s_ps[j + 1], sub_end) - base;
#pragma unroll 4
                    for (int qq = lo; qq < hi; qq++) {
                        const double w = sw[qq * SWS + c];
                        const float4 y = *(const float4 *)&vs[qq * PX_VS + 4 * rq];
                        acc[0] += w * y.x;
                        acc[1] += w * y.y;
                        acc[2] += w * y.z;
                        acc[3] += w * y.w;
                    }
                    if (s_ps[j + 1] <= sub_end) {  // the box ends in this pass
                        if (c < cn && 4 * rq < a.ldp) {  // rows past the stride are zero (ranks past the kept ones)
                            const float4 v = float4{(float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]};
                            *(float4 *)&a.P[((a.off + jb + j) * a.ncols + c0 + c) * a.ldp + 4 * rq] = v;
                            pmax = fmaxf(pmax, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
                        }
                        acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
                        done++;
                    }
                }
                jb += done;
            }
        }
    }
    if (a.xmax) {  // the level's scale for the two-part FP16 proxies
        for (int o = 32; o > 0; o >>= 1) pmax = fmaxf(pmax, __shfl_xor(pmax, o));
        if ((tid & 63) == 0) atomicMax(a.xmax, __float_as_uint(pmax));
    }
}

template <int N>
__global__ __launch_bounds__(256) void l2p_proxy(ProxyArgs a)
{
    typedef PxDims<N> D;
    __shared__ __attribute__((aligned(16))) float vs[PX_PT * PX_VS];
    __shared__ float tab[D::TAB];
    __shared__ int s_ps[PX_PT + 1];
    __shared__ int s_bx[PX_PT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const __bf16 *Ob = (const __bf16 *)a.ob;
    const int64_t nranges = (a.npts + PX_PT - 1) / PX_PT;
    const int cl = lane >> 1, hf = lane & 1;  // output phase: column cl, ranks 16hf..16hf+15
    for (int64_t k = blockIdx.x; k < nranges; k += gridDim.x) {
        int bA, bB;
        px_range(a, k, bA, bB);
        if (bA >= bB) continue;
        const int p_begin = a.pstart[bA], p_end = a.pstart[bB];
        for (int c0 = 0; c0 < a.ncols; c0 += PX_K) {
            const int cn = min(PX_K, a.ncols - c0);
            int jb = bA;
            for (int base = p_begin; base < p_end; base += PX_PT) {
                const int sub_end = min(base + PX_PT, p_end), np = sub_end - base;
                __syncthreads();
                const int nbx = px_pass_boxes(a, s_ps, jb, bB, sub_end);
                px_point_vec<N>(a, Ob, s_ps, nbx, jb, base, np, vs, tab, s_bx);
                // A pass of few boxes (more than 16 points each on average): every
                // wave walks all of them and takes every 4th point of each, so a
                // big box keeps all waves busy.  A pass of many small boxes: one
                // wave per box (the waves do not all load every box's row).
                // Lane (cl, hf) holds half of the box's Pf row cl, and the wave's
                // next row is loaded while this box's points run.  Each output
                // element gets one addition per launch, so the no-return atomic
                // adds leave the sums deterministic.
                auto load_row = [&](int j, float (&pr)[16]) __attribute__((always_inline)) {
                    const int jj = min(j, nbx - 1);
                    const float *row = a.Pf + ((a.off + jb + jj) * a.ncols + c0 + min(cl, cn - 1)) * a.ldf + 16 * hf;
                    if (a.pf_vec) {
#pragma unroll
                        for (int k4 = 0; k4 < 4; k4++) {
                            const float4 v = ((const float4 *)row)[k4];
                            pr[4 * k4] = v.x;
                            pr[4 * k4 + 1] = v.y;
                            pr[4 * k4 + 2] = v.z;
                            pr[4 * k4 + 3] = v.w;
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < 16; k++) pr[k] = row[k];
                    }
                };
                const bool split = nbx < 8;  // the same for the whole workgroup
                const int j0 = split ? 0 : wave, js = split ? 1 : 4, q0 = split ? wave : 0, qs = split ? 4 : 1;
                float nxt[16];
                if (j0 < nbx) load_row(j0, nxt);
                for (int j = j0; j < nbx; j += js) {
                    const int lo = max(s_ps[j], base) - base, hi = min(s_ps[j + 1], sub_end) - base;
                    double pf[16];  // rows past the rank are not written by the M2L
#pragma unroll
                    for (int k = 0; k < 16; k++) pf[k] = 16 * hf + k < a.rank ? (double)nxt[k] : 0.0;
                    if (j + js < nbx) load_row(j + js, nxt);
                    for (int qq = lo + q0; qq < hi; qq += qs) {
                        const float4 *zv = (const float4 *)&vs[qq * PX_VS + 16 * hf];
                        double s = 0.0;
#pragma unroll
                        for (int k4 = 0; k4 < 4; k4++) {
                            const float4 z = zv[k4];
                            s += (double)z.x * pf[4 * k4] + (double)z.y * pf[4 * k4 + 1] + (double)z.z * pf[4 * k4 + 2] +
                                 (double)z.w * pf[4 * k4 + 3];
                        }
                        s += __shfl_xor(s, 1);
                        if (hf == 0 && cl < cn) {
                            double *dst = a.lp + (int64_t)(base + qq) * a.ncols + c0 + cl;
                            if (a.accumulate) unsafeAtomicAdd(dst, a.alpha * s);
                            else *dst = a.alpha * s;
                        }
                    }
                }
                jb += nbx - (s_ps[nbx] > sub_end ? 1 : 0);  // boxes that end in this pass
            }
        }
    }
}

static unsigned proxy_grid(const ProxyArgs &a, int64_t npts)
{
    if (a.rank < 1 || a.rank > PX_K) fail(PBB_ERR_INTERNAL, "proxy pass rank out of range");
    if (npts <= 0 || npts != a.npts) fail(PBB_ERR_INTERNAL, "proxy pass size");
    return (unsigned)std::min<int64_t>((npts + PX_PT - 1) / PX_PT, 1024);  // persistent workgroups
}

bool launch_p2m_proxy(const ProxyArgs &a, int n, int64_t npts, hipStream_t st)
{
    if (n < 4 || n > 8) return false;
    const unsigned g = proxy_grid(a, npts);
    if (a.ldp < 16 || a.ldp % 4) fail(PBB_ERR_INTERNAL, "proxy stride must hold 16 rows in 16-byte units");
    switch (n) {
    case 4: p2m_proxy<4><<<g, 256, 0, st>>>(a); break;
    case 5: p2m_proxy<5><<<g, 256, 0, st>>>(a); break;
    case 6: p2m_proxy<6><<<g, 256, 0, st>>>(a); break;
    case 7: p2m_proxy<7><<<g, 256, 0, st>>>(a); break;
    default: p2m_proxy<8><<<g, 256, 0, st>>>(a); break;
    }
    check(hipGetLastError(), "p2m_proxy");
    return true;
}

bool launch_l2p_proxy(const ProxyArgs &a, int n, int64_t npts, hipStream_t st)
{
    if (n < 4 || n > 8) return false;
    const unsigned g = proxy_grid(a, npts);
    if (a.ldf < a.rank) fail(PBB_ERR_INTERNAL, "M2L result stride below the rank");
    if (a.pf_vec && a.ldf % 4) fail(PBB_ERR_INTERNAL, "whole-row loads need a stride of 4 floats");
    switch (n) {
    case 4: l2p_proxy<4><<<g, 256, 0, st>>>(a); break;
    case 5: l2p_proxy<5><<<g, 256, 0, st>>>(a); break;
    case 6: l2p_proxy<6><<<g, 256, 0, st>>>(a); break;
    case 7: l2p_proxy<7><<<g, 256, 0, st>>>(a); break;
    default: l2p_proxy<8><<<g, 256, 0, st>>>(a); break;
    }
    check(hipGetLastError(), "l2p_proxy");
    return true;
}

}  // namespace dev
}  // namespace pbb
