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
//   P32 [box][col][kpad]               FP32 proxies, k padded with zeros
//   Pf32[box][col][rf]                 FP32 result
// Fragment layout of v_mfma_f32_16x16x32_{bf16,f16} (lane l of 64, e < 8):
//   A[i][k]: i = l % 16, k = 8*(l/16) + e     B[k][j]: j = l % 16, k = 8*(l/16) + e
//   D[i][j]: j = l % 16, i = 4*(l/16) + v     (v = 0..3, FP32)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace pbb {
namespace dev {

__constant__ int c_dirs[8][189][4];  // far directions of each octant class (this module's copy)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

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

// max |x| over n floats into *out (as ordered bits; out cleared by caller)
__global__ void absmax_f32(const float *__restrict__ x, int64_t n, unsigned *__restrict__ out)
{
    float m = 0.0f;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(x[q]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// ---------------------------------------------------------------------------
// One workgroup owns TB (<= 16) target boxes of one octant class x cpt
// columns (W = 64*NWC columns).  Its 2*NWC waves form a 2 x NWC grid: wave
// (wr, wc) holds row blocks [wr*RBW, wr*RBW + RBW) x the 4 16-column blocks
// of column group wc in FP32 accumulators.  The (direction, 32-deep k chunk)
// steps of the tile's non-empty directions form one software pipeline with
// one barrier per step:
//   E chunks (pre-split) are staged in a double-buffered LDS tile, the next
//   one written while the current one feeds the MFMAs;
//   each lane loads the proxies of its own B fragments straight from global
//   memory into registers one step ahead, and splits them into 16-bit parts.
constexpr int TB_MAX = 16;

template <typename T, int NSPLIT, int RB, int NWC>
__global__ __launch_bounds__(NWC * 128) void m2l_mfma_split(M2LMixArgs a)
{
    typedef typename Vec8<T>::type V8;
    constexpr int NW = 2 * NWC, NT = NW * 64, R = 16 * RB, KS = 40, RBW = (RB + 1) / 2;
    constexpr int NA = NSPLIT * R * 4;  // 16-byte units of one E chunk
    constexpr int AR = (NA + NT - 1) / NT;
    __shared__ __attribute__((aligned(16))) T As[2][NSPLIT * R * KS];
    __shared__ int s_sid[189 * TB_MAX];
    __shared__ short s_dir[189];
    __shared__ int s_nd;
    __shared__ int s_tgt[TB_MAX];
    __shared__ int s_xyz[TB_MAX][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;
    const int rf = a.rf, TB = a.TB, nch = a.nch;
    const int lin = xcd_tile(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    // the column chunks of one tile are adjacent in launch order: they run
    // side by side on one XCD through the same (direction, k-chunk) steps and
    // share each E chunk through its L2
    const int tile = lin / gridDim.y;
    int o = 0;
    while (o < 7 && tile >= a.tile_prefix[o + 1]) o++;
    const int64_t first = a.class_start[o] + (int64_t)(tile - a.tile_prefix[o]) * TB;
    const int64_t cend = a.class_start[o + 1];
    const int c0 = (lin - tile * (int)gridDim.y) * a.cpt;
    const int cpt = min(a.cpt, a.ncols - c0);
    const int S = a.side;
    // FP16 operands are scaled into range by 2^xs (P) and 2^escale (E)
    int xs = 0;
    if (a.xmax) {
        float mx = __uint_as_float(*a.xmax);
        if (mx > 0.0f) {
            int e;
            frexpf(mx, &e);
            xs = 12 - e;
        }
    }
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
        if (s_tgt[tb] >= 0) {
            int sx = s_xyz[tb][0] + c_dirs[o][k][0], sy = s_xyz[tb][1] + c_dirs[o][k][1],
                sz = s_xyz[tb][2] + c_dirs[o][k][2];
            if (sx >= 0 && sy >= 0 && sz >= 0 && sx < S && sy < S && sz < S)
                sid = a.map[((int64_t)sx * S + sy) * S + sz];
        }
        s_sid[k * TB_MAX + tb] = sid;
    }
    __syncthreads();
    for (int k = tid; k < 189; k += NT) {  // does direction k reach any source box?
        bool any = false;
        for (int tb = 0; tb < TB; tb++) any |= s_sid[k * TB_MAX + tb] >= 0;
        s_dir[k] = any ? 1 : 0;
    }
    __syncthreads();
    if (tid == 0) {  // compact the non-empty directions in class order
        int nd = 0;
        for (int k = 0; k < 189; k++)
            if (s_dir[k]) s_dir[nd++] = (short)k;
        s_nd = nd;
    }
    __syncthreads();
    // steps run k-chunk-major: all directions of one chunk, then the next
    // chunk, so that consecutive steps gather the same proxy rows of largely
    // the same source boxes (see class_directions)
    const int nd = s_nd, nsteps = nd * nch;

    const int fr = lane % 16, fq = lane / 16;
    int btb[4], bcl[4];  // box slot and column of this lane's B column in each block
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        const int col = wc * 64 + cb * 16 + fr;
        const int tb = col / cpt;
        btb[cb] = tb < TB ? tb : -1;
        bcl[cb] = col - tb * cpt;
    }
    uint4 ra[AR];
    float4 bn[4][2];
    auto loadA = [&](int step) {
        const int ch = step / nd, k = s_dir[step - ch * nd];
        const uint4 *src = (const uint4 *)((const T *)a.Es + ((size_t)c_dirs[o][k][3] * nch + ch) * NSPLIT * R * 32);
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int u = tid + r * NT;
            ra[r] = u < NA ? src[u] : uint4{0, 0, 0, 0};
        }
    };
    auto storeA = [&](int buf) {
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int u = tid + r * NT;
            if (u < NA) {
                const int p = u / (R * 4), row = (u / 4) % R, q = u % 4;
                *(uint4 *)&As[buf][(p * R + row) * KS + 8 * q] = ra[r];
            }
        }
    };
    auto loadB = [&](int step) {
        const int ch = step / nd, k = s_dir[step - ch * nd];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            float4 v0 = {0, 0, 0, 0}, v1 = {0, 0, 0, 0};
            const int sid = btb[cb] >= 0 ? s_sid[k * TB_MAX + btb[cb]] : -1;
            if (sid >= 0) {
                const float4 *src = (const float4 *)(a.P + ((size_t)(a.off + sid) * a.ncols + c0 + bcl[cb]) * a.kpad +
                                                     ch * 32 + 8 * fq);
                v0 = src[0];
                v1 = src[1];
            }
            bn[cb][0] = v0;
            bn[cb][1] = v1;
        }
    };
    f32x4 acc[RBW][4];
#pragma unroll
    for (int i = 0; i < RBW; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0, 0, 0, 0};

    if (nsteps > 0) {
        loadA(0);
        storeA(0);
        if (nsteps > 1) loadA(1);
        loadB(0);
    }
    __syncthreads();
    for (int step = 0; step < nsteps; step++) {
        const int buf = step & 1;
        V8 bfr[4][NSPLIT];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            float x[8] = {bn[cb][0].x, bn[cb][0].y, bn[cb][0].z, bn[cb][0].w,
                          bn[cb][1].x, bn[cb][1].y, bn[cb][1].z, bn[cb][1].w};
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
        // the other LDS buffer was last read in step - 1, before the barrier
        if (step + 1 < nsteps) storeA(buf ^ 1);
        if (step + 2 < nsteps) loadA(step + 2);  // in flight during the MFMAs
        if (step + 1 < nsteps) loadB(step + 1);
        const T *Ab = As[buf];
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int rb = wr * RBW + i;
            if (rb >= RB) break;
            V8 afr[NSPLIT];
#pragma unroll
            for (int p = 0; p < NSPLIT; p++) afr[p] = *(const V8 *)&Ab[(p * R + rb * 16 + fr) * KS + 8 * fq];
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                // smallest terms first
                if constexpr (NSPLIT == 3) {
                    acc[i][cb] = mfma16(afr[2], bfr[cb][0], acc[i][cb]);
                    acc[i][cb] = mfma16(afr[1], bfr[cb][1], acc[i][cb]);
                    acc[i][cb] = mfma16(afr[0], bfr[cb][2], acc[i][cb]);
                }
                acc[i][cb] = mfma16(afr[1], bfr[cb][0], acc[i][cb]);
                acc[i][cb] = mfma16(afr[0], bfr[cb][1], acc[i][cb]);
                acc[i][cb] = mfma16(afr[0], bfr[cb][0], acc[i][cb]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        if (btb[cb] < 0) continue;
        const int t = s_tgt[btb[cb]];
        if (t < 0 || bcl[cb] >= cpt) continue;
        float *dst = a.Pf + ((size_t)(a.off + t) * a.ncols + c0 + bcl[cb]) * rf;
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int rb = wr * RBW + i;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rb * 16 + 4 * fq + v;
                if (rb < RB && row < rf) dst[row] = acc[i][cb][v] * unscale;
            }
        }
    }
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

void launch_absmax(const float *x, int64_t n, unsigned *out, hipStream_t st)
{
    if (n <= 0) return;
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    absmax_f32<<<blocks, 256, 0, st>>>(x, n, out);
    check(hipGetLastError(), "absmax_f32");
}

template <typename T, int NSPLIT>
static void m2l_split_t(const M2LMixArgs &a, dim3 grid, hipStream_t st)
{
    const int rb = (a.rf + 15) / 16;
    const int w = a.TB * a.cpt;  // tile width: 64, 128 or 256 columns
    if (w > 256) fail(PBB_ERR_INTERNAL, "mixed-precision M2L tile wider than 256 columns");
#define PBB_CASE(RB)                                                                   \
    case RB:                                                                           \
        if (w > 128) m2l_mfma_split<T, NSPLIT, RB, 4><<<grid, 512, 0, st>>>(a);        \
        else if (w > 64) m2l_mfma_split<T, NSPLIT, RB, 2><<<grid, 256, 0, st>>>(a);    \
        else m2l_mfma_split<T, NSPLIT, RB, 1><<<grid, 128, 0, st>>>(a);                \
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
    dim3 grid((unsigned)ntiles, (unsigned)ncolchunks);
    switch (variant) {
    case PBB_PREC_MIXED: m2l_split_t<__bf16, 3>(a, grid, st); break;
    case PBB_PREC_MIXED_F16X2: m2l_split_t<_Float16, 2>(a, grid, st); break;
    case PBB_PREC_MIXED_BF16X2: m2l_split_t<__bf16, 2>(a, grid, st); break;
    default: fail(PBB_ERR_INVALID_ARG, "unknown mixed-precision variant");
    }
    check(hipGetLastError(), "m2l_mfma_split");
}

}  // namespace dev
}  // namespace pbb
