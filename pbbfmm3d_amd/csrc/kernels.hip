// kernels.hip -- gfx950 kernels of the FP64 black-box FMM matvec.
//
// Data layout in HBM (one plan):
//   points    sorted by leaf (Morton order), re-centred, SoA x[N] y[N] z[N]
//   perm      sorted position -> caller's index (weights/outputs stay in the
//             caller's column-major order w[c*N + i])
//   ws        weights gathered into sorted order, ws[c*N + k]
//   S / F     per box (global id = level offset + compact id) a column-major
//             n3 x ncols block: S[(g*ncols + c)*n3 + l].  The upward pass's
//             S (sourceval) and the downward pass's F (targetval) share it.
//   P         proxies (proxysval) r_s x ncols per box, P[(g*ncols + c)*rs + a]
//   Pf        M2L result r_f x ncols per box
// Operators: VTw = VT diag(Kweights) (r_s x n3), Uw = diag(Kweights) U
// (n3 x r_f), E (316 blocks r_f x r_s), all column-major FP64.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace pbb {
namespace dev {

__constant__ int c_dirs[8][189][4];  // far directions of each octant class

typedef double dbl4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// device kernel functors (formulas of pbb_kernel_id); r2 = |s - t|^2 with
// the differences taken as source - target (kernel_Types.cpp:21-23)
template <int KID>
__device__ __forceinline__ double keval(double dx, double dy, double dz, const KParams &p)
{
    double r2 = dx * dx + dy * dy + dz * dz;
    if constexpr (KID == PBB_KERNEL_LAPLACIAN) return 1.0 / sqrt(r2);
    else if constexpr (KID == PBB_KERNEL_ONEOVERR2) { double ri = 1.0 / sqrt(r2); return ri * ri; }
    else if constexpr (KID == PBB_KERNEL_GAUSSIAN) return exp(-r2 / (p.p[0] * p.p[0]));
    else if constexpr (KID == PBB_KERNEL_EXPONENTIAL) return p.p[1] * exp(-sqrt(r2) / p.p[0]);
    else if constexpr (KID == PBB_KERNEL_SOFT_LAPLACIAN) return 1.0 / sqrt(r2 + p.p[0]);
    else if constexpr (KID == PBB_KERNEL_QUADRIC) return 1.0 + r2;
    else if constexpr (KID == PBB_KERNEL_INV_QUADRIC) return 1.0 / (1.0 + r2);
    else if constexpr (KID == PBB_KERNEL_LOGARITHM) return r2 == 0 ? 0.0 : 0.5 * log(r2);
    else if constexpr (KID == PBB_KERNEL_THIN_PLATE) return r2 == 0 ? 0.0 : 0.5 * r2 * log(r2);
    else if constexpr (KID == PBB_KERNEL_R2) return r2;
    else if constexpr (KID == PBB_KERNEL_CONSTANT) return 1.0;
    else if constexpr (KID == PBB_KERNEL_EXP_ANISO)
        return p.p[0] * exp(-sqrt(p.p[1] * dx * dx + p.p[2] * dy * dy + p.p[3] * dz * dz));
    else return 0.0;
}

// The same functors with the transcendental part in FP32 (the mixed modes'
// near field): r^2 from the FP64 differences, then the FP32 library
// function; relative error a few 1e-7 per value
template <int KID>
__device__ __forceinline__ double keval_f32(double dx, double dy, double dz, const KParams &p)
{
    const double r2 = dx * dx + dy * dy + dz * dz;
    const float f = (float)r2;
    if constexpr (KID == PBB_KERNEL_LAPLACIAN) return (double)(1.0f / sqrtf(f));
    else if constexpr (KID == PBB_KERNEL_ONEOVERR2) return (double)(1.0f / f);
    else if constexpr (KID == PBB_KERNEL_GAUSSIAN) return (double)expf(-f / (float)(p.p[0] * p.p[0]));
    else if constexpr (KID == PBB_KERNEL_EXPONENTIAL) return p.p[1] * (double)expf(-sqrtf(f) / (float)p.p[0]);
    else if constexpr (KID == PBB_KERNEL_SOFT_LAPLACIAN) return (double)(1.0f / sqrtf(f + (float)p.p[0]));
    else if constexpr (KID == PBB_KERNEL_LOGARITHM) return r2 == 0 ? 0.0 : (double)(0.5f * logf(f));
    else if constexpr (KID == PBB_KERNEL_THIN_PLATE) return r2 == 0 ? 0.0 : r2 * (double)(0.5f * logf(f));
    else if constexpr (KID == PBB_KERNEL_EXP_ANISO)
        return p.p[0] * (double)expf(-sqrtf((float)(p.p[1] * dx * dx + p.p[2] * dy * dy + p.p[3] * dz * dz)));
    else return keval<KID>(dx, dy, dz, p);  // polynomial and rational ones: FP64 is as cheap
}

// S_m(u), m < n: Chebyshev interpolant by Clenshaw (H2_3D_Tree.cpp:830-879)
__device__ __forceinline__ void sn_eval(int n, const double *__restrict__ Tkz, double u, double *S)
{
    const double pfac = 2.0 / n;
    for (int m = 0; m < n; m++) {
        const double *Tk = Tkz + m * n;
        double d0 = 0, d1 = 0;
        for (int j = n - 1; j > 0; j--) {
            double d2 = d0;
            d0 = 2.0 * u * d0 - d1 + Tk[j];
            d1 = d2;
        }
        S[m] = pfac * (u * d0 - d1 + 0.5 * Tk[0]);
    }
}

// S_m(u) for m < n written to dst[m * stride]
__device__ __forceinline__ void sn_eval_to(int n, const double *__restrict__ Tkz, double u, double *dst, int stride)
{
    const double pfac = 2.0 / n;
    for (int m = 0; m < n; m++) {
        const double *Tk = Tkz + m * n;
        double d0 = 0, d1 = 0;
        for (int j = n - 1; j > 0; j--) {
            double d2 = d0;
            d0 = 2.0 * u * d0 - d1 + Tk[j];
            d1 = d2;
        }
        dst[m * stride] = pfac * (u * d0 - d1 + 0.5 * Tk[0]);
    }
}

// ---------------------------------------------------------------------------
// ws[c*N + k] = w[c*w_n + j] with j = perm[k] - w_lo when 0 <= j < w_n,
// else 0: the points that carry no weight (the targets of a plan whose
// sources are not its targets) enter the sums with weight zero
__global__ void gather_weights(const double *__restrict__ w, const int32_t *__restrict__ perm, int64_t N,
                               int ncols, double *__restrict__ ws, int64_t w_lo, int64_t w_n)
{
    int64_t total = N * ncols;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = q / N, k = q - c * N;
        const int64_t j = perm[k] - w_lo;
        ws[q] = (uint64_t)j < (uint64_t)w_n ? w[c * w_n + j] : 0.0;
    }
}

// wsT[k*ncols + c], the same values point-major (for the near field)
__global__ void gather_weights_t(const double *__restrict__ w, const int32_t *__restrict__ perm, int64_t N,
                                 int ncols, double *__restrict__ wsT, int64_t w_lo, int64_t w_n)
{
    int64_t total = N * ncols;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = q / ncols, c = q - k * ncols;
        const int64_t j = perm[k] - w_lo;
        wsT[q] = (uint64_t)j < (uint64_t)w_n ? w[c * w_n + j] : 0.0;
    }
}

// The same wsT in two coalesced steps: wT[j*ncols + c] = w[c*w_n + j] (a
// transpose through LDS tiles of 32 points x 32 columns), then whole rows
// wsT[k] = wT[perm[k] - w_lo] (zero rows for points outside the range).
__global__ __launch_bounds__(256) void transpose_weights(const double *__restrict__ w, int64_t w_n, int ncols,
                                                         double *__restrict__ wT)
{
    __shared__ double tile[32][33];
    const int64_t j0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows of 32 lanes
    for (int c0 = 0; c0 < ncols; c0 += 32) {
        __syncthreads();
        for (int c = ty; c < 32; c += 8)
            tile[c][tx] = c0 + c < ncols && j0 + tx < w_n ? w[(int64_t)(c0 + c) * w_n + j0 + tx] : 0.0;
        __syncthreads();
        for (int j = ty; j < 32; j += 8)
            if (j0 + j < w_n && c0 + tx < ncols) wT[(j0 + j) * ncols + c0 + tx] = tile[tx][j];
    }
}

__global__ void gather_rows(const double *__restrict__ wT, const int32_t *__restrict__ perm, int64_t N, int ncols,
                            double *__restrict__ wsT, int64_t w_lo, int64_t w_n)
{
    const int64_t total = N * ncols;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = q / ncols, c = q - k * ncols;
        const int64_t j = perm[k] - w_lo;
        wsT[q] = (uint64_t)j < (uint64_t)w_n ? wT[j * ncols + c] : 0.0;
    }
}

// ---------------------------------------------------------------------------
// P2M, compute.hpp:287-323: one workgroup per non-empty leaf.
// S[(g*ncols + c)*n3 + (l1 n + l2) n + l3] = sum_j w_jc Sx_l1(j) Sy_l2(j) Sz_l3(j)
constexpr int P2M_PTS = 32;
constexpr int P2M_R = 16;

__global__ __launch_bounds__(256) void p2m_fp64(P2MArgs a)
{
    __shared__ double sS[3][16][P2M_PTS];
    __shared__ double sW[64][P2M_PTS];
    const int b = blockIdx.x;
    const int n = a.n, n3 = n * n * n, tid = threadIdx.x;
    const int t0 = a.leaf_start[b], nt = a.leaf_start[b + 1] - t0;
    const uint32_t cc = (uint32_t)a.coord[b];
    const double cen[3] = {a.ctab[cc & 1023], a.ctab[(cc >> 10) & 1023], a.ctab[(cc >> 20) & 1023]};
    const double ihalfL = 2.0 / a.len;
    const double *xyz[3] = {a.x, a.y, a.z};
    double *Sout = a.S + (size_t)(a.off + b) * a.ncols * n3;

    for (int c0 = 0; c0 < a.ncols; c0 += a.cc) {
        const int cn = min(a.cc, a.ncols - c0);
        double acc[P2M_R];
#pragma unroll
        for (int r = 0; r < P2M_R; r++) acc[r] = 0;
        for (int j0 = 0; j0 < nt; j0 += P2M_PTS) {
            const int jn = min(P2M_PTS, nt - j0);
            if (tid < 3 * P2M_PTS) {
                int p = tid % P2M_PTS, ax = tid / P2M_PTS;
                double S[16];
                double u = p < jn ? ihalfL * (xyz[ax][t0 + j0 + p] - cen[ax]) : 0.0;
                sn_eval(n, a.Tkz, u, S);
                for (int m = 0; m < n; m++) sS[ax][m][p] = p < jn ? S[m] : 0.0;
            }
            for (int q = tid; q < cn * P2M_PTS; q += 256) {
                int c = q / P2M_PTS, p = q % P2M_PTS;
                sW[c][p] = p < jn ? a.ws[(int64_t)(c0 + c) * a.N + t0 + j0 + p] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < P2M_R; r++) {
                int q = tid + 256 * r;
                if (q < cn * n3) {
                    int c = q / n3, l = q - c * n3;
                    int l1 = l / (n * n), l2 = (l / n) % n, l3 = l % n;
                    double s = 0;
                    for (int p = 0; p < jn; p++) s += sW[c][p] * sS[0][l1][p] * sS[1][l2][p] * sS[2][l3][p];
                    acc[r] += s;
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < P2M_R; r++) {
            int q = tid + 256 * r;
            if (q < cn * n3) Sout[(size_t)c0 * n3 + q] = acc[r];
        }
    }
}

// ---------------------------------------------------------------------------
// M2M, compute.hpp:234-283 + :956-1034: one workgroup per parent box; the
// children's expansions are contracted mode by mode (z, y, x) in LDS.
// TS: the storage type of the expansions (FP64, or FP32 in the mixed modes);
// the transfers run in FP64 either way.
template <typename TS>
__global__ __launch_bounds__(256) void m2m_fp64(M2MArgs a)
{
    extern __shared__ double lds[];
    const int n = a.n, n2 = n * n, n3 = n2 * n, tid = threadIdx.x;
    const int CC = a.cc;
    double *A = lds, *B = lds + CC * n3, *Cw = lds + 2 * CC * n3;
    for (int q = tid; q < 2 * n2; q += 256) Cw[q] = a.Cw[q];
    const int b = blockIdx.x;
    TS *Sp = (TS *)a.Sparent + (size_t)(a.off_parent + b) * a.ncols * n3;
    for (int c0 = 0; c0 < a.ncols; c0 += CC) {
        const int cn = min(CC, a.ncols - c0);
        const int tot = cn * n3;
        double acc[P2M_R];
#pragma unroll
        for (int r = 0; r < P2M_R; r++) acc[r] = 0;
        for (int o = 0; o < 8; o++) {
            const int ch = a.child[(int64_t)b * 8 + o];
            if (ch < 0) continue;
            const TS *Sc = (const TS *)a.Schild + ((size_t)(a.off_child + ch) * a.ncols + c0) * n3;
            const double *Cx = Cw + ((o >> 2) & 1) * n2, *Cy = Cw + ((o >> 1) & 1) * n2, *Cz = Cw + (o & 1) * n2;
            __syncthreads();
            for (int q = tid; q < tot; q += 256) A[q] = Sc[q];
            __syncthreads();
            for (int q = tid; q < tot; q += 256) {  // z: B[ab n + l3] = sum_k A[ab n + k] Cz[l3 n + k]
                int c = q / n3, l = q - c * n3, ab = l / n, l3 = l % n;
                const double *src = A + c * n3 + ab * n;
                double s = 0;
                for (int k = 0; k < n; k++) s += src[k] * Cz[l3 * n + k];
                B[q] = s;
            }
            __syncthreads();
            for (int q = tid; q < tot; q += 256) {  // y: A[l1,l2,l3] = sum_k B[l1,k,l3] Cy[l2 n + k]
                int c = q / n3, l = q - c * n3, l1 = l / n2, l2 = (l / n) % n, l3 = l % n;
                const double *src = B + c * n3 + l1 * n2 + l3;
                double s = 0;
                for (int k = 0; k < n; k++) s += src[k * n] * Cy[l2 * n + k];
                A[q] = s;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < P2M_R; r++) {  // x: acc[l1,l2,l3] += sum_k A[k,l2,l3] Cx[l1 n + k]
                int q = tid + 256 * r;
                if (q < tot) {
                    int c = q / n3, l = q - c * n3, l1 = l / n2, l23 = l % n2;
                    const double *src = A + c * n3 + l23;
                    double s = 0;
                    for (int k = 0; k < n; k++) s += src[k * n2] * Cx[l1 * n + k];
                    acc[r] += s;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < P2M_R; r++) {
            int q = tid + 256 * r;
            if (q < tot) Sp[(size_t)c0 * n3 + q] = (TS)acc[r];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// L2L, compute.hpp:551-587 + :798-878: one workgroup per parent box, the
// transpose of M2M added into each non-empty child.
template <typename TS>
__global__ __launch_bounds__(256) void l2l_fp64(M2MArgs a)
{
    extern __shared__ double lds[];
    const int n = a.n, n2 = n * n, n3 = n2 * n, tid = threadIdx.x;
    const int CC = a.cc;
    double *Pp = lds, *X = lds + CC * n3, *Y = lds + 2 * CC * n3, *Cw = lds + 3 * CC * n3;
    for (int q = tid; q < 2 * n2; q += 256) Cw[q] = a.Cw[q];
    const int b = blockIdx.x;
    const TS *Fp = (const TS *)a.Sparent + (size_t)(a.off_parent + b) * a.ncols * n3;
    for (int c0 = 0; c0 < a.ncols; c0 += CC) {
        const int cn = min(CC, a.ncols - c0);
        const int tot = cn * n3;
        __syncthreads();
        for (int q = tid; q < tot; q += 256) Pp[q] = Fp[(size_t)c0 * n3 + q];
        for (int o = 0; o < 8; o++) {
            const int ch = a.child[(int64_t)b * 8 + o];
            if (ch < 0) continue;
            TS *Fc = (TS *)a.Schild + ((size_t)(a.off_child + ch) * a.ncols + c0) * n3;
            const double *Cx = Cw + ((o >> 2) & 1) * n2, *Cy = Cw + ((o >> 1) & 1) * n2, *Cz = Cw + (o & 1) * n2;
            __syncthreads();
            for (int q = tid; q < tot; q += 256) {  // x: X[l1,l2,l3] = sum_k P[k,l2,l3] Cx[k n + l1]
                int c = q / n3, l = q - c * n3, l1 = l / n2, l23 = l % n2;
                const double *src = Pp + c * n3 + l23;
                double s = 0;
                for (int k = 0; k < n; k++) s += src[k * n2] * Cx[k * n + l1];
                X[q] = s;
            }
            __syncthreads();
            for (int q = tid; q < tot; q += 256) {  // y: Y[l1,l2,l3] = sum_k X[l1,k,l3] Cy[k n + l2]
                int c = q / n3, l = q - c * n3, l1 = l / n2, l2 = (l / n) % n, l3 = l % n;
                const double *src = X + c * n3 + l1 * n2 + l3;
                double s = 0;
                for (int k = 0; k < n; k++) s += src[k * n] * Cy[k * n + l2];
                Y[q] = s;
            }
            __syncthreads();
            for (int q = tid; q < tot; q += 256) {  // z: Fc[ab n + l3] += sum_k Y[ab n + k] Cz[k n + l3]
                int c = q / n3, l = q - c * n3, ab = l / n, l3 = l % n;
                const double *src = Y + c * n3 + ab * n;
                double s = 0;
                for (int k = 0; k < n; k++) s += src[k] * Cz[k * n + l3];
                Fc[q] = (TS)((double)Fc[q] + s);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// C[:, j] = alpha * A B[:, j] for j < ncol: A (M x K, lda M, FP64), B
// (K x ncol, ldb >= K), C (ldc >= M rows per column; rows M..ldc-1 are
// written as zeros), column-major.  64 x 64 output tiles, K in chunks of 16
// through LDS, 4 x 4 outputs per thread, FP64 arithmetic; B and C may be
// FP32.  Serves the proxy projection P = VTw S (compute.hpp:376-395) and
// the local expansion F = scale Uw Pf (compute.hpp:497-508).
template <typename TB, typename TC>
__global__ __launch_bounds__(256) void gemm_small_a(GemmArgs g)
{
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    const TB *B = (const TB *)g.B;
    TC *C = (TC *)g.C;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int64_t n0 = (int64_t)blockIdx.x * 64;
    const int m0 = blockIdx.y * 64;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        for (int q = tid; q < 16 * 64; q += 256) {
            int row = q % 64, k = q / 64;
            int gr = m0 + row, gk = k0 + k;
            As[k][row] = (gr < g.M && gk < g.K) ? g.A[gr + (int64_t)gk * g.M] : 0.0;
        }
        for (int q = tid; q < 16 * 64; q += 256) {
            int k = q % 16, col = q / 16;
            int gk = k0 + k;
            int64_t gc = n0 + col;
            Bs[k][col] = (gk < g.K && gc < g.ncol) ? (double)B[gk + gc * g.ldb] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[k][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[k][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int64_t gc = n0 + ty + 16 * j;
        if (gc >= g.ncol) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int gr = m0 + tx + 16 * i;
            if (gr < g.M) C[gr + gc * g.ldc] = (TC)(g.alpha * acc[i][j]);
            else if (gr < g.ldc) C[gr + gc * g.ldc] = (TC)0;
        }
    }
}

// ---------------------------------------------------------------------------
// The same GEMM on the FP64 matrix cores: a workgroup of 4 waves computes all
// RB 16-row blocks of M x 64 columns (one 16-column block per wave), K in
// chunks of 16.  A (small, shared by every workgroup) goes through a
// double-buffered LDS tile; each lane loads its own B fragment straight from
// memory, one chunk ahead, into two register sets used in turn.  Within a
// chunk, lane group fk supplies k = 4 fk + ks at MFMA sub-step ks, so its
// four B values are consecutive in memory.  Loads are branch-free (clamped
// addresses, masked values) so that the waits on loads in flight are exact.
template <typename TB, typename TC, int RB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void gemm_mfma_f64(GemmArgs g)
{
    constexpr int NT = 256, W = 64, R = 16 * RB, KC = 16, AS = R + 2;
    constexpr int AR = (KC * R + NT - 1) / NT;
    __shared__ double As[2][KC * AS];
    const TB *B = (const TB *)g.B;
    TC *C = (TC *)g.C;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t n0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * W;
    const int nsteps = (g.K + KC - 1) / KC;
    const int fr = lane % 16, fk = lane / 16;
    const int64_t gc = n0 + wave * 16 + fr;  // this lane's B and C column
    const TB *bcol = B + (gc < g.ncol ? gc : 0) * g.ldb;
    double ra[AR];
    auto loadA = [&](int step) {
        const int k0 = step * KC;
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int q = tid + r * NT, row = q % R, gk = k0 + q / R;
            const bool ok = q < KC * R && row < g.M && gk < g.K;
            const double v = g.A[ok ? row + (int64_t)gk * g.M : 0];
            ra[r] = ok ? v : 0.0;
        }
    };
    auto storeA = [&](int buf) {
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int q = tid + r * NT;
            if (AR * NT == KC * R || q < KC * R) As[buf][(q / R) * AS + q % R] = ra[r];
        }
    };
    auto loadB = [&](int step, double (&b)[4]) {
        const int k = step * KC + 4 * fk;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const double v = (double)bcol[min(k + e, g.K - 1)];
            b[e] = k + e < g.K ? v : 0.0;
        }
    };
    dbl4 acc[RB];
#pragma unroll
    for (int i = 0; i < RB; i++) acc[i] = dbl4{0, 0, 0, 0};
    auto run_step = [&](int step, const double (&bcur)[4], double (&bnext)[4]) {
        const int buf = step & 1;
        storeA(buf ^ 1);  // last read in step - 1, before the barrier
        loadA(min(step + 2, nsteps - 1));
        loadB(min(step + 1, nsteps - 1), bnext);
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const double *a = &As[buf][(4 * fk + ks) * AS + fr];
#pragma unroll
            for (int rb_ = 0; rb_ < RB; rb_++)
                acc[rb_] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb_ * 16], bcur[ks], acc[rb_], 0, 0, 0);
        }
        __syncthreads();
    };
    double bA[4], bB[4];
    loadA(0);
    storeA(0);
    loadA(min(1, nsteps - 1));
    loadB(0, bA);
    __syncthreads();
    for (int step = 0; step < nsteps; step += 2) {
        run_step(step, bA, bB);
        if (step + 1 < nsteps) run_step(step + 1, bB, bA);
    }
    if (gc < g.ncol) {
#pragma unroll
        for (int rb_ = 0; rb_ < RB; rb_++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rb_ * 16 + fk + 4 * v;
                if (row < g.M) C[row + gc * g.ldc] = (TC)(g.alpha * acc[rb_][v]);
                else if (row < g.ldc) C[row + gc * g.ldc] = (TC)0;
            }
        for (int row = R; row < g.ldc; row++)  // padding rows beyond the MFMA tile
            if (fk == 0) C[row + gc * g.ldc] = (TC)0;
    }
}

// ---------------------------------------------------------------------------
// M2L, compute.hpp:401-493 + Moment2Local :880-923.  A workgroup owns TB
// target boxes of one octant class (Morton-adjacent) and the same column
// chunk of each: a W-wide output tile Pf (all r_f rows) that stays in
// registers while the 189 far directions of the class stream through:
//   Pf[:, tile] += E_d P[:, tile + d]     (empty source boxes give zeros)
// ---------------------------------------------------------------------------
// On the FP64 matrix cores (v_mfma_f64_16x16x4_f64).  The tile is W = 32*NWC
// columns; its 2*NWC waves form a 2 x NWC grid: wave (wr, wc) holds row
// blocks [wr*RBW, wr*RBW + RBW) x the two 16-column blocks of column group
// wc.  The (direction, 16-deep k chunk) steps form one software pipeline
// with one barrier per step: E chunks go through a double-buffered LDS tile;
// each lane loads the proxies of its own B fragments straight from memory,
// one step ahead, into two register sets used in turn.  Within a chunk, lane
// group fk supplies k = 4 fk + ks at MFMA sub-step ks, so its four proxy
// values are consecutive in memory.  Loads are branch-free (clamped
// addresses, masked values) so that the waits on loads in flight are exact.
// Fragment layout of the 16x16x4 FP64 MFMA (lane l of 64):
//   A[i][k]: i = l % 16, k = l / 16      B[k][j]: j = l % 16, k = l / 16
//   D[i][j]: j = l % 16, i = l / 16 + 4*v     (v = 0..3, one double each)

constexpr int TB_MAX = 16;

template <int RB, int NWC>
__global__ __launch_bounds__(NWC * 128) void m2l_mfma_f64(M2LArgs a)
{
    constexpr int NT = NWC * 128, R = 16 * RB, KC = 16, AS = R + 2, RBW = (RB + 1) / 2;
    constexpr int AR = (KC * R + NT - 1) / NT;
    __shared__ double As[2][KC * AS];
    __shared__ int s_sid[189 * TB_MAX];
    __shared__ short s_dir[189];
    __shared__ int s_nd;
    __shared__ int s_tgt[TB_MAX];
    __shared__ int s_xyz[TB_MAX][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;
    const int rf = a.rf, rs = a.rs, TB = a.TB;
    const int nch = (rs + KC - 1) / KC;
    // gridDim.z row groups of R rows each (ranks above one group's rows).
    // The row groups and column chunks of one tile are adjacent in launch
    // order: they run side by side on one XCD through the same (direction,
    // k-chunk) steps and share each proxy and E chunk through its L2
    const int gyz = (int)(gridDim.y * gridDim.z);
    const int lin = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x * gyz);
    const int tile = lin / gyz;
    const int rem = lin - tile * gyz, chunk = rem / (int)gridDim.z;
    const int row0 = (rem - chunk * (int)gridDim.z) * R;
    int o = 0;
    while (o < 7 && tile >= a.tile_prefix[o + 1]) o++;
    const int64_t first = a.class_start[o] + (int64_t)(tile - a.tile_prefix[o]) * TB;
    const int64_t cend = a.class_start[o + 1];
    const int c0 = chunk * a.cpt;
    const int cpt = min(a.cpt, a.ncols - c0);
    const int S = a.side;
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
    // step order: a.chunk_major (see m2l_step); in chunk-major order
    // consecutive steps gather the same proxy rows of largely the same source
    // boxes (see class_directions)
    const int nd = s_nd, nsteps = nd * nch;

    const int fr = lane % 16, fk = lane / 16;
    int btb[2], bcl[2];  // box slot and column of this lane's B column in each block
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        const int col = wc * 32 + cb * 16 + fr;
        const int tb = col / cpt;
        btb[cb] = tb < TB ? tb : -1;
        bcl[cb] = col - tb * cpt;
    }
    double ra[AR];
    auto loadA = [&](int step) {
        int di, ch;
        m2l_step(step, nd, nch, a.chunk_major, di, ch);
        const int k = s_dir[di], k0 = ch * KC;
        const double *Ed = a.E + (size_t)c_dirs[o][k][3] * rf * rs;
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int q = tid + r * NT, row = row0 + q % R, gk = k0 + q / R;
            const bool ok = q < KC * R && row < rf && gk < rs;
            const double v = Ed[ok ? row + (size_t)gk * rf : 0];
            ra[r] = ok ? v : 0.0;
        }
    };
    auto storeA = [&](int buf) {
#pragma unroll
        for (int r = 0; r < AR; r++) {
            const int q = tid + r * NT;
            if (AR * NT == KC * R || q < KC * R) As[buf][(q / R) * AS + q % R] = ra[r];
        }
    };
    // a column without a source box loads box 0 of the level (always valid)
    // and is zeroed through the mask at its use
    auto loadB = [&](int step, double (&bn)[2][4], int &okm) {
        int di, ch;
        m2l_step(step, nd, nch, a.chunk_major, di, ch);
        const int k = s_dir[di], k0 = ch * KC + 4 * fk;
        okm = 0;
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            const int sid0 = btb[cb] >= 0 ? s_sid[k * TB_MAX + btb[cb]] : -1;
            okm |= (sid0 >= 0) << cb;
            const double *col = a.P + ((size_t)(a.off + max(sid0, 0)) * a.ncols + c0 + bcl[cb]) * a.ldp;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const double v = col[min(k0 + e, rs - 1)];
                bn[cb][e] = k0 + e < rs ? v : 0.0;
            }
        }
    };
    dbl4 acc[RBW][2];
#pragma unroll
    for (int i = 0; i < RBW; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = dbl4{0, 0, 0, 0};
    auto run_step = [&](int step, const double (&bcur)[2][4], int okcur, double (&bnext)[2][4], int &oknext) {
        const int buf = step & 1;
        double b[2][4];
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            const bool ok = (okcur >> cb) & 1;
#pragma unroll
            for (int e = 0; e < 4; e++) b[cb][e] = ok ? bcur[cb][e] : 0.0;
        }
        storeA(buf ^ 1);  // last read in step - 1, before the barrier
        loadA(min(step + 2, nsteps - 1));
        loadB(min(step + 1, nsteps - 1), bnext, oknext);
        // a wave whose two column blocks have no source box in this direction
        // (zero proxies) skips the step's MFMAs; the sums are unchanged
        if (__ballot(okcur != 0) != 0) {
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                const double *ap = &As[buf][(4 * fk + ks) * AS + fr];
#pragma unroll
                for (int i = 0; i < RBW; i++) {
                    const int rb_ = wr * RBW + i;
                    if (rb_ >= RB) break;
                    const double av = ap[rb_ * 16];
#pragma unroll
                    for (int cb = 0; cb < 2; cb++)
                        acc[i][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[cb][ks], acc[i][cb], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    };
    double bA[2][4], bB[2][4];
    int okA = 0, okB = 0;
    if (nsteps > 0) {
        loadA(0);
        storeA(0);
        loadA(min(1, nsteps - 1));
        loadB(0, bA, okA);
    }
    __syncthreads();
    for (int step = 0; step < nsteps; step += 2) {
        run_step(step, bA, okA, bB, okB);
        if (step + 1 < nsteps) run_step(step + 1, bB, okB, bA, okA);
    }
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        if (btb[cb] < 0) continue;
        const int t = s_tgt[btb[cb]];
        if (t < 0) continue;
        double *dst = a.Pf + ((size_t)(a.off + t) * a.ncols + c0 + bcl[cb]) * a.ldf;
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int rb_ = wr * RBW + i;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = row0 + rb_ * 16 + fk + 4 * v;
                if (rb_ < RB && row < rf) dst[row] = acc[i][cb][v];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Near field, compute.hpp:590-751, as a target-owner kernel: one wavefront
// per non-empty leaf computes, for each of its targets, the L2P of the
// leaf's local expansion plus the direct sum over the points of the 27
// neighbour leaves (self included, isinf -> 0), and adds both into the
// caller's output through perm.  No partner writes, no atomics.
// The neighbours' point ranges are concatenated into one index space and
// streamed in 64-point chunks; each chunk's kernel values K(t_i, s_j) are
// evaluated once into LDS (targets in passes of up to 16), then lane
// (i, g) accumulates columns g, g+G, ... of target i (G = 64 / TT).
constexpr int NV_SRC = 32;  // source points per chunk (LDS per wave bounds occupancy)
constexpr int NV_TGT = 16;  // targets per pass
constexpr int NV_COL = 32;  // columns per pass

template <int KID>
__global__ __launch_bounds__(64) void near_fp64(NearArgs a)
{
    __shared__ double sK[NV_TGT][NV_SRC + 1];
    __shared__ double sw[NV_COL][NV_SRC + 1];
    __shared__ double sx[NV_SRC], sy[NV_SRC], sz[NV_SRC];
    __shared__ double tx[NV_TGT], ty[NV_TGT], tz[NV_TGT];
    __shared__ double sS[3][16][NV_TGT];  // S_m of each target coordinate
    __shared__ int s_start[27], s_pre[28];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int n = a.n, n2 = n * n, n3 = n2 * n;
    const int t0 = a.leaf_start[b], nt = a.leaf_start[b + 1] - t0;
    const uint32_t cc = (uint32_t)a.coord[b];
    const int ix = cc & 1023, iy = (cc >> 10) & 1023, iz = (cc >> 20) & 1023;
    const int S = a.side;
    const double cen[3] = {a.ctab[ix], a.ctab[iy], a.ctab[iz]};
    const double ihalfL = 2.0 / a.len;
    if (lane < 27) {
        int x = ix + lane / 9 - 1, y = iy + (lane / 3) % 3 - 1, z = iz + lane % 3 - 1;
        int sb = (x >= 0 && y >= 0 && z >= 0 && x < S && y < S && z < S) ? a.map[((int64_t)x * S + y) * S + z] : -1;
        s_start[lane] = sb >= 0 ? a.leaf_start[sb] : 0;
        s_pre[lane + 1] = sb >= 0 ? a.leaf_start[sb + 1] - a.leaf_start[sb] : 0;
    }
    __syncthreads();
    if (lane == 0) {
        s_pre[0] = 0;
        for (int m = 0; m < 27; m++) s_pre[m + 1] += s_pre[m];
    }
    __syncthreads();
    const int ns = s_pre[27];

    for (int c0 = 0; c0 < a.ncols; c0 += NV_COL) {
        const int cn = min(NV_COL, a.ncols - c0);
        for (int i0 = 0; i0 < nt; i0 += NV_TGT) {
            const int tn = min(NV_TGT, nt - i0);
            const int TT = tn <= 1 ? 1 : tn <= 2 ? 2 : tn <= 4 ? 4 : tn <= 8 ? 8 : 16;
            const int G = 64 / TT;
            const int i = lane / G, g = lane % G;
            const bool active = i < tn;
            __syncthreads();
            if (lane < tn) {
                tx[lane] = a.x[t0 + i0 + lane];
                ty[lane] = a.y[t0 + i0 + lane];
                tz[lane] = a.z[t0 + i0 + lane];
            }
            __syncthreads();
            if (a.F != nullptr && lane < 3 * tn) {
                const int ti = lane / 3, ax = lane % 3;
                const double xv = ax == 0 ? tx[ti] : ax == 1 ? ty[ti] : tz[ti];
                sn_eval_to(n, a.Tkz, ihalfL * (xv - cen[ax]), &sS[ax][0][ti], NV_TGT);
            }
            __syncthreads();
            double acc[NV_COL / 4];
#pragma unroll
            for (int k = 0; k < NV_COL / 4; k++) acc[k] = 0;
            // L2P (compute.hpp:594-648)
            if (a.F != nullptr && active) {
                const double *Fb = a.F + ((size_t)(a.off + b) * a.ncols + c0) * n3;
#pragma unroll
                for (int k = 0; k < NV_COL / 4; k++) {
                    const int cl = g + G * k;
                    if (cl >= cn) break;
                    const double *F = Fb + (size_t)cl * n3;
                    double sum = 0;
                    int l = 0;
                    for (int l1 = 0; l1 < n; l1++)
                        for (int l2 = 0; l2 < n; l2++) {
                            const double t2 = sS[0][l1][i] * sS[1][l2][i];
                            for (int l3 = 0; l3 < n; l3++) sum += F[l++] * t2 * sS[2][l3][i];
                        }
                    acc[k] += sum;
                }
            }
            // P2P over the concatenated neighbour points (compute.hpp:650-737)
            for (int j0 = 0; j0 < ns; j0 += NV_SRC) {
                const int jn = min(NV_SRC, ns - j0);
                __syncthreads();
                if (lane < jn) {
                    const int q = j0 + lane;
                    int m = 0;
                    while (s_pre[m + 1] <= q) m++;
                    const int gi = s_start[m] + q - s_pre[m];
                    sx[lane] = a.x[gi];
                    sy[lane] = a.y[gi];
                    sz[lane] = a.z[gi];
                    for (int c = 0; c < cn; c++) sw[c][lane] = a.ws[(int64_t)(c0 + c) * a.N + gi];
                }
                __syncthreads();
                for (int q = lane; q < TT * NV_SRC; q += 64) {
                    const int ii = q / NV_SRC, jj = q % NV_SRC;
                    double kv = 0.0;
                    if (ii < tn && jj < jn) {
                        kv = keval<KID>(sx[jj] - tx[ii], sy[jj] - ty[ii], sz[jj] - tz[ii], a.kp);
                        if (isinf(kv)) kv = 0.0;
                    }
                    sK[ii][jj] = kv;
                }
                __syncthreads();
                if (active) {
                    for (int jj = 0; jj < jn; jj++) {
                        const double kv = sK[i][jj];
#pragma unroll
                        for (int k = 0; k < NV_COL / 4; k++) {
                            const int cl = g + G * k;
                            if (cl < cn) acc[k] += kv * sw[cl][jj];
                        }
                    }
                }
            }
            const int64_t oi = active ? a.perm[t0 + i0 + i] - a.out_lo : -1;
            if ((uint64_t)oi < (uint64_t)a.out_n) {
#pragma unroll
                for (int k = 0; k < NV_COL / 4; k++) {
                    const int cl = g + G * k;
                    if (cl < cn) a.out[(int64_t)(c0 + cl) * a.out_n + oi] += acc[k];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Near field in two kernels, for Chebyshev orders known at compile time.
//   l2p_fp64  L2P (compute.hpp:594-648): lp[k*ncols + c] = sum_l S_l(x_k)
//             F[c][l] for the targets k of each leaf, point-major, sorted
//             order.  8 columns of the leaf's F at a time are staged in LDS
//             with coalesced loads; lanes (target, column, part) split the
//             n^3 sum and reduce across the parts, so that leaves of one or
//             two points keep every lane busy.
//   p2p_fp64  P2P (compute.hpp:650-737) + scatter (:739-749) on top of lp.
//             One wave per leaf; lane (i, g) holds columns g, g+G, ... of
//             target i (G = 64 / targets per pass) in registers.  Source
//             chunks of 32 points (positions, point-major weights) are staged
//             in LDS while the next chunk is loaded.  Passes of up to 16
//             targets evaluate each K(t, s) once into LDS; wider passes
//             evaluate it per lane (at most 2 lanes per target).

constexpr int L2P_CC = 8;    // columns of F staged per step
constexpr int L2P_TGT = 32;  // targets per pass

// one level of the direct downward pass: one workgroup per segment of at
// most L2P_SEG points of one box; writes (or adds to) lp[k*ncols + c]
template <int N, typename TF>
__global__ __launch_bounds__(256) void l2p_fp64(DownArgs a)
{
    constexpr int N2 = N * N, N3 = N2 * N, FS = N3 + 1;
    __shared__ double Fs[L2P_CC * FS];
    __shared__ double sS[3][N][L2P_TGT];
    const int tid = threadIdx.x;
    const int b = a.seg_box[blockIdx.x], p0 = a.seg_p0[blockIdx.x];
    const int p1 = min(p0 + L2P_SEG, a.pstart[b + 1]);
    const uint32_t cc = (uint32_t)a.coord[b];
    const double ihalfL = 2.0 / a.len;
    const TF *Fb = (const TF *)a.F + (size_t)(a.off + b) * a.ncols * N3;
    for (int c0 = 0; c0 < a.ncols; c0 += L2P_CC) {
        const int cn = min(L2P_CC, a.ncols - c0);
        __syncthreads();  // the previous step's readers of Fs and sS are done
        for (int q = tid; q < cn * N3; q += 256) {
            const int qc = q / N3;
            Fs[qc * FS + q - qc * N3] = (double)Fb[(size_t)c0 * N3 + q];
        }
        for (int i0 = p0; i0 < p1; i0 += L2P_TGT) {
            const int tn = min(L2P_TGT, p1 - i0);
            const int tp = tn <= 1 ? 1 : tn <= 2 ? 2 : tn <= 4 ? 4 : tn <= 8 ? 8 : tn <= 16 ? 16 : 32;
            const int P = 32 / tp;  // parts per (target, column): tp * 8 * P = 256
            const int p = tid % P, c = (tid / P) % L2P_CC, i = tid / (P * L2P_CC);
            if (i0 != p0) __syncthreads();  // the previous chunk's readers of sS are done
            if (tid < 3 * L2P_TGT) {
                const int ti = tid % L2P_TGT, ax = tid / L2P_TGT;
                if (ti < tn) {
                    const double *xa = ax == 0 ? a.x : ax == 1 ? a.y : a.z;
                    const double cen = a.ctab[(cc >> (10 * ax)) & 1023];
                    double S[N];
                    sn_eval_n<N>(a.Tkz, ihalfL * (xa[i0 + ti] - cen), S);
#pragma unroll
                    for (int m = 0; m < N; m++) sS[ax][m][ti] = S[m];
                }
            }
            __syncthreads();
            const bool active = i < tn && c < cn;
            double s = 0.0;
            if (active) {
                double Sz[N];
#pragma unroll
                for (int m = 0; m < N; m++) Sz[m] = sS[2][m][i];
                for (int l12 = p; l12 < N2; l12 += P) {
                    const int l1 = l12 / N, l2 = l12 - l1 * N;
                    const double *f = &Fs[c * FS + l12 * N];
                    double s3 = 0.0;
#pragma unroll
                    for (int l3 = 0; l3 < N; l3++) s3 += f[l3] * Sz[l3];
                    s += sS[0][l1][i] * sS[1][l2][i] * s3;
                }
            }
            for (int o = 1; o < P; o <<= 1) s += __shfl_xor(s, o);
            if (active && p == 0) {
                double *dst = a.lp + (int64_t)(i0 + i) * a.ncols + c0 + c;
                *dst = a.accumulate ? *dst + s : s;
            }
        }
    }
}

// The same evaluation with FP32 local expansions (the mixed modes): all
// columns of a box's F staged at once (32 per step, 16 from order 7), as
// FP32 in LDS, so that a box with few points costs one load of F and two
// barriers; lanes take (target, column) pairs of 8 targets at a time and
// each sums the whole n^3 expansion in FP64.
template <int N>
__global__ __launch_bounds__(256, 5) void l2p_f32(DownArgs a)
{
    constexpr int N2 = N * N, N3 = N2 * N, FS = N3 + 1;
    constexpr int CC = N <= 6 ? 32 : 16, TG = 256 / CC;
    __shared__ float Fs[CC * FS];
    __shared__ double sS[3][N][TG];
    const int tid = threadIdx.x;
    const int b = a.seg_box[blockIdx.x], p0 = a.seg_p0[blockIdx.x];
    const int p1 = min(p0 + L2P_SEG, a.pstart[b + 1]);
    const uint32_t cc = (uint32_t)a.coord[b];
    const double ihalfL = 2.0 / a.len;
    const float *Fb = (const float *)a.F + (size_t)(a.off + b) * a.ncols * N3;
    const int i = tid / CC, c = tid % CC;
    for (int c0 = 0; c0 < a.ncols; c0 += CC) {
        const int cn = min(CC, a.ncols - c0);
        __syncthreads();  // the previous step's readers of Fs and sS are done
        for (int q = tid; q < cn * N3; q += 256) {
            const int qc = q / N3;
            Fs[qc * FS + q - qc * N3] = Fb[(size_t)c0 * N3 + q];
        }
        for (int i0 = p0; i0 < p1; i0 += TG) {
            const int tn = min(TG, p1 - i0);
            if (i0 != p0) __syncthreads();  // the previous chunk's readers of sS are done
            if (tid < 3 * TG) {
                const int ti = tid % TG, ax = tid / TG;
                if (ti < tn) {
                    const double *xa = ax == 0 ? a.x : ax == 1 ? a.y : a.z;
                    const double cen = a.ctab[(cc >> (10 * ax)) & 1023];
                    double S[N];
                    sn_eval_n<N>(a.Tkz, ihalfL * (xa[i0 + ti] - cen), S);
#pragma unroll
                    for (int m = 0; m < N; m++) sS[ax][m][ti] = S[m];
                }
            }
            __syncthreads();
            if (i < tn && c < cn) {
                double Sz[N];
#pragma unroll
                for (int m = 0; m < N; m++) Sz[m] = sS[2][m][i];
                const float *f = &Fs[c * FS];
                double s = 0.0;
                for (int l1 = 0; l1 < N; l1++) {
                    const double s1 = sS[0][l1][i];
                    double s2 = 0.0;
#pragma unroll
                    for (int l2 = 0; l2 < N; l2++) {
                        double s3 = 0.0;
#pragma unroll
                        for (int l3 = 0; l3 < N; l3++) s3 += (double)f[(l1 * N + l2) * N + l3] * Sz[l3];
                        s2 += sS[1][l2][i] * s3;
                    }
                    s += s1 * s2;
                }
                double *dst = a.lp + (int64_t)(i0 + i) * a.ncols + c0 + c;
                *dst = a.accumulate ? *dst + s : s;
            }
        }
    }
}

// Direct upward pass.  P2M at the leaves followed by M2M (compute.hpp:
// 234-333, :956-1034) equals P2M at every level straight from the points:
// the child-to-parent transfer interpolates the parent's Chebyshev
// polynomials S_m, which have degree n-1, at the child's nodes, which is
// exact.  One workgroup per work item (a box of levels 2..L, or a segment of
// P2M_SEG points of a larger box); thread pencils (column c, l1, l2) hold
// the n values of l3 in registers.
template <int N, typename TS>
__global__ __launch_bounds__(256) void p2m_levels(UpArgs a)
{
    constexpr int N2 = N * N, N3 = N2 * N, PTS = 64, PP = PTS + 1;
    constexpr int CC = N <= 6 ? 32 : 16;        // columns per pass
    constexpr int PR = (CC * N2 + 255) / 256;   // pencils per thread
    // LDS: the point tables of a pass, then (FP32 expansions) the pass's
    // expansion, staged so that it is stored with whole 16-byte writes
    constexpr int U_SS = 3 * N * PP, U_SW = PTS * (CC + 1), U_SF = (CC * N3 + 1) / 2;
    constexpr int U_ALL = U_SS + U_SW > U_SF ? U_SS + U_SW : U_SF;
    __shared__ double smem[U_ALL];
    double(*sS)[N][PP] = reinterpret_cast<double(*)[N][PP]>(smem);
    double(*sw)[CC + 1] = reinterpret_cast<double(*)[CC + 1]>(smem + U_SS);
    float *sF = reinterpret_cast<float *>(smem);
    const int tid = threadIdx.x;
    const int4 it = a.items[blockIdx.x];
    const int li = it.x, b = it.y, p0 = it.z;
    const int p1 = min(p0 + P2M_SEG, a.lt.pstart[li][b + 1]);
    const uint32_t cc = (uint32_t)a.lt.coord[li][b];
    const double *ctab = a.lt.ctab[li];
    const double cen[3] = {ctab[cc & 1023], ctab[(cc >> 10) & 1023], ctab[(cc >> 20) & 1023]};
    const double ihalfL = 2.0 / a.lt.len[li];
    // split boxes write FP64 partial sums to scratch slots, others S directly
    double *Sp = it.w >= 0 ? a.scratch + (size_t)it.w * a.ncols * N3 : nullptr;
    TS *Sb = it.w >= 0 ? nullptr : (TS *)a.S + (size_t)(a.lt.off[li] + b) * a.ncols * N3;
    for (int c0 = 0; c0 < a.ncols; c0 += CC) {
        const int cn = min(CC, a.ncols - c0);
        double acc[PR][N];
#pragma unroll
        for (int r = 0; r < PR; r++)
#pragma unroll
            for (int m = 0; m < N; m++) acc[r][m] = 0.0;
        for (int j0 = p0; j0 < p1; j0 += PTS) {
            const int jn = min(PTS, p1 - j0);
            __syncthreads();  // the previous chunk is consumed
            if (tid < 3 * PTS) {
                const int pp = tid % PTS, ax = tid / PTS;
                double S[N];
                if (pp < jn) {
                    const double *xa = ax == 0 ? a.x : ax == 1 ? a.y : a.z;
                    sn_eval_n<N>(a.Tkz, ihalfL * (xa[j0 + pp] - cen[ax]), S);
                } else {
#pragma unroll
                    for (int m = 0; m < N; m++) S[m] = 0.0;
                }
#pragma unroll
                for (int m = 0; m < N; m++) sS[ax][m][pp] = S[m];
            }
            for (int q = tid; q < PTS * CC; q += 256) {
                const int pp = q / CC, c = q - pp * CC;
                sw[pp][c] = (pp < jn && c < cn) ? a.wsT[(int64_t)(j0 + pp) * a.ncols + c0 + c] : 0.0;
            }
            __syncthreads();
            for (int pp = 0; pp < jn; pp++) {
                double sz[N];
#pragma unroll
                for (int m = 0; m < N; m++) sz[m] = sS[2][m][pp];
#pragma unroll
                for (int r = 0; r < PR; r++) {
                    const int q = tid + 256 * r;
                    if (q < CC * N2) {
                        const int c = q / N2, l12 = q - c * N2, l1 = l12 / N, l2 = l12 - l1 * N;
                        const double t = sw[pp][c] * sS[0][l1][pp] * sS[1][l2][pp];
#pragma unroll
                        for (int m = 0; m < N; m++) acc[r][m] += t * sz[m];
                    }
                }
            }
        }
        if constexpr (sizeof(TS) == 4) {
            if (!Sp) {  // FP32 S: through LDS, then contiguous (16-byte where aligned) stores
                __syncthreads();  // the pass's point tables are consumed
#pragma unroll
                for (int r = 0; r < PR; r++) {
                    const int q = tid + 256 * r;
                    if (q < CC * N2) {
                        const int c = q / N2, l12 = q - c * N2;
#pragma unroll
                        for (int m = 0; m < N; m++) sF[c * N3 + l12 * N + m] = (float)acc[r][m];
                    }
                }
                __syncthreads();
                float *dst = (float *)Sb + (size_t)c0 * N3;
                const int n = cn * N3;
                if ((n & 3) == 0 && ((uintptr_t)dst & 15) == 0) {
                    for (int q = tid; q < n / 4; q += 256) ((float4 *)dst)[q] = ((const float4 *)sF)[q];
                } else {
                    for (int q = tid; q < n; q += 256) dst[q] = sF[q];
                }
                continue;
            }
        }
#pragma unroll
        for (int r = 0; r < PR; r++) {
            const int q = tid + 256 * r;
            if (q < CC * N2) {
                const int c = q / N2, l12 = q - c * N2;
                if (c < cn) {
                    const size_t at = (size_t)(c0 + c) * N3 + l12 * N;
                    if (Sp) {
#pragma unroll
                        for (int m = 0; m < N; m++) Sp[at + m] = acc[r][m];
                    } else {
#pragma unroll
                        for (int m = 0; m < N; m++) Sb[at + m] = (TS)acc[r][m];
                    }
                }
            }
        }
    }
}

// S of the split boxes: the sum of their segments' partial expansions, in
// segment order
template <typename TS>
__global__ void p2m_reduce(UpReduce r)
{
    const int64_t len = (int64_t)r.ncols * r.n3;
    const int64_t total = (int64_t)r.nred * len;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = q / len, e = q - k * len;
        const int4 rd = r.reds[k];
        double s = 0.0;
        for (int i = 0; i < rd.z; i++) s += r.scratch[(int64_t)(rd.y + i) * len + e];
        ((TS *)r.S)[(int64_t)rd.x * len + e] = (TS)s;
    }
}

constexpr int P2P_SC = 32;  // source points per LDS chunk
constexpr int P2P_NC = 32;  // columns per pass
constexpr int P2P_SK = 16;  // passes of at most this many targets share K through LDS

struct P2PShared {
    double sw[P2P_SC][P2P_NC + 1];
    double sx[P2P_SC], sy[P2P_SC], sz[P2P_SC];
    double sK[P2P_SK][P2P_SC + 1];
    double tx[P2P_SK], ty[P2P_SK], tz[P2P_SK];
    int s_start[27], s_pre[28];
};

template <int KID, int TT>
__device__ __forceinline__ void p2p_pass(const NearArgs &a, P2PShared &sh, int t0, int i0, int tn, int c0,
                                         int cn, int ns)
{
    constexpr int G = 64 / TT, CPL = P2P_NC / G, WR = P2P_SC * P2P_NC / 64;
    constexpr bool SHARED_K = TT <= P2P_SK;
    const int lane = threadIdx.x, i = lane / G, g = lane % G;
    const bool active = i < tn;
    const int64_t ti = t0 + i0 + (active ? i : 0);
    const double tx = a.x[ti], ty = a.y[ti], tz = a.z[ti];
    if constexpr (SHARED_K) {
        __syncthreads();
        if (lane < tn) {
            sh.tx[lane] = a.x[t0 + i0 + lane];
            sh.ty[lane] = a.y[t0 + i0 + lane];
            sh.tz[lane] = a.z[t0 + i0 + lane];
        }
    }
    double acc[CPL];
#pragma unroll
    for (int k = 0; k < CPL; k++) {
        const int cl = g + G * k;
        acc[k] = (a.lp && active && cl < cn) ? a.lp[ti * a.ncols + c0 + cl] : 0.0;
    }
    double pw[WR], px = 0, py = 0, pz = 0;
    auto load = [&](int j0) {
        const int jn = min(P2P_SC, ns - j0);
        int gi = 0;
        if (lane < jn) {
            const int q = j0 + lane;
            int m = 0;
            while (sh.s_pre[m + 1] <= q) m++;
            gi = sh.s_start[m] + q - sh.s_pre[m];
            px = a.x[gi];
            py = a.y[gi];
            pz = a.z[gi];
        }
#pragma unroll
        for (int r = 0; r < WR; r++) {
            const int u = lane + 64 * r, j = u / P2P_NC, c = u % P2P_NC;
            const int gj = __shfl(gi, j);
            pw[r] = (j < jn && c < cn) ? a.wsT[(int64_t)gj * a.ncols + c0 + c] : 0.0;
        }
    };
    if (ns > 0) load(0);
    for (int j0 = 0; j0 < ns; j0 += P2P_SC) {
        const int jn = min(P2P_SC, ns - j0);
        __syncthreads();  // the previous chunk is consumed
#pragma unroll
        for (int r = 0; r < WR; r++) {
            const int u = lane + 64 * r;
            sh.sw[u / P2P_NC][u % P2P_NC] = pw[r];
        }
        if (lane < P2P_SC) {
            sh.sx[lane] = px;
            sh.sy[lane] = py;
            sh.sz[lane] = pz;
        }
        __syncthreads();
        if (j0 + P2P_SC < ns) load(j0 + P2P_SC);  // in flight during the sums
        if constexpr (SHARED_K) {
            for (int q = lane; q < TT * P2P_SC; q += 64) {
                const int ii = q / P2P_SC, jj = q % P2P_SC;
                double kv = 0.0;
                if (ii < tn && jj < jn) {
                    kv = keval<KID>(sh.sx[jj] - sh.tx[ii], sh.sy[jj] - sh.ty[ii], sh.sz[jj] - sh.tz[ii], a.kp);
                    if (isinf(kv)) kv = 0.0;
                }
                sh.sK[ii][jj] = kv;
            }
            __syncthreads();
            if (active)
                for (int jj = 0; jj < jn; jj++) {
                    const double kv = sh.sK[i][jj];
#pragma unroll
                    for (int k = 0; k < CPL; k++) acc[k] += kv * sh.sw[jj][g + G * k];
                }
        } else {
            if (active)
                for (int jj = 0; jj < jn; jj++) {
                    double kv = keval<KID>(sh.sx[jj] - tx, sh.sy[jj] - ty, sh.sz[jj] - tz, a.kp);
                    if (isinf(kv)) kv = 0.0;
#pragma unroll
                    for (int k = 0; k < CPL; k++) acc[k] += kv * sh.sw[jj][g + G * k];
                }
        }
    }
    const int64_t oi = active ? a.perm[ti] - a.out_lo : -1;
    if ((uint64_t)oi < (uint64_t)a.out_n) {
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            const int cl = g + G * k;
            if (cl < cn) a.out[(int64_t)(c0 + cl) * a.out_n + oi] += acc[k];
        }
    }
}

// one workgroup per leaf of a list whose leaves hold at most TT targets (or
// more, in passes of TT, for the widest class)
template <int KID, int TT>
__global__ __launch_bounds__(64) void p2p_fp64(NearArgs a, const int32_t *__restrict__ leaves)
{
    __shared__ P2PShared sh;
    const int b = leaves[blockIdx.x], lane = threadIdx.x;
    const int t0 = a.leaf_start[b], nt = a.leaf_start[b + 1] - t0;
    const uint32_t cc = (uint32_t)a.coord[b];
    const int ix = cc & 1023, iy = (cc >> 10) & 1023, iz = (cc >> 20) & 1023;
    const int S = a.side;
    if (lane < 27) {
        const int x = ix + lane / 9 - 1, y = iy + (lane / 3) % 3 - 1, z = iz + lane % 3 - 1;
        const int sb = (x >= 0 && y >= 0 && z >= 0 && x < S && y < S && z < S) ? a.map[((int64_t)x * S + y) * S + z] : -1;
        sh.s_start[lane] = sb >= 0 ? a.leaf_start[sb] : 0;
        sh.s_pre[lane + 1] = sb >= 0 ? a.leaf_start[sb + 1] - a.leaf_start[sb] : 0;
    }
    __syncthreads();
    if (lane == 0) {
        sh.s_pre[0] = 0;
        for (int m = 0; m < 27; m++) sh.s_pre[m + 1] += sh.s_pre[m];
    }
    __syncthreads();
    const int ns = sh.s_pre[27];
    for (int c0 = 0; c0 < a.ncols; c0 += P2P_NC) {
        const int cn = min(P2P_NC, a.ncols - c0);
        for (int i0 = 0; i0 < nt; i0 += TT) p2p_pass<KID, TT>(a, sh, t0, i0, min(TT, nt - i0), c0, cn, ns);
    }
}

// Near field of the dense leaves on the FP64 matrix cores (leaves of 16 or
// more targets).  K (targets x sources) times W (sources x columns) is a
// product of 16 x 4 blocks of K with 4 x 16 blocks of W: the A fragment of
// v_mfma_f64_16x16x4f64 holds one value per lane, so each lane evaluates
// the kernel for its (target, source) entry of every block and the matrix
// cores accumulate 16*TB targets x 32 columns in FP64; the VALU only
// evaluates kernels.  Sources are staged 32 at a time in LDS as in
// p2p_fp64; the result (with the leaf's L2P sums) is added to the output
// through perm, in FP64 as there.  KF32 (mixed modes): the kernel values in
// FP32 (keval_f32), the sums still FP64.
constexpr int PM_SC = 32;

template <int KID, int TB, bool KF32>
__global__ __launch_bounds__(64) void p2p_mfma(NearArgs a, const int32_t *__restrict__ leaves)
{
    __shared__ double sw[PM_SC][33];
    __shared__ double sx[PM_SC], sy[PM_SC], sz[PM_SC];
    __shared__ int s_start[27], s_pre[28];
    const int lane = threadIdx.x, fr = lane & 15, fk = lane >> 4;
    const int b = leaves[blockIdx.x];
    const int t0 = a.leaf_start[b], nt = a.leaf_start[b + 1] - t0;
    const uint32_t cc = (uint32_t)a.coord[b];
    const int ix = cc & 1023, iy = (cc >> 10) & 1023, iz = (cc >> 20) & 1023;
    const int S = a.side;
    if (lane < 27) {
        const int x = ix + lane / 9 - 1, y = iy + (lane / 3) % 3 - 1, z = iz + lane % 3 - 1;
        const int sb = (x >= 0 && y >= 0 && z >= 0 && x < S && y < S && z < S) ? a.map[((int64_t)x * S + y) * S + z] : -1;
        s_start[lane] = sb >= 0 ? a.leaf_start[sb] : 0;
        s_pre[lane + 1] = sb >= 0 ? a.leaf_start[sb + 1] - a.leaf_start[sb] : 0;
    }
    __syncthreads();
    if (lane == 0) {
        s_pre[0] = 0;
        for (int m = 0; m < 27; m++) s_pre[m + 1] += s_pre[m];
    }
    __syncthreads();
    const int ns = s_pre[27];
    for (int c0 = 0; c0 < a.ncols; c0 += 32) {
        const int cn = min(32, a.ncols - c0);
        for (int i0 = 0; i0 < nt; i0 += 16 * TB) {
            // this lane's target row fr of each 16-target block (rows past
            // the leaf repeat its last target; their sums are not stored)
            double tx[TB], ty[TB], tz[TB];
#pragma unroll
            for (int tb = 0; tb < TB; tb++) {
                const int64_t ti = t0 + min(i0 + tb * 16 + fr, nt - 1);
                tx[tb] = a.x[ti];
                ty[tb] = a.y[ti];
                tz[tb] = a.z[ti];
            }
            // accumulators: D[fk + 4v][fr] of block (tb, cb), with the L2P sums
            dbl4 acc[TB][2];
#pragma unroll
            for (int tb = 0; tb < TB; tb++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int ti = i0 + tb * 16 + fk + 4 * v, col = cb * 16 + fr;
                        acc[tb][cb][v] = (a.lp && ti < nt && col < cn) ? a.lp[(int64_t)(t0 + ti) * a.ncols + c0 + col] : 0.0;
                    }
            for (int j0 = 0; j0 < ns; j0 += PM_SC) {
                const int jn = min(PM_SC, ns - j0);
                __syncthreads();  // the previous chunk is consumed
                int gi = 0;
                if (lane < jn) {
                    const int q = j0 + lane;
                    int m = 0;
                    while (s_pre[m + 1] <= q) m++;
                    gi = s_start[m] + q - s_pre[m];
                    sx[lane] = a.x[gi];
                    sy[lane] = a.y[gi];
                    sz[lane] = a.z[gi];
                }
#pragma unroll
                for (int r = 0; r < PM_SC * 32 / 64; r++) {
                    const int u = lane + 64 * r, j = u / 32, c = u % 32;
                    const int gj = __shfl(gi, j);
                    sw[j][c] = (j < jn && c < cn) ? a.wsT[(int64_t)gj * a.ncols + c0 + c] : 0.0;
                }
                __syncthreads();
                for (int k0 = 0; k0 < jn; k0 += 4) {
                    const int jj = k0 + fk;  // this lane's source of the 4-block (rows past jn carry zero weights)
                    const int js = min(jj, jn - 1);
                    const double px = sx[js], py = sy[js], pz = sz[js];
                    const double w0 = sw[jj & (PM_SC - 1)][fr], w1 = sw[jj & (PM_SC - 1)][16 + fr];
                    const double b0 = jj < jn ? w0 : 0.0, b1 = jj < jn ? w1 : 0.0;
#pragma unroll
                    for (int tb = 0; tb < TB; tb++) {
                        double kv = KF32 ? keval_f32<KID>(px - tx[tb], py - ty[tb], pz - tz[tb], a.kp)
                                         : keval<KID>(px - tx[tb], py - ty[tb], pz - tz[tb], a.kp);
                        if (isinf(kv) || jj >= jn) kv = 0.0;  // (padding sources add nothing)
                        acc[tb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv, b0, acc[tb][0], 0, 0, 0);
                        acc[tb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv, b1, acc[tb][1], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int tb = 0; tb < TB; tb++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int ti = i0 + tb * 16 + fk + 4 * v;
                    if (ti >= nt) continue;
                    const int64_t oi = a.perm[t0 + ti] - a.out_lo;
                    if ((uint64_t)oi >= (uint64_t)a.out_n) continue;
#pragma unroll
                    for (int cb = 0; cb < 2; cb++) {
                        const int col = cb * 16 + fr;
                        if (col < cn) a.out[(int64_t)(c0 + col) * a.out_n + oi] += acc[tb][cb][v];
                    }
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Position gradients of the FMM matvec (no reference counterpart: the
// reference has no derivatives).  For each target i of a leaf:
//   g_i += sum_c v_ic [ sum_l F_c[l] grad S_l(x_i) + sum_{j near} grad_1 K(x_i, x_j) u_jc ]
// where grad S_l is the derivative of the tensor-product Chebyshev
// interpolant (S_m'(u) = (2/n) sum_k T_k(x_m) k U_{k-1}(u), times 2/len) and
// grad_1 K the kernel's gradient in its first (target) argument.  Called
// with (v, F, u) = (dL/dout, forward leaf expansions, weights) for the
// target side and (weights, transposed-FMM leaf expansions, dL/dout) for the
// source side (symmetric kernels: d/dy K(x, y) = grad_1 K(y, x)).
template <int KID>
__device__ __forceinline__ void kgrad(double dx, double dy, double dz, const KParams &p, double &gx, double &gy,
                                      double &gz)
{
    // d = source - target; returns grad wrt the target position
    const double r2 = dx * dx + dy * dy + dz * dz;
    double s = 0.0;  // grad = s * d
    if (r2 == 0.0) {
        gx = gy = gz = 0.0;
        return;
    }
    if constexpr (KID == PBB_KERNEL_LAPLACIAN) { double ri = 1.0 / sqrt(r2); s = ri * ri * ri; }
    else if constexpr (KID == PBB_KERNEL_ONEOVERR2) { s = 2.0 / (r2 * r2); }
    else if constexpr (KID == PBB_KERNEL_GAUSSIAN) { double a2 = p.p[0] * p.p[0]; s = 2.0 / a2 * exp(-r2 / a2); }
    else if constexpr (KID == PBB_KERNEL_EXPONENTIAL) { double r = sqrt(r2); s = p.p[1] * exp(-r / p.p[0]) / (p.p[0] * r); }
    else if constexpr (KID == PBB_KERNEL_SOFT_LAPLACIAN) { double q = 1.0 / sqrt(r2 + p.p[0]); s = q * q * q; }
    else if constexpr (KID == PBB_KERNEL_QUADRIC || KID == PBB_KERNEL_R2) { s = -2.0; }
    else if constexpr (KID == PBB_KERNEL_INV_QUADRIC) { double q = 1.0 / (1.0 + r2); s = 2.0 * q * q; }
    else if constexpr (KID == PBB_KERNEL_LOGARITHM) { s = -1.0 / r2; }
    else if constexpr (KID == PBB_KERNEL_THIN_PLATE) { s = -(log(r2) + 1.0); }
    else if constexpr (KID == PBB_KERNEL_EXP_ANISO) {
        const double rho = sqrt(p.p[1] * dx * dx + p.p[2] * dy * dy + p.p[3] * dz * dz);
        const double f = rho > 0 ? p.p[0] * exp(-rho) / rho : 0.0;
        gx = f * p.p[1] * dx; gy = f * p.p[2] * dy; gz = f * p.p[3] * dz;
        return;
    }
    gx = s * dx; gy = s * dy; gz = s * dz;
}

// S_m(u) and S_m'(u), m < n
__device__ __forceinline__ void sn_eval_d(int n, const double *__restrict__ Tkz, double u, double *S, double *dS)
{
    double T[16], dT[16];
    T[0] = 1.0; dT[0] = 0.0;
    if (n > 1) { T[1] = u; dT[1] = 1.0; }
    double Um2 = 1.0, Um1 = 2.0 * u;  // U_0, U_1
    for (int k = 2; k < n; k++) {
        T[k] = 2.0 * u * T[k - 1] - T[k - 2];
        dT[k] = k * Um1;  // T_k' = k U_{k-1}
        const double Uk = 2.0 * u * Um1 - Um2;
        Um2 = Um1;
        Um1 = Uk;
    }
    const double pfac = 2.0 / n;
    for (int m = 0; m < n; m++) {
        const double *Tk = Tkz + m * n;
        double s = 0.5 * Tk[0], d = 0.0;
        for (int k = 1; k < n; k++) {
            s += Tk[k] * T[k];
            d += Tk[k] * dT[k];
        }
        S[m] = pfac * s;
        dS[m] = pfac * d;
    }
}

template <int KID>
__global__ __launch_bounds__(64) void near_grad_fp64(NearGradArgs a)
{
    __shared__ double sG[3][NV_TGT][NV_SRC + 1];
    __shared__ double su[NV_COL][NV_SRC + 1];
    __shared__ double sx[NV_SRC], sy[NV_SRC], sz[NV_SRC];
    __shared__ double tx[NV_TGT], ty[NV_TGT], tz[NV_TGT];
    __shared__ double red[3][64];
    __shared__ double sS[6][16][NV_TGT];  // S_m and S_m' of each target coordinate
    __shared__ int s_start[27], s_pre[28];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int n = a.n, n2 = n * n, n3 = n2 * n;
    const int t0 = a.leaf_start[b], nt = a.leaf_start[b + 1] - t0;
    const uint32_t cc = (uint32_t)a.coord[b];
    const int ix = cc & 1023, iy = (cc >> 10) & 1023, iz = (cc >> 20) & 1023;
    const int S = a.side;
    const double cen[3] = {a.ctab[ix], a.ctab[iy], a.ctab[iz]};
    const double ihalfL = 2.0 / a.len;
    if (lane < 27) {
        int x = ix + lane / 9 - 1, y = iy + (lane / 3) % 3 - 1, z = iz + lane % 3 - 1;
        int sb = (x >= 0 && y >= 0 && z >= 0 && x < S && y < S && z < S) ? a.map[((int64_t)x * S + y) * S + z] : -1;
        s_start[lane] = sb >= 0 ? a.leaf_start[sb] : 0;
        s_pre[lane + 1] = sb >= 0 ? a.leaf_start[sb + 1] - a.leaf_start[sb] : 0;
    }
    __syncthreads();
    if (lane == 0) {
        s_pre[0] = 0;
        for (int m = 0; m < 27; m++) s_pre[m + 1] += s_pre[m];
    }
    __syncthreads();
    const int ns = s_pre[27];
    for (int i0 = 0; i0 < nt; i0 += NV_TGT) {
        const int tn = min(NV_TGT, nt - i0);
        const int TT = tn <= 1 ? 1 : tn <= 2 ? 2 : tn <= 4 ? 4 : tn <= 8 ? 8 : 16;
        const int G = 64 / TT;
        const int i = lane / G, g = lane % G;
        const bool active = i < tn;
        __syncthreads();
        if (lane < tn) {
            tx[lane] = a.x[t0 + i0 + lane];
            ty[lane] = a.y[t0 + i0 + lane];
            tz[lane] = a.z[t0 + i0 + lane];
        }
        __syncthreads();
        if (a.F != nullptr && lane < 3 * tn) {
            const int ti = lane / 3, ax = lane % 3;
            const double xv = ax == 0 ? tx[ti] : ax == 1 ? ty[ti] : tz[ti];
            double Sv[16], dSv[16];
            sn_eval_d(n, a.Tkz, ihalfL * (xv - cen[ax]), Sv, dSv);
            for (int m = 0; m < n; m++) {
                sS[ax][m][ti] = Sv[m];
                sS[3 + ax][m][ti] = dSv[m];
            }
        }
        __syncthreads();
        double gx = 0, gy = 0, gz = 0;
        for (int c0 = 0; c0 < a.ncols; c0 += NV_COL) {
            const int cn = min(NV_COL, a.ncols - c0);
            double vv[NV_COL / 4];
#pragma unroll
            for (int k = 0; k < NV_COL / 4; k++) {
                const int cl = g + G * k;
                vv[k] = (active && cl < cn) ? a.v[(int64_t)(c0 + cl) * a.N + t0 + i0 + i] : 0.0;
            }
            if (a.F != nullptr && active) {
                const double *Sx = &sS[0][0][i], *Sy = &sS[1][0][i], *Sz = &sS[2][0][i];
                const double *dSx = &sS[3][0][i], *dSy = &sS[4][0][i], *dSz = &sS[5][0][i];
                const double *Fb = a.F + ((size_t)b * a.ncols + c0) * n3;
#pragma unroll
                for (int k = 0; k < NV_COL / 4; k++) {
                    const int cl = g + G * k;
                    if (cl >= cn) break;
                    const double *F = Fb + (size_t)cl * n3;
                    double ax = 0, ay = 0, az = 0;
                    int l = 0;
                    for (int l1 = 0; l1 < n; l1++)
                        for (int l2 = 0; l2 < n; l2++)
                            for (int l3 = 0; l3 < n; l3++) {
                                const double f = F[l++];
                                const int o1 = l1 * NV_TGT, o2 = l2 * NV_TGT, o3 = l3 * NV_TGT;
                                ax += f * dSx[o1] * Sy[o2] * Sz[o3];
                                ay += f * Sx[o1] * dSy[o2] * Sz[o3];
                                az += f * Sx[o1] * Sy[o2] * dSz[o3];
                            }
                    gx += vv[k] * ax * ihalfL;
                    gy += vv[k] * ay * ihalfL;
                    gz += vv[k] * az * ihalfL;
                }
            }
            for (int j0 = 0; j0 < ns; j0 += NV_SRC) {
                const int jn = min(NV_SRC, ns - j0);
                __syncthreads();
                if (lane < jn) {
                    const int q = j0 + lane;
                    int m = 0;
                    while (s_pre[m + 1] <= q) m++;
                    const int gi = s_start[m] + q - s_pre[m];
                    sx[lane] = a.x[gi];
                    sy[lane] = a.y[gi];
                    sz[lane] = a.z[gi];
                    for (int c = 0; c < cn; c++) su[c][lane] = a.u[(int64_t)(c0 + c) * a.N + gi];
                }
                __syncthreads();
                for (int q = lane; q < TT * NV_SRC; q += 64) {
                    const int ii = q / NV_SRC, jj = q % NV_SRC;
                    double kx = 0, ky = 0, kz = 0;
                    if (ii < tn && jj < jn) kgrad<KID>(sx[jj] - tx[ii], sy[jj] - ty[ii], sz[jj] - tz[ii], a.kp, kx, ky, kz);
                    sG[0][ii][jj] = kx;
                    sG[1][ii][jj] = ky;
                    sG[2][ii][jj] = kz;
                }
                __syncthreads();
                if (active) {
                    for (int jj = 0; jj < jn; jj++) {
                        double w = 0;
#pragma unroll
                        for (int k = 0; k < NV_COL / 4; k++) {
                            const int cl = g + G * k;
                            if (cl < cn) w += vv[k] * su[cl][jj];
                        }
                        gx += w * sG[0][i][jj];
                        gy += w * sG[1][i][jj];
                        gz += w * sG[2][i][jj];
                    }
                }
            }
        }
        // reduce the G lanes of each target
        __syncthreads();
        red[0][lane] = gx;
        red[1][lane] = gy;
        red[2][lane] = gz;
        __syncthreads();
        if (lane < tn) {
            double sxg = 0, syg = 0, szg = 0;
            for (int q = 0; q < G; q++) {
                sxg += red[0][lane * G + q];
                syg += red[1][lane * G + q];
                szg += red[2][lane * G + q];
            }
            const int64_t oi = a.perm[t0 + i0 + lane];
            a.grad[3 * oi] += sxg;
            a.grad[3 * oi + 1] += syg;
            a.grad[3 * oi + 2] += szg;
        }
    }
}

// ---------------------------------------------------------------------------
// Direct sum (DirectCalc3D, test.hpp:18-34, isinf -> 0): partial sums of
// one slice of the sources per workgroup row, reduced by a second kernel.
//   part[(split*ncols + c)*T + i] = sum_{j in slice} K(t_i, s_j) w[c*Ns + j]
constexpr int DS_TGT = 64, DS_SRC = 64, DS_CPT = 8;

template <int KID>
__global__ __launch_bounds__(256) void direct_partial(DirectArgs a)
{
    __shared__ double sx[DS_SRC], sy[DS_SRC], sz[DS_SRC];
    __shared__ double sw[4 * DS_CPT][DS_SRC];
    const int tid = threadIdx.x, slot = tid / 4, g = tid % 4;
    const int64_t i = (int64_t)blockIdx.x * DS_TGT + slot;
    const int split = blockIdx.y;
    const int64_t per = (a.Ns + a.nsplit - 1) / a.nsplit;
    const int64_t j0s = split * per, j1s = min(a.Ns, j0s + per);
    const bool active = i < a.T;
    double xi = 0, yi = 0, zi = 0;
    if (active) { xi = a.tgt[3 * i]; yi = a.tgt[3 * i + 1]; zi = a.tgt[3 * i + 2]; }
    for (int c0 = 0; c0 < a.ncols; c0 += 4 * DS_CPT) {
        const int cn = min(4 * DS_CPT, a.ncols - c0);
        double acc[DS_CPT];
#pragma unroll
        for (int k = 0; k < DS_CPT; k++) acc[k] = 0;
        for (int64_t j0 = j0s; j0 < j1s; j0 += DS_SRC) {
            const int jn = (int)min((int64_t)DS_SRC, j1s - j0);
            __syncthreads();
            if (tid < jn) {
                sx[tid] = a.src[3 * (j0 + tid)];
                sy[tid] = a.src[3 * (j0 + tid) + 1];
                sz[tid] = a.src[3 * (j0 + tid) + 2];
            }
            for (int q = tid; q < cn * DS_SRC; q += 256) {
                int c = q / DS_SRC, p = q % DS_SRC;
                sw[c][p] = p < jn ? a.w[(int64_t)(c0 + c) * a.Ns + j0 + p] : 0.0;
            }
            __syncthreads();
            if (active) {
                for (int p = 0; p < jn; p++) {
                    double kv = keval<KID>(sx[p] - xi, sy[p] - yi, sz[p] - zi, a.kp);
                    if (isinf(kv)) kv = 0.0;
#pragma unroll
                    for (int k = 0; k < DS_CPT; k++) {
                        int cl = g + 4 * k;
                        if (cl < cn) acc[k] += kv * sw[cl][p];
                    }
                }
            }
        }
        if (active) {
#pragma unroll
            for (int k = 0; k < DS_CPT; k++) {
                int cl = g + 4 * k;
                if (cl < cn) a.part[((int64_t)split * a.ncols + c0 + cl) * a.T + i] = acc[k];
            }
        }
    }
}

__global__ void direct_reduce(const double *__restrict__ part, int nsplit, int64_t total,
                              double *__restrict__ out)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        double s = 0;
        for (int k = 0; k < nsplit; k++) s += part[k * total + q];
        out[q] = s;
    }
}

// ---------------------------------------------------------------------------
// launchers

void upload_dirs(const int dirs[8][189][4])
{
    check(hipMemcpyToSymbol(HIP_SYMBOL(c_dirs), dirs, sizeof(int) * 8 * 189 * 4), "hipMemcpyToSymbol(c_dirs)");
}

void launch_gather_weights(const double *w, const int32_t *perm, int64_t N, int ncols, double *ws, int64_t w_lo,
                           int64_t w_n, hipStream_t st)
{
    int64_t total = N * ncols;
    int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    gather_weights<<<blocks, 256, 0, st>>>(w, perm, N, ncols, ws, w_lo, w_n);
    check(hipGetLastError(), "gather_weights");
}

void launch_gather_weights_t(const double *w, const int32_t *perm, int64_t N, int ncols, double *wsT,
                             int64_t w_lo, int64_t w_n, hipStream_t st)
{
    int64_t total = N * ncols;
    if (total <= 0) return;
    int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    gather_weights_t<<<blocks, 256, 0, st>>>(w, perm, N, ncols, wsT, w_lo, w_n);
    check(hipGetLastError(), "gather_weights_t");
}

void launch_gather_rows_t(const double *w, const int32_t *perm, int64_t N, int ncols, double *wT, double *wsT,
                          int64_t w_lo, int64_t w_n, hipStream_t st)
{
    if (N * ncols <= 0 || w_n <= 0) return;
    const int64_t tb = (w_n + 31) / 32;
    if (tb > INT32_MAX) fail(PBB_ERR_UNSUPPORTED, "too many points for the weight transpose");
    transpose_weights<<<(unsigned)tb, 256, 0, st>>>(w, w_n, ncols, wT);
    check(hipGetLastError(), "transpose_weights");
    const int blocks = (int)std::min<int64_t>((N * ncols + 255) / 256, 65536);
    gather_rows<<<blocks, 256, 0, st>>>(wT, perm, N, ncols, wsT, w_lo, w_n);
    check(hipGetLastError(), "gather_rows");
}

void launch_p2m(const P2MArgs &a, int nleaf, hipStream_t st)
{
    if (nleaf <= 0) return;
    p2m_fp64<<<nleaf, 256, 0, st>>>(a);
    check(hipGetLastError(), "p2m_fp64");
}

void launch_m2m(const M2MArgs &a, int nparent, hipStream_t st)
{
    if (nparent <= 0) return;
    size_t lds = sizeof(double) * (2 * (size_t)a.cc * a.n * a.n * a.n + 2 * a.n * a.n);
    if (a.s32) m2m_fp64<float><<<nparent, 256, lds, st>>>(a);
    else m2m_fp64<double><<<nparent, 256, lds, st>>>(a);
    check(hipGetLastError(), "m2m_fp64");
}

void launch_l2l(const M2MArgs &a, int nparent, hipStream_t st)
{
    if (nparent <= 0) return;
    size_t lds = sizeof(double) * (3 * (size_t)a.cc * a.n * a.n * a.n + 2 * a.n * a.n);
    if (a.s32) l2l_fp64<float><<<nparent, 256, lds, st>>>(a);
    else l2l_fp64<double><<<nparent, 256, lds, st>>>(a);
    check(hipGetLastError(), "l2l_fp64");
}

void launch_gemm(const GemmArgs &g, hipStream_t st)
{
    if (g.ncol <= 0) return;
    if (g.ldb < g.K || g.ldc < g.M) fail(PBB_ERR_INTERNAL, "gemm: leading dimension too small");
    const int rb = (g.M + 15) / 16;
    if (rb <= 16) {
        dim3 grid((unsigned)((g.ncol + 63) / 64));
#define PBB_G(R)                                                                                   \
    case R:                                                                                        \
        if (!g.b_f32 && !g.c_f32) gemm_mfma_f64<double, double, R><<<grid, 256, 0, st>>>(g);       \
        else if (!g.b_f32 && g.c_f32) gemm_mfma_f64<double, float, R><<<grid, 256, 0, st>>>(g);    \
        else if (g.b_f32 && !g.c_f32) gemm_mfma_f64<float, double, R><<<grid, 256, 0, st>>>(g);    \
        else gemm_mfma_f64<float, float, R><<<grid, 256, 0, st>>>(g);                              \
        break;
        switch (rb) {
            PBB_G(1) PBB_G(2) PBB_G(3) PBB_G(4) PBB_G(5) PBB_G(6) PBB_G(7) PBB_G(8) PBB_G(9) PBB_G(10) PBB_G(11)
            PBB_G(12) PBB_G(13) PBB_G(14) PBB_G(15) PBB_G(16)
        }
#undef PBB_G
        check(hipGetLastError(), "gemm_mfma_f64");
        return;
    }
    dim3 grid((unsigned)((g.ncol + 63) / 64), (unsigned)((std::max(g.M, g.ldc) + 63) / 64));
    if (!g.b_f32 && !g.c_f32) gemm_small_a<double, double><<<grid, 256, 0, st>>>(g);
    else if (!g.b_f32 && g.c_f32) gemm_small_a<double, float><<<grid, 256, 0, st>>>(g);
    else if (g.b_f32 && !g.c_f32) gemm_small_a<float, double><<<grid, 256, 0, st>>>(g);
    else gemm_small_a<float, float><<<grid, 256, 0, st>>>(g);
    check(hipGetLastError(), "gemm_small_a");
}

void launch_m2l(const M2LArgs &a, int ntiles, int ncolchunks, hipStream_t st)
{
    if (ntiles <= 0) return;
    // at most 16 row blocks (256 rows) per workgroup: larger ranks run as
    // row groups of equal size
    const int rball = (a.rf + 15) / 16, ngroups = (rball + 15) / 16, rb = (rball + ngroups - 1) / ngroups;
    dim3 grid((unsigned)ntiles, (unsigned)ncolchunks, (unsigned)ngroups);
    const int w = a.TB * a.cpt;  // tile width chosen by the caller: at most 128 columns
    if (w > 128) fail(PBB_ERR_INTERNAL, "FP64 M2L tile wider than 128 columns");
#define PBB_M2L_CASE(R)                                                        \
    case R:                                                                    \
        if (w > 64) m2l_mfma_f64<R, 4><<<grid, 512, 0, st>>>(a);               \
        else if (w > 32) m2l_mfma_f64<R, 2><<<grid, 256, 0, st>>>(a);          \
        else m2l_mfma_f64<R, 1><<<grid, 128, 0, st>>>(a);                      \
        break;
    switch (rb) {
        PBB_M2L_CASE(1) PBB_M2L_CASE(2) PBB_M2L_CASE(3) PBB_M2L_CASE(4) PBB_M2L_CASE(5) PBB_M2L_CASE(6)
        PBB_M2L_CASE(7) PBB_M2L_CASE(8) PBB_M2L_CASE(9) PBB_M2L_CASE(10) PBB_M2L_CASE(11)
        PBB_M2L_CASE(12) PBB_M2L_CASE(13) PBB_M2L_CASE(14) PBB_M2L_CASE(15) PBB_M2L_CASE(16)
    default: fail(PBB_ERR_INTERNAL, "M2L row group above 256 rows");
    }
#undef PBB_M2L_CASE
    check(hipGetLastError(), "m2l_mfma_f64");
}

void launch_near(const NearArgs &a, int kid, int nleaf, hipStream_t st)
{
    if (nleaf <= 0) return;
#define PBB_NEAR_CASE(K)                                  \
    case K:                                               \
        near_fp64<K><<<nleaf, 64, 0, st>>>(a);            \
        break;
    switch (kid) {
        PBB_NEAR_CASE(PBB_KERNEL_LAPLACIAN) PBB_NEAR_CASE(PBB_KERNEL_ONEOVERR2) PBB_NEAR_CASE(PBB_KERNEL_GAUSSIAN)
        PBB_NEAR_CASE(PBB_KERNEL_EXPONENTIAL) PBB_NEAR_CASE(PBB_KERNEL_SOFT_LAPLACIAN)
        PBB_NEAR_CASE(PBB_KERNEL_QUADRIC) PBB_NEAR_CASE(PBB_KERNEL_INV_QUADRIC)
        PBB_NEAR_CASE(PBB_KERNEL_LOGARITHM) PBB_NEAR_CASE(PBB_KERNEL_THIN_PLATE) PBB_NEAR_CASE(PBB_KERNEL_R2)
        PBB_NEAR_CASE(PBB_KERNEL_CONSTANT) PBB_NEAR_CASE(PBB_KERNEL_EXP_ANISO)
    default: fail(PBB_ERR_UNSUPPORTED, "no device functor for this kernel id");
    }
#undef PBB_NEAR_CASE
    check(hipGetLastError(), "near_fp64");
}

bool launch_near_split(const NearArgs &a, int kid, int nleaf, hipStream_t st)
{
    if (a.n < 4 || a.n > 8) return false;
    if (nleaf <= 0) return true;
    const NearArgs &b = a;
    for (int c = 0; c < NEAR_CLASSES; c++) {
        const int cnt = a.leaf_class_off[c + 1] - a.leaf_class_off[c];
        if (cnt <= 0) continue;
        const int32_t *lst = a.leaf_class + a.leaf_class_off[c];
#define PBB_P2P_TT(K, T)                                                    \
    case T:                                                                 \
        if constexpr (T >= 32) {                                            \
            if (a.kf32) p2p_mfma<K, T / 16, true><<<cnt, 64, 0, st>>>(b, lst);  \
            else p2p_mfma<K, T / 16, false><<<cnt, 64, 0, st>>>(b, lst);        \
        } else {                                                            \
            p2p_fp64<K, T><<<cnt, 64, 0, st>>>(b, lst);                     \
        }                                                                   \
        break;
#define PBB_P2P_CASE(K)                                                     \
    case K:                                                                 \
        switch (2 << c) { PBB_P2P_TT(K, 2) PBB_P2P_TT(K, 4) PBB_P2P_TT(K, 8) \
                          PBB_P2P_TT(K, 16) PBB_P2P_TT(K, 32) PBB_P2P_TT(K, 64) } \
        break;
        switch (kid) {
            PBB_P2P_CASE(PBB_KERNEL_LAPLACIAN) PBB_P2P_CASE(PBB_KERNEL_ONEOVERR2) PBB_P2P_CASE(PBB_KERNEL_GAUSSIAN)
            PBB_P2P_CASE(PBB_KERNEL_EXPONENTIAL) PBB_P2P_CASE(PBB_KERNEL_SOFT_LAPLACIAN)
            PBB_P2P_CASE(PBB_KERNEL_QUADRIC) PBB_P2P_CASE(PBB_KERNEL_INV_QUADRIC)
            PBB_P2P_CASE(PBB_KERNEL_LOGARITHM) PBB_P2P_CASE(PBB_KERNEL_THIN_PLATE) PBB_P2P_CASE(PBB_KERNEL_R2)
            PBB_P2P_CASE(PBB_KERNEL_CONSTANT) PBB_P2P_CASE(PBB_KERNEL_EXP_ANISO)
        default: fail(PBB_ERR_UNSUPPORTED, "no device functor for this kernel id");
        }
#undef PBB_P2P_CASE
#undef PBB_P2P_TT
    }
    check(hipGetLastError(), "p2p_fp64");
    return true;
}

bool launch_p2m_levels(const UpArgs &a, int nitems, const UpReduce &r, hipStream_t st)
{
    if (a.n < 4 || a.n > 8) return false;
    if (nitems <= 0) return true;
#define PBB_P2M_N(TS)                                                      \
    switch (a.n) {                                                         \
    case 4: p2m_levels<4, TS><<<nitems, 256, 0, st>>>(a); break;           \
    case 5: p2m_levels<5, TS><<<nitems, 256, 0, st>>>(a); break;           \
    case 6: p2m_levels<6, TS><<<nitems, 256, 0, st>>>(a); break;           \
    case 7: p2m_levels<7, TS><<<nitems, 256, 0, st>>>(a); break;           \
    default: p2m_levels<8, TS><<<nitems, 256, 0, st>>>(a); break;          \
    }
    if (a.s32) PBB_P2M_N(float) else PBB_P2M_N(double)
#undef PBB_P2M_N
    check(hipGetLastError(), "p2m_levels");
    if (r.nred > 0) {
        const int64_t total = (int64_t)r.nred * r.ncols * r.n3;
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
        if (r.s32) p2m_reduce<float><<<blocks, 256, 0, st>>>(r);
        else p2m_reduce<double><<<blocks, 256, 0, st>>>(r);
        check(hipGetLastError(), "p2m_reduce");
    }
    return true;
}

bool launch_l2p_level(const DownArgs &a, int nseg, hipStream_t st)
{
    if (a.n < 4 || a.n > 8) return false;
    if (nseg <= 0) return true;
    if (a.f32) {
        switch (a.n) {
        case 4: l2p_f32<4><<<nseg, 256, 0, st>>>(a); break;
        case 5: l2p_f32<5><<<nseg, 256, 0, st>>>(a); break;
        case 6: l2p_f32<6><<<nseg, 256, 0, st>>>(a); break;
        case 7: l2p_f32<7><<<nseg, 256, 0, st>>>(a); break;
        default: l2p_f32<8><<<nseg, 256, 0, st>>>(a); break;
        }
        check(hipGetLastError(), "l2p_f32");
        return true;
    }
    switch (a.n) {
    case 4: l2p_fp64<4, double><<<nseg, 256, 0, st>>>(a); break;
    case 5: l2p_fp64<5, double><<<nseg, 256, 0, st>>>(a); break;
    case 6: l2p_fp64<6, double><<<nseg, 256, 0, st>>>(a); break;
    case 7: l2p_fp64<7, double><<<nseg, 256, 0, st>>>(a); break;
    default: l2p_fp64<8, double><<<nseg, 256, 0, st>>>(a); break;
    }
    check(hipGetLastError(), "l2p_fp64");
    return true;
}

void launch_direct(const DirectArgs &a, int kid, hipStream_t st)
{
    if (a.T <= 0) return;
    dim3 grid((unsigned)((a.T + DS_TGT - 1) / DS_TGT), (unsigned)a.nsplit);
#define PBB_DIRECT_CASE(K)                                \
    case K:                                               \
        direct_partial<K><<<grid, 256, 0, st>>>(a);       \
        break;
    switch (kid) {
        PBB_DIRECT_CASE(PBB_KERNEL_LAPLACIAN) PBB_DIRECT_CASE(PBB_KERNEL_ONEOVERR2)
        PBB_DIRECT_CASE(PBB_KERNEL_GAUSSIAN) PBB_DIRECT_CASE(PBB_KERNEL_EXPONENTIAL)
        PBB_DIRECT_CASE(PBB_KERNEL_SOFT_LAPLACIAN) PBB_DIRECT_CASE(PBB_KERNEL_QUADRIC)
        PBB_DIRECT_CASE(PBB_KERNEL_INV_QUADRIC) PBB_DIRECT_CASE(PBB_KERNEL_LOGARITHM)
        PBB_DIRECT_CASE(PBB_KERNEL_THIN_PLATE) PBB_DIRECT_CASE(PBB_KERNEL_R2)
        PBB_DIRECT_CASE(PBB_KERNEL_CONSTANT) PBB_DIRECT_CASE(PBB_KERNEL_EXP_ANISO)
    default: fail(PBB_ERR_UNSUPPORTED, "no device functor for this kernel id");
    }
#undef PBB_DIRECT_CASE
    check(hipGetLastError(), "direct_partial");
    int64_t total = a.T * a.ncols;
    int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    direct_reduce<<<blocks, 256, 0, st>>>(a.part, a.nsplit, total, a.out);
    check(hipGetLastError(), "direct_reduce");
}

void launch_near_grad(const NearGradArgs &a, int kid, int nleaf, hipStream_t st)
{
    if (nleaf <= 0) return;
#define PBB_GRAD_CASE(K)                                  \
    case K:                                               \
        near_grad_fp64<K><<<nleaf, 64, 0, st>>>(a);       \
        break;
    switch (kid) {
        PBB_GRAD_CASE(PBB_KERNEL_LAPLACIAN) PBB_GRAD_CASE(PBB_KERNEL_ONEOVERR2) PBB_GRAD_CASE(PBB_KERNEL_GAUSSIAN)
        PBB_GRAD_CASE(PBB_KERNEL_EXPONENTIAL) PBB_GRAD_CASE(PBB_KERNEL_SOFT_LAPLACIAN)
        PBB_GRAD_CASE(PBB_KERNEL_QUADRIC) PBB_GRAD_CASE(PBB_KERNEL_INV_QUADRIC)
        PBB_GRAD_CASE(PBB_KERNEL_LOGARITHM) PBB_GRAD_CASE(PBB_KERNEL_THIN_PLATE) PBB_GRAD_CASE(PBB_KERNEL_R2)
        PBB_GRAD_CASE(PBB_KERNEL_CONSTANT) PBB_GRAD_CASE(PBB_KERNEL_EXP_ANISO)
    default: fail(PBB_ERR_UNSUPPORTED, "no device functor for this kernel id");
    }
#undef PBB_GRAD_CASE
    check(hipGetLastError(), "near_grad_fp64");
}

// out[0..n) = 0, as a kernel so that it can sit in a captured graph
__global__ void fill_zero_f64(double *__restrict__ out, int64_t n)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = 0.0;
}

template <int KID>
__global__ __launch_bounds__(256) void kernel_blocks_dev(KernelBlockArgs a)
{
    const int64_t n6 = (int64_t)a.n3 * a.n3, total = a.symm_num * n6;
    const int n = a.n;
    for (int64_t q = blockIdx.x * (int64_t)256 + threadIdx.x; q < total; q += (int64_t)gridDim.x * 256) {
        const int d = (int)(q / n6);
        const int64_t rem = q - d * n6;
        const int j = (int)(rem / a.n3), i = (int)(rem - (int64_t)j * a.n3);
        const int i1 = i / (n * n), i2 = (i / n) % n, i3 = i % n;
        const int j1 = j / (n * n), j2 = (j / n) % n, j3 = j % n;
        const double t0 = 0.5 * a.nodes[i1] * a.box_len, t1 = 0.5 * a.nodes[i2] * a.box_len,
                     t2 = 0.5 * a.nodes[i3] * a.box_len;
        const double s0 = (a.dirs[3 * d] + 0.5 * a.nodes[j1]) * a.box_len,
                     s1 = (a.dirs[3 * d + 1] + 0.5 * a.nodes[j2]) * a.box_len,
                     s2 = (a.dirs[3 * d + 2] + 0.5 * a.nodes[j3]) * a.box_len;
        a.Kcell[q] = keval<KID>(s0 - t0, s1 - t1, s2 - t2, a.kp) / (a.Kw[i] * a.Kw[j]);
    }
}

// K_d = symm * K_{315-d}^T for the directions past symm_num
__global__ __launch_bounds__(256) void kernel_blocks_symm(KernelBlockArgs a)
{
    const int64_t n6 = (int64_t)a.n3 * a.n3, total = (316 - a.symm_num) * n6;
    for (int64_t q = blockIdx.x * (int64_t)256 + threadIdx.x; q < total; q += (int64_t)gridDim.x * 256) {
        const int d = a.symm_num + (int)(q / n6);
        const int64_t rem = q - (int64_t)(d - a.symm_num) * n6;
        const int j = (int)(rem / a.n3), i = (int)(rem - (int64_t)j * a.n3);
        a.Kcell[d * n6 + i + (int64_t)j * a.n3] = a.symm * a.Kcell[(315 - d) * n6 + j + (int64_t)i * a.n3];
    }
}

void launch_kernel_blocks(const KernelBlockArgs &a, int kid, hipStream_t st)
{
    const int64_t n6 = (int64_t)a.n3 * a.n3;
    if (a.n3 <= 0 || a.n3 != a.n * a.n * a.n || a.kcell_len < 316 * n6 || a.symm_num < 0 || a.symm_num > 316)
        fail(PBB_ERR_INTERNAL, "kernel block arguments do not fit the buffer");
    const int blocks = (int)std::min<int64_t>((a.symm_num * n6 + 255) / 256, 65536);
#define PBB_KB_CASE(K)                                                  \
    case K:                                                             \
        kernel_blocks_dev<K><<<blocks, 256, 0, st>>>(a);                \
        break;
    switch (kid) {
        PBB_KB_CASE(PBB_KERNEL_LAPLACIAN) PBB_KB_CASE(PBB_KERNEL_ONEOVERR2) PBB_KB_CASE(PBB_KERNEL_GAUSSIAN)
        PBB_KB_CASE(PBB_KERNEL_EXPONENTIAL) PBB_KB_CASE(PBB_KERNEL_SOFT_LAPLACIAN) PBB_KB_CASE(PBB_KERNEL_QUADRIC)
        PBB_KB_CASE(PBB_KERNEL_INV_QUADRIC) PBB_KB_CASE(PBB_KERNEL_LOGARITHM) PBB_KB_CASE(PBB_KERNEL_THIN_PLATE)
        PBB_KB_CASE(PBB_KERNEL_R2) PBB_KB_CASE(PBB_KERNEL_CONSTANT) PBB_KB_CASE(PBB_KERNEL_EXP_ANISO)
    default: fail(PBB_ERR_UNSUPPORTED, "no device functor for this kernel id");
    }
#undef PBB_KB_CASE
    check(hipGetLastError(), "kernel_blocks_dev");
    if (a.symm_num < 316) {
        const int b2 = (int)std::min<int64_t>(((316 - a.symm_num) * n6 + 255) / 256, 65536);
        kernel_blocks_symm<<<b2, 256, 0, st>>>(a);
        check(hipGetLastError(), "kernel_blocks_symm");
    }
}

void launch_fill_zero(double *out, int64_t n, hipStream_t st)
{
    if (n <= 0) return;
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    fill_zero_f64<<<blocks, 256, 0, st>>>(out, n);
    check(hipGetLastError(), "fill_zero_f64");
}

// Per-workgroup partial sums of (a - b)^2 and b^2 over n doubles, written to
// part[2*g] and part[2*g + 1]; the host adds the pairs (no atomics, so the
// result does not depend on the schedule).
__global__ __launch_bounds__(256) void diff_norms_f64(const double *__restrict__ a, const double *__restrict__ b,
                                                      int64_t n, double *__restrict__ part)
{
    __shared__ double sd[256], sb[256];
    double d2 = 0, b2 = 0;
    for (int64_t q = blockIdx.x * (int64_t)256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
        const double bv = b[q], dv = a[q] - bv;
        d2 += dv * dv;
        b2 += bv * bv;
    }
    sd[threadIdx.x] = d2;
    sb[threadIdx.x] = b2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sd[threadIdx.x] += sd[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

void launch_diff_norms(const double *a, const double *b, int64_t n, double *part, int nblocks, hipStream_t st)
{
    diff_norms_f64<<<nblocks, 256, 0, st>>>(a, b, n, part);
    check(hipGetLastError(), "diff_norms_f64");
}

}  // namespace dev
}  // namespace pbb
