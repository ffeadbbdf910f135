// kernels.hpp -- argument blocks and launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "host.hpp"

namespace pbb {

inline void check(hipError_t e, const char *what)
{
    if (e != hipSuccess) fail(PBB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace dev {

// Workgroups are dispatched to the 8 XCDs round-robin by linear id; this
// maps linear id b of G so that each XCD receives one contiguous range of
// logical tiles (Morton-adjacent targets share sources and operators in
// that XCD's L2).
__device__ __forceinline__ int xcd_tile(int b, int G)
{
    const int x = b & 7, q = b >> 3, base = G >> 3, rem = G & 7;
    return x * base + (x < rem ? x : rem) + q;
}

struct KParams {
    double p[4];
};

struct P2MArgs {
    int n, ncols, cc;
    int64_t N, off;            // off: global box offset of the leaf level
    double len;
    const int32_t *leaf_start, *coord;
    const double *ctab, *Tkz;
    const double *x, *y, *z, *ws;
    double *S;
};

struct M2MArgs {
    int n, ncols, cc;
    int64_t off_parent, off_child;
    const int32_t *child;       // nparent*8
    const double *Cw;
    double *Sparent;            // M2M: written; L2L: read (parent F)
    double *Schild;             // M2M: read;    L2L: accumulated
    int s32;                    // the expansions are stored in FP32 (same offsets in floats)
};

// per-level tables of the direct upward / downward passes (levels 2..L)
constexpr int MAX_FAR_LEVELS = 9;
struct LevelTab {
    int nlev;                                  // levels 2 .. 1 + nlev
    int64_t off[MAX_FAR_LEVELS + 1];           // global box offsets; off[nlev] = total boxes
    const int32_t *pstart[MAX_FAR_LEVELS];     // box -> first sorted point (nb + 1 entries)
    const int32_t *coord[MAX_FAR_LEVELS];
    const double *ctab[MAX_FAR_LEVELS];
    double len[MAX_FAR_LEVELS];
};

constexpr int P2M_SEG = 2048;  // points per workgroup of the direct upward pass

struct UpArgs {
    int n, ncols;
    LevelTab lt;
    // work items (level index, box, first point, scratch slot or -1), coarse
    // levels first; boxes of more than P2M_SEG points are split into items
    // whose partial expansions go to scratch slots and are summed afterwards
    const int4 *items;
    const double *Tkz, *x, *y, *z, *wsT;
    void *S;                   // expansions: FP64, or FP32 with s32
    double *scratch;           // partial expansions of split boxes (FP64)
    int s32;
};

struct UpReduce {
    int ncols, n3, nred;
    const int4 *reds;          // (global box id, first slot, slot count, 0)
    const double *scratch;
    void *S;                   // FP64, or FP32 with s32
    int s32;
};

constexpr int L2P_SEG = 256;  // targets per workgroup of the per-level L2P

struct DownArgs {
    int n, ncols, accumulate;
    int64_t off;                               // global box offset of the level
    double len;
    const int32_t *seg_box, *seg_p0, *pstart, *coord;
    const double *ctab, *Tkz, *x, *y, *z;
    const void *F;                             // local expansions: FP64, or FP32 with f32
    double *lp;                                // point-major L2P results
    int f32;
};

// Proxy passes of the levels whose operator ranks are at most 32 (mixed
// modes): P2M straight to the proxies P = VT' S, and L2P straight from the
// M2L results Pf, without the n^3 expansions S and F of the level.  Range k
// holds the boxes whose first point lies in [k*PX_PT, (k+1)*PX_PT);
// persistent workgroups walk the ranges in passes of PX_PT points.
constexpr int PX_PT = 128;
constexpr int PX_K = 32;       // ranks held per point (zero-padded)
struct ProxyArgs {
    int ncols, nb, accumulate;
    int rank;                  // ks (P2M) or kf (L2P), <= PX_K
    int64_t npts;              // points of the tree (pstart[nb])
    int ldp, ldf;              // P and Pf strides per (box, column)
    int pf_vec;                // L2P: Pf rows 16-byte aligned (ldf = PX_K)
    int64_t off;               // global box offset of the level
    double len, alpha;         // box length; L2P scale (homogeneous kernels)
    const int32_t *pstart, *coord;
    const double *ctab, *Tkz, *x, *y, *z;
    const double *wsT;         // P2M: point-major sorted weights
    const void *ob;            // operator split into 3 BF16 parts, [part][rank < 32][k] (row stride
                               // 32*ceil(n^3/32) + 8): P2M: VT' rows, L2P: U' columns
    float *P;                  // P2M: proxies, rows 0..min(ldp, PX_K)-1 written (those past are zero by construction)
    unsigned *xmax;            // P2M: atomic max of |P| as FP32 bits (cleared by the caller), or nullptr
    const float *Pf;           // L2P: M2L results
    double *lp;                // L2P: point-major results
};

struct GemmArgs {
    int M, K, ldb, ldc;
    int b_f32, c_f32;          // B / C element type: 0 FP64, 1 FP32
    int64_t ncol;
    double alpha;
    const double *A;
    const void *B;
    void *C;
};

struct M2LArgs {
    int rf, rs, ncols, TB, cpt, side, chunk_major;  // step order: see m2l_step
    int ldp, ldf;                                   // P and Pf strides per (box, column), >= rs, rf
    int64_t off;
    int tile_prefix[9];
    int64_t class_start[9];
    const int32_t *order, *coord, *map;
    const double *E, *P;
    double *Pf;
};

constexpr int NEAR_CLASSES = 6;

// Step s of an M2L tile with nd non-empty directions and nch k chunks:
// direction-major (the chunks of one E_d, then the next direction) or
// chunk-major (one chunk of every direction, then the next chunk).
__device__ __forceinline__ void m2l_step(int s, int nd, int nch, int chunk_major, int &di, int &ch)
{
    if (chunk_major) {
        ch = s / nd;
        di = s - ch * nd;
    } else {
        di = s / nch;
        ch = s - di * nch;
    }
}

// S_m(u), m = 0..N-1: the Chebyshev interpolation weights of a point at
// u in [-1, 1] of its box (ComputeSn + ClenshawSum, H2_3D_Tree.cpp:830-879),
// from Tkz[m*N + j] = T_j(node_m).
template <int N>
__device__ __forceinline__ void sn_eval_n(const double *__restrict__ Tkz, double u, double *S)
{
    const double pfac = 2.0 / N;
#pragma unroll
    for (int m = 0; m < N; m++) {
        double d0 = 0, d1 = 0;
#pragma unroll
        for (int j = N - 1; j > 0; j--) {
            const double d2 = d0;
            d0 = 2.0 * u * d0 - d1 + Tkz[m * N + j];
            d1 = d2;
        }
        S[m] = pfac * (u * d0 - d1 + 0.5 * Tkz[m * N]);
    }
}

struct NearArgs {
    int n, ncols, side;
    int64_t N, off;
    double len;
    KParams kp;
    const int32_t *leaf_start, *coord, *map, *perm;
    const double *ctab, *Tkz;
    const double *x, *y, *z, *ws;
    const double *F;            // leaf-level local expansions (nullptr: no far field)
    // out[c*out_n + perm[k] - out_lo] += ... for the sorted points k whose
    // input index perm[k] lies in [out_lo, out_lo + out_n); the others
    // (sources of a plan whose sources are not its targets) write nothing
    double *out;
    int64_t out_lo, out_n;
    const double *wsT;          // sorted weights, point-major wsT[k*ncols + c]
    double *lp;                 // L2P results, point-major (nullptr: no far field)
    // leaves grouped by size: class c holds leaves of at most 2 << c points
    // (the last class also the larger ones)
    const int32_t *leaf_class;
    int leaf_class_off[NEAR_CLASSES + 1];
    int kf32;                   // dense leaves: kernel values in FP32 (mixed modes)
};

struct DirectArgs {
    int64_t T, Ns;
    int ncols, nsplit;
    KParams kp;
    const double *tgt, *src, *w;   // AoS targets / sources, column-major weights
    double *part, *out;            // part: nsplit x ncols x T;  out: ncols x T
};

struct M2LMixArgs {
    int rf, rs, ncols, TB, cpt, side, kpad, nch, escale, R;  // R: rows of the split E (16-multiple)
    int ldf;                                                 // Pf stride per (box, column), >= rf
    int chunk_major;                                         // step order: see m2l_step
    int row_split;                                           // rows of r_f split over two workgroups
    int wave_rows;    // row-wave kernel: the waves split the rows, E straight to registers
    int dlo, dhi;     // only directions whose operator index (0..315) lies in [dlo, dhi)
    int accumulate;   // add to Pf (an earlier direction group's partial sums) instead of overwriting
    int split_acc;    // row-wave kernel, two-part modes: cross products summed apart (64-column tiles)
    int64_t off;
    int tile_prefix[9];
    int64_t class_start[9];
    const int32_t *order, *coord, *map;
    const void *Es;             // split E of this level's operator set
    const float *P;             // FP32 proxies, kpad per column
    const unsigned *xmax;       // max |P| of the level (FP16 scaling) or nullptr
    float *Pf;
    int k16;          // f16 x2, ks at most 16: two directions' 16-deep chunks per 32-deep step, 16-deep pre-split proxies
};

struct NearGradArgs {
    int n, ncols, side;
    int64_t N;
    double len;
    KParams kp;
    const int32_t *leaf_start, *coord, *map, *perm;
    const double *ctab, *Tkz;
    const double *x, *y, *z;
    const double *u, *v;        // sorted source weights / target coefficients (column-major N)
    const double *F;            // leaf expansions [leaf][col][n3] (nullptr: none)
    double *grad;               // AoS N x 3, caller's order, accumulated
};

void upload_dirs(const int dirs[8][189][4]);
void launch_near_grad(const NearGradArgs &a, int kid, int nleaf, hipStream_t st);
void launch_fill_zero(double *out, int64_t n, hipStream_t st);
// part[2g], part[2g+1] = sums of (a-b)^2 and b^2 over workgroup g's share of n
void launch_diff_norms(const double *a, const double *b, int64_t n, double *part, int nblocks, hipStream_t st);
void launch_split_e(int variant, const double *E, int rf, int rs, int R, int nch, int escale, void *Es,
                    hipStream_t st);
// max |x| over the first `used` floats of rows rows of stride ld
void launch_absmax_rows(const float *x, int64_t rows, int ld, int used, unsigned *out, hipStream_t st);
// two-part modes: rewrite the first cpr 32-float k-chunks of each of rows
// rows of P (stride ld floats) as scaled 16-bit parts
// kd: k-chunk depth (32; 16 for the one chunk of a set whose ks is at most 16, f16 x2)
void launch_split_p2(int variant, float *P, int64_t rows, int ld, int cpr, const unsigned *xmax, hipStream_t st,
                     int kd = 32);
void upload_dirs_mixed(const int dirs[8][189][4]);
void launch_zero_u32(unsigned *p, int n, hipStream_t st);
// widest row-wave M2L tile (PBB_M2L_WAVE_ROWS) for a variant and split-E row count
int m2l_rows_max_width(int variant, int R);
void launch_m2l_mixed(int variant, const M2LMixArgs &a, int ntiles, int ncolchunks, hipStream_t st);
// ws[c*N + k] = w[c*w_n + perm[k] - w_lo] for perm[k] in [w_lo, w_lo + w_n),
// else 0 (w_n = N, w_lo = 0 for a plan whose sources are its targets)
void launch_gather_weights(const double *w, const int32_t *perm, int64_t N, int ncols, double *ws, int64_t w_lo,
                           int64_t w_n, hipStream_t st);
void launch_p2m(const P2MArgs &a, int nleaf, hipStream_t st);
void launch_m2m(const M2MArgs &a, int nparent, hipStream_t st);
void launch_l2l(const M2MArgs &a, int nparent, hipStream_t st);
void launch_gemm(const GemmArgs &g, hipStream_t st);

// C = alpha * A B on the BF16 matrix cores with three-part operands (mixed
// modes): A pre-split by launch_split_a3 into [3][R][32*nch] BF16
struct GemmSplitArgs {
    int M, K, ldb, ldc, b_f32, c_f32, nch;
    int64_t ncol;
    double alpha;
    const void *As, *B;
    void *C;
};
void launch_split_a3(const double *A, int M, int K, int R, int Kp, void *As, hipStream_t st);
void launch_gemm_split3(const GemmSplitArgs &g, hipStream_t st);
void launch_m2l(const M2LArgs &a, int ntiles, int ncolchunks, hipStream_t st);
void launch_near(const NearArgs &a, int kid, int nleaf, hipStream_t st);
// the split near field for orders 4..8 (false: order not covered)
bool launch_near_split(const NearArgs &a, int kid, int nleaf, hipStream_t st);
// direct upward pass: S of every box of levels 2..L from its points
bool launch_p2m_levels(const UpArgs &a, int nitems, const UpReduce &r, hipStream_t st);
// direct downward pass, one level: lp (+)= L2P of the level's F at its points
bool launch_l2p_level(const DownArgs &a, int nseg, hipStream_t st);
// the proxy passes of one level (orders 4..8; false for other orders);
// npts: points of the tree
bool launch_p2m_proxy(const ProxyArgs &a, int n, int64_t npts, hipStream_t st);
bool launch_l2p_proxy(const ProxyArgs &a, int n, int64_t npts, hipStream_t st);
// the point-major copy: wsT[k*ncols + c], the same rule
// wsT through a point-major copy wT of the caller's weights (w_n x ncols
// doubles of scratch): a coalesced transpose, then whole-row gathers
void launch_gather_rows_t(const double *w, const int32_t *perm, int64_t N, int ncols, double *wT, double *wsT,
                          int64_t w_lo, int64_t w_n, hipStream_t st);
void launch_gather_weights_t(const double *w, const int32_t *perm, int64_t N, int ncols, double *wsT,
                             int64_t w_lo, int64_t w_n, hipStream_t st);
void launch_direct(const DirectArgs &a, int kid, hipStream_t st);

// Kernel blocks of one operator set on the device (H2_3D_Tree.cpp:277-374):
// Kcell[d*n6 + i + j*n3] = K(t_i, s_j^(d)) / (Kw_i Kw_j) for the 316
// directions (dirs: 316 x 3 offsets in Ktable order), the first symm_num
// evaluated, the rest filled as symm * K_{315-d}^T.
struct KernelBlockArgs {
    int n, n3, symm_num, symm;
    double box_len;
    KParams kp;
    const double *nodes, *Kw;
    const int *dirs;
    double *Kcell;
    int64_t kcell_len;          // doubles at Kcell (316 n3^2 needed)
};
void launch_kernel_blocks(const KernelBlockArgs &a, int kid, hipStream_t st);

}  // namespace dev
}  // namespace pbb
