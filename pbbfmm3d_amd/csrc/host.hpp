// host.hpp -- host-side (CPU, FP64) pieces of the FMM: Chebyshev weights,
// kernel evaluation, operator precompute (TSQR + one-sided Jacobi SVD), and
// the tree geometry of a point set.  No HIP types here.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "pbbfmm3d_amd.h"

namespace pbb {

// Error carrying a status code across the C ABI.
struct Error {
    pbb_status status;
    std::string msg;
};

[[noreturn]] void fail(pbb_status st, const std::string &msg);

// Runs body(i) for i in [0, n) on up to hardware_concurrency() threads.
void parallel_for(int64_t n, const std::function<void(int64_t)> &body);

// ---- kernels ---------------------------------------------------------------
// Host FP64 evaluation of the functor `id` (formulas of pbb_kernel_id).
double kernel_eval_host(int id, const double *p, const double *t, const double *s);
bool kernel_id_valid(int id);

// ---- Chebyshev pieces (H2_3D_Tree.cpp:657-909) ----------------------------
struct ChebWeights {
    int n = 0;
    std::vector<double> nodes;   // n
    std::vector<double> Tkz;     // n*n, Tkz[m*n+k] = T_k(node_m)
    std::vector<double> Kw;      // n^3
    std::vector<double> Cw;      // 2*n*n
    std::vector<int> Ktable;     // 343
    void compute(int order);
};
// S_m(x), m = 0..n-1 (ComputeSn + ClenshawSum, H2_3D_Tree.cpp:830-879)
void cheb_sn(int n, const double *Tkz, double x, double *S);

// ---- SVD ---------------------------------------------------------------------
// Upper-triangular factor R (n x n, column-major) of a tall m x n matrix A
// (column-major, ld m) by a two-level Householder TSQR.  A is destroyed.
void tsqr_r(double *A, int64_t m, int n, std::vector<double> &R);
// One-sided Jacobi SVD of a square n x n matrix A (column-major):
// sigma (descending) and the right singular vectors V (n x n column-major,
// column k belongs to sigma[k]).  A is destroyed.
void jacobi_svd_right(double *A, int n, std::vector<double> &sigma, std::vector<double> &V);
// Cutoff rule of H2_3D_Tree.cpp:406-421.
int cutoff_rank(const std::vector<double> &sigma, double eps);

// ---- operators ----------------------------------------------------------------
struct HostOps {
    pbb_kernel_desc kernel{};
    double L = 1;
    int tree_level = 0;
    int n = 0, n3 = 0;
    double eps = 0;
    int rf = 0, rs = 0;
    int nsets = 0;
    bool homog = false;
    double homogen_flipped = 0;  // H2_3D_Tree.cpp:45
    ChebWeights w;
    std::vector<double> U;       // nsets x n3 x rf   (col-major)
    std::vector<double> VT;      // nsets x rs x n3   (col-major)
    std::vector<double> E;       // nsets x 316 x rf x rs (col-major)
    std::vector<double> sigma_f, sigma_s;  // first set
    int ops_set(int level) const { return homog ? 0 : (level >= 2 ? level - 2 : 0); }
};

using HostEval = std::function<double(const double *, const double *)>;

// Optional device help for the precompute: kernel blocks of one operator
// set (box length, Kcell as kernel_blocks lays it out) and the compression
// E_d = U^T K_d V of all 316 directions (Kcell of the set's last blocks
// call).  Either may be empty; the host loops run where it is.
struct PrecomputeDevice {
    std::function<void(double box_len, std::vector<double> &Kcell)> blocks;
    std::function<void(int rf, int rs, const double *U, const double *VT, double *E)> compress;
};

// Full precompute: ComputeWeights + PrecomputeM2L (H2_3D_Tree.cpp:40-66).
void precompute_ops(HostOps &ops, const HostEval &eval, const PrecomputeDevice *dev = nullptr);
// One operator set re-compressed for the mixed modes.  The reference keeps
// the level-2 cutoff at every level of a non-homogeneous kernel
// (H2_3D_Tree.cpp:406-436), but a smooth kernel's far field is of much lower
// numerical rank on small boxes.  With Q_L (r_f x kf) and Q_R (r_s x ks) the
// leading left and right singular vectors of all 316 E_d together
// ([E_0 .. E_315] and [E_0; ..; E_315]), E'_d = Q_L^T E_d Q_R, U' = Uw Q_L
// and VT' = Q_R^T VTw apply the same far field within tau (the relative
// Frobenius norm of the discarded singular values of the stacked E, each
// side), with kf a multiple of 16 and ks 16 or a multiple of 32 (chunk
// sizes of the M2L's matrix-core steps).
// When that saves nothing the set keeps its own factors (kf = r_f, ks = r_s).
struct ReducedSet {
    int kf = 0, ks = 0;
    std::vector<double> E;    // 316 x kf x ks (col-major per direction)
    std::vector<double> VTw;  // ks x n3 (col-major), Kweights applied
    std::vector<double> Uw;   // n3 x kf (col-major), Kweights applied
    double tail_f = 0, tail_s = 0;  // discarded relative Frobenius norms
};
void reduce_set(const HostOps &ops, int set, const double *VTw, const double *Uw, double tau, ReducedSet &out);

// Kernel blocks of one operator set (H2_3D_Tree.cpp:277-374):
// Kcell[d*n6 + i + j*n3] = K(t_i, s_j^(d)) / (Kw_i Kw_j).
void kernel_blocks(const HostOps &ops, const HostEval &eval, double boxLen,
                   std::vector<double> &Kcell);

// ---- tree geometry ----------------------------------------------------------
struct Level {
    int64_t nb = 0;                 // non-empty boxes
    std::vector<int32_t> coord;     // packed ix | iy << 10 | iz << 20
    std::vector<int32_t> map;       // dense (ix*S + iy)*S + iz -> compact id or -1
    std::vector<int32_t> child;     // nb*8 compact ids at level+1 or -1
    std::vector<double> ctab;       // per-axis centres, 2^level entries
    double len = 0;
    // M2L target order: compact ids grouped by octant class, Morton order
    std::vector<int32_t> order;
    int64_t class_count[8] = {0};
};

struct HostTree {
    int Lv = 0;
    int64_t N = 0;
    double ctr[3] = {0, 0, 0};
    std::vector<double> xs, ys, zs;      // sorted, re-centred
    std::vector<int32_t> perm;           // sorted position -> input index
    std::vector<int32_t> leaf_start;     // nb[Lv]+1
    std::vector<Level> lv;               // 0..Lv
    int64_t m2l_pairs = 0, near_pairs = 0, max_leaf = 0, outside = 0;
    int64_t m2l_pairs_level[16] = {0};
};

// Replicates compute.hpp:92-127 (re-centring on the bbox centre) and
// DistributeFieldPoints (compute.hpp:1126-1203) with the node centres of
// BuildFMMHierarchy (H2_3D_Tree.cpp:917-977).  The bbox (and so the centre)
// is that of the first nbox points: the targets, when the point set is the
// targets followed by separate sources; nbox = N when they are the same set.
void build_tree(HostTree &T, const double *xyz, int64_t N, double L, int levels, int64_t nbox);

// Leaf grid coordinates, packed ix | iy << 10 | iz << 20, of a point p
// already re-centred, by build_tree's comparison chain over the per-axis
// node centres lv[l].ctab (strict '<', compute.hpp:1149-1173).
uint32_t leaf_coord(const std::vector<Level> &lv, int levels, const double p[3]);

// Far-field directions of an octant class (x bit 4, y bit 2, z bit 1):
// dirs[k] = {dx, dy, dz, E index}, 189 entries.
void class_directions(int octant, const std::vector<int> &Ktable, int dirs[189][4]);

}  // namespace pbb
