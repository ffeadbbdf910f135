/*
 * pbbfmm3d_amd.h -- C ABI of the MI355X (gfx950) black-box FMM matvec.
 *
 * Plain C types only: pointers, sizes and status codes.  Every entry point
 * names the PBBFMM3D interface it replaces (paths relative to the PBBFMM3D
 * source tree).  The C++ drop-in headers (H2_3D_Tree.hpp, kernel_Types.hpp,
 * compute.hpp) are thin wrappers over these calls; Python binds them with
 * ctypes.
 *
 * Model: one point set is both the sources and the targets (the reference
 * supports nothing else: compute.hpp:1195-1196).  Weights and outputs are
 * column-major, w[c*N + j] and out[c*N + i] (compute.hpp:316, :747), and
 * outputs are ACCUMULATED (+=), as compute.hpp:747 does.
 *
 * Errors: every call returns a pbb_status; pbb_last_error() returns a
 * thread-local message for the last failing call on that thread.  Nothing
 * calls exit() or abort().
 */
#ifndef PBBFMM3D_AMD_H
#define PBBFMM3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBB_API __attribute__((visibility("default")))

#define PBB_ABI_VERSION 2

typedef enum pbb_status {
    PBB_OK = 0,
    PBB_ERR_INVALID_ARG = 1,  /* bad size, null pointer, bad option      */
    PBB_ERR_UNSUPPORTED = 2,  /* valid but not implemented here          */
    PBB_ERR_HIP = 3,          /* a HIP runtime call failed               */
    PBB_ERR_OUT_OF_MEMORY = 4,
    PBB_ERR_NO_DEVICE = 5,    /* no usable gfx950 device                 */
    PBB_ERR_INTERNAL = 6,
    PBB_ERR_NOT_FOUND = 7     /* a file to read does not exist           */
} pbb_status;

/* Device kernel functors (the P2P near field runs these on the GPU; the
 * operator precompute evaluates them on the host in FP64 unless a host
 * callback is given).  r is |source - target|.                          */
typedef enum pbb_kernel_id {
    PBB_KERNEL_NONE = 0,           /* host callback only: no device P2P     */
    PBB_KERNEL_LAPLACIAN = 1,      /* 1/r          kernel_Types.cpp:15-28   */
    PBB_KERNEL_ONEOVERR2 = 2,      /* 1/r^2        kernel_Types.cpp:109-121 */
    PBB_KERNEL_GAUSSIAN = 3,       /* exp(-r^2/p0^2) (p0=1: :64-74)         */
    PBB_KERNEL_EXPONENTIAL = 4,    /* p1*exp(-r/p0)                         */
    PBB_KERNEL_SOFT_LAPLACIAN = 5, /* 1/sqrt(r^2+p0)                        */
    PBB_KERNEL_QUADRIC = 6,        /* 1+r^2        kernel_Types.cpp:131-138 */
    PBB_KERNEL_INV_QUADRIC = 7,    /* 1/(1+r^2)    kernel_Types.cpp:148-155 */
    PBB_KERNEL_LOGARITHM = 8,      /* 0.5 log r^2  kernel_Types.cpp:85-99   */
    PBB_KERNEL_THIN_PLATE = 9,     /* 0.5 r^2 log r^2  :165-178             */
    PBB_KERNEL_R2 = 10,            /* r^2 (degree-2 homogeneous test kernel)*/
    PBB_KERNEL_CONSTANT = 11,      /* 1 (test kernel)                       */
    PBB_KERNEL_POLY_XY = 12,       /* (tx-sx)(ty+sy)+p0 (non-symmetric test,
                                      host only: no device functor)        */
    PBB_KERNEL_EXP_ANISO = 13      /* p0*exp(-sqrt(p1 dx^2 + p2 dy^2 + p3 dz^2))
                                      (python/mykernel_python.hpp:10-17)   */
} pbb_kernel_id;

/* What H2_3D_Tree::SetKernelProperty sets (H2_3D_Tree.hpp:96,
 * kernel_Types.cpp:7-13) plus the device functor.                        */
typedef struct pbb_kernel_desc {
    int32_t id;        /* pbb_kernel_id                                    */
    int32_t symmetry;  /* 1 symmetric, -1 antisymmetric, 0 neither         */
    double homogen;    /* degree as declared (1/r: -1); |h| > 0.1 selects
                          one unit-length operator set (H2_3D_Tree.cpp:8,
                          :134-136), else one set per level 2..tree_level  */
    double param[4];   /* functor parameters (see pbb_kernel_id)           */
} pbb_kernel_desc;

/* Host FP64 kernel evaluation K(target, source); stands in for the virtual
 * H2_3D_Tree::EvaluateKernel (H2_3D_Tree.hpp:94) during the precompute.   */
typedef double (*pbb_host_kernel_fn)(void *ctx, const double *target3,
                                     const double *source3);

typedef struct pbb_ops pbb_ops;   /* operators: weights, U, VT, E          */
typedef struct pbb_plan pbb_plan; /* one point set's tree, device-resident */

typedef struct pbb_ops_info {
    int32_t order;        /* n                                            */
    int32_t n3;           /* n^3                                          */
    int32_t rank_f;       /* cutoff.f                                     */
    int32_t rank_s;       /* cutoff.s                                     */
    int32_t nsets;        /* operator sets (1 or tree_level-1)            */
    int32_t homogeneous;  /* 1 if one unit-length set                     */
    int32_t tree_level;
    int32_t reserved;
    double L;
    double homogen;       /* as declared                                  */
    double eps;
} pbb_ops_info;

/* Precision modes of pbb_plan_matvec.                                    */
typedef enum pbb_precision {
    PBB_PREC_FP64 = 0,          /* every phase in FP64 (the drop-in path)  */
    PBB_PREC_MIXED = 1,         /* M2L on the BF16 matrix cores, operands
                                   split in three BF16 parts (6 products),
                                   FP32 accumulation                        */
    PBB_PREC_MIXED_F16X2 = 2,   /* M2L with operands split in two scaled
                                   FP16 parts (3 products)                  */
    PBB_PREC_MIXED_BF16X2 = 3   /* M2L with two BF16 parts (3 products)     */
} pbb_precision;

#define PBB_NUM_PHASES 8
/* phase order of pbb_plan_timings: P2M, M2M, VT proxy, M2L, U expand, L2L,
 * L2P+P2P, total (milliseconds, HIP events on the plan's stream)         */

typedef struct pbb_plan_info {
    int64_t npoints;
    int32_t tree_level;
    int32_t nonempty_levels;
    int64_t boxes[16];      /* non-empty boxes per level 0..tree_level     */
    int64_t m2l_pairs;      /* far-field (target, source) box pairs        */
    int64_t near_pairs;     /* ordered near-field point pairs incl. self   */
    int64_t max_leaf_points;
    double center[3];       /* bbox centre subtracted (compute.hpp:108-111)*/
    int64_t outside_points; /* points outside the root cube (extrapolated
                               as the reference does)                      */
    int64_t ntargets;       /* separate sets (pbb_plan_create_ts): the
                               targets and the sources; one set: both N    */
    int64_t nsources;
    int64_t m2l_pairs_level[16]; /* m2l_pairs by level (0 at levels 0, 1)  */
} pbb_plan_info;

/* ---- library ---------------------------------------------------------- */

/* Thread-local message for the last failing call on this thread.          */
PBB_API const char *pbb_last_error(void);
PBB_API int pbb_abi_version(void);
/* Number of visible gfx950 devices (0 without a GPU; never fails).        */
PBB_API int pbb_device_count(void);

/* ---- operators -------------------------------------------------------- */

/* Replaces H2_3D_Tree::H2_3D_Tree + buildFMMTree (H2_3D_Tree.cpp:12-66):
 * Chebyshev weights (ComputeWeights :715-826) and the SVD-compressed M2L
 * operators (PrecomputeM2L :74-99, ComputeKernelCheb :251-575), computed on
 * the host in FP64, then uploaded to the current device.  host_eval may be
 * NULL, then kernel->id must name a functor.  use_chebyshev must be 1: the
 * uniform-grid/FFT branch is not provided.                                */
PBB_API pbb_status pbb_ops_create(const pbb_kernel_desc *kernel,
                                  pbb_host_kernel_fn host_eval, void *ctx,
                                  double L, int32_t tree_level, int32_t order,
                                  double eps, int32_t use_chebyshev,
                                  pbb_ops **out);

/* pbb_ops_create with flags.  By default (flags 0) the kernel blocks (the
 * 316 n^6 evaluations of ComputeKernelCheb, H2_3D_Tree.cpp:277-374) and the
 * compression E_d = U^T K_d V (:529-556) run on the current GPU when one is
 * visible and the kernel has a device functor (no host callback); the two
 * SVDs (:400, :471) always run on the host.  PBB_OPS_HOST_PRECOMPUTE keeps
 * every step on the host.                                                 */
#define PBB_OPS_HOST_PRECOMPUTE 1
PBB_API pbb_status pbb_ops_create_ex(const pbb_kernel_desc *kernel,
                                     pbb_host_kernel_fn host_eval, void *ctx,
                                     double L, int32_t tree_level, int32_t order,
                                     double eps, int32_t use_chebyshev, int32_t flags,
                                     pbb_ops **out);

/* Builds operators from given arrays (the layout FMMReadMatrices reads,
 * H2_3D_Tree.cpp:186-239): U nsets x (n3 x rank_f), VT nsets x (rank_s x
 * n3), E nsets x 316 x (rank_f x rank_s), all column-major FP64.          */
PBB_API pbb_status pbb_ops_create_from_arrays(
    const pbb_kernel_desc *kernel, double L, int32_t tree_level, int32_t order,
    double eps, int32_t rank_f, int32_t rank_s, int32_t nsets, const double *U,
    const double *VT, const double *E, pbb_ops **out);

/* Operator cache files in the reference's format (PrecomputeM2L /
 * StartPrecompute / ComputeKernelCheb / FMMReadMatrices,
 * H2_3D_Tree.cpp:74-99, :138-162, :424-441, :496-515, :528-558, :186-239):
 *   K file ({kernelType}ChebK{n}.bin): double L, int32 tree_level, then
 *     per operator set 316 x (rank_f x rank_s) column-major E blocks;
 *   U file ({kernelType}ChebU{n}.bin): int32 rank_f, then per set U;
 *   V file ({kernelType}ChebV{n}.bin): int32 rank_s, then per set VT.
 * One operator set for homogeneous kernels, else one per level 2..L.
 * pbb_ops_read_cache returns PBB_ERR_NOT_FOUND when a file is missing and
 * PBB_ERR_INVALID_ARG when the files do not fit the order, or (per-level
 * operators) were written for another L or tree_level, as
 * PrecomputeAvailable (:102-131) decides.  pbb_ops_write_cache also writes
 * "<k_path>.params", the kernel id, symmetry, homogeneity, parameters,
 * order, eps (and L, tree_level for per-level operators) the files hold;
 * pbb_ops_read_cache refuses files whose parameter file names others, and
 * reads files without one (the reference's) as the reference does.       */
PBB_API pbb_status pbb_ops_write_cache(const pbb_ops *ops, const char *k_path,
                                       const char *u_path, const char *v_path);
PBB_API pbb_status pbb_ops_read_cache(const pbb_kernel_desc *kernel, double L,
                                      int32_t tree_level, int32_t order,
                                      double eps, const char *k_path,
                                      const char *u_path, const char *v_path,
                                      pbb_ops **out);

PBB_API pbb_status pbb_ops_query(const pbb_ops *ops, pbb_ops_info *info);

/* Copies the host operators out (any pointer may be NULL).  sigma_f and
 * sigma_s receive the n3 singular values of K_fat and K_thin of the first
 * operator set (H2_3D_Tree.cpp:400, :471).                               */
PBB_API pbb_status pbb_ops_export(const pbb_ops *ops, double *U, double *VT,
                                  double *E, double *sigma_f, double *sigma_s);

/* The mixed-precision modes run each operator set re-compressed: the
 * reference keeps the level-2 SVD cutoff at every level of a
 * non-homogeneous kernel (H2_3D_Tree.cpp:406-436), while on small boxes a
 * smooth kernel's 316 M2L blocks E_d share far fewer singular directions.
 * Each set's E_d are projected on the leading left / right singular
 * vectors of all 316 blocks together (Q_L, Q_R: E'_d = Q_L^T E_d Q_R, U' =
 * U Q_L, VT' = Q_R^T VT), keeping kf (a multiple of 16) and ks (16 or a
 * multiple of 32) directions so that the discarded singular values of the
 * stacked blocks hold at most tol of their Frobenius norm on each side;
 * where that saves nothing the set keeps its own factors.  No reference
 * counterpart (the reference has no mixed mode).  tol in [0, 1e-6],
 * default 2e-8 (far below the mixed modes' 1e-6 bar); 0 keeps
 * every set's own factors.  Set it before the first mixed-precision call
 * on these operators (afterwards PBB_ERR_INVALID_ARG).  The FP64 mode
 * always runs the reference's operators.                                  */
PBB_API pbb_status pbb_ops_set_mixed_tolerance(pbb_ops *ops, double tol);
/* enable != 0: the FP64 mode runs the re-compressed operator sets too
 * (faster on smooth non-homogeneous kernels; its result then differs from
 * the reference's by about the tolerance); 0 (default): the reference's
 * operators.  Takes effect from the next call on these operators; a batch
 * captured before keeps the operators it was captured with.               */
PBB_API pbb_status pbb_ops_set_fp64_recompressed(pbb_ops *ops, int32_t enable);
/* The ranks the mixed modes keep per operator set (arrays of nsets: sets in
 * level order 2..L, or one set for a homogeneous kernel; ks as kept, before
 * the padding of the proxies to 32-deep chunks) and the discarded relative
 * Frobenius norm (tail may be NULL).  Computes them on the host if needed;
 * no device work.                                                         */
PBB_API pbb_status pbb_ops_mixed_ranks(pbb_ops *ops, int32_t *kf, int32_t *ks, double *tail);

PBB_API void pbb_ops_destroy(pbb_ops *ops);

/* ---- plans (prepared tree, device resident) --------------------------- */

/* Replaces FMMDistribute + InteractionList (compute.hpp:157-190,
 * :1042-1203) and the bbox re-centring (compute.hpp:92-127) for one point
 * set (host AoS xyz, N points).  The tree is built on the device (Morton
 * keys, radix sort, per-level box lists) and kept there; unlike the
 * reference a plan can serve any number of matvecs.                      */
PBB_API pbb_status pbb_plan_create(pbb_ops *ops, const double *xyz,
                                   int64_t npoints, pbb_plan **out);

/* pbb_plan_create with flags: PBB_PLAN_HOST_TREE builds the tree on the
 * host (a CPU restatement of FMMDistribute / InteractionList that the GPU
 * build reproduces array for array, kept as its check); flags 0 builds it
 * on the GPU, as pbb_plan_create does.  PBB_PLAN_REJECT_OUTSIDE fails with
 * PBB_ERR_INVALID_ARG when a point lies outside the root cube (side L
 * around the bbox centre); without it such points are extrapolated, as in
 * the reference, and counted in pbb_plan_info.outside_points.
 * PBB_PLAN_EXPANSION_ROUTE keeps the n^3-coefficient expansions S and F of
 * every level in the mixed modes: without it, the levels whose re-compressed
 * ranks are at most 32 form their proxies straight from the points and
 * evaluate the M2L results straight at the points (the same sums in
 * another order, without the expansions in device memory).               */
#define PBB_PLAN_HOST_TREE 1
#define PBB_PLAN_REJECT_OUTSIDE 2
#define PBB_PLAN_EXPANSION_ROUTE 4
PBB_API pbb_status pbb_plan_create_ex(pbb_ops *ops, const double *xyz,
                                      int64_t npoints, int32_t flags,
                                      pbb_plan **out);

/* A plan of separate targets and sources: f(x_i) = sum_j K(x_i, y_j) w_j for
 * ntargets targets x and nsources sources y (host AoS).  Replaces
 * H2_3D_Compute's separate target and source arguments (compute.hpp:20),
 * which the reference re-centres together on the targets' bbox
 * (compute.hpp:92-127) but distributes as one set (targetlist,
 * compute.hpp:1134-1139), so that it needs them equal.  Here one tree holds
 * the targets followed by the sources, centred on the targets' bbox; the
 * sources carry the weights and the targets receive the outputs.  Then
 * pbb_plan_matvec takes w_dev column-major nsources x ncols and accumulates
 * into out_dev column-major ntargets x ncols, and pbb_plan_matvec_transpose
 * the reverse (weights on the targets, outputs at the sources).  Position
 * gradients and pbb_plan_update_points need one point set
 * (PBB_ERR_UNSUPPORTED).  flags: as pbb_plan_create_ex.                  */
PBB_API pbb_status pbb_plan_create_ts(pbb_ops *ops, const double *targets,
                                      int64_t ntargets, const double *sources,
                                      int64_t nsources, int32_t flags,
                                      pbb_plan **out);

PBB_API pbb_status pbb_plan_query(const pbb_plan *plan, pbb_plan_info *info);

/* Moves the plan's points without rebuilding the tree: xyz (host AoS, the
 * plan's N points in the caller's order) replaces the coordinates; box
 * membership, the sorted order and the bbox centre subtracted at creation
 * stay as built (the reference builds a new tree per H2_3D_Compute, so it
 * has no counterpart; this is the tree a finite difference or an
 * optimisation step with small moves needs).  Every point must stay in its
 * leaf box, else PBB_ERR_INVALID_ARG and the plan is unchanged.  The
 * copies run on `stream` after the work queued there, and the call returns
 * once they are done.                                                     */
PBB_API pbb_status pbb_plan_update_points(pbb_plan *plan, const double *xyz, int64_t npoints,
                                          void *stream);

/* Replaces FMMCompute (compute.hpp:192-222): upward pass, M2L, downward
 * pass, near field.  w_dev: device FP64 weights, column-major N x ncols;
 * out_dev: device FP64, column-major N x ncols, accumulated (+=) (with
 * separate sets: nsources x ncols in, ntargets x ncols out).
 * stream: a hipStream_t (NULL = the null stream), on which all kernels run
 * in order.  Asynchronous with respect to the host; the call returns once
 * all work is enqueued.  Every call on one plan (matvec, transpose,
 * gradients, tuning) uses the plan's one device workspace: issue them on
 * one stream, or order the streams yourself; calls on different plans are
 * independent and may run on any streams and host threads.              */
PBB_API pbb_status pbb_plan_matvec(pbb_plan *plan, const double *w_dev,
                                   int32_t ncols, double *out_dev,
                                   int32_t precision, void *stream);

/* pbb_plan_matvec without any device allocation: no reference counterpart.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as pbb_plan_matvec; the workspace comes from a
 * fixed array of PBB_NOALLOC_SCRATCH_BYTES bytes per GPU that the library
 * holds until the process ends, instead of the plan's own workspace.
 * Fails with PBB_ERR_INVALID_ARG, before any launch, when the matvec needs
 * more workspace than that (roughly 3*N*ncols*8 bytes of sorted weights
 * plus, per non-empty box of levels >= 2, (n^3 + 2*rank)*ncols*8 bytes in
 * FP64); a mixed precision needs its split operators made first, by one
 * pbb_plan_matvec or pbb_plan_tune_m2l in that precision on a plan of the
 * same operators.  Every call on one GPU shares that GPU's array, which
 * keeps the last call's contents: calls that overlap in time on one GPU
 * (other streams, other host threads, other plans) race; serialise them. */
#define PBB_NOALLOC_SCRATCH_BYTES (64u << 20)
PBB_API pbb_status pbb_plan_matvec_noalloc(pbb_plan *plan, const double *w_dev,
                                           int32_t ncols, double *out_dev,
                                           int32_t precision, void *stream);

/* The adjoint: out += A~^T w for the same tree, where A~ is the matrix the
 * FMM applies.  Runs the transposed FMM (P2M <-> L2P and M2M <-> L2L swap
 * roles, the M2L uses V^T/U swapped and E_{-d}^T) with the same kernels;
 * this is the gradient of a loss with respect to the weights.  No reference
 * counterpart (the reference has no adjoint); used by the autograd
 * function.  Requires a symmetric kernel for the near field.  With
 * separate sets (pbb_plan_create_ts) w_dev holds ntargets x ncols and
 * out_dev nsources x ncols.                                               */
PBB_API pbb_status pbb_plan_matvec_transpose(pbb_plan *plan, const double *w_dev,
                                             int32_t ncols, double *out_dev,
                                             int32_t precision, void *stream);

/* Position gradients of L(out), out = A~(x) sigma, given psi = dL/dout:
 * grad[3*i + d] += dL/dx_i[d] (AoS N x 3, caller's order) for the tree
 * fixed (box membership and the bbox centre are piecewise constant in x).
 * Target side: the derivative of the L2P interpolant and of the P2P kernel
 * at x_i weighted by psi; source side: the same for the transposed FMM run
 * on psi, weighted by sigma.  sigma, psi: device FP64, column-major N x
 * ncols.  No reference counterpart; used by the autograd function.        */
PBB_API pbb_status pbb_plan_grad_positions(pbb_plan *plan, const double *sigma_dev,
                                           const double *psi_dev, int32_t ncols,
                                           double *grad_dev, int32_t precision,
                                           void *stream);

/* Caps the plan's matvec workspace at `bytes` (0: what the device holds).
 * A matvec whose workspace for all its columns does not fit runs as
 * chunks of columns, each a complete FMM pass with the same sums (phase
 * times then cover the last chunk).  No reference counterpart.           */
PBB_API pbb_status pbb_plan_set_workspace_limit(pbb_plan *plan, int64_t bytes);

/* Phase times of the last pbb_plan_matvec on this plan, in ms
 * (PBB_NUM_PHASES values).  Synchronises the plan's events.               */
PBB_API pbb_status pbb_plan_timings(pbb_plan *plan, double *ms);

/* Enables (1) or disables (0, default) per-phase event timing.            */
PBB_API pbb_status pbb_plan_set_timing(pbb_plan *plan, int32_t enable);

/* M2L tile of the FP64 mode (precision PBB_PREC_FP64) or of the mixed
 * modes (any mixed precision): `width` columns per workgroup tile (FP64 32,
 * 64 or 128; mixed 64, 128 or 256), `cols_per_box` weight columns of each
 * target box in a tile (1..64), `flags`: PBB_M2L_CHUNK_MAJOR for chunk-major
 * steps (one k chunk of every direction, then the next chunk; otherwise the
 * k chunks of one E_d, then the next direction), PBB_M2L_ROW_SPLIT (mixed
 * modes) to split the rows of r_f over two workgroups per tile,
 * PBB_M2L_WAVE_ROWS (mixed modes, no row split) for the kernel whose
 * waves split the rows of r_f and load E_d into registers (ranks above 192,
 * or above 128 in PBB_PREC_MIXED, cap its width at 128; ranks above 224
 * always run it, as row groups of workgroups above 256 rows),
 * PBB_M2L_SPLIT_ACC (with PBB_M2L_WAVE_ROWS; PBB_PREC_MIXED_F16X2 and
 * PBB_PREC_MIXED_BF16X2, ignored in PBB_PREC_MIXED) to sum the split's
 * cross products apart from the leading products, so that the FP32 sums
 * round a third as often, with the VT and U products in FP64: a smaller
 * difference from the FP64 mode at a cost (tiles capped at 64 columns),
 * PBB_M2L_DIR_GROUPS(g) (mixed modes, g <= 64) to split the 316 M2L
 * operators into g groups of consecutive index, one launch per group on the
 * large levels, so that a launch's operators stay in L2.  width 0 restores
 * the built-in choice.  No reference counterpart: tuning of this build's
 * M2L.                                                                    */
#define PBB_M2L_CHUNK_MAJOR 1
#define PBB_M2L_ROW_SPLIT 2
#define PBB_M2L_WAVE_ROWS 4
#define PBB_M2L_SPLIT_ACC 8
#define PBB_M2L_DIR_GROUPS(g) ((g) << 8)
PBB_API pbb_status pbb_plan_set_m2l_tile(pbb_plan *plan, int32_t precision, int32_t width,
                                         int32_t cols_per_box, int32_t flags);

/* The M2L tile of one tree level (2..tree_level) of the precision's class,
 * overriding the plan-wide tile there (pbb_plan_set_m2l_tile clears the
 * overrides of its class).  PBB_M2L_SPLIT_ACC stays a plan-wide choice: it
 * also sets the VT / U precision of the matvec.  pbb_plan_get_m2l_tile
 * returns the tile in force at a level (level -1: the plan-wide tile) as
 * tile[3] = {width, cols_per_box, flags}.  No reference counterpart.     */
PBB_API pbb_status pbb_plan_set_m2l_tile_level(pbb_plan *plan, int32_t precision, int32_t level,
                                               int32_t width, int32_t cols_per_box, int32_t flags);
PBB_API pbb_status pbb_plan_get_m2l_tile(const pbb_plan *plan, int32_t precision, int32_t level,
                                         int32_t *tile);

/* Autotuner: times the M2L of one matvec on w_dev (device FP64, N x ncols)
 * for every candidate tile of the precision's class, keeps the fastest on
 * the plan and returns it in best[3] = {width, cols_per_box, flags} (may be
 * NULL).  From the same runs it also picks the fastest tile of each tree
 * level and keeps those per-level tiles when the combination is faster
 * (pbb_plan_get_m2l_tile reads them).  Runs on `stream`, waits for it, and
 * uses a scratch output of its own.  No reference counterpart.           */
PBB_API pbb_status pbb_plan_tune_m2l(pbb_plan *plan, const double *w_dev, int32_t ncols,
                                     int32_t precision, void *stream, int32_t *best);

/* The autotuner with an accuracy bound (mixed modes): runs the FP64 matvec
 * on w_dev once, and keeps the fastest candidate whose matvec differs from
 * it by at most max_rel_diff (||y - y64|| / ||y64||); when no candidate
 * does, it keeps the most accurate one.  *achieved (may be NULL) receives
 * the kept candidate's difference, so the caller can tell the two cases
 * apart.  max_rel_diff 0 (or the FP64 mode) tunes for speed alone.  No
 * reference counterpart.                                                  */
PBB_API pbb_status pbb_plan_tune_m2l_within(pbb_plan *plan, const double *w_dev, int32_t ncols,
                                            int32_t precision, double max_rel_diff, void *stream,
                                            int32_t *best, double *achieved);

PBB_API void pbb_plan_destroy(pbb_plan *plan);

/* ---- batches of clouds (one HIP graph) -------------------------------- */

typedef struct pbb_batch pbb_batch;

#define PBB_BATCH_TRANSPOSE 1  /* run the transposed FMM of every plan      */
#define PBB_BATCH_OVERWRITE 2  /* clear each output first (else +=)         */

/* Captures the matvecs of nplans plans (any sizes, orders and tree levels,
 * one device) on weights w_dev[i] / outputs out_dev[i] (device FP64,
 * column-major N_i x ncols) into one HIP graph.  The buffers are bound at
 * capture: refill them in place between launches.  The batch allocates
 * its own workspaces here (one per distinct plan), none during a launch,
 * so calls on the plans themselves never disturb it; the plans must
 * outlive the batch.  The entries run as up to four concurrent graph
 * branches; entries that share a plan, or whose weights or outputs overlap
 * another entry's output (e.g. a chain A(A w)), are put in one branch and
 * run there in batch order.  No reference counterpart: the reference runs
 * one cloud per H2_3D_Compute.                                           */
PBB_API pbb_status pbb_batch_create(pbb_plan *const *plans, int32_t nplans,
                                    const double *const *w_dev, double *const *out_dev,
                                    int32_t ncols, int32_t precision, int32_t flags,
                                    pbb_batch **out);
/* One entry of a batch with flags of its own: out_dev (+)= A w_dev, or the
 * transposed FMM with PBB_BATCH_TRANSPOSE (e.g. a forward matvec and an
 * adjoint of the same plan in one graph launch).                         */
typedef struct pbb_batch_entry {
    pbb_plan *plan;
    const double *w_dev;   /* device FP64, column-major N x ncols          */
    double *out_dev;       /* device FP64, column-major N x ncols          */
    int32_t flags;         /* PBB_BATCH_TRANSPOSE | PBB_BATCH_OVERWRITE    */
    int32_t reserved;      /* 0                                            */
} pbb_batch_entry;

/* pbb_batch_create over entries that carry their own flags; same rules
 * (workspaces, branches, conflicts: an entry and the transposed entry of
 * the same plan share its workspace, so they run in one branch in batch
 * order).  No reference counterpart.                                     */
PBB_API pbb_status pbb_batch_create_entries(const pbb_batch_entry *entries, int32_t nentries,
                                            int32_t ncols, int32_t precision, pbb_batch **out);

/* Replays the graph on stream (NULL = the null stream). */
PBB_API pbb_status pbb_batch_launch(pbb_batch *batch, void *stream);
PBB_API void pbb_batch_destroy(pbb_batch *batch);

/* ---- accuracy harness ------------------------------------------------- */

/* Replaces DirectCalc3D (test.hpp:18-34): the O(T*N) FP64 direct sum
 * out[c*T + i] = sum_j K(t_i, s_j) w[c*N + j] with isinf -> 0, on device
 * buffers (targets and sources AoS xyz).  Overwrites out; synchronous.     */
PBB_API pbb_status pbb_direct_sum(const pbb_kernel_desc *kernel, const double *tgt_dev,
                                  int64_t ntargets, const double *src_dev, int64_t nsources,
                                  const double *w_dev, int32_t ncols, double *out_dev,
                                  void *stream);

/* ---- one-shot drop-in ------------------------------------------------- */

/* Replaces the H2_3D_Compute<T> constructor (compute.hpp:67-149): host
 * FP64 buffers of N targets and N sources, weights column-major N x ncols,
 * output accumulated.  Builds a plan (one point set when target and source
 * hold the same points, else pbb_plan_create_ts), runs the FP64 matvec,
 * copies back, frees the plan.                                            */
PBB_API pbb_status pbb_matvec(pbb_ops *ops, const double *target,
                              const double *source, int64_t npoints,
                              const double *weight, int32_t ncols,
                              double *output);

/* pbb_matvec with ntargets targets and nsources sources: weights
 * column-major nsources x ncols, output ntargets x ncols, accumulated.    */
PBB_API pbb_status pbb_matvec_ts(pbb_ops *ops, const double *target,
                                 int64_t ntargets, const double *source,
                                 int64_t nsources, const double *weight,
                                 int32_t ncols, double *output);

/* The H2_3D_Compute<T> constructor (compute.hpp:67-149) in a chosen
 * precision mode (pbb_precision): the reference computes in FP64 only, so
 * pbb_matvec and pbb_matvec_ts do too; this entry lets callers of the
 * reference-shaped surfaces (H2_3D_Tree::precision, FMMCompute.Compute's
 * precision argument) run the mixed modes.  Targets and sources holding
 * the same points run as one point set.  Each call builds a fresh plan, so
 * the mixed M2L runs its default tiles (no tuning); a plan kept across
 * calls (pbb_plan_create + pbb_plan_tune_m2l) is the way to tuned tiles.  */
PBB_API pbb_status pbb_matvec_prec(pbb_ops *ops, const double *target,
                                   int64_t ntargets, const double *source,
                                   int64_t nsources, const double *weight,
                                   int32_t ncols, double *output,
                                   int32_t precision);

#ifdef __cplusplus
}
#endif

#endif /* PBBFMM3D_AMD_H */
