// precompute.cpp -- host FP64 operator precompute of the Chebyshev bbFMM.
//
// Follows H2_3D_Tree::buildFMMTree (H2_3D_Tree.cpp:40-66): ComputeWeights
// (:715-826), then for each operator set the 316 direction blocks of the
// weighted kernel (:277-374), the SVD of K_fat (n3 x 316 n3) for U and of
// K_thin (316 n3 x n3) for VT with the tail-energy cutoff (:376-516), and
// E_d = U^T K_d V (:518-556).  The SVD is this library's own: a Householder
// TSQR reduces each wide/tall matrix to an n3 x n3 triangle, and a one-sided
// Jacobi sweep on that triangle gives singular values and right singular
// vectors to full FP64 accuracy (svd.cpp).  No LAPACK is needed.
#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <cmath>
#include <exception>
#include <cstring>
#include <thread>

#include "host.hpp"

namespace pbb {

void fail(pbb_status st, const std::string &msg) { throw Error{st, msg}; }

void parallel_for(int64_t n, const std::function<void(int64_t)> &body)
{
    if (n <= 0) return;
    // at most the hardware threads, and at most OMP_NUM_THREADS when it is set
    // (a host's share of its cores for this process)
    int nt = (int)std::max(1u, std::thread::hardware_concurrency());
    if (const char *e = std::getenv("OMP_NUM_THREADS")) {
        const int cap = std::atoi(e);
        if (cap > 0) nt = std::min(nt, cap);
    }
    nt = (int)std::min<int64_t>(nt, n);
    if (nt == 1) {
        for (int64_t i = 0; i < n; i++) body(i);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> errs(nt);
    std::atomic<int64_t> next{0};
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            try {
                for (;;) {
                    int64_t i = next.fetch_add(1);
                    if (i >= n) break;
                    body(i);
                }
            } catch (...) {
                errs[t] = std::current_exception();
                next.store(n);
            }
        });
    for (auto &x : th) x.join();
    for (auto &e : errs)
        if (e) std::rethrow_exception(e);
}

bool kernel_id_valid(int id) { return id >= PBB_KERNEL_LAPLACIAN && id <= PBB_KERNEL_EXP_ANISO; }

// Formulas of pbb_kernel_id (kernel_Types.cpp:15-178 for the built-ins).
double kernel_eval_host(int id, const double *p, const double *t, const double *s)
{
    double dx = s[0] - t[0], dy = s[1] - t[1], dz = s[2] - t[2];
    double r2 = dx * dx + dy * dy + dz * dz;
    switch (id) {
    case PBB_KERNEL_LAPLACIAN: return 1. / std::sqrt(r2);
    case PBB_KERNEL_ONEOVERR2: { double ri = 1. / std::sqrt(r2); return ri * ri; }
    case PBB_KERNEL_GAUSSIAN: return std::exp(-r2 / (p[0] * p[0]));
    case PBB_KERNEL_EXPONENTIAL: return p[1] * std::exp(-std::sqrt(r2) / p[0]);
    case PBB_KERNEL_SOFT_LAPLACIAN: return 1. / std::sqrt(r2 + p[0]);
    case PBB_KERNEL_QUADRIC: return 1.0 + r2;
    case PBB_KERNEL_INV_QUADRIC: return 1.0 / (1.0 + r2);
    case PBB_KERNEL_LOGARITHM: return r2 == 0 ? 0. : 0.5 * std::log(r2);
    case PBB_KERNEL_THIN_PLATE: return r2 == 0 ? 0. : 0.5 * r2 * std::log(r2);
    case PBB_KERNEL_R2: return r2;
    case PBB_KERNEL_CONSTANT: return 1.0;
    case PBB_KERNEL_POLY_XY: return (t[0] - s[0]) * (t[1] + s[1]) + p[0];
    case PBB_KERNEL_EXP_ANISO: return p[0] * std::exp(-std::sqrt(p[1] * dx * dx + p[2] * dy * dy + p[3] * dz * dz));
    default: return NAN;
    }
}

void cheb_sn(int n, const double *Tkz, double x, double *S)
{
    const double pfac = 2. / n;
    for (int m = 0; m < n; m++) {
        const double *Tk = Tkz + m * n;
        double d0 = 0, d1 = 0;
        for (int j = n - 1; j > 0; j--) {
            double d2 = d0;
            d0 = 2.0 * x * d0 - d1 + Tk[j];
            d1 = d2;
        }
        S[m] = pfac * (x * d0 - d1 + 0.5 * Tk[0]);
    }
}

void ChebWeights::compute(int order)
{
    n = order;
    int n2 = n * n, n3 = n2 * n;
    nodes.resize(n);
    for (int m = 0; m < n; m++) nodes[m] = std::cos(M_PI * ((double)m + 0.5) / (double)n);
    Tkz.assign(n2, 0);
    for (int m = 0; m < n; m++) {
        double x = nodes[m], *v = &Tkz[m * n];
        v[0] = 1;
        if (n > 1) v[1] = x;
        for (int k = 2; k < n; k++) v[k] = 2.0 * x * v[k - 1] - v[k - 2];
    }
    Kw.resize(n3);
    int c = 0;
    for (int l1 = 0; l1 < n; l1++) {
        double t1 = 1. / std::sqrt(1 - nodes[l1] * nodes[l1]);
        for (int l2 = 0; l2 < n; l2++) {
            double t2 = t1 / std::sqrt(1 - nodes[l2] * nodes[l2]);
            for (int l3 = 0; l3 < n; l3++) Kw[c++] = t2 / std::sqrt(1 - nodes[l3] * nodes[l3]);
        }
    }
    // Cweights: child nodes mapped into the parent frame (H2_3D_Tree.cpp:775-819)
    Cw.assign(2 * n2, 0);
    std::vector<double> S(n);
    for (int half = 0; half < 2; half++)
        for (int j = 0; j < n; j++) {
            cheb_sn(n, Tkz.data(), nodes[j] * .5 + (half ? 0.5 : -0.5), S.data());
            for (int m = 0; m < n; m++) Cw[half * n2 + m * n + j] = S[m];
        }
    Ktable.assign(343, -1);
    int ncell = 0, ninter = 0;
    for (int l1 = -3; l1 < 4; l1++)
        for (int l2 = -3; l2 < 4; l2++)
            for (int l3 = -3; l3 < 4; l3++, ncell++)
                if (std::abs(l1) > 1 || std::abs(l2) > 1 || std::abs(l3) > 1) Ktable[ncell] = ninter++;
}

int cutoff_rank(const std::vector<double> &sigma, double eps)
{
    int m = (int)sigma.size();
    double sumsqr = 0;
    for (int i = m - 1; i >= 0; --i) sumsqr += sigma[i] * sigma[i];
    double sum = 0, epsqr = sumsqr * eps * eps;
    int cut = m;
    for (int i = m - 1; i >= 0; --i) {
        sum += sigma[i] * sigma[i];
        if (sum < epsqr) --cut;
        else break;
    }
    return cut;
}

void kernel_blocks(const HostOps &ops, const HostEval &eval, double boxLen, std::vector<double> &Kcell)
{
    const int n = ops.n, n3 = ops.n3;
    const size_t n6 = (size_t)n3 * n3;
    const int symm = ops.kernel.symmetry;
    const int symmNum = 158 * (2 - std::abs(symm));
    const auto &nodes = ops.w.nodes;
    const auto &Kw = ops.w.Kw;
    int dirs[316][3], c = 0;
    for (int k1 = -3; k1 < 4; k1++)
        for (int k2 = -3; k2 < 4; k2++)
            for (int k3 = -3; k3 < 4; k3++)
                if (std::abs(k1) > 1 || std::abs(k2) > 1 || std::abs(k3) > 1) {
                    dirs[c][0] = k1; dirs[c][1] = k2; dirs[c][2] = k3; c++;
                }
    Kcell.assign(316 * n6, 0.0);
    parallel_for(symmNum, [&](int64_t d) {
        double *K = &Kcell[d * n6];
        for (int j = 0; j < n3; j++) {
            int j1 = j / (n * n), j2 = (j / n) % n, j3 = j % n;
            double s[3] = {(dirs[d][0] + 0.5 * nodes[j1]) * boxLen, (dirs[d][1] + 0.5 * nodes[j2]) * boxLen,
                           (dirs[d][2] + 0.5 * nodes[j3]) * boxLen};
            for (int i = 0; i < n3; i++) {
                int i1 = i / (n * n), i2 = (i / n) % n, i3 = i % n;
                double t[3] = {0.5 * nodes[i1] * boxLen, 0.5 * nodes[i2] * boxLen, 0.5 * nodes[i3] * boxLen};
                K[i + (size_t)j * n3] = eval(t, s) / (Kw[i] * Kw[j]);
            }
        }
    });
    if (symm) {
        for (int d = symmNum; d < 316; d++) {
            double *K = &Kcell[d * n6];
            const double *Ks = &Kcell[(size_t)(315 - d) * n6];
            for (int j = 0; j < n3; j++)
                for (int i = 0; i < n3; i++) K[i + (size_t)j * n3] = symm * Ks[j + (size_t)i * n3];
        }
    }
}

// E_d = U^T K_d V for the 316 directions (H2_3D_Tree.cpp:529-556).
static void compress(int n3, int rf, int rs, const double *U, const double *VT, const std::vector<double> &Kcell,
                     double *E)
{
    parallel_for(316, [&](int64_t d) {
        const double *K = &Kcell[d * (size_t)n3 * n3];
        std::vector<double> KV((size_t)n3 * rs, 0.0);
        // KV = K * V, V(j, b) = VT[b + j*rs]
        for (int j = 0; j < n3; j++)
            for (int b = 0; b < rs; b++) {
                double v = VT[b + (size_t)j * rs];
                const double *kc = K + (size_t)j * n3;
                double *kv = &KV[(size_t)b * n3];
                for (int i = 0; i < n3; i++) kv[i] += kc[i] * v;
            }
        double *Ed = E + d * (size_t)rf * rs;
        for (int b = 0; b < rs; b++)
            for (int a = 0; a < rf; a++) {
                const double *u = U + (size_t)a * n3, *kv = &KV[(size_t)b * n3];
                double s = 0;
                for (int i = 0; i < n3; i++) s += u[i] * kv[i];
                Ed[a + (size_t)b * rf] = s;
            }
    });
}

void precompute_ops(HostOps &ops, const HostEval &eval, const PrecomputeDevice *dev)
{
    const int n = ops.n;
    ops.n3 = n * n * n;
    const int n3 = ops.n3;
    ops.w.compute(n);
    ops.homogen_flipped = -ops.kernel.homogen;
    ops.homog = std::fabs(ops.homogen_flipped) > 0.1;  // IsHomoKernel, H2_3D_Tree.cpp:134-136
    ops.nsets = ops.homog ? 1 : std::max(1, ops.tree_level - 1);
    const int64_t m = 316LL * n3;

    std::vector<double> Kcell, A, R, sig, V;
    std::vector<double> Ufull, VTfull;
    for (int set = 0; set < ops.nsets; set++) {
        // homogeneous: one unit-length set; otherwise levels 2.. with L/4, L/8, ...
        // (StartPrecompute, H2_3D_Tree.cpp:138-162)
        double boxLen = ops.homog ? 1.0 : ops.L / 4;
        if (!ops.homog)
            for (int s = 0; s < set; s++) boxLen /= 2;
        if (dev && dev->blocks) dev->blocks(boxLen, Kcell);
        else kernel_blocks(ops, eval, boxLen, Kcell);
        const size_t n6 = (size_t)n3 * n3;

        // K_fat^T: row (d, j), column i -> value K_d(i, j); left singular
        // vectors of K_fat are the right singular vectors of its R factor.
        A.assign((size_t)m * n3, 0.0);
        parallel_for(n3, [&](int64_t i) {
            double *col = &A[(size_t)i * m];
            for (int d = 0; d < 316; d++)
                for (int j = 0; j < n3; j++) col[(size_t)d * n3 + j] = Kcell[d * n6 + i + (size_t)j * n3];
        });
        tsqr_r(A.data(), m, n3, R);
        jacobi_svd_right(R.data(), n3, sig, V);
        if (set == 0) {
            ops.sigma_f = sig;
            ops.rf = cutoff_rank(sig, ops.eps);
        }
        Ufull = V;  // n3 x n3, columns = U

        // K_thin: row (d, i), column j -> value K_d(i, j)
        A.assign((size_t)m * n3, 0.0);
        parallel_for(n3, [&](int64_t j) {
            double *col = &A[(size_t)j * m];
            for (int d = 0; d < 316; d++)
                std::memcpy(col + (size_t)d * n3, &Kcell[d * n6 + (size_t)j * n3], sizeof(double) * n3);
        });
        tsqr_r(A.data(), m, n3, R);
        jacobi_svd_right(R.data(), n3, sig, V);
        if (set == 0) {
            ops.sigma_s = sig;
            ops.rs = cutoff_rank(sig, ops.eps);
            ops.U.assign((size_t)ops.nsets * n3 * ops.rf, 0.0);
            ops.VT.assign((size_t)ops.nsets * ops.rs * n3, 0.0);
            ops.E.assign((size_t)ops.nsets * 316 * ops.rf * ops.rs, 0.0);
        }
        const int rf = ops.rf, rs = ops.rs;
        double *U = &ops.U[(size_t)set * n3 * rf];
        double *VT = &ops.VT[(size_t)set * rs * n3];
        std::memcpy(U, Ufull.data(), sizeof(double) * (size_t)n3 * rf);
        for (int j = 0; j < n3; j++)
            for (int b = 0; b < rs; b++) VT[b + (size_t)j * rs] = V[j + (size_t)b * n3];
        if (dev && dev->compress) dev->compress(rf, rs, U, VT, &ops.E[(size_t)set * 316 * rf * rs]);
        else compress(n3, rf, rs, U, VT, Kcell, &ops.E[(size_t)set * 316 * rf * rs]);
    }
}

// Far directions of an octant class: the children of the parent's 27
// neighbours that are not among the target's own 27 neighbours
// (InteractionList, compute.hpp:1062-1100, on a uniform grid).  They are
// listed in the order the M2L kernels visit them: grouped by the parity of
// the offset and, inside a group, as a snake over its 3x3x3 grid of step 2
// with z fastest.  Boxes of one class sit 2 apart, so consecutive directions
// of a group reach largely the same source boxes for a tile of neighbouring
// targets.
void class_directions(int octant, const std::vector<int> &Ktable, int dirs[189][4])
{
    const int o[3] = {(octant >> 2) & 1, (octant >> 1) & 1, octant & 1};
    int k = 0;
    for (int par = 0; par < 8; par++) {
        const int s0[3] = {-2 - o[0] + ((par >> 2) & 1), -2 - o[1] + ((par >> 1) & 1), -2 - o[2] + (par & 1)};
        for (int i = 0; i < 3; i++)
            for (int jj = 0; jj < 3; jj++) {
                const int j = (i & 1) ? 2 - jj : jj;
                for (int ll = 0; ll < 3; ll++) {
                    const int l = ((i * 3 + jj) & 1) ? 2 - ll : ll;
                    const int dx = s0[0] + 2 * i, dy = s0[1] + 2 * j, dz = s0[2] + 2 * l;
                    if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
                    dirs[k][0] = dx; dirs[k][1] = dy; dirs[k][2] = dz;
                    dirs[k][3] = Ktable[(dx + 3) * 49 + (dy + 3) * 7 + (dz + 3)];
                    k++;
                }
            }
    }
    if (k != 189) fail(PBB_ERR_INTERNAL, "far-direction count is not 189");
}

// Right singular vectors and values of a tall m x n matrix A (col-major,
// destroyed): TSQR, then the one-sided Jacobi SVD of R.
static void right_svd(std::vector<double> &A, int64_t m, int n, std::vector<double> &sigma,
                      std::vector<double> &V)
{
    std::vector<double> R;
    tsqr_r(A.data(), m, n, R);
    jacobi_svd_right(R.data(), n, sigma, V);
}

// Smallest k (a multiple of `mult`, at most r) whose discarded singular
// values hold at most tau of the Frobenius norm; tail: what they hold.
// the smallest candidate rank (first, then the multiples of mult) whose
// discarded singular values hold at most tau of the Frobenius norm
static int reduced_rank(const std::vector<double> &sigma, int r, int first, int mult, double tau, double &tail)
{
    double tot = 0;
    for (double v : sigma) tot += v * v;
    std::vector<double> rest(sigma.size() + 1, 0.0);
    for (int i = (int)sigma.size() - 1; i >= 0; i--) rest[i] = rest[i + 1] + sigma[i] * sigma[i];
    for (int k = first; k < r; k = k < mult ? mult : k + mult)
        if (rest[k] <= tau * tau * tot) {
            tail = tot > 0 ? std::sqrt(rest[k] / tot) : 0.0;
            return k;
        }
    tail = 0;
    return r;
}

void reduce_set(const HostOps &h, int set, const double *VTw, const double *Uw, double tau, ReducedSet &out)
{
    const int rf = h.rf, rs = h.rs, n3 = h.n3;
    const double *E = &h.E[(size_t)set * 316 * rf * rs];
    const int64_t mR = 316LL * rf, mL = 316LL * rs;
    // [E_0; ..; E_315] (316 rf x rs) for Q_R, [E_0^T; ..; E_315^T] (316 rs x rf) for Q_L
    std::vector<double> BR((size_t)mR * rs), BL((size_t)mL * rf);
    parallel_for(316, [&](int64_t d) {
        const double *Ed = E + (size_t)d * rf * rs;
        for (int j = 0; j < rs; j++)
            for (int i = 0; i < rf; i++) {
                const double v = Ed[i + (size_t)j * rf];
                BR[(size_t)(d * rf + i) + (size_t)j * mR] = v;
                BL[(size_t)(d * rs + j) + (size_t)i * mL] = v;
            }
    });
    std::vector<double> sR, VR, sL, VL;
    right_svd(BR, mR, rs, sR, VR);
    right_svd(BL, mL, rf, sL, VL);
    // ks: 16 (one 16-deep k chunk of the matrix cores) or whole 32-deep chunks
    const int ks = reduced_rank(sR, rs, 16, 32, tau, out.tail_s);
    const int kf = reduced_rank(sL, rf, 16, 16, tau, out.tail_f);
    out.kf = kf;
    out.ks = ks;
    if (kf >= rf && ks >= rs) {  // nothing to gain: the set's own factors
        out.kf = rf;
        out.ks = rs;
        out.tail_f = out.tail_s = 0;
        out.E.assign(E, E + (size_t)316 * rf * rs);
        out.VTw.assign(VTw, VTw + (size_t)rs * n3);
        out.Uw.assign(Uw, Uw + (size_t)n3 * rf);
        return;
    }
    // Q_L = VL[:, :kf], Q_R = VR[:, :ks] (column-major, ld rf / rs)
    out.E.assign((size_t)316 * kf * ks, 0.0);
    parallel_for(316, [&](int64_t d) {
        const double *Ed = E + (size_t)d * rf * rs;
        std::vector<double> T((size_t)rf * ks, 0.0);  // E_d Q_R
        for (int j = 0; j < ks; j++)
            for (int q = 0; q < rs; q++) {
                const double v = VR[q + (size_t)j * rs];
                for (int i = 0; i < rf; i++) T[i + (size_t)j * rf] += Ed[i + (size_t)q * rf] * v;
            }
        double *Eo = &out.E[(size_t)d * kf * ks];
        for (int j = 0; j < ks; j++)
            for (int a = 0; a < kf; a++) {
                double acc = 0;
                for (int i = 0; i < rf; i++) acc += VL[i + (size_t)a * rf] * T[i + (size_t)j * rf];
                Eo[a + (size_t)j * kf] = acc;
            }
    });
    out.VTw.assign((size_t)ks * n3, 0.0);  // Q_R^T VTw
    out.Uw.assign((size_t)n3 * kf, 0.0);   // Uw Q_L
    parallel_for(n3, [&](int64_t q) {
        for (int a = 0; a < ks; a++) {
            double acc = 0;
            for (int i = 0; i < rs; i++) acc += VR[i + (size_t)a * rs] * VTw[i + (size_t)q * rs];
            out.VTw[a + (size_t)q * ks] = acc;
        }
        for (int a = 0; a < kf; a++) {
            double acc = 0;
            for (int i = 0; i < rf; i++) acc += Uw[q + (size_t)i * n3] * VL[i + (size_t)a * rf];
            out.Uw[q + (size_t)a * n3] = acc;
        }
    });
}

}  // namespace pbb
