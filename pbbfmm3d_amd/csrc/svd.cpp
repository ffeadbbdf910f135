// svd.cpp -- FP64 SVD building blocks for the M2L operator precompute.
//
// The reference calls LAPACK dgesvd on K_fat (n3 x 316 n3) and K_thin
// (316 n3 x n3) (H2_3D_Tree.cpp:400, :471).  Both are tall/wide with only
// n3 (<= a few hundred) singular values, so this library reduces each to an
// n3 x n3 triangle R with a Householder TSQR (backward stable: the singular
// values of R are those of the input to within eps * ||K||), then runs the
// one-sided (Hestenes) Jacobi method on R.  The accumulated rotations are
// the right singular vectors of R, orthogonal to working precision however
// small the singular value; the left vectors are never formed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "host.hpp"

namespace pbb {

// In-place Householder QR of a rows x n column-major block (ld rows); on
// return the upper triangle (first min(rows, n) rows) holds R.
static void householder_qr(double *B, int64_t rows, int n)
{
    const int kmax = (int)std::min<int64_t>(rows, n);
    std::vector<double> v(rows);
    for (int k = 0; k < kmax; k++) {
        double *ck = B + (size_t)k * rows;
        double norm2 = 0;
        for (int64_t i = k; i < rows; i++) norm2 += ck[i] * ck[i];
        if (norm2 == 0) continue;
        double norm = std::sqrt(norm2);
        double alpha = ck[k] > 0 ? -norm : norm;
        int64_t len = rows - k;
        for (int64_t i = 0; i < len; i++) v[i] = ck[k + i];
        v[0] -= alpha;
        double vn2 = 0;
        for (int64_t i = 0; i < len; i++) vn2 += v[i] * v[i];
        ck[k] = alpha;
        for (int64_t i = k + 1; i < rows; i++) ck[i] = 0;
        if (vn2 == 0) continue;
        double f = 2.0 / vn2;
        for (int j = k + 1; j < n; j++) {
            double *cj = B + (size_t)j * rows + k;
            double s = 0;
            for (int64_t i = 0; i < len; i++) s += v[i] * cj[i];
            s *= f;
            for (int64_t i = 0; i < len; i++) cj[i] -= s * v[i];
        }
    }
}

// R (n x n, zero below the diagonal) of the rows x n matrix at A (ld lda).
static void block_r(const double *A, int64_t lda, int64_t rows, int n, double *R)
{
    std::vector<double> B((size_t)rows * n);
    for (int j = 0; j < n; j++) std::memcpy(&B[(size_t)j * rows], A + (size_t)j * lda, sizeof(double) * rows);
    householder_qr(B.data(), rows, n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) R[i + (size_t)j * n] = (i <= j && i < rows) ? B[i + (size_t)j * rows] : 0.0;
}

void tsqr_r(double *A, int64_t m, int n, std::vector<double> &R)
{
    R.assign((size_t)n * n, 0.0);
    const int64_t minrows = std::max<int64_t>(4LL * n, 512);
    if (m <= 2 * minrows) {
        block_r(A, m, m, n, R.data());
        return;
    }
    int64_t nb = std::min<int64_t>(m / minrows, 256);
    int64_t bs = (m + nb - 1) / nb;
    nb = (m + bs - 1) / bs;
    // stacked block triangles, (nb*n) x n column-major
    std::vector<double> S((size_t)nb * n * n, 0.0);
    const int64_t ldS = nb * n;
    parallel_for(nb, [&](int64_t b) {
        int64_t r0 = b * bs, rows = std::min(bs, m - r0);
        std::vector<double> Rb((size_t)n * n);
        block_r(A + r0, m, rows, n, Rb.data());
        for (int j = 0; j < n; j++)
            std::memcpy(&S[(size_t)j * ldS + b * n], &Rb[(size_t)j * n], sizeof(double) * n);
    });
    tsqr_r(S.data(), ldS, n, R);
}

void jacobi_svd_right(double *A, int n, std::vector<double> &sigma, std::vector<double> &V)
{
    std::vector<double> W(A, A + (size_t)n * n);
    std::vector<double> Vr((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Vr[i + (size_t)i * n] = 1.0;
    std::vector<double> nrm(n);
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += W[i + (size_t)j * n] * W[i + (size_t)j * n];
        nrm[j] = s;
    }
    const double tol = 1e-15;
    for (int sweep = 0; sweep < 80; sweep++) {
        int rotations = 0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double *wp = &W[(size_t)p * n], *wq = &W[(size_t)q * n];
                double alpha = nrm[p], beta = nrm[q], gamma = 0;
                for (int i = 0; i < n; i++) gamma += wp[i] * wq[i];
                if (alpha == 0 || beta == 0) continue;
                if (std::fabs(gamma) <= tol * std::sqrt(alpha * beta)) continue;
                rotations++;
                double zeta = (beta - alpha) / (2 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                double c = 1 / std::sqrt(1 + t * t), s = c * t;
                double na = 0, nb = 0;
                for (int i = 0; i < n; i++) {
                    double a = wp[i], b = wq[i];
                    double x = c * a - s * b, y = s * a + c * b;
                    wp[i] = x; wq[i] = y;
                    na += x * x; nb += y * y;
                }
                nrm[p] = na; nrm[q] = nb;
                double *vp = &Vr[(size_t)p * n], *vq = &Vr[(size_t)q * n];
                for (int i = 0; i < n; i++) {
                    double a = vp[i], b = vq[i];
                    vp[i] = c * a - s * b;
                    vq[i] = s * a + c * b;
                }
            }
        if (rotations == 0) break;
    }
    std::vector<double> sv(n);
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += W[i + (size_t)j * n] * W[i + (size_t)j * n];
        sv[j] = std::sqrt(s);
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return sv[a] > sv[b]; });
    sigma.resize(n);
    V.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; k++) {
        sigma[k] = sv[idx[k]];
        std::memcpy(&V[(size_t)k * n], &Vr[(size_t)idx[k] * n], sizeof(double) * n);
    }
}

}  // namespace pbb
