/*
 * pbb_oracle.c -- CPU restatement of PBBFMM3D's Chebyshev black-box FMM
 * matvec (H2_3D_Compute) and of the closed-form pieces of its operator
 * precompute.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the test suite, the
 * smoke entry point and the benchmark's CPU-baseline leg are the only users.
 * The product library never links, loads or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the PBBFMM3D source tree).  Plain C99 + OpenMP, FP64 throughout, as the
 * reference is.
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   pbbo_cheb_nodes / pbbo_tkz / pbbo_kweights   closed forms
 *   pbbo_sn                                       interpolation identity,
 *                                                 partition of unity,
 *                                                 polynomial reproduction
 *   pbbo_cweights                                 C[m*n+j] = S_m((x_j -+ 1)/2)
 *   pbbo_ktable                                   316 entries, lexicographic
 *   pbbo_kernel_eval                              closed forms per kernel
 *   pbbo_kernel_blocks                            brute force + symmetric fill
 *   pbbo_compress                                 U^T K V identity
 *   pbbo_matvec                                   exact (to rounding) against
 *                                                 the direct sum for kernels
 *                                                 that are polynomials of
 *                                                 degree < n; approximation-
 *                                                 level against the direct
 *                                                 sum for 1/r and Gaussian;
 *                                                 brute-force near field on
 *                                                 tiny inputs
 *   pbbo_direct                                   brute force (numpy)
 *   pbbo_matvec with a target mask                equal, on the masked
 *                                                 targets, to the full
 *                                                 pbbo_matvec (tests)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------
 * Kernels.  Ids match the product's include/pbbfmm3d_amd.h.
 * kernel_Types.cpp:15-28 (1/r), :56-74 (exp(-r^2)), :77-178 (others);
 * parameterised variants are the user kernels that README.md:174-230
 * describes.  The distance uses diff = source - target as the reference.
 * ---------------------------------------------------------------------- */
enum {
    PBBO_K_LAPLACIAN = 1,      /* 1/r                                    */
    PBBO_K_ONEOVERR2 = 2,      /* 1/r^2                                  */
    PBBO_K_GAUSSIAN = 3,       /* exp(-r^2 / p0^2)                       */
    PBBO_K_EXPONENTIAL = 4,    /* p1 * exp(-r / p0)                      */
    PBBO_K_SOFT_LAPLACIAN = 5, /* 1/sqrt(r^2 + p0)                       */
    PBBO_K_QUADRIC = 6,        /* 1 + r^2                                */
    PBBO_K_INV_QUADRIC = 7,    /* 1 / (1 + r^2)                          */
    PBBO_K_LOGARITHM = 8,      /* 0.5 log r^2, 0 at r = 0                */
    PBBO_K_THIN_PLATE = 9,     /* 0.5 r^2 log r^2, 0 at r = 0            */
    PBBO_K_R2 = 10,            /* r^2 (homogeneous of degree 2)          */
    PBBO_K_CONSTANT = 11,      /* 1                                      */
    PBBO_K_POLY_XY = 12,       /* (tx - sx) * (ty + sy) + p0 (not symm.) */
    PBBO_K_EXP_ANISO = 13      /* p0 exp(-sqrt(p1 dx^2 + p2 dy^2 + p3 dz^2)) */
};

double pbbo_kernel_eval(int id, const double *p, double tx, double ty, double tz,
                        double sx, double sy, double sz)
{
    double dx = sx - tx, dy = sy - ty, dz = sz - tz;
    double r2 = dx * dx + dy * dy + dz * dz;
    switch (id) {
    case PBBO_K_LAPLACIAN: return 1. / sqrt(r2);
    case PBBO_K_ONEOVERR2: { double ri = 1. / sqrt(r2); return ri * ri; }
    case PBBO_K_GAUSSIAN: return exp(-r2 / (p[0] * p[0]));
    case PBBO_K_EXPONENTIAL: return p[1] * exp(-sqrt(r2) / p[0]);
    case PBBO_K_SOFT_LAPLACIAN: return 1. / sqrt(r2 + p[0]);
    case PBBO_K_QUADRIC: return 1.0 + r2;
    case PBBO_K_INV_QUADRIC: return 1.0 / (1.0 + r2);
    case PBBO_K_LOGARITHM: return r2 == 0 ? 0. : 0.5 * log(r2);
    case PBBO_K_THIN_PLATE: return r2 == 0 ? 0. : 0.5 * r2 * log(r2);
    case PBBO_K_R2: return r2;
    case PBBO_K_CONSTANT: return 1.0;
    case PBBO_K_POLY_XY: return (tx - sx) * (ty + sy) + p[0];
    case PBBO_K_EXP_ANISO: return p[0] * exp(-sqrt(p[1] * dx * dx + p[2] * dy * dy + p[3] * dz * dz));
    default: return NAN;
    }
}

/* isinf -> 0 as compute.hpp:767-768 and test.hpp:29-30 */
static inline double kernel_p2p(int id, const double *p, double tx, double ty,
                                double tz, double sx, double sy, double sz)
{
    double v = pbbo_kernel_eval(id, p, tx, ty, tz, sx, sy, sz);
    return isinf(v) ? 0. : v;
}

/* ------------------------------------------------------------------------
 * Closed-form Chebyshev pieces.
 * ---------------------------------------------------------------------- */

/* H2_3D_Tree.cpp:657-666 (Chebyshev branch, alpha = 0, len = 1) */
void pbbo_cheb_nodes(int n, double *nodes)
{
    for (int m = 0; m < n; m++)
        nodes[m] = cos(M_PI * ((double)m + 0.5) / (double)n);
}

/* H2_3D_Tree.cpp:751-757 + ComputeTk :901-909: Tkz[m*n+k] = T_k(node_m) */
void pbbo_tkz(int n, double *Tkz)
{
    double *nodes = malloc(sizeof(double) * n);
    pbbo_cheb_nodes(n, nodes);
    for (int m = 0; m < n; m++) {
        double x = nodes[m];
        double *v = Tkz + m * n;
        v[0] = 1;
        if (n > 1) v[1] = x;
        for (int k = 2; k < n; k++) v[k] = 2.0 * x * v[k - 1] - v[k - 2];
    }
    free(nodes);
}

/* H2_3D_Tree.cpp:759-772: Kweights[(l1 n + l2) n + l3] = prod (1-x^2)^-1/2 */
void pbbo_kweights(int n, double *Kw)
{
    double *nodes = malloc(sizeof(double) * n);
    pbbo_cheb_nodes(n, nodes);
    int c = 0;
    for (int l1 = 0; l1 < n; l1++) {
        double t1 = 1. / sqrt(1 - nodes[l1] * nodes[l1]);
        for (int l2 = 0; l2 < n; l2++) {
            double t2 = t1 / sqrt(1 - nodes[l2] * nodes[l2]);
            for (int l3 = 0; l3 < n; l3++)
                Kw[c++] = t2 / sqrt(1 - nodes[l3] * nodes[l3]);
        }
    }
    free(nodes);
}

/* ClenshawSum, H2_3D_Tree.cpp:869-879 */
static double clenshaw(int n, double x, const double *Tk)
{
    double d0 = 0, d1 = 0, d2;
    for (int j = n - 1; j > 0; j--) {
        d2 = d0;
        d0 = 2.0 * x * d0 - d1 + Tk[j];
        d1 = d2;
    }
    return x * d0 - d1 + 0.5 * Tk[0];
}

/* ComputeSn (Chebyshev branch), H2_3D_Tree.cpp:830-851, one coordinate:
 * S[m] = (2/n) * Clenshaw(x; T_k(node_m)) */
void pbbo_sn(int n, const double *Tkz, double x, double *S)
{
    double pfac = 2. / n;
    for (int m = 0; m < n; m++) S[m] = pfac * clenshaw(n, x, Tkz + m * n);
}

/* H2_3D_Tree.cpp:774-819 (alpha = 0):
 * Cw[m*n + j]       = S_m(node_j/2 - 1/2)   ("-" child)
 * Cw[n*n + m*n + j] = S_m(node_j/2 + 1/2)   ("+" child) */
void pbbo_cweights(int n, double *Cw)
{
    double *Tkz = malloc(sizeof(double) * n * n);
    double *nodes = malloc(sizeof(double) * n);
    double *S = malloc(sizeof(double) * n);
    pbbo_tkz(n, Tkz);
    pbbo_cheb_nodes(n, nodes);
    for (int half = 0; half < 2; half++) {
        double shift = half == 0 ? -0.5 : 0.5;
        for (int j = 0; j < n; j++) {
            pbbo_sn(n, Tkz, nodes[j] * .5 + shift, S);
            for (int m = 0; m < n; m++) Cw[half * n * n + m * n + j] = S[m];
        }
    }
    free(Tkz); free(nodes); free(S);
}

/* H2_3D_Tree.cpp:731-747: Ktable over (l1,l2,l3) in [-3,3]^3 lexicographic */
void pbbo_ktable(int *Ktable)
{
    int ncell = 0, ninteract = 0;
    for (int l1 = -3; l1 < 4; l1++)
        for (int l2 = -3; l2 < 4; l2++)
            for (int l3 = -3; l3 < 4; l3++) {
                Ktable[ncell] = -1;
                if (abs(l1) > 1 || abs(l2) > 1 || abs(l3) > 1)
                    Ktable[ncell] = ninteract++;
                ncell++;
            }
}

/* ------------------------------------------------------------------------
 * M2L kernel blocks, H2_3D_Tree.cpp:277-374.
 * Kcell[d*n6 + i + j*n3] = K(t_i, s_j^(d)) / (Kw_i Kw_j), t_i = node_i*L/2,
 * s_j^(d) = (k + node_j/2)*L, d the lexicographic far-direction index.
 * For symm != 0 only d < 158 is evaluated and Kcell[d] = symm*Kcell[315-d]^T.
 * ---------------------------------------------------------------------- */
void pbbo_kernel_blocks(int id, const double *p, int symm, int n, double boxLen,
                        const double *Kw, double *Kcell)
{
    int n3 = n * n * n;
    size_t n6 = (size_t)n3 * n3;
    double *nodes = malloc(sizeof(double) * n);
    pbbo_cheb_nodes(n, nodes);
    int symmNum = 158 * (2 - abs(symm));
    int dirs[316][3];
    int c = 0;
    for (int k1 = -3; k1 < 4; k1++)
        for (int k2 = -3; k2 < 4; k2++)
            for (int k3 = -3; k3 < 4; k3++)
                if (abs(k1) > 1 || abs(k2) > 1 || abs(k3) > 1) {
                    dirs[c][0] = k1; dirs[c][1] = k2; dirs[c][2] = k3; c++;
                }
#pragma omp parallel for schedule(dynamic)
    for (int d = 0; d < symmNum; d++) {
        double *K = Kcell + (size_t)d * n6;
        for (int j = 0; j < n3; j++) {
            int j1 = j / (n * n), j2 = (j / n) % n, j3 = j % n;
            double sx = (dirs[d][0] + 0.5 * nodes[j1]) * boxLen;
            double sy = (dirs[d][1] + 0.5 * nodes[j2]) * boxLen;
            double sz = (dirs[d][2] + 0.5 * nodes[j3]) * boxLen;
            for (int i = 0; i < n3; i++) {
                int i1 = i / (n * n), i2 = (i / n) % n, i3 = i % n;
                double tx = 0.5 * nodes[i1] * boxLen;
                double ty = 0.5 * nodes[i2] * boxLen;
                double tz = 0.5 * nodes[i3] * boxLen;
                K[i + (size_t)j * n3] =
                    pbbo_kernel_eval(id, p, tx, ty, tz, sx, sy, sz) / (Kw[i] * Kw[j]);
            }
        }
    }
    if (symm) {
        for (int d = symmNum; d < 316; d++) {
            double *K = Kcell + (size_t)d * n6;
            const double *Ks = Kcell + (size_t)(315 - d) * n6;
            for (int j = 0; j < n3; j++)
                for (int i = 0; i < n3; i++)
                    K[i + (size_t)j * n3] = symm * Ks[j + (size_t)i * n3];
        }
    }
    free(nodes);
}

/* Cutoff rule, H2_3D_Tree.cpp:406-421 / :474-491: drop trailing singular
 * values while their accumulated energy stays below eps^2 * total. */
int pbbo_cutoff(const double *sigma, int m, double eps)
{
    double sumsqr = 0;
    for (int i = m - 1; i >= 0; --i) sumsqr += sigma[i] * sigma[i];
    double sum = 0, epsqr = sumsqr * eps * eps;
    int cut = m;
    for (int i = m - 1; i >= 0; --i) {
        sum += sigma[i] * sigma[i];
        if (sum < epsqr) --cut; else break;
    }
    return cut;
}

/* H2_3D_Tree.cpp:523-556: E[d] = U^T Kcell[d] V, U n3 x rf (col-major, first
 * rf columns of U0), VT rs x n3 (col-major), E[d] rf x rs col-major. */
void pbbo_compress(int n3, int rf, int rs, const double *U, const double *VT,
                   const double *Kcell, double *E)
{
#pragma omp parallel for schedule(dynamic)
    for (int d = 0; d < 316; d++) {
        const double *K = Kcell + (size_t)d * n3 * n3;
        double *KV = malloc(sizeof(double) * n3 * rs);
        for (int b = 0; b < rs; b++)
            for (int i = 0; i < n3; i++) {
                double s = 0;
                for (int j = 0; j < n3; j++) s += K[i + (size_t)j * n3] * VT[b + (size_t)j * rs];
                KV[i + (size_t)b * n3] = s;
            }
        double *Ed = E + (size_t)d * rf * rs;
        for (int b = 0; b < rs; b++)
            for (int a = 0; a < rf; a++) {
                double s = 0;
                for (int i = 0; i < n3; i++) s += U[i + (size_t)a * n3] * KV[i + (size_t)b * n3];
                Ed[a + (size_t)b * rf] = s;
            }
        free(KV);
    }
}

/* ------------------------------------------------------------------------
 * Direct sum, test.hpp:18-34 (isinf -> 0), generalised to Nt targets.
 * out[c*Nt + i] = sum_j K(t_i, s_j) w[c*Ns + j]   (overwrites out)
 * ---------------------------------------------------------------------- */
void pbbo_direct(int id, const double *p, const double *tgt, int64_t Nt,
                 const double *src, int64_t Ns, const double *w, int nCols,
                 double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Nt; i++) {
        double acc[64];
        for (int c0 = 0; c0 < nCols; c0 += 64) {
            int cn = nCols - c0 < 64 ? nCols - c0 : 64;
            for (int c = 0; c < cn; c++) acc[c] = 0;
            for (int64_t j = 0; j < Ns; j++) {
                double k = kernel_p2p(id, p, tgt[3 * i], tgt[3 * i + 1], tgt[3 * i + 2],
                                      src[3 * j], src[3 * j + 1], src[3 * j + 2]);
                for (int c = 0; c < cn; c++) acc[c] += k * w[(int64_t)(c0 + c) * Ns + j];
            }
            for (int c = 0; c < cn; c++) out[(int64_t)(c0 + c) * Nt + i] = acc[c];
        }
    }
}

/* ------------------------------------------------------------------------
 * The matvec.
 * ---------------------------------------------------------------------- */
typedef struct {
    int n;            /* interpolation order                              */
    int rf, rs;       /* cutoff.f, cutoff.s                               */
    int nops;         /* operator sets: 1 (homogeneous) or levels-1       */
    int homog;        /* 1 if |homogen| > 0.1 (H2_3D_Tree.cpp:134-136)    */
    double homogen;   /* the flipped value (H2_3D_Tree.cpp:45)            */
    double L;
    int levels;
    int kid;          /* P2P kernel id + params                           */
    double kp[4];
    const double *U;  /* nops x (n3 x rf)       col-major                 */
    const double *VT; /* nops x (rs x n3)       col-major                 */
    const double *E;  /* nops x 316 x (rf x rs) col-major                 */
} pbbo_ops;

typedef struct {
    double cx, cy, cz, len;
    int Nf;                  /* == Ns (sources are targets)             */
    int cid;                 /* compact id among non-empty, -1 if empty */
} onode;

typedef struct {
    int neighbors[27];       /* node index at the same level or -1      */
    int ninter;
    int inter[189];
    double *S;               /* sourceval  n3 x nCols                   */
    double *P;               /* proxysval  rs x nCols                   */
    double *F;               /* targetval  n3 x nCols                   */
    int *list;               /* leaves: point indices                   */
    double *phi;             /* leaves: nodePhi Nf x nCols              */
} onode_data;

typedef struct {
    int levels;
    onode *nodes[16];        /* dense per level, index = Morton code    */
    onode_data *data[16];    /* compact per level                       */
    int ncompact[16];
} otree;

static inline int64_t pow8(int l) { return (int64_t)1 << (3 * l); }

/* BuildFMMHierarchy + NewNode centres, H2_3D_Tree.cpp:917-1019: child i
 * gets x from i>=4, y from (i mod 4)>=2, z from i odd, each +- L/4 of the
 * parent centre, length L/2.  Index of child i of p is 8p+i (leafIndex). */
static void otree_geometry(otree *T, double L)
{
    T->nodes[0] = calloc(1, sizeof(onode));
    T->nodes[0]->cx = T->nodes[0]->cy = T->nodes[0]->cz = 0;
    T->nodes[0]->len = L;
    for (int l = 0; l < T->levels; l++) {
        int64_t np = pow8(l);
        T->nodes[l + 1] = calloc((size_t)pow8(l + 1), sizeof(onode));
        for (int64_t p = 0; p < np; p++) {
            onode *A = &T->nodes[l][p];
            double q = 0.25 * A->len, h = 0.5 * A->len;
            double lx = A->cx - q, ly = A->cy - q, lz = A->cz - q;
            double rx = A->cx + q, ry = A->cy + q, rz = A->cz + q;
            for (int i = 0; i < 8; i++) {
                onode *C = &T->nodes[l + 1][8 * p + i];
                C->cx = i < 4 ? lx : rx;
                C->cy = (i % 4) < 2 ? ly : ry;
                C->cz = (i % 2) == 0 ? lz : rz;
                C->len = h;
            }
        }
    }
}

/* DistributeFieldPoints, compute.hpp:1126-1203: octant by strict '<'
 * against the node centre; leaf lists keep the input order. */
static int otree_distribute(otree *T, const double *x, int N)
{
    int Lv = T->levels;
    int *leaf = malloc(sizeof(int) * N);
    for (int i = 0; i < N; i++) {
        int64_t node = 0;
        for (int l = 0; l < Lv; l++) {
            onode *A = &T->nodes[l][node];
            int j = (x[3 * i] < A->cx ? 0 : 4) + (x[3 * i + 1] < A->cy ? 0 : 2) +
                    (x[3 * i + 2] < A->cz ? 0 : 1);
            node = 8 * node + j;
        }
        leaf[i] = (int)node;
        T->nodes[Lv][node].Nf++;
    }
    for (int l = Lv - 1; l >= 0; l--)
        for (int64_t p = 0; p < pow8(l); p++) {
            int s = 0;
            for (int i = 0; i < 8; i++) s += T->nodes[l + 1][8 * p + i].Nf;
            T->nodes[l][p].Nf = s;
        }
    for (int l = 0; l <= Lv; l++) {
        int c = 0;
        for (int64_t p = 0; p < pow8(l); p++)
            T->nodes[l][p].cid = T->nodes[l][p].Nf > 0 ? c++ : -1;
        T->ncompact[l] = c;
        T->data[l] = calloc(c > 0 ? c : 1, sizeof(onode_data));
        for (int k = 0; k < c; k++)
            for (int m = 0; m < 27; m++) T->data[l][k].neighbors[m] = -1;
    }
    /* leaf lists in increasing input index (stable distribution) */
    int *fill = calloc(T->ncompact[Lv] + 1, sizeof(int));
    for (int64_t p = 0; p < pow8(Lv); p++) {
        onode *A = &T->nodes[Lv][p];
        if (A->cid >= 0) T->data[Lv][A->cid].list = malloc(sizeof(int) * A->Nf);
    }
    for (int i = 0; i < N; i++) {
        int cid = T->nodes[Lv][leaf[i]].cid;
        T->data[Lv][cid].list[fill[cid]++] = i;
    }
    free(fill);
    free(leaf);
    return 0;
}

static inline onode_data *ND(otree *T, int l, int64_t p)
{
    int c = T->nodes[l][p].cid;
    return c >= 0 ? &T->data[l][c] : NULL;
}

/* InteractionList, compute.hpp:1042-1118 (root: neighbors[0] = root,
 * compute.hpp:176).  Children of the parent's neighbours are near when
 * max |round(diff / childlen)| <= 1, else far. */
static void otree_interactions(otree *T)
{
    ND(T, 0, 0)->neighbors[0] = 0;
    for (int l = 0; l < T->levels; l++) {
        for (int64_t a = 0; a < pow8(l); a++) {
            onode *A = &T->nodes[l][a];
            if (A->Nf == 0) continue;
            onode_data *Ad = ND(T, l, a);
            double cutoff = A->len / 2;
            for (int i = 0; i < 27; i++) {
                int b = Ad->neighbors[i];
                if (b < 0) continue;
                for (int j = 0; j < 8; j++) {
                    int64_t bc = 8 * (int64_t)b + j;
                    onode *Bc = &T->nodes[l + 1][bc];
                    if (Bc->Nf == 0) continue;
                    for (int k = 0; k < 8; k++) {
                        int64_t cc = 8 * a + k;
                        onode *C = &T->nodes[l + 1][cc];
                        if (C->Nf == 0) continue;
                        onode_data *Cd = ND(T, l + 1, cc);
                        int x = (int)round((Bc->cx - C->cx) / cutoff);
                        int y = (int)round((Bc->cy - C->cy) / cutoff);
                        int z = (int)round((Bc->cz - C->cz) / cutoff);
                        if (!(abs(x) > 1 || abs(y) > 1 || abs(z) > 1))
                            Cd->neighbors[9 * (x + 1) + 3 * (y + 1) + (z + 1)] = (int)bc;
                        else
                            Cd->inter[Cd->ninter++] = (int)bc;
                    }
                }
            }
        }
    }
}

/* Moment2Moment, compute.hpp:956-1034: z, then y, then x mode products
 * with the Cweights block chosen by the child octant.  SS is overwritten. */
static void m2m(int n, int nCols, int octant, const double *Cw, const double *Sc,
                double *SS, double *Sz, double *Sy)
{
    int n2 = n * n, n3 = n2 * n;
    const double *Cx = Cw + ((octant >> 2) & 1) * n2;
    const double *Cy = Cw + ((octant >> 1) & 1) * n2;
    const double *Cz = Cw + (octant & 1) * n2;
    for (int c = 0; c < nCols; c++) {
        const double *S = Sc + (size_t)c * n3;
        double *z = Sz + (size_t)c * n3, *y = Sy + (size_t)c * n3, *o = SS + (size_t)c * n3;
        for (int ab = 0; ab < n2; ab++)
            for (int l3 = 0; l3 < n; l3++) {
                double s = 0;
                for (int k = 0; k < n; k++) s += S[ab * n + k] * Cz[l3 * n + k];
                z[ab * n + l3] = s;
            }
        for (int l1 = 0; l1 < n; l1++)
            for (int l2 = 0; l2 < n; l2++)
                for (int l3 = 0; l3 < n; l3++) {
                    double s = 0;
                    for (int k = 0; k < n; k++) s += z[l1 * n2 + k * n + l3] * Cy[l2 * n + k];
                    y[l1 * n2 + l2 * n + l3] = s;
                }
        for (int l1 = 0; l1 < n; l1++)
            for (int l2 = 0; l2 < n; l2++)
                for (int l3 = 0; l3 < n; l3++) {
                    double s = 0;
                    for (int k = 0; k < n; k++) s += y[k * n2 + l2 * n + l3] * Cx[l1 * n + k];
                    o[l1 * n2 + l2 * n + l3] = s;
                }
    }
}

/* Local2Local, compute.hpp:798-878: the transpose of m2m, accumulated. */
static void l2l(int n, int nCols, int octant, const double *Cw, const double *F,
                double *Fchild, double *Fx, double *Fy)
{
    int n2 = n * n, n3 = n2 * n;
    const double *Cx = Cw + ((octant >> 2) & 1) * n2;
    const double *Cy = Cw + ((octant >> 1) & 1) * n2;
    const double *Cz = Cw + (octant & 1) * n2;
    for (int c = 0; c < nCols; c++) {
        const double *f = F + (size_t)c * n3;
        double *x = Fx + (size_t)c * n3, *y = Fy + (size_t)c * n3, *o = Fchild + (size_t)c * n3;
        for (int l1 = 0; l1 < n; l1++)
            for (int l2 = 0; l2 < n; l2++)
                for (int l3 = 0; l3 < n; l3++) {
                    double s = 0;
                    for (int k = 0; k < n; k++) s += f[k * n2 + l2 * n + l3] * Cx[k * n + l1];
                    x[l1 * n2 + l2 * n + l3] = s;
                }
        for (int l1 = 0; l1 < n; l1++)
            for (int l2 = 0; l2 < n; l2++)
                for (int l3 = 0; l3 < n; l3++) {
                    double s = 0;
                    for (int k = 0; k < n; k++) s += x[l1 * n2 + k * n + l3] * Cy[k * n + l2];
                    y[l1 * n2 + l2 * n + l3] = s;
                }
        for (int ab = 0; ab < n2; ab++)
            for (int l3 = 0; l3 < n; l3++) {
                double s = 0;
                for (int k = 0; k < n; k++) s += y[ab * n + k] * Cz[k * n + l3];
                o[ab * n + l3] += s;
            }
    }
}

/* ops level index: lvl_shift = max(l-2,0) for non-homogeneous kernels,
 * compute.hpp:387-389, :419-420 */
static inline int ops_level(const pbbo_ops *o, int l)
{
    if (o->homog) return 0;
    return l >= 2 ? l - 2 : 0;
}

/*
 * pbbo_matvec: H2_3D_Compute<T>::H2_3D_Compute, compute.hpp:67-149, with
 * FMMDistribute (:157-190), UpwardPass (:224-397), FarFieldInteractions
 * (:401-549), DownwardPass (:551-587), NearFieldInteractions (:590-751).
 * Sources are the targets (one point array, compute.hpp:1195-1196).
 * xyz: AoS N x 3; w: column-major N x nCols; out += (compute.hpp:747).
 * Levels 0 and 1 carry no far field (every level-1 box is a neighbour of
 * every other), so their expansions are not formed.
 *
 * tmask (optional, N bytes): when given, only the outputs of the points with
 * tmask[i] != 0 are formed (other entries of out are left as they are).  The
 * upward pass and the proxies stay complete; the far field and the downward
 * pass run only for the boxes holding a masked point and their ancestors,
 * and those leaves take their near field from all 27 neighbours directly
 * (K(x_i, x_j) w_j for every neighbour point j) instead of sharing each
 * pair with the partner leaf as compute.hpp:680-725 does.  Same terms, so
 * the masked outputs equal the full ones to rounding; it only makes the
 * oracle affordable at a million points.
 */
int pbbo_matvec(const pbbo_ops *o, const double *xyz_in, int N, const double *w,
                int nCols, double *out, const unsigned char *tmask, int nbox)
{
    int n = o->n, n2 = n * n, n3 = n2 * n, Lv = o->levels;
    int rf = o->rf, rs = o->rs;
    if (Lv < 1 || Lv > 10 || N <= 0 || nCols <= 0 || nbox < 0 || nbox > N) return -1;

    /* bbox of the targets and re-centring, compute.hpp:92-127: the bbox is
     * that of the targets only (compute.hpp:97-104) and the sources take
     * the same shift (:121-126).  nbox > 0: the targets are the first nbox
     * points (the rest are sources placed in the same tree); nbox = 0: all
     * N points are the targets. */
    if (nbox == 0) nbox = N;
    double mn[3] = {1e32, 1e32, 1e32}, mx[3] = {-1e32, -1e32, -1e32};
    for (int i = 0; i < nbox; i++)
        for (int d = 0; d < 3; d++) {
            double v = xyz_in[3 * i + d];
            mn[d] = v < mn[d] ? v : mn[d];
            mx[d] = v > mx[d] ? v : mx[d];
        }
    double ctr[3];
    for (int d = 0; d < 3; d++) ctr[d] = 0.5 * (mn[d] + mx[d]);
    double *x = malloc(sizeof(double) * 3 * (size_t)N);
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) x[3 * i + d] = xyz_in[3 * i + d] - ctr[d];

    double Tkz[256], Kw[4096], Cw[512];
    pbbo_tkz(n, Tkz);
    pbbo_kweights(n, Kw);
    pbbo_cweights(n, Cw);
    int Ktable[343];
    pbbo_ktable(Ktable);

    otree T;
    memset(&T, 0, sizeof(T));
    T.levels = Lv;
    otree_geometry(&T, o->L);
    otree_distribute(&T, x, N);
    otree_interactions(&T);

    /* boxes whose outputs are formed: leaves holding a masked point and
     * their ancestors (every box when there is no mask) */
    unsigned char *mark[16] = {0};
    if (tmask) {
        for (int l = 0; l <= Lv; l++) mark[l] = calloc((size_t)pow8(l), 1);
        for (int64_t p = 0; p < pow8(Lv); p++) {
            onode_data *Ad = ND(&T, Lv, p);
            if (!Ad) continue;
            for (int j = 0; j < T.nodes[Lv][p].Nf; j++)
                if (tmask[Ad->list[j]]) mark[Lv][p] = 1;
        }
        for (int l = Lv - 1; l >= 0; l--)
            for (int64_t p = 0; p < pow8(l); p++)
                for (int i = 0; i < 8; i++)
                    if (mark[l + 1][8 * p + i]) mark[l][p] = 1;
    }
#define MARKED(l, p) (!tmask || mark[l][p])

    /* ---- upward pass: P2M at leaves (compute.hpp:287-333) ---- */
    int64_t nleafall = pow8(Lv);
#pragma omp parallel for schedule(dynamic)
    for (int64_t p = 0; p < nleafall; p++) {
        onode *A = &T.nodes[Lv][p];
        if (A->Nf == 0) continue;
        onode_data *Ad = ND(&T, Lv, p);
        int Ns = A->Nf;
        double ihalfL = 2. / A->len;
        double *Ss = malloc(sizeof(double) * 3 * n * Ns);
        for (int j = 0; j < Ns; j++) {
            int k = Ad->list[j];
            double sx = ihalfL * (x[3 * k] - A->cx);
            double sy = ihalfL * (x[3 * k + 1] - A->cy);
            double sz = ihalfL * (x[3 * k + 2] - A->cz);
            double t[16];
            pbbo_sn(n, Tkz, sx, t);
            for (int m = 0; m < n; m++) Ss[(m * Ns + j) * 3 + 0] = t[m];
            pbbo_sn(n, Tkz, sy, t);
            for (int m = 0; m < n; m++) Ss[(m * Ns + j) * 3 + 1] = t[m];
            pbbo_sn(n, Tkz, sz, t);
            for (int m = 0; m < n; m++) Ss[(m * Ns + j) * 3 + 2] = t[m];
        }
        Ad->S = malloc(sizeof(double) * n3 * nCols);
        for (int col = 0; col < nCols; col++) {
            int l = 0;
            for (int l1 = 0; l1 < n; l1++)
                for (int l2 = 0; l2 < n; l2++)
                    for (int l3 = 0; l3 < n; l3++) {
                        double sum = 0;
                        for (int j = 0; j < Ns; j++) {
                            int k = Ad->list[j];
                            sum += w[(int64_t)col * N + k] * Ss[(l1 * Ns + j) * 3] *
                                   Ss[(l2 * Ns + j) * 3 + 1] * Ss[(l3 * Ns + j) * 3 + 2];
                        }
                        Ad->S[n3 * col + l++] = sum;
                    }
        }
        free(Ss);
    }

    /* ---- M2M, children in octant order (compute.hpp:234-283) ---- */
    for (int l = Lv - 1; l >= 2; l--) {
        int64_t np = pow8(l);
#pragma omp parallel for schedule(dynamic)
        for (int64_t p = 0; p < np; p++) {
            onode *A = &T.nodes[l][p];
            if (A->Nf == 0) continue;
            onode_data *Ad = ND(&T, l, p);
            Ad->S = calloc((size_t)n3 * nCols, sizeof(double));
            double *SS = malloc(sizeof(double) * n3 * nCols * 3);
            for (int i = 0; i < 8; i++) {
                onode_data *Cd = ND(&T, l + 1, 8 * p + i);
                if (!Cd) continue;
                m2m(n, nCols, i, Cw, Cd->S, SS, SS + n3 * nCols, SS + 2 * n3 * nCols);
                for (int q = 0; q < n3 * nCols; q++) Ad->S[q] += SS[q];
            }
            free(SS);
        }
    }

    /* ---- proxies P = VT (Kweights o S), compute.hpp:373-396 ---- */
    for (int l = 2; l <= Lv; l++) {
        int ol = ops_level(o, l);
        const double *VT = o->VT + (size_t)ol * rs * n3;
#pragma omp parallel for schedule(dynamic)
        for (int k = 0; k < T.ncompact[l]; k++) {
            onode_data *Ad = &T.data[l][k];
            Ad->P = malloc(sizeof(double) * rs * nCols);
            double *Sw = malloc(sizeof(double) * n3);
            for (int col = 0; col < nCols; col++) {
                for (int q = 0; q < n3; q++) Sw[q] = Ad->S[col * n3 + q] * Kw[q];
                for (int a = 0; a < rs; a++) {
                    double s = 0;
                    for (int q = 0; q < n3; q++) s += VT[a + (size_t)q * rs] * Sw[q];
                    Ad->P[col * rs + a] = s;
                }
            }
            free(Sw);
        }
    }

    /* ---- far field: Pf = sum_B E_d P_B; F = Kw o (scale U Pf),
     *      compute.hpp:401-549, Moment2Local :880-923 ---- */
    for (int l = 2; l <= Lv; l++) {
        int ol = ops_level(o, l);
        const double *U = o->U + (size_t)ol * n3 * rf;
        const double *E = o->E + (size_t)ol * 316 * rf * rs;
        int64_t np = pow8(l);
#pragma omp parallel for schedule(dynamic)
        for (int64_t p = 0; p < np; p++) {
            onode *A = &T.nodes[l][p];
            if (A->Nf == 0 || !MARKED(l, p)) continue;
            onode_data *Ad = ND(&T, l, p);
            double iL = 1.0 / A->len;
            double scale = pow(iL, o->homogen);
            double *Pf = calloc((size_t)rf * nCols, sizeof(double));
            for (int i = 0; i < Ad->ninter; i++) {
                int64_t b = Ad->inter[i];
                onode *B = &T.nodes[l][b];
                onode_data *Bd = ND(&T, l, b);
                int k1 = (int)round(iL * (B->cx - A->cx)) + 3;
                int k2 = (int)round(iL * (B->cy - A->cy)) + 3;
                int k3 = (int)round(iL * (B->cz - A->cz)) + 3;
                int d = Ktable[49 * k1 + 7 * k2 + k3];
                const double *Ed = E + (size_t)d * rf * rs;
                for (int col = 0; col < nCols; col++) {
                    const double *pb = Bd->P + (size_t)col * rs;
                    double *pf = Pf + (size_t)col * rf;
                    for (int bb = 0; bb < rs; bb++) {
                        double v = pb[bb];
                        const double *e = Ed + (size_t)bb * rf;
                        for (int a = 0; a < rf; a++) pf[a] += e[a] * v;
                    }
                }
            }
            Ad->F = malloc(sizeof(double) * n3 * nCols);
            for (int col = 0; col < nCols; col++)
                for (int q = 0; q < n3; q++) {
                    double s = 0;
                    for (int a = 0; a < rf; a++) s += U[q + (size_t)a * n3] * Pf[col * rf + a];
                    Ad->F[col * n3 + q] = scale * s * Kw[q];
                }
            free(Pf);
        }
    }

    /* ---- downward pass, compute.hpp:551-587 ---- */
    for (int l = 2; l < Lv; l++) {
        int64_t np = pow8(l);
#pragma omp parallel for schedule(dynamic)
        for (int64_t p = 0; p < np; p++) {
            onode *A = &T.nodes[l][p];
            if (A->Nf == 0 || !MARKED(l, p)) continue;
            onode_data *Ad = ND(&T, l, p);
            double *tmp = malloc(sizeof(double) * n3 * nCols * 2);
            for (int i = 0; i < 8; i++) {
                onode_data *Cd = ND(&T, l + 1, 8 * p + i);
                if (!Cd || !MARKED(l + 1, 8 * p + i)) continue;
                l2l(n, nCols, i, Cw, Ad->F, Cd->F, tmp, tmp + n3 * nCols);
            }
            free(tmp);
        }
    }

    /* ---- near field loop 1: L2P, compute.hpp:594-648 ---- */
#pragma omp parallel for schedule(dynamic)
    for (int64_t p = 0; p < nleafall; p++) {
        onode *A = &T.nodes[Lv][p];
        if (A->Nf == 0 || !MARKED(Lv, p)) continue;
        onode_data *Ad = ND(&T, Lv, p);
        int Nf = A->Nf;
        Ad->phi = calloc((size_t)Nf * nCols, sizeof(double));
        if (Lv < 2) continue; /* no far field at all */
        double ihalfL = 2. / A->len;
        double *Sf = malloc(sizeof(double) * 3 * n * Nf);
        for (int i = 0; i < Nf; i++) {
            int k = Ad->list[i];
            double t[16];
            pbbo_sn(n, Tkz, ihalfL * (x[3 * k] - A->cx), t);
            for (int m = 0; m < n; m++) Sf[(m * Nf + i) * 3] = t[m];
            pbbo_sn(n, Tkz, ihalfL * (x[3 * k + 1] - A->cy), t);
            for (int m = 0; m < n; m++) Sf[(m * Nf + i) * 3 + 1] = t[m];
            pbbo_sn(n, Tkz, ihalfL * (x[3 * k + 2] - A->cz), t);
            for (int m = 0; m < n; m++) Sf[(m * Nf + i) * 3 + 2] = t[m];
        }
        for (int col = 0; col < nCols; col++)
            for (int i = 0; i < Nf; i++) {
                int l = 0;
                double sum = 0;
                for (int l1 = 0; l1 < n; l1++) {
                    double t1 = Sf[(l1 * Nf + i) * 3];
                    for (int l2 = 0; l2 < n; l2++) {
                        double t2 = t1 * Sf[(l2 * Nf + i) * 3 + 1];
                        for (int l3 = 0; l3 < n; l3++)
                            sum += Ad->F[col * n3 + l++] * t2 * Sf[(l3 * Nf + i) * 3 + 2];
                    }
                }
                Ad->phi[col * Nf + i] += sum;
            }
        free(Sf);
    }

    /* ---- near field loop 2: symmetric P2P over the 27 neighbours with
     *      leafIndex_A <= leafIndex_B, compute.hpp:650-737.  The partner's
     *      share is accumulated under a lock as the reference's critical
     *      section does. ---- */
#pragma omp parallel for schedule(dynamic)
    for (int64_t p = 0; p < nleafall; p++) {
        onode *A = &T.nodes[Lv][p];
        if (A->Nf == 0 || !MARKED(Lv, p)) continue;
        onode_data *Ad = ND(&T, Lv, p);
        int Nf = A->Nf;
        double *Aout = calloc((size_t)Nf * nCols, sizeof(double));
        for (int m = 0; m < 27; m++) {
            int b = Ad->neighbors[m];
            if (b < 0) continue;
            onode_data *Bd = ND(&T, Lv, b);
            int Ns = T.nodes[Lv][b].Nf;
            if (m == 13) {
                /* EvaluateField_self, compute.hpp:773-785 */
                for (int j = 0; j < Nf; j++) {
                    int kj = Ad->list[j];
                    for (int i = 0; i <= j; i++) {
                        int ki = Ad->list[i];
                        double kv = kernel_p2p(o->kid, o->kp, x[3 * ki], x[3 * ki + 1], x[3 * ki + 2],
                                               x[3 * kj], x[3 * kj + 1], x[3 * kj + 2]);
                        for (int col = 0; col < nCols; col++) {
                            Aout[col * Nf + i] += kv * w[(int64_t)col * N + kj];
                            if (i != j) Aout[col * Nf + j] += kv * w[(int64_t)col * N + ki];
                        }
                    }
                }
                continue;
            }
            if (tmask) {
                /* masked run: this leaf takes every neighbour's share itself */
                for (int j = 0; j < Ns; j++) {
                    int kj = Bd->list[j];
                    for (int i = 0; i < Nf; i++) {
                        int ki = Ad->list[i];
                        double kv = kernel_p2p(o->kid, o->kp, x[3 * ki], x[3 * ki + 1], x[3 * ki + 2],
                                               x[3 * kj], x[3 * kj + 1], x[3 * kj + 2]);
                        for (int col = 0; col < nCols; col++)
                            Aout[col * Nf + i] += kv * w[(int64_t)col * N + kj];
                    }
                }
                continue;
            }
            if (p > b) continue;
            double *Bout = calloc((size_t)Ns * nCols, sizeof(double));
            for (int j = 0; j < Ns; j++) {
                int kj = Bd->list[j];
                for (int i = 0; i < Nf; i++) {
                    int ki = Ad->list[i];
                    double kv = kernel_p2p(o->kid, o->kp, x[3 * ki], x[3 * ki + 1], x[3 * ki + 2],
                                           x[3 * kj], x[3 * kj + 1], x[3 * kj + 2]);
                    for (int col = 0; col < nCols; col++) {
                        Aout[col * Nf + i] += kv * w[(int64_t)col * N + kj];
                        Bout[col * Ns + j] += kv * w[(int64_t)col * N + ki];
                    }
                }
            }
#pragma omp critical(pbbo_nearfield)
            for (int q = 0; q < Ns * nCols; q++) Bd->phi[q] += Bout[q];
            free(Bout);
        }
#pragma omp critical(pbbo_nearfield)
        for (int q = 0; q < Nf * nCols; q++) Ad->phi[q] += Aout[q];
        free(Aout);
    }

    /* ---- scatter, compute.hpp:739-749 ---- */
    for (int64_t p = 0; p < nleafall; p++) {
        onode *A = &T.nodes[Lv][p];
        if (A->Nf == 0 || !MARKED(Lv, p)) continue;
        onode_data *Ad = ND(&T, Lv, p);
        for (int col = 0; col < nCols; col++)
            for (int i = 0; i < A->Nf; i++)
                out[(int64_t)col * N + Ad->list[i]] += Ad->phi[col * A->Nf + i];
    }

    for (int l = 0; l <= Lv; l++) {
        for (int k = 0; k < T.ncompact[l]; k++) {
            onode_data *d = &T.data[l][k];
            free(d->S); free(d->P); free(d->F); free(d->list); free(d->phi);
        }
        free(T.data[l]);
        free(T.nodes[l]);
        free(mark[l]);
    }
#undef MARKED
    free(x);
    return 0;
}

/* Flat-argument wrapper for ctypes. */
int pbbo_matvec_flat(int n, int rf, int rs, int nops, int homog, double homogen,
                     double L, int levels, int kid, const double *kp,
                     const double *U, const double *VT, const double *E,
                     const double *xyz, int N, const double *w, int nCols,
                     double *out, const unsigned char *tmask, int nbox)
{
    pbbo_ops o;
    o.n = n; o.rf = rf; o.rs = rs; o.nops = nops; o.homog = homog;
    o.homogen = homogen; o.L = L; o.levels = levels; o.kid = kid;
    for (int i = 0; i < 4; i++) o.kp[i] = kp ? kp[i] : 0;
    o.U = U; o.VT = VT; o.E = E;
    return pbbo_matvec(&o, xyz, N, w, nCols, out, tmask, nbox);
}

int pbbo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
