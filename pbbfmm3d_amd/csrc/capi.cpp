// capi.cpp -- the C ABI (include/pbbfmm3d_amd.h): operator objects, plans
// (device-resident trees) and the launch sequence of one matvec.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <functional>
#include <mutex>

#include "host.hpp"
#include "kernels.hpp"
#include "tree_gpu.hpp"

using namespace pbb;

namespace {

thread_local std::string g_last_error;

pbb_status set_error(pbb_status st, const std::string &msg)
{
    g_last_error = msg;
    return st;
}

template <typename F>
pbb_status guarded(F &&f)
{
    try {
        f();
        return PBB_OK;
    } catch (const Error &e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc &) {
        return set_error(PBB_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return set_error(PBB_ERR_INTERNAL, e.what());
    }
}

template <typename T>
T *dalloc(size_t count, const char *what)
{
    if (count == 0) count = 1;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(e == hipErrorOutOfMemory ? PBB_ERR_OUT_OF_MEMORY : PBB_ERR_HIP,
             std::string("hipMalloc(") + what + ", " + std::to_string(count * sizeof(T)) + " bytes): " +
                 hipGetErrorString(e));
    }
    return (T *)p;
}

template <typename T>
T *dupload(const std::vector<T> &v, const char *what)
{
    T *p = dalloc<T>(v.size(), what);
    if (!v.empty()) check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), what);
    return p;
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

int current_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        fail(PBB_ERR_NO_DEVICE, "no HIP device is visible");
    }
    int d = 0;
    check(hipGetDevice(&d), "hipGetDevice");
    return d;
}

// Makes `dev` current for the scope, restoring the caller's device after.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        check(hipGetDevice(&prev), "hipGetDevice");
        if (prev != dev) check(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace

// E'_d of every operator set split into 16-bit parts for one mixed
// variant: set s at element off[s], [316][nch[s]][nsplit][R[s]][32]
struct SplitE {
    void *d = nullptr;
    std::vector<size_t> off;
    std::vector<int> R, nch, escale;  // per set
};

// The operator sets of the mixed modes (reduce_set): set s has kf x ks
// blocks E'_d, VT'w (ks x n3) and U'w (n3 x kf), at the given element
// offsets of the FP64 device arrays.
struct MixedSet {
    int kf = 0, ks = 0;
    int ks_used = 0;  // ks before the padding to 32-deep chunks (rows / columns past it are zero)
    size_t e_off = 0, vt_off = 0, u_off = 0;
    double tail = 0;  // discarded relative Frobenius norm (larger side)
};

struct pbb_ops {
    HostOps h;
    std::mutex mu;
    int device = -1;
    std::unique_ptr<pbb_ops> adj;  // transposed operator set, built on demand
    double *d_Tkz = nullptr, *d_Cw = nullptr, *d_VTw = nullptr, *d_Uw = nullptr, *d_E = nullptr;
    // The mixed modes' operator sets, each re-compressed within level_tol
    // (0: the sets' own factors); set before their first use.  fp64_reduced:
    // the FP64 mode runs them too (off: the reference's operators)
    double level_tol = 2e-8;
    bool fp64_reduced = false;
    std::vector<MixedSet> mset;
    std::vector<double> mset_emax;  // max |E'_d| per set
    std::vector<double> hm_E, hm_VTw, hm_Uw;  // host copies until uploaded
    double *dm_E = nullptr, *dm_VTw = nullptr, *dm_Uw = nullptr;
    // the proxy passes' operators, split into three BF16 parts: per set, VT'
    // then U', each [part][rank < 32][k < px_kp] (row stride px_kp + 8)
    uint16_t *dm_px = nullptr;
    int px_ks = 0;
    size_t px_off(int set, int which) const { return ((size_t)set * 2 + which) * 3 * 32 * px_ks; }
    SplitE split[4];
    // VT'w and U'w of every set split into three BF16 parts for the mixed
    // modes' proxy / expansion products: set s at off[s], [3][R][32*nch[s]]
    struct SplitA {
        void *d = nullptr;
        std::vector<size_t> off;
        std::vector<int> nch;
    } vt3, u3;

    // The mixed operator sets on the host (caller holds mu)
    void compute_mixed_sets()
    {
        if (!mset.empty()) return;
        const int n3 = h.n3, rf = h.rf, rs = h.rs;
        std::vector<std::vector<double>> VTw(h.nsets), Uw(h.nsets);
        for (int s = 0; s < h.nsets; s++) {
            VTw[s].resize((size_t)rs * n3);
            Uw[s].resize((size_t)n3 * rf);
            for (int q = 0; q < n3; q++) {
                for (int a = 0; a < rs; a++)
                    VTw[s][a + (size_t)q * rs] = h.VT[(size_t)s * rs * n3 + a + (size_t)q * rs] * h.w.Kw[q];
                for (int a = 0; a < rf; a++)
                    Uw[s][q + (size_t)a * n3] = h.U[(size_t)s * n3 * rf + q + (size_t)a * n3] * h.w.Kw[q];
            }
        }
        std::vector<ReducedSet> red(h.nsets);
        for (int s = 0; s < h.nsets; s++) {
            if (level_tol > 0) {
                reduce_set(h, s, VTw[s].data(), Uw[s].data(), level_tol, red[s]);
            } else {
                red[s].kf = rf;
                red[s].ks = rs;
                red[s].E.assign(h.E.begin() + (size_t)s * 316 * rf * rs, h.E.begin() + (size_t)(s + 1) * 316 * rf * rs);
                red[s].VTw = VTw[s];
                red[s].Uw = Uw[s];
            }
        }
        // ks padded to whole k chunks with zero columns of E' and zero rows of
        // VT'w, so that the proxy GEMM writes every row the M2L reads
        std::vector<int> ks_used(h.nsets);
        for (int s = 0; s < h.nsets; s++) ks_used[s] = red[s].ks;
        for (auto &r : red) {
            const int ks32 = 32 * ((r.ks + 31) / 32);
            if (ks32 == r.ks) continue;
            std::vector<double> E2((size_t)316 * r.kf * ks32, 0.0), V2((size_t)ks32 * n3, 0.0);
            for (int d = 0; d < 316; d++)
                std::copy(r.E.begin() + (size_t)d * r.kf * r.ks, r.E.begin() + (size_t)(d + 1) * r.kf * r.ks,
                          E2.begin() + (size_t)d * r.kf * ks32);
            for (int q = 0; q < n3; q++)
                for (int a = 0; a < r.ks; a++) V2[a + (size_t)q * ks32] = r.VTw[a + (size_t)q * r.ks];
            r.E.swap(E2);
            r.VTw.swap(V2);
            r.ks = ks32;
        }
        std::vector<double> E, V, U;
        mset.assign(h.nsets, MixedSet{});
        mset_emax.assign(h.nsets, 0.0);
        for (int s = 0; s < h.nsets; s++) {
            MixedSet &m = mset[s];
            m.kf = red[s].kf;
            m.ks = red[s].ks;
            m.ks_used = ks_used[s];
            m.tail = std::max(red[s].tail_f, red[s].tail_s);
            m.e_off = E.size();
            m.vt_off = V.size();
            m.u_off = U.size();
            E.insert(E.end(), red[s].E.begin(), red[s].E.end());
            V.insert(V.end(), red[s].VTw.begin(), red[s].VTw.end());
            U.insert(U.end(), red[s].Uw.begin(), red[s].Uw.end());
            for (double x : red[s].E) mset_emax[s] = std::max(mset_emax[s], std::fabs(x));
        }
        hm_E.swap(E);
        hm_VTw.swap(V);
        hm_Uw.swap(U);
    }

    // ... and on the device (caller holds mu; the operators are uploaded)
    void make_mixed_sets()
    {
        if (dm_E) return;
        compute_mixed_sets();
        dm_E = dupload(hm_E, "mixed E");
        dm_VTw = dupload(hm_VTw, "mixed VT");
        dm_Uw = dupload(hm_Uw, "mixed U");
        {
            // BF16 with round-to-nearest-even, and back
            auto f2bf = [](float f) -> uint16_t {
                uint32_t u;
                std::memcpy(&u, &f, 4);
                if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
                u += 0x7fffu + ((u >> 16) & 1u);
                return (uint16_t)(u >> 16);
            };
            auto bf2f = [](uint16_t b) -> float {
                const uint32_t u = (uint32_t)b << 16;
                float f;
                std::memcpy(&f, &u, 4);
                return f;
            };
            const int n3 = h.n3, kp = 32 * ((n3 + 31) / 32);
            px_ks = kp + 8;
            std::vector<uint16_t> px((size_t)h.nsets * 2 * 3 * 32 * px_ks, 0);
            for (int s = 0; s < h.nsets; s++) {
                const MixedSet &m = mset[s];
                for (int which = 0; which < 2; which++) {
                    const int rank = which == 0 ? m.ks : m.kf;
                    uint16_t *dst = px.data() + px_off(s, which);
                    for (int r = 0; r < 32 && r < rank; r++)
                        for (int k = 0; k < n3; k++) {
                            double x = which == 0 ? hm_VTw[m.vt_off + r + (size_t)k * m.ks]
                                                  : hm_Uw[m.u_off + k + (size_t)r * n3];
                            for (int p = 0; p < 3; p++) {
                                const uint16_t b = f2bf((float)x);
                                dst[((size_t)p * 32 + r) * px_ks + k] = b;
                                x -= (double)bf2f(b);
                            }
                        }
                }
            }
            dm_px = dupload(px, "proxy operators");
        }
        std::vector<double>().swap(hm_E);
        std::vector<double>().swap(hm_VTw);
        std::vector<double>().swap(hm_Uw);
    }

    const SplitA *ensure_split3(hipStream_t st)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (vt3.d) return &vt3;
        make_mixed_sets();
        const int n3 = h.n3, ns = h.nsets;
        // VT'w: M = ks rows, K = n3; U'w: M = n3 rows, K = kf
        size_t nv = 0, nu = 0;
        vt3.off.assign(ns, 0);
        vt3.nch.assign(ns, 0);
        u3.off.assign(ns, 0);
        u3.nch.assign(ns, 0);
        for (int s = 0; s < ns; s++) {
            vt3.off[s] = nv;
            vt3.nch[s] = (n3 + 31) / 32;
            nv += (size_t)3 * (16 * ((mset[s].ks + 15) / 16)) * vt3.nch[s] * 32;
            u3.off[s] = nu;
            u3.nch[s] = (mset[s].kf + 31) / 32;
            nu += (size_t)3 * (16 * ((n3 + 15) / 16)) * u3.nch[s] * 32;
        }
        vt3.d = dalloc<uint16_t>(nv, "split VT");
        u3.d = dalloc<uint16_t>(nu, "split U");
        for (int s = 0; s < ns; s++) {
            const MixedSet &m = mset[s];
            dev::launch_split_a3(dm_VTw + m.vt_off, m.ks, n3, 16 * ((m.ks + 15) / 16), vt3.nch[s] * 32,
                                 (uint16_t *)vt3.d + vt3.off[s], st);
            dev::launch_split_a3(dm_Uw + m.u_off, n3, m.kf, 16 * ((n3 + 15) / 16), u3.nch[s] * 32,
                                 (uint16_t *)u3.d + u3.off[s], st);
        }
        check(hipStreamSynchronize(st), "split VT / U");
        return &vt3;
    }

    // E'_d split into 16-bit parts for a mixed-precision variant (built once)
    const SplitE &ensure_split(int variant, hipStream_t st)
    {
        std::lock_guard<std::mutex> lk(mu);
        SplitE &S = split[variant];
        if (S.d) return S;
        make_mixed_sets();
        const int nsplit = variant == PBB_PREC_MIXED ? 3 : 2, ns = h.nsets;
        S.off.assign(ns, 0);
        S.R.assign(ns, 0);
        S.nch.assign(ns, 0);
        S.escale.assign(ns, 0);
        size_t total = 0;
        for (int s = 0; s < ns; s++) {
            S.R[s] = 16 * ((mset[s].kf + 15) / 16);
            S.nch[s] = (mset[s].ks + 31) / 32;
            S.off[s] = total;
            total += (size_t)316 * S.nch[s] * nsplit * S.R[s] * 32;
        }
        S.d = dalloc<uint16_t>(total, "split E");
        for (int s = 0; s < ns; s++) {
            const MixedSet &m = mset[s];
            if (variant == PBB_PREC_MIXED_F16X2) {
                int e = 0;
                if (mset_emax[s] > 0) std::frexp(mset_emax[s], &e);
                S.escale[s] = mset_emax[s] > 0 ? 12 - e : 0;
            }
            dev::launch_split_e(variant, dm_E + m.e_off, m.kf, m.ks, S.R[s], S.nch[s], S.escale[s],
                                (uint16_t *)S.d + S.off[s], st);
        }
        check(hipStreamSynchronize(st), "split E");
        return S;
    }

    void upload(int dev)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (device == dev) return;
        if (device >= 0) fail(PBB_ERR_UNSUPPORTED, "operators are already resident on another device");
        const int n3 = h.n3, rf = h.rf, rs = h.rs;
        std::vector<double> VTw(h.VT.size()), Uw(h.U.size());
        for (int s = 0; s < h.nsets; s++)
            for (int q = 0; q < n3; q++) {
                for (int a = 0; a < rs; a++) {
                    size_t i = (size_t)s * rs * n3 + a + (size_t)q * rs;
                    VTw[i] = h.VT[i] * h.w.Kw[q];
                }
                for (int a = 0; a < rf; a++) {
                    size_t i = (size_t)s * n3 * rf + q + (size_t)a * n3;
                    Uw[i] = h.U[i] * h.w.Kw[q];
                }
            }
        d_Tkz = dupload(h.w.Tkz, "Tkz");
        d_Cw = dupload(h.w.Cw, "Cweights");
        d_VTw = dupload(VTw, "VT");
        d_Uw = dupload(Uw, "U");
        d_E = dupload(h.E, "E");
        device = dev;
    }
    // The operators of the transposed FMM (P2M <-> L2P, M2M <-> L2L): U' = VT^T,
    // VT' = U^T, E'_d = E_{-d}^T with -d at index 315 - d (Ktable order is
    // symmetric under negation).  Same weights, kernel and level scales.
    pbb_ops *adjoint()
    {
        std::lock_guard<std::mutex> lk(mu);
        if (adj) return adj.get();
        auto a = std::make_unique<pbb_ops>();
        a->level_tol = level_tol;
        a->fp64_reduced = fp64_reduced;
        HostOps &g = a->h;
        g = h;
        g.rf = h.rs;
        g.rs = h.rf;
        g.sigma_f = h.sigma_s;
        g.sigma_s = h.sigma_f;
        const int n3 = h.n3, rf = h.rf, rs = h.rs;
        g.U.assign((size_t)h.nsets * n3 * rs, 0.0);
        g.VT.assign((size_t)h.nsets * rf * n3, 0.0);
        g.E.assign(h.E.size(), 0.0);
        for (int s = 0; s < h.nsets; s++) {
            const double *U = &h.U[(size_t)s * n3 * rf], *VT = &h.VT[(size_t)s * rs * n3];
            double *U2 = &g.U[(size_t)s * n3 * rs], *VT2 = &g.VT[(size_t)s * rf * n3];
            for (int q = 0; q < n3; q++) {
                for (int b = 0; b < rs; b++) U2[q + (size_t)b * n3] = VT[b + (size_t)q * rs];
                for (int a2 = 0; a2 < rf; a2++) VT2[a2 + (size_t)q * rf] = U[q + (size_t)a2 * n3];
            }
            for (int d = 0; d < 316; d++) {
                const double *Ed = &h.E[((size_t)s * 316 + (315 - d)) * rf * rs];
                double *E2 = &g.E[((size_t)s * 316 + d) * rf * rs];
                for (int a2 = 0; a2 < rf; a2++)
                    for (int b = 0; b < rs; b++) E2[b + (size_t)a2 * rs] = Ed[a2 + (size_t)b * rf];
            }
        }
        adj = std::move(a);
        return adj.get();
    }

    ~pbb_ops()
    {
        if (device >= 0) {
            int prev = -1;
            (void)hipGetDevice(&prev);
            (void)hipSetDevice(device);
            dfree(d_Tkz); dfree(d_Cw); dfree(d_VTw); dfree(d_Uw); dfree(d_E);
            dfree(dm_E); dfree(dm_VTw); dfree(dm_Uw); dfree(dm_px);
            for (auto &x : split) dfree(x.d);
            dfree(vt3.d); dfree(u3.d);
            if (prev >= 0) (void)hipSetDevice(prev);
        }
    }
};

// Device buffers of one matvec.  The FP64 mode keeps the proxies P and the
// M2L results Pf in FP64; the mixed modes keep them in FP32 (P32 k-padded to
// kpad, a multiple of 32).  The forward and the transposed operator sets of
// one plan fit the same layout (same n3, max(rank_f, rank_s) and kpad).
struct Workspace {
    double *ws = nullptr, *wsT = nullptr;  // sorted weights, column- and point-major
    double *lp = nullptr;                  // per-level L2P results, point-major
    double *up = nullptr;                  // partial expansions of split boxes
    double *S = nullptr;                   // expansions S / F, [box][col][n3]
    double *P = nullptr, *Pf = nullptr;    // FP64 mode
    float *P32 = nullptr, *Pf32 = nullptr; // mixed modes
    unsigned *xmax = nullptr;              // per-level max |P| (FP16 scaling)
    int kpad = 0;
    int ncols = 0, mixed = -1;             // what the layout was made for
};

struct pbb_plan;
static size_t layout_workspace(const pbb_plan &p, int ncols, bool mixed, char *base, Workspace &W);

struct DevLevel {
    int64_t nb = 0, off = 0;
    int32_t *coord = nullptr, *map = nullptr, *child = nullptr, *order = nullptr;
    int32_t *pstart = nullptr;                       // box -> first sorted point (levels >= 2)
    int32_t *seg_box = nullptr, *seg_p0 = nullptr;   // L2P segments of at most L2P_SEG points
    int nseg = 0;
    int up_item0 = 0, up_red0 = 0;                  // the level's first direct upward work item and reduction
    double *ctab = nullptr;
    double len = 0;
    int64_t class_count[8] = {0};
};

struct pbb_plan {
    pbb_ops *ops = nullptr;
    int device = -1;
    hipStream_t stream = nullptr;
    int Lv = 0;
    int64_t N = 0;
    // A plan of separate targets and sources holds the nt targets followed
    // by the ns sources (N = nt + ns; sep); otherwise nt = ns = N, one set.
    bool sep = false;
    int64_t nt = 0, ns = 0;
    pbb_plan_info info{};
    double *x = nullptr, *y = nullptr, *z = nullptr;
    int32_t *perm = nullptr, *leaf_start = nullptr;
    int32_t *leaf_class = nullptr;              // leaves grouped by point count
    int4 *up_items = nullptr, *up_reds = nullptr;  // direct upward pass work list
    int n_up_items = 0, n_up_reds = 0, n_up_slots = 0;
    bool expansion_route = false;               // PBB_PLAN_EXPANSION_ROUTE: no proxy passes
    int leaf_class_off[dev::NEAR_CLASSES + 1] = {0};
    std::vector<DevLevel> lv;
    int64_t Btot = 0;
    Workspace W;                  // the plan's own matvec buffers (one block)
    char *ws_block = nullptr;
    size_t ws_bytes = 0;
    double *gF = nullptr, *gU = nullptr;  // position gradients: saved leaf F, sorted weights
    // for pbb_plan_update_points: per-level centres, and (fetched from the
    // device on first use when the tree was built there) sorted position ->
    // input index and the leaf coordinates of every sorted point
    std::vector<int32_t> h_perm;
    std::vector<uint32_t> h_leaf;
    std::vector<Level> h_geom;
    int32_t *pt_leaf = nullptr;  // device: leaf coordinates of every sorted point
    int g_ncols = 0;
    bool timing = false;
    struct M2LTile {
        int width = 0, cpt = 0, flags = 0;  // width 0 = built-in choice
    } m2l_tile[2];                          // FP64, mixed
    M2LTile m2l_level[2][16];               // per-level tiles (width 0: m2l_tile)
    hipEvent_t lev_ev[17] = {};             // M2L level boundaries (autotuner)
    bool lev_made = false, lev_timing = false;
    size_t ws_limit = 0;                    // workspace cap in bytes (0: device memory)
    hipEvent_t ev[PBB_NUM_PHASES] = {};
    bool events_made = false;
    double last_ms[PBB_NUM_PHASES] = {0};

    ~pbb_plan()
    {
        if (device < 0) return;
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        dfree(x); dfree(y); dfree(z); dfree(perm); dfree(leaf_start); dfree(leaf_class);
        dfree(up_items); dfree(up_reds); dfree(pt_leaf);
        for (auto &L : lv) {
            dfree(L.coord); dfree(L.map); dfree(L.child); dfree(L.order); dfree(L.ctab);
            dfree(L.pstart); dfree(L.seg_box); dfree(L.seg_p0);
        }
        dfree(ws_block); dfree(gF); dfree(gU);
        if (events_made)
            for (auto &e : ev) (void)hipEventDestroy(e);
        if (lev_made)
            for (auto &e : lev_ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        if (prev >= 0) (void)hipSetDevice(prev);
    }

    // The plan's own matvec buffers, grown (never shrunk) to fit ncols
    // columns in the precision class; one block, laid out by layout_workspace.
    void ensure_workspace(int ncols, bool mixed)
    {
        const size_t need = layout_workspace(*this, ncols, mixed, nullptr, W);
        if (need > ws_bytes) {
            dfree(ws_block);
            ws_block = nullptr;
            ws_bytes = 0;
            ws_block = dalloc<char>(need, "matvec workspace");
            ws_bytes = need;
        }
        layout_workspace(*this, ncols, mixed, ws_block, W);
    }
};

// Lays out the buffers of one matvec of p with ncols columns from base
// (nullptr: sizes only, pointers left null) and returns the bytes used; every
// buffer starts on a 256-byte boundary.
static size_t layout_workspace(const pbb_plan &p, int ncols, bool mixed, char *base, Workspace &W)
{
    const HostOps &h = p.ops->h;
    size_t off = 0;
    auto take = [&](size_t bytes) -> void * {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return base ? (void *)(base + at) : nullptr;
    };
    const size_t Nc = (size_t)p.N * ncols, Bc = (size_t)p.Btot * ncols;
    const int rmax = std::max(h.rf, h.rs);
    W = Workspace{};
    W.ncols = ncols;
    W.mixed = mixed ? 1 : 0;
    W.kpad = 32 * ((rmax + 31) / 32);
    W.ws = (double *)take(Nc * sizeof(double));
    W.wsT = (double *)take(Nc * sizeof(double));
    if (p.Lv >= 2) {
        W.lp = (double *)take(Nc * sizeof(double));
        W.S = (double *)take(Bc * h.n3 * sizeof(double));
        if (mixed) {
            W.P32 = (float *)take(Bc * W.kpad * sizeof(float));
            W.Pf32 = (float *)take(Bc * rmax * sizeof(float));
            W.xmax = (unsigned *)take((size_t)(p.Lv + 1) * sizeof(unsigned));
        } else {
            W.P = (double *)take(Bc * rmax * sizeof(double));
            W.Pf = (double *)take(Bc * rmax * sizeof(double));
        }
    }
    if (p.n_up_slots > 0) W.up = (double *)take((size_t)p.n_up_slots * ncols * h.n3 * sizeof(double));
    return off;
}

// ---------------------------------------------------------------------------

extern "C" {

PBB_API const char *pbb_last_error(void) { return g_last_error.c_str(); }

PBB_API int pbb_abi_version(void) { return PBB_ABI_VERSION; }

PBB_API int pbb_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

static void validate_params(const pbb_kernel_desc *kernel, double L, int tree_level, int order, double eps,
                            int use_chebyshev)
{
    if (!kernel) fail(PBB_ERR_INVALID_ARG, "kernel descriptor is NULL");
    if (use_chebyshev != 1)
        fail(PBB_ERR_UNSUPPORTED, "only the Chebyshev branch (use_chebyshev = 1) is provided");
    if (!(L > 0) || !std::isfinite(L)) fail(PBB_ERR_INVALID_ARG, "L must be positive");
    if (tree_level < 1 || tree_level > 9) fail(PBB_ERR_INVALID_ARG, "tree_level must be in [1, 9]");
    if (order < 2 || order > 12) fail(PBB_ERR_INVALID_ARG, "interpolation order must be in [2, 12]");
    if (!(eps >= 0) || !(eps < 1)) fail(PBB_ERR_INVALID_ARG, "eps must be in [0, 1)");
    if (kernel->symmetry < -1 || kernel->symmetry > 1) fail(PBB_ERR_INVALID_ARG, "symmetry must be -1, 0 or 1");
    if (kernel->id != PBB_KERNEL_NONE && !kernel_id_valid(kernel->id))
        fail(PBB_ERR_INVALID_ARG, "unknown kernel id");
}

}  // extern "C"

namespace {
// Device buffers of one operator precompute on the GPU (kernel blocks and
// the compression), freed with it.
struct DevicePrecompute {
    int device = -1;
    hipStream_t st = nullptr;
    const HostOps *h = nullptr;
    double *nodes = nullptr, *Kw = nullptr, *Kcell = nullptr, *X = nullptr, *Ut = nullptr, *V = nullptr,
           *E = nullptr;
    int *dirs = nullptr;
    size_t kcell_n = 0, x_n = 0, e_n = 0;

    void start(const HostOps &ops, int dev)
    {
        h = &ops;
        device = dev;
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    }
    // K blocks of one operator set on the device; the host copy feeds the
    // SVDs.  The buffers are made at the first call, once precompute_ops has
    // set the order's nodes, weights and n3.
    void blocks(double box_len, std::vector<double> &Kc)
    {
        const HostOps &o = *h;
        const size_t kn = (size_t)316 * o.n3 * o.n3;
        if (o.n3 <= 0 || o.w.nodes.size() != (size_t)o.n || o.w.Kw.size() != (size_t)o.n3)
            fail(PBB_ERR_INTERNAL, "kernel blocks requested before the Chebyshev weights");
        if (!Kcell) {
            nodes = dupload(o.w.nodes, "Chebyshev nodes");
            Kw = dupload(o.w.Kw, "Kweights");
            std::vector<int> d3;
            for (int k1 = -3; k1 < 4; k1++)
                for (int k2 = -3; k2 < 4; k2++)
                    for (int k3 = -3; k3 < 4; k3++)
                        if (std::abs(k1) > 1 || std::abs(k2) > 1 || std::abs(k3) > 1)
                            d3.insert(d3.end(), {k1, k2, k3});
            dirs = dupload(d3, "directions");
            Kcell = dalloc<double>(kn, "kernel blocks");
            kcell_n = kn;
        }
        if (kcell_n != kn) fail(PBB_ERR_INTERNAL, "kernel block buffer made for another order");
        dev::KernelBlockArgs a{};
        a.kcell_len = (int64_t)kcell_n;
        a.n = o.n; a.n3 = o.n3; a.symm = o.kernel.symmetry; a.symm_num = 158 * (2 - std::abs(a.symm));
        a.box_len = box_len;
        std::memcpy(a.kp.p, o.kernel.param, sizeof(a.kp.p));
        a.nodes = nodes; a.Kw = Kw; a.dirs = dirs; a.Kcell = Kcell;
        dev::launch_kernel_blocks(a, o.kernel.id, st);
        Kc.resize((size_t)316 * o.n3 * o.n3);
        check(hipMemcpyAsync(Kc.data(), Kcell, sizeof(double) * Kc.size(), hipMemcpyDeviceToHost, st), "copy blocks");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    // E_d = U^T K_d V: X = U^T [K_0 .. K_315] in one GEMM, then E_d = X_d V
    void compress(int rf, int rs, const double *U, const double *VT, double *Eh)
    {
        const int n3 = h->n3;
        if (!Kcell || kcell_n != (size_t)316 * n3 * n3) fail(PBB_ERR_INTERNAL, "compression before the kernel blocks");
        std::vector<double> ut((size_t)rf * n3), v((size_t)n3 * rs);
        for (int a = 0; a < rf; a++)
            for (int i = 0; i < n3; i++) ut[a + (size_t)i * rf] = U[i + (size_t)a * n3];
        for (int j = 0; j < n3; j++)
            for (int b = 0; b < rs; b++) v[j + (size_t)b * n3] = VT[b + (size_t)j * rs];
        dfree(Ut);
        dfree(V);
        Ut = V = nullptr;
        Ut = dupload(ut, "U^T");
        V = dupload(v, "V");
        if (x_n < (size_t)rf * 316 * n3) {
            dfree(X);
            X = nullptr;
            X = dalloc<double>((size_t)rf * 316 * n3, "U^T K");
            x_n = (size_t)rf * 316 * n3;
        }
        if (e_n < (size_t)316 * rf * rs) {
            dfree(E);
            E = nullptr;
            E = dalloc<double>((size_t)316 * rf * rs, "E");
            e_n = (size_t)316 * rf * rs;
        }
        dev::GemmArgs g{};
        g.M = rf; g.K = n3; g.ldb = n3; g.ldc = rf; g.ncol = (int64_t)316 * n3; g.alpha = 1.0;
        g.A = Ut; g.B = Kcell; g.C = X;
        dev::launch_gemm(g, st);
        for (int d = 0; d < 316; d++) {
            dev::GemmArgs e{};
            e.M = rf; e.K = n3; e.ldb = n3; e.ldc = rf; e.ncol = rs; e.alpha = 1.0;
            e.A = X + (size_t)d * rf * n3; e.B = V; e.C = E + (size_t)d * rf * rs;
            dev::launch_gemm(e, st);
        }
        check(hipMemcpyAsync(Eh, E, sizeof(double) * (size_t)316 * rf * rs, hipMemcpyDeviceToHost, st), "copy E");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    ~DevicePrecompute()
    {
        if (device < 0) return;
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(device);
        if (st) (void)hipStreamSynchronize(st);
        dfree(nodes); dfree(Kw); dfree(dirs); dfree(Kcell); dfree(X); dfree(Ut); dfree(V); dfree(E);
        if (st) (void)hipStreamDestroy(st);
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};
}  // namespace

extern "C" {

PBB_API pbb_status pbb_ops_create(const pbb_kernel_desc *kernel, pbb_host_kernel_fn host_eval, void *ctx, double L,
                                  int32_t tree_level, int32_t order, double eps, int32_t use_chebyshev,
                                  pbb_ops **out)
{
    return pbb_ops_create_ex(kernel, host_eval, ctx, L, tree_level, order, eps, use_chebyshev, 0, out);
}

PBB_API pbb_status pbb_ops_create_ex(const pbb_kernel_desc *kernel, pbb_host_kernel_fn host_eval, void *ctx,
                                     double L, int32_t tree_level, int32_t order, double eps,
                                     int32_t use_chebyshev, int32_t flags, pbb_ops **out)
{
    return guarded([&] {
        if (!out) fail(PBB_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (flags & ~PBB_OPS_HOST_PRECOMPUTE) fail(PBB_ERR_INVALID_ARG, "unknown operator flags");
        validate_params(kernel, L, tree_level, order, eps, use_chebyshev);
        if (!host_eval && !kernel_id_valid(kernel->id))
            fail(PBB_ERR_INVALID_ARG, "no host callback and no kernel functor id");
        auto ops = std::make_unique<pbb_ops>();
        HostOps &h = ops->h;
        h.kernel = *kernel;
        h.L = L;
        h.tree_level = tree_level;
        h.n = order;
        h.eps = eps;
        HostEval eval;
        if (host_eval)
            eval = [host_eval, ctx](const double *t, const double *s) { return host_eval(ctx, t, s); };
        else {
            int id = kernel->id;
            double p[4];
            std::memcpy(p, kernel->param, sizeof(p));
            eval = [id, p](const double *t, const double *s) { return kernel_eval_host(id, p, t, s); };
        }
        // the kernel blocks and the compression run on the current GPU when
        // there is one and the kernel has a device functor; the SVDs stay on
        // the host
        int ndev = 0;
        const bool on_gpu = !(flags & PBB_OPS_HOST_PRECOMPUTE) && !host_eval && kernel->id != PBB_KERNEL_POLY_XY &&
                            kernel_id_valid(kernel->id) && hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0;
        (void)hipGetLastError();
        if (on_gpu) {
            DevicePrecompute dp;
            dp.start(h, current_device());
            PrecomputeDevice pd;
            pd.blocks = [&](double box_len, std::vector<double> &Kc) { dp.blocks(box_len, Kc); };
            pd.compress = [&](int rf, int rs, const double *U, const double *VT, double *E) {
                dp.compress(rf, rs, U, VT, E);
            };
            precompute_ops(h, eval, &pd);
        } else {
            precompute_ops(h, eval);
        }
        *out = ops.release();
    });
}

// operators of a kernel and tree shape from given U, VT, E (any source);
// the weights come from the order, the operator sets from the homogeneity
static std::unique_ptr<pbb_ops> ops_from_arrays(const pbb_kernel_desc *kernel, double L, int tree_level, int order,
                                                double eps, int rank_f, int rank_s, int nsets, const double *U,
                                                const double *VT, const double *E)
{
    validate_params(kernel, L, tree_level, order, eps, 1);
    if (!U || !VT || !E) fail(PBB_ERR_INVALID_ARG, "operator arrays must not be NULL");
    auto ops = std::make_unique<pbb_ops>();
    HostOps &h = ops->h;
    h.kernel = *kernel;
    h.L = L;
    h.tree_level = tree_level;
    h.n = order;
    h.n3 = order * order * order;
    h.eps = eps;
    h.w.compute(order);
    h.homogen_flipped = -kernel->homogen;
    h.homog = std::fabs(h.homogen_flipped) > 0.1;
    int want = h.homog ? 1 : std::max(1, tree_level - 1);
    if (nsets != want) fail(PBB_ERR_INVALID_ARG, "nsets does not match the kernel's homogeneity and tree_level");
    if (rank_f < 1 || rank_f > h.n3 || rank_s < 1 || rank_s > h.n3) fail(PBB_ERR_INVALID_ARG, "bad ranks");
    h.rf = rank_f;
    h.rs = rank_s;
    h.nsets = nsets;
    h.U.assign(U, U + (size_t)nsets * h.n3 * rank_f);
    h.VT.assign(VT, VT + (size_t)nsets * rank_s * h.n3);
    h.E.assign(E, E + (size_t)nsets * 316 * rank_f * rank_s);
    return ops;
}

PBB_API pbb_status pbb_ops_create_from_arrays(const pbb_kernel_desc *kernel, double L, int32_t tree_level,
                                              int32_t order, double eps, int32_t rank_f, int32_t rank_s,
                                              int32_t nsets, const double *U, const double *VT, const double *E,
                                              pbb_ops **out)
{
    return guarded([&] {
        if (!out) fail(PBB_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        *out = ops_from_arrays(kernel, L, tree_level, order, eps, rank_f, rank_s, nsets, U, VT, E).release();
    });
}

}  // extern "C"

// ---- operator cache files in the reference's format ----------------------
//   K file: double box length L, int32 tree level (StartPrecompute,
//           H2_3D_Tree.cpp:141-143), then per operator set the 316 E blocks,
//           rank_f x rank_s column-major each, in Ktable order (:528-558)
//   U file: int32 rank_f (:424-427), then per set U, n3 x rank_f
//           column-major (:440-441)
//   V file: int32 rank_s (:496-498), then per set VT, rank_s x n3
//           column-major (:510-515)
// Operator sets: one for homogeneous kernels (unit box), else levels 2..L
// (:146-160); read back as FMMReadMatrices does (:186-239).
namespace {
struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(path ? std::fopen(path, mode) : nullptr) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};
long file_size(FILE *f)
{
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    return n;
}
void write_all(FILE *f, const void *p, size_t size, size_t n, const char *what)
{
    if (std::fwrite(p, size, n, f) != n) fail(PBB_ERR_INVALID_ARG, std::string("cannot write the ") + what);
}
void read_all(FILE *f, void *p, size_t size, size_t n, const char *what)
{
    if (std::fread(p, size, n, f) != n) fail(PBB_ERR_INVALID_ARG, std::string("cannot read the ") + what);
}
}  // namespace

namespace {
// The parameters operator cache files were made for, kept beside the K file
// as "<K file>.params" (the reference names its files by kernel type and
// order only, H2_3D_Tree.cpp:81-83, so a cache made with another eps or
// kernel parameter would be read silently).
std::string cache_params(const pbb_kernel_desc &k, double L, int tree_level, int order, double eps, bool homog)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "pbbfmm3d_amd operator cache: kernel %d symmetry %d homogen %.17g params %.17g %.17g %.17g %.17g "
                  "order %d eps %.17g L %.17g tree_level %d\n",
                  k.id, k.symmetry, k.homogen, k.param[0], k.param[1], k.param[2], k.param[3], order, eps,
                  homog ? 1.0 : L, homog ? 0 : tree_level);
    return buf;
}
}  // namespace
extern "C" {

PBB_API pbb_status pbb_ops_write_cache(const pbb_ops *ops, const char *k_path, const char *u_path, const char *v_path)
{
    return guarded([&] {
        if (!ops || !k_path || !u_path || !v_path) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        const HostOps &h = ops->h;
        {
            File fp((std::string(k_path) + ".params").c_str(), "w");
            if (!fp.f) fail(PBB_ERR_INVALID_ARG, "cannot open the cache parameter file for writing");
            const std::string t = cache_params(h.kernel, h.L, h.tree_level, h.n, h.eps, h.homog);
            write_all(fp.f, t.data(), 1, t.size(), "parameter file");
        }
        File fk(k_path, "wb"), fu(u_path, "wb"), fv(v_path, "wb");
        if (!fk.f || !fu.f || !fv.f) fail(PBB_ERR_INVALID_ARG, "cannot open the cache files for writing");
        const double L = h.L;
        const int32_t lvl = h.tree_level, rf = h.rf, rs = h.rs;
        write_all(fk.f, &L, sizeof(double), 1, "K file");
        write_all(fk.f, &lvl, sizeof(int32_t), 1, "K file");
        write_all(fk.f, h.E.data(), sizeof(double), h.E.size(), "K file");
        write_all(fu.f, &rf, sizeof(int32_t), 1, "U file");
        write_all(fu.f, h.U.data(), sizeof(double), h.U.size(), "U file");
        write_all(fv.f, &rs, sizeof(int32_t), 1, "V file");
        write_all(fv.f, h.VT.data(), sizeof(double), h.VT.size(), "V file");
    });
}

PBB_API pbb_status pbb_ops_read_cache(const pbb_kernel_desc *kernel, double L, int32_t tree_level, int32_t order,
                                      double eps, const char *k_path, const char *u_path, const char *v_path,
                                      pbb_ops **out)
{
    return guarded([&] {
        if (!out) fail(PBB_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        validate_params(kernel, L, tree_level, order, eps, 1);
        File fk(k_path, "rb"), fu(u_path, "rb"), fv(v_path, "rb");
        if (!fk.f || !fu.f || !fv.f) fail(PBB_ERR_NOT_FOUND, "cache files not found");
        {  // files written here carry their parameters; the reference's do not
            File fp((std::string(k_path) + ".params").c_str(), "r");
            if (fp.f) {
                char buf[512] = {0};
                const size_t got = std::fread(buf, 1, sizeof(buf) - 1, fp.f);
                const bool homog = std::fabs(kernel->homogen) > 0.1;
                if (std::string(buf, got) != cache_params(*kernel, L, tree_level, order, eps, homog))
                    fail(PBB_ERR_INVALID_ARG, "the cache files were written for other kernel parameters, eps, order "
                                              "or tree (see " + std::string(k_path) + ".params)");
            }
        }
        const long nk = file_size(fk.f), nu = file_size(fu.f), nv = file_size(fv.f);
        double len = 0;
        int32_t lvl = 0, rf = 0, rs = 0;
        read_all(fk.f, &len, sizeof(double), 1, "K file header");
        read_all(fk.f, &lvl, sizeof(int32_t), 1, "K file header");
        read_all(fu.f, &rf, sizeof(int32_t), 1, "U file header");
        read_all(fv.f, &rs, sizeof(int32_t), 1, "V file header");
        const int n3 = order * order * order;
        if (rf < 1 || rf > n3 || rs < 1 || rs > n3) fail(PBB_ERR_INVALID_ARG, "cache ranks do not fit the order");
        const bool homog = std::fabs(kernel->homogen) > 0.1;
        // PrecomputeAvailable (H2_3D_Tree.cpp:114-122): per-level operators
        // belong to one root length and depth
        if (!homog && (lvl != tree_level || std::fabs(len - L) >= 1e-6))
            fail(PBB_ERR_INVALID_ARG, "cache was written for another root length or tree level");
        const int nsets = homog ? 1 : std::max(1, tree_level - 1);
        const size_t su = (size_t)nsets * n3 * rf, sv = (size_t)nsets * rs * n3, se = (size_t)nsets * 316 * rf * rs;
        if ((size_t)nu != 4 + 8 * su || (size_t)nv != 4 + 8 * sv || (size_t)nk != 12 + 8 * se)
            fail(PBB_ERR_INVALID_ARG, "cache file sizes do not match the order, ranks and operator sets");
        std::vector<double> U(su), VT(sv), E(se);
        read_all(fu.f, U.data(), sizeof(double), su, "U file");
        read_all(fv.f, VT.data(), sizeof(double), sv, "V file");
        read_all(fk.f, E.data(), sizeof(double), se, "K file");
        *out = ops_from_arrays(kernel, L, tree_level, order, eps, rf, rs, nsets, U.data(), VT.data(), E.data())
                   .release();
    });
}

PBB_API pbb_status pbb_ops_query(const pbb_ops *ops, pbb_ops_info *info)
{
    return guarded([&] {
        if (!ops || !info) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        const HostOps &h = ops->h;
        std::memset(info, 0, sizeof(*info));
        info->order = h.n;
        info->n3 = h.n3;
        info->rank_f = h.rf;
        info->rank_s = h.rs;
        info->nsets = h.nsets;
        info->homogeneous = h.homog ? 1 : 0;
        info->tree_level = h.tree_level;
        info->L = h.L;
        info->homogen = h.kernel.homogen;
        info->eps = h.eps;
    });
}

PBB_API pbb_status pbb_ops_set_mixed_tolerance(pbb_ops *ops, double tol)
{
    return guarded([&] {
        if (!ops) fail(PBB_ERR_INVALID_ARG, "ops is NULL");
        if (!(tol >= 0) || tol > 1e-6) fail(PBB_ERR_INVALID_ARG, "tol must be in [0, 1e-6]");
        std::lock_guard<std::mutex> lk(ops->mu);
        if (!ops->mset.empty() || (ops->adj && !ops->adj->mset.empty()))
            fail(PBB_ERR_INVALID_ARG, "the mixed-precision operators are prepared already: set the tolerance "
                                      "before the first mixed-precision call");
        ops->level_tol = tol;
        if (ops->adj) ops->adj->level_tol = tol;
    });
}

PBB_API pbb_status pbb_ops_set_fp64_recompressed(pbb_ops *ops, int32_t enable)
{
    return guarded([&] {
        if (!ops) fail(PBB_ERR_INVALID_ARG, "ops is NULL");
        std::lock_guard<std::mutex> lk(ops->mu);
        ops->fp64_reduced = enable != 0;
        if (ops->adj) ops->adj->fp64_reduced = enable != 0;
    });
}

PBB_API pbb_status pbb_ops_mixed_ranks(pbb_ops *ops, int32_t *kf, int32_t *ks, double *tail)
{
    return guarded([&] {
        if (!ops || !kf || !ks) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        std::lock_guard<std::mutex> lk(ops->mu);
        ops->compute_mixed_sets();
        for (int s = 0; s < ops->h.nsets; s++) {
            kf[s] = ops->mset[s].kf;
            ks[s] = ops->mset[s].ks_used;
            if (tail) tail[s] = ops->mset[s].tail;
        }
    });
}

PBB_API pbb_status pbb_ops_export(const pbb_ops *ops, double *U, double *VT, double *E, double *sigma_f,
                                  double *sigma_s)
{
    return guarded([&] {
        if (!ops) fail(PBB_ERR_INVALID_ARG, "ops is NULL");
        const HostOps &h = ops->h;
        if (U) std::memcpy(U, h.U.data(), sizeof(double) * h.U.size());
        if (VT) std::memcpy(VT, h.VT.data(), sizeof(double) * h.VT.size());
        if (E) std::memcpy(E, h.E.data(), sizeof(double) * h.E.size());
        if (sigma_f) {
            if (h.sigma_f.empty()) std::fill(sigma_f, sigma_f + h.n3, NAN);
            else std::memcpy(sigma_f, h.sigma_f.data(), sizeof(double) * h.sigma_f.size());
        }
        if (sigma_s) {
            if (h.sigma_s.empty()) std::fill(sigma_s, sigma_s + h.n3, NAN);
            else std::memcpy(sigma_s, h.sigma_s.data(), sizeof(double) * h.sigma_s.size());
        }
    });
}

PBB_API void pbb_ops_destroy(pbb_ops *ops) { delete ops; }

}  // extern "C"

static void plan_create_host(pbb_ops *ops, const double *xyz, int64_t npoints, int64_t nbox, pbb_plan **out)
{
    {
        const HostOps &h = ops->h;
        int dev = current_device();
        HostTree T;
        build_tree(T, xyz, npoints, h.L, h.tree_level, nbox);
        ops->upload(dev);
        auto plan = std::make_unique<pbb_plan>();
        plan->ops = ops;
        plan->device = dev;
        check(hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking), "hipStreamCreate");
        plan->Lv = T.Lv;
        plan->N = plan->nt = plan->ns = T.N;
        plan->x = dupload(T.xs, "x");
        plan->y = dupload(T.ys, "y");
        plan->z = dupload(T.zs, "z");
        plan->perm = dupload(T.perm, "perm");
        plan->leaf_start = dupload(T.leaf_start, "leaf_start");
        {  // class c: leaves of at most 2 << c points; the last class takes the rest
            const int64_t nbl = T.lv[T.Lv].nb;
            std::vector<int32_t> cls[dev::NEAR_CLASSES], all;
            for (int64_t b = 0; b < nbl; b++) {
                const int nt = T.leaf_start[b + 1] - T.leaf_start[b];
                int c = 0;
                while (c < dev::NEAR_CLASSES - 1 && nt > (2 << c)) c++;
                cls[c].push_back((int32_t)b);
            }
            all.reserve(nbl);
            for (int c = 0; c < dev::NEAR_CLASSES; c++) {
                plan->leaf_class_off[c] = (int)all.size();
                all.insert(all.end(), cls[c].begin(), cls[c].end());
            }
            plan->leaf_class_off[dev::NEAR_CLASSES] = (int)all.size();
            plan->leaf_class = dupload(all, "leaf classes");
        }
        plan->lv.resize(T.Lv + 1);
        int64_t off = 0;
        for (int l = 0; l <= T.Lv; l++) {
            DevLevel &D = plan->lv[l];
            const Level &V = T.lv[l];
            D.nb = V.nb;
            D.len = V.len;
            D.coord = dupload(V.coord, "coord");
            D.map = dupload(V.map, "map");
            D.ctab = dupload(V.ctab, "ctab");
            if (!V.child.empty()) D.child = dupload(V.child, "child");
            if (!V.order.empty()) D.order = dupload(V.order, "order");
            std::copy(std::begin(V.class_count), std::end(V.class_count), D.class_count);
            if (l >= 2) {
                D.off = off;
                off += V.nb;
            }
        }
        plan->Btot = off;
        // point ranges of the boxes of levels 2..L: boxes and points are both
        // in Morton order, so every box owns a contiguous range
        if (off >= (1LL << 31)) fail(PBB_ERR_UNSUPPORTED, "too many boxes");
        std::vector<int4> up_items, up_reds;
        int n_slots = 0;
        for (int l = 2; l <= T.Lv; l++) {
            const Level &V = T.lv[l], &LL = T.lv[T.Lv];
            const int sh = T.Lv - l, S = 1 << l;
            std::vector<int32_t> ps(V.nb + 1, -1), sb, s0;
            for (int64_t b = 0; b < LL.nb; b++) {
                const uint32_t c = (uint32_t)LL.coord[b];
                const int ix = (c & 1023) >> sh, iy = ((c >> 10) & 1023) >> sh, iz = ((c >> 20) & 1023) >> sh;
                const int32_t a = V.map[((int64_t)ix * S + iy) * S + iz];
                if (a < 0 || a >= V.nb) fail(PBB_ERR_INTERNAL, "leaf without an ancestor box");
                if (ps[a] < 0) ps[a] = T.leaf_start[b];
            }
            ps[V.nb] = (int32_t)T.N;
            for (int64_t b = 0; b < V.nb; b++) {
                if (ps[b] < 0 || (b > 0 && ps[b] < ps[b - 1])) fail(PBB_ERR_INTERNAL, "box point ranges out of order");
                for (int32_t q = ps[b]; q < ps[b + 1]; q += dev::L2P_SEG) {
                    sb.push_back((int32_t)b);
                    s0.push_back(q);
                }
            }
            plan->lv[l].up_item0 = (int)up_items.size();
            plan->lv[l].up_red0 = (int)up_reds.size();
            for (int64_t b = 0; b < V.nb; b++) {  // direct upward pass work items
                const int32_t np = ps[b + 1] - ps[b];
                if (np <= dev::P2M_SEG) {
                    up_items.push_back(int4{l - 2, (int)b, ps[b], -1});
                    continue;
                }
                const int first = n_slots;
                for (int32_t q = ps[b]; q < ps[b + 1]; q += dev::P2M_SEG) up_items.push_back(int4{l - 2, (int)b, q, n_slots++});
                up_reds.push_back(int4{(int)(plan->lv[l].off + b), first, n_slots - first, 0});
            }
            DevLevel &D = plan->lv[l];
            D.pstart = dupload(ps, "box point ranges");
            D.seg_box = dupload(sb, "L2P segments");
            D.seg_p0 = dupload(s0, "L2P segments");
            D.nseg = (int)sb.size();
        }
        plan->h_perm = T.perm;
        plan->h_leaf.resize(T.N);
        for (int64_t b = 0; b < T.lv[T.Lv].nb; b++)
            for (int32_t q = T.leaf_start[b]; q < T.leaf_start[b + 1]; q++)
                plan->h_leaf[q] = (uint32_t)T.lv[T.Lv].coord[b];
        plan->h_geom.resize(T.Lv + 1);
        for (int l = 0; l <= T.Lv; l++) {
            plan->h_geom[l].ctab = T.lv[l].ctab;
            plan->h_geom[l].len = T.lv[l].len;
        }
        plan->up_items = dupload(up_items, "upward work items");
        plan->up_reds = dupload(up_reds, "upward reductions");
        plan->n_up_items = (int)up_items.size();
        plan->n_up_reds = (int)up_reds.size();
        plan->n_up_slots = n_slots;
        int dirs[8][189][4];
        for (int o = 0; o < 8; o++) class_directions(o, h.w.Ktable, dirs[o]);
        dev::upload_dirs(dirs);
        dev::upload_dirs_mixed(dirs);

        pbb_plan_info &I = plan->info;
        std::memset(&I, 0, sizeof(I));
        I.npoints = T.N;
        I.tree_level = T.Lv;
        for (int l = 0; l <= T.Lv && l < 16; l++) I.boxes[l] = T.lv[l].nb;
        I.nonempty_levels = T.Lv + 1;
        I.m2l_pairs = T.m2l_pairs;
        std::copy(std::begin(T.m2l_pairs_level), std::end(T.m2l_pairs_level), I.m2l_pairs_level);
        I.near_pairs = T.near_pairs;
        I.max_leaf_points = T.max_leaf;
        for (int d = 0; d < 3; d++) I.center[d] = T.ctr[d];
        I.outside_points = T.outside;
        I.ntargets = I.nsources = T.N;
        *out = plan.release();
    }
}

extern "C" {

// The plan's tree built on the GPU (tree_gpu.hip): the same arrays as the
// host build, with the plan's derived lists made there too.
static void plan_create_gpu(pbb_ops *ops, const double *xyz, int64_t npoints, int64_t nbox, pbb_plan **out)
{
    const HostOps &h = ops->h;
    const int dev = current_device();
    const int Lv = h.tree_level;
    ops->upload(dev);
    auto plan = std::make_unique<pbb_plan>();
    plan->ops = ops;
    plan->device = dev;
    check(hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking), "hipStreamCreate");
    // per-axis node centres with the arithmetic of BuildFMMHierarchy, as
    // build_tree computes them
    plan->h_geom.assign(Lv + 1, Level());
    plan->h_geom[0].ctab = {0.0};
    plan->h_geom[0].len = h.L;
    for (int l = 0; l < Lv; l++) {
        const Level &P = plan->h_geom[l];
        Level &C = plan->h_geom[l + 1];
        C.ctab.resize((size_t)1 << (l + 1));
        const double q = 0.25 * P.len;
        for (size_t i = 0; i < P.ctab.size(); i++) {
            C.ctab[2 * i] = P.ctab[i] - q;
            C.ctab[2 * i + 1] = P.ctab[i] + q;
        }
        C.len = 0.5 * P.len;
    }
    std::vector<std::vector<double>> ctab(Lv + 1);
    for (int l = 0; l <= Lv; l++) ctab[l] = plan->h_geom[l].ctab;
    int dirs[8][189][4];
    for (int o = 0; o < 8; o++) class_directions(o, h.w.Ktable, dirs[o]);
    dev::GpuTree T;
    dev::build_tree_gpu(T, xyz, npoints, nbox, h.L, Lv, ctab, dirs, plan->stream);
    // the plan owns the arrays from here on (its destructor frees them)
    plan->Lv = Lv;
    plan->N = plan->nt = plan->ns = T.N;
    plan->x = T.x; plan->y = T.y; plan->z = T.z;
    plan->perm = T.perm; plan->leaf_start = T.leaf_start; plan->leaf_class = T.leaf_class;
    plan->pt_leaf = T.pt_leaf;
    plan->up_items = T.up_items; plan->up_reds = T.up_reds;
    plan->lv.resize(Lv + 1);
    for (int l = 0; l <= Lv; l++) {
        DevLevel &D = plan->lv[l];
        const dev::GpuLevel &V = T.lv[l];
        D.nb = V.nb; D.off = V.off; D.coord = V.coord; D.map = V.map; D.child = V.child; D.order = V.order;
        D.pstart = V.pstart; D.seg_box = V.seg_box; D.seg_p0 = V.seg_p0; D.nseg = V.nseg;
        D.up_item0 = V.up_item0; D.up_red0 = V.up_red0;
        std::copy(std::begin(V.class_count), std::end(V.class_count), D.class_count);
        D.len = plan->h_geom[l].len;
    }
    for (int l = 0; l <= Lv; l++) plan->lv[l].ctab = dupload(ctab[l], "ctab");
    std::copy(std::begin(T.leaf_class_off), std::end(T.leaf_class_off), plan->leaf_class_off);
    plan->n_up_items = T.n_up_items;
    plan->n_up_reds = T.n_up_reds;
    plan->n_up_slots = T.n_up_slots;
    plan->Btot = T.Btot;
    dev::upload_dirs(dirs);
    dev::upload_dirs_mixed(dirs);
    pbb_plan_info &I = plan->info;
    std::memset(&I, 0, sizeof(I));
    I.npoints = T.N;
    I.tree_level = Lv;
    for (int l = 0; l <= Lv && l < 16; l++) I.boxes[l] = T.lv[l].nb;
    I.nonempty_levels = Lv + 1;
    I.m2l_pairs = T.m2l_pairs;
    std::copy(std::begin(T.m2l_pairs_level), std::end(T.m2l_pairs_level), I.m2l_pairs_level);
    I.near_pairs = T.near_pairs;
    I.max_leaf_points = T.max_leaf;
    for (int d = 0; d < 3; d++) I.center[d] = T.ctr[d];
    I.outside_points = T.outside;
    I.ntargets = I.nsources = T.N;
    *out = plan.release();
}

// A plan of npoints points whose first nbox define the bbox centre, built
// as flags say; with PBB_PLAN_REJECT_OUTSIDE, fails when a point lies
// outside the root cube.
static void plan_create(pbb_ops *ops, const double *xyz, int64_t npoints, int64_t nbox, int32_t flags,
                        pbb_plan **out)
{
    if (flags & ~(PBB_PLAN_HOST_TREE | PBB_PLAN_REJECT_OUTSIDE | PBB_PLAN_EXPANSION_ROUTE))
        fail(PBB_ERR_INVALID_ARG, "unknown plan flags");
    const HostOps &h = ops->h;
    if (!kernel_id_valid(h.kernel.id) || h.kernel.id == PBB_KERNEL_POLY_XY)
        fail(PBB_ERR_UNSUPPORTED, "the kernel has no device functor for the near field");
    pbb_plan *p = nullptr;
    if (flags & PBB_PLAN_HOST_TREE) plan_create_host(ops, xyz, npoints, nbox, &p);
    else plan_create_gpu(ops, xyz, npoints, nbox, &p);
    std::unique_ptr<pbb_plan> plan(p);
    plan->expansion_route = (flags & PBB_PLAN_EXPANSION_ROUTE) != 0;
    if ((flags & PBB_PLAN_REJECT_OUTSIDE) && plan->info.outside_points > 0)
        fail(PBB_ERR_INVALID_ARG, std::to_string(plan->info.outside_points) +
                                      " points lie outside the root cube of side L around the targets' bbox "
                                      "centre: build the operators with a larger L");
    *out = plan.release();
}

PBB_API pbb_status pbb_plan_create_ex(pbb_ops *ops, const double *xyz, int64_t npoints, int32_t flags,
                                      pbb_plan **out)
{
    return guarded([&] {
        if (!out) fail(PBB_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (!ops || !xyz) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        plan_create(ops, xyz, npoints, npoints, flags, out);
    });
}

PBB_API pbb_status pbb_plan_create_ts(pbb_ops *ops, const double *targets, int64_t ntargets, const double *sources,
                                      int64_t nsources, int32_t flags, pbb_plan **out)
{
    return guarded([&] {
        if (!out) fail(PBB_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (!ops || !targets || !sources) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ntargets <= 0 || nsources <= 0) fail(PBB_ERR_INVALID_ARG, "ntargets and nsources must be positive");
        if (ntargets + nsources >= (1LL << 31)) fail(PBB_ERR_INVALID_ARG, "ntargets + nsources must be < 2^31");
        // one tree over [targets; sources], centred on the targets' bbox
        // (compute.hpp:92-127 centres on the targets and shifts the sources
        // by the same vector)
        std::vector<double> both((size_t)3 * (ntargets + nsources));
        std::memcpy(both.data(), targets, sizeof(double) * 3 * (size_t)ntargets);
        std::memcpy(both.data() + 3 * (size_t)ntargets, sources, sizeof(double) * 3 * (size_t)nsources);
        pbb_plan *p = nullptr;
        plan_create(ops, both.data(), ntargets + nsources, ntargets, flags, &p);
        p->sep = true;
        p->nt = ntargets;
        p->ns = nsources;
        p->info.ntargets = ntargets;
        p->info.nsources = nsources;
        *out = p;
    });
}

PBB_API pbb_status pbb_plan_create(pbb_ops *ops, const double *xyz, int64_t npoints, pbb_plan **out)
{
    return pbb_plan_create_ex(ops, xyz, npoints, 0, out);
}

PBB_API pbb_status pbb_plan_update_points(pbb_plan *plan, const double *xyz, int64_t npoints, void *stream)
{
    return guarded([&] {
        if (!plan || !xyz) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (plan->sep) fail(PBB_ERR_UNSUPPORTED, "moving points needs a plan whose sources are its targets");
        if (npoints != plan->N) fail(PBB_ERR_INVALID_ARG, "npoints must be the plan's point count");
        const int64_t N = plan->N;
        const double *c = plan->info.center;
        DeviceGuard g(plan->device);
        if (plan->h_perm.empty()) {  // a tree built on the GPU: its ordering, once
            std::vector<int32_t> perm(N), leaf(N);
            check(hipMemcpy(perm.data(), plan->perm, N * sizeof(int32_t), hipMemcpyDeviceToHost), "copy perm");
            check(hipMemcpy(leaf.data(), plan->pt_leaf, N * sizeof(int32_t), hipMemcpyDeviceToHost), "copy leaves");
            plan->h_leaf.assign(leaf.begin(), leaf.end());
            plan->h_perm.swap(perm);
        }
        std::vector<double> xs(N), ys(N), zs(N);
        std::vector<char> moved_out(std::max<int64_t>(1, (N + 65535) / 65536), 0);
        parallel_for((N + 65535) / 65536, [&](int64_t blk) {
            for (int64_t k = blk * 65536, e = std::min(N, k + 65536); k < e; k++) {
                const int64_t i = plan->h_perm[k];
                const double p[3] = {xyz[3 * i] - c[0], xyz[3 * i + 1] - c[1], xyz[3 * i + 2] - c[2]};
                if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) ||
                    leaf_coord(plan->h_geom, plan->Lv, p) != plan->h_leaf[k])
                    moved_out[blk] = 1;
                xs[k] = p[0];
                ys[k] = p[1];
                zs[k] = p[2];
            }
        });
        for (char m : moved_out)
            if (m) fail(PBB_ERR_INVALID_ARG, "a point left its leaf box (or is not finite): build a new plan");
        hipStream_t st = (hipStream_t)stream;
        check(hipMemcpyAsync(plan->x, xs.data(), N * sizeof(double), hipMemcpyHostToDevice, st), "upload x");
        check(hipMemcpyAsync(plan->y, ys.data(), N * sizeof(double), hipMemcpyHostToDevice, st), "upload y");
        check(hipMemcpyAsync(plan->z, zs.data(), N * sizeof(double), hipMemcpyHostToDevice, st), "upload z");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
    });
}

PBB_API pbb_status pbb_plan_query(const pbb_plan *plan, pbb_plan_info *info)
{
    return guarded([&] {
        if (!plan || !info) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        *info = plan->info;
    });
}

PBB_API pbb_status pbb_plan_set_timing(pbb_plan *plan, int32_t enable)
{
    return guarded([&] {
        if (!plan) fail(PBB_ERR_INVALID_ARG, "plan is NULL");
        if (enable && !plan->events_made) {
            DeviceGuard g(plan->device);
            for (auto &e : plan->ev) check(hipEventCreate(&e), "hipEventCreate");
            plan->events_made = true;
        }
        plan->timing = enable != 0;
    });
}

PBB_API pbb_status pbb_plan_set_m2l_tile(pbb_plan *plan, int32_t precision, int32_t width, int32_t cols_per_box,
                                         int32_t flags)
{
    return guarded([&] {
        if (!plan) fail(PBB_ERR_INVALID_ARG, "plan is NULL");
        if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
            fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
        const bool mixed = precision != PBB_PREC_FP64;
        const bool wok = width == 0 || (mixed ? (width == 64 || width == 128 || width == 256)
                                              : (width == 32 || width == 64 || width == 128));
        if (!wok) fail(PBB_ERR_INVALID_ARG, "M2L width must be 0, 64, 128 or 256 (mixed) / 32, 64 or 128 (FP64)");
        if (width && (cols_per_box < 1 || cols_per_box > 64)) fail(PBB_ERR_INVALID_ARG, "cols_per_box must be 1..64");
        const int known = PBB_M2L_CHUNK_MAJOR | (mixed ? PBB_M2L_ROW_SPLIT | PBB_M2L_WAVE_ROWS | PBB_M2L_SPLIT_ACC |
                                                             PBB_M2L_DIR_GROUPS(0xff)
                                                       : 0);
        if (flags & ~known) fail(PBB_ERR_INVALID_ARG, "unknown M2L tile flags");
        if ((flags & PBB_M2L_WAVE_ROWS) && (flags & PBB_M2L_ROW_SPLIT))
            fail(PBB_ERR_INVALID_ARG, "PBB_M2L_WAVE_ROWS takes no row split");
        if ((flags & PBB_M2L_SPLIT_ACC) && !(flags & PBB_M2L_WAVE_ROWS))
            fail(PBB_ERR_INVALID_ARG, "PBB_M2L_SPLIT_ACC needs PBB_M2L_WAVE_ROWS");
        if (((flags >> 8) & 0xff) > 64) fail(PBB_ERR_INVALID_ARG, "at most 64 M2L direction groups");
        plan->m2l_tile[mixed ? 1 : 0] = {width, width ? cols_per_box : 0, flags};
        for (auto &t : plan->m2l_level[mixed ? 1 : 0]) t = pbb_plan::M2LTile{};
    });
}

PBB_API pbb_status pbb_plan_set_m2l_tile_level(pbb_plan *plan, int32_t precision, int32_t level, int32_t width,
                                               int32_t cols_per_box, int32_t flags)
{
    return guarded([&] {
        if (!plan) fail(PBB_ERR_INVALID_ARG, "plan is NULL");
        if (level < 2 || level > plan->Lv) fail(PBB_ERR_INVALID_ARG, "level must be in [2, tree_level]");
        const bool mixed = precision != PBB_PREC_FP64;
        // validate as the plan-wide setter does, then keep it for one level
        const pbb_plan::M2LTile keep = plan->m2l_tile[mixed ? 1 : 0];
        pbb_plan::M2LTile levels[16];
        std::copy(std::begin(plan->m2l_level[mixed ? 1 : 0]), std::end(plan->m2l_level[mixed ? 1 : 0]), levels);
        const pbb_status st = pbb_plan_set_m2l_tile(plan, precision, width, cols_per_box, flags);
        const pbb_plan::M2LTile t = plan->m2l_tile[mixed ? 1 : 0];
        plan->m2l_tile[mixed ? 1 : 0] = keep;
        std::copy(std::begin(levels), std::end(levels), plan->m2l_level[mixed ? 1 : 0]);
        if (st != PBB_OK) fail(st, g_last_error);
        plan->m2l_level[mixed ? 1 : 0][level] = t;
    });
}

PBB_API pbb_status pbb_plan_get_m2l_tile(const pbb_plan *plan, int32_t precision, int32_t level, int32_t *tile)
{
    return guarded([&] {
        if (!plan || !tile) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
            fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
        if (level != -1 && (level < 2 || level > plan->Lv)) fail(PBB_ERR_INVALID_ARG, "level must be -1 or in [2, tree_level]");
        const int c = precision != PBB_PREC_FP64 ? 1 : 0;
        const pbb_plan::M2LTile &t = level >= 0 && plan->m2l_level[c][level].width ? plan->m2l_level[c][level]
                                                                                   : plan->m2l_tile[c];
        tile[0] = t.width;
        tile[1] = t.cpt;
        tile[2] = t.flags;
    });
}

PBB_API pbb_status pbb_plan_set_workspace_limit(pbb_plan *plan, int64_t bytes)
{
    return guarded([&] {
        if (!plan) fail(PBB_ERR_INVALID_ARG, "plan is NULL");
        if (bytes < 0) fail(PBB_ERR_INVALID_ARG, "bytes must be >= 0");
        plan->ws_limit = (size_t)bytes;
    });
}

PBB_API pbb_status pbb_plan_timings(pbb_plan *plan, double *ms)
{
    return guarded([&] {
        if (!plan || !ms) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (!plan->events_made) fail(PBB_ERR_INVALID_ARG, "timing is not enabled on this plan");
        DeviceGuard g(plan->device);
        check(hipEventSynchronize(plan->ev[PBB_NUM_PHASES - 1]), "hipEventSynchronize");
        for (int p = 0; p < PBB_NUM_PHASES - 1; p++) {
            float t = 0;
            check(hipEventElapsedTime(&t, plan->ev[p], plan->ev[p + 1]), "hipEventElapsedTime");
            ms[p] = t;
        }
        float tot = 0;
        check(hipEventElapsedTime(&tot, plan->ev[0], plan->ev[PBB_NUM_PHASES - 1]), "hipEventElapsedTime");
        ms[PBB_NUM_PHASES - 1] = tot;
    });
}

// Which of the plan's sorted points carry the weights and which receive the
// outputs: weights w[c*in_n + j] belong to the points of input index
// in_lo + j, outputs out[c*out_n + i] to those of index out_lo + i.  One
// point set: both ranges are all N points.  Separate sets: the FMM (o =
// p->ops) takes weights on the sources and writes the targets; the
// transposed FMM (the adjoint operators) the other way round.
struct PointIo {
    int64_t in_lo, in_n, out_lo, out_n;
};
static PointIo point_io(const pbb_plan *p, const pbb_ops *o)
{
    if (!p->sep) return {0, p->N, 0, p->N};
    if (o == p->ops) return {p->nt, p->ns, 0, p->nt};
    return {0, p->nt, p->nt, p->ns};
}

static void matvec_run(pbb_plan *p, pbb_ops *o, const Workspace &W, const double *w, int ncols, double *out,
                       hipStream_t st, int prec, bool build);

// One matvec of the plan's operator set o in its own workspace.
// Columns are independent: when the workspace of all ncols columns does not
// fit beside what the device already holds, the matvec runs over chunks of
// columns, each a complete FMM pass (the same sums, so the same results).
// A pass that keeps its expansions for the caller (out == nullptr) is not
// split.
static void matvec_own(pbb_plan *p, pbb_ops *o, const double *w, int ncols, double *out, hipStream_t st, int prec)
{
    const bool mixed = prec != PBB_PREC_FP64;
    const PointIo io = point_io(p, o);
    int chunk = ncols;
    if (out) {
        Workspace tmp;
        size_t need = layout_workspace(*p, chunk, mixed, nullptr, tmp);
        if (need > p->ws_bytes || (p->ws_limit && need > p->ws_limit)) {
            size_t free_b = 0, total_b = 0;
            check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            // The workspace stays with the plan after the call, so leave the
            // caller a sixteenth of the device (at least 1 GiB) for its own
            // allocations.
            const size_t margin = std::max((size_t)(1ull << 30), total_b / 16);
            size_t room = p->ws_bytes + free_b - std::min(free_b, margin);
            if (p->ws_limit) room = std::min(room, p->ws_limit);
            while (chunk > 1 && need > room) {
                chunk = (chunk + 1) / 2;
                need = layout_workspace(*p, chunk, mixed, nullptr, tmp);
            }
        }
    }
    for (int c0 = 0; c0 < ncols; c0 += chunk) {
        const int cn = std::min(chunk, ncols - c0);
        p->ensure_workspace(cn, mixed);
        matvec_run(p, o, p->W, w + (size_t)c0 * io.in_n, cn, out ? out + (size_t)c0 * io.out_n : nullptr, st, prec,
                   true);
    }
}

// Relative difference ||a - b|| / ||b|| of two device FP64 arrays of n
// values, reduced on the device (part: 2 * 1024 doubles) and summed here.
static double device_rel_diff(const double *a, const double *b, int64_t n, double *part, hipStream_t st)
{
    constexpr int nblk = 1024;
    dev::launch_diff_norms(a, b, n, part, nblk, st);
    double h[2 * nblk];
    check(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, st), "copy norms");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    double d2 = 0, b2 = 0;
    for (int i = 0; i < nblk; i++) {
        d2 += h[2 * i];
        b2 += h[2 * i + 1];
    }
    return b2 > 0 ? std::sqrt(d2 / b2) : (d2 > 0 ? INFINITY : 0.0);
}

// The M2L tile autotuner.  max_rel_diff > 0 (mixed modes): a candidate is
// kept only if its matvec differs from the FP64 matvec on the same weights
// by at most max_rel_diff (relative 2-norm); when none does, the most
// accurate one is kept.  achieved: that candidate's difference (0 without
// a bound).
static void tune_m2l(pbb_plan *plan, const double *w_dev, int ncols, int precision, hipStream_t st,
                     double max_rel_diff, int32_t *best, double *achieved)
{
    if (!plan || !w_dev) fail(PBB_ERR_INVALID_ARG, "NULL argument");
    if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
    if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
        fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
    if (!(max_rel_diff >= 0) || !std::isfinite(max_rel_diff))
        fail(PBB_ERR_INVALID_ARG, "max_rel_diff must be finite and >= 0");
    const bool mixed = precision != PBB_PREC_FP64;
    const bool bounded = mixed && max_rel_diff > 0;
    DeviceGuard g(plan->device);
    const bool was_timing = plan->timing;
    check(pbb_plan_set_timing(plan, 1) == PBB_OK ? hipSuccess : hipErrorUnknown, "enable timing");
    const int64_t nout = point_io(plan, plan->ops).out_n * (int64_t)ncols;
    double *scratch = nullptr, *ref = nullptr, *part = nullptr;
    // candidates {width, columns per box, flags}
    static const int c64[][3] = {{32, 32, 0}, {64, 32, 0}, {128, 32, 0}, {64, 32, 1}, {128, 32, 1}};
    // (the column-wave tiles of 128 and 256 columns suit the small ranks of
    // re-compressed deep levels, whose rows would leave row-wave waves idle)
    static const int cmx[][3] = {{128, 16, 5}, {128, 32, 5}, {128, 8, 5}, {256, 8, 5}, {128, 16, 4},
                                 {64, 16, 5}, {128, 32, 1}, {64, 16, 0}, {128, 16, 5 | PBB_M2L_DIR_GROUPS(4)},
                                 {64, 16, 13}, {64, 8, 13}, {64, 32, 13}, {128, 16, 1}, {256, 16, 1},
                                 {256, 32, 1}};
    const int(*cands)[3] = mixed ? cmx : c64;
    const int ncand = mixed ? (int)(sizeof(cmx) / sizeof(cmx[0])) : (int)(sizeof(c64) / sizeof(c64[0]));
    pbb_plan::M2LTile bestt{}, closest{};
    double best_ms = 0, best_err = 0, closest_err = INFINITY;
    const int cls = mixed ? 1 : 0, Lv = plan->Lv;
    std::vector<std::vector<float>> lev_ms(ncand, std::vector<float>(Lv + 1, 0.0f));
    std::vector<bool> eligible(ncand, false);
    for (auto &t : plan->m2l_level[cls]) t = pbb_plan::M2LTile{};
    if (!plan->lev_made) {
        for (auto &e : plan->lev_ev) check(hipEventCreate(&e), "hipEventCreate");
        plan->lev_made = true;
    }
    plan->lev_timing = true;
    try {
        scratch = dalloc<double>((size_t)nout, "tuning output");
        if (bounded) {
            ref = dalloc<double>((size_t)nout, "tuning reference");
            part = dalloc<double>(2 * 1024, "tuning norms");
            dev::launch_fill_zero(ref, nout, st);
            matvec_own(plan, plan->ops, w_dev, ncols, ref, st, PBB_PREC_FP64);
        }
        // one untimed matvec builds the operator splits and workspaces
        matvec_own(plan, plan->ops, w_dev, ncols, scratch, st, precision);
        for (int c = 0; c < ncand; c++) {
            plan->m2l_tile[mixed ? 1 : 0] = {cands[c][0], std::min(cands[c][1], ncols), cands[c][2]};
            if (bounded) dev::launch_fill_zero(scratch, nout, st);
            matvec_own(plan, plan->ops, w_dev, ncols, scratch, st, precision);
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            float t = 0;
            check(hipEventElapsedTime(&t, plan->ev[3], plan->ev[4]), "hipEventElapsedTime");
            for (int l = 2; l <= Lv; l++)
                check(hipEventElapsedTime(&lev_ms[c][l], plan->lev_ev[l], plan->lev_ev[l + 1]), "hipEventElapsedTime");
            const double err = bounded ? device_rel_diff(scratch, ref, nout, part, st) : 0.0;
            if (err < closest_err) {
                closest = plan->m2l_tile[mixed ? 1 : 0];
                closest_err = err;
            }
            if (bounded && !(err <= max_rel_diff)) continue;
            eligible[c] = true;
            if (bestt.width == 0 || t < best_ms) {
                bestt = plan->m2l_tile[mixed ? 1 : 0];
                best_ms = t;
                best_err = err;
            }
        }
        if (bestt.width == 0) {  // nothing within the bound: the most accurate
            bestt = closest;
            best_err = closest_err;
        }
        plan->m2l_tile[cls] = bestt;
        // per level, the fastest eligible candidate with the same
        // cross-product-sum choice (which decides the VT / U precision of
        // the whole matvec); kept if the combination is faster still, and,
        // with a bound, within it
        const int sacc = bestt.flags & PBB_M2L_SPLIT_ACC;
        pbb_plan::M2LTile lev[16];
        bool differs = false;
        int bi = -1;
        for (int c = 0; c < ncand; c++)
            if (cands[c][0] == bestt.width && std::min(cands[c][1], ncols) == bestt.cpt && cands[c][2] == bestt.flags)
                bi = c;
        for (int l = 2; l <= Lv && bi >= 0; l++) {
            int k = bi;
            for (int c = 0; c < ncand; c++)
                if (eligible[c] && (cands[c][2] & PBB_M2L_SPLIT_ACC) == sacc && lev_ms[c][l] < lev_ms[k][l]) k = c;
            lev[l] = {cands[k][0], std::min(cands[k][1], ncols), cands[k][2]};
            differs |= k != bi;
        }
        if (differs) {
            std::copy(std::begin(lev), std::end(lev), plan->m2l_level[cls]);
            if (bounded) dev::launch_fill_zero(scratch, nout, st);
            matvec_own(plan, plan->ops, w_dev, ncols, scratch, st, precision);
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            float t = 0;
            check(hipEventElapsedTime(&t, plan->ev[3], plan->ev[4]), "hipEventElapsedTime");
            const double err = bounded ? device_rel_diff(scratch, ref, nout, part, st) : 0.0;
            if (t < best_ms && (!bounded || err <= max_rel_diff)) best_err = err;
            else
                for (auto &x : plan->m2l_level[cls]) x = pbb_plan::M2LTile{};
        }
    } catch (...) {
        dfree(scratch);
        dfree(ref);
        dfree(part);
        plan->timing = was_timing;
        plan->lev_timing = false;
        throw;
    }
    dfree(scratch);
    dfree(ref);
    dfree(part);
    plan->timing = was_timing;
    plan->lev_timing = false;
    if (best) {
        best[0] = bestt.width;
        best[1] = bestt.cpt;
        best[2] = bestt.flags;
    }
    if (achieved) *achieved = best_err;
}

PBB_API pbb_status pbb_plan_tune_m2l(pbb_plan *plan, const double *w_dev, int32_t ncols, int32_t precision,
                                     void *stream, int32_t *best)
{
    return guarded([&] { tune_m2l(plan, w_dev, ncols, precision, (hipStream_t)stream, 0.0, best, nullptr); });
}

PBB_API pbb_status pbb_plan_tune_m2l_within(pbb_plan *plan, const double *w_dev, int32_t ncols, int32_t precision,
                                            double max_rel_diff, void *stream, int32_t *best, double *achieved)
{
    return guarded(
        [&] { tune_m2l(plan, w_dev, ncols, precision, (hipStream_t)stream, max_rel_diff, best, achieved); });
}

// One FMM application.  out == nullptr stops after the downward pass, leaving
// the leaf local expansions in W.S and the sorted weights in W.ws.
static void fill_levels(const pbb_plan *p, dev::LevelTab &t)
{
    t.nlev = p->Lv - 1;
    if (t.nlev > dev::MAX_FAR_LEVELS) fail(PBB_ERR_UNSUPPORTED, "tree too deep for the direct passes");
    for (int i = 0; i < t.nlev; i++) {
        const DevLevel &D = p->lv[i + 2];
        t.off[i] = D.off;
        t.pstart[i] = D.pstart;
        t.coord[i] = D.coord;
        t.ctab[i] = D.ctab;
        t.len[i] = D.len;
    }
    t.off[t.nlev] = p->Btot;
}

// PBB_M2L_SPLIT_ACC in force for precision prec (two-part modes only)
static bool split_acc_on(const pbb_plan *p, int prec)
{
    const pbb_plan::M2LTile &tmx = p->m2l_tile[1];
    return (prec == PBB_PREC_MIXED_F16X2 || prec == PBB_PREC_MIXED_BF16X2) && tmx.width &&
           (tmx.flags & PBB_M2L_WAVE_ROWS) && (tmx.flags & PBB_M2L_SPLIT_ACC);
}

// The split operators a matvec of p with operator set o in precision prec
// uses (nullptr where none): E_d in 16-bit parts for the mixed M2L, and the
// three-part BF16 VT / U when both have at most 224 rows (PBB_M2L_SPLIT_ACC
// keeps VT / U in FP64).  build: made on first use (allocates and
// synchronises the stream); otherwise they must exist already.
static void split_ops(const pbb_plan *p, pbb_ops *o, int prec, bool build, hipStream_t st, const SplitE *&SE,
                      const pbb_ops::SplitA *&SA)
{
    SE = nullptr;
    SA = nullptr;
    const HostOps &h = o->h;
    if (prec == PBB_PREC_FP64 && o->fp64_reduced && p->Lv >= 2) {
        std::lock_guard<std::mutex> lk(o->mu);
        if (!o->dm_E) {
            if (!build)
                fail(PBB_ERR_INVALID_ARG, "the re-compressed FP64 operators are not prepared yet: one "
                                          "pbb_plan_matvec prepares them");
            o->make_mixed_sets();
        }
        return;
    }
    if (prec == PBB_PREC_FP64 || p->Lv < 2) return;
    const bool want3 = !split_acc_on(p, prec) && h.rs <= 224 && h.n3 <= 224;
    if (build) {
        SE = &o->ensure_split(prec, st);
        if (want3) SA = o->ensure_split3(st);
        return;
    }
    std::lock_guard<std::mutex> lk(o->mu);
    if (!o->split[prec].d || (want3 && !o->vt3.d))
        fail(PBB_ERR_INVALID_ARG, "the operators of this precision are not split yet: one pbb_plan_matvec "
                                  "in this precision (or pbb_plan_tune_m2l) prepares them");
    SE = &o->split[prec];
    if (want3) SA = &o->vt3;
}

static void matvec_run(pbb_plan *p, pbb_ops *o, const Workspace &W, const double *w, int ncols, double *out,
                       hipStream_t st, int prec, bool build)
{
    const bool mixed = prec != PBB_PREC_FP64;
    const HostOps &h = o->h;
    const int n = h.n, n3 = h.n3, Lv = p->Lv;
    auto mark = [&](int k) {
        if (p->timing) check(hipEventRecord(p->ev[k], st), "hipEventRecord");
    };
    if (W.ncols != ncols || W.mixed != (mixed ? 1 : 0)) fail(PBB_ERR_INTERNAL, "workspace laid out for another call");
    const SplitE *SE = nullptr;
    const pbb_ops::SplitA *SA = nullptr;
    split_ops(p, o, prec, build, st, SE, SA);
    const bool split_acc = split_acc_on(p, prec);
    const pbb_plan::M2LTile &tmx = p->m2l_tile[1];
    const int ldP = mixed ? W.kpad : h.rs;
    if (mixed && h.rs > W.kpad) fail(PBB_ERR_INTERNAL, "proxy padding too small");
    mark(0);
    // orders 4..8 run the direct passes: P2M at every level instead of
    // P2M + M2M, and (with a near field to add into) L2P at every level
    // instead of L2L + L2P; both are exact rewritings (see p2m_levels)
    const bool direct = n >= 4 && n <= 8;
    const bool direct_down = direct && out != nullptr && Lv >= 2;
    // the mixed modes keep S / F in FP32 when the direct passes read them
    // (not with the FP64 VT / U of the cross-product-sum tile)
    const bool s32 = mixed && direct_down && !split_acc;
    // levels Lp..Lv run the proxy passes (P2M straight to P, L2P straight
    // from Pf, no S / F of the level): the deepest levels whose re-compressed
    // ranks are at most 32.  Not with points outside the root cube: their
    // extrapolated S_n terms are large and cancel, which the proxy passes'
    // FP32 rank vectors do not resolve (the expansion route sums S in FP64).
    int Lp = Lv + 1;
    if (s32 && !p->expansion_route && p->info.outside_points == 0 && n <= 8 && ldP % 4 == 0 &&
        ldP >= dev::PX_K && !o->mset.empty())
        while (Lp > 2 && o->mset[h.ops_set(Lp - 1)].kf <= dev::PX_K && o->mset[h.ops_set(Lp - 1)].ks <= dev::PX_K &&
               h.rf >= dev::PX_K)
            Lp--;
    // the M2L results of a proxy level keep rows of PX_K floats (128 bytes)
    // inside the level's own region of Pf (rows of r_f), for whole-row loads
    auto pf_level = [&](int l, int &ldf) -> float * {
        if (l < Lp) {
            ldf = h.rf;
            return W.Pf32;
        }
        ldf = dev::PX_K;
        return W.Pf32 + (size_t)p->lv[l].off * ncols * (h.rf - dev::PX_K);
    };
    // With proxy levels, the expansion levels 2..Le (Le = Lp - 1) take the
    // reference's structure: P2M straight from the points at Le only, M2M up
    // to level 2, L2L back down to Le, L2P at Le only.  The transfers are
    // exact (the direct passes' rewriting in reverse) and cost a few boxes'
    // worth of work instead of every point at every level.
    const bool hyb = direct_down && Lp <= Lv;
    const int Le = hyb ? Lp - 1 : Lv;
    const PointIo io = point_io(p, o);
    // the direct passes and the split near field read only the point-major
    // sorted weights wsT; W.ws (column-major sorted) serves the other paths
    // and, on the direct path, holds the transposed input on the way to wsT
    const bool ws_needed = !direct_down;  // (the split near field covers the direct orders 4..8)
    if (ws_needed) dev::launch_gather_weights(w, p->perm, p->N, ncols, W.ws, io.in_lo, io.in_n, st);
    if (direct && ws_needed) dev::launch_gather_weights_t(w, p->perm, p->N, ncols, W.wsT, io.in_lo, io.in_n, st);
    else if (direct) dev::launch_gather_rows_t(w, p->perm, p->N, ncols, W.ws, W.wsT, io.in_lo, io.in_n, st);
    const int cc_p2m = std::max(1, std::min({ncols, 64, 4096 / n3}));
    const int cc_m2m = std::max(1, std::min(ncols, 2048 / n3));
    if (Lv >= 2 && direct) {
        dev::UpArgs a{};
        a.n = n; a.ncols = ncols;
        fill_levels(p, a.lt);
        a.items = p->up_items;
        a.Tkz = o->d_Tkz; a.x = p->x; a.y = p->y; a.z = p->z; a.wsT = W.wsT; a.S = W.S;
        a.scratch = W.up; a.s32 = s32;
        // the work items and reductions are in level order: levels 2..Lp-1,
        // or only Le with the M2M above it
        const int i1 = Lp <= Lv ? p->lv[Lp].up_item0 : p->n_up_items;
        const int r1 = Lp <= Lv ? p->lv[Lp].up_red0 : p->n_up_reds;
        const int i0 = hyb ? (Le >= 2 ? p->lv[Le].up_item0 : i1) : 0;
        const int r0 = hyb ? (Le >= 2 ? p->lv[Le].up_red0 : r1) : 0;
        a.items = p->up_items + i0;
        dev::UpReduce r{};
        r.ncols = ncols; r.n3 = n3; r.reds = p->up_reds + r0; r.nred = r1 - r0;
        r.scratch = W.up; r.S = W.S; r.s32 = s32;
        dev::launch_p2m_levels(a, i1 - i0, r, st);
    } else if (Lv >= 2) {
        const DevLevel &LL = p->lv[Lv];
        dev::P2MArgs a{};
        a.n = n; a.ncols = ncols; a.cc = cc_p2m; a.N = p->N; a.off = LL.off; a.len = LL.len;
        a.leaf_start = p->leaf_start; a.coord = LL.coord; a.ctab = LL.ctab; a.Tkz = o->d_Tkz;
        a.x = p->x; a.y = p->y; a.z = p->z; a.ws = W.ws; a.S = W.S;
        dev::launch_p2m(a, (int)LL.nb, st);
    }
    mark(1);
    for (int l = (direct ? Le : Lv) - 1; l >= 2 && (!direct || hyb); l--) {
        dev::M2MArgs a{};
        a.n = n; a.ncols = ncols; a.cc = cc_m2m;
        a.off_parent = p->lv[l].off; a.off_child = p->lv[l + 1].off;
        a.child = p->lv[l].child; a.Cw = o->d_Cw; a.Sparent = W.S; a.Schild = W.S; a.s32 = s32;
        dev::launch_m2m(a, (int)p->lv[l].nb, st);
    }
    mark(2);
    // the mixed modes (and FP64 with fp64_reduced) run each level with its
    // re-compressed operator set (kf x ks, pbb_ops::mset); P and Pf keep
    // their strides
    const bool reduced = mixed || o->fp64_reduced;
    auto mset = [&](int l) -> const MixedSet & { return o->mset[h.ops_set(l)]; };
    // FP64 keeps a set's own factors where its padded ks would pass the
    // FP64 proxy stride r_s (a set that was not re-compressed)
    auto red = [&](int l) { return mixed || (reduced && mset(l).ks <= h.rs); };
    // f16 x2 sets of ks <= 16 (stored padded to one 32-deep chunk): the M2L
    // runs 16-deep steps on proxies pre-split to 16 + 16 parts (not with the
    // cross-product-sum tile, which keeps its own kernel for its accuracy)
    auto k16 = [&](int l) {
        return prec == PBB_PREC_MIXED_F16X2 && !split_acc && SE && mset(l).ks_used <= 16 && mset(l).kf <= 32 &&
               SE->nch[h.ops_set(l)] == 1;
    };
    // the proxies of a proxy level keep compact rows (16 floats where the M2L
    // takes 16-deep steps, else 32) inside the level's own region of P (rows
    // of ldP), so that a box's proxies fill whole cache lines
    auto p_level = [&](int l, int &ldp) -> float * {
        if (l < Lp) {
            ldp = ldP;
            return W.P32;
        }
        ldp = k16(l) ? 16 : dev::PX_K;
        return W.P32 + (size_t)p->lv[l].off * ncols * (ldP - ldp);
    };
    for (int l = 2; l <= Lv; l++) {
        const DevLevel &D = p->lv[l];
        const int sl = h.ops_set(l);
        dev::GemmArgs g{};
        g.M = h.rs; g.K = n3; g.ldb = n3; g.ldc = ldP; g.ncol = D.nb * ncols; g.alpha = 1.0;
        g.A = o->d_VTw + (size_t)sl * h.rs * n3;
        if (red(l)) {
            g.M = mset(l).ks;
            g.A = o->dm_VTw + mset(l).vt_off;
        }
        g.b_f32 = s32 ? 1 : 0;
        g.B = s32 ? (const void *)((float *)W.S + (size_t)D.off * ncols * n3)
                  : (const void *)(W.S + (size_t)D.off * ncols * n3);
        if (mixed) {
            g.c_f32 = 1;
            g.C = W.P32 + (size_t)D.off * ncols * ldP;
        } else {
            g.C = W.P + (size_t)D.off * ncols * h.rs;
        }
        if (l >= Lp) {  // proxies straight from the points (rows past ks are zero)
            dev::ProxyArgs x{};
            x.ncols = ncols; x.nb = (int)D.nb; x.rank = mset(l).ks; x.off = D.off; x.len = D.len;
            x.pstart = D.pstart; x.coord = D.coord; x.ctab = D.ctab; x.Tkz = o->d_Tkz;
            x.x = p->x; x.y = p->y; x.z = p->z; x.wsT = W.wsT;
            x.ob = o->dm_px + o->px_off(sl, 0); x.npts = p->N;
            x.P = p_level(l, x.ldp);
            if (prec == PBB_PREC_MIXED_F16X2) {  // the proxy pass also finds the level's max |P|
                dev::launch_zero_u32(W.xmax + l, 1, st);
                x.xmax = W.xmax + l;
            }
            dev::launch_p2m_proxy(x, n, p->N, st);
        } else if (SA) {  // three-part BF16 matrix cores (P is FP32 in the mixed modes)
            dev::GemmSplitArgs q{};
            q.M = g.M; q.K = g.K; q.ldb = g.ldb; q.ldc = g.ldc; q.b_f32 = g.b_f32; q.c_f32 = 1; q.nch = SA->nch[sl];
            q.ncol = g.ncol; q.alpha = 1.0; q.As = (const uint16_t *)SA->d + SA->off[sl];
            q.B = g.B; q.C = g.C;
            dev::launch_gemm_split3(q, st);
        } else {
            dev::launch_gemm(g, st);
        }
        if (mixed) {
            // the k chunks the M2L reads: the GEMM writes every row of them
            // (rows past ks come from zero rows of the split operator)
            const int used = 32 * SE->nch[sl];
            if (used > ldP) fail(PBB_ERR_INTERNAL, "proxy chunks past the padding");
            int ldl = 0;
            float *Pl = p_level(l, ldl) + (size_t)D.off * ncols * ldl;  // this level's first row
            if (prec == PBB_PREC_MIXED_F16X2 && l < Lp) {
                dev::launch_zero_u32(W.xmax + l, 1, st);
                dev::launch_absmax_rows(Pl, D.nb * ncols, ldl, used, W.xmax + l, st);
            }
            if (prec == PBB_PREC_MIXED_F16X2 || prec == PBB_PREC_MIXED_BF16X2)
                dev::launch_split_p2(prec, Pl, D.nb * ncols, ldl,
                                     k16(l) ? 1 : used / 32, prec == PBB_PREC_MIXED_F16X2 ? W.xmax + l : nullptr,
                                     st, k16(l) ? 16 : 32);
        }
    }
    mark(3);
    // M2L tiles: W columns = TB target boxes x cpt columns.
    // FP64: W = 128 (64 for 17..31 columns), matrix-core bound, so few boxes
    // per tile keep the share of empty (box, direction) slots low.
    // Mixed: the row-wave kernel, W = 128 (8 boxes x 16 columns), chunk-major
    // steps, two workgroups per CU.
    auto lev_mark = [&](int l) {
        if (p->timing && p->lev_timing) check(hipEventRecord(p->lev_ev[l], st), "hipEventRecord");
    };
    for (int l = 2; l <= Lv && mixed; l++) {
        lev_mark(l);
        const DevLevel &D = p->lv[l];
        const pbb_plan::M2LTile &tl = p->m2l_level[1][l].width ? p->m2l_level[1][l] : tmx;
        const int cptm = tl.width ? std::min(ncols, tl.cpt) : std::min(ncols, 16);
        const int nchunks_m = (ncols + cptm - 1) / cptm;
        const int width_m = tl.width ? tl.width : 128;
        const int groups_m = tl.width ? std::max(1, (tl.flags >> 8) & 0xff) : 1;
        const int sl = h.ops_set(l);
        dev::M2LMixArgs a{};
        a.rf = mset(l).kf; a.rs = mset(l).ks;
        a.ncols = ncols; a.TB = std::max(1, std::min(16, width_m / cptm)); a.cpt = cptm;
        a.chunk_major = tl.width ? (tl.flags & PBB_M2L_CHUNK_MAJOR) : 1;
        a.row_split = tl.width ? (tl.flags & PBB_M2L_ROW_SPLIT) != 0 : 0;
        a.wave_rows = tl.width ? (tl.flags & PBB_M2L_WAVE_ROWS) != 0 : 1;
        if (a.rf > 224) {  // the column-wave tile takes at most 224 rows
            a.wave_rows = 1;
            a.row_split = 0;
        }
        if (a.wave_rows) a.TB = std::max(1, std::min(a.TB, dev::m2l_rows_max_width(prec, SE->R[sl]) / cptm));
        a.split_acc = split_acc;
        if (split_acc) a.TB = std::max(1, std::min(a.TB, 64 / cptm));
        a.k16 = k16(l) ? 1 : 0;
        if (a.k16) {  // the column-wave kernel with all rows
            a.wave_rows = 0;
            a.row_split = 0;
            a.split_acc = 0;
            a.TB = std::max(1, std::min(16, width_m / cptm));
        }
        a.side = 1 << l;
        a.nch = SE->nch[sl]; a.escale = SE->escale[sl]; a.off = D.off; a.R = SE->R[sl];
        a.tile_prefix[0] = 0;
        a.class_start[0] = 0;
        for (int c = 0; c < 8; c++) {
            a.class_start[c + 1] = a.class_start[c] + D.class_count[c];
            a.tile_prefix[c + 1] = a.tile_prefix[c] + (int)((D.class_count[c] + a.TB - 1) / a.TB);
        }
        a.order = D.order; a.coord = D.coord; a.map = D.map;
        a.Es = (const uint16_t *)SE->d + SE->off[sl];
        a.P = p_level(l, a.kpad); a.Pf = pf_level(l, a.ldf);
        a.xmax = prec == PBB_PREC_MIXED_F16X2 ? W.xmax + l : nullptr;
        // direction groups only where the launch fills the GPU many times over
        const int64_t wgs = (int64_t)a.tile_prefix[8] * nchunks_m * (a.row_split ? 2 : 1);
        const int ng = wgs >= 2048 ? groups_m : 1, per = (316 + ng - 1) / ng;
        for (int g = 0; g < ng; g++) {
            a.dlo = g * per;
            a.dhi = std::min(316, (g + 1) * per);
            a.accumulate = g > 0;
            dev::launch_m2l_mixed(prec, a, a.tile_prefix[8], nchunks_m, st);
        }
    }
    for (int l = 2; l <= Lv && !mixed; l++) {
        lev_mark(l);
        const DevLevel &D = p->lv[l];
        const pbb_plan::M2LTile &t64 = p->m2l_level[0][l].width ? p->m2l_level[0][l] : p->m2l_tile[0];
        const int cpt = t64.width ? std::min(ncols, t64.cpt) : std::min(ncols, 64);
        const int width = t64.width ? t64.width : ncols >= 32 || ncols <= 16 ? 128 : 64;
        const int TB = std::max(1, std::min(16, width / cpt));
        const int nchunks = (ncols + cpt - 1) / cpt;
        dev::M2LArgs a{};
        a.rf = h.rf; a.rs = h.rs; a.ldp = h.rs; a.ldf = h.rf;
        a.ncols = ncols; a.TB = TB; a.cpt = cpt; a.side = 1 << l; a.off = D.off;
        a.chunk_major = t64.width ? (t64.flags & PBB_M2L_CHUNK_MAJOR) : 1;
        a.tile_prefix[0] = 0;
        a.class_start[0] = 0;
        for (int c = 0; c < 8; c++) {
            a.class_start[c + 1] = a.class_start[c] + D.class_count[c];
            a.tile_prefix[c + 1] = a.tile_prefix[c] + (int)((D.class_count[c] + TB - 1) / TB);
        }
        a.order = D.order; a.coord = D.coord; a.map = D.map;
        a.E = o->d_E + (size_t)h.ops_set(l) * 316 * h.rf * h.rs;
        if (red(l)) {
            a.rf = mset(l).kf;
            a.rs = mset(l).ks;
            a.E = o->dm_E + mset(l).e_off;
        }
        a.P = W.P; a.Pf = W.Pf;
        dev::launch_m2l(a, a.tile_prefix[8], nchunks, st);
    }
    if (Lv >= 2) lev_mark(Lv + 1);
    mark(4);
    // local expansions F_l = scale_l * Uw Pf_l overwrite S_l (no longer needed);
    // the proxy-pass levels evaluate Pf_l at their points instead
    for (int l = 2; l <= Lv && l < Lp; l++) {
        const DevLevel &D = p->lv[l];
        const int sl = h.ops_set(l);
        dev::GemmArgs g{};
        g.M = n3; g.K = h.rf; g.ldb = h.rf; g.ldc = n3; g.ncol = D.nb * ncols;
        g.alpha = h.homog ? std::pow(1.0 / D.len, h.homogen_flipped) : 1.0;  // compute.hpp:433-435
        g.A = o->d_Uw + (size_t)sl * n3 * h.rf;
        if (red(l)) {
            g.K = mset(l).kf;
            g.A = o->dm_Uw + mset(l).u_off;
        }
        if (mixed) {
            g.b_f32 = 1;
            g.B = W.Pf32 + (size_t)D.off * ncols * h.rf;
        } else {
            g.B = W.Pf + (size_t)D.off * ncols * h.rf;
        }
        g.c_f32 = s32 ? 1 : 0;
        g.C = s32 ? (void *)((float *)W.S + (size_t)D.off * ncols * n3) : (void *)(W.S + (size_t)D.off * ncols * n3);
        if (SA) {  // three-part BF16 matrix cores (Pf is FP32 in the mixed modes)
            const auto &U3 = o->u3;
            dev::GemmSplitArgs q{};
            q.M = g.M; q.K = g.K; q.ldb = g.ldb; q.ldc = g.ldc; q.b_f32 = 1; q.c_f32 = g.c_f32; q.nch = U3.nch[sl];
            q.ncol = g.ncol; q.alpha = g.alpha; q.As = (const uint16_t *)U3.d + U3.off[sl];
            q.B = g.B; q.C = g.C;
            dev::launch_gemm_split3(q, st);
        } else {
            dev::launch_gemm(g, st);
        }
    }
    mark(5);
    // downward pass, top-down: F_l += L2L(F_{l-1}), F_{l-1} already complete;
    // the direct pass evaluates every level's F at the points instead
    for (int l = 3; l <= Le && hyb; l++) {  // F_l += L2L(F_{l-1}) down to Le
        dev::M2MArgs a{};
        a.n = n; a.ncols = ncols; a.cc = cc_m2m;
        a.off_parent = p->lv[l - 1].off; a.off_child = p->lv[l].off;
        a.child = p->lv[l - 1].child; a.Cw = o->d_Cw; a.Sparent = W.S; a.Schild = W.S; a.s32 = s32;
        dev::launch_l2l(a, (int)p->lv[l - 1].nb, st);
    }
    for (int l = 2; l <= Lv && direct_down; l++) {
        const DevLevel &D = p->lv[l];
        if (hyb && l < Le) continue;  // reaches the points through L2L into level Le
        if (l >= Lp) {
            dev::ProxyArgs x{};
            x.ncols = ncols; x.nb = (int)D.nb; x.accumulate = l > 2; x.rank = mset(l).kf;
            x.off = D.off; x.len = D.len;
            x.alpha = h.homog ? std::pow(1.0 / D.len, h.homogen_flipped) : 1.0;  // as the U product's scale
            x.pstart = D.pstart; x.coord = D.coord; x.ctab = D.ctab; x.Tkz = o->d_Tkz;
            x.x = p->x; x.y = p->y; x.z = p->z;
            x.ob = o->dm_px + o->px_off(h.ops_set(l), 1); x.lp = W.lp; x.npts = p->N;
            x.Pf = pf_level(l, x.ldf);
            x.pf_vec = ((uintptr_t)(x.Pf + (size_t)D.off * ncols * x.ldf) & 15) == 0;
            dev::launch_l2p_proxy(x, n, p->N, st);
            continue;
        }
        dev::DownArgs a{};
        a.n = n; a.ncols = ncols; a.accumulate = hyb ? 0 : l > 2; a.off = D.off; a.len = D.len;
        a.seg_box = D.seg_box; a.seg_p0 = D.seg_p0; a.pstart = D.pstart; a.coord = D.coord; a.ctab = D.ctab;
        a.Tkz = o->d_Tkz; a.x = p->x; a.y = p->y; a.z = p->z; a.F = W.S; a.lp = W.lp; a.f32 = s32;
        dev::launch_l2p_level(a, D.nseg, st);
    }
    for (int l = 3; l <= Lv && !direct_down; l++) {
        dev::M2MArgs a{};
        a.n = n; a.ncols = ncols; a.cc = cc_m2m;
        a.off_parent = p->lv[l - 1].off; a.off_child = p->lv[l].off;
        a.child = p->lv[l - 1].child; a.Cw = o->d_Cw; a.Sparent = W.S; a.Schild = W.S;
        dev::launch_l2l(a, (int)p->lv[l - 1].nb, st);
    }
    mark(6);
    if (out) {
        const DevLevel &LL = p->lv[Lv];
        dev::NearArgs a{};
        a.n = n; a.ncols = ncols; a.side = 1 << Lv; a.N = p->N; a.off = LL.off; a.len = LL.len;
        std::memcpy(a.kp.p, h.kernel.param, sizeof(a.kp.p));
        a.leaf_start = p->leaf_start; a.coord = LL.coord; a.map = LL.map; a.perm = p->perm;
        a.ctab = LL.ctab; a.Tkz = o->d_Tkz; a.x = p->x; a.y = p->y; a.z = p->z; a.ws = W.ws;
        a.F = Lv >= 2 ? W.S : nullptr;
        a.out = out;
        a.out_lo = io.out_lo;
        a.out_n = io.out_n;
        a.wsT = W.wsT;
        a.lp = direct_down ? W.lp : nullptr;
        a.leaf_class = p->leaf_class;
        std::copy(std::begin(p->leaf_class_off), std::end(p->leaf_class_off), a.leaf_class_off);
        a.kf32 = mixed ? 1 : 0;
        if (!(direct && dev::launch_near_split(a, h.kernel.id, (int)LL.nb, st)))
            dev::launch_near(a, h.kernel.id, (int)LL.nb, st);
    }
    mark(7);
}

PBB_API pbb_status pbb_plan_matvec(pbb_plan *plan, const double *w_dev, int32_t ncols, double *out_dev,
                                   int32_t precision, void *stream)
{
    return guarded([&] {
        if (!plan || !w_dev || !out_dev) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
        if ((int64_t)ncols * plan->N >= (1LL << 40)) fail(PBB_ERR_INVALID_ARG, "ncols * npoints too large");
        if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
            fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
        DeviceGuard g(plan->device);
        matvec_own(plan, plan->ops, w_dev, ncols, out_dev, (hipStream_t)stream, precision);
    });
}

}  // extern "C"

// ---- the matvec that allocates nothing ------------------------------------
// pbb_plan_matvec_noalloc lays its workspace out in this file-scope array
// instead of device memory the plan allocates.  Every GPU the library runs
// on holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel or a copy: the name is a host-side stand-in, so
// the copy's device address is looked up once per device (with that device
// current) and kept in g_scratch_ptr.
static __device__ __align__(256) char scratch[PBB_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device dev, which must be current.
char *scratch_ptr(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) fail(PBB_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            fail(PBB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = (char *)p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

PBB_API pbb_status pbb_plan_matvec_noalloc(pbb_plan *plan, const double *w_dev, int32_t ncols, double *out_dev,
                                           int32_t precision, void *stream)
{
    return guarded([&] {
        if (!plan || !w_dev || !out_dev) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
        if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
            fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
        const bool mixed = precision != PBB_PREC_FP64;
        Workspace W;
        const size_t need = layout_workspace(*plan, ncols, mixed, nullptr, W);
        if (need > (size_t)PBB_NOALLOC_SCRATCH_BYTES)
            fail(PBB_ERR_INVALID_ARG, "this matvec needs " + std::to_string(need) +
                                          " bytes of workspace; the allocation-free matvec has " +
                                          std::to_string((size_t)PBB_NOALLOC_SCRATCH_BYTES));
        DeviceGuard g(plan->device);
        char *base = scratch_ptr(plan->device);
        layout_workspace(*plan, ncols, mixed, base, W);
        // build = false: the split operators of a mixed mode must exist already
        matvec_run(plan, plan->ops, W, w_dev, ncols, out_dev, (hipStream_t)stream, precision, false);
    });
}

PBB_API pbb_status pbb_plan_matvec_transpose(pbb_plan *plan, const double *w_dev, int32_t ncols, double *out_dev,
                                             int32_t precision, void *stream)
{
    return guarded([&] {
        if (!plan || !w_dev || !out_dev) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
        if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
            fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
        if (plan->ops->h.kernel.symmetry != 1)
            fail(PBB_ERR_UNSUPPORTED, "the transposed near field needs a symmetric kernel");
        pbb_ops *adj = plan->ops->adjoint();
        DeviceGuard g(plan->device);
        adj->upload(plan->device);
        matvec_own(plan, adj, w_dev, ncols, out_dev, (hipStream_t)stream, precision);
    });
}

PBB_API pbb_status pbb_plan_grad_positions(pbb_plan *plan, const double *sigma_dev, const double *psi_dev,
                                           int32_t ncols, double *grad_dev, int32_t precision, void *stream)
{
    return guarded([&] {
        if (!plan || !sigma_dev || !psi_dev || !grad_dev) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
        if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
            fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
        if (plan->ops->h.kernel.symmetry != 1)
            fail(PBB_ERR_UNSUPPORTED, "position gradients need a symmetric kernel");
        if (plan->sep) fail(PBB_ERR_UNSUPPORTED, "position gradients need a plan whose sources are its targets");
        pbb_ops *adj = plan->ops->adjoint();
        DeviceGuard g(plan->device);
        adj->upload(plan->device);
        hipStream_t st = (hipStream_t)stream;
        pbb_plan *p = plan;
        const HostOps &h = p->ops->h;
        const int Lv = p->Lv, n3 = h.n3;
        const DevLevel &LL = p->lv[Lv];
        if (ncols > p->g_ncols) {
            dfree(p->gF); dfree(p->gU);
            p->gF = p->gU = nullptr;
            p->g_ncols = 0;
            if (Lv >= 2) p->gF = dalloc<double>((size_t)LL.nb * ncols * n3, "saved leaf expansions");
            p->gU = dalloc<double>((size_t)p->N * ncols, "sorted weights (gradient)");
            p->g_ncols = ncols;
        }
        const size_t leafF = (size_t)LL.nb * ncols * n3;
        // forward FMM with sigma: leaf expansions F and sorted sigma
        matvec_own(p, p->ops, sigma_dev, ncols, nullptr, st, precision);
        const Workspace &W = p->W;
        if (Lv >= 2)
            check(hipMemcpyAsync(p->gF, W.S + (size_t)LL.off * ncols * n3, leafF * sizeof(double),
                                 hipMemcpyDeviceToDevice, st), "save leaf expansions");
        check(hipMemcpyAsync(p->gU, W.ws, (size_t)p->N * ncols * sizeof(double), hipMemcpyDeviceToDevice, st),
              "save sorted weights");
        // transposed FMM with psi: its leaf expansions and sorted psi
        matvec_own(p, adj, psi_dev, ncols, nullptr, st, precision);
        dev::NearGradArgs a{};
        a.n = h.n; a.ncols = ncols; a.side = 1 << Lv; a.N = p->N; a.len = LL.len;
        std::memcpy(a.kp.p, h.kernel.param, sizeof(a.kp.p));
        a.leaf_start = p->leaf_start; a.coord = LL.coord; a.map = LL.map; a.perm = p->perm;
        a.ctab = LL.ctab; a.Tkz = p->ops->d_Tkz; a.x = p->x; a.y = p->y; a.z = p->z;
        a.grad = grad_dev;
        // target side: d/dx_i of (A sigma)_i weighted by psi_i
        a.v = W.ws; a.u = p->gU; a.F = Lv >= 2 ? p->gF : nullptr;
        dev::launch_near_grad(a, h.kernel.id, (int)LL.nb, st);
        // source side: d/dx_j of (A^T psi)_j weighted by sigma_j
        a.v = p->gU; a.u = W.ws; a.F = Lv >= 2 ? W.S + (size_t)LL.off * ncols * n3 : nullptr;
        dev::launch_near_grad(a, h.kernel.id, (int)LL.nb, st);
    });
}

PBB_API void pbb_plan_destroy(pbb_plan *plan) { delete plan; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many plans' matvecs captured once into a HIP graph and replayed.
struct pbb_batch {
    int device = -1;
    std::vector<char *> blocks;     // the batch's own workspaces, one per distinct plan
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> side;  // capture branches
    std::vector<hipEvent_t> ev;     // fork and join points
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~pbb_batch()
    {
        if (device < 0) return;
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        if (stream) (void)hipStreamDestroy(stream);
        for (char *b : blocks) dfree(b);
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

extern "C" {

// Captures the entries' matvecs into one HIP graph (see pbb_batch_create).
static void batch_create(const pbb_batch_entry *E, int nplans, int ncols, int precision, pbb_batch **out)
{
    if (!out) fail(PBB_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!E || nplans < 1) fail(PBB_ERR_INVALID_ARG, "empty batch");
    if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
    if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
        fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
    const bool mixed = precision != PBB_PREC_FP64;
    std::vector<pbb_plan *> plans(nplans);
    std::vector<const double *> w_dev(nplans);
    std::vector<double *> out_dev(nplans);
    for (int i = 0; i < nplans; i++) {
        plans[i] = E[i].plan;
        w_dev[i] = E[i].w_dev;
        out_dev[i] = E[i].out_dev;
        if (!plans[i] || !w_dev[i] || !out_dev[i]) fail(PBB_ERR_INVALID_ARG, "NULL plan or buffer in batch");
        if (E[i].flags & ~(PBB_BATCH_TRANSPOSE | PBB_BATCH_OVERWRITE))
            fail(PBB_ERR_INVALID_ARG, "unknown batch entry flags");
    }
    const int dev = plans[0]->device;
    for (int i = 0; i < nplans; i++) {
        if (plans[i]->device != dev) fail(PBB_ERR_INVALID_ARG, "batch plans must share one device");
        if ((E[i].flags & PBB_BATCH_TRANSPOSE) && plans[i]->ops->h.kernel.symmetry != 1)
            fail(PBB_ERR_UNSUPPORTED, "the transposed near field needs a symmetric kernel");
    }
    DeviceGuard g(dev);
    auto b = std::make_unique<pbb_batch>();
    b->device = dev;
    check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    // everything that allocates or synchronises happens before the capture:
    // the operator splits, and one workspace per distinct plan owned by the
    // batch (calls on the plans themselves never touch it).  The forward and
    // the transposed operator sets of a plan fit one workspace layout.
    std::vector<pbb_ops *> ops(nplans);
    std::vector<Workspace> wsp(nplans);
    for (int i = 0; i < nplans; i++) {
        pbb_plan *p = plans[i];
        ops[i] = (E[i].flags & PBB_BATCH_TRANSPOSE) ? p->ops->adjoint() : p->ops;
        ops[i]->upload(dev);
        const SplitE *se = nullptr;
        const pbb_ops::SplitA *sa = nullptr;
        split_ops(p, ops[i], precision, true, b->stream, se, sa);
        int first = i;
        for (int k = 0; k < i; k++)
            if (plans[k] == p) {
                first = k;
                break;
            }
        if (first < i) {
            wsp[i] = wsp[first];
            continue;
        }
        Workspace tmp;
        const size_t need = layout_workspace(*p, ncols, mixed, nullptr, tmp);
        b->blocks.push_back(dalloc<char>(need, "batch workspace"));
        layout_workspace(*p, ncols, mixed, b->blocks.back(), wsp[i]);
    }
    check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    // independent entries run as concurrent graph branches.  Entries
    // conflict when they share a plan (its workspace) or when one's input
    // or output overlaps another's output; conflicting entries (closed
    // transitively) go to one branch and run there in batch order.
    std::vector<int> branch(nplans, 0);
    int nbr = 1;
    {
        std::vector<int> root(nplans);
        for (int i = 0; i < nplans; i++) root[i] = i;
        std::function<int(int)> find = [&](int i) { return root[i] == i ? i : root[i] = find(root[i]); };
        auto overlap = [&](const void *a, int64_t na, const void *b2, int64_t nb2) {
            const char *x = (const char *)a, *y = (const char *)b2;
            return x < y + nb2 * (int64_t)sizeof(double) && y < x + na * (int64_t)sizeof(double);
        };
        std::vector<PointIo> io(nplans);
        for (int i = 0; i < nplans; i++) io[i] = point_io(plans[i], ops[i]);
        for (int i = 0; i < nplans; i++) {
            const int64_t wi = io[i].in_n * (int64_t)ncols, oi = io[i].out_n * (int64_t)ncols;
            for (int j = 0; j < i; j++) {
                const int64_t wj = io[j].in_n * (int64_t)ncols, oj = io[j].out_n * (int64_t)ncols;
                const bool c = plans[i] == plans[j] || overlap(out_dev[i], oi, out_dev[j], oj) ||
                               overlap(w_dev[i], wi, out_dev[j], oj) || overlap(w_dev[j], wj, out_dev[i], oi);
                if (c) root[find(i)] = find(j);
            }
        }
        const int maxbr = std::min(nplans, 4);
        std::vector<int> comp_branch(nplans, -1);
        int next = 0;
        for (int i = 0; i < nplans; i++) {
            const int r = find(i);
            if (comp_branch[r] < 0) comp_branch[r] = next++ % maxbr;
            branch[i] = comp_branch[r];
        }
        nbr = std::max(1, std::min(next, maxbr));
    }
    b->side.resize(nbr, nullptr);
    b->ev.resize(nbr + 1, nullptr);
    for (int k = 0; k < nbr; k++) check(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking), "hipStreamCreate");
    for (auto &e : b->ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    std::vector<bool> timing(nplans);
    for (int i = 0; i < nplans; i++) timing[i] = plans[i]->timing;
    for (int i = 0; i < nplans; i++) plans[i]->timing = false;
    check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        check(hipEventRecord(b->ev[nbr], b->stream), "fork");
        for (int k = 0; k < nbr; k++) check(hipStreamWaitEvent(b->side[k], b->ev[nbr], 0), "fork");
        for (int i = 0; i < nplans; i++) {
            hipStream_t s = b->side[branch[i]];
            if (E[i].flags & PBB_BATCH_OVERWRITE)
                dev::launch_fill_zero(out_dev[i], point_io(plans[i], ops[i]).out_n * (int64_t)ncols, s);
            matvec_run(plans[i], ops[i], wsp[i], w_dev[i], ncols, out_dev[i], s, precision, false);
        }
        for (int k = 0; k < nbr; k++) {
            check(hipEventRecord(b->ev[k], b->side[k]), "join");
            check(hipStreamWaitEvent(b->stream, b->ev[k], 0), "join");
        }
    } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(b->stream, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        for (int i = 0; i < nplans; i++) plans[i]->timing = timing[i];
        throw;
    }
    check(hipStreamEndCapture(b->stream, &b->graph), "hipStreamEndCapture");
    for (int i = 0; i < nplans; i++) plans[i]->timing = timing[i];
    check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    *out = b.release();
}

PBB_API pbb_status pbb_batch_create(pbb_plan *const *plans, int32_t nplans, const double *const *w_dev,
                                    double *const *out_dev, int32_t ncols, int32_t precision, int32_t flags,
                                    pbb_batch **out)
{
    return guarded([&] {
        if (out) *out = nullptr;
        if (!plans || !w_dev || !out_dev || nplans < 1) fail(PBB_ERR_INVALID_ARG, "empty batch");
        std::vector<pbb_batch_entry> E(nplans);
        for (int i = 0; i < nplans; i++) E[i] = pbb_batch_entry{plans[i], w_dev[i], out_dev[i], flags, 0};
        batch_create(E.data(), nplans, ncols, precision, out);
    });
}

PBB_API pbb_status pbb_batch_create_entries(const pbb_batch_entry *entries, int32_t nentries, int32_t ncols,
                                            int32_t precision, pbb_batch **out)
{
    return guarded([&] { batch_create(entries, nentries, ncols, precision, out); });
}

PBB_API pbb_status pbb_batch_launch(pbb_batch *batch, void *stream)
{
    return guarded([&] {
        if (!batch) fail(PBB_ERR_INVALID_ARG, "batch is NULL");
        DeviceGuard g(batch->device);
        check(hipGraphLaunch(batch->exec, (hipStream_t)stream), "hipGraphLaunch");
    });
}

PBB_API void pbb_batch_destroy(pbb_batch *batch) { delete batch; }

PBB_API pbb_status pbb_direct_sum(const pbb_kernel_desc *kernel, const double *tgt_dev, int64_t ntargets,
                                  const double *src_dev, int64_t nsources, const double *w_dev, int32_t ncols,
                                  double *out_dev, void *stream)
{
    return guarded([&] {
        if (!kernel || !tgt_dev || !src_dev || !w_dev || !out_dev) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ntargets < 0 || nsources < 0 || ncols < 1) fail(PBB_ERR_INVALID_ARG, "bad sizes");
        if (!kernel_id_valid(kernel->id) || kernel->id == PBB_KERNEL_POLY_XY)
            fail(PBB_ERR_UNSUPPORTED, "the kernel has no device functor");
        (void)current_device();
        hipStream_t st = (hipStream_t)stream;
        dev::DirectArgs a{};
        a.T = ntargets; a.Ns = nsources; a.ncols = ncols;
        int64_t tblocks = (ntargets + 63) / 64;
        a.nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(256, (4096 + tblocks - 1) / std::max<int64_t>(1, tblocks)));
        a.nsplit = (int)std::min<int64_t>(a.nsplit, std::max<int64_t>(1, nsources / 256));
        std::memcpy(a.kp.p, kernel->param, sizeof(a.kp.p));
        a.tgt = tgt_dev; a.src = src_dev; a.w = w_dev; a.out = out_dev;
        if (ntargets == 0) return;
        if (nsources == 0) {
            check(hipMemsetAsync(out_dev, 0, sizeof(double) * ntargets * ncols, st), "clear output");
            return;
        }
        a.part = dalloc<double>((size_t)a.nsplit * ncols * ntargets, "direct-sum partials");
        try {
            dev::launch_direct(a, kernel->id, st);
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
        } catch (...) {
            dfree(a.part);
            throw;
        }
        dfree(a.part);
    });
}

// One matvec from host buffers: a plan (one point set when same, else the
// targets followed by the sources), the weights uploaded, out += A w.
static pbb_status matvec_host(pbb_ops *ops, const double *target, int64_t nt, const double *source, int64_t ns,
                              bool same, const double *weight, int32_t ncols, double *output, int32_t precision)
{
    pbb_plan *plan = nullptr;
    pbb_status st = same ? pbb_plan_create(ops, target, nt, &plan)
                         : pbb_plan_create_ts(ops, target, nt, source, ns, 0, &plan);
    if (st != PBB_OK) return st;
    double *dw = nullptr, *dout = nullptr;
    st = guarded([&] {
        DeviceGuard g(plan->device);
        const size_t nw = (size_t)ns * ncols, no = (size_t)nt * ncols;
        dw = dalloc<double>(nw, "weights");
        dout = dalloc<double>(no, "output");
        check(hipMemcpyAsync(dw, weight, nw * sizeof(double), hipMemcpyHostToDevice, plan->stream), "copy weights");
        check(hipMemsetAsync(dout, 0, no * sizeof(double), plan->stream), "clear output");
        matvec_own(plan, plan->ops, dw, ncols, dout, plan->stream, precision);
        std::vector<double> res(no);
        check(hipMemcpyAsync(res.data(), dout, no * sizeof(double), hipMemcpyDeviceToHost, plan->stream),
              "copy output");
        check(hipStreamSynchronize(plan->stream), "hipStreamSynchronize");
        for (size_t i = 0; i < no; i++) output[i] += res[i];
    });
    {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(plan->device);
        dfree(dw);
        dfree(dout);
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    pbb_plan_destroy(plan);
    return st;
}

PBB_API pbb_status pbb_matvec(pbb_ops *ops, const double *target, const double *source, int64_t npoints,
                              const double *weight, int32_t ncols, double *output)
{
    bool same = true;
    pbb_status st = guarded([&] {
        if (!ops || !target || !source || !weight || !output) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
        if (npoints <= 0) fail(PBB_ERR_INVALID_ARG, "npoints must be positive");
        same = target == source || std::memcmp(target, source, sizeof(double) * 3 * (size_t)npoints) == 0;
    });
    if (st != PBB_OK) return st;
    return matvec_host(ops, target, npoints, source, npoints, same, weight, ncols, output, PBB_PREC_FP64);
}

PBB_API pbb_status pbb_matvec_ts(pbb_ops *ops, const double *target, int64_t ntargets, const double *source,
                                 int64_t nsources, const double *weight, int32_t ncols, double *output)
{
    pbb_status st = guarded([&] {
        if (!ops || !target || !source || !weight || !output) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
        if (ntargets <= 0 || nsources <= 0) fail(PBB_ERR_INVALID_ARG, "ntargets and nsources must be positive");
    });
    if (st != PBB_OK) return st;
    return matvec_host(ops, target, ntargets, source, nsources, false, weight, ncols, output, PBB_PREC_FP64);
}

PBB_API pbb_status pbb_matvec_prec(pbb_ops *ops, const double *target, int64_t ntargets, const double *source,
                                   int64_t nsources, const double *weight, int32_t ncols, double *output,
                                   int32_t precision)
{
    bool same = false;
    pbb_status st = guarded([&] {
        if (!ops || !target || !source || !weight || !output) fail(PBB_ERR_INVALID_ARG, "NULL argument");
        if (ncols < 1) fail(PBB_ERR_INVALID_ARG, "ncols must be >= 1");
        if (ntargets <= 0 || nsources <= 0) fail(PBB_ERR_INVALID_ARG, "ntargets and nsources must be positive");
        if (precision < PBB_PREC_FP64 || precision > PBB_PREC_MIXED_BF16X2)
            fail(PBB_ERR_INVALID_ARG, "unknown precision mode");
        same = ntargets == nsources &&
               (target == source || std::memcmp(target, source, sizeof(double) * 3 * (size_t)ntargets) == 0);
    });
    if (st != PBB_OK) return st;
    return matvec_host(ops, target, ntargets, source, nsources, same, weight, ncols, output, precision);
}

}  // extern "C"
