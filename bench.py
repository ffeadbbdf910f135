#!/usr/bin/env python
"""Benchmark of the black-box FMM matvec on one MI355X.

Workload (default): 4,000,000 points, a mixture of 8 isotropic Gaussians
(sigma 0.05, centres U[0.2,0.8]^3) plus 20% uniform background clipped to
the unit cube; weights N(0,1), 32 columns; Gaussian kernel exp(-r^2/0.3^2);
tree level 7, Chebyshev order 6, SVD eps 1e-8.  One step = one matvec
(upward pass, M2L, downward pass, near field) from device-resident weights
to device output on a prepared tree.  Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--precision f16x2|mixed|bf16x2|fp64]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "targets/s at N=4M nCols=32, rel-err<=1e-6 vs FP64: mixed MFMA vs FP64 GPU"
PEAK = {"fp64": ("mfma", 78.6, "TFLOP/s"), "mixed": ("mfma", 2500.0, "TFLOP/s"),
        "f16x2": ("mfma", 2500.0, "TFLOP/s"), "bf16x2": ("mfma", 2500.0, "TFLOP/s")}
M2L_SPLIT = {"mixed": "bf16x3", "f16x2": "f16x2", "bf16x2": "bf16x2"}


def dtype_label(precision, flags):
    """What each phase computes in.  Mixed modes: M2L on split 16-bit MFMA
    with FP32 sums (each level's re-compressed operator set; 16-deep steps
    where the kept source rank is at most 16), proxies P and M2L results Pf
    stored FP32; on the levels of rank <= 32 the proxy passes (rank vectors
    on three-part BF16 MFMA, FP64 box sums and Pf products); above them VT /
    U on three-part BF16 MFMA (FP64 with the cross-product-sum M2L tile),
    expansions S / F stored FP32 with FP64 arithmetic (P2M, M2M, L2L, L2P);
    the near field of dense leaves with FP32 kernel values and FP64 MFMA
    sums, the rest of the near field in FP64."""
    if precision == "fp64":
        return "f64"
    split_acc = precision in ("f16x2", "bf16x2") and flags is not None and (flags & 8)
    vtu = "f64 (P, Pf f32)" if split_acc else "bf16x3->f32 (P, Pf f32)"
    return (f"m2l {M2L_SPLIT[precision]}->f32{' (cross products summed apart)' if split_acc else ''}; "
            f"levels of rank<=32: proxy passes bf16x3->f32 rank vectors, f64 box sums; "
            f"other levels: vt/u {vtu}, p2m/m2m/l2l/l2p f64 (S, F f32); "
            f"p2p f64 (dense leaves: f32 kernel values, f64 mfma sums)")


# bytes per M2L launch (FETCH_SIZE x 2 + WRITE_SIZE x 1, calibrated on a plain
# copy), averaged over the 6 launches of one f16x2 matvec of the default
# workload (45.09 GB per matvec)
M2L_TRAFFIC_F16X2_4M_CLUSTERED = 7.516e9
TRAFFIC_SOURCE = "profiles/m2l_f16x2_4m_clustered_gaussian_traffic.txt"


def clustered_points(n, seed):
    rng = np.random.default_rng(seed)
    nb = int(round(0.2 * n))
    nc = n - nb
    centres = rng.uniform(0.2, 0.8, (8, 3))
    lab = rng.integers(0, 8, nc)
    pts = centres[lab] + 0.05 * rng.standard_normal((nc, 3))
    return np.clip(np.concatenate([pts, rng.random((nb, 3))]), 0.0, 1.0)


def make_kernel(name):
    import pbbfmm3d_amd as pbb
    return {"gaussian": pbb.gaussian(0.3), "laplacian": pbb.laplacian()}[name]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(torch, pbb, capi):
    """The oracle (plain-C restatement of the reference's CPU matvec,
    OpenMP on the host's threads) on two whole problems of the reference's
    CPU-runnable shapes, both 1/r: 1,000,000 uniform points at tree level 6,
    order 5, eps 1e-7, 8 columns (the drop-in configuration), and 20,000
    uniform points at level 4, order 4, eps 1e-5, one column.  The GPU FP64
    path runs the same million-point problem (same operators, weights and
    points), so the two rates compare like for like, and its output is
    checked against the oracle's over all targets."""
    from oracle import oracle as orc
    res = {}
    for tag, N, nc, lvl, n, eps in (("1m", 1_000_000, 8, 6, 5, 1e-7), ("20k", 20_000, 1, 4, 4, 1e-5)):
        rng = np.random.default_rng(5)
        xyz = rng.random((N, 3))
        w = rng.uniform(-1.0, 1.0, N * nc)
        ops = pbb.Operators(pbb.laplacian(), 1.0, lvl, n, eps)
        U, VT, E, _, _ = ops.export()
        oops = orc.Ops(orc.LAPLACIAN, np.zeros(4), 1, -1.0, 1.0, lvl, n, eps, ops.rank_f, ops.rank_s,
                       ops.nsets, U, VT, E)
        reps = 1 if N > 100_000 else 5
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ref = orc.matvec(oops, xyz, w, nc)
            times.append(time.perf_counter() - t0)
        dt = float(np.median(times))
        plan = pbb.Plan(ops, xyz)
        wd = torch.tensor(w.reshape(nc, N), device="cuda:0")
        out = torch.zeros_like(wd)
        plan.matvec(wd, out)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out.zero_()
        plan.matvec(wd, out)
        torch.cuda.synchronize()
        gt = time.perf_counter() - t0
        diff = float(np.linalg.norm(out.cpu().numpy().reshape(-1) - ref) / np.linalg.norm(ref))
        res[tag] = {"npoints": N, "ncols": nc, "tree_level": lvl, "order": n, "eps": eps,
                    "cpu_s": dt, "cpu_targets_per_s": N / dt, "gpu_fp64_s": gt, "gpu_fp64_targets_per_s": N / gt,
                    "gpu_over_cpu": dt / gt, "gpu_vs_oracle_rel_diff": diff}
        log(f"cpu baseline {tag}: oracle {dt:.2f} s, GPU FP64 {1e3 * gt:.1f} ms, rel diff {diff:.2e}")
        del plan, wd, out
    big = res["1m"]
    return {"value": big["cpu_targets_per_s"], "unit": "targets/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"the whole 1,000,000-point problem of the reference's drop-in configuration (1/r, uniform "
                      f"points, tree level 6, order 5, eps 1e-7, 8 columns; the headline workload itself needs "
                      f"~170 GB for the oracle), oracle matvec {big['cpu_s']:.1f} s on {orc.num_threads()} "
                      f"threads; the GPU FP64 path takes {1e3 * big['gpu_fp64_s']:.1f} ms on it "
                      f"({big['gpu_over_cpu']:.0f}x)",
            "problems": res}


def phase_names(order):
    """The library's eight phase slots under the names of what runs in
    them.  Orders 4..8 run the direct passes: P2M at every level replaces
    P2M + M2M (the M2M slot stays empty) and the per-level L2P replaces L2L
    (the near slot then holds P2P, the sum of the levels' L2P and the
    scatter)."""
    if 4 <= order <= 8:
        return {"p2m": "p2m_all_levels(=p2m+m2m)", "m2m": None, "vt_proxy": "vt_proxy", "m2l": "m2l",
                "u_expand": "u_expand", "l2l": "l2p_all_levels(=l2l+l2p far)", "near": "p2p+scatter",
                "total": "total"}
    return {"p2m": "p2m", "m2m": "m2m", "vt_proxy": "vt_proxy", "m2l": "m2l", "u_expand": "u_expand",
            "l2l": "l2l", "near": "l2p+p2p+scatter", "total": "total"}


def named_phases(phase, order):
    out = {}
    for k, v in phase.items():
        nm = phase_names(order)[k]
        if nm is not None:
            out[nm] = round(v, 3)
    return out


def phase_rooflines(phase, info, ops, nc, precision, kernel_id, ranks=None, proxy_from=None):
    """Algorithmic bytes or flops of each phase over its measured time, as a
    fraction of MI355X's peak (HBM 8 TB/s, FP64 78.6 TF/s): HBM-bound
    expansion passes count the bytes they must move (S / P / Pf / F once,
    points and weights once per level that reads them), P2P counts
    (c_K + 2 nCols) flops per ordered pair (c_K = 9 for 1/r, 11 for the
    Gaussian).  ranks: {level: (kf, ks)} the path runs (default r_f x r_s);
    S and F are FP64 in the FP64 mode and FP32 in the mixed modes, P and Pf
    likewise.  proxy_from: the first level of the proxy passes (levels of
    rank <= 32, mixed modes), whose boxes have no S / F; above it P2M and
    L2P run at its parent level only, with M2M / L2L to the coarser ones."""
    n3, rf, rs = ops.n3, ops.rank_f, ops.rank_s
    L = info["tree_level"]
    N = info["npoints"]
    boxes = info["boxes"]
    B = sum(boxes[2:L + 1])
    eb = 8 if precision == "fp64" else 4  # S / F and P / Pf element bytes
    ranks = ranks or {l: (rf, rs) for l in range(2, L + 1)}
    S = B * n3 * nc * eb
    P = sum(boxes[l] * nc * ranks[l][1] for l in range(2, L + 1)) * eb
    Pf = sum(boxes[l] * nc * ranks[l][0] for l in range(2, L + 1)) * eb
    nlev = L - 1
    hbm = 8.0e12
    ck = 11 if kernel_id == 3 else 9
    res = {}

    def hbm_entry(name, t_ms, by):
        t = t_ms * 1e-3
        if t > 0:
            res[name] = {"bound": "hbm", "bytes": by, "GB/s": round(by / t / 1e9, 1),
                         "frac": round(by / t / hbm, 3)}

    Lp = proxy_from if proxy_from is not None else L + 1
    if Lp <= L:
        pts, lpw = N * (24 + 8 * nc), N * nc * 8  # points + weights; one level's point-major results
        nx = L - Lp + 1                           # proxy levels
        Se = sum(boxes[l] for l in range(2, Lp)) * n3 * nc * eb  # S (then F) of the expansion levels
        Px = sum(boxes[l] * nc * ranks[l][1] for l in range(Lp, L + 1)) * eb
        Pfx = sum(boxes[l] * nc * ranks[l][0] for l in range(Lp, L + 1)) * eb
        hbm_entry("p2m_all_levels", phase["p2m"], (pts + 2 * Se) if Lp > 2 else 0)
        hbm_entry("vt_proxy", phase["vt_proxy"], Se + (P - Px) + nx * pts + Px)
        hbm_entry("u_expand", phase["u_expand"], (Pf - Pfx) + Se)
        hbm_entry("l2p_all_levels", phase["l2l"],
                  (2 * Se + N * 24 + lpw if Lp > 2 else 0) + Pfx + nx * (N * 24 + 2 * lpw))
    else:
        hbm_entry("p2m_all_levels", phase["p2m"], S + nlev * N * (24 + 8 * nc))
        hbm_entry("vt_proxy", phase["vt_proxy"], S + P)
        hbm_entry("u_expand", phase["u_expand"], Pf + S)
        hbm_entry("l2p_all_levels", phase["l2l"], S + nlev * N * 24 + (2 * (nlev - 1) + 1) * N * nc * 8)
    t = phase["near"] * 1e-3
    if t > 0:
        fl = info["near_pairs"] * (ck + 2 * nc)
        res["p2p"] = {"bound": "valu_fp64", "flops": fl, "TFLOP/s": round(fl / t / 1e12, 2),
                      "frac": round(fl / t / 78.6e12, 3)}
    return res


def proxy_from(ops, mranks, L, args):
    """First level of the proxy passes: the deepest levels whose kept ranks
    are at most 32 (mixed modes without the cross-product-sum tile, orders
    4..8, r_f >= 32); L + 1 when there are none."""
    if mranks is None or not 4 <= args.order <= 8 or ops.rank_f < 32:
        return L + 1
    lp = L + 1
    while lp > 2:
        kf, ks, _ = mranks[lp - 3 if ops.nsets > 1 else 0]
        if kf > 32 or ks > 32:
            break
        lp -= 1
    return lp


def kernel_name(args):
    return {"gaussian": "exp(-r^2/0.3^2)", "laplacian": "1/r"}[args.kernel]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="f16x2", choices=["fp64", "mixed", "f16x2", "bf16x2"],
                    help="f16x2 = M2L on FP16 MFMA, operands split in two scaled FP16 parts "
                         "(the fastest variant within 1e-6 of the FP64 path); mixed = BF16 x3")
    ap.add_argument("--fp64-steps", type=int, default=1,
                    help="timed FP64 steps run beside a mixed run (speedup and error)")
    ap.add_argument("--kernel", default="gaussian", choices=["gaussian", "laplacian"])
    ap.add_argument("--npoints", type=int, default=4_000_000)
    ap.add_argument("--ncols", type=int, default=32)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--err-targets", type=int, default=10000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--order-sweep", default="7",
                    help="comma-separated orders measured beside the headline at the same points, weights and "
                         "sampled targets: FP64 and mixed errors against the direct sum, speed ratio; the "
                         "line's accuracy_gate is the lowest of them whose FP64 path is within 1e-6 of the "
                         "direct sum ('' to skip)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-cloud forward+adjoint batch measured beside the headline")
    ap.add_argument("--workload", default="clustered_4m", choices=["clustered_4m", "batch_256_clouds"],
                    help="clustered_4m: the headline single cloud of 4M clustered points (default); "
                         "batch_256_clouds: 256 independent clouds in one HIP-graph batch (a separate "
                         "line, not the headline metric)")
    ap.add_argument("--m2l-tiles", default="",
                    help="JSON {level: [width, cols_per_box, flags]} of the headline precision's M2L tiles "
                         "(as the tuner reports them) set instead of running the tuner (profiled runs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in M2L tiles instead of running the plan's autotuner first")
    ap.add_argument("--max-rel-diff", type=float, default=1e-6,
                    help="accuracy bound of the autotuner in the mixed modes: keep the fastest M2L tile "
                         "whose matvec differs from the FP64 matvec by at most this much (0: speed only)")
    ap.add_argument("--m2l-split-acc", action="store_true",
                    help="two-part modes: run the M2L with the split's cross products summed apart and "
                         "VT / U in FP64 (64-column tiles; closer to the FP64 path, slower)")
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="first run one device copy of this many GiB (a known-size kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of profiled runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (a cloud would "
              "shard by Morton ranges of boxes with a proxy halo exchange; batches by cloud)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import pbbfmm3d_amd as pbb
    from pbbfmm3d_amd import _capi as capi

    if args.workload == "batch_256_clouds":
        run_batch_line(args, torch, pbb, capi)
        return

    dev = torch.device("cuda:0")
    if args.calibrate_copy_gib > 0:
        n = int(args.calibrate_copy_gib * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy of {n * 8} bytes done")
        del src, dst
    kernel = make_kernel(args.kernel)
    N, nc = args.npoints, args.ncols
    log(f"workload N={N} ncols={nc} level={args.level} order={args.order} eps={args.eps} "
        f"kernel={args.kernel} precision={args.precision}")
    xyz = clustered_points(N, seed=0)
    w_host = np.random.default_rng(1).standard_normal((nc, N))

    t0 = time.perf_counter()
    ops = pbb.Operators(kernel, 1.0, args.level, args.order, args.eps)
    t_pre = time.perf_counter() - t0
    log(f"precompute {t_pre:.2f} s, ranks {ops.rank_f}/{ops.rank_s}, sets {ops.nsets}")
    t0 = time.perf_counter()
    plan = pbb.Plan(ops, xyz)
    t_plan = time.perf_counter() - t0
    info = plan.info()
    log(f"plan {t_plan:.2f} s, boxes {info['boxes']}, m2l pairs {info['m2l_pairs']}, "
        f"near pairs {info['near_pairs']}")

    precs = {"fp64": capi.PREC_FP64, "mixed": capi.PREC_MIXED, "f16x2": capi.PREC_MIXED_F16X2,
             "bf16x2": capi.PREC_MIXED_BF16X2}
    prec = precs[args.precision]
    w = torch.tensor(w_host, dtype=torch.float64, device=dev)
    out = torch.zeros_like(w)
    # M2L tile widths picked by the plan's autotuner (untimed, like the tree)
    t0 = time.perf_counter()
    widths = None
    if args.m2l_tiles:
        tiles = {int(k): v for k, v in json.loads(args.m2l_tiles).items()}
        for lvl, (wd, cpb, fl) in sorted(tiles.items()):
            plan.set_m2l_tile_level(lvl, wd, cpb, fl, prec)
        widths = {"m2l_tile": {"levels": tiles, "set_from": "--m2l-tiles"}}
    elif not args.no_tune:
        widths = {"m2l_tile": plan.tune_m2l(w, prec, args.max_rel_diff if prec != capi.PREC_FP64 else 0.0),
                  "max_rel_diff": args.max_rel_diff if prec != capi.PREC_FP64 else None}
        if prec != capi.PREC_FP64 and args.fp64_steps > 0:
            widths["m2l_tile_fp64"] = plan.tune_m2l(w, capi.PREC_FP64)
        widths["tune_s"] = round(time.perf_counter() - t0, 3)
    if args.m2l_split_acc:  # a fixed tile with the cross products summed apart (two-part modes)
        flags = capi.M2L_CHUNK_MAJOR | capi.M2L_WAVE_ROWS | capi.M2L_SPLIT_ACC
        plan.set_m2l_tile(64, 16, flags, prec)
        widths = dict(widths or {}, m2l_tile={"width": 64, "cols_per_box": 16, "flags": flags})
    log(f"m2l tiles {widths or 'built-in'}")
    plan.set_timing(True)

    def timed(p, warmup, steps, tag):
        for i in range(warmup):
            plan.matvec(w, out, precision=p)
            torch.cuda.synchronize()
            log(f"{tag} warmup {i} done")
        phase = {k: 0.0 for k in capi.PHASE_NAMES}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            plan.matvec(w, out, precision=p)
            for k, v in plan.timings().items():  # waits for this step's events
                phase[k] += v
            log(f"{tag} step {i} done")
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        return elapsed / steps, {k: v / steps for k, v in phase.items()}

    def one(p):
        o = torch.zeros_like(w)
        plan.matvec(w, o, precision=p)
        torch.cuda.synchronize()
        return o

    fp64 = None
    if prec != capi.PREC_FP64 and args.fp64_steps > 0:
        t64, ph64 = timed(capi.PREC_FP64, 1, args.fp64_steps, "fp64")
        out64 = one(capi.PREC_FP64)
        fp64 = {"value": N / t64, "ms_per_step": 1e3 * t64, "phase_ms": named_phases(ph64, args.order),
                "phase_rooflines": phase_rooflines(ph64, info, ops, nc, "fp64", kernel.id)}
    fp64r = None
    if prec != capi.PREC_FP64 and args.fp64_steps > 0:
        # the FP64 mode with the same re-compressed operator sets as the mixed
        # path (the speed-up from the arithmetic alone)
        ops.set_fp64_recompressed(True)
        t64r, ph64r = timed(capi.PREC_FP64, 1, 1, "fp64 re-compressed")
        out64r = one(capi.PREC_FP64)
        ops.set_fp64_recompressed(False)
        fp64r = {"value": N / t64r, "ms_per_step": 1e3 * t64r, "phase_ms": named_phases(ph64r, args.order),
                 "rel_diff_vs_fp64_path": float(torch.linalg.norm(out64r - out64) / torch.linalg.norm(out64))}
        del out64r
    t_step, phase = timed(prec, args.warmup, args.steps, args.precision)
    ms_per_step = 1e3 * t_step
    value = N / t_step
    out = one(prec)

    # accuracy against the FP64 direct sum on sampled targets
    ts = np.sort(np.random.default_rng(2).choice(N, size=min(args.err_targets, N), replace=False))
    tgt = torch.tensor(xyz[ts], dtype=torch.float64, device=dev)
    src = torch.tensor(xyz, dtype=torch.float64, device=dev)
    ref = torch.zeros((nc, len(ts)), dtype=torch.float64, device=dev)
    d = kernel.desc()
    import ctypes as C
    capi.check(capi.lib().pbb_direct_sum(C.byref(d), tgt.data_ptr(), len(ts), src.data_ptr(), N,
                                         w.data_ptr(), nc, ref.data_ptr(), None))
    torch.cuda.synchronize()
    tsd = torch.tensor(ts, device=dev)

    def err(o):
        return float(torch.linalg.norm(o[:, tsd] - ref) / torch.linalg.norm(ref))

    rel_err = err(out)
    log(f"rel err vs FP64 direct sum on {len(ts)} targets: {rel_err:.3e}")
    if fp64 is not None:
        fp64["rel_err_vs_direct"] = err(out64)
        fp64["speedup"] = value / fp64["value"]
        fp64["rel_diff_vs_fp64_path"] = float(torch.linalg.norm(out - out64) / torch.linalg.norm(out64))
        log(f"fp64 {fp64['value']:.0f} targets/s, speedup {fp64['speedup']:.2f}, "
            f"rel diff vs FP64 path {fp64['rel_diff_vs_fp64_path']:.3e}")
        if fp64r is not None:
            fp64r["speedup_of_headline"] = value / fp64r["value"]
            fp64["recompressed"] = fp64r
            log(f"fp64 re-compressed {fp64r['value']:.0f} targets/s (headline {fp64r['speedup_of_headline']:.2f}x), "
                f"rel diff vs FP64 path {fp64r['rel_diff_vs_fp64_path']:.3e}")
        del out64

    # roofline of the dominant kernel (M2L): algorithmic flops 2*kf*ks*ncols
    # per far-field box pair of each level, with the ranks the path runs (the
    # reference's r_f x r_s in FP64, each level's re-compressed operator set
    # in the mixed modes), over the M2L phase time (HIP events on the stream
    # the M2L launches run on)
    L = info["tree_level"]
    pairs = info["m2l_pairs_level"]
    mranks = ops.mixed_ranks() if prec != capi.PREC_FP64 else None

    def level_ranks(l):
        if mranks is None:
            return ops.rank_f, ops.rank_s
        kf, ks, tail = mranks[l - 2 if ops.nsets > 1 else 0]
        # a set that keeps its own factors runs r_f x r_s (ks is only padded)
        return (ops.rank_f, ops.rank_s) if (kf == ops.rank_f and tail == 0.0) else (kf, ks)

    flops = sum(2.0 * level_ranks(l)[0] * level_ranks(l)[1] * nc * pairs[l] for l in range(2, L + 1))
    flops_ref = 2.0 * ops.rank_f * ops.rank_s * nc * info["m2l_pairs"]
    achieved = flops / (phase["m2l"] * 1e-3) / 1e12
    bound, peak, unit = PEAK[args.precision]
    # HBM bytes per M2L launch (FETCH_SIZE x 2 + WRITE_SIZE x 1, calibrated
    # on a plain copy), measured by rocprofv3 on this workload and kept in
    # profiles/; null for other workloads
    traffic = None
    if (args.precision, N, nc, args.level, args.order, args.kernel) == ("f16x2", 4_000_000, 32, 7, 6, "gaussian"):
        traffic = M2L_TRAFFIC_F16X2_4M_CLUSTERED
    roofline = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 4), "traffic": traffic,
                "traffic_source": TRAFFIC_SOURCE if traffic else None,
                "kernel": "m2l", "launches_per_step": args.level - 1,
                "algorithmic_flops_per_step": flops,
                "ranks_per_level": {l: list(level_ranks(l)) for l in range(2, L + 1)},
                "flops_per_step_at_reference_ranks": flops_ref}

    sweep = []
    for order in [int(x) for x in args.order_sweep.split(",") if x.strip()]:
        if "plan" in locals():
            del plan, out
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        ops_o = pbb.Operators(kernel, 1.0, args.level, order, args.eps)
        t_pre_o = time.perf_counter() - t0
        plan_o = pbb.Plan(ops_o, xyz)
        log(f"order {order}: precompute {t_pre_o:.1f} s, ranks {ops_o.rank_f}/{ops_o.rank_s}")

        def run_o(p):
            o = torch.zeros_like(w)
            plan_o.matvec(w, o, precision=p)  # untimed: workspace and operator splits
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            o.zero_()
            plan_o.matvec(w, o, precision=p)
            torch.cuda.synchronize()
            return time.perf_counter() - t0, o

        t64_o, o64 = run_o(capi.PREC_FP64)
        e64 = err(o64)
        log(f"order {order}: fp64 {1e3 * t64_o:.0f} ms, rel err vs direct {e64:.3e}")
        # the mixed tile's bound: where the FP64 path is within 1e-6 of the
        # direct sum, the room its error leaves under 1e-6 (the arithmetic
        # difference and the truncation error are independent and add in
        # quadrature; a tenth is kept in reserve for the sampled targets),
        # else the headline bound against the FP64 path
        bound = 0.9 * math.sqrt(1e-12 - e64 * e64) if e64 < 1e-6 else args.max_rel_diff
        tile_o = plan_o.tune_m2l(w, prec, bound)
        tm_o, om = run_o(prec)
        sweep.append({"order": order, "ranks": [ops_o.rank_f, ops_o.rank_s], "precompute_s": round(t_pre_o, 2),
                      "fp64_ms": 1e3 * t64_o, "mixed_ms": 1e3 * tm_o,
                      "fp64_targets_per_s": N / t64_o, "mixed_targets_per_s": N / tm_o,
                      "speedup": t64_o / tm_o,
                      "fp64_rel_err_vs_direct": e64, "mixed_rel_err_vs_direct": err(om),
                      "mixed_rel_diff_vs_fp64_path": float(torch.linalg.norm(om - o64) / torch.linalg.norm(o64)),
                      "tuner_bound": bound, "m2l_tile": tile_o})
        log(f"order {order}: {sweep[-1]}")
        del plan_o, o64, om
    # the lowest measured order at which the FP64 path is within 1e-6 of the
    # direct sum, and the mixed path there
    gate = None
    for r in sorted(sweep, key=lambda r: r["order"]):
        if r["fp64_rel_err_vs_direct"] <= 1e-6:
            gate = {k: r[k] for k in ("order", "fp64_rel_err_vs_direct", "mixed_rel_err_vs_direct",
                                      "fp64_targets_per_s", "mixed_targets_per_s", "speedup")}
            gate["mixed_within_1e-6_of_direct"] = r["mixed_rel_err_vs_direct"] <= 1e-6
            break

    batch = None
    if not args.no_batch:
        if "plan" in locals():
            del plan, w, out
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        log("256-cloud batch (forward + adjoint in one graph) ...")
        batch = batch_clouds(args, torch, pbb, capi)

    cpu = None
    if not args.no_cpu_baseline:
        if "plan" in locals():
            del plan, w, out
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        log("cpu baseline (oracle on whole problems of the reference's CPU-runnable shapes) ...")
        cpu = cpu_baseline(torch, pbb, capi)
        log(f"cpu baseline {cpu['value']:.1f} targets/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC, "value": value, "unit": "targets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype_label(args.precision, (widths or {}).get("m2l_tile", {}).get("flags")),
        "data": "synthetic",
        "config": {"workload": f"{N} clustered points (8 Gaussians sigma 0.05 + 20% uniform), "
                               f"{kernel_name(args)}, {nc} columns, level {args.level}, order "
                               f"{args.order}, eps {args.eps}",
                   "npoints": N, "ncols": nc, "tree_level": args.level, "order": args.order,
                   "eps": args.eps, "kernel": args.kernel, "precision": args.precision},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "rel_err_vs_direct": rel_err,
        "fp64_path": fp64,
        "phase_ms": named_phases(phase, args.order),
        "phase_rooflines": phase_rooflines(phase, info, ops, nc, args.precision, kernel.id,
                                           {l: level_ranks(l) for l in range(2, info["tree_level"] + 1)},
                                           proxy_from(ops, mranks, info["tree_level"], args)),
        "ranks": [ops.rank_f, ops.rank_s],
        "m2l_pairs": info["m2l_pairs"], "near_pairs": info["near_pairs"],
        "precompute_s": round(t_pre, 3), "plan_s": round(t_plan, 3),
        # as the reference times its constructor (distribution + matvec,
        # compute.hpp:140-146): upload, GPU tree build and one matvec
        "with_plan_targets_per_s": N / (t_step + t_plan),
        "tuning": widths,
        "batch_256_clouds": batch,
        "order_sweep": sweep or None,
        "accuracy_gate": gate,
    }
    print(json.dumps(line), flush=True)


def batch_clouds(args, torch, pbb, capi, log_prefix="batch"):
    """256 independent clouds (sizes log-uniform in [2e3, 2e5], U[0,1)^3
    positions, seed 1000 + k), tree level round(log8(N/60)) clamped to 3..6,
    order 4..6 in turn, exp(-r/0.5), 16 columns, eps 1e-7: the forward
    matvec AND the adjoint of every cloud captured into one HIP graph
    (pbb_batch_create_entries), M2L tiles tuned per operator set (within
    1e-6 of FP64 in the mixed mode), FP64 and the headline mixed mode.
    targets/s = total points / seconds per launch (each launch applies A
    and A^T to every cloud)."""
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(999)
    ncl, nc, eps = 256, 16, 1e-7
    sizes = np.exp(rng.uniform(np.log(2e3), np.log(2e5), ncl)).astype(int)
    kernel = pbb.exponential(0.5)
    t0 = time.perf_counter()
    ops, plans, ws, ps = {}, [], [], []
    for k in range(ncl):
        n = int(sizes[k])
        lvl = int(min(6, max(3, round(np.log(n / 60.0) / np.log(8.0)))))
        order = 4 + k % 3
        if (lvl, order) not in ops:
            ops[(lvl, order)] = pbb.Operators(kernel, 1.0, lvl, order, eps)
        xyz = np.random.default_rng(1000 + k).random((n, 3))
        plans.append(pbb.Plan(ops[(lvl, order)], xyz))
        ws.append(torch.tensor(np.random.default_rng(2000 + k).standard_normal((nc, n)), dtype=torch.float64,
                               device=dev))
        ps.append(torch.tensor(np.random.default_rng(3000 + k).standard_normal((nc, n)), dtype=torch.float64,
                               device=dev))
    t_setup = time.perf_counter() - t0
    total = int(sizes.sum())
    log(f"{log_prefix}: {ncl} clouds, {total} points, {len(ops)} operator sets, setup {t_setup:.1f} s")
    precs = {"fp64": capi.PREC_FP64, "mixed": capi.PREC_MIXED, "f16x2": capi.PREC_MIXED_F16X2,
             "bf16x2": capi.PREC_MIXED_BF16X2}
    res = {}
    t0 = time.perf_counter()
    for name in ("fp64", args.precision):
        pbb.Batch.tune(plans, ws, precs[name], args.max_rel_diff if name != "fp64" else 0.0)
    t_tune = time.perf_counter() - t0
    ent_plans = [p for p in plans for _ in (0, 1)]
    ent_w = [w for pair in zip(ws, ps) for w in pair]
    for name in ("fp64", args.precision):
        prec = precs[name]
        outs = [torch.zeros_like(w) for w in ent_w]
        b = pbb.Batch(ent_plans, ent_w, outs, nc, precision=prec, transpose=[False, True] * ncl)
        for i in range(max(1, args.warmup)):
            b.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            b.launch()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        res[name] = (dt, outs)
        log(f"{log_prefix} {name}: {1e3 * dt:.1f} ms per launch")
        del b
    # the graph equals the eager calls (FP64, sampled clouds); the mixed
    # batch stays within the mixed bar of the FP64 one
    _, o64 = res["fp64"]
    same = 0.0
    for k in range(0, ncl, 37):
        e = plans[k].matvec(ws[k])
        same = max(same, float(torch.linalg.norm(e - o64[2 * k]) / torch.linalg.norm(o64[2 * k])))
        e = plans[k].matvec_transpose(ps[k])
        same = max(same, float(torch.linalg.norm(e - o64[2 * k + 1]) / torch.linalg.norm(o64[2 * k + 1])))
    dt, om = res[args.precision]
    num = sum(float(torch.linalg.norm(a - b) ** 2) for a, b in zip(om, o64))
    den = sum(float(torch.linalg.norm(b) ** 2) for b in o64)
    return {"workload": f"{ncl} clouds of U[0,1)^3 points, sizes log-uniform in [2e3, 2e5] ({total} points), "
                        f"levels 3..6, orders 4..6, exp(-r/0.5), {nc} columns, eps {eps}; forward and adjoint of "
                        f"every cloud in one graph launch",
            "clouds": ncl, "npoints": total, "ncols": nc, "precision": args.precision,
            "targets_per_s": total / dt, "ms_per_launch": 1e3 * dt,
            "fp64": {"targets_per_s": total / res["fp64"][0], "ms_per_launch": 1e3 * res["fp64"][0],
                     "max_rel_diff_vs_eager": same},
            "rel_diff_vs_fp64_batch": (num / den) ** 0.5, "setup_s": round(t_setup, 2),
            "tune_s": round(t_tune, 2)}


def run_batch_line(args, torch, pbb, capi):
    """The 256-cloud batch as a line of its own (not the headline metric)."""
    r = batch_clouds(args, torch, pbb, capi, "batch_256_clouds")
    line = {
        "metric": "targets/s, 256-cloud batch, forward + adjoint in one HIP graph", "value": r["targets_per_s"],
        "unit": "targets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_launch"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype_label(args.precision, None), "data": "synthetic",
        "config": {k: r[k] for k in ("workload", "clouds", "npoints", "ncols", "precision")},
        "fp64_batch": r["fp64"], "rel_diff_vs_fp64_batch": r["rel_diff_vs_fp64_batch"], "setup_s": r["setup_s"],
        "tune_s": r["tune_s"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
