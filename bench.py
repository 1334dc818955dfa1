"""Benchmark of the MI355X box FGT: leaf points per second of one boxfgt call.

One "step" is one evaluation of the resident plan (leaf values -> PW ->
translation -> pot/grad -> near field) over the workload's leaf values,
already in HBM.  The default workload is the adaptive configuration the
metric is quoted on: a periodic adaptive quadtree of ~20k leaves (levels
7-10, n=16 Legendre, ~5.1M points), nd=1, delta=1e-4, eps=1e-9, pot+grad.

Prints ONE JSON line.  ``value`` = nleaf*npbox*nd / (time per step).
``roofline`` describes the kernel with the largest share of the step, its
algorithmic FLOPs per launch divided by its average duration measured with
HIP events on the plan's stream over the timed steps.  ``cpu_baseline``
times the reference Fortran boxfgt (compiled from its sources into
oracle/_ref) -- or, where that library is absent, the C oracle -- on the
host cores, on the same tree and density.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
# The bound of a leaf pass counted in FP64 products: the matrix peak of the
# arithmetic divided by the matrix products one FP64 product takes there
# (split-FP32: at most three FP32 hi/lo products; three-slice bf16: six
# slice products)
MATRIX_PEAK = {"fp64": (FP64_PEAK_TFLOPS, "FP64 matrix peak"),
               "split32": (157.3 / 3, "FP32 matrix peak 157.3 / 3 hi-lo products"),
               "bf16x3": (2500.0 / 6, "bf16 matrix peak 2500 / 6 slice products")}
DTYPE = {"fp64": "f64", "split32": "f32x2+f64", "bf16x3": "bf16x3+f32+f64"}


def stage_peak(precision, stage):
    """(peak, note) of a stage's products: the three-slice bf16 plans run the
    near field in split-FP32."""
    return MATRIX_PEAK["split32" if precision == "bf16x3" and stage == "near_field" else precision]
SETTLE_S = 0.25  # seconds of untimed executions before the warm-up
HBM_PEAK_GBS = 8000.0


def algorithmic_work(plan, w):
    """Per-launch algorithmic FLOPs (FP64) or bytes of each stage, from the plan's
    schedule: the upward pass 4n^2H per PW leaf (y contraction) + 8nPH per x
    contraction, the evaluation 8*ntab*nPH per x contraction + 4c n^2 H per
    leaf (c = 1, 3, 6 products for pot, pg, pgh); an x contraction serves a
    whole column of leaves (same cutoff box, level and x cell) where the
    kernels share it (plan.info xc_up / xc_down), else one leaf.  Near field
    {4, 10, 18} n^3 per source/target pair.  The merge reads the chunk sums
    it adds and writes the sums (16 B per complex coefficient); the
    translation reads each cutoff box's expansion once (the partners' reads
    are meant to hit in L2) and writes the H x H folded pairs (32 B each)."""
    info = plan.info
    n, nd = w.norder, w.nd
    P, nexp = info["npw"], info["nexp"]
    H = (P + 1) // 2
    L = info["n_pw_leaves"]
    ntab = w.ifpgh
    c = {1: 1, 2: 3, 3: 6}[w.ifpgh]
    up = nd * (L * 4 * n * n * H + info["xc_up"] * 8 * n * P * H)
    down = nd * (info["xc_down"] * 8 * ntab * n * P * H + L * 4 * c * n * n * H)
    near = nd * info["n_near_pairs"] * {1: 4, 2: 10, 3: 18}[w.ifpgh] * n**3
    merge_bytes = 16 * nd * nexp * (info["merge_reads"] + info["merge_writes"])
    translate_bytes = info["n_ancestors"] * nd * (16 * nexp + 32 * H * H)
    return {"leaf_to_pw": ("flop", up), "merge": ("byte", merge_bytes), "translate": ("byte", translate_bytes),
            "pw_to_values": ("flop", down), "near_field": ("flop", near)}


def per_leaf_work(plan, w):
    """The same two leaf passes counted with one x contraction per PW leaf
    (nd * L * [4nH(n + 2P)] upward, nd * L * [8 ntab nPH + 4c n^2 H] downward):
    the work of the leaf-by-leaf algorithm, which the column sharing avoids."""
    info = plan.info
    n, nd = w.norder, w.nd
    P = info["npw"]
    H = (P + 1) // 2
    L = info["n_pw_leaves"]
    c = {1: 1, 2: 3, 3: 6}[w.ifpgh]
    return {"leaf_to_pw": nd * L * 4 * n * H * (n + 2 * P),
            "pw_to_values": nd * L * (8 * w.ifpgh * n * P * H + 4 * c * n * n * H)}


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` recorded from rocprofv3 counters in
    profiles/boxfgt_traffic.txt for this workload and precision, else None."""
    path = os.path.join(ROOT, "profiles", "boxfgt_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and not line.startswith("#") and f[:3] == [workload, precision, kernel]:
            return {"bytes": int(f[3]) + int(f[4]), "fetch_bytes": int(f[3]), "write_bytes": int(f[4]),
                    "source": "profiles/boxfgt_traffic.txt (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)"}
    return None


TREE_WORKLOAD = ("64 moving Gaussians a*exp(-|x-c(t)|^2/1e-5) (seed 7), n=16, eps=1e-9, periodic, l2 tail rule: "
                 "tree build at t=0, then one re-mesh with the values carried after the Gaussians move to t=0.002")


def tree_bumps():
    """(a, cx, cy, vx, vy, s) of the tree workload's Gaussians."""
    rng = np.random.default_rng(7)
    return [(1.0 + rng.random(), rng.uniform(-.45, .45), rng.uniform(-.45, .45), rng.uniform(-1, 1),
             rng.uniform(-1, 1), 1e-5) for _ in range(64)]


def tree_fun_gpu(bumps, t):
    import torch

    col = lambda j: torch.tensor([b[j] for b in bumps], dtype=torch.float64, device="cuda")  # noqa: E731
    a, cx, cy, vx, vy, s = (col(j) for j in range(6))
    cx, cy = cx + vx * t, cy + vy * t

    def f(xy):
        dx = xy[:, 0:1] - cx[None, :]
        dy = xy[:, 1:2] - cy[None, :]
        return (a[None, :] * torch.exp(-(dx * dx + dy * dy) / s[None, :])).sum(dim=1)

    return f


def tree_leg():
    """The adaptive tree on the GPU (bfgt_tree_build + bfgt_tree_adapt) with
    the user function evaluated on the GPU: best of 3 after a warm-up."""
    import torch

    from heat2d_per_box_amd.devtree import DeviceTree

    b = tree_bumps()
    f0, f1 = tree_fun_gpu(b, 0.0), tree_fun_gpu(b, 0.002)
    builds, adapts, sizes = [], [], None
    for rep in range(4):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        D = DeviceTree.build(f0, 1, 16, 1e-9, 0, 1, max_levels=20, max_boxes=1 << 17)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        nb0 = D.info["nboxes"]
        D.adapt(f1, 1e-9, max_levels=20)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        sizes = (nb0, D.info["nboxes"], D.info["nleaves"])
        D.close()
        if rep:
            builds.append(t1 - t0)
            adapts.append(t2 - t1)
    return {"workload": TREE_WORKLOAD, "boxes_built": sizes[0], "boxes_after_remesh": sizes[1],
            "leaves_after_remesh": sizes[2], "gpu_build_s": min(builds), "gpu_remesh_s": min(adapts),
            "user_function": "torch on the GPU (the reference calls its Fortran function per point on the host)"}


NEWTREE_WORKLOAD = ("the nd=3 spot workload on its 4,017-box tree (n=8, delta=1e-3, eps=1e-9, pot+grad+hess), "
                    "boxfgt with ifnewtree=1: transform, refinement of the leaves whose pot is not resolved, "
                    "coarsening, 2:1 rebalance")


def _newtree_case():
    from heat2d_per_box_amd.densities import make_workload
    from heat2d_per_box_amd.tree import with_capacity

    w = make_workload("c3", "small", delta=1e-3, eps=1e-9)
    return w, with_capacity(w.tree, 6 * w.tree.nboxes)


def newtree_leg():
    """boxfgt with ifnewtree=1 through the host-array drop-in: best of 3 calls
    after a warm-up, each on a fresh copy of the tree (a refined tree is a new
    tree, so every call builds its plan)."""
    from heat2d_per_box_amd.boxfgt import boxfgt as boxfgt_host

    w, tc = _newtree_case()
    nb, npb, nd = tc.centers.shape[0], w.norder ** 2, w.nd
    fv = np.zeros((nb, npb, nd))
    fv[: w.fvals.shape[0]] = w.fvals
    best = None
    for rep in range(4):
        t = tc.copy()
        pot, grad, hess = np.zeros((nb, npb, nd)), np.zeros((nb, npb, 2, nd)), np.zeros((nb, npb, 3, nd))
        ti = np.zeros(20)
        t0 = time.perf_counter()
        nl = boxfgt_host(nd, 2, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, npb, nb, t.nlevels, t.ltree, t.itree,
                         t.iptr, t.centers, t.boxsize, fv, w.ifpgh, pot, grad, hess, ifnewtree=1, timeinfo=ti)
        dt = time.perf_counter() - t0
        t.nlevels = nl
        if rep and (best is None or dt < best[0]):
            best = (dt, ti.copy(), t.nboxes)
    dt, ti, nb1 = best
    return {"workload": NEWTREE_WORKLOAD, "boxes_before": int(w.tree.nboxes), "boxes_after": int(nb1),
            "boxes_created": int(ti[18]), "gpu_call_s": dt,
            "refine_s": float(ti[8]), "coarsen_s": float(ti[9]), "rebalance_s": float(ti[10])}


def newtree_reference(nthreads):
    """The reference boxfgt with ifnewtree=1 on the same case (best of 3)."""
    from oracle import oracle as O

    w, tc = _newtree_case()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        r = O.reference_boxfgt(tc, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh,
                               nthreads=nthreads, ifnewtree=1)
        times.append(time.perf_counter() - t0)
    return {"call_s": min(times), "boxes_after": int(r.tree.nboxes), "threads": nthreads,
            "timeinfo_9_11_s": [float(x) for x in r.timeinfo[8:11]]}


def tree_reference(nthreads):
    """The reference's vol_tree_mem2 + vol_tree_build2 and vol_tree_adap_fv on
    the tree workload (one call each, C test function per point)."""
    from oracle import oracle as O

    if not O.reference_available():
        return None
    O.set_reference_threads(nthreads)
    b = tree_bumps()
    dp0, dp1 = O.tree_fun_dpars(b, 0.0), O.tree_fun_dpars(b, 0.002)
    t0 = time.perf_counter()
    mb, ml, lt, rintl = O.reference_tree_mem(1, 0, 1, 1e-9, 1.0, 16, 2, 1.0, dp0)
    caps = (2 * mb, ml + 5, 2 * lt)
    T = O.reference_tree_build(1, 0, 1, 1e-9, 1.0, 16, 2, 1.0, dp0, rintl, *caps)
    t1 = time.perf_counter()
    A = O.reference_tree_adapt(T, T.vals.copy(), 1, 0, 1, 1e-9, 16, 2, rintl, dp1, *caps)
    t2 = time.perf_counter()
    return {"build_s": t1 - t0, "remesh_s": t2 - t1, "boxes_built": T.nboxes, "boxes_after_remesh": A.nboxes,
            "threads": nthreads}


def cpu_baseline(w, gpu_out, max_threads, extra=None):
    """Time the reference Fortran boxfgt (or the C oracle) on the host cores, and
    report the relative L2 difference (leaf points) of the GPU outputs from it.
    The reference runs on min(max_threads, CPUs this process may use) OpenMP
    threads bound close (OMP_PROC_BIND=close, OMP_PLACES=cores, set before its
    OpenMP runtime starts); its own stage times (timeinfo(1..12)) are kept."""
    from oracle import oracle as O

    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    nthreads = max(1, min(affinity, max_threads))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if O.reference_available():
        kind = "reference"
        run = lambda: O.reference_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder,  # noqa: E731
                                         w.ifpgh, nthreads=nthreads)
    else:
        kind = "port"
        nthreads = 1
        run = lambda: O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder,  # noqa: E731
                                      w.ifpgh)
    r = run()  # warm-up
    # full calls for about 10 s of host work (at least 5), median reported
    times, tinfo = [], []
    t_start = time.time()
    while len(times) < 5 or (time.time() - t_start < 10.0 and len(times) < 1000):
        t0 = time.perf_counter()
        r = run()
        times.append(time.perf_counter() - t0)
        tinfo.append(np.asarray(r.timeinfo[:12], dtype=float))
    t_total = time.time() - t_start
    tmed = float(np.median(times))
    ti_med = np.median(np.array(tinfo), axis=0)
    lv = w.tree.nchild == 0

    def rel(outs):
        e = {}
        for name, ref in (("pot", r.pot), ("grad", r.grad), ("hess", r.hess)):
            g = outs.get(name)
            if g is None or ref is None:
                continue
            den = np.linalg.norm(ref[lv])
            e[name] = float(np.linalg.norm((g - ref)[lv]) / den) if den > 0 else 0.0
        return e

    err = rel(gpu_out)
    err_extra = {k: rel(v) for k, v in (extra or {}).items()}
    tree_ref = tree_reference(nthreads) if kind == "reference" else None
    newtree_ref = newtree_reference(nthreads) if kind == "reference" else None
    return {
        "value": w.npts * w.nd / tmed,
        "unit": "leaf pts/s",
        "cores": nthreads,
        "kind": kind,
        "sample": f"one full boxfgt call on the same workload ({w.tree.nleaves} leaves, {w.npts} pts), "
                  f"median of {len(times)} calls over {t_total:.1f} s after one warm-up; "
                  f"{'amdflang -O3 -fopenmp reference' if kind == 'reference' else 'C oracle, 1 thread'}",
        "seconds_per_call": tmed,
        "host_cpus": os.cpu_count(),
        "affinity_cpus": affinity,
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "timeinfo_median_s": [float(x) for x in ti_med] if kind == "reference" else None,
        "steps_1_to_6_s": float(ti_med[1:7].sum()) if kind == "reference" else None,
        "timeinfo_note": "reference timeinfo(1..12): 1 set-up (tables), 2-6 PW steps 1-5, 7 near field, "
                         "8 asymptotic, 9-11 tree changes, 12 targets; steps_1_to_6_s = sum of slots 2-7, "
                         "the part a kept GPU plan runs (compare with ms_per_step)",
        "gpu_vs_cpu_rel_err": err,
        **{f"gpu_{k}_vs_cpu_rel_err": v for k, v in err_extra.items()},
        "tree_remesh": tree_ref,
        "newtree": newtree_ref,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--scale", default="full", choices=["full", "mixed", "small", "tiny"],
                    help="mixed: c2 or c5 at full size with near-field leaves (floor below the cutoff level)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "split32", "bf16x3"])
    ap.add_argument("--delta", type=float, default=None, help="override delta (config c3 only)")
    ap.add_argument("--eps", type=float, default=None, help="override eps (config c3 only)")
    ap.add_argument("--densities", type=int, default=None,
                    help="keep the first K densities of the workload (e.g. c3 with one density)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tree", action="store_true", help="skip timing the adaptive tree build and re-mesh")
    ap.add_argument("--no-split", action="store_true", help="skip timing the other precisions")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--settle-s", type=float, default=SETTLE_S,
                    help="seconds of untimed executions before the warm-up (shorter under a profiler)")
    ap.add_argument("--accumulate", action="store_true",
                    help="time the accumulating execution (pot += A f) instead of the overwriting one "
                         "(pot = A f on the leaves); both are reported")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a device copy of this size first (calibrates rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the box FGT here has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import torch

    from heat2d_per_box_amd.boxfgt import BoxFGT
    from heat2d_per_box_amd.densities import make_workload

    t0 = time.time()
    kw = {}
    if args.config == "c3":
        if args.delta is not None:
            kw["delta"] = args.delta
        if args.eps is not None:
            kw["eps"] = args.eps
    w = make_workload(args.config, args.scale, **kw)
    if args.densities is not None and args.densities < w.nd:
        k = max(args.densities, 1)
        w = dataclasses.replace(w, fvals=np.ascontiguousarray(w.fvals[..., :k]), nd=k,
                                description=f"{w.description}; the first {k} of its {w.nd} densities")
    t_workload = time.time() - t0
    print(f"[bench] workload {args.config}/{args.scale}: {w.tree.nleaves} leaves, {w.npts} pts, built in "
          f"{t_workload:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    if args.copy_probe_mib:
        # a plain streaming kernel of known traffic (reads and writes copy_probe_mib MiB each):
        # calibrates the FETCH_SIZE / WRITE_SIZE counters of a profiled run
        a = torch.empty(args.copy_probe_mib * (1 << 20) // 8, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        del a, b
    tree = w.tree.copy()
    t0 = time.time()
    plan = BoxFGT(tree, w.nd, w.delta, w.eps, w.ipoly, w.norder, w.ifpgh, precision=args.precision)
    torch.cuda.synchronize()
    t_plan = time.time() - t0
    nb, npb, nd = w.fvals.shape
    fv = torch.from_numpy(w.fvals).to(dev)
    pot = torch.zeros((nb, npb, nd), dtype=torch.float64, device=dev)
    grad = torch.zeros((nb, npb, 2, nd), dtype=torch.float64, device=dev) if w.ifpgh >= 2 else None
    hess = torch.zeros((nb, npb, 3, nd), dtype=torch.float64, device=dev) if w.ifpgh >= 3 else None

    names = ["leaf_to_pw", "merge", "translate", "split", "pw_to_values", "near_field"]
    work = algorithmic_work(plan, w)
    # the overwriting execution (pot = A f on the leaves, as the reference's
    # callers use boxfgt: clear, call, add the scaled result) unless asked
    ow = not args.accumulate

    # one checked call for the error columns (its first launches also load the
    # kernels), then the warm-up, which records events around every stage to
    # pick the dominant one (one such call even with --warmup 0)
    plan(fv, pot, grad, hess, overwrite=ow)
    torch.cuda.synchronize()
    gpu_out = {"pot": pot.cpu().numpy(), "grad": grad.cpu().numpy() if grad is not None else None,
               "hess": hess.cpu().numpy() if hess is not None else None}
    # settle: untimed executions for SETTLE_S seconds of wall time (at least
    # 20), so that the GPU runs at its steady clocks when the warm-up starts
    # (from idle, the first ~100 executions of the c2 step run up to 15 %
    # slower)
    t_settle = time.time()
    n_settle = 0
    while n_settle < 20 or time.time() - t_settle < args.settle_s:
        for _ in range(10):
            plan(fv, pot, grad, hess, overwrite=ow)
        n_settle += 10
        torch.cuda.synchronize()
    t_settle = time.time() - t_settle
    plan.profile_begin(max(args.warmup, 1))
    for _ in range(max(args.warmup, 1)):
        plan(fv, pot, grad, hess, overwrite=ow)
    torch.cuda.synchronize()
    warm_sum, nwarm = plan.profile_end()
    warm_ms = {k: 1e3 * v / max(nwarm, 1) for k, v in zip(names, warm_sum)}
    dom = max((k for k in names if k in work and work[k][1]), key=lambda k: warm_ms[k])
    print(f"[bench] plan {t_plan:.2f}s, warm-up done; timing {args.steps} steps", file=sys.stderr, flush=True)

    # the timed region: K executions back to back; HIP events on the plan's
    # stream around the dominant stage only (events between the other stages
    # would add their own gaps), in the last tenth of them (at least 5): each
    # event costs device time that a production call, which records none,
    # does not pay.  The window's events exist before the region starts.
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    n_dom = min(args.steps, max(5, args.steps // 10))
    plan.profile_begin(n_dom, names.index(dom))
    plan.profile_end()
    torch.cuda.synchronize()
    e0.record(stream)
    for i in range(args.steps):
        if i == args.steps - n_dom:
            plan.profile_begin(n_dom, names.index(dom))
        plan(fv, pot, grad, hess, overwrite=ow)
    e1.record(stream)
    torch.cuda.synchronize()
    dom_sum, ndom = plan.profile_end()
    ms = e0.elapsed_time(e1) / args.steps
    npts = w.npts * w.nd
    value = npts / (ms * 1e-3)

    # every stage's time from K more executions with events between all the
    # stages (outside the timed region)
    plan.profile_begin(args.steps)
    for _ in range(args.steps):
        plan(fv, pot, grad, hess, overwrite=ow)
    torch.cuda.synchronize()
    stage_sum, nprof = plan.profile_end()
    stage_ms = {k: 1e3 * v / max(nprof, 1) for k, v in zip(names, stage_sum)}
    dom_ms = 1e3 * dom_sum[names.index(dom)] / max(ndom, 1)
    # the other execution mode over K more calls (outside the timed region),
    # after one untimed call in that mode
    plan(fv, pot, grad, hess, overwrite=not ow)
    torch.cuda.synchronize()
    f0 = torch.cuda.Event(enable_timing=True)
    f1 = torch.cuda.Event(enable_timing=True)
    f0.record(stream)
    for _ in range(args.steps):
        plan(fv, pot, grad, hess, overwrite=not ow)
    f1.record(stream)
    torch.cuda.synchronize()
    ms_other_mode = f0.elapsed_time(f1) / args.steps
    kind, amount = work[dom]
    dur = dom_ms * 1e-3
    if kind == "flop":
        pk = stage_peak(args.precision, dom)
        roof = {"bound": "mfma", "achieved": amount / dur / 1e12, "peak": pk[0], "unit": "TFLOP/s",
                "peak_note": pk[1]}
    else:
        roof = {"bound": "hbm", "achieved": amount / dur / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    kern = {"leaf_to_pw": "k_up", "merge": "k_merge", "translate": "k_m2l", "pw_to_values": "k_down",
            "near_field": "k_near"}
    roof["kernel"] = kern[dom]
    # the recorded counters are of the overwriting execution
    tr = measured_traffic(f"{args.config}/{args.scale}", args.precision, kern[dom]) if ow else None
    roof["traffic"] = tr["bytes"] if tr else None
    if tr:
        roof["traffic_detail"] = tr
    roof["algorithmic_per_launch"] = amount
    roof["avg_launch_ms"] = dom_ms
    roof["timing"] = (f"HIP events on the plan's stream around this stage in the last {ndom} of the "
                      f"{args.steps} timed executions")
    plw = per_leaf_work(plan, w)
    if dom in plw:
        # the leaf-by-leaf operation count over the same time: how fast the
        # kernel does the per-leaf algorithm's work (above the peak when the
        # column sharing removes more than the remaining inefficiency)
        roof["per_leaf_formula_per_launch"] = plw[dom]
        roof["per_leaf_formula_tflops"] = plw[dom] / dur / 1e12
    # every stage against its own bound
    stages_roof = {}
    for k in names:
        if k not in work or not work[k][1] or stage_ms[k] <= 0:
            continue
        knd, amt = work[k]
        t = stage_ms[k] * 1e-3
        if knd == "flop":
            stages_roof[k] = {"bound": "mfma", "achieved_tflops": amt / t / 1e12,
                              "frac": amt / t / 1e12 / stage_peak(args.precision, k)[0]}
        else:
            stages_roof[k] = {"bound": "hbm", "achieved_gbs": amt / t / 1e9, "frac": amt / t / 1e9 / HBM_PEAK_GBS}

    # the other arithmetics on the same tree (split-FP32 and three-slice bf16
    # leaf transforms vs all-FP64)
    others = []
    for oprec in ([] if args.no_split else [x for x in ("fp64", "split32", "bf16x3") if x != args.precision]):
        plan2 = BoxFGT(w.tree.copy(), w.nd, w.delta, w.eps, w.ipoly, w.norder, w.ifpgh, precision=oprec)
        pot2 = torch.zeros_like(pot)
        grad2 = torch.zeros_like(grad) if grad is not None else None
        hess2 = torch.zeros_like(hess) if hess is not None else None
        plan2(fv, pot2, grad2, hess2, overwrite=ow)
        torch.cuda.synchronize()
        out2 = {"pot": pot2.cpu().numpy(), "grad": grad2.cpu().numpy() if grad2 is not None else None,
                "hess": hess2.cpu().numpy() if hess2 is not None else None}
        for _ in range(args.warmup):
            plan2(fv, pot2, grad2, hess2, overwrite=ow)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            plan2(fv, pot2, grad2, hess2, overwrite=ow)
        e1.record(stream)
        torch.cuda.synchronize()
        ms2 = e0.elapsed_time(e1) / args.steps
        plan2.profile_begin(args.steps)
        for _ in range(args.steps):
            plan2(fv, pot2, grad2, hess2, overwrite=ow)
        torch.cuda.synchronize()
        ssum2, np2 = plan2.profile_end()
        lv = w.tree.nchild == 0
        diff = {}
        for k in ("pot", "grad", "hess"):
            if gpu_out[k] is not None:
                den = np.linalg.norm(gpu_out[k][lv])
                diff[k] = float(np.linalg.norm((out2[k] - gpu_out[k])[lv]) / den) if den > 0 else 0.0
        others.append({"precision": oprec, "ran_precision": ["fp64", "split32", "bf16x3"][plan2.info["precision"]],
                       "ms_per_step": ms2, "value": npts / (ms2 * 1e-3),
                       "time_ratio_vs_" + args.precision: ms2 / ms,
                       "stage_ms": {k: 1e3 * v / max(np2, 1) for k, v in zip(names, ssum2)},
                       "rel_diff_vs_" + args.precision + "_gpu": diff, "_out": out2})
        plan2.close()

    # adjoint (transposed operator) on the same plan: cotangents of every output
    adj = None
    if plan.info["mfma"] or not plan.info["has_pw"]:
        gen = torch.Generator(device=dev).manual_seed(7)
        gp = torch.randn((nb, npb, nd), dtype=torch.float64, device=dev, generator=gen)
        gg = torch.randn((nb, npb, 2, nd), dtype=torch.float64, device=dev, generator=gen) if w.ifpgh >= 2 else None
        gh = torch.randn((nb, npb, 3, nd), dtype=torch.float64, device=dev, generator=gen) if w.ifpgh >= 3 else None
        fbar = torch.zeros((nb, npb, nd), dtype=torch.float64, device=dev)
        for _ in range(max(args.warmup, 1)):
            plan.adjoint(gp, fbar, gg, gh)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            plan.adjoint(gp, fbar, gg, gh)
        e1.record(stream)
        torch.cuda.synchronize()
        adj_ms = e0.elapsed_time(e1) / args.steps
        adj = {"ms_per_call": adj_ms, "over_forward": adj_ms / ms,
               "note": "fbar += A^T (gpot, ggrad, ghess) on the same resident plan, FP64"}

    # the host-array drop-in: a cold call (plan + H2D + kernels + D2H, and the
    # pinned staging buffers set up), then the same call again, which reuses
    # the kept plan (H2D + kernels + D2H); then calls that build a new plan
    # every time with the staging buffers in place (plan keeping switched
    # off), as a driver that re-meshes every step makes them
    from heat2d_per_box_amd.boxfgt import boxfgt as boxfgt_host, release_plans

    def host_call():
        t2 = w.tree.copy()
        hp = np.zeros((nb, npb, nd))
        hg = np.zeros((nb, npb, 2, nd)) if w.ifpgh >= 2 else None
        hh = np.zeros((nb, npb, 3, nd)) if w.ifpgh >= 3 else None
        ti = np.zeros(20)
        tc0 = time.perf_counter()
        boxfgt_host(nd, 2, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, npb, nb, t2.nlevels, t2.ltree, t2.itree,
                    t2.iptr, t2.centers, t2.boxsize, w.fvals, w.ifpgh, hp, hg, hh, timeinfo=ti,
                    precision=args.precision)
        wall = time.perf_counter() - tc0
        parts = {"plan_host_s": ti[13], "tables_s": ti[14], "copies_in_s": ti[15], "kernels_s": float(ti[1:7].sum()),
                 "copies_out_s": ti[16]}
        return wall, parts

    release_plans()
    t_cold, cold_parts = host_call()
    t_kept, kept_parts = min((host_call() for _ in range(3)), key=lambda r: r[0])
    os.environ["BFGT_HOST_PLAN_CACHE"] = "0"
    host_call()  # runs on the kept plan one last time and drops it
    t_remesh, remesh_parts = min((host_call() for _ in range(3)), key=lambda r: r[0])
    del os.environ["BFGT_HOST_PLAN_CACHE"]
    release_plans()

    out = {
        "metric": "boxfgt leaf pts/s at eps=1e-9: split-FP32 vs FP64 GPU; adjoint/forward time",
        "value": value,
        "unit": "leaf pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}/{args.scale}: {w.description}",
            "leaves": int(w.tree.nleaves),
            "points": int(w.npts),
            "nd": nd,
            "delta": w.delta,
            "eps": w.eps,
            "ifpgh": w.ifpgh,
            "precision": args.precision,
            "npwlevel": plan.info["npwlevel"],
            "npw": plan.info["npw"],
            "near_pairs": plan.info["n_near_pairs"],
        },
        "settle": {"calls": n_settle, "s": t_settle,
                   "note": "untimed executions before the warm-up, until the GPU runs at its steady clocks"},
        "execution": "overwrite" if ow else "accumulate",
        "ms_per_step_" + ("accumulate" if ow else "overwrite"): ms_other_mode,
        "execution_note": "overwrite: pot (grad, hess) = A fvals on the leaves (bfgt_plan_execute_overwrite); "
                          "accumulate: pot += A fvals (bfgt_plan_execute, the reference's contract, which reads "
                          "the old outputs); same results on zeroed outputs",
        "stage_ms": stage_ms,
        "stage_ms_note": "K more executions, untimed, with HIP events between every stage",
        "t_cold_s": t_remesh,
        "cold_pts_per_s": npts / t_remesh,
        "t_cold_note": "a host-array call that builds its plan (lists, schedules, tables), copies the inputs in, "
                       "runs the kernels and copies the outputs back -- the scope of the reference's call; "
                       "t_first_call_of_process_s adds the one-time set-up of a process (pinned staging buffers) "
                       "and the first touch of freshly allocated host arrays",
        "t_first_call_of_process_s": t_cold,
        "t_host_kept_plan_s": t_kept,
        "host_kept_plan_pts_per_s": npts / t_kept,
        "t_host_new_plan_s": t_remesh,
        "host_call_parts": {"cold": cold_parts, "kept_plan": kept_parts, "new_plan": remesh_parts,
                            "note": "host-array calls (timeinfo slots 14-17 and 2-7): a new plan's input copies "
                                    "run beside its build, so copies_in_s is the wait for them after it"},
        "t_plan_s": t_plan,
        "roofline": roof,
        "stage_roofline": stages_roof,
        "adjoint": adj,
    }
    if not args.no_tree:
        print("[bench] timing the adaptive tree", file=sys.stderr, flush=True)
        out["tree_remesh"] = tree_leg()
        print("[bench] timing ifnewtree=1", file=sys.stderr, flush=True)
        out["newtree"] = newtree_leg()
    outs2 = {}
    for other in others:
        outs2[other["precision"]] = other.pop("_out")
        out[other["precision"]] = other
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, gpu_out, args.cpu_threads,
                                           outs2 or None)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
