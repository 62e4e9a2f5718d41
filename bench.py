#!/usr/bin/env python3
"""Benchmark of the BVH hot path on one MI355X.

Workload (one "step"): over N = 10M points in [0,1]^3 (float32, resident in
HBM before timing starts) build the BVH, run the radius query with about 32
expected neighbours (r = (3*32/(4 pi N))^(1/3), make_intersects(points, r) ->
CRS) and the kNN query k = 16 (make_nearest(points, 16) -> CRS), queries =
the points.  Timed for the quantised (Q16) and FP32 internal-node formats on
the uniform set and on the 1000-component Gaussian mixture.

`value` = whole-step throughput of the headline case (uniform, Q16 nodes):
N / (build + radius + kNN time), in queries per second (each point is
inserted, radius-queried and kNN-queried once per step).  Per-phase rates
(build Mpts/s, radius q/s, kNN q/s) of every case are reported beside it.

The roofline object is for the dominant kernel, with its duration taken from
HIP events the library records around each launch on its own stream and its
algorithmic bytes from the node records the traversal reads (counted by the
kernel itself in an untimed pass).  cpu_baseline times the CPU oracle
(C++/OpenMP restatement of ArborX's Serial/OpenMP path) on bounded samples
of the same workload, on every OpenMP thread and on one thread (median of 5
steps after a warm-up).  The 100M-point moving-points loop (3 steps per node
format by default) follows the 10M cases.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kNN(k=16)+radius queries/s & build Mpts/s, 10M 3D fp32: quantised vs FP32 nodes"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--neighbours", type=int, default=32)
    ap.add_argument("--datasets", default="uniform,gaussian")
    ap.add_argument("--formats", default="q16,fp32")
    ap.add_argument("--cpu-sample", type=int, default=2_000_000,
                    help="points of the CPU baseline on all OpenMP threads")
    ap.add_argument("--cpu-serial-sample", type=int, default=200_000,
                    help="points of the CPU baseline on one thread")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--box-n", type=int, default=5_000_000,
                    help="data and query boxes of the box-primitive case (0 = skip)")
    ap.add_argument("--ray-n", type=int, default=1_000_000,
                    help="rays over the box-primitive case's data boxes (0 = skip)")
    ap.add_argument("--fcc-cells", default="80,80,78",
                    help="FCC cubic cells of the pair-energy case ('' = skip)")
    ap.add_argument("--e2e-steps", type=int, default=3,
                    help="timed steps of the moving-points loop, after one untimed step "
                         "(0 = skip)")
    ap.add_argument("--e2e-n", type=int, default=100_000_000)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this size first, as a known-size "
                         "kernel for scaling hardware byte counters")
    ap.add_argument("--json-out", default=None)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibrate_radius(ax, p, r, target, sample=97, iters=24):
    """Radius with `target` neighbours per point on average (over every
    `sample`-th point, untimed; 97 is prime to the mixture's 1000 centres,
    which point i takes as i % 1000): bracket by doubling / halving from r,
    then bisect on the measured mean until the bracket is one float apart
    or the mean is within 0.05% of the target."""
    import numpy as np
    bvh = ax.BoundingVolumeHierarchy(p)
    q = p[::sample].contiguous()

    def mean(rr):
        _, off = bvh.query(ax.make_intersects(q, float(np.float32(rr))))
        return float(off[-1].item()) / q.shape[0]

    lo, hi = r, r
    while mean(lo) > target:
        lo *= 0.5
    while mean(hi) < target:
        hi *= 2.0
    for _ in range(iters):
        mid = float(np.float32(0.5 * (lo + hi)))
        if not (lo < mid < hi):
            break
        m = mean(mid)
        if abs(m - target) <= 5e-4 * target:
            lo = hi = mid
            break
        if m < target:
            lo = mid
        else:
            hi = mid
    best = min((lo, hi), key=lambda rr: abs(mean(rr) - target))
    return float(np.float32(best))


PMC = os.path.join(ROOT, "profiles", "pmc_kernels_uniform10M.json")
# VALU issue capacity of MI355X: 256 CUs x 4 SIMDs, one wave64 instruction
# per SIMD every 4 cycles (16 lanes wide), 2.4 GHz peak engine clock
VALU_PEAK_INSTS_S = 256 * 4 * 2.4e9 / 4


def measured(kernel, case):
    """The committed rocprofv3 PMC figures of `kernel` (its base name) in
    `case` ("uniform/q16", "boxes/fp32", ...): HBM bytes per dispatch
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known size) and
    SQ instruction counts; None when the workload was not profiled."""
    if not os.path.exists(PMC):
        return None
    with open(PMC) as f:
        data = json.load(f)
    base = kernel.strip("()").split("<")[0]
    return data.get("cases", {}).get(case, {}).get(base)


def timed(fn, steps, warmup, stream):
    """Mean ms of fn() over `steps` runs after `warmup`, from HIP events on
    the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for s in range(steps):
        fn()
        ev[s + 1].record(stream)
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[-1]) / steps


def kernel_ms(ax, fn, match):
    """Total ms of the library's kernels whose name contains one of `match`
    during one fn() call (HIP events around each launch)."""
    with ax.profile(timing=True, visits=False) as prof:
        fn()
    return sum(sum(v) for k, v in prof.kernels.items() if any(m in k for m in match))


def bench_boxes(ax, args, torch, synthetic, stream):
    """5M lognormal boxes, 5M query boxes drawn the same way: intersects()
    CRS with FP32 and Q16 internal nodes; identical results are checked;
    bytes per query from the node records the traversal reads."""
    n = args.box_n
    data = torch.from_numpy(synthetic.lognormal_boxes(n, 4)).cuda()
    qb = torch.from_numpy(synthetic.lognormal_boxes(n, 5)).cuda()
    pred = ax.intersects(ax.Boxes(qb))
    out, ref = {}, None
    for fmt in ("fp32", "q16"):
        tb = timed(lambda: ax.BoundingVolumeHierarchy(data, node_format=fmt), args.steps,
                   args.warmup, stream)
        bvh = ax.BoundingVolumeHierarchy(data, node_format=fmt)
        tq = timed(lambda: bvh.query(pred), args.steps, args.warmup, stream)
        trav = kernel_ms(ax, lambda: bvh.query(pred), ("k_spatial",))
        idx, off = bvh.query(pred)
        if ref is None:
            ref = (idx, off)
        identical = bool(torch.equal(off, ref[1]) and torch.equal(idx, ref[0]))
        with ax.profile(timing=False, visits=True) as pv:
            bvh.query(pred)
        v = pv.visits
        s_int = 16 if fmt == "q16" else 32
        # query box 24 B + order 4 B + count 8 B; node records (box leaf 32 B);
        # 4 B per result
        alg = v["queries"] * 36 + v["internal"] * s_int + v["leaf"] * 32 + v["results"] * 4
        out[fmt] = dict(build_ms=tb, query_ms=tq, traversal_kernel_ms=trav,
                        query_qps=n / tq * 1e3, hits_per_query=v["results"] / n,
                        internal_visits_per_query=v["internal"] / n,
                        leaf_visits_per_query=v["leaf"] / n,
                        bytes_per_query=alg / n, traversal_alg_GBps=alg / (trav * 1e-3) / 1e9,
                        identical_to_fp32=identical)
        pm = measured("k_spatial", f"boxes/{fmt}")
        if pm:
            # measured HBM bytes of the traversal (rocprofv3 PMC, committed)
            out[fmt]["traversal_hbm_bytes_per_query"] = pm["hbm_bytes"] / n
            out[fmt]["traversal_hbm_GBps"] = pm["hbm_bytes"] / (trav * 1e-3) / 1e9
        log(f"[bench] boxes/{fmt}: build {tb:.2f} ms | query {tq:.2f} ms "
            f"({n / tq / 1e3:.1f} Mq/s, {v['results'] / n:.3f} hits/q) | traversal "
            f"{trav:.2f} ms, {alg / n:.0f} B/q, {out[fmt]['traversal_alg_GBps']:.0f} GB/s "
            "(algorithmic) | "
            f"identical {identical}")
    out["q16_traversal_speedup"] = out["fp32"]["traversal_kernel_ms"] / out["q16"]["traversal_kernel_ms"]
    out["config"] = f"{n} data boxes (seed 4) and {n} query boxes (seed 5): centres U[0,1)^3, " \
                    "sides lognormal(ln 0.002, 0.5)"
    return out


def bench_rays(ax, args, torch, synthetic, stream):
    """Rays over the box case's data boxes: Intersects<Ray>, ordered rows
    (ordered_intersects) and nearest(ray, k) for k = 1 and 8, FP32 and Q16
    nodes; the two formats' rows are checked identical."""
    n, nr = args.box_n, args.ray_n
    data = torch.from_numpy(synthetic.lognormal_boxes(n, 4)).cuda()
    o, d = synthetic.rays(nr, 9)
    rays = ax.Rays(torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda())
    preds = {"intersects": ax.intersects(rays), "ordered": ax.ordered_intersects(rays),
             "nearest1": ax.nearest(rays, 1), "nearest8": ax.nearest(rays, 8)}
    out, ref = {}, {}
    for fmt in ("fp32", "q16"):
        bvh = ax.BoundingVolumeHierarchy(data, node_format=fmt)
        res = {}
        for name, p in preds.items():
            t = timed(lambda: bvh.query(p), args.steps, args.warmup, stream)
            idx, off = bvh.query(p)
            if fmt == "fp32" and name != "intersects":
                ref[name] = (idx, off)
            same = name == "intersects" or fmt == "fp32" or bool(
                torch.equal(off, ref[name][1]) and torch.equal(idx, ref[name][0]))
            res[name] = dict(ms=t, rays_per_s=nr / t * 1e3, hits=int(idx.numel()),
                             identical_to_fp32=same)
        out[fmt] = res
        log(f"[bench] rays/{fmt}: " + " | ".join(
            f"{k} {v['ms']:.2f} ms ({v['rays_per_s'] / 1e6:.1f} Mrays/s, {v['hits']} hits)"
            for k, v in res.items()) + " | identical " +
            str(all(v["identical_to_fp32"] for v in res.values())))
        del bvh
    out["config"] = f"{nr} rays (origins U[-0.2,1.2)^3, Gaussian directions, every fifth " \
                    f"along an axis, seed 9) over the {n} data boxes of the box case"
    return out


def bench_energy(ax, args, torch, synthetic, stream):
    """FCC lattice (nearest-neighbour distance 1, jitter +-0.05), cutoff 2.5:
    pair energy forward only vs fused forward + backward, and the gradient
    against FP64 central differences on a 1000-atom subregion."""
    import numpy as np
    cells = tuple(int(c) for c in args.fcc_cells.split(","))
    x_h = synthetic.fcc_lattice(cells, seed=6)
    n = x_h.shape[0]
    x = torch.from_numpy(x_h).cuda()
    rc = 2.5
    off, idx = ax.energy.neighbour_csr(x, rc)
    nnz = int(off[-1].item())
    t_csr = timed(lambda: ax.energy.neighbour_csr(x, rc), args.steps, args.warmup, stream)
    t_f = timed(lambda: ax.pair_energy(x, off, idx, rc, want_grad=False), args.steps,
                args.warmup, stream)
    t_fb = timed(lambda: ax.pair_energy(x, off, idx, rc), args.steps, args.warmup, stream)
    k_f = kernel_ms(ax, lambda: ax.pair_energy(x, off, idx, rc, want_grad=False),
                    ("k_energy", "k_reduce"))
    k_fb = kernel_ms(ax, lambda: ax.pair_energy(x, off, idx, rc), ("k_energy", "k_reduce"))

    def autograd_step():
        xr = x.detach().requires_grad_(True)
        e = ax.PairEnergy.apply(xr, rc, off, idx, False)
        e.backward()

    def autograd_recompute():
        xr = x.detach().requires_grad_(True)
        e = ax.PairEnergy.apply(xr, rc, None, None, True)
        e.backward()

    t_ag = timed(autograd_step, args.steps, args.warmup, stream)
    t_agr = timed(autograd_recompute, max(1, args.steps // 2), 1, stream)
    e, g = ax.pair_energy(x, off, idx, rc)
    # 1000 atoms nearest the lattice centre
    centre = x_h.mean(axis=0)
    atoms = np.argsort(((x_h - centre) ** 2).sum(1), kind="stable")[:1000].astype(np.int32)
    fd = ax.pair_energy_fd(x, off, idx, rc, atoms, h=1e-6).cpu().numpy()
    ga = g.cpu().numpy().astype(np.float64)[atoms]
    rel = float(np.linalg.norm(ga - fd) / np.linalg.norm(fd))
    alg_f = n * 12 + nnz * 16 + (n + 1) * 8
    out = dict(n_points=n, cells=list(cells), cutoff=rc, mean_neighbours=nnz / n,
               energy=float(e.item()), csr_ms=t_csr, forward_ms=t_f, fwd_bwd_ms=t_fb,
               fwd_bwd_over_forward=t_fb / t_f, forward_kernel_ms=k_f, fwd_bwd_kernel_ms=k_fb,
               forward_alg_GBps=alg_f / (k_f * 1e-3) / 1e9,
               fwd_bwd_alg_GBps=(alg_f + n * 12) / (k_fb * 1e-3) / 1e9,
               autograd_step_ms=t_ag, autograd_recompute_csr_step_ms=t_agr,
               grad_rel_err_vs_fp64_fd=rel, fd_atoms=1000, fd_h=1e-6)
    log(f"[bench] energy: N={n}, {nnz / n:.1f} nbrs | CSR {t_csr:.2f} ms | fwd {t_f:.3f} ms "
        f"| fwd+bwd {t_fb:.3f} ms (x{t_fb / t_f:.2f}) | autograd {t_ag:.2f} ms, with CSR "
        f"recompute {t_agr:.2f} ms | grad rel err vs FP64 FD {rel:.2e}")
    return out


def bench_e2e(ax, args, torch, synthetic, stream, fmt):
    """Moving-points loop: each step displaces every point by N(0, 1e-4) per
    axis, rebuilds the tree, runs kNN k=32 and the pair energy forward +
    backward with the CRS recomputed: built from the moved points inside the
    forward pass (its own Q16 tree and full neighbour list) and freed when
    the forward pass returns; the fused kernel gives E and dE/dx in that one
    pass, and backward applies the kept gradient.  The energy sees the coordinates scaled
    by 2.5 / r32 (r32 = radius of about 32 neighbours), i.e. LJ with
    sigma = r32 / 2.5 and cutoff 2.5 sigma, so that pair terms stay finite."""
    n = args.e2e_n
    x = torch.from_numpy(synthetic.uniform_points(n, 7)).cuda()
    r32 = float(synthetic.radius_for_neighbours(n, 32))
    rc = 2.5
    scale = rc / r32
    gen = torch.Generator(device="cuda")
    torch.cuda.synchronize()
    ax.trim_memory()  # what the previous case left cached does not count here
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    free0, total = torch.cuda.mem_get_info()
    tb = tk = te = 0.0
    finite = True
    t0 = time.perf_counter()
    # step 0 is untimed: it grows the library's pools and torch's cache from
    # empty (the device-memory peak below still includes it)
    for s in range(args.e2e_steps + 1):
        gen.manual_seed(8 + s)
        x.add_(torch.randn(x.shape, generator=gen, device="cuda") * 1e-4)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        bvh = ax.BoundingVolumeHierarchy(x, node_format=fmt)
        ev[1].record(stream)
        ki, ko = bvh.query(ax.make_nearest(x, 32))
        del ki, ko
        ev[2].record(stream)
        del bvh
        xr = x.detach().requires_grad_(True)
        e = ax.PairEnergy.apply(xr * scale, rc, None, None, False)
        e.backward()
        ev[3].record(stream)
        torch.cuda.synchronize()
        finite = finite and bool(torch.isfinite(e).item())
        if s > 0:
            tb += ev[0].elapsed_time(ev[1])
            tk += ev[1].elapsed_time(ev[2])
            te += ev[2].elapsed_time(ev[3])
        del xr, e
        if s % 10 == 0:
            log(f"[bench] e2e/{fmt} step {s}: {time.perf_counter() - t0:.1f}s")
    k = args.e2e_steps
    peak = torch.cuda.max_memory_allocated() - base
    # the library's stream-ordered pools keep their high-water mark, so the
    # device memory in use after the loop is the loop's peak (plus torch's
    # cache); torch's own counter sees only the result arrays
    free1, _ = torch.cuda.mem_get_info()
    dev_peak = free0 - free1
    out = dict(n_points=n, steps=k, cutoff=r32, build_ms=tb / k, knn_ms=tk / k,
               energy_fwd_bwd_ms=te / k, step_ms=(tb + tk + te) / k,
               peak_hbm_GB=dev_peak / 1e9, peak_torch_allocated_GB=peak / 1e9,
               energy_finite=finite)
    log(f"[bench] e2e/{fmt}: build {tb / k:.1f} ms | kNN k=32 {tk / k:.1f} ms | energy fwd+bwd "
        f"(CRS recomputed) {te / k:.1f} ms | step {(tb + tk + te) / k:.1f} ms | peak device "
        f"memory {dev_peak / 1e9:.1f} GB (torch arrays {peak / 1e9:.1f} GB)")
    return out


def sig(v, digits=6):
    """Floats of a JSON tree to `digits` significant digits (small values
    such as relative errors keep their magnitude)."""
    if isinstance(v, float):
        return float(f"{v:.{digits}g}") if math.isfinite(v) else v
    if isinstance(v, dict):
        return {k: sig(x, digits) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [sig(x, digits) for x in v]
    return v


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_step_times(oracle, np, synthetic, n, neighbours, k, reps):
    """Per-phase seconds of `reps` repetitions (after one warm-up) of one
    step -- build, two-pass radius CRS, kNN -- of the oracle over the first
    n points of the uniform set, with the radius for `neighbours` at that
    N.  Wall clock around each phase (the oracle returns when its OpenMP
    loops have joined)."""
    pts = synthetic.uniform_points(n, 2)
    r = np.float32(synthetic.radius_for_neighbours(n, neighbours))
    sph = np.concatenate([pts, np.full((n, 1), r, np.float32)], 1)
    out = []
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        t = oracle.Tree(pts)
        t1 = time.perf_counter()
        t.query_spatial(0, sph)
        t2 = time.perf_counter()
        t.query_knn(pts, k)
        t3 = time.perf_counter()
        del t
        if rep > 0:
            out.append((t1 - t0, t2 - t1, t3 - t2))
    return out


def cpu_baseline(args, K, synthetic):
    """The oracle (C++/OpenMP restatement of ArborX's Serial/OpenMP path) on
    the host: median of 5 steps after a warm-up, on every OpenMP thread and
    on one thread ("Serial"), each on a bounded sample of the uniform
    workload at the same neighbour count and k."""
    import numpy as np

    import oracle

    threads = oracle.num_threads()
    res = {}
    for label, nt, n in (("all_threads", threads, min(args.cpu_sample, args.n)),
                         ("serial", 1, min(args.cpu_serial_sample, args.n))):
        oracle.set_num_threads(nt)
        reps = cpu_step_times(oracle, np, synthetic, n, args.neighbours, K, 5)
        step = sorted(sum(r) for r in reps)
        med = step[len(step) // 2]
        res[label] = dict(threads=nt, n_points=n, median_step_s=med,
                          queries_per_s=n / med,
                          median_build_s=sorted(r[0] for r in reps)[2],
                          median_radius_s=sorted(r[1] for r in reps)[2],
                          median_knn_s=sorted(r[2] for r in reps)[2],
                          steps_s=[sum(r) for r in reps])
    oracle.set_num_threads(threads)
    a, s1 = res["all_threads"], res["serial"]
    return dict(value=round(a["queries_per_s"], 1), unit="queries/s", cores=threads,
                kind="port", cpu_model=cpu_model(),
                serial_value=round(s1["queries_per_s"], 1),
                sample=(f"CPU restatement of ArborX Serial/OpenMP (the oracle): one step = "
                        f"build + two-pass radius CRS ({args.neighbours} nbrs) + kNN k={K} over "
                        f"the first N points of the uniform set, queries = points; median of 5 "
                        f"after 1 warm-up; value: N={a['n_points']} on {threads} OpenMP threads "
                        f"({a['median_step_s']:.2f} s/step); serial_value: "
                        f"N={s1['n_points']} on 1 thread ({s1['median_step_s']:.2f} s/step)"),
                runs=res)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the index "
              "is built and queried on one device; sharding over several GPUs is a design, "
              "not code, here", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import arborx_amd as ax
    from arborx_amd import synthetic

    if args.copy_calibration_mib > 0:
        words = args.copy_calibration_mib * (1 << 18)
        src = torch.ones(words, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: {args.copy_calibration_mib} MiB read + "
            f"{args.copy_calibration_mib} MiB written")
        del src, dst

    N, K = args.n, args.k
    r = float(np.float32(synthetic.radius_for_neighbours(N, args.neighbours)))
    gens = {
        "uniform": lambda: synthetic.uniform_points(N, 2),
        "gaussian": lambda: synthetic.gaussian_mixture_points(N, 3, n_centres=1000, sigma=0.01),
    }
    datasets = [d for d in args.datasets.split(",") if d]
    formats = [f for f in args.formats.split(",") if f]
    pts = {}
    radius = {}
    for d in datasets:
        t0 = time.time()
        pts[d] = torch.from_numpy(gens[d]()).cuda()
        log(f"[bench] generated {d} N={N} in {time.time() - t0:.1f}s")
        radius[d] = r if d == "uniform" else calibrate_radius(ax, pts[d], r, args.neighbours)
        log(f"[bench] {d}: radius {radius[d]:.7g} for ~{args.neighbours} neighbours")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def step(p, fmt, ev=None, rq=r):
        if ev:
            ev[0].record(stream)
        bvh = ax.BoundingVolumeHierarchy(p, node_format=fmt)
        if ev:
            ev[1].record(stream)
        idx, off = bvh.query(ax.make_intersects(p, rq))
        if ev:
            ev[2].record(stream)
        ki, ko = bvh.query(ax.make_nearest(p, K))
        if ev:
            ev[3].record(stream)
        return int(off[-1].item()), ki.numel()

    results = {}
    # Untimed steps before the first case: the GPU leaves its idle clocks
    # during the first tens of milliseconds of work, which would otherwise
    # fall on the first case's warm-up (its kernels then run ~8% slower
    # there, which the kernel-trace averages would include).
    if datasets and formats:
        for _ in range(2):
            step(pts[datasets[0]], formats[0], rq=radius[datasets[0]])
        torch.cuda.synchronize()
    for d in datasets:
        for fmt in formats:
            p = pts[d]
            for _ in range(args.warmup):
                step(p, fmt, rq=radius[d])
            torch.cuda.synchronize()
            evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)]
                   for _ in range(args.steps)]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            nnz = 0
            for s in range(args.steps):
                nnz, knn_nnz = step(p, fmt, evs[s], rq=radius[d])
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t0) / args.steps
            tb = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
            tr = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps
            tk = sum(e[2].elapsed_time(e[3]) for e in evs) / args.steps
            res = dict(ms_per_step=wall * 1e3, build_ms=tb, radius_ms=tr, knn_ms=tk,
                       build_Mpts_s=N / tb / 1e3, radius_qps=N / tr * 1e3, knn_qps=N / tk * 1e3,
                       knn_plus_radius_qps=N / (tr + tk) * 1e3,
                       step_qps=N / (wall * 1e3) * 1e3, radius_results=nnz,
                       mean_neighbours=nnz / N, radius=radius[d])
            results[f"{d}/{fmt}"] = res
            log(f"[bench] {d}/{fmt}: step {wall * 1e3:.2f} ms | build {tb:.2f} ms "
                f"({res['build_Mpts_s']:.0f} Mpts/s) | radius {tr:.2f} ms "
                f"({res['radius_qps'] / 1e6:.1f} Mq/s, {nnz / N:.1f} nbrs) | kNN {tk:.2f} ms "
                f"({res['knn_qps'] / 1e6:.1f} Mq/s)")

    head_key = "uniform/q16" if "uniform/q16" in results else next(iter(results))
    head = results[head_key]

    # ---- per-kernel profile of one step of the headline case + visit counts
    roofline = None
    kernels = {}
    if not args.no_profile:
        d, fmt = head_key.split("/")
        p = pts[d]
        with ax.profile(timing=True, visits=False) as prof:
            step(p, fmt, rq=radius[d])
        kernels = {k: dict(launches=len(v), total_ms=sum(v)) for k, v in prof.kernels.items()
                   if not k.startswith("host:")}
        host = {k: dict(calls=len(v), total_ms=sum(v)) for k, v in prof.kernels.items()
                if k.startswith("host:")}
        log("[bench] host time (one step): " +
            ", ".join(f"{k}: {v['total_ms']:.3f} ms x{v['calls']}" for k, v in host.items()))
        dom = max(kernels.items(), key=lambda kv: kv[1]["total_ms"])
        log("[bench] kernel times (one step): " +
            ", ".join(f"{k}: {v['total_ms']:.3f} ms x{v['launches']}"
                      for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["total_ms"])[:8]))
        # algorithmic bytes of the traversal kernels, from the node records
        # the kernels themselves count in an untimed pass
        bvh = ax.BoundingVolumeHierarchy(p, node_format=fmt)
        s_int = 16 if fmt == "q16" else 32
        with ax.profile(timing=False, visits=True) as pv:
            bvh.query(ax.make_intersects(p, r))
        vr = pv.visits
        with ax.profile(timing=False, visits=True) as pv:
            bvh.query(ax.make_nearest(p, K))
        vk = pv.visits

        def alg_bytes(name):
            """Bytes the algorithm must move for all launches of one call."""
            if "k_knn_collect" in name:
                # query 12 B + order 4 B; node records; each candidate written
                # to the query's row and read back (8 B each way); k indices out
                return (vk["queries"] * (16 + 4 * K) + vk["internal"] * s_int
                        + vk["leaf"] * 16 + vk["results"] * 16), "kNN query"
            if "MODE_COUNT" in name:
                # query 12 B + order 4 B + count 8 B; node records
                return (vr["queries"] * 24 + vr["internal"] * s_int + vr["leaf"] * 16,
                        "radius query (count pass)")
            if "MODE_BUFFER" in name:
                # query 12 B + order 4 B + count 8 B; node records; 4 B per
                # result kept in the query's scratch row
                return (vr["queries"] * 24 + vr["internal"] * s_int + vr["leaf"] * 16
                        + vr["results"] * 4, "radius query (buffered pass)")
            if "MODE_FILL" in name:
                # query 12 B + order 4 B + offsets 16 B; node records; 4 B per result
                return (vr["queries"] * 32 + vr["internal"] * s_int + vr["leaf"] * 16
                        + vr["results"] * 4, "radius query (fill pass)")
            return None, None

        name = dom[0]
        alg, per = alg_bytes(name)
        tot_ms = dom[1]["total_ms"]
        if alg is None:  # fall back to the largest traversal kernel
            cands = [(k, v) for k, v in kernels.items() if alg_bytes(k)[0] is not None]
            name, v = max(cands, key=lambda kv: kv[1]["total_ms"])
            alg, per = alg_bytes(name)
            tot_ms = v["total_ms"]
        launches = kernels[name]["launches"]
        achieved = alg / (tot_ms * 1e-3) / 1e9
        avg_ms = tot_ms / launches
        pm = measured(name, head_key)
        traffic = int(pm["hbm_bytes"]) if pm else None
        roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                        unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                        kernel=name, launches_per_call=launches,
                        avg_launch_ms=round(avg_ms, 4),
                        alg_bytes_per_launch=int(alg / launches),
                        alg_bytes_per_unit=round(alg / N, 1), unit_of_work=per,
                        dominant_kernel=dom[0], visits_radius=vr, visits_knn=vk)
        if pm:
            # What actually binds the kernel, from the committed rocprofv3
            # counters of the same workload: measured HBM bytes per launch
            # against the 8 TB/s peak, and VALU wave-instructions per launch
            # against the device's issue rate.  Algorithmic bytes exceed the
            # measured ones: node records hit in L2 and the Infinity Cache.
            hbm = pm["hbm_bytes"] / (avg_ms * 1e-3) / 1e9
            issue_ms = pm["SQ_INSTS_VALU"] / VALU_PEAK_INSTS_S * 1e3
            roofline.update(
                hbm_measured_GBps=round(hbm, 1), hbm_measured_frac=round(hbm / HBM_PEAK_GBS, 4),
                valu=dict(insts_per_launch=int(pm["SQ_INSTS_VALU"]),
                          peak_insts_per_s=VALU_PEAK_INSTS_S, issue_ms=round(issue_ms, 4),
                          issue_frac=round(issue_ms / avg_ms, 4)),
                binding=("VALU issue" if issue_ms / avg_ms >= 0.7 else "memory latency"),
                counters_source=os.path.relpath(PMC, ROOT))
        log(f"[bench] roofline: {roofline}")

    # ---- CPU baseline: oracle on host cores, bounded samples
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, K, synthetic)
        log(f"[bench] cpu baseline: {cpu}")

    extra = {}
    if args.box_n > 0:
        extra["boxes"] = bench_boxes(ax, args, torch, synthetic, stream)
        if args.ray_n > 0:
            extra["rays"] = bench_rays(ax, args, torch, synthetic, stream)
    if args.fcc_cells:
        extra["energy"] = bench_energy(ax, args, torch, synthetic, stream)
    if args.e2e_steps > 0:
        del pts
        torch.cuda.empty_cache()
        extra["e2e"] = {fmt: bench_e2e(ax, args, torch, synthetic, stream, fmt)
                        for fmt in ("q16", "fp32")}

    line = {
        "metric": METRIC,
        "value": round(head["step_qps"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "10M uniform points in [0,1]^3 (and 10M 1000-component "
                               "Gaussian mixture): BVH build + radius (~32 nbrs) + kNN k=16, "
                               "queries = points",
                   "n_points": N, "radius": radius, "k": K, "headline": head_key},
        "cases": sig(results),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    for key, v in extra.items():
        line[key] = sig(v)
    if kernels:
        line["kernels_one_step_ms"] = sig({k: v["total_ms"] for k, v in kernels.items()})
    s = json.dumps(line)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(s + "\n")
    print(s, flush=True)


if __name__ == "__main__":
    main()
