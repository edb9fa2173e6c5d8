#!/usr/bin/env python3
"""Headline benchmark of the restricted power diagram hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 1,000,000 seeds
i.i.d. uniform in [0,1]^3 with power weights w ~ U[0, 0.25 h^2], h = N^-1/3,
in the [0,1000]^3 frame; domain = solid torus R=0.3 r=0.12 (axis z, centred)
voxelized at 128^3; caps K=64 P=40 V=60 T=100.  One step = one full solve
(binning + kNN + clip/restrict/integrate + every re-queue pass) from seeds
already resident in HBM, for the mixed form and for the all-FP64 form.

Prints ONE JSON line.  `value` = mixed-form cells/s.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cells/sec @1M seeds mixed vs FP64 at ≤1e-6 rel err; p50/p99 req latency"
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (published)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_flops(st, n_cells, mixed):
    """Algorithmic flops per cell of the two clip kernels, from the oracle's
    event counters on the same inputs (single pass, with the library's culls:
    cells proven outside the domain are not clipped, so they add no work):
        F = c_pred*n_pred + 50*n_radius_vertex + 25*n_planes + 41*n_bbox_vertex
            + 420*n_integrated_vertex
    (op counts of ConvexCell.cl:79-97, 416-448, 344-373, 314-328, 930-935,
    577-619; c_pred = 109 for the filtered FP32 test, 45 for FP64).  The
    bisector kernel does the predicates, radius checks and bisector planes of
    the neighbour loop; the finish kernel the cone-plane predicates and
    planes, the bbox and the integration.  Returns {kernel: (fp32, fp64)}."""
    pred_a = st["n_pred"] - st["n_pred_restrict"]
    pred_b = st["n_pred_restrict"]
    planes_a = st["n_planes"] - st["n_planes_restrict"]
    planes_b = st["n_planes_restrict"]
    if mixed:
        a = (109.0 * pred_a + 50.0 * st["n_radius_vertex"], 25.0 * planes_a)
        b = (109.0 * pred_b, 25.0 * planes_b + 41.0 * st["n_bbox_vertex"] + 420.0 * st["n_integrated_vertex"])
    else:
        a = (0.0, 45.0 * pred_a + 50.0 * st["n_radius_vertex"] + 25.0 * planes_a)
        b = (0.0, 45.0 * pred_b + 25.0 * planes_b + 41.0 * st["n_bbox_vertex"] + 420.0 * st["n_integrated_vertex"])
    return {"cells": (a[0] / n_cells, a[1] / n_cells), "finish": (b[0] / n_cells, b[1] / n_cells)}


def valu_roofline(f32, f64, n, seconds, kernel, traffic):
    achieved = (f32 + f64) * n / seconds / 1e12
    peak = (f32 + f64) / (f32 / PEAK_FP32_TFLOPS + f64 / PEAK_FP64_TFLOPS)
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": traffic, "kernel": kernel, "flops_per_cell": {"fp32": f32, "fp64": f64},
            "launch_ms": seconds * 1e3}


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel on this workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes kept in profiles/ (scaled by the counters'
    reading for a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_1M_torus.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["bytes_per_launch"]


def rel_errors(got, ref, n):
    """Errors of moments got against ref: of the mean cell volume (all cells),
    and per cell on the cells above 1% of the mean cell volume (relative
    volume, barycentre displacement over the cell's size)."""
    v = 1e9 / n
    ev = float(np.abs(got[:, 3] - ref[:, 3]).max() / v)
    em = float(np.abs(got[:, :3] - ref[:, :3]).max() / (v * 1000.0))
    big = ref[:, 3] > 0.01 * v
    if big.any():
        rv = float((np.abs(got[big, 3] - ref[big, 3]) / ref[big, 3]).max())
        b_got = got[big, :3] / got[big, 3:4]
        b_ref = ref[big, :3] / ref[big, 3:4]
        rb = float((np.abs(b_got - b_ref).max(1) / np.cbrt(ref[big, 3])).max())
    else:
        rv = rb = 0.0
    return {"vol_over_mean_cell": ev, "moment_over_mean_cell_x1000": em, "vol_rel_cells_gt_1pct": rv,
            "bary_over_cell_size": rb, "cells_gt_1pct": int(big.sum())}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


SLOTS = 2  # batches in flight in the serving stream (2 streams each: the 4 hardware queues)


def serving(rpd, ctx, n_req, log, slots=SLOTS):
    """Serving stream: Poisson arrivals, seed counts log-uniform in [1e3, 1e6],
    uniform seeds with w ~ U[0, 0.25 h^2], one of 4 resident domains at 96^3,
    successive requests on one seed set with weights perturbed by
    N(0, 0.01 h^2); mixed form, K50 P32 V48.  The arrival rate is set to 70%
    of the throughput measured by a back-to-back pass of the same requests."""
    rng = np.random.default_rng(1)
    doms = [ctx.domain(rpd.Grid.shape(code, 96, prm)) for code, prm in
            ((rpd.SHAPE_SPHERE, (0.5, 0.5, 0.5, 0.45)), (rpd.SHAPE_TORUS, (0.5, 0.5, 0.5, 0.3, 0.12)),
             (rpd.SHAPE_CUBE, ()), (rpd.SHAPE_CUBE_MINUS_SPHERE, (0.5, 0.5, 0.5, 0.3)))]
    n_sets = 24
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n_sets)).astype(np.int64)
    sets = []
    for i, n in enumerate(sizes):
        h = (1.0 / n) ** (1.0 / 3.0)
        sets.append((rpd.generate_seeds(rpd.SEEDS_UNIFORM, int(n), seed=1000 + i, wmax=0.25 * h * h * 1e6), h))
    reqs = []
    for r in range(n_req):
        k = int(rng.integers(n_sets))
        x, h = sets[k]
        x = x.copy()
        x[:, 3] = np.maximum(0.0, x[:, 3] + rng.normal(0.0, 0.01 * h * h * 1e6, x.shape[0]))
        sets[k] = (x, h)
        reqs.append((int(rng.integers(len(doms))), x))
    caps = dict(K=50, P=32, V=48, T=100, precision=rpd.PREC_MIXED)
    cells = sum(x.shape[0] for _, x in reqs)

    def run(rate):
        srv = rpd.Server(ctx, max_inflight=slots)
        lat = []
        done = threading.Event()

        err = []

        def poller():
            try:
                while len(lat) < len(reqs):
                    r = srv.poll(wait=True)
                    if r is None:
                        time.sleep(0.0005)
                        continue
                    lat.append(r[2])
            except Exception as e:  # surfaced by the main thread
                err.append(e)
            done.set()

        th = threading.Thread(target=poller)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
        for (d, x), gap in zip(reqs, gaps):
            t_next += gap
            while True:
                now = time.perf_counter()
                if now >= t_next:
                    break
                time.sleep(min(0.001, t_next - now))
            srv.submit(doms[d], x, **caps)
        if not done.wait(timeout=300):
            raise RuntimeError("serving stream did not finish")
        t1 = time.perf_counter()
        th.join()
        srv.close()
        if err:
            raise err[0]
        return t1 - t0, np.array(lat)

    run(0)  # warm the slots
    t_b2b, _ = run(0)
    rate = 0.7 * len(reqs) / t_b2b
    t_str, lat = run(rate)
    for d in doms:
        d.close()
    res = {"requests": len(reqs), "cells": int(cells), "arrival_rate_per_s": rate,
           "backtoback_cells_per_s": cells / t_b2b, "stream_cells_per_s": cells / t_str,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)), "slots": slots,
           "batching": "packed batches (rpd_solve_batch) of the queued requests, up to 2^19 seeds each",
           "workload": "Poisson arrivals at 70% of back-to-back throughput; n log-uniform [1e3,1e6]; 4 domains "
                       "@96^3; weights perturbed N(0,0.01h^2) per reuse; mixed K50 P32 V48"}
    log("bench: serving", res)
    return res


def mesh_grid_bench(args):
    """Mesh-to-grid generator (the mesh path of utils/createGrid): a torus
    mesh R=3 r=1.1 of 800x400 quads (640,000 triangles) voxelized at
    --mesh-division.  GPU: rpd_grid_from_mesh wall time, host arrays in and a
    host grid out.  CPU leg: the oracle's per-voxel list loop (the
    reference's main.cpp:178-219 without its tetrahedralization), one thread,
    and a bit-exact comparison of the lists."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import __graft_entry__ as ge
    from meshes import torus

    rpd = ge.load_package()
    g = args.mesh_division
    P, T = torus(3.0, 1.1, 800, 400, rot_seed=12)
    times = []
    for s in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        grid = rpd.Grid.mesh(P, T, g)
        dt = time.perf_counter() - t0
        if s >= args.warmup:
            times.append(dt)
        log("bench: mesh grid step %d %.1f ms" % (s, 1e3 * dt))
    out = {"workload": "mesh-grid", "triangles": int(T.shape[0]), "division": g, "entries": int(grid.tri_list.size),
           "ms_min": 1e3 * min(times), "ms_median": 1e3 * float(np.median(times)),
           "corner_flags": {str(k): int(v) for k, v in zip(*np.unique(grid.in_domain, return_counts=True))}}
    if not args.no_cpu:
        orc = ge.load_oracle()
        xyz, tris = orc.mesh_points(P, T)
        t0 = time.perf_counter()
        offs, lst = orc.mesh_lists(xyz, tris, g)
        out["cpu_lists_ms"] = 1e3 * (time.perf_counter() - t0)
        out["cpu_cores"] = 1
        out["lists_bit_exact"] = bool(np.array_equal(offs, grid.offsets) and np.array_equal(lst, grid.tri_list))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000, help="cells timed on the CPU oracle")
    ap.add_argument("--serve-requests", type=int, default=240, help="requests in the serving stream")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-only", action="store_true",
                    help="only run the serving stream (prints its JSON object)")
    ap.add_argument("--serve-slots", type=int, default=SLOTS, help="batches in flight in the serving stream")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (reference kernel for HBM counter scaling)")
    ap.add_argument("--mesh-grid", action="store_true",
                    help="only time the mesh-to-grid generator (torus mesh of 640k triangles, --mesh-division)")
    ap.add_argument("--mesh-division", type=int, default=256)
    args = ap.parse_args()

    if args.mesh_grid:
        mesh_grid_bench(args)
        return

    if args.serving_only:
        import __graft_entry__ as ge
        rpd = ge.load_package()
        ctx = rpd.Context(0)
        print(json.dumps(serving(rpd, ctx, args.serve_requests, log, args.serve_slots)), flush=True)
        ctx.close()
        return

    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("bench: copied %d bytes" % (a.numel() * 4))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the hot path is a single-device kernel; "
              "multi-GPU sharding is described in DESIGN.md and not run). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge

    rpd = ge.load_package()
    n = args.n
    h = (1.0 / n) ** (1.0 / 3.0)
    wmax = 0.25 * h * h * 1e6  # frame units: lengths x1e3, weights x1e6
    K, P, V, T = 64, 40, 60, 100
    log("bench: generating %d seeds and the %d^3 torus" % (n, args.grid))
    seeds = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=0, wmax=wmax)
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, args.grid, (0.5, 0.5, 0.5, 0.3, 0.12))

    ctx = rpd.Context(0)
    dom = ctx.domain(grid)
    dseeds = torch.from_numpy(seeds).to("cuda:0")
    torch.cuda.synchronize()

    def step(prec):
        ctx.set_seeds_device(dseeds.data_ptr(), n)
        return ctx.solve(dom, K=K, P=P, V=V, T=T, precision=prec)

    res = {}
    for name, prec in (("mixed", rpd.PREC_MIXED), ("fp64", rpd.PREC_FP64)):
        for _ in range(args.warmup):
            step(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [step(prec) for _ in range(args.steps)]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) / args.steps * 1e3
        res[name] = {"ms": ms, "stats": stats, "out": ctx.moments(), "status": ctx.status()}
        s = stats[-1]
        log("bench: %s %.2f ms/solve  pass1 knn %.2f ms clip %.2f ms  passes %d fallback %d hist %s" %
            (name, ms, s["ms_knn_first"], s["ms_clip_first"], s["passes"], s["fallback_cells"],
             s["status_hist"]))

    err_mixed_vs_fp64 = rel_errors(res["mixed"]["out"], res["fp64"]["out"], n)
    log("bench: mixed vs fp64", err_mixed_vs_fp64)

    # host buffers on both sides (rpd_set_seeds from host memory, moments
    # read back to host memory): the PCIe-inclusive rate of the mixed form
    host_out = np.empty((n, 4), np.float64)
    pcie_ms = []
    for i in range(max(1, min(args.steps, 10))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.set_seeds(seeds)
        ctx.solve(dom, K=K, P=P, V=V, T=T, precision=rpd.PREC_MIXED)
        host_out[:] = ctx.moments()
        pcie_ms.append((time.perf_counter() - t0) * 1e3)
    pcie = {"cells_per_s": n / (float(np.median(pcie_ms)) * 1e-3), "ms_per_step": float(np.median(pcie_ms)),
            "what": "seeds copied from host memory, solve, moments copied to host memory (median of %d)" %
                    len(pcie_ms)}
    log("bench: PCIe-inclusive", pcie)

    # ---- CPU oracle on a bounded sample of the same workload ----------------
    cpu = None
    flop = None
    err_vs_oracle = None
    if not args.no_cpu and args.cpu_sample > 0:
        orc = ge.load_oracle()
        og = orc.Grid(grid.size, grid.in_domain, grid.points, grid.triangles, grid.tri_list, grid.offsets,
                      grid.centre)
        stride = max(1, n // args.cpu_sample)
        ids = np.arange(0, n, stride, dtype=np.int32)[: args.cpu_sample]
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, min(avail, int(os.environ.get("OMP_NUM_THREADS", str(avail)))))

        def timed(sel, **kw):
            t0 = time.perf_counter()
            out = orc.solve_ids(seeds, sel, og, K=K, P=P, V=V, T=T, **kw)
            return out, time.perf_counter() - t0

        # the reference's algorithm (every cell clipped, reference frame), FP64
        (ref, ref_st, _, ost), dt = timed(ids, threads=threads)
        # the same on one core, on every 16th cell of the sample
        ids1 = ids[::16].copy()
        _, dt1 = timed(ids1, threads=1)
        # the same with the library's culls (cells provably outside the domain
        # get volume 0 without being clipped): the algorithmic part of the gain
        _, dtc = timed(ids, threads=threads, cull=1)
        model = cpu_model()
        cpu = {"value": float(ids.size / dt), "unit": "cells/s", "cores": int(threads), "kind": "port",
               "sample": "%d cells (every %d-th seed) of the 1M-seed torus workload, all seeds as neighbours: "
                         "binning + kNN + clip + every re-queue pass of the FP64 (USE_DOUBLE) oracle, the "
                         "reference's algorithm (every cell clipped), on %d threads of %s; "
                         "%.1f s" % (ids.size, stride, threads, model, dt),
               "cpu_model": model, "host_threads_available": int(avail), "host_cpus": int(os.cpu_count() or 0),
               "one_core": {"value": float(ids1.size / dt1), "cells": int(ids1.size), "seconds": dt1},
               "with_culls": {"value": float(ids.size / dtc), "cores": int(threads), "seconds": dtc}}
        log("bench: cpu oracle", cpu)
        _, _, _, ost1 = orc.solve_ids(seeds, ids, og, K=K, P=P, V=V, T=T, threads=threads, single_pass=1, cull=1)
        flop = {"mixed": kernel_flops(ost1, ids.size, True), "fp64": kernel_flops(ost1, ids.size, False)}
        # parity: the FP64 oracle evaluated in each seed's frame, as the kernels
        # clip and integrate (the reference-frame oracle's own rounding error
        # reaches 1e-12 per cell, reported beside it)
        ref_seed, _, _, _ = orc.solve_ids(seeds, ids, og, K=K, P=P, V=V, T=T, threads=threads, frame=1)
        err_vs_oracle = {k: rel_errors(res[k]["out"][ids], ref_seed, n) for k in ("mixed", "fp64")}
        err_vs_oracle["reference_frame_fp64"] = rel_errors(res["fp64"]["out"][ids], ref, n)
        err_vs_oracle["oracle"] = ("FP64 oracle in each seed's frame; reference_frame_fp64: the FP64 form against "
                                   "the oracle in the reference's [0,1000]^3 frame")
        log("bench: vs oracle", err_vs_oracle)

    # ---- rooflines: the dominant kernel (pass-1 finish kernel, mixed form),
    # the bisector kernel and the kNN; launch times from HIP events recorded on
    # the solve's stream ------------------------------------------------------
    sm = res["mixed"]["stats"]
    t_cells = float(np.mean([s["ms_cells_first"] for s in sm])) * 1e-3
    t_finish = float(np.mean([s["ms_finish_first"] for s in sm])) * 1e-3
    roofline = roofline_cells = None
    if flop is not None:
        ka = "clip_cells_kernel<float, double, true, 1, false>"
        kb = "finish_cells_kernel<float, double, true, 1, false, false>"
        roofline_cells = valu_roofline(*flop["mixed"]["cells"], n, t_cells, ka + " (mixed form, pass 1)",
                                       measured_traffic(ka))
        roofline = valu_roofline(*flop["mixed"]["finish"], n, t_finish, kb + " (mixed form, pass 1)",
                                 measured_traffic(kb))
    # kNN launch: each query reads its seed and writes K ids; the sorted seed
    # array (FP32 xyzw) is read at least once
    t_knn = float(np.mean([s["ms_knn_first"] for s in sm])) * 1e-3
    n_query = n - int(sm[-1]["skipped_cells"])
    knn_bytes = (16 + 4 * K) * n_query + 16 * n
    roofline_knn = {"bound": "hbm", "achieved": knn_bytes / t_knn / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": knn_bytes / t_knn / 1e9 / PEAK_HBM_GBS, "queries": n_query,
                    "bytes": "(16 + 4K) per query + 16 per seed", "launch_ms": t_knn * 1e3,
                    "kernel": "knn_kernel<float, 1, 4>", "traffic": measured_traffic("knn_kernel<float, 1, 4>")}

    serve = None
    if not args.no_serving and args.serve_requests > 0:
        serve = serving(rpd, ctx, args.serve_requests, log)

    ms_mixed = res["mixed"]["ms"]
    ms_fp64 = res["fp64"]["ms"]
    line = {
        "metric": METRIC,
        "value": n / (ms_mixed * 1e-3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (SplitMix64 seeds, analytic voxel torus)",
        "config": {"workload": "1M uniform power-weighted seeds (w~U[0,0.25h^2]), torus R0.3 r0.12 @128^3, "
                               "K64 P40 V60 T100, mixed form; FP64 form timed beside it",
                   "n": n, "grid": args.grid, "K": K, "P": P, "V": V, "T": T},
        "fp64_cells_per_s": n / (ms_fp64 * 1e-3),
        "fp64_ms_per_step": ms_fp64,
        "mixed_over_fp64": ms_fp64 / ms_mixed,
        "err_mixed_vs_fp64": err_mixed_vs_fp64,
        "pcie_inclusive": pcie,
        "err_vs_oracle_sample": err_vs_oracle,
        "fallback_fraction": res["mixed"]["stats"][-1]["fallback_cells"] / n,
        "status_hist_mixed": res["mixed"]["stats"][-1]["status_hist"],
        "skipped_cells": int(res["mixed"]["stats"][-1]["skipped_cells"]),
        "stage_ms": {k: {"cover": st["ms_cover"], "knn": st["ms_knn_first"], "cells": st["ms_cells_first"],
                         "finish": st["ms_finish_first"], "pass_cells": st["pass_cells"],
                         "pass_ms": [a + b for a, b in zip(st["pass_ms_knn"], st["pass_ms_clip"])]}
                     for k, st in (("mixed", res["mixed"]["stats"][-1]), ("fp64", res["fp64"]["stats"][-1]))},
        "roofline": roofline,
        "roofline_cells": roofline_cells,
        "roofline_knn": roofline_knn,
        "cpu_baseline": cpu,
        "serving": serve,
    }
    dom.close()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
