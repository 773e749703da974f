"""Benchmark: poses/s of energy + DOF gradient for flexible scoring on one MI355X.

The measured path (`value`) is the fused adjoint: energy and the exact
gradient with respect to the rigid DOFs, all mode amplitudes and the 12x12
pair-parameter table (rc, ac, emin of every type pair, summed over poses),
DOFs resident in HBM.  One step = k_adjoint over all poses + the fixed-order
reduction of the per-workgroup parameter rows.  Reported beside it: the
adjoint without parameter gradients, the fused forward kernel with the
reference's force semantics, the plain per-item path (host buffers, 100 poses
per call) and the drop-in as::Server path.

Workload (default): the flexible-scoring configuration — a 3000-bead receptor
and an 800-bead ligand (seeded synthetic bodies, 12 bead types, 20% charged),
5 + 5 orthonormal modes, receptor and ligand grids at 0.875 A (2.6 A outer),
neighbour list with a 10 A plateau, 100,000 poses.  One step scores all poses
with DOFs already resident in HBM.

The same line carries, under "batched", the 64 uneven pairs scored by one
graph replay (against one launch per pair and against the drop-in Server
path), and under "serving", the docking-job stream's p50/p99 latency per
arrival rate (--no-batched / --no-serve skip them).

Prints one JSON line with the metric, a roofline object for the dominant
kernel (algorithmic bytes per pose of the scoring path x poses / launch time
from HIP events on the launch stream), the CPU baseline — the reference's own
CPU path (compiled from its sources into oracle/_ref) on every host core,
worker threads taking items of 100 poses as the reference's server does, over
a bounded sample; the oracle restatement stands in when the compiled
reference is absent — and "parity": the timed adjoint's results against the
reference's energies and the oracle's gradients.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "gpuattract_2.0_amd")
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "poses/sec (energy+DOF gradient) on 1 MI355X, batched+adjoint vs plain, at 1e-5 rel err"


def _package():
    if "gpuattract_amd" in sys.modules:
        return sys.modules["gpuattract_amd"]
    spec = importlib.util.spec_from_file_location(
        "gpuattract_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpuattract_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    return oracle_py


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (as lscpu prints it) and its logical cores."""
    name = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                name = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return name, os.cpu_count() or 1


def plain_rel(np, x, y):
    """|x - y| / |y| elementwise, 0 where the bits agree (inf and nan included)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.abs(x - y) / np.abs(y)
    r[x == y] = 0.0
    r[np.isnan(x) & np.isnan(y)] = 0.0
    return r


def make(args, W):
    if args.workload == "modes":
        return W.make_workload("flexible scoring 3000/800, 5+5 modes", 20261203, 3000, 800,
                               args.poses, 5, 5, with_basis=True)
    return W.make_workload("rigid scoring 1000/300", 20261202, 1000, 300, args.poses)


def cpu_baseline(wl, args):
    """The reference's CPU path (or, without it, the oracle restatement) on
    every host core over a bounded sample of the same poses."""
    O = _oracle()
    model, threads = cpu_model()
    n = min(args.cpu_poses, wl.dofs.shape[0])
    d = wl.dofs[:n]
    if O.ref_available():
        kind = "reference"
        res, dt = O.ref_score_mt(wl, d, nthreads=threads)
    else:
        kind = "port"
        f = O.score_modes if wl.modes else O.score_rigid
        t0 = time.perf_counter()
        res = f(wl, d, nthreads=threads)
        dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="poses/s", cores=threads, kind=kind, cpu_model=model,
                sample=f"first {n} poses of the same workload (energy + forces, the reference's "
                       f"gradient), {threads} worker threads, items of 100 poses, {dt:.1f} s"), res


def adjoint_parity(np, O, wl, got, ref, k):
    """The timed adjoint's output `got` (n, 7 | 27) against the reference's
    forward energies `ref` (its first rows), against the oracle's FP32 adjoint
    in the kernel's summation order (bit for bit) and against the FP64 adjoint
    (errors scaled by the sum of the absolute terms of each component)."""
    ow = got.shape[1]
    m = ref.shape[0]
    e_rel = plain_rel(np, got[:m, 0], ref[:, 0])
    t0 = time.perf_counter()
    ko = O.adjoint_kernel_order(wl, wl.dofs[:k], nthreads=os.cpu_count() or 1)
    ko = ko[:, :ow].astype(np.float32)
    bits = got[:k].view(np.uint32) == ko.view(np.uint32)
    r64, sc = O.adjoint(wl, wl.dofs[:k], scale=True)
    cols = [c for c in range(1, ow) if sc[:, c].max() > 0]
    g = got[:k, cols].astype(np.float64)
    scaled = np.abs(g - r64[:, cols]) / np.maximum(sc[:, cols], 1e-30)
    e64 = plain_rel(np, got[:k, 0], r64[:, 0])
    out = dict(energy_vs_reference_poses=int(m),
               energy_vs_reference_max_rel_err=float(e_rel.max()),
               energy_vs_reference_bit_exact_share=float(np.mean(got[:m, 0].view(np.uint32) ==
                                                                 ref[:, 0].astype(np.float32).view(np.uint32))),
               kernel_order_oracle_poses=int(k),
               kernel_order_oracle_bit_exact=bool(bits.all()),
               kernel_order_oracle_bit_exact_share=float(bits.mean()),
               energy_vs_fp64_max_rel_err=float(e64.max()),
               gradient_vs_fp64_max_scaled_err=float(scaled.max()),
               gradient_vs_fp64_median_scaled_err=float(np.median(scaled)))
    log(f"parity: energy vs reference ({m} poses) max rel {out['energy_vs_reference_max_rel_err']:.2e}; "
        f"kernel-order oracle ({k} poses) bit-exact {out['kernel_order_oracle_bit_exact']}; gradient "
        f"vs FP64 max scaled {out['gradient_vs_fp64_max_scaled_err']:.2e} "
        f"({time.perf_counter() - t0:.1f} s)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["modes", "rigid", "batched", "minimize", "serve"], default="modes")
    ap.add_argument("--min-poses", type=int, default=50000, help="minimize: starting poses")
    ap.add_argument("--min-evals", type=int, default=500, help="minimize: VA13 evaluation limit")
    ap.add_argument("--min-check", type=int, default=1000,
                    help="minimize: poses compared with the oracle's VA13 on the CPU")
    ap.add_argument("--pairs", type=int, default=64, help="batched workload: number of pairs")
    ap.add_argument("--batched-steps", type=int, default=10, help="graph replays timed (batched)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-jobs", type=int, default=40, help="serve: jobs per arrival rate")
    ap.add_argument("--serve-rates", default="5,20,50,200", help="serve: arrival rates, jobs/s")
    ap.add_argument("--serve-evals", type=int, default=200, help="serve: VA13 evaluation limit")
    ap.add_argument("--poses", type=int, default=100000)
    ap.add_argument("--cpu-poses", type=int, default=50000)
    ap.add_argument("--check-poses", type=int, default=512,
                    help="poses of the timed adjoint compared with the oracle's adjoints")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-server", action="store_true",
                    help="skip timing the drop-in Server path")
    ap.add_argument("--no-pgrad", action="store_true",
                    help="DOF and mode gradients only (no pair-parameter gradient)")
    ap.add_argument("--plain-poses", type=int, default=5000)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="also run the plain copy kernel over this many bytes "
                         "(calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only (the path runs on one GPU's "
                                   "share of the poses); run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    _package()
    from gpuattract_amd import hip, traffic, workload as W
    if args.workload == "batched":
        return run_batched(args, np, hip, traffic, W)
    if args.workload == "minimize":
        return run_minimize(args, np, hip, W)
    if args.workload == "serve":
        return run_serve(args, np, W)

    t0 = time.perf_counter()
    wl = make(args, W)
    log(f"workload built in {time.perf_counter() - t0:.1f} s: {wl.name}, {wl.dofs.shape[0]} poses")
    ctx = hip.Context(0)
    hs = ctx.upload_workload(wl)
    pgrad = not args.no_pgrad and wl.rec_inner_basis is not None
    if pgrad:
        ctx.upload_basis(hs["grid"], wl)
    T = wl.params.T
    n = wl.dofs.shape[0]
    out_w = 27 if wl.modes else 7
    d_in = ctx.to_device(wl.dofs)
    d_out = ctx.dev_alloc(n * out_w * 4)
    sh = ctx.stream()  # the stream every launch below is issued on
    score = ctx.score_modes_dev if wl.modes else ctx.score_rigid_dev
    d_grad = ctx.dev_alloc(n * out_w * 4)
    d_pg = ctx.dev_alloc(T * T * 3 * 4) if pgrad else None

    def fwd_step():
        score(hs, d_in.ptr, n, d_out.ptr, sh)

    def step():
        ctx.adjoint_dev(hs, d_in.ptr, wl.modes, n, None, d_grad.ptr,
                        d_pg.ptr if pgrad else None, sh)

    def dof_step():
        ctx.adjoint_dev(hs, d_in.ptr, wl.modes, n, None, d_grad.ptr, None, sh)

    def rate(fn, reps):
        fn()
        a, b = ctx.event(), ctx.event()
        ctx.record(a, sh)
        for _ in range(reps):
            fn()
        ctx.record(b, sh)
        return ctx.elapsed_ms(a, b) / reps

    # forward kernel and the adjoint without parameter gradients, reported beside
    reps = max(1, args.steps // 2)
    fwd_ms = rate(fwd_step, reps)
    log(f"forward kernel: {fwd_ms:.3f} ms/step, {n / fwd_ms * 1e3:.4g} poses/s")
    dof_ms = rate(dof_step, reps)
    log(f"adjoint, DOF + mode gradients only: {dof_ms:.3f} ms/step, {n / dof_ms * 1e3:.4g} poses/s")
    ctx.set_pose_order(False)
    unordered_ms = rate(step, reps)
    ctx.set_pose_order(True)
    log(f"adjoint in input pose order: {unordered_ms:.3f} ms/step, {n / unordered_ms * 1e3:.4g} poses/s")

    for _ in range(args.warmup):
        step()
    ctx.sync(sh)
    log(f"warm-up done ({args.warmup} steps)")
    starts = [ctx.event() for _ in range(args.steps)]
    ends = [ctx.event() for _ in range(args.steps)]
    ctx.sync(sh)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.record(starts[k], sh)
        step()
        ctx.record(ends[k], sh)
    ctx.sync(sh)
    wall = time.perf_counter() - t0
    launch_ms = [ctx.elapsed_ms(s, e) for s, e in zip(starts, ends)]
    avg_launch = float(np.mean(launch_ms)) / 1e3
    value = n * args.steps / wall
    log(f"timed {args.steps} steps: {wall:.3f} s, {value:.4g} poses/s, launch avg {avg_launch*1e3:.3f} ms")
    timed = ctx.to_host(d_grad, (n, out_w))

    copy_calibration(args, ctx, sh)
    # roofline numerator: the scoring path's algorithmic bytes per pose (bead
    # attributes, grid corners, neighbour-list descriptors and entries, mode
    # vectors, DOFs in and results out); the adjoint's own reads (it skips the
    # receptor-on-ligand-grid lookups) are reported beside it, and the
    # parameter gradient's grid-cell records (16 bytes per ligand bead without
    # pose weights: key and position, written, sorted and reduced by the cell
    # pass)
    bpp, bpp_sd = traffic.bytes_per_pose(wl)
    adj_bpp, _ = traffic.bytes_per_pose(wl, adjoint=True)
    rec_bpp = 16.0 * wl.lig.n if pgrad else 0.0
    achieved = bpp * n / avg_launch
    prof = profiled_traffic(args.workload) if pgrad else None
    roofline = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                    frac=achieved / HBM_PEAK, traffic=prof["bytes"] if prof else None,
                    traffic_source=prof["source"] if prof else None,
                    kernel=(f"k_adjoint<{5 if wl.modes else 0}>" +
                            (" + the parameter gradient's cell pass and row reduction" if pgrad else "")),
                    bytes_per_pose=bpp, bytes_per_pose_sd=bpp_sd, poses_per_launch=n,
                    avg_launch_ms=avg_launch * 1e3,
                    adjoint_read_bytes_per_pose=adj_bpp,
                    param_grad_record_bytes_per_pose=rec_bpp,
                    frac_adjoint_reads=adj_bpp * n / avg_launch / HBM_PEAK)

    # the same adjoint from host buffers in one call (DOFs over PCIe in, results
    # and parameter gradients out): the PCIe-inclusive rate, never `value`
    ctx.adjoint(hs, wl.dofs[:1000], pgrad=pgrad, T=T)
    t0 = time.perf_counter()
    ctx.adjoint(hs, wl.dofs, pgrad=pgrad, T=T)
    host_adj = n / (time.perf_counter() - t0)
    log(f"adjoint from host buffers, one call: {host_adj:.4g} poses/s")

    # The plain per-item path: host buffers, 100 poses per call (the reference's
    # item size, server/Server.h:39), as the drop-in service issues them.
    m = min(args.plain_poses, n)
    host_score = ctx.score_modes if wl.modes else ctx.score_rigid
    host_score(hs, wl.dofs[:100])
    t0 = time.perf_counter()
    for a in range(0, m, 100):
        host_score(hs, wl.dofs[a:a + 100])
    plain = m / (time.perf_counter() - t0)
    log(f"plain item-100 host path: {plain:.4g} poses/s")
    # the drop-in path itself (as::Server, items of 100), best of 1 and 2 workers
    ms = min(n, 20000)
    server_by_workers = {}
    for workers in (() if args.no_server else (1, 2)):
        sec = server_pass_seconds(np, hip, wl, wl.dofs[:ms], workers)
        if sec:
            server_by_workers[workers] = ms / sec
    server = max(server_by_workers.values()) if server_by_workers else None
    if server:
        log("drop-in Server path, items of 100, 8 requests in flight: " + ", ".join(
            f"{w} worker(s) {r:.4g} poses/s" for w, r in server_by_workers.items()))

    cpu = parity = None
    if not args.no_cpu:
        cpu, ref = cpu_baseline(wl, args)
        fwd = ctx.to_host(d_out, (n, out_w))[:ref.shape[0]]
        cpu["forward_vs_reference_energy_max_rel_err"] = float(plain_rel(np, fwd[:, 0], ref[:, 0]).max())
        log(f"cpu baseline: {cpu['value']:.4g} poses/s on {cpu['cores']} threads ({cpu['cpu_model']})")
        parity = adjoint_parity(np, _oracle(), wl, timed, ref, min(args.check_poses, n))
    ctx.close()

    batched = None
    if not args.no_batched:
        batched = batched_leg(args, np, hip, traffic, W)
    serving = None
    if not args.no_serve:
        rows, st = serve_rows(args, np, W)
        serving = dict(stream=f"{st['jobs']} docking jobs, {st['receptors']} receptors, {st['poses']} "
                              f"poses, VA13 <= {args.serve_evals} evaluations",
                       rates=[{k: r[k] for k in ("rate_jobs_per_s", "p50_ms", "p99_ms", "poses_per_s",
                                                 "pose_evaluations_per_s", "jobs_per_s", "max_running")}
                              for r in rows])

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator: bodies, pair table, grids, poses, modes)",
        "config": {
            "workload": wl.name,
            "n_rec": wl.rec.n, "n_lig": wl.lig.n, "modes_rec": wl.rec.M, "modes_lig": wl.lig.M,
            "poses": n,
            "path": ("fused adjoint (energy + exact DOF, mode" +
                     (" and 12x12 pair-parameter" if pgrad else "") +
                     " gradients), device-resident"),
            "adjoint_dof_modes_only_poses_per_s": n / dof_ms * 1e3,
            "adjoint_input_order_poses_per_s": n / unordered_ms * 1e3,
            "forward_kernel_poses_per_s": n / fwd_ms * 1e3,
            "adjoint_host_buffers_poses_per_s": host_adj,
            "plain_item100_poses_per_s": plain,
            "adjoint_vs_plain": value / plain,
            "server_item100_poses_per_s": server,
            "server_item100_by_workers": server_by_workers,
            "adjoint_vs_server": value / server if server else None,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


def run_minimize(args, np, hip, W):
    """Docking minimization: random starting poses of a 4000-bead receptor and
    a 1000-bead ligand with 10 + 10 modes, each minimized by VA13 (at most
    --min-evals line-search evaluations) driven by the adjoint, minimizer
    states resident on the device.  A step is one complete minimization run."""
    t0 = time.perf_counter()
    wl = W.make_workload("minimization 4000/1000, 10+10 modes", 20261205, 4000, 1000,
                         args.min_poses, 10, 10, with_lig_grid=False)
    n = wl.dofs.shape[0]
    log(f"workload built in {time.perf_counter() - t0:.1f} s: {wl.name}, {n} starting poses")
    ctx = hip.Context(0)
    hs = ctx.upload_workload(wl)
    sh = ctx.stream()
    d_in = ctx.to_device(wl.dofs)
    d_out = ctx.dev_alloc(n * 26 * 4)
    d_e, d_eb, d_nf = ctx.dev_alloc(n * 4), ctx.dev_alloc(n * 4), ctx.dev_alloc(n * 4)
    # warm-up on a few poses
    ctx.minimize_dev(hs, d_in.ptr, True, min(n, 256), args.min_evals, d_out.ptr, d_e.ptr, d_eb.ptr,
                     d_nf.ptr, sh)
    ctx.sync(sh)
    e0 = ctx.adjoint(hs, wl.dofs[:min(n, 4096)])[:, 0]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.minimize_dev(hs, d_in.ptr, True, n, args.min_evals, d_out.ptr, d_e.ptr, d_eb.ptr,
                         d_nf.ptr, sh)
    ctx.sync(sh)
    wall = (time.perf_counter() - t0) / args.steps
    nf = ctx.to_host(d_nf, (n,), np.int32)
    eb = ctx.to_host(d_eb, (n,))
    el = ctx.to_host(d_e, (n,))
    evals = float(nf.sum() + n)  # line-search evaluations plus each start point
    log(f"{n} minimizations in {wall:.2f} s: {evals / wall:.4g} pose evaluations/s, mean "
        f"{nf.mean():.1f} line-search evaluations per pose, {np.mean(nf >= args.min_evals):.2%} at the limit")
    check = None
    if args.min_check > 0 and not args.no_cpu:
        # the oracle's VA13 on its FP32 adjoint in the kernel's summation
        # order: the device minimizer's trajectories, pose for pose
        O = _oracle()
        m = min(n, args.min_check)
        threads = os.cpu_count() or 1
        t1 = time.perf_counter()
        parts = []
        for a in range(0, m, 128):  # a progress line per part
            parts.append(O.minimize_kernel_order(wl, wl.dofs[a:min(m, a + 128)],
                                                 ivmax=args.min_evals, nthreads=threads))
            log(f"  oracle VA13: {min(m, a + 128)} of {m} poses ({time.perf_counter() - t1:.0f} s)")
        od, oel, oeb, onf = (np.concatenate([p[i] for p in parts]) for i in range(4))
        dt = time.perf_counter() - t1
        dd = ctx.to_host(d_out, (n, 26))[:m]
        same_dofs = np.all(dd.view(np.uint32) == od.astype(np.float32).view(np.uint32), axis=1)
        same = same_dofs & (nf[:m] == onf) & (eb[:m].astype(np.float64) == oeb) & \
            (el[:m].astype(np.float64) == oel)
        rel = np.abs(eb[:m].astype(np.float64) - oeb) / np.maximum(np.abs(oeb), 1e-30)
        check = dict(poses=m, identical_share=float(np.mean(same)),
                     best_energy_max_rel_err=float(rel.max()),
                     oracle_mean_evaluations=float(onf.mean()),
                     oracle_pose_evaluations_per_s=float((onf.sum() + m) / dt),
                     oracle_threads=threads, oracle_seconds=dt)
        log(f"oracle VA13 (kernel order) on {m} poses ({threads} threads, {dt:.1f} s): "
            f"{check['identical_share']:.4f} identical (DOFs, energies, evaluation counts), "
            f"max relative best-energy difference {check['best_energy_max_rel_err']:.1e}")
    line = {
        "metric": "poses/sec (energy+DOF gradient) on 1 MI355X, batched+adjoint vs plain, at 1e-5 rel err",
        "value": evals / wall, "unit": "pose evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 scoring, fp64 minimizer",
        "data": "synthetic (seeded generator: bodies, pair table, grids, poses, modes)",
        "config": {"workload": wl.name, "starting_poses": n, "max_fun_eval": args.min_evals,
                   "minimizations_per_s": n / wall, "mean_line_search_evaluations": float(nf.mean()),
                   "share_at_limit": float(np.mean(nf >= args.min_evals)),
                   "mean_energy_start": float(np.mean(e0)), "mean_energy_best": float(np.mean(eb)),
                   "mean_energy_last": float(np.mean(el)), "oracle_check": check},
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def serve_rows(args, np, W):
    """Serving stream: docking jobs (receptor 500-8000 beads, ligand 100-2000,
    0-10 modes per body, 500-20000 poses, VA13 capped at --serve-evals
    evaluations; 30% of jobs reuse an earlier job's receptor) arrive as a
    Poisson process; the server builds a receptor's grids on the device at its
    first job and keeps them cached, admits jobs by HBM footprint and runs the
    admitted ones concurrently.  Each arrival rate runs the same job list on a
    fresh server (cold cache).  Reported: p50/p99 job latency (submission to
    results on the host) and sustained throughput (minimized poses and pose
    evaluations per second from the first submission to the last result)."""
    import threading
    from gpuattract_amd import serve
    t0 = time.perf_counter()
    params, recs, jobs = W.make_serve_stream(20261206, args.serve_jobs,
                                             max_fun_eval=args.serve_evals)
    total_poses = sum(j.dofs.shape[0] for j in jobs)
    log(f"stream built in {time.perf_counter() - t0:.1f} s: {len(jobs)} jobs, {len(recs)} receptors, "
        f"{total_poses} poses")
    rates = [float(r) for r in args.serve_rates.split(",")]
    per_rate = []
    for rate in rates:
        rng = np.random.default_rng(int(rate * 1000) + 7)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
        with serve.Server(0) as srv:
            pid = srv.add_params(params)
            rid = [srv.add_receptor_body(r.body, pid, r.inner, r.outer, 10.0, 10.8) for r in recs]
            lid = [srv.add_ligand(j.lig) for j in jobs]
            ids = [None] * len(jobs)
            start = time.perf_counter()

            def submitter():
                for k, j in enumerate(jobs):
                    delay = start + arrivals[k] - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                    ids[k] = srv.submit(dict(rec=rid[j.receptor], lig=lid[k], params=pid), j.dofs,
                                        j.max_fun_eval)

            th = threading.Thread(target=submitter)
            th.start()
            lat, evals, cached = [], 0, 0
            for k in range(len(jobs)):
                while ids[k] is None:
                    time.sleep(0.0005)
                r = srv.wait(ids[k])
                lat.append(r[4]["latency_ms"])
                evals += int(r[3].sum()) + r[3].shape[0]
                cached += r[4]["grid_cached"]
            wall = time.perf_counter() - start
            th.join()
            st = srv.stats()
        lat = np.array(lat)
        row = dict(rate_jobs_per_s=rate, jobs=len(jobs), wall_s=wall,
                   poses_per_s=total_poses / wall, pose_evaluations_per_s=evals / wall,
                   jobs_per_s=len(jobs) / wall, p50_ms=float(np.percentile(lat, 50)),
                   p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
                   max_running=st["max_running"], grid_hits=st["grid_hits"],
                   grid_misses=st["grid_misses"], evictions=st["evictions"],
                   peak_gb=st["peak_bytes"] / 1e9, busy_s=st["busy_s"])
        per_rate.append(row)
        log(f"rate {rate:g} jobs/s: {wall:.2f} s, {row['poses_per_s']:.4g} poses/s, "
            f"{row['pose_evaluations_per_s']:.4g} evaluations/s, p50 {row['p50_ms']:.0f} ms, "
            f"p99 {row['p99_ms']:.0f} ms, up to {row['max_running']} jobs at once, "
            f"{row['grid_hits']} grid hits / {row['grid_misses']} builds")
    return per_rate, dict(jobs=len(jobs), receptors=len(recs), poses=total_poses)


def run_serve(args, np, W):
    """The serving stream on its own (serve_rows), the best rate as `value`."""
    per_rate, st = serve_rows(args, np, W)
    best = max(per_rate, key=lambda r: r["poses_per_s"])
    line = {
        "metric": "docking jobs served: sustained minimized poses/s with p50/p99 job latency",
        "value": best["poses_per_s"], "unit": "minimized poses/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": best["wall_s"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 scoring, fp64 minimizer",
        "data": "synthetic (seeded stream of bodies, poses, modes; grids built on the device)",
        "config": {"workload": f"serving stream, {st['jobs']} jobs, {st['receptors']} receptors, "
                               f"{st['poses']} poses, VA13 <= {args.serve_evals} evaluations",
                   "rates": per_rate},
    }
    print(json.dumps(line), flush=True)


_DRIVER = None


def _driver():
    global _DRIVER
    if _DRIVER is None:
        import ctypes as C
        path = os.path.join(ROOT, "tests", "cpp", "_bin", "libservice_driver.so")
        if not os.path.exists(path):
            return None
        _DRIVER = C.CDLL(path)
        _DRIVER.drv_error.restype = C.c_char_p
    return _DRIVER


def server_pass_seconds(np, hip, wl, dofs, workers, requests=8, item=100, reps=3):
    """Wall time of one pass of the drop-in path over `dofs`: as::Server with
    `workers` workers and the MI355 scoring service of the pair (items of
    `item` poses; 100 is the reference's default, server/Server.h:39), the
    poses split into `requests` requests submitted together; mean of `reps`
    timed passes after one untimed pass.  Driven through the C++ driver
    (tests/cpp/service_driver.cpp); None when it is not built."""
    import ctypes as C
    L = _driver()
    if L is None:
        return None
    p = wl.params
    tab = hip.param_table(p)
    rec, lig = hip.protein_desc(wl.rec), hip.protein_desc(wl.lig)
    gi, go, nl = hip.grid_desc(wl.rec_inner), hip.grid_desc(wl.rec_outer), hip.nl_desc(wl.rec_nl)
    d = np.ascontiguousarray(dofs, np.float32)
    m = d.shape[0]
    out = np.zeros((m, 27 if wl.modes else 7), np.float32)
    tail = (hip._p(tab), p.T, p.shape, p.dielec, C.c_float(p.epsilon), C.c_float(p.ffelec),
            hip._p(d), C.c_size_t(m), hip._p(out), workers, item, requests, reps)
    sec = C.c_double(0.0)
    if wl.modes:
        li, lo = hip.grid_desc(wl.lig_inner), hip.grid_desc(wl.lig_outer)
        st = L.drv_time_modes(C.byref(rec), C.byref(lig), C.byref(gi), C.byref(go), C.byref(nl),
                              C.byref(li), C.byref(lo), *tail, C.byref(sec))
    else:
        st = L.drv_time_rigid(C.byref(rec), C.byref(lig), C.byref(gi), C.byref(go), C.byref(nl),
                              *tail, C.byref(sec))
    if st != 0:
        raise RuntimeError("drop-in Server path: " + L.drv_error().decode())
    return sec.value


def copy_calibration(args, ctx, sh):
    """The plain copy kernel over --copy-calibration bytes (read once, written
    once), so that a counter run can scale FETCH_SIZE / WRITE_SIZE."""
    if not args.copy_calibration:
        return
    nb = args.copy_calibration // 16 * 16
    a, bb = ctx.dev_alloc(nb), ctx.dev_alloc(nb)
    ctx.copy(bb, a, nb, sh)
    ctx.sync(sh)
    log(f"copy calibration: {nb} bytes read + {nb} bytes written")
    a.release()
    bb.release()


def profiled_traffic(workload):
    """HBM bytes per launch (per graph replay for the batched workload) from
    the calibrated FETCH_SIZE / WRITE_SIZE counter runs kept in profiles/:
    {bytes, source} or None."""
    name = f"traffic_{workload}.json"
    try:
        rec = json.load(open(os.path.join(ROOT, "profiles", name)))
    except (OSError, ValueError):
        return None
    return dict(bytes=rec.get("hbm_bytes_per_launch"),
                source=f"profiles/{name}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of "
                       f"{rec.get('kernel', 'the kernel')} on {rec.get('workload', 'this workload')}, "
                       f"scaled by the counts of a copy kernel of known size "
                       f"(factors {rec.get('fetch_factor')} / {rec.get('write_factor')})")


def batched_leg(args, np, hip, traffic, W, check=0):
    """The 64 uneven pairs (receptor 500-5000 beads, ligand 100-1500, 0-10
    modes per body, 1000-50000 poses, all log-uniform except modes) scored by
    the batched adjoint: buckets tuned by the autotuner, launches captured in
    a hipGraph; one step replays the graph over every pose of every pair.
    Receptor grids are built on the device.  Beside it: one device launch per
    pair, the plain item-100 path and the drop-in Server path pair after pair
    (each pair's grids fetched to the host for it, one pair at a time).  With
    check > 0, the first `check` poses of every pair are compared with the
    oracle's FP32 adjoint in the kernel's summation order (bit for bit)."""
    t0 = time.perf_counter()
    pairs = W.make_batch(20261204, args.pairs, host_grids=False)
    total = sum(p.dofs.shape[0] for p in pairs)
    log(f"batched: {len(pairs)} pairs made in {time.perf_counter() - t0:.1f} s, {total} poses")
    ctx = hip.Context(0)
    t0 = time.perf_counter()
    staged = []
    for wl in pairs:
        hs = ctx.upload_workload(wl)
        n, ow = wl.dofs.shape[0], (27 if wl.modes else 7)
        staged.append(dict(wl=wl, hs=hs, d_dofs=ctx.to_device(wl.dofs), n=n, ow=ow,
                           d_out=ctx.dev_alloc(n * ow * 4)))
    log(f"batched: grids built on the device and poses uploaded in {time.perf_counter() - t0:.1f} s")
    b = ctx.batch([dict(hs=s["hs"], d_dofs=s["d_dofs"].ptr, d_out=s["d_out"].ptr, n=s["n"],
                        modes=s["wl"].modes) for s in staged])
    sh = ctx.stream()
    t0 = time.perf_counter()
    b.tune(sh)
    buckets = b.buckets()
    log(f"batched: tuned in {time.perf_counter() - t0:.1f} s: " + ", ".join(
        f"M{k['modeCapacity']}/size{k['sizeClass']}: {k['numPairs']} pairs, {k['wavesPerPose']} "
        f"waves/pose, {k['posesPerTile']} poses/tile, {k['ms']:.2f} ms" for k in buckets))
    for _ in range(max(1, args.warmup)):
        b.run(sh)
    ctx.sync(sh)
    steps = args.batched_steps
    starts = [ctx.event() for _ in range(steps)]
    ends = [ctx.event() for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.record(starts[k], sh)
        b.run(sh)
        ctx.record(ends[k], sh)
    ctx.sync(sh)
    wall = time.perf_counter() - t0
    launch = float(np.mean([ctx.elapsed_ms(a, c) for a, c in zip(starts, ends)])) / 1e3
    value = total * steps / wall
    log(f"batched: timed {steps} graph replays: {value:.4g} poses/s, {launch * 1e3:.3f} ms per replay")
    copy_calibration(args, ctx, sh)
    outs = [ctx.to_host(s["d_out"], (s["n"], s["ow"]))[:max(check, 1)] for s in staged]

    # the same work pair by pair: one device adjoint launch per pair
    e0, e1 = ctx.event(), ctx.event()
    ctx.record(e0, sh)
    for s in staged:
        ctx.adjoint_dev(s["hs"], s["d_dofs"].ptr, s["wl"].modes, s["n"], None, s["d_out"].ptr, None, sh)
    ctx.record(e1, sh)
    per_pair = total / (ctx.elapsed_ms(e0, e1) / 1e3)
    # the plain path: host buffers, 100 poses per call, every pair
    m = max(100, args.plain_poses // len(staged) // 100 * 100)
    t0 = time.perf_counter()
    done = 0
    for s in staged:
        d = s["wl"].dofs[:m]
        for a in range(0, d.shape[0], 100):
            ctx.adjoint(s["hs"], d[a:a + 100])
        done += d.shape[0]
    plain = done / (time.perf_counter() - t0)
    log(f"batched: per-pair device launches {per_pair:.4g} poses/s; plain item-100 path {plain:.4g} poses/s")

    # pair after pair, with the pair's grids on the host: algorithmic bytes,
    # the drop-in Server path and the oracle check
    O = _oracle() if check else None
    algo = algo_fwd = secs = 0.0
    worst, exact, checked = 0.0, True, 0
    t0 = time.perf_counter()
    for k, s in enumerate(staged):
        wl = s["wl"]
        W.fill_grids_from_device(ctx, wl)
        T = wl.params.T
        gli, glo = W.pair_lig_grid_geoms(wl)
        if wl.modes:
            # the forward path's ligand grids (the adjoint does not read them)
            wl.lig_inner = W.IGrid(gli, T + 1, ctx.generate_grid(wl.lig, wl.params, gli, src_is_ligand=True))
            wl.lig_outer = W.IGrid(glo, T + 1, ctx.generate_grid(wl.lig, wl.params, glo, src_is_ligand=True))
        algo += traffic.bytes_per_pose(wl, sample=64, adjoint=True)[0] * s["n"]
        if wl.modes:
            algo_fwd += traffic.bytes_per_pose(wl, sample=64)[0] * s["n"]
        else:
            algo_fwd += traffic.bytes_per_pose(wl, sample=64, adjoint=True)[0] * s["n"]
        if not args.no_server and secs is not None:
            sec = server_pass_seconds(np, hip, wl, wl.dofs, 2, reps=1)
            secs = None if sec is None else secs + sec
        if check:
            waves = buckets[b.pair_bucket(k)]["wavesPerPose"]
            c = min(check, s["n"])
            ko = O.adjoint_kernel_order(wl, wl.dofs[:c], waves=waves,
                                        nthreads=os.cpu_count() or 1)[:, :s["ow"]].astype(np.float32)
            g = outs[k][:c]
            exact &= bool(np.array_equal(g.view(np.uint32), ko.view(np.uint32)))
            worst = max(worst, float(plain_rel(np, g[:, 0], ko[:, 0]).max()))
            checked += c
        # host memory: one pair's grids at a time
        wl.rec_inner = W.IGrid(wl.rec_inner.geom, T + 1, None)
        wl.rec_outer = W.IGrid(wl.rec_outer.geom, T + 1, None)
        wl.rec_nl = W.NLGrid(wl.rec_nl.geom, wl.rec_nl.plateau, None, None)
        wl.lig_inner = wl.lig_outer = None
        if k % 16 == 15:
            log(f"batched: {k + 1} pairs through the host-side passes ({time.perf_counter() - t0:.0f} s)")
    server = total / secs if secs else None
    if server:
        log(f"batched: drop-in Server path (2 workers, items of 100), pair after pair: {server:.4g} poses/s")
    # the scoring path's bytes per pose as for the single pair (the forward
    # formula: receptor beads on the ligand grids too); the adjoint's own
    # reads (it builds and reads no ligand grids) reported beside them
    roofline = dict(bound="hbm", achieved=algo_fwd / launch / 1e9, peak=HBM_PEAK / 1e9,
                    unit="GB/s", frac=algo_fwd / launch / HBM_PEAK, kernel="k_adjoint_batch",
                    bytes_per_pose=algo_fwd / total, adjoint_read_bytes_per_pose=algo / total,
                    frac_adjoint_reads=algo / launch / HBM_PEAK)
    prof = profiled_traffic("batched")
    if prof:
        roofline.update(traffic=prof["bytes"], traffic_source=prof["source"])
    out = dict(workload=f"batched {len(staged)} uneven pairs", pairs=len(staged), poses=total,
               poses_per_s=value, ms_per_replay=launch * 1e3, replays=steps,
               path="batched adjoint (energy + exact DOF and mode gradients), autotuned buckets, "
                    "hipGraph replay",
               buckets=buckets, per_pair_launches_poses_per_s=per_pair,
               plain_item100_poses_per_s=plain, batched_vs_plain=value / plain,
               server_item100_poses_per_s=server,
               batched_vs_server=value / server if server else None,
               batched_vs_per_pair=value / per_pair, roofline=roofline)
    if check:
        out["parity"] = dict(poses=checked, kernel_order_oracle_bit_exact=exact,
                             energy_max_rel_err=worst)
        log(f"batched: {checked} poses against the kernel-order oracle: bit-exact {exact}")
    b.close()
    ctx.close()
    return out


def run_batched(args, np, hip, traffic, W):
    """The batched leg on its own line, the graph replay rate as `value`, with
    the oracle's FP32 adjoint timed on the host as the CPU baseline."""
    args.batched_steps = args.steps
    res = batched_leg(args, np, hip, traffic, W, check=0 if args.no_cpu else 160)
    cpu = None
    if not args.no_cpu:
        O = _oracle()
        from concurrent.futures import ThreadPoolExecutor
        model, threads = cpu_model()
        pairs = W.make_batch(20261204, args.pairs, host_grids=False)
        ctx = hip.Context(0)
        k = 160
        dt = 0.0
        for wl in pairs:  # grids from the device, one pair at a time
            W.fill_grids_from_device(ctx, wl)
            d = wl.dofs[:k]
            t0 = time.perf_counter()
            with ThreadPoolExecutor(threads) as ex:
                list(ex.map(lambda c: O.adjoint(wl, d[c:c + 10], f32=True), range(0, d.shape[0], 10)))
            dt += time.perf_counter() - t0
            wl.rec_inner = wl.rec_outer = wl.rec_nl = None
        ctx.close()
        cpu = dict(value=k * len(pairs) / dt, unit="poses/s", cores=threads, kind="port",
                   cpu_model=model,
                   sample=f"first {k} poses of every pair, the oracle's FP32 adjoint (energy + exact "
                          f"gradient), items of 10 poses on {threads} threads, {dt:.1f} s")
    line = {
        "metric": METRIC, "value": res["poses_per_s"], "unit": "poses/s", "n_gpus": 1,
        "steps": res["replays"], "warmup": args.warmup, "ms_per_step": res["ms_per_replay"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded generator: bodies, pair tables, grids, poses, modes)",
        "config": {k: v for k, v in res.items() if k not in ("roofline", "parity")},
        "roofline": res["roofline"], "cpu_baseline": cpu, "parity": res.get("parity"),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
