"""Full PIC cycle throughput on one MI355X.

    python bench.py [--steps K] [--warmup W] [--config prod|small] [--cpu-seconds S]

Workload (default, "prod"): 1024x512 cells, ds 0.5, dt 0.05, 200 particles
per cell per species (2 x 104,857,600), mi/me 1836, Te/Ti 10, probe bias
-10 kTe/e, Jacobi with FP64 residual to 1e-6 (at most 500 sweeps), FP32
particles and fields, int64 fixed-point charge.  Synthetic seeded initial
load; the first steps after init are warm-up.  One step is one pass of the
cycle deposit -> solve -> field -> move -> cc over both species.

Prints one JSON line: particle-steps/s over the timed steps (inputs resident
in HBM), the roofline of the dominant kernel (HIP events on the run's
stream), the full-cycle algorithmic-byte roofline, and a CPU baseline (the
oracle's OpenMP FP64 build on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
FP64_MFMA_PEAK = 78.6  # MI355X dense FP64 matrix, TFLOP/s
BREAKDOWN_STEPS = 3    # steps after the timed ones for the per-class kernel times

CONFIGS = {
    "prod": dict(ncx=1024, ncy=512, ppc=200, mi=1836.0, kti=0.1, phi_p=-10.0),
    "small": dict(ncx=128, ncy=64, ppc=20, mi=100.0, kti=1.0, phi_p=-5.0),
}


def workload_name(cfg, p):
    return (f"{cfg['ncx']}x{cfg['ncy']} cells, {cfg['ppc']} ppc/species "
            f"(2x{p.per_row * p.ncy} particles), mi/me={cfg['mi']:g}, "
            f"Te/Ti={1.0 / cfg['kti']:g}, phi_p={cfg['phi_p']:g} kTe/e, "
            f"Jacobi FP64 residual {p.jacobi_tol:g} (max {p.jacobi_max_sweeps} sweeps)")


def cpu_baseline(cfg, min_seconds, max_steps):
    """The oracle's OpenMP FP64 cycle on the same workload (the CPU checker,
    never the measured path): full cycles after its own init, repeated until
    min_seconds of stepping have passed (at least 2, at most max_steps)."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    p = O.make_params(**cfg)
    sim = O.Sim(p, omp=True)
    sim.set_solver(1, 1e-6, 1e-6, 0, 500, 50)
    sim.init()
    t = p.dt
    n = 0
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and (steps < 2 or time.perf_counter() - t0 < min_seconds):
        n += sim.count(0) + sim.count(1)
        sim.step(t)
        t += p.dt
        steps += 1
    el = time.perf_counter() - t0
    del sim
    return {"value": n / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} full cycles of the same workload after init "
                      f"(FP64, OpenMP, {threads} threads, Jacobi capped at 500 sweeps), "
                      f"{el:.1f} s"}


def profiled_traffic(kernel):
    """HBM bytes per particle of a kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "scatter_push_1024x512_200ppc_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic_bytes_per_particle" and w[1] == kernel:
                    return float(w[2]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None

def direct_leg(cp, cfg, steps, warmup, tile):
    """The same cycle with the Poisson system solved directly in FP64 (the
    converged fixed point of the Jacobi iteration) instead of 500 sweeps."""
    p = cp.params(**cfg, jacobi_mode=cp.POISSON_DIRECT, tile_cells=tile)
    run = cp.Run(p)
    run.init_sim()
    t = p.dt
    for _ in range(warmup):
        run.step(t)
        t += p.dt
    run.synchronize()
    n0 = run.number_of_particles(0) + run.number_of_particles(1)
    run.timing(True, classes=["solve"])
    run.timing_read()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step(t)
        t += p.dt
    run.synchronize()
    wall = time.perf_counter() - t0
    kern = run.timing_read()
    n1 = run.number_of_particles(0) + run.number_of_particles(1)
    _, resid = run.last_solve()
    run.close()
    solve_ms = kern["solve"][0] / max(kern["solve"][1], 1)
    # the solve's four FP64 products: 2 x (m x ncx x ncx) for the DFT along x
    # and 2 x (m x m x ncx) for the sine transform along y, m = ncy - 1
    m, nx = p.ncy - 1, p.ncx
    flops = 2.0 * (2.0 * m * nx * nx + 2.0 * m * m * nx)
    achieved = flops / (solve_ms * 1e-3) / 1e12 if solve_ms > 0 else 0.0
    return {"value": 0.5 * (n0 + n1) * steps / wall, "unit": "particle-steps/s",
            "ms_per_step": 1e3 * wall / steps, "poisson_fp64_residual": resid,
            "solve_ms_per_launch": solve_ms,
            "solve_roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK,
                               "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK,
                               "flops_per_solve": flops,
                               "note": "whole solve (four products, slice sums, FP64 "
                                       "residual) against the dense FP64 matrix peak"}}


def draw_batched(n_runs, seed):
    """256 independent runs: ncx ~ U{64..512}, ncy ~ U{32..256}, ppc ~ U{20..100},
    bias ~ U[-10, 2] kTe/e; other parameters as the small configuration."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_runs):
        out.append(dict(ncx=int(rng.integers(64, 513)), ncy=int(rng.integers(32, 257)),
                        ppc=int(rng.integers(20, 101)), phi_p=float(rng.uniform(-10.0, 2.0)),
                        mi=100.0, kti=1.0, run_id=k))
    return out


def batched_leg(cp, n_runs, steps, seq_steps, seed, fixed_tile=0):
    """One HIP-graph step advances all runs; baseline: the same runs stepped
    one after another through the plain path (cupic_step per run), skipped
    when seq_steps is 0.  The tile (cells per wave strip) is tuned in place
    on the batch itself: each candidate is set on every run and timed over 2
    batched steps, unless fixed_tile names one."""
    draws = draw_batched(n_runs, seed)
    runs = []
    for d in draws:
        p = cp.params(**d, jacobi_mode=cp.JACOBI_RESIDUAL, jacobi_tol=1e-6,
                      jacobi_tol_init=1e-6, jacobi_max_sweeps=500, jacobi_check_every=50)
        r = cp.Run(p)
        r.init_sim()
        runs.append(r)
    for r in runs:
        r.synchronize()
    npart = sum(r.number_of_particles(0) + r.number_of_particles(1) for r in runs)
    print(f"# batched: {n_runs} runs, {npart} particles", file=sys.stderr, flush=True)
    tune = {}
    done = 0
    for tile in ((fixed_tile,) if fixed_tile else (0, 2, 4, 6, 8, 12, 16)):
        for r in runs:
            r.set_tile(tile)
        b = cp.Batch(runs, graph=True)
        b.step(1)
        t0 = time.perf_counter()
        b.step(2)
        tune[tile] = 1e3 * (time.perf_counter() - t0) / 2
        b.close()
        done += 3
    tile = min(tune, key=tune.get)
    print(f"# batched: tile tuning (ms per batched step) {tune} -> {tile}", file=sys.stderr,
          flush=True)
    for r in runs:
        r.set_tile(tile)
    b = cp.Batch(runs, graph=True)
    b.step(1)
    done += 1
    t0 = time.perf_counter()
    b.step(steps)
    wall_b = time.perf_counter() - t0
    b.close()
    done += steps
    # the runs continue on the plain path, one after another, from the time
    # the batch reached (t accumulated as t += dt, like the device clock)
    t_next = []
    for r in (runs if seq_steps > 0 else ()):
        t = 0.0
        for _ in range(1 + done):
            t += r.p.dt
        r.step(t)                     # warm-up of the plain path
        r.synchronize()
        t_next.append(t + r.p.dt)
    t0 = time.perf_counter()
    for r, t in zip(runs, t_next):
        for _ in range(seq_steps):
            r.step(t)
            t += r.p.dt
        r.synchronize()
    wall_s = time.perf_counter() - t0
    for r in runs:
        r.close()
    rs_b = n_runs * steps / wall_b
    rs_s = n_runs * seq_steps / wall_s if seq_steps > 0 else None
    return {"runs": n_runs, "particles": int(npart), "steps": steps,
            "tile_cells": tile, "tile_tuning_ms_per_step": tune,
            "ms_per_batched_step": 1e3 * wall_b / steps,
            "run_steps_per_s": rs_b, "particle_steps_per_s": npart * steps / wall_b,
            "sequential_plain_run_steps_per_s": rs_s,
            "sequential_plain_steps_per_run": seq_steps,
            "speedup_vs_sequential_plain": rs_b / rs_s if rs_s else None,
            "workload": "256 runs, ncx~U{64..512}, ncy~U{32..256}, ppc~U{20..100}, "
                        "phi_p~U[-10,2], mi/me 100, Te/Ti 1, Jacobi FP64 residual 1e-6 "
                        "(max 500 sweeps), one HIP graph per batched step"}


def draw_requests(n, rate, seed):
    """I-V points: Poisson arrivals at `rate` per second, grid ncx ~ U{128..1024},
    ncy ~ U{64..512}, ppc ~ U{20..200}, bias ~ U[-10, 2] kTe/e."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        out.append((t, dict(ncx=int(rng.integers(128, 1025)), ncy=int(rng.integers(64, 513)),
                            ppc=int(rng.integers(20, 201)),
                            phi_p=float(rng.uniform(-10.0, 2.0)), run_id=1000 + k)))
    return out


def converged_rule(prev_net, ie, ii, rel):
    gross = ii + ie
    return gross > 0.0 and abs((ii - ie) - prev_net) / gross < rel


def serving_leg(cp, n_req, rate, max_steps, seed):
    """Continuous batching (cupic_server) against the same points served one
    at a time through the plain path with the same convergence rule."""
    import numpy as np
    reqs = draw_requests(n_req, rate, seed)
    window, rnd, rel = 500, 50, 0.005
    srv_cfg = cp.server_config(max_steps=max_steps, window_steps=window, round_steps=rnd,
                               rel_change=rel)
    srv = cp.Server(srv_cfg)
    start = srv.stats()["clock"]
    pending = list(range(n_req))
    ids = {}
    arrival = {}
    while True:
        now = srv.stats()["clock"] - start
        while pending and reqs[pending[0]][0] <= now:
            k = pending.pop(0)
            ids[k] = srv.submit(cp.params(**reqs[k][1]))
            arrival[k] = reqs[k][0]
        active, queued = srv.poll()
        if not pending and active == 0 and queued == 0:
            break
        if active == 0 and queued == 0 and pending:
            time.sleep(max(0.0, reqs[pending[0]][0] - (srv.stats()["clock"] - start)))
    res = {k: srv.result(i) for k, i in ids.items()}
    wall = max(r["t_done"] for r in res.values()) - start
    st = srv.stats()
    srv.close()
    lat = np.array([res[k]["t_done"] - start - arrival[k] for k in res])
    conv = sum(r["status"] == cp.POINT_CONVERGED for r in res.values())
    served = {"points": n_req, "converged": int(conv), "capped": int(n_req - conv),
              "points_per_s": n_req / wall, "wall_s": wall,
              "latency_p50_s": float(np.percentile(lat, 50)),
              "latency_p99_s": float(np.percentile(lat, 99)),
              "mean_steps": float(np.mean([r["steps"] for r in res.values()])),
              "rounds": st["rounds"], "batch_rebuilds": st["rebuilds"]}
    print(f"# served: {served}", file=sys.stderr, flush=True)
    # plain path: one point at a time, in arrival order, same rule
    t_clock = 0.0
    lat_p = []
    same = True
    t0 = time.perf_counter()
    for k in range(n_req):
        arr, d = reqs[k]
        t_clock = max(t_clock, arr)
        c = srv_cfg
        p = cp.params(**d, jacobi_mode=c.jacobi_mode, jacobi_tol=c.jacobi_tol,
                      jacobi_tol_init=c.jacobi_tol, jacobi_max_sweeps=c.jacobi_max_sweeps,
                      jacobi_check_every=c.jacobi_check_every)
        s0 = time.perf_counter()
        r = cp.Run(p)
        r.init_sim()
        t = p.dt
        steps, prev, have, ws = 0, 0.0, False, [0, 0]
        status = cp.POINT_CAPPED
        while steps < max_steps:
            r.step(t)
            t += p.dt
            steps += 1
            if steps % window == 0:
                tot = r.probe_tally()["total_probe"]
                ie = (tot[0] - ws[0]) / (window * p.dt)
                ii = (tot[1] - ws[1]) / (window * p.dt)
                ws = [int(tot[0]), int(tot[1])]
                if have and converged_rule(prev, ie, ii, rel):
                    status = cp.POINT_CONVERGED
                    break
                prev, have = ii - ie, True
        r.synchronize()
        tot = r.probe_tally()["total_probe"]
        r.close()
        t_clock += time.perf_counter() - s0
        lat_p.append(t_clock - arr)
        same &= (steps == res[k]["steps"] and status == res[k]["status"]
                 and [int(v) for v in tot] == res[k]["total_probe"])
    lat_p = np.array(lat_p)
    plain = {"points_per_s": n_req / max(t_clock, 1e-9),
             "latency_p50_s": float(np.percentile(lat_p, 50)),
             "latency_p99_s": float(np.percentile(lat_p, 99)),
             "busy_s": time.perf_counter() - t0}
    return {"workload": f"{n_req} I-V points, Poisson arrivals {rate:g}/s, ncx~U{{128..1024}}, "
                        f"ncy~U{{64..512}}, ppc~U{{20..200}}, phi_p~U[-10,2]; converged when "
                        f"the net probe current changes < {rel:g} of the gross current "
                        f"between {window}-step windows (cap {max_steps} steps); direct FP64 "
                        f"Poisson solve",
            "served": served, "plain_one_at_a_time": plain,
            "points_per_s_vs_plain": served["points_per_s"] / plain["points_per_s"],
            "results_identical_to_plain": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="prod", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-max-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile", type=int, default=0,
                    help="cells per wave strip in the push passes (1..16); skips the tuning")
    ap.add_argument("--batched-runs", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched leg (tile from --tile, else tuned; "
                         "sequential baseline from --batched-seq-steps) and print its JSON")
    ap.add_argument("--batched-seq-steps", type=int, default=3)
    ap.add_argument("--serve-requests", type=int, default=12)
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-max-steps", type=int, default=2000)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device with torch: a plain copy "
                         "kernel of known size that scales profiler byte counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: a single run of this path lives on one device "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    if a.copy_probe_mib > 0:
        import torch
        n = a.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy probe: {a.copy_probe_mib} MiB read + {a.copy_probe_mib} MiB written",
              file=sys.stderr, flush=True)
        # achievable HBM bandwidth of plain streaming kernels, for reference
        nb = n * 4
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        for label, fn, nbytes in (("copy", lambda: dst.copy_(src), 2 * nb),
                                  ("read (sum)", lambda: src.sum(), nb),
                                  ("write (fill)", lambda: dst.fill_(2.0), nb)):
            fn()
            ev[0].record()
            for _ in range(5):
                fn()
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1]) / 5
            print(f"# {label}: {nbytes / ms / 1e6:.0f} GB/s", file=sys.stderr, flush=True)
        del src, dst

    import cupic_amd as cp
    name, cus = cp.init_dev(0)
    if a.batched_only:
        res = batched_leg(cp, a.batched_runs, a.batched_steps, a.batched_seq_steps,
                          20261203, fixed_tile=a.tile)
        print(json.dumps({"batched": res}), flush=True)
        return
    cfg = CONFIGS[a.config]
    p = cp.params(**cfg, jacobi_mode=cp.JACOBI_RESIDUAL, jacobi_tol=1e-6,
                  jacobi_tol_init=1e-6, jacobi_max_sweeps=500, jacobi_check_every=50)
    print(f"# {name}, {cus} CUs; {workload_name(cfg, p)}", file=sys.stderr, flush=True)
    run = cp.Run(p)
    t0 = time.perf_counter()
    run.init_sim()
    run.synchronize()
    print(f"# init {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    t = p.dt
    for _ in range(a.warmup):
        run.step(t)
        t += p.dt
    tile = a.tile
    if tile:
        run.set_tile(tile)
    elif not a.no_autotune:
        # in-situ: each candidate tile steps this run (untimed warm-up steps)
        t, tile, _, ttimes = run.autotune(t, steps=2)
        print(f"# autotuned tile {tile}: {ttimes}", file=sys.stderr, flush=True)
    run.synchronize()
    print("# warm-up done", file=sys.stderr, flush=True)
    n0 = run.number_of_particles(0) + run.number_of_particles(1)
    sw0, _ = run.solver_totals()
    # The timed steps carry HIP events around the scatter only (the kernel
    # the roofline is quoted for); events around every kernel class cost
    # about 3 % of a step, so the per-class breakdown comes from a few more
    # steps after the timed ones.
    if not a.no_timing:
        run.timing(True, classes=["scatter"])
        run.timing_read()
    run.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run.step(t)
        t += p.dt
    run.synchronize()
    wall = time.perf_counter() - t0
    n1 = run.number_of_particles(0) + run.number_of_particles(1)
    sw1, _ = run.solver_totals()
    kern = {}
    if not a.no_timing:
        kern_timed = run.timing_read()
        run.timing(True)
        run.timing_read()
        for _ in range(BREAKDOWN_STEPS):
            run.step(t)
            t += p.dt
        run.synchronize()
        kern = run.timing_read()
        kern["scatter"] = kern_timed["scatter"]
    run.timing(False)
    ps = 0.5 * (n0 + n1) * a.steps
    value = ps / wall
    ms_step = 1e3 * wall / a.steps
    S = (sw1 - sw0) / a.steps
    nodes = p.ncx * (p.ncy + 1)
    npart = 0.5 * (n0 + n1)
    cycle_bytes = 84.0 * npart + nodes * (32.0 + 12.0 * S)
    tally = run.probe_tally()
    out = {
        "metric": "particle-steps/s (full PIC cycle)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload_name(cfg, p), "grid": [p.ncx, p.ncy],
                   "ppc": cfg["ppc"], "particles": int(n1), "jacobi_mode": "fp64-residual",
                   "jacobi_sweeps_per_step": S, "tile_cells": tile or "auto",
                   "device": name},
    }
    if kern:
        per = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in kern.items()}
        out["kernels_ms_per_launch"] = per
        out["kernels_ms_per_launch_source"] = (
            f"scatter: HIP events around each k_scatter of the {a.steps} timed steps; "
            f"the other classes: {BREAKDOWN_STEPS} further steps with events around "
            f"every kernel class")
        # the dominant kernel: the scatter reads the 16-byte state and writes
        # it and its 1-byte push hint
        dom = "scatter"
        bytes_per_particle = 33.0
        dur = per[dom] * 1e-3
        ach = bytes_per_particle * npart / dur if dur > 0 else 0.0
        tpp, src = profiled_traffic("k_" + dom) if a.config == "prod" else (None, None)
        out["roofline"] = {"bound": "hbm", "kernel": "k_" + dom, "achieved": ach / 1e9,
                           "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
                           "traffic": tpp * npart if tpp else None,
                           "traffic_unit": "HBM bytes per launch",
                           "traffic_source": src,
                           "bytes_per_particle": bytes_per_particle,
                           "avg_launch_ms": per[dom]}
    out["cycle_roofline"] = {"bytes_per_step": cycle_bytes,
                             "achieved_GBps": cycle_bytes / (wall / a.steps) / 1e9,
                             "frac": cycle_bytes / (wall / a.steps) / HBM_PEAK,
                             "formula": "84*(Ne+Ni) + nodes*(32 + 12*S)"}
    design_bytes = 34.0 * npart + nodes * (32.0 + 12.0 * S)
    out["cycle_roofline_design"] = {"bytes_per_step": design_bytes,
                                    "frac": design_bytes / (wall / a.steps) / HBM_PEAK,
                                    "formula": "34*(Ne+Ni) + nodes*(32 + 12*S): push_rank reads "
                                               "the 1 B push hint, scatter reads 16 B and writes "
                                               "16 B + the 1 B hint"}
    if kern:
        # the same bytes split between the particle kernels and the grid
        # kernels, each over its own kernels' device time per step (the
        # per-class times of the breakdown steps)
        per = out["kernels_ms_per_launch"]
        part_ms = sum(per[k] for k in ("push_rank", "inject", "scan", "scatter", "finish"))
        grid_ms = sum(per[k] for k in ("deposit", "solve", "field"))
        grid_bytes = nodes * (32.0 + 12.0 * S)
        out["particle_kernels_roofline"] = {
            "ms_per_step": part_ms,
            "frac": 84.0 * npart / (part_ms * 1e-3) / HBM_PEAK,
            "frac_design": 34.0 * npart / (part_ms * 1e-3) / HBM_PEAK,
            "formula": "84*(Ne+Ni) (design: 34*(Ne+Ni)) over push_rank + inject + scan + "
                       "scatter + finish"}
        out["grid_kernels_roofline"] = {
            "ms_per_step": grid_ms,
            "frac": grid_bytes / (grid_ms * 1e-3) / HBM_PEAK,
            "formula": "nodes*(32 + 12*S) over deposit + solve + field, against the HBM "
                       "peak; the Jacobi grids stay in the 256 MB Infinity Cache, which "
                       "serves most of these bytes"}
    out["probe"] = {"total_absorbed": [int(v) for v in tally["total_probe"]],
                    "charge_collected": tally["charge"]}
    run.close()
    if not a.no_direct:
        out["plain_direct_poisson"] = direct_leg(cp, cfg, a.steps, a.warmup, tile)
        print(f"# direct: {out['plain_direct_poisson']}", file=sys.stderr, flush=True)
    if a.batched_runs > 0:
        out["batched"] = batched_leg(cp, a.batched_runs, a.batched_steps,
                                     a.batched_seq_steps, 20261203)
        print(f"# batched: {out['batched']}", file=sys.stderr, flush=True)
    if a.serve_requests > 0:
        out["serving"] = serving_leg(cp, a.serve_requests, a.serve_rate, a.serve_max_steps,
                                     20261203)
    if not a.no_cpu and a.cpu_max_steps > 0:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, a.cpu_seconds, a.cpu_max_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
