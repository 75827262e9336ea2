"""Full PIC cycle throughput on one MI355X.

    python bench.py [--steps K] [--warmup W] [--config prod|small] [--solver direct|jacobi]

Workload (default, "prod"): 1024x512 cells, ds 0.5, dt 0.05, 200 particles
per cell per species (2 x 104,857,600), mi/me 1836, Te/Ti 10, probe bias
-10 kTe/e, FP32 particles and fields, int64 fixed-point charge, the Poisson
system solved to an FP64 residual below 1e-6 every step, 5000 steps (the
whole run).  The solve is the fixed point of the Jacobi iteration computed
directly in FP64 (DFT along x, sine transform along y); Jacobi capped at
500 sweeps cannot reach 1e-6 on this grid and drives the potential out of
[phi_p, 0] within a few hundred steps, which the "jacobi_capped" leg
measures.  Synthetic seeded initial load; the first steps after init are
warm-up.  One step is one pass of the cycle deposit -> solve -> field ->
move -> cc over both species.

Prints one JSON line: particle-steps/s over the timed steps (inputs resident
in HBM), the roofline of the dominant kernel (HIP events on the run's
stream), the measured HBM traffic of a whole step, and CPU baselines (the
oracle's OpenMP FP64 build: a bounded sample of the same workload, and
the small configuration's 1000 steps).
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


def workload_name(cfg, p, solver):
    how = "FFT" if direct_uses_fft(cfg["ncx"], cfg["ncy"]) else "dense products"
    solve = ("Poisson to an FP64 residual < 1e-6 each step: the Jacobi fixed point solved "
             f"directly in FP64 ({how})" if solver == "direct" else
             f"Jacobi FP64 residual {p.jacobi_tol:g} (max {p.jacobi_max_sweeps} sweeps)")
    return (f"{cfg['ncx']}x{cfg['ncy']} cells, {cfg['ppc']} ppc/species "
            f"(2x{p.per_row * p.ncy} particles), mi/me={cfg['mi']:g}, "
            f"Te/Ti={1.0 / cfg['kti']:g}, phi_p={cfg['phi_p']:g} kTe/e, {solve}")


def physical_cores():
    """Physical cores among the CPUs this process may run on, from
    /proc/cpuinfo (distinct (package, core) pairs), and that CPU count."""
    cpus = sorted(os.sched_getaffinity(0))
    pairs = set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in list(f) + [""]:
                if not line.strip():
                    if "processor" in cur and int(cur["processor"]) in cpus:
                        pairs.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
    except OSError:
        pass
    return (len(pairs) or None), len(cpus)


def cpu_baseline(cfg, min_seconds, max_steps, solver="jacobi"):
    """The oracle's OpenMP FP64 cycle on the same workload (the CPU checker,
    never the measured path): full cycles after its own init, repeated until
    min_seconds of stepping have passed (at least 2, at most max_steps).  The
    oracle solves Poisson as the GPU run does: directly in FP64 (FFT along
    x, Thomas along y), or with FP64 Jacobi sweeps (500 at most per step)."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    phys, ncpu = physical_cores()
    p = O.make_params(**cfg)
    sim = O.Sim(p, omp=True)
    if solver == "direct":
        sim.set_solver(3, 1e-6, 1e-6, 0, 0, 50)
        how = "Poisson solved directly"
    else:
        sim.set_solver(1, 1e-6, 1e-6, 0, 500, 50)
        how = "Jacobi capped at 500 sweeps"
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
            "physical_cores": phys, "cpus_available": ncpu, "steps": steps,
            "ms_per_step": 1e3 * el / steps,
            "sample": f"{steps} full cycles of the same workload after init "
                      f"(FP64, OpenMP, {threads} threads on {phys} physical cores, {how}), "
                      f"{el:.1f} s"}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "cycle_1024x512_200ppc_hbm_traffic.md")


def profiled_traffic(kernel):
    """HBM bytes per particle of a kernel ("step": all kernels of one step),
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-calibrated), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic_bytes_per_particle" and w[1] == kernel:
                    return float(w[2]), os.path.relpath(TRAFFIC_PROFILE, ROOT)
    except OSError:
        pass
    return None, None

def run_params(cp, cfg, solver, tile=0):
    mode = cp.POISSON_DIRECT if solver == "direct" else cp.JACOBI_RESIDUAL
    return cp.params(**cfg, jacobi_mode=mode, jacobi_tol=1e-6, jacobi_tol_init=1e-6,
                     jacobi_max_sweeps=500, jacobi_check_every=50, tile_cells=tile)


def end_state(cp, run):
    """Solver residual, potential range, probe and sort statistics of a run
    (read after its timed steps)."""
    sweeps, resid = run.last_solve()
    phi = run.mesh(cp.MESH_PHI)
    tally = run.probe_tally()
    return {"last_solve_sweeps": sweeps, "last_solve_fp64_residual": resid,
            "phi_min": float(phi.min()), "phi_max": float(phi.max()),
            "phi_p": run.p.phi_p,
            "particles": [run.number_of_particles(0), run.number_of_particles(1)],
            "probe_absorbed_total": [int(v) for v in tally["total_probe"]],
            "probe_charge_total": tally["charge"]}


def jacobi_capped_leg(cp, cfg, steps, tile):
    """The same workload with FP32 Jacobi capped at 500 sweeps per step (FP64
    residual checked every 50): its throughput over `steps` steps from init,
    and where the potential has gone by then."""
    p = run_params(cp, cfg, "jacobi", tile)
    run = cp.Run(p)
    run.init_sim()
    t = p.dt
    run.step(t)
    t += p.dt
    run.synchronize()
    c0 = run.counters()
    t0 = time.perf_counter()
    for _ in range(steps - 1):
        run.step(t)
        t += p.dt
    run.synchronize()
    wall = time.perf_counter() - t0
    c1 = run.counters()
    out = {"steps": steps, "ms_per_step": 1e3 * wall / (steps - 1),
           "value": (c1["particle_steps"] - c0["particle_steps"]) / wall,
           "unit": "particle-steps/s", "far_movers_total": c1["far"]}
    try:
        out.update(end_state(cp, run))
    except cp.CupicError as e:
        out["error"] = str(e)
    out["note"] = ("500 sweeps remove about 1 % of the long-wavelength error per step on "
                   "513 rows, so the FP64 residual stays far above 1e-6 and the potential "
                   "drifts out of [phi_p, 0]: not the physics of the configuration")
    run.close()
    return out


def small_leg(cp, steps):
    """The small configuration (128x64 cells, 20 ppc, mi/me 100, Te/Ti 1,
    phi_p -5, Jacobi FP64 residual 1e-6 with at most 500 sweeps) for its
    whole run on the GPU."""
    p = run_params(cp, CONFIGS["small"], "jacobi")
    run = cp.Run(p)
    run.init_sim()
    run.synchronize()
    c0 = run.counters()
    t = p.dt
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step(t)
        t += p.dt
    run.synchronize()
    wall = time.perf_counter() - t0
    c1 = run.counters()
    sw, _ = run.solver_totals()
    out = {"steps": steps, "ms_per_step": 1e3 * wall / steps,
           "value": (c1["particle_steps"] - c0["particle_steps"]) / wall,
           "unit": "particle-steps/s", "jacobi_sweeps_per_step": sw / (steps + 1)}
    out.update(end_state(cp, run))
    run.close()
    return out


def cpu_small(steps):
    """The oracle's OpenMP FP64 cycle on the small configuration, all steps."""
    r = cpu_baseline(CONFIGS["small"], 1e9, steps)
    r["sample"] = (f"the whole small-configuration run, {steps} steps after init (FP64, "
                   f"OpenMP, {r['cores']} threads, Jacobi FP64 residual 1e-6, max 500 sweeps)")
    return r


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


def direct_uses_fft(ncx, ncy):
    """The library's choice for the direct solve: FFT kernels when ncx and
    2 ncy are powers of two up to 4096 (and ncy >= 2), dense products
    otherwise."""
    def ok(n):
        return 2 <= n <= 4096 and n & (n - 1) == 0
    return ncy >= 2 and ok(ncx) and ok(2 * ncy)


# direct_method of the batched and served runs (--direct-method)
DIRECT_METHOD = 0


def batched_leg(cp, n_runs, steps, seq_steps, seed, fixed_tile=0, solver="direct"):
    """One HIP-graph step advances all runs; baseline: the same runs stepped
    one after another through the plain path (cupic_step per run), skipped
    when seq_steps is 0.  The tile (cells per wave strip) is tuned in place
    on the batch itself: each candidate is set on every run and timed over 2
    batched steps, unless fixed_tile names one.  The Poisson system is solved
    to the FP64 residual 1e-6 (directly); Jacobi capped at 500 sweeps lets
    the larger grids' potentials run away within a few hundred steps."""
    draws = draw_batched(n_runs, seed)
    runs = []
    mode = cp.POISSON_DIRECT if solver == "direct" else cp.JACOBI_RESIDUAL
    for d in draws:
        p = cp.params(**d, jacobi_mode=mode, jacobi_tol=1e-6,
                      jacobi_tol_init=1e-6, jacobi_max_sweeps=500, jacobi_check_every=50,
                      direct_method=DIRECT_METHOD)
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
    # where the runs' potentials are after the window: runs whose potential
    # left the range spanned by the bias and the grounded top by more than
    # 10 kTe have not stayed physical under the capped solve
    wild = 0
    for r in runs:
        ph = r.mesh(cp.MESH_PHI)
        lo, hi = min(r.p.phi_p, 0.0) - 10.0, max(r.p.phi_p, 0.0) + 10.0
        wild += int(ph.min() < lo or ph.max() > hi)
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
            "direct_method": ("auto", "dense", "fft")[DIRECT_METHOD],
            "ms_per_batched_step": 1e3 * wall_b / steps,
            "run_steps_per_s": rs_b, "particle_steps_per_s": npart * steps / wall_b,
            "sequential_plain_run_steps_per_s": rs_s,
            "sequential_plain_steps_per_run": seq_steps,
            "speedup_vs_sequential_plain": rs_b / rs_s if rs_s else None,
            "runs_with_phi_beyond_10kTe_of_range": wild,
            "workload": "256 runs, ncx~U{64..512}, ncy~U{32..256}, ppc~U{20..100}, "
                        "phi_p~U[-10,2], mi/me 100, Te/Ti 1, Poisson to the FP64 residual "
                        + ("1e-6 solved directly" if solver == "direct" else
                           "1e-6 by Jacobi (max 500 sweeps)")
                        + f"; batch timed over {steps} steps, the plain path over "
                          f"{seq_steps} steps of every run; one HIP graph per batched step"}


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
                            phi_p=float(rng.uniform(-10.0, 2.0)), run_id=1000 + k,
                            direct_method=DIRECT_METHOD)))
    return out


def converged_rule(prev_net, ie, ii, rel):
    gross = ii + ie
    return gross > 0.0 and abs((ii - ie) - prev_net) / gross < rel


def serving_leg(cp, n_req, rate, max_steps, seed, plain="full", max_particles=0, policy=0):
    """Continuous batching (cupic_server) against the same points served one
    at a time through the plain path with the same convergence rule.  plain
    "full" steps every point through the plain path and checks that the
    results are the same; "estimate" times each point's init and 100 plain
    steps and prices the steps the point needed (the alone-and-in-a-mix
    identity is a GPU test)."""
    import numpy as np
    reqs = draw_requests(n_req, rate, seed)
    window, rnd, rel = 500, 50, 0.005
    srv_cfg = cp.server_config(max_steps=max_steps, window_steps=window, round_steps=rnd,
                               rel_change=rel, max_particles=max_particles, policy=policy)
    srv = cp.Server(srv_cfg)
    start = srv.stats()["clock"]
    pending = list(range(n_req))
    ids = {}
    arrival = {}
    resident = []          # (clock, points resident) after every poll
    while True:
        now = srv.stats()["clock"] - start
        while pending and reqs[pending[0]][0] <= now:
            k = pending.pop(0)
            ids[k] = srv.submit(cp.params(**reqs[k][1]))
            arrival[k] = reqs[k][0]
        active, queued = srv.poll()
        resident.append((srv.stats()["clock"] - start, active))
        if not pending and active == 0 and queued == 0:
            break
        if active == 0 and queued == 0 and pending:
            time.sleep(max(0.0, reqs[pending[0]][0] - (srv.stats()["clock"] - start)))
    res = {k: srv.result(i) for k, i in ids.items()}
    wall = max(r["t_done"] for r in res.values()) - start
    st = srv.stats()
    srv.close()
    lat = {k: res[k]["t_done"] - start - arrival[k] for k in res}
    conv = [k for k in res if res[k]["status"] == cp.POINT_CONVERGED]
    # particles moved, approximated by each point's steps x its particles at the end
    psteps = sum(r["steps"] * sum(r["particles"]) for r in res.values())
    occ = np.array(resident)
    mean_res = float(np.sum(occ[1:, 1] * np.diff(occ[:, 0])) / max(occ[-1, 0] - occ[0, 0], 1e-9)) \
        if len(occ) > 1 else 0.0

    def pct(v, q):
        return float(np.percentile(v, q)) if len(v) else None
    served = {"points": n_req, "converged": len(conv), "capped": n_req - len(conv),
              "points_per_s": n_req / wall,
              "converged_points_per_s": len(conv) / wall, "wall_s": wall,
              "latency_converged_p50_s": pct([lat[k] for k in conv], 50),
              "latency_converged_p99_s": pct([lat[k] for k in conv], 99),
              "latency_all_p50_s": pct(list(lat.values()), 50),
              "latency_all_p99_s": pct(list(lat.values()), 99),
              "steps_converged": sorted(int(res[k]["steps"]) for k in conv),
              "mean_steps": float(np.mean([r["steps"] for r in res.values()])),
              "particle_steps_per_s": psteps / wall,
              "mean_resident_points": mean_res,
              "rounds": st["rounds"], "batch_rebuilds": st["rebuilds"]}
    print(f"# served: {served}", file=sys.stderr, flush=True)
    c = srv_cfg

    def plain_params(d):
        return cp.params(**d, jacobi_mode=c.jacobi_mode, jacobi_tol=c.jacobi_tol,
                         jacobi_tol_init=c.jacobi_tol, jacobi_max_sweeps=c.jacobi_max_sweeps,
                         jacobi_check_every=c.jacobi_check_every)
    t_clock = 0.0
    lat_p = []
    same = True
    busy = 0.0
    for k in range(n_req):
        arr, d = reqs[k]
        t_clock = max(t_clock, arr)
        s0 = time.perf_counter()
        r = cp.Run(plain_params(d))
        r.init_sim()
        t = r.p.dt
        if plain == "estimate":
            r.synchronize()
            t_init = time.perf_counter() - s0
            n_probe = min(100, int(res[k]["steps"]))
            s1 = time.perf_counter()
            for _ in range(n_probe):
                r.step(t)
                t += r.p.dt
            r.synchronize()
            per = (time.perf_counter() - s1) / max(n_probe, 1)
            dt_point = t_init + per * res[k]["steps"]
            r.close()
        else:
            steps, prev, have, ws = 0, 0.0, False, [0, 0]
            status = cp.POINT_CAPPED
            while steps < max_steps:
                r.step(t)
                t += r.p.dt
                steps += 1
                if steps % window == 0:
                    tot = r.probe_tally()["total_probe"]
                    ie = (tot[0] - ws[0]) / (window * r.p.dt)
                    ii = (tot[1] - ws[1]) / (window * r.p.dt)
                    ws = [int(tot[0]), int(tot[1])]
                    if have and converged_rule(prev, ie, ii, rel):
                        status = cp.POINT_CONVERGED
                        break
                    prev, have = ii - ie, True
            r.synchronize()
            tot = r.probe_tally()["total_probe"]
            r.close()
            dt_point = time.perf_counter() - s0
            same &= (steps == res[k]["steps"] and status == res[k]["status"]
                     and [int(v) for v in tot] == res[k]["total_probe"])
        busy += dt_point
        t_clock += dt_point
        lat_p.append(t_clock - arr)
    plain_res = {"mode": plain, "points_per_s": n_req / max(t_clock, 1e-9),
                 "converged_points_per_s": len(conv) / max(t_clock, 1e-9),
                 "latency_converged_p50_s": pct([lat_p[k] for k in conv], 50),
                 "latency_converged_p99_s": pct([lat_p[k] for k in conv], 99),
                 "busy_s": busy, "particle_steps_per_s": psteps / max(busy, 1e-9)}
    out = {"workload": f"{n_req} I-V points, Poisson arrivals {rate:g}/s, ncx~U{{128..1024}}, "
                       f"ncy~U{{64..512}}, ppc~U{{20..200}}, phi_p~U[-10,2], mi/me 100, "
                       f"Te/Ti 1; converged when the net probe current changes < {rel:g} of "
                       f"the gross current between {window}-step windows (cap {max_steps} "
                       f"steps); direct FP64 Poisson solve",
           "direct_method": ("auto", "dense", "fft")[DIRECT_METHOD],
           "admission": {"policy": ("fcfs", "smallest_first")[policy],
                         "max_resident_particles": max_particles or None},
           "served": served, "plain_one_at_a_time": plain_res,
           "points_per_s_vs_plain": served["points_per_s"] / plain_res["points_per_s"]}
    if plain == "full":
        out["results_identical_to_plain"] = bool(same)
    return out

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000,
                    help="timed steps (default: the configuration's whole 5000-step run)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="prod", choices=sorted(CONFIGS))
    ap.add_argument("--solver", default=None, choices=["direct", "jacobi"],
                    help="Poisson solve of the headline run (prod: direct, small: jacobi)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-max-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile", type=int, default=0,
                    help="cells per wave strip in the push passes (1..16); skips the tuning")
    ap.add_argument("--jacobi-steps", type=int, default=300,
                    help="steps of the capped-Jacobi leg (0: skip)")
    ap.add_argument("--small-steps", type=int, default=1000,
                    help="steps of the small-configuration leg, GPU and CPU (0: skip)")
    ap.add_argument("--batched-runs", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=500,
                    help="timed batched steps (default: one 500-step convergence window)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched leg (tile from --tile, else tuned; "
                         "sequential baseline from --batched-seq-steps) and print its JSON")
    ap.add_argument("--batched-seq-steps", type=int, default=50,
                    help="timed plain steps of every run for the sequential baseline")
    ap.add_argument("--direct-method", default="auto", choices=["auto", "dense", "fft"],
                    help="direct Poisson solve of the batched and served runs: FFT for "
                         "power-of-two lengths else dense products (auto), dense products "
                         "always, or FFT for any length (Bluestein)")
    ap.add_argument("--serve-requests", type=int, default=50)
    ap.add_argument("--serve-max-particles", type=float, default=3e8,
                    help="server's resident-particle budget (0: none)")
    ap.add_argument("--serve-policy", default="fcfs", choices=["fcfs", "smallest"])
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-max-steps", type=int, default=10000)
    ap.add_argument("--serve-plain", default="estimate", choices=["full", "estimate"],
                    help="plain baseline: every point stepped through the plain path, or "
                         "each point's init and 100 steps timed and its steps priced")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving leg and print its JSON")
    ap.add_argument("--copy-probe-mib", type=int, default=1024,
                    help="size of the plain copy/read/write kernels run first: the achievable "
                         "HBM bandwidth, and the dispatch that scales profiler byte counters "
                         "(0: skip)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: a single run of this path lives on one device "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)
    solver = a.solver or ("direct" if a.config == "prod" else "jacobi")
    global DIRECT_METHOD
    DIRECT_METHOD = {"auto": 0, "dense": 1, "fft": 2}[a.direct_method]

    import cupic_amd as cp
    name, cus = cp.init_dev(0)
    hbm_probe = {}
    if a.copy_probe_mib > 0:
        nb = a.copy_probe_mib << 20
        for kind in ("copy", "read", "write"):
            ms, rate = cp.bandwidth_probe(nb, kind)
            hbm_probe[kind + "_GBps"] = rate / 1e9
        print(f"# plain streaming kernels over {a.copy_probe_mib} MiB: {hbm_probe}",
              file=sys.stderr, flush=True)
    if a.serving_only:
        res = serving_leg(cp, a.serve_requests, a.serve_rate, a.serve_max_steps, 20261203,
                          a.serve_plain, int(a.serve_max_particles),
                          1 if a.serve_policy == "smallest" else 0)
        print(json.dumps({"serving": res}), flush=True)
        return
    if a.batched_only:
        res = batched_leg(cp, a.batched_runs, a.batched_steps, a.batched_seq_steps,
                          20261203, fixed_tile=a.tile)
        print(json.dumps({"batched": res}), flush=True)
        return
    cfg = CONFIGS[a.config]
    p = run_params(cp, cfg, solver)
    print(f"# {name}, {cus} CUs; {workload_name(cfg, p, solver)}", file=sys.stderr, flush=True)
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
    c0 = run.counters()
    print(f"# warm-up done ({c0['steps']} steps)", file=sys.stderr, flush=True)
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
    c1 = run.counters()
    sw1, _ = run.solver_totals()
    psteps = c1["particle_steps"] - c0["particle_steps"]
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
    # share of particles the counting pass pushed in full (hint missing or
    # overturnable), from a few steps with the counter on
    run.set_statistics(True)
    ca = run.counters()
    for _ in range(BREAKDOWN_STEPS):
        run.step(t)
        t += p.dt
    cb = run.counters()
    run.set_statistics(False)
    und = (sum(cb["undecided"]) - sum(ca["undecided"])) / \
        max(cb["particle_steps"] - ca["particle_steps"], 1)
    value = psteps / wall
    ms_step = 1e3 * wall / a.steps
    npart = psteps / a.steps            # mean particles moved per step
    state = end_state(cp, run)
    far = [c1["far"][s] - c0["far"][s] for s in (0, 1)]
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
        "config": {"workload": workload_name(cfg, p, solver), "grid": [p.ncx, p.ncy],
                   "ppc": cfg["ppc"], "poisson": solver,
                   "mean_particles_per_step": npart,
                   "untimed_steps_before": c0["steps"],
                   "jacobi_sweeps_per_step": (sw1 - sw0) / a.steps,
                   "tile_cells": tile or "auto", "device": name, **state,
                   "far_movers_per_step": [f / a.steps for f in far],
                   "undecided_hint_share_end": und},
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
    if kern and solver == "direct":
        solve_ms = out["kernels_ms_per_launch"]["solve"]
        m, nx = p.ncy - 1, p.ncx
        if direct_uses_fft(p.ncx, p.ncy):
            # the FFT path in a step: the three transform kernels (the solve
            # is started by the charge conversion, its FP64 residual is
            # reduced by the field kernel); algorithmic bytes: b and the
            # Dirichlet rows in (8 B per node), the x spectra out and in
            # (2 x 8), the y solve's result out and in (2 x 8), the FP64 and
            # FP32 potential out (12): 52 B per interior node
            nbytes = 52.0 * m * nx
            ach = nbytes / (solve_ms * 1e-3) if solve_ms > 0 else 0.0
            out["solve_roofline"] = {"bound": "hbm", "achieved": ach / 1e9,
                                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                     "frac": ach / HBM_PEAK, "bytes_per_solve": nbytes,
                                     "avg_launch_ms": solve_ms,
                                     "note": "the solve's three FFT kernels: three dependent "
                                             "launches of about 12 us each, launch-latency "
                                             "bound"}
        else:
            # the dense path's four FP64 products: 2 x (m x ncx x ncx) for the
            # DFT along x and 2 x (m x m x ncx) for the sine transform along
            # y, against the dense FP64 matrix peak
            flops = 2.0 * (2.0 * m * nx * nx + 2.0 * m * m * nx)
            ach = flops / (solve_ms * 1e-3) / 1e12 if solve_ms > 0 else 0.0
            out["solve_roofline"] = {"bound": "mfma", "achieved": ach, "peak": FP64_MFMA_PEAK,
                                     "unit": "TFLOP/s", "frac": ach / FP64_MFMA_PEAK,
                                     "flops_per_solve": flops, "avg_launch_ms": solve_ms,
                                     "note": "the solve's four products and slice sums "
                                             "against the dense FP64 matrix peak"}
    # measured HBM bytes of a whole step (every kernel, rocprofv3 counters,
    # copy-calibrated) over the measured step time
    tstep, src = profiled_traffic("step") if a.config == "prod" else (None, None)
    if tstep:
        bps = tstep * npart
        out["step_hbm_traffic"] = {"bytes_per_step": bps,
                                   "bytes_per_particle_step": tstep,
                                   "achieved_GBps": bps / (ms_step * 1e-3) / 1e9,
                                   "frac_of_peak": bps / (ms_step * 1e-3) / HBM_PEAK,
                                   "source": src}
    if hbm_probe:
        out["hbm_probe"] = hbm_probe
    run.close()
    if a.jacobi_steps > 0 and a.config == "prod" and solver == "direct":
        out["jacobi_capped"] = jacobi_capped_leg(cp, cfg, a.jacobi_steps, tile)
        print(f"# jacobi capped: {out['jacobi_capped']}", file=sys.stderr, flush=True)
    if a.small_steps > 0:
        out["small_config"] = small_leg(cp, a.small_steps)
        print(f"# small config: {out['small_config']}", file=sys.stderr, flush=True)
    if a.batched_runs > 0:
        out["batched"] = batched_leg(cp, a.batched_runs, a.batched_steps,
                                     a.batched_seq_steps, 20261203)
        print(f"# batched: {out['batched']}", file=sys.stderr, flush=True)
    if a.serve_requests > 0:
        out["serving"] = serving_leg(cp, a.serve_requests, a.serve_rate, a.serve_max_steps,
                                     20261203, a.serve_plain, int(a.serve_max_particles),
                                     1 if a.serve_policy == "smallest" else 0)
    if not a.no_cpu and a.cpu_max_steps > 0:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, a.cpu_seconds, a.cpu_max_steps, solver)
        if a.small_steps > 0:
            out["cpu_baseline_small_config"] = cpu_small(a.small_steps)
            g = out.get("small_config", {}).get("value")
            if g:
                out["cpu_baseline_small_config"]["gpu_over_cpu"] = \
                    g / out["cpu_baseline_small_config"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
