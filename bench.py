"""Throughput of the FP64 MPM step at 16M particles on one MI355X.

Workload (the 16M-particle throughput configuration): a 200x100x100-cell
block, 8 points per cell (16,000,000), on the 512x256x256 grid with
h = 1/64 m, E = 9e9 Pa, nu = 0.3, friction angle 60 deg, cohesion 2e6 Pa,
sphere indenter r = 0.1 m at 1 m/s. The configuration fixes layout,
material and indenter but no time step; the headline uses dt = 2e-6 s
(elastic CFL number 0.47: the P-wave speed sqrt((lambda + 2 mu)/rho) =
3634 m/s times dt/h). The same workload at dt = 1e-5 s (CFL 2.33: the
block blows apart, every point crushes and spreads over ~4x more grid
bricks) is timed after it and reported as "unstable_dt".

The same workload at 27 and 64 points per cell (134x67x66 and 100x50x50
cells on the same grid, 200 timed steps each) is reported as
"points_per_cell".

The 2M-point block with the damage model (256x128x128 grid, dt = 5e-6 s,
2000 substeps, the indenter force read back every 10 substeps) is reported
as "damage".

The calibration adjoint (reverse mode of the elastic step on the
200,000-point calibration scene: 40x25x25 cells on a 128x64x64 grid, 400
substeps, force sampled every 10, checkpoints every 20) is timed after it and
reported as "adjoint": the backward pass per substep (replay from the
checkpoints included) over a plain forward substep of the same scene.

The batched ensemble (64 independent scenes of a material sweep, counts
log-uniform in [20k, 500k], each on its own 128x64x64 sub-grid, advanced in
the same launches, launch settings picked by icempm_autotune) is timed last
and reported as "ensemble", with its throughput as a fraction of the
single-scene headline: launched substep by substep ("launched_value") and
replayed from hipGraphs captured per run between re-sorts ("value").

One step = one substep of the reference's Model3D::Step loop (reset, P2G,
grid update, G2P with the return mapping), with the G2P of a step fused with
the P2G of the next and the brick re-sort every --sort-period steps inside
the timed region. Points are uploaded before timing (HBM-resident).

Prints one JSON line: particle-updates/s, the HBM roofline of the dominant
kernel (algorithmic bytes / HIP-event duration on the library's stream,
measured traffic from the committed rocprofv3 counter summary) and the
host-core baseline (the oracle on this host's cores, a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes (FP64 fields of the reference SoA):
# P2G reads utility, pos, vel, Fe, Bp = 25 doubles; G2P reads 14 and writes 26;
# per node holding mass: flush 32 B (P2G), update 32 B read + 24 B write,
# G2P 24 B read, reset 32 B.
B_P2G_PT, B_G2P_PT = 200, 320
B_P2G_NODE, B_UPD_NODE, B_G2P_NODE, B_RESET_NODE = 32, 56, 24, 32
B_SORT_PT = 24 + 8 + 4 * 16 + 416
# HBM traffic per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by the same counters on a copy of known size), keyed by kernel phase
# and workload; written by hand from the profile summary under profiles/.
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cpus() -> dict:
    """The host's processors: the CPU model, os.cpu_count() and the CPUs this
    process may run on (its affinity mask, the share of the host it is given)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        share = os.cpu_count() or 1
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "process_cpus": share}


def default_cpu_threads() -> int:
    """Every CPU of this process's share, or OMP_NUM_THREADS where that is set lower."""
    share = host_cpus()["process_cpus"]
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(share, int(env)) if env.isdigit() and int(env) > 0 else share


def cpu_baseline(scene, prms, hs, max_seconds: float, max_steps: int, threads: int) -> dict:
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc
    from icempm_3dm_amd import capi

    shape, center, vel = scene.indenter()
    ind = capi.Indenter.make(shape, center, vel)
    soa = hs.copy()
    grid = orc.grid_zeros(prms)
    bins = np.zeros(prms.IndenterArraySize())
    t_sim = 0.0
    # one untimed step (page-in), then timed steps until the budget is spent
    _, t_sim = orc.step(prms, ind, soa.host_buffer, soa.capacity, soa.size, grid, bins, 1, t_sim, 0, threads)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        _, t_sim = orc.step(prms, ind, soa.host_buffer, soa.capacity, soa.size, grid, bins, 1, t_sim, 1 + steps,
                            threads)
        steps += 1
        if time.perf_counter() - t0 >= max_seconds:
            break
    dt = time.perf_counter() - t0
    hc = host_cpus()
    return {
        "value": soa.size * steps / dt,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps of the same {soa.size}-point workload after 1 warm-up step "
                  f"(oracle/mpm_oracle.cpp, OpenMP, {threads} threads = every CPU of this process's share: "
                  f"{hc['process_cpus']} of the host's {hc['host_cpus']}, {hc['cpu_model']}); {dt:.1f} s",
        **hc,
    }


def run_unstable(scene, hs, args) -> dict:
    """The same points at the larger time step args.unstable_dt, where every
    point crushes and the cloud spreads over more, sparser brick segments."""
    sub = argparse.Namespace(**vars(args))
    sub.steps = args.unstable_steps
    log(f"bench: dt={args.unstable_dt}: {args.warmup} warm-up + {sub.steps} timed steps")
    el2, _, err2, _, stats2 = run_steps(scene, hs, sub, args.unstable_dt, False)
    n = hs.size
    return {"dt": args.unstable_dt, "steps": sub.steps, "value": n * sub.steps / el2,
            "ms_per_step": el2 * 1e3 / sub.steps, "device_error": err2,
            "brick_segments_at_end": stats2["segments"], "crushed_fraction": stats2["crushed_fraction"],
            "launch": stats2["launch"]}


def run_steps(scene, hs, args, dt: float, timed: bool):
    """Uploads the points, runs warm-up + timed steps at time step dt;
    returns (elapsed s, phase timings, device error, active nodes)."""
    from dataclasses import replace
    from icempm_3dm_amd import GPUImplementation6, scenes

    sc = replace(scene, dt=dt)
    prms = scenes.make_params(sc)
    g = GPUImplementation6(0)
    g.hssoa = hs.copy()
    g.allocate_arrays(prms, hs.capacity)
    g.update_constants(prms)
    shape, center, vel = sc.indenter()
    g.set_indenter(shape, center, vel)
    g.set_sort_period(args.sort_period)
    g.set_fusion(args.fuse != 0)
    g.set_blocks_per_cu(args.blocks_per_cu)
    g.transfer_ponts_to_device()
    t_sim, step = 0.0, 0
    t_sim = g.step(args.warmup, t_sim, step)
    step += args.warmup
    g.synchronize()
    launch = {"fuse": args.fuse != 0, "sort_period": args.sort_period, "blocks_per_cu": args.blocks_per_cu,
              "autotuned": False}
    if args.autotune:
        # launch settings from the autotuner, timed on the warmed-up state (which it restores)
        launch = dict(g.autotune(50), autotuned=True)
    if timed:
        g.enable_timing(True)
        g.reset_timings()
    g.synchronize()
    t0 = time.perf_counter()
    # the window opens with a brick re-sort, so that it holds at least its
    # share of the re-sorts at the chosen period (1 + (steps - 1) // period)
    g.sort_by_brick()
    t_sim = g.step(args.steps, t_sim, step)
    g.synchronize()
    elapsed = time.perf_counter() - t0
    timings = g.read_timings() if timed else {}
    if timed:
        # one brick re-sort timed on its own (the timed window may hold none)
        timings["sort_in_run"] = timings.pop("sort")
        g.reset_timings()
        g.sort_by_brick()
        timings["sort"] = g.read_timings()["sort"]
    g.enable_timing(False)
    err = g.device_error_code()
    stats = g.read_stats()
    g.transfer_from_device()
    crushed = int(np.count_nonzero(g.hssoa.utility()[:g.hssoa.size] & (1 << 16)))
    stats["crushed_fraction"] = crushed / max(g.hssoa.size, 1)
    active_nodes = 0
    if timed:
        # nodes holding mass after the last step (for the per-node algorithmic bytes)
        g.reset_grid()
        g.p2g()
        grid = g.download_grid()
        active_nodes = int(np.count_nonzero(grid[0]))
        del grid
    g.close()
    stats["launch"] = launch
    return elapsed, timings, err, active_nodes, stats


def run_ensemble(args) -> dict:
    """Warm-up + timed steps of the sweep ensemble; particle-updates/s."""
    from icempm_3dm_amd.ensemble import BatchedEnsemble, sweep_scenes

    sw = sweep_scenes(args.ensemble, seed=2024, dt=args.dt)
    ens = BatchedEnsemble(sw)
    tuning = ens.autotune(50)
    n = ens.n_points

    def timed(steps):
        ens.step(args.warmup)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.step(steps)
        ens.synchronize()
        return time.perf_counter() - t0

    # launched substep by substep, then replayed from captured hipGraphs
    # (one graph per run between re-sorts, captured during the warm-up); the
    # ensemble's value is that of the faster mode, both are reported
    elapsed_launch = timed(args.steps)
    elapsed_graph, captures = None, 0
    if args.graph:
        ens.set_graph(True)
        elapsed_graph = timed(args.steps)
        captures = ens.graph_captures()[0]
    graph_faster = elapsed_graph is not None and elapsed_graph < elapsed_launch
    elapsed = elapsed_graph if graph_faster else elapsed_launch
    import ctypes as C
    from icempm_3dm_amd import capi
    c = C.c_int(0)
    capi.check(capi.lib().icempm_error_code(ens.ctx, C.byref(c)))
    code = c.value
    counts = [s.n_points for s in sw]
    ens.close()
    return {"scenes": len(sw), "points": n, "points_min": min(counts), "points_max": max(counts),
            "value": n * args.steps / elapsed, "ms_per_step": elapsed * 1e3 / args.steps,
            "mode": "hipGraph replay" if graph_faster else "launched",
            "launched_value": n * args.steps / elapsed_launch,
            "launched_ms_per_step": elapsed_launch * 1e3 / args.steps,
            "graph_value": n * args.steps / elapsed_graph if elapsed_graph else None,
            "graph_ms_per_step": elapsed_graph * 1e3 / args.steps if elapsed_graph else None,
            "graph_captures": captures, "device_error": code, "autotuned": tuning}


def run_ppc(args, name: str) -> dict:
    """The 16M-point throughput workload at another point density (same
    grid, fewer cells): particle-updates/s and the fused kernel's HBM
    roofline fraction."""
    from dataclasses import replace
    from icempm_3dm_amd import scenes

    scene = replace(scenes.SCENES[name], dt=args.dt)
    hs = scenes.make_points(scene)
    n = hs.size
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup = min(args.steps, 200), min(args.warmup, 20)
    elapsed, timings, err, active_nodes, stats = run_steps(scene, hs, sub, args.dt, True)
    ms, cnt = timings.get("g2p2g", (0.0, 0))
    out = {"workload": f"{name}: {n} points ({scene.ppc_axis ** 3}/cell) on a 512x256x256 grid, dt={args.dt:g}",
           "particles": n, "steps": sub.steps, "value": n * sub.steps / elapsed,
           "ms_per_step": elapsed * 1e3 / sub.steps, "active_nodes": active_nodes, "device_error": err}
    if cnt:
        b = (B_G2P_PT + B_P2G_PT) * n + (B_G2P_NODE + B_P2G_NODE) * active_nodes
        gbs = b / (ms / cnt * 1e-3) / 1e9
        out.update({"k_g2p2g_avg_ms": ms / cnt, "k_g2p2g_alg_GBs": gbs, "k_g2p2g_frac": gbs / HBM_PEAK_GBS})
    return out


def run_damage(args) -> dict:
    """The 2M-point block with the damage model (256x128x128 grid, dt 5e-6,
    2000 substeps, the indenter force read back every 10 substeps):
    particle-updates/s including the read-backs, the force curve's peak and
    end, and the damage field at the end."""
    from icempm_3dm_amd import GPUImplementation6, capi, scenes

    sc = scenes.BLOCK_2M
    hs = scenes.make_points(sc)
    n = hs.size
    g = GPUImplementation6(0)
    g.hssoa = hs
    prms = scenes.make_params(sc)
    g.allocate_arrays(prms, hs.capacity)
    g.update_constants(prms)
    g.set_indenter(*sc.indenter())
    g.set_damage(capi.Damage.make(**scenes.BLOCK_2M_DAMAGE))
    g.set_sort_period(50)
    g.set_blocks_per_cu(16)
    g.transfer_ponts_to_device()
    forces = []
    t_sim = 0.0
    g.synchronize()
    t0 = time.perf_counter()
    for k in range(sc.steps // 10):
        g.reset_indenter_force_accumulator()
        t_sim = g.step(10, t_sim, 10 * k)
        forces.append(g.read_indenter_bins().reshape(-1, 3).sum(0) / 10)
    elapsed = time.perf_counter() - t0
    err = g.device_error_code()
    g.transfer_from_device()
    D = capi.damage_of(g.hssoa.utility()[:n])
    crushed = int(np.count_nonzero(g.hssoa.utility()[:n] & capi.STATUS_CRUSHED))
    g.close()
    fmag = np.linalg.norm(np.array(forces), axis=1)
    return {"workload": f"{sc.name}: {n} points on a 256x128x128 grid, dt={sc.dt:g}, {sc.steps} substeps, "
                        f"damage (softening strain {scenes.BLOCK_2M_DAMAGE['softening_strain']}, residual cohesion "
                        f"{scenes.BLOCK_2M_DAMAGE['residual_cohesion']}), force read back every 10 substeps",
            "value": n * sc.steps / elapsed, "ms_per_step": elapsed * 1e3 / sc.steps, "device_error": err,
            "force_peak_N": float(fmag.max()), "force_end_N": float(fmag[-1]), "crushed_points": crushed,
            "damaged_points": int(np.count_nonzero(D > 0)), "damage_max": float(D.max()),
            "fully_damaged_points": int(np.count_nonzero(D >= 1.0))}


def run_adjoint(args) -> dict:
    """Forward with checkpoints and backward of the calibration run, and the
    plain step on the same scene; times per substep."""
    from icempm_3dm_amd.calibration import CALIBRATION_SCENE, TARGET_THETA, ForceCurveModel

    n_steps, S, K = 400, 10, 20
    m = ForceCurveModel(CALIBRATION_SCENE, n_steps, S, K, theta=TARGET_THETA)
    w = np.ones((n_steps // S, 3))
    m.forward()
    m.backward(w)
    t0 = time.perf_counter()
    F = m.forward()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = m.backward(w)
    tb = time.perf_counter() - t0
    gpu = m.gpu
    gpu.set_sort_period(args.sort_period)
    gpu.set_fusion(args.fuse != 0)
    t = gpu.step(50, 0.0, 0)
    gpu.synchronize()
    t0 = time.perf_counter()
    t = gpu.step(n_steps, t, 50)
    gpu.synchronize()
    ts = time.perf_counter() - t0
    # the same substeps replayed from captured hipGraphs
    tg = None
    if args.graph:
        gpu.set_graph(True)
        t = gpu.step(50, t, 50 + n_steps)
        gpu.synchronize()
        t0 = time.perf_counter()
        gpu.step(n_steps, t, 100 + n_steps)
        gpu.synchronize()
        tg = time.perf_counter() - t0
    n = m.gpu.hssoa.size
    m.close()
    return {"workload": f"{CALIBRATION_SCENE.name}: {n} points on a 128x64x64 grid, {n_steps} substeps, dt=2e-6, "
                        f"force every {S}, checkpoints every {K}",
            "ratio": (tb / n_steps) / (ts / n_steps),
            "backward_ms_per_step": tb * 1e3 / n_steps, "forward_ms_per_step": tf * 1e3 / n_steps,
            "step_ms_per_step": ts * 1e3 / n_steps,
            "step_graph_ms_per_step": tg * 1e3 / n_steps if tg is not None else None,
            "dL_dE_ones": float(g[0]), "force_last": [float(x) for x in F[-1]]}


def copy_calibration(nbytes: int) -> int:
    """One device-to-device hipMemcpy of nbytes (HIP runtime of the library)."""
    import ctypes as C

    from icempm_3dm_amd import capi

    capi.lib()  # the library's HIP runtime is the one found by soname below
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), nbytes) or hip.hipMalloc(C.byref(dst), nbytes):
        print("copy calibration: hipMalloc failed", file=sys.stderr)
        return 1
    rc = hip.hipMemset(src, 1, nbytes) or hip.hipMemcpy(dst, src, nbytes, 3) or hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "rc": rc}))
    return 1 if rc else 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--scene", default="block_16m")
    ap.add_argument("--dt", type=float, default=2e-6)
    ap.add_argument("--unstable-dt", type=float, default=1e-5, help="0: skip the second run")
    ap.add_argument("--unstable-steps", type=int, default=500,
                    help="timed steps of the dt=1e-5 run (long enough for the points to crush and spread)")
    ap.add_argument("--fuse", type=int, default=1)
    ap.add_argument("--sort-period", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-max-steps", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host-core baseline threads (0: every CPU of this process's share)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble", type=int, default=64, help="scenes of the ensemble run; 0: skip it")
    ap.add_argument("--adjoint", type=int, default=1, help="0: skip the adjoint timing")
    ap.add_argument("--damage", type=int, default=1, help="0: skip the 2M-point damage run")
    ap.add_argument("--graph", type=int, default=1,
                    help="0: skip the hipGraph replays (rocprofv3's kernel trace does not follow stream capture)")
    ap.add_argument("--ppc-sweep", type=int, default=1, help="0: skip the 27 and 64 points-per-cell runs")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: launch settings from icempm_autotune (else --fuse / --sort-period / --blocks-per-cu)")
    ap.add_argument("--blocks-per-cu", type=int, default=8)
    ap.add_argument("--only", choices=("ensemble", "adjoint", "damage", "unstable"), default=None,
                    help="run only this part and print its result (not the bench line)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES between two device buffers with hipMemcpy and exit (the copy of known "
                         "size that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)

    if args.cpu_threads <= 0:
        args.cpu_threads = default_cpu_threads()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs the "
              "single-partition step on one device)", file=sys.stderr)
        return 2

    from dataclasses import replace
    from icempm_3dm_amd import capi, scenes

    if capi.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    if args.only == "ensemble":
        log(f"bench: ensemble of {args.ensemble} scenes: {args.warmup} warm-up + {args.steps} timed steps")
        print(json.dumps({"ensemble": run_ensemble(args)}), flush=True)
        return 0
    if args.only == "damage":
        log("bench: 2M-point block with damage (2000 substeps)")
        print(json.dumps({"damage": run_damage(args)}), flush=True)
        return 0
    if args.only == "adjoint":
        log("bench: calibration adjoint (200k points, 400 substeps)")
        print(json.dumps({"adjoint": run_adjoint(args)}), flush=True)
        return 0
    scene = replace(scenes.SCENES[args.scene], dt=args.dt)
    prms = scenes.make_params(scene)
    log(f"bench: generating {scene.name} ({scene.n_points} points)")
    hs = scenes.make_points(scene)
    n = hs.size
    if args.only == "unstable":
        print(json.dumps({"unstable_dt": run_unstable(scene, hs, args)}), flush=True)
        return 0

    log(f"bench: dt={args.dt}: {args.warmup} warm-up + {args.steps} timed steps")
    elapsed, timings, err, active_nodes, stats = run_steps(scene, hs, args, args.dt, True)
    if err != 0:
        print(f"bench: device error {err}", file=sys.stderr)

    ms_per_step = elapsed * 1e3 / args.steps
    value = n * args.steps / elapsed
    # the re-sort's share per step at the run's period, whether or not the
    # timed window happened to hold one
    period = stats["launch"]["sort_period"]
    sorts_in_window = timings.pop("sort_in_run", (0.0, 0))[1]
    sort_ms = timings["sort"][0] / max(timings["sort"][1], 1)
    owed = (args.steps / period if period > 0 else 0.0) - sorts_in_window
    value_amortized = n * args.steps / (elapsed + owed * sort_ms * 1e-3)
    node_bytes = {"p2g": B_P2G_NODE, "update": B_UPD_NODE, "g2p": B_G2P_NODE, "reset": B_RESET_NODE,
                  "g2p2g": B_G2P_NODE + B_P2G_NODE, "sort": 0}
    pt_bytes = {"p2g": B_P2G_PT, "update": 0, "g2p": B_G2P_PT, "reset": 0, "g2p2g": B_G2P_PT + B_P2G_PT,
                "sort": B_SORT_PT}
    kernel_of = {"p2g": "k_p2g", "update": "k_update_nodes", "g2p": "k_g2p", "reset": "k_reset_grid",
                 "g2p2g": "k_g2p2g", "sort": "sort kernels"}
    phases = {}
    for name, (ms, cnt) in timings.items():
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        b = pt_bytes[name] * n + node_bytes[name] * active_nodes
        phases[name] = {"kernel": kernel_of[name], "launches": cnt, "avg_ms": avg_ms, "alg_bytes": b,
                        "alg_GBs": b / (avg_ms * 1e-3) / 1e9}
    dominant = max((k for k in phases if k != "sort"), key=lambda k: phases[k]["avg_ms"] * phases[k]["launches"])
    d = phases[dominant]
    traffic = None
    if TRAFFIC_FILE.exists():
        tr = json.loads(TRAFFIC_FILE.read_text())
        key = f"{d['kernel']}@{scene.name}@dt={args.dt:g}"
        if key in tr:
            traffic = tr[key]["hbm_bytes_per_launch"]
    traffic_gbs = traffic / (d["avg_ms"] * 1e-3) / 1e9 if traffic else None
    step_bytes = (B_P2G_PT + B_G2P_PT) * n + (B_P2G_NODE + B_UPD_NODE + B_G2P_NODE + B_RESET_NODE) * active_nodes
    roofline = {
        "bound": "hbm",
        "kernel": d["kernel"],
        "achieved": round(d["alg_GBs"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(d["alg_GBs"] / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_GBs": round(traffic_gbs, 1) if traffic_gbs else None,
        "traffic_frac": round(traffic_gbs / HBM_PEAK_GBS, 4) if traffic_gbs else None,
        "alg_bytes_per_launch": d["alg_bytes"],
        "step_alg_GBs": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "step_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }

    unstable = run_unstable(scene, hs, args) if args.unstable_dt > 0 else None

    ppc = None
    if args.ppc_sweep:
        ppc = {}
        for name in ("block_16m_ppc27", "block_16m_ppc64"):
            log(f"bench: {name}")
            ppc[name] = run_ppc(args, name)

    damage = None
    if args.damage:
        log("bench: 2M-point block with damage (2000 substeps)")
        damage = run_damage(args)

    adjoint = None
    if args.adjoint:
        log("bench: calibration adjoint (200k points, 400 substeps)")
        adjoint = run_adjoint(args)

    ensemble = None
    if args.ensemble > 0:
        log(f"bench: ensemble of {args.ensemble} scenes: {args.warmup} warm-up + {args.steps} timed steps")
        ensemble = run_ensemble(args)
        ensemble["fraction_of_single_scene"] = ensemble["value"] / value

    cpu = None
    if not args.no_cpu:
        log(f"bench: host-core baseline ({args.cpu_threads} threads, <= {args.cpu_seconds} s)")
        cpu = cpu_baseline(scene, prms, hs, args.cpu_seconds, args.cpu_max_steps, args.cpu_threads)

    out = {
        "metric": "particle-updates/s and % of HBM roofline at 16M particles; adjoint/forward step-time ratio",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{scene.name}: {n} points ({scene.ppc_axis ** 3}/cell) on a "
                        f"{scene.grid[0]}x{scene.grid[1]}x{scene.grid[2]} grid, h=1/64, dt={args.dt:g}, "
                        f"sphere indenter, G2P+P2G fusion {'on' if stats['launch']['fuse'] else 'off'}, "
                        f"brick re-sort every {stats['launch']['sort_period']} steps, "
                        f"{stats['launch']['blocks_per_cu']} workgroups per CU",
            "particles": n,
            "grid": list(scene.grid),
            "dt": args.dt,
            "active_nodes": active_nodes,
            "brick_segments_at_end": stats["segments"],
            "launch": stats["launch"],
            "device_error": err,
        },
        "value_with_amortized_sort": value_amortized,
        "sorts_in_timed_window": sorts_in_window,
        "crushed_fraction": stats["crushed_fraction"],
        "roofline": roofline,
        "phases": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                   for k, v in phases.items()},
        "unstable_dt": unstable,
        "points_per_cell": ppc,
        "ensemble": ensemble,
        "adjoint": adjoint,
        "damage": damage,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0 if err == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
