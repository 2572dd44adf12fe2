#!/usr/bin/env python3
"""Benchmark: atom-steps/s of the 3D Lennard-Jones liquid at N = 4,194,304 on one MI355X.

Workload (the large single system): FCC lattice at density 0.8442 melted at T = 1.2, rc 2.5
sigma, Verlet skin 0.3 sigma, dt 0.005, integer positions/velocities, f32 pair forces rounded per
pair to integer kicks.  A "step" is one velocity-Verlet step of all N atoms, list rebuilds
included.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/sec (3D LJ N=4M) + batched systems/sec and p99 latency vs plain"
FLOP_PER_PAIR = 23  # real.rs:145-156 counted: r2 5, s2 1, s6 2, fr 6, u 4, a 1, kicks 3, pe 1
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
BYTES_PER_ATOM_STEP = 72  # read+write 3 x i64 positions and 3 x i32 velocities
# HBM bytes per k_step_half launch at N = 4,194,304 measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, each calibrated against a plain 1 GiB copy (profiles/ keeps the summary and the
# commands).  Keyed by (variant, block); a counter pass cannot run inside this process.
MEASURED_TRAFFIC = {
    (1, 2048): (1.577e9, "profiles/k_step_half_lj3d_n4194304_hbm_traffic.md"),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def ensure_built() -> None:
    from micrographia_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        from micrographia_amd import build

        build.build()


def make_liquid(n: int, density: float, t: float, seed: int):
    import numpy as np

    import micrographia_amd as rh

    sys_ = rh.System3.new(rh.cube_for(n, density), 0.005, rh.Pair.lj(2.5), [1.0])
    pos = rh.fcc(n, sys_.box_len, rh.fcc_cells(n, sys_.box_len), 0.05, seed)
    sp = np.zeros(n, np.uint32)
    vel = rh.maxwell(sys_, sp, t, seed)
    return sys_, rh.State3(pos, vel, sp)


def pairs_per_atom(density: float, rc: float = 2.5) -> float:
    import math

    return 4.0 / 3.0 * math.pi * rc ** 3 * density


def _cpu_3d(n: int, density: float, t: float, seed: int, prec: int, budget_s: float):
    """Whole steps of the oracle's IntEngine3<f32> (prec 32) or <f64> (prec 64) on the liquid,
    one thread, as many as fit the time budget (at least one)."""
    from oracle import oracle as O

    sys_, state = make_liquid(n, density, t, seed)
    osys = O.system3(sys_.box_len)
    pos, vel, sp = state.pos.copy(), state.vel.copy(), state.species.copy()
    steps, t0 = 0, time.perf_counter()
    while True:
        O.step3(osys, pos, vel, sp, 1, prec)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or el / steps * (steps + 1) > budget_s * 1.5:
            break
    return steps, el


def cpu_baseline(n: int, density: float, t: float, seed: int, budget_s: float = 10.0):
    """The oracle (C restatement of md-ref's IntEngine3), one thread, on this host, on the same
    liquid (density, temperature, cutoff) at n atoms, for both of md-ref's 3D precisions:
    IntEngine3<f32> (the reported value; the GPU engine's arithmetic) and IntEngine3<f64>
    beside it, each over as many whole steps as fit the time budget.  One step of the 4M-atom
    system alone takes ~50 s on one core."""
    from oracle import oracle as O

    if not os.path.exists(O.LIB):
        O.build()
    s32, e32 = _cpu_3d(n, density, t, seed, 32, budget_s)
    s64, e64 = _cpu_3d(n, density, t, seed, 64, budget_s)
    return {
        "value": n * s32 / e32,
        "unit": "atom-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{s32} step(s) of the same liquid at N={n} (rho {density}, T {t}, rc 2.5) with the oracle's IntEngine3<f32> (md-ref's cell-list loop, one thread), {e32:.1f} s",
        "ns_per_particle_step": e32 / (n * s32) * 1e9,
        "f64": {
            "value": n * s64 / e64,
            "unit": "atom-steps/s",
            "sample": f"{s64} step(s) of the same liquid with the oracle's IntEngine3<f64>, {e64:.1f} s",
            "ns_per_particle_step": e64 / (n * s64) * 1e9,
        },
        "host_cpu": _cpu_model(),
        "host_nproc": os.cpu_count(),
    }


def c1_fluid(seed: int):
    """The 2D LJ fluid of the reference's CPU configuration: N = 4096 on a 64 x 64 lattice in a
    periodic box at density 0.7, Maxwell T = 1.0, rc 2.5 shifted, dt 0.005 (init.rs:105-121)."""
    import micrographia_amd as rh

    sys_ = rh.System.new(rh.box_for(4096, 0.7, 1.0), 0.005, rh.Pair.lj(2.5), [1.0])
    return sys_, rh.fluid(sys_, 4096, 1.0, seed)


def c1_gpu_leg(seeds, steps: int = 1000):
    """The same 2D fluid on the 2D engine (Verlet lists, skin 0.4), 1000 steps per seed."""
    import torch

    import micrographia_amd as rh

    out = {}
    for seed in seeds:
        sys_, st = c1_fluid(seed)
        e = rh.GpuEngine(sys_, st)
        e.step(10)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.step(steps)
        torch.cuda.synchronize()
        out[str(seed)] = 4096 * steps / (time.perf_counter() - t0)
        e.close()
    return {"workload": f"2D LJ fluid, N=4096, rho 0.7, T 1.0, rc 2.5, dt 0.005, {steps} steps, Verlet skin 0.4", "atom_steps_per_s_by_seed": out}


def cpu_baseline_c1(seed: int, steps: int = 1000):
    """The oracle's IntEngine<f64> (md-ref's 2D engine with f64 forces), one thread, on the same
    2D fluid: the reference's CPU configuration as specified, on md-ref's default pair path
    (fast.rs, half shell and branch-free rows; int_engine.rs:89, 136-141), with the reference
    loop beside it."""
    from oracle import oracle as O

    if not os.path.exists(O.LIB):
        O.build()
    sys_, st = c1_fluid(seed)
    osys = O.system2(sys_.box_len)
    out = {}
    for path in ("fast", "loop"):
        pos, vel, sp = st.pos.copy(), st.vel.copy(), st.species.copy()
        t0 = time.perf_counter()
        if path == "fast":
            with O.fast_path():
                O.step2(osys, pos, vel, sp, steps, 64)
        else:
            O.step2(osys, pos, vel, sp, steps, 64)
        out[path] = (time.perf_counter() - t0, O.hash2(pos, vel, sp))
    el, h = out["fast"]
    assert h == out["loop"][1], "the fast pair path and the reference loop must give the same trajectory"
    return {
        "value": 4096 * steps / el,
        "unit": "atom-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} steps of the 2D fluid (N=4096, rho 0.7, T 1.0, seed {seed}) with the oracle's IntEngine<f64> on md-ref's fast pair path, {el:.1f} s",
        "ns_per_particle_step": el / (4096 * steps) * 1e9,
        "reference_loop": {"value": 4096 * steps / out["loop"][0], "unit": "atom-steps/s", "ns_per_particle_step": out["loop"][0] / (4096 * steps) * 1e9},
    }


def batched_leg(count: int, steps: int, chunk: int, meta_seed: int, plain: bool, table: str = None, default_tiles: bool = True):
    """The batched ensemble: `count` independent LJ systems (2D/3D 50/50, N log-uniform in
    [256, 65536], density U[0.3, 0.9], T U[0.5, 2.0], own seeds) advanced together by the batched
    engine in hipGraph chunks of `chunk` steps, against the same systems stepped one at a time by
    the single-system engines (one system per launch).  `table`: a tuning file to read (buckets
    found there are not timed again) and write; `default_tiles`: also time the untuned tiles."""
    import torch

    import micrographia_amd as rh
    from micrographia_amd.batch import BatchEngine
    from micrographia_amd.workloads import ensemble

    from micrographia_amd.autotune import BatchTuner

    specs = ensemble(count, meta_seed=meta_seed)
    t0 = time.perf_counter()
    built = [s.build() for s in specs]
    atoms = sum(st.n() for _, st in built)
    log(f"[bench] batched: {count} systems, {atoms} atoms, host init {time.perf_counter() - t0:.1f} s")
    # the autotuner: per (dimension, size class), list variant x cell occupancy, timed on up to 8
    # of the ensemble's own systems per bucket
    tuner = BatchTuner(path=table, per_bucket=8)
    tuner.tune(built)
    log(f"[bench] batched: tuned {len(tuner.table)} buckets in {tuner.seconds:.1f} s")

    def run_ensemble(tuned: bool):
        b = BatchEngine(budget_bytes=0, max_systems=max(count, 1), graph_chunk=chunk)
        if tuned:
            tuner.install(b)
        need = sum(b.bytes_needed(sys_, st.n()) for sys_, st in built)
        ids = [b.add(sys_, st, steps + chunk) for sys_, st in built]
        assert all(i >= 0 for i in ids), "the ensemble does not fit the arena"
        b.step_all(chunk)  # warm-up chunk: first lists, graph capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.step_all(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        left = b.poll()
        assert all(v == 0 for v in left.values())
        # list rebuilds over all steps run (warm-up chunk included), weighted by system size
        rebuilt = {3: 0, 2: 0}
        for i, (sys_, st) in zip(ids, built):
            rebuilt[len(sys_.box_len)] += b.list_stats(i).rebuilds * st.n()
        b.close()
        return el, need, rebuilt

    el_default = run_ensemble(False)[0] if default_tiles else None
    el, need, rebuilt = run_ensemble(True)
    atoms_by_dim = {d: sum(st.n() for sys_, st in built if len(sys_.box_len) == d) for d in (2, 3)}
    out = {
        "workload": f"{count} independent LJ systems (2D/3D 50/50, N log-uniform [256, 65536], rho U[0.3,0.9], T U[0.5,2.0], rc 2.5, skin 0.3, dt 0.005), meta-seed {meta_seed}, hipGraph chunks of {chunk} steps",
        "systems": count,
        "atoms": atoms,
        "steps": steps,
        "seconds": el,
        "atom_steps_per_s": atoms * steps / el,
        "system_steps_per_s": count * steps / el,
        "systems_per_s": count * steps / chunk / el,
        "systems_per_s_unit": f"systems advanced by one {chunk}-step chunk per second",
        "arena_bytes_needed": need,
        "rebuilt_fraction_per_step": {f"{d}d": rebuilt[d] / max(atoms_by_dim[d], 1) / (steps + chunk) for d in (2, 3)},
        "autotuned": {
            "table": {k: {"variant": v[0], "cell_factor": v[1], "block": v[2] if len(v) > 2 else 256} for k, v in sorted(tuner.table.items())},
            "tuning_seconds": tuner.seconds,
            "atom_steps_per_s_default_tiles": atoms * steps / el_default if el_default else None,
        },
    }
    if plain:
        # the same systems, one engine at a time: stepping timed, engine creation not
        tot = 0.0
        for sys_, st in built:
            if isinstance(sys_, rh.System3):
                e = rh.GpuEngine3(sys_, st, rh.verlet3_with_skin(sys_, st.n(), 0.3))
            else:
                e = rh.GpuEngine(sys_, st, rh.verlet_with_skin(sys_, st.n(), 0.3))
            e.step(chunk)  # same warm-up as the batch
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            e.step(steps)
            torch.cuda.synchronize()
            tot += time.perf_counter() - t1
            e.close()
        out["plain_seconds"] = tot
        out["plain_atom_steps_per_s"] = atoms * steps / tot
        out["speedup_vs_plain"] = tot / el
    log(f"[bench] batched: {out['atom_steps_per_s']:.4g} atom-steps/s" + (f", {out['speedup_vs_plain']:.1f}x plain" if plain else ""))
    return out


def serving_leg(rate: float, duration: float, meta_seed: int, plain: bool):
    """The serving loop: Poisson arrivals at `rate`/s over `duration` s of LJ requests (N
    log-uniform in [1k, 1M], 2D/3D 50/50, steps U[100, 10k], density U[0.3, 0.9], T U[0.5, 2.0],
    own seeds), admitted against the device-memory budget, continuously batched."""
    import torch

    from micrographia_amd.serve import Server, serve_plain
    from micrographia_amd.workloads import request_stream

    reqs = request_stream(rate, duration, meta_seed=meta_seed)
    free, total = torch.cuda.mem_get_info()
    budget = int(free * 0.85)
    log(f"[bench] serving: {len(reqs)} requests, budget {budget / 2**30:.1f} GiB")
    srv = Server(budget_bytes=budget, max_systems=4096, chunk_s=0.02)
    rep = srv.serve(reqs, realtime=True)
    srv.close()
    offered = sum(r.n * r.steps for r in reqs) / duration if duration > 0 else 0.0
    out = {"workload": f"Poisson {rate}/s over {duration} s (meta-seed {meta_seed}): N log-uniform [1k, 1M], 2D/3D 50/50, steps U[100, 10k], rho U[0.3, 0.9], T U[0.5, 2.0]; admission budget {budget} B of device memory",
           **rep.summary(),
           # the work arriving per second of the arrival window, against what the GPU served
           "offered_atom_steps_per_s": offered,
           "offered_over_served": offered / rep.summary()["atom_steps_per_s"] if rep.completed else None}
    if plain:
        p = serve_plain(reqs, realtime=True)
        out["plain"] = p.summary()
        for c, d in zip(sorted(rep.completed, key=lambda c: c.request.index), sorted(p.completed, key=lambda c: c.request.index)):
            assert c.state_hash == d.state_hash, f"request {c.request.index} differs between the server and the plain engine"
        out["identical_to_plain"] = True
    log(f"[bench] serving: {out['atom_steps_per_s']:.4g} atom-steps/s, p50 {out['p50_latency_s']:.3f} s, p99 {out['p99_latency_s']:.3f} s")
    return out


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--density", type=float, default=0.8442)
    ap.add_argument("--temperature", type=float, default=1.2)
    ap.add_argument("--skin", type=float, default=0.3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--cpu-n", type=int, default=262144, help="atoms in the CPU-baseline sample")
    ap.add_argument("--variant", type=int, default=1, help="0 full Verlet list, 1 half list in blocks")
    ap.add_argument("--block", type=int, default=2048, help="half-list block size (256, 512, 1024, 2048)")
    ap.add_argument("--batch-systems", type=int, default=512, help="systems in the batched leg (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-table", default=None, help="batched tuning table to read and write (JSON)")
    ap.add_argument("--no-batch-default", dest="batch_default", action="store_false", help="skip the batched run with untuned tiles")
    ap.add_argument("--serve-duration", type=float, default=1.15, help="seconds of Poisson arrivals in the serving leg (0: skip); 1.15 s at 200/s gives 202 requests, enough for a p99 below the maximum")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--no-plain", action="store_true", help="skip the one-system-at-a-time comparisons")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="also run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--seeds", default="2,3", help="further seeds of the headline liquid to time (comma list; empty: none)")
    ap.add_argument("--no-c1", dest="c1", action="store_false", help="skip the 2D N=4096 fluid (GPU and CPU)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the engine shards "
              "independent systems across GPUs by construction and is measured per GPU.", file=sys.stderr)
        return 2

    import torch  # before the engine library: one HIP runtime per process

    ensure_built()
    import micrographia_amd as rh

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    log(f"[bench] building N={args.n} FCC liquid, rho={args.density}, T={args.temperature}")
    t0 = time.perf_counter()
    sys_, state = make_liquid(args.n, args.density, args.temperature, args.seed)
    log(f"[bench] init {time.perf_counter() - t0:.1f} s; creating engine")
    mode = rh.verlet3_with_skin(sys_, args.n, args.skin)
    eng = rh.GpuEngine3(sys_, state, mode)
    eng.tune(args.variant, args.block)
    log(f"[bench] warmup {args.warmup} steps (variant {args.variant}, block {args.block})")
    eng.step(args.warmup)
    torch.cuda.synchronize()
    s0 = eng.read_list_stats()
    eng.set_timing(True)
    f0 = eng.kernel_ms()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0

    f1 = eng.kernel_ms()
    s1 = eng.read_list_stats()
    eng.set_timing(False)
    n = args.n
    value = n * args.steps / el
    log(f"[bench] {args.steps} steps in {el:.3f} s -> {value:.4g} atom-steps/s")

    force_ms = f1[0] - f0[0]
    launches = f1[1] - f0[1]
    rebuild_ms = f1[2] - f0[2]
    rebuilds = s1.rebuilds - s0.rebuilds
    executed = s1.steps_done - s0.steps_done
    skipped = s1.launches_skipped - s0.launches_skipped
    per_step_ms = force_ms / max(executed, 1)  # executed launches carry all but microseconds
    flop = n * pairs_per_atom(args.density) * FLOP_PER_PAIR
    achieved_tflops = flop / (per_step_ms * 1e-3) / 1e12
    hbm_gbs = n * BYTES_PER_ATOM_STEP / (per_step_ms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded FCC lattice + Maxwell-Boltzmann velocities, generated on the host)",
        "config": {
            "workload": f"3D LJ liquid, N={n}, FCC rho={args.density}, T={args.temperature}, rc=2.5, skin={args.skin}, dt=0.005, Verlet lists rebuilt on the device",
            "n_atoms": n,
            "density": args.density,
            "temperature": args.temperature,
            "rc": 2.5,
            "skin": args.skin,
            "dt": 0.005,
            "kernel_variant": args.variant,
            "block": args.block,
        },
        "roofline": {
            "bound": "valu",
            "kernel": ("k_step_half (half list in blocks of %d, LDS partner kicks; leapfrog; skin check)" % args.block)
            if args.variant == 1 else "k_step_list (full list; leapfrog; skin check)",
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "traffic": MEASURED_TRAFFIC.get((args.variant, args.block), (None,))[0] if n == 4194304 else None,
            "traffic_unit": "bytes per launch (HBM fetch + write)",
            "traffic_source": MEASURED_TRAFFIC.get((args.variant, args.block), (None, None))[1] if n == 4194304 else None,
            "flop_per_atom_step": pairs_per_atom(args.density) * FLOP_PER_PAIR,
            "avg_launch_ms": per_step_ms,
            "avg_launch_ms_all_launches": force_ms / max(launches, 1),
            "launches": launches,
            "skipped_launches": skipped,
            "hbm_algorithmic_gbs": hbm_gbs,
            "hbm_frac": hbm_gbs / PEAK_HBM_GBS,
        },
        "rebuilds": rebuilds,
        "rebuild_ms_per_rebuild": rebuild_ms / max(rebuilds, 1),
        "pair_interactions_per_s": value * pairs_per_atom(args.density),
    }
    eng.close()
    del eng
    extra_seeds = [int(x) for x in args.seeds.split(",") if x.strip()] if args.seeds else []
    if extra_seeds:
        by_seed = {str(args.seed): value}
        for sd in extra_seeds:
            log(f"[bench] seed {sd}")
            s2, st2 = make_liquid(args.n, args.density, args.temperature, sd)
            e2 = rh.GpuEngine3(s2, st2, rh.verlet3_with_skin(s2, args.n, args.skin))
            e2.tune(args.variant, args.block)
            e2.step(args.warmup)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e2.step(args.steps)
            torch.cuda.synchronize()
            by_seed[str(sd)] = args.n * args.steps / (time.perf_counter() - t0)
            e2.close()
            del e2
        out["value_by_seed"] = by_seed
    if args.c1:
        log("[bench] 2D fluid, N=4096")
        out["c1"] = c1_gpu_leg([1, 2, 3])
    if args.calibrate_copy:
        x = torch.empty(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = args.calibrate_copy << 20
    if args.batch_systems:
        out["batched"] = batched_leg(args.batch_systems, args.batch_steps, 100, 2026, not args.no_plain, args.batch_table, args.batch_default)
    if args.serve_duration > 0:
        out["serving"] = serving_leg(args.serve_rate, args.serve_duration, 7, not args.no_plain)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle, one thread)")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.density, args.temperature, args.seed, args.cpu_budget)
        if args.c1:
            out["cpu_baseline"]["c1"] = cpu_baseline_c1(1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
