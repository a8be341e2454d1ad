"""Benchmark: chromo MC move-evaluations per second on one MI355X, batched vs plain.

Headline workload (``value``): 512 independent replicas advanced together by
the persistent kernel, chain lengths N ~ U{2000..50000}, 32^3..64^3 voxels of
30 nm, per-replica chi ~ U[0,2], eps ~ U[-5,-1], mu ~ U[-2,1], Bernoulli(0.3)
marks, Philox streams.  One *step* is one mc_sim sweep of every replica
(all_moves: 161 proposals per replica).

The chains are first run ``--equilibrate`` sweeps (default 2,500) so that the
move controllers reach their steady-state amplitudes: the controllers raise the
crank-shaft, slide and end-pivot windows by one bead every ~19 sweeps until they
reach their 150-bead bound (~2,200 sweeps), and the mean move grows from ~3.7
beads over the first 200 sweeps to ~20.6.  The timed window is the sweep range
reported in ``config.sweeps_timed``; a sweep there costs ~1.8x one of the first
400 sweeps, so the first sweeps would overstate the rate by that much.  After the warmup, chromo_amd.autotune times every
pair (wavefronts per replica, sweeps per graph launch) and keeps the fastest;
all settings produce the same chains.  The timed steps replay a captured HIP
graph of that many sweeps (``--no-graph``: direct launches).

Also reported: the plain path (one kernel launch per proposal) on the
100,000-bead / 64^3 configuration, the ratio of the two, a dE / decision check
of four replicas against the CPU oracle, near-tie counters of the timed window,
the roofline of the dominant kernel (algorithmic HBM bytes per launch over its
HIP-event-timed duration), and the CPU oracle on the host cores: the same 512
replicas from the same state over the same sweep window (lean and
reference-faithful builds), and the 1,000-bead / 30^3 configuration.

``--adjoint`` instead times the energy + adjoint evaluation (E, dE/dr, dE/dtheta)
on the 20,000-bead and 10^6-bead configurations and prints its own JSON line.
"""
from __future__ import annotations

import argparse
import copy
import dataclasses
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MC move-evals/s on 1 GPU, batched vs plain; dE err <=1e-9 kT; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_replicas(systems, rng_mode, seeds, streams, log_cap=0):
    from chromo_amd.engine import DeviceReplicas
    return DeviceReplicas([s.chain for s in systems], [s.field for s in systems],
                          [copy.deepcopy(s.ctrls) for s in systems], rng_mode=rng_mode,
                          seeds=seeds, streams=streams, log_cap=log_cap)


def timed(fn, n):
    import torch
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(n):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1000.0


def _divisors(k, cands=(1, 2, 4, 5, 10, 20, 25, 40, 50)):
    return [c for c in cands if c <= k and k % c == 0]


def bench_batched(args):
    from chromo_amd import _lib as L
    from chromo_amd import workloads as W
    t0 = time.time()
    systems = W.config_batched(args.replicas, seed=3)
    dr = build_replicas(systems, L.RNG_PHILOX, [20240601] * len(systems),
                        list(range(len(systems))))
    log(f"batched: {len(systems)} replicas built in {time.time() - t0:.1f}s "
        f"(beads {sum(s.chain.N for s in systems)})")
    done = 0
    # steady state: the controllers raise the move amplitudes over the first few
    # hundred sweeps; run them as part of preparing the workload
    while done < args.equilibrate:
        k = min(500, args.equilibrate - done)
        dr.run(k, sync=False, waves=2)
        done += k
    dr.raise_on_error()
    for _ in range(args.warmup):
        dr.run(1, sync=False, waves=2)
    done += args.warmup
    dr.raise_on_error()
    if args.waves == "auto":
        from chromo_amd.autotune import autotune
        spl = _divisors(args.steps) if args.sweeps_per_launch == "auto" else [int(args.sweeps_per_launch)]
        c_before = int(dr.counter_array()[:, L.CNT_PROPOSALS].sum())
        tune = autotune(dr, sweeps=20, sweeps_per_launch=spl)
        # every candidate's graph ran sweeps (ordinary sweeps of the same chains)
        done += sum(s_ + max(1, 20 // s_) * s_ for (_, s_) in tune)
        log("autotune (ms/sweep): " + ", ".join(f"W={w} S={s_}: {t * 1e3:.3f}"
                                               for (w, s_), t in tune.items())
            + f" -> W={dr.waves}, S={dr.sweeps_per_launch}")
    else:
        tune = None
        dr.waves = int(args.waves)
        dr.sweeps_per_launch = 1 if args.sweeps_per_launch == "auto" else int(args.sweeps_per_launch)
    spl = dr.sweeps_per_launch
    if args.steps % spl:
        raise SystemExit(f"--steps {args.steps} is not a multiple of {spl} sweeps per launch")
    if args.no_graph:
        g = None
        step = lambda: dr.run(spl, sync=False)
    else:
        g = dr.capture(spl)             # spl sweeps of every replica per graph launch
        step = g.replay
    torch_sync()
    snap = snapshot(dr, systems) if not args.no_cpu_baseline else None
    c0 = dr.counter_array()
    secs = timed(step, args.steps // spl)
    dr.raise_on_error()
    c1 = dr.counter_array()
    d = c1 - c0
    props = int(d[:, L.CNT_PROPOSALS].sum())
    nbytes = int(d[:, L.CNT_BYTES].sum())
    return dict(secs=secs, proposals=props, bytes=nbytes, accepted=int(d[:, L.CNT_ACCEPTED].sum()),
                near_ties={"metropolis": int(d[:, L.CNT_NEAR_TIES].sum()),
                           "voxel_floor": int(d[:, L.CNT_TIE_VOXEL].sum()),
                           "over_vf_limit_voxels": int(d[:, L.CNT_TIE_HUGE].sum())},
                replicas=len(systems), beads=int(sum(s.chain.N for s in systems)), systems=systems,
                waves=dr.waves, spl=spl, tune=tune, first_sweep=done, snap=snap,
                launches=args.steps // spl, mean_n=float(d[:, L.CNT_SUM_N].sum()) / max(1, props))


def torch_sync():
    import torch
    torch.cuda.synchronize()


def snapshot(dr, systems):
    """Every replica's chain, binder states and controllers at the start of the
    timed window (host copies), for the CPU leg."""
    out = []
    for i, s in enumerate(systems):
        r, t3, t2 = dr.positions(i)
        out.append((dataclasses.replace(s.chain, r=r, t3=t3, t2=t2, states=dr.binder_states(i)),
                    dr.controllers(i)))
    return out


def bench_plain(args):
    from chromo_amd import _lib as L
    from chromo_amd import workloads as W
    s = W.config_plain_gpu(seed=2)
    dr = build_replicas([s], L.RNG_GLIBC, [1], None)
    dr.run_plain(1)
    c0 = dr.counter_array()
    secs = timed(lambda: dr.run_plain(1, sync=False), args.plain_steps)
    dr.raise_on_error()
    c1 = dr.counter_array()
    d = (c1 - c0)[0]
    return dict(secs=secs, proposals=int(d[L.CNT_PROPOSALS]), launches=args.plain_steps * (161 + 5),
                bytes=int(d[L.CNT_BYTES]))


def dE_check(systems, sweeps=20, n=4, waves=1):
    """GPU (launch shape ``waves``) vs oracle on the first n replicas of the batched workload."""
    from chromo_amd import _lib as L
    from oracle import oracle as O
    worst, mism, total = 0.0, 0, 0
    for i in range(n):
        s = systems[i]
        o = O.OracleSim(s.chain, s.field, rng_mode=O.RNG_PHILOX, seed=20240601, stream=i)
        lo = o.mc_sim(sweeps, copy.deepcopy(s.ctrls), max_logs=10 ** 6)
        dr = build_replicas([s], L.RNG_PHILOX, [20240601], [i], log_cap=len(lo) + 16)
        dr.run(sweeps, waves=waves)
        lg = dr.logs(0)
        if len(lg) != len(lo):
            mism += abs(len(lg) - len(lo))
        k = min(len(lg), len(lo))
        total += k
        mism += int(np.sum(lg["accepted"][:k] != lo["accepted"][:k]))
        dg = lg["dE_poly"][:k] + lg["dE_field"][:k]
        do = lo["dE_poly"][:k] + lo["dE_field"][:k]
        ok = np.abs(do) < 1e50
        if ok.any():
            worst = max(worst, float(np.max(np.abs(dg[ok] - do[ok]))))
    return dict(proposals=total, max_abs_dE_err=worst, decision_mismatches=mism)


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """The host cores this job may use: the affinity set, capped by OMP_NUM_THREADS
    where the machine sets it to the share of cores that goes with one GPU."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(b, steps, config1_sweeps=1000):
    """The C oracle on the host cores, one replica per thread at a time:

    * the same 512 replicas, from the same chains, binder states and controller
      amplitudes as the GPU's timed window, over the same number of sweeps, in
      the lean build (affected voxels only) and the reference-faithful build
      (the reference's O(n_bins) reset and accept scans per field move);
    * the 1,000-bead / 30^3 configuration, one replica per thread.
    """
    from chromo_amd import workloads as W
    from oracle import oracle as O
    threads = cpu_threads()
    snap, systems = b["snap"], b["systems"]
    out = {}
    for kind, faithful in (("lean", False), ("faithful", True)):
        t0 = time.time()
        sims = [O.OracleSim(ch, s.field, rng_mode=O.RNG_PHILOX, seed=20240601, stream=i, faithful=faithful)
                for i, ((ch, _), s) in enumerate(zip(snap, systems))]
        ctrls = [copy.deepcopy(c) for _, c in snap]
        log(f"cpu {kind}: {len(sims)} oracle replicas built in {time.time() - t0:.1f}s")
        p, sec = O.run_batch(sims, ctrls, steps, threads)
        out[kind] = dict(value=p / sec, proposals=p, seconds=sec)
        log(f"cpu {kind}: {p} proposals in {sec:.2f}s -> {p / sec:.4g}/s")
        del sims
    s1 = W.config_plain_cpu()
    c1 = {}
    for kind, faithful in (("lean", False), ("faithful", True)):
        sweeps = config1_sweeps if not faithful else max(1, config1_sweeps // 5)
        sims = [O.OracleSim(s1.chain, s1.field, rng_mode=O.RNG_PHILOX, seed=1, stream=i, faithful=faithful)
                for i in range(threads)]
        ctrls = [copy.deepcopy(s1.ctrls) for _ in range(threads)]
        p, sec = O.run_batch(sims, ctrls, sweeps, threads)
        c1[kind] = dict(value=p / sec, per_core=p / sec / threads, sweeps=sweeps, proposals=p)
        log(f"cpu config 1 {kind}: {p / sec:.4g}/s on {threads} threads")
    first = b["first_sweep"]
    return dict(value=out["lean"]["value"], unit="move-evals/s", cores=threads, kind="port",
                cpu_model=cpu_model(),
                sample=(f"all {len(snap)} batched replicas from the GPU's state at the start of its "
                        f"timed window, sweeps {first}..{first + steps} ({out['lean']['proposals']} "
                        f"proposals in {out['lean']['seconds']:.2f} s), oracle lean build, one replica "
                        f"per thread at a time on {threads} threads"),
                faithful=out["faithful"], config1={"workload": "1,000 beads, 30^3 voxels, B=2", **c1})


def bench_adjoint(args):
    """Energy + adjoint (E, dE/dr, dE/dtheta through chromo_energy_adjoint) per second
    on the 20,000-bead / B = 3 / 48^3 configuration and on one 10^6-bead / B = 4 /
    128^3 chain, inputs resident in HBM.  Algorithmic bytes per evaluation:
    N (48 + 2B) (positions, states, marks) + 2 * 8 (B+1) V (density write and read)
    + 24 N (gradient write)."""
    import torch
    from chromo_amd import workloads as W
    from chromo_amd.energy import EnergyModel, energy
    out = []
    for name, s in (("20,000 beads, B=3, 48^3 voxels", W.config_adjoint()),
                    ("1,000,000 beads, B=4, 128^3 voxels", W.config_large(1)[0])):
        model = EnergyModel(s.chain, s.field, s.binders)
        r = torch.tensor(s.chain.r, dtype=torch.float64, device="cuda")
        th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
        for _ in range(max(1, args.warmup)):
            energy(model, r, th, grad=True)
        k = args.adjoint_steps
        secs = timed(lambda: energy(model, r, th, grad=True), k)
        N, B = s.chain.N, s.chain.B
        V = s.field.nx * s.field.ny * s.field.nz
        nbytes = N * (48 + 2 * B) + 2 * 8 * (B + 1) * V + 24 * N
        gbs = nbytes * k / secs / 1e9
        out.append({"workload": name, "N": N, "B": B, "voxels": V, "evaluations": k,
                    "ms_per_evaluation": 1e3 * secs / k, "evaluations_per_s": k / secs,
                    "algorithmic_bytes_per_evaluation": nbytes, "achieved_GBs": gbs,
                    "frac_of_hbm_peak": gbs / HBM_PEAK_GBS})
        log(f"adjoint {name}: {1e3 * secs / k:.4f} ms per evaluation, {gbs:.1f} GB/s")
        del model, r, th
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each sweep directly instead of replaying a captured HIP graph")
    ap.add_argument("--waves", default="auto",
                    help="wavefronts per replica (1 or 2) or 'auto' to time each and keep the fastest")
    ap.add_argument("--sweeps-per-launch", default="auto",
                    help="sweeps per graph launch, or 'auto' (with --waves auto: time the divisors of "
                         "--steps among 1, 2, 4, 5, 10, 20, ...)")
    ap.add_argument("--equilibrate", type=int, default=2500,
                    help="sweeps run before the warmup so that the move amplitudes are at steady state "
                         "(crank-shaft, slide and end-pivot windows reach their 150-bead bound after "
                         "~2,200 sweeps; the mean move then holds ~20.6 beads)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 256 MiB device to device 5 times (calibrates rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count)")
    ap.add_argument("--adjoint", action="store_true",
                    help="only time the energy + adjoint evaluation (its own JSON line)")
    ap.add_argument("--adjoint-steps", type=int, default=50)
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 256 * 1024 * 1024 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        c = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            c.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes_per_launch": n * 8, "launches": 5}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: replicas are independent chains; multi-GPU runs "
              "would shard replicas with no exchange and are not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "needs a ROCm GPU"
    if args.adjoint:
        print(json.dumps({"metric": "energy + adjoint evaluations/s (E, dE/dr, dE/dtheta), FP64",
                          "unit": "evaluations/s", "n_gpus": 1, "dtype": "f64", "data": "synthetic",
                          "higher_is_better": True, "hbm_peak_GBs": HBM_PEAK_GBS,
                          "workloads": bench_adjoint(args)}), flush=True)
        return

    b = bench_batched(args)
    value = b["proposals"] / b["secs"]
    log(f"batched: {b['proposals']} proposals in {b['secs']:.3f}s -> {value:.4g}/s")
    p = bench_plain(args)
    plain_value = p["proposals"] / p["secs"]
    log(f"plain: {p['proposals']} proposals in {p['secs']:.3f}s -> {plain_value:.4g}/s")
    chk = None if args.no_check else dE_check(b["systems"], waves=b["waves"])
    if chk:
        log(f"check: {chk}")
    cpu = None if args.no_cpu_baseline else cpu_baseline(b, args.steps)
    if cpu:
        log(f"cpu: {cpu}")

    achieved = b["bytes"] / b["secs"] / 1e9
    traffic = None
    kname = "k_mc_run" if b["waves"] == 1 else f"k_mc_spec_w{b['waves']}"
    prof = ROOT / "profiles" / f"traffic_{kname}.json"
    if prof.exists():
        try:
            # measured on this workload's steady-state timed window, per sweep; a
            # launch of S sweeps moves S times that
            tj = json.loads(prof.read_text())
            traffic = float(tj["hbm_bytes_per_sweep"]) * b["spl"]
        except Exception:
            traffic = None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "move-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * b["secs"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {b['replicas']} replicas, N~U{{2000..50000}} "
                        f"({b['beads']} beads), 32^3-64^3 voxels of 30 nm, B=2, all_moves + "
                        "SimpleControl, Philox; one step = one sweep (161 proposals per replica)",
            "replicas": b["replicas"],
            "sweeps_timed": [b["first_sweep"], b["first_sweep"] + args.steps],
            "mean_beads_per_move": b["mean_n"],
            "waves_per_replica": b["waves"],
            "sweeps_per_launch": b["spl"],
            "parallelism": (f"replicas, one workgroup each; {b['waves']} wavefront(s) per replica "
                            "evaluating proposals of a controller block at once; "
                            f"{b['spl']} sweep(s) per graph launch"),
            "autotune_ms_per_sweep": ({f"W={w},S={s_}": t * 1e3 for (w, s_), t in b["tune"].items()}
                                      if b["tune"] else None),
        },
        "plain": {"value": plain_value, "unit": "move-evals/s",
                  "workload": "100,000 beads, 64^3 voxels, B=2, one kernel launch per proposal",
                  "proposals": p["proposals"], "seconds": p["secs"]},
        "batched_over_plain": value / plain_value,
        "acceptance": b["accepted"] / max(1, b["proposals"]),
        "near_ties": b["near_ties"],
        "dE_check": chk,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_mc_run" if b["waves"] == 1 else f"k_mc_spec<{b['waves']}, 2>",
                     "algorithmic_bytes_per_launch": b["bytes"] / b["launches"],
                     "launch_ms": 1000.0 * b["secs"] / b["launches"]},
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        "gpu_over_cpu_faithful": (value / cpu["faithful"]["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
