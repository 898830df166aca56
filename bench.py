"""Throughput benchmark of the Metropolis + parallel-tempering hot path.

Workload (one GPU): 64^3 periodic simple-cubic lattice, 64 replicas on a
geometric temperature ladder 0.2..2.0 J, J = 1, bulk D = 0.5 J, Bz = 0.2 J,
fixed cone radius 0.5, seeded random-sphere start (synthetic data).
One step = 10 exact checkerboard sweeps of all 64 replicas followed by one
replica-exchange round; E and M are recorded every 2 sweeps and the
topological charge of the middle layer every 2, all on device inside one
replayed hipGraph.

Before the timed runs (untimed): 2000 plain sweeps collect training records,
the cone-width network is trained on them, 4000 more sweeps equilibrate and
the state is checkpointed on the device. Then the same schedule is timed
twice from that checkpoint, with the same random-number counters: fixed cone
r = 0.5, and learned cone.

value  = site updates/s of the fixed-cone run (every replica's sites x sweeps
/ wall time); also reported per sampler: tau_int-corrected independent
samples/s of E and of Q (middle layer) and their learned/fixed ratio, the
plain sweep kernel's HBM roofline fraction (the fused sweep when the
autotuner keeps a fused flag set, else the half sweep; the learned sampler
always runs half sweeps) from hipEvent pairs around its launches inside the
same schedule (ising_time_pt_sweeps, the second of two 200-sweep windows
after each timed run), and the FP64 CPU oracle's rate on this host's cores.

The tau_int half of the metric comes from a record window of its own,
10,000 sweeps per sampler with the same schedule (timed: samples/s uses
that run's wall time), so it does not depend on --steps. Sokal's window is
trusted only for replicas whose series spans at least 50 tau_int; the
learned/fixed ratio sums replicas converged in both runs, with a 1-sigma
bootstrap band over those replicas, and the unconverged ones are counted.
200 untimed sweeps and the --warmup steps follow the window before the
timed steps (the device's clocks drop while the window's records are
copied out).

`--workload c3` measures the batched uneven configuration instead: 256
replicas = 8 field values Bz = 0, 0.05, .., 0.35 x 32 temperatures (0.2..2.0
geometric), lattices 24^3, 32^3, 48^3 by field (b mod 3), J = 1, bulk D = 0.5,
all in one launch per half sweep; both autotuners run (launch shape, then the
sweeps-per-swap interval), then 2000 sweeps are timed with E/M every 10.

`--workload bcc` and `--workload hk` measure the reference's two other
lattices under the same PT schedule shape (64 replicas, T 0.2..2.0
geometric, J = 1, bulk D = 0.5, Bz = 0.2, swap every 10 sweeps, E/M every
10): body-centred cubic 50^3 cells (250,000 sites per replica, the cell of
the reference's PT driver; two colour launches per sweep) and hyper-kagome
25^3 cells (187,500 sites per replica, the cell of its ISING driver; three
colour launches per sweep, table-driven neighbours); 2000 timed sweeps,
in-schedule kernel timing, and the FP64 CPU oracle on one replica of the
same cell. With `--profile-sweeps N` they run only N plain sweeps.

`--workload demag` measures Metropolis_mcs with the demagnetising term on:
64 replicas of 64^3 (T 0.2..2.0 geometric, J = 1, bulk D = 0.5, Bz = 0.2),
the field of every replica recomputed before the first sweep and after every
7th (the reference's int(pi sqrt 5 / (2 r)) at r = 0.5); 140 timed sweeps in
one ising_sweep call. The roofline line is the field recomputation as a
whole (deposit, x and y transforms, the fused z-transform/projection pass,
inverse y and x transforms, gather), timed with a hipEvent pair
(ising_time_demag), against the bytes its seven passes read and write once
each.

`--workload mix` shows the replica exchange mixing at the headline lattice
size: the same 64^3 x 64 schedule on a ladder narrow enough for exchanges to
be accepted (T 0.90..0.935 geometric; on the 0.2..2.0 ladder neighbouring
temperatures are ~20 energy standard deviations apart and no exchange is
accepted). 2000 equilibration sweeps, 2000 timed sweeps with the headline's
records, then 2000 swap intervals with the temperature map read after each:
acceptance per neighbouring pair and the replicas' round trips between the
coldest and the hottest temperature.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c3|bcc|hk|demag|mix]
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

L = 64
N_REP = 64
T_MIN, T_MAX = 0.2, 2.0
SWAP_EVERY = 10
MEASURE_EVERY = 2
WINDING_EVERY = 2            # Q of the middle layer only
TRAIN_SWEEPS = 2000          # plain sweeps that generate the training records
COLLECT_EVERY = 20
RECORDS_PER_REPLICA = 2048
TRAIN_STEPS = 2000
TRAIN_BATCH = 65536
EQUIL_SWEEPS = 4000          # after training, learned sampler, before the checkpoint
TRAIN_LR = 3e-3
MLP_FLOPS_PER_UPDATE = 2 * (21 * 32 + 32 * 32 + 32)
CONE = 0.5
SEED = 20261203
BYTES_PER_UPDATE = 24  # FP32 (x,y,z) read once + written once per site update
LAUNCH_TIMING_SWEEPS = 200   # sweeps of the schedule whose plain half-sweep launches are timed
TAU_SWEEPS = 10000           # record window of the tau_int-corrected metric, per sampler
REWARM_SWEEPS = 200          # untimed, between that window's copy-out and the timed steps
HBM_PEAK_GBS = 8000.0


def host_cores() -> tuple[int, int]:
    """(threads the CPU baseline uses, CPUs this process may run on): every
    CPU of the process's affinity set, capped by OMP_NUM_THREADS where the
    host sets it (a shared GPU host gives each GPU's jobs that many)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    use = min(avail, int(cap)) if cap.isdigit() and int(cap) > 0 else avail
    return use, avail


def cpu_baseline(seconds: float = 12.0) -> dict:
    """FP64 OpenMP oracle (exact checkerboard sweep, same Hamiltonian) on one
    64^3 replica at T = 0.5 J, timed on this host's cores."""
    from oracle import pyoracle as orc
    cores, avail = host_cores()
    orc.set_threads(cores)
    lat = orc.lattice(L, L, L)
    h = orc.ham(1.0, 0.5, "bulk", 0.2, 0.0)
    s = orc.init_random(lat, SEED, 0)
    t0 = time.perf_counter()
    orc.sweeps(lat, h, s, 2.0, CONE, SEED, 0, 0, 1)
    one = time.perf_counter() - t0
    n = max(1, int(seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    orc.sweeps(lat, h, s, 2.0, CONE, SEED, 0, 1, n)
    dt = time.perf_counter() - t0
    return {"value": L ** 3 * n / dt, "unit": "site updates/s", "cores": cores, "host_cpus": avail, "kind": "port",
            "sample": f"FP64 C/OpenMP oracle, 1 replica 64^3, T=0.5, r=0.5, {n} sweeps in {dt:.1f} s "
                      f"on {cores} threads ({avail} CPUs available)"}


def progress(msg: str) -> None:
    """One line per finished stage on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def counters_from_profiles() -> dict:
    """DRAM bytes per half-sweep launch and VALU busy of the sweep kernels,
    from the rocprofv3 counter runs summarised in profiles/."""
    p = ROOT / "profiles" / "sweep_traffic.json"
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tuning", type=int, default=None,
                    help="with --profile-sweeps: use these launch-tuning flags instead of autotuning")
    ap.add_argument("--eager", action="store_true",
                    help="launch every period eagerly instead of replaying a captured hipGraph (same kernels)")
    ap.add_argument("--workload", choices=("c5", "c3", "bcc", "hk", "demag", "mix"), default="c5",
                    help="c5: 64^3 x 64 PT, learned vs fixed (default); c3: 256 replicas of mixed lattices; "
                         "bcc / hk: 64-replica PT on body-centred-cubic / hyper-kagome cells; "
                         "mix: replica-exchange mixing on a narrow 64^3 x 64 ladder")
    ap.add_argument("--profile-sweeps", type=int, default=0,
                    help="only run N plain sweeps, N learned sweeps (untrained network) and one 256 MiB "
                         "copy probe (for rocprofv3 counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the path shards by PT ladder with no data exchange, "
              "so --gpus N > 1 is not measured here.", file=sys.stderr)
        return 2

    from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
    from ising_amd.analysis import independent_samples, ratio_bootstrap
    from ising_amd.learned import generate_data, train

    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "mix":
        return bench_mix(args)
    if args.workload in ("bcc", "hk"):
        return bench_cell(args, args.workload)
    if args.workload == "demag":
        return bench_demag(args)

    T = geometric_temperatures(T_MIN, T_MAX, N_REP)
    ham = Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.0)
    eng = Engine([Ladder(Lattice(L, L, L), ham, list(1.0 / T), CONE)], seed=SEED)
    eng.init_random()
    n_sites = L ** 3 * N_REP

    if args.profile_sweeps:
        from ising_amd.learned import initial_params
        if args.tuning is None:
            flags, _ = eng.autotune(5)
        else:
            flags = args.tuning
            eng.set_tuning(flags)
        ms = eng.time_sweeps(args.profile_sweeps)
        eng.mlp_set_params(initial_params(scale=0.3))
        eng.mlp_enable(True)
        ms_learned = eng.time_sweeps(args.profile_sweeps)
        copy_ms = eng.bandwidth_probe(256 << 20)
        print(json.dumps({"tuning": flags, "ms_per_half_sweep": ms, "ms_per_half_sweep_learned": ms_learned,
                          "copy_probe_bytes": 2 * (256 << 20), "copy_probe_ms": copy_ms}), flush=True)
        eng.close()
        return 0

    # Untimed preparation: launch-tuning flags chosen by timing (real sweeps);
    # 2000 plain sweeps that equilibrate and collect training records; Adam
    # training of the cone-width network on them.
    tune_flags, tune_ms = eng.autotune(10)
    progress(f"autotune: flags {tune_flags}")
    t0 = time.perf_counter()
    n_records = generate_data(eng, sweeps=TRAIN_SWEEPS, collect_every=COLLECT_EVERY, per_replica=RECORDS_PER_REPLICA,
                              swap_every=SWAP_EVERY, use_graph=not args.eager)
    progress(f"training data: {n_records} records")
    loss = train(eng, steps=TRAIN_STEPS, batch=TRAIN_BATCH, lr=TRAIN_LR)
    progress(f"training: loss {float(loss[:10].mean()):.4f} -> {float(loss[-10:].mean()):.4f}")
    # Further equilibration (the trained sampler relaxes the cold replicas
    # fastest), then one device checkpoint that both timed runs start from
    # with the same random-number counters.
    eng.mlp_enable(True)
    eng.pt_run(EQUIL_SWEEPS, SWAP_EVERY, 0, 0, use_graph=not args.eager)
    eng.save_state()
    prep_s = time.perf_counter() - t0
    progress("equilibrated and checkpointed")

    n_sweeps = args.steps * SWAP_EVERY

    def timed_run(learned: bool) -> dict:
        eng.restore_state()
        eng.mlp_enable(learned)
        # The record window of the tau_int-corrected metric: TAU_SWEEPS sweeps
        # of the same schedule, timed on their own, whatever --steps is, after
        # two untimed periods. It runs first, so the series buffers and the
        # captured graph it leaves on the context already fit the throughput
        # run below.
        eng.pt_run(2 * SWAP_EVERY, SWAP_EVERY, MEASURE_EVERY, WINDING_EVERY, use_graph=not args.eager,
                   winding_layer=L // 2)
        eng.synchronize()
        t1 = time.perf_counter()
        out = eng.pt_run(TAU_SWEEPS, SWAP_EVERY, MEASURE_EVERY, WINDING_EVERY, use_graph=not args.eager,
                         winding_layer=L // 2)
        eng.synchronize()
        tau_wall = time.perf_counter() - t1
        # The device idles while the window's records (328 MB with the charge
        # series of every layer) are copied out, and its clocks drop: a short
        # timed run straight after that starts ~4% slow. REWARM_SWEEPS untimed
        # sweeps bring it back, then the --warmup steps run right before the
        # timed ones.
        eng.pt_run(REWARM_SWEEPS, SWAP_EVERY, MEASURE_EVERY, WINDING_EVERY, use_graph=not args.eager,
                   winding_layer=L // 2, want=())
        if args.warmup > 0:
            eng.pt_run(args.warmup * SWAP_EVERY, SWAP_EVERY, MEASURE_EVERY, WINDING_EVERY, use_graph=not args.eager,
                       winding_layer=L // 2, want=())
        # Device-wide synchronisation around the timed region goes through the
        # library (torch is not imported: it bundles its own HIP runtime).
        # The charge records are computed and stored on the device as in the
        # window above; only E and M are copied out (the charge series of the
        # window is what the metric uses).
        eng.synchronize()
        t0 = time.perf_counter()
        eng.pt_run(n_sweeps, SWAP_EVERY, MEASURE_EVERY, WINDING_EVERY, use_graph=not args.eager, winding_layer=L // 2,
                   want=("energy", "magnetisation"))
        eng.synchronize()
        wall = time.perf_counter() - t0
        e_series = out["energy"]                                  # [n_meas, R]
        q_series = out["winding_raw"]                             # middle layer, [n_q, R]
        _, tau_e, conv_e = independent_samples(e_series)
        _, tau_q, conv_q = independent_samples(q_series)
        # the plain sweep launches timed inside this schedule (hipEvents), per
        # half sweep; the first window only brings the device back from the
        # host-side pause above (its clocks drop while it idles)
        for _ in range(2):
            ms_launch = eng.time_pt_sweeps(LAUNCH_TIMING_SWEEPS, SWAP_EVERY, MEASURE_EVERY, WINDING_EVERY,
                                           winding_layer=L // 2)
        return {"wall": wall, "updates_per_s": n_sites * n_sweeps / wall, "ms_per_launch": ms_launch,
                "tau_wall": tau_wall,
                "rate_E": e_series.shape[0] / (2.0 * tau_e) / tau_wall, "conv_E": conv_e,
                "rate_Q": q_series.shape[0] / (2.0 * tau_q) / tau_wall, "conv_Q": conv_q,
                "tau_int_E_sweeps": tau_e * MEASURE_EVERY, "tau_int_Q_sweeps": tau_q * WINDING_EVERY,
                "mean_E_per_site": (e_series.mean(axis=0) / L ** 3).tolist(),
                "mean_absM": np.linalg.norm(out["magnetisation"], axis=2).mean(axis=0).tolist()}

    fixed = timed_run(False)
    progress(f"fixed: {fixed['wall'] * 1e3 / args.steps:.3f} ms/step")
    learned = timed_run(True)
    progress(f"learned: {learned['wall'] * 1e3 / args.steps:.3f} ms/step")
    att, acc = eng.swap_stats()

    # The dominant kernel of the fixed-cone run: one fused launch per sweep
    # when the autotuner kept a fused flag set, else the plain half sweep.
    # time_pt_sweeps reports per half-sweep equivalent; a fused launch is two.
    fused = (tune_flags >> 8) & 0xFF > 0
    halves = 2 if fused else 1
    ms_launch = fixed["ms_per_launch"] * halves
    bytes_launch = BYTES_PER_UPDATE * n_sites / 2 * halves
    achieved = bytes_launch / (ms_launch * 1e-3) / 1e9
    counters = counters_from_profiles()
    traffic = counters.get("fused", {}).get("hbm_bytes_per_launch") if fused else counters.get("hbm_bytes_per_launch")
    copy_ms = eng.bandwidth_probe(256 << 20)
    mlp_tflops = MLP_FLOPS_PER_UPDATE * n_sites / 2 / (learned["ms_per_launch"] * 1e-3) / 1e12

    cpu = None if args.no_cpu else cpu_baseline()
    if cpu:
        progress(f"cpu baseline: {cpu['value']:.3g} updates/s on {cpu['cores']} threads")

    def summary(r: dict) -> dict:
        out = {"site_updates_per_s": r["updates_per_s"], "ms_per_step": r["wall"] * 1e3 / args.steps,
               "ms_per_half_sweep": r["ms_per_launch"], "tau_window_sweeps": TAU_SWEEPS,
               "tau_window_s": r["tau_wall"]}
        for o, key in (("E", "samples_per_s_E"), ("Q", "samples_per_s_Q_middle_layer")):
            conv, tau = r[f"conv_{o}"], r[f"tau_int_{o}_sweeps"]
            out[key] = float(r[f"rate_{o}"][conv].sum())
            out[f"unconverged_{o}"] = int((~conv).sum())
            t = tau[conv] if conv.any() else tau
            out[f"tau_int_{o}_sweeps"] = {"min": float(t.min()), "median": float(np.median(t)),
                                         "max": float(t.max()), "converged_replicas": int(conv.sum())}
        return out

    def gain(o: str) -> dict:
        """learned / fixed samples/s over the replicas converged in both runs,
        with the 16-84% bootstrap band over those replicas."""
        both = fixed[f"conv_{o}"] & learned[f"conv_{o}"]
        if not both.any():
            return {"ratio": None, "replicas": 0}
        r, lo, hi = ratio_bootstrap(learned[f"rate_{o}"][both], fixed[f"rate_{o}"][both])
        per_sweep = r * learned["tau_wall"] / fixed["tau_wall"]
        return {"ratio": r, "band_16_84": [lo, hi], "replicas": int(both.sum()), "samples_per_sweep_ratio": per_sweep}

    line = {
        "metric": "site updates/s; tau_int-corrected samples/s at 64^3x64 PT, learned vs fixed",
        "value": fixed["updates_per_s"],
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fixed["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 spins, fp64 reductions, bf16 MFMA network",
        "data": "synthetic (seeded random-sphere start)",
        "config": {"workload": "PT 64^3 x 64 replicas, T 0.2-2.0 geometric, J=1 D=0.5 (bulk) Bz=0.2; "
                               "swap every 10 sweeps, E/M every 2, middle-layer Q every 2; value = fixed cone r=0.5",
                   "sweeps_per_step": SWAP_EVERY},
        "fixed": summary(fixed),
        "learned": summary(learned),
        "learned_over_fixed": {"samples_per_s_E": gain("E"), "samples_per_s_Q": gain("Q"),
                               "site_updates_per_s": learned["updates_per_s"] / fixed["updates_per_s"]},
        "training": {"records": n_records, "steps": TRAIN_STEPS, "batch": TRAIN_BATCH,
                     "loss_first": float(loss[:10].mean()), "loss_last": float(loss[-10:].mean()),
                     "prep_seconds": prep_s},
        "observables": {"fixed_E_per_site": fixed["mean_E_per_site"][::8], "learned_E_per_site": learned["mean_E_per_site"][::8]},
        # tau_int(E) in sweeps at every 8th temperature position (negative: the window did not close)
        "tau_int_E_by_temperature": {
            "T": [round(float(x), 4) for x in T[::8]],
            "fixed": [round(float(t if c else -t), 2) for t, c in zip(fixed["tau_int_E_sweeps"][::8], fixed["conv_E"][::8])],
            "learned": [round(float(t if c else -t), 2)
                        for t, c in zip(learned["tau_int_E_sweeps"][::8], learned["conv_E"][::8])]},
        "swap_acceptance": {"min": float((acc[:-1] / np.maximum(att[:-1], 1)).min()),
                            "mean": float((acc[:-1] / np.maximum(att[:-1], 1)).mean())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "valu_busy": counters.get("fused", {}).get("valu_busy") if fused else counters.get("valu_busy"),
                     "learned_valu_busy": counters.get("learned_valu_busy"),
                     "kernel": "k_sweep_fused<false, false, false>" if fused else "k_half_sweep<false, false, false, false>",
                     "ms_per_launch": ms_launch, "site_updates_per_launch": n_sites // 2 * halves,
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "copy_probe_GBs": 2 * (256 << 20) / (copy_ms * 1e-3) / 1e9,
                     "tuning_flags": tune_flags,
                     "tuning_ms_per_half_sweep": {str(f): float(t) for f, t in tune_ms.items()},
                     "learned_kernel_mfma_TFLOPs": mlp_tflops, "mfma_bf16_peak_TFLOPs": 2500.0},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


C3_SIZES = (24, 32, 48)
C3_SWEEPS = 2000
C3_MEASURE_EVERY = 10
C3_SWAP_CANDIDATES = (1, 2, 5, 10, 20, 50)
C3_MAX_OVERHEAD = 0.05


def bench_c3(args) -> int:
    """Batched uneven PT: 8 ladders x 32 temperatures of mixed lattice sizes in
    one launch; launch shape and sweeps-per-swap interval chosen by the
    autotuners; 2000 timed sweeps."""
    from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
    T = geometric_temperatures(T_MIN, T_MAX, 32)
    ladders = [Ladder(Lattice(C3_SIZES[b % 3], C3_SIZES[b % 3], C3_SIZES[b % 3]),
                      Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.05 * b), list(1.0 / T), CONE) for b in range(8)]
    n_sites = sum(len(l.betas) * l.lattice.sites for l in ladders)
    eng = Engine(ladders, seed=SEED)
    eng.init_random()
    tune_flags, tune_ms = eng.autotune(10)
    progress(f"c3 autotune: flags {tune_flags}")
    swap_every, swap_ms = eng.autotune_schedule(C3_SWAP_CANDIDATES, n_sweeps=20, max_overhead=C3_MAX_OVERHEAD,
                                                measure_every=C3_MEASURE_EVERY)
    progress(f"c3 schedule: swap every {swap_every} ({swap_ms})")
    eng.pt_run(200, swap_every, 0, 0, use_graph=not args.eager)
    eng.synchronize()
    t0 = time.perf_counter()
    out = eng.pt_run(C3_SWEEPS, swap_every, C3_MEASURE_EVERY, 0, use_graph=not args.eager)
    eng.synchronize()
    wall = time.perf_counter() - t0
    progress(f"c3 timed: {wall * 1e3:.1f} ms for {C3_SWEEPS} sweeps")
    ms_launch = eng.time_pt_sweeps(200, swap_every, C3_MEASURE_EVERY, 0)
    att, acc = eng.swap_stats()
    bytes_launch = BYTES_PER_UPDATE * n_sites / 2
    achieved = bytes_launch / (ms_launch * 1e-3) / 1e9
    line = {
        "metric": "site updates/s, batched uneven PT (256 replicas on 24^3/32^3/48^3)",
        "value": n_sites * C3_SWEEPS / wall,
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": C3_SWEEPS,
        "warmup": 200,
        "ms_per_step": wall * 1e3 / C3_SWEEPS,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 spins, fp64 reductions",
        "data": "synthetic (seeded random-sphere start)",
        "config": {"workload": "batched uneven PT: 256 replicas = 8 Bz (0..0.35) x 32 T (0.2-2.0 geometric), "
                               "lattices 24^3/32^3/48^3 by field, J=1 D=0.5 (bulk); one launch per half sweep; "
                               f"2000 sweeps, E/M every {C3_MEASURE_EVERY}, swap interval from the schedule tuner",
                   "sites": n_sites, "step": "one sweep of every replica"},
        "schedule": {"swap_every": swap_every, "candidates_ms_per_sweep": {str(k): v for k, v in swap_ms.items()},
                     "max_overhead": C3_MAX_OVERHEAD},
        "swap_acceptance": {"mean": float((acc / np.maximum(att, 1))[att > 0].mean()),
                            "min": float((acc / np.maximum(att, 1))[att > 0].min())},
        "energy_per_site_first_ladder": (out["energy"][-1][:32:8] / ladders[0].lattice.sites).tolist(),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_half_sweep",
                     "ms_per_launch": ms_launch, "algorithmic_bytes_per_launch": bytes_launch,
                     "tuning_flags": tune_flags,
                     "tuning_ms_per_half_sweep": {str(f): float(t) for f, t in tune_ms.items()}},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


MIX_T = (0.90, 0.935)
MIX_EQUIL_SWEEPS = 2000
MIX_TIMED_SWEEPS = 2000
MIX_ROUNDS = 2000


def round_trips(slot_of_temp: np.ndarray) -> np.ndarray:
    """Per replica slot, the completed round trips coldest -> hottest ->
    coldest in a sequence of temperature maps ([n, R], slot at each position):
    a trip ends whenever a slot reaches the end opposite to the last end it
    visited, and two trips make a round trip."""
    n, r = slot_of_temp.shape
    last = np.full(r, -1)
    trips = np.zeros(r, dtype=np.int64)
    for m in slot_of_temp:
        for end, slot in ((0, m[0]), (1, m[r - 1])):
            if last[slot] != end:
                if last[slot] >= 0:
                    trips[slot] += 1
                last[slot] = end
    return trips // 2


def bench_mix(args) -> int:
    """The headline schedule on a narrow ladder: swap acceptance per pair and
    round trips of the replicas through the ladder."""
    from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
    T = geometric_temperatures(MIX_T[0], MIX_T[1], N_REP)
    lat = Lattice(L, L, L)
    eng = Engine([Ladder(lat, Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2), list(1.0 / T), CONE)], seed=SEED)
    eng.init_random()
    tune_flags, _ = eng.autotune(10)
    progress(f"mix autotune: flags {tune_flags}")
    eng.pt_run(MIX_EQUIL_SWEEPS, SWAP_EVERY, 0, 0, use_graph=not args.eager, want=())
    att0, acc0 = eng.swap_stats()
    eng.synchronize()
    t0 = time.perf_counter()
    out = eng.pt_run(MIX_TIMED_SWEEPS, SWAP_EVERY, MEASURE_EVERY, WINDING_EVERY, use_graph=not args.eager,
                     winding_layer=L // 2, want=("energy",))
    eng.synchronize()
    wall = time.perf_counter() - t0
    progress(f"mix timed: {wall * 1e3:.1f} ms for {MIX_TIMED_SWEEPS} sweeps")
    maps = [eng.temperature_map()]
    for _ in range(MIX_ROUNDS):
        eng.pt_run(SWAP_EVERY, SWAP_EVERY, 0, 0, use_graph=not args.eager, want=())
        maps.append(eng.temperature_map())
    att, acc = eng.swap_stats()
    eng.close()
    att, acc = att - att0, acc - acc0
    rate = acc[:-1] / np.maximum(att[:-1], 1)
    trips = round_trips(np.array(maps))
    moved = np.array(maps)
    n_sites = lat.sites * N_REP
    line = {
        "metric": "site updates/s and replica-exchange acceptance, 64^3x64 PT on a narrow ladder",
        "value": n_sites * MIX_TIMED_SWEEPS / wall,
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": MIX_TIMED_SWEEPS,
        "warmup": MIX_EQUIL_SWEEPS,
        "ms_per_step": wall * 1e3 / MIX_TIMED_SWEEPS,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 spins, fp64 reductions",
        "data": "synthetic (seeded random-sphere start)",
        "config": {"workload": f"PT mixing: 64^3 x 64 replicas, T {MIX_T[0]}-{MIX_T[1]} geometric, J=1 D=0.5 (bulk) "
                               f"Bz=0.2, cone 0.5, swap every {SWAP_EVERY}, E/M and middle-layer Q every 2 in the "
                               f"timed run; then {MIX_ROUNDS} swap intervals with the temperature map read after each",
                   "step": "one sweep of every replica"},
        "swap_acceptance": {"min": float(rate.min()), "mean": float(rate.mean()), "max": float(rate.max()),
                            "per_pair": [round(float(x), 4) for x in rate],
                            "attempts_per_pair": int(att[:-1].min()),
                            "rounds": "timed run and mixing run"},
        "round_trips": {"total": int(trips.sum()), "replicas_with_one": int((trips > 0).sum()),
                        "swap_rounds": MIX_ROUNDS,
                        "replicas_reaching_both_ends": int(sum(
                            1 for s in range(N_REP)
                            if (moved[:, 0] == s).any() and (moved[:, N_REP - 1] == s).any()))},
        "energy_per_site": (out["energy"][-1][::16] / lat.sites).tolist(),
        "tuning_flags": tune_flags,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


CELL_WORKLOADS = {
    # cell name, cells per axis, colour launches per sweep, description
    "bcc": ("bcc", 50, 2, "body-centred cubic 50^3 cells (2 atoms each; the reference's PT lattice)"),
    "hk": ("hyperkagome", 25, 3, "hyper-kagome 25^3 cells (12 atoms each; the reference's ISING lattice)"),
}
CELL_SWEEPS = 2000
CELL_MEASURE_EVERY = 10


def cpu_baseline_cell(cell: str, n_cells: int, seconds: float = 10.0) -> dict:
    """FP64 OpenMP oracle sweeps on one replica of the cell at T = 0.5 J."""
    from oracle import pyoracle as orc
    cores, _ = host_cores()
    orc.set_threads(cores)
    lat = orc.lattice(n_cells, n_cells, n_cells, cell=cell)
    n = orc.nsites(lat)
    h = orc.ham(1.0, 0.5, "bulk", 0.2, 0.0)
    s = orc.init_random(lat, SEED, 0)
    t0 = time.perf_counter()
    orc.sweeps(lat, h, s, 2.0, CONE, SEED, 0, 0, 1)
    one = time.perf_counter() - t0
    k = max(1, int(seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    orc.sweeps(lat, h, s, 2.0, CONE, SEED, 0, 1, k)
    dt = time.perf_counter() - t0
    return {"value": n * k / dt, "unit": "site updates/s", "cores": cores, "kind": "port",
            "sample": f"FP64 C/OpenMP oracle, 1 replica of {n_cells}^3 {cell} cells ({n} sites), T=0.5, r=0.5, "
                      f"{k} sweeps in {dt:.1f} s"}


def bench_cell(args, name: str) -> int:
    """64-replica PT on a multi-atom cell: autotuned launch shape, 2000 timed
    sweeps (swap every 10, E/M every 10), the colour launches timed inside
    the schedule."""
    from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
    cell, nc, launches, desc = CELL_WORKLOADS[name]
    T = geometric_temperatures(T_MIN, T_MAX, N_REP)
    lat = Lattice(nc, nc, nc, cell=cell)
    eng = Engine([Ladder(lat, Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2), list(1.0 / T), CONE)], seed=SEED)
    eng.init_random()
    n_sites = lat.sites * N_REP
    if args.profile_sweeps:
        if args.tuning is not None:
            eng.set_tuning(args.tuning)
        eng.sweep(args.profile_sweeps)
        eng.synchronize()
        progress(f"{name}: {args.profile_sweeps} plain sweeps done")
        eng.close()
        return 0
    tune_flags, tune_ms = eng.autotune(10)
    progress(f"{name} autotune: flags {tune_flags}")
    eng.pt_run(200, SWAP_EVERY, 0, 0, use_graph=not args.eager)
    eng.synchronize()
    t0 = time.perf_counter()
    out = eng.pt_run(CELL_SWEEPS, SWAP_EVERY, CELL_MEASURE_EVERY, 0, use_graph=not args.eager)
    eng.synchronize()
    wall = time.perf_counter() - t0
    progress(f"{name} timed: {wall * 1e3:.1f} ms for {CELL_SWEEPS} sweeps")
    ms_launch = eng.time_pt_sweeps(200, SWAP_EVERY, CELL_MEASURE_EVERY, 0)
    att, acc = eng.swap_stats()
    eng.close()
    bytes_launch = BYTES_PER_UPDATE * n_sites / launches
    achieved = bytes_launch / (ms_launch * 1e-3) / 1e9
    traffic = counters_from_profiles().get("cells", {}).get(name, {}).get("hbm_bytes_per_launch")
    cpu = None if args.no_cpu else cpu_baseline_cell(cell, {"bcc": 24, "hk": 12}[name])
    line = {
        "metric": f"site updates/s, 64-replica PT on {desc}",
        "value": n_sites * CELL_SWEEPS / wall,
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": CELL_SWEEPS,
        "warmup": 200,
        "ms_per_step": wall * 1e3 / CELL_SWEEPS,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 spins, fp64 reductions",
        "data": "synthetic (seeded random-sphere start)",
        "config": {"workload": f"PT on {desc}: 64 replicas, T 0.2-2.0 geometric, J=1 D=0.5 (bulk) Bz=0.2, "
                               f"swap every {SWAP_EVERY}, E/M every {CELL_MEASURE_EVERY}; {launches} colour launches "
                               "per sweep", "sites": n_sites, "step": "one sweep of every replica"},
        "swap_acceptance": {"mean": float((acc / np.maximum(att, 1))[att > 0].mean()),
                            "min": float((acc / np.maximum(att, 1))[att > 0].min())},
        "energy_per_site": (out["energy"][-1][::16] / lat.sites).tolist(),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_half_sweep",
                     "ms_per_launch": ms_launch, "algorithmic_bytes_per_launch": bytes_launch,
                     "traffic_source": "profiles/sweep_multiatom_cells.md (FETCH_SIZE x2, WRITE_SIZE x1)",
                     "tuning_flags": tune_flags,
                     "tuning_ms_per_launch": {str(f): float(t) for f, t in tune_ms.items()}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


DEMAG_INTERVAL = 7           # int(pi sqrt(5) / (2 r)) at r = 0.5 (src/energyMin.f90:258-259,362-369)
DEMAG_SWEEPS = 140


def demag_bytes_per_field(n_sites: int, lx: int, ly: int, lz: int) -> int:
    """HBM bytes of one field recomputation of one replica when each of its
    seven passes reads its input once and writes its output once: FP32 spins
    in (12 B per site) and the FP64 field out (24 B per site); the FP64 real
    mesh written by the deposit, read by the forward x transform, written by
    the inverse one and read by the gather (4 x 24 B per cell); the
    half-complex mesh written by the forward x transform, read and written by
    the forward y pass, the fused z-transform/projection pass and the
    inverse y pass, and read by the inverse x transform (8 x 48 B per mode,
    (lx/2+1) ly lz modes)."""
    cells = lx * ly * lz
    modes = (lx // 2 + 1) * ly * lz
    return 36 * n_sites + 96 * cells + 384 * modes


def cpu_baseline_demag(seconds_hint: float = 10.0) -> dict:
    """FP64 OpenMP oracle: Metropolis_mcs units of one field + 7 sweeps on one
    64^3 replica (the field of the oracle is direct per-axis DFTs)."""
    from oracle import pyoracle as orc
    cores, _ = host_cores()
    orc.set_threads(cores)
    lat = orc.lattice(L, L, L)
    h = orc.ham(1.0, 0.5, "bulk", 0.2, 0.0)
    s = orc.init_random(lat, SEED, 0)
    n = orc.nsites(lat)
    units, dt = 0, 0.0
    while dt < seconds_hint or units < 1:
        t0 = time.perf_counter()
        orc.sweeps_demag(lat, h, s, 2.0, CONE, SEED, 0, units * DEMAG_INTERVAL, DEMAG_INTERVAL, interval=0)
        dt += time.perf_counter() - t0
        units += 1
    return {"value": n * DEMAG_INTERVAL * units / dt, "unit": "site updates/s", "cores": cores, "kind": "port",
            "sample": f"FP64 C/OpenMP oracle, 1 replica of 64^3, T=0.5, r=0.5: {units} x (field + "
                      f"{DEMAG_INTERVAL} sweeps) in {dt:.1f} s"}


def bench_demag(args) -> int:
    """64 replicas of 64^3 with the demagnetising term: 140 timed sweeps with
    the field recomputed every 7, and the field recomputation timed alone."""
    from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
    T = geometric_temperatures(T_MIN, T_MAX, N_REP)
    lat = Lattice(L, L, L)
    eng = Engine([Ladder(lat, Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, demag=True), list(1.0 / T), CONE)],
                 seed=SEED)
    eng.init_random()
    eng.set_demag_interval(DEMAG_INTERVAL)
    n_sites = lat.sites * N_REP
    eng.sweep(2 * DEMAG_INTERVAL)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.sweep(DEMAG_SWEEPS)
    eng.synchronize()
    wall = time.perf_counter() - t0
    progress(f"demag timed: {wall * 1e3:.1f} ms for {DEMAG_SWEEPS} sweeps")
    eng.time_demag(2)
    ms_field = eng.time_demag(10)
    e = eng.energy()
    eng.close()
    bytes_field = demag_bytes_per_field(lat.sites, L, L, L) * N_REP
    achieved = bytes_field / (ms_field * 1e-3) / 1e9
    traffic = counters_from_profiles().get("demag", {}).get("hbm_bytes_per_field")
    cpu = None if args.no_cpu else cpu_baseline_demag()
    line = {
        "metric": "site updates/s with the demagnetising field, 64^3x64 Metropolis_mcs",
        "value": n_sites * DEMAG_SWEEPS / wall,
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": DEMAG_SWEEPS,
        "warmup": 2 * DEMAG_INTERVAL,
        "ms_per_step": wall * 1e3 / DEMAG_SWEEPS,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 spins, fp64 field and reductions",
        "data": "synthetic (seeded random-sphere start)",
        "config": {"workload": "Metropolis_mcs with the demagnetising term: 64 replicas of 64^3, T 0.2-2.0 "
                               "geometric, J=1 D=0.5 (bulk) Bz=0.2, r=0.5, field recomputed before the first "
                               f"sweep and after every {DEMAG_INTERVAL}th", "sites": n_sites,
                   "step": "one sweep of every replica with its share of field recomputations"},
        "ms_per_field": ms_field,
        "energy_per_site": (np.asarray(e)[::16] / lat.sites).tolist(),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/demag_field_pt64x64.md (FETCH_SIZE x2, WRITE_SIZE x1)",
                     "kernel": "field recomputation (k_demag_deposit, k_dft_lines_r x4, k_demag_zproject, "
                               "k_demag_gather)", "ms_per_launch": ms_field,
                     "algorithmic_bytes_per_launch": bytes_field},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
