#!/usr/bin/env python3
"""Benchmark: 30-qubit complex128 brick circuit on one MI355X (plain per-gate method).

One step = zero state + 24 brick layers (U3 on all 30 qubits, CZ on even/odd neighbour
pairs alternating) with 4 three-control X gates (at least one operand among the top 6
qubits) + full-register probability reduction + 4096 seeded samples.  That is 1072 gate
calls on a 16 GiB state.  `value` is gates per second over the timed steps; the state is
resident in HBM throughout.

Also reported:
  roofline      -- the dominant kernel (the 2x2 pair kernel that runs U3 and MCX), its
                   algorithmic bytes per launch (2 * 16 B * amplitudes it may change) over
                   its average duration from HIP events recorded around each launch on the
                   simulator's stream during the timed steps, against 8 TB/s;
  cpu_baseline  -- the oracle (the reference's plain method restated in C++/OpenMP) on the
                   host cores, on a bounded sample of the same workload;
  fused_c64     -- the precision form on the 32-qubit brick circuit of depth 32 (complex64
                   state, gates fused into k=5 blocks on FP32 MFMA) against the plain FP64
                   per-gate path on the same GPU: time of each, speedup, and fidelity /
                   max |d amp| of the fused state against the FP64 state (both resident in
                   HBM, compared on the GPU in FP64).
"""
import argparse
import json
import os
import pathlib
import sys
import time

REPO = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "rqs-svg_amd"))

import pyqcs  # noqa: E402
from pyqcs import circuits  # noqa: E402

HBM_PEAK = 8.0e12
CIRCUIT_SEED = 20250101
SAMPLE_SEED = 4242


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per pair-kernel launch measured with rocprofv3 PMC counters, if committed."""
    f = REPO / "profiles" / "pair_kernel_brick30_traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(n, circ, budget_s=20.0):
    """Time the oracle on the first gates of the same circuit (30 qubits, host cores)."""
    lib = pyqcs.oracle_library()
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    t_alloc = time.perf_counter()
    sim = pyqcs.Simulator(lib, n, seed=1)
    t0 = time.perf_counter()
    done = 0
    for op in circ:
        sim.gate(*op)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    sim.close()
    log(f"cpu baseline: {done} gates in {dt:.2f}s (alloc+init {t0 - t_alloc:.2f}s)")
    return {
        "value": done / dt,
        "unit": "gates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {done} gates of the same 30-qubit brick circuit (U3 on q0.., then CZ), "
                  f"complex128, oracle/qcs_oracle.cpp with OpenMP, {budget_s:.0f}s budget",
    }


def fused_leg(n, depth, k, shots):
    """Config: 32 qubits, brick depth 32; fused complex64 vs plain complex128 on one GPU."""
    lib = pyqcs.gpu_library()
    circ = circuits.brick(n, depth, seed=CIRCUIT_SEED + 3)
    warm = circuits.brick(n, 1, seed=1)[: n + 2]
    res = {"config": f"brick{n}: {n} qubits, {depth} brick layers (U3 all + CZ even/odd), "
                     f"{len(circ)} gates, {shots} samples",
           "gates": len(circ), "block_qubits": k}
    plain = pyqcs.Simulator(lib, n, seed=1)
    plain.run(warm)
    plain.call("set_zero_state")
    plain.norm()
    t0 = time.perf_counter()
    plain.run(circ)
    plain.probability_total()
    t_plain = time.perf_counter() - t0
    log(f"fused leg: plain c128 {t_plain:.3f}s")
    fused = pyqcs.Simulator(lib, n, seed=1, precision=1, block_qubits=k)
    fused.run(warm)
    fused.call("set_zero_state")
    fused.norm()
    g0, b0 = fused.fusion_stats()
    t0 = time.perf_counter()
    fused.run(circ)
    total = fused.probability_total()
    t_fused = time.perf_counter() - t0
    g1, b1 = fused.fusion_stats()
    ts = time.perf_counter()
    fused.sample(shots, seed=SAMPLE_SEED)
    t_sample = time.perf_counter() - ts
    ov, maxdiff = fused.overlap(plain)
    fid = abs(ov) ** 2 / (fused.norm() * plain.norm())
    fused.close()
    plain.close()
    log(f"fused leg: fused c64 {t_fused:.3f}s ({b1 - b0} blocks), fidelity {fid:.12f}, max|d| {maxdiff:.2e}")
    res.update({"plain_c128_s": t_plain, "fused_c64_s": t_fused, "speedup": t_plain / t_fused,
                "blocks": b1 - b0,
                "fidelity": fid, "infidelity": 1 - fid, "max_abs_diff": maxdiff,
                "probability_total": total, "sample_s": t_sample})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=24)
    ap.add_argument("--shots", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--fused-qubits", type=int, default=32)
    ap.add_argument("--fused-depth", type=int, default=32)
    ap.add_argument("--fused-k", type=int, default=5)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one write-only fill and one read-only norm pass of the "
                         "full state (known HBM bytes, for scaling PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding is a design, not part of this build", file=sys.stderr)
        sys.exit(2)

    n = args.qubits
    circ = circuits.brick(n, args.depth, seed=CIRCUIT_SEED, mcx=4)
    ngates = len(circ)
    alg_bytes = sum(circuits.algorithmic_bytes(n, op) for op in circ)
    log(f"workload: {n} qubits, depth {args.depth}, {ngates} gates, {alg_bytes / 1e12:.2f} TB algorithmic")

    lib = pyqcs.gpu_library()
    sim = pyqcs.Simulator(lib, n, seed=1)

    def step():
        sim.call("set_zero_state")
        sim.run(circ)
        total = sim.probability_total()
        out = sim.sample(args.shots, seed=SAMPLE_SEED)
        return total, out

    for i in range(args.warmup):
        t0 = time.perf_counter()
        total, _ = step()
        log(f"warmup {i}: {time.perf_counter() - t0:.3f}s total prob {total:.15f}")

    sim.profile_reset()
    sim.profile(True)
    e0, e1 = sim.event_create(), sim.event_create()
    sim.flush()
    t0 = time.perf_counter()
    sim.event_record(e0)
    for i in range(args.steps):
        ts = time.perf_counter()
        total, out = step()
        log(f"step {i}: {time.perf_counter() - ts:.3f}s")
    sim.event_record(e1)
    gpu_s = sim.elapsed(e0, e1)
    wall = time.perf_counter() - t0
    prof = sim.profile_get()
    sim.profile(False)
    log(f"timed: wall {wall:.3f}s, stream events {gpu_s:.3f}s")
    for k, v in prof.items():
        gbs = v["bytes"] / v["seconds"] / 1e9 if v["seconds"] else 0
        log(f"  {k:9s} launches {v['launches']:6d}  {v['seconds']:.4f}s  avg {v['seconds'] / v['launches'] * 1e3:.3f} ms  {gbs:.1f} GB/s")
    if args.calibrate:
        sim.call("set_zero_state")  # writes 16 * 2^n bytes
        log(f"calibration norm {sim.norm():.3f}")  # reads 16 * 2^n bytes
    sim.close()

    pair = prof.get("pair", {"launches": 1, "seconds": 1, "bytes": 0})
    avg_s = pair["seconds"] / pair["launches"]
    bytes_per_launch = pair["bytes"] / pair["launches"]
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = traffic_from_profiles()

    result = {
        "metric": "gates/s + % HBM roofline, 30q complex128; fused c64 speedup at fidelity>=1-1e-6",
        "value": ngates * args.steps / wall,
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 circuit parameters and sampling uniforms)",
        "config": {
            "workload": f"brick{n}_c128: {n} qubits complex128, {args.depth} brick layers "
                        f"(U3 all + CZ even/odd), 4 MCX with 3 controls, full probability "
                        f"reduction + {args.shots} samples",
            "qubits": n, "depth": args.depth, "gates": ngates, "shots": args.shots,
            "circuit_seed": CIRCUIT_SEED, "sample_seed": SAMPLE_SEED,
            "method": "plain: one kernel per gate",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_pair (U3 / MCX)",
            "achieved": achieved,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / (HBM_PEAK / 1e9),
            "traffic": traffic,
        },
        "sweeps_per_s": alg_bytes / (2 * 16 * (1 << n)) * args.steps / wall,
        "hbm_frac_whole_step": alg_bytes * args.steps / wall / HBM_PEAK,
    }
    if not args.no_fused:
        result["fused_c64"] = fused_leg(args.fused_qubits, args.fused_depth, args.fused_k, args.shots)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(n, circ, args.cpu_budget)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
