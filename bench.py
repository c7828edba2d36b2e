#!/usr/bin/env python3
"""Headline benchmark: HMult(relin+rescale)/s + rotations/s at logN=16, L=34,
batched vs plain, with p50/p99 latency.

Workload (uneven batch): C ciphertext pairs (default 256) with levels
l = 34 - d, d ~ U{2..34} (so every rescale is legal, l in 0..32), slots
uniform[-1, 1] at scale 2^40; each pair does HMult (NTT conversion of both
inputs, multiply, relinearize, rescale) and then a rotation by its own step
s ~ U{1..1024} with that step's direct key (all keys HBM-resident).
  * plain:   the drop-in per-op path (the reference's op sequence, one
             launch per op), one ciphertext per call sequence;
  * unbatched fused: the fused kernels, one ciphertext per call sequence;
  * batched: ciphertexts of equal level processed as one batched call
             (shared relinearization key, per-entry rotation keys); the
             independent level groups are issued on two streams, largest
             and smallest side by side.
One "step" is one pass over all C pairs; value = (HMults + rotations) per
second in batched mode.  The key-switch roofline and the per-phase rates
are timed in a run of the same step with one level group at a time, so
each call has the GPU to itself.  Inputs are encrypted and resident in HBM before
timing.  Prints one JSON line.

After the timed steps, a serving run (``serving`` in the JSON line) feeds
Poisson arrivals of encrypted evaluations (a degree-7 Chebyshev polynomial
and a 64-slot window sum) to the continuous-batching loop and reports
throughput and p50/p99 latency, and a diagonal-method matrix-vector run
(``matvec``: 128 diagonals at level 0) reports its time.  The CPU baseline
(``cpu_baseline``, rank 0 only) times HMult + rotation at five levels on the
host's cores: on the reference's own CPU kernels when oracle/_ref holds them
(kind "reference"), with the oracle's C restatement beside it (kind "port").
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import statistics
import subprocess
import sys
import time

import numpy as np
import torch

METRIC = "rotations/s + HMult(relin+rescale)/s at logN=16 L=34, batched vs plain; p99"
HBM_PEAK_GBS = 8000.0
# Measured HBM bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated) per algorithmic
# byte of the fused key switch on this workload; the rocprofv3 --pmc passes
# and the arithmetic are recorded in profiles/keyswitch_fused_traffic.md
# (last re-measured on the current key-switch kernels).
KS_TRAFFIC_BYTES_PER_ALG_BYTE = 2.956
# Measured VALU wave-instructions (SQ_INSTS_VALU) per algorithmic byte of the
# fused key switch on this workload (profiles/keyswitch_fused_valu_issue.md).
KS_VALU_INSTS_PER_ALG_BYTE = 0.33187
# gfx950 issues one wave64 VALU instruction per CU per cycle for full-rate
# operations (the rate behind the 78.6 TFLOP/s binary64 FMA peak):
# 256 CUs x 2.4 GHz.
VALU_PEAK_GINSTS = 256 * 2.4


def parse_args():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--count", type=int, default=256, help="ciphertext pairs per step")
    p.add_argument("--seed", type=int, default=20260807)
    p.add_argument("--ntt-mode", default="fast", choices=["fast", "exact"])
    p.add_argument("--skip-plain", action="store_true", help="time only the batched mode (profiling)")
    p.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    p.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    p.add_argument("--serve-requests", type=int, default=600,
                   help="requests in the serving run (0 skips it)")
    p.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    p.add_argument("--serve-plan", default="whole", choices=["fine", "split", "whole"],
                   help="stage plan of the serving program (steps per serving-loop iteration)")
    p.add_argument("--serve-inflight", type=int, default=2,
                   help="serving iterations the host may run ahead of the device")
    p.add_argument("--matvec-diagonals", type=int, default=128,
                   help="diagonals of the matrix-vector run (0 skips it)")
    p.add_argument("--workload", default="batched", choices=["batched", "ntt", "matvec", "serving"],
                   help="batched: the headline line (uneven HMult + rotation batch, with every side "
                        "measurement); ntt / matvec / serving: only that configuration's measurement")
    p.add_argument("--streams", type=int, default=2,
                   help="streams the batched step issues its independent level groups on (1: one stream)")
    p.add_argument("--graph", dest="graph", action="store_true", default=True,
                   help="also time the batched step as one captured HIP graph (default on)")
    p.add_argument("--no-graph", dest="graph", action="store_false")
    p.add_argument("--calibration-copy", action="store_true",
                   help="run one 256 MiB device copy (torch.neg) before timing, as the byte "
                        "reference for hardware-counter traffic figures")
    return p.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(eng, count: int, seed: int):
    rng = np.random.default_rng(seed)
    levels = 34 - rng.integers(2, 35, size=count)
    steps = rng.integers(1, 1025, size=count)
    slots = eng.config.slots
    groups = {}
    for lvl in sorted(set(levels.tolist())):
        idx = np.nonzero(levels == lvl)[0]
        x = rng.uniform(-1, 1, size=(len(idx), slots))
        y = rng.uniform(-1, 1, size=(len(idx), slots))
        groups[lvl] = {
            "index": idx,
            "steps": [int(steps[i]) for i in idx],
            "a": eng.encrypt(x, level=lvl),
            "b": eng.encrypt(y, level=lvl),
            "x": x,
            "y": y,
        }
    return levels, steps, groups


def entry(ct, j):
    out = ct.clone()
    out.data = ct.data[:, j: j + 1]
    return out


def run_plain(eng, groups, lat_h, lat_r):
    outs = []
    for lvl, g in groups.items():
        for j, s in enumerate(g["steps"]):
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
            h = eng.hmult(entry(g["a"], j), entry(g["b"], j))
            e1.record()
            r = eng.rotate_by_step(h, s)
            e2.record()
            if lat_h is not None:
                lat_h.append((e0, e1))
                lat_r.append((e1, e2))
            outs.append(r)
    return outs


def run_batched(eng, groups, lat_h, lat_r, streams: int = 1):
    """One batched HMult and one batched rotation per level group.  With
    streams > 1 the independent level groups are issued on that many side
    streams (fhelium_amd.schedule.run_on_streams: with two, one stream walks
    from the largest group and one from the smallest)."""

    from fhelium_amd.schedule import run_on_streams

    items = list(groups.values())

    def task(g):
        def run():
            # timing events only when latencies are recorded
            if lat_h is None:
                return eng.rotate_by_steps(eng.hmult(g["a"], g["b"]), g["steps"])
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
            h = eng.hmult(g["a"], g["b"])
            e1.record()
            out = eng.rotate_by_steps(h, g["steps"])
            e2.record()
            lat_h.append((e0, e1, len(g["steps"])))
            lat_r.append((e1, e2, len(g["steps"])))
            return out
        return run

    return run_on_streams([task(g) for g in items], streams)


def graph_leg(eng, groups, steps: int, eager_outs) -> dict:
    """The whole batched step (every level group's HMult and rotations, one
    level group at a time on one stream) captured as one HIP graph
    (fhelium_amd.graph.CapturedStep) and replayed on the resident inputs;
    the first replay must equal the eager outputs bit for bit.  The
    two-stream schedule is not captured: capturing it ended the process with
    a segmentation fault, so the graph holds the one-stream schedule."""

    from fhelium_amd.graph import CapturedStep

    items = list(groups.items())
    inputs = [ct for _, g in items for ct in (g["a"], g["b"])]

    def step(*cts):
        return [eng.rotate_by_steps(eng.hmult(cts[2 * i], cts[2 * i + 1]), g["steps"])
                for i, (_, g) in enumerate(items)]

    captured = CapturedStep(step, *inputs)
    captured.graph.replay()
    torch.cuda.synchronize()
    exact = all(torch.equal(o.data, e.data) for o, e in zip(captured.static_output, eager_outs))
    t0 = time.perf_counter()
    for _ in range(steps):
        captured.graph.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    del captured
    return {"path": "whole batched step as one captured HIP graph, replayed on resident inputs",
            "ms_per_step": ms, "bit_exact_vs_eager": exact}


def timed(fn, steps: int, warmup: int, before_timing=None):
    for _ in range(warmup):
        fn(None, None)
    torch.cuda.synchronize()
    if before_timing is not None:
        before_timing()
    lat_h, lat_r = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        fn(lat_h, lat_r)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1000.0 / steps, lat_h, lat_r


def percentile(values, q):
    return float(np.percentile(np.asarray(values), q)) if values else None


def step_algorithmic_bytes(eng, groups) -> int:
    """Compulsory HBM bytes of one step: per HMult at level l in a group of
    B sharing the relinearization key, 4R in (two NTT-form inputs) + 2(R-1)
    out + dnum 2 (39-l) / B key limbs; per rotation 2R in + 2R out + its own
    key's dnum 2 (39-l) limbs (R = 35 - l Q rows, limb = N int64)."""

    cfg = eng.config
    limb = cfg.N * 8
    total = 0
    for lvl, g in groups.items():
        b = len(g["steps"])
        r = cfg.num_q_primes - lvl
        qp = r + cfg.num_p_primes
        dnum = len(eng.key_switcher.layout.digit_specs(lvl))
        total += limb * (b * (4 * r + 2 * (r - 1)) + dnum * 2 * qp)
        total += limb * b * (4 * r + dnum * 2 * qp)
    return total


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(eng, seconds_budget: float = 20.0):
    """The oracle (CPU restatement of the reference kernels) on the host's
    cores: HMult (NTT conversion of both inputs, multiply, relinearize,
    rescale) plus a rotation at levels 0, 8, 16, 24 and 32 of the same
    preset, repeated while the budget lasts.  Threads: every core this
    process may run on, capped by OMP_NUM_THREADS when the host sets it (the
    share of a shared GPU host given to one GPU)."""

    from oracle.engine_ref import EngineRef

    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = max(1, min(affinity, int(omp))) if omp and omp.isdigit() else max(1, affinity)
    os.environ["FH_ORACLE_THREADS"] = str(threads)
    ref = EngineRef(eng.runtime.host)
    levels = (0, 8, 16, 24, 32)
    rng = np.random.default_rng(1)
    rlk = eng.relinearization_key.data.cpu().numpy()
    rot = eng.rotation_key(1)
    rotk = rot.data.cpu().numpy()
    inputs = {}
    for level in levels:
        x = rng.uniform(-1, 1, size=(1, eng.config.slots))
        inputs[level] = (eng.encrypt(x, level=level).data.cpu().numpy(),
                         eng.encrypt(x, level=level).data.cpu().numpy())
    ops = 0
    per_level = {lvl: [] for lvl in levels}
    t0 = time.perf_counter()
    rounds = 0
    while True:
        for level in levels:
            an, bn = inputs[level]
            t1 = time.perf_counter()
            na = ref.coefficient_domain_to_ntt_domain(an, level)
            nb = ref.coefficient_domain_to_ntt_domain(bn, level)
            r = ref.rescale(ref.relinearize(ref.multiply(na, nb, level), rlk, level), level)
            ref.rotate(r, rotk, rot.galois_element, level + 1)
            per_level[level].append(time.perf_counter() - t1)
            ops += 2
        rounds += 1
        if time.perf_counter() - t0 > seconds_budget:
            break
    dt = time.perf_counter() - t0
    port = {"value": ops / dt, "unit": "ops/s", "cores": threads, "kind": "port",
            "sample": f"{rounds} round(s) of HMult + rotation at each of levels {list(levels)} "
                      "(logN=16, L34): the oracle's C restatement of the reference CPU kernels, "
                      "OpenMP over RNS rows",
            "ops_per_s_by_level": {str(k): 2 / float(np.median(v)) for k, v in per_level.items()}}
    host = {"os_cpu_count": os.cpu_count(), "affinity_cpus": affinity, "omp_num_threads": omp,
            "torch_num_threads": torch.get_num_threads(), "cpu_model": _cpu_model()}
    ref = reference_cpu_baseline(threads, seconds_budget, levels)
    if ref is None:
        return {**port, **host}
    return {"value": ref["value"], "unit": "ops/s", "cores": threads, "kind": "reference",
            "sample": f"{ref['rounds']} round(s) of HMult + rotation at each of levels {list(levels)} "
                      "(logN=16, L34) on the reference's own CPU kernels (csrc/ops/*/cpu, built from "
                      "source into oracle/_ref) composed as its engine composes them, indexed radix-2 "
                      "NTT, ATen intra-op threads",
            "ops_per_s_by_level": ref["ops_per_s_by_level"], "port": port, **host}


def reference_cpu_baseline(threads: int, seconds: float, levels) -> dict | None:
    """Times the reference's CPU kernels (oracle/_ref, when it was built) in a
    child process: that library defines the same operator names as this
    package's registrations.  None when it is absent or fails."""

    script = pathlib.Path(__file__).resolve().parent / "oracle" / "ref_baseline.py"
    lib = script.parent / "_ref" / "libfhelium_ref_cpu.so"
    if not lib.exists():
        return None
    try:
        proc = subprocess.run([sys.executable, str(script), "--seconds", str(seconds), "--threads", str(threads),
                               "--levels", ",".join(str(x) for x in levels)],
                              capture_output=True, text=True, timeout=10 * seconds + 120)
        if proc.returncode != 0:
            log(f"reference cpu baseline failed: {proc.stderr.strip()[-300:]}")
            return None
        return json.loads(proc.stdout.strip().splitlines()[-1])
    except (subprocess.TimeoutExpired, ValueError, IndexError) as exc:
        log(f"reference cpu baseline failed: {exc}")
        return None


def copy_ceiling(reps: int = 5) -> dict:
    """HBM copy rate of the library's own 16-byte-vector copy kernel
    (fh_stream_copy): 1 GiB (HBM-resident) and 64 MiB (fits the 256 MB
    Infinity Cache), bytes read + written per second."""

    from fhelium_amd import device as D

    out = {}
    for label, nbytes in (("hbm_1GiB", 1 << 30), ("cache_64MiB", 64 << 20)):
        src = torch.randint(0, 1 << 60, (nbytes // 8,), dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        D.stream_copy(dst, src)
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        for i in range(reps):
            evs[2 * i].record()
            D.stream_copy(dst, src)
            evs[2 * i + 1].record()
        torch.cuda.synchronize()
        ms = float(np.median([evs[2 * i].elapsed_time(evs[2 * i + 1]) for i in range(reps)]))
        out[label] = {"bytes_moved": 2 * nbytes, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
        del src, dst
    return {"kernel": "fh_stream_copy (16-byte vector loads and stores)", **out}


def ntt_measure(eng, reps: int = 5) -> dict:
    """Forward (to Montgomery) and inverse (to standard) NTT of the ring-kernel
    shape [64 ciphertexts, 2 polys, 35 limbs, 2^16], uniform residues,
    out of place; algorithmic bytes = one read and one write of every limb.
    The launch schedule of this shape is autotuned first (fhelium_amd/
    autotune.py; the chosen schedules are reported)."""

    from fhelium_amd import autotune

    cfg = eng.config
    be = eng.runtime.ntt
    if be.tuning is None:
        be.tuning = autotune.TuningCache()
    tuned = {d: autotune.tune_ntt(be, 128, cfg.num_q_primes, 0, forward=(d == "forward"))
             for d in ("forward", "inverse")}
    log(f"ntt autotune: forward {tuned['forward']['choice']}, inverse {tuned['inverse']['choice']}")
    q = torch.tensor(cfg.moduli[: cfg.num_q_primes], dtype=torch.int64, device="cuda").view(1, 1, -1, 1)
    x = torch.randint(0, 2**62, (64, 2, cfg.num_q_primes, cfg.N), dtype=torch.int64, device="cuda") % q
    ntt = eng.runtime.ntt
    limb_bytes = 2 * cfg.N * 8
    nbytes = x.numel() // cfg.N * limb_bytes
    out = {}
    for name, fn in (("forward", lambda t: ntt.forward_to_montgomery(t, 0)),
                     ("inverse", lambda t: ntt.inverse_to_standard(t, 0))):
        y = fn(x)  # warm-up
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        for i in range(reps):
            evs[2 * i].record()
            y = fn(x)
            evs[2 * i + 1].record()
        torch.cuda.synchronize()
        ms = float(np.median([evs[2 * i].elapsed_time(evs[2 * i + 1]) for i in range(reps)]))
        out[name] = {"ms": ms, "achieved": nbytes / (ms * 1e-3) / 1e9}
        del y
    del x
    achieved = 2 * nbytes / ((out["forward"]["ms"] + out["inverse"]["ms"]) * 1e-3) / 1e9
    return {
        "bound": "hbm", "kernel": "fh_ntt_forward / fh_ntt_inverse (two-pass radix-16, fast arithmetic)",
        "workload": "[64, 2, 35, 65536] int64 residues, forward to Montgomery + inverse to standard",
        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "algorithmic_bytes_per_call": nbytes,
        "forward_ms": out["forward"]["ms"], "inverse_ms": out["inverse"]["ms"],
        "autotune": {d: {"choice": r["choice"], "candidates_ms": r["candidates_ms"]} for d, r in tuned.items()},
    }


def matvec_run(eng, args, reps: int = 3) -> dict:
    """Diagonal-method matrix-vector product at level 0: one ciphertext of a
    uniform[-1, 1] 32768-slot vector times a K-diagonal uniform[-1, 1]
    plaintext matrix (rotations 1..K-1 with their own keys, one rescale)."""

    from fhelium_amd.linear import diagonal_matvec, prepare_diagonals

    k = args.matvec_diagonals
    rng = np.random.default_rng(args.seed + 9)
    for s in range(1, k):
        eng.rotation_key(s)
    x = rng.uniform(-1, 1, eng.config.slots)
    diags = rng.uniform(-1, 1, (k, eng.config.slots))
    ct = eng.encrypt(x, level=0)
    pts = prepare_diagonals(eng, diags, 0)
    out = diagonal_matvec(eng, ct, pts)  # warm-up
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    for i in range(reps):
        evs[2 * i].record()
        out = diagonal_matvec(eng, ct, pts)
        evs[2 * i + 1].record()
    torch.cuda.synchronize()
    ms = float(np.median([evs[2 * i].elapsed_time(evs[2 * i + 1]) for i in range(reps)]))
    want = sum(diags[i] * np.roll(x, i) for i in range(k))
    err = float(np.abs(eng.decrypt(out).real - want).max())
    log(f"matvec: {k} diagonals at level 0, {ms:.1f} ms, max-abs error {err:.2e}")
    return {
        "workload": f"diagonal-method matrix-vector product, {k} diagonals, level 0 (logN=16, L=34), "
                    f"rotations 1..{k - 1} as batched key switches with per-entry keys, one rescale",
        "ms": ms, "rotations_per_s": (k - 1) / (ms * 1e-3), "max_abs_error": err,
        "error_bound": k * 2e-5,
    }


def serving_run(eng, args) -> dict:
    """Serving loop: Poisson arrivals of degree-7 Chebyshev (sigmoid-like,
    seeded) evaluations on uniform[-4, 4] inputs plus a 64-slot window sum,
    10% of them at a reduced start level; keys resident; memory admission."""

    from fhelium_amd.serving import (PolynomialWindowSum, Request, ServingLoop, logistic_chebyshev,
                                     poisson_arrivals)

    rng = np.random.default_rng(args.seed + 5)
    prog = PolynomialWindowSum(eng, logistic_chebyshev(args.seed), radius=4.0, window=64,
                               stage_plan=args.serve_plan)
    reduced_level = 16
    pool = {}
    for lvl, k in ((0, 12), (reduced_level, 4)):
        xs = rng.uniform(-4, 4, size=(k, eng.config.slots))
        pool[lvl] = [(x, eng.encrypt(x, level=lvl)) for x in xs]
    n = args.serve_requests
    levels = np.where(rng.uniform(size=n) < 0.1, reduced_level, 0)
    arrivals = poisson_arrivals(n, args.serve_rate, rng)

    def requests(count, times):
        return [Request(i, pool[int(levels[i])][i % len(pool[int(levels[i])])][1], float(times[i]))
                for i in range(count)]

    loop = ServingLoop(prog, max_batch=32, max_active=64, max_inflight=args.serve_inflight)
    loop.run(requests(min(n, 8), np.zeros(n)))  # warm-up: kernels and constant plaintexts
    torch.cuda.synchronize()
    reqs = requests(n, arrivals)
    stats = loop.run(reqs)
    # service capacity: the same requests all arriving at once (the loop is
    # never idle, so completions per second measure what one GPU sustains)
    sat = requests(n, np.zeros(n))
    cap = loop.run(sat)
    log(f"serving capacity: {cap.completed} requests at once in {cap.wall_seconds:.2f}s, "
        f"{cap.throughput:.1f} req/s, max batch {cap.max_batch_seen}")
    errs = []
    for r in reqs[:4]:
        x = pool[r.start_level][r.rid % len(pool[r.start_level])][0]
        errs.append(float(np.abs(eng.decrypt(r.output).real - prog.expected(x)).max()))
    log(f"serving: {stats.completed} requests in {stats.wall_seconds:.2f}s, "
        f"{stats.throughput:.1f} req/s, p50 {stats.percentile(50) * 1e3:.1f} ms, "
        f"p99 {stats.percentile(99) * 1e3:.1f} ms, max active {stats.max_active}, arrivals realized "
        f"{(n - 1) / float(arrivals[-1] - arrivals[0]) if n > 1 else float('nan'):.1f} req/s")
    return {
        "workload": "Poisson arrivals of degree-7 Chebyshev (sigmoid-like, seeded) on uniform[-4,4] "
                    "plus a 64-slot window sum (6 rotate-and-add steps); 10% at start level 16",
        "requests": stats.completed, "arrival_rate_per_s": args.serve_rate,
        "arrival_rate_realized_per_s": (n - 1) / float(arrivals[-1] - arrivals[0]) if n > 1 else None,
        "completed_per_arrived": (stats.throughput * float(arrivals[-1] - arrivals[0]) / (n - 1)) if n > 1 else None,
        "throughput_per_s": stats.throughput, "p50_ms": stats.percentile(50) * 1e3,
        "p99_ms": stats.percentile(99) * 1e3, "max_active": stats.max_active,
        "capacity_per_s": cap.throughput, "capacity_note": "all requests arriving at once: completions per "
                                                            "second the loop sustains on one GPU",
        "stage_plan": args.serve_plan, "max_batch": stats.max_batch_seen, "iterations": stats.iterations,
        "key_switches_per_request": 12, "max_abs_error_sample": max(errs),
        "error_bound": 64 * 2e-5,
    }


def side_workload(eng, args) -> int:
    """One configuration on its own (profiling runs): prints its JSON line."""

    if args.workload == "ntt":
        copy = copy_ceiling()
        ntt = ntt_measure(eng)
        ntt["frac_of_copy_ceiling"] = ntt["achieved"] / copy["hbm_1GiB"]["GBps"]
        result = {"metric": "forward + inverse NTT throughput, [64, 2, 35, 2^16] int64 residues",
                  "value": ntt["achieved"], "unit": "GB/s (algorithmic)", "n_gpus": 1, "dtype": "int64",
                  "roofline": ntt, "hbm_copy_ceiling": copy}
    elif args.workload == "matvec":
        mv = matvec_run(eng, args)
        result = {"metric": "128-diagonal matrix-vector product (hoisted rotations), logN=16 L34",
                  "value": mv["ms"], "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "int64",
                  "matvec": mv}
    else:
        sv = serving_run(eng, args)
        result = {"metric": "serving throughput at Poisson arrivals", "value": sv["throughput_per_s"],
                  "unit": "requests/s", "n_gpus": 1, "dtype": "int64", "serving": sv}
    print(json.dumps(result), flush=True)
    return 0


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(the batch would shard across GPUs with no per-operation communication); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm device", file=sys.stderr)
        return 2
    torch.cuda.set_device(0)
    from fhelium_amd import _native as nat
    from fhelium_amd.config import Preset
    from fhelium_amd.engine import CkksEngine

    log(nat.device_info())
    t_setup = time.perf_counter()
    eng = CkksEngine(Preset.slots32768_scale40_levels34_int64, ntt_mode=args.ntt_mode, rng_seed=args.seed)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    if args.workload != "batched":
        return side_workload(eng, args)
    levels, steps, groups = build_workload(eng, args.count, args.seed)
    for i, s in enumerate(sorted(set(steps.tolist()))):
        eng.generate_rotation_key(int(s))
        if i % 50 == 49:
            torch.cuda.synchronize()
            log(f"{i + 1} rotation keys generated")
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t_setup:.1f}s: {args.count} pairs, {len(groups)} levels, "
        f"{len(eng.rotation_keys)} rotation keys, {torch.cuda.memory_allocated() / 2**30:.1f} GiB resident")

    # correctness sample (outside timing): decrypt a few batched results
    outs = run_batched(eng, groups, None, None)
    max_err = 0.0
    for (lvl, g), out in list(zip(groups.items(), outs))[:3]:
        got = eng.decrypt(entry(out, 0)).real[0]
        want = np.roll(g["x"][0] * g["y"][0], g["steps"][0])
        max_err = max(max_err, float(np.abs(got - want).max()))
    log(f"decrypted max-abs error (sample) {max_err:.3e}")

    if args.skip_plain:
        ms_plain, plh, plr = float("nan"), [], []
        ms_unb, ubh, ubr = float("nan"), [], []
    else:
        # plain: the drop-in path -- the reference's op sequence on the
        # per-op kernels (one launch per op), one ciphertext per call
        eng.fused_keyswitch = False
        ms_plain, plh, plr = timed(lambda h, r: run_plain(eng, groups, h, r), args.steps, args.warmup)
        eng.fused_keyswitch = True
        # unbatched fused: the fused kernels, still one ciphertext per call
        ms_unb, ubh, ubr = timed(lambda h, r: run_plain(eng, groups, h, r), args.steps, args.warmup)
    log(f"plain   {ms_plain:.2f} ms/step (per-op drop-in path)")
    log(f"unbatched fused {ms_unb:.2f} ms/step")
    if args.calibration_copy:
        src = torch.randint(0, 1 << 40, (1 << 25,), dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log("calibration copy: 268435456 B read + 268435456 B written (torch.neg, int64 x 2^25)")
        del src, dst
    def start_recorders():
        eng.key_switcher.recorder = []

    # one level group at a time: the per-call timings of the key-switch
    # roofline and the per-phase rates come from this run, each call alone
    # on the GPU
    ms_serial, bh, br = timed(lambda h, r: run_batched(eng, groups, h, r), args.steps, args.warmup,
                              start_recorders)
    ks_rec = eng.key_switcher.recorder
    eng.key_switcher.recorder = None
    log(f"batched, one level group at a time {ms_serial:.2f} ms/step")
    if args.streams > 1:
        outs_c = run_batched(eng, groups, None, None, args.streams)
        torch.cuda.synchronize()
        if not all(torch.equal(a.data, b.data) for a, b in zip(outs_c, outs)):
            raise RuntimeError("level groups on several streams differ from the one-stream step")
        ms_batch, ch, cr = timed(lambda h, r: run_batched(eng, groups, h, r, args.streams), args.steps,
                                 args.warmup)
        log(f"batched, level groups on {args.streams} streams {ms_batch:.2f} ms/step")
    else:
        ms_batch, ch, cr = ms_serial, bh, br
    graph = None
    if args.graph:
        graph = graph_leg(eng, groups, args.steps, outs)
        graph["value"] = 2 * args.count / (graph["ms_per_step"] * 1e-3)
        graph["vs_eager_one_stream"] = ms_serial / graph["ms_per_step"]
        graph["vs_timed_schedule"] = ms_batch / graph["ms_per_step"]
        log(f"graph replay {graph['ms_per_step']:.2f} ms/step (bit-exact vs eager: {graph['bit_exact_vs_eager']})")

    n_ops = 2 * args.count
    plain_h = [a.elapsed_time(b) for a, b in plh]
    plain_r = [a.elapsed_time(b) for a, b in plr]
    batch_h = [a.elapsed_time(b) for a, b, _ in bh]
    batch_r = [a.elapsed_time(b) for a, b, _ in br]
    conc_calls = [a.elapsed_time(b) for a, b, _ in ch + cr]
    ntt = ntt_measure(eng)
    copy = copy_ceiling()
    ntt["frac_of_copy_ceiling"] = ntt["achieved"] / copy["hbm_1GiB"]["GBps"]
    # timed region only: the recorders start after the warmup steps
    ks_ms = sum(a.elapsed_time(b) for a, b, _ in ks_rec)
    ks_bytes = sum(nb for _, _, nb in ks_rec)
    ks_calls = len(ks_rec)
    ks_achieved = ks_bytes / (ks_ms * 1e-3) / 1e9 if ks_ms > 0 else None
    step_bytes = step_algorithmic_bytes(eng, groups)
    step_achieved = step_bytes / (ms_batch * 1e-3) / 1e9

    result = {
        "metric": METRIC,
        "value": n_ops / (ms_batch * 1e-3),
        "unit": "ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic: seeded uniform[-1,1] slot vectors encrypted at scale 2^40, random keys",
        "config": {
            "workload": "batched uneven CKKS evaluation: {} ciphertext pairs, levels 0..32, "
                        "HMult+relin+rescale then rotate by per-pair step 1..1024".format(args.count),
            "logN": 16, "levels": 34, "pairs": args.count, "ntt_mode": args.ntt_mode,
        },
        "hmult_per_s": args.count / (sum(batch_h) / args.steps * 1e-3),
        "rotations_per_s": args.count / (sum(batch_r) / args.steps * 1e-3),
        "phase_rates_note": "hmult_per_s and rotations_per_s: each phase's calls timed in the run with one "
                            "level group at a time",
        "schedule": {"streams": args.streams,
                     "order": "level groups from both ends on two streams" if args.streams == 2 else
                              ("round robin" if args.streams > 2 else "one stream"),
                     "one_stream_ms_per_step": ms_serial,
                     "vs_one_stream": ms_serial / ms_batch},
        "plain": None if args.skip_plain else {
            "path": "drop-in per-op path: the reference's op sequence (one launch per op, ~100 per "
                    "key switch), one ciphertext per call",
            "value": n_ops / (ms_plain * 1e-3),
            "ms_per_step": ms_plain,
            "hmult_per_s": args.count / (sum(plain_h) / args.steps * 1e-3),
            "rotations_per_s": args.count / (sum(plain_r) / args.steps * 1e-3),
            "hmult_p50_ms": percentile(plain_h, 50), "hmult_p99_ms": percentile(plain_h, 99),
            "rotation_p50_ms": percentile(plain_r, 50), "rotation_p99_ms": percentile(plain_r, 99),
        },
        "unbatched_fused": None if args.skip_plain else {
            "path": "fused kernels (8-launch key switch, fused HMult front), one ciphertext per call",
            "value": n_ops / (ms_unb * 1e-3),
            "ms_per_step": ms_unb,
            "hmult_per_s": args.count / (sum(a.elapsed_time(b) for a, b in ubh) / args.steps * 1e-3),
            "rotations_per_s": args.count / (sum(a.elapsed_time(b) for a, b in ubr) / args.steps * 1e-3),
        },
        "graph_replay": graph,
        "batched_vs_plain": None if args.skip_plain else ms_plain / ms_batch,
        "batched_vs_unbatched_fused": None if args.skip_plain else ms_unb / ms_batch,
        "p50_ms": percentile(conc_calls, 50),
        "p99_ms": percentile(conc_calls, 99),
        "latency_note": "batched p50/p99 are per batched call (one level group, HMult or rotation phase) "
                        "in the timed schedule; one_stream_p50_ms/p99_ms with one level group at a time",
        "one_stream_p50_ms": percentile(batch_h + batch_r, 50),
        "one_stream_p99_ms": percentile(batch_h + batch_r, 99),
        "max_abs_error": max_err,
        "roofline": {
            "bound": "hbm",
            "kernel": "fused hybrid key switch (fh_keyswitch_fused: k_ks_decompose, k_ks_modup_columns, "
                      "k_ks_tiles_mac (with the inverse tile stages), k_inv_columns, k_ks_moddown_add)",
            "achieved": ks_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (ks_achieved / HBM_PEAK_GBS) if ks_achieved else None,
            "traffic": (KS_TRAFFIC_BYTES_PER_ALG_BYTE * ks_bytes / ks_calls
                        if ks_calls and KS_TRAFFIC_BYTES_PER_ALG_BYTE else None),
            "traffic_source": "estimate: measured HBM bytes per algorithmic byte "
                              f"({KS_TRAFFIC_BYTES_PER_ALG_BYTE}) from the rocprofv3 --pmc FETCH_SIZE and "
                              "WRITE_SIZE passes in profiles/keyswitch_fused_traffic.md, times this "
                              "run's algorithmic bytes",
            "calls": ks_calls,
            "timing": "HIP events around each fh_keyswitch_fused call on its launch stream, over the "
                      "timed run with one level group at a time (each call alone on the GPU)",
            "avg_call_ms": ks_ms / ks_calls if ks_calls else None,
            "algorithmic_bytes_per_call": (ks_bytes / ks_calls) if ks_calls else None,
        },
        "valu_roofline": {
            "bound": "valu",
            "kernel": "fused hybrid key switch (all eight launches)",
            "achieved": (KS_VALU_INSTS_PER_ALG_BYTE * ks_bytes / (ks_ms * 1e-3) / 1e9) if ks_ms > 0 else None,
            "peak": VALU_PEAK_GINSTS,
            "unit": "G VALU wave-instructions/s",
            "frac": (KS_VALU_INSTS_PER_ALG_BYTE * ks_bytes / (ks_ms * 1e-3) / 1e9 / VALU_PEAK_GINSTS)
                    if ks_ms > 0 else None,
            "valu_instructions_per_algorithmic_byte": KS_VALU_INSTS_PER_ALG_BYTE,
        },
        "ntt_roofline": ntt,
        "hbm_copy_ceiling": copy,
        "step_roofline": {
            "bound": "hbm", "algorithmic_bytes_per_step": step_bytes, "achieved": step_achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_achieved / HBM_PEAK_GBS,
        },
    }
    if args.serve_requests > 0:
        result["serving"] = serving_run(eng, args)
    if args.matvec_diagonals > 0:
        result["matvec"] = matvec_run(eng, args)
    if args.cpu_baseline:
        log("cpu baseline (oracle, then the reference's own CPU kernels when built) ...")
        result["cpu_baseline"] = cpu_baseline(eng)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
