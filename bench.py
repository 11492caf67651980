#!/usr/bin/env python3
"""Benchmark: 28-qubit complex64 brickwork circuit (depth 20: U3 on every
qubit, then CZ on alternating nearest-neighbour pairs) on one MI355X.

A step is one simulate() of the whole circuit from |0..0> with the plain
per-gate HIP kernels (830 gate passes over the 2 GiB state).  The JSON line
reports gate-applications/s for that path, the HBM roofline of its dominant
kernel (the one-qubit U3 pass), the compile()-fused form of the same circuit,
infidelity against a complex128 run of the same circuit, and the FP64 CPU
oracle timed on the host as a baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from afquantumsim_amd import circuits, hip, ops  # noqa: E402

N_QUBITS = 28
DEPTH = 20
SEED = 2
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(lowered, n):
    return sum(16.0 * (1 << n) * ops.touched_fraction(o, n) for o in lowered)


def host_cpu():
    """nproc, the CPU model and the threads this process may use."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        import subprocess
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20)
        for line in r.stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_threads": usable, "model": model}


def cpu_baseline(gates, n, budget_s=10.0):
    """The plain complex64 CPU state vector (oracle/cpu_sv.c: in-place passes,
    OpenMP over every thread this process may use) on the same workload.

    value: gate-apps/s on a bounded sample of the 28-qubit brickwork -- whole
    layers from the start of the circuit until budget_s has passed.  Beside
    it: the 20-qubit random-layer circuit (1200 gates) plus 1000 sampled shots
    in full, and the reference's own method restated (4^n extrapolation)."""
    import oracle
    host = host_cpu()
    threads = host["usable_threads"]
    sv = np.zeros(1 << n, dtype=np.complex64)
    sv[0] = 1.0
    done = layers = 0
    t0 = time.perf_counter()
    while done < len(gates):
        layer_end = done + n
        # the CZ pairs of this layer follow its U3s
        while layer_end < len(gates) and gates[layer_end]["kind"] == "CZ":
            layer_end += 1
        oracle.cpu_sv_run(sv, n, gates[done:layer_end], threads=threads)
        done = layer_end
        layers += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    del sv
    # the 20-qubit random-layer workload with 1000 shots, whole circuit
    n20, g20 = 20, circuits.random_layered(20, 40, seed=1)
    u = np.random.default_rng(12345).uniform(size=1000)
    sv20 = np.zeros(1 << n20, dtype=np.complex64)
    cdf = np.empty(1 << n20, dtype=np.float64)
    times = []
    for _ in range(3):
        sv20[:] = 0
        sv20[0] = 1.0
        t1 = time.perf_counter()
        oracle.cpu_sv_run(sv20, n20, g20, threads=threads)
        oracle.cpu_sv_sample(sv20, n20, u, threads=threads, cdf=cdf)
        times.append(time.perf_counter() - t1)
    t20 = float(np.median(times))
    return {"value": round(done / dt, 3), "unit": "gate-apps/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of {len(gates)} gates ({layers} whole layer(s)) of the "
                      f"{n}-qubit brickwork, complex64 in place, {threads} OpenMP threads (oracle/cpu_sv.c)",
            "host": host,
            "random20_with_1000_shots": {"ms_per_circuit": round(1e3 * t20, 2),
                                         "gate_apps_per_s": round(len(g20) / t20, 1),
                                         "sample": "whole 1200-gate circuit + 1000 shots, median of 3"},
            "reference_method": reference_method_baseline(n, threads)}


def reference_method_baseline(n_target, threads, sizes=(10, 11, 12)):
    """The reference's own algorithm restated on the CPU (oracle.refmethod_run:
    an N x N identity product, then a dense N x N product per dense gate and an
    O(N) pass per CSR gate) on depth-2 brickwork circuits at small n; the time
    per gate application grows as 4^n, so the fit t(n) = c 4^n extrapolates
    it to the n_target-qubit workload."""
    import oracle
    rates, per_gate = {}, {}
    for n in sizes:
        gates = circuits.brickwork(n, 2, seed=SEED)
        sv = oracle.basis_state(n, 0)
        t0 = time.perf_counter()
        oracle.refmethod_run(sv, n, gates, threads)
        dt = time.perf_counter() - t0
        rates[n] = round(len(gates) / dt, 2)
        per_gate[n] = dt / len(gates)
    c = sum(per_gate[n] / 4.0 ** n for n in sizes) / len(sizes)
    return {"unit": "gate-apps/s", "cores": threads, "kind": "reference-method restatement",
            "measured": rates, "fit": "seconds per gate = c * 4^n", "c": c,
            "extrapolated_value": 1.0 / (c * 4.0 ** n_target),
            "sample": f"depth-2 brickwork at n = {list(sizes)}; the reference itself needs ArrayFire, "
                      "absent from this image"}


def batched(args):
    """4096 variational states (n ~ U{10..16}, L ~ U{4..12} layers of RY/RZ
    + CX ladder), <sum Z> in FP64, one persistent launch from a hipGraph.
    Compared with the plain per-gate path on a sample of the same states."""
    spec = circuits.variational_batch(4096, seed=4)
    lowered = [(n, ops.lower_all(g)) for n, g in spec]
    n_gates = sum(len(o) for _, o in lowered)
    t0 = time.perf_counter()
    b = hip.Batch(lowered)
    setup_s = time.perf_counter() - t0
    tuned = b.autotune()
    times = []
    expect = None
    for _ in range(max(2, args.steps)):
        expect, ms = b.run()
        times.append(ms)
    ms = float(np.median(times))
    info = b.info()
    # FP64 check of a few states against their own complex128 single-state run
    worst = 0.0
    for i in (0, 1, 2, 4095):
        n, g = spec[i]
        st = hip.State(n, hip.C128)
        st.apply(ops.lower_all(g, exact=True))
        worst = max(worst, abs(st.expect_sum_z() - expect[i]))
        st.close()
    # plain per-gate path on a sample (one state at a time)
    # the same states one at a time through aqs_apply (states of at most 12
    # qubits run there as LDS-resident op runs, larger ones one launch per op)
    sample = lowered[:256]
    sg = sum(len(o) for _, o in sample)
    st_cache = {}
    t0 = time.perf_counter()
    for n, o in sample:
        st = st_cache.get(n)
        if st is None:
            st = st_cache[n] = hip.State(n, hip.C64)
        st.set_basis(0)
        st.apply(o)
        st.expect_sum_z()
    plain_s = time.perf_counter() - t0
    for st in st_cache.values():
        st.close()
    b.close()
    out = {"states": len(spec), "gates": n_gates, "passes_after_absorption": info["passes"],
           "launch_ms": round(ms, 3), "gate_apps_per_s": round(n_gates / (ms * 1e-3), 1),
           "autotuned": tuned, "setup_s": round(setup_s, 3),
           "max_abs_err_sum_z_vs_fp64": worst,
           "plain_per_gate_gate_apps_per_s": round(sg / plain_s, 1),
           "speedup_vs_plain": round((n_gates / (ms * 1e-3)) / (sg / plain_s), 1)}
    log(f"batched: {out}")
    return out


def random20(args, n=20, depth=40, shots=1000):
    """20-qubit random layered circuit (H/X/RZ/RY per qubit, CX on a random
    matching; 1200 gates), simulate() from |0..0>, then 1000 sampled shots
    (profile_measure_all); plain per-gate and compile()-fused."""
    gates = circuits.random_layered(n, depth, seed=1)
    lowered = ops.lower_all(gates)
    u = np.random.default_rng(12345).uniform(size=shots)
    st = hip.State(n, hip.C64)
    out = {"qubits": n, "depth": depth, "gates": len(lowered), "shots": shots}
    reps = max(3, args.steps)
    for label, sch in (("plain", None), ("fused_k5", hip.Schedule(lowered, n, 5)),
                       ("tiled", hip.Schedule(lowered, n, mode="tiled"))):
        def run_once():
            st.set_basis(0)
            if sch is None:
                st.apply(lowered)
            else:
                st.apply_schedule(sch)
            return st.sample(u)
        run_once()
        st.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            idx = run_once()
        dt = (time.perf_counter() - t0) / reps
        entry = {"ms_per_circuit_with_sampling": round(1e3 * dt, 3),
                 "gate_apps_per_s": round(len(lowered) / dt, 1)}
        if sch is not None:
            entry["passes"] = sch.info()[0]
        out[label] = entry
    # sampled indices against the oracle's rule on the same amplitudes
    import oracle
    psi = st.download()
    out["samples_match_oracle"] = bool(np.array_equal(idx.astype(np.uint32), oracle.sample_c64(psi, n, u)))
    ref = hip.State(n, hip.C128)
    ref.apply(ops.lower_all(gates, exact=True))
    ov = st.inner(ref)
    out["infidelity_vs_fp64"] = max(0.0, 1.0 - abs(ov) ** 2 / (st.norm2() * ref.norm2()))
    ref.close()
    st.close()
    log(f"random20: {out}")
    return out


# The reference's published algorithm timings (benchmark/results.md:33-35:
# QFT 10 qubits, Grover 10 qubits, Shor N = 15 on 12 qubits; build +
# simulate per repetition, mean of 100, ArrayFire 3.8.2 OpenCL on a Radeon
# Pro 5500M).
REFERENCE_ALGO_MS = {"qft10": 6.05, "grover10": 192.0, "shor12": 106.0}


def reference_algorithms(args):
    """The reference's own benchmark program (benchmark/benchmark.cpp)
    restated over this library's aqs:: C++ API (afquantumsim_amd/algo_bench,
    a child process), against the published numbers."""
    import subprocess
    exe = os.path.join(ROOT, "afquantumsim_amd", "algo_bench")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build it with `python -m afquantumsim_amd.build`")
    r = subprocess.run([exe, str(args.algo_reps)], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"algo_bench failed ({r.returncode}): {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    res["reference_published_ms"] = REFERENCE_ALGO_MS
    res["speedup_vs_published"] = {k: round(REFERENCE_ALGO_MS[k] / res["uncompiled"][f"{k}_ms"][0], 1)
                                   for k in REFERENCE_ALGO_MS}
    log(f"reference algorithms: {res}")
    return res


def qft_closed_form(n=26, seed=26):
    """The compiled QFT (fourier_transform + final swaps, and the inverse
    QFT) on a seeded Gaussian complex64 state, against the closed forms
    evaluated with numpy's FP64 FFT on the host:
      fourier_transform(n) + swaps = bitrev(ifft(bitrev(v))) sqrt(N)
      inverse_fourier_transform(n) = bitrev(fft(v)) / sqrt(N)
    26 qubits keeps the host FFT to seconds; the 30-qubit runs are compared
    with the complex128 path on the device."""
    out = {"qubits": n}
    st = hip.State(n, hip.C64)

    def bitrev(x):
        return x.reshape([2] * n).transpose(list(reversed(range(n)))).reshape(-1)

    for label, gates, form in (
            ("qft", circuits.fourier_transform(n, final_swaps=True),
             lambda v: bitrev(np.fft.ifft(bitrev(v), norm="ortho"))),
            ("iqft", circuits.inverse_fourier_transform(n), lambda v: bitrev(np.fft.fft(v, norm="ortho")))):
        sch = hip.Schedule(ops.lower_all(gates), n, mode="auto")
        st.gaussian(seed)
        v = st.download().astype(np.complex128)
        st.apply_schedule(sch)
        got = st.download().astype(np.complex128)
        sch.close()
        want = form(v)
        ov = np.vdot(want, got)
        infid = max(0.0, 1.0 - abs(ov) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real))
        out[f"{label}_infidelity_vs_fft"] = infid
        if not infid < 1e-9:
            raise RuntimeError(f"{label}({n}) departs from the FFT closed form: infidelity {infid}")
    st.close()
    log(f"QFT closed form: {out}")
    return out


def qft30(args, n=30, seed=30, inverse=False):
    """30-qubit QFT (fourier_transform + final swaps, or
    inverse_fourier_transform) on a seeded Gaussian complex64 state
    (normalised in FP64): unfused per-gate path, compile() fused into <= 5-qubit
    MFMA blocks, the tiled schedule, and a complex128 run as reference."""
    gates = (circuits.inverse_fourier_transform(n) if inverse
             else circuits.fourier_transform(n, final_swaps=True))
    lowered = ops.lower_all(gates)
    alg = algorithmic_bytes(lowered, n)
    st = hip.State(n, hip.C64)
    ref = hip.State(n, hip.C128)
    ref.gaussian(seed)
    ref.apply(ops.lower_all(gates, exact=True))
    out = {"qubits": n, "gates": len(lowered), "algorithmic_bytes_unfused": alg}

    def infid():
        ov = st.inner(ref)
        return max(0.0, 1.0 - abs(ov) ** 2 / (st.norm2() * ref.norm2()))

    st.gaussian(seed)
    st.sync()
    t0 = time.perf_counter()
    st.apply(lowered)
    st.sync()
    dt = time.perf_counter() - t0
    out["unfused"] = {"ms": round(1e3 * dt, 2), "gate_apps_per_s": round(len(lowered) / dt, 1),
                      "hbm_gbs": round(alg / dt / 1e9, 1), "infidelity": infid()}
    for kmax in (5, "tiled"):
        t0 = time.perf_counter()
        sch = hip.Schedule(lowered, n, kmax) if kmax != "tiled" else hip.Schedule(lowered, n, mode="tiled")
        fuse_s = time.perf_counter() - t0
        passes, _ = sch.info()
        st.gaussian(seed)
        st.apply_schedule(sch)  # uploads the block matrices once
        st.gaussian(seed)
        st.sync()
        st.profile(True)
        t0 = time.perf_counter()
        st.apply_schedule(sch)
        st.sync()
        dt = time.perf_counter() - t0
        prof = st.profile_read()
        st.profile(False)
        mf_ms, mf_n = prof.get("MATK_MFMA", (0.0, 0))
        inf_vs_c128 = infid()
        out["tiled" if kmax == "tiled" else f"fused_k{kmax}"] = {"passes": passes, "ms": round(1e3 * dt, 2),
                                 "gate_apps_per_s": round(len(lowered) / dt, 1), "infidelity": inf_vs_c128,
                                 "fuse_host_s": round(fuse_s, 3), "mfma_launches": mf_n,
                                 "mfma_avg_ms": round(mf_ms / max(1, mf_n), 4),
                                 "mfma_tflops": round(8 * 32 * (1 << n) / (mf_ms / max(1, mf_n) * 1e-3) / 1e12, 1)
                                 if mf_n else None,
                                 "speedup_vs_unfused": round(out["unfused"]["ms"] / (1e3 * dt), 2),
                                 "profile_ms": {k: [round(v[0], 3), v[1]] for k, v in prof.items()}}
        sch.close()
    st.close()
    ref.close()
    log(f"{'iqft30' if inverse else 'qft30'}: {out}")
    return out


def served_requests(count, seed):
    """Brickwork requests: n log-uniform over 12..28 (P(n) ~ 1/n on the
    integers), depth ~ U{10..100}, 1000 shots with seeded uniforms."""
    rng = np.random.default_rng(seed)
    ns = np.arange(12, 29)
    p = (1.0 / ns) / np.sum(1.0 / ns)
    out = []
    for i in range(count):
        n = int(rng.choice(ns, p=p))
        depth = int(rng.integers(10, 101))
        gates = circuits.brickwork(n, depth, seed=50_000 + i)
        lowered = ops.lower_all(gates)
        out.append({"id": i, "n": n, "depth": depth, "ops": lowered, "packed": hip.PackedOps(lowered), "gates": gates,
                    "u": rng.uniform(size=1000)})
    return out


def serve_stream(reqs, rate, budget, workers, seed, coalesce=True):
    """Poisson arrivals at `rate` per second into a fresh server (requests
    packed beforehand); returns per-request latency (arrival -> sampled
    indices) and queueing, wall time and server statistics."""
    import threading
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
    srv = hip.Server(budget, workers=workers)
    if not coalesce:
        srv.set_grouping(0)
    lat = {}
    submitted = threading.Semaphore(0)

    def collector():
        for r in reqs:
            submitted.acquire()
            res, _, _ = srv.wait(r["id"])
            lat[r["id"]] = (res["t_done"] - res["t_submit"], res["t_admit"] - res["t_submit"])

    th = threading.Thread(target=collector)
    th.start()
    start = time.perf_counter()
    for r, at in zip(reqs, arrivals):
        delay = start + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        srv.submit(r["id"], r["n"], r["packed"], r["u"])
        submitted.release()
    th.join()
    wall = time.perf_counter() - start
    stats = srv.stats()
    srv.close()
    l = np.array([lat[r["id"]][0] for r in reqs])
    q = np.array([lat[r["id"]][1] for r in reqs])
    gates = sum(len(r["ops"]) for r in reqs)
    return {"arrival_rate_per_s": round(rate, 2), "p50_ms": round(1e3 * float(np.percentile(l, 50)), 2),
            "p99_ms": round(1e3 * float(np.percentile(l, 99)), 2), "mean_queue_ms": round(1e3 * float(q.mean()), 2),
            "throughput_req_per_s": round(len(reqs) / wall, 2), "gate_apps_per_s": round(gates / wall, 1),
            "max_concurrent": stats["max_concurrent"], "peak_resident_bytes": stats["peak_resident_bytes"],
            "coalesced_launches": stats["groups"], "coalesced_requests": stats["grouped"]}


def served(args):
    """Poisson arrivals of mixed-size brickwork requests into the request
    server (4 workers; compile() fusion per request above 16 qubits, queued
    requests of at most 16 qubits coalesced into batched launches): p50/p99
    latency (arrival -> sampled indices) and throughput at arrival rates from
    0.7x one worker's capacity up to and past the saturation of all four, and
    at the two highest rates with coalescing off; the same requests run one
    at a time with the plain per-gate kernels as the comparator; infidelity
    of served results against complex128."""
    budget = 64 << 30
    workers = 4
    # service-time calibration: requests alone, one after another
    calib = served_requests(24, seed=77)
    srv = hip.Server(budget, workers=1)
    t0 = time.perf_counter()
    for r in calib:
        srv.submit(r["id"], r["n"], r["packed"], r["u"])
        srv.wait(r["id"])
    alone_s = (time.perf_counter() - t0) / len(calib)
    srv.close()
    reqs = served_requests(args.serve_requests, seed=78)
    sweep = []
    for mult in (0.7, 2.0, 4.0, 8.0):
        entry = serve_stream(reqs, mult / alone_s, budget, workers, seed=79)
        entry["rate_vs_one_worker_capacity"] = mult
        sweep.append(entry)
        log(f"served at {mult}x one worker's capacity: {entry}")
    no_coalescing = []
    for mult in (4.0, 8.0):
        entry = serve_stream(reqs, mult / alone_s, budget, workers, seed=79, coalesce=False)
        entry["rate_vs_one_worker_capacity"] = mult
        no_coalescing.append(entry)
        log(f"served at {mult}x, coalescing off: {entry}")
    main = sweep[0]
    # plain comparator: the calibration requests one at a time with the
    # per-gate kernels (no fusion, no concurrency), same uniforms; the fused
    # one-worker service time above is on the same requests
    plain = calib
    st_cache = {}
    t0 = time.perf_counter()
    worst_plain = 0.0
    for r in plain:
        st = st_cache.get(r["n"])
        if st is None:
            st = st_cache[r["n"]] = hip.State(r["n"], hip.C64)
        st.set_basis(0)
        st.apply(r["packed"])
        st.sample(r["u"])
    plain_s = time.perf_counter() - t0
    for st in st_cache.values():
        st.close()
    plain_gates = sum(len(r["ops"]) for r in plain)
    # accuracy of served results (every request of at most 24 qubits in the
    # plain sample and beyond, up to 16) against complex128, and of the
    # plain path on the same requests
    worst = 0.0
    checked = 0
    chk = hip.Server(budget, workers=2)
    for r in [x for x in reqs if x["n"] <= 24][:16]:
        chk.submit(10_000 + r["id"], r["n"], r["packed"], r["u"][:10], keep_state=True)
        _, _, psi = chk.wait(10_000 + r["id"])
        ref = hip.State(r["n"], hip.C128)
        ref.apply(ops.lower_all(r["gates"], exact=True))
        x = ref.download()
        ov = np.vdot(x, psi.astype(np.complex128))
        worst = max(worst, 1.0 - abs(ov) ** 2 / (np.vdot(psi, psi).real * np.vdot(x, x).real))
        st = hip.State(r["n"], hip.C64)
        st.set_basis(0)
        st.apply(r["packed"])
        ovp = st.inner(ref)
        worst_plain = max(worst_plain, 1.0 - abs(ovp) ** 2 / (st.norm2() * ref.norm2()))
        st.close()
        ref.close()
        checked += 1
    chk.close()
    out = {"requests": len(reqs), "workers": workers, "hbm_budget_bytes": budget,
           "arrival_rate_per_s": main["arrival_rate_per_s"], "p50_ms": main["p50_ms"], "p99_ms": main["p99_ms"],
           "mean_queue_ms": main["mean_queue_ms"], "throughput_req_per_s": main["throughput_req_per_s"],
           "gate_apps_per_s": main["gate_apps_per_s"], "max_concurrent": main["max_concurrent"],
           "alone_mean_service_ms": round(1e3 * alone_s, 2), "rate_sweep": sweep,
           "rate_sweep_coalescing_off": no_coalescing,
           "plain_per_gate_one_at_a_time": {
               "requests": len(plain), "mean_service_ms": round(1e3 * plain_s / len(plain), 2),
               "throughput_req_per_s": round(len(plain) / plain_s, 2),
               "gate_apps_per_s": round(plain_gates / plain_s, 1), "max_infidelity_checked": worst_plain,
               "fused_one_worker_mean_service_ms": round(1e3 * alone_s, 2),
               "sample": f"the {len(plain)} calibration requests (same distribution), back to back, no queueing"},
           "max_infidelity_checked": worst, "requests_checked": checked, "shots_per_request": 1000,
           "qubits": "log-uniform 12..28", "depth": "U{10..100}"}
    log(f"served: {out}")
    return out


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic_u3_28q.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def probe(args):
    """Short fixed kernel sequence for counter collection (rocprofv3 --pmc):
    every pass kind the benchmarks run, on a state of --probe-qubits qubits.

      copy x3           k_copy16, the calibration of FETCH_SIZE / WRITE_SIZE
      U3 x3             one-qubit passes on target bits n-1, 2 and 0 (tuned shapes)
      CZ x3             high bits (k_phase_v4), bits 0,1 and 2,3 (k_phase_lines)
      brickwork k=5     dense 5-qubit blocks (k_matk_mfma, k_matk)
      brickwork tiled   LDS tile passes (k_tile)
      QFT k=5           MFMA blocks carrying diagonal tables and the folded
                        bit reversal (k_matk_mfma*, k_diagt_v4)
      QFT tiled         tile passes and the restoring bit permutation (k_bitperm)
      sample            1000 shots (k_block_sums, k_scan_chunks, k_scan_top,
                        k_scan_fixup, k_sample)
    """
    hip.init(0)
    n = args.probe_qubits
    a = hip.State(n, hip.C64)
    b = hip.State(n, hip.C64)
    a.gaussian(1)
    a.autotune(reps=1)
    log(f"probe: state of {n} qubits, launch shapes tuned")
    brick = ops.lower_all(circuits.brickwork(n, 2, seed=SEED))
    u3 = [o for o in brick if o["kind"] == hip.OP_MAT1 and o["targets"][0] in (0, n - 3, n - 1)][:3]
    cz = ops.lower_all([circuits.gate("CZ", 0, 1), circuits.gate("CZ", n - 2, n - 1),
                        circuits.gate("CZ", n - 4, n - 3)])
    qft = ops.lower_all(circuits.fourier_transform(n, final_swaps=True))
    schedules = [("brickwork k=5", hip.Schedule(brick, n, 5)),
                 ("brickwork tiled", hip.Schedule(brick, n, mode="tiled")),
                 ("QFT k=5", hip.Schedule(qft, n, 5)),
                 ("QFT tiled", hip.Schedule(qft, n, mode="tiled"))]
    for _ in range(3):
        b.copy_from(a)
    b.sync()
    log(f"probe: copy x3 ({8 << n} B each)")
    a.apply(u3)
    a.apply(cz)
    a.sync()
    log("probe: U3 x3, CZ x3")
    for label, sch in schedules:
        a.apply_schedule(sch)
        a.sync()
        log(f"probe: {label}: {[k for k, _ in sch.passes()]}")
        sch.close()
    a.normalize()
    idx = a.sample(np.random.default_rng(5).uniform(size=1000))
    log(f"probe: sampled {len(idx)} shots; done")


def run(args):
    hip.init(0)
    dev = hip.device_info()
    log(f"device {dev}")
    n = N_QUBITS
    gates = circuits.brickwork(n, DEPTH, seed=SEED)
    lowered = ops.lower_all(gates)
    n_gates = len(lowered)
    alg_bytes = algorithmic_bytes(lowered, n)

    st = hip.State(n, hip.C64)
    t0 = time.perf_counter()
    tuned = st.autotune(reps=3)
    log(f"autotuned one-qubit launch shapes in {time.perf_counter() - t0:.2f} s: {tuned}")

    def step():
        st.set_basis(0)
        st.apply(lowered)

    for _ in range(args.warmup):
        step()
    st.sync()
    log("warmup done")

    st.profile(True)
    st.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    st.sync()
    dt = time.perf_counter() - t0
    prof = st.profile_read()
    st.profile(False)
    log(f"timed {args.steps} steps in {dt:.3f} s; profile {prof}")
    ms_per_step = 1e3 * dt / args.steps
    gate_apps = n_gates * args.steps / dt

    # dominant kernel: the one-qubit U3 pass (MAT1)
    mat1_ms, mat1_launches = prof.get("MAT1", (0.0, 0))
    mat1_avg_ms = mat1_ms / max(1, mat1_launches)
    mat1_bytes = 16.0 * (1 << n)
    achieved = mat1_bytes / (mat1_avg_ms * 1e-3) / 1e9 if mat1_avg_ms > 0 else 0.0
    traffic = load_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic.get("bytes_per_launch") if traffic else None,
                "kernel": "one-qubit U3 pass (k_mat1_v4 / k_mat1_fronts / k_mat1_lx / k_mat1_b0, launch shape tuned per target)",
                "avg_launch_ms": round(mat1_avg_ms, 4),
                "algorithmic_bytes_per_launch": mat1_bytes}

    # FP64 reference of the same circuit (complex128 path) -> infidelity
    ref = hip.State(n, hip.C128)
    ref.set_basis(0)
    ref.apply(ops.lower_all(gates, exact=True))
    st.set_basis(0)
    st.apply(lowered)
    ov = st.inner(ref)
    n64 = st.norm2()
    infid = max(0.0, 1.0 - abs(ov) ** 2 / n64)
    log(f"infidelity plain c64 vs c128: {infid:.3e}")

    # compile(): fused 5-qubit (and smaller) blocks on the same circuit
    fused = {}
    for kmax in (4, 5, "tiled"):
        if kmax == "tiled":
            sch = hip.Schedule(lowered, n, mode="tiled")
        else:
            sch = hip.Schedule(lowered, n, kmax)
        passes, _ = sch.info()
        st.set_basis(0)
        st.apply_schedule(sch)
        st.sync()
        t0 = time.perf_counter()
        reps = max(1, args.steps)
        for _ in range(reps):
            st.set_basis(0)
            st.apply_schedule(sch)
        st.sync()
        fdt = (time.perf_counter() - t0) / reps
        ovf = st.inner(ref)
        inf_f = max(0.0, 1.0 - abs(ovf) ** 2 / st.norm2())
        key = "tiled" if kmax == "tiled" else f"k{kmax}"
        fused[key] = {"passes": passes, "ms_per_step": round(1e3 * fdt, 3),
                      "gate_apps_per_s": round(n_gates / fdt, 1), "infidelity": inf_f,
                      "speedup_vs_plain": round((dt / args.steps) / fdt, 3), "model_cost": round(sch.cost(), 2)}
        log(f"fused {key}: {fused[key]}")
        sch.close()
    ref.close()
    st.close()

    batch = batched(args) if not args.no_batch else None
    r20 = random20(args)
    qft = qft30(args) if not args.no_qft else None
    iqft = qft30(args, inverse=True) if not args.no_qft else None
    qft_check = qft_closed_form() if not args.no_qft else None
    serve = served(args) if args.serve_requests > 0 else None
    algos = reference_algorithms(args) if args.algo_reps > 0 else None

    cpu = cpu_baseline(gates, n) if not args.no_cpu else None
    if cpu:
        log(f"cpu baseline {cpu}")

    out = {
        "metric": "gate-apps/s + HBM GB/s at 28q c64 vs plain; served p50/p99 at <=1e-5 infidelity",
        "value": round(gate_apps, 1),
        "unit": "gate-apps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded brickwork circuit, angles ~ U[0, 2pi))",
        "config": {"workload": "brickwork-28q-depth20 (U3 per qubit + CZ alternating pairs), plain per-gate",
                   "qubits": n, "depth": DEPTH, "gates": n_gates, "state_bytes": 8 << n,
                   "algorithmic_bytes_per_step": alg_bytes, "seed": SEED},
        "hbm_gbs_per_step": round(alg_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "infidelity_vs_fp64": infid,
        "roofline": roofline,
        "fused": fused,
        "batched": batch,
        "random20": r20,
        "qft30": qft,
        "iqft30": iqft,
        "qft_closed_form": qft_check,
        "served": serve,
        "reference_algorithms": algos,
        "cpu_baseline": cpu,
        "device": dev,
    }
    print(json.dumps(out), flush=True)


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--no-batch", action="store_true", help="skip the batched-states section")
    p.add_argument("--no-qft", action="store_true", help="skip the 30-qubit QFT section")
    p.add_argument("--serve-requests", type=int, default=120,
                   help="requests in the served-stream section (0 skips it)")
    p.add_argument("--algo-reps", type=int, default=100,
                   help="repetitions of the reference's algorithm benchmarks (0 skips them)")
    p.add_argument("--probe", action="store_true",
                   help="run a short fixed kernel sequence for hardware-counter collection and exit")
    p.add_argument("--probe-qubits", type=int, default=N_QUBITS, help="state size of the --probe sequence")
    args = p.parse_args()
    if args.probe:
        probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the state-vector path here runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
