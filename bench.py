"""Benchmark: WHIR commit + prove of one 2^24-variable KoalaBear polynomial on one MI355X.

Workload (the single-GPU configuration of the headline metric): num_variables
24, folding factor 4, log inverse rate 1 (rate 1/2, 2^25-entry codeword), RS
domain initial reduction 3, 100-bit security with 16 PoW bits, CapacityBound,
Poseidon2 width-16 Merkle tree, two seeded extension-field evaluation
constraints.  The polynomial's 2^24 coefficients are drawn from SmallRng seed 0
and uploaded once, and the two constraints are evaluated once before the timed
steps, as the reference's benchmark does in its setup (benches/whir.rs:103-104);
every timed step clones that statement and runs commit and prove from the
device-resident input, as the reference's "prove" benchmark does
(benches/whir.rs:157-183), and returns the proof bytes to the host.

Prints one JSON line.  --gpus N > 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

# MI355X peaks: HBM3E bandwidth, and the VALU issue rate: 256 CUs x 4 SIMDs x
# 16 lanes x 2.4 GHz = 39.3 T lane-instructions/s.  On gfx950 the instructions
# the hashing kernels are made of (32-bit add/min, v_mul_lo_u32,
# v_mad_u64_u32, v_mad_i64_i32, v_lshl_add_u64) all issue at that rate
# (measured with long independent chains of each).
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 39.3
# VALU instructions per width-16 Poseidon2 permutation in the gfx950 build:
# 4485 in the compiled grinding kernel's loop (one permutation per candidate
# plus a few instructions of candidate setup), down from 7496 for the textbook
# form (one modular reduction per operation).  Hashing stages are issue-bound,
# so permutations x 4485 / time against the issue rate is their roofline.
VALU_OPS_PER_PERM = 4485
# A grinding candidate costs 123 fewer (4331 against 4454 in the lockstep
# kernel's compiled loop): its initial linear layer is a precomputed vector
# plus the witness times one column, and only element 7 of the last layer is
# formed.
VALU_OPS_PER_CANDIDATE = 4485 - 123
VALU_OPS = {"merkle_leaves": VALU_OPS_PER_PERM, "merkle_tree": VALU_OPS_PER_PERM, "grind": VALU_OPS_PER_CANDIDATE}
VALU_STAGES = tuple(VALU_OPS)
# Kernels of each stage, for the HBM traffic measured with rocprofv3 --pmc
# (FETCH_SIZE and WRITE_SIZE per dispatch, calibrated on a copy of known size).
STAGE_KERNELS = {   # per stage: the kernels one launch group runs, each a list of alternatives
    "grind": [["k_grind", "k_grind_wide", "k_grind_lock<512>"]],
    "merkle_leaves": [["k_leaf_hash<16>"]],
    "lde_base": [["k_lde_pass1<unsigned int, 10>"], ["k_lde_pass2<unsigned int, 10, false, true>"]],
}
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic_n24.json")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay on their cores (read when the oracle loads)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")


def progress(msg):
    """A progress line on stderr (the JSON result alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def traffic_summary(fetch_csv, write_csv):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE; KiB per
    dispatch), scaled by what each counter reads for the 256 MiB calibration copy, into TRAFFIC_FILE."""
    import csv

    def load(path, counter):
        per = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("kb::", "").strip()
                per.setdefault(name, []).append(float(r["Counter_Value"]))
        return per

    fetch, write = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    cal = "__amd_rocclr_copyBuffer"
    moved_kib = 256 * 1024
    f_cal, w_cal = max(fetch[cal]), max(write[cal])   # the calibration copy is the largest copy dispatch
    ff, wf = moved_kib / f_cal, moved_kib / w_cal
    kernels = {}
    for name in fetch:
        if name == cal or name not in write:
            continue
        fk = sum(fetch[name]) / len(fetch[name])
        wk = sum(write[name]) / len(write[name])
        kernels[name] = {"dispatches": len(fetch[name]), "fetch_kib_raw": round(fk, 1), "write_kib_raw": round(wk, 1),
                         "fetch_bytes": round(fk * 1024 * ff), "write_bytes": round(wk * 1024 * wf)}
    out = {"workload": "WHIR commit+prove n=24 (bench.py --steps 1 --warmup 1 --copy-calibration)",
           "calibration": {"kernel": cal + ", 256 MiB device to device", "bytes": 256 << 20,
                           "fetch_kib_read": f_cal, "write_kib_read": w_cal, "fetch_factor": ff, "write_factor": wf},
           "kernels": dict(sorted(kernels.items(), key=lambda kv: -(kv[1]["fetch_bytes"] + kv[1]["write_bytes"])))}
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(out, f, indent=1)
    print(f"fetch factor {ff:.3f}, write factor {wf:.3f}")
    print("| kernel | dispatches | FETCH raw KiB | WRITE raw KiB | fetch MiB (scaled) | write MiB (scaled) |")
    print("|---|---:|---:|---:|---:|---:|")
    for name, k in out["kernels"].items():
        print(f"| `{name}` | {k['dispatches']} | {k['fetch_kib_raw']} | {k['write_kib_raw']} | "
              f"{k['fetch_bytes'] / 2**20:.2f} | {k['write_bytes'] / 2**20:.2f} |")


def hbm_roof(stage, name):
    if not stage["launches"] or stage["ms"] <= 0:
        return None
    avg_s = stage["ms"] / stage["launches"] * 1e-3
    per = stage["work"] / stage["launches"]
    gbs = per / avg_s / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": round(per),
            "avg_launch_ms": round(stage["ms"] / stage["launches"], 4), "launches": stage["launches"]}


def sumcheck_eq_bench(W, args, n=26, l0=3, reps=3):
    """eq(z, .) expansion over 26 extension variables (2^26 EF4 = 1 GiB) and a
    26-round sumcheck of f * eq(z, .), f uniform over KoalaBear, the first 3
    rounds by the split (SVO) route, challenges from a duplex challenger."""
    import ctypes
    cfg = W.WhirConfig(W.params(n), device=0)
    f = W.sample_field(10, 1 << n)
    z = W.sample_field(11, 4 * n)
    d_f = W.DeviceBuffer.from_array(f)
    d_eq = W.DeviceBuffer(16 << n)
    one = np.array([0x01FFFFFE, 0, 0, 0], np.uint32)   # 1 in Montgomery form
    polys = np.zeros(8 * n, np.uint32)
    fin = np.zeros(8, np.uint32)

    def run():
        W.check(W.lib().whirgpu_eq_expand(cfg.handle, W.u32p(z), n, W.u32p(one), d_eq.ptr))
        ch = cfg.challenger(observe_domain_separator=False)
        W.check(W.lib().whirgpu_sumcheck_eq(cfg.handle, d_f.ptr, n, W.u32p(z), l0, None, ctypes.byref(ch),
                                            W.u32p(polys), W.u32p(fin)))

    run()
    cfg.profile(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    wall = (time.perf_counter() - t0) * 1e3 / reps
    recs = cfg.profile_records()
    st = cfg.profile_read()
    cfg.profile(False)

    def large(stage, min_bytes=64 << 20):
        """The stage's launch groups moving at least min_bytes (the rounds whose
        time is bandwidth, not launch latency), against the HBM peak."""
        sel = [(ms, w) for s_, ms, w in recs if s_ == stage and w >= min_bytes]
        if not sel:
            return None
        ms = sum(m for m, _ in sel)
        by = sum(w for _, w in sel)
        gbs = by / (ms * 1e-3) / 1e9
        return {"launches": len(sel), "bytes": round(by), "ms": round(ms, 4), "achieved": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    sc_ms = sum(st[k]["ms"] for k in ("sumcheck", "eq_dot", "combine", "svo_fold")) / reps
    out = {
        "workload": "eq expansion 2^26 EF4 + 26-round sumcheck of f*eq (f base field 2^26), 3 SVO rounds",
        "eq_expand": hbm_roof(st["eq_expand"], "eq_expand"),
        "sumcheck_rounds": hbm_roof(st["sumcheck"], "sumcheck fold+coefficients"),
        "sumcheck_rounds_64MiB_plus": large("sumcheck"),
        "svo_partials": hbm_roof(st["eq_dot"], "SVO partial sums"),
        "svo_fold": hbm_roof(st["svo_fold"], "SVO 3-variable fold"),
        "weights": hbm_roof(st["combine"], "eq weights after SVO"),
        "sumcheck_device_ms": round(sc_ms, 3),
        "eq_expand_ms": round(st["eq_expand"]["ms"] / reps, 3),
        "wall_ms_per_rep": round(wall, 3),
    }
    d_f.free()
    d_eq.free()
    cfg.close()
    return out


def batched_bench(W, workers=12, count=64, seed=2024, group_max=16):
    """64 proofs with num_variables uniform in {16..22}, one constraint each,
    polynomials resident in HBM: proofs/s one at a time (one context per size,
    proofs back to back; median of five passes) against proofs/s through the
    batching server, all jobs submitted at once so that same-size jobs are
    proved as groups (median of five batches; host CPU time of the process per
    batched proof)."""
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in rng.integers(16, 23, size=count)]
    bufs = [W.DeviceBuffer.from_array(W.sample_field(2000 + i, 1 << n)) for i, n in enumerate(ns)]
    pts = [[W.sample_field(3000 + i, 4 * n).reshape(n, 4)] for i, n in enumerate(ns)]
    cfgs = {n: W.WhirConfig(W.params(n), device=0) for n in sorted(set(ns))}
    tuned = {n: W.autotune.apply_cached(c) is not None for n, c in cfgs.items()}   # the same shapes both legs
    for i, n in enumerate(ns):   # warm every context once
        if n in cfgs and not getattr(cfgs[n], "_warm", False):
            W.prove_once(cfgs[n], None, pts[i], device_ptr=bufs[i].ptr)
            cfgs[n]._warm = True
    seq_times = []
    for _ in range(5):            # median of five passes, as for the batches
        t = time.perf_counter()
        seq = [W.prove_once(cfgs[n], None, pts[i], device_ptr=bufs[i].ptr)[0] for i, n in enumerate(ns)]
        seq_times.append(time.perf_counter() - t)
    seq_s = sorted(seq_times)[len(seq_times) // 2]
    for c in cfgs.values():
        c.close()
    srv = W.Server(W.params(0), workers=workers, group_max=group_max, tuning_cache=W.autotune.DEFAULT_CACHE)

    def run_batch():
        reqs = [W.ProofRequest(None, pts[i], device_ptr=bufs[i].ptr, num_variables=n) for i, n in enumerate(ns)]
        t0 = time.perf_counter()
        srv.submit_many(reqs)
        for r in reqs:
            assert srv.wait(r) == 0, W.lib().whirgpu_last_error()
        return reqs, time.perf_counter() - t0

    run_batch()                   # warm the workers' contexts
    times = []
    cpu0 = time.process_time()
    for _ in range(5):            # median of five batches
        reqs, s = run_batch()
        times.append(s)
    cpu_ms = (time.process_time() - cpu0) * 1e3 / (5 * count)
    bat_s = sorted(times)[len(times) // 2]
    srv.close()
    same = all(r.proof() == s for r, s in zip(reqs, seq))
    lat = sorted(r.job.latency_ms for r in reqs)
    for b in bufs:
        b.free()
    return {"workload": f"{count} proofs, num_variables uniform in {{16..22}} (seed {seed}), 1 constraint each, "
                        f"folding 4, rate 1/2, rs-reduction 3, lambda 100",
            "one_at_a_time_proofs_per_s": round(count / seq_s, 2), "batched_proofs_per_s": round(count / bat_s, 2),
            "speedup": round(seq_s / bat_s, 2), "workers": workers, "group_max": group_max,
            "tuned_sizes": sorted(n for n, t in tuned.items() if t),
            "bit_identical_to_one_at_a_time": same,
            "passes_timed": len(seq_times), "batches_timed": len(times), "host_cpu_ms_per_proof": round(cpu_ms, 3),
            "batch_latency_ms_p50": round(lat[len(lat) // 2], 3), "batch_latency_ms_max": round(lat[-1], 3)}


def serving_bench(W, workers=8, rate=200.0, count=200, seed=7):
    """Poisson arrivals at `rate` req/s, num_variables log-uniform in [16, 24],
    1-4 constraints each, host-resident inputs uploaded per request, admission
    by free HBM; latency is submit to proof bytes on the host."""
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in np.rint(np.exp(rng.uniform(np.log(16), np.log(24), count)))]
    ks = [int(x) for x in rng.integers(1, 5, count)]
    gaps = rng.exponential(1.0 / rate, count)
    polys = [W.sample_field(5000 + i, 1 << n) for i, n in enumerate(ns)]
    pts = [[W.sample_field(6000 + 10 * i + j, 4 * n).reshape(n, 4) for j in range(k)]
           for i, (n, k) in enumerate(zip(ns, ks))]
    srv = W.Server(W.params(0), workers=workers, tuning_cache=W.autotune.DEFAULT_CACHE)
    warm = [W.ProofRequest(polys[i], pts[i]) for i in range(count)]
    for r in warm:
        srv.submit(r)
    for r in warm:
        assert srv.wait(r) == 0
    reqs = [W.ProofRequest(polys[i], pts[i]) for i in range(count)]
    sub = []
    t0 = time.perf_counter()
    due = t0
    for r, g in zip(reqs, gaps):
        due += g
        d = due - time.perf_counter()
        if d > 0:
            time.sleep(d)
        sub.append(time.perf_counter())
        srv.submit(r)
    for r in reqs:
        assert srv.wait(r) == 0
    srv.close()
    lat = np.array([r.job.latency_ms for r in reqs])
    done = max(s + r.job.latency_ms / 1e3 for s, r in zip(sub, reqs))
    same = all(a.proof() == b.proof() for a, b in zip(warm, reqs))
    return {"workload": f"Poisson {rate:.0f} req/s, {count} requests, num_variables log-uniform in [16,24] "
                        f"(seed {seed}), 1-4 constraints, host inputs, admission by free HBM",
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "proofs_per_s": round(count / (done - t0), 2), "offered_rate": rate, "workers": workers,
            "mean_queue_ms": round(float(np.mean([r.job.queue_ms for r in reqs])), 3),
            "same_alone_and_in_mix": same}


def timed_proofs(W, cfg, d_ptr, pts, steps, warmup=1):
    """ms per commit + prove, as the headline step: the statement's constraints
    evaluated once, cloned per step, proof bytes returned to the host."""
    st0 = W.InitialStatement(cfg, None, device_ptr=d_ptr)
    for p in pts:
        st0.evaluate(p)
    times = []
    for i in range(warmup + steps):
        t = time.perf_counter()
        st = st0.clone()
        ch = cfg.challenger()
        proof = W.WhirProof(cfg)
        pd = W.CommitmentWriter(cfg).commit(proof, ch, st)
        W.Prover(cfg).prove(proof, ch, st, pd)
        data = proof.to_bytes()
        proof.close()
        st.close()
        if i >= warmup:
            times.append((time.perf_counter() - t) * 1e3)
    st0.close()
    return float(np.median(times)), data


def schedule_bench(W, n, d_ptr, pts, benchmark_ms, steps=5):
    """The autotuner's cached best folding schedule for this n (same security
    level and PoW bits as the benchmark), proved now beside the benchmark's
    Constant(4) schedule.  A different schedule is a different proof."""
    dev = W.autotune.load_cache()["devices"].get(W.device_name(0))
    key = f"n={n},lambda=100,pow=16,CapacityBound,rate=1"
    if not dev or key not in dev["schedules"]:
        return None
    best = dev["schedules"][key]["best"]
    prm = W.params(n, folding_factor=best["folding_factor"], folding_factor_first=best["folding_factor_first"],
                   rs_domain_initial_reduction_factor=best["rs_reduction"])
    cfg = W.WhirConfig(prm, device=0)
    W.autotune.apply_cached(cfg)
    ms, data = timed_proofs(W, cfg, d_ptr, pts, steps)
    cfg.close()
    return {"schedule": {k: best[k] for k in ("folding_factor", "folding_factor_first", "rs_reduction", "rounds")},
            "prove_ms": round(ms, 3), "proof_bytes": len(data), "benchmark_schedule_prove_ms": round(benchmark_ms, 3),
            "ratio": round(benchmark_ms / ms, 3),
            "note": "same security level and PoW bits; the benchmark schedule is the reference's Constant(4), rs 3"}


def verifier_bench(W, cfg, proof, pts, evals, runs=7):
    """The product's host verifier (Verifier::verify after parse_commitment,
    one thread) on the headline proof: decode + verify, median of `runs`."""
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        pr = W.WhirProof.from_bytes(cfg, proof)
        W.Verifier(cfg).verify(pr, cfg.challenger(), pts, evals)
        times.append((time.perf_counter() - t) * 1e3)
        pr.close()
    times.sort()
    return {"what": "decode + verify the n=24 proof on the host (1 thread)", "ms": round(times[len(times) // 2], 3),
            "runs": runs, "proof_bytes": len(proof)}


def keccak_bench(W, n, d_ptr, pts, benchmark_ms, steps=5):
    """The Keccak MMCS variant (src/whir/mod.rs:236-383: Keccak Merkle tree,
    Keccak-256 byte challenger) at the benchmark's parameters: commit + prove
    and the host verifier's decode + verify, beside the Poseidon2 benchmark."""
    cfg = W.WhirConfig(W.params(n, mmcs="keccak"), device=0)
    ms, data = timed_proofs(W, cfg, d_ptr, pts, steps)
    st = W.InitialStatement(cfg, None, device_ptr=d_ptr)
    evals = np.array([st.evaluate(p) for p in pts], dtype=np.uint32).reshape(len(pts), 4)
    st.close()
    times = []
    for _ in range(5):
        t = time.perf_counter()
        W.Verifier(cfg).verify(data, cfg.challenger(), pts, evals)
        times.append((time.perf_counter() - t) * 1e3)
    cfg.close()
    return {"prove_ms": round(ms, 3), "verify_ms": round(float(np.median(times)), 3), "proof_bytes": len(data),
            "poseidon2_prove_ms": round(benchmark_ms, 3),
            "note": "same parameters as the benchmark, Keccak Merkle tree and byte challenger; transcript host-driven"}


def cpu_baseline(n, poly, pts, gpu_proof, runs=3):
    """The same commit + prove on the CPU oracle (a naive port: canonical 64-bit
    modular arithmetic, OpenMP over Merkle leaves, DFT columns and grinding
    candidates), threads pinned to cores, one warm-up run, then the median of
    `runs` runs, with per-stage medians under the reference's span names."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import ORACLE_SPANS, oracle_params, oracle_prove
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cproof, _, _ = oracle_prove(oracle_params(n), poly, pts)   # warm-up
    totals, stages = [], []
    for _ in range(runs):
        _, _, ms = oracle_prove(oracle_params(n), poly, pts)
        totals.append(ms[0] + ms[1])
        stages.append(ms[2:])
    med = float(np.median(totals))
    per = {name: round(float(np.median([s_[i] for s_ in stages])), 1) for i, name in enumerate(ORACLE_SPANS)}
    return {"value": round(med, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"full n={n} commit+prove of the same inputs on the CPU oracle (naive port: canonical 64-bit "
                      f"% arithmetic), OMP threads pinned to cores, 1 warm-up + median of {runs} runs",
            "runs_ms": [round(t, 1) for t in totals], "stage_ms": per, "bit_exact": cproof == gpu_proof}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3, help="timed oracle runs after one warm-up (median reported)")
    ap.add_argument("--skip-kernel-bench", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the n-variable commit+prove steps (no verifier, schedule, Keccak, 2^26, batched or "
                         "serving workloads): the run the committed kernel profiles cover")
    ap.add_argument("--workers", type=int, default=12, help="batching server workers")
    ap.add_argument("--group-max", type=int, default=16, help="largest group a batching server worker forms")
    ap.add_argument("--serving-workers", type=int, default=8)
    ap.add_argument("--only-kernel-bench", action="store_true",
                    help="run only the 2^26 eq-expansion + sumcheck workload and print its JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device to device (reference dispatch for rocprofv3 --pmc byte counters)")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the per-dispatch HBM traffic file from the counter_collection.csv of a FETCH_SIZE "
                         "and a WRITE_SIZE pass over --copy-calibration runs, print its table, and exit")
    args = ap.parse_args()
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge

    W = ge._package()
    if args.copy_calibration:
        src, dst = W.DeviceBuffer(256 << 20), W.DeviceBuffer(256 << 20)
        for _ in range(2):
            W.check(W.lib().whirgpu_memcpy(0, W.ctypes.c_void_p(dst.ptr), W.ctypes.c_void_p(src.ptr), 256 << 20, 2))
        del src, dst
    if args.only_kernel_bench:
        print(json.dumps({"sumcheck_eq_2^26": sumcheck_eq_bench(W, args)}))
        return
    traffic = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("kernels", {})
    n = args.n
    prm = W.params(n)
    cfg = W.WhirConfig(prm, device=0)
    tuning = W.autotune.apply_cached(cfg)   # launch shapes only: the proof bytes do not change
    poly = W.sample_field(0, 1 << n)
    pts = [W.sample_field(1 + j, 4 * n).reshape(n, 4) for j in range(2)]
    d_poly = W.DeviceBuffer.from_array(poly)

    st0 = W.InitialStatement(cfg, None, device_ptr=d_poly.ptr)
    evals0 = np.array([st0.evaluate(p) for p in pts], dtype=np.uint32).reshape(len(pts), 4)

    def step():
        st = st0.clone()
        ch = cfg.challenger()
        proof = W.WhirProof(cfg)
        pd = W.CommitmentWriter(cfg).commit(proof, ch, st)
        W.Prover(cfg).prove(proof, ch, st, pd)
        out = proof.to_bytes()
        proof.close()
        st.close()
        return out

    progress(f"setup done (n={n}); warm-up")
    for _ in range(args.warmup):
        proof = step()
    # Timed steps: no profiling events, so `value` is the plain proof time.
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        proof = step()
        times.append((time.perf_counter() - s) * 1e3)
    total_ms = (time.perf_counter() - t0) * 1e3
    ms = total_ms / args.steps
    # The same number of steps again with HIP events recorded on the library's
    # stream around every kernel group: per-stage device time for the rooflines.
    cfg.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        proof = step()
    ms_profiled = (time.perf_counter() - t0) * 1e3 / args.steps
    stages = cfg.profile_read()
    cfg.profile(False)

    def stage_traffic(name):
        """HBM bytes per launch of the stage: per-dispatch traffic of its kernels
        (a launch group runs each once; alternatives such as the one-lane and
        lane-parallel grinds are averaged by their dispatch counts)."""
        groups = STAGE_KERNELS.get(name)
        if not groups:
            return None
        total = 0.0
        for alts in groups:
            alts = [k for k in alts if k in traffic]
            if not alts:
                return None
            n = sum(traffic[k]["dispatches"] for k in alts)
            total += sum((traffic[k]["fetch_bytes"] + traffic[k]["write_bytes"]) * traffic[k]["dispatches"]
                         for k in alts) / n
        return round(total)

    def roof(name):
        s = stages[name]
        if not s["launches"] or s["ms"] <= 0:
            return None
        avg_s = s["ms"] / s["launches"] * 1e-3
        if name in VALU_STAGES:
            per_launch = s["work"] / s["launches"] * VALU_OPS[name]
            achieved = per_launch / avg_s / 1e12
            return {"bound": "valu", "kernel": name, "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS,
                    "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": stage_traffic(name),
                    "work_per_launch": round(s["work"] / s["launches"]), "work_unit": "Poseidon2-16 permutations",
                    "ops_per_unit": VALU_OPS[name], "avg_launch_ms": round(s["ms"] / s["launches"], 4)}
        per_launch = s["work"] / s["launches"]
        achieved = per_launch / avg_s / 1e9
        return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": stage_traffic(name),
                "bytes_per_launch": round(per_launch), "avg_launch_ms": round(s["ms"] / s["launches"], 4)}

    dominant = max(stages, key=lambda k: stages[k]["ms"])
    verifier = schedule = keccak = None
    if not args.headline_only:
        verifier = verifier_bench(W, cfg, proof, pts, evals0)
        schedule = schedule_bench(W, n, d_poly.ptr, pts, ms)
        keccak = keccak_bench(W, n, d_poly.ptr, pts, ms)
        progress(f"keccak variant: {keccak['prove_ms']} ms per proof")

    progress(f"n={n}: {ms:.3f} ms per proof; profiled steps done")
    kernels = None if args.skip_kernel_bench or args.headline_only else sumcheck_eq_bench(W, args)
    if kernels:
        progress("2^26 sumcheck/eq workload done")
    batched = None if args.skip_batch or args.headline_only else batched_bench(W, workers=args.workers,
                                                                                group_max=args.group_max)
    if batched:
        progress(f"batched: {batched['speedup']}x")
    serving = None if args.skip_batch or args.headline_only else serving_bench(W, workers=args.serving_workers)
    if serving:
        progress(f"serving: p99 {serving['p99_ms']} ms")

    cpu = None
    if not args.no_cpu_baseline:
        progress("CPU baseline")
        cpu = cpu_baseline(n, poly, pts, proof, runs=args.cpu_runs)

    out = {
        "metric": "prove ms @2^24 vars" if n == 24 else f"prove ms @2^{n} vars",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (KoalaBear, Montgomery form)",
        "data": "synthetic (SmallRng seed 0 coefficients, seeds 1-2 points)",
        "config": {"workload": f"WHIR commit+prove n={n} KoalaBear/EF4, folding 4, rate 1/2, rs-reduction 3, "
                               f"lambda 100, pow 16, CapacityBound, Poseidon2-16 Merkle, 2 eval constraints",
                   "num_variables": n, "proofs_per_step": 1},
        "proofs_per_s": round(1e3 / ms, 2),
        "step_ms": [round(t, 3) for t in times],
        "ms_per_step_profiled": round(ms_profiled, 3),
        "proof_bytes": len(proof),
        "roofline": roof(dominant),
        "roofline_by_stage": {k: roof(k) for k in stages if stages[k]["launches"]},
        "stage_ms_per_step": {k: round(v["ms"] / args.steps, 4) for k, v in stages.items()},
        "cpu_baseline": cpu,
        "verifier": verifier,
        "launch_shapes": "autotuned (cache)" if tuning is not None else "built-in",
        "schedule_autotune": schedule,
        "keccak_mmcs": keccak,
        "sumcheck_eq_2^26": kernels,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
