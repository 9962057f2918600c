"""Benchmark of the RNS key-switching hot path on one MI355X.

A step is one batched relinearization of 512 degree-2 BGV ciphertexts at
n = 2^14 with L = 8 primes (the largest below 2^50 with q = 1 mod 2^15) and
gadget base 2^10 (k = 40 digits), with the relinearization key resident in
HBM.  Inputs are synthetic seeded residues, already in HBM when the timed
region starts.  The line also reports NTTs/s at (n = 2^14, L = 8), the plain
per-call key-switch rate, Galois key-switch and modulus-switch rates, the
per-step latency p50/p99, the roofline of the dominant kernel, and the CPU
oracle timed on this host as a reported baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shell-encryption_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
FP64_OPS_PEAK_T = 39.3  # 78.6 TFLOP/s FP64 vector, an FMA counted as two flops
# HBM bytes per launch of the key-switch kernel, measured by rocprofv3
# FETCH_SIZE + WRITE_SIZE and scaled by the same counters' reading for a
# streaming kernel of known size (bench.py --copy-probe).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "keyswitch_relin_n16384_L8_b512_pmc.md")
LOG_N = 14
L = 8
LOG_B = 10
BATCH = 512
NTT_BATCH = 1024
# Untimed forward + inverse pairs before each timed NTT measurement.
NTT_WARMUP = 50


def primes_for(log_n, count, log_bound=50):
    from common import ntt_primes
    return ntt_primes(log_n, count, log_bound)


def uniform(rng, qs, shape_prefix, n):
    """Residues uniform in [0, q_i) for limb i; shape prefix + (L, n)."""
    out = np.empty(tuple(shape_prefix) + (len(qs), n), dtype=np.uint64)
    for i, q in enumerate(qs):
        out[..., i, :] = rng.integers(0, q, size=tuple(shape_prefix) + (n,), dtype=np.uint64)
    return out


def percentile(v, p):
    return float(np.percentile(np.asarray(v), p))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Worker threads for the CPU baseline: the cores this process may use,
    at most 16 (the host share of one GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def _timed(fn, items, threads, budget_s, cap):
    """Runs fn(item) over items (cycled) on `threads` host threads until
    budget_s elapses or `cap` calls are done; returns (calls, seconds)."""
    import threading
    lock = threading.Lock()
    state = {"next": 0}
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                i = state["next"]
                if i >= cap or time.perf_counter() - t0 >= budget_s:
                    return
                state["next"] = i + 1
            fn(items[i % len(items)])

    th = [threading.Thread(target=worker) for _ in range(threads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return state["next"], time.perf_counter() - t0


def cpu_isa():
    """The widest vector extension the host reports (the vectorised
    baseline's clones dispatch on it)."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((l for l in f if l.startswith("flags")), "")
    except OSError:
        return "unknown"
    for isa in ("avx512f", "avx2"):
        if " %s" % isa in flags:
            return isa
    return "x86-64"


def cpu_baseline(qs, comps, kas, kbs, log_bs, ntt_inputs):
    """The oracle's C restatements of the reference's RnsRelinKey::ApplyTo
    and forward transform on this host (their calls release the interpreter
    lock), in two forms: scalar (IterativeCooleyTukey and scalar
    FusedMulAddInPlace) and vectorised (the Highway transform,
    dft_transformations_hwy.cc:64-158, with eight-lane FusedMulAddInPlace).
    Each on one thread and on every available core (independent ciphertexts
    per thread, as the reference has no threading of its own).  Bounded
    samples of a few seconds each; `value` is the fastest configuration."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rns = oracle.Rns(qs, LOG_N)
    threads = host_threads()
    res = {}
    for form, vec in (("scalar", False), ("vectorised", True)):
        relin = lambda c: rns.relin_apply(c, kas, kbs, log_bs, vectorised=vec)  # noqa: E731
        ntt = (lambda x: rns.hwy_ntt_fwd(x)) if vec else (lambda x: rns.ntt_fwd(x))
        r1, s1 = _timed(relin, comps, 1, 4.0, 48)
        rt, st = _timed(relin, comps, threads, 4.0, 48 * threads)
        n1, ns1 = _timed(ntt, ntt_inputs, 1, 2.0, 400)
        nt, nst = _timed(ntt, ntt_inputs, threads, 2.0, 400 * threads)
        res[form] = {"key_switches_per_s": {"threads_1": round(r1 / s1, 2),
                                            "threads_%d" % threads: round(rt / st, 2)},
                     "ntts_per_s": {"threads_1": round(n1 / ns1, 1),
                                    "threads_%d" % threads: round(nt / nst, 1)},
                     "sample": "%d relinearizations in %.1f s on 1 thread, %d in %.1f s on %d "
                               "threads; %d forward NTTs in %.1f s on 1 thread, %d in %.1f s on "
                               "%d threads" % (r1, s1, rt, st, threads, n1, ns1, nt, nst, threads)}
    best = max(res, key=lambda f: res[f]["key_switches_per_s"]["threads_%d" % threads])
    return {"value": res[best]["key_switches_per_s"]["threads_%d" % threads],
            "unit": "key-switches/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "isa": cpu_isa(), "form": best,
            "scalar": res["scalar"], "vectorised": res["vectorised"],
            "sample": "oracle relinearizations of degree-2 ciphertexts at n=2^14, L=8, k=40 and "
                      "forward NTTs of RNS polynomials at (2^14, 8), scalar and vectorised, 1 "
                      "thread and %d threads, about 4 s (relinearization) and 2 s (NTT) each" %
                      threads}


def measured_traffic():
    """traffic_bytes_per_launch from the committed counter summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def stream_copy_gbs():
    """Read+write bandwidth of a plain streaming kernel (torch bitwise_not
    over 2^28 int64 words), measured by `bench.py --copy-probe` in a child
    process (torch's HIP runtime and this library's are kept apart): the
    practical ceiling next to the 8 TB/s peak.  None if the child fails."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-probe", "--copy-words",
                            str(1 << 28)], capture_output=True, text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1])["gbs"]
    except Exception:  # a missing figure must not fail the bench line
        return None


def ntt_probe():
    """Only the FP64 transforms of the NTT leg's shape (1024 RNS polynomials
    at n=2^14, L=8): NTT_WARMUP forward + inverse pairs, then four forward
    and four inverse launches (the workload of the NTT profiles; their last
    four dispatches of each kind are the measured ones), then the modulus
    switch of the same polynomials, warmed up the same way."""
    import shell_gpu as sg
    qs = primes_for(LOG_N, L)
    rng = np.random.default_rng(20261203)
    ctx = sg.Context(0)
    pids = ctx.primes(qs, LOG_N)
    d_x = ctx.array(uniform(rng, qs, (NTT_BATCH,), 1 << LOG_N))
    for _ in range(NTT_WARMUP):
        ctx.ntt_fwd(d_x, NTT_BATCH, pids, sg.NTT_F64)
        ctx.ntt_inv(d_x, NTT_BATCH, pids, sg.NTT_F64)
    for _ in range(4):
        ctx.ntt_fwd(d_x, NTT_BATCH, pids, sg.NTT_F64)
    for _ in range(4):
        ctx.ntt_inv(d_x, NTT_BATCH, pids, sg.NTT_F64)
    ql_inv = ctx.ql_inv(pids)
    for _ in range(NTT_WARMUP // 2 + 4):
        ctx.mod_reduce_lsb(d_x, 65537, ql_inv, True, NTT_BATCH, pids)
    ctx.sync()
    print(json.dumps({"ntt_probe": "forward and inverse FP64, then ModReduceLsb",
                      "polys": NTT_BATCH, "n": 1 << LOG_N, "L": L,
                      "launches": {"ntt_fwd": NTT_WARMUP + 4, "ntt_inv": NTT_WARMUP + 4,
                                   "mod_reduce_lsb": NTT_WARMUP // 2 + 4}}))
    return 0


def copy_probe(words):
    """A streaming kernel of known size for scaling HBM byte counters:
    bitwise_not over `words` int64 words (default 2^25: 256 MiB read and 256
    MiB written per launch), five timed launches after one warm-up."""
    import torch
    a = torch.arange(words, dtype=torch.int64, device="cuda")
    b = torch.empty_like(a)
    torch.bitwise_not(a, out=b)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(5):
        torch.bitwise_not(a, out=b)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / 5
    print(json.dumps({"copy_probe": "bitwise_not int64", "read_bytes": a.numel() * 8,
                      "write_bytes": b.numel() * 8, "launches": 6,
                      "gbs": round(2 * words * 8 / (ms * 1e-3) / 1e9, 1)}))
    return 0


UNEVEN_POLYS = 4096


def uneven_batch(ctx, rng):
    """The uneven-batching workload: 4096 polynomials with n uniform in
    {2^11..2^15} and L uniform in 1..12 (the L largest primes below 2^50 with
    q = 1 mod 2n), coefficients uniform mod each prime; forward NTT, pointwise
    multiply and inverse NTT each as one persistent launch over all limbs,
    the chain captured in a HIP graph, the modulus switch of every polynomial
    with L >= 2 as one launch; beside the plain form of one call per
    polynomial and operation."""
    import shell_gpu as sg
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    print("[bench] uneven batch: generating", file=sys.stderr, flush=True)
    log_ns = rng.integers(11, 16, size=UNEVEN_POLYS)
    Ls = rng.integers(1, 13, size=UNEVEN_POLYS)
    primes = {ln: primes_for(ln, 12) for ln in range(11, 16)}
    pids = {ln: ctx.primes(primes[ln], ln) for ln in range(11, 16)}
    sizes = [int(L) << int(ln) for ln, L in zip(log_ns, Ls)]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    total = int(offs[-1])
    host = np.empty(total, dtype=np.uint64)
    other = np.empty(total, dtype=np.uint64)
    for ln in range(11, 16):
        n = 1 << ln
        for i, q in enumerate(primes[ln]):
            rows = [int(offs[p]) + i * n for p in range(UNEVEN_POLYS)
                    if log_ns[p] == ln and Ls[p] > i]
            if not rows:
                continue
            idx = (np.asarray(rows)[:, None] + np.arange(n)[None, :]).ravel()
            host[idx] = rng.integers(0, q, size=idx.size, dtype=np.uint64)
            other[idx] = rng.integers(0, q, size=idx.size, dtype=np.uint64)
    d = ctx.array(host)
    d_other = ctx.array(other)
    jobs = []
    for p in range(UNEVEN_POLYS):
        ln, L, n = int(log_ns[p]), int(Ls[p]), 1 << int(log_ns[p])
        for i in range(L):
            off = 8 * (int(offs[p]) + i * n)
            jobs.append((d.ptr + off, d_other.ptr + off, pids[ln][i], ln))
    plan = ctx.batch_plan(jobs)
    ops = [sg.BATCH_NTT_FWD, sg.BATCH_MUL, sg.BATCH_NTT_INV]
    for op in ops:
        plan.run(op)
    ctx.sync()
    res = {"polys": UNEVEN_POLYS, "limbs": len(jobs), "words": total}
    for name, op in (("ntt_fwd", sg.BATCH_NTT_FWD), ("mul", sg.BATCH_MUL),
                     ("ntt_inv", sg.BATCH_NTT_INV)):
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(3):
            plan.run(op)
        ctx.sync()
        ms, k = ctx.profile_query("batch")
        ctx.profile(False)
        res[name + "_ms"] = round(ms / k, 4)
    butterflies = sum((L << ln) // 2 * ln for ln, L in zip(log_ns.tolist(), Ls.tolist()))
    res["limb_ntts_per_s"] = round(len(jobs) / (res["ntt_fwd_ms"] * 1e-3), 1)
    res["fwd_hbm_frac"] = round(total * 16 / (res["ntt_fwd_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    res["fwd_butterflies_per_s"] = round(butterflies / (res["ntt_fwd_ms"] * 1e-3), 1)
    # The chain as one HIP graph, replayed.
    d.upload(host)
    ctx.sync()
    ctx.graph_begin()
    for op in ops:
        plan.run(op)
    graph = ctx.graph_end()
    graph.launch()
    ctx.sync()
    got = d.download()
    t0 = time.perf_counter()
    for _ in range(5):
        graph.launch()
    ctx.sync()
    res["graph_chain_ms"] = round((time.perf_counter() - t0) * 1e3 / 5, 4)
    graph.close()
    # Bit-exactness of sampled polynomials (the first replay's output).
    ok = True
    for p in list(range(0, UNEVEN_POLYS, 512)) + [UNEVEN_POLYS - 1]:
        ln, L = int(log_ns[p]), int(Ls[p])
        rns = oracle.Rns(primes[ln][:L], ln)
        n = 1 << ln
        x = host[offs[p]:offs[p + 1]].reshape(L, n)
        y = other[offs[p]:offs[p + 1]].reshape(L, n)
        e = rns.ntt_inv(rns.pointwise("mul", rns.ntt_fwd(x), y))
        ok &= bool(np.array_equal(got[offs[p]:offs[p + 1]].reshape(L, n), e))
    res["bit_exact"] = ok
    # Plain per-call form: one call per polynomial and operation.
    d.upload(host)
    ctx.sync()
    t0 = time.perf_counter()
    for p in range(UNEVEN_POLYS):
        ln, L = int(log_ns[p]), int(Ls[p])
        v = d.view(int(offs[p]), (1, L, 1 << ln))
        vo = d_other.view(int(offs[p]), (1, L, 1 << ln))
        pl = pids[ln][:L]
        ctx.ntt_fwd(v, 1, pl)
        ctx.pointwise(sg.OP_MUL, v, vo, None, 1, pl)
        ctx.ntt_inv(v, 1, pl)
    ctx.sync()
    res["plain_chain_ms"] = round((time.perf_counter() - t0) * 1e3, 4)
    res["graph_vs_plain"] = round(res["plain_chain_ms"] / res["graph_chain_ms"], 2)
    plan.close()
    # Modulus switch of every polynomial with at least two limbs, one launch.
    mjobs = []
    for p in range(UNEVEN_POLYS):
        ln, L = int(log_ns[p]), int(Ls[p])
        if L < 2:
            continue
        ql = ctx.ql_inv(pids[ln][:L])
        mjobs.append((d.ptr + 8 * int(offs[p]), pids[ln][:L], list(ql), ln))
    mplan = ctx.mod_reduce_plan(mjobs, 65537, True)
    mplan.run(sg.BATCH_MOD_REDUCE)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(3):
        mplan.run(sg.BATCH_MOD_REDUCE)
    ctx.sync()
    ms, k = ctx.profile_query("batch")
    ctx.profile(False)
    res["modswitch_polys"] = len(mjobs)
    res["modswitch_ms"] = round(ms / k, 4)
    mplan.close()
    del d, d_other
    return res


SERVING_RATES = (25.0, 50.0, 100.0, 200.0, 400.0, 4000.0)


def serving_checks(ctx, srv, db_id, db_dev, cols, pids, qs, keys, query, log_n, levels, log_bs):
    """Bit-exactness of the serving leg: the server's answer to one query
    equals the same query expanded and multiplied by direct calls; the
    inner product matches a big-integer sum over all rows at sampled
    positions; three expansion levels at this shape match the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from common import expand_oracle
    L, n = len(qs), 1 << log_n
    rows = 1 << levels
    answer, _ = srv.wait(srv.submit(db_id, query))
    buf = ctx.empty((rows, 2, L, n))
    init = np.zeros((rows, 2, L, n), dtype=np.uint64)
    init[0] = query
    buf.upload(init)
    ctx.expand([buf], levels, keys, log_bs, pids)
    out = ctx.empty((cols, 2, L, n))
    ctx.pir_inner_product([buf], levels, db_dev, cols, [out], pids)
    direct = out.download()
    server_equals_direct = bool(np.array_equal(direct.ravel(), answer))
    expd = buf.download()
    recs = db_dev.download()  # exported records: x * 2^-64, so e * x is the Montgomery product
    prng = np.random.default_rng(7)
    ok = True
    for _ in range(16):
        limb, pos = int(prng.integers(0, L)), int(prng.integers(0, n))
        q = qs[limb]
        for comp in range(2):
            e = expd[:, comp, limb, pos].astype(object)
            for c in range(cols):
                x = recs[c::cols, limb, pos][:rows].astype(object)
                ok &= int(np.dot(e, x)) % q == int(direct[c, comp, limb, pos])
    rns = oracle.Rns(qs, log_n)
    host_keys = [(k[0].download(), k[1].download()) for k in keys[:3]]
    small = ctx.empty((8, 2, L, n))
    init8 = np.zeros((8, 2, L, n), dtype=np.uint64)
    init8[0] = query
    small.upload(init8)
    ctx.expand([small], 3, keys[:3], log_bs, pids)
    exp_ok = bool(np.array_equal(small.download(), expand_oracle(rns, query, 3, host_keys, log_bs)))
    return {"server_equals_direct_calls": server_equals_direct,
            "inner_product_sampled_bit_exact": bool(ok),
            "inner_product_sample": "16 (limb, position) pairs x 2 components x %d columns, "
                                    "sums over all %d rows with big integers" % (cols, rows),
            "expansion_3_levels_vs_oracle": exp_ok}


def serving(ctx, rng):
    """The serving workload: encrypted queries (one ciphertext at n=2^12,
    L=2 primes below 2^50) against three device-resident synthetic databases
    of 2^14, 2^16 and 2^18 records (plaintext polynomials, coefficients
    uniform mod t=65537), each query expanded by 12 levels of oblivious
    expansion (Galois substitutions + key switching, base 2^10) into 4096
    row selectors and then multiplied into the database (4096 rows x 4, 16
    or 64 columns).  Poisson arrivals swept over offered rates through the
    knee; the server admits by HBM footprint, starts newcomers in cohorts
    of four while others run, and batches requests of one stage."""
    import shell_gpu as sg
    log_n, Lq, levels, t, log_b = 12, 2, 12, 65537, 10
    qs = primes_for(log_n, Lq)
    pids = ctx.primes(qs, log_n)
    n = 1 << log_n
    k = sum((q.bit_length() + log_b - 1) // log_b for q in qs)
    print("[bench] serving: databases and keys", file=sys.stderr, flush=True)
    keys = [(ctx.array(uniform(rng, qs, (k,), n)), ctx.array(uniform(rng, qs, (k,), n)))
            for _ in range(levels)]
    rows = 1 << levels
    db_sizes = [1 << 14, 1 << 16, 1 << 18]
    dbs = []
    for i, records in enumerate(db_sizes):
        d = ctx.empty((records, Lq, n))
        ctx.pir_db_synthesize(1000 + i, t, records, d, pids)
        ctx.pir_db_export(d, records, d, pids)  # in place: the server reads exported records
        dbs.append((d, records // rows))
    ctx.sync()
    queries = uniform(rng, qs, (8, 2), n)
    budget = 16 << 30
    cohort = 4
    srv = sg.PirServer(ctx, pids, log_n, [log_b] * Lq, keys, max_active=16, hbm_budget=budget,
                       cohort=cohort)
    ids = [srv.add_db(d, c) for d, c in dbs]
    srv.start()
    srv.wait(srv.submit(ids[0], queries[0]))  # warm-up: first-use allocations
    print("[bench] serving: bit-exactness checks", file=sys.stderr, flush=True)
    res = {"workload": "n=2^12, L=2, 12 expansion levels (4096 rows), t=65537, base 2^10",
           "capacity": srv.stats()["capacity"], "cohort": cohort,
           "per_query_hbm_bytes": ((rows + 64) * 2 * Lq * n * 8), "dbs": {}}
    res["checks"] = serving_checks(ctx, srv, ids[0], dbs[0][0], dbs[0][1], pids, qs, keys,
                                   queries[0], log_n, levels, [log_b] * Lq)
    st0 = srv.stats()
    for (d, cols), records, db in zip(dbs, db_sizes, ids):
        entry = {"records": records, "columns": cols, "db_bytes": records * Lq * n * 8,
                 "rates": []}
        for rate in SERVING_RATES:
            count = int(min(96, max(16, rate * 0.4)))
            print("[bench] serving: %d records, %.0f/s offered" % (records, rate),
                  file=sys.stderr, flush=True)
            lat, wall = srv.load_test(db, queries, count, rate, seed=records + int(rate))
            entry["rates"].append({"offered_qps": rate, "requests": count,
                                   "queries_per_s": round(count / wall, 2),
                                   "p50_ms": round(percentile(lat, 50), 3),
                                   "p99_ms": round(percentile(lat, 99), 3)})
        res["dbs"][str(records)] = entry
    st = srv.stats()
    res["mean_batch"] = round((st["batched_items"] - st0["batched_items"]) /
                              max(st["stage_runs"] - st0["stage_runs"], 1), 2)
    srv.close()
    del dbs
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the known-size streaming kernel used to scale byte counters")
    ap.add_argument("--copy-words", type=int, default=1 << 25)
    ap.add_argument("--ntt-probe", action="store_true",
                    help="only run four forward and four inverse transforms of the NTT leg (for counter runs)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_words)
    if args.ntt_probe:
        return ntt_probe()
    if args.gpus != 1:
        print("bench.py measures one GPU only (--gpus 1): the multi-GPU path, which "
              "would shard the batch of ciphertexts with a key replica per GPU, is not "
              "built in this repository.", file=sys.stderr)
        return 2

    import shell_gpu as sg

    rng = np.random.default_rng(20261203)
    qs = primes_for(LOG_N, L)
    n = 1 << LOG_N
    log_bs = [LOG_B] * L
    k = sum((q.bit_length() + LOG_B - 1) // LOG_B for q in qs)
    poly_words = L * n

    ctx = sg.Context(0)
    pids = ctx.primes(qs, LOG_N)
    print("[bench] generating inputs", file=sys.stderr, flush=True)
    comps = uniform(rng, qs, (BATCH, 3), n)
    kas = uniform(rng, qs, (k,), n)
    kbs = uniform(rng, qs, (k,), n)
    d_comps = ctx.array(comps)
    d_ka, d_kb = ctx.array(kas), ctx.array(kbs)
    d_o0, d_o1 = ctx.empty((BATCH, L, n)), ctx.empty((BATCH, L, n))

    # The key is prepared once, as a key object is created once and applied
    # to many ciphertexts: shg_ks_key holds FP64-form copies of its residues
    # for the key-switch kernel (outside the timed region, like key upload).
    ks_key = ctx.ks_key(d_ka, d_kb, k, pids)

    def relin_step():
        ctx.relin_apply_key(d_comps, 2, 0b111, ks_key, log_bs, d_o0, d_o1, BATCH, pids)

    print("[bench] warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        relin_step()
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        relin_step()
        ctx.sync()
        lat.append((time.perf_counter() - t0) * 1e3)
    elapsed = time.perf_counter() - t_start
    ks_ms, ks_launches = ctx.profile_query("keyswitch")
    intt_ms, intt_launches = ctx.profile_query("keyswitch_intt")
    ctx.profile(False)
    ms_per_step = elapsed * 1e3 / args.steps
    value = BATCH / (elapsed / args.steps)
    print("[bench] relin batched: %.3f ms/step, %.0f key-switches/s" % (ms_per_step, value),
          file=sys.stderr, flush=True)
    # The same step through the plain entry point (keys as Montgomery words),
    # outside the timed region: its kernel time, and the same output words.
    prepared_out = (d_o0.download(), d_o1.download())
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(3):
        ctx.relin_apply(d_comps, 2, 0b111, d_ka, d_kb, k, log_bs, d_o0, d_o1, BATCH, pids)
    ctx.sync()
    raw_ms, raw_n = ctx.profile_query("keyswitch")
    ctx.profile(False)
    raw_same = bool(np.array_equal(d_o0.download(), prepared_out[0]) and
                    np.array_equal(d_o1.download(), prepared_out[1]))
    del prepared_out

    # Bit-exactness of two ciphertexts of the batch against the oracle.
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rns = oracle.Rns(qs, LOG_N)
    o0, o1 = d_o0.download(), d_o1.download()
    bit_exact = True
    for c in (0, BATCH - 1):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        bit_exact &= bool(np.array_equal(o0[c], e0) and np.array_equal(o1[c], e1))

    # Dominant kernel: keyswitch_n14_kernel, bound by FP64 vector issue.  Work
    # per key switch (relinearization of one degree-2 ciphertext at n=2^14,
    # L=8, k=40): 320 forward limb transforms of 114,688 butterflies, each
    # butterfly one exact FMA-based modular product (6 FP64 operations) and
    # two additions = 8 operations, plus 2*k*L*n = 10,485,760 key
    # multiply-accumulate terms of 7 operations (the product and the add).
    # Peak: MI355X's 78.6 TFLOP/s FP64 vector rate counts an FMA as two
    # flops, i.e. 39.3 T FP64 operations issued per second.
    ks_avg_ms = ks_ms / max(ks_launches, 1)
    bfly_per_ks = L * k * (n // 2) * LOG_N
    mac_per_ks = 2 * k * L * n
    ops_per_ks = bfly_per_ks * 8 + mac_per_ks * 7
    ops = BATCH * ops_per_ks
    valu_tops = ops / (ks_avg_ms * 1e-3) / 1e12
    # HBM, secondary: per key switch c2 (1 MiB) + c0, c1 (2 MiB) + output
    # (2 MiB), and the 80 MiB key once per launch.
    ks_bytes = BATCH * 5 * poly_words * 8 + 2 * k * poly_words * 8
    ks_gbs = ks_bytes / (ks_avg_ms * 1e-3) / 1e9
    roofline = {"bound": "valu", "achieved": round(valu_tops, 2), "peak": FP64_OPS_PEAK_T,
                "unit": "T FP64 ops/s", "frac": round(valu_tops / FP64_OPS_PEAK_T, 4),
                "traffic": measured_traffic(),
                "kernel": "keyswitch_n14_kernel<F64KsKeyArith, 10>",
                "avg_launch_ms": round(ks_avg_ms, 4),
                "ops_per_launch": ops,
                "op_costs": {"butterflies_per_key_switch": bfly_per_ks, "ops_per_butterfly": 8,
                             "macs_per_key_switch": mac_per_ks, "ops_per_mac": 7},
                "hbm": {"achieved": round(ks_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ks_gbs / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes_per_launch": ks_bytes}}

    out = {
        "metric": "key-switches/s & NTTs/s at n=2^14,L=8 batched vs plain (bit-exact); p50/p99",
        "value": round(value, 1), "unit": "key-switches/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded uniform residues and key matrices)",
        "config": {"workload": "BGV relinearization of 512 degree-2 ciphertexts, n=2^14, L=8 "
                               "primes < 2^50, gadget base 2^10 (k=40), key resident in HBM",
                   "n": n, "L": L, "log_b": LOG_B, "k": k, "batch": BATCH},
        "latency_ms": {"p50": round(percentile(lat, 50), 4), "p99": round(percentile(lat, 99), 4)},
        "bit_exact": bit_exact,
        "roofline": roofline,
        "keyswitch_intt_ms_per_step": round(intt_ms / max(intt_launches, 1), 4),
        "key": "prepared once (shg_ks_key: FP64-form residues) and resident in HBM",
        "keyswitch_plain_key_ms": round(raw_ms / max(raw_n, 1), 4),
        "plain_key_same_words": raw_same,
    }

    if not args.quick:
        # NTTs/s: forward transform of NTT_BATCH RNS polynomials at (2^14, 8).
        x = uniform(rng, qs, (NTT_BATCH,), n)
        d_x = ctx.array(x)
        res = {}
        for name, v in (("f64", sg.NTT_F64), ("int", sg.NTT_INT)):
            # Warm-up: the GPU idled while the input was generated, and its
            # clock takes tens of milliseconds of load to ramp back up.
            for _ in range(NTT_WARMUP):
                ctx.ntt_fwd(d_x, NTT_BATCH, pids, v)
                ctx.ntt_inv(d_x, NTT_BATCH, pids, v)
            ctx.sync()
            ctx.profile(True)
            ctx.profile_reset()
            reps = 20
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.ntt_fwd(d_x, NTT_BATCH, pids, v)
            ctx.sync()
            el = time.perf_counter() - t0
            f_ms, f_n = ctx.profile_query("ntt_fwd")
            for _ in range(reps):
                ctx.ntt_inv(d_x, NTT_BATCH, pids, v)
            ctx.sync()
            i_ms, i_n = ctx.profile_query("ntt_inv")
            ctx.profile(False)
            avg = f_ms / f_n
            by = NTT_BATCH * 2 * poly_words * 8
            res[name] = {"ntts_per_s": round(NTT_BATCH / (avg * 1e-3), 1),
                         "fwd_avg_launch_ms": round(avg, 4),
                         "inv_avg_launch_ms": round(i_ms / i_n, 4),
                         "hbm_gbs": round(by / (avg * 1e-3) / 1e9, 1),
                         "hbm_frac": round(by / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "wall_ntts_per_s": round(NTT_BATCH * reps / el, 1)}
        copy_gbs = stream_copy_gbs()
        res["stream_copy_gbs"] = copy_gbs
        if copy_gbs:
            res["f64"]["frac_of_stream_copy"] = round(res["f64"]["hbm_gbs"] / copy_gbs, 4)
        out["ntts_per_s"] = res["f64"]["ntts_per_s"]
        out["ntt"] = res
        # round trip stays bit-exact
        d_x2 = ctx.array(x[:4])
        ctx.ntt_fwd(d_x2, 4, pids)
        ctx.ntt_inv(d_x2, 4, pids)
        out["ntt_round_trip_identity"] = bool(np.array_equal(d_x2.download(), x[:4]))
        del d_x

        # Plain per-call form: one API call per ciphertext, synchronous.
        calls = 64
        one0, one1 = ctx.empty((1, L, n)), ctx.empty((1, L, n))
        plain_lat = []
        for c in range(calls + 4):
            t0 = time.perf_counter()
            sub = d_comps.view(c * 3 * poly_words, (1, 3, L, n))
            ctx.relin_apply_key(sub, 2, 0b111, ks_key, log_bs, one0, one1, 1, pids)
            ctx.sync()
            if c >= 4:
                plain_lat.append((time.perf_counter() - t0) * 1e3)
        plain = 1e3 / np.mean(plain_lat)
        out["plain_keyswitches_per_s"] = round(plain, 1)
        out["plain_latency_ms"] = {"p50": round(percentile(plain_lat, 50), 4),
                                   "p99": round(percentile(plain_lat, 99), 4)}
        out["batched_vs_plain"] = round(value / plain, 2)

        # Galois substitution x -> x^5 and key switch, then modulus switch.
        gal_in = ctx.array(comps[:, :2])
        d_sub = ctx.empty((BATCH, 2, L, n))
        ctx.substitute(gal_in, d_sub, 5, BATCH * 2, pids)
        ctx.galois_apply_key(d_sub, True, ks_key, log_bs, d_o0, d_o1, BATCH, pids)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.substitute(gal_in, d_sub, 5, BATCH * 2, pids)
            ctx.galois_apply_key(d_sub, True, ks_key, log_bs, d_o0, d_o1, BATCH, pids)
        ctx.sync()
        out["galois_keyswitches_per_s"] = round(3 * BATCH / (time.perf_counter() - t0), 1)
        del gal_in, d_sub
        ql_inv = ctx.ql_inv(pids)
        ms_in = ctx.array(comps[:, :2])
        ctx.mod_reduce_lsb(ms_in, 65537, ql_inv, True, BATCH * 2, pids)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.mod_reduce_lsb(ms_in, 65537, ql_inv, True, BATCH * 2, pids)
        ctx.sync()
        out["modswitch_ciphertexts_per_s"] = round(3 * BATCH / (time.perf_counter() - t0), 1)
        del ms_in

        # Autotuned transform configuration for the NTT workload's shape
        # (its own moduli and batch), timed alternately with the explicit
        # FP64 default on the same buffer so that clock drift and buffer
        # placement affect both alike.
        chosen = ctx.autotune_ntt_moduli(pids, NTT_BATCH)
        d_x = ctx.array(uniform(rng, qs, (NTT_BATCH,), n))
        for _ in range(NTT_WARMUP):
            ctx.ntt_fwd(d_x, NTT_BATCH, pids)
            ctx.ntt_inv(d_x, NTT_BATCH, pids)
        ctx.sync()
        acc = {"auto": [0.0, 0, 0.0, 0], "f64": [0.0, 0, 0.0, 0]}
        ctx.profile(True)
        for _ in range(5):
            for name, v in (("auto", sg.NTT_AUTO), ("f64", sg.NTT_F64)):
                ctx.profile_reset()
                for _ in range(2):
                    ctx.ntt_fwd(d_x, NTT_BATCH, pids, v)
                for _ in range(2):
                    ctx.ntt_inv(d_x, NTT_BATCH, pids, v)
                ctx.sync()
                f_ms, f_n = ctx.profile_query("ntt_fwd")
                i_ms, i_n = ctx.profile_query("ntt_inv")
                a = acc[name]
                a[0] += f_ms; a[1] += f_n; a[2] += i_ms; a[3] += i_n
        ctx.profile(False)
        by = NTT_BATCH * 2 * poly_words * 8
        fa = acc["auto"][0] / acc["auto"][1]
        out["ntt"]["autotuned"] = {"config": chosen, "fwd_avg_launch_ms": round(fa, 4),
                                   "inv_avg_launch_ms": round(acc["auto"][2] / acc["auto"][3], 4),
                                   "ntts_per_s": round(NTT_BATCH / (fa * 1e-3), 1),
                                   "hbm_frac": round(by / (fa * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                   "f64_same_buffer_fwd_ms": round(acc["f64"][0] / acc["f64"][1], 4),
                                   "f64_same_buffer_inv_ms": round(acc["f64"][2] / acc["f64"][3], 4)}
        del d_x
        out["uneven_batch"] = uneven_batch(ctx, rng)
        out["serving"] = serving(ctx, rng)

    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(qs, comps[:16], kas, kbs, log_bs,
                                           uniform(rng, qs, (8,), n))
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
