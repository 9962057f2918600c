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
# HBM bytes per launch of the key-switch kernel, measured by rocprofv3
# FETCH_SIZE + WRITE_SIZE and scaled by the same counters' reading for a
# streaming kernel of known size (bench.py --copy-probe).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "keyswitch_relin_n16384_L8_b512_pmc.md")
LOG_N = 14
L = 8
LOG_B = 10
BATCH = 512
NTT_BATCH = 1024


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


def cpu_baseline(qs, comps, kas, kbs, log_bs, budget_s=12.0):
    """The oracle's relinearization (a scalar restatement of the reference's
    RnsRelinKey::ApplyTo) on one host core, over a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rns = oracle.Rns(qs, LOG_N)
    done, t0 = 0, time.perf_counter()
    while True:
        rns.relin_apply(comps[done % len(comps)], kas, kbs, log_bs)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= 64:
            break
    return {"value": done / el, "unit": "key-switches/s", "cores": 1, "kind": "port",
            "sample": "%d relinearizations of degree-2 ciphertexts at n=2^14, L=8, k=40 by the "
                      "oracle (single-threaded C restatement), %.1f s" % (done, el)}


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
    at n=2^14, L=8), four forward then four inverse launches: the workload of
    the NTT counter runs."""
    import shell_gpu as sg
    qs = primes_for(LOG_N, L)
    rng = np.random.default_rng(20261203)
    ctx = sg.Context(0)
    pids = ctx.primes(qs, LOG_N)
    d_x = ctx.array(uniform(rng, qs, (NTT_BATCH,), 1 << LOG_N))
    for _ in range(4):
        ctx.ntt_fwd(d_x, NTT_BATCH, pids, sg.NTT_F64)
    for _ in range(4):
        ctx.ntt_inv(d_x, NTT_BATCH, pids, sg.NTT_F64)
    ctx.sync()
    print(json.dumps({"ntt_probe": "forward and inverse FP64", "polys": NTT_BATCH,
                      "n": 1 << LOG_N, "L": L, "launches": 8}))
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
        rns_ql = oracle.Rns(primes[ln][:L], ln).ql_inv()
        mjobs.append((d.ptr + 8 * int(offs[p]), pids[ln][:L], list(rns_ql), ln))
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


def serving(ctx, rng):
    """The serving workload: encrypted queries (one ciphertext at n=2^12,
    L=2 primes below 2^50) against three device-resident synthetic databases
    of 2^14, 2^16 and 2^18 records (plaintext polynomials, coefficients
    uniform mod t=65537), each query expanded by 12 levels of oblivious
    expansion (Galois substitutions + key switching, base 2^10) into 4096
    row selectors and then multiplied into the database (4096 rows x 4, 16
    or 64 columns).  Poisson arrivals at a low and a saturating rate; the
    server admits by HBM footprint and batches requests of one stage."""
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
    srv = sg.PirServer(ctx, pids, log_n, [log_b] * Lq, keys, max_active=16, hbm_budget=budget)
    ids = [srv.add_db(d, c) for d, c in dbs]
    srv.start()
    srv.wait(srv.submit(ids[0], queries[0]))  # warm-up: first-use allocations
    res = {"workload": "n=2^12, L=2, 12 expansion levels (4096 rows), t=65537, base 2^10",
           "capacity": srv.stats()["capacity"],
           "per_query_hbm_bytes": ((rows + 64) * 2 * Lq * n * 8), "dbs": {}}
    for (d, cols), records, db in zip(dbs, db_sizes, ids):
        entry = {"records": records, "columns": cols, "db_bytes": records * Lq * n * 8}
        for name, rate, count in (("low_rate", 20.0, 12), ("high_rate", 4000.0, 64)):
            print("[bench] serving: %d records, %s" % (records, name), file=sys.stderr, flush=True)
            lat, wall = srv.load_test(db, queries, count, rate, seed=records + int(rate))
            entry[name] = {"offered_qps": rate, "requests": count,
                           "queries_per_s": round(count / wall, 2),
                           "p50_ms": round(percentile(lat, 50), 3),
                           "p99_ms": round(percentile(lat, 99), 3)}
        res["dbs"][str(records)] = entry
    st = srv.stats()
    res["mean_batch"] = round(st["batched_items"] / max(st["stage_runs"], 1), 2)
    srv.close()
    del dbs
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
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

    def relin_step():
        ctx.relin_apply(d_comps, 2, 0b111, d_ka, d_kb, k, log_bs, d_o0, d_o1, BATCH, pids)

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

    # Bit-exactness of two ciphertexts of the batch against the oracle.
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rns = oracle.Rns(qs, LOG_N)
    o0, o1 = d_o0.download(), d_o1.download()
    bit_exact = True
    for c in (0, BATCH - 1):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        bit_exact &= bool(np.array_equal(o0[c], e0) and np.array_equal(o1[c], e1))

    # Dominant kernel roofline: algorithmic bytes per ciphertext = its gadget
    # digits as 16-bit planes (k*n*2, written by the inverse-transform step)
    # + c0, c1 (2*L*n*8) + output (2*L*n*8), plus the key (2*k*L*n*8) shared
    # by the whole launch.
    ks_avg_ms = ks_ms / max(ks_launches, 1)
    ks_bytes = BATCH * (k * n * 2 + 4 * poly_words * 8) + 2 * k * poly_words * 8
    ks_gbs = ks_bytes / (ks_avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(ks_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ks_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                "kernel": "keyswitch_kernel<13,4,1,F64KsArith,planes>",
                "avg_launch_ms": round(ks_avg_ms, 4),
                "algorithmic_bytes_per_launch": ks_bytes}

    # The key switch is bound by FP64 vector arithmetic, not HBM: its
    # algorithmic FP64 operations per launch (a Cooley-Tukey butterfly = one
    # exact FMA-based product of 6 operations + 2 additions; stage 0 of a
    # half-limb = product + 1 addition; a key multiply-accumulate term = 7)
    # against MI355X's FP64 vector issue rate (78.6 TFLOP/s counting an FMA
    # as two flops = 39.3 T operations/s).
    half = n // 2
    bfly = BATCH * L * k * (2 * (half // 2) * (LOG_N - 1) * 8 + 2 * half * 7)
    mac = BATCH * L * k * n * 2 * 7
    ops = bfly + mac
    fp64_valu = {"achieved": round(ops / (ks_avg_ms * 1e-3) / 1e12, 2), "peak": 39.3,
                 "unit": "T FP64 ops/s", "frac": round(ops / (ks_avg_ms * 1e-3) / 39.3e12, 4),
                 "ops_per_launch": ops}
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
        "fp64_valu": fp64_valu,
        "keyswitch_intt_ms_per_step": round(intt_ms / max(intt_launches, 1), 4),
    }

    if not args.quick:
        # NTTs/s: forward transform of NTT_BATCH RNS polynomials at (2^14, 8).
        x = uniform(rng, qs, (NTT_BATCH,), n)
        d_x = ctx.array(x)
        for v in (sg.NTT_F64, sg.NTT_INT):
            ctx.ntt_fwd(d_x, NTT_BATCH, pids, v)
            ctx.ntt_inv(d_x, NTT_BATCH, pids, v)
        ctx.sync()
        res = {}
        for name, v in (("f64", sg.NTT_F64), ("int", sg.NTT_INT)):
            ctx.profile(True)
            ctx.profile_reset()
            reps = 10
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
            ctx.relin_apply(sub, 2, 0b111, d_ka, d_kb, k, log_bs, one0, one1, 1, pids)
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
        ctx.galois_apply(d_sub, True, d_ka, d_kb, log_bs, d_o0, d_o1, BATCH, pids)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.substitute(gal_in, d_sub, 5, BATCH * 2, pids)
            ctx.galois_apply(d_sub, True, d_ka, d_kb, log_bs, d_o0, d_o1, BATCH, pids)
        ctx.sync()
        out["galois_keyswitches_per_s"] = round(3 * BATCH / (time.perf_counter() - t0), 1)
        del gal_in, d_sub
        ql_inv = rns.ql_inv()
        ms_in = ctx.array(comps[:, :2])
        ctx.mod_reduce_lsb(ms_in, 65537, ql_inv, True, BATCH * 2, pids)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.mod_reduce_lsb(ms_in, 65537, ql_inv, True, BATCH * 2, pids)
        ctx.sync()
        out["modswitch_ciphertexts_per_s"] = round(3 * BATCH / (time.perf_counter() - t0), 1)
        del ms_in

        # Autotuned transform configuration for the NTT workload's shape.
        chosen = ctx.autotune_ntt(LOG_N, NTT_BATCH * L)
        d_x = ctx.array(uniform(rng, qs, (NTT_BATCH,), n))
        ctx.ntt_fwd(d_x, NTT_BATCH, pids)
        ctx.sync()
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(10):
            ctx.ntt_fwd(d_x, NTT_BATCH, pids)
        for _ in range(10):
            ctx.ntt_inv(d_x, NTT_BATCH, pids)
        ctx.sync()
        f_ms, f_n = ctx.profile_query("ntt_fwd")
        i_ms, i_n = ctx.profile_query("ntt_inv")
        ctx.profile(False)
        by = NTT_BATCH * 2 * poly_words * 8
        out["ntt"]["autotuned"] = {"config": chosen, "fwd_avg_launch_ms": round(f_ms / f_n, 4),
                                   "inv_avg_launch_ms": round(i_ms / i_n, 4),
                                   "ntts_per_s": round(NTT_BATCH / (f_ms / f_n * 1e-3), 1),
                                   "hbm_frac": round(by / (f_ms / f_n * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        del d_x
        out["uneven_batch"] = uneven_batch(ctx, rng)
        out["serving"] = serving(ctx, rng)

    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(qs, comps[:8], kas, kbs, log_bs)
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
