"""Batched uneven launches and HIP graphs of op chains, bit-exact against the
oracle.  The uneven batch mirrors the shape of the uneven-batching workload
(rings 2^11..2^15, 1..L limbs per polynomial, primes chosen per ring) at a
size the oracle checks in seconds."""
import numpy as np
import pytest

import oracle
import shell_gpu as sg
from common import ntt_primes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sg.Context(0)
    yield c
    c.close()


def _uneven(ctx, seed, num_polys, max_limbs, log_bound=50, wide_every=0):
    """Random polynomials (log_n in 11..15, 1..max_limbs limbs) as separate
    device buffers; returns per-limb records and the job tuples."""
    rng = np.random.default_rng(seed)
    recs, jobs = [], []
    for p in range(num_polys):
        log_n = int(rng.integers(11, 16))
        L = int(rng.integers(1, max_limbs + 1))
        bound = 60 if (wide_every and p % wide_every == 0) else log_bound
        qs = ntt_primes(log_n, L, bound)
        n = 1 << log_n
        x = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in qs])
        y = np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in qs])
        dx, dy = ctx.array(x), ctx.array(y)
        pids = ctx.primes(qs, log_n)
        for i in range(L):
            recs.append(dict(q=qs[i], log_n=log_n, x=x[i], y=y[i], dx=dx, dy=dy, i=i, n=n))
            jobs.append((dx.ptr + 8 * i * n, dy.ptr + 8 * i * n, pids[i], log_n))
    return recs, jobs


def _expected(recs, ops):
    out = []
    for r in recs:
        rns = oracle.Rns([r["q"]], r["log_n"])
        v = r["x"][None].copy()
        for op in ops:
            if op == sg.BATCH_NTT_FWD:
                v = rns.ntt_fwd(v)
            elif op == sg.BATCH_NTT_INV:
                v = rns.ntt_inv(v)
            else:
                v = rns.pointwise("mul", v, r["y"][None])
        out.append(v[0])
    return out


def _got(recs):
    cache = {}
    out = []
    for r in recs:
        key = id(r["dx"])
        if key not in cache:
            cache[key] = r["dx"].download()
        out.append(cache[key][r["i"]])
    return out


def _reset(recs):
    seen = set()
    for r in recs:
        if id(r["dx"]) in seen:
            continue
        seen.add(id(r["dx"]))
        full = np.stack([rr["x"] for rr in recs if rr["dx"] is r["dx"]])
        r["dx"].upload(full)


@pytest.mark.parametrize("wide_every", [0, 5])
def test_batch_uneven_chain(ctx, wide_every):
    """NTT, pointwise multiply, INTT over uneven jobs, one launch each; with
    a 60-bit prime in the mix the plan takes the integer policy."""
    recs, jobs = _uneven(ctx, 11 + wide_every, 18, 4, wide_every=wide_every)
    plan = ctx.batch_plan(jobs)
    ops = [sg.BATCH_NTT_FWD, sg.BATCH_MUL, sg.BATCH_NTT_INV]
    for k in range(1, len(ops) + 1):
        _reset(recs)
        for op in ops[:k]:
            plan.run(op)
        ctx.sync()
        got, exp = _got(recs), _expected(recs, ops[:k])
        for g, e, r in zip(got, exp, recs):
            assert np.array_equal(g, e), (k, r["q"], r["log_n"])
    plan.close()


def test_batch_one_call_and_errors(ctx):
    recs, jobs = _uneven(ctx, 3, 4, 2)
    ctx.batch_limbs(jobs, sg.BATCH_NTT_FWD)
    for g, e in zip(_got(recs), _expected(recs, [sg.BATCH_NTT_FWD])):
        assert np.array_equal(g, e)
    d, _, p, ln = jobs[0]
    with pytest.raises(sg.ShgError):
        ctx.batch_plan([(d, None, p, ln + 1 if ln < 15 else ln - 1)])  # prime of another degree
    with pytest.raises(sg.ShgError):
        ctx.batch_plan([(d + 8, None, p, ln)])  # misaligned
    q10 = ntt_primes(10, 1)[0]
    with pytest.raises(sg.ShgError):
        ctx.batch_plan([(d, None, ctx.prime(q10, 10), 10)])  # below 2^11
    plan = ctx.batch_plan([(d, None, p, ln)])
    with pytest.raises(sg.ShgError):
        plan.run(sg.BATCH_MUL)  # no second operand
    plan.close()


def test_graph_batch_chain_replays(ctx):
    """A captured NTT -> multiply -> INTT chain replays bit-exactly, three
    times, with eager uploads and synchronisation between the replays."""
    recs, jobs = _uneven(ctx, 21, 10, 3)
    plan = ctx.batch_plan(jobs)
    ops = [sg.BATCH_NTT_FWD, sg.BATCH_MUL, sg.BATCH_NTT_INV]
    _reset(recs)
    for op in ops:
        plan.run(op)
    ctx.sync()
    eager = _got(recs)
    for g, e in zip(eager, _expected(recs, ops)):
        assert np.array_equal(g, e)
    _reset(recs)
    ctx.sync()
    ctx.graph_begin()
    for op in ops:
        plan.run(op)
    graph = ctx.graph_end()
    for _ in range(3):
        _reset(recs)
        ctx.sync()
        graph.launch()
        ctx.sync()
        for g, e in zip(_got(recs), eager):
            assert np.array_equal(g, e)
    graph.close()
    plan.close()


def test_graph_tensor_relin_modswitch(ctx):
    """The ciphertext chain tensor product -> relinearization -> modulus
    switch captured once and replayed, against the eager run and the
    oracle."""
    log_n, L, cnt = 12, 3, 4
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    n = 1 << log_n
    a = np.stack([rns.uniform(300 + c, 2) for c in range(cnt)])
    b = np.stack([rns.uniform(400 + c, 2) for c in range(cnt)])
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    kas, kbs = rns.uniform(7, k), rns.uniform(8, k)
    da, db = ctx.array(a), ctx.array(b)
    dka, dkb = ctx.array(kas), ctx.array(kbs)
    t3 = ctx.empty((cnt, 3, L, n))
    o0, o1 = ctx.empty((cnt, L, n)), ctx.empty((cnt, L, n))
    ql_inv = rns.ql_inv()

    def chain():
        ctx.tensor_product(da, 2, db, 2, t3, cnt, pids)
        ctx.relin_apply(t3, 2, 0b111, dka, dkb, k, log_bs, o0, o1, cnt, pids)
        ctx.mod_reduce_lsb(o0, 65537, ql_inv, True, cnt, pids)
        ctx.mod_reduce_lsb(o1, 65537, ql_inv, True, cnt, pids)

    chain()
    ctx.sync()
    e0, e1 = o0.download()[:, : L - 1], o1.download()[:, : L - 1]
    for c in range(cnt):
        tp = rns.tensor_product(a[c], b[c])
        r0, r1 = rns.relin_apply(tp, kas, kbs, log_bs)
        assert np.array_equal(e0[c], rns.mod_reduce_lsb(r0, 65537, ql_inv, True))
        assert np.array_equal(e1[c], rns.mod_reduce_lsb(r1, 65537, ql_inv, True))
    ctx.graph_begin()
    chain()
    graph = ctx.graph_end()
    for rep in range(3):
        o0.upload(np.zeros((cnt, L, n), dtype=np.uint64))
        o1.upload(np.zeros((cnt, L, n), dtype=np.uint64))
        ctx.sync()
        graph.launch()
        ctx.sync()
        assert np.array_equal(o0.download()[:, : L - 1], e0), rep
        assert np.array_equal(o1.download()[:, : L - 1], e1), rep
    graph.close()


@pytest.mark.parametrize("is_ntt", [True, False])
def test_batch_mod_reduce_uneven(ctx, is_ntt):
    """ModReduceLsb of polynomials of uneven ring degree (2^11..2^15) and
    level (2..6 limbs) in one launch, against the oracle per polynomial."""
    rng = np.random.default_rng(5 + is_ntt)
    t = 65537
    polys, jobs = [], []
    for p in range(14):
        log_n = int(rng.integers(11, 16))
        L = int(rng.integers(2, 7))
        qs = ntt_primes(log_n, L)
        rns = oracle.Rns(qs, log_n)
        x = np.stack([rng.integers(0, q, size=1 << log_n, dtype=np.uint64) for q in qs])
        d = ctx.array(x)
        ql_inv = rns.ql_inv()
        polys.append((x, d, rns, ql_inv))
        jobs.append((d.ptr, ctx.primes(qs, log_n), list(ql_inv), log_n))
    plan = ctx.mod_reduce_plan(jobs, t, is_ntt)
    plan.run(sg.BATCH_MOD_REDUCE)
    ctx.sync()
    for x, d, rns, ql_inv in polys:
        L = x.shape[0]
        got = d.download()[: L - 1]
        assert np.array_equal(got, rns.mod_reduce_lsb(x, t, ql_inv, is_ntt))
        assert np.array_equal(d.download()[L - 1], x[L - 1])  # last limb untouched
    with pytest.raises(sg.ShgError):
        plan.run(sg.BATCH_NTT_FWD)
    plan.close()


def test_graph_prepared_key_relin_and_galois_n16384(ctx):
    """Key switches with a prepared key (the FP64-key kernel at 2^14)
    captured after one eager call at the same size and replayed three times,
    with eager work and synchronisation between replays, against the eager
    result and the oracle."""
    log_n, L, cnt = 14, 3, 3
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    n = 1 << log_n
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(900 + c, 3) for c in range(cnt)])
    kas, kbs = rns.uniform(17, k), rns.uniform(18, k)
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    key = ctx.ks_key(dka, dkb, k, pids)
    o0, o1 = ctx.empty((cnt, L, n)), ctx.empty((cnt, L, n))
    g0, g1 = ctx.empty((cnt, L, n)), ctx.empty((cnt, L, n))
    gal = ctx.array(np.ascontiguousarray(comps[:, :2]))

    def chain():
        ctx.relin_apply_key(dc, 2, 0b111, key, log_bs, o0, o1, cnt, pids)
        ctx.galois_apply_key(gal, True, key, log_bs, g0, g1, cnt, pids)

    chain()
    ctx.sync()
    eager = [b.download() for b in (o0, o1, g0, g1)]
    r0, r1 = rns.relin_apply(comps[0], kas, kbs, log_bs)
    assert np.array_equal(eager[0][0], r0) and np.array_equal(eager[1][0], r1)
    s0, s1 = rns.galois_apply(comps[cnt - 1, 0], comps[cnt - 1, 1], kas, kbs, log_bs)
    assert np.array_equal(eager[2][cnt - 1], s0) and np.array_equal(eager[3][cnt - 1], s1)
    ctx.graph_begin()
    chain()
    graph = ctx.graph_end()
    zero = np.zeros((cnt, L, n), dtype=np.uint64)
    for _ in range(3):
        for b in (o0, o1, g0, g1):
            b.upload(zero)
        ctx.sync()
        graph.launch()
        ctx.sync()
        for b, e in zip((o0, o1, g0, g1), eager):
            assert np.array_equal(b.download(), e)
    graph.close()
    key.close()


@pytest.mark.gpu
def test_graph_two_per_cu_ntt_and_modswitch_n16384(ctx):
    """FP64 transforms at 2^14 (the two-workgroups-per-CU kernels) and the
    two-launch modulus switch (inverse of the dropped limbs into scratch,
    then one workgroup per remaining limb) captured after one eager call at
    the same size and replayed three times, with eager uploads and
    synchronisation between replays, against the eager result and the
    oracle."""
    log_n, L, cnt = 14, 4, 5
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    x = rns.uniform(4242, cnt)
    ql_inv = ctx.ql_inv(pids)
    d = ctx.array(x)

    def chain():
        ctx.ntt_fwd(d, cnt, pids)
        ctx.ntt_inv(d, cnt, pids)
        ctx.ntt_fwd(d, cnt, pids)
        ctx.mod_reduce_lsb(d, 65537, ql_inv, True, cnt, pids)

    chain()
    ctx.sync()
    eager = d.download()
    # The first L - 1 limbs hold the result; compare them with the oracle.
    for c in (0, cnt - 1):
        exp = rns.mod_reduce_lsb(rns.ntt_fwd(x[c]), 65537, ql_inv, True)
        assert np.array_equal(eager[c, :L - 1], exp), c
    ctx.graph_begin()
    chain()
    graph = ctx.graph_end()
    for _ in range(3):
        d.upload(x)
        ctx.sync()
        graph.launch()
        ctx.sync()
        assert np.array_equal(d.download(), eager)
    graph.close()
