"""Parity of the workloads at their own shapes, on the launch grids the
benchmark uses:

* relinearization and Galois key switching of 512 degree-2 ciphertexts at
  n = 2^14, L = 8, base 2^10 (k = 40): 8,192 workgroups with the XCD remap,
  the first, a middle and the last ciphertext against the oracle;
* uneven batches with 7..12-limb polynomials at n = 2^14 and 2^15 (forward,
  multiply, inverse and ModReduceLsb plans);
* encrypted queries at n = 2^12, L = 2 with 6 expansion levels (64 rows),
  several concurrent queries and 9 columns, which takes the split-sum inner
  product, directly and through the serving loop;
* the lazy multiply-accumulate with a 62-bit modulus, whose 128-bit
  accumulator is folded every 8 terms (rns/lazy_rns_polynomial.h:156-181).
"""
import numpy as np
import pytest

import oracle
import shell_gpu as sg
from common import expand_oracle, inner_oracle, ntt_primes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sg.Context(0)
    yield c
    c.close()


def _uniform(rng, qs, prefix, n):
    out = np.empty(tuple(prefix) + (len(qs), n), dtype=np.uint64)
    for i, q in enumerate(qs):
        out[..., i, :] = rng.integers(0, q, size=tuple(prefix) + (n,), dtype=np.uint64)
    return out


def test_keyswitch_512_ciphertexts_n16384_L8(ctx):
    log_n, L, B, log_b = 14, 8, 512, 10
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    n = 1 << log_n
    log_bs = [log_b] * L
    k, _ = rns.gadget_dims(log_bs)
    assert k == 40
    rng = np.random.default_rng(512)
    comps = _uniform(rng, qs, (B, 3), n)
    kas, kbs = _uniform(rng, qs, (k,), n), _uniform(rng, qs, (k,), n)
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((B, L, n)), ctx.empty((B, L, n))
    ctx.relin_apply(dc, 2, 0b111, dka, dkb, k, log_bs, o0, o1, B, pids)
    g0, g1 = o0.download(), o1.download()
    for c in (0, B // 2 + 3, B - 1):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), ("relin", c)
    # The prepared key (FP64-form words, the bench's timed kernel) gives the
    # same words for every one of the 512 ciphertexts.
    key = ctx.ks_key(dka, dkb, k, pids)
    p0, p1 = ctx.empty((B, L, n)), ctx.empty((B, L, n))
    ctx.relin_apply_key(dc, 2, 0b111, key, log_bs, p0, p1, B, pids)
    assert np.array_equal(p0.download(), g0) and np.array_equal(p1.download(), g1)
    del p0, p1
    # Galois x -> x^5 of (c0, c1): substitution then the key switch.
    gal = ctx.array(np.ascontiguousarray(comps[:, :2]))
    sub = ctx.empty((B, 2, L, n))
    ctx.substitute(gal, sub, 5, 2 * B, pids)
    ctx.galois_apply(sub, True, dka, dkb, log_bs, o0, o1, B, pids)
    g0, g1 = o0.download(), o1.download()
    for c in (0, B // 2 + 3, B - 1):
        s0, s1 = rns.substitute(comps[c, 0], 5), rns.substitute(comps[c, 1], 5)
        e0, e1 = rns.galois_apply(s0, s1, kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), ("galois", c)


def test_uneven_batch_with_12_limb_polynomials(ctx):
    """Forward, multiply and inverse plans plus a ModReduceLsb plan over
    polynomials of 7..12 limbs at n = 2^14 and 2^15 mixed with smaller ones."""
    rng = np.random.default_rng(4096)
    shapes = [(15, 12), (15, 11), (15, 7), (14, 12), (14, 9), (13, 12), (12, 10), (11, 8),
              (15, 1), (14, 2)]
    polys = []
    jobs = []
    for p, (log_n, L) in enumerate(shapes):
        qs = ntt_primes(log_n, L)
        rns = oracle.Rns(qs, log_n)
        pids = ctx.primes(qs, log_n)
        n = 1 << log_n
        x = _uniform(rng, qs, (), n)
        y = _uniform(rng, qs, (), n)
        dx, dy = ctx.array(x), ctx.array(y)
        polys.append((x, y, dx, dy, rns, pids, log_n))
        for i in range(L):
            jobs.append((dx.ptr + 8 * i * n, dy.ptr + 8 * i * n, pids[i], log_n))
    plan = ctx.batch_plan(jobs)
    for op in (sg.BATCH_NTT_FWD, sg.BATCH_MUL, sg.BATCH_NTT_INV):
        plan.run(op)
    ctx.sync()
    for x, y, dx, dy, rns, pids, log_n in polys:
        e = rns.ntt_inv(rns.pointwise("mul", rns.ntt_fwd(x), y))
        assert np.array_equal(dx.download(), e), (log_n, x.shape[0])
    plan.close()
    mjobs, expect = [], []
    for x, y, dx, dy, rns, pids, log_n in polys:
        if x.shape[0] < 2:
            continue
        dy.upload(y)
        ql_inv = ctx.ql_inv(pids)
        assert np.array_equal(ql_inv, rns.ql_inv())
        mjobs.append((dy.ptr, pids, list(ql_inv), log_n))
        expect.append((dy, rns.mod_reduce_lsb(y, 65537, ql_inv, True), y.shape[0]))
    mplan = ctx.mod_reduce_plan(mjobs, 65537, True)
    mplan.run(sg.BATCH_MOD_REDUCE)
    ctx.sync()
    for dy, e, L in expect:
        assert np.array_equal(dy.download()[: L - 1], e), L
    mplan.close()


LOG_N, LQ, LEVELS, COLS, QUERIES = 12, 2, 6, 9, 5


@pytest.fixture(scope="module")
def pir_case(ctx):
    qs = ntt_primes(LOG_N, LQ)
    rns = oracle.Rns(qs, LOG_N)
    pids = ctx.primes(qs, LOG_N)
    log_bs = [10] * LQ
    k, _ = rns.gadget_dims(log_bs)
    keys = [(rns.uniform(2100 + j, k), rns.uniform(2200 + j, k)) for j in range(LEVELS)]
    dkeys = [(ctx.array(a), ctx.array(b)) for a, b in keys]
    K = 1 << LEVELS
    db = rns.uniform(2300, K * COLS)
    dbx = ctx.empty(db.shape)
    ctx.pir_db_export(ctx.array(db), K * COLS, dbx, pids)
    queries = [rns.uniform(2400 + g, 2) for g in range(QUERIES)]
    expanded = [expand_oracle(rns, q, LEVELS, keys, log_bs) for q in queries]
    answers = [inner_oracle(rns, e, db, COLS) for e in expanded]
    return dict(rns=rns, pids=pids, log_bs=log_bs, dkeys=dkeys, dbx=dbx, queries=queries,
                expanded=expanded, answers=answers, K=K)


# Five queries form two groups of four (the second one partly filled); with
# 9 columns and 64 rows the inner product splits the sum over the rows into
# 4 partial sums and adds them in a second kernel, as it does for one query.
@pytest.mark.parametrize("group", [[0, 1, 2, 3, 4], [2]])
def test_expand_and_inner_product_n4096_six_levels(ctx, pir_case, group):
    s = pir_case
    K, n = s["K"], 1 << LOG_N
    bufs = []
    for g in group:
        init = np.zeros((K, 2, LQ, n), dtype=np.uint64)
        init[0] = s["queries"][g]
        bufs.append(ctx.array(init))
    ctx.expand(bufs, LEVELS, s["dkeys"], s["log_bs"], s["pids"])
    outs = [ctx.empty((COLS, 2, LQ, n)) for _ in group]
    ctx.pir_inner_product(bufs, LEVELS, s["dbx"], COLS, outs, s["pids"])
    ctx.sync()
    for g, b, o in zip(group, bufs, outs):
        assert np.array_equal(b.download(), s["expanded"][g]), g
        assert np.array_equal(o.download(), s["answers"][g]), g


def test_server_n4096_six_levels_alone_equals_mixed(ctx, pir_case):
    s = pir_case
    srv = sg.PirServer(ctx, s["pids"], LOG_N, s["log_bs"], s["dkeys"], max_active=8)
    db = srv.add_db(s["dbx"], COLS)
    srv.start()
    try:
        alone = srv.wait(srv.submit(db, s["queries"][1]))[0]
        tickets = [srv.submit(db, s["queries"][g]) for g in range(QUERIES)]
        mixed = [srv.wait(t)[0] for t in tickets]
    finally:
        srv.close()
    assert np.array_equal(alone, mixed[1])
    for g in range(QUERIES):
        assert np.array_equal(mixed[g], s["answers"][g].ravel()), g


def test_lazy_fma_62bit_modulus_folds_accumulator(ctx):
    log_n = 10
    qs = [ntt_primes(log_n, 1, 62)[0], ntt_primes(log_n, 1, 50)[0]]
    assert qs[0].bit_length() == 62
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    words = len(qs) * rns.n
    for terms in (7, 8, 9, 64, 200):
        as_, bs = rns.uniform(3100 + terms, terms), rns.uniform(3200 + terms, terms)
        # The largest residues maximise every 128-bit partial sum.
        as_[:, 0, :8] = qs[0] - 1
        bs[:, 0, :8] = qs[0] - 1
        out = ctx.empty((len(qs), rns.n))
        ctx.lazy_fma(ctx.array(as_), words, ctx.array(bs), words, terms, out, pids)
        assert np.array_equal(out.download(), rns.lazy_fma(as_, bs)), terms
