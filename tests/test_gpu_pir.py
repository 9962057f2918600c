"""Oblivious expansion and the inner product against a database, bit-exact
against the oracle composed step by step as ObliviousExpander does
(oblivious_expand.h:173-227), and end to end: an expanded BGV query selects
one database row, which decrypts to the plaintext records."""
import numpy as np
import pytest

import oracle
import shell_gpu as sg
from common import Bgv, expand_oracle, inner_oracle, ntt_primes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sg.Context(0)
    yield c
    c.close()


def test_expand_and_inner_product_bit_exact(ctx):
    log_n, L, levels, C = 10, 2, 3, 3
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    keys = [(rns.uniform(700 + j, k), rns.uniform(800 + j, k)) for j in range(levels)]
    dkeys = [(ctx.array(a), ctx.array(b)) for a, b in keys]
    K = 1 << levels
    queries = [rns.uniform(900 + g, 2) for g in range(3)]
    bufs = []
    for q in queries:
        b = ctx.array(np.concatenate([q[None], np.zeros((K - 1, 2, L, rns.n), dtype=np.uint64)]))
        bufs.append(b)
    ctx.expand(bufs, levels, dkeys, log_bs, pids)
    ctx.sync()
    exps = []
    for q, b in zip(queries, bufs):
        e = expand_oracle(rns, q, levels, keys, log_bs)
        assert np.array_equal(b.download(), e)
        exps.append(e)
    db = rns.uniform(55, K * C)
    ddb = ctx.array(db)
    dbx = ctx.empty(db.shape)
    ctx.pir_db_export(ddb, K * C, dbx, pids)
    outs = [ctx.empty((C, 2, L, rns.n)) for _ in queries]
    ctx.pir_inner_product(bufs, levels, dbx, C, outs, pids)
    ctx.sync()
    for e, o in zip(exps, outs):
        assert np.array_equal(o.download(), inner_oracle(rns, e, db, C))


def test_expansion_selects_a_row_and_decrypts(ctx):
    log_n, L, levels, C, t = 10, 2, 3, 2, 65537
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    bgv = Bgv(rns, t, seed=17)
    log_bs = [10] * L
    n, K = rns.n, 1 << levels
    keys = [bgv.galois_key(n // (1 << j) + 1, log_bs) for j in range(levels)]
    dkeys = [(ctx.array(a), ctx.array(b)) for a, b in keys]
    rng = np.random.default_rng(3)
    rows = rng.integers(0, t, size=(K * C, n))
    db = np.stack([bgv.encode(r) for r in rows])
    dbx = ctx.empty(db.shape)
    ctx.pir_db_export(ctx.array(db), K * C, dbx, pids)
    inv = pow(K, -1, t)
    for sel in (0, 5):
        m = np.zeros(n, dtype=np.int64)
        m[sel] = inv
        ct = bgv.encrypt(m)
        buf = ctx.array(np.concatenate([ct[None], np.zeros((K - 1, 2, L, n), dtype=np.uint64)]))
        ctx.expand([buf], levels, dkeys, log_bs, pids)
        out = ctx.empty((C, 2, L, n))
        ctx.pir_inner_product([buf], levels, dbx, C, [out], pids)
        ctx.sync()
        expd = buf.download()
        for r in range(K):
            dec = bgv.decrypt(expd[r])
            assert dec[0] == (1 if r == sel else 0) and not dec[1:].any(), (sel, r)
        ans = out.download()
        for c in range(C):
            assert np.array_equal(bgv.decrypt(ans[c]), rows[sel * C + c] % t), (sel, c)


@pytest.mark.parametrize("log_n,levels", [(12, 5), (13, 3), (10, 3)])
def test_expand_with_prepared_keys(ctx, log_n, levels):
    """Expansion levels with prepared Galois keys (the FP64-key kernels at
    2^12 and 2^13, the Montgomery words below) equal the oracle's
    composition word for word."""
    L = 2
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    keys = [(rns.uniform(1700 + j, k), rns.uniform(1800 + j, k)) for j in range(levels)]
    dkeys = [(ctx.array(a), ctx.array(b)) for a, b in keys]
    pkeys = [ctx.ks_key(a, b, k, pids) for a, b in dkeys]
    K = 1 << levels
    queries = [rns.uniform(1900 + g, 2) for g in range(2)]
    bufs = [ctx.array(np.concatenate([q[None], np.zeros((K - 1, 2, L, rns.n), dtype=np.uint64)]))
            for q in queries]
    for j in range(levels):
        ctx.expand_level_key(bufs, j, pkeys[j], log_bs, pids)
    ctx.sync()
    for q, b in zip(queries, bufs):
        assert np.array_equal(b.download(), expand_oracle(rns, q, levels, keys, log_bs))
    for key in pkeys:
        key.close()
