"""shg_relin_apply_noalloc (the key switch that allocates nothing, with its
work space in a fixed per-GPU area) against shg_relin_apply on the same
inputs, word for word, for two calls in a row with the second one smaller,
plus one ciphertext per case against the oracle."""
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


def _case(ctx, log_n, L, log_b, degree, ntt_mask, seed):
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    log_bs = [log_b] * L
    k, _ = rns.gadget_dims(log_bs)
    nkeys = (degree - 2) * degree + k
    kas, kbs = rns.uniform(seed, nkeys), rns.uniform(seed + 1, nkeys)
    return dict(rns=rns, pids=pids, log_bs=log_bs, nkeys=nkeys, kas=kas, kbs=kbs,
                dka=ctx.array(kas), dkb=ctx.array(kbs), degree=degree, mask=ntt_mask, L=L,
                n=1 << log_n, seed=seed)


def _both(ctx, c, count):
    rng = np.random.default_rng(c["seed"] + count)
    comps = np.empty((count, c["degree"] + 1, c["L"], c["n"]), dtype=np.uint64)
    for i, q in enumerate(c["rns"].qs):
        comps[:, :, i, :] = rng.integers(0, q, size=(count, c["degree"] + 1, c["n"]),
                                         dtype=np.uint64)
    dc = ctx.array(comps)
    outs = []
    for fn in (ctx.relin_apply, ctx.relin_apply_noalloc):
        o0 = ctx.array(np.full((count, c["L"], c["n"]), 7, dtype=np.uint64))
        o1 = ctx.array(np.full((count, c["L"], c["n"]), 7, dtype=np.uint64))
        fn(dc, c["degree"], c["mask"], c["dka"], c["dkb"], c["nkeys"], c["log_bs"], o0, o1, count,
           c["pids"])
        outs.append((o0.download(), o1.download()))
    (a0, a1), (b0, b1) = outs
    assert np.array_equal(a0, b0) and np.array_equal(a1, b1)
    e0, e1 = c["rns"].relin_apply(comps[count - 1], c["kas"], c["kbs"], c["log_bs"],
                                  ntt_mask=c["mask"])
    assert np.array_equal(b0[count - 1], e0) and np.array_equal(b1[count - 1], e1)


# (log n, L, log base, degree, NTT mask, counts): digit planes at the
# benchmark's shape in several chunks of the work area, the exported-copy
# path of bases above 2^16, and a degree-3 ciphertext whose component 2 is in
# coefficient form (both paths in one call).
CASES = [
    (14, 8, 10, 2, 0b111, (120, 20)),
    (13, 4, 20, 2, 0b111, (300, 5)),
    (12, 3, 10, 3, 0b1011, (40, 3)),
]


@pytest.mark.parametrize("log_n,L,log_b,degree,mask,counts", CASES)
def test_noalloc_matches_relin_apply(ctx, log_n, L, log_b, degree, mask, counts):
    c = _case(ctx, log_n, L, log_b, degree, mask, 7000 + log_n)
    for count in counts:
        _both(ctx, c, count)


def test_noalloc_rejects_unsupported_ring_and_leaves_outputs(ctx):
    c = _case(ctx, 15, 2, 10, 2, 0b111, 7100)
    comps = ctx.array(np.zeros((1, 3, 2, c["n"]), dtype=np.uint64))
    o0 = ctx.array(np.full((1, 2, c["n"]), 5, dtype=np.uint64))
    o1 = ctx.array(np.full((1, 2, c["n"]), 5, dtype=np.uint64))
    with pytest.raises(sg.ShgError) as e:
        ctx.relin_apply_noalloc(comps, 2, 0b111, c["dka"], c["dkb"], c["nkeys"], c["log_bs"], o0,
                                o1, 1, c["pids"])
    assert e.value.code == -6 and "2^14" in e.value.msg
    assert (o0.download() == 5).all() and (o1.download() == 5).all()
