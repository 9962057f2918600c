"""The lazy absorb path on the GPU against the oracle, word for word:
LazyRnsPolynomial's FusedMulAddInPlace, FusedMulSumAddInPlace and
FusedMulDifferenceAddInPlace (rns/lazy_rns_polynomial.cc:89-254) as the
device applies them (SHG_OP_FMA and shg_fused_mul_sum_add on one
accumulator), against the oracle's 128-bit lazy sum and Export, including
moduli near 2^60 where the reference refreshes every 2^7 terms."""
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


@pytest.mark.parametrize("log_n,L,log_bound,num_terms,count", [(4, 2, 60, 150, 1),
                                                                (12, 4, 50, 9, 3),
                                                                (14, 8, 50, 6, 2)])
def test_lazy_mixed_updates(ctx, log_n, L, log_bound, num_terms, count):
    qs = ntt_primes(log_n, L, log_bound)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    kinds = ["fma", "sum", "diff"]
    acc = ctx.array(np.zeros((count, L, rns.n), dtype=np.uint64))
    expect_terms = [[] for _ in range(count)]
    for t in range(num_terms):
        a, b, c = (rns.uniform(100 + 3 * t + j, count) for j in range(3))
        if t % 5 == 0:
            for i, q in enumerate(qs):
                a[:, i, 0] = q  # NegateInPlace(0) leaves q
                b[:, i, 1] = q
        kind = kinds[t % 3]
        da, db, dc = ctx.array(a), ctx.array(b), ctx.array(c)
        if kind == "fma":
            ctx.pointwise(sg.OP_FMA, acc, da, db, count, pids)
        else:
            ctx.fused_mul_sum_add(acc, da, db, dc, kind == "diff", count, pids)
        for k in range(count):
            expect_terms[k].append((kind, a[k], b[k]) if kind == "fma" else (kind, a[k], b[k], c[k]))
    got = acc.download()
    for k in range(count):
        assert np.array_equal(got[k], rns.lazy_mixed(expect_terms[k])), k


def test_fused_mul_sum_add_refuses_wide_moduli(ctx):
    qs = ntt_primes(4, 1, 62)
    pids = ctx.primes(qs, 4)
    x = ctx.array(np.zeros((1, 1, 16), dtype=np.uint64))
    with pytest.raises(sg.ShgError, match="too large to perform fused multiply-sum-add"):
        ctx.fused_mul_sum_add(x, x, x, x, False, 1, pids)
