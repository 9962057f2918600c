"""Pins the oracle's LazyRnsPolynomial restatement over mixed updates
(rns/lazy_rns_polynomial.h:44-115, rns/lazy_rns_polynomial.cc:89-254) with
closed forms in Python integers: with R = 2^64, Export of the lazy sum is

    sum_t  a_t b_t R^-1               (FusedMulAddInPlace)
         + (a_t + b_t) c_t R^-1       (FusedMulSumAddInPlace)
         + (a_t - b_t) c_t R^-1       (FusedMulDifferenceAddInPlace)   mod q,

canonical.  Moduli near 2^60 hold the lazy level at 2^7, so 200 terms pass
through several refreshes; the reference's own lazy tests compare with eager
FusedMulAddInPlace the same way (rns/lazy_rns_polynomial_test.cc:393-485).
Entries equal to q (NegateInPlace(0)) are included in the operands."""
import numpy as np
import pytest

import oracle
from common import ntt_primes


def _closed(rns, terms):
    out = np.zeros((rns.L, rns.n), dtype=np.uint64)
    for i, q in enumerate(rns.qs):
        rinv = pow(1 << 64, -1, q)
        for k in range(rns.n):
            s = 0
            for t in terms:
                a, b = int(t[1][i, k]), int(t[2][i, k])
                if t[0] == "fma":
                    s += a * b
                elif t[0] == "sum":
                    s += (a + b) * int(t[3][i, k])
                else:
                    s += (a - b) * int(t[3][i, k])
            out[i, k] = s * rinv % q
    return out


@pytest.mark.parametrize("log_bound,num_terms", [(50, 5), (60, 200)])
def test_lazy_mixed_closed_form(log_bound, num_terms):
    log_n = 4
    rns = oracle.Rns(ntt_primes(log_n, 2, log_bound), log_n)
    kinds = ["fma", "sum", "diff"]
    terms = []
    for t in range(num_terms):
        a, b, c = rns.uniform(10 + 3 * t), rns.uniform(11 + 3 * t), rns.uniform(12 + 3 * t)
        if t % 7 == 0:
            for i, q in enumerate(rns.qs):
                a[i, 0] = q  # NegateInPlace(0) leaves q
                b[i, 1] = q
        kind = kinds[t % 3]
        terms.append((kind, a, b) if kind == "fma" else (kind, a, b, c))
    got = rns.lazy_mixed(terms)
    assert np.array_equal(got, _closed(rns, terms))
    assert all(int(v) < q for i, q in enumerate(rns.qs) for v in got[i])
    if log_bound == 60:
        assert oracle.lib().so_lazy_log_max_level(rns.primes[0].mont.ref()) == 7


def test_lazy_mixed_fma_only_matches_lazy_fma():
    rns = oracle.Rns(ntt_primes(6, 3), 6)
    as_, bs = rns.uniform(1, 9), rns.uniform(2, 9)
    terms = [("fma", as_[t], bs[t]) for t in range(9)]
    assert np.array_equal(rns.lazy_mixed(terms), rns.lazy_fma(as_, bs))
