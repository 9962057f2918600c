"""Pins the oracle's BFV-side RNS restatements with closed forms in big
integers, coefficient by coefficient (A = the CRT value in [0, Q)):

* ModReduceMsb: (A - [A]_{q_l}) / q_l with the balanced remainder, in both
  forms (rns/rns_polynomial.cc:349-417);
* SwitchRnsBasis: the centered representative of A (A - Q for A >= Q/2,
  the rounding of sum y_i / q_i picks it) mod p_j (rns/rns_polynomial.cc:419-517);
* ScaleAndSwitchRnsBasis: round(P * A / Q) mod p_j (:519-636);
* ApproxSwitchRnsBasis: A + e * Q mod p_j with one integer e per
  coefficient, 0 <= e < L (or |e| <= L/2 balanced) (:638-701).
"""
import numpy as np
import pytest

import oracle
from common import ntt_primes


def _values(rns, x):
    """Exported values of a (L, n) Montgomery array as Python ints."""
    return [[rns.primes[i].mont.exp(int(v)) for v in x[i]] for i in range(rns.L)]


def _crt(rns, vals, k):
    Q = 1
    for q in rns.qs:
        Q *= q
    A = 0
    for i, q in enumerate(rns.qs):
        Qi = Q // q
        A += vals[i][k] * Qi * pow(Qi, -1, q)
    return A % Q, Q


def _chains(log_n, L, K):
    qs = ntt_primes(log_n, L + K)
    return oracle.Rns(qs[:L], log_n), oracle.Rns(qs[L:], log_n)


@pytest.mark.parametrize("log_n,L", [(4, 2), (5, 3), (6, 5)])
@pytest.mark.parametrize("is_ntt", [False, True])
def test_mod_reduce_msb_closed_form(log_n, L, is_ntt):
    rns = oracle.Rns(ntt_primes(log_n, L), log_n)
    x = rns.uniform(40 + log_n)
    ql_inv = rns.ql_inv()
    out = rns.mod_reduce_msb(x, ql_inv, is_ntt)
    sub = rns.sub(L - 1)
    coeff_in = rns.ntt_inv(x) if is_ntt else x
    coeff_out = sub.ntt_inv(out) if is_ntt else out
    vin, vout = _values(rns, coeff_in), _values(sub, coeff_out)
    ql = rns.qs[-1]
    for k in range(rns.n):
        A, _ = _crt(rns, vin, k)
        al = A % ql
        if al > ql // 2:
            al -= ql
        b = (A - al) // ql
        for i, q in enumerate(sub.qs):
            assert vout[i][k] == b % q, (k, i)


@pytest.mark.parametrize("log_n,L,K", [(4, 2, 1), (5, 3, 2), (6, 4, 4)])
def test_switch_and_scale_rns_basis_closed_form(log_n, L, K):
    rns, out = _chains(log_n, L, K)
    c = rns.basis_constants(out)
    x = rns.uniform(70 + L)
    x[:, 0] = 0  # A = 0 and A = Q - 1 at the first two coefficients
    for i, q in enumerate(rns.qs):
        x[i, 1] = rns.primes[i].mont.imp(q - 1)
    sw = rns.switch_rns_basis(x, out, c)
    sc = rns.scale_and_switch_rns_basis(x, out, c)
    vin, vsw, vsc = _values(rns, x), _values(out, sw), _values(out, sc)
    P = 1
    for p in out.qs:
        P *= p
    for k in range(rns.n):
        A, Q = _crt(rns, vin, k)
        r = (2 * P * A + Q) // (2 * Q)  # round(P * A / Q), halves away from zero (A >= 0)
        Ac = A - Q if 2 * A >= Q else A
        for j, p in enumerate(out.qs):
            assert vsw[j][k] == Ac % p, (k, j)
            assert vsc[j][k] == r % p, (k, j)


@pytest.mark.parametrize("balanced", [False, True])
@pytest.mark.parametrize("is_ntt", [False, True])
def test_approx_switch_rns_basis_closed_form(balanced, is_ntt):
    rns, out = _chains(5, 4, 3)
    c = rns.basis_constants(out)
    x = rns.uniform(91)
    y = rns.approx_switch_rns_basis(x, out, c, is_ntt, balanced)
    vin = _values(rns, rns.ntt_inv(x) if is_ntt else x)
    vout = _values(out, y)
    es = range(-2, 3) if balanced else range(0, 4)
    for k in range(rns.n):
        A, Q = _crt(rns, vin, k)
        found = [e for e in es if all(vout[j][k] == (A + e * Q) % p for j, p in enumerate(out.qs))]
        assert len(found) == 1, k
