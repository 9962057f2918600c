"""Pins the CPU oracle (oracle/shell_oracle.c) before it is trusted as the
checker of the HIP path.  The reference ships no golden vectors for this path
(its tests are property tests over random data), so each oracle function is
pinned by closed forms, invariants, Python big-integer arithmetic and brute
force, chosen so that a plausible mistake (a dropped term, a wrong index or
sign, a transposed operand) fails at least one check.  The mirrored reference
tests are cited beside each check.
"""
import numpy as np
import pytest

from common import (Bgv, negacyclic_mul_mod_q, negacyclic_mul_mod_t,
                    ntt_primes, substitute_coeffs)
import oracle

R = 2**64
Q50 = 1125899906826241  # largest prime < 2^50 with q = 1 mod 8192


def brv(i, log_n):
    return int(format(i, "0%db" % log_n)[::-1], 2) if log_n else 0


# ---------------------------------------------------------------------------
# Montgomery arithmetic (montgomery_test.cc:115-127,342-534,411-439)
# ---------------------------------------------------------------------------

MODULI = [17, 97, 257, 12289, 65537, 0x3FFFFFFFFFFFFFFF & ~2 | 1, Q50,
          1125899904679937, 4611686018427387847, 576460752303423433]


@pytest.mark.parametrize("q", [m for m in MODULI if m < 2**62 and m % 2 == 1])
def test_mont_params_closed_form(q):
    m = oracle.Mont(q)
    s = m.s
    assert s.modulus == q
    assert s.log_modulus == q.bit_length()
    assert s.r_mod_modulus == R % q
    assert R * s.inv_r - q * s.inv_modulus == 1
    assert 0 <= s.inv_r < q
    assert s.barrett_numerator == R // q
    assert s.barrett_numerator_big_lo + (s.barrett_numerator_big_hi << 64) == 2**127 // q
    assert s.r_mod_modulus_barrett == ((R % q) << 64) // q
    assert s.inv_r_barrett == (s.inv_r << 64) // q


def test_mont_rejects_bad_moduli():
    with pytest.raises(ValueError):
        oracle.Mont(2**62 + 1)
    with pytest.raises(ValueError):
        oracle.Mont(1 << 40)


@pytest.mark.parametrize("q", [17, 97, 257])
def test_mont_exhaustive_small(q):
    m = oracle.Mont(q)
    rinv = pow(R, -1, q)
    for a in range(q):
        assert m.imp(a) == a * R % q
        assert m.exp(m.imp(a)) == a
        for b in range(q):
            assert m.mul(a, b) == a * b * rinv % q
            assert m.add(a, b) == (a + b) % q
            assert m.sub(a, b) == (a - b) % q
    assert m.neg(0) == q  # NegateInPlace quirk (montgomery.h:396-399)


@pytest.mark.parametrize("q", [Q50, 1125899904679937, 4611686018427387847])
def test_mont_random_vs_bigint(q):
    m = oracle.Mont(q)
    r = np.random.default_rng(q % 1000)
    rinv = pow(R, -1, q)
    for _ in range(2000):
        a = int(r.integers(0, q))
        b = int(r.integers(0, q))
        acc = int(r.integers(0, q))
        assert m.mul(a, b) == a * b * rinv % q
        assert m.fma(acc, a, b) == (acc + a * b * rinv) % q
        c, cb = m.get_constant(b)
        assert c == b * rinv % q and cb == (c << 64) // q
        # MulConstant == Mul (montgomery_test.cc:411-439), also for x < 4q
        assert m.mul_const(a, c, cb) == m.mul(a, b)
        if 4 * q < 2**64:
            x = a + 3 * q
            assert m.mul_const(x, c, cb) == x * c % q
        assert m.imp(a) == a * R % q
        assert m.exp(a) == a * rinv % q
        assert m.add(a, b) == (a + b) % q
        assert m.sub(a, b) == (a - b) % q
        big = int(r.integers(0, 2**62)) * int(r.integers(0, 2**62))
        assert m.barrett_reduce_big(big) == big % q
    assert m.inverse(m.imp(5)) == m.imp(pow(5, -1, q))
    assert m.modexp(m.imp(3), 1000) == m.imp(pow(3, 1000, q))


# ---------------------------------------------------------------------------
# NTT parameters (ntt_parameters_test.cc:71-286)
# ---------------------------------------------------------------------------

def psi_by_rule(q, log_n):
    """Smallest t >= 2 whose power t^((q-1)/2n) has order exactly 2n."""
    two_n = 2 << log_n
    k = (q - 1) // two_n
    t = 2
    while True:
        c = pow(t, k, q)
        if pow(c, two_n // 2, q) != 1:
            return t, c
        t += 1


@pytest.mark.parametrize("log_n,q", [(3, 17), (4, 97), (11, Q50), (12, Q50),
                                     (13, 1125899905351681), (14, 1125899904679937)])
def test_ntt_tables(log_n, q):
    p = oracle.Prime(q, log_n)
    tb = p.tables()
    n = 1 << log_n
    rinv = pow(R, -1, q)
    t, psi = psi_by_rule(q, log_n)
    assert tb["psi"] * rinv % q == psi
    assert pow(psi, n, q) == q - 1
    for i in range(n):
        assert tb["bitrevs"][i] == brv(i, log_n)
    pb = [int(v) * rinv % q for v in tb["psis_bitrev"]]
    pib = [int(v) * rinv % q for v in tb["psis_inv_bitrev"]]
    for i in range(n):
        assert pb[i] == pow(psi, brv(i, log_n), q)
        assert pib[i] * pow(psi, brv(i, log_n) + 1, q) % q == 1
        c = int(tb["psis_bitrev_c"][i])
        assert c == pb[i] and int(tb["psis_bitrev_cb"][i]) == (c << 64) // q
        c2 = int(tb["psis_inv_bitrev_c"][i])
        assert c2 == pib[i] and int(tb["psis_inv_bitrev_cb"][i]) == (c2 << 64) // q
    assert tb["n_inv"] * rinv % q == pow(n, -1, q)


def test_psi_appendix_values():
    # Values recorded for the configs' first primes (computed independently).
    for log_n, q, t_exp, psi_exp in [(12, 1125899906826241, 11, 1094367314627419),
                                     (13, 1125899906826241, 11, 971709613313032),
                                     (14, 1125899904679937, 3, 974161070679573)]:
        t, psi = psi_by_rule(q, log_n)
        assert (t, psi) == (t_exp, psi_exp)
        p = oracle.Prime(q, log_n)
        assert p.tables()["psi"] * pow(R, -1, q) % q == psi_exp


def test_prime_rejects():
    with pytest.raises(ValueError):
        oracle.Prime(Q50, 16)
    with pytest.raises(ValueError):
        oracle.Prime(97, 6)  # 97 != 1 mod 128


# ---------------------------------------------------------------------------
# NTT / INTT (dft_transformations_test.cc:84-140, polynomial_test.cc:310-336)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("log_n,q", [(1, 17), (2, 17), (4, 97), (6, 257), (8, 12289),
                                     (10, Q50)])
def test_ntt_brute_force(log_n, q):
    """Output i is the evaluation at psi^(2*brv(i)+1) (bit-reversed order)."""
    p = oracle.Prime(q, log_n)
    n = 1 << log_n
    psi = p.tables()["psi"] * pow(R, -1, q) % q
    r = np.random.default_rng(log_n)
    x = np.array([int(r.integers(0, q)) for _ in range(n)], dtype=np.uint64)
    y = p.ntt_fwd(x)
    for i in range(n):
        e = 2 * brv(i, log_n) + 1
        ev = sum(int(x[j]) * pow(psi, e * j, q) for j in range(n)) % q
        assert int(y[i]) == ev
    assert np.array_equal(p.ntt_inv(y), x)


@pytest.mark.parametrize("log_n,q", [(11, Q50), (12, Q50), (13, 1125899905351681),
                                     (14, 1125899904679937), (15, 1125899904679937)])
def test_ntt_closed_forms_and_round_trip(log_n, q):
    p = oracle.Prime(q, log_n)
    n = 1 << log_n
    a = 123456789 % q
    delta = np.zeros(n, dtype=np.uint64)
    delta[0] = a
    assert np.all(p.ntt_fwd(delta) == a)
    assert np.array_equal(p.ntt_inv(np.full(n, a, dtype=np.uint64)), delta)
    x = np.random.default_rng(7).integers(0, q, size=n, dtype=np.uint64)
    y = p.ntt_fwd(x)
    assert np.all(y < q)
    assert np.array_equal(p.ntt_inv(y), x)


def test_negacyclic_product_schoolbook():
    """Config-1 shape: n=4096 product via NTT * NTT -> INTT equals the
    schoolbook negacyclic product (values in the Montgomery domain)."""
    q = Q50
    p = oracle.Prime(q, 12)
    m = p.mont
    r = np.random.default_rng(11)
    a = r.integers(0, q, size=4096, dtype=np.uint64)
    b = r.integers(0, q, size=4096, dtype=np.uint64)
    fa, fb = p.ntt_fwd(a), p.ntt_fwd(b)
    prod = np.array([m.mul(int(u), int(v)) for u, v in zip(fa, fb)], dtype=np.uint64)
    c = p.ntt_inv(prod)
    rinv = pow(R, -1, q)
    # Montgomery reps: rep(a)*rep(b)*R^-1 = rep(a*b), so the product of the
    # representations, scaled by R^-1, is expected.
    expect = negacyclic_mul_mod_q(a, b, q)
    assert [int(v) for v in c] == [v * rinv % q for v in expect]


# ---------------------------------------------------------------------------
# Substitute (rns/rns_polynomial_test.cc:1088-1109)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("log_n,power", [(4, 1), (4, 3), (4, 31), (11, 5), (11, 2047)])
def test_substitute_closed_form(log_n, power):
    qs = ntt_primes(log_n, 3)
    rns = oracle.Rns(qs, log_n)
    x = rns.uniform(99)
    got = rns.substitute(rns.ntt_fwd(x), power)
    for i, q in enumerate(qs):
        sub = substitute_coeffs([int(v) for v in x[i]], power, q)
        exp = rns.primes[i].ntt_fwd(np.array(sub, dtype=np.uint64))
        assert np.array_equal(got[i], exp)
    with pytest.raises(ValueError):
        rns.substitute(x, 2)
    with pytest.raises(ValueError):
        rns.substitute(x, 2 * rns.n + 1)


def test_substitute_composition():
    rns = oracle.Rns(ntt_primes(10, 2), 10)
    x = rns.ntt_fwd(rns.uniform(5))
    a = rns.substitute(rns.substitute(x, 5), 7)
    b = rns.substitute(x, 35)
    assert np.array_equal(a, b)
    assert np.array_equal(rns.substitute(x, 1), x)


# ---------------------------------------------------------------------------
# Modulus conversion, gadget (rns/rns_gadget_test.cc:174-242)
# ---------------------------------------------------------------------------

def test_convert_modulus_balanced_closed_form():
    qa, qb = ntt_primes(11, 2)
    ma, mb = oracle.Mont(qa), oracle.Mont(qb)
    r = np.random.default_rng(3)
    vals = r.integers(0, qa, size=512, dtype=np.uint64)
    vals[:4] = [0, qa // 2, qa // 2 + 1, qa - 1]
    for src, dst in ((ma, mb), (mb, ma)):
        reps = np.array([src.imp(int(v) % src.q) for v in vals], dtype=np.uint64)
        out = np.zeros_like(reps)
        oracle.lib().so_convert_modulus_balanced(src.ref(), dst.ref(), oracle._p64(reps),
                                                 oracle._p64(out), len(reps))
        for v, o in zip(vals, out):
            v = int(v) % src.q
            signed = v - src.q if v > src.q // 2 else v
            assert int(o) == (signed % dst.q) * R % dst.q


@pytest.mark.parametrize("log_bs", [[10, 10, 10], [4, 7, 20]])
def test_gadget_reconstruction(log_bs):
    qs = ntt_primes(9, 3)
    rns = oracle.Rns(qs, 9)
    a = rns.uniform(17)
    k, dims = rns.gadget_dims(log_bs)
    assert dims == [-(-q.bit_length() // b) for q, b in zip(qs, log_bs)]
    digits = rns.gadget_decompose(a, log_bs)
    gs = rns.gadget_components(log_bs, rns.q_hats(), rns.q_hat_invs())
    # digits are unsigned base-B digits of the exported residue, lifted to
    # every limb
    row = 0
    for i, q in enumerate(qs):
        vals = rns.export_values(a)[i]
        for d in range(dims[i]):
            dv = rns.export_values(digits[row])
            for l in range(len(qs)):
                assert all(int(dv[l][j]) == (int(vals[j]) >> (log_bs[i] * d)) & ((1 << log_bs[i]) - 1)
                           for j in range(rns.n))
            row += 1
    # <g, g^-1(a)> = a
    acc = np.zeros_like(a)
    for j in range(k):
        acc = rns.pointwise("add", acc, rns.mul_scalar(digits[j], gs[j]))
    assert np.array_equal(acc, a)


# ---------------------------------------------------------------------------
# ModReduceLsb (rns/rns_polynomial_test.cc:1150-1332)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("is_ntt", [True, False])
def test_mod_reduce_lsb_closed_form(is_ntt):
    log_n, t = 8, 65537
    qs = ntt_primes(log_n, 4)
    rns = oracle.Rns(qs, log_n)
    a = rns.uniform(23)
    x = rns.ntt_fwd(a) if is_ntt else a
    out = rns.mod_reduce_lsb(x, t, rns.ql_inv(), is_ntt=is_ntt)
    red = rns.sub(3)
    if is_ntt:
        out = red.ntt_inv(out)
    Q = qs[0] * qs[1] * qs[2] * qs[3]
    Qr = Q // qs[3]
    ql = qs[3]
    r_, k_ = ql % t, (ql - ql % t) // t
    ints = rns.crt_centered(a)
    lv = rns.export_values(a)[3]
    got = red.crt_centered(out)
    for j in range(rns.n):
        d = (k_ * int(lv[j])) % ql
        d = d - ql if d > ql // 2 else d
        A = r_ * ints[j] + t * d
        assert A % ql == 0
        expect = (A // ql) % Qr
        assert got[j] % Qr == expect


def test_mod_reduce_small_to_zero():
    """t * small polynomial reduces to a multiple of t that stays small."""
    log_n, t = 6, 65537
    rns = oracle.Rns(ntt_primes(log_n, 3), log_n)
    e = np.random.default_rng(1).integers(-5, 6, size=rns.n)
    x = rns.mul_int(rns.import_signed(e), t)
    out = rns.mod_reduce_lsb(x, t, rns.ql_inv(), is_ntt=False)
    got = rns.sub(2).crt_centered(out)
    assert all(v % t == 0 and abs(v) < 10 * t for v in got)


# ---------------------------------------------------------------------------
# Lazy FMA (rns/lazy_rns_polynomial_test.cc:393-485)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("terms", [1, 2, 17, 64])
def test_lazy_equals_eager(terms):
    rns = oracle.Rns(ntt_primes(8, 2), 8)
    as_ = rns.uniform(3, terms)
    bs = rns.uniform(4, terms)
    eager = np.zeros((rns.L, rns.n), dtype=np.uint64)
    for i in range(terms):
        eager = rns.pointwise("fma", eager, as_[i], bs[i])
    assert np.array_equal(rns.lazy_fma(as_, bs), eager)
    assert rns.primes[0].mont.s.log_modulus == 50
    assert oracle.lib().so_lazy_log_max_level(rns.primes[0].mont.ref()) == 27


# ---------------------------------------------------------------------------
# Ciphertext-level: tensor product, relinearization, Galois, mod switch
# (rns_bgv_ciphertext_test.cc:293-329,660-713;
#  rns_relinearization_key_test.cc:424-500; rns_galois_key_test.cc:401-462)
# ---------------------------------------------------------------------------

def test_bgv_pipeline_decrypts():
    log_n, t = 9, 65537
    qs = ntt_primes(log_n, 4)
    rns = oracle.Rns(qs, log_n)
    bgv = Bgv(rns, t, seed=5)
    r = np.random.default_rng(8)
    m1 = r.integers(0, t, size=rns.n)
    m2 = r.integers(0, t, size=rns.n)
    c1, c2 = bgv.encrypt(m1), bgv.encrypt(m2)
    assert np.array_equal(bgv.decrypt(c1), m1 % t)
    prod = rns.tensor_product(c1, c2)
    expect = negacyclic_mul_mod_t(m1, m2, t)
    assert np.array_equal(bgv.decrypt(prod), expect)
    log_bs = [10] * len(qs)
    kas, kbs = bgv.relin_key(log_bs)
    r0, r1 = rns.relin_apply(prod, kas, kbs, log_bs)
    relin = np.stack([r0, r1])
    assert np.array_equal(bgv.decrypt(relin), expect)
    # Galois x -> x^5 then key switch back to s
    sub = np.stack([rns.substitute(c, 5) for c in relin])
    assert np.array_equal(bgv.decrypt(sub, power=5),
                          np.array(substitute_coeffs(expect, 5, t), dtype=np.int64))
    gas, gbs = bgv.galois_key(5, log_bs)
    g0, g1 = rns.galois_apply(sub[0], sub[1], gas, gbs, log_bs)
    assert np.array_equal(bgv.decrypt(np.stack([g0, g1])),
                          np.array(substitute_coeffs(expect, 5, t), dtype=np.int64))
    # mod switch dropping the last prime preserves the plaintext
    ms = np.stack([rns.mod_reduce_lsb(c, t, rns.ql_inv(), True) for c in (g0, g1)])
    red = rns.sub(3)
    bgv_r = Bgv.__new__(Bgv)
    bgv_r.rns, bgv_r.t = red, t
    bgv_r.s = bgv.s[:3].copy()
    assert np.array_equal(bgv_r.decrypt(ms),
                          np.array(substitute_coeffs(expect, 5, t), dtype=np.int64))


def test_relin_key_index_quirk_degree3():
    """Degree-3 relinearization uses key index (i-2)*l + j
    (rns/rns_relinearization_key.cc:222); with l = 3, the second block
    starts at index 3, not k."""
    log_n = 5
    qs = ntt_primes(log_n, 2)
    rns = oracle.Rns(qs, log_n)
    log_bs = [25, 25]
    k, _ = rns.gadget_dims(log_bs)
    comps = rns.uniform(1, 4)
    kas = rns.uniform(2, 2 * k)
    kbs = rns.uniform(3, 2 * k)
    c0, c1 = rns.relin_apply(comps, kas, kbs, log_bs)
    e0, e1 = comps[0].copy(), comps[1].copy()
    for i in (2, 3):
        ci = rns.ntt_inv(comps[i])
        dg = rns.gadget_decompose(ci, log_bs)
        for j in range(k):
            d = rns.ntt_fwd(dg[j])
            idx = (i - 2) * 3 + j
            e0 = rns.pointwise("fma", e0, d, kbs[idx])
            e1 = rns.pointwise("fma", e1, d, kas[idx])
    assert np.array_equal(c0, e0) and np.array_equal(c1, e1)


# ---------------------------------------------------------------------------
# The reference's own test moduli and Montgomery constant (tests/golden/
# reference_moduli.json, copied from shell_encryption/constants.h and
# shell_encryption/rns/testing/parameters.h).
# ---------------------------------------------------------------------------

import json
import os

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_moduli.json")))


def test_reference_montgomery_kat():
    for e in GOLDEN["montgomery_kat"]:
        assert oracle.Mont(e["q"]).s.inv_modulus == e["inv_modulus"], e["source"]


@pytest.mark.parametrize("entry", GOLDEN["moduli"], ids=lambda e: "%d-%d" % (e["q"], e["log_n"]))
def test_reference_moduli_ntt(entry):
    """Tables follow the psi rule and sampled outputs are the evaluations at
    psi^(2*brv(i)+1), for every (q, log n) pair the reference tests use."""
    q, log_n = entry["q"], entry["log_n"]
    p = oracle.Prime(q, log_n)
    n = 1 << log_n
    t, psi = psi_by_rule(q, log_n)
    assert p.tables()["psi"] * pow(R, -1, q) % q == psi
    x = np.random.default_rng(q % 1000).integers(0, q, size=n, dtype=np.uint64)
    y = p.ntt_fwd(x)
    for i in (0, 1, n // 2 - 1, n // 2, n - 1):
        e = 2 * brv(i, log_n) + 1
        w = pow(psi, e, q)
        acc, wp = 0, 1
        for j in range(n):
            acc += int(x[j]) * wp
            wp = wp * w % q
        assert int(y[i]) == acc % q
    assert np.array_equal(p.ntt_inv(y), x)
