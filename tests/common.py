"""Shared helpers for the test suite: NTT-friendly primes, seeded inputs, and a
small BGV harness (key generation, encryption, decryption) built on the CPU
oracle, used to check that homomorphic results decrypt correctly.

The BGV conventions follow the reference:
  encryption  c0 = t*e + m + a*s, c1 = -a      (rns/rns_secret_key.h:188-231)
  decryption  sum_i c_i * s^i, then centered mod Q, then mod t
                                               (rns/rns_secret_key.cc:59-84)
  relin key   b_j = -a_j*s + t*e_j + g_j*s^2   (rns/rns_relinearization_key.cc:149-172)
  galois key  b_j = -a_j*s + t*e_j + g_j*s(X^p) (rns/rns_galois_key.cc:146-199)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if os.path.join(ROOT, "oracle") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402


def is_prime(n):
    """Deterministic Miller-Rabin for n < 3.3e24."""
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41]
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


_PRIME_CACHE = {}


def ntt_primes(log_n, count, log_bound=50):
    """The `count` largest primes below 2^log_bound with q = 1 mod 2^(log_n+1),
    in descending order (the configs' modulus chains)."""
    key = (log_n, count, log_bound)
    if key not in _PRIME_CACHE:
        step = 1 << (log_n + 1)
        q = ((1 << log_bound) - 1) // step * step + 1
        if q >= (1 << log_bound):
            q -= step
        out = []
        while len(out) < count:
            if is_prime(q):
                out.append(q)
            q -= step
        _PRIME_CACHE[key] = out
    return list(_PRIME_CACHE[key])


def rng(seed):
    return np.random.default_rng(seed)


def centered_binomial(r, n, variance):
    """Centred binomial samples of the given variance (sum of 2*variance
    fair bits minus variance), as in sample_error.h:43-100."""
    bits = r.integers(0, 2, size=(n, 2 * variance))
    return bits[:, :variance].sum(axis=1) - bits[:, variance:].sum(axis=1)


class Bgv:
    """BGV over an oracle RNS chain.  Everything lives in NTT form."""

    def __init__(self, rns, t, seed, variance=8):
        self.rns = rns
        self.t = int(t)
        self.variance = variance
        self.r = rng(seed)
        s = centered_binomial(self.r, rns.n, variance)
        self.s_coeffs = s
        self.s = rns.ntt_fwd(rns.import_signed(s))
        self._seed = seed

    def error(self):
        e = centered_binomial(self.r, self.rns.n, self.variance)
        return self.rns.ntt_fwd(self.rns.import_signed(e))

    def uniform(self):
        return self.rns.uniform(int(self.r.integers(0, 2**62)))

    def encode(self, m):
        """Plaintext coefficients in [0, t) -> NTT-form RNS polynomial,
        reading values above t/2 as negative (balanced)."""
        m = np.asarray(m, dtype=np.int64) % self.t
        bal = np.where(m > self.t // 2, m - self.t, m)
        return self.rns.ntt_fwd(self.rns.import_signed(bal))

    def encrypt(self, m):
        rns = self.rns
        a = self.uniform()
        e = self.error()
        c0 = rns.mul_int(e, self.t)
        c0 = rns.pointwise("add", c0, self.encode(m))
        c0 = rns.pointwise("fma", c0, a, self.s)
        c1 = rns.pointwise("neg", a)
        return np.stack([c0, c1])

    def decrypt(self, comps, power=1):
        rns = self.rns
        s = self.s if power == 1 else rns.substitute(self.s, power)
        acc = np.array(comps[0], dtype=np.uint64, copy=True)
        sp = s.copy()
        for i in range(1, len(comps)):
            if i > 1:
                sp = rns.pointwise("mul", sp, s)
            acc = rns.pointwise("add", acc, rns.pointwise("mul", comps[i], sp))
        coeffs = rns.crt_centered(rns.ntt_inv(acc))
        return np.array([c % self.t for c in coeffs], dtype=np.int64)

    def _keyswitch_key(self, target, log_bs):
        rns = self.rns
        k, _ = rns.gadget_dims(log_bs)
        gs = rns.gadget_components(log_bs, rns.q_hats(), rns.q_hat_invs())
        kas, kbs = [], []
        for j in range(k):
            a = self.uniform()
            e = self.error()
            b = rns.mul_int(e, self.t)
            b = rns.pointwise("add", b, rns.mul_scalar(target, gs[j]))
            b = rns.pointwise("fma", b, rns.pointwise("neg", a), self.s)
            kas.append(a)
            kbs.append(b)
        return np.stack(kas), np.stack(kbs)

    def relin_key(self, log_bs):
        s2 = self.rns.pointwise("mul", self.s, self.s)
        return self._keyswitch_key(s2, log_bs)

    def galois_key(self, power, log_bs):
        sp = self.rns.substitute(self.s, power)
        return self._keyswitch_key(sp, log_bs)


def negacyclic_mul_mod_t(a, b, t):
    """Schoolbook product in Z_t[X]/(X^n+1) (testing/coefficient_polynomial.h:144-175)."""
    n = len(a)
    full = np.convolve(np.asarray(a, dtype=object), np.asarray(b, dtype=object))
    out = [0] * n
    for i, v in enumerate(full):
        if i < n:
            out[i] += v
        else:
            out[i - n] -= v
    return np.array([int(v) % t for v in out], dtype=np.int64)


def negacyclic_mul_mod_q(a, b, q):
    """Exact schoolbook negacyclic product of residue vectors modulo q, using
    17-bit splits so that every partial convolution fits in int64."""
    n = len(a)
    a = [int(v) for v in a]
    b = [int(v) for v in b]
    mask = (1 << 17) - 1
    ap = [np.array([(v >> (17 * k)) & mask for v in a], dtype=np.int64) for k in range(3)]
    bp = [np.array([(v >> (17 * k)) & mask for v in b], dtype=np.int64) for k in range(3)]
    acc = [0] * (2 * n - 1)
    for i in range(3):
        for j in range(3):
            c = np.convolve(ap[i], bp[j])
            sh = 17 * (i + j)
            for idx in range(2 * n - 1):
                acc[idx] += int(c[idx]) << sh
    out = []
    for i in range(n):
        v = acc[i] - (acc[i + n] if i + n < 2 * n - 1 else 0)
        out.append(v % q)
    return out


def substitute_coeffs(vals, power, q):
    """f(X) -> f(X^power) mod (X^n + 1, q) on coefficient values."""
    n = len(vals)
    out = [0] * n
    for j, v in enumerate(vals):
        e = (j * power) % (2 * n)
        if e < n:
            out[e] = (out[e] + int(v)) % q
        else:
            out[e - n] = (out[e - n] - int(v)) % q
    return out


# Oblivious expansion and the database inner product composed from oracle
# steps exactly as ObliviousExpander::ExpandCiphertext orders them
# (oblivious_expand.h:173-227).

def monomial_ntt(rns, j):
    """-X^(n - 2^j) = X^(-2^j) in NTT form (Montgomery representations)."""
    x = np.zeros((len(rns.qs), rns.n), dtype=np.uint64)
    for i, p in enumerate(rns.primes):
        x[i, rns.n - (1 << j)] = p.mont.imp(rns.qs[i] - 1)
    return rns.ntt_fwd(x)


def expand_oracle(rns, ct, levels, keys, log_bs):
    n = rns.n
    res = [None] * (1 << levels)
    res[0] = [ct[0], ct[1]]
    for j in range(levels):
        p = n // (1 << j) + 1
        mono = monomial_ntt(rns, j)
        kas, kbs = keys[j]
        for k in range(1 << j):
            s0, s1 = rns.substitute(res[k][0], p), rns.substitute(res[k][1], p)
            g0, g1 = rns.galois_apply(s0, s1, kas, kbs, log_bs)
            d = [rns.pointwise("sub", res[k][0], g0), rns.pointwise("sub", res[k][1], g1)]
            res[k + (1 << j)] = [rns.pointwise("mul", d[0], mono), rns.pointwise("mul", d[1], mono)]
            res[k] = [rns.pointwise("add", res[k][0], g0), rns.pointwise("add", res[k][1], g1)]
    return np.stack([np.stack(r) for r in res])


def inner_oracle(rns, expd, db, C):
    K = expd.shape[0]
    out = []
    for c in range(C):
        acc = [np.zeros_like(expd[0, 0]), np.zeros_like(expd[0, 0])]
        for r in range(K):
            for comp in range(2):
                acc[comp] = rns.pointwise("add", acc[comp],
                                          rns.pointwise("mul", expd[r, comp], db[r * C + c]))
        out.append(np.stack(acc))
    return np.stack(out)
