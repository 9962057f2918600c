"""ctypes front-end of the CPU oracle (oracle/shell_oracle.c).

TEST INFRASTRUCTURE ONLY.  The HIP path is tested against the oracle: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  The product library never routes through it.

Arrays are numpy uint64, limb-major: an RNS polynomial with L limbs and n
coefficients is an (L, n) array; a batch of polynomials adds leading axes.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

U64P = ctypes.POINTER(ctypes.c_uint64)
U32P = ctypes.POINTER(ctypes.c_uint32)
INTP = ctypes.POINTER(ctypes.c_int)


def build():
    """Compiles liboracle.so from shell_oracle.c (make in this directory)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


class _Mont(ctypes.Structure):
    _fields_ = [
        ("modulus", ctypes.c_uint64),
        ("log_modulus", ctypes.c_uint),
        ("r_mod_modulus", ctypes.c_uint64),
        ("r_mod_modulus_barrett", ctypes.c_uint64),
        ("inv_modulus", ctypes.c_uint64),
        ("inv_r", ctypes.c_uint64),
        ("inv_r_barrett", ctypes.c_uint64),
        ("barrett_numerator", ctypes.c_uint64),
        ("barrett_numerator_big_lo", ctypes.c_uint64),
        ("barrett_numerator_big_hi", ctypes.c_uint64),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        MP = ctypes.POINTER(_Mont)
        u64 = ctypes.c_uint64
        L.so_mont_init.argtypes = [u64, MP]
        for f in ("so_import", "so_export", "so_barrett_reduce", "so_neg"):
            getattr(L, f).argtypes = [MP, u64]
            getattr(L, f).restype = u64
        for f in ("so_mul", "so_add", "so_sub", "so_modexp"):
            getattr(L, f).argtypes = [MP, u64, u64]
            getattr(L, f).restype = u64
        L.so_inverse.argtypes = [MP, u64]
        L.so_inverse.restype = u64
        L.so_mul_const.argtypes = [MP, u64, u64, u64]
        L.so_mul_const.restype = u64
        L.so_fma.argtypes = [MP, u64, u64, u64]
        L.so_fma.restype = u64
        L.so_fma_const.argtypes = [MP, u64, u64, u64, u64]
        L.so_fma_const.restype = u64
        L.so_barrett_reduce_big.argtypes = [MP, u64, u64]
        L.so_barrett_reduce_big.restype = u64
        L.so_get_constant.argtypes = [MP, u64, U64P, U64P]
        L.so_prime_create.argtypes = [u64, ctypes.c_int, INTP]
        L.so_prime_create.restype = ctypes.c_void_p
        L.so_prime_destroy.argtypes = [ctypes.c_void_p]
        L.so_prime_mont.argtypes = [ctypes.c_void_p]
        L.so_prime_mont.restype = MP
        L.so_prime_tables.argtypes = [ctypes.c_void_p, U64P, U32P, U64P]
        L.so_ntt_fwd.argtypes = [ctypes.c_void_p, U64P]
        L.so_ntt_inv.argtypes = [ctypes.c_void_p, U64P]
        PP = ctypes.POINTER(ctypes.c_void_p)
        L.so_poly_pointwise.argtypes = [PP, ctypes.c_int, ctypes.c_int, U64P, U64P, U64P]
        L.so_poly_mul_scalar.argtypes = [PP, ctypes.c_int, U64P, U64P]
        L.so_poly_mul_int.argtypes = [PP, ctypes.c_int, U64P, u64]
        L.so_poly_ntt_fwd.argtypes = [PP, ctypes.c_int, U64P]
        L.so_poly_ntt_inv.argtypes = [PP, ctypes.c_int, U64P]
        L.so_poly_substitute.argtypes = [PP, ctypes.c_int, ctypes.c_int, U64P, U64P]
        L.so_poly_substitute.restype = ctypes.c_int
        L.so_convert_modulus_balanced.argtypes = [MP, MP, U64P, U64P, ctypes.c_int]
        L.so_poly_mod_reduce_lsb.argtypes = [PP, ctypes.c_int, u64, U64P, ctypes.c_int, U64P]
        L.so_poly_mod_reduce_lsb.restype = ctypes.c_int
        L.so_gadget_dims.argtypes = [PP, ctypes.c_int, INTP, INTP]
        L.so_gadget_dims.restype = ctypes.c_int
        L.so_gadget_components.argtypes = [PP, ctypes.c_int, INTP, U64P, U64P, U64P]
        L.so_gadget_decompose.argtypes = [PP, ctypes.c_int, INTP, U64P, U64P]
        L.so_relin_apply.argtypes = [PP, ctypes.c_int, INTP, ctypes.c_int, U64P,
                                     ctypes.c_int, U64P, U64P, U64P, U64P]
        L.so_relin_apply.restype = ctypes.c_int
        L.so_hwy_relin_apply.argtypes = L.so_relin_apply.argtypes
        L.so_hwy_relin_apply.restype = ctypes.c_int
        L.so_hwy_poly_ntt_fwd.argtypes = [PP, ctypes.c_int, U64P]
        L.so_galois_apply.argtypes = [PP, ctypes.c_int, INTP, U64P, U64P, ctypes.c_int,
                                      U64P, U64P, U64P, U64P]
        L.so_galois_apply.restype = ctypes.c_int
        L.so_tensor_product.argtypes = [PP, ctypes.c_int, ctypes.c_int, U64P,
                                        ctypes.c_int, U64P, U64P]
        L.so_lazy_log_max_level.argtypes = [MP]
        L.so_lazy_log_max_level.restype = ctypes.c_int
        L.so_lazy_fma.argtypes = [PP, ctypes.c_int, ctypes.c_int, U64P, U64P, U64P]
        L.so_lazy_mixed.argtypes = [PP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                    U64P, U64P, U64P, U64P]
        L.so_fill_uniform.argtypes = [PP, ctypes.c_int, u64, U64P]
        L.so_now_seconds.restype = ctypes.c_double
        U8P = ctypes.POINTER(ctypes.c_uint8)
        sz = ctypes.c_size_t
        L.so_transcribe_bits.argtypes = [U64P, sz, ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, U64P, sz]
        L.so_transcribe_bits.restype = ctypes.c_longlong
        L.so_serialize_vector.argtypes = [MP, U64P, sz, U8P, sz]
        L.so_serialize_vector.restype = ctypes.c_longlong
        L.so_deserialize_vector.argtypes = [MP, ctypes.c_int, U8P, sz, U64P]
        L.so_deserialize_vector.restype = ctypes.c_int
        L.so_chacha20_block.argtypes = [U8P, ctypes.c_uint32, U8P, U8P]
        L.so_chacha_prng_init.argtypes = [ctypes.c_void_p, U8P, sz]
        L.so_chacha_prng_init.restype = ctypes.c_int
        L.so_chacha_rand8.argtypes = [ctypes.c_void_p]
        L.so_chacha_rand8.restype = ctypes.c_uint8
        L.so_chacha_rand64.argtypes = [ctypes.c_void_p]
        L.so_chacha_rand64.restype = u64
        L.so_sample_uniform_chacha.argtypes = [PP, ctypes.c_int, ctypes.c_void_p, U64P]
        L.so_poly_mod_reduce_msb.argtypes = [PP, ctypes.c_int, U64P, ctypes.c_int, U64P]
        L.so_poly_mod_reduce_msb.restype = ctypes.c_int
        L.so_switch_rns_basis.argtypes = [PP, ctypes.c_int, PP, ctypes.c_int, U64P, U64P, U64P,
                                          U64P, U64P]
        L.so_scale_and_switch_rns_basis.argtypes = [PP, ctypes.c_int, PP, ctypes.c_int, U64P, U64P,
                                                    U64P, U64P, U64P]
        L.so_approx_switch_rns_basis.argtypes = [PP, ctypes.c_int, PP, ctypes.c_int, U64P, U64P,
                                                 ctypes.c_int, ctypes.c_int, U64P, U64P]
        _lib = L
    return _lib


U8P_ = ctypes.POINTER(ctypes.c_uint8)
CHACHA_OUTPUT_BYTES = 255 * 32


def _p8(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U8P_)


def transcribe_bits(values, in_type_bits, bit_length, in_bits, out_bits, out_type_bits):
    """TranscribeBits over integers given as a uint64 array."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    cap = max(1, -(-int(bit_length) // int(out_bits)))
    out = np.zeros(cap, dtype=np.uint64)
    r = lib().so_transcribe_bits(v.ctypes.data_as(U64P), v.size, int(in_type_bits),
                                 int(bit_length), int(in_bits), int(out_bits), int(out_type_bits),
                                 out.ctypes.data_as(U64P), cap)
    if r < 0:
        raise ValueError("TranscribeBits error %d" % r)
    return out[:r]


def chacha20_block(key, counter, nonce):
    k = np.frombuffer(bytes(key), dtype=np.uint8).copy()
    nn = np.frombuffer(bytes(nonce), dtype=np.uint8).copy()
    out = np.zeros(64, dtype=np.uint8)
    lib().so_chacha20_block(_p8(k), int(counter), _p8(nn), _p8(out))
    return out.tobytes()


class ChaChaPrng:
    """SingleThreadChaChaPrng restated (prng/single_thread_chacha_prng.cc)."""

    def __init__(self, key):
        self._buf = ctypes.create_string_buffer(32 + 8 + CHACHA_OUTPUT_BYTES + 64)
        k = np.frombuffer(bytes(key), dtype=np.uint8).copy()
        if lib().so_chacha_prng_init(ctypes.addressof(self._buf), _p8(k), k.size) != 0:
            raise ValueError("the ChaCha PRNG takes a 32-byte key")

    @property
    def h(self):
        return ctypes.addressof(self._buf)

    def rand8(self):
        return int(lib().so_chacha_rand8(self.h))

    def rand64(self):
        return int(lib().so_chacha_rand64(self.h))


def _p64(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U64P)


def _ints(v):
    arr = (ctypes.c_int * len(v))(*v)
    return arr


class Mont:
    """MontgomeryIntParams<Uint64> (montgomery.h:77-192)."""

    def __init__(self, q):
        self.s = _Mont()
        rc = lib().so_mont_init(q, ctypes.byref(self.s))
        if rc != 0:
            raise ValueError("invalid modulus %d (code %d)" % (q, rc))
        self.q = int(q)

    def ref(self):
        return ctypes.byref(self.s)

    def imp(self, x):
        return lib().so_import(self.ref(), int(x))

    def exp(self, x):
        return lib().so_export(self.ref(), int(x))

    def mul(self, a, b):
        return lib().so_mul(self.ref(), int(a), int(b))

    def add(self, a, b):
        return lib().so_add(self.ref(), int(a), int(b))

    def sub(self, a, b):
        return lib().so_sub(self.ref(), int(a), int(b))

    def neg(self, a):
        return lib().so_neg(self.ref(), int(a))

    def fma(self, acc, a, b):
        return lib().so_fma(self.ref(), int(acc), int(a), int(b))

    def mul_const(self, x, c, cb):
        return lib().so_mul_const(self.ref(), int(x), int(c), int(cb))

    def get_constant(self, w):
        c = ctypes.c_uint64()
        cb = ctypes.c_uint64()
        lib().so_get_constant(self.ref(), int(w), ctypes.byref(c), ctypes.byref(cb))
        return c.value, cb.value

    def modexp(self, b, e):
        return lib().so_modexp(self.ref(), int(b), int(e))

    def inverse(self, a):
        return lib().so_inverse(self.ref(), int(a))

    def barrett_reduce_big(self, v):
        v = int(v)
        return lib().so_barrett_reduce_big(self.ref(), v & (2**64 - 1), v >> 64)


class Prime:
    """PrimeModulus: Montgomery params + NttParameters for (q, log_n)."""

    def __init__(self, q, log_n):
        st = ctypes.c_int()
        self.h = lib().so_prime_create(int(q), int(log_n), ctypes.byref(st))
        if not self.h:
            raise ValueError("so_prime_create(%d, %d) failed: %d" % (q, log_n, st.value))
        self.q = int(q)
        self.log_n = int(log_n)
        self.n = 1 << log_n
        self.mont = Mont(q)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.so_prime_destroy(h)
            self.h = None

    def tables(self):
        n = self.n
        t6 = np.zeros(6 * n, dtype=np.uint64)
        br = np.zeros(n, dtype=np.uint32)
        pn = np.zeros(2, dtype=np.uint64)
        lib().so_prime_tables(self.h, _p64(t6), br.ctypes.data_as(U32P), _p64(pn))
        names = ["psis_bitrev", "psis_bitrev_c", "psis_bitrev_cb",
                 "psis_inv_bitrev", "psis_inv_bitrev_c", "psis_inv_bitrev_cb"]
        out = {nm: t6[i * n:(i + 1) * n].copy() for i, nm in enumerate(names)}
        out["bitrevs"] = br
        out["psi"] = int(pn[0])
        out["n_inv"] = int(pn[1])
        return out

    def ntt_fwd(self, x):
        y = np.ascontiguousarray(x, dtype=np.uint64).copy()
        assert y.size == self.n
        lib().so_ntt_fwd(self.h, _p64(y))
        return y

    def ntt_inv(self, x):
        y = np.ascontiguousarray(x, dtype=np.uint64).copy()
        assert y.size == self.n
        lib().so_ntt_inv(self.h, _p64(y))
        return y


class Rns:
    """An RNS modulus chain at one ring degree: the `moduli` span."""

    def __init__(self, qs, log_n):
        self.qs = [int(q) for q in qs]
        self.log_n = int(log_n)
        self.n = 1 << self.log_n
        self.primes = [Prime(q, log_n) for q in self.qs]
        self._arr = (ctypes.c_void_p * len(self.primes))(*[p.h for p in self.primes])

    @property
    def L(self):
        return len(self.primes)

    def pp(self, nl=None):
        return self._arr

    def sub(self, nl):
        """The first `nl` moduli (a prefix of the chain)."""
        r = Rns.__new__(Rns)
        r.qs = self.qs[:nl]
        r.log_n = self.log_n
        r.n = self.n
        r.primes = self.primes[:nl]
        r._arr = (ctypes.c_void_p * nl)(*[p.h for p in r.primes])
        return r

    def _poly(self, x, nl=None):
        nl = self.L if nl is None else nl
        y = np.ascontiguousarray(x, dtype=np.uint64).copy()
        assert y.size == nl * self.n, (y.shape, nl, self.n)
        return y

    def uniform(self, seed, count=None):
        """Uniform residues (as Montgomery representations) from a seed."""
        if count is None:
            x = np.zeros((self.L, self.n), dtype=np.uint64)
            lib().so_fill_uniform(self._arr, self.L, int(seed), _p64(x))
            return x
        out = np.zeros((count, self.L, self.n), dtype=np.uint64)
        for c in range(count):
            lib().so_fill_uniform(self._arr, self.L, int(seed) * 1000003 + c, _p64(out[c]))
        return out

    def serialize(self, x):
        """RnsPolynomial::Serialize's coefficient vectors: one byte string per
        limb (MontgomeryInt::SerializeVector)."""
        x = self._poly(x).reshape(self.L, self.n)
        out = []
        for i, p in enumerate(self.primes):
            b = p.mont.s.log_modulus
            buf = np.zeros((self.n * b + 7) // 8, dtype=np.uint8)
            r = lib().so_serialize_vector(p.mont.ref(), _p64(np.ascontiguousarray(x[i])), self.n,
                                          _p8(buf), buf.size)
            if r < 0:
                raise ValueError("SerializeVector error %d" % r)
            out.append(buf[:r].tobytes())
        return out

    def deserialize(self, vectors):
        """RnsPolynomial::Deserialize's coefficients from per-limb strings."""
        assert len(vectors) == self.L
        out = np.zeros((self.L, self.n), dtype=np.uint64)
        for i, (p, v) in enumerate(zip(self.primes, vectors)):
            b = np.frombuffer(bytes(v), dtype=np.uint8).copy() if len(v) else np.zeros(1, np.uint8)
            row = np.zeros(self.n, dtype=np.uint64)
            r = lib().so_deserialize_vector(p.mont.ref(), self.n, _p8(b), len(v), _p64(row))
            if r != 0:
                raise ValueError("DeserializeVector error %d" % r)
            out[i] = row
        return out

    def mod_reduce_msb(self, x, ql_inv, is_ntt=True):
        """RnsPolynomial::ModReduceMsb: the first L-1 limbs of the result."""
        y = self._poly(x).reshape(self.L, self.n)
        qi = np.ascontiguousarray(ql_inv, dtype=np.uint64)
        rc = lib().so_poly_mod_reduce_msb(self._arr, self.L, _p64(qi), int(bool(is_ntt)), _p64(y))
        assert rc == 0
        return y[: self.L - 1]

    def basis_constants(self, out):
        """CRT constants for switching from this chain (Q) to `out` (P), as
        Montgomery representations: q_hat_inv[i] = (Q/q_i)^-1 mod q_i,
        q_hat_mod_ps[i, j] = (Q/q_i) mod p_j, q_mod_ps[j] = Q mod p_j,
        q_inv_mod_ps[j, i] = q_i^-1 mod p_j, p_mod_qs[i] = P mod q_i."""
        Q, P = 1, 1
        for q in self.qs:
            Q *= q
        for p in out.qs:
            P *= p
        imp = lambda m, v: m.imp(v % m.s.modulus)  # noqa: E731
        qm = [p.mont for p in self.primes]
        pm = [p.mont for p in out.primes]
        return {
            "q_hat_inv": np.array([imp(qm[i], pow(Q // q, -1, q)) for i, q in enumerate(self.qs)],
                                  dtype=np.uint64),
            "q_hat_mod_ps": np.array([[imp(pm[j], Q // q) for j in range(out.L)] for q in self.qs],
                                     dtype=np.uint64),
            "q_mod_ps": np.array([imp(pm[j], Q) for j in range(out.L)], dtype=np.uint64),
            "q_inv_mod_ps": np.array([[imp(pm[j], pow(q, -1, out.qs[j])) for q in self.qs]
                                      for j in range(out.L)], dtype=np.uint64),
            "p_mod_qs": np.array([imp(qm[i], P) for i in range(self.L)], dtype=np.uint64),
        }

    def switch_rns_basis(self, x, out, c):
        y = np.zeros((out.L, self.n), dtype=np.uint64)
        lib().so_switch_rns_basis(self._arr, self.L, out._arr, out.L, _p64(c["q_hat_inv"]),
                                  _p64(np.ascontiguousarray(c["q_hat_mod_ps"])), _p64(c["q_mod_ps"]),
                                  _p64(self._poly(x)), _p64(y))
        return y

    def scale_and_switch_rns_basis(self, x, out, c):
        y = np.zeros((out.L, self.n), dtype=np.uint64)
        lib().so_scale_and_switch_rns_basis(self._arr, self.L, out._arr, out.L,
                                            _p64(c["q_hat_inv"]),
                                            _p64(np.ascontiguousarray(c["q_inv_mod_ps"])),
                                            _p64(c["p_mod_qs"]), _p64(self._poly(x)), _p64(y))
        return y

    def approx_switch_rns_basis(self, x, out, c, is_ntt, balanced):
        y = np.zeros((out.L, self.n), dtype=np.uint64)
        lib().so_approx_switch_rns_basis(self._arr, self.L, out._arr, out.L, _p64(c["q_hat_inv"]),
                                         _p64(np.ascontiguousarray(c["q_hat_mod_ps"])),
                                         int(bool(is_ntt)), int(bool(balanced)),
                                         _p64(self._poly(x)), _p64(y))
        return y

    def sample_uniform_chacha(self, prng, count):
        """`count` RnsPolynomial::SampleUniform draws from one ChaCha PRNG."""
        out = np.zeros((count, self.L, self.n), dtype=np.uint64)
        for c in range(count):
            lib().so_sample_uniform_chacha(self._arr, self.L, prng.h, _p64(out[c]))
        return out

    def ntt_fwd(self, x):
        y = self._poly(x)
        lib().so_poly_ntt_fwd(self._arr, self.L, _p64(y))
        return y.reshape(np.shape(x))

    def ntt_inv(self, x):
        y = self._poly(x)
        lib().so_poly_ntt_inv(self._arr, self.L, _p64(y))
        return y.reshape(np.shape(x))

    def pointwise(self, op, x, a=None, b=None):
        ops = {"add": 0, "sub": 1, "mul": 2, "neg": 3, "fma": 4}
        y = self._poly(x)
        aa = self._poly(a) if a is not None else None
        bb = self._poly(b) if b is not None else None
        lib().so_poly_pointwise(self._arr, self.L, ops[op], _p64(y),
                                _p64(aa) if aa is not None else None,
                                _p64(bb) if bb is not None else None)
        return y.reshape(np.shape(x))

    def mul_scalar(self, x, scalars):
        y = self._poly(x)
        s = np.ascontiguousarray(scalars, dtype=np.uint64)
        lib().so_poly_mul_scalar(self._arr, self.L, _p64(y), _p64(s))
        return y.reshape(np.shape(x))

    def mul_int(self, x, scalar):
        y = self._poly(x)
        lib().so_poly_mul_int(self._arr, self.L, _p64(y), int(scalar))
        return y.reshape(np.shape(x))

    def substitute(self, x, power):
        y = self._poly(x)
        out = np.zeros_like(y)
        rc = lib().so_poly_substitute(self._arr, self.L, int(power), _p64(y), _p64(out))
        if rc != 0:
            raise ValueError("substitute failed: %d" % rc)
        return out.reshape(np.shape(x))

    def mod_reduce_lsb(self, x, t, ql_inv, is_ntt=True):
        y = self._poly(x)
        qi = np.ascontiguousarray(ql_inv, dtype=np.uint64)
        rc = lib().so_poly_mod_reduce_lsb(self._arr, self.L, int(t), _p64(qi),
                                          int(bool(is_ntt)), _p64(y))
        if rc != 0:
            raise ValueError("mod_reduce_lsb failed: %d" % rc)
        return y.reshape(self.L, self.n)[: self.L - 1].copy()

    def gadget_dims(self, log_bs):
        dims = (ctypes.c_int * self.L)()
        k = lib().so_gadget_dims(self._arr, self.L, _ints(log_bs), dims)
        return k, list(dims)

    def gadget_components(self, log_bs, q_hats, q_hat_invs):
        k, _ = self.gadget_dims(log_bs)
        gs = np.zeros((k, self.L), dtype=np.uint64)
        lib().so_gadget_components(self._arr, self.L, _ints(log_bs),
                                   _p64(np.ascontiguousarray(q_hats, dtype=np.uint64)),
                                   _p64(np.ascontiguousarray(q_hat_invs, dtype=np.uint64)),
                                   _p64(gs))
        return gs

    def gadget_decompose(self, a, log_bs):
        k, _ = self.gadget_dims(log_bs)
        aa = self._poly(a)
        out = np.zeros((k, self.L, self.n), dtype=np.uint64)
        lib().so_gadget_decompose(self._arr, self.L, _ints(log_bs), _p64(aa), _p64(out))
        return out

    def hwy_ntt_fwd(self, x):
        """The vectorised host transform (Highway algorithm, shell_cpu_vec.c)."""
        y = self._poly(x)
        lib().so_hwy_poly_ntt_fwd(self._arr, self.L, _p64(y))
        return y.reshape(np.shape(x))

    def relin_apply(self, comps, key_as, key_bs, log_bs, ntt_mask=None, vectorised=False):
        comps = np.ascontiguousarray(comps, dtype=np.uint64)
        degree = comps.shape[0] - 1
        if ntt_mask is None:
            ntt_mask = (1 << (degree + 1)) - 1
        c0 = np.zeros((self.L, self.n), dtype=np.uint64)
        c1 = np.zeros_like(c0)
        fn = lib().so_hwy_relin_apply if vectorised else lib().so_relin_apply
        rc = fn(self._arr, self.L, _ints(log_bs), degree, _p64(comps),
                int(ntt_mask), _p64(np.ascontiguousarray(key_as)),
                _p64(np.ascontiguousarray(key_bs)), _p64(c0), _p64(c1))
        if rc != 0:
            raise ValueError("relin_apply failed: %d" % rc)
        return c0, c1

    def galois_apply(self, c0, c1, key_as, key_bs, log_bs, c1_is_ntt=True):
        o0 = np.zeros((self.L, self.n), dtype=np.uint64)
        o1 = np.zeros_like(o0)
        rc = lib().so_galois_apply(self._arr, self.L, _ints(log_bs),
                                   _p64(self._poly(c0)), _p64(self._poly(c1)),
                                   int(bool(c1_is_ntt)),
                                   _p64(np.ascontiguousarray(key_as)),
                                   _p64(np.ascontiguousarray(key_bs)), _p64(o0), _p64(o1))
        if rc != 0:
            raise ValueError("galois_apply failed: %d" % rc)
        return o0, o1

    def tensor_product(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        la, lb = a.shape[0], b.shape[0]
        out = np.zeros((la + lb - 1, self.L, self.n), dtype=np.uint64)
        lib().so_tensor_product(self._arr, self.L, la, _p64(a), lb, _p64(b), _p64(out))
        return out

    def lazy_fma(self, as_, bs):
        as_ = np.ascontiguousarray(as_, dtype=np.uint64)
        bs = np.ascontiguousarray(bs, dtype=np.uint64)
        out = np.zeros((self.L, self.n), dtype=np.uint64)
        lib().so_lazy_fma(self._arr, self.L, as_.shape[0], _p64(as_), _p64(bs), _p64(out))
        return out

    def lazy_mixed(self, terms):
        """LazyRnsPolynomial from CreateZero through a list of updates, then
        Export (rns/lazy_rns_polynomial.h:44-115, .cc:89-254): each term is
        ("fma", a, b), ("sum", a, b, c) or ("diff", a, b, c) with (L, n)
        Montgomery arrays."""
        kinds = {"fma": 0, "sum": 1, "diff": 2}
        T = len(terms)
        zero = np.zeros((self.L, self.n), dtype=np.uint64)
        ks = (ctypes.c_int * max(T, 1))(*[kinds[t[0]] for t in terms])
        pick = lambda k: np.ascontiguousarray(
            [np.asarray(t[k], dtype=np.uint64) if len(t) > k else zero for t in terms],
            dtype=np.uint64).reshape(T, self.L, self.n)
        as_, bs, cs = pick(1), pick(2), pick(3)
        out = np.zeros((self.L, self.n), dtype=np.uint64)
        lib().so_lazy_mixed(self._arr, self.L, T, ks, _p64(as_), _p64(bs), _p64(cs), _p64(out))
        return out

    # ---- RnsContext constants (rns/rns_context.cc:169-282) ----
    def q_inv_mod(self, i, j):
        """q_i^-1 mod q_j (Montgomery rep), prime_q_invs_[i].Component(j)."""
        m = self.primes[j].mont
        if i == j:
            return m.imp(1)
        return m.inverse(m.imp(self.qs[i]))

    def ql_inv(self):
        """MainPrimeModulusInverseResidues()[L-1] restricted to the first L-1."""
        last = self.L - 1
        return np.array([self.q_inv_mod(last, j) for j in range(last)], dtype=np.uint64)

    def q_hats(self):
        """MainPrimeModulusComplements(L-1) (rns/rns_context.cc:234-257)."""
        out = []
        for i in range(self.L):
            m = self.primes[i].mont
            v = m.imp(1)
            for j in range(self.L):
                if j != i:
                    v = m.mul(v, m.imp(self.qs[j]))
            out.append(v)
        return np.array(out, dtype=np.uint64)

    def q_hat_invs(self):
        """MainPrimeModulusCrtFactors(L-1) (rns/rns_context.cc:259-282)."""
        out = []
        for i in range(self.L):
            m = self.primes[i].mont
            v = m.imp(1)
            for j in range(self.L):
                if j != i:
                    v = m.mul(v, self.q_inv_mod(j, i))
            out.append(v)
        return np.array(out, dtype=np.uint64)

    # ---- helpers for decrypt checks (value domain, Python ints) ----
    def import_signed(self, vals):
        """Small signed integers -> RNS poly (coeff form) of Montgomery reps."""
        vals = np.asarray(vals, dtype=np.int64)
        out = np.zeros((self.L, self.n), dtype=np.uint64)
        for i, p in enumerate(self.primes):
            q = p.q
            r = np.mod(vals.astype(object), q).astype(np.uint64)
            # Montgomery rep = v * 2^64 mod q
            out[i] = np.array([(int(v) << 64) % q for v in r], dtype=np.uint64)
        return out

    def export_values(self, x):
        """Montgomery reps -> values (per limb), as Python-int object array."""
        x = np.asarray(x).reshape(self.L, self.n)
        out = np.empty((self.L, self.n), dtype=object)
        for i, p in enumerate(self.primes):
            rinv = pow(2**64, -1, p.q)
            out[i] = [(int(v) * rinv) % p.q for v in x[i]]
        return out

    def crt_centered(self, x):
        """Coefficient-form RNS poly -> centered integers mod Q (Python ints)."""
        vals = self.export_values(x)
        Q = 1
        for q in self.qs:
            Q *= q
        res = [0] * self.n
        for i, q in enumerate(self.qs):
            qh = Q // q
            f = qh * pow(qh, -1, q)
            for j in range(self.n):
                res[j] += int(vals[i][j]) * f
        out = []
        for j in range(self.n):
            v = res[j] % Q
            if v > Q // 2:
                v -= Q
            out.append(v)
        return out


def splitmix64_stream(seed, count):
    """SplitMix64 outputs, bit-identical to so_splitmix64."""
    mask = 2**64 - 1
    s = int(seed) & mask
    out = []
    for _ in range(count):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out.append(z ^ (z >> 31))
    return out


def now():
    return lib().so_now_seconds()
