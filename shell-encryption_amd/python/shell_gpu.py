"""ctypes binding of libshellgpu.so (include/shell_gpu.h) for the tests, the
smoke check and the benchmark.

The binding fails loudly when the library is missing: there is no CPU
fallback anywhere on this path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libshellgpu.so")

U64P = ctypes.POINTER(ctypes.c_uint64)
INTP = ctypes.POINTER(ctypes.c_int)
VP = ctypes.c_void_p

OP_ADD, OP_SUB, OP_MUL, OP_NEG, OP_FMA = 0, 1, 2, 3, 4
BATCH_NTT_FWD, BATCH_NTT_INV, BATCH_MUL, BATCH_MOD_REDUCE = 0, 1, 2, 3
NTT_AUTO, NTT_INT, NTT_F64 = 0, 1, 2
NTT_TILING_ALT, NTT_PREFETCH, NTT_SPLIT_LIMB, NTT_WAVE_EXCHANGE = 4, 8, 16, 32
NTT_STAGE_SPLIT = 64
NTT_TWO_PER_CU = 128
BASIS_EXACT, BASIS_SCALE, BASIS_APPROX = 0, 1, 2

_lib = None


class ShgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("shell_gpu error %d: %s" % (code, msg))
        self.code = code
        self.msg = msg


class LimbJob(ctypes.Structure):
    _fields_ = [("data", VP), ("other", VP), ("prime", ctypes.c_int), ("log_n", ctypes.c_int)]


class PirServerConfig(ctypes.Structure):
    _fields_ = [("log_n", ctypes.c_int), ("levels", ctypes.c_int), ("max_active", ctypes.c_int),
                ("hbm_budget_bytes", ctypes.c_size_t), ("cohort", ctypes.c_int)]


class PolyJob(ctypes.Structure):
    _fields_ = [("data", VP), ("primes", ctypes.POINTER(ctypes.c_int)),
                ("host_ql_inv", ctypes.POINTER(ctypes.c_uint64)), ("num_limbs", ctypes.c_int),
                ("log_n", ctypes.c_int)]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libshellgpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    i = ctypes.c_int
    sz = ctypes.c_size_t
    sig = {
        "shg_create": [i, ctypes.POINTER(VP)],
        "shg_destroy": [VP],
        "shg_set_stream": [VP, VP],
        "shg_sync": [VP],
        "shg_alloc": [VP, sz, ctypes.POINTER(VP)],
        "shg_free": [VP, VP],
        "shg_upload": [VP, VP, VP, sz],
        "shg_download": [VP, VP, VP, sz],
        "shg_prime_create": [VP, ctypes.c_uint64, i, INTP],
        "shg_prime_info": [VP, i, U64P, U64P],
        "shg_rns_ql_inv": [VP, INTP, i, U64P],
        "shg_ntt_fwd": [VP, VP, i, i, INTP, i],
        "shg_ntt_inv": [VP, VP, i, i, INTP, i],
        "shg_pointwise": [VP, i, VP, VP, sz, VP, sz, i, i, INTP],
        "shg_mul_scalar": [VP, VP, U64P, i, i, INTP],
        "shg_mul_int": [VP, VP, ctypes.c_uint64, i, i, INTP],
        "shg_substitute": [VP, VP, VP, i, i, i, INTP],
        "shg_mod_reduce_lsb": [VP, VP, ctypes.c_uint64, U64P, i, i, i, INTP],
        "shg_tensor_product": [VP, VP, i, VP, i, VP, i, i, INTP],
        "shg_mod_reduce_msb": [VP, VP, U64P, i, i, i, INTP],
        "shg_switch_rns_basis": [VP, i, VP, i, i, i, i, INTP, VP, i, INTP, U64P, U64P, U64P],
        "shg_relin_apply": [VP, VP, i, i, VP, VP, i, INTP, VP, VP, i, i, INTP],
        "shg_relin_apply_noalloc": [VP, VP, i, i, VP, VP, i, INTP, VP, VP, i, i, INTP],
        "shg_galois_apply": [VP, VP, i, VP, VP, INTP, VP, VP, i, i, INTP],
        "shg_gadget_decompose": [VP, VP, INTP, VP, i, i, INTP],
        "shg_lazy_fma": [VP, VP, sz, VP, sz, i, VP, i, INTP],
        "shg_fused_mul_sum_add": [VP, i, VP, VP, VP, VP, i, i, INTP],
        "shg_ks_key_create": [VP, VP, VP, i, i, INTP, ctypes.POINTER(VP)],
        "shg_ks_key_destroy": [VP, VP],
        "shg_relin_apply_key": [VP, VP, i, i, VP, INTP, VP, VP, i, i, INTP],
        "shg_galois_apply_key": [VP, VP, i, VP, INTP, VP, VP, i, i, INTP],
        "shg_oblivious_expand_level_key": [VP, ctypes.POINTER(VP), i, i, VP, INTP, i, INTP],
        "shg_batch_limbs": [VP, ctypes.POINTER(LimbJob), i, i],
        "shg_batch_plan_create": [VP, ctypes.POINTER(LimbJob), i, i, ctypes.POINTER(VP)],
        "shg_batch_plan_destroy": [VP, VP],
        "shg_batch_mod_reduce_plan_create": [VP, ctypes.POINTER(PolyJob), i, ctypes.c_uint64, i, i,
                                             ctypes.POINTER(VP)],
        "shg_batch_run": [VP, VP, i],
        "shg_oblivious_expand_level": [VP, ctypes.POINTER(VP), i, i, VP, VP, INTP, i, INTP],
        "shg_oblivious_expand": [VP, ctypes.POINTER(VP), i, i, ctypes.POINTER(VP),
                                 ctypes.POINTER(VP), INTP, i, INTP],
        "shg_pir_db_export": [VP, VP, ctypes.c_longlong, VP, i, INTP],
        "shg_pir_db_synthesize": [VP, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_longlong, VP, i,
                                  INTP],
        "shg_pir_inner_product": [VP, ctypes.POINTER(VP), i, i, VP, i, ctypes.POINTER(VP), i, INTP],
        "shg_pir_server_create": [VP, INTP, i, INTP, ctypes.POINTER(VP), ctypes.POINTER(VP),
                                  ctypes.POINTER(PirServerConfig), ctypes.POINTER(VP)],
        "shg_pir_server_add_db": [VP, VP, i, INTP],
        "shg_pir_server_start": [VP],
        "shg_pir_server_submit": [VP, i, U64P, ctypes.POINTER(ctypes.c_longlong)],
        "shg_pir_server_wait": [VP, ctypes.c_longlong, U64P, ctypes.POINTER(ctypes.c_double)],
        "shg_pir_server_stats": [VP, ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_longlong), INTP],
        "shg_pir_server_load_test": [VP, i, U64P, i, i, ctypes.c_double, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)],
        "shg_pir_server_destroy": [VP],
        "shg_serialized_bytes": [VP, INTP, i, ctypes.POINTER(sz), ctypes.POINTER(sz)],
        "shg_serialize_polys": [VP, VP, i, i, INTP, VP],
        "shg_deserialize_polys": [VP, VP, i, i, INTP, VP],
        "shg_sample_uniform_chacha": [VP, ctypes.c_char_p, sz, i, i, INTP, VP],
        "shg_graph_begin": [VP],
        "shg_graph_end": [VP, ctypes.POINTER(VP)],
        "shg_graph_launch": [VP, VP],
        "shg_graph_destroy": [VP, VP],
        "shg_autotune_ntt": [VP, i, i, INTP],
        "shg_autotune_ntt_moduli": [VP, INTP, i, i, INTP],
        "shg_profile_enable": [VP, i],
        "shg_profile_reset": [VP],
        "shg_profile_query": [VP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_longlong)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.shg_last_error.restype = ctypes.c_char_p
    L.shg_get_stream.argtypes = [VP]
    L.shg_get_stream.restype = VP
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise ShgError(rc, lib().shg_last_error().decode())


def _ints(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


class DeviceArray:
    """A device buffer of uint64 words with a numpy-like shape."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * 8
        p = VP()
        _check(lib().shg_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @property
    def p(self):
        return VP(self.ptr)

    def offset(self, words):
        return VP(self.ptr + 8 * int(words))

    def view(self, offset_words, shape):
        """A non-owning view of part of this buffer."""
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.shape = self.ctx, tuple(shape)
        v.size = int(np.prod(v.shape))
        v.nbytes = v.size * 8
        assert offset_words + v.size <= self.size
        v.ptr = self.ptr + 8 * int(offset_words)
        v.owner = self
        v.is_view = True
        return v

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        assert a.size == self.size, (a.shape, self.shape)
        _check(lib().shg_upload(self.ctx.h, self.p, a.ctypes.data_as(VP), self.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.uint64)
        _check(lib().shg_download(self.ctx.h, out.ctypes.data_as(VP), self.p, self.nbytes))
        return out

    def free(self):
        if getattr(self, "is_view", False):
            self.ptr = 0
            return
        if self.ptr:
            lib().shg_free(self.ctx.h, self.p)
            self.ptr = 0

    def __del__(self):
        try:
            if not getattr(self, "is_view", False) and self.ptr and self.ctx.h:
                lib().shg_free(self.ctx.h, VP(self.ptr))
                self.ptr = 0
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        h = VP()
        _check(lib().shg_create(int(device), ctypes.byref(h)))
        self.h = h.value
        self._primes = {}

    def close(self):
        if self.h:
            lib().shg_destroy(VP(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().shg_sync(self.h))

    def array(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        return DeviceArray(self, a.shape).upload(a)

    def empty(self, shape):
        return DeviceArray(self, shape)

    def prime(self, q, log_n):
        key = (int(q), int(log_n))
        if key not in self._primes:
            pid = ctypes.c_int()
            _check(lib().shg_prime_create(self.h, int(q), int(log_n), ctypes.byref(pid)))
            self._primes[key] = pid.value
        return self._primes[key]

    def primes(self, qs, log_n):
        return [self.prime(q, log_n) for q in qs]

    def prime_info(self, pid, log_n):
        params = np.zeros(10, dtype=np.uint64)
        tables = np.zeros(4 << log_n, dtype=np.uint64)
        _check(lib().shg_prime_info(self.h, int(pid), params.ctypes.data_as(U64P),
                                    tables.ctypes.data_as(U64P)))
        return params, tables

    def ql_inv(self, primes):
        """Montgomery reps of q_l^-1 mod q_i (i < L-1) for dropping primes[-1]."""
        out = np.zeros(max(len(primes) - 1, 1), dtype=np.uint64)
        _check(lib().shg_rns_ql_inv(self.h, _ints(primes), len(primes), out.ctypes.data_as(U64P)))
        return out[:len(primes) - 1]

    # ---- ops ----------------------------------------------------------------
    def ntt_fwd(self, buf, count, primes, variant=NTT_AUTO, offset=0):
        _check(lib().shg_ntt_fwd(self.h, buf.offset(offset), int(count), len(primes),
                                 _ints(primes), int(variant)))

    def ntt_inv(self, buf, count, primes, variant=NTT_AUTO, offset=0):
        _check(lib().shg_ntt_inv(self.h, buf.offset(offset), int(count), len(primes),
                                 _ints(primes), int(variant)))

    def pointwise(self, op, x, a, b, count, primes, a_stride=None, b_stride=None):
        words = len(primes) * (x.size // (count * len(primes)))
        a_stride = words if a_stride is None else a_stride
        b_stride = words if b_stride is None else b_stride
        _check(lib().shg_pointwise(self.h, int(op), x.p, a.p if a is not None else None,
                                   int(a_stride), b.p if b is not None else None, int(b_stride),
                                   int(count), len(primes), _ints(primes)))

    def mul_scalar(self, x, scalars, count, primes):
        s = np.ascontiguousarray(scalars, dtype=np.uint64)
        _check(lib().shg_mul_scalar(self.h, x.p, s.ctypes.data_as(U64P), int(count),
                                    len(primes), _ints(primes)))

    def mul_int(self, x, scalar, count, primes):
        _check(lib().shg_mul_int(self.h, x.p, int(scalar), int(count), len(primes),
                                 _ints(primes)))

    def substitute(self, src, dst, power, count, primes):
        _check(lib().shg_substitute(self.h, src.p, dst.p, int(power), int(count), len(primes),
                                    _ints(primes)))

    def mod_reduce_lsb(self, x, t, ql_inv, is_ntt, count, primes):
        qi = np.ascontiguousarray(ql_inv, dtype=np.uint64)
        _check(lib().shg_mod_reduce_lsb(self.h, x.p, int(t), qi.ctypes.data_as(U64P),
                                        int(bool(is_ntt)), int(count), len(primes),
                                        _ints(primes)))

    def mod_reduce_msb(self, x, ql_inv, is_ntt, count, primes):
        qi = np.ascontiguousarray(ql_inv, dtype=np.uint64)
        _check(lib().shg_mod_reduce_msb(self.h, x.p, qi.ctypes.data_as(U64P), int(bool(is_ntt)),
                                        int(count), len(primes), _ints(primes)))

    def switch_rns_basis(self, mode, src, in_is_ntt, balanced, count, in_primes, out, out_primes,
                         q_hat_inv, matrix, vec=None):
        qh = np.ascontiguousarray(q_hat_inv, dtype=np.uint64)
        m = np.ascontiguousarray(matrix, dtype=np.uint64)
        v = np.ascontiguousarray(vec if vec is not None else np.zeros(1), dtype=np.uint64)
        _check(lib().shg_switch_rns_basis(self.h, int(mode), src.p, int(bool(in_is_ntt)),
                                          int(bool(balanced)), int(count), len(in_primes),
                                          _ints(in_primes), out.p, len(out_primes),
                                          _ints(out_primes), qh.ctypes.data_as(U64P),
                                          m.ctypes.data_as(U64P), v.ctypes.data_as(U64P)))

    def tensor_product(self, a, len_a, b, len_b, out, count, primes):
        _check(lib().shg_tensor_product(self.h, a.p, int(len_a), b.p, int(len_b), out.p,
                                        int(count), len(primes), _ints(primes)))

    def relin_apply(self, comps, degree, ntt_mask, key_as, key_bs, num_key_polys, log_bs,
                    out0, out1, count, primes):
        _check(lib().shg_relin_apply(self.h, comps.p, int(degree), int(ntt_mask), key_as.p,
                                     key_bs.p, int(num_key_polys), _ints(log_bs), out0.p,
                                     out1.p, int(count), len(primes), _ints(primes)))

    def relin_apply_noalloc(self, comps, degree, ntt_mask, key_as, key_bs, num_key_polys, log_bs,
                            out0, out1, count, primes):
        _check(lib().shg_relin_apply_noalloc(self.h, comps.p, int(degree), int(ntt_mask), key_as.p,
                                             key_bs.p, int(num_key_polys), _ints(log_bs), out0.p,
                                             out1.p, int(count), len(primes), _ints(primes)))

    def galois_apply(self, comps, c1_is_ntt, key_as, key_bs, log_bs, out0, out1, count, primes):
        _check(lib().shg_galois_apply(self.h, comps.p, int(bool(c1_is_ntt)), key_as.p,
                                      key_bs.p, _ints(log_bs), out0.p, out1.p, int(count),
                                      len(primes), _ints(primes)))

    def gadget_decompose(self, a, log_bs, digits, count, primes):
        _check(lib().shg_gadget_decompose(self.h, a.p, _ints(log_bs), digits.p, int(count),
                                          len(primes), _ints(primes)))

    def lazy_fma(self, as_, as_stride, bs, bs_stride, num_terms, out, primes):
        _check(lib().shg_lazy_fma(self.h, as_.p, int(as_stride), bs.p, int(bs_stride),
                                  int(num_terms), out.p, len(primes), _ints(primes)))

    def fused_mul_sum_add(self, x, a, b, c, subtract, count, primes):
        """x <- x + (a +/- b) * c (LazyRnsPolynomial::FusedMulSum/DifferenceAddInPlace)."""
        _check(lib().shg_fused_mul_sum_add(self.h, int(bool(subtract)), x.p, a.p, b.p, c.p,
                                           int(count), len(primes), _ints(primes)))

    def ks_key(self, key_as, key_bs, num_key_polys, primes):
        """A prepared key-switching key (shg_ks_key_create); key_as/key_bs
        must outlive it."""
        return KsKey(self, key_as, key_bs, num_key_polys, primes)

    def relin_apply_key(self, comps, degree, ntt_mask, key, log_bs, out0, out1, count, primes):
        _check(lib().shg_relin_apply_key(self.h, comps.p, int(degree), int(ntt_mask), key.h,
                                         _ints(log_bs), out0.p, out1.p, int(count), len(primes),
                                         _ints(primes)))

    def galois_apply_key(self, comps, c1_is_ntt, key, log_bs, out0, out1, count, primes):
        _check(lib().shg_galois_apply_key(self.h, comps.p, int(bool(c1_is_ntt)), key.h,
                                          _ints(log_bs), out0.p, out1.p, int(count), len(primes),
                                          _ints(primes)))

    def batch_limbs(self, jobs, op):
        """jobs: (data_ptr, other_ptr or None, prime, log_n) tuples."""
        arr = (LimbJob * len(jobs))(*[LimbJob(d, o, p, ln) for (d, o, p, ln) in jobs])
        _check(lib().shg_batch_limbs(self.h, arr, len(jobs), int(op)))

    def batch_plan(self, jobs, variant=NTT_AUTO):
        return BatchPlan(self, jobs, variant)

    def mod_reduce_plan(self, jobs, t, is_ntt, variant=NTT_AUTO):
        """jobs: (data_ptr, prime list, ql_inv list, log_n) per polynomial."""
        return BatchPlan(self, None, variant, mod_reduce=(jobs, t, is_ntt))

    @staticmethod
    def _ptrs(bufs):
        vals = [b.ptr if isinstance(b, DeviceArray) else int(b) for b in bufs]
        return (VP * max(len(vals), 1))(*vals)

    def expand_level(self, bufs, level, key_as, key_bs, log_bs, primes):
        """Oblivious expansion level `level` for every query buffer in bufs."""
        _check(lib().shg_oblivious_expand_level(self.h, self._ptrs(bufs), len(bufs), int(level),
                                                key_as.p, key_bs.p, _ints(log_bs), len(primes),
                                                _ints(primes)))

    def expand_level_key(self, bufs, level, key, log_bs, primes):
        """expand_level with a prepared Galois key (KsKey)."""
        _check(lib().shg_oblivious_expand_level_key(self.h, self._ptrs(bufs), len(bufs),
                                                    int(level), key.h, _ints(log_bs),
                                                    len(primes), _ints(primes)))

    def expand(self, bufs, levels, keys, log_bs, primes):
        """keys: one (key_as, key_bs) DeviceArray pair per level."""
        ka = self._ptrs([k[0] for k in keys])
        kb = self._ptrs([k[1] for k in keys])
        _check(lib().shg_oblivious_expand(self.h, self._ptrs(bufs), len(bufs), int(levels), ka, kb,
                                          _ints(log_bs), len(primes), _ints(primes)))

    def pir_db_export(self, db, num_records, out, primes):
        _check(lib().shg_pir_db_export(self.h, db.p, int(num_records), out.p, len(primes),
                                       _ints(primes)))

    def pir_db_synthesize(self, seed, t, num_records, out, primes):
        _check(lib().shg_pir_db_synthesize(self.h, int(seed), int(t), int(num_records), out.p,
                                           len(primes), _ints(primes)))

    def pir_inner_product(self, expanded, levels, db_x, num_columns, outs, primes):
        _check(lib().shg_pir_inner_product(self.h, self._ptrs(expanded), len(expanded), int(levels),
                                           db_x.p, int(num_columns), self._ptrs(outs), len(primes),
                                           _ints(primes)))

    def serialized_bytes(self, primes):
        """(per-limb byte lengths, polynomial byte length) of the wire form."""
        limbs = (ctypes.c_size_t * len(primes))()
        total = ctypes.c_size_t()
        _check(lib().shg_serialized_bytes(self.h, _ints(primes), len(primes), limbs,
                                          ctypes.byref(total)))
        return list(limbs), total.value

    def serialize(self, polys, count, primes, out, offset_bytes=0):
        """Packs count polynomials into the device byte buffer `out`
        (a DeviceArray, used as bytes from offset_bytes)."""
        _check(lib().shg_serialize_polys(self.h, polys.p, int(count), len(primes), _ints(primes),
                                         VP(out.ptr + int(offset_bytes))))

    def deserialize(self, src, count, primes, polys, offset_bytes=0):
        _check(lib().shg_deserialize_polys(self.h, VP(src.ptr + int(offset_bytes)), int(count),
                                           len(primes), _ints(primes), polys.p))

    def sample_uniform_chacha(self, seed, count, primes, out):
        seed = bytes(seed)
        _check(lib().shg_sample_uniform_chacha(self.h, seed, len(seed), int(count), len(primes),
                                               _ints(primes), out.p))

    def graph_begin(self):
        _check(lib().shg_graph_begin(self.h))

    def graph_end(self):
        g = VP()
        _check(lib().shg_graph_end(self.h, ctypes.byref(g)))
        return Graph(self, g)

    def profile(self, enable=True):
        _check(lib().shg_profile_enable(self.h, int(bool(enable))))

    def profile_reset(self):
        _check(lib().shg_profile_reset(self.h))

    def profile_query(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _check(lib().shg_profile_query(self.h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def autotune_ntt_moduli(self, primes, count):
        c = ctypes.c_int()
        _check(lib().shg_autotune_ntt_moduli(self.h, _ints(primes), len(primes), int(count),
                                             ctypes.byref(c)))
        return c.value

    def autotune_ntt(self, log_n, limbs):
        c = ctypes.c_int()
        _check(lib().shg_autotune_ntt(self.h, int(log_n), int(limbs), ctypes.byref(c)))
        return c.value


class KsKey:
    """shg_ks_key: the caller's key arrays plus FP64-form copies."""

    def __init__(self, ctx, key_as, key_bs, num_key_polys, primes):
        self.ctx = ctx
        self.keep = (key_as, key_bs)
        self.h = VP()
        _check(lib().shg_ks_key_create(ctx.h, key_as.p, key_bs.p, int(num_key_polys), len(primes),
                                       _ints(primes), ctypes.byref(self.h)))

    def close(self):
        if self.h and self.ctx.h:
            _check(lib().shg_ks_key_destroy(self.ctx.h, self.h))
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the context may be gone
            pass


class BatchPlan:
    """A device-resident job table for shg_batch_run."""

    def __init__(self, ctx, jobs, variant=NTT_AUTO, mod_reduce=None):
        self.ctx = ctx
        self.h = VP()
        if mod_reduce is None:
            arr = (LimbJob * max(len(jobs), 1))(*[LimbJob(d, o, p, ln) for (d, o, p, ln) in jobs])
            _check(lib().shg_batch_plan_create(ctx.h, arr, len(jobs), int(variant),
                                               ctypes.byref(self.h)))
            return
        pjobs, t, is_ntt = mod_reduce
        keep = []
        arr = (PolyJob * max(len(pjobs), 1))()
        for k, (d, primes, ql_inv, ln) in enumerate(pjobs):
            pa = (ctypes.c_int * len(primes))(*primes)
            qa = (ctypes.c_uint64 * max(len(ql_inv), 1))(*[int(v) for v in ql_inv])
            keep += [pa, qa]
            arr[k] = PolyJob(d, pa, qa, len(primes), ln)
        _check(lib().shg_batch_mod_reduce_plan_create(ctx.h, arr, len(pjobs), int(t),
                                                      int(bool(is_ntt)), int(variant),
                                                      ctypes.byref(self.h)))

    def run(self, op):
        _check(lib().shg_batch_run(self.ctx.h, self.h, int(op)))

    def close(self):
        if self.h:
            lib().shg_batch_plan_destroy(self.ctx.h, self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """An instantiated HIP graph captured from the context's stream."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    def launch(self):
        _check(lib().shg_graph_launch(self.ctx.h, self.h))

    def close(self):
        if self.h:
            lib().shg_graph_destroy(self.ctx.h, self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PirServer:
    """Continuous-batching server of encrypted queries (shg_pir_server_*)."""

    def __init__(self, ctx, primes, log_n, log_bs, keys, max_active, hbm_budget=0, cohort=0):
        self.ctx = ctx
        self.levels = len(keys)
        self.keys = keys  # keep the key arrays alive
        ka = Context._ptrs([k[0] for k in keys])
        kb = Context._ptrs([k[1] for k in keys])
        cfg = PirServerConfig(int(log_n), self.levels, int(max_active), int(hbm_budget),
                              int(cohort))
        self.h = VP()
        _check(lib().shg_pir_server_create(ctx.h, _ints(primes), len(primes), _ints(log_bs), ka, kb,
                                           ctypes.byref(cfg), ctypes.byref(self.h)))
        self.ct_words = 2 * len(primes) << int(log_n)
        self.columns = {}
        self.dbs = []

    def add_db(self, db_exported, num_columns):
        d = ctypes.c_int()
        _check(lib().shg_pir_server_add_db(self.h, db_exported.p, int(num_columns), ctypes.byref(d)))
        self.columns[d.value] = int(num_columns)
        self.dbs.append(db_exported)
        return d.value

    def start(self):
        _check(lib().shg_pir_server_start(self.h))

    def submit(self, db, query):
        q = np.ascontiguousarray(query, dtype=np.uint64).ravel()
        assert q.size == self.ct_words
        r = ctypes.c_longlong()
        _check(lib().shg_pir_server_submit(self.h, int(db), q.ctypes.data_as(U64P), ctypes.byref(r)))
        return r.value, int(db)

    def wait(self, ticket):
        rid, db = ticket
        out = np.empty(self.columns[db] * self.ct_words, dtype=np.uint64)
        ms = ctypes.c_double()
        _check(lib().shg_pir_server_wait(self.h, rid, out.ctypes.data_as(U64P), ctypes.byref(ms)))
        return out, ms.value

    def stats(self):
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        cap = ctypes.c_int()
        _check(lib().shg_pir_server_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                          ctypes.byref(cap)))
        return {"steps": a.value, "stage_runs": b.value, "batched_items": c.value,
                "capacity": cap.value}

    def load_test(self, db, queries, num_requests, rate, seed=1):
        q = np.ascontiguousarray(queries, dtype=np.uint64)
        lat = (ctypes.c_double * num_requests)()
        wall = ctypes.c_double()
        _check(lib().shg_pir_server_load_test(self.h, int(db), q.ctypes.data_as(U64P),
                                              q.size // self.ct_words, int(num_requests),
                                              float(rate), int(seed), lat, ctypes.byref(wall)))
        return np.array(lat[:]), wall.value

    def close(self):
        if self.h:
            lib().shg_pir_server_destroy(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
