"""Bit-exact parity of the HIP path (through the C-ABI) against the CPU
oracle on identical seeded inputs.  Every output word is compared."""
import json
import os

import numpy as np
import pytest

import oracle
import shell_gpu as sg
from common import Bgv, negacyclic_mul_mod_t, ntt_primes, substitute_coeffs

pytestmark = pytest.mark.gpu

VARIANTS = [sg.NTT_INT, sg.NTT_F64]


@pytest.fixture(scope="module")
def ctx():
    c = sg.Context(0)
    yield c
    c.close()


def _rns(log_n, L, log_bound=50):
    qs = ntt_primes(log_n, L, log_bound)
    return qs, oracle.Rns(qs, log_n)


def test_prime_tables_match_oracle(ctx):
    for log_n, L in [(3, 1), (12, 2), (14, 2)]:
        qs, rns = _rns(log_n, L)
        for q, p in zip(qs, rns.primes):
            pid = ctx.prime(q, log_n)
            params, tables = ctx.prime_info(pid, log_n)
            tb = p.tables()
            s = p.mont.s
            assert list(params[:8]) == [s.modulus, s.log_modulus, s.r_mod_modulus,
                                        s.r_mod_modulus_barrett, s.inv_modulus, s.inv_r,
                                        s.inv_r_barrett, s.barrett_numerator]
            assert params[8] == tb["psi"] and params[9] == tb["n_inv"]
            n = 1 << log_n
            assert np.array_equal(tables[:n], tb["psis_bitrev_c"])
            assert np.array_equal(tables[n:2 * n], tb["psis_bitrev_cb"])
            assert np.array_equal(tables[2 * n:3 * n], tb["psis_inv_bitrev_c"])
            assert np.array_equal(tables[3 * n:], tb["psis_inv_bitrev_cb"])


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("log_n,L,count", [(1, 1, 3), (2, 2, 2), (4, 1, 5), (5, 2, 3),
                                           (8, 3, 2), (11, 2, 3), (12, 1, 4), (13, 4, 2),
                                           (14, 8, 2), (15, 2, 2)])
def test_ntt_round_trip_bit_exact(ctx, variant, log_n, L, count):
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    x = rns.uniform(1000 + log_n, count)
    # extreme residues: 0, 1, q-1 and q (the image of NegateInPlace(0))
    for i, q in enumerate(qs):
        x[0, i, :4] = [0, 1, q - 1, q][: min(4, rns.n)]
    d = ctx.array(x)
    ctx.ntt_fwd(d, count, pids, variant)
    got = d.download()
    exp = np.stack([rns.ntt_fwd(x[c]) for c in range(count)])
    assert np.array_equal(got, exp)
    ctx.ntt_inv(d, count, pids, variant)
    back = d.download()
    exp_back = np.stack([rns.ntt_inv(exp[c]) for c in range(count)])
    assert np.array_equal(back, exp_back)
    xr = x.copy()
    for i, q in enumerate(qs):
        xr[0, i, :4][xr[0, i, :4] == q] = 0
    assert np.array_equal(back, xr)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("log_n", [11, 14, 15])
def test_ntt_extreme_inputs(ctx, variant, log_n):
    """Constant and alternating extreme residues (moduli just below 2^50)
    push the lazy-integer and FP64 growth bounds hardest; forward and
    inverse transforms of each pattern must match the oracle."""
    qs, rns = _rns(log_n, 2)
    pids = ctx.primes(qs, log_n)
    n = rns.n
    pats = []
    for pat in range(4):
        x = np.zeros((2, n), dtype=np.uint64)
        for i, q in enumerate(qs):
            if pat == 0:
                x[i] = q - 1
            elif pat == 1:
                x[i] = q
            elif pat == 2:
                x[i, ::2] = q - 1
            else:
                x[i, ::2] = 1
                x[i, 1::2] = q - 1
        pats.append(x)
    x = np.stack(pats)
    d = ctx.array(x)
    ctx.ntt_fwd(d, len(pats), pids, variant)
    assert np.array_equal(d.download(), np.stack([rns.ntt_fwd(v) for v in x]))
    d = ctx.array(x)
    ctx.ntt_inv(d, len(pats), pids, variant)
    assert np.array_equal(d.download(), np.stack([rns.ntt_inv(v) for v in x]))


def test_ntt_wide_modulus_int_only(ctx):
    """60-bit primes take the integer path (FP64 needs q < 2^50)."""
    qs, rns = _rns(12, 2, log_bound=60)
    pids = ctx.primes(qs, 12)
    x = rns.uniform(77, 3)
    d = ctx.array(x)
    ctx.ntt_fwd(d, 3, pids, sg.NTT_AUTO)
    assert np.array_equal(d.download(), np.stack([rns.ntt_fwd(v) for v in x]))
    with pytest.raises(sg.ShgError):
        ctx.ntt_fwd(d, 3, pids, sg.NTT_F64)


@pytest.mark.parametrize("op", ["add", "sub", "mul", "neg", "fma"])
def test_pointwise(ctx, op):
    qs, rns = _rns(11, 3)
    pids = ctx.primes(qs, 11)
    cnt = 3
    x, a, b = rns.uniform(1, cnt), rns.uniform(2, cnt), rns.uniform(3, cnt)
    x[0, :, 0] = 0
    a[0, :, 1] = [q for q in qs]  # the Negate(0) = q quirk as an operand
    dx, da, db = ctx.array(x), ctx.array(a), ctx.array(b)
    opc = {"add": sg.OP_ADD, "sub": sg.OP_SUB, "mul": sg.OP_MUL, "neg": sg.OP_NEG,
           "fma": sg.OP_FMA}[op]
    ctx.pointwise(opc, dx, da, db, cnt, pids)
    exp = np.stack([rns.pointwise(op, x[c], a[c], b[c]) for c in range(cnt)])
    assert np.array_equal(dx.download(), exp)
    if op == "neg":
        assert np.all(dx.download()[0, :, 0] == np.array(qs, dtype=np.uint64))


def test_pointwise_broadcast_and_scalars(ctx):
    qs, rns = _rns(10, 2)
    pids = ctx.primes(qs, 10)
    x, a = rns.uniform(5, 4), rns.uniform(6, 1)
    dx, da = ctx.array(x), ctx.array(a)
    ctx.pointwise(sg.OP_MUL, dx, da, None, 4, pids, a_stride=0)
    exp = np.stack([rns.pointwise("mul", x[c], a[0]) for c in range(4)])
    assert np.array_equal(dx.download(), exp)
    sc = np.array([rns.primes[i].mont.imp(12345 + i) for i in range(2)], dtype=np.uint64)
    ctx.mul_scalar(dx, sc, 4, pids)
    exp = np.stack([rns.mul_scalar(exp[c], sc) for c in range(4)])
    assert np.array_equal(dx.download(), exp)
    ctx.mul_int(dx, 65537, 4, pids)
    exp = np.stack([rns.mul_int(exp[c], 65537) for c in range(4)])
    assert np.array_equal(dx.download(), exp)


@pytest.mark.parametrize("log_n,power", [(5, 1), (5, 63), (11, 5), (14, 5), (14, 2 * 16384 - 1)])
def test_substitute(ctx, log_n, power):
    qs, rns = _rns(log_n, 3)
    pids = ctx.primes(qs, log_n)
    x = rns.uniform(9, 2)
    src, dst = ctx.array(x), ctx.empty(x.shape)
    ctx.substitute(src, dst, power, 2, pids)
    exp = np.stack([rns.substitute(x[c], power) for c in range(2)])
    assert np.array_equal(dst.download(), exp)
    with pytest.raises(sg.ShgError):
        ctx.substitute(src, dst, 2, 2, pids)


@pytest.mark.parametrize("is_ntt", [True, False])
@pytest.mark.parametrize("log_n,L", [(9, 3), (11, 2), (12, 4), (14, 8), (15, 3)])
def test_mod_reduce_lsb(ctx, is_ntt, log_n, L):
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    t = 65537
    x = rns.uniform(21, 2)
    d = ctx.array(x)
    ql_inv = rns.ql_inv()
    ctx.mod_reduce_lsb(d, t, ql_inv, is_ntt, 2, pids)
    got = d.download()[:, : L - 1]
    exp = np.stack([rns.mod_reduce_lsb(x[c], t, ql_inv, is_ntt) for c in range(2)])
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("is_ntt", [True, False])
def test_mod_reduce_lsb_many(ctx, is_ntt):
    """More polynomials than resident workgroups: the fused per-polynomial
    kernel loops; first, middle and last polynomials against the oracle."""
    log_n, L, count = 12, 4, 600
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    t = 65537
    x = rns.uniform(22 + int(is_ntt), count)
    d = ctx.array(x)
    ql_inv = rns.ql_inv()
    ctx.mod_reduce_lsb(d, t, ql_inv, is_ntt, count, pids)
    got = d.download()
    for c in (0, count // 2, count - 1):
        assert np.array_equal(got[c, : L - 1], rns.mod_reduce_lsb(x[c], t, ql_inv, is_ntt)), c
        assert np.array_equal(got[c, L - 1], x[c, L - 1]), c  # the dropped limb is left as it was


def test_tensor_product(ctx):
    qs, rns = _rns(13, 4)
    pids = ctx.primes(qs, 13)
    cnt = 3
    a = np.stack([rns.uniform(100 + c, 2) for c in range(cnt)])
    b = np.stack([rns.uniform(200 + c, 2) for c in range(cnt)])
    da, db = ctx.array(a), ctx.array(b)
    out = ctx.empty((cnt, 3, 4, rns.n))
    ctx.tensor_product(da, 2, db, 2, out, cnt, pids)
    exp = np.stack([rns.tensor_product(a[c], b[c]) for c in range(cnt)])
    assert np.array_equal(out.download(), exp)


def test_ntt_n32768_more_limbs_than_workgroups(ctx):
    """2^15 limbs (FP64: one persistent kernel with both halves held in
    VGPRs) with 320 limbs, more than one per resident workgroup: sampled
    forward outputs and every inverse word against the oracle."""
    qs, rns = _rns(15, 4)
    pids = ctx.primes(qs, 15)
    count = 80
    x = rns.uniform(1515, count)
    d = ctx.array(x)
    ctx.ntt_fwd(d, count, pids, sg.NTT_F64)
    got = d.download()
    for c in (0, 41, count - 1):
        assert np.array_equal(got[c], rns.ntt_fwd(x[c])), c
    y = rns.uniform(1516, count)
    d2 = ctx.array(y)
    ctx.ntt_inv(d2, count, pids, sg.NTT_F64)
    got2 = d2.download()
    for c in (0, 17, count - 1):
        assert np.array_equal(got2[c], rns.ntt_inv(y[c])), c
    ctx.ntt_inv(d, count, pids, sg.NTT_F64)
    assert np.array_equal(d.download(), x)


@pytest.mark.parametrize("variant", [sg.NTT_F64, sg.NTT_INT])
def test_polynomial_product_n4096(ctx, variant):
    """The plain negacyclic product at its own shape: n = 2^12, one prime
    (the largest below 2^50 with q = 1 mod 2^13), 64 pairs: forward NTT of
    both, pointwise Montgomery product, inverse NTT, every word against the
    oracle; and the NTT -> INTT round trip is the identity."""
    qs, rns = _rns(12, 1)
    pids = ctx.primes(qs, 12)
    pairs = 64
    a, b = rns.uniform(41, pairs), rns.uniform(42, pairs)
    da, db = ctx.array(a), ctx.array(b)
    ctx.ntt_fwd(da, pairs, pids, variant)
    ctx.ntt_fwd(db, pairs, pids, variant)
    assert np.array_equal(db.download()[pairs - 1], rns.ntt_fwd(b[pairs - 1]))
    ctx.pointwise(sg.OP_MUL, da, db, None, pairs, pids)  # MulInPlace
    ctx.ntt_inv(da, pairs, pids, variant)
    got = da.download()
    for c in range(pairs):
        exp = rns.ntt_inv(rns.pointwise("mul", rns.ntt_fwd(a[c]), rns.ntt_fwd(b[c])))
        assert np.array_equal(got[c], exp), c
    ctx.ntt_inv(db, pairs, pids, variant)  # round trip of b
    assert np.array_equal(db.download(), b)


def test_tensor_product_batch_n8192_decrypts(ctx):
    """The batched tensor-product workload at its own shape: n = 2^13, L = 4
    primes, t = 65537, 1024 two-component ciphertexts (eight encrypted pairs
    tiled) multiplied in NTT form and brought back by the inverse NTT on the
    GPU; sampled ciphertexts bit-exact against the oracle and decrypting to
    the negacyclic product of the messages mod t."""
    log_n, L, t, B, distinct = 13, 4, 65537, 1024, 8
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    bgv = Bgv(rns, t, seed=11)
    r = np.random.default_rng(12)
    ms = [(r.integers(0, t, rns.n), r.integers(0, t, rns.n)) for _ in range(distinct)]
    c1 = np.stack([bgv.encrypt(m[0]) for m in ms])
    c2 = np.stack([bgv.encrypt(m[1]) for m in ms])
    reps = B // distinct
    da, db = ctx.array(np.tile(c1, (reps, 1, 1, 1))), ctx.array(np.tile(c2, (reps, 1, 1, 1)))
    prod = ctx.empty((B, 3, L, rns.n))
    ctx.tensor_product(da, 2, db, 2, prod, B, pids)
    ntt_form = prod.download()
    ctx.ntt_inv(prod, B * 3, pids)
    coeff = prod.download()
    for c in (0, 5, B - 1):
        i = c % distinct
        exp = rns.tensor_product(c1[i], c2[i])
        assert np.array_equal(ntt_form[c], exp), c
        for j in range(3):
            assert np.array_equal(coeff[c, j], rns.ntt_inv(exp[j])), (c, j)
    for c in (3, B - 2):
        i = c % distinct
        expect = np.array(negacyclic_mul_mod_t(ms[i][0], ms[i][1], t), dtype=np.int64)
        assert np.array_equal(bgv.decrypt(ntt_form[c]), expect), c


def test_gadget_decompose(ctx):
    qs, rns = _rns(10, 3)
    pids = ctx.primes(qs, 10)
    log_bs = [10, 7, 20]
    k, _ = rns.gadget_dims(log_bs)
    a = rns.uniform(4, 2)
    da = ctx.array(a)
    dg = ctx.empty((2, k, 3, rns.n))
    ctx.gadget_decompose(da, log_bs, dg, 2, pids)
    exp = np.stack([rns.gadget_decompose(a[c], log_bs) for c in range(2)])
    assert np.array_equal(dg.download(), exp)


# Bases above 2^16 take the path without digit planes: an inverse transform
# with the ExportInt scaling fused (n^-1 2^-64), then the key switch.
# At 2^14 (split limbs) bases up to 2^11 take stage 0 from a table of
# products (7, 8, 10, 11 bits), 12 and 16 bits the products themselves; a
# common base of 8, 10 or 12 bits packs the digit planes at that width.
@pytest.mark.parametrize("log_n,L,log_b,count", [(6, 2, 10, 3), (12, 3, 10, 2), (12, 2, 10, 64),
                                                 (13, 4, 10, 2), (14, 8, 10, 2), (15, 2, 25, 1),
                                                 (13, 2, 20, 2), (14, 3, 20, 1), (14, 3, 7, 2),
                                                 (14, 2, 11, 2), (14, 2, 12, 1), (14, 2, 16, 1),
                                                 (14, 3, 8, 2)])
def test_relin_apply(ctx, log_n, L, log_b, count):
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    log_bs = [log_b] * L
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(300 + c, 3) for c in range(count)])
    kas, kbs = rns.uniform(1, k), rns.uniform(2, k)
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((count, L, rns.n)), ctx.empty((count, L, rns.n))
    ctx.relin_apply(dc, 2, 0b111, dka, dkb, k, log_bs, o0, o1, count, pids)
    g0, g1 = o0.download(), o1.download()
    for c in range(count):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0), c
        assert np.array_equal(g1[c], e1), c


def test_relin_degree3_and_coeff_form_component(ctx):
    qs, rns = _rns(8, 2)
    pids = ctx.primes(qs, 8)
    log_bs = [25, 25]
    k, _ = rns.gadget_dims(log_bs)
    comps = rns.uniform(5, 4)[None]
    kas, kbs = rns.uniform(6, 2 * k), rns.uniform(7, 2 * k)
    mask = 0b1011  # component 2 in coefficient form
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((1, 2, rns.n)), ctx.empty((1, 2, rns.n))
    ctx.relin_apply(dc, 3, mask, dka, dkb, 2 * k, log_bs, o0, o1, 1, pids)
    e0, e1 = rns.relin_apply(comps[0], kas, kbs, log_bs, ntt_mask=mask)
    assert np.array_equal(o0.download()[0], e0) and np.array_equal(o1.download()[0], e1)


@pytest.mark.parametrize("log_b", [10, 16])
def test_relin_degree3_n16384_accumulates(ctx, log_b):
    """Degree 3 at 2^14: the second source component's key switch adds onto
    the first one's output (digit planes packed at 10 bits, or 16-bit rows),
    for two ciphertexts."""
    qs, rns = _rns(14, 2)
    pids = ctx.primes(qs, 14)
    log_bs = [log_b, log_b]
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(40 + c, 4) for c in range(2)])
    kas, kbs = rns.uniform(41, 2 * k), rns.uniform(42, 2 * k)
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((2, 2, rns.n)), ctx.empty((2, 2, rns.n))
    ctx.relin_apply(dc, 3, 0b1111, dka, dkb, 2 * k, log_bs, o0, o1, 2, pids)
    g0, g1 = o0.download(), o1.download()
    for c in range(2):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), c


@pytest.mark.parametrize("log_n,L,count", [(7, 2, 2), (12, 2, 16), (14, 8, 2)])
def test_galois_apply(ctx, log_n, L, count):
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(400 + c, 2) for c in range(count)])
    kas, kbs = rns.uniform(11, k), rns.uniform(12, k)
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((count, L, rns.n)), ctx.empty((count, L, rns.n))
    ctx.galois_apply(dc, True, dka, dkb, log_bs, o0, o1, count, pids)
    g0, g1 = o0.download(), o1.download()
    for c in range(count):
        e0, e1 = rns.galois_apply(comps[c, 0], comps[c, 1], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0), c
        assert np.array_equal(g1[c], e1), c


@pytest.mark.parametrize("terms", [1, 5, 64])
def test_lazy_fma(ctx, terms):
    qs, rns = _rns(12, 2)
    pids = ctx.primes(qs, 12)
    as_, bs = rns.uniform(31, terms), rns.uniform(32, terms)
    da, db = ctx.array(as_), ctx.array(bs)
    out = ctx.empty((2, rns.n))
    words = 2 * rns.n
    ctx.lazy_fma(da, words, db, words, terms, out, pids)
    assert np.array_equal(out.download(), rns.lazy_fma(as_, bs))


def test_bgv_chain_decrypts_on_gpu(ctx):
    """Tensor product -> relinearize -> Galois x->x^5 -> mod switch on the
    GPU, each step bit-exact against the oracle and the result decrypting to
    the plaintext computation."""
    log_n, L, t = 10, 4, 65537
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    bgv = Bgv(rns, t, seed=3)
    r = np.random.default_rng(4)
    m1, m2 = r.integers(0, t, rns.n), r.integers(0, t, rns.n)
    c1, c2 = bgv.encrypt(m1), bgv.encrypt(m2)
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    rka, rkb = bgv.relin_key(log_bs)
    gka, gkb = bgv.galois_key(5, log_bs)
    da, db = ctx.array(c1[None]), ctx.array(c2[None])
    prod = ctx.empty((1, 3, L, rns.n))
    ctx.tensor_product(da, 2, db, 2, prod, 1, pids)
    exp_prod = rns.tensor_product(c1, c2)
    assert np.array_equal(prod.download()[0], exp_prod)
    r0, r1 = ctx.empty((1, L, rns.n)), ctx.empty((1, L, rns.n))
    ctx.relin_apply(prod, 2, 0b111, ctx.array(rka), ctx.array(rkb), k, log_bs, r0, r1, 1, pids)
    e0, e1 = rns.relin_apply(exp_prod, rka, rkb, log_bs)
    assert np.array_equal(r0.download()[0], e0) and np.array_equal(r1.download()[0], e1)
    relin = ctx.array(np.stack([e0, e1])[None])
    sub = ctx.empty((1, 2, L, rns.n))
    ctx.substitute(relin, sub, 5, 2, pids)
    g0, g1 = ctx.empty((1, L, rns.n)), ctx.empty((1, L, rns.n))
    ctx.galois_apply(sub, True, ctx.array(gka), ctx.array(gkb), log_bs, g0, g1, 1, pids)
    gs = np.stack([g0.download()[0], g1.download()[0]])
    expect = np.array(substitute_coeffs(negacyclic_mul_mod_t(m1, m2, t), 5, t), dtype=np.int64)
    assert np.array_equal(bgv.decrypt(gs), expect)
    ms = ctx.array(gs)
    ctx.mod_reduce_lsb(ms, t, rns.ql_inv(), True, 2, pids)
    red = rns.sub(L - 1)
    out = ms.download()[:, : L - 1]
    bgv_r = Bgv.__new__(Bgv)
    bgv_r.rns, bgv_r.t, bgv_r.s = red, t, bgv.s[: L - 1].copy()
    assert np.array_equal(bgv_r.decrypt(out), expect)


def test_errors_leave_outputs_untouched(ctx):
    qs, rns = _rns(8, 2)
    pids = ctx.primes(qs, 8)
    x = rns.uniform(1, 1)
    d = ctx.array(x)
    with pytest.raises(sg.ShgError):
        ctx.mod_reduce_lsb(d, 65537, [], True, 1, pids[:1])
    with pytest.raises(sg.ShgError):
        ctx.ntt_fwd(d, 1, [pids[0], 999])
    assert np.array_equal(d.download(), x)


_GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "reference_moduli.json")))


@pytest.mark.parametrize("entry", _GOLDEN["moduli"], ids=lambda e: "%d-%d" % (e["q"], e["log_n"]))
def test_ntt_reference_moduli(ctx, entry):
    """Every (q, log n) pair of the reference's own tests, through both
    transform variants that accept it (FP64 needs q < 2^50)."""
    q, log_n = entry["q"], entry["log_n"]
    rns = oracle.Rns([q], log_n)
    pids = ctx.primes([q], log_n)
    x = rns.uniform(q % 997, 3)
    exp = np.stack([rns.ntt_fwd(v) for v in x])
    for variant in VARIANTS:
        if variant == sg.NTT_F64 and q >= 2**50:
            continue
        d = ctx.array(x)
        ctx.ntt_fwd(d, 3, pids, variant)
        assert np.array_equal(d.download(), exp)
        ctx.ntt_inv(d, 3, pids, variant)
        assert np.array_equal(d.download(), x)


CONFIGS = [sg.NTT_F64 | sg.NTT_TILING_ALT, sg.NTT_F64 | sg.NTT_PREFETCH,
           sg.NTT_F64 | sg.NTT_TILING_ALT | sg.NTT_PREFETCH, sg.NTT_INT | sg.NTT_TILING_ALT,
           sg.NTT_F64 | sg.NTT_SPLIT_LIMB, sg.NTT_INT | sg.NTT_SPLIT_LIMB,
           sg.NTT_F64 | sg.NTT_WAVE_EXCHANGE, sg.NTT_INT | sg.NTT_WAVE_EXCHANGE,
           sg.NTT_F64 | sg.NTT_STAGE_SPLIT, sg.NTT_INT | sg.NTT_STAGE_SPLIT,
           sg.NTT_F64 | sg.NTT_STAGE_SPLIT | sg.NTT_TILING_ALT, sg.NTT_F64 | sg.NTT_TWO_PER_CU,
           sg.NTT_INT | sg.NTT_TWO_PER_CU]


@pytest.mark.parametrize("cfg", [sg.NTT_F64, sg.NTT_F64 | sg.NTT_TWO_PER_CU])
def test_ntt_fwd_n16384_every_limb(ctx, cfg):
    """Every one of 640 limbs of a forward transform at 2^14 (more limbs than
    resident workgroups) equals the oracle's, not only a sample."""
    log_n, L, count = 14, 4, 160
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    x = rns.uniform(777, count)
    d = ctx.array(x)
    ctx.ntt_fwd(d, count, pids, cfg)
    got = d.download()
    bad = [c for c in range(count) if not np.array_equal(got[c], rns.ntt_fwd(x[c]))]
    assert not bad, bad


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("log_n,L,count", [(12, 2, 3), (13, 3, 2), (14, 4, 3), (15, 2, 2)])
def test_ntt_launch_configs(ctx, cfg, log_n, L, count):
    """Every register tiling, the prefetching loop, the split forward
    transform (2^14 only), the wave-exchange forward and inverse
    transforms (FP64, 2^12..2^14; other shapes ignore the flags) and the
    two-workgroups-per-CU transforms (FP64, 2^14; also the default there),
    with more limbs than resident workgroups at 2^14 so the persistent loops
    iterate.
    The inverse is checked against the oracle on its own input as well as by
    the round trip."""
    if log_n == 14:
        count = 160  # 640 limbs > 256 CUs at one workgroup each
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    x = rns.uniform(500 + log_n + cfg, count)
    d = ctx.array(x)
    if cfg & sg.NTT_STAGE_SPLIT and log_n < 13:  # half-limb slices start at 2^13
        with pytest.raises(sg.ShgError):
            ctx.ntt_fwd(d, count, pids, cfg)
        return
    ctx.ntt_fwd(d, count, pids, cfg)
    got = d.download()
    for c in sorted({0, 1, count - 1}):
        assert np.array_equal(got[c], rns.ntt_fwd(x[c])), c
    ctx.ntt_inv(d, count, pids, cfg)
    assert np.array_equal(d.download(), x)
    # The inverse of uniform (not transform-output) words.
    y = rns.uniform(900 + log_n + cfg, 2)
    d2 = ctx.array(y)
    ctx.ntt_inv(d2, 2, pids, cfg)
    got = d2.download()
    for c in range(2):
        assert np.array_equal(got[c], rns.ntt_inv(y[c])), c


def test_autotune_then_auto(ctx):
    qs, rns = _rns(13, 4)
    pids = ctx.primes(qs, 13)
    chosen = ctx.autotune_ntt(13, 256)
    assert chosen & 3 in (sg.NTT_INT, sg.NTT_F64) and chosen & ~255 == 0
    x = rns.uniform(77, 64)  # 256 limbs: the tuned shape
    d = ctx.array(x)
    ctx.ntt_fwd(d, 64, pids)
    got = d.download()
    for c in (0, 63):
        assert np.array_equal(got[c], rns.ntt_fwd(x[c]))
    ctx.ntt_inv(d, 64, pids)
    assert np.array_equal(d.download(), x)
    with pytest.raises(sg.ShgError):
        ctx.ntt_fwd(d, 64, pids, 3)


@pytest.mark.parametrize("log_n,L,count,log_bound", [(14, 8, 4, 50), (14, 3, 2, 60), (13, 2, 1, 50),
                                                     (15, 2, 3, 62)])
def test_autotune_moduli_then_auto(ctx, log_n, L, count, log_bound):
    """The autotuner over the caller's moduli: wide moduli (2^50 and above)
    get an integer configuration, and the AUTO calls of that shape that
    follow use the table's pick, bit-exact against the oracle."""
    qs = ntt_primes(log_n, L, log_bound)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    chosen = ctx.autotune_ntt_moduli(pids, count)
    assert chosen & ~255 == 0
    if log_bound > 50:
        assert chosen & 3 == sg.NTT_INT
    x = rns.uniform(31 + log_n, count)
    d = ctx.array(x)
    ctx.ntt_fwd(d, count, pids)
    got = d.download()
    for c in range(count):
        assert np.array_equal(got[c], rns.ntt_fwd(x[c])), c
    ctx.ntt_inv(d, count, pids)
    assert np.array_equal(d.download(), x)


def test_relin_apply_mixed_bases_n16384(ctx):
    """Different gadget bases per limb (rns/rns_gadget.cc:42-106 takes one
    per modulus): the stage-0 table covers the largest, 2^11, and digits of
    the 2^5- and 2^9-base limbs index its low entries."""
    qs, rns = _rns(14, 3)
    pids = ctx.primes(qs, 14)
    log_bs = [11, 5, 9]
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(700 + c, 3) for c in range(2)])
    kas, kbs = rns.uniform(21, k), rns.uniform(22, k)
    dc = ctx.array(comps)
    dka, dkb = ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((2, 3, rns.n)), ctx.empty((2, 3, rns.n))
    ctx.relin_apply(dc, 2, 0b111, dka, dkb, k, log_bs, o0, o1, 2, pids)
    g0, g1 = o0.download(), o1.download()
    for c in range(2):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), c


@pytest.mark.parametrize("log_n,L,log_b,count", [(14, 8, 10, 3), (13, 3, 10, 2), (12, 2, 10, 4),
                                                 (14, 2, 20, 1), (14, 2, 16, 2), (14, 2, 8, 1),
                                                 (14, 3, 12, 2)])
def test_prepared_key_relin_and_galois(ctx, log_n, L, log_b, count):
    """Key switching with a prepared key (shg_ks_key: FP64-form copies read
    by the digit-plane kernel at 2^13 and 2^14; the Montgomery words on every
    other path) equals the oracle and the plain entry points word for word."""
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    log_bs = [log_b] * L
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(800 + c, 3) for c in range(count)])
    kas, kbs = rns.uniform(31, k), rns.uniform(32, k)
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    key = ctx.ks_key(dka, dkb, k, pids)
    o0, o1 = ctx.empty((count, L, rns.n)), ctx.empty((count, L, rns.n))
    ctx.relin_apply_key(dc, 2, 0b111, key, log_bs, o0, o1, count, pids)
    g0, g1 = o0.download(), o1.download()
    for c in sorted({0, count - 1}):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), c
    p0, p1 = ctx.empty((count, L, rns.n)), ctx.empty((count, L, rns.n))
    ctx.relin_apply(dc, 2, 0b111, dka, dkb, k, log_bs, p0, p1, count, pids)
    assert np.array_equal(p0.download(), g0) and np.array_equal(p1.download(), g1)
    gal = np.ascontiguousarray(comps[:, :2])
    dg = ctx.array(gal)
    ctx.galois_apply_key(dg, True, key, log_bs, o0, o1, count, pids)
    g0, g1 = o0.download(), o1.download()
    e0, e1 = rns.galois_apply(gal[0, 0], gal[0, 1], kas, kbs, log_bs)
    assert np.array_equal(g0[0], e0) and np.array_equal(g1[0], e1)
    key.close()


def test_prepared_key_wide_moduli_and_checks(ctx):
    """Moduli of 2^50 and above keep no FP64 copy (the integer key switch
    reads the Montgomery words); a key is tied to its moduli."""
    log_n, L = 13, 2
    qs = ntt_primes(log_n, L, 60)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    log_bs = [20] * L
    k, _ = rns.gadget_dims(log_bs)
    comps = rns.uniform(41, 3)[None]
    kas, kbs = rns.uniform(42, k), rns.uniform(43, k)
    dka, dkb = ctx.array(kas), ctx.array(kbs)
    key = ctx.ks_key(dka, dkb, k, pids)
    o0, o1 = ctx.empty((1, L, rns.n)), ctx.empty((1, L, rns.n))
    ctx.relin_apply_key(ctx.array(comps), 2, 0b111, key, log_bs, o0, o1, 1, pids)
    e0, e1 = rns.relin_apply(comps[0], kas, kbs, log_bs)
    assert np.array_equal(o0.download()[0], e0) and np.array_equal(o1.download()[0], e1)
    other = ctx.primes(ntt_primes(log_n, L), log_n)
    with pytest.raises(sg.ShgError, match="other RNS moduli"):
        ctx.relin_apply_key(ctx.array(comps), 2, 0b111, key, log_bs, o0, o1, 1, other)
    key.close()


@pytest.mark.parametrize("log_n,log_b", [(14, 16), (14, 10), (13, 16)])
def test_prepared_key_extreme_inputs(ctx, log_n, log_b):
    """The prepared-key schedule's bounds at their edge: c2 whose every
    coefficient exports to q - 1 (for these moduli just below 2^50, digits
    of all ones in nearly every position), c0 and c1 at q - 1, and key
    residues at floor(q/2), floor(q/2) + 1, q - 1 and 0 (the largest
    balanced magnitudes and both signs)."""
    L = 3
    qs, rns = _rns(log_n, L)
    pids = ctx.primes(qs, log_n)
    log_bs = [log_b] * L
    k, _ = rns.gadget_dims(log_bs)
    comps = np.empty((1, 3, L, rns.n), dtype=np.uint64)
    kas = np.empty((k, L, rns.n), dtype=np.uint64)
    kbs = np.empty_like(kas)
    for i, q in enumerate(qs):
        comps[0, :2, i, :] = q - 1
        comps[0, 2, i, :] = rns.primes[i].mont.imp(q - 1)
        pat = np.array([q // 2, q // 2 + 1, q - 1, 0], dtype=np.uint64)
        kas[:, i, :] = np.resize(pat, rns.n)
        kbs[:, i, :] = np.resize(pat[::-1], rns.n)
    comps[0, 2] = rns.ntt_fwd(comps[0, 2])  # NTT form: the digit-plane path
    dka, dkb = ctx.array(kas), ctx.array(kbs)
    key = ctx.ks_key(dka, dkb, k, pids)
    o0, o1 = ctx.empty((1, L, rns.n)), ctx.empty((1, L, rns.n))
    ctx.relin_apply_key(ctx.array(comps), 2, 0b111, key, log_bs, o0, o1, 1, pids)
    e0, e1 = rns.relin_apply(comps[0], kas, kbs, log_bs)
    assert np.array_equal(o0.download()[0], e0) and np.array_equal(o1.download()[0], e1)
    key.close()


@pytest.mark.parametrize("seed", range(6))
def test_relin_random_shapes(ctx, seed):
    """Relinearization on randomly drawn shapes against the oracle: ring
    degree 2^4..2^15, 1..8 moduli of 30..50 bits, a gadget base per limb
    from 2^2 to 2^24 (digit planes, the stage-0 table and the unfused form
    as the draw falls), degree 2 or 3, NTT or coefficient-form sources, and
    the plain or prepared key."""
    r = np.random.default_rng(1000 + seed)
    log_n = int(r.integers(4, 16))
    L = int(r.integers(1, 9 if log_n <= 14 else 3))
    bound = int(r.integers(30, 51))
    qs = ntt_primes(log_n, L, bound)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    log_bs = [int(b) for b in r.integers(2, 25, size=L)]
    k, _ = rns.gadget_dims(log_bs)
    degree = int(r.integers(2, 4))
    mask = 0b11 | (int(r.integers(0, 1 << (degree - 1))) << 2)
    count = int(r.integers(1, 4))
    comps = np.stack([rns.uniform(2000 + 7 * seed + c, degree + 1) for c in range(count)])
    nk = k * (degree - 1)
    kas, kbs = rns.uniform(3000 + seed, nk), rns.uniform(4000 + seed, nk)
    dc, dka, dkb = ctx.array(comps), ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((count, L, rns.n)), ctx.empty((count, L, rns.n))
    if seed % 2:
        key = ctx.ks_key(dka, dkb, nk, pids)
        ctx.relin_apply_key(dc, degree, mask, key, log_bs, o0, o1, count, pids)
        key.close()
    else:
        ctx.relin_apply(dc, degree, mask, dka, dkb, nk, log_bs, o0, o1, count, pids)
    g0, g1 = o0.download(), o1.download()
    info = (log_n, L, bound, log_bs, degree, mask, count)
    for c in range(count):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs, ntt_mask=mask)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), (info, c)


@pytest.mark.gpu
@pytest.mark.parametrize("L,count,other_bits,last_bits",
                         [(2, 1, 50, 50), (12, 3, 50, 50), (5, 2, 30, 50), (5, 2, 50, 30)])
def test_mod_reduce_lsb_n16384_two_launch_shapes(ctx, L, count, other_bits, last_bits):
    """ModReduceLsb at n = 2^14 in NTT form with FP64 arithmetic runs as two
    launches (the dropped limbs' inverse into scratch, then one workgroup per
    remaining limb).  Its edges: two limbs and a single polynomial (one
    workgroup in each launch), twelve limbs, and moduli of different widths
    either way round, so that the balanced lift from q_l to q_i crosses
    sizes (rns_polynomial.cc:262-300, modulus_conversion.h:30-83)."""
    log_n = 14
    if other_bits == last_bits:
        qs = ntt_primes(log_n, L, other_bits)
    else:
        qs = ntt_primes(log_n, L - 1, other_bits) + ntt_primes(log_n, 1, last_bits)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    x = rns.uniform(7100 + L + other_bits, count)
    ql_inv = ctx.ql_inv(pids)
    d = ctx.array(x)
    ctx.mod_reduce_lsb(d, 65537, ql_inv, True, count, pids)
    got = d.download()
    for c in range(count):
        assert np.array_equal(got[c, :L - 1], rns.mod_reduce_lsb(x[c], 65537, ql_inv, True)), c
        assert np.array_equal(got[c, L - 1], x[c, L - 1])
