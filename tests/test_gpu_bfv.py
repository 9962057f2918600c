"""The BFV-side RNS operations on the GPU against the oracle, word for word:
ModReduceMsb (rns/rns_polynomial.cc:349-417) in both forms on the fused
per-polynomial kernel (n >= 2^11) and the five-kernel form, and the basis
conversions SwitchRnsBasis, ScaleAndSwitchRnsBasis and ApproxSwitchRnsBasis
(:419-701), including NTT-form input and the balanced lift, and the
composition ApproxModReduceMsb performs (:799-852)."""
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


@pytest.mark.parametrize("log_n,L,count", [(5, 3, 3), (10, 2, 4), (11, 4, 2), (13, 6, 3),
                                           (14, 8, 4), (15, 3, 2)])
@pytest.mark.parametrize("is_ntt", [True, False])
def test_mod_reduce_msb(ctx, log_n, L, count, is_ntt):
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    x = rns.uniform(60 + log_n, count)
    ql_inv = ctx.ql_inv(pids)
    d = ctx.array(x)
    ctx.mod_reduce_msb(d, ql_inv, is_ntt, count, pids)
    got = d.download()
    for c in range(count):
        assert np.array_equal(got[c, : L - 1], rns.mod_reduce_msb(x[c], ql_inv, is_ntt)), c
        assert np.array_equal(got[c, L - 1], x[c, L - 1])


def _chains(log_n, L, K):
    qs = ntt_primes(log_n, L + K)
    return qs[:L], qs[L:]


@pytest.mark.parametrize("log_n,L,K,count", [(4, 2, 1, 3), (10, 3, 2, 2), (12, 4, 4, 2),
                                             (14, 8, 3, 1), (13, 16, 16, 1)])
def test_switch_and_scale_rns_basis(ctx, log_n, L, K, count):
    qs, ps = _chains(log_n, L, K)
    rq, rp = oracle.Rns(qs, log_n), oracle.Rns(ps, log_n)
    iq, ip = ctx.primes(qs, log_n), ctx.primes(ps, log_n)
    c = rq.basis_constants(rp)
    x = rq.uniform(80 + L, count)
    x[0, :, 0] = 0
    dx = ctx.array(x)
    out = ctx.empty((count, K, rq.n))
    ctx.switch_rns_basis(sg.BASIS_EXACT, dx, False, False, count, iq, out, ip, c["q_hat_inv"],
                         c["q_hat_mod_ps"], c["q_mod_ps"])
    got = out.download()
    for k in range(count):
        assert np.array_equal(got[k], rq.switch_rns_basis(x[k], rp, c)), k
    ctx.switch_rns_basis(sg.BASIS_SCALE, dx, False, False, count, iq, out, ip, c["q_hat_inv"],
                         c["q_inv_mod_ps"], c["p_mod_qs"])
    got = out.download()
    for k in range(count):
        assert np.array_equal(got[k], rq.scale_and_switch_rns_basis(x[k], rp, c)), k
    with pytest.raises(sg.ShgError):  # coefficient form only, as in the reference
        ctx.switch_rns_basis(sg.BASIS_EXACT, dx, True, False, count, iq, out, ip, c["q_hat_inv"],
                             c["q_hat_mod_ps"], c["q_mod_ps"])


@pytest.mark.parametrize("log_n,L,K", [(10, 3, 2), (12, 5, 3), (14, 8, 4)])
@pytest.mark.parametrize("is_ntt", [False, True])
@pytest.mark.parametrize("balanced", [False, True])
def test_approx_switch_rns_basis(ctx, log_n, L, K, is_ntt, balanced):
    qs, ps = _chains(log_n, L, K)
    rq, rp = oracle.Rns(qs, log_n), oracle.Rns(ps, log_n)
    iq, ip = ctx.primes(qs, log_n), ctx.primes(ps, log_n)
    c = rq.basis_constants(rp)
    count = 2
    x = rq.uniform(90 + L, count)
    dx = ctx.array(x)
    out = ctx.empty((count, K, rq.n))
    ctx.switch_rns_basis(sg.BASIS_APPROX, dx, is_ntt, balanced, count, iq, out, ip,
                         c["q_hat_inv"], c["q_hat_mod_ps"])
    got = out.download()
    for k in range(count):
        assert np.array_equal(got[k], rq.approx_switch_rns_basis(x[k], rp, c, is_ntt, balanced)), k
    assert np.array_equal(dx.download(), x)  # input untouched


def test_approx_mod_reduce_msb_composition(ctx):
    """ApproxModReduceMsb (rns/rns_polynomial.cc:799-852): delta =
    ApproxSwitch(aux -> Q), NTT, (a - delta) * P^-1, on the device and in the
    oracle."""
    log_n, L, K = 12, 4, 2
    qs, ps = _chains(log_n, L, K)
    rq, rp = oracle.Rns(qs, log_n), oracle.Rns(ps, log_n)
    iq, ip = ctx.primes(qs, log_n), ctx.primes(ps, log_n)
    cp = rp.basis_constants(rq)  # aux chain P -> main chain Q
    P = 1
    for p in ps:
        P *= p
    p_inv = np.array([rq.primes[i].mont.imp(pow(P, -1, q)) for i, q in enumerate(qs)],
                     dtype=np.uint64)
    a = rq.uniform(5)                    # NTT form mod Q
    aux = rp.uniform(6)                  # coefficient form mod P
    delta_o = rq.ntt_fwd(rp.approx_switch_rns_basis(aux, rq, cp, False, False))
    expect = rq.mul_scalar(rq.pointwise("sub", a, delta_o), p_inv)
    da, daux = ctx.array(a), ctx.array(aux)
    delta = ctx.empty((1, L, rq.n))
    ctx.switch_rns_basis(sg.BASIS_APPROX, daux, False, False, 1, ip, delta, iq, cp["q_hat_inv"],
                         cp["q_hat_mod_ps"])
    ctx.ntt_fwd(delta, 1, iq)
    ctx.pointwise(sg.OP_SUB, da, delta, None, 1, iq)
    ctx.mul_scalar(da, p_inv, 1, iq)
    assert np.array_equal(da.download(), expect)
