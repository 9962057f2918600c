"""Cases at the edges of the key-switch contract: gadget digits above q/2
(bases as wide as the modulus), graph capture of calls that use the
context's work buffers, and two threads expanding queries on one context.
Every output is compared word for word with the oracle."""
import threading

import numpy as np
import pytest

import oracle
import shell_gpu as sg
from common import expand_oracle, ntt_primes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sg.Context(0)
    yield c
    c.close()


def _chain(log_n, spec):
    """spec: (count, log_bound) pairs, concatenated into one modulus chain."""
    qs = []
    for count, bound in spec:
        qs += [q for q in ntt_primes(log_n, count + len(qs), bound) if q not in qs][:count]
    return qs, oracle.Rns(qs, log_n)


# A digit of limb i exceeds q_i / 2 when 2^b - 1 > q_i / 2, i.e. for bases
# 2^(log q_i) and 2^(log q_i - 1); its lift to the other limbs is then
# digit - q_i (rns/rns_gadget.cc:125-151 with modulus_conversion.h:49-81).
# 30-bit chains take the FP64 key switch, 55-bit chains the integer one, and
# the mixed chain lifts both from a smaller and from a larger modulus.
BALANCED = [
    ([(3, 30)], [30, 30, 30]),
    ([(3, 30)], [29, 29, 29]),
    ([(2, 55)], [55, 54]),
    ([(1, 30), (1, 50)], [30, 49]),
    ([(1, 50), (1, 30)], [50, 29]),
]


@pytest.mark.parametrize("spec,log_bs", BALANCED)
def test_gadget_decompose_balanced_digits(ctx, spec, log_bs):
    log_n = 10
    qs, rns = _chain(log_n, spec)
    pids = ctx.primes(qs, log_n)
    k, _ = rns.gadget_dims(log_bs)
    a = rns.uniform(41, 3)
    # Force digits just above and below q_i / 2 on a few coefficients.
    for i, q in enumerate(qs):
        a[0, i, :4] = [q // 2, q // 2 + 1, q - 1, 0]
    dg = ctx.empty((3, k, len(qs), rns.n))
    ctx.gadget_decompose(ctx.array(a), log_bs, dg, 3, pids)
    got = dg.download()
    for c in range(3):
        assert np.array_equal(got[c], rns.gadget_decompose(a[c], log_bs)), c


@pytest.mark.parametrize("spec,log_bs", BALANCED)
def test_relin_and_galois_balanced_digits(ctx, spec, log_bs):
    log_n, count = 10, 3
    qs, rns = _chain(log_n, spec)
    L = len(qs)
    pids = ctx.primes(qs, log_n)
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(50 + c, 3) for c in range(count)])
    kas, kbs = rns.uniform(61, k), rns.uniform(62, k)
    dka, dkb = ctx.array(kas), ctx.array(kbs)
    o0, o1 = ctx.empty((count, L, rns.n)), ctx.empty((count, L, rns.n))
    ctx.relin_apply(ctx.array(comps), 2, 0b111, dka, dkb, k, log_bs, o0, o1, count, pids)
    g0, g1 = o0.download(), o1.download()
    for c in range(count):
        e0, e1 = rns.relin_apply(comps[c], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), c
    ctx.galois_apply(ctx.array(comps[:, :2]), True, dka, dkb, log_bs, o0, o1, count, pids)
    g0, g1 = o0.download(), o1.download()
    for c in range(count):
        e0, e1 = rns.galois_apply(comps[c, 0], comps[c, 1], kas, kbs, log_bs)
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), c


def _relin_case(ctx, log_n, L, count, seed):
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    comps = np.stack([rns.uniform(seed + c, 3) for c in range(count)])
    kas, kbs = rns.uniform(seed + 500, k), rns.uniform(seed + 501, k)
    d = dict(rns=rns, pids=pids, log_bs=log_bs, k=k, comps=comps, kas=kas, kbs=kbs,
             dc=ctx.array(comps), dka=ctx.array(kas), dkb=ctx.array(kbs),
             o0=ctx.empty((count, L, rns.n)), o1=ctx.empty((count, L, rns.n)), count=count)
    return d

def _run_relin(ctx, d):
    ctx.relin_apply(d["dc"], 2, 0b111, d["dka"], d["dkb"], d["k"], d["log_bs"], d["o0"], d["o1"],
                    d["count"], d["pids"])


def _check_relin(d):
    g0, g1 = d["o0"].download(), d["o1"].download()
    for c in range(d["count"]):
        e0, e1 = d["rns"].relin_apply(d["comps"][c], d["kas"], d["kbs"], d["log_bs"])
        assert np.array_equal(g0[c], e0) and np.array_equal(g1[c], e1), c


def test_graph_outlives_buffer_growth():
    """A relinearization captured at a small batch keeps replaying correctly
    after an eager call at a larger batch has grown the context's work
    buffers (the captured ones are retired, not freed)."""
    c = sg.Context(0)
    try:
        small = _relin_case(c, 12, 3, 2, 100)
        big = _relin_case(c, 12, 3, 24, 200)
        _run_relin(c, small)
        c.sync()
        c.graph_begin()
        _run_relin(c, small)
        graph = c.graph_end()
        _run_relin(c, big)  # grows the digit planes past the captured buffer
        c.sync()
        _check_relin(big)
        for rep in range(3):
            small["o0"].upload(np.zeros(small["o0"].shape, dtype=np.uint64))
            small["o1"].upload(np.zeros(small["o1"].shape, dtype=np.uint64))
            graph.launch()
            c.sync()
            _check_relin(small)
            _run_relin(c, big)  # eager work between replays
            c.sync()
        graph.close()
        _run_relin(c, small)
        c.sync()
        _check_relin(small)
    finally:
        c.close()


def test_capture_refuses_growth_and_host_calls():
    c = sg.Context(0)
    try:
        case = _relin_case(c, 12, 2, 3, 300)
        c.graph_begin()
        with pytest.raises(sg.ShgError) as e:
            _run_relin(c, case)  # no eager call at this size yet: buffers would grow
        assert e.value.code == -2 and "capture" in e.value.msg
        with pytest.raises(sg.ShgError) as e:
            c.sync()
        assert e.value.code == -2
        with pytest.raises(sg.ShgError) as e:
            case["o0"].upload(np.zeros(case["o0"].shape, dtype=np.uint64))
        assert e.value.code == -2
        graph = c.graph_end()
        graph.close()
        # Outside capture the same call works and matches the oracle.
        _run_relin(c, case)
        c.sync()
        _check_relin(case)
    finally:
        c.close()


def test_concurrent_expansion_on_one_context(ctx):
    """Two threads expand different queries on one context at once; each
    level holds the context's lock across its three launches."""
    log_n, L, levels = 10, 2, 3
    qs = ntt_primes(log_n, L)
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    keys = [(rns.uniform(1700 + j, k), rns.uniform(1800 + j, k)) for j in range(levels)]
    dkeys = [(ctx.array(a), ctx.array(b)) for a, b in keys]
    K = 1 << levels
    groups = [[rns.uniform(1900 + 10 * t + g, 2) for g in range(1 + t)] for t in range(2)]
    bufs = [[ctx.array(np.concatenate([q[None], np.zeros((K - 1, 2, L, rns.n), np.uint64)]))
             for q in grp] for grp in groups]
    errors = []

    def work(t):
        try:
            for _ in range(3):
                for b, q in zip(bufs[t], groups[t]):
                    b.upload(np.concatenate([q[None], np.zeros((K - 1, 2, L, rns.n), np.uint64)]))
                ctx.expand(bufs[t], levels, dkeys, log_bs, pids)
                ctx.sync()
        except Exception as ex:  # reported below
            errors.append(ex)

    th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for grp, bs in zip(groups, bufs):
        for q, b in zip(grp, bs):
            assert np.array_equal(b.download(), expand_oracle(rns, q, levels, keys, log_bs))
