"""Wire formats on the GPU against the oracle: bit-packed coefficient
vectors (MontgomeryInt::SerializeVector per limb, montgomery.cc:143-219)
written and read by device kernels, and the seeded random pad
(RnsPolynomial::SampleUniform over SingleThreadChaChaPrng) generated on the
device, word for word."""
import numpy as np
import pytest

import oracle
import shell_gpu as sg
from common import is_prime, ntt_primes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sg.Context(0)
    yield c
    c.close()


def low_primes(log_n, count, bits):
    """The smallest primes above 2^(bits-1) with q = 1 mod 2n: about half of
    all log2(q)-bit draws are rejected."""
    step = 2 << log_n
    q = ((1 << (bits - 1)) // step + 1) * step + 1
    out = []
    while len(out) < count:
        if is_prime(q):
            out.append(q)
        q += step
    return out


# (log n, moduli): byte strings of n * log2(q) bits that are and are not
# multiples of 8 bytes, mixed widths in one chain, and a chain wider than 50
# bits.
SHAPES = [(2, ntt_primes(2, 2)), (5, ntt_primes(5, 3)), (6, ntt_primes(6, 2)),
          (10, ntt_primes(10, 1, 30) + ntt_primes(10, 2, 50)), (12, ntt_primes(12, 2, 61)),
          (14, ntt_primes(14, 8)), (15, ntt_primes(15, 3, 45))]


@pytest.mark.parametrize("log_n,qs", SHAPES)
@pytest.mark.parametrize("offset", [0, 3])
def test_serialize_round_trip_matches_oracle(ctx, log_n, qs, offset):
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    count = 3
    x = rns.uniform(50 + log_n, count)
    x[0, 0, 0] = qs[0]  # NegateInPlace(0) = q also fits log2(q) bits
    limb_bytes, poly_bytes = ctx.serialized_bytes(pids)
    assert limb_bytes == [(rns.n * q.bit_length() + 7) // 8 for q in qs]
    dx = ctx.array(x)
    words = (count * poly_bytes + offset + 7) // 8 + 1
    buf = ctx.array(np.zeros(words, dtype=np.uint64))
    ctx.serialize(dx, count, pids, buf, offset)
    raw = buf.download().view(np.uint8)
    assert not raw[:offset].any()
    for c in range(count):
        expect = b"".join(rns.serialize(x[c]))
        start = offset + c * poly_bytes
        assert raw[start:start + poly_bytes].tobytes() == expect, c
    assert not raw[offset + count * poly_bytes:].any()
    back = ctx.empty(x.shape)
    ctx.deserialize(buf, count, pids, back, offset)
    assert np.array_equal(back.download(), x)
    # Bytes written by the oracle load the same words.
    host = np.zeros(words * 8, dtype=np.uint8)
    for c in range(count):
        host[c * poly_bytes:(c + 1) * poly_bytes] = np.frombuffer(
            b"".join(rns.serialize(x[c])), dtype=np.uint8)
    src = ctx.array(host.view(np.uint64))
    ctx.deserialize(src, count, pids, back)
    assert np.array_equal(back.download(), x)


@pytest.mark.parametrize("log_n,qs,count", [
    (10, ntt_primes(10, 3), 3),
    (10, low_primes(10, 2, 30), 3),          # ~50% of draws rejected
    (11, low_primes(11, 1, 20) + ntt_primes(11, 1, 62), 2),
    (14, ntt_primes(14, 8), 40),             # a Galois key's pad at n=2^14, L=8, k=40
])
def test_sample_uniform_chacha_matches_oracle(ctx, log_n, qs, count):
    rns = oracle.Rns(qs, log_n)
    pids = ctx.primes(qs, log_n)
    seed = bytes((7 * i + log_n) % 256 for i in range(32))
    out = ctx.empty((count, len(qs), rns.n))
    ctx.sample_uniform_chacha(seed, count, pids, out)
    expect = rns.sample_uniform_chacha(oracle.ChaChaPrng(seed), count)
    got = out.download()
    for c in range(count):
        assert np.array_equal(got[c], expect[c]), c
    for i, q in enumerate(qs):
        assert int(got[:, i].max()) < q


def test_sample_uniform_chacha_rejects_bad_seed(ctx):
    pids = ctx.primes(ntt_primes(10, 1), 10)
    out = ctx.empty((1, 1, 1 << 10))
    with pytest.raises(sg.ShgError) as e:
        ctx.sample_uniform_chacha(bytes(16), 1, pids, out)
    assert "key length of 16" in e.value.msg
