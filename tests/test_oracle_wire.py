"""Pins the oracle's wire-format and PRNG restatements:

* TranscribeBits and MontgomeryInt::SerializeVector: the property the
  reference's own test states (transcription_test.cc:175-216: bit k of the
  stream is bit k % i of input k / i and bit k % j of output k / j, all other
  output bits zero), a hand-worked example, the closed form of the packed
  stream as one little-endian integer, and the round trip;
* ChaCha20 against the RFC 7539 Sec. 2.3.2 block and keystreams written by
  the openssl tool (tests/golden/chacha_vectors.json), and the PRNG built on
  it (prng/chacha_prng_util.cc) against the same streams;
* SampleUniform's rejection rule: values are the low log2(q) bits of
  consecutive little-endian 64-bit draws, skipping those >= q.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from common import ntt_primes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chacha_vectors.json")


@pytest.mark.parametrize("i,j", [(1, 8), (3, 8), (8, 3), (50, 8), (8, 50), (62, 8), (8, 62),
                                 (13, 64), (64, 13), (64, 64), (7, 5)])
def test_transcribe_bits_property(i, j):
    rng = np.random.default_rng(i * 100 + j)
    for length in (1, 2, 5, 17):
        hi = (1 << i) - 1 if i < 64 else (1 << 64) - 1
        vals = rng.integers(0, hi, size=length, dtype=np.uint64, endpoint=True)
        in_type = 8 if i <= 8 else 64
        out_type = 8 if j <= 8 else 64
        out = oracle.transcribe_bits(vals, in_type, length * i, i, j, out_type)
        assert out.size == (length * i + j - 1) // j
        for bit in range(length * i):
            assert (int(vals[bit // i]) >> (bit % i)) & 1 == (int(out[bit // j]) >> (bit % j)) & 1
        if j < 64:
            assert all(int(v) >> j == 0 for v in out)


def test_serialize_vector_example_and_closed_form():
    # Two 3-bit-style values would not fit a real modulus; use q = 13 (log 4):
    # values 5 = 0101, 12 = 1100 -> stream bits 0-3 = 5, 4-7 = 12 -> byte 0xc5.
    m = oracle.Mont(13)
    assert m.s.log_modulus == 4
    buf = np.zeros(4, dtype=np.uint8)
    r = oracle.lib().so_serialize_vector(m.ref(), oracle._p64(np.array([5, 12], dtype=np.uint64)),
                                         2, oracle._p8(buf), 4)
    assert r == 1 and buf[0] == 0xC5
    for log_n, bound in [(4, 50), (10, 50), (12, 30), (3, 62), (11, 55)]:
        qs = ntt_primes(log_n, 2, bound)
        rns = oracle.Rns(qs, log_n)
        x = rns.uniform(log_n * 7 + bound)
        x[0, 0], x[1, -1] = qs[0] - 1, qs[1]  # q itself (NegateInPlace(0)) fits log q bits
        vecs = rns.serialize(x)
        for i, q in enumerate(qs):
            b = q.bit_length()
            stream = sum(int(v) << (k * b) for k, v in enumerate(x[i]))
            assert vecs[i] == stream.to_bytes((rns.n * b + 7) // 8, "little")
        assert np.array_equal(rns.deserialize(vecs), x)


def test_deserialize_rejects_short_input():
    qs = ntt_primes(5, 1)
    rns = oracle.Rns(qs, 5)
    v = rns.serialize(rns.uniform(3))
    with pytest.raises(ValueError):
        rns.deserialize([v[0][:-1]])


def test_chacha20_block_matches_rfc_and_openssl():
    g = json.load(open(GOLDEN))
    assert g["blocks"][0]["keystream"].startswith("10f1e7e4d13b5915500fdd1f")  # RFC 7539 2.3.2
    for b in g["blocks"]:
        ks = oracle.chacha20_block(bytes.fromhex(b["key"]), b["counter"], bytes.fromhex(b["nonce"]))
        assert ks.hex() == b["keystream"]


def test_chacha_prng_stream_matches_openssl():
    g = json.load(open(GOLDEN))
    for s in g["prng_streams"]:
        prng = oracle.ChaChaPrng(bytes.fromhex(s["key"]))
        for buf in s["buffers"]:
            data = bytes(prng.rand8() for _ in range(255 * 32))
            assert hashlib.sha256(data).hexdigest() == buf["sha256"], buf["salt"]
            assert data[:16].hex() == buf["head"] and data[-16:].hex() == buf["tail"]
    # Rand64 is eight Rand8 little-endian.
    a, b = oracle.ChaChaPrng(bytes(32)), oracle.ChaChaPrng(bytes(32))
    for _ in range(300):
        assert a.rand64() == sum(b.rand8() << (8 * k) for k in range(8))


@pytest.mark.parametrize("log_n,bound", [(6, 50), (5, 30), (4, 62), (7, 20)])
def test_sample_uniform_rejection_rule(log_n, bound):
    qs = ntt_primes(log_n, 2, bound)
    rns = oracle.Rns(qs, log_n)
    key = bytes(range(7, 39))
    got = rns.sample_uniform_chacha(oracle.ChaChaPrng(key), 2)
    src = oracle.ChaChaPrng(key)
    for c in range(2):
        for i, q in enumerate(qs):
            b = q.bit_length()
            for k in range(rns.n):
                while True:
                    v = src.rand64() & ((1 << b) - 1)
                    if v < q:
                        break
                assert int(got[c, i, k]) == v
