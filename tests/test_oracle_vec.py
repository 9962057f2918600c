"""The vectorised host baseline (oracle/shell_cpu_vec.c: the reference's
Highway forward transform, dft_transformations_hwy.cc:64-158, and
RnsRelinKey::ApplyToRlweCiphertext over it) against the scalar
restatements, word for word: the same transform and key switch computed
another way, including moduli near 2^62 where the lazy [0, 4q) ranges are
tightest and ring degrees whose every layer is scalar or vector."""
import numpy as np
import pytest

import oracle
from common import ntt_primes


@pytest.mark.parametrize("log_n,L,log_bound", [(3, 2, 50), (4, 2, 50), (10, 3, 50),
                                               (12, 2, 62), (14, 2, 50), (15, 1, 61)])
def test_hwy_ntt_matches_scalar(log_n, L, log_bound):
    rns = oracle.Rns(ntt_primes(log_n, L, log_bound), log_n)
    x = rns.uniform(5, 3)
    for i, q in enumerate(rns.qs):
        x[0, i, : 8] = q - 1  # extreme residues
        x[1, i, : 8] = 0
    for c in range(3):
        assert np.array_equal(rns.hwy_ntt_fwd(x[c]), rns.ntt_fwd(x[c])), c


@pytest.mark.parametrize("log_n,L,log_b", [(6, 3, 10), (10, 2, 20)])
def test_hwy_relin_matches_scalar(log_n, L, log_b):
    rns = oracle.Rns(ntt_primes(log_n, L), log_n)
    log_bs = [log_b] * L
    k = sum((q.bit_length() + log_b - 1) // log_b for q in rns.qs)
    comps = rns.uniform(7, 3)
    kas, kbs = rns.uniform(8, k), rns.uniform(9, k)
    s0, s1 = rns.relin_apply(comps, kas, kbs, log_bs)
    v0, v1 = rns.relin_apply(comps, kas, kbs, log_bs, vectorised=True)
    assert np.array_equal(s0, v0) and np.array_equal(s1, v1)
