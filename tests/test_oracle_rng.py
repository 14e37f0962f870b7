"""Pins the oracle's draw protocol: Philox4x32-10 known answers, exact uniforms, the deterministic
log/sqrt and the inverse-normal mapping against independent references."""
import math

import numpy as np
import pytest


# Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11)
KATS = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KATS)
def test_philox_known_answers(orc, ctr, key, expect):
    assert orc.philox(ctr, key).tolist() == expect


def test_uniform_is_open_closed_and_exact(orc):
    L = orc.lib()
    assert L.oro_u01(0) == 2.0 ** -24
    assert L.oro_u01(0xFFFFFFFF) == 1.0
    assert L.oro_u01(0x80000000) == 0.5 + 2.0 ** -24
    for x in (1, 255, 256, 0x12345678):
        assert L.oro_u01(x) == ((x >> 8) + 1) / 2.0 ** 24


def test_deterministic_log_and_sqrt(orc):
    L = orc.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.random(2000), [2.0 ** -52, 1e-300, 0.5, 1.0, 0.999999]])
    for x in xs:
        x = float(x)
        if x <= 0:
            continue
        assert abs(L.oro_det_log(x) - math.log(x)) <= 4e-16 * max(1.0, abs(math.log(x)))
        y = L.oro_det_sqrt(-2.0 * math.log(x)) if x < 1 else None
        if y is not None:
            assert abs(y - math.sqrt(-2.0 * math.log(x))) <= 2.5e-16 * y  # within one ulp


def test_normal_mapping_matches_inverse_cdf(orc):
    """The normal of words (a, b) equals Phi^-1((k + 0.5) 2^-52) with k = (a>>6)<<26 | (b>>6)."""
    from scipy.special import ndtri
    L = orc.lib()
    rng = np.random.default_rng(2)
    words = rng.integers(0, 2 ** 32, size=(20000, 2), dtype=np.uint64)
    # include the far tails and the centre
    words = np.concatenate([words, [[0, 0], [0xFFFFFFFF, 0xFFFFFFFF], [0x80000000, 0], [0x7FFFFFFF, 0xFFFFFFFF]]])
    for a, b in words:
        k = (int(a) >> 6 << 26) | (int(b) >> 6)
        u = (k + 0.5) * 2.0 ** -52
        z = L.oro_normal_from_words(int(a), int(b))
        ref = np.float32(ndtri(u))
        assert abs(z - ref) <= 3e-7 * max(1.0, abs(float(ref))), (a, b, z, ref)


def test_normal_moments(orc):
    L = orc.lib()
    rng = np.random.default_rng(3)
    w = rng.integers(0, 2 ** 32, size=(200000, 2), dtype=np.uint64)
    z = np.array([L.oro_normal_from_words(int(a), int(b)) for a, b in w[:50000]])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02
    assert abs(np.mean(z ** 3)) < 0.05 and abs(np.mean(z ** 4) - 3.0) < 0.1
