"""The CPU restatement of the device noise generator against published known
answers: Random123's Philox4x32-10 known-answer vectors, and the law of the
Box-Muller deviates (mean, variance, independence of the two components)."""
import numpy as np

import noise_ref as nr

# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key -> output
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_known_answers():
    for ctr, key, want in KAT:
        got = nr.philox4x32_10(*ctr, *key)
        assert [int(np.asarray(v)) for v in got] == list(want), (ctr, key)


def test_philox_vectorized_equals_scalar():
    modes = np.arange(0, 5000, 7, dtype=np.uint64)
    w = nr.philox4x32_10(modes, 3 << 4, 11, 0, 0x1234, 0x5678)
    for i in (0, 17, len(modes) - 1):
        s = nr.philox4x32_10(int(modes[i]), 3 << 4, 11, 0, 0x1234, 0x5678)
        assert [int(v[i]) for v in w] == [int(np.asarray(v)) for v in s]


def test_box_muller_law():
    z0, z1 = nr.gauss_quad(np.arange(200000, dtype=np.uint64), 2, 5, 99)
    z = np.concatenate([z0, z1])
    for part in (z.real, z.imag):
        assert abs(part.mean()) < 0.01
        assert abs(part.var() - 1.0) < 0.01
    assert abs(np.mean(z.real * z.imag)) < 0.01
    # the radius uniform stays inside (0, 1]: the deepest tail is sqrt(2 ln 2^33)
    assert np.abs(z).max() <= np.sqrt(2 * 33 * np.log(2)) + 1e-12
