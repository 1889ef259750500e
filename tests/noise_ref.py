"""CPU restatement of the device noise generator, for the tests only.

Philox4x32-10 is the counter-based generator of Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3" (SC'11), as published with
Random123: ten rounds of two 32x32->64 multiplies with the constants
M0 = 0xD2511F53, M1 = 0xCD9E8D57 and the Weyl key bumps W0 = 0x9E3779B9,
W1 = 0xBB67AE85.  Its known-answer vectors are checked in
tests/test_noise_ref.py.

The device draws two complex deviates per call: counter (mode, stream << 4,
step_lo, step_hi), key (seed_lo, seed_hi ^ 0x5348), and a Box-Muller pair from
each two words, u1 = (w + 1/2) 2^-32, u2 = w' 2^-32, r = sqrt(-2 ln u1),
z = r (cos 2 pi u2, sin 2 pi u2).  Here everything after the integer rounds is
evaluated in float64 as the reference value for the device's single-precision
evaluation.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = 0x9E3779B9
W1 = 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorized Philox4x32-10; counters are arrays (any broadcastable
    shapes), keys are ints.  Returns the four output words as uint64 arrays
    holding 32-bit values."""
    c = [np.asarray(v, dtype=np.uint64) & MASK for v in (c0, c1, c2, c3)]
    c = np.broadcast_arrays(*c)
    c = [v.copy() for v in c]
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c


def box_muller(a, b):
    u1 = (a.astype(np.float64) + 0.5) * 2.0 ** -32
    u2 = b.astype(np.float64) * 2.0 ** -32
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)


def gauss_quad(mode_id, stream, step, seed):
    """The two complex deviates the device draws for (mode, stream, step)."""
    k0 = seed & 0xFFFFFFFF
    k1 = ((seed >> 32) & 0xFFFFFFFF) ^ 0x5348
    w = philox4x32_10(mode_id, np.uint64(stream) << np.uint64(4), step & 0xFFFFFFFF, step >> 32, k0, k1)
    return box_muller(w[0], w[1]), box_muller(w[2], w[3])


def mode_table(cfg):
    """Per reference-layout slot [gy, gx]: the device's noise mode id, whether
    the draw is conjugated (kx = 0, ky < 0 takes its ky > 0 partner's draw) and
    whether the slot is an active mode that draws at all."""
    ny, hx = cfg.ny, cfg.nx // 2 + 1
    nay, hxa = cfg.ny_active, cfg.nx_active // 2   # active kx: gx < Nax/2
    half = nay // 2
    gy, gx = np.meshgrid(np.arange(ny), np.arange(hx), indexing="ij")
    ky = np.where(gy <= ny // 2, gy, gy - ny)
    live = (gx < hxa) & (ky < half) & (-ky < half) & ((gx != 0) | (ky != 0))
    cidx = np.where(ky >= 0, ky, ky + nay)
    conj = (gx == 0) & (ky < 0)
    mode = gx * nay + np.where(conj, -ky, cidx)
    return np.where(live, mode, 0).astype(np.uint64), conj, live
