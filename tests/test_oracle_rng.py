"""Pins the oracle's restatement of the reference RNG (include/dmhm/core/random.hpp,
src/core/random.cpp) against independent closed forms."""
import math

import numpy as np

import oracle as orc

A = 6364136223846793005
C = 1442695040888963407
M = 1 << 64


def limb_mul(a, b):
    """MultiplyWith64BitMod as described at random.hpp:146-216: 16-bit limbs,
    only the four lowest partial-product rows kept, then carries."""
    al = [(a >> (16 * i)) & 0xFFFF for i in range(4)]
    bl = [(b >> (16 * i)) & 0xFFFF for i in range(4)]
    c = [0, 0, 0, 0]
    for i in range(4):
        for j in range(4 - i):
            t = al[i] * bl[j]
            c[i + j] += t & 0xFFFF
            if i + j + 1 < 4:
                c[i + j + 1] += t >> 16
    for i in range(3):                      # CarryUpper16Bits
        c[i + 1] += c[i] >> 16
        c[i] &= 0xFFFF
    c[3] &= 0xFFFF
    return sum(c[i] << (16 * i) for i in range(4))


def test_limb_arithmetic_equals_native_modular_product():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        b = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        assert limb_mul(a, b) == (a * b) % M


def test_serial_lcg_stream_closed_form():
    orc.rng_initialize()
    got = orc.rng_serial_raw(64)
    x = 17 + (21 << 32)                    # Initialize seed {17, 21} (global.cpp:86-93)
    for i in range(64):
        assert int(got[i]) == x            # SerialLcg returns the state, then advances
        x = (limb_mul(A, x) + C) % M


def test_uniform_is_upper_word_plus_one():
    orc.rng_initialize()
    u = orc.rng_serial_uniform(32)
    x = 17 + (21 << 32)
    for i in range(32):
        assert u[i] == ((x >> 32) + 1.0) / 4294967296.0
        assert 0.0 < u[i] <= 1.0
        x = (A * x + C) % M


def _uniforms(n):
    x = 17 + (21 << 32)
    out = []
    for _ in range(n):
        out.append(((x >> 32) + 1.0) / 4294967296.0)
        x = (A * x + C) % M
    return out


def test_gaussian_vector_box_muller_pairs():
    orc.rng_initialize()
    g = orc.gaussian_vector(7)                # odd length: 3 pairs + one half Box-Muller
    u = _uniforms(8)
    exp = []
    for p in range(3):
        U, V = u[2 * p], u[2 * p + 1]
        r = math.sqrt(-2 * math.log(U))
        exp += [r * math.cos(2 * math.pi * V), r * math.sin(2 * math.pi * V)]
    exp.append(math.sqrt(-2 * math.log(u[6])) * math.cos(2 * math.pi * u[7]))
    np.testing.assert_allclose(g, exp, rtol=0, atol=1e-14)


def test_gaussian_vectors_row_quirk():
    """SerialGaussianRandomVectors writes the last pair of each column at rows
    n-2, n-1 (n = number of columns, random.hpp:441,443)."""
    orc.set_options(ref_quirk=True)
    orc.rng_initialize()
    G = orc.gaussian_vectors(10, 3)
    assert np.all(G[8:, :] == 0.0)            # rows m-2, m-1 never written
    orc.set_options(ref_quirk=False)
    orc.rng_initialize()
    H = orc.gaussian_vectors(10, 3)
    assert np.all(H[8:, :] != 0.0)
    # same stream consumption per column: the quirk only moves the last pair
    for j in range(3):
        keep = [0] + list(range(3, 8))
        np.testing.assert_array_equal(G[keep, j], H[keep, j])
        np.testing.assert_array_equal(G[1:3, j], H[8:10, j])
    orc.set_options()
