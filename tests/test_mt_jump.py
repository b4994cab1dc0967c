"""CPU tests of the mt19937 jump-ahead that splits the shared camera stream
(UniformSample, CameraSampler.h:19-27) over workgroups: the host routine
behind sppm_mt19937_jump must land on the same words as sequential
generation.  The sequential generator here is pinned by the C++ standard's
known answer (std::mt19937 default-seeded, 10000th output = 4123659995)."""
import numpy as np
import pytest

import sppm_amd

UPPER, LOWER, MATRIX_A = np.uint32(0x80000000), np.uint32(0x7FFFFFFF), np.uint32(0x9908B0DF)


def seed_window(seed):
    w = np.zeros(624, np.uint64)
    w[0] = seed
    for i in range(1, 624):
        w[i] = (1812433253 * (int(w[i - 1]) ^ (int(w[i - 1]) >> 30)) + i) & 0xFFFFFFFF
    return w.astype(np.uint32)


def raw_words(window, n):
    """x[0 .. n) with x[0..623] = window and x[m] = x[m-227] ^ mix(x[m-624], x[m-623]),
    227 words at a time (each block depends on earlier words only)."""
    x = np.zeros(max(n, 624), np.uint32)
    x[:624] = window
    m = 624
    while m < n:
        k = min(227, n - m)
        a, b, c = x[m - 624:m - 624 + k], x[m - 623:m - 623 + k], x[m - 227:m - 227 + k]
        y = (a & UPPER) | (b & LOWER)
        x[m:m + k] = c ^ (y >> np.uint32(1)) ^ np.where(y & np.uint32(1), MATRIX_A, np.uint32(0))
        m += k
    return x[:n]


def temper(z):
    z = z ^ (z >> np.uint32(11))
    z = z ^ ((z << np.uint32(7)) & np.uint32(0x9D2C5680))
    z = z ^ ((z << np.uint32(15)) & np.uint32(0xEFC60000))
    return z ^ (z >> np.uint32(18))


def test_sequential_generator_known_answer():
    x = raw_words(seed_window(5489), 624 + 10000)
    assert int(temper(x[624 + 9999])) == 4123659995


@pytest.mark.parametrize("steps", [0, 1, 226, 624, 1000, 19937, 98304, 6 * 64 * 64 * 63 // 64])
@pytest.mark.parametrize("seed", [0, 5489])
def test_jump_matches_sequential(steps, seed):
    lib = sppm_amd.load_library()
    w0 = seed_window(seed)
    seq = raw_words(w0, steps + 624 + 800)
    got = sppm_amd.mt19937_jump(w0, steps, lib)
    # the window itself: words 1..623 and the top bit of word 0 (the only part of word 0 later read)
    np.testing.assert_array_equal(got[1:], seq[steps + 1:steps + 624])
    assert (int(got[0]) >> 31) == (int(seq[steps]) >> 31)
    # the outputs that follow
    np.testing.assert_array_equal(temper(raw_words(got, 624 + 800)[624:]), temper(seq[steps + 624:steps + 1424]))


def test_jump_known_answer_and_composition():
    lib = sppm_amd.load_library()
    w = sppm_amd.mt19937_jump(seed_window(5489), 9999, lib)
    assert int(temper(raw_words(w, 625)[624])) == 4123659995
    # jumps compose: 4000 + 5999 = 9999
    w2 = sppm_amd.mt19937_jump(sppm_amd.mt19937_jump(seed_window(5489), 4000, lib), 5999, lib)
    assert int(temper(raw_words(w2, 625)[624])) == 4123659995
