"""Random streams of the oracle (and the host seeding helper) against the real
libraries: glibc rand() (the stream behind every proposal draw and the
Metropolis draw, chromo/mc/mc_sim.pyx:171), numpy's legacy RandomState.randint
(binding states, chromo/mc/move_funcs.pyx:819) and Philox4x32-10 known answers."""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def libc():
    return ctypes.CDLL("libc.so.6")


@pytest.mark.parametrize("seed", [1, 2, 7, 12345, 0, 2 ** 31 - 1, 2 ** 31 + 5, 2 ** 32 - 1])
def test_glibc_stream_matches_libc(oracle, libc, seed):
    libc.srand(ctypes.c_uint(seed))
    ref = np.array([libc.rand() for _ in range(2000)], dtype=np.int32)
    assert np.array_equal(oracle.glibc_stream(seed, 2000), ref)


def test_unseeded_rand_is_srand1(oracle, libc):
    # the reference never seeds libc: its stream is srand(1)'s
    libc.srand(1)
    ref = [libc.rand() for _ in range(50)]
    assert list(oracle.glibc_stream(1, 50)) == ref


@pytest.mark.parametrize("seed", [0, 1, 987654321, 2 ** 32 - 1])
@pytest.mark.parametrize("high", [1, 2, 3, 4, 5, 17])
def test_mt_randint_matches_numpy(oracle, seed, high):
    rs = np.random.RandomState(seed)
    ref = np.array([rs.randint(0, high) for _ in range(1500)])
    assert np.array_equal(oracle.mt_randint_stream(seed, high, 1500), ref)


def test_host_srand_words_match_libc(libc):
    from chromo_amd.mc.mc_sim import _srand_words
    for seed in (1, 3, 99, 2 ** 31 + 7):
        w = _srand_words(seed)
        st, f, r = [int(x) for x in w[:31]], int(w[31]), int(w[32])
        out = []
        for _ in range(200):
            v = (st[f] + st[r]) & 0xffffffff
            st[f] = v
            out.append(v >> 1)
            f += 1
            if f >= 31:
                f = 0
                r += 1
            else:
                r += 1
                if r >= 31:
                    r = 0
        libc.srand(ctypes.c_uint(seed))
        assert out == [libc.rand() for _ in range(200)]


# Known-answer vectors of Philox4x32-10 (Random123 kat_vectors).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    assert [int(x) for x in oracle.philox4x32_10(ctr, key)] == expect


def test_philox_counter_sensitivity(oracle):
    a = oracle.philox4x32_10([1, 0, 0, 0], [5, 6])
    b = oracle.philox4x32_10([2, 0, 0, 0], [5, 6])
    c = oracle.philox4x32_10([1, 0, 0, 0], [5, 7])
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
