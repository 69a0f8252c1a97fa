"""Random-number generators on the path.

* Philox4x32-10 (per-insertion streams of the batched Widom kernel): the
  oracle's restatement against the published known-answer vectors of
  Salmon et al., SC'11 (Random123 kat_vectors, philox4x32 R=10).
* glibc rand() (the reference's generator, data_struct.cpp:10-15): the
  engine's restatement against the C library itself.
"""
import ctypes as C

import numpy as np
import pytest

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat(ctr, key, expect, oracle_lib):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    oracle_lib.or_philox4x32_10(c, k, o)
    assert tuple(o) == expect


def test_philox_uniform_range_and_moments(oracle_lib):
    u = np.array([oracle_lib.or_philox_uniform(12345, 7, i, d) for i in range(2000) for d in range(8)])
    assert np.all((u > 0.0) & (u < 1.0))
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1.0 / 12.0) < 0.005
    # distinct streams / indices differ
    assert oracle_lib.or_philox_uniform(1, 0, 0, 0) != oracle_lib.or_philox_uniform(1, 1, 0, 0)
    assert oracle_lib.or_philox_uniform(1, 0, 0, 0) != oracle_lib.or_philox_uniform(1, 0, 1, 0)


@pytest.mark.parametrize("seed", [0, 1, 7, 12345, 2 ** 31 - 1, 4000000000])
def test_glibc_rand_restatement(seed):
    from graspa_amd import engine
    libc = C.CDLL("libc.so.6")
    libc.srand.argtypes = [C.c_uint]
    n = 5000
    out = (C.c_int * n)()
    assert engine.lib().graspa_glibc_rand_sequence(seed, n, out) == 0
    libc.srand(seed)
    ref = [libc.rand() for _ in range(n)]
    assert list(out) == ref
