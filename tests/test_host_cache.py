"""CPU tests of the host block cache behind the library's host CSR arrays
(igc_host.cpp: host_alloc / host_free / host_cache_trim), compiled with g++ on
its own: blocks of 4 MiB and more are reused for requests of at least half their
size, smaller ones are plain heap memory, and a trim releases the cached blocks."""
import ctypes
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def hc(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("hc") / "host_cache.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"), "-o", so,
                    os.path.join(HERE, "host_cache_harness.cpp"),
                    os.path.join(ROOT, "indexgen_amd", "csrc", "igc_host.cpp"), "-lpthread"], check=True)
    L = ctypes.CDLL(so)
    L.hc_alloc.restype = ctypes.c_void_p
    L.hc_alloc.argtypes = [ctypes.c_size_t]
    L.hc_free.argtypes = [ctypes.c_void_p]
    return L


MiB = 1 << 20


def test_large_blocks_are_reused(hc):
    hc.hc_trim()
    a = hc.hc_alloc(8 * MiB)
    assert a and a % (2 * MiB) == 0          # 2 MiB aligned
    ctypes.memset(a, 7, 8 * MiB)             # the block is usable in full
    hc.hc_free(a)
    b = hc.hc_alloc(6 * MiB)                 # at least half of the cached block: reused
    assert b == a
    c = hc.hc_alloc(6 * MiB)                 # the cache is empty again: a new block
    assert c and c != a
    hc.hc_free(b)
    hc.hc_free(c)
    d = hc.hc_alloc(3 * MiB + 1)             # under 4 MiB: heap memory, never a cached block
    assert d not in (a, c)
    hc.hc_free(d)
    e = hc.hc_alloc(2 * MiB)                 # small request: never a cached block
    assert e not in (a, c)
    hc.hc_free(e)
    hc.hc_trim()


def test_blocks_too_large_for_a_request_are_not_lent(hc):
    hc.hc_trim()
    big = hc.hc_alloc(64 * MiB)
    hc.hc_free(big)
    small = hc.hc_alloc(8 * MiB)             # the 64 MiB block is more than twice the request
    assert small != big
    again = hc.hc_alloc(40 * MiB)            # 64 MiB <= 2 x 40 MiB: reused
    assert again == big
    hc.hc_free(small)
    hc.hc_free(again)
    hc.hc_trim()
    fresh = hc.hc_alloc(40 * MiB)            # after a trim nothing is cached
    assert fresh
    hc.hc_free(fresh)
    hc.hc_trim()


def test_free_of_null_and_many_cycles(hc):
    hc.hc_free(None)
    hc.hc_trim()
    seen = set()
    for _ in range(20):
        p = hc.hc_alloc(12 * MiB)
        seen.add(p)
        hc.hc_free(p)
    assert len(seen) == 1                    # the same faulted-in block every time
    hc.hc_trim()
