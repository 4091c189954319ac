"""The oracle's candidate generators (LinearCode, VTCode, DifferentialVTCode,
AllStrings, each followed by applyFilters) against the lists the compiled
reference's own generators wrote (tests/golden/gen_golden.npz), including for
LinearCode the unfiltered list (enumeration order and GF(4) arithmetic)."""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "gen_golden.npz"))


def digest(syms: np.ndarray) -> str:
    rows = (np.ascontiguousarray(syms, np.uint8) + 48).tobytes()
    n = syms.shape[1] if syms.ndim == 2 else 0
    words = [rows[i * n:(i + 1) * n].decode() for i in range(syms.shape[0])]
    return hashlib.sha256("\n".join(words).encode()).hexdigest()


def cases(prefix):
    return sorted({k.split("_")[0] for k in GOLD.files if k.startswith(prefix) and k[len(prefix)].isdigit()})


def check_kept(key, got):
    assert got.shape[0] == int(GOLD[f"{key}_count"]), key
    if f"{key}_syms" in GOLD.files:
        assert np.array_equal(got, GOLD[f"{key}_syms"]), key
    assert digest(got) == GOLD[f"{key}_digest"].item().decode(), key


def filt(key):
    mr, lo, hi = GOLD[f"{key}_filter"]
    return int(mr), float(lo), float(hi)


def test_fixture_cases_present():
    assert len(cases("linear")) == 9 and len(cases("vt")) == 4 and len(cases("dvt")) == 4 and len(cases("all")) == 3


@pytest.mark.parametrize("key", cases("linear"))
def test_linear_matches_reference(key):
    G, b = GOLD[f"{key}_G"], GOLD[f"{key}_bias"]
    full = O.gen_linear(G, b)
    assert full.shape[0] == 4 ** G.shape[0]
    assert digest(full) == GOLD[f"{key}_full_digest"].item().decode()
    check_kept(key, O.gen_linear(G, b, *filt(key)))


@pytest.mark.parametrize("key", cases("vt"))
def test_vt_matches_reference(key):
    n, a, b = (int(x) for x in GOLD[f"{key}_params"])
    check_kept(key, O.gen_vt(n, a, b, *filt(key)))


@pytest.mark.parametrize("key", cases("dvt"))
def test_dvt_matches_reference(key):
    n, s = (int(x) for x in GOLD[f"{key}_params"])
    check_kept(key, O.gen_dvt(n, s, *filt(key)))


@pytest.mark.parametrize("key", cases("all"))
def test_all_strings_match_reference(key):
    (n,) = (int(x) for x in GOLD[f"{key}_params"])
    check_kept(key, O.gen_all(n, *filt(key)))


def test_linear_code_properties():
    """Closed forms the fixtures do not cover: the code is linear (the sum of
    two codewords minus the bias is a codeword minus the bias) and its size is
    4^k; a zero generator with bias gives one word repeated."""
    key = cases("linear")[2]
    G, b = GOLD[f"{key}_G"], GOLD[f"{key}_bias"]
    full = O.gen_linear(G, b)
    words = {bytes(r ^ b) for r in full}
    rng = np.random.default_rng(3)
    for _ in range(50):
        x, y = full[rng.integers(0, len(full), 2)]
        assert bytes(x ^ y) in words  # (x-b) + (y-b) = x ^ y ^ b ^ b
    z = O.gen_linear(np.zeros((2, 5), np.uint8), np.array([1, 2, 3, 0, 1], np.uint8))
    assert z.shape == (16, 5) and (z == [1, 2, 3, 0, 1]).all()
