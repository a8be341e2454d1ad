"""Geometry and bead selection of the oracle.

Pinned by the reference's own known-answer tests (tests/golden/
reference_known_answers.json, transcribed from tests/test_moves.py and
tests/test_linalg.py), by the Rodrigues closed form and orthogonality, and by
the selection rules of chromo/util/bead_selection.pyx recomputed from libc
draws, plus the distribution check of tests/test_bead_selection.py.
"""
import ctypes
import json
import math
from pathlib import Path

import numpy as np
import pytest

FX = json.loads((Path(__file__).parent / "golden" / "reference_known_answers.json").read_text())


def _apply(oracle, T, case):
    r = np.asarray(case["r"], dtype=float)
    t3 = np.asarray(case.get("t3", np.zeros_like(r)), dtype=float)
    t2 = np.asarray(case.get("t2", np.zeros_like(r)), dtype=float)
    inds = case["inds"]
    ro, t3o, t2o = oracle.transform(T, r[inds], t3[inds], t2[inds])
    R, T3, T2 = r.copy(), t3.copy(), t2.copy()
    R[inds], T3[inds], T2[inds] = ro, t3o, t2o
    return R, T3, T2


def test_end_pivot_known_answer(oracle):
    c = FX["end_pivot"]
    T = oracle.axis_rotation(c["axis"], c["r"][c["point_index"]], c["angle"])
    r, t3, t2 = _apply(oracle, T, c)
    assert np.allclose(r, c["r_expected"]) and np.allclose(t3, c["t3_expected"])
    assert np.allclose(t2, c["t2_expected"])


def test_slide_known_answer(oracle):
    c = FX["slide"]
    T = oracle.translation(c["delta"])
    r, _, _ = _apply(oracle, T, c)
    assert np.allclose(r, c["r_expected"])


def test_tangent_rotation_known_answer(oracle):
    c = FX["tangent_rotation"]
    T = oracle.axis_rotation(c["axis"], c["point"], c["angle"])
    r, t3, t2 = _apply(oracle, T, c)
    assert np.allclose(r, c["r_expected"]) and np.allclose(t3, c["t3_expected"])
    assert np.allclose(t2, c["t2_expected"])


def test_crank_shaft_known_answer(oracle):
    c = FX["crank_shaft"]
    rr = np.asarray(c["r"], dtype=float)
    axis = rr[c["axis_from"][0]] - rr[c["axis_from"][1]]
    axis = axis / np.linalg.norm(axis)
    T = oracle.axis_rotation(axis, rr[c["point_index"]], c["angle"])
    r, t3, _ = _apply(oracle, T, c)
    assert np.allclose(r, c["r_expected"]) and np.allclose(t3, c["t3_expected"])


def test_z_rotation_and_inverse(oracle):
    c = FX["z_rotation"]
    T = oracle.axis_rotation(c["axis"], c["point"], c["angle"])
    Tinv = np.linalg.inv(oracle.axis_rotation(c["axis"], c["point"], -c["angle"]))
    assert np.allclose(T, c["expected"]) and np.allclose(T, Tinv)


def test_translation_matrix(oracle):
    c = FX["translation"]
    T = oracle.translation(c["delta"])
    assert np.allclose(T, c["expected"])
    assert np.allclose(T @ np.asarray(c["r0"], dtype=float), c["r1_expected"])


@pytest.mark.parametrize("seed", range(20))
def test_rotation_closed_form(oracle, seed):
    rs = np.random.RandomState(seed)
    a = rs.normal(size=3)
    a /= np.linalg.norm(a)
    p = rs.normal(size=3) * 50
    ang = rs.uniform(-np.pi, np.pi)
    T = oracle.axis_rotation(a, p, ang)
    R = T[:3, :3]
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    Rf = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
    assert np.allclose(R, Rf, atol=1e-14)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-13
    # points on the axis through p are fixed: T @ (p + s a) == p + s a
    for s in (-3.0, 0.0, 7.5):
        q = np.append(p + s * a, 1.0)
        assert np.allclose(T @ q, q, atol=1e-10)


def test_translation_leaves_orientation(oracle):
    T = oracle.translation([1.5, -2.0, 0.25])
    r = np.random.RandomState(0).normal(size=(4, 3))
    t3 = np.random.RandomState(1).normal(size=(4, 3))
    ro, t3o, t2o = oracle.transform(T, r, t3, t3)
    assert np.array_equal(t3o, t3) and np.allclose(ro - r, [1.5, -2.0, 0.25])


# ---------------------------------------------------------------------------
# bead selection: chromo/util/bead_selection.pyx
# ---------------------------------------------------------------------------

CBB_CASES = [  # (b0, b1, N) -> (ind0, indf), check_bead_bounds :157-192
    (5, 9, 20, (5, 9)), (9, 5, 20, (5, 9)), (4, 4, 20, (4, 5)), (3, 25, 20, (3, 20)),
    (7, -2, 20, (0, 8)), (25, 3, 20, (3, 20)), (-4, 6, 20, (0, 6)), (20, 21, 20, (20, 20)),
    (0, 0, 1, (0, 1)), (19, 20, 20, (19, 20)),
]


@pytest.mark.parametrize("b0,b1,N,expect", CBB_CASES)
def test_check_bead_bounds(oracle, b0, b1, N, expect):
    assert oracle.check_bead_bounds(b0, b1, N) == expect


@pytest.mark.parametrize("w", [1, 2, 10, 30, 150])
def test_capped_exponential_formula(oracle, w):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(77)
    draws = [libc.rand() for _ in range(20000)]
    got = oracle.select_stream(77, 3, w, w, 0, 200)   # capped_exponential(w, w)
    pos = 0
    exp = []
    for _ in range(200):
        while True:
            u = draws[pos] / 2147483647
            pos += 1
            r = int(-math.log10(u + 0.00001) * w * 0.45 + 1.0001)
            if r <= w:
                exp.append(r)
                break
    assert list(got) == exp
    assert got.min() >= 1 and got.max() <= w


def test_from_left_right_point_rules(oracle):
    N = 500
    left = oracle.select_stream(11, 0, 40, N, 0, 3000)
    right = oracle.select_stream(11, 1, 40, N, 0, 3000)
    assert left.min() >= 1 and left.max() <= 40
    assert np.array_equal(right, N - left)   # same draws, from_right = N - from_left
    pt = oracle.select_stream(12, 2, 40, N, 250, 3000)
    assert pt.min() >= 250 - 40 and pt.max() <= 250 + 40
    # window < 1 returns ind0 without drawing
    assert np.all(oracle.select_stream(12, 2, 0, N, 33, 10) == 33)
    with pytest.raises(ValueError):
        oracle.select_stream(1, 0, N + 1, N, 0, 1)
    with pytest.raises(ValueError):
        oracle.select_stream(1, 2, N + 1, N, 0, 1)


def test_from_left_distribution(oracle):
    """Shape check after tests/test_bead_selection.py: P(r) follows the capped
    exponential density; compare the empirical histogram with the exact pmf."""
    w = 200
    x = oracle.select_stream(5, 0, w, 10000, 0, 40000)
    # exact pmf of r = floor(-log10(U + 1e-5) * 0.45 w + 1.0001), U ~ U[0,1], capped by rejection
    k = np.arange(1, w + 1)
    lo = np.clip(10.0 ** (-(k + 1 - 1.0001) / (0.45 * w)) - 1e-5, 0, 1)
    hi = np.clip(10.0 ** (-(k - 1.0001) / (0.45 * w)) - 1e-5, 0, 1)
    p = hi - lo
    p /= p.sum()
    obs = np.bincount(x, minlength=w + 1)[1:]
    chi2 = np.sum((obs - 40000 * p) ** 2 / np.maximum(40000 * p, 1e-12))
    assert chi2 < 300  # ~200 dof
    assert obs[0] > obs[-1]
