"""Pins the CPU oracle (oracle/rpy_oracle.c, oracle/rpy_scalar.py).

Against: the reference's golden fixtures (tests/ref/*.npz of the reference,
copied to tests/golden), closed forms of the self mobilities, continuity of
the RPY branches at r = 2a and of the wall damping at z = a, symmetry and
positive definiteness, and the second (vector-form) restatement.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import rpy_scalar
from tests.conftest import GOLDEN


def test_freespace_pair_golden():
    ref = np.load(os.path.join(GOLDEN, "pair_mobility_nbody_freespace.npz"))
    for i in range(len(ref["r_vecs"])):
        a = ref["a"].ravel()[i]
        eta = ref["eta"].ravel()[i]
        M = oracle.dense_matrix(ref["r_vecs"][i], eta, a, angular=True)
        assert np.abs(M - ref["M"][i]).max() <= 1e-14 * np.abs(ref["M"][i]).max()


def _wall_norm(nP, eta):
    n = 3 * nP
    norm = np.zeros((2 * n, 2 * n))
    norm[:n, :n] = 1 / (6 * np.pi * eta)
    norm[n:, n:] = 1 / (8 * np.pi * eta)
    norm[n:, :n] = 1 / (6 * np.pi * eta)
    norm[:n, n:] = 1 / (6 * np.pi * eta)
    return norm


def test_wall_self_golden():
    ref = np.load(os.path.join(GOLDEN, "self_mobility_bw_ref_noimg.npz"))
    eta = 1 / 4 / np.sqrt(np.pi)
    norm = _wall_norm(1, eta)
    for k, h in enumerate(ref["heights"].ravel()):
        M = oracle.dense_matrix([38.4, 38.4, h], eta, 1.0, angular=True, wall=True) / norm
        assert np.allclose(M, ref["M"][k], atol=1e-15, rtol=1e-14)


@pytest.mark.parametrize("offset", ["x", "y"])
def test_wall_pair_golden(offset):
    ref = np.load(os.path.join(GOLDEN, f"pair_mobility_bw_ref_noimg_offset_{offset}.npz"))
    eta = 1 / 4 / np.sqrt(np.pi)
    norm = _wall_norm(2, eta)
    ov = np.array([1.0, 0, 0]) if offset == "x" else np.array([0, 1.0, 0])
    for i, sep in enumerate([3.0, 4.0, 8.0]):
        for k, h in enumerate(np.asarray(ref["heights"]).ravel()):
            c = np.array([38.4, 38.4, h])
            M = oracle.dense_matrix(np.array([c + sep * ov / 2, c - sep * ov / 2]), eta, 1.0,
                                    angular=True, wall=True) / norm
            assert np.allclose(M, ref["M"][i, k], atol=1e-15, rtol=1e-14)


def test_self_closed_forms():
    a, eta = 0.9123, 1.123
    M = oracle.dense_matrix([0.0, 0.0, 0.0], eta, a, angular=True)
    assert np.allclose(M[:3, :3], np.eye(3) / (6 * np.pi * eta * a), rtol=1e-15, atol=0)
    assert np.allclose(M[3:, 3:], np.eye(3) / (8 * np.pi * eta * a ** 3), rtol=1e-15, atol=0)
    assert np.all(M[:3, 3:] == 0) and np.all(M[3:, :3] == 0)


def test_wall_self_swan_brady_closed_form():
    """Swan & Brady self mobility near a wall, written in normalised form."""
    for h in [1.0, 1.7, 3.3, 9.6]:
        M = oracle.dense_matrix([0.0, 0.0, h], 1.0, 1.0, angular=True, wall=True)
        t0 = 1 / (6 * np.pi)
        r0 = 1 / (8 * np.pi)
        par = 1 - 9 / (16 * h) + 1 / (8 * h ** 3) - 1 / (16 * h ** 5)
        per = 1 - 9 / (8 * h) + 1 / (2 * h ** 3) - 1 / (8 * h ** 5)
        assert np.isclose(M[0, 0] / t0, par, rtol=1e-14)
        assert np.isclose(M[2, 2] / t0, per, rtol=1e-14)
        assert np.isclose(M[3, 3] / r0, 1 - 5 / (16 * h ** 3), rtol=1e-14)
        assert np.isclose(M[5, 5] / r0, 1 - 1 / (8 * h ** 3), rtol=1e-14)
        # translation-rotation coupling 3/(32 h^4) in units of 1/(6 pi eta a^2)
        assert np.isclose(abs(M[0, 4]) / t0, 3 / (32 * h ** 4), rtol=1e-13)


def test_rpy_textbook_form_and_continuity():
    """Open UF block against the Rotne-Prager-Yamakawa tensor written in the
    rr^T form of the literature (both branches), continuity at r = 2a."""
    a = 1.0
    for r in [0.0, 0.3, 1.0, 1.999999, 2.0, 2.000001, 3.5, 20.0]:
        d = np.array([0.3, -0.5, 0.81])
        d = d / np.linalg.norm(d) * r if r > 0 else np.zeros(3)
        B = oracle.pair_block(0, d, np.zeros(3), a)[:3, :3]
        rh = d / r if r > 0 else np.zeros(3)
        if r >= 2 * a:
            ref = (0.75 * a / r * (1 + 2 * a ** 2 / (3 * r ** 2))) * np.eye(3) + \
                  (0.75 * a / r * (1 - 2 * a ** 2 / r ** 2)) * np.outer(rh, rh)
        else:
            ref = (1 - 9 * r / (32 * a)) * np.eye(3) + (3 * r / (32 * a)) * np.outer(rh, rh)
        assert np.allclose(B, ref, rtol=1e-14, atol=1e-15)
    lo = oracle.pair_block(0, [2 * a - 1e-12, 0, 0], [0, 0, 0], a)
    hi = oracle.pair_block(0, [2 * a + 1e-12, 0, 0], [0, 0, 0], a)
    assert np.abs(lo - hi).max() < 1e-10


def test_rotational_rpy_limits():
    """WT and UT/WF continuity at r = 2a; r -> 0 limits equal the self terms."""
    a = 1.0
    d = np.array([0.6, 0.0, 0.8])
    lo = oracle.pair_block(0, d * (2 - 1e-12), [0, 0, 0], a)
    hi = oracle.pair_block(0, d * (2 + 1e-12), [0, 0, 0], a)
    assert np.abs(lo - hi).max() < 1e-10
    tiny = oracle.pair_block(0, d * 1e-9, [0, 0, 0], a)
    self_ = oracle.pair_block(0, [0, 0, 0], [0, 0, 0], a)
    assert np.abs(tiny - self_).max() < 1e-8


def test_wall_damping_continuity_and_limit():
    a = 1.0
    pj = np.array([2.0, 0.3, 3.0])
    lo = oracle.pair_block(1, [0, 0, a - 1e-12], pj, a)
    hi = oracle.pair_block(1, [0, 0, a + 1e-12], pj, a)
    assert np.abs(lo - hi).max() < 1e-10
    z0 = oracle.pair_block(1, [0, 0, 0.0], pj, a)
    assert np.abs(z0).max() == 0.0
    half = oracle.pair_block(1, [0, 0, 0.5], pj, a)
    assert np.allclose(half, 0.5 * oracle.pair_block(1, [0, 0, a], pj, a), rtol=1e-14)


@pytest.mark.parametrize("wall", [False, True])
def test_symmetric_positive_definite(wall):
    rng = np.random.default_rng(1)
    for trial in range(20):
        n = 10
        pos = rng.random((n, 3)) * 10 - np.array([5, 5, 0])  # tests/utils.py:87-103
        if wall:
            pos[:, 2] = rng.random(n) * 10
        M = oracle.dense_matrix(pos, 1.0, 1.0, angular=True, wall=wall)
        assert np.abs(M - M.T).max() < 1e-13 * np.abs(M).max()
        assert np.all(np.linalg.eigvalsh(M) > 0)


def test_scalar_restatement_agrees():
    rng = np.random.default_rng(2)
    for wall in (0, 1):
        for t in range(400):
            pi = rng.uniform(-3, 3, 3)
            pj = rng.uniform(-3, 3, 3)
            if wall:
                pi[2] = rng.uniform(0.0, 5)
                pj[2] = rng.uniform(0.0, 5)
            if t % 20 == 0:
                pj = pi.copy()
            F = rng.normal(size=3)
            T = rng.normal(size=3)
            B = oracle.pair_block(wall, pi, pj, 1.0)
            res = rpy_scalar.pair_apply(wall, pi, pj, F, T, {"UF", "UT", "WT", "WF"})
            pairs = {"UF": B[:3, :3] @ F, "UT": B[:3, 3:] @ T, "WF": B[3:, :3] @ F, "WT": B[3:, 3:] @ T}
            for k, ref in pairs.items():
                assert np.abs(np.array(res[k]) - ref).max() <= 1e-13 * max(1, np.abs(ref).max())


def test_matrix_free_equals_dense_with_batches():
    rng = np.random.default_rng(3)
    sizes = [7, 1, 20]
    pos = rng.uniform(0, 6, (28, 3))
    pos[:, 2] += 0.5
    F = rng.normal(size=84)
    T = rng.normal(size=84)
    for wall in (False, True):
        M = oracle.dense_matrix(pos, 0.8, 1.1, angular=True, wall=wall, sys_sizes=sizes)
        U, W = oracle.mdot(pos, 0.8, 1.1, F=F, T=T, angular=True, wall=wall, sys_sizes=sizes)
        ref = M @ np.concatenate([F, T])
        assert np.allclose(np.concatenate([U, W]), ref, rtol=1e-13, atol=1e-15)
        # block-diagonal over systems
        assert np.all(M[:21, 21:24] == 0)
    Uo = oracle.mdot_open_uf(pos, 0.8, 1.1, F)
    U, _ = oracle.mdot(pos, 0.8, 1.1, F=F)
    assert np.allclose(Uo, U, rtol=1e-14, atol=1e-16)


def test_row_sampled_product_equals_full():
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 7, (60, 3))
    pos[:, 2] += 0.4
    F = rng.normal(size=180)
    T = rng.normal(size=180)
    U, W = oracle.mdot(pos, 1.0, 1.0, F=F, T=T, angular=True, wall=True)
    rows = np.array([3, 17, 59, 0])
    Ur, Wr = oracle.mdot_rows(pos, 1.0, 1.0, rows, F=F, T=T, angular=True, wall=True)
    assert np.array_equal(Ur, U.reshape(-1, 3)[rows]) and np.array_equal(Wr, W.reshape(-1, 3)[rows])
