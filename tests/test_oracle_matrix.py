"""Known-answer checks that pin the oracle's single-matrix operations.

Every check compares against an independent formulation: explicit 2x2/3x3
complex matrices built from Pauli/Gell-Mann generators, scipy's expm/logm,
or the published Philox-4x32-10 test vectors.  None of them re-types the
oracle's own formula.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import _oracle as O

RNG = np.random.default_rng(12345)


def rand_link(nc, scale=1.0):
    a = RNG.normal(0, scale, O.ne(nc))
    return O.from_matrix(nc, sla.expm(1j * O.alg_matrix(nc, a)))


@pytest.mark.parametrize("nc", [2, 3])
def test_products_match_matrix_algebra(nc):
    for _ in range(20):
        A, B = rand_link(nc), rand_link(nc)
        MA, MB = O.to_matrix(nc, A), O.to_matrix(nc, B)
        np.testing.assert_allclose(O.to_matrix(nc, O.mul(nc, A, B)), MA @ MB, atol=1e-14)
        np.testing.assert_allclose(O.to_matrix(nc, O.mul(nc, A, B, "nd")), MA @ MB.conj().T, atol=1e-14)
        np.testing.assert_allclose(O.to_matrix(nc, O.mul(nc, A, B, "dn")), MA.conj().T @ MB, atol=1e-14)


@pytest.mark.parametrize("nc", [2, 3])
def test_projection_is_retr_i_ta_M(nc):
    # SU(2) projections are defined on the quaternion span; SU(3) on any matrix
    for _ in range(10):
        if nc == 2:
            M = RNG.normal(size=4)
        else:
            M = RNG.normal(size=18)
        MM = O.to_matrix(nc, M)
        want = [np.real(np.trace(1j * t @ MM)) * 1.7 for t in O.generators(nc)]
        np.testing.assert_allclose(O.proj(nc, 1.7, M), want, atol=1e-14)


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("scale", [1e-9, 1e-3, 0.3, 2.0, 7.0])
def test_exp_matches_scipy_expm(nc, scale):
    for _ in range(25):
        a = RNG.normal(0, scale, O.ne(nc))
        c = RNG.uniform(-1.5, 1.5)
        want = sla.expm(1j * c * O.alg_matrix(nc, a))
        got = O.to_matrix(nc, O.expi(nc, c, a))
        np.testing.assert_allclose(got, want, atol=2e-14 * max(1.0, scale))


def test_su3_exp_degenerate_spectra():
    # alpha along lambda_8 or lambda_3 alone gives doubly degenerate or
    # symmetric spectra: the Jacobi fallback must handle them
    cases = [np.eye(8)[7] * 0.9, np.eye(8)[2] * 1.3, np.zeros(8), np.eye(8)[0] * 1e-12,
             np.array([0, 0, 0.5, 0, 0, 0, 0, 0.5 * np.sqrt(3)])]
    for a in cases:
        want = sla.expm(1j * O.alg_matrix(3, a))
        got = O.to_matrix(3, O.expi(3, 1.0, a))
        np.testing.assert_allclose(got, want, atol=3e-15)


def test_su3_exp_degenerate_spectra_all_angles():
    # doubly degenerate spectra at every angle: along lambda_8, along
    # (lambda_3 +- sqrt(3) lambda_8)/2 and along random conjugates V lambda_8 V^+
    # (a Newton step on the double root of the characteristic cubic once threw
    # the eigenvalue off here)
    t = O.generators(3)
    rng = np.random.default_rng(13)
    dirs = [np.eye(8)[7], np.array([0, 0, 0.5, 0, 0, 0, 0, 0.5 * np.sqrt(3)]),
            np.array([0, 0, 0.5, 0, 0, 0, 0, -0.5 * np.sqrt(3)])]
    for _ in range(4):
        V = sla.expm(1j * O.alg_matrix(3, rng.normal(0, 1.0, 8)))
        L = V @ (2 * t[7]) @ V.conj().T
        dirs.append(np.array([2 * np.real(np.trace(ta @ L)) for ta in t]))
    for d in dirs:
        for th in np.concatenate([np.linspace(-3, 3, 601), rng.normal(0, 1e-3, 50)]):
            want = sla.expm(1j * O.alg_matrix(3, th * d))
            got = O.to_matrix(3, O.expi(3, 1.0, th * d))
            np.testing.assert_allclose(got, want, atol=2e-14)


@pytest.mark.parametrize("nc", [2, 3])
def test_log_inverts_exp_and_matches_logm(nc):
    # a generator of its own, so the draws do not depend on which tests ran
    # before in the process; the log inverts exp only on the principal
    # branch, so draws with an eigenvalue of a.t beyond pi - 0.05 are skipped
    rng = np.random.default_rng(4000 + nc)
    for scale in [1e-6, 0.05, 0.5, 1.5]:
        n = 0
        while n < 20:
            a = rng.normal(0, scale, O.ne(nc))
            if np.max(np.abs(np.linalg.eigvalsh(O.alg_matrix(nc, a)))) > np.pi - 0.05:
                continue
            n += 1
            U = O.expi(nc, 1.0, a)
            # the hot path calls the log with c = -1 (SUNc/AdvancedOperations.cpp:437)
            b = O.logi(nc, -1.0, U)
            np.testing.assert_allclose(-b, a, atol=1e-13)
            H = -1j * sla.logm(O.to_matrix(nc, U))
            want = [2 * np.real(np.trace(t @ H)) for t in O.generators(nc)]
            np.testing.assert_allclose(-b, want, atol=1e-12)


def test_su3_log_large_angles_principal_branch():
    # eigen-phases anywhere in (-pi, pi] exercise the phase-shifted Cayley
    # transform.  The reference takes the principal arg of each eigenvalue and
    # projects onto the generators (SUNc/SU3/AlgebraOperations.cpp:213-230), so
    # when the principal phases sum to +-2 pi, exp(log U) = U times a cube
    # root of unity; scipy's principal logm gives the same projection.
    for _ in range(200):
        a = RNG.normal(0, 2.0, 8)
        U = O.expi(3, 1.0, a)
        b = O.logi(3, 1.0, U)
        H = -1j * sla.logm(O.to_matrix(3, U))
        want = [2 * np.real(np.trace(t @ H)) for t in O.generators(3)]
        np.testing.assert_allclose(b, want, atol=1e-11)
        back = O.to_matrix(3, O.expi(3, 1.0, b)) @ O.to_matrix(3, U).conj().T
        w = back[0, 0]
        np.testing.assert_allclose(back, w * np.eye(3), atol=1e-11)
        assert abs(w ** 3 - 1) < 1e-11


@pytest.mark.parametrize("nc", [2, 3])
def test_adjoint_action(nc):
    for _ in range(20):
        V = rand_link(nc)
        E = RNG.normal(size=O.ne(nc))
        MV, ME = O.to_matrix(nc, V), O.alg_matrix(nc, E)
        X = MV @ ME @ MV.conj().T
        Y = MV.conj().T @ ME @ MV
        np.testing.assert_allclose(O.adj_act(nc, V, E), [2 * np.real(np.trace(t @ X)) for t in O.generators(nc)], atol=1e-14)
        np.testing.assert_allclose(O.adj_act(nc, V, E, True), [2 * np.real(np.trace(t @ Y)) for t in O.generators(nc)], atol=1e-14)


@pytest.mark.parametrize("nc", [2, 3])
def test_unitarity_norm(nc):
    U = rand_link(nc)
    assert O.unitarity_norm(nc, U) < 1e-14
    W = U * 1.01
    if nc == 3:
        want = np.linalg.norm(O.to_matrix(3, W) @ O.to_matrix(3, W).conj().T - np.eye(3))
    else:  # the reference norms the 4 quaternion components (SUNc/SU2/GroupOperations.cpp:172-184)
        want = abs(1.01 ** 2 - 1)
    assert abs(O.unitarity_norm(nc, W) - want) < 1e-13


def test_su3_structure_constants_from_commutators():
    # [t^a, t^b] = i f^{abc} t^c, f computed from Gell-Mann matrices
    import ctypes
    t = O.generators(3)
    f = np.zeros((8, 8, 8))
    for a in range(8):
        for b in range(8):
            C = t[a] @ t[b] - t[b] @ t[a]
            for c in range(8):
                f[a, b, c] = np.real(-2j * np.trace(C @ t[c]))
    # the oracle's vacuum estimator uses f via su3_f; check through a lattice
    # where only f enters: a single site with A_x, A_y, A_z along 3 generators
    assert abs(f[0, 1, 2] - 1) < 1e-14 and abs(f[3, 4, 7] - np.sqrt(3) / 2) < 1e-14
    lat = O.Lattice(3, (1, 1, 1))
    for (a, b, c) in [(0, 1, 2), (0, 3, 6), (0, 4, 5), (3, 4, 7), (5, 6, 7), (1, 4, 6)]:
        lat.set_identity()
        u = lat.U.reshape(3, 18)
        eps = 1e-4
        for mu, g in enumerate((a, b, c)):
            al = np.zeros(8)
            al[g] = eps
            u[mu] = O.expi(3, 1.0, al)
        A = [O.proj(3, 2.0, u[m]) for m in range(3)]
        want = sum(f[i, j, k] * A[0][i] * A[1][j] * A[2][k] for i in range(8) for j in range(8) for k in range(8)) / (16 * np.pi ** 2)
        assert abs(lat.vacuum_ncs() - want) < 1e-20 + 1e-10 * abs(want)
        assert abs(want) > 0


def test_philox_known_answers():
    # Philox-4x32-10 test vectors published with Random123 (kat_vectors)
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normals_are_standard():
    z = np.concatenate([O.link_normals(7, 2, s, l, 8) for s in range(20) for l in range(500)])
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    # different streams/steps are different numbers
    assert not np.allclose(O.link_normals(7, 1, 0, 3, 8), O.link_normals(7, 2, 0, 3, 8))
