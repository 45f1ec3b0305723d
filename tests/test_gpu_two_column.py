"""Between the steps of one update call the SU(3) update kernels store only
columns 0 and 1 of the new links (the next step loads those and rebuilds
column 2); the last step of the call stores complete links.  Checked here:

- n steps in one call give bit for bit the links and fields of n calls of one
  step (every step of which stores complete links), for leapfrog, Langevin
  and the gradient flow, in both precisions, single lattices and batches;
- the links an entry point leaves behind are complete: column 2 agrees with
  conj(col0 x col1) to the storage precision (a stale column 2 from two steps
  back is off by the step size), and the measurement kernels, which read
  whole links, agree between the two runs."""
import numpy as np
import pytest

import topologicalcymsolver_amd as T

pytestmark = pytest.mark.gpu


def bits(A):
    return np.ascontiguousarray(A, dtype=np.float64).view(np.uint64)


def column2_error(U):
    """max |col2 - conj(col0 x col1)| over the links of AoS SU(3) links."""
    M = U.reshape(-1, 9, 2)
    M = M[..., 0] + 1j * M[..., 1]              # element (row, col) at row + 3 col
    c0, c1, c2 = M[:, 0:3], M[:, 3:6], M[:, 6:9]
    return np.abs(c2 - np.conj(np.cross(c0, c1))).max()


def solver(N, prec, seed):
    g = T.Solver(3, N, prec)
    g.random_links(seed, 0.3)
    g.random_efield(seed, 0.4)
    return g


def advance(g, mode, k0, k):
    if mode == "leapfrog":
        g.dynamics_update(k, 0.05)
    elif mode == "langevin":
        g.langevin_update(k, 0.05, 0.3, 0.5, seed=7, step0=k0)
    else:
        g.gradient_flow(k, 0.02, u=T.U_HOT, e=T.E_HOT)


@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
@pytest.mark.parametrize("mode", ["leapfrog", "langevin", "flow"])
def test_steps_in_one_call_equal_single_steps(prec, mode):
    N, n = (8, 6, 5), 5
    a, b = solver(N, prec, 31), solver(N, prec, 31)
    U0 = a.get_links()
    advance(a, mode, 0, n)
    for k in range(n):
        advance(b, mode, k, 1)
    Ua, Ub = a.get_links(), b.get_links()
    assert np.array_equal(bits(Ua), bits(Ub))
    assert np.array_equal(bits(a.get_efield()), bits(b.get_efield()))
    tol = 2e-6 if prec == T.FP32 else 1e-13
    assert column2_error(Ua) < tol
    assert np.abs(Ua - U0).max() > 100 * tol          # the links moved
    assert a.unitarity_violation() == b.unitarity_violation()
    assert np.array_equal(a.bulk_update(), b.bulk_update())
    assert a.ncs_dot() == b.ncs_dot()
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
def test_cooling_and_hot_ncs_see_complete_links(prec):
    """The cooled copy (one call of 12 flow steps) and the per-step E.B of
    the hot-trajectory integral read whole links."""
    N = (6, 6, 6)
    a, b = solver(N, prec, 57), solver(N, prec, 57)
    a.cool_n_save(12, 0.02)
    b.copy_links(T.U_NEW, T.U_HOT)
    for _ in range(12):
        b.gradient_flow(1, 0.02, u=T.U_NEW, e=T.E_COOL)
    assert np.array_equal(bits(a.get_links(T.U_NEW)), bits(b.get_links(T.U_NEW)))
    assert a.vacuum_ncs(u=T.U_NEW) == b.vacuum_ncs(u=T.U_NEW)
    da = a.evolve_ncs(0, 4, 0.05)
    db = 0.0
    for _ in range(4):
        db += b.evolve_ncs(0, 1, 0.05)
    assert abs(da - db) <= 1e-12 * max(1.0, abs(da))
    assert np.array_equal(bits(a.get_links()), bits(b.get_links()))
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
def test_batch_steps_in_one_call_equal_single_steps(prec):
    sizes = [(4, 4, 4), (6, 5, 4), (8, 8, 8)]
    a, b = T.Batch(3, sizes, prec), T.Batch(3, sizes, prec)
    a.initial_condition(808, 0.3, 0.4, maxiter=2000, tol=1e-14)
    b.initial_condition(808, 0.3, 0.4, maxiter=2000, tol=1e-14)
    a.leapfrog(4, 0.05)
    a.flow(3, 0.02)
    for _ in range(4):
        b.leapfrog(1, 0.05)
    for _ in range(3):
        b.flow(1, 0.02)
    tol = 2e-6 if prec == T.FP32 else 1e-13
    for e in range(len(sizes)):
        Ua = a.get_links(e)
        assert np.array_equal(bits(Ua), bits(b.get_links(e))), e
        assert np.array_equal(bits(a.get_efield(e)), bits(b.get_efield(e))), e
        assert column2_error(Ua) < tol, e
    assert np.array_equal(a.unitarity(), b.unitarity())
    a.close()
    b.close()
