"""GPU: the cooling method's Delta N_CS over several measurement intervals
(ChernSimonsNumber::CoolingMethod::Update, TOPOLOGY/CoolingMethod.cpp:127-165).

tcym_delta_ncs_cooling keeps the field strength B of its U_NEW; when the
links are then copied to U_OLD (tcym_cooling_update does, and so does the
drop-in's CoolingMethod::Update), the next measurement reads B for its E.B
at U_OLD, so each measurement after the first forms two field strengths
instead of three.  The kept B is used only while nothing has accessed those
links since; these tests check that the kept path gives the same sums to
the bit as the path that forms all three, that any access to U_OLD (a read
or a write) between measurements makes the next one form B again, and that
the sequence matches the oracle.
"""
import numpy as np
import pytest

import _oracle as O
import topologicalcymsolver_amd as T

pytestmark = pytest.mark.gpu

CASES = [(2, (8, 6, 8)), (3, (6, 6, 4))]


def solver(nc, N, seed, precision):
    lat = O.Lattice(nc, N)
    lat.random(seed, 0.5, 0.4)
    s = T.Solver(nc, N, precision)
    s.set_links(lat.U)
    s.set_efield(lat.E)
    return lat, s


@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
@pytest.mark.parametrize("nc,N", CASES)
def test_kept_field_strength_is_bit_identical(nc, N, precision):
    _, a = solver(nc, N, 31, precision)
    _, b = solver(nc, N, 31, precision)
    a.cooling_method_start(8, 0.05)
    b.cooling_method_start(8, 0.05)
    da, db = [], []
    for _ in range(4):
        a.dynamics_update(5, 0.05)
        b.dynamics_update(5, 0.05)
        b.unitarity_violation(u=T.U_OLD)      # a read of U_OLD: b forms all three B
        da.append(a.cooling_method_update(8, 0.05, measure=True))
        db.append(b.cooling_method_update(8, 0.05, measure=True))
    assert da == db, (da, db)
    np.testing.assert_array_equal(a.get_links(T.U_OLD), b.get_links(T.U_OLD))


@pytest.mark.parametrize("nc,N", CASES)
def test_write_to_u_old_discards_kept_field_strength(nc, N):
    lat, a = solver(nc, N, 32, T.FP64)
    _, c = solver(nc, N, 32, T.FP64)
    other = lat.copy()
    other.flow(4, 0.05)       # links written into U_OLD: the hot links flowed 4 steps
    a.cooling_method_start(8, 0.05)
    c.cooling_method_start(8, 0.05)
    a.cooling_method_update(8, 0.05, measure=True)     # a keeps B of U_OLD
    c.cooling_method_update(8, 0.05, measure=False)    # c keeps nothing
    a.set_links(other.U, slot=T.U_OLD)
    c.set_links(other.U, slot=T.U_OLD)
    d_a = a.cooling_method_update(8, 0.05, measure=True)
    d_c = c.cooling_method_update(8, 0.05, measure=True)
    assert d_a == d_c, (d_a, d_c)
    # the oracle: U_NEW is the hot links flowed 8 steps (no dynamics between the calls)
    cold = lat.copy()
    cold.flow(8, 0.05)
    d_or = lat.delta_ncs_cooling(other.U, cold.U)
    assert abs(d_a - d_or) < 1e-11 * max(1.0, abs(d_or)), (d_a, d_or)


@pytest.mark.parametrize("nc,N", CASES)
def test_measurement_sequence_matches_oracle(nc, N):
    lat, gpu = solver(nc, N, 34, T.FP64)
    gpu.cooling_method_start(8, 0.05)
    cold = lat.copy()
    cold.flow(8, 0.05)
    u_old = cold.U.copy()
    for k in range(3):
        lat.leapfrog(5, 0.05)
        gpu.dynamics_update(5, 0.05)
        d_gpu = gpu.cooling_method_update(8, 0.05, measure=True)
        cold = lat.copy()
        cold.flow(8, 0.05)
        d_or = lat.delta_ncs_cooling(u_old, cold.U)
        assert abs(d_gpu - d_or) < 1e-11 * max(1.0, abs(d_or)), (k, d_gpu, d_or)
        u_old = cold.U.copy()


@pytest.mark.parametrize("nc,N", CASES)
def test_kept_field_strength_follows_the_copy(nc, N):
    # the drop-in's order: cool into U_NEW, Delta N_CS, then U_OLD <- U_NEW
    _, a = solver(nc, N, 35, T.FP32)
    _, b = solver(nc, N, 35, T.FP32)
    da, db = [], []
    for s in (a, b):
        s.copy_links(T.U_AUX, T.U_HOT)
        s.cool_n_save(8, 0.05, u_cold=T.U_NEW)
        s.copy_links(T.U_OLD, T.U_NEW)
    for _ in range(3):
        for s, d in ((a, da), (b, db)):
            s.dynamics_update(5, 0.05)
            s.cool_n_save(8, 0.05, u_cold=T.U_NEW)
            if s is b:
                s.get_links(T.U_OLD)                # b forms all three B
            d.append(s.delta_ncs_cooling(T.U_OLD, T.U_NEW))
            s.copy_links(T.U_OLD, T.U_NEW)
    assert da == db, (da, db)
    # a copy into U_OLD from links without a kept B, then a measurement
    for s in (a, b):
        s.copy_links(T.U_OLD, T.U_AUX)
        s.cool_n_save(8, 0.05, u_cold=T.U_NEW)
    b.unitarity_violation(u=T.U_OLD)
    assert a.delta_ncs_cooling(T.U_OLD, T.U_NEW) == b.delta_ncs_cooling(T.U_OLD, T.U_NEW)
