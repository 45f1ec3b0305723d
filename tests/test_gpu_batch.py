"""Batches of uneven lattices (include/tcym_batch.h) against single-lattice
contexts on the same seeds: every field and every per-lattice sum must be
bit-identical, through eager steps, graph-replayed segments and after the
autotuner has changed the update kernels' walk shape and grid.

The reference runs such ensembles as independent MPI ranks with seed + rank
(SU2Solver.cpp:161-164); the single-lattice path is itself checked against the
oracle in test_gpu_parity.py."""
import numpy as np
import pytest

import topologicalcymsolver_amd as T

pytestmark = pytest.mark.gpu

# linear walk (4^3, 6^3, 5x7x9, 8^3) and patch walk (16^3) lattices
SU2_SIZES = [(4, 4, 4), (6, 6, 6), (5, 7, 9), (16, 16, 16), (8, 8, 8)]
SU3_SIZES = [(4, 4, 4), (6, 5, 4), (8, 8, 8)]
SEED0 = 4004
NOISE0 = 910


def build(nc, sizes, prec, su=0.2, se=0.3, a=(1.0, 1.0, 1.0), aScale=1.0):
    """A batch and one Solver per lattice with the same initial condition."""
    b = T.Batch(nc, sizes, prec, a=a, aScale=aScale)
    iters = b.initial_condition(SEED0, su, se, maxiter=3000, tol=1e-14)
    singles = []
    for e, n in enumerate(sizes):
        g64 = T.Solver(nc, n, T.FP64, a=a, aScale=aScale)
        g64.random_links(SEED0 + e, su)
        g64.random_efield(SEED0 + e, se)
        it, _ = g64.gauss_law_restore(maxiter=3000, tol=1e-14)
        assert it == iters[e]
        s = T.Solver(nc, n, prec, a=a, aScale=aScale)
        s.set_links(g64.get_links())
        s.set_efield(g64.get_efield())
        g64.close()
        singles.append(s)
    return b, singles


def assert_same_fields(b, singles, slot_u=T.U_HOT, slot_e=T.E_HOT):
    for e, s in enumerate(singles):
        assert np.array_equal(b.get_links(e, slot_u), s.get_links(slot_u)), ("links", e)
        if slot_e is not None:
            assert np.array_equal(b.get_efield(e, slot_e), s.get_efield(slot_e)), ("efield", e)


def assert_same_measurements(b, singles):
    nd = b.ncs_dot()
    bu = b.bulk()
    ga = b.gauss()
    un = b.unitarity()
    for e, s in enumerate(singles):
        assert nd[e] == s.ncs_dot(), ("ncsdot", e)
        assert np.array_equal(bu[e], s.bulk_update()), ("bulk", e)
        assert tuple(ga[e]) == s.gauss_violation(), ("gauss", e)
        assert tuple(un[e]) == s.unitarity_violation(), ("unitarity", e)


@pytest.mark.parametrize("nc,sizes", [(2, SU2_SIZES), (3, SU3_SIZES)])
@pytest.mark.parametrize("prec", [T.FP32, T.FP64])
@pytest.mark.parametrize("spacing", [((1.0, 1.0, 1.0), 1.0), ((1.0, 1.3, 0.7), 0.9)])
def test_batch_matches_single_lattices(nc, sizes, prec, spacing):
    b, singles = build(nc, sizes, prec, a=spacing[0], aScale=spacing[1])
    assert_same_fields(b, singles)
    assert_same_measurements(b, singles)
    # Langevin: key NOISE0 + e, Philox steps continue across calls
    b.langevin(3, 0.05, 0.1, 0.5, NOISE0)
    b.langevin(2, 0.05, 0.1, 0.5, NOISE0)
    assert b.get_step() == 5
    for e, s in enumerate(singles):
        s.langevin_update(3, 0.05, 0.1, 0.5, NOISE0 + e, 0)
        s.langevin_update(2, 0.05, 0.1, 0.5, NOISE0 + e, 3)
    assert_same_fields(b, singles)
    # leapfrog
    b.leapfrog(3, 0.05)
    for s in singles:
        s.dynamics_update(3, 0.05)
    assert_same_fields(b, singles)
    assert_same_measurements(b, singles)


@pytest.mark.parametrize("nc,sizes", [(2, SU2_SIZES), (3, SU3_SIZES[:2])])
def test_batch_cooling_matches_single_lattices(nc, sizes):
    b, singles = build(nc, sizes, T.FP32)
    b.cooling_start(4, 0.02)
    for s in singles:
        s.cooling_method_start(4, 0.02)
    assert_same_fields(b, singles, T.U_OLD, None)
    b.langevin(3, 0.05, 0.1, 0.5, NOISE0)
    d = b.cooling_update(4, 0.02)
    for e, s in enumerate(singles):
        s.langevin_update(3, 0.05, 0.1, 0.5, NOISE0 + e, 0)
        assert d[e] == s.cooling_method_update(4, 0.02), e
    assert_same_fields(b, singles, T.U_OLD, None)
    assert_same_fields(b, singles, T.U_MID, T.E_MID)
    # the fused flow of a slot
    b.flow(3, 0.02, u=T.U_HOT)
    for s in singles:
        s.gradient_flow(3, 0.02, u=T.U_HOT, e=-1)
    assert_same_fields(b, singles, T.U_HOT, None)


@pytest.mark.parametrize("mode", [0, 1])
def test_graph_segments_match_eager_steps(mode):
    """evolve() replays a captured graph; every replay must equal the same
    steps and measurements run one call at a time (and a single lattice)."""
    sizes = SU2_SIZES
    b1, singles = build(2, sizes, T.FP32)
    b2, _ = build(2, sizes, T.FP32)
    nseg, steps = 4, 3
    traj = b1.evolve(mode, nseg, steps, 0.05, 0.1, 0.5, NOISE0)
    want = np.empty((len(sizes), nseg))
    for j in range(nseg):
        if mode == 1:
            b2.langevin(steps, 0.05, 0.1, 0.5, NOISE0)
        else:
            b2.leapfrog(steps, 0.05)
        want[:, j] = b2.ncs_dot()
        for e, s in enumerate(singles):
            if mode == 1:
                s.langevin_update(steps, 0.05, 0.1, 0.5, NOISE0 + e, j * steps)
            else:
                s.dynamics_update(steps, 0.05)
            assert traj[e, j] == s.ncs_dot(), (e, j)
    assert np.array_equal(traj, want)
    for e in range(len(sizes)):
        assert np.array_equal(b1.get_links(e), b2.get_links(e))
        assert np.array_equal(b1.get_efield(e), b2.get_efield(e))
    assert_same_fields(b1, singles)
    # a second evolve call continues the same Philox stream and reuses the graph
    b1.evolve(mode, 2, steps, 0.05, 0.1, 0.5, NOISE0)
    for j in range(2):
        if mode == 1:
            b2.langevin(steps, 0.05, 0.1, 0.5, NOISE0)
        else:
            b2.leapfrog(steps, 0.05)
    for e in range(len(sizes)):
        assert np.array_equal(b1.get_links(e), b2.get_links(e))


def test_odd_segment_length_alternates_buffers():
    b1, _ = build(2, SU2_SIZES[:3], T.FP64)
    b2, _ = build(2, SU2_SIZES[:3], T.FP64)
    b1.evolve(1, 3, 1, 0.05, 0.1, 0.5, NOISE0)
    for j in range(3):
        b2.langevin(1, 0.05, 0.1, 0.5, NOISE0)
    for e in range(3):
        assert np.array_equal(b1.get_links(e), b2.get_links(e))


@pytest.mark.parametrize("nc,sizes", [(2, SU2_SIZES), (3, SU3_SIZES)])
def test_autotuned_batch_is_bit_identical(nc, sizes):
    b, singles = build(nc, sizes, T.FP32)
    choice, us = b.tune(mode=1, dt=0.05, reps=2)
    assert choice[0] in (-1, 0, 4, 5, 6, 7, 8, 9) and choice[1] in (2, 4, 8, 16) and us > 0
    # the tuner ran on copies: the fields are untouched
    assert_same_fields(b, singles)
    # the single lattices run the form the tuner picked
    for s in singles:
        s.set_update_form(T.FORM_PLAQUETTE if choice[0] >= 7 else T.FORM_LINK)
    b.langevin(3, 0.05, 0.1, 0.5, NOISE0)
    b.leapfrog(2, 0.05)
    for e, s in enumerate(singles):
        s.langevin_update(3, 0.05, 0.1, 0.5, NOISE0 + e, 0)
        s.dynamics_update(2, 0.05)
    assert_same_fields(b, singles)
    assert_same_measurements(b, singles)
    traj = b.evolve(1, 2, 2, 0.05, 0.1, 0.5, NOISE0)
    for e, s in enumerate(singles):
        s.langevin_update(2, 0.05, 0.1, 0.5, NOISE0 + e, 3)
        assert traj[e, 0] == s.ncs_dot()
    for e, s in enumerate(singles):                   # the second segment of evolve
        s.langevin_update(2, 0.05, 0.1, 0.5, NOISE0 + e, 5)
    assert_same_fields(b, singles)
    # the flow pass keeps a choice of its own
    fchoice, fus = b.tune(mode=3, dt=0.02, reps=2)
    assert fchoice[0] in (-1, 0, 4, 5, 6, 7, 8, 9) and fchoice[1] in (2, 4, 8, 16) and fus > 0
    assert_same_fields(b, singles)
    for s in singles:
        s.set_update_form(T.FORM_PLAQUETTE if fchoice[0] >= 7 else T.FORM_LINK)
    b.flow(5, 0.02, u=T.U_HOT)
    for s in singles:
        s.gradient_flow(5, 0.02, u=T.U_HOT, e=-1)
    assert_same_fields(b, singles)


@pytest.mark.parametrize("nc,sizes", [(2, SU2_SIZES + [(16, 24, 12), (48, 48, 20), (40, 10, 6)]), (3, SU3_SIZES)])
@pytest.mark.parametrize("shape", [7, 8, 9])
def test_plaquette_batch_matches_single_lattices(nc, sizes, shape):
    """The plaquette form in a batch (256-thread blocks: a whole plane per
    block where it fits, else bands of a few rows with a recomputed halo row
    and idle threads beyond them; units of 4, 8 or 16 planes) against each
    lattice alone in the plaquette form (768-thread blocks, other bands and
    slabs): every site's force is the same sum in the same order, so the
    fields and measurements are the same to the bit, through Langevin,
    leapfrog, flow and graph-replayed segments."""
    b, singles = build(nc, sizes, T.FP32)
    for s in singles:
        s.set_update_form(T.FORM_PLAQUETTE)
    b.set_schedule(1, shape, 4, True)
    b.set_schedule(3, shape, 4, True)
    b.timing(True)
    b.langevin(3, 0.05, 0.1, 0.5, NOISE0)
    b.leapfrog(2, 0.05)
    for e, s in enumerate(singles):
        s.langevin_update(3, 0.05, 0.1, 0.5, NOISE0 + e, 0)
        s.dynamics_update(2, 0.05)
    assert_same_fields(b, singles)
    assert_same_measurements(b, singles)
    traj = b.evolve(1, 2, 2, 0.05, 0.1, 0.5, NOISE0)
    for k in range(2):
        for e, s in enumerate(singles):
            s.langevin_update(2, 0.05, 0.1, 0.5, NOISE0 + e, 3 + 2 * k)
            assert traj[e, k] == s.ncs_dot(), (e, k)
    assert_same_fields(b, singles)
    b.flow(5, 0.02, u=T.U_HOT)
    for s in singles:
        s.gradient_flow(5, 0.02, u=T.U_HOT, e=-1)
    assert_same_fields(b, singles, slot_e=None)
    # the plaquette kernels ran, and only they
    assert b.kernel_time("batch_langevin_plaq")[1] == 3 and b.kernel_time("batch_leapfrog_plaq")[1] == 2
    assert b.kernel_time("batch_flow_plaq")[1] == 5 and b.kernel_time("batch_flow")[1] == 0
    assert b.kernel_time("batch_langevin")[1] == 0
    b.close()
    for s in singles:
        s.close()


def test_plaquette_batch_schedule_checked():
    b = T.Batch(2, [(4, 4, 4), (6, 6, 6)], T.FP64)
    with pytest.raises(T.TcymError):
        b.set_schedule(3, 7)            # FP64 storage: per-link form only
    b.close()
    b = T.Batch(2, [(4, 4, 4), (200, 7, 4)], T.FP32)
    with pytest.raises(T.TcymError):
        b.set_schedule(1, 8)            # rows of 200 sites do not fit a 256-thread block in bands
    for bad in ((1, 10), (1, 3), (2, 7)):
        with pytest.raises(T.TcymError):
            b.set_schedule(*bad)
    b.set_schedule(1, 0, 4, False)
    b.close()


def test_batch_rejects_bad_arguments():
    with pytest.raises(T.TcymError):
        T.Batch(2, [(2, 4, 4)], T.FP32)
    b = T.Batch(2, [(4, 4, 4)], T.FP32)
    buf = np.empty(3 * 64 * 4)
    assert b.L.tcym_batch_get_links(b.h, T.U_HOT, 3, buf.ctypes.data_as(T._dp)) == T.E_ARG
    assert b.L.tcym_batch_get_links(b.h, 99, 0, buf.ctypes.data_as(T._dp)) == T.E_ARG
    with pytest.raises(T.TcymError):
        b.evolve(2, 1, 1, 0.05)


def test_batched_geodesic_at_scale_matches_single_lattices():
    """The cooled Delta N_CS of a batch of 15 lattices from 16^3 to 48^3 in
    FP32: the batched geodesic interpolation (U_mid, E_mid) must equal the
    single-lattice one on the same cooled configurations, bit for bit, at
    every site (a scalar buffer offset once corrupted a few sites per run at
    this scale, which the small batches above did not show)."""
    sizes = [(L, L, L) for L in (16, 24, 32, 40, 48) * 3]
    b = T.Batch(2, sizes, T.FP32)
    b.initial_condition(SEED0, 0.2, 0.3, maxiter=3000, tol=1e-14)
    b.cooling_start(6, 0.02)
    old = [b.get_links(e, T.U_OLD) for e in range(len(sizes))]
    b.langevin(5, 0.05, 0.1, 0.5, NOISE0)
    d = b.cooling_update(6, 0.02)
    assert np.isfinite(d).all()
    for e, n in enumerate(sizes):
        s = T.Solver(2, n, T.FP32)
        s.set_links(old[e], T.U_OLD)
        s.set_links(b.get_links(e, T.U_NEW), T.U_NEW)
        s.geodesic_interpolation(0.5)
        assert np.array_equal(b.get_links(e, T.U_MID), s.get_links(T.U_MID)), e
        assert np.array_equal(b.get_efield(e, T.E_MID), s.get_efield(T.E_MID)), e
        s.close()
