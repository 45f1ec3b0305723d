"""GPU parity: the HIP library (through the C ABI) against the oracle on the
same seeded inputs.

FP64 form: per-kernel results agree with the oracle to <= 1e-12 absolute on
O(1) quantities (the two sides use different but exact exp/log algorithms and
a different summation order, so agreement is at rounding level, not bitwise);
trajectories to relative 1e-10; Gauss and unitarity violations, which are
rounding-level quantities, to absolute 1e-12.
"""
import numpy as np
import pytest

import _oracle as O
import topologicalcymsolver_amd as T

pytestmark = pytest.mark.gpu

CASES = [(2, (6, 7, 8)), (3, (5, 6, 4))]


def pair(nc, N, seed, su, se, restore=False, precision=T.FP64):
    lat = O.Lattice(nc, N)
    lat.random(seed, su, se, restore=restore)
    gpu = T.Solver(nc, N, precision)
    gpu.set_links(lat.U)
    gpu.set_efield(lat.E)
    return lat, gpu


@pytest.mark.parametrize("nc,N", CASES)
def test_upload_download_round_trip(nc, N):
    lat, gpu = pair(nc, N, 3, 0.7, 0.4)
    np.testing.assert_array_equal(gpu.get_links(), lat.U)
    np.testing.assert_array_equal(gpu.get_efield(), lat.E)
    gpu.copy_links(T.U_AUX, T.U_HOT)
    np.testing.assert_array_equal(gpu.get_links(T.U_AUX), lat.U)
    gpu.zero_efield(T.E_AUX)
    assert not gpu.get_efield(T.E_AUX).any()


@pytest.mark.parametrize("sigma", [0.2, 1.5])
@pytest.mark.parametrize("nc,N", CASES)
def test_random_initial_conditions_match_oracle(nc, N, sigma):
    # sigma 1.5 takes the SU(3) exponential through its scaling and squaring
    # (c1 = tr(Q^2)/2 well above 1/4)
    lat = O.Lattice(nc, N)
    lat.random(1234, sigma, 0.3, restore=False)
    gpu = T.Solver(nc, N)
    gpu.random_links(1234, sigma)
    gpu.random_efield(1234, 0.3)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-14 if sigma < 1 else 1e-13)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-14)


@pytest.mark.parametrize("nc,N", CASES)
def test_leapfrog_matches_oracle(nc, N):
    lat, gpu = pair(nc, N, 7, 0.6, 0.5)
    lat.leapfrog(10, 0.05)
    gpu.dynamics_update(10, 0.05)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)


@pytest.mark.parametrize("nc,N", CASES)
def test_split_updates_match_oracle(nc, N):
    lat, gpu = pair(nc, N, 8, 0.6, 0.5)
    lat.leapfrog(1, 0.03)
    gpu.update_electric_fields(0.03)
    gpu.update_gauge_links(0.03)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-13)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-13)


@pytest.mark.parametrize("nc,N", CASES)
def test_langevin_matches_oracle(nc, N):
    lat, gpu = pair(nc, N, 9, 0.5, 0.3)
    lat.langevin(6, 0.05, 0.1, 0.5, seed=77, step0=4)
    gpu.langevin_update(6, 0.05, 0.1, 0.5, seed=77, step0=4)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)


@pytest.mark.parametrize("nc,N", CASES)
def test_langevin_with_noise_field_matches_in_kernel_noise(nc, N):
    lat, gpu = pair(nc, N, 10, 0.5, 0.3)
    gpu.copy_links(T.U_AUX, T.U_HOT)
    gpu.copy_efield(T.E_AUX, T.E_HOT)
    gpu.langevin_update(1, 0.05, 1.0, 0.5, seed=5, step0=11)
    gpu.random_gaussian_noise(5, 11, np.sqrt(2 * 1.0 * 0.5))
    Xi = lat.noise(5, 11, np.sqrt(2 * 1.0 * 0.5))
    np.testing.assert_allclose(gpu.get_efield(T.E_XI), Xi, atol=1e-14)
    gpu.langevin_update_fields(0.05, 1.0, u=T.U_AUX, e=T.E_AUX)
    np.testing.assert_allclose(gpu.get_links(T.U_AUX), gpu.get_links(), atol=1e-15)
    np.testing.assert_allclose(gpu.get_efield(T.E_AUX), gpu.get_efield(), atol=1e-15)


@pytest.mark.parametrize("nc,N", CASES)
def test_flow_matches_oracle(nc, N):
    lat, gpu = pair(nc, N, 11, 0.6, 0.0)
    lat.flow(5, 0.02)
    gpu.gradient_flow(5, 0.02, u=T.U_HOT, e=T.E_HOT)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)
    # fused flow without storing the force gives the same links
    lat2, gpu2 = pair(nc, N, 11, 0.6, 0.0)
    gpu2.gradient_flow(5, 0.02, u=T.U_HOT, e=-1)
    np.testing.assert_allclose(gpu2.get_links(), lat.U, atol=1e-12)


@pytest.mark.parametrize("nc,N", CASES)
def test_calibration_flow_matches_oracle(nc, N):
    lat, gpu = pair(nc, N, 12, 0.6, 0.0)
    d_or = lat.flow_calibration(4, 0.05)
    d_gpu = gpu.gradient_flow(4, 0.05, calibrate=True, u=T.U_HOT, e=T.E_HOT)
    assert abs(d_gpu - d_or) < 1e-12 * max(1.0, abs(d_or))
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)


@pytest.mark.parametrize("nc,N", CASES)
def test_observables_match_oracle(nc, N):
    lat, gpu = pair(nc, N, 13, 0.7, 0.6)
    for improved in (True, False):
        want = lat.edotb() if improved else lat.edotb_unimproved()
        got = gpu.edotb(improved)
        assert abs(got - want) < 1e-12 * max(1.0, abs(want)), (improved, got, want)
    np.testing.assert_allclose(gpu.edotb_map(), lat.edotb_map(), atol=1e-12)
    assert abs(gpu.ncs_dot() - lat.ncsdot()) < 1e-13
    np.testing.assert_allclose(gpu.bulk_update(), lat.bulk(), rtol=1e-13, atol=1e-14)
    gm, gr = gpu.gauss_violation()
    om, orr = lat.gauss()
    assert abs(gm - om) < 1e-12 and abs(gr - orr) < 1e-12
    um, ur = gpu.unitarity_violation()
    vm, vr = lat.unitarity()
    assert abs(um - vm) < 1e-12 and abs(ur - vr) < 1e-12
    assert abs(gpu.vacuum_ncs() - lat.vacuum_ncs()) < 1e-13


@pytest.mark.parametrize("nc,N", CASES)
def test_anisotropic_lattice_matches_oracle(nc, N):
    """Spacings (1, 1.3, 0.7) and aScale 0.9 enter every coefficient (the
    oracle's anisotropic force and energy are pinned in test_oracle_lattice.py):
    leapfrog, Langevin, flow, the observables and GaussRestore."""
    a, asc = (1.0, 1.3, 0.7), 0.9

    def make(seed, su, se, restore=False):
        lat = O.Lattice(nc, N, a=a, aScale=asc)
        lat.random(seed, su, se, restore=restore)
        gpu = T.Solver(nc, N, T.FP64, a=a, aScale=asc)
        gpu.set_links(lat.U)
        gpu.set_efield(lat.E)
        return lat, gpu

    lat, gpu = make(41, 0.5, 0.4)
    lat.leapfrog(6, 0.05)
    gpu.dynamics_update(6, 0.05)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)
    lat.langevin(4, 0.05, 0.2, 0.5, seed=3, step0=2)
    gpu.langevin_update(4, 0.05, 0.2, 0.5, seed=3, step0=2)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)
    for improved in (True, False):
        want = lat.edotb() if improved else lat.edotb_unimproved()
        assert abs(gpu.edotb(improved) - want) < 1e-12 * max(1.0, abs(want)), improved
    np.testing.assert_allclose(gpu.edotb_map(), lat.edotb_map(), atol=1e-12)
    assert abs(gpu.ncs_dot() - lat.ncsdot()) < 1e-13
    np.testing.assert_allclose(gpu.bulk_update(), lat.bulk(), rtol=1e-13, atol=1e-14)
    gm, gr = gpu.gauss_violation()
    om, orr = lat.gauss()
    assert abs(gm - om) < 1e-12 and abs(gr - orr) < 1e-12
    assert abs(gpu.vacuum_ncs() - lat.vacuum_ncs()) < 1e-13
    lat, gpu = make(42, 0.6, 0.0)
    lat.flow(4, 0.03)
    gpu.gradient_flow(4, 0.03, u=T.U_HOT, e=T.E_HOT)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)
    lat, gpu = make(43, 0.5, 0.5)
    it_o, _ = lat.gauss_restore()
    it_g, _ = gpu.gauss_law_restore()
    assert it_g == it_o, (it_g, it_o)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)


@pytest.mark.parametrize("nc,N", CASES)
def test_gauss_restore_matches_oracle(nc, N):
    lat, gpu = pair(nc, N, 14, 0.5, 0.5)
    it_o, fm_o = lat.gauss_restore()
    it_g, fm_g = gpu.gauss_law_restore()
    assert it_g == it_o, (it_g, it_o)
    assert fm_g <= 1e-14 and fm_o <= 1e-14
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)


def test_gauss_restore_reports_iteration_cap():
    lat, gpu = pair(2, (6, 6, 6), 15, 0.5, 0.5)
    it, fm = gpu.gauss_law_restore(maxiter=2, check=False)
    assert it == 2 and fm > 1e-14
    with pytest.raises(T.TcymError) as ei:
        gpu.gauss_law_restore(maxiter=1)
    assert ei.value.code == T.E_NOCONV


def test_geodesic_large_and_degenerate_angles_match_oracle():
    """U_new = exp(i A) U_old per link with the eigenvalues of A spread over
    (-pi + 0.1, pi - 0.1), and with doubly degenerate spectra (A along
    lambda_8 or lambda_3 with equal phases): the log's Cayley shift and its
    Jacobi rotations (hardware reciprocal estimates with Newton steps) against
    the oracle's eigen-decomposition."""
    import scipy.linalg as sla
    nc, N = 3, (4, 4, 4)
    lat, gpu = pair(nc, N, 41, 0.5, 0.0)
    Uold = lat.U.copy()
    rng = np.random.default_rng(43)
    t = O.generators(3)
    nl = Uold.size // 18
    Unew = np.empty_like(Uold)
    for k in range(nl):
        if k % 4 == 0:
            a = np.zeros(8)
            a[7 if k % 8 == 0 else 2] = rng.uniform(-2.5, 2.5)
        else:
            while True:
                a = rng.normal(0, 1.5, 8)
                if np.max(np.abs(np.linalg.eigvalsh(O.alg_matrix(3, a)))) < np.pi - 0.1:
                    break
        W = sla.expm(1j * O.alg_matrix(3, a))
        Unew[18*k:18*k + 18] = O.from_matrix(3, W @ O.to_matrix(3, Uold[18*k:18*k + 18]))
    gpu.set_links(Uold, T.U_OLD)
    gpu.set_links(Unew, T.U_NEW)
    gpu.geodesic_interpolation(0.5)
    Um, Em = lat.geodesic(Uold, Unew, 0.5)
    np.testing.assert_allclose(gpu.get_links(T.U_MID), Um, atol=1e-11)
    np.testing.assert_allclose(gpu.get_efield(T.E_MID), Em, atol=1e-11)


@pytest.mark.parametrize("nc,N", CASES)
def test_geodesic_and_cooling_delta_ncs_match_oracle(nc, N):
    lat, gpu = pair(nc, N, 16, 0.5, 0.0)
    Uold = lat.U.copy()
    lat2 = O.Lattice(nc, N)
    lat2.random(17, 0.5, 0.0, restore=False)
    Unew = lat2.U.copy()
    gpu.set_links(Uold, T.U_OLD)
    gpu.set_links(Unew, T.U_NEW)
    gpu.geodesic_interpolation(0.3)
    Um, Em = lat.geodesic(Uold, Unew, 0.3)
    np.testing.assert_allclose(gpu.get_links(T.U_MID), Um, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(T.E_MID), Em, atol=1e-12)
    d_or = lat.delta_ncs_cooling(Uold, Unew)
    d_gpu = gpu.delta_ncs_cooling()
    assert abs(d_gpu - d_or) < 1e-12 * max(1.0, abs(d_or))


@pytest.mark.parametrize("nc,N", CASES)
def test_cooling_method_update_matches_oracle(nc, N):
    # ChernSimonsNumber::CoolingMethod: cool hot links 8 flow steps -> UNew,
    # Delta N_CS between consecutive cooled configurations, UOld <- UNew
    lat, gpu = pair(nc, N, 18, 0.5, 0.4)
    gpu.cooling_method_start(8, 0.05)
    cold = lat.copy()
    cold.flow(8, 0.05)
    Uold = cold.U.copy()
    np.testing.assert_allclose(gpu.get_links(T.U_OLD), Uold, atol=1e-12)
    lat.leapfrog(5, 0.05)
    gpu.dynamics_update(5, 0.05)
    d_gpu = gpu.cooling_method_update(8, 0.05, measure=True)
    cold = lat.copy()
    cold.flow(8, 0.05)
    d_or = lat.delta_ncs_cooling(Uold, cold.U)
    assert abs(d_gpu - d_or) < 1e-11 * max(1.0, abs(d_or)), (d_gpu, d_or)
    np.testing.assert_allclose(gpu.get_links(T.U_OLD), cold.U, atol=1e-12)


@pytest.mark.parametrize("nc", [2, 3])
def test_fp32_link_update_at_large_angles_tracks_fp64(nc):
    # Dynamics::UpdateGaugeLinks with E large enough that the exponent's
    # c1 = tr(Q^2)/2 exceeds 1/4 (SU(3): scaling and squaring in FP32; SU(2):
    # the sin/cos branch), FP32 against FP64 on the same fields
    N = (6, 6, 4)
    lat = O.Lattice(nc, N)
    lat.random(29, 0.4, 2.5, restore=False)
    g64 = T.Solver(nc, N, T.FP64)
    g32 = T.Solver(nc, N, T.FP32)
    for g in (g64, g32):
        g.set_links(lat.U)
        g.set_efield(lat.E)
        g.update_gauge_links(1.0)
    np.testing.assert_allclose(g32.get_links(), g64.get_links(), atol=3e-6)
    assert g32.unitarity_violation()[0] < 2e-6


@pytest.mark.parametrize("nc", [2, 3])
def test_fp32_form_tracks_fp64(nc):
    N = (8, 8, 8)
    lat, g64 = pair(nc, N, 19, 0.3, 0.3, restore=True)
    g32 = T.Solver(nc, N, T.FP32)
    g32.set_links(lat.U)
    g32.set_efield(lat.E)
    g64.dynamics_update(40, 0.05)
    g32.dynamics_update(40, 0.05)
    assert g32.unitarity_violation()[0] < 2e-6          # FP64 reunitarisation after every step
    assert abs(g32.energy() - g64.energy()) < 1e-5 * g64.energy()
    np.testing.assert_allclose(g32.get_links(), g64.get_links(), atol=2e-4)
    assert abs(g32.ncs_dot() - g64.ncs_dot()) < 1e-5
    # FP64 Gauss restoration of the FP32 fields
    g32.random_efield(5, 0.3)
    it, fm = g32.gauss_law_restore()
    assert fm <= 1e-14
    assert g32.gauss_violation()[1] < 1e-6               # stored back in FP32


def test_trajectory_su2_16_matches_oracle():
    """SU(2) 16^3, sigma 0.2, sigma_E 0.3 Gauss-projected, 200 leapfrog steps
    dt = 0.05 (the reference's CPU-runnable configuration): energy and the
    hot-trajectory N_CS integral dN += dt/2 [NCsDot(U_n, E_n+1/2) + NCsDot(U_n+1, E_n+1/2)]
    agree with the oracle to relative 1e-10; Gauss/unitarity to 1e-12 absolute."""
    nc, N, dt = 2, (16, 16, 16), 0.05
    lat = O.Lattice(nc, N)
    lat.random(1001, 0.2, 0.3, restore=True)
    gpu = T.Solver(nc, N)
    gpu.set_links(lat.U)
    gpu.set_efield(lat.E)
    n_or = n_gpu = 0.0
    e_or, e_gpu, ncs_or, ncs_gpu = [], [], [], []
    for step in range(200):
        U0 = lat.U.copy()
        lat.leapfrog(1, dt)
        n_or += 0.5 * dt * (lat.ncsdot(U0, lat.E) + lat.ncsdot())
        gpu.copy_links(T.U_AUX, T.U_HOT)
        gpu.dynamics_update(1, dt)
        n_gpu += 0.5 * dt * (gpu.ncs_dot(u=T.U_AUX) + gpu.ncs_dot())
        if step % 10 == 9:
            e_or.append(lat.energy()); e_gpu.append(gpu.energy())
            ncs_or.append(n_or); ncs_gpu.append(n_gpu)
    e_or, e_gpu = np.array(e_or), np.array(e_gpu)
    assert np.abs(e_gpu - e_or).max() <= 1e-10 * np.abs(e_or).max()
    scale = max(np.abs(ncs_or).max(), 1e-12)
    assert np.abs(np.array(ncs_gpu) - np.array(ncs_or)).max() <= 1e-10 * scale, (ncs_gpu[-1], ncs_or[-1])
    assert abs(gpu.gauss_violation()[1] - lat.gauss()[1]) < 1e-12
    assert abs(gpu.unitarity_violation()[0] - lat.unitarity()[0]) < 1e-12


@pytest.mark.parametrize("nc,N", CASES)
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_hot_trajectory_ncs_on_device(nc, N, mode, precision):
    """tcym_evolve_ncs keeps the hot-trajectory integral on the device; it must
    equal the same sum composed on the host from per-step NCsDot calls (the
    composition that test_trajectory_su2_16_matches_oracle pins to the oracle).
    From the second step on its first E.B takes the field strength B of the
    links that the previous step's second E.B formed and kept: the sum is
    still the same to the bit."""
    lat, gpu = pair(nc, N, 21, 0.5, 0.4, restore=True, precision=precision)
    ref = T.Solver(nc, N, precision)
    ref.set_links(lat.U)
    ref.set_efield(lat.E)
    dt, steps = 0.05, 12
    d_dev = gpu.evolve_ncs(mode, steps, dt, 0.3, 0.5, 77, 5)
    d_host = 0.0
    for k in range(steps):
        ref.copy_links(T.U_AUX, T.U_HOT)
        if mode == 1:
            ref.langevin_update(1, dt, 0.3, 0.5, 77, 5 + k)
        else:
            ref.dynamics_update(1, dt)
        d_host += 0.5 * dt * (ref.ncs_dot(u=T.U_AUX) + ref.ncs_dot())
    assert d_dev == d_host
    assert np.array_equal(gpu.get_links(), ref.get_links())
    assert np.array_equal(gpu.get_efield(), ref.get_efield())
    # leapfrog against the oracle as well
    if mode == 0 and precision == T.FP64:
        n_or = 0.0
        for k in range(steps):
            U0 = lat.U.copy()
            lat.leapfrog(1, dt)
            n_or += 0.5 * dt * (lat.ncsdot(U0, lat.E) + lat.ncsdot())
        assert abs(d_dev - n_or) <= 1e-10 * max(abs(n_or), 1e-12)


@pytest.mark.parametrize("nc,N", CASES)
def test_fp32_langevin_noise(nc, N):
    """FP32 form: the in-kernel Langevin noise (Box-Muller in FP32 from the
    same 53-bit Philox uniforms) equals the noise-field path bit for bit and
    the oracle's FP64 draw to FP32 rounding."""
    lat, gpu = pair(nc, N, 10, 0.5, 0.3, precision=T.FP32)
    sd = np.sqrt(2 * 1.0 * 0.5)
    gpu.copy_links(T.U_AUX, T.U_HOT)
    gpu.copy_efield(T.E_AUX, T.E_HOT)
    gpu.langevin_update(1, 0.05, 1.0, 0.5, seed=5, step0=11)
    gpu.random_gaussian_noise(5, 11, sd)
    Xi = lat.noise(5, 11, sd)
    xi32 = gpu.get_efield(T.E_XI)
    np.testing.assert_allclose(xi32, Xi, rtol=5e-6, atol=5e-6 * sd)
    gpu.langevin_update_fields(0.05, 1.0, u=T.U_AUX, e=T.E_AUX)
    assert np.array_equal(gpu.get_links(T.U_AUX), gpu.get_links())
    assert np.array_equal(gpu.get_efield(T.E_AUX), gpu.get_efield())


@pytest.mark.parametrize("nc,N", CASES)
def test_hard_scales_match_oracle(nc, N):
    lat, gpu = pair(nc, N, 14, 0.7, 0.6)
    np.testing.assert_allclose(gpu.hard_scales(), lat.hard_scales(), rtol=1e-12, atol=1e-14)


BLOCK_CASES = [(2, (8, 6, 4)), (3, (4, 8, 4))]


@pytest.mark.parametrize("nc,N", BLOCK_CASES)
def test_blocking_matches_oracle(nc, N):
    lat, gpu = pair(nc, N, 31, 0.6, 0.4)
    gpu.block(T.U_HOT, T.E_HOT)
    coarse = lat.blocked()
    assert gpu.level(u=T.U_HOT) == (1, coarse.N) and gpu.level(e=T.E_HOT) == (1, coarse.N)
    np.testing.assert_allclose(gpu.get_links(), coarse.U, atol=1e-13)
    assert not gpu.get_efield().any()
    # measurements at the coarse resolution (spacing 2)
    np.testing.assert_allclose(gpu.bulk_update(), coarse.bulk(), rtol=1e-12, atol=1e-14)
    assert abs(gpu.vacuum_ncs() - coarse.vacuum_ncs()) < 1e-13
    # a slot at another resolution cannot be mixed in
    gpu.set_efield(lat.E, T.E_AUX)
    with pytest.raises(T.TcymError):
        gpu.bulk_update(u=T.U_HOT, e=T.E_AUX)


@pytest.mark.parametrize("nc,N", BLOCK_CASES)
def test_cool_blocked_matches_oracle(nc, N):
    """Cooling::CoolNSave with BlockFrequency 3 and MaxBlockingNumber 1: the
    cold links are blocked after flow step 3 and flowed on at spacing 2."""
    lat, gpu = pair(nc, N, 32, 0.6, 0.0)
    nb = gpu.cool_blocked(8, 0.02, 3, 1, u_hot=T.U_HOT, u_cold=T.U_NEW)
    cold, nbo = lat.copy(), 0
    for k in range(1, 9):
        cold.flow(1, 0.02)
        if k % (3 * 8 ** nbo) == 0 and nbo < 1:
            cold = cold.blocked()
            nbo += 1
    assert nb == nbo == 1
    np.testing.assert_allclose(gpu.get_links(T.U_NEW), cold.U, atol=1e-12)
    # hot links untouched, at full resolution
    np.testing.assert_allclose(gpu.get_links(T.U_HOT), lat.U, atol=0)


@pytest.mark.parametrize("nc,N", BLOCK_CASES)
def test_continue_cooling_matches_oracle(nc, N):
    """Cooling::ContinueCooling in calibration mode: Delta N_CS along the flow,
    the per-step log {t, MAGNETIC, ELECTRIC, DeltaNCs, VacEst} and blocking."""
    lat, gpu = pair(nc, N, 33, 0.6, 0.0)
    step0, nsteps, dtau, ctau0 = 2, 7, 0.05, 0.3
    d, nb, log = gpu.continue_cooling(step0, nsteps, dtau, ctau0, block_frequency=4, max_blocking=1,
                                      u_src=T.U_HOT, u_work=T.U_COOL, e_work=T.E_COOL)
    work = lat.copy()
    work.E[:] = 0.0
    acc, ctau, nbo, rows = 0.0, ctau0, 0, []

    def row(w):
        b = w.bulk()
        a = w.p.a[0]
        cb = (a / a ** 3) ** 2
        return [ctau, 0.5 * cb * b[3:].sum(), 0.5 * cb * b[:3].sum(), acc, w.vacuum_ncs()]

    rows.append(row(work))
    for k in range(step0 + 1, nsteps + 1):
        acc += work.flow_calibration(1, dtau)
        ctau += dtau * (work.p.a[0] ** 3) ** (2.0 / 3.0)
        rows.append(row(work))
        if k % (4 * 8 ** nbo) == 0 and nbo < 1:
            work = work.blocked()
            nbo += 1
    assert nb == nbo == 1
    assert abs(d - acc) <= 1e-10 * max(abs(acc), 1e-12)
    np.testing.assert_allclose(log, np.array(rows), rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(gpu.get_links(T.U_COOL), work.U, atol=1e-12)


@pytest.mark.parametrize("nc,N", [(2, (6, 5, 4)), (3, (4, 5, 6))])
def test_thermalize_matches_oracle(nc, N):
    """ThermalDynamics::Thermalize: thermal E reassignment (Philox stream 3),
    Gauss-law restoration, leapfrog and the ThermalEnergy rows every 20 steps,
    from unit links; then the FP32 form against the FP64 one."""
    lat = O.Lattice(nc, N)
    gpu = T.Solver(nc, N, T.FP64)
    gpu.unit_links()
    gpu.zero_efield()
    args = (2, 1.6, 0.0375, 4.0, 2024)   # 2 cycles of 43 steps, rows at steps 20 and 40
    rows_o = lat.thermalize(*args)
    rows_g = gpu.thermalize(*args)
    assert rows_g.shape == rows_o.shape == (4, 7)
    np.testing.assert_allclose(rows_g, rows_o, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-11)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-11)
    g32 = T.Solver(nc, N, T.FP32)
    g32.unit_links()
    g32.zero_efield()
    rows_32 = g32.thermalize(*args)
    np.testing.assert_allclose(rows_32[:, 1:], rows_g[:, 1:], rtol=2e-5, atol=1e-7)
    np.testing.assert_array_equal(rows_32[:, 0], rows_g[:, 0])


@pytest.mark.parametrize("nc,N,a", [(2, (6, 5, 4), (1.0, 1.0, 1.0)), (3, (4, 6, 5), (1.0, 1.25, 0.8))])
def test_quasi_particles_match_oracle(nc, N, a):
    """InitialConditions::SetQuasiParticles: momentum modes, rocFFT inverse
    transform, exponentiated links and E against the numpy restatement; then
    the FP32 form and the Gauss-law restoration the reference applies."""
    lat = O.Lattice(nc, N, a=a, aScale=0.9)
    O.quasi_particles(lat, 1.5, 1.0, 77)
    gpu = T.Solver(nc, N, T.FP64, a=a, aScale=0.9)
    gpu.quasi_particles(1.5, 1.0, 77)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)
    assert gpu.unitarity_violation()[0] < 1e-13
    g32 = T.Solver(nc, N, T.FP32, a=a, aScale=0.9)
    g32.quasi_particles(1.5, 1.0, 77)
    np.testing.assert_allclose(g32.get_links(), lat.U, atol=2e-7)
    np.testing.assert_allclose(g32.get_efield(), lat.E, atol=1e-6 * np.abs(lat.E).max())
    it, fm = gpu.gauss_law_restore()
    it_o, fm_o = lat.gauss_restore()
    assert it == it_o and fm <= 1e-14
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)
