"""The C++ drop-in (include/tcym_dropin.hpp): a program written against the
reference's Dynamics::Update / NCsDot / Bulk / HardScales / GaussLaw /
Unitarity / Cooling::CoolNSave / CoolingMethod::Calibrate /
IO::SaveConfiguration / IO::LoadConfiguration / ThermalDynamics::Thermalize /
InitialConditions::SetQuasiParticles entry points over
GaugeLinks/ElectricFields runs on the GPU and matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import topologicalcymsolver_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "dropin", "_build", "dropin_check")


def test_dropin_check_is_built():
    assert os.path.exists(EXE), "run __graft_entry__.build() first"


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [2, 3])
def test_dropin_matches_oracle(nc, tmp_path):
    N = 6
    lat = O.Lattice(nc, N)
    lat.random(2024, 0.4, 0.3, restore=True)
    lat.U.tofile(tmp_path / "U.bin")
    lat.E.tofile(tmp_path / "E.bin")
    r = subprocess.run([EXE, str(nc), str(N), str(tmp_path / "U.bin"), str(tmp_path / "E.bin"), str(tmp_path / "out"),
                        str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    ncs, t00, gmax, umax, txx, tzz, lxx, lzz = map(float, lines[0].split())
    lat.leapfrog(4, 0.05)
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_U.bin"), lat.U, atol=1e-12)
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_E.bin"), lat.E, atol=1e-12)
    assert abs(ncs - lat.ncsdot()) < 1e-13
    assert abs(t00 - lat.energy()) < 1e-12 * lat.energy()
    assert abs(gmax - lat.gauss()[0]) < 1e-12 and abs(umax - lat.unitarity()[0]) < 1e-12
    # T_ii and the hard scales (a = aScale = 1: cE = cB = 1)
    b = lat.bulk()
    assert abs(txx - 0.5 * ((b[1] + b[2] + b[4] + b[5]) - (b[0] + b[3]))) < 1e-12 * lat.energy()
    assert abs(tzz - 0.5 * ((b[0] + b[1] + b[3] + b[4]) - (b[2] + b[5]))) < 1e-12 * lat.energy()
    H = lat.hard_scales()
    assert abs(lxx - (H[1] + H[2] - H[0])) < 1e-12 * H.sum()
    assert abs(lzz - (H[0] + H[1] - H[2])) < 1e-12 * H.sum()
    # 4 more steps inside a Resident scope, downloaded when the scope ends
    lat.leapfrog(4, 0.05)
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_U2.bin"), lat.U, atol=1e-12)
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_E2.bin"), lat.E, atol=1e-12)

    # Cooling::CoolNSave(8 steps, MaxBlockingNumber 1) with BlockFrequency 3:
    # blocked after step 3, the cold copy is the 3^3 lattice at spacing 2
    dtau = 0.2 * 0.2
    cold, nb = lat.copy(), 0
    for k in range(1, 9):
        cold.flow(1, dtau)
        if k % (3 * 8 ** nb) == 0 and nb < 1:
            cold = cold.blocked()
            nb += 1
    n_c, a_c, nb_c, steps_c, ctau_c = lines[1].split()
    assert (int(n_c), float(a_c), int(nb_c), int(steps_c)) == (N // 2, 2.0, 1, 8)
    # 3 steps at a = 1 and 5 at a = 2: (a^3)^(2/3) = 1 and 4
    assert abs(float(ctau_c) - dtau * (3 + 5 * 4)) < 1e-14
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_C.bin"), cold.U, atol=1e-12)
    # CoolingMethod::Start (4 steps) then Calibrate: calibration steps 5..9 from
    # the cooled copy with E zero, blocked after step 6 (BlockFrequency 6)
    work = lat.copy()
    work.flow(4, dtau)
    work.E[:] = 0.0
    acc, ctau, nb, rows = 0.0, 4 * dtau, 0, []

    def row(w):
        b = w.bulk()
        cb = (w.p.a[0] / w.p.a[0] ** 3) ** 2
        return [ctau, 0.5 * cb * b[3:].sum(), 0.5 * cb * b[:3].sum(), acc, w.vacuum_ncs()]

    rows.append(row(work))
    for k in range(5, 10):
        acc += work.flow_calibration(1, dtau)
        ctau += dtau * (work.p.a[0] ** 3) ** (2.0 / 3.0)
        rows.append(row(work))
        if k % (6 * 8 ** nb) == 0 and nb < 1:
            work = work.blocked()
            nb += 1
    d_cal, nb_cal, steps_cal, ctau_cal = lines[2].split()
    assert (int(nb_cal), int(steps_cal)) == (1, 9)
    assert abs(float(d_cal) - acc) <= 1e-10 * max(abs(acc), 1e-12)
    assert abs(float(ctau_cal) - ctau) < 1e-13
    # the log file, written with the stream's default 6 significant digits
    log = np.loadtxt(tmp_path / "VacuumCoolingT0.4ID7.log")
    np.testing.assert_allclose(log, np.array(rows), rtol=1e-5, atol=1e-9)
    # IO::SaveConfiguration with the reference's 15 digits, IO::LoadConfiguration back
    U2, E2 = np.fromfile(tmp_path / "out_U2.bin"), np.fromfile(tmp_path / "out_E2.bin")
    U15 = np.array([float("%.15g" % v) for v in U2])
    E15 = np.array([float("%.15g" % v) for v in E2])
    assert np.array_equal(T.read_text("links", nc, (N, N, N), tmp_path / "UOutT0.4ID7.txt"), U15)
    assert np.array_equal(T.read_text("efield", nc, (N, N, N), tmp_path / "EOutT0.4ID7.txt"), E15)
    back = lat.copy()
    back.U[:], back.E[:] = U15, E15
    assert abs(float(lines[3]) - back.ncsdot()) < 1e-13
    # InitialConditions::SetZero + ThermalDynamics::Thermalize(2, 1.6), beta 4, key MySEED = 7
    th = O.Lattice(nc, N)
    rows = th.thermalize(2, 1.6, 0.0375, 4.0, 7)
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_TU.bin"), th.U, atol=1e-11)
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_TE.bin"), th.E, atol=1e-11)
    np.testing.assert_allclose(np.loadtxt(tmp_path / "ThermalEnergyID7.txt"), rows, rtol=1e-5, atol=1e-9)
    # InitialConditions::SetQuasiParticles (Qs 1.5, n0 1, key 7) with its Gauss-law restoration
    qp = O.Lattice(nc, N)
    O.quasi_particles(qp, 1.5, 1.0, 7)
    qp.gauss_restore()
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_QU.bin"), qp.U, atol=1e-12)
    np.testing.assert_allclose(np.fromfile(tmp_path / "out_QE.bin"), qp.E, atol=1e-12)


DRIVER = os.path.join(ROOT, "tests", "dropin", "_build", "sphaleron_driver")


def test_sphaleron_driver_is_built():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"


def oracle_sphaleron_run(nc, N, U0, E0):
    """The sphaleron driver's run on the oracle: 40 leapfrog steps of 0.025,
    cooled copies of 12 flow steps (dtau 1/8) every 5 steps with the geodesic
    Delta N_CS, calibration (flow steps 13..18, blocked after 16) at the start
    and every 20 steps, energy rows every 20 steps; then 3 Langevin steps and a
    noise-field step."""
    dt, dtau = 0.025, 0.125
    lat = O.Lattice(nc, N)
    lat.U[:], lat.E[:] = U0, E0
    out = {"energy": [], "cool": [], "calib": [], "logs": {}}

    def energy(t):
        b, H = lat.bulk(), lat.hard_scales()
        el, mg = 0.5 * b[:3].sum(), 0.5 * b[3:].sum()
        tii = [0.5 * (b.sum() - 2 * (b[m] + b[3 + m])) for m in range(3)]
        lam = [H.sum() - 2 * H[m] for m in range(3)]
        out["energy"].append([t, el + mg] + tii + [el, mg] + lam)

    def cool(U):
        c = lat.copy()
        c.U[:], c.E[:] = U, 0.0
        c.flow(12, dtau)
        return c.U.copy()

    state = {"calib": 0.0, "rt": 0.0}

    def calibrate(t, Unew):
        w = lat.copy()
        w.U[:], w.E[:] = Unew, 0.0
        acc, ctau, nb, rows = 0.0, 12 * dtau, 0, []

        def row(w):
            b = w.bulk()
            cb = (w.p.a[0] / w.p.a[0] ** 3) ** 2
            return [ctau, 0.5 * cb * b[3:].sum(), 0.5 * cb * b[:3].sum(), acc, w.vacuum_ncs()]
        rows.append(row(w))
        for k in range(13, 19):
            acc += w.flow_calibration(1, dtau)
            ctau += dtau * (w.p.a[0] ** 3) ** (2.0 / 3.0)
            rows.append(row(w))
            if k == 16 and nb < 1:
                w = w.blocked()
                nb += 1
        out["logs"]["%g" % t] = np.array(rows)
        out["calib"].append([t, acc, state["calib"], state["rt"], state.get("rtprev", 0.0)])
        state["rtprev"] = state["rt"]
        state["calib"] = acc

    energy(0.0)
    Uold = cool(lat.U)
    out["cool"].append([0.0, 0.0, 0.0])
    calibrate(0.0, Uold)
    for step in range(1, 41):
        lat.leapfrog(1, dt)
        t = step * dt
        if step % 20 == 0:
            energy(t)
        if step % 5 == 0:
            Unew = cool(lat.U)
            if step == 5:
                out["hot_map"] = lat.edotb_map()
                Um, Em = lat.geodesic(Uold, Unew, 0.5)
                out["cool_map"] = _with(lat, Um, Em).edotb_map()
            state["rt"] += lat.delta_ncs_cooling(Uold, Unew)
            out["cool"].append([t, state["rt"], 0.0])
            Uold = Unew
        if step % 20 == 0:
            calibrate(t, Uold)
    lat.langevin(3, 0.01, 1.0, 0.5, 7, 0)
    Xi = lat.noise(7, 3, 1.0)
    lat.langevin_xi(Xi, 0.01, 1.0)
    out["U"], out["E"], out["box"] = lat.U.copy(), lat.E.copy(), lat.edotb_map()
    return out


def _with(lat, U, E):
    c = lat.copy()
    c.U[:], c.E[:] = U, E
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("resident", [0, 1])
def test_sphaleron_driver_matches_oracle(nc, resident, tmp_path):
    N = 8
    lat = O.Lattice(nc, N)
    lat.random(2025, 0.4, 0.3, restore=True)
    lat.U.tofile(tmp_path / "U.bin")
    lat.E.tofile(tmp_path / "E.bin")
    print("sphaleron driver SU(%d) %d^3 resident %d" % (nc, N, resident), flush=True)
    r = subprocess.run([DRIVER, str(nc), str(N), str(tmp_path / "U.bin"), str(tmp_path / "E.bin"), str(tmp_path),
                        str(resident)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].split() == ["refused", "3", "tSteps", "42", "tauLSteps", "4"], r.stdout
    want = oracle_sphaleron_run(nc, N, lat.U, lat.E)
    # CoolNSave (6 flow steps) + NCsDot of the cold copy with the hot E, twice,
    # a leapfrog step between: the second value must come from the second copy
    hot = _with(lat, want["U"], want["E"])
    for k, got in enumerate(map(float, lines[1].split())):
        cold = _with(hot, hot.U, hot.E)
        cold.flow(6, 0.125)
        assert abs(got - cold.ncsdot(cold.U, hot.E)) < 1e-12, (k, got)
        hot.leapfrog(1, 0.025)
    np.testing.assert_allclose(np.fromfile(tmp_path / "final_U.bin"), want["U"], atol=1e-11)
    np.testing.assert_allclose(np.fromfile(tmp_path / "final_E.bin"), want["E"], atol=1e-11)
    scale = abs(want["energy"][0][1])
    np.testing.assert_allclose(np.loadtxt(tmp_path / "EnergyID7.txt"), np.array(want["energy"]), rtol=1e-10,
                               atol=1e-12 * scale)
    np.testing.assert_allclose(np.loadtxt(tmp_path / "CoolNCsID7.txt"), np.array(want["cool"]), rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(np.loadtxt(tmp_path / "CalibNCsID7.txt"), np.array(want["calib"]), rtol=1e-9,
                               atol=1e-13)
    # the calibration logs, the maps and the box map in the stream's default 6 digits
    for t, rows in want["logs"].items():
        np.testing.assert_allclose(np.loadtxt(tmp_path / ("VacuumCoolingT%sID7.log" % t)), rows, rtol=1e-5, atol=1e-9)
    for name, m in (("HotEdotBMapC0", want["hot_map"]), ("CoolEdotBMapC0", want["cool_map"])):
        got = np.loadtxt(tmp_path / (name + "ID7.txt"))
        assert got.shape == (N ** 3, 4)
        xyz = got[:, :3].astype(int)
        assert [tuple(v) for v in xyz[:3]] == [(0, 0, 0), (1, 0, 0), (2, 0, 0)]
        np.testing.assert_allclose(got[:, 3], m[xyz[:, 0] + N * (xyz[:, 1] + N * xyz[:, 2])], rtol=1e-5,
                                   atol=1e-6 * np.abs(m).max())
    box = np.loadtxt(tmp_path / "BoxMapID7.txt")
    assert [tuple(v) for v in box[:, :3].astype(int)] == [(1, 0, 3), (2, 0, 3)]
    np.testing.assert_allclose(box[:, 3], want["box"][[1 + N * N * 3, 2 + N * N * 3]], rtol=1e-5, atol=1e-9)
    # the precision checks of the loop
    assert "#GAUSS LAW VIOLATION AT T=0.625 MAX=" in r.stderr and "#UNITARITY VIOLATION AT T=0.625 MAX=" in r.stderr
    assert r.stderr.count("#CALIBRATION COMPLETED") == 3
