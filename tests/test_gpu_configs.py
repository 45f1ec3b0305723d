"""The FP32-mixed form against the FP64 form at the full sizes of the
reference's parity configurations, through properties that do not need the
oracle at that size:

- SU(2) 64^3: links exp(i eps^a sigma^a/2), eps ~ N(0, 0.2^2), E ~ N(0, 0.3^2)
  Gauss-projected; 1000 leapfrog steps dt = 0.05, then 1000 Langevin steps
  (damping 0.1, T = 0.5);
- SU(3) 96^3: eps ~ N(0, 0.15^2), E ~ N(0, 0.25^2) Gauss-projected; 500 leapfrog
  steps dt = 0.04, then 500 gradient-flow steps dtau = 0.02 on a copy;
- 64 batched SU(2) lattices, and SU(3) 192^3 leapfrog with cooling (below).

Checked: Gauss law and unitarity conserved to rounding in FP64 and their
growth in FP32, relative energy difference FP32 vs FP64 <= 1e-5, the
hot-trajectory Delta N_CS (trapezoid of NCsDot over every step) within 1e-3,
equipartition <E^a E^a> = n_E T per link after the Langevin steps, the
Langevin N_CS over its first 200 steps and, beyond them, FP32's separation
from FP64 growing at the rate of an FP64 twin perturbed by one ulp (the
Langevin phase is chaotic), the magnetic energy falling monotonically under
the flow, and the cooled fields' magnetic energy and vacuum N_CS estimator
agreeing between the forms.  The FP64 form itself is checked against the
oracle in test_gpu_parity.py."""
import numpy as np
import pytest

import topologicalcymsolver_amd as T

pytestmark = pytest.mark.gpu


def initial_condition(nc, N, seed, sigma_u, sigma_e):
    g = T.Solver(nc, N, T.FP64)
    g.random_links(seed, sigma_u)
    g.random_efield(seed, sigma_e)
    it, fm = g.gauss_law_restore(maxiter=10000, tol=1e-14)
    assert fm <= 1e-14, (it, fm)
    U, E = g.get_links(), g.get_efield()
    g.close()
    return U, E


def solver(nc, N, prec, U, E):
    g = T.Solver(nc, N, prec)
    g.set_links(U)
    g.set_efield(E)
    return g


def su2_64_run(prec, U, E):
    """1000 leapfrog steps (dt 0.05) then 1000 Langevin steps (damping 0.1,
    T 0.5): hot-trajectory N_CS every 100 steps of each phase, E at the start
    and at the end of the Langevin phase"""
    g = solver(2, (64, 64, 64), prec, U, E)
    energy, ncs = [g.energy()], [0.0]
    for _ in range(10):
        ncs.append(ncs[-1] + g.evolve_ncs(0, 100, 0.05))
        energy.append(g.energy())
    gauss, unit = g.gauss_violation()[0], g.unitarity_violation()[0]
    E0 = g.get_efield()
    lncs = [0.0]
    for k in range(10):
        lncs.append(lncs[-1] + g.evolve_ncs(1, 100, 0.05, damping=0.1, T=0.5, seed=77, step0=100 * k))
    bulk = g.bulk_update()
    r = dict(energy=np.array(energy), ncs=np.array(ncs), gauss=abs(gauss), unit=unit, esq=bulk[:3],
             lncs=np.array(lncs), E0=E0, E1=g.get_efield())
    g.close()
    return r


def rel_dist(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2) / np.mean(b ** 2)))


def test_su2_64_leapfrog_then_langevin():
    """The Langevin phase is chaotic: an FP64 twin whose initial E differs by
    one ulp separates from the FP64 run as exp(lambda t), lambda ~ 0.14 per
    unit time (x900 over the 50 time units of the 1000 Langevin steps: fields
    6e-14 -> 5e-11 relative, N_CS integral 8e-10), and a single FP32 rounding
    of the initial E in FP64 arithmetic ends 8.5e-4 away in N_CS.  The FP32
    form carries its rounding from every step: its fields separate from FP64
    at the same lambda (1.5e-5 -> 1.3e-2), and its Langevin N_CS integral
    ends ~0.1 away from the FP64 one.  So the Langevin N_CS is compared over
    the first 200 steps (within 1e-3, measured 1.2e-4), and after that FP32's
    separation must grow at the twins' rate: no faster, i.e. no error of its
    own beyond rounding."""
    N = (64, 64, 64)
    U, E = initial_condition(2, N, 2002, 0.2, 0.3)
    d = su2_64_run(T.FP64, U, E)
    f = su2_64_run(T.FP32, U, E)
    w = su2_64_run(T.FP64, U, np.nextafter(E, np.inf))
    # FP64: Gauss law and unitarity conserved by the leapfrog to rounding;
    # the energy oscillates at O(dt^2) without drift
    assert d["gauss"] < 1e-12 and d["unit"] < 1e-12, (d["gauss"], d["unit"])
    assert np.ptp(d["energy"]) < 1e-2 * d["energy"][0]
    # FP32: growth of the violations over 1000 steps (FP64 reunitarisation each step)
    assert f["gauss"] < 5e-5 and f["unit"] < 1e-6, (f["gauss"], f["unit"])
    assert np.abs(f["energy"] - d["energy"]).max() <= 1e-5 * np.abs(d["energy"]).max()
    assert np.abs(f["ncs"] - d["ncs"]).max() <= 1e-3, (f["ncs"][-1], d["ncs"][-1])
    # Langevin at T = 0.5 after 50 time units (5 damping times): <E^a E^a> = n_E T = 1.5 per link
    for r in (d, f):
        assert np.all(np.abs(r["esq"] - 1.5) < 0.03 * 1.5), r["esq"]
    assert np.abs(f["esq"] - d["esq"]).max() < 2e-3 * 1.5
    # the chaos floor and the FP32 form against it
    tw = [rel_dist(w["E0"], d["E0"]), rel_dist(w["E1"], d["E1"])]
    t32 = [rel_dist(f["E0"], d["E0"]), rel_dist(f["E1"], d["E1"])]
    lam_twin, lam32 = np.log(tw[1] / tw[0]) / 50.0, np.log(t32[1] / t32[0]) / 50.0
    dl32, dltw = np.abs(f["lncs"] - d["lncs"]), np.abs(w["lncs"] - d["lncs"])
    print("one-ulp twin: fields %.2e -> %.2e, N_CS %.2e; FP32: fields %.2e -> %.2e, N_CS %.2e (first 200 steps %.2e); "
          "lambda %.3f / %.3f" % (tw[0], tw[1], dltw.max(), t32[0], t32[1], dl32.max(), dl32[:3].max(), lam_twin, lam32))
    assert 0.05 < lam_twin < 0.5, lam_twin
    assert abs(lam32 / lam_twin - 1.0) < 0.25, (lam32, lam_twin)
    assert dltw.max() < 1e-8 and t32[0] < 1e-4
    assert dl32[:3].max() <= 1e-3, dl32
def test_su3_96_leapfrog_then_cooling():
    N = (96, 96, 96)
    U, E = initial_condition(3, N, 3003, 0.15, 0.25)
    run = {}
    for prec in (T.FP64, T.FP32):
        g = solver(3, N, prec, U, E)
        energy, ncs = [g.energy()], [0.0]
        for _ in range(5):
            ncs.append(ncs[-1] + g.evolve_ncs(0, 100, 0.04))
            energy.append(g.energy())
        gauss, unit = g.gauss_violation()[0], g.unitarity_violation()[0]
        g.copy_links(T.U_COOL, T.U_HOT)
        bmag = [g.bulk_update(u=T.U_COOL, e=T.E_COOL)[3:].sum()]
        for _ in range(5):
            g.gradient_flow(100, 0.02, u=T.U_COOL, e=T.E_COOL)
            bmag.append(g.bulk_update(u=T.U_COOL, e=T.E_COOL)[3:].sum())
        run[prec] = dict(energy=np.array(energy), ncs=np.array(ncs), gauss=abs(gauss), unit=unit, bmag=np.array(bmag),
                         vac=g.vacuum_ncs(u=T.U_COOL))
        g.close()
    d, f = run[T.FP64], run[T.FP32]
    assert d["gauss"] < 1e-12 and d["unit"] < 1e-12, (d["gauss"], d["unit"])
    assert np.ptp(d["energy"]) < 1e-2 * d["energy"][0]
    assert f["gauss"] < 5e-5 and f["unit"] < 1e-6, (f["gauss"], f["unit"])
    assert np.abs(f["energy"] - d["energy"]).max() <= 1e-5 * np.abs(d["energy"]).max()
    assert np.abs(f["ncs"] - d["ncs"]).max() <= 1e-3, (f["ncs"][-1], d["ncs"][-1])
    # the flow lowers the magnetic energy monotonically, in both forms alike
    for r in (d, f):
        assert np.all(np.diff(r["bmag"]) < 0), r["bmag"]
    assert np.abs(f["bmag"] - d["bmag"]).max() <= 1e-4 * d["bmag"].max()
    assert abs(f["vac"] - d["vac"]) <= 1e-3, (f["vac"], d["vac"])


def test_batched_su2_64_ensembles():
    """64 SU(2) lattices L^3, L drawn from {16, 24, 32, 40, 48} (numpy Philox,
    seed 4004), links eps ~ N(0, 0.2^2), E ~ N(0, 0.3^2) Gauss-projected; 500
    Langevin steps (dt 0.05, damping 0.1, T 0.5) in graph-replayed segments of
    10 steps with NCsDot of every lattice, and the cooled Delta N_CS (200 flow
    steps, dtau 0.02) every 100 steps: per lattice, FP32 against FP64, and an
    FP64 twin whose initial E differs by one ulp (the chaos floor: 3.9e-13 in
    Delta N_CS after the 500 steps, against FP32's 4.6e-4)."""
    rng = np.random.Generator(np.random.Philox(4004))
    sizes = [int(L) for L in rng.choice([16, 24, 32, 40, 48], size=64)]
    run = {}
    for prec in (T.FP64, T.FP32, "twin"):
        b = T.Batch(2, sizes, T.FP64 if prec == "twin" else prec)
        it = b.initial_condition(4004, 0.2, 0.3, maxiter=10000, tol=1e-14)
        assert np.all(np.asarray(it) < 10000)
        if prec == "twin":
            for e in range(len(sizes)):
                b.set_efield(e, np.nextafter(b.get_efield(e), np.inf))
        b.cooling_start(200, 0.02)
        ncsdot, dncs = [], []
        for _ in range(5):
            ncsdot.append(b.evolve(1, 10, 10, 0.05, 0.1, 0.5, 5004))
            dncs.append(b.cooling_update(200, 0.02))
        run[prec] = dict(ncsdot=np.concatenate(ncsdot, axis=1), dncs=np.cumsum(np.array(dncs).T, axis=1),
                         unit=b.unitarity()[:, 0].max())
        b.close()
    d, f, w = run[T.FP64], run[T.FP32], run["twin"]
    print("batched: one-ulp twin Delta N_CS %.2e, FP32 %.2e" % (np.abs(w["dncs"] - d["dncs"]).max(),
                                                                 np.abs(f["dncs"] - d["dncs"]).max()))
    assert np.abs(w["dncs"] - d["dncs"]).max() < 1e-9
    assert d["unit"] < 1e-12 and f["unit"] < 1e-6, (d["unit"], f["unit"])
    assert np.abs(f["dncs"] - d["dncs"]).max() <= 1e-3, np.abs(f["dncs"] - d["dncs"]).max()
    assert np.abs(f["ncsdot"] - d["ncsdot"]).max() <= 1e-3, np.abs(f["ncsdot"] - d["ncsdot"]).max()


def test_su3_192_leapfrog_with_cooling():
    """SU(3) 192^3 (eps ~ N(0, 0.15^2), E ~ N(0, 0.25^2) Gauss-projected): 1000
    leapfrog steps dt = 0.04, every 50 steps 200 flow steps dtau = 0.02 on a
    copy and the geodesic Delta N_CS between consecutive cooled copies.  FP32
    against FP64 over the run: |Delta N_CS| <= 1e-3 and relative energy
    difference <= 1e-5 at every measurement."""
    N = (192, 192, 192)
    U, E = initial_condition(3, N, 5005, 0.15, 0.25)
    run = {}
    for prec in (T.FP64, T.FP32):
        g = solver(3, N, prec, U, E)
        g.cooling_method_start(200, 0.02)
        energy, ncs, total = [], [], 0.0
        for _ in range(20):
            g.dynamics_update(50, 0.04)
            total += g.cooling_method_update(200, 0.02)
            ncs.append(total)
            energy.append(g.energy())
        run[prec] = dict(energy=np.array(energy), ncs=np.array(ncs), gauss=abs(g.gauss_violation()[0]),
                         unit=g.unitarity_violation()[0])
        g.close()
    d, f = run[T.FP64], run[T.FP32]
    assert d["gauss"] < 1e-12 and d["unit"] < 1e-12, (d["gauss"], d["unit"])
    assert f["gauss"] < 5e-5 and f["unit"] < 1e-6, (f["gauss"], f["unit"])
    assert np.abs(f["ncs"] - d["ncs"]).max() <= 1e-3, np.abs(f["ncs"] - d["ncs"]).max()
    assert np.abs(f["energy"] - d["energy"]).max() <= 1e-5 * np.abs(d["energy"]).max()
