"""The GPU leapfrog (tcym_leapfrog, tcym_update_efield; Dynamics::Update,
DYNAMICS/Dynamics.cpp:204-218) against what the mathematics fixes, through
the C ABI: time reversibility and the lattice dispersion of linear waves (the
closed forms are derived in tests/test_oracle_dynamics.py).  FP64 to
rounding; the FP32 form (per-link and plaquette kernels) to its storage
precision."""
import numpy as np
import pytest

import _oracle as O
import topologicalcymsolver_amd as T
from test_oracle_dynamics import SPACINGS, set_wave, wave_amplitudes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("spacing", SPACINGS)
@pytest.mark.parametrize("prec,form,N,tol", [(T.FP64, T.FORM_AUTO, (4, 3, 5), 1e-12),
                                             (T.FP32, T.FORM_LINK, (4, 3, 5), 2e-5),
                                             (T.FP32, T.FORM_PLAQUETTE, (64, 4, 3), 2e-5)])
def test_leapfrog_is_time_reversible(nc, spacing, prec, form, N, tol):
    """n steps, one kick, E -> -E, n steps, E -> -E, a kick with -dt: back at
    the start (FP32: the links are rounded and reunitarised at every step)."""
    a, aScale = spacing
    lat = O.Lattice(nc, N, a, aScale)
    lat.random(17, 0.5, 0.6, restore=False)
    g = T.Solver(nc, N, prec, a=a, aScale=aScale)
    g.set_update_form(form)
    g.set_links(lat.U)
    g.set_efield(lat.E)
    U0, E0 = g.get_links(), g.get_efield()
    n, dt = 25, 0.05
    g.dynamics_update(n, dt)
    g.update_electric_fields(dt)
    g.set_efield(-g.get_efield())
    g.dynamics_update(n, dt)
    g.set_efield(-g.get_efield())
    g.update_electric_fields(-dt)
    du, de = np.abs(g.get_links() - U0).max(), np.abs(g.get_efield() - E0).max()
    g.close()
    print("reversal: max|dU| %.2e max|dE| %.2e" % (du, de))
    assert du < tol and de < tol


@pytest.mark.parametrize("nc,colour", [(2, 2), (3, 7)])
@pytest.mark.parametrize("spacing,direction,polarisation", [(SPACINGS[0], 0, 1), (SPACINGS[1], 2, 0)])
def test_linear_waves_follow_the_lattice_dispersion_fp64(nc, colour, spacing, direction, polarisation):
    a, aScale = spacing
    N = [4, 4, 4]
    N[direction] = 8
    lat = O.Lattice(nc, tuple(N), a, aScale)
    eps, dt, block = 1e-4, 0.05, 20
    k = set_wave(lat, direction, polarisation, colour, eps, 1)
    g = T.Solver(nc, tuple(N), T.FP64, a=a, aScale=aScale)
    g.set_links(lat.U)
    g.set_efield(lat.E)
    w2 = 4.0 * np.sin(k / 2.0) ** 2 / a[direction] ** 2
    theta = np.arccos(1.0 - dt * dt * w2 / 2.0)
    xs = np.arange(N[direction])
    for b in range(1, 9):
        g.dynamics_update(block, dt)
        lat.U = g.get_links()
        al = wave_amplitudes(lat, direction, polarisation)
        want = eps * np.cos(k * xs) * np.cos((b * block + 0.5) * theta) / np.cos(theta / 2.0)
        assert np.abs(al[:, colour] - want).max() < 1e-6 * eps
        assert np.abs(np.delete(al, colour, axis=1)).max() < 1e-6 * eps
    g.close()


@pytest.mark.parametrize("nc,colour", [(2, 2), (3, 7)])
@pytest.mark.parametrize("form", [T.FORM_LINK, T.FORM_PLAQUETTE])
def test_linear_waves_follow_the_lattice_dispersion_fp32(nc, colour, form):
    """A long wave (k = 2 pi / 64 along x) at eps = 1e-2, where the plaquette
    angles (k eps ~ 1e-3) keep the nonlinear shift near 1e-7 and the FP32
    rounding of the links (6e-8) accumulates to about 1e-4 of eps over the
    160 steps."""
    N = (64, 4, 4)
    lat = O.Lattice(nc, N)
    eps, dt, block = 1e-2, 0.2, 20
    k = set_wave(lat, 0, 1, colour, eps, 1)
    g = T.Solver(nc, N, T.FP32)
    g.set_update_form(form)
    g.timing(True)
    g.set_links(lat.U)
    g.set_efield(lat.E)
    w2 = 4.0 * np.sin(k / 2.0) ** 2
    theta = np.arccos(1.0 - dt * dt * w2 / 2.0)
    xs = np.arange(N[0])
    for b in range(1, 9):
        g.dynamics_update(block, dt)
        lat.U = g.get_links()
        al = wave_amplitudes(lat, 0, 1)
        want = eps * np.cos(k * xs) * np.cos((b * block + 0.5) * theta) / np.cos(theta / 2.0)
        assert np.abs(al[:, colour] - want).max() < 1e-3 * eps
    plaq = g.kernel_time("leapfrog_plaq")[1]
    g.close()
    assert plaq == (8 * block if form == T.FORM_PLAQUETTE else 0)
