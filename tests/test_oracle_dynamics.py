"""Checks that pin the oracle's leapfrog (Dynamics::Update, DYNAMICS/Dynamics.cpp:204-218:
UpdateElectricFields :145-202, then UpdateGaugeLinks :97-138) to what the
mathematics fixes, independently of the oracle's own coefficients:

- Time reversibility.  With the kick K (E += gamma F(U)) before the drift D
  (U <- exp(i cU E) U), n steps from (U_0, E_0) followed by one kick, E -> -E,
  n steps, E -> -E and a kick with -dt return to (U_0, E_0) exactly (to
  rounding): the drift is undone by the drift of -E because the exponential
  is exact, and each kick by the kick of the same links because the force
  depends on the links only.  The same protocol on the drift-before-kick
  ordering does not return.
- The lattice dispersion of linear waves.  An abelian transverse wave
  alpha(x) = eps cos(k x) in one colour of the y links, E = 0, evolves in the
  linear regime as alpha_{n+1} - 2 alpha_n + alpha_{n-1} = -dt^2 w^2 alpha_n
  with w^2 = 4 sin^2(k a_x / 2) / a_x^2 (the lattice Laplacian, speed of
  light 1), so alpha_n = eps cos(k x) cos((n + 1/2) theta) / cos(theta / 2),
  cos theta = 1 - dt^2 w^2 / 2.  This fixes the product of the force
  normalisation, gamma = -2 dt a^3 and cU = -dt a_mu^2 / a^3 (a factor 2 or a
  spacing power off changes w, a sign turns the oscillation into growth);
  the finite-difference force test fixes the force against the energy.
"""
import numpy as np
import pytest

import _oracle as O

SPACINGS = [((1.0, 1.0, 1.0), 1.0), ((1.3, 1.0, 0.7), 0.9)]


def reversal(lat, n, dt, order="kick-drift"):
    """n steps, kick, E -> -E, n steps, E -> -E, kick with -dt."""
    def step(k):
        for _ in range(k):
            if order == "kick-drift":
                lat.leapfrog(1, dt)
            else:
                lat.update_links(dt)
                lat.update_efield(dt)
    step(n)
    lat.update_efield(dt)
    lat.E *= -1.0
    step(n)
    lat.E *= -1.0
    lat.update_efield(-dt)


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("spacing", SPACINGS)
def test_leapfrog_is_time_reversible(nc, spacing):
    a, aScale = spacing
    lat = O.Lattice(nc, (4, 3, 5), a, aScale)
    lat.random(17, 0.5, 0.6, restore=False)
    U0, E0 = lat.U.copy(), lat.E.copy()
    reversal(lat, 25, 0.05)
    assert np.abs(lat.U - U0).max() < 1e-12 and np.abs(lat.E - E0).max() < 1e-12
    # the other order of the two half steps is not reversed by this protocol
    bad = O.Lattice(nc, (4, 3, 5), a, aScale)
    bad.U, bad.E = U0.copy(), E0.copy()
    reversal(bad, 25, 0.05, order="drift-kick")
    assert max(np.abs(bad.U - U0).max(), np.abs(bad.E - E0).max()) > 1e-3


def set_wave(lat, direction, polarisation, colour, eps, m):
    """Links of the polarisation direction exp(i alpha t^colour), alpha = eps cos(k x_direction),
    k = 2 pi m / N_direction; every other link 1, E = 0."""
    nc, N = lat.nc, lat.N
    lat.set_identity()
    lat.E[:] = 0.0
    s = O.su(nc)
    k = 2.0 * np.pi * m / N[direction]
    for z in range(N[2]):
        for y in range(N[1]):
            for x in range(N[0]):
                c = (x, y, z)[direction]
                al = np.zeros(O.ne(nc))
                al[colour] = eps * np.cos(k * c)
                site = x + N[0] * (y + N[1] * z)
                lat.U[s * (polarisation + 3 * site):s * (polarisation + 3 * site + 1)] = O.expi(nc, 1.0, al)
    return k


def wave_amplitudes(lat, direction, polarisation):
    """alpha^a of the polarisation links along the line y = z = 0 (or the
    direction's axis), one row per site of that line."""
    nc, N = lat.nc, lat.N
    s = O.su(nc)
    out = []
    for c in range(N[direction]):
        q = [0, 0, 0]
        q[direction] = c
        site = q[0] + N[0] * (q[1] + N[1] * q[2])
        out.append(O.logi(nc, 1.0, lat.U[s * (polarisation + 3 * site):s * (polarisation + 3 * site + 1)]))
    return np.array(out)


@pytest.mark.parametrize("nc,colour", [(2, 2), (2, 0), (3, 2), (3, 7)])
@pytest.mark.parametrize("spacing,direction,polarisation", [(SPACINGS[0], 0, 1), (SPACINGS[1], 0, 1), (SPACINGS[1], 2, 0)])
def test_linear_waves_follow_the_lattice_dispersion(nc, colour, spacing, direction, polarisation):
    a, aScale = spacing
    N = [4, 4, 4]
    N[direction] = 8
    lat = O.Lattice(nc, tuple(N), a, aScale)
    eps, m, dt, block = 1e-4, 1, 0.05, 20
    k = set_wave(lat, direction, polarisation, colour, eps, m)
    ad = a[direction]
    w2 = 4.0 * np.sin(k / 2.0) ** 2 / ad ** 2
    theta = np.arccos(1.0 - dt * dt * w2 / 2.0)
    xs = np.arange(N[direction])
    for b in range(1, 9):
        lat.leapfrog(block, dt)
        n = b * block
        al = wave_amplitudes(lat, direction, polarisation)
        want = eps * np.cos(k * xs) * np.cos((n + 0.5) * theta) / np.cos(theta / 2.0)
        assert np.abs(al[:, colour] - want).max() < 1e-6 * eps, (n, al[:, colour], want)
        others = np.delete(al, colour, axis=1)
        assert np.abs(others).max() < 1e-6 * eps
        # a wave at half the frequency squared (a dropped factor 2) is far off
        th2 = np.arccos(1.0 - dt * dt * w2 / 4.0)
        if b == 8:
            off = eps * np.cos(k * xs) * np.cos((n + 0.5) * th2) / np.cos(th2 / 2.0)
            assert np.abs(al[:, colour] - off).max() > 0.1 * eps
    # the other links stay 1 (no longitudinal or other-polarisation response)
    for mu in range(3):
        if mu == polarisation:
            continue
        u = lat.U.reshape(-1, 3, O.su(nc))[:, mu, :]
        one = O.Lattice(nc, (1, 1, 1)).U.reshape(3, -1)[0]
        assert np.abs(u - one).max() < 1e-12
