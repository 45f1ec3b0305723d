"""The plaquette form of the update kernels (k_update_plaq: each plaquette
multiplied out once and shared among its four links, FP32 storage; by
default on lattices whose x extent is a multiple of 64, forced on any whose
rows fit a block) against the oracle and against the
per-link form (TCYM_FORM_LINK) on the same seeded inputs.

Both forms sum the same force terms in FP32, in different orders, so they
agree with each other and with the FP64 oracle to FP32 rounding (atol 5e-5
after a few steps, the bar of the per-link form's own FP32 tests); the
plaquette form must not be further from the oracle than the per-link form
is.  Shapes cover one block, several blocks along y and z (the halo row and
the z slab boundaries, periodic wrap included), a z extent that the slab
length does not divide, nx = 64 to 384 (2 to 8 rows per block), and
anisotropic spacings."""
import numpy as np
import pytest

import _oracle as O
import topologicalcymsolver_amd as T

pytestmark = pytest.mark.gpu

# PY (rows per block) 8, 8, 7, 4, 4, 2, 3, 2: with PY = 2 the second row builds
# a halo plaquette and is also the last row.  Rows that are not whole waves
# (forced plaquette form): (16, 16, 5) and (24, 6, 7) a whole plane per block,
# (40, 30, 4) and (48, 18, 5) bands of 15 and 9 rows with a halo row
SHAPES = [(64, 8, 8), (64, 16, 24), (64, 7, 6), (128, 4, 5), (192, 8, 13), (256, 4, 6), (256, 6, 4), (384, 4, 3),
          (16, 16, 5), (24, 6, 7), (40, 30, 4), (48, 18, 5)]
SPACINGS = [((1.0, 1.0, 1.0), 1.0), ((1.0, 1.3, 0.7), 0.9)]
KERNEL = {"leapfrog": "leapfrog", "langevin": "langevin", "flow": "flow"}


def run(nc, N, a, aScale, kind, form, lat0, steps):
    g = T.Solver(nc, N, T.FP32, a=a, aScale=aScale)
    g.set_update_form(form)
    g.set_links(lat0.U)
    g.set_efield(lat0.E)
    g.timing(True)
    if kind == "leapfrog":
        g.dynamics_update(steps, 0.05)
    elif kind == "langevin":
        g.langevin_update(steps, 0.05, 0.5, 0.4, 77, 3)
    else:
        g.gradient_flow(steps, 0.02, u=T.U_HOT, e=T.E_HOT)
    plaq = g.kernel_time(KERNEL[kind] + "_plaq")[1]
    link = g.kernel_time(KERNEL[kind])[1]
    out = g.get_links(), g.get_efield(), g.energy(), g.ncs_dot()
    g.close()
    return out, plaq, link


def oracle(nc, N, a, aScale, kind, lat0, steps):
    lat = lat0.copy()
    if kind == "leapfrog":
        lat.leapfrog(steps, 0.05)
    elif kind == "langevin":
        lat.langevin(steps, 0.05, 0.5, 0.4, 77, 3)
    else:
        lat.flow(steps, 0.02)
    return lat


@pytest.mark.parametrize("nc", [3, 2])
@pytest.mark.parametrize("N", SHAPES)
@pytest.mark.parametrize("kind", ["leapfrog", "langevin", "flow"])
@pytest.mark.parametrize("spacing", SPACINGS)
def test_plaquette_form_matches_oracle_and_link_form(nc, N, kind, spacing):
    a, aScale = spacing
    steps = 4
    print("plaquette form: SU(%d) %s %s a=%s aScale=%g" % (nc, N, kind, a, aScale), flush=True)
    lat0 = O.Lattice(nc, N, a, aScale)
    lat0.random(31, 0.3, 0.3, restore=False)
    ref = oracle(nc, N, a, aScale, kind, lat0, steps)
    (Up, Ep, enp, ndp), plaq_p, link_p = run(nc, N, a, aScale, kind, T.FORM_PLAQUETTE, lat0, steps)
    (Ul, El, enl, ndl), plaq_l, link_l = run(nc, N, a, aScale, kind, T.FORM_LINK, lat0, steps)
    # the plaquette kernel ran in the plaquette form and only there
    assert (plaq_p, link_p) == (steps, 0)
    assert (plaq_l, link_l) == (0, steps)
    errp = max(np.abs(Up - ref.U).max(), np.abs(Ep - ref.E).max())
    errl = max(np.abs(Ul - ref.U).max(), np.abs(El - ref.E).max())
    assert errp < 5e-5 and errl < 5e-5, (errp, errl)
    assert errp <= 2.0 * errl + 1e-6, (errp, errl)
    assert np.abs(Up - Ul).max() < 5e-5 and np.abs(Ep - El).max() < 5e-5
    en = ref.energy()
    assert abs(enp - en) < 1e-5 * abs(en) and abs(enl - en) < 1e-5 * abs(en)
    nd = ref.ncsdot()
    assert abs(ndp - nd) < 1e-5 and abs(ndl - nd) < 1e-5


def test_plaquette_form_long_flow_tracks_link_form():
    """200 flow steps (a cooling episode) at SU(3) 192x8x13: both FP32 forms
    stay within FP32 rounding of the FP64 form, the plaquette form no further
    from it than 1.5 times the per-link form (both sum the same force in
    another order), and both keep the links unitary.

    The flowed E is the force of the cooled links, a sum of plaquette terms
    that nearly cancel: its rms is 2e-3 of the initial one and its FP32
    rounding error is 7.5e-5 of its rms (bar 2e-4).  The energy, a sum of
    squares of it and of the plaquettes, moves with that error by up to twice
    the relative rms error (Cauchy-Schwarz): bar 4e-4 relative."""
    N = (192, 8, 13)
    lat0 = O.Lattice(3, N)
    lat0.random(5, 0.3, 0.3, restore=False)
    res = []
    for prec, form in ((T.FP32, T.FORM_PLAQUETTE), (T.FP32, T.FORM_LINK), (T.FP64, T.FORM_LINK)):
        g = T.Solver(3, N, prec)
        g.set_update_form(form)
        g.set_links(lat0.U)
        g.set_efield(lat0.E)
        g.gradient_flow(200, 0.02, u=T.U_HOT, e=T.E_HOT)
        assert g.unitarity_violation()[0] < 2e-6
        res.append((g.get_links(), g.get_efield(), g.energy()))
        g.close()
    (up, fp, ep), (ul, fl, el), (u64, f64, e64) = res
    du_p, du_l = np.abs(up - u64).max(), np.abs(ul - u64).max()
    rms = lambda a: float(np.sqrt(np.mean(np.square(a))))
    df_p, df_l = rms(fp - f64) / rms(f64), rms(fl - f64) / rms(f64)
    de_p, de_l = abs(ep - e64) / abs(e64), abs(el - e64) / abs(e64)
    print("vs FP64: links max|d| plaquette %.2e, per-link %.2e; flowed E rms rel. d %.2e, %.2e; energy rel. d %.2e, %.2e"
          % (du_p, du_l, df_p, df_l, de_p, de_l))
    assert du_p < 1e-4 and du_l < 1e-4
    assert df_p < 2e-4 and df_l < 2e-4
    assert de_p < 4e-4 and de_l < 4e-4
    assert du_p < 1.5 * du_l and df_p < 1.5 * df_l


def test_update_form_argument_checked():
    g = T.Solver(3, (64, 4, 4), T.FP32)
    with pytest.raises(T.TcymError):
        g.set_update_form(7)
    g.close()
    # shapes the plaquette form does not take run the per-link kernel
    for N in [(400, 4, 4), (64, 13, 8), (448, 4, 4)]:
        g = T.Solver(3, N, T.FP32)
        g.set_update_form(T.FORM_PLAQUETTE)
        g.timing(True)
        g.random_links(3, 0.2)
        g.dynamics_update(1, 0.05)
        assert g.kernel_time("leapfrog")[1] == 1 and g.kernel_time("leapfrog_plaq")[1] == 0
        g.close()
    # FP64 storage: per-link form
    g = T.Solver(3, (64, 8, 8), T.FP64)
    g.timing(True)
    g.random_links(3, 0.2)
    g.dynamics_update(1, 0.05)
    assert g.kernel_time("leapfrog")[1] == 1 and g.kernel_time("leapfrog_plaq")[1] == 0
    g.close()


@pytest.mark.parametrize("nc,N,plaq", [(3, (64, 64, 64), False), (3, (64, 16, 24), False), (3, (192, 192, 192), True),
                                       (3, (128, 128, 128), False), (2, (128, 128, 128), True), (2, (64, 64, 64), False)])
def test_auto_form_takes_the_plaquette_form_where_it_is_faster(nc, N, plaq):
    """TCYM_FORM_AUTO: the plaquette kernel only where at least two of its
    blocks per compute unit exist (SU(3) 192^3: 768 blocks) and, for SU(3),
    the x rows are at least 192 sites long; elsewhere the per-link kernel."""
    g = T.Solver(nc, N, T.FP32)
    g.timing(True)
    g.unit_links()
    g.dynamics_update(1, 0.05)
    assert g.kernel_time("leapfrog_plaq")[1] == (1 if plaq else 0)
    assert g.kernel_time("leapfrog")[1] == (0 if plaq else 1)
    g.close()


@pytest.mark.parametrize("nc", [3, 2])
def test_plaquette_form_langevin_with_noise_field(nc):
    """Mode 2 (LangevinDynamics::UpdateFields with the Xi field) in the
    plaquette form: the same step as the in-kernel noise of mode 1 on the
    noise field that tcym_noise draws for that step, and within FP32 rounding
    of the per-link form's mode 2."""
    N = (64, 8, 8)
    lat0 = O.Lattice(nc, N)
    lat0.random(41, 0.3, 0.3, restore=False)
    sd = np.sqrt(2 * 0.5 * 0.4)
    res = {}
    for form in (T.FORM_PLAQUETTE, T.FORM_LINK):
        g = T.Solver(nc, N, T.FP32)
        g.set_update_form(form)
        g.set_links(lat0.U)
        g.set_efield(lat0.E)
        g.copy_links(T.U_AUX, T.U_HOT)
        g.copy_efield(T.E_AUX, T.E_HOT)
        g.timing(True)
        g.langevin_update(1, 0.05, 0.5, 0.4, seed=5, step0=11)
        g.random_gaussian_noise(5, 11, sd)
        g.langevin_update_fields(0.05, 0.5, u=T.U_AUX, e=T.E_AUX)
        plaq = g.kernel_time("langevin_xi_plaq")[1]
        assert plaq == (1 if form == T.FORM_PLAQUETTE else 0)
        np.testing.assert_allclose(g.get_links(T.U_AUX), g.get_links(), atol=2e-6)
        np.testing.assert_allclose(g.get_efield(T.E_AUX), g.get_efield(), atol=2e-6)
        res[form] = (g.get_links(T.U_AUX), g.get_efield(T.E_AUX))
        g.close()
    np.testing.assert_allclose(res[T.FORM_PLAQUETTE][0], res[T.FORM_LINK][0], atol=1e-5)
    np.testing.assert_allclose(res[T.FORM_PLAQUETTE][1], res[T.FORM_LINK][1], atol=1e-5)
    ref = lat0.copy()
    ref.langevin_xi(ref.noise(5, 11, sd), 0.05, 0.5)
    np.testing.assert_allclose(res[T.FORM_PLAQUETTE][0], ref.U, atol=5e-5)
    np.testing.assert_allclose(res[T.FORM_PLAQUETTE][1], ref.E, atol=5e-5)


@pytest.mark.parametrize("N,pz", [((192, 192, 96), 8), ((192, 192, 72), 6)])
def test_auto_form_with_shortened_slabs(N, pz):
    """On flatter lattices TCYM_FORM_AUTO shortens the z slab (12 -> 8 or 6
    planes) to keep two blocks per compute unit: the flow and leapfrog steps
    there agree with the per-link form to FP32 rounding."""
    print("auto form, slab of %d planes: SU(3) %s" % (pz, N), flush=True)
    res = {}
    for form in (T.FORM_AUTO, T.FORM_LINK):
        g = T.Solver(3, N, T.FP32)
        g.set_update_form(form)
        g.random_links(17, 0.3)
        g.random_efield(17, 0.3)
        g.timing(True)
        g.dynamics_update(2, 0.05)
        g.gradient_flow(3, 0.02, u=T.U_HOT, e=T.E_HOT)
        assert g.kernel_time("flow_plaq")[1] == (3 if form == T.FORM_AUTO else 0)
        assert g.kernel_time("leapfrog_plaq")[1] == (2 if form == T.FORM_AUTO else 0)
        res[form] = (g.get_links(), g.get_efield(), g.energy())
        g.close()
    assert np.abs(res[T.FORM_AUTO][0] - res[T.FORM_LINK][0]).max() < 5e-5
    assert np.abs(res[T.FORM_AUTO][1] - res[T.FORM_LINK][1]).max() < 5e-5
    assert abs(res[T.FORM_AUTO][2] - res[T.FORM_LINK][2]) < 1e-5 * abs(res[T.FORM_LINK][2])
