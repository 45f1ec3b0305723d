"""Edge cases of the single-lattice library through the C ABI:

- the smallest lattices it accepts (extent 3, where the improved E.B's
  +-2 reach wraps onto the other side) and ragged extents, against the oracle;
- update calls of zero steps leave every field bit for bit as it was, and a
  zero-step cooling is a copy;
- the errors the entry points return: extents below 3, spacings <= 0, fields
  beyond 4 GiB, negative step counts, unknown slots, negative temperature."""
import numpy as np
import pytest

import topologicalcymsolver_amd as T
import _oracle as O

pytestmark = pytest.mark.gpu


def bits(A):
    return np.ascontiguousarray(A, dtype=np.float64).view(np.uint64)


def pair(nc, N, seed, su, se):
    lat = O.Lattice(nc, N)
    lat.random(seed, su, se, restore=False)
    gpu = T.Solver(nc, N)
    gpu.set_links(lat.U)
    gpu.set_efield(lat.E)
    return lat, gpu


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("N", [(3, 3, 3), (3, 5, 4), (7, 3, 3)])
def test_smallest_lattices_match_oracle(nc, N):
    lat, gpu = pair(nc, N, 61, 0.6, 0.5)
    lat.leapfrog(4, 0.05)
    gpu.dynamics_update(4, 0.05)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    np.testing.assert_allclose(gpu.get_efield(), lat.E, atol=1e-12)
    want = lat.edotb()
    assert abs(gpu.edotb(True) - want) < 1e-12 * max(1.0, abs(want))
    np.testing.assert_allclose(gpu.edotb_map(), lat.edotb_map(), atol=1e-12)
    np.testing.assert_allclose(gpu.bulk_update(), lat.bulk(), rtol=1e-13, atol=1e-14)
    gm, gr = gpu.gauss_violation()
    om, orr = lat.gauss()
    assert abs(gm - om) < 1e-12 and abs(gr - orr) < 1e-12
    lat.flow(3, 0.02)
    gpu.gradient_flow(3, 0.02, u=T.U_HOT, e=T.E_HOT)
    np.testing.assert_allclose(gpu.get_links(), lat.U, atol=1e-12)
    gpu.close()


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
def test_zero_steps_change_nothing(nc, prec):
    g = T.Solver(nc, (5, 4, 6), prec)
    g.random_links(71, 0.4)
    g.random_efield(71, 0.3)
    U, E = g.get_links(), g.get_efield()
    g.dynamics_update(0, 0.05)
    g.langevin_update(0, 0.05, 0.1, 0.5, seed=1)
    g.gradient_flow(0, 0.02, u=T.U_HOT, e=T.E_HOT)
    assert g.evolve_ncs(0, 0, 0.05) == 0.0
    assert np.array_equal(bits(g.get_links()), bits(U))
    assert np.array_equal(bits(g.get_efield()), bits(E))
    g.cool_n_save(0, 0.02)
    assert np.array_equal(bits(g.get_links(T.U_NEW)), bits(U))
    g.close()


def test_argument_errors():
    for N, a in (((2, 5, 5), (1.0, 1.0, 1.0)), ((5, 5, 5), (1.0, 0.0, 1.0))):
        with pytest.raises(T.TcymError) as ex:
            T.Solver(3, N, a=a)
        assert ex.value.code == T.E_ARG
    # one SU(3) FP64 link field of 256^3 is 7.2 GiB: beyond the 32-bit buffer offsets
    with pytest.raises(T.TcymError) as ex:
        T.Solver(3, (256, 256, 256), T.FP64)
    assert ex.value.code == T.E_UNSUPPORTED
    g = T.Solver(2, (4, 4, 4))
    g.unit_links()
    g.zero_efield()
    for call in (lambda: g.dynamics_update(-1, 0.05),
                 lambda: g.dynamics_update(1, 0.05, u=99),
                 lambda: g.dynamics_update(1, 0.05, e=-7),
                 lambda: g.langevin_update(1, 0.05, 0.1, -0.5),
                 lambda: g.gradient_flow(-2, 0.02, u=T.U_HOT, e=T.E_HOT),
                 lambda: g.get_links(slot=6)):
        with pytest.raises(T.TcymError) as ex:
            call()
        assert ex.value.code == T.E_ARG
    # the context stays usable after refused calls
    g.dynamics_update(2, 0.05)
    assert g.unitarity_violation()[0] < 1e-14
    g.close()
