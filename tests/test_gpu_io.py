"""Configuration files through a device context (include/tcym_io.h): the
reference's text files written from and read into device slots, and binary
checkpoints from which a run resumes bit-identically."""
import numpy as np
import pytest

import topologicalcymsolver_amd as T

pytestmark = pytest.mark.gpu


def bits(A):
    return np.ascontiguousarray(A, dtype=np.float64).view(np.uint64)


def solver(nc, N, prec, seed=606):
    g = T.Solver(nc, N, prec)
    g.random_links(seed, 0.4)
    g.random_efield(seed, 0.3)
    return g


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
def test_text_files_from_and_into_slots(nc, prec, tmp_path):
    N = (6, 5, 4)
    g = solver(nc, N, prec)
    g.dynamics_update(3, 0.05)
    U, E = g.get_links(), g.get_efield()
    # 17 digits: exact round trip through the files
    g.save_configuration(tmp_path / "U17.txt", tmp_path / "E17.txt", digits=17)
    h = T.Solver(nc, N, prec)
    h.load_configuration(tmp_path / "U17.txt", tmp_path / "E17.txt", u=T.U_AUX, e=T.E_AUX)
    assert np.array_equal(bits(h.get_links(T.U_AUX)), bits(U))
    assert np.array_equal(bits(h.get_efield(T.E_AUX)), bits(E))
    # the reference's 15 digits: the file is what the host writer makes of the same arrays
    g.save_configuration(tmp_path / "U15.txt", tmp_path / "E15.txt")
    T.write_text("links", nc, N, U, tmp_path / "Uh.txt")
    T.write_text("efield", nc, N, E, tmp_path / "Eh.txt")
    assert (tmp_path / "U15.txt").read_bytes() == (tmp_path / "Uh.txt").read_bytes()
    assert (tmp_path / "E15.txt").read_bytes() == (tmp_path / "Eh.txt").read_bytes()
    h.load_configuration(tmp_path / "U15.txt", None)
    want = T.read_text("links", nc, N, tmp_path / "U15.txt")
    if prec == T.FP32:
        want = want.astype(np.float32).astype(np.float64)
    assert np.array_equal(bits(h.get_links()), bits(want))
    g.close()
    h.close()


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
def test_checkpoint_resume_is_bit_identical(nc, prec, tmp_path):
    """Leapfrog and Langevin continued from a checkpoint equal the run that
    never stopped; the Philox step and key travel in the header."""
    N = (6, 4, 5)
    g = solver(nc, N, prec, seed=77)
    g.dynamics_update(4, 0.05)
    g.langevin_update(3, 0.05, 0.2, 0.4, seed=9, step0=0)
    g.save_checkpoint(tmp_path / "c.bin", step=3, seed=9, time=0.35, acc=(0.125,))
    g.langevin_update(4, 0.05, 0.2, 0.4, seed=9, step0=3)
    g.dynamics_update(3, 0.05)
    h = T.Solver(nc, N, prec)
    info = h.load_checkpoint(tmp_path / "c.bin")
    assert info["step"] == 3 and info["seed"] == 9 and info["time"] == 0.35 and info["acc"][0] == 0.125
    assert info["precision"] == prec and info["N"] == N and info["has_efield"]
    h.langevin_update(4, 0.05, 0.2, 0.4, seed=info["seed"], step0=info["step"])
    h.dynamics_update(3, 0.05)
    assert np.array_equal(bits(h.get_links()), bits(g.get_links()))
    assert np.array_equal(bits(h.get_efield()), bits(g.get_efield()))
    g.close()
    h.close()


def test_checkpoint_of_a_blocked_slot_and_mismatches(tmp_path):
    nc, N = 2, (8, 8, 4)
    g = solver(nc, N, T.FP64)
    g.copy_links(T.U_NEW, T.U_HOT)
    g.block(T.U_NEW)
    g.save_checkpoint(tmp_path / "coarse.bin", u=T.U_NEW, e=-1)
    info, U, E = T.read_checkpoint(tmp_path / "coarse.bin")
    assert info["N"] == (4, 4, 2) and info["a"] == (2.0, 2.0, 2.0) and E is None
    assert np.array_equal(bits(U), bits(g.get_links(T.U_NEW)))
    g.save_configuration(tmp_path / "Uc.txt", None, u=T.U_NEW, digits=17)
    assert np.array_equal(bits(T.read_text("links", nc, (4, 4, 2), tmp_path / "Uc.txt")), bits(U))
    # a checkpoint of other extents, or without E when E is asked for, is refused
    with pytest.raises(T.TcymError) as ex:
        g.load_checkpoint(tmp_path / "coarse.bin", e=-1)
    assert ex.value.code == T.E_ARG
    g.save_checkpoint(tmp_path / "links.bin", e=-1)
    with pytest.raises(T.TcymError) as ex:
        g.load_checkpoint(tmp_path / "links.bin", e=T.E_AUX)
    assert ex.value.code == T.E_IO
    g.load_checkpoint(tmp_path / "links.bin", u=T.U_AUX, e=-1)
    assert np.array_equal(bits(g.get_links(T.U_AUX)), bits(g.get_links()))
    g.close()
