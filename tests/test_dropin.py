"""The C++ drop-in (include/tcym_dropin.hpp): a program written against the
reference's Dynamics::Update / NCsDot / Bulk / HardScales / GaussLaw /
Unitarity entry points over GaugeLinks/ElectricFields runs on the GPU and
matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

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
    r = subprocess.run([EXE, str(nc), str(N), str(tmp_path / "U.bin"), str(tmp_path / "E.bin"), str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ncs, t00, gmax, umax, txx, tzz, lxx, lzz = map(float, r.stdout.split())
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
