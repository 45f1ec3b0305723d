"""Configuration files (include/tcym_io.h) on the host, no device needed.

The reference's text format (IO/SaveConfiguration.cpp:15-51,
IO/LoadConfiguration.cpp:46-122) is checked byte for byte against an
iostream restatement (oracle/or_text_io.cpp), which uses the reference's own
formatting mechanism (operator<< with precision 15) and parser (operator>>),
so the library's threaded to_chars / from_chars code is checked against an
independent implementation.  Binary checkpoints are checked by round trips,
their header and their corruption checks."""
import os
import random
import subprocess

import numpy as np
import pytest

import topologicalcymsolver_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_IO = os.path.join(ROOT, "oracle", "_build", "or_text_io")
# values whose shortest forms differ in shape ("-0", "1e-05", "1e+23", ...); the
# ends of the double range are left out: written with 15 digits they round to
# values outside the normal range, which no stream parser reads back as written
SPECIALS = [0.0, -0.0, 1.0, -1.0, 0.1, 1e-5, 1e-4, 123456789012345678.0, 9.9999999999999999e22,
            1.7e308, 2.3e-308, 0.99999999999999995, 1e15, 1e16, 123.456,
            -0.000123456789012345678, 0.5, 1.0 / 3.0]


def sample(nc, N, kind, seed):
    rng = np.random.default_rng(seed)
    n = 3 * N[0] * N[1] * N[2] * T._record_size(nc, kind)
    A = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n).astype(float)
    A[:len(SPECIALS)] = SPECIALS
    return A


def bits(A):
    return np.ascontiguousarray(A, dtype=np.float64).view(np.uint64)


def oracle(*args):
    r = subprocess.run([ORACLE_IO] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_oracle_is_built():
    assert os.path.exists(ORACLE_IO), "run make -C oracle"


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("kind", ["links", "efield"])
@pytest.mark.parametrize("digits", [15, 17])
def test_writer_matches_iostream(nc, kind, digits, tmp_path):
    N = (3, 4, 5)
    A = sample(nc, N, kind, 11 + nc + digits)
    A.tofile(tmp_path / "a.bin")
    oracle("write", nc, *N, digits, kind, tmp_path / "a.bin", tmp_path / "ref.txt")
    T.write_text(kind, nc, N, A, tmp_path / "lib.txt", digits=digits, nthreads=3)
    ref = (tmp_path / "ref.txt").read_bytes()
    assert (tmp_path / "lib.txt").read_bytes() == ref
    # the reference's line shape: links end with a space, fields do not
    first = ref.split(b"\n")[0]
    assert first.startswith(b"0 0 0 0 ") and first.endswith(b" ") == (kind == "links")
    assert len(first.split()) == 4 + T._record_size(nc, kind)


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("kind", ["links", "efield"])
def test_reader_matches_iostream(nc, kind, tmp_path):
    N = (4, 3, 5)
    A = sample(nc, N, kind, 23 + nc)
    A.tofile(tmp_path / "a.bin")
    oracle("write", nc, *N, 15, kind, tmp_path / "a.bin", tmp_path / "ref.txt")
    oracle("read", nc, *N, kind, tmp_path / "ref.txt", tmp_path / "back.bin")
    want = np.fromfile(tmp_path / "back.bin")
    got = T.read_text(kind, nc, N, tmp_path / "ref.txt", nthreads=2)
    assert np.array_equal(bits(got), bits(want))


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("kind", ["links", "efield"])
def test_17_digits_round_trip_exactly_and_15_digits_round_like_printf(nc, kind, tmp_path):
    N = (3, 3, 4)
    A = sample(nc, N, kind, 31 + nc)
    T.write_text(kind, nc, N, A, tmp_path / "f17.txt", digits=17)
    assert np.array_equal(bits(T.read_text(kind, nc, N, tmp_path / "f17.txt")), bits(A))
    T.write_text(kind, nc, N, A, tmp_path / "f15.txt", digits=15)
    want = np.array([float("%.15g" % v) for v in A])
    assert np.array_equal(bits(T.read_text(kind, nc, N, tmp_path / "f15.txt")), bits(want))


def test_large_file_over_many_chunks_and_threads(tmp_path):
    """73 728 records: several 32 768-record chunks per thread batch when
    writing, several byte ranges when reading."""
    nc, kind, N = 2, "efield", (32, 32, 24)
    A = sample(nc, N, kind, 5)
    A.tofile(tmp_path / "a.bin")
    oracle("write", nc, *N, 15, kind, tmp_path / "a.bin", tmp_path / "ref.txt")
    for t in (1, 2, 5):
        T.write_text(kind, nc, N, A, tmp_path / ("lib%d.txt" % t), nthreads=t)
        assert (tmp_path / ("lib%d.txt" % t)).read_bytes() == (tmp_path / "ref.txt").read_bytes()
    B = [T.read_text(kind, nc, N, tmp_path / "ref.txt", nthreads=t) for t in (1, 3, 16)]
    assert all(np.array_equal(bits(b), bits(B[0])) for b in B)


def test_reader_places_lines_by_coordinates(tmp_path):
    """Any line order, blank lines, CRLF line ends and '+' signs read as the
    reference's stream parser reads them."""
    nc, kind, N = 3, "links", (3, 4, 3)
    A = sample(nc, N, kind, 41)
    T.write_text(kind, nc, N, A, tmp_path / "a.txt", digits=17)
    lines = (tmp_path / "a.txt").read_text().splitlines()
    random.Random(3).shuffle(lines)
    out = []
    for k, ln in enumerate(lines):
        if k % 5 == 0:
            out.append("")
        if k % 7 == 0:
            ln = " ".join(("+" + t if not t.startswith("-") else t) for t in ln.split())
        out.append(ln + ("\r" if k % 3 == 0 else ""))
    (tmp_path / "b.txt").write_text("\n".join(out) + "\n\n")
    assert np.array_equal(bits(T.read_text(kind, nc, N, tmp_path / "b.txt", nthreads=4)), bits(A))


def _mutate(lines, case):
    if case == "missing":
        return lines[:-1]
    if case == "repeated":
        return lines[:-1] + [lines[0]]
    if case == "outside":
        t = lines[5].split()
        t[1] = "7"
        return lines[:5] + [" ".join(t)] + lines[6:]
    if case == "few_values":
        return lines[:3] + [" ".join(lines[3].split()[:-1])] + lines[4:]
    if case == "many_values":
        return lines[:3] + [lines[3] + " 1.5"] + lines[4:]
    if case == "garbage":
        t = lines[2].split()
        t[6] = "1.2.3"
        return lines[:2] + [" ".join(t)] + lines[3:]
    if case == "bad_mu":
        t = lines[2].split()
        t[3] = "3"
        return lines[:2] + [" ".join(t)] + lines[3:]
    if case == "empty":
        return []
    raise ValueError(case)


@pytest.mark.parametrize("case", ["missing", "repeated", "outside", "few_values", "many_values", "garbage",
                                  "bad_mu", "empty"])
def test_reader_rejects_malformed_files(case, tmp_path):
    nc, kind, N = 2, "links", (3, 4, 5)
    T.write_text(kind, nc, N, sample(nc, N, kind, 1), tmp_path / "a.txt")
    lines = (tmp_path / "a.txt").read_text().splitlines()
    (tmp_path / "b.txt").write_text("\n".join(_mutate(lines, case)) + "\n")
    with pytest.raises(T.TcymError) as ex:
        T.read_text(kind, nc, N, tmp_path / "b.txt")
    assert ex.value.code == T.E_IO


def test_missing_file_and_bad_arguments(tmp_path):
    with pytest.raises(T.TcymError) as ex:
        T.read_text("efield", 2, (3, 3, 3), tmp_path / "nope.txt")
    assert ex.value.code == T.E_IO
    with pytest.raises(T.TcymError) as ex:
        T.write_text("efield", 2, (3, 3, 3), np.zeros(243), tmp_path / "x.txt", digits=0)
    assert ex.value.code == T.E_ARG
    with pytest.raises(T.TcymError) as ex:
        T.write_text("efield", 2, (3, 3, 3), np.zeros(243), tmp_path / "no_dir" / "x.txt")
    assert ex.value.code == T.E_IO


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_checkpoint_round_trip(nc, precision, tmp_path):
    N = (4, 3, 5)
    U, E = sample(nc, N, "links", 7), sample(nc, N, "efield", 8)
    U[np.abs(U) > 1e30] = 1.0   # representable in FP32
    E[np.abs(E) > 1e30] = 1.0
    info = dict(nc=nc, precision=precision, N=N, a=(1.0, 0.5, 2.0), aScale=0.7, step=123456789012, seed=2 ** 63 + 5,
                time=12.5, acc=(1.0, -2.0, 3.5))
    T.write_checkpoint(tmp_path / "c.bin", U, E, **info)
    h, U2, E2 = T.read_checkpoint(tmp_path / "c.bin")
    assert h["nc"] == nc and h["precision"] == precision and h["N"] == N and h["a"] == (1.0, 0.5, 2.0)
    assert h["aScale"] == 0.7 and h["step"] == 123456789012 and h["seed"] == 2 ** 63 + 5 and h["time"] == 12.5
    assert h["acc"] == (1.0, -2.0, 3.5, 0.0, 0.0, 0.0, 0.0, 0.0) and h["has_efield"]
    if precision == T.FP64:
        assert np.array_equal(bits(U2), bits(U)) and np.array_equal(bits(E2), bits(E))
    else:
        assert np.array_equal(bits(U2), bits(U.astype(np.float32).astype(np.float64)))
        assert np.array_equal(bits(E2), bits(E.astype(np.float32).astype(np.float64)))
    size = 512 + (U.size + E.size) * (8 if precision == T.FP64 else 4)
    assert os.path.getsize(tmp_path / "c.bin") == size
    # links only
    T.write_checkpoint(tmp_path / "u.bin", U, None, **info)
    h, U3, E3 = T.read_checkpoint(tmp_path / "u.bin")
    assert not h["has_efield"] and E3 is None and np.array_equal(bits(U3), bits(U2))


@pytest.mark.parametrize("where", ["payload_links", "payload_efield", "header", "magic", "truncated"])
def test_checkpoint_detects_corruption(where, tmp_path):
    nc, N = 2, (3, 3, 3)
    U, E = sample(nc, N, "links", 9), sample(nc, N, "efield", 10)
    T.write_checkpoint(tmp_path / "c.bin", U, E, nc=nc, precision=T.FP64, N=N)
    b = bytearray((tmp_path / "c.bin").read_bytes())
    if where == "payload_links":
        b[512 + 100] ^= 0x10
    elif where == "payload_efield":
        b[512 + U.size * 8 + 17] ^= 0x01
    elif where == "header":
        b[88] ^= 0x01   # the time field
    elif where == "magic":
        b[0] = ord("X")
    else:
        b = b[:-8]
    (tmp_path / "c.bin").write_bytes(bytes(b))
    with pytest.raises(T.TcymError) as ex:
        T.read_checkpoint(tmp_path / "c.bin")
    assert ex.value.code == T.E_IO
