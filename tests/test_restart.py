"""SALMON restart data (salmon2_amd/restart.py) against files written by a Fortran compiler
with the I/O statements of src/io/checkpoint_restart.f90 (tests/golden/write_restart.f90,
compiled here with the image's flang): the reader recovers the closed-form values the
program wrote, and the writer reproduces its files byte for byte."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from salmon2_amd import restart as R

N = (5, 4, 3)
NO, NK = 3, 2
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "write_restart.f90")


def _flang():
    from salmon2_amd import build

    fc = build.flang()
    return fc if shutil.which(fc) or os.path.exists(fc) else None


@pytest.fixture(scope="module")
def fortran_dirs(tmp_path_factory):
    fc = _flang()
    if fc is None:
        pytest.skip("no Fortran compiler in this image")
    d = tmp_path_factory.mktemp("restart")
    exe = str(d / "write_restart")
    r = subprocess.run([fc, "-O1", "-o", exe, GOLDEN], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip(f"flang could not build the writer: {r.stderr[-300:]}")
    out = {}
    for mode in ("complex", "real"):
        for nspin in (1, 2):
            p = d / f"{mode}{nspin}"
            p.mkdir()
            subprocess.run([exe, str(p), mode, str(nspin)], check=True)
            out[mode, nspin] = str(p)
    return out


def _closed_form(real, nspin=1):
    nx, ny, nz = N
    ik, io, isp, iz, iy, ix = np.meshgrid(np.arange(1, NK + 1), np.arange(1, NO + 1), np.arange(1, nspin + 1),
                                          np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                                          indexing="ij")
    if real:
        psi = ix - 2 * iy + 3 * iz + 0.125 * io - 7 * ik + 0j
    else:
        psi = (ix + 10 * iy + 100 * iz + 0.5 * isp) + 1j * (io + 10 * ik + 0.25 * (ix - iy))
    rocc = np.array([[[2.0 - 0.1 * io - 0.01 * ik - 0.001 * (s - 1) for io in range(1, NO + 1)]
                      for ik in range(1, NK + 1)] for s in range(1, nspin + 1)])
    z, y, x = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    vh1 = 0.5 * x - y + 0.25 * z
    return psi, rocc, (vh1, -vh1 + 1.0 / 3.0)


@pytest.mark.parametrize("nspin", [1, 2])
@pytest.mark.parametrize("mode", ["complex", "real"])
def test_reads_fortran_written_restart(fortran_dirs, mode, nspin):
    d = R.read_restart(fortran_dirs[mode, nspin], N, nspin=nspin, ly=N[1] + 1)
    psi, rocc, vh = _closed_form(mode == "real", nspin)
    assert d["iter"] == 1234 and d["nprocs"] == 4 and d["if_real_orbital"] == (mode == "real")
    assert (d["mk"], d["mo"]) == (NK, NO)
    assert np.array_equal(d["psi"][..., : N[1], :], psi)
    assert not d["psi"][..., N[1], :].any()  # the padding row stays zero
    assert np.array_equal(d["rocc"], rocc)
    v1, v2 = R.read_vh_stock(fortran_dirs[mode, nspin], N)
    assert np.array_equal(v1, vh[0]) and np.array_equal(v2, vh[1])
    ac_tot, ac_ind, ac_ind0 = R.read_rtdata(fortran_dirs[mode, nspin])
    ref_tot, ref_ind = _rtdata_closed_form()
    assert np.array_equal(ac_tot, ref_tot[1234:1236]) and np.array_equal(ac_ind, ref_ind[1234:1236])
    assert np.array_equal(ac_ind0, ref_ind[0])


def _rtdata_closed_form():
    t = np.arange(0, 1237, dtype=np.float64)[:, None]
    d = np.arange(1, 4, dtype=np.float64)[None, :]
    return 1e-3 * t + 0.1 * d, -2e-4 * t + 0.01 * d * d


@pytest.mark.parametrize("nspin", [1, 2])
@pytest.mark.parametrize("mode", ["complex", "real"])
def test_writer_reproduces_fortran_bytes(fortran_dirs, mode, nspin, tmp_path):
    psi, rocc, vh = _closed_form(mode == "real", nspin)
    R.write_checkpoint(str(tmp_path), psi, N[1], rocc, 1234, nprocs=4, if_real_orbital=(mode == "real"), vh_stock=vh)
    tot, ind = _rtdata_closed_form()
    R.write_rtdata(str(tmp_path), tot[1234:1236], ind[1234:1236], ind[0])
    for name in ("info.bin", "occupation.bin", "wfn.bin", "Vh_stock.bin", "rtdata.bin"):
        a = open(os.path.join(fortran_dirs[mode, nspin], name), "rb").read()
        b = open(os.path.join(tmp_path, name), "rb").read()
        assert a == b, name


def test_subset_and_sliced_round_trip(tmp_path):
    rng = np.random.default_rng(9)
    nk, no, ny = 3, 5, 6
    psi = rng.standard_normal((nk, no, 1, 4, ny + 1, 7)) + 1j * rng.standard_normal((nk, no, 1, 4, ny + 1, 7))
    psi[..., ny, :] = 0.0
    rocc = rng.random((1, nk, no))
    R.write_checkpoint(str(tmp_path / "s"), psi, ny, rocc, 7, sliced_block=2)
    names = sorted(os.listdir(tmp_path / "s"))
    assert "k_000001_ob_000001" in names and "k_000003_ob_000005" in names and "wfn.bin" not in names
    assert os.path.exists(tmp_path / "s" / "k_000002_ob_000003" / "wfn_ob_000004.dat")
    d = R.read_restart(str(tmp_path / "s"), (7, ny, 4), ly=ny + 1)
    assert np.array_equal(d["psi"], psi) and np.array_equal(d["rocc"], rocc) and d["iter"] == 7
    R.write_checkpoint(str(tmp_path / "w"), psi, ny, rocc, 7)
    d = R.read_restart(str(tmp_path / "w"), (7, ny, 4), no=2, nk=1, ly=ny + 1)
    assert np.array_equal(d["psi"], psi[:1, :2]) and np.array_equal(d["rocc"], rocc[:, :1, :2])
    with pytest.raises(ValueError, match="asked for"):
        R.read_restart(str(tmp_path / "w"), (7, ny, 4), no=6)


def test_long_records_split_into_subrecords(tmp_path, monkeypatch):
    monkeypatch.setattr(R, "_MAXREC", 16)
    payload = bytes(range(40))
    with open(tmp_path / "f.bin", "wb") as f:
        R.write_record(f, payload)
        R.write_record(f, b"abcd")
    raw = open(tmp_path / "f.bin", "rb").read()
    marks = np.frombuffer(raw[:4], "<i4")[0], np.frombuffer(raw[20:24], "<i4")[0]
    assert marks == (-16, -16)  # continued sub-records carry negative markers
    assert R.read_records(str(tmp_path / "f.bin")) == [payload, b"abcd"]


def test_rejects_truncated_files(tmp_path):
    psi = np.ones((1, 2, 1, 2, 3, 2), dtype=np.complex128)
    R.write_checkpoint(str(tmp_path), psi, 3, np.ones((1, 1, 2)), 1)
    with open(tmp_path / "wfn.bin", "r+b") as f:
        f.truncate(16)
    with pytest.raises(ValueError, match="bytes, expected"):
        R.read_restart(str(tmp_path), (2, 3, 2))
