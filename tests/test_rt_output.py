"""SYSname_rt.data writer (write_rt_data_3d, src/io/write.f90:686-790): layout and round trip."""
import numpy as np

from salmon2_amd.rt import _fe, write_rt_data


def test_fortran_e_edit():
    assert _fe(6.51045729817e-4) == " 0.651045729817000E-003"
    assert _fe(-2.97090762249e-4) == "-0.297090762249000E-003"
    assert _fe(0.0) == " 0.000000000000000E+000"
    assert _fe(9.99999999999999999e-4) == " 0.100000000000000E-002"
    rng = np.random.default_rng(1)
    for x in rng.standard_normal(200) * 10.0 ** rng.integers(-30, 30, 200):
        s = _fe(float(x))
        assert len(s) == 23 and abs(float(s) - x) <= 5e-15 * abs(x)  # 15 digits of a mantissa in [0.1, 1)


def test_rt_data_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    nt, dt = 7, 0.16
    res = {k: rng.standard_normal((nt, 3)) * 1e-3 for k in ("ac_ext", "e_ext", "ac_tot", "e_tot", "curr")}
    p = tmp_path / "Si_rt.data"
    write_rt_data(p, dt, res)
    lines = p.read_text().splitlines()
    assert lines[0] == "# Real time calculation: "
    assert lines[5] == "# Jm: Matter current density (electrons)"
    assert lines[6].startswith("# 1:Time[a.u.] 2:Ac_ext_x[a.u.]") and lines[6].endswith("16:Jm_z[a.u.]")
    assert len(lines) == 7 + nt
    for i, ln in enumerate(lines[7:]):
        f = ln.split()
        assert len(f) == 16 and len(ln) == 16 + 15 * 24
        assert float(f[0]) == round((i + 1) * dt, 8)
        got = np.array([float(v) for v in f[1:]])
        ref = np.concatenate([res[k][i] for k in ("ac_ext", "e_ext", "ac_tot", "e_tot", "curr")])
        assert np.all(np.abs(got - ref) <= 5e-15 * np.abs(ref))
    # the reference's verification reads column 16 (index 15) as Jm_z
    assert float(lines[7 + nt - 1].split()[15]) == float(_fe(res["curr"][-1][2]))


def test_rt_data_polarized(tmp_path):
    """nspin = 2: the Jm_u / Jm_d header lines and columns 14-19 of spin = 'polarized'
    (write.f90:713-719, 744-751, 775-778)."""
    rng = np.random.default_rng(3)
    nt, dt = 5, 0.08
    res = {k: rng.standard_normal((nt, 3)) * 1e-3 for k in ("ac_ext", "e_ext", "ac_tot", "e_tot")}
    res["curr_s"] = rng.standard_normal((nt, 2, 3)) * 1e-4
    res["curr"] = res["curr_s"].sum(axis=1)
    p = tmp_path / "Si_rt.data"
    write_rt_data(p, dt, res)
    lines = p.read_text().splitlines()
    assert lines[5] == "# Jm_u: Matter current density for spin-up electrons"
    assert lines[6] == "# Jm_d: Matter current density for spin-down electrons"
    assert lines[7].endswith("14:Jm_u_x[a.u.] 15:Jm_u_y[a.u.] 16:Jm_u_z[a.u.] 17:Jm_d_x[a.u.] 18:Jm_d_y[a.u.] "
                             "19:Jm_d_z[a.u.]")
    assert len(lines) == 8 + nt
    for i, ln in enumerate(lines[8:]):
        f = [float(v) for v in ln.split()]
        assert len(f) == 19
        ref = np.concatenate([res["curr_s"][i, 0], res["curr_s"][i, 1]])
        assert np.all(np.abs(np.array(f[13:]) - ref) <= 5e-15 * np.abs(ref))
    # one channel: the unpolarised layout
    res1 = dict(res, curr_s=res["curr_s"][:, :1])
    write_rt_data(p, dt, res1)
    assert len(p.read_text().splitlines()[7].split()) == 16  # first row after the 7 header lines
