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
