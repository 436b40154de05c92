"""Table generator pinned against the tables FLINT ships (tests/golden/tables_ref.npz)."""
import numpy as np

from conftest import golden
from flint_amd import tables as TB


def test_wd_arrhenius_matches_shipped_table():
    g = golden("tables_ref.npz")
    d = TB.load_descriptor("WD")
    k = TB.kinetics_tables(d, 1, 15000)
    T = g["wd_T"]
    rows = (T - 1).astype(int)
    ref = g["wd_arrh"]  # [nrow, 3 reactions, (kf, kb)]
    kf = k["kf"][rows]
    m = ref[:, :, 0] > 1e-280
    rel = np.abs(kf[m] - ref[:, :, 0][m]) / ref[:, :, 0][m]
    assert rel.max() < 1e-12  # the file prints 13 significant digits
    assert np.all(k["kb"] == 0.0) and np.all(ref[:, :, 1] == 0.0)  # irreversible


def test_pelucchi_thermo_matches_shipped_table():
    g = golden("tables_ref.npz")
    d = TB.load_descriptor("Pelucchi")
    th = TB.thermo_tables(d, 1, 3000)
    rows = (g["pel_T"] - 1).astype(int)
    ref = g["pel_thermo"]  # [nrow, ns, (cp, h, s, dcp)]
    for j, key in enumerate(("cp", "h", "s")):
        # printed with 6 decimals: absolute 5e-7 plus 1e-12 relative
        err = np.abs(th[key][rows] - ref[:, :, j])
        assert np.all(err <= 5e-7 + 1e-12 * np.abs(ref[:, :, j])), key


def test_pelucchi_falloff_tables_match_shipped_tables():
    g = golden("tables_ref.npz")
    d = TB.load_descriptor("Pelucchi")
    k = TB.kinetics_tables(d, 1, 3000)
    rows = (g["pel_T"] - 1).astype(int)
    for key, ref in (("troe", g["pel_troe"]), ("lind", g["pel_lind"])):
        gen = k[key][rows]
        m = (np.abs(ref) > 1e-280) & (np.abs(ref) < 1e299)
        rel = np.abs(gen[m] - ref[m]) / np.abs(ref[m])
        assert rel.max() < 2e-12, key


def test_tecplot_roundtrip(tmp_path):
    d = TB.load_descriptor("TSR-GP-24")
    k = TB.kinetics_tables(d, 900, 1000)
    p = tmp_path / "chemistry-Arrhenius.dat"
    TB.write_tecplot(str(p), "Arrhenius data", ["Temperature", "kf", "kb"], [f"Reaction {i+1}" for i in range(3)],
                     k["T"], [np.stack([k["kf"][:, i], k["kb"][:, i]], 1) for i in range(3)], fmt="%.17E")
    T, z = TB.read_tecplot(str(p))
    assert np.array_equal(T, k["T"])
    assert np.array_equal(z[2][:, 0], k["kf"][:, 2])
