"""Table generator pinned against the tables FLINT ships (tests/golden/tables_ref.npz)."""
import os

import numpy as np
import pytest

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


def test_write_native_loads_back_exactly(tmp_path):
    """A native folder written for Pelucchi (Arrhenius, three-body, Troe and Lindemann reactions)
    parses back into exactly the generated tables, with FLINT's file names, headers and the
    (A17,I4) reaction-count line that read_chemistry reads (Load_Chemistry.f90:38)."""
    from flint_amd.mechdata import MechData
    folder = tmp_path / "pel"
    TB.write_native("Pelucchi", str(folder))
    md = MechData("Pelucchi")
    names = sorted(p.name for p in folder.iterdir())
    assert names == ["chemistry-Arrhenius.dat", "chemistry-Lindemann.dat", "chemistry-Troe.dat",
                     "chemistry-info.txt", "composition.txt", "phase.txt", "thermo.dat"]
    info = (folder / "chemistry-info.txt").read_text().splitlines()
    assert info[0] == "Pelucchi" and int(info[2][17:21]) == 103
    kinds = [l.split()[1] for l in info[7:7 + 103]]
    assert kinds.count("falloff-Troe") == md.n_troe and kinds.count("falloff-Lindemann") == md.n_lind
    T, zones = TB.read_tecplot(str(folder / "chemistry-Arrhenius.dat"))
    assert T[0] == md.Tk_min and len(zones) == md.n_arrh
    assert np.array_equal(np.stack([z[:, 0] for z in zones], 1), md.kf)
    assert np.array_equal(np.stack([z[:, 1] for z in zones], 1), md.kb)
    _, zones = TB.read_tecplot(str(folder / "chemistry-Troe.dat"))
    assert np.array_equal(np.stack(zones, 1), md.troe)
    _, zones = TB.read_tecplot(str(folder / "chemistry-Lindemann.dat"))
    assert np.array_equal(np.stack(zones, 1), md.lind)
    T, zones = TB.read_tecplot(str(folder / "thermo.dat"))
    assert T[0] == 1 and md.Th_min == 0  # the loader adds the T = 0 row (Load_ThermoTransport.f90:80-85)
    for j, tab in enumerate((md.cp, md.h, md.s)):
        assert np.array_equal(np.stack([z[:, j] for z in zones], 1), tab[1:])
    phase = (folder / "phase.txt").read_text().splitlines()
    assert phase[0] == "ideal-gas phase" and [l.split()[0] for l in phase[1:]] == md.species
    comp = (folder / "composition.txt").read_text().split()
    ne = int(comp[0])
    assert comp[1:1 + ne] == list(md.elements) and comp[1 + ne] == "composition"
    assert np.array_equal(np.array(comp[2 + ne:], dtype=float).reshape(md.ns, ne).T, md.composition)


@pytest.mark.skipif(not os.path.isdir("/root/reference/test"), reason="reference checkout absent")
@pytest.mark.parametrize("mech", ["WD", "TSR-GP-24", "TSR-Rich-31", "Pelucchi"])
def test_written_chemistry_info_equals_shipped_file(tmp_path, mech):
    """chemistry-info.txt written from the descriptor (with the general routine's Reaction
    definition block) is byte-identical to the file the reference ships."""
    from flint_amd.tables import write_native
    folder = write_native(mech, str(tmp_path / mech))
    with open(os.path.join(folder, "chemistry-info.txt")) as f:
        mine = f.read()
    with open(f"/root/reference/test/{mech}/INPUT/chemistry-info.txt") as f:
        ref = f.read()
    assert mine == ref
