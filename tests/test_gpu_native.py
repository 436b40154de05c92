"""The native input folder and the Fortran drop-in, end to end on the GPU.

* flint_mech_load on a folder written by flint_amd.tables.write_native (FLINT's phase.txt,
  composition.txt, chemistry-info.txt and Tecplot tables, Load_ThermoTransport.f90:7-126,
  Load_Chemistry.f90:7-164) gives a mechanism whose source terms, Jacobians, equilibria and
  reactor steps are bit-identical to the one built from the same tables in memory;
* the Fortran drop-in (flint_amd/fortran/flint_amd.f90: FLINT's modules FLINT_Lib_Thermodynamic,
  FLINT_Load_ThermoTransport, FLINT_Load_Chemistry, FLINT_Lib_Chemistry_data,
  FLINT_Lib_Chemistry_wdot, FLINT_Lib_Chemistry_rhs, FLINT_CEA_setup, FLINT_CEA_solver), used by a
  program with FLINT's own use lines and call sequence, fills the module arrays a FLINT caller
  reads and deallocates, falls back to the general routine for an unknown mechanism name with
  FLINT's warning, and returns, per cell, exactly what the C ABI returns for the same cells."""
import os
import subprocess

import numpy as np
import pytest

from flint_amd import states as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "flint_amd", "fortran", "_build", "drop_in_check")
NAME = "TSR-Rich-31"


@pytest.fixture(scope="module")
def native(tmp_path_factory, mechs, cuda):
    from flint_amd.native import Mechanism
    from flint_amd.tables import write_native
    folder = str(tmp_path_factory.mktemp("native") / NAME)
    print(f"[gpu] write_native {NAME}", flush=True)
    from flint_amd.transport import species_tables
    md = mechs[NAME]
    tb = species_tables(md)
    write_native(NAME, folder, transport=(tb[2], tb[3]))
    mm = Mechanism.from_data(md)
    mm.transport_setup(tb)
    return folder, md, Mechanism.load(folder), mm


def _eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def test_loaded_folder_matches_in_memory_tables(native, cuda):
    import torch
    folder, md, ml, mm = native
    assert ml.ns == md.ns and ml.query() == mm.query()
    n = 2048
    st = S.radical_states(md.species, md.W, 91, 0, n)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    rho, T = dev(S.soa(st["rho"])), dev(st["T"])
    out = []
    for m in (ml, mm):
        w = torch.empty_like(rho)
        m.source(rho, T, w)
        Z = torch.cat([rho, T[None]], 0)
        F, J = torch.empty_like(Z), torch.empty((md.ns + 1) ** 2, n, dtype=torch.float64, device=cuda)
        m.rhs(Z, F, J)
        X = dev(np.vstack([S.soa(st["Y"]), st["T"][None]]))
        m.advance(X, dev(st["p"]), mode=1, t_end=1e-5)
        out.append([x.cpu().numpy() for x in (w, F, J, X)])
    torch.cuda.synchronize()
    print("[gpu] native folder vs in-memory tables", flush=True)
    for a, b in zip(*out):
        assert _eq(a, b)
    # the general routine's tables come with the folder (chemistry-info.txt's Reaction definition)
    mm.general_setup(md.general)
    rho_d, T_d = dev(S.soa(st["rho"])), dev(st["T"])
    wg = [torch.empty_like(rho_d) for _ in range(2)]
    ml.source_general(rho_d, T_d, wg[0])
    mm.source_general(rho_d, T_d, wg[1])
    torch.cuda.synchronize()
    assert torch.equal(wg[0], wg[1])
    cs = S.radical_states(md.species, md.W, 92, 0, 512, T_lo=600.0, T_hi=900.0)
    r1 = ml.equilibrium_host(cs["T"], S.soa(cs["Y"]), p=cs["p"] / 1e5, hp=True)
    r2 = mm.equilibrium_host(cs["T"], S.soa(cs["Y"]), p=cs["p"] / 1e5, hp=True)
    assert all(_eq(a, b) for a, b in zip(r1, r2))
    assert np.mean(r1[2] == 0) > 0.9


def _blocks(path):
    """'# name rows cols' numeric blocks and '# name count' blocks of text lines."""
    out, name, rows, text = {}, None, [], False
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                if name:
                    out[name] = rows if text else np.array(rows)
                head = line.split()
                name, text, rows = head[1], len(head) == 3, []
            else:
                rows.append(line.strip() if text else [float(v) for v in line.split()])
    out[name] = rows if text else np.array(rows)
    return out


def test_fortran_drop_in_matches_c_abi(native, tmp_path):
    folder, md, ml, mm = native
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: build with python -m flint_amd.build (needs flang)")
    ns, nz = md.ns, md.ns + 1
    n, m, k = 64, 32, 48
    st = S.radical_states(md.species, md.W, 93, 0, n)
    cs = S.radical_states(md.species, md.W, 94, 0, m, T_lo=600.0, T_hi=900.0)
    rs = S.radical_states(md.species, md.W, 95, 0, k, ignition_half=True)
    t_end, rtol, atol = 2e-5, 1e-6, 1e-12
    q, nd = 40, 3
    fs = S.flow_states(md.species, md.W, 96, 0, q, nd=nd)
    fin, fout = tmp_path / "states.txt", tmp_path / "results.txt"
    with open(fin, "w") as f:
        f.write(f"{n}\n")
        for i in range(n):
            f.write(" ".join("%.17e" % v for v in [st["T"][i], *st["rho"][i]]) + "\n")
        f.write(f"{m}\n")
        for i in range(m):
            f.write(" ".join("%.17e" % v for v in [cs["T"][i], cs["p"][i] / 1e5, *cs["Y"][i]]) + "\n")
        f.write(f"{k}\n")
        for i in range(k):
            f.write(" ".join("%.17e" % v for v in [rs["p"][i], *rs["Y"][i], rs["T"][i]]) + "\n")
        f.write("%.17e %.17e %.17e\n" % (t_end, rtol, atol))
        f.write(f"{q} {nd}\n")
        for i in range(q):
            f.write(" ".join("%.17e" % v for v in [fs["T0"][i], *fs["prim"][:, i]]) + "\n")
    print("[gpu] fortran drop_in_check", flush=True)
    r = subprocess.run([EXE, folder, str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _blocks(fout)
    # FLINT_Lib_Thermodynamic's arrays as the caller reads them: names, W, R_k, thermo tables over
    # Tmin:Tmax, composition, and the accessors over them
    assert got["species_names"] == list(md.species)
    assert _eq(got["wm_tab"], np.stack([md.W, 8314.51 / md.W], 1))
    r0 = lambda T0: T0 - md.Th_min
    th = got["thermo_rows"]
    assert _eq(th[:, 0], md.cp[0]) and _eq(th[:, 1], md.cp[r0(1000)]) and _eq(th[:, 2], md.cp[-1])
    assert _eq(th[:, 3], md.h[r0(1000)]) and _eq(th[:, 4], md.s[r0(1000)]) and np.all(np.isfinite(th[:, 5]))
    assert _eq(got["composition"], np.asarray(md.composition).T)
    rho, T = S.soa(st["rho"]), st["T"]
    Rk = 8314.51 / md.W
    Ti = T.astype(int)
    fr = (T - Ti)[:, None]
    cpT = md.cp[Ti - md.Th_min] + (md.cp[Ti + 1 - md.Th_min] - md.cp[Ti - md.Th_min]) * fr
    h1 = md.h[Ti - md.Th_min, 0] + (md.h[Ti + 1 - md.Th_min, 0] - md.h[Ti - md.Th_min, 0]) * fr[:, 0]
    ref = np.stack([(Rk[:, None] * rho / rho.sum(0)).sum(0), (rho / rho.sum(0) * cpT.T).sum(0), h1], 1)
    assert np.allclose(got["f_Rtot_f_cp_H0i"], ref, rtol=1e-14, atol=0)
    from flint_amd.elements import element_weights
    assert _eq(got["el_weight"][:, 0], element_weights(md.elements))
    # the same cells through the C ABI (host-memory entry points, in-memory tables)
    w = mm.source_host(rho, T)
    mm.general_setup(md.general)
    assert _eq(got["general"], mm.source_general_host(rho, T).T)
    assert "defaulting to the general procedure" in r.stdout
    F, J = mm.rhs_host(np.vstack([rho, T[None]]), jac=True)
    F1 = np.stack([mm.rhs_host(np.concatenate([rho[:, i], T[i:i + 1]])[:, None])[:, 0] for i in range(n)], 1)
    assert _eq(got["chemistry_source"], w.T) and _eq(got["chemistry_source_batch"], w.T)
    assert _eq(got["rhs_native"], F1.T) and _eq(got["rhs_native_batch"], F.T)
    assert _eq(got["rhs_native_batch_J"], J.T)
    Tu, Yu, su = mm.equilibrium_host(cs["T"], S.soa(cs["Y"]), hp=False)
    assert _eq(got["CEA_solve_uv"], np.vstack([Tu[None], Yu]).T)
    Th, Yh, sh = mm.equilibrium_host(cs["T"], S.soa(cs["Y"]), p=cs["p"] / 1e5, hp=True)
    assert _eq(got["CEA_solve_batch_hp"], np.vstack([Th[None], Yh, sh[None]]).T)
    X, sx = mm.advance_host(np.vstack([S.soa(rs["Y"]), rs["T"][None]]), rs["p"], mode=1, t_end=t_end, rtol=rtol,
                            atol=atol)
    assert _eq(got["reactor_advance_batch"], np.vstack([X, sx[None]]).T)
    assert np.mean(sx == 0) > 0.9 and np.mean(sh == 0) > 0.9
    cons = mm.prim2cons_host(fs["prim"])
    prim, Tq, sq = mm.cons2prim_host(cons, fs["T0"])
    assert _eq(got["prim2cons"], cons.T) and _eq(got["prim2cons_batch"], cons.T)
    assert _eq(got["cons2prim"], prim.T)
    assert _eq(got["cons2prim_batch"], np.vstack([prim, Tq[None], sq[None]]).T)
    assert np.all(sq == 0)
    r = fs["prim"][:ns]
    mu, lam, stt = mm.transport_host(r, r.sum(0), fs["T0"])
    assert _eq(got["co_k_mi_lam_Wilke"], np.vstack([mu, lam]).T)
    assert _eq(got["transport_batch"], np.vstack([mu, lam, stt]).T) and np.all(stt == 0)
