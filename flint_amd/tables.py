"""Table generator: species thermo and rate-constant tables on a 1 K grid.

FLINT never evaluates NASA polynomials or Arrhenius expressions at run time; every property
is linearly interpolated from a 1 K table (Lib_Chemistry_data.f90:36-48,
Lib_ThermoTransport.f90:516-542).  The tables are produced offline by a generator that is not
part of FLINT; this module is this project's generator.  Its conventions were fixed by matching
the tables FLINT does ship:

* gas constant 8314.46261815324 J/(kmol K) and 4184 J/kcal (test/WD/INPUT/chemistry-Arrhenius.dat
  is reproduced to all 13 printed digits);
* per-mass cp, h, s and dcp/dT from NASA-7 or NASA-9 polynomials; below the lowest polynomial
  temperature cp is held constant, h grows linearly and s logarithmically
  (test/Pelucchi/INPUT/thermo.dat);
* kf in kmol/m^3/s units without the third-body concentration, kb = kf/Kc for reversible
  reactions and 0 for irreversible ones, Kc at the 1 atm standard state;
* falloff tables hold k_inf, k_0, k_c = Kc (capped at 1e300) and the Troe F_cent
  (test/Pelucchi/INPUT/chemistry-{Troe,Lindemann}.dat).

Tables are returned in FLINT's own orientation, ``tab[T - Tmin, column]`` with one column per
species or per reaction of the given kind.
"""
import json
import os

import numpy as np

R_GEN = 8314.46261815324     # J/(kmol K), generator gas constant
P_STD = 101325.0             # Pa, standard-state pressure of the NASA fits
KC_CAP = 1.0e300

_MECH_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mechanisms")


def load_descriptor(name):
    with open(os.path.join(_MECH_DIR, name + ".json")) as f:
        return json.load(f)


# ----------------------------------------------------------------------------------------------
# NASA polynomials (dimensionless cp/R, h/RT, s/R)
# ----------------------------------------------------------------------------------------------
def _nasa7(a, T):
    cp = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])))
    h = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T
    s = a[0] * np.log(T) + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6]
    dcp = a[1] + T * (2 * a[2] + T * (3 * a[3] + T * 4 * a[4]))
    return cp, h, s, dcp


def _nasa9(a, T):
    lnT = np.log(T)
    cp = a[0] / T ** 2 + a[1] / T + a[2] + T * (a[3] + T * (a[4] + T * (a[5] + T * a[6])))
    h = (-a[0] / T ** 2 + a[1] * lnT / T + a[2] + T * (a[3] / 2 + T * (a[4] / 3 + T * (a[5] / 4 + T * a[6] / 5)))
         + a[7] / T)
    s = (-a[0] / (2 * T ** 2) - a[1] / T + a[2] * lnT + T * (a[3] + T * (a[4] / 2 + T * (a[5] / 3 + T * a[6] / 4)))
         + a[8])
    dcp = -2 * a[0] / T ** 3 - a[1] / T ** 2 + a[3] + T * (2 * a[4] + T * (3 * a[5] + T * 4 * a[6]))
    return cp, h, s, dcp


def _species_nondim(th, T, extend=True):
    """cp/R, h/RT, s/R, d(cp/R)/dT of one species on the array T.

    With ``extend`` the fit is continued outside its range with constant cp (the thermo tables);
    without it the end polynomials are evaluated as they are (the equilibrium constants).
    """
    T = np.asarray(T, dtype=np.float64)
    ranges = th["ranges"]
    f = _nasa7 if th["model"] == "NASA7" else _nasa9
    tlo, thi = ranges[0], ranges[-1]
    Tc = np.clip(T, tlo, thi) if extend else T
    cp = np.empty_like(T)
    h = np.empty_like(T)
    s = np.empty_like(T)
    dcp = np.empty_like(T)
    nreg = len(th["data"])
    for i in range(nreg):
        lo, hi = ranges[i], ranges[i + 1]
        m = np.ones(T.shape, dtype=bool)
        if i > 0:
            m &= Tc > lo
        if i < nreg - 1:
            m &= Tc <= hi
        if not np.any(m):
            continue
        c, hh, ss, d = f(th["data"][i], Tc[m])
        cp[m], h[m], s[m], dcp[m] = c, hh, ss, d
    # constant cp outside the fitted range: h linear, s logarithmic
    if extend:
        dcp[(T <= tlo) | (T >= thi)] = 0.0
    out = T != Tc
    if np.any(out):
        h[out] = (h[out] * Tc[out] + cp[out] * (T[out] - Tc[out])) / T[out]
        s[out] = s[out] + cp[out] * np.log(T[out] / Tc[out])
    return cp, h, s, dcp


def thermo_tables(desc, Tmin, Tmax):
    """Per-mass cp [J/kg/K], h [J/kg], s [J/kg/K], dcp/dT on T = Tmin..Tmax (rows)."""
    T = np.arange(Tmin, Tmax + 1, dtype=np.float64)
    ns = len(desc["species"])
    cp = np.empty((T.size, ns))
    h = np.empty_like(cp)
    s = np.empty_like(cp)
    dcp = np.empty_like(cp)
    for k, th in enumerate(desc["thermo"]):
        c, hh, ss, d = _species_nondim(th, T)
        Rk = R_GEN / desc["W"][k]
        cp[:, k] = c * Rk
        h[:, k] = hh * Rk * T
        s[:, k] = ss * Rk
        dcp[:, k] = d * Rk
    return dict(T=T, cp=cp, h=h, s=s, dcp=dcp)


# ----------------------------------------------------------------------------------------------
# Rate constants
# ----------------------------------------------------------------------------------------------
def _arrhenius(p, T):
    with np.errstate(over="ignore", under="ignore"):
        return p["A"] * T ** p["b"] * np.exp(-p["Ea"] / (R_GEN * T))


def _kc(desc, r, T):
    """Equilibrium constant in kmol/m^3 units at 1 atm standard state."""
    ns = len(desc["species"])
    g = np.zeros(T.size)
    dn = 0.0
    for k, v in r["nu_p"].items():
        c, h, s, _ = _species_nondim(desc["thermo"][int(k)], T, extend=False)
        g += v * (h - s)
        dn += v
    for k, v in r["nu_r"].items():
        c, h, s, _ = _species_nondim(desc["thermo"][int(k)], T, extend=False)
        g -= v * (h - s)
        dn -= v
    with np.errstate(over="ignore", under="ignore"):
        kc = np.exp(-g) * (P_STD / (R_GEN * T)) ** dn
    return np.minimum(kc, KC_CAP)


def _fcent(tp, T):
    with np.errstate(over="ignore", under="ignore"):
        f = (1.0 - tp["A"]) * np.exp(-T / tp["T3"]) + tp["A"] * np.exp(-T / tp["T1"])
        if tp["T2"] is not None:
            f = f + np.exp(-tp["T2"] / T)
    return f


def kinetics_tables(desc, Tmin, Tmax):
    """Rate tables on T = Tmin..Tmax.

    Returns kf, kb [nT, n_arrh] for the Arrhenius and three-body reactions (in mechanism order),
    troe [nT, n_troe, 4] = (k_inf, k_0, k_c, F_cent) and lind [nT, n_lind, 3] = (k_inf, k_0, k_c).
    """
    T = np.arange(Tmin, Tmax + 1, dtype=np.float64)
    kf, kb, troe, lind = [], [], [], []
    for r in desc["reactions"]:
        if r["type"] in ("arrhenius", "three-body"):
            f = _arrhenius(r["rate"], T)
            kf.append(f)
            if r["reversible"]:
                kc = _kc(desc, r, T)
                with np.errstate(divide="ignore", invalid="ignore", over="ignore", under="ignore"):
                    kb.append(np.where(kc > 0, f / kc, 0.0))
            else:
                kb.append(np.zeros_like(T))
        else:
            kinf = _arrhenius(r["high"], T)
            k0 = _arrhenius(r["low"], T)
            kc = _kc(desc, r, T) if r["reversible"] else np.full_like(T, KC_CAP)
            if r["type"] == "troe":
                troe.append(np.stack([kinf, k0, kc, _fcent(r["troe"], T)], axis=1))
            else:
                lind.append(np.stack([kinf, k0, kc], axis=1))
    out = dict(T=T, kf=np.stack(kf, axis=1), kb=np.stack(kb, axis=1))
    out["troe"] = np.stack(troe, axis=1) if troe else np.zeros((T.size, 0, 4))
    out["lind"] = np.stack(lind, axis=1) if lind else np.zeros((T.size, 0, 3))
    return out


# ----------------------------------------------------------------------------------------------
# Tecplot-ASCII writer / reader for FLINT's native table files (docs/input_native.md)
# ----------------------------------------------------------------------------------------------
def write_tecplot(path, title, varnames, zone_names, T, blocks, fmt="%.12E"):
    """Write one zone per column block: blocks is [nzone][nT, nvar]."""
    with open(path, "w") as f:
        f.write('TITLE = "%s"\n' % title)
        f.write("VARIABLES = " + ", ".join('"%s"' % v for v in varnames) + "\n")
        for name, blk in zip(zone_names, blocks):
            f.write('ZONE T="%s"\nI=%d, F=POINT\n' % (name, T.size))
            data = np.column_stack([T, blk])
            np.savetxt(f, data, fmt=["%12.3f"] + [fmt] * blk.shape[1])


def read_tecplot(path):
    """Return (T, [zone arrays nT x nvar]) for a POINT-format Tecplot ASCII file."""
    zones = []
    cur = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            if s.startswith(("TITLE", "VARIABLES", "I=")):
                continue
            if s.startswith("ZONE"):
                if cur:
                    zones.append(np.array(cur))
                cur = []
                continue
            cur.append([float(v) for v in s.split()])
    if cur:
        zones.append(np.array(cur))
    T = zones[0][:, 0]
    return T, [z[:, 1:] for z in zones]


# ----------------------------------------------------------------------------------------------
# Native input folder (docs/input_native.md): what read_idealgas_thermo and read_chemistry read
# ----------------------------------------------------------------------------------------------
_KIND = {"arrhenius": "Arrhenius", "three-body": "three-body-Arrhenius", "troe": "falloff-Troe",
         "lindemann": "falloff-Lindemann"}


def write_native(name, folder, prefix="", fmt="%.16E", transport=None):
    """Write the native input folder of mechanism ``name`` (phase.txt, composition.txt,
    chemistry-info.txt, thermo.dat and chemistry-{Arrhenius,Troe,Lindemann}.dat) in the layout
    FLINT's loaders read (Load_ThermoTransport.f90:7-126, Load_Chemistry.f90:7-164).

    The reference ships such folders with 13 significant digits; the default ``fmt`` prints 17
    so that a folder written here loads into exactly the tables generated in memory.  The
    chemistry-info.txt carries the Reaction definition block of the data-driven ``general``
    routine (Load_Chemistry.f90:55-89), from :func:`flint_amd.mechanism.general_tables`.
    ``transport=(mu_tab, k_tab)`` ([ns, nT] on the thermo rows, as
    :func:`flint_amd.transport.species_tables` returns them) also writes transport.dat, one zone
    per species with (T, mu, k), the layout read_idealgas_transport reads
    (Load_ThermoTransport.f90:129-174)."""
    desc = load_descriptor(name)
    Tmin, Tmax = desc["T_range"]
    os.makedirs(folder, exist_ok=True)
    path = lambda f: os.path.join(folder, prefix + f)
    sp, W = desc["species"], desc["W"]
    with open(path("phase.txt"), "w") as f:
        f.write("ideal-gas phase\n")
        for s, w in zip(sp, W):
            f.write("%s %f\n" % (s, w))
    el, comp = desc["elements"], np.asarray(desc["composition"])
    with open(path("composition.txt"), "w") as f:
        f.write("%d\n" % len(el))
        f.write("".join("%s\n" % e for e in el))
        f.write("\ncomposition\n")
        for k in range(len(sp)):
            f.write(" ".join("%d" % round(v) for v in comp[:, k]) + "\n")
    rx = desc["reactions"]
    with open(path("chemistry-info.txt"), "w") as f:
        f.write("%s\nN.ro species = %d\nN.ro reactions = %d\n\n" % (desc["name"], len(sp), len(rx)))
        f.write("General loop info. They are not used if the mechanism is exlplicitly defined\n\n")
        f.write("Reaction type\n")
        for j, r in enumerate(rx):
            f.write("%d %s\n" % (j + 1, _KIND[r["type"]]))
        f.write("\nReaction definition\n")
        from .mechanism import general_tables  # mechanism imports this module
        na, nt, nl, ni1, ni2, eps = general_tables(desc)
        kinds = [0 if r["type"] in ("arrhenius", "three-body") else (1 if r["type"] == "troe" else 2) for r in rx]
        row_of = {}
        for q in (0, 1, 2):
            for j in [j for j in range(len(rx)) if kinds[j] == q]:
                row_of[j] = len(row_of)
        names = list(sp) + ["M"]
        for j in range(len(rx)):
            a = row_of[j]
            for k, nm in enumerate(names):
                f.write("%d %s %r %r %r\n" % (j + 1, nm, float(ni1[a, k]), float(ni2[a, k]), float(eps[a, k])))
    th = thermo_tables(desc, Tmin, Tmax)
    T = th["T"]
    write_tecplot(path("thermo.dat"), "Mass Thermodynamic Properties",
                  ["Temperature", "Cp", "Enthalpy", "Entropy", "dCp"], sp, T,
                  [np.column_stack([th[v][:, k] for v in ("cp", "h", "s", "dcp")]) for k in range(len(sp))], fmt)
    if transport is not None:
        mu, kt = transport
        write_tecplot(path("transport.dat"), "Transport Properties", ["Temperature", "mu", "k"], sp, T,
                      [np.column_stack([mu[k], kt[k]]) for k in range(len(sp))], fmt)
    kin = kinetics_tables(desc, Tmin, Tmax)
    na = kin["kf"].shape[1]
    write_tecplot(path("chemistry-Arrhenius.dat"), "Arrhenius data", ["Temperature", "kf", "kb"],
                  ["Reaction %d" % (a + 1) for a in range(na)], T,
                  [np.column_stack([kin["kf"][:, a], kin["kb"][:, a]]) for a in range(na)], fmt)
    for key, title, var in (("troe", "Fall-off Troe data", ["k_inf", "k_0", "k_c", "F_cent"]),
                            ("lind", "Fall-off Lindemann data", ["k_inf", "k_0", "k_c"])):
        tab = kin[key]
        if tab.shape[1]:
            fname = "chemistry-Troe.dat" if key == "troe" else "chemistry-Lindemann.dat"
            write_tecplot(path(fname), title, ["Temperature"] + var,
                          ["Reaction %d" % (q + 1) for q in range(tab.shape[1])], T,
                          [tab[:, q, :] for q in range(tab.shape[1])], fmt)
    return folder
