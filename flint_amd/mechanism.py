"""Mechanism compiler: turns a mechanism descriptor into the straight-line rate program that the
FLINT explicit routines evaluate, with their exact arithmetic.

FLINT's explicit routines (e.g. src/lib/Lib_ChemMech/TSR-GP-24.f90:4-415) are emitted by
utils/YTF.py.  Their arithmetic is fixed by three properties of that emitter which this module
reproduces from the descriptor alone:

* third-body sums ``M`` are written left to right over the species in phase order, with the
  efficiency printed as ``*{eff:.8g}`` and omitted when it is 1 (utils/YTF.py:29-34,125-145).
  A printed literal with a decimal point is a default-REAL (float32) Fortran constant, so the
  efficiency enters the FP64 sum rounded through float32;
  three-body reactions drop zero-efficiency species; falloff reactions default absent
  species to 1 (utils/YTF.py:150-163); an empty efficiency map becomes ``sum(coi(1:ns))``;
* rates of progress are ``k*(m_1)*(m_2)...[*M]`` with the reactant (product) factors in
  Cantera's alphabetical species order and integer powers expanded as repeated products
  (utils/YTF.py:20-27,100-123);
* the net production rate of species k is ``W_k*(+(nu''-nu')*(prodf(r)-prodb(r))+...)`` over the
  reactions that name k, in reaction order, with the coefficient a float32 difference of the two
  printed stoichiometric coefficients (utils/YTF.py:235-260).

The compiled program is consumed by the HIP code generator (flint_amd/codegen.py) and, as flat
arrays, by the CPU oracle.
"""
from dataclasses import dataclass, field

import numpy as np

from .tables import load_descriptor

RUNIV = 8314.51  # J/(kmol K), FLINT's gas constant (Lib_ThermoTransport.f90:13)


def f32(x):
    """Value of a printed default-REAL Fortran literal, promoted to real(8)."""
    return float(np.float32(x))


def fortran_literal(val):
    """The real(8) value the Fortran compiler gives to the text ``f'{val:.8g}'``."""
    s = f"{val:.8g}"
    if any(c in s for c in ".eE"):
        return f32(float(s))
    return float(int(s))


@dataclass
class Reaction:
    index: int                 # 0-based mechanism order
    kind: str                  # arrhenius | three-body | troe | lindemann
    table: int                 # 0-based column in its table family
    reversible: bool
    fwd: list                  # [(species, integer power)] in emitted order
    rev: list
    m_terms: object = None     # None (no M), "sum", or [(species, eff or None for bare)]
    m_set: int = -1            # index of the distinct M sum this reaction uses
    eq: str = ""


@dataclass
class Program:
    name: str
    species: list
    W: list
    ns: int
    reactions: list
    m_sets: list               # distinct M sums: "sum" or tuple of (species, eff|None)
    wdot_terms: list           # per species: [(reaction, coefficient)] in reaction order
    n_arrh: int = 0
    n_troe: int = 0
    n_lind: int = 0
    T_range: list = field(default_factory=list)
    desc: dict = field(default_factory=dict)
    handwritten: bool = False  # fractional orders/stoichiometry: needs a hand-written routine


def _alpha(species, stoich):
    """(species index, integer power) pairs in Cantera's std::map order (byte-wise names)."""
    items = sorted(((species[int(k)], int(k), v) for k, v in stoich.items()), key=lambda t: t[0])
    out = []
    for _, k, v in items:
        if abs(v - round(v)) > 1e-12:
            raise ValueError("non-integer stoichiometry needs a hand-written routine")
        n = int(round(v))
        if n > 0:
            out.append((k, n))
    return out


def compile_mechanism(name_or_desc):
    desc = load_descriptor(name_or_desc) if isinstance(name_or_desc, str) else name_or_desc
    species = desc["species"]
    ns = len(species)
    reactions = []
    m_sets = []
    handwritten = False
    counters = dict(arr=0, troe=0, lind=0)
    for j, r in enumerate(desc["reactions"]):
        kind = r["type"]
        fam = {"arrhenius": "arr", "three-body": "arr", "troe": "troe", "lindemann": "lind"}[kind]
        table = counters[fam]
        counters[fam] += 1
        m = None
        if kind == "three-body":
            eff = r["eff"]
            if all(abs(e - 1.0) < 1e-8 for e in eff) and not r.get("eff_listed", True):
                m = "sum"
            elif all(abs(e - 1.0) < 1e-8 for e in eff):
                m = "sum"
            else:
                m = []
                for k, e in enumerate(eff):
                    if f"{e:.8g}" == "0":
                        continue
                    m.append((k, None if abs(e - 1.0) < 1e-8 else fortran_literal(e)))
                m = tuple(m)
        elif kind in ("troe", "lindemann"):
            eff = r.get("eff_ytf")
            if eff is None:
                m = "sum"
            else:
                m = tuple((k, None if abs(e - 1.0) < 1e-8 else fortran_literal(e)) for k, e in enumerate(eff))
        try:
            fwd, rev = _alpha(species, r["nu_r"]), _alpha(species, r["nu_p"])
        except ValueError:
            fwd, rev = [], []   # fractional stoichiometry: only a hand-written routine (WD.f90)
            handwritten = True
        if "orders" in r:
            fwd, rev = [], []
            handwritten = True
        rx = Reaction(index=j, kind=kind, table=table, reversible=r["reversible"], fwd=fwd, rev=rev,
                      m_terms=m, eq=r["equation"])
        if m is not None:
            if m not in m_sets:
                m_sets.append(m)
            rx.m_set = m_sets.index(m)
        reactions.append(rx)
    wdot_terms = []
    for k in range(ns):
        terms = []
        for j, r in enumerate(desc["reactions"]):
            if str(k) in r["nu_r"] or str(k) in r["nu_p"]:
                nip = r["nu_p"].get(str(k), 0.0)
                nir = r["nu_r"].get(str(k), 0.0)
                coef = float(np.float32(np.float32(str(nip)) - np.float32(str(nir))))
                terms.append((j, coef))
        wdot_terms.append(terms)
    return Program(name=desc["name"], species=species, W=desc["W"], ns=ns, reactions=reactions,
                   m_sets=m_sets, wdot_terms=wdot_terms, n_arrh=counters["arr"],
                   n_troe=counters["troe"], n_lind=counters["lind"], T_range=desc["T_range"], desc=desc,
                   handwritten=handwritten)


# ----------------------------------------------------------------------------------------------
# Flat integer/float arrays for the C oracle (oracle/flint_oracle.c, struct orc_mech)
# ----------------------------------------------------------------------------------------------
KIND_CODE = {"arrhenius": 0, "three-body": 1, "troe": 2, "lindemann": 3}


def general_tables(desc):
    """Tables of FLINT's data-driven ``general`` routine (Lib_Chemistry_wdot.f90:96-221), as the
    Reaction definition block of chemistry-info.txt holds them (Load_Chemistry.f90:55-89).

    Returns (na, nt, nl, ni1, ni2, eps): for every reaction of each kind (Arrhenius and three-body,
    then Troe, then Lindemann, each in file order) ns+1 values of the reactant and product
    coefficients and the third-body efficiencies, the last entry being the third body M
    (ni1 = ni2 = 1 for three-body and fall-off reactions; its efficiency 0).  Efficiencies are the
    exact doubles of the mechanism file (not the float32-rounded literals of the generated
    routines).  Arrays are [na+nt+nl][ns+1], the memory order of FLINT's (ns+1, nrc) tables."""
    sp = desc["species"]
    ns = len(sp)
    rx = desc["reactions"]
    rows = [r for r in rx if r["type"] in ("arrhenius", "three-body")] + \
           [r for r in rx if r["type"] == "troe"] + [r for r in rx if r["type"] == "lindemann"]
    ni1 = np.zeros((len(rows), ns + 1))
    ni2 = np.zeros_like(ni1)
    eps = np.zeros_like(ni1)
    for j, r in enumerate(rows):
        for k, v in r["nu_r"].items():
            ni1[j, int(k)] = v
        for k, v in r["nu_p"].items():
            ni2[j, int(k)] = v
        if r["type"] != "arrhenius":
            ni1[j, ns] = ni2[j, ns] = 1.0
            eps[j, :ns] = r.get("eff_ytf") or r["eff"]
    na = sum(r["type"] in ("arrhenius", "three-body") for r in rows)
    nt = sum(r["type"] == "troe" for r in rows)
    return na, nt, len(rows) - na - nt, ni1, ni2, eps


def flatten(prog):
    """Encode the program as flat arrays (all int32 unless noted).

    rx_kind[nr], rx_table[nr], rx_mset[nr]
    rx_fwd_ptr[nr+1], rx_fwd_sp[], rx_fwd_pow[]  (same for rev)
    ms_ptr[nms+1], ms_sp[], ms_eff[] (float64; NaN = bare term), ms_is_sum[nms]
    wd_ptr[ns+1], wd_rx[], wd_coef[] (float64)
    W[ns] (float64)
    """
    nr = len(prog.reactions)
    out = {}
    out["rx_kind"] = np.array([KIND_CODE[r.kind] for r in prog.reactions], np.int32)
    out["rx_table"] = np.array([r.table for r in prog.reactions], np.int32)
    out["rx_mset"] = np.array([r.m_set for r in prog.reactions], np.int32)
    for side in ("fwd", "rev"):
        ptr = [0]
        sp, pw = [], []
        for r in prog.reactions:
            for k, n in getattr(r, side):
                sp.append(k)
                pw.append(n)
            ptr.append(len(sp))
        out[f"rx_{side}_ptr"] = np.array(ptr, np.int32)
        out[f"rx_{side}_sp"] = np.array(sp or [0], np.int32)
        out[f"rx_{side}_pow"] = np.array(pw or [0], np.int32)
    ptr = [0]
    sp, eff, is_sum = [], [], []
    for m in prog.m_sets:
        if m == "sum":
            is_sum.append(1)
        else:
            is_sum.append(0)
            for k, e in m:
                sp.append(k)
                eff.append(np.nan if e is None else e)
        ptr.append(len(sp))
    out["ms_ptr"] = np.array(ptr, np.int32)
    out["ms_sp"] = np.array(sp or [0], np.int32)
    out["ms_eff"] = np.array(eff or [0.0], np.float64)
    out["ms_is_sum"] = np.array(is_sum or [0], np.int32)
    ptr = [0]
    rx, cf = [], []
    for terms in prog.wdot_terms:
        for j, c in terms:
            rx.append(j)
            cf.append(c)
        ptr.append(len(rx))
    out["wd_ptr"] = np.array(ptr, np.int32)
    out["wd_rx"] = np.array(rx or [0], np.int32)
    out["wd_coef"] = np.array(cf or [0.0], np.float64)
    out["W"] = np.array(prog.W, np.float64)
    return out
