"""Mechanism data: compiled rate program plus the 1 K tables, in FLINT's and in device layout.

Host-side equivalent of what FLINT's loaders put into module globals
(read_idealgas_thermo, Load_ThermoTransport.f90:7-89; read_chemistry, Load_Chemistry.f90:7-164):
molecular weights, thermo tables with their [Tmin, Tmax] (with the Tmin = 1 -> 0 row of
Load_ThermoTransport.f90:70,80-85), and the kinetics tables with their own [Ti1, Ti2].
"""
import functools

import numpy as np

from . import tables as _tables
from .mechanism import compile_mechanism, general_tables


class MechData:
    def __init__(self, name):
        self.name = name
        self.prog = compile_mechanism(name)
        self.desc = self.prog.desc
        self.species = list(self.prog.species)
        self.ns = self.prog.ns
        self.W = np.array(self.prog.W, dtype=np.float64)
        self.is_wd = name == "WD"
        Tmin, Tmax = self.desc["T_range"]
        kin = _tables.kinetics_tables(self.desc, Tmin, Tmax)
        self.Tk_min, self.Tk_max = Tmin, Tmax
        self.kf = np.ascontiguousarray(kin["kf"])
        self.kb = np.ascontiguousarray(kin["kb"])
        self.troe = np.ascontiguousarray(kin["troe"])
        self.lind = np.ascontiguousarray(kin["lind"])
        th = _tables.thermo_tables(self.desc, Tmin, Tmax)
        cp, h, s = th["cp"], th["h"], th["s"]
        if Tmin == 1:  # Load_ThermoTransport.f90:80-85
            cp = np.vstack([cp[:1], cp])
            h = np.vstack([np.zeros((1, self.ns)), h])
            s = np.vstack([s[:1], s])
            self.Th_min = 0
        else:
            self.Th_min = Tmin
        self.Th_max = Tmax
        self.cp, self.h, self.s = (np.ascontiguousarray(a) for a in (cp, h, s))
        self.elements = self.desc["elements"]
        self.composition = np.array(self.desc["composition"], dtype=np.float64)  # [ne, ns]

    @functools.cached_property
    def general(self):
        """(na, nt, nl, ni1, ni2, eps) of FLINT's data-driven general routine (general_tables)."""
        return general_tables(self.desc)

    @property
    def n_arrh(self):
        return self.kf.shape[1]

    @property
    def n_troe(self):
        return self.troe.shape[1]

    @property
    def n_lind(self):
        return self.lind.shape[1]

    def device_tables(self):
        """Row-major interleaved tables with one guard row (copy of the last row).

        tabk[row][2a+0/1]                      kf, kb of Arrhenius/three-body reaction a
        tabk[row][2*na + 4t + 0..3]           k_inf, k_0, k_c, F_cent of Troe reaction t
        tabk[row][2*na + 4*nt + 3l + 0..2]    k_inf, k_0, k_c of Lindemann reaction l
        tabh[row][2k+0/1]                      cp_k, h_k
        (padded to an even width so that every row starts 16-byte aligned)
        """
        na, nt, nl = self.n_arrh, self.n_troe, self.n_lind
        nT = self.kf.shape[0]
        wk = 2 * na + 4 * nt + 3 * nl
        wk += wk % 2
        tabk = np.zeros((nT + 1, wk))
        tabk[:nT, 0:2 * na:2] = self.kf
        tabk[:nT, 1:2 * na:2] = self.kb
        if nt:
            tabk[:nT, 2 * na:2 * na + 4 * nt] = self.troe.reshape(nT, 4 * nt)
        if nl:
            tabk[:nT, 2 * na + 4 * nt:2 * na + 4 * nt + 3 * nl] = self.lind.reshape(nT, 3 * nl)
        tabk[nT] = tabk[nT - 1]
        nTh = self.cp.shape[0]
        tabh = np.zeros((nTh + 1, 2 * self.ns))
        tabh[:nTh, 0::2] = self.cp
        tabh[:nTh, 1::2] = self.h
        tabh[nTh] = tabh[nTh - 1]
        return np.ascontiguousarray(tabk), np.ascontiguousarray(tabh), wk


@functools.lru_cache(maxsize=8)
def mech_data(name):
    return MechData(name)
