"""ctypes binding of the CPU oracle (oracle/_build/libfcfd_oracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfcfd_oracle.so")
# the same restatement with the C library's exp / log / log10 / pow (what the reference calls)
LIBM_PATH = os.path.join(HERE, "_build", "libfcfd_oracle_libm.so")
REPO = os.path.dirname(HERE)
MECH_DIR = os.path.join(REPO, "fortrancfd_amd", "data", "GRI3.0_CH4")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def _lib(libm=False):
    path = LIBM_PATH if libm else LIB_PATH
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    lib.fo_mech_load.restype = C.c_void_p
    lib.fo_mech_load.argtypes = [C.c_char_p, C.c_char_p]
    for name in ["fo_radiation_heat_loss", "fo_wall_temperature", "fo_wall_heat", "fo_ignition_power"]:
        getattr(lib, name).restype = C.c_double
    for name in ["fo_pressure", "fo_density", "fo_specific_energy", "fo_cp", "fo_cv",
                 "fo_sonic_speed", "fo_temperature_con", "fo_heat_release_rate",
                 "fo_heat_release_rate_omegadot", "fo_mcavis", "fo_mcacon", "fo_calc_dt", "fo_step"]:
        getattr(lib, name).restype = C.c_double
    lib.fo_domain_create.restype = C.c_void_p
    lib.fo_domain_create_smr.restype = C.c_void_p
    return lib


class FoOpts(C.Structure):
    _fields_ = [("flag_acu", C.c_int), ("flag_weno", C.c_int), ("flag_muscl", C.c_int),
                ("select_riemann", C.c_int), ("select_time", C.c_int), ("is_spherical", C.c_int),
                ("do_hydro", C.c_int), ("do_tran", C.c_int), ("dx_comp", C.c_double)]


class FoDomain(C.Structure):
    _fields_ = [("m", C.c_void_p), ("o", FoOpts), ("nblk", C.c_int), ("ncell", C.c_int),
                ("lv", C.c_int), ("nequ", C.c_int), ("dx", C.c_double),
                ("Qc", _D), ("qf", _D), ("Fplus", _D), ("godunov", _D)]


class FoDomainStats(C.Structure):
    """fo_domain up to its per-cell integrator counters."""
    _fields_ = FoDomain._fields_ + [("Qc0", _D), ("k1", _D), ("k2", _D), ("k3", _D), ("k4", _D),
                                    ("hmem", _D), ("cstats", _I)]


class FoAux(C.Structure):
    """fo_aux: open boundaries and auxiliary sources of a domain."""
    _fields_ = [("lbound", C.c_int), ("rbound", C.c_int), ("T_in", C.c_double), ("u_in", C.c_double),
                ("rho_u_in", C.c_double), ("p_out", C.c_double), ("x_in", C.c_double * 64),
                ("y_in", C.c_double * 64), ("nrad", C.c_int), ("rad_idx", C.c_int * 16),
                ("pmac", (C.c_double * 6) * 16), ("wall_heat", C.c_int), ("Nu", C.c_double),
                ("d_m", C.c_double), ("Tw_min", C.c_double), ("Tw_max", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("ig_type", C.c_int), ("n_ig", C.c_int),
                ("E_j", C.c_double * 16), ("r_m", C.c_double * 16), ("t_s", C.c_double * 16),
                ("t0_s", C.c_double * 16), ("vol", C.c_double * 16)]


class FoDomainAux(C.Structure):
    """fo_domain in full."""
    _fields_ = FoDomainStats._fields_ + [("aux", FoAux), ("xm", _D), ("qrad", _D), ("qw", _D),
                                         ("time", C.c_double), ("dxb", _D)]


def _p(a):
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """CPU restatement of the reference path on one mechanism."""

    def __init__(self, mech_dir: str = MECH_DIR, transport: bool = True, libm: bool = False):
        """libm=False: transcendental functions from the library's fcfd_math.h (the GPU's
        operations); libm=True: the C library's (the reference's)."""
        self.lib = _lib(libm)
        chem = os.path.join(mech_dir, "chem.bin").encode()
        tran = os.path.join(mech_dir, "tran.bin").encode() if transport else None
        self.m = C.c_void_p(self.lib.fo_mech_load(chem, tran))
        if not self.m:
            raise RuntimeError("oracle: mechanism load failed")
        self.kk = self.lib.fo_mech_kk(self.m)
        self.ii = self.lib.fo_mech_ii(self.m)
        self.nequ = 3 + self.kk

    def __del__(self):
        try:
            self.lib.fo_mech_free(self.m)
        except Exception:
            pass

    # ---- tables ---------------------------------------------------------------------------
    def weights(self):
        w = np.zeros(self.kk)
        self.lib.fo_mech_weights(self.m, _p(w))
        return w

    def nasa(self):
        tt = np.zeros(self.kk * 3)
        aa = np.zeros(self.kk * 2 * 7)
        self.lib.fo_mech_nasa(self.m, _p(tt), _p(aa))
        return tt.reshape(self.kk, 3), aa.reshape(self.kk, 2, 7)

    def arrhenius(self):
        par = np.zeros(self.ii * 4)
        nu = np.zeros(self.ii * 6, np.int32)
        nk = np.zeros(self.ii * 6, np.int32)
        ns = np.zeros(self.ii, np.int32)
        nr = np.zeros(self.ii, np.int32)
        self.lib.fo_mech_arrhenius(self.m, _p(par), nu.ctypes.data_as(_I), nk.ctypes.data_as(_I),
                                   ns.ctypes.data_as(_I), nr.ctypes.data_as(_I))
        return par.reshape(self.ii, 4), nu.reshape(self.ii, 6), nk.reshape(self.ii, 6), ns, nr

    # ---- species / mixture thermo -----------------------------------------------------------
    def _species(self, fn, T):
        out = np.zeros(self.kk)
        getattr(self.lib, fn)(self.m, C.c_double(T), _p(out))
        return out

    def hml(self, T): return self._species("fo_ckhml", T)
    def uml(self, T): return self._species("fo_ckuml", T)
    def cpml(self, T): return self._species("fo_ckcpml", T)
    def smh(self, T): return self._species("fo_cksmh", T)
    def hms(self, T): return self._species("fo_ckhms", T)

    def ytx(self, Y):
        Y = _f64(Y)
        X = np.zeros(self.kk)
        self.lib.fo_ckytx(self.m, _p(Y), _p(X))
        return X

    def xty(self, X):
        X = _f64(X)
        Y = np.zeros(self.kk)
        self.lib.fo_ckxty(self.m, _p(X), _p(Y))
        return Y

    def adjust_sum(self, v):
        v = _f64(v).copy()
        self.lib.fo_adjust_sum(C.c_int(v.size), _p(v))
        return v

    def pressure(self, rho, T, X):
        return self.lib.fo_pressure(self.m, C.c_double(rho), C.c_double(T), _p(_f64(X)))

    def density(self, p, T, X):
        return self.lib.fo_density(self.m, C.c_double(p), C.c_double(T), _p(_f64(X)))

    def specific_energy(self, T, X):
        return self.lib.fo_specific_energy(self.m, C.c_double(T), _p(_f64(X)))

    def cp(self, T, X): return self.lib.fo_cp(self.m, C.c_double(T), _p(_f64(X)))
    def cv(self, T, X): return self.lib.fo_cv(self.m, C.c_double(T), _p(_f64(X)))

    def sonic_speed(self, rho, p, T, X):
        return self.lib.fo_sonic_speed(self.m, C.c_double(rho), C.c_double(p), C.c_double(T),
                                       _p(_f64(X)))

    def temperature_con(self, e, X):
        return self.lib.fo_temperature_con(self.m, C.c_double(e), _p(_f64(X)))

    def heat_release_rate(self, Y0, Y, rho, dt):
        return self.lib.fo_heat_release_rate(self.m, _p(_f64(Y0)), _p(_f64(Y)), C.c_double(rho),
                                             C.c_double(dt))

    def heat_release_rate_omegadot(self, Y, rho, T):
        return self.lib.fo_heat_release_rate_omegadot(self.m, _p(_f64(Y)), C.c_double(rho),
                                                      C.c_double(T))

    # ---- eos ------------------------------------------------------------------------------
    def qf_to_qc(self, qf):
        qf = _f64(qf)
        out = np.zeros_like(qf)
        for r in range(qf.shape[0]):
            self.lib.fo_qf_to_qc(self.m, _p(qf[r]), out[r].ctypes.data_as(_D))
        return out

    def qc_to_qf(self, Qc):
        Qc = _f64(Qc)
        out = np.zeros_like(Qc)
        for r in range(Qc.shape[0]):
            self.lib.fo_qc_to_qf(self.m, _p(Qc[r]), out[r].ctypes.data_as(_D))
        return out

    # ---- reconstruction -------------------------------------------------------------------
    def recon(self, s):
        """s: (n, 8) stencils -> (n, 18) in the harness order (weno3z, weno5 js, weno5 z, weno7z,
        muscl3 flags 1..4; left then right)."""
        s = _f64(s)
        out = np.zeros((s.shape[0], 18))
        l, r = C.c_double(), C.c_double()
        for i in range(s.shape[0]):
            row = s[i]
            self.lib.fo_weno3z(_p(row), C.byref(l), C.byref(r)); out[i, 0:2] = l.value, r.value
            self.lib.fo_weno5(_p(row), C.byref(l), C.byref(r), 2); out[i, 2:4] = l.value, r.value
            self.lib.fo_weno5(_p(row), C.byref(l), C.byref(r), 3); out[i, 4:6] = l.value, r.value
            self.lib.fo_weno7z(_p(row), C.byref(l), C.byref(r)); out[i, 6:8] = l.value, r.value
            for f in range(1, 5):
                self.lib.fo_muscl3(_p(row), C.byref(l), C.byref(r), f)
                out[i, 6 + 2 * f:8 + 2 * f] = l.value, r.value
        return out

    # ---- flux -----------------------------------------------------------------------------
    @staticmethod
    def opts(flag_acu=2, flag_weno=2, flag_muscl=0, select_riemann=2, select_time=1, dx_comp=1e-3,
             do_tran=0, is_spherical=0):
        return FoOpts(flag_acu, flag_weno, flag_muscl, select_riemann, select_time, is_spherical, 1,
                      do_tran, dx_comp)

    def flux_advection(self, qf, n, lv, dx, **kw):
        """dx: the block's cell size; dx_comp (default dx) the fine mesh size of the domain."""
        kw.setdefault("dx_comp", dx)
        o = self.opts(**kw)
        qf = _f64(qf)
        F = np.zeros((n + 1, self.nequ))
        G = np.zeros((n + 1, 3))
        self.lib.fo_flux_advection(self.m, C.byref(o), C.c_int(n), C.c_int(lv), C.c_double(dx),
                                   _p(qf), _p(F), _p(G))
        return F, G

    def flux_tran(self, Qc, n, lv, dx):
        Qc = _f64(Qc)
        F = np.zeros((n + 1, self.nequ))
        self.lib.fo_flux_tran(self.m, C.c_int(n), C.c_int(lv), C.c_double(dx), _p(Qc), _p(F))
        return F

    def mcavis(self, T, X): return self.lib.fo_mcavis(self.m, C.c_double(T), _p(_f64(X)))
    def mcacon(self, T, X): return self.lib.fo_mcacon(self.m, C.c_double(T), _p(_f64(X)))

    def mcadif(self, p, T, X):
        D = np.zeros(self.kk)
        self.lib.fo_mcadif(self.m, C.c_double(p), C.c_double(T), _p(_f64(X)), _p(D))
        return D

    # ---- kinetics -------------------------------------------------------------------------
    def ckwyr(self, rho_cgs, T, Y):
        w = np.zeros(self.kk)
        qf = np.zeros(self.ii)
        qr = np.zeros(self.ii)
        self.lib.fo_ckwyr(self.m, C.c_double(rho_cgs), C.c_double(T), _p(_f64(Y)), _p(w), _p(qf),
                          _p(qr))
        return w, qf, qr

    def ckcdyr(self, rho_cgs, T, Y):
        c = np.zeros(self.kk)
        d = np.zeros(self.kk)
        self.lib.fo_ckcdyr(self.m, C.c_double(rho_cgs), C.c_double(T), _p(_f64(Y)), _p(c), _p(d))
        return c, d


class ChemStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("rejects", C.c_int), ("newton", C.c_int),
                ("jacobians", C.c_int), ("failed", C.c_int)]


def _chem_methods():
    def ncf(self):
        mm = self.lib.fo_mech_mm(self.m)
        out = np.zeros(self.kk * mm, np.int32)
        self.lib.fo_mech_ncf(self.m, out.ctypes.data_as(_I))
        return out.reshape(self.kk, mm)

    def chem_rhs(self, rho_cgs, yT):
        f = np.zeros(self.kk + 1)
        self.lib.fo_chem_rhs(self.m, C.c_double(rho_cgs), _p(_f64(yT)), _p(f))
        return f

    def chem_jac(self, rho_cgs, yT):
        n = self.kk + 1
        J = np.zeros(n * n)
        self.lib.fo_chem_jac(self.m, C.c_double(rho_cgs), _p(_f64(yT)), _p(J))
        return J.reshape(n, n)

    def chem_step(self, dt, rho_si, yT, rtol=1e-5, atol=1e-13):
        """macks(dt, nspe+1, yiT, rho): returns (yT_new, stats dict, rc)."""
        y = _f64(yT).copy()
        st = ChemStats()
        rc = self.lib.fo_chem_step(self.m, C.c_double(dt), C.c_double(rho_si), _p(y),
                                   C.c_double(rtol), C.c_double(atol), C.byref(st))
        return y, {k: getattr(st, k) for k, _ in ChemStats._fields_}, rc

    return ncf, chem_rhs, chem_jac, chem_step


Oracle.ncf, Oracle.chem_rhs, Oracle.chem_jac, Oracle.chem_step = _chem_methods()


class Domain:
    """Multi-block wall-bounded domain stepped by the oracle (reference storage layout)."""

    def __init__(self, orc: Oracle, cells, nblk, ncell, dx, smr=None, **opts):
        """smr = (lev_smr_0, num_smr_0, lev_smr_n, num_smr_n): the reference's stretched-mesh
        layout around nblk blocks at dx (cells then covers every block)."""
        self.orc = orc
        self.o = orc.opts(dx_comp=dx, **opts)
        cells = _f64(cells)
        if smr is None:
            ptr = orc.lib.fo_domain_create(orc.m, C.byref(self.o), C.c_int(nblk), C.c_int(ncell),
                                           C.c_double(dx), _p(cells))
        else:
            l0, n0, ln, nn = smr
            nblk = nblk + (l0 - 1) * n0 + (ln - 1) * nn
            assert cells.shape[0] == nblk * ncell
            ptr = orc.lib.fo_domain_create_smr(orc.m, C.byref(self.o), C.c_int(nblk - (l0 - 1) * n0 - (ln - 1) * nn),
                                               C.c_int(ncell), C.c_double(dx), C.c_int(l0), C.c_int(n0),
                                               C.c_int(ln), C.c_int(nn), _p(cells))
            if not ptr:
                raise ValueError("oracle: invalid refinement levels")
        self.ptr = C.c_void_p(ptr)
        self.d = C.cast(self.ptr, C.POINTER(FoDomain)).contents
        self.nblk, self.ncell, self.lv, self.nequ = nblk, ncell, self.d.lv, orc.nequ

    def __del__(self):
        try:
            self.orc.lib.fo_domain_free(self.ptr)
        except Exception:
            pass

    def _arr(self, ptr):
        n = self.nblk * (self.ncell + 2 * self.lv) * self.nequ
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.nblk, self.ncell + 2 * self.lv,
                                                              self.nequ)

    @property
    def Qc(self): return self._arr(self.d.Qc).copy()

    @property
    def qf(self): return self._arr(self.d.qf).copy()

    def calc_dt(self, cfl):
        return self.orc.lib.fo_calc_dt(self.ptr, C.c_double(cfl))

    def solve_fluid(self, dt):
        self.orc.lib.fo_solve_fluid(self.ptr, C.c_double(dt))

    def refresh(self):
        self.orc.lib.fo_refresh(self.ptr)

    def solve_source(self, dt, rtol=1e-5, atol=1e-13):
        n = self.nblk * self.ncell
        hrr, hrd = np.zeros(n), np.zeros(n)
        nf = C.c_int()
        self.orc.lib.fo_solve_source(self.ptr, C.c_double(dt), C.c_double(rtol), C.c_double(atol),
                                     _p(hrr), _p(hrd), C.byref(nf))
        return hrr, hrd, nf.value

    def step(self, cfl, react=True, strang=False, rtol=1e-5, atol=1e-13, diagnostics=False):
        """One driver step (fo_step): Lie (the reference's order) or Strang splitting; returns
        dt, or (dt, HRR, HRR_dot) of the last source call with diagnostics."""
        n = self.nblk * self.ncell
        hrr, hrd = (np.zeros(n), np.zeros(n)) if diagnostics else (None, None)
        nf = C.c_int()
        dt = self.orc.lib.fo_step(self.ptr, C.c_double(cfl), C.c_int(int(react)), C.c_int(int(strang)),
                                  C.c_double(rtol), C.c_double(atol),
                                  _p(hrr) if diagnostics else None, _p(hrd) if diagnostics else None,
                                  C.byref(nf))
        if nf.value:
            raise RuntimeError(f"oracle: {nf.value} cells failed to integrate")
        return (dt, hrr, hrd) if diagnostics else dt

    # ---- open boundaries and auxiliary sources ----------------------------------------------
    @property
    def full(self):
        return FoDomainAux.from_address(self.ptr.value)

    def set_flow_boundary(self, flag_lbound, flag_rbound, T_in=300.0, u_in_cm_s=0.0, x_in=None,
                          p_out_atm=1.0):
        x = _f64(np.zeros(self.nequ - 3) if x_in is None else x_in)
        rc = self.orc.lib.fo_domain_set_flow_boundary(self.ptr, C.c_int(flag_lbound), C.c_int(flag_rbound),
                                                      C.c_double(T_in), C.c_double(u_in_cm_s), _p(x),
                                                      C.c_double(p_out_atm))
        if rc:
            raise ValueError("oracle: invalid boundary flags")

    def set_radiation(self, species, pmac):
        sp = np.ascontiguousarray(species, np.int32)
        pm = _f64(pmac).reshape(-1)
        if self.orc.lib.fo_domain_set_radiation(self.ptr, C.c_int(sp.size), sp.ctypes.data_as(_I), _p(pm)):
            raise ValueError("oracle: invalid radiation parameters")

    def set_wall_heat(self, Nu, d_mm, Tw_min, Tw_max, alpha, beta, on=True):
        self.orc.lib.fo_domain_set_wall_heat(self.ptr, C.c_int(int(on)), *(C.c_double(v) for v in
                                             (Nu, d_mm, Tw_min, Tw_max, alpha, beta)))

    def set_ignition(self, kind, E_mJ, r_um, t_us, t0_us):
        arrs = [_f64(np.atleast_1d(v)) for v in (E_mJ, r_um, t_us, t0_us)]
        if self.orc.lib.fo_domain_set_ignition(self.ptr, C.c_int(kind), C.c_int(arrs[0].size),
                                               *(_p(a) for a in arrs)):
            raise ValueError("oracle: invalid ignition parameters")

    @property
    def xm(self):
        return np.ctypeslib.as_array(self.full.xm, shape=(self.nblk * self.ncell,)).copy()

    def set_xm(self, xm):
        x = _f64(xm)
        assert x.size == self.nblk * self.ncell
        self.orc.lib.fo_domain_set_xm(self.ptr, _p(x))

    @property
    def dxb(self):
        return np.ctypeslib.as_array(self.full.dxb, shape=(self.nblk,)).copy()

    @property
    def time(self):
        return self.full.time

    def set_time(self, t):
        self.full.time = t

    def aux_sources(self, dt, time):
        self.orc.lib.fo_aux_sources(self.ptr, C.c_double(dt), C.c_double(time))

    def aux_outputs(self):
        """Qrad and Qw (W/m^3) of the last auxiliary-source call, per interior cell."""
        n = self.nblk * self.ncell
        f = self.full
        return (np.ctypeslib.as_array(f.qrad, shape=(n,)).copy(),
                np.ctypeslib.as_array(f.qw, shape=(n,)).copy())

    def eos_fields(self, hrr=None, hrr_dot=None, idx0=1):
        """(15 + KK, n) output rows (fo_eos_fields) with the given last HRR / HRR_dot."""
        n = self.nblk * self.ncell
        out = np.zeros((self.nequ + 12, n))
        h = None if hrr is None else _f64(hrr)
        hd = None if hrr_dot is None else _f64(hrr_dot)
        self.orc.lib.fo_eos_fields(self.ptr, None if h is None else _p(h), None if hd is None else _p(hd),
                                   C.c_int(idx0), _p(out))
        return out

    def restart(self, rows):
        """fo_domain_restart from rows rho, u, T, e_total, X_1..X_KK."""
        self.orc.lib.fo_domain_restart(self.ptr, _p(_f64(rows)))

    def chem_stats(self):
        """Per interior cell (block-major): steps, rejects, newton, failed of the last
        solve_source, as the library's fcfd_chem_stats."""
        d = FoDomainStats.from_address(self.ptr.value)
        n = self.nblk * self.ncell
        return np.ctypeslib.as_array(d.cstats, shape=(n * 4,)).reshape(n, 4).copy()
