"""ctypes binding of libfcfd_hip.so (include/fcfd.h).

The library is the product: every compute call below runs HIP kernels on the current device.  If
the shared object is missing, loading fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_lib", "libfcfd_hip.so")
DATA_DIR = os.path.join(PKG, "data")
GRI30 = os.path.join(DATA_DIR, "GRI3.0_CH4")
LIDRYER_H2 = os.path.join(DATA_DIR, "LiDryer_H2")

_D = C.POINTER(C.c_double)
_lib = None


class FcfdError(RuntimeError):
    pass


class Opts(C.Structure):
    """fcfd_opts: the reference's &params switches (SRC/params.f90:106-111)."""
    _fields_ = [("do_hydro", C.c_int), ("do_tran", C.c_int), ("do_reaction", C.c_int),
                ("select_time", C.c_int), ("select_riemann", C.c_int), ("flag_acu", C.c_int),
                ("flag_muscl", C.c_int), ("flag_weno", C.c_int), ("is_spherical", C.c_int),
                ("flag_lbound", C.c_int), ("flag_rbound", C.c_int), ("atol", C.c_double),
                ("rtol", C.c_double)]


def opts(do_hydro=1, do_tran=0, do_reaction=0, select_time=1, select_riemann=2, flag_acu=2,
         flag_muscl=0, flag_weno=2, atol=1e-13, rtol=1e-5, is_spherical=0, flag_lbound=0,
         flag_rbound=0) -> Opts:
    return Opts(do_hydro, do_tran, do_reaction, select_time, select_riemann, flag_acu, flag_muscl,
                flag_weno, is_spherical, flag_lbound, flag_rbound, atol, rtol)


def lib():
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7 / HSA runtime.
        # Loading torch first lets this library bind to the runtime already in the process
        # (same soname); loading it first would bring in a second HSA runtime that torch then
        # cannot initialise.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise FcfdError(f"{LIB_PATH} is missing: build it with `python -m fortrancfd_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.fcfd_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise FcfdError(f"fcfd error {rc}: {lib().fcfd_last_error().decode()}")


def _p(a):
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Mechanism:
    """CHEMKIN-II link files (chem.bin, tran.bin) as read by CKINIT / MCINIT."""

    def __init__(self, mech_dir: str = GRI30, transport: bool = True, text: bool = False):
        """text=False: the link file chem.bin (CKINIT's input); text=True: the CHEMKIN text
        chem.inp + therm.dat read directly.  Transport from tran.bin either way."""
        L = lib()
        self.handle = C.c_void_p()
        tran = os.path.join(mech_dir, "tran.bin").encode() if transport else None
        if text:
            therm = os.path.join(mech_dir, "therm.dat")
            _check(L.fcfd_mech_load_text(os.path.join(mech_dir, "chem.inp").encode(),
                                         therm.encode() if os.path.exists(therm) else None, tran,
                                         C.byref(self.handle)))
        else:
            chem = os.path.join(mech_dir, "chem.bin").encode()
            _check(L.fcfd_mech_load(chem, tran, C.byref(self.handle)))
        kk, ii = C.c_int(), C.c_int()
        _check(L.fcfd_mech_dims(self.handle, C.byref(kk), C.byref(ii)))
        self.kk, self.ii, self.nequ = kk.value, ii.value, 3 + kk.value
        mm = C.c_int()
        _check(L.fcfd_mech_elements(self.handle, C.byref(mm)))
        self.mm = mm.value
        names, wts, ncf = [], np.zeros(self.kk), np.zeros((self.kk, self.mm), np.int32)
        buf = C.create_string_buffer(32)
        for k in range(self.kk):
            w = C.c_double()
            row = np.zeros(self.mm, np.int32)
            _check(L.fcfd_mech_species(self.handle, C.c_int(k), buf, C.c_int(32), C.byref(w),
                                       row.ctypes.data_as(C.POINTER(C.c_int))))
            names.append(buf.value.decode())
            wts[k] = w.value
            ncf[k] = row
        self.names, self.weights, self.ncf = names, wts, ncf

    def __del__(self):
        try:
            lib().fcfd_mech_free(self.handle)
        except Exception:
            pass

    # ---- tables (host) -------------------------------------------------------------------
    def thermo(self, k):
        t, a = np.zeros(3), np.zeros(14)
        _check(lib().fcfd_mech_thermo(self.handle, C.c_int(k), _p(t), _p(a)))
        return t, a

    def reaction(self, i):
        par = np.zeros(3)
        nu, nk = np.zeros(6, np.int32), np.zeros(6, np.int32)
        ns, nr = C.c_int(), C.c_int()
        P = C.POINTER(C.c_int)
        _check(lib().fcfd_mech_reaction(self.handle, C.c_int(i), _p(par), nu.ctypes.data_as(P),
                                        nk.ctypes.data_as(P), C.byref(ns), C.byref(nr)))
        return par, nu, nk, ns.value, nr.value

    def falloff(self, i):
        ifo, kfl = C.c_int(), C.c_int()
        fp = np.zeros(8)
        _check(lib().fcfd_mech_falloff(self.handle, C.c_int(i), C.byref(ifo), C.byref(kfl), _p(fp)))
        return ifo.value, kfl.value, fp

    def thirdbody(self, i):
        n = C.c_int()
        nk, a = np.zeros(16, np.int32), np.zeros(16)
        _check(lib().fcfd_mech_thirdbody(self.handle, C.c_int(i), C.byref(n),
                                         nk.ctypes.data_as(C.POINTER(C.c_int)), _p(a)))
        return n.value, nk[:max(n.value, 0)], a[:max(n.value, 0)]

    # ---- stateless kernels on host arrays ----------------------------------------------------
    def calc_flux_advection(self, qf, n, lv, dx, **kw):
        """calc_flux_advection (SRC/flux.f90:130-373) on one block: qf rows (n+2lv, nequ) ->
        (Fplus (n+1, nequ), godunov_state (n+1, 3))."""
        o = opts(**kw)
        qf = _f64(qf)
        F = np.zeros((n + 1, self.nequ))
        G = np.zeros((n + 1, 3))
        _check(lib().fcfd_flux_advection_host(self.handle, C.byref(o), C.c_int(n), C.c_int(lv),
                                              C.c_double(dx), _p(qf), _p(F), _p(G)))
        return F, G

    def calc_flux_tran(self, Qc, n, lv, dx):
        """calc_flux_tran (SRC/flux.f90:18-127) on one block: conserved rows (n+2lv, nequ) ->
        Fplus (n+1, nequ) of the transport fluxes."""
        Qc = _f64(Qc)
        F = np.zeros((n + 1, self.nequ))
        _check(lib().fcfd_flux_tran_host(self.handle, C.c_int(n), C.c_int(lv), C.c_double(dx),
                                         _p(Qc), _p(F)))
        return F

    def qfToQc(self, qf):
        qf = _f64(np.atleast_2d(qf))
        out = np.zeros_like(qf)
        _check(lib().fcfd_qf_to_qc_host(self.handle, C.c_int(qf.shape[0]), _p(qf), _p(out)))
        return out

    def QcToqf(self, Qc):
        Qc = _f64(np.atleast_2d(Qc))
        out = np.zeros_like(Qc)
        _check(lib().fcfd_qc_to_qf_host(self.handle, C.c_int(Qc.shape[0]), _p(Qc), _p(out)))
        return out

    def rates(self, rho, T, Y):
        """CKWYR + CKCDYR on the device (the integrator's rate code): rho [g/cm^3] and T of shape
        [n], Y [n, KK] -> (wdot, cdot, ddot [n, KK], qf, qr [n, II])."""
        rho, T = _f64(np.atleast_1d(rho)), _f64(np.atleast_1d(T))
        Y = _f64(np.atleast_2d(Y))
        n = T.size
        assert rho.size == n and Y.shape == (n, self.kk)
        w, c, d = (np.zeros((n, self.kk)) for _ in range(3))
        qf, qr = np.zeros((n, self.ii)), np.zeros((n, self.ii))
        _check(lib().fcfd_rates_host(self.handle, C.c_int(n), _p(rho), _p(T), _p(Y), _p(w), _p(c),
                                     _p(d), _p(qf), _p(qr)))
        return w, c, d, qf, qr

    def macks(self, dt, rho, yT, rtol=1e-5, atol=1e-13):
        """macks(dt, nspe+1, yiT, rho) (SRC/source.f90:58) on n cells: yT rows (Y..., T), dt and
        rho [kg/m^3] per cell.  Returns (yT_new, stats[n, 4] = steps, rejects, newton, failed)."""
        yT = _f64(np.atleast_2d(yT)).copy()
        n = yT.shape[0]
        dt = _f64(np.broadcast_to(dt, (n,)))
        rho = _f64(np.broadcast_to(rho, (n,)))
        st = np.zeros((n, 4), np.int32)
        o = opts(atol=atol, rtol=rtol, do_reaction=1)
        _check(lib().fcfd_chem_step_host(self.handle, C.byref(o), C.c_int(n), _p(dt), _p(rho),
                                         _p(yT), st.ctypes.data_as(C.POINTER(C.c_int))))
        return yT, st

    def chem_batch_device(self, rho, dt, T, Y, stats=None, rtol=1e-5, atol=1e-13, stream=None):
        """Batched chemistry on device tensors (torch): rho, dt, T of shape [n], Y of shape [KK, n]
        (float64, contiguous, same device); T and Y are updated in place."""
        n = int(T.numel())
        o = opts(atol=atol, rtol=rtol, do_reaction=1)
        ptr = lambda t: C.c_void_p(t.data_ptr())
        _check(lib().fcfd_chem_batch(self.handle, C.byref(o), C.c_int(n), ptr(rho), ptr(dt), ptr(T),
                                     ptr(Y), ptr(stats) if stats is not None else None,
                                     C.c_void_p(stream) if stream else None))


def recon(stencils):
    """weno3z / weno5 (JS, Z) / weno7z / muscl3 (flags 1-4) on (n, 8) stencils -> (n, 18)."""
    s = _f64(stencils)
    out = np.zeros((s.shape[0], 18))
    _check(lib().fcfd_recon_host(C.c_int(s.shape[0]), _p(s), _p(out)))
    return out


class Solver:
    """A multi-block wall-bounded domain on the GPU: the state b_obj(:)%Qc/%qf of the reference
    (SRC/block.f90:10-23) held on the device, advanced by calc_dt / solve_fluid / refresh."""

    def __init__(self, mech: Mechanism, nblk: int, ncell: int, dx: float, smr=None, **kw):
        """smr = (lev_smr_0, num_smr_0, lev_smr_n, num_smr_n): the reference's stretched-mesh
        layout around the nblk blocks at dx (self.nblk then counts every block)."""
        self.mech = mech
        self.o = opts(**kw)
        self.handle = C.c_void_p()
        if smr is None:
            _check(lib().fcfd_create(mech.handle, C.byref(self.o), C.c_int(nblk), C.c_int(ncell),
                                     C.c_double(dx), C.byref(self.handle)))
        else:
            _check(lib().fcfd_create_smr(mech.handle, C.byref(self.o), C.c_int(nblk), C.c_int(ncell),
                                         C.c_double(dx), *(C.c_int(v) for v in smr), C.byref(self.handle)))
            nblk = lib().fcfd_nblocks(self.handle)
        self.nblk, self.ncell, self.dx = nblk, ncell, dx
        self.lv = lib().fcfd_lvir(self.handle)
        self.nequ = mech.nequ

    def __del__(self):
        try:
            lib().fcfd_destroy(self.handle)
        except Exception:
            pass

    def init_cells(self, cells):
        """cells: (nblk*ncell, 3+KK) rows of T [K], p [Pa], u [m/s], X."""
        cells = _f64(cells)
        assert cells.shape == (self.nblk * self.ncell, self.nequ)
        _check(lib().fcfd_init_cells(self.handle, _p(cells)))

    def upload(self, Qc, qf=None):
        """Qc, qf: (nblk, ncell+2lv, nequ) in the reference's per-block layout."""
        Qc = _f64(Qc)
        qf = None if qf is None else _f64(qf)
        for b in range(self.nblk):
            _check(lib().fcfd_upload(self.handle, C.c_int(b), _p(Qc[b]),
                                     None if qf is None else _p(qf[b])))

    def download(self):
        shape = (self.nblk, self.ncell + 2 * self.lv, self.nequ)
        Qc, qf = np.zeros(shape), np.zeros(shape)
        for b in range(self.nblk):
            _check(lib().fcfd_download(self.handle, C.c_int(b), _p(Qc[b]), _p(qf[b])))
        return Qc, qf

    def calc_dt(self, cfl):
        dt = C.c_double()
        _check(lib().fcfd_calc_dt(self.handle, C.c_double(cfl), C.byref(dt)))
        return dt.value

    def solve_fluid(self, dt):
        _check(lib().fcfd_solve_fluid(self.handle, C.c_double(dt)))

    def refresh(self):
        _check(lib().fcfd_refresh(self.handle))

    def solve_source(self, dt, time=0.0, diagnostics=True):
        """reaction_macks on every interior cell; returns (HRR, HRR_dot) per interior cell when
        diagnostics is set, else None."""
        n = self.nblk * self.ncell
        if not diagnostics:
            _check(lib().fcfd_solve_source(self.handle, C.c_double(dt), C.c_double(time), None,
                                           None))
            return None
        hrr, hrd = np.zeros(n), np.zeros(n)
        _check(lib().fcfd_solve_source(self.handle, C.c_double(dt), C.c_double(time), _p(hrr),
                                       _p(hrd)))
        return hrr, hrd

    def set_step(self, strang=False, hrr_each_step=False):
        """Step composition of advance(): Lie (default, the reference's) or Strang splitting, and
        whether each step's source call computes HRR / HRR_dot (read with get_hrr)."""
        _check(lib().fcfd_set_step(self.handle, C.c_int(int(strang)), C.c_int(int(hrr_each_step))))

    def get_hrr(self):
        n = self.nblk * self.ncell
        hrr, hrd = np.zeros(n), np.zeros(n)
        _check(lib().fcfd_get_hrr(self.handle, _p(hrr), _p(hrd)))
        return hrr, hrd

    def set_surrogate(self, sur):
        """Route solve_source through a Surrogate (None: the FP64 integrator everywhere)."""
        _check(lib().fcfd_set_surrogate(self.handle, sur.handle if sur is not None else None))
        self._sur = sur

    def set_chem_workgroups(self, n):
        """Grid of the implicit integrator (0: default): n one-wave workgroups for lists longer
        than 8 n cells, else n / 4 four-wave workgroups."""
        _check(lib().fcfd_set_chem_workgroups(self.handle, C.c_int(n)))

    def source_accepted(self):
        """(per-cell route of the last solve_source: 2 explicit step over dt, 3 two explicit half
        steps, 1 surrogate, 0 implicit FP64 integrator; number of implicit-integrator cells)."""
        acc = np.zeros(self.nblk * self.ncell, np.int32)
        nf = C.c_int()
        _check(lib().fcfd_source_accepted(self.handle, acc.ctypes.data_as(C.POINTER(C.c_int)),
                                          C.byref(nf)))
        return acc, nf.value

    def chem_stats(self):
        st = np.zeros((self.nblk * self.ncell, 4), np.int32)
        _check(lib().fcfd_chem_stats(self.handle, st.ctypes.data_as(C.POINTER(C.c_int))))
        return st

    def advance(self, cfl, nsteps, time=0.0):
        t = C.c_double(time)
        _check(lib().fcfd_advance(self.handle, C.c_double(cfl), C.c_int(nsteps), C.byref(t)))
        return t.value

    # ---- open boundaries and auxiliary per-cell sources ----------------------------------------
    def set_flow_boundary(self, T_in=300.0, u_in_cm_s=0.0, x_in=None, p_out_atm=1.0):
        """Inflow / outflow state for flag_lbound = 1 / flag_rbound = 1 (the values of the
        reference's flow-boundary file, fortrancfd_amd.inputs.read_flow_boundary)."""
        x = None if x_in is None else _f64(x_in)
        _check(lib().fcfd_set_flow_boundary(self.handle, C.c_double(T_in), C.c_double(u_in_cm_s),
                                            None if x is None else _p(x), C.c_double(p_out_atm)))

    def set_radiation(self, species, pmac):
        """Optically thin radiation loss of the given species (0-based) with their six absorption
        coefficients each; an empty list turns it off."""
        sp = np.ascontiguousarray(species, np.int32)
        pm = _f64(pmac).reshape(-1)
        _check(lib().fcfd_set_radiation(self.handle, C.c_int(sp.size),
                                        sp.ctypes.data_as(C.POINTER(C.c_int)) if sp.size else None,
                                        _p(pm) if sp.size else None))

    def set_wall_heat(self, Nu, d_mm, Tw_min, Tw_max, alpha, beta, on=True):
        """Micro-flow-reactor wall heat loss (the &mfrwalltemp values)."""
        _check(lib().fcfd_set_wall_heat(self.handle, C.c_int(int(on)), *(C.c_double(v) for v in
                                        (Nu, d_mm, Tw_min, Tw_max, alpha, beta))))

    def set_ignition(self, kind, E_mJ=(), r_um=(), t_us=(), t0_us=()):
        """Ignition deposit: kind 1 Frendi, 2 Zhang, 0 off; per source energy (mJ), radius (um),
        duration and start (us)."""
        arrs = [_f64(np.atleast_1d(v)) for v in (E_mJ, r_um, t_us, t0_us)]
        n = arrs[0].size
        _check(lib().fcfd_set_ignition(self.handle, C.c_int(kind), C.c_int(n),
                                       *((_p(a) if n else None) for a in arrs)))

    def set_cell_centres(self, xm):
        xm = _f64(xm)
        assert xm.size == self.nblk * self.ncell
        _check(lib().fcfd_set_cell_centres(self.handle, _p(xm)))

    def block_dx(self):
        out = np.zeros(self.nblk)
        _check(lib().fcfd_block_dx(self.handle, _p(out)))
        return out

    def cell_centres(self):
        out = np.zeros(self.nblk * self.ncell)
        _check(lib().fcfd_get_cell_centres(self.handle, _p(out)))
        return out

    def get_aux(self):
        """(Qrad, Qw) of the last auxiliary-source call, W/m^3 per interior cell."""
        n = self.nblk * self.ncell
        qr, qw = np.zeros(n), np.zeros(n)
        _check(lib().fcfd_get_aux(self.handle, _p(qr), _p(qw)))
        return qr, qw

    def set_timing(self, on: bool):
        _check(lib().fcfd_set_timing(self.handle, C.c_int(1 if on else 0)))

    def timing(self, kernel_id):
        ms, n = C.c_double(), C.c_longlong()
        _check(lib().fcfd_get_timing(self.handle, C.c_int(kernel_id), C.byref(ms), C.byref(n)))
        return ms.value, n.value


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def bf16_gemm(A, sam, sak, B, sbk, sbn, M, N, K, C32, ldc, ksplit=1, stream=None):
    """C32 = A B with the surrogate's BF16 MFMA kernel; A, B, C32 are device tensors (A, B
    bfloat16 or int16 bit patterns) addressed through the given element strides."""
    L = lib()
    L.fcfd_bf16_gemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_long, C.c_long,
                                 C.c_void_p, C.c_long, C.c_long, C.c_void_p, C.c_long, C.c_int,
                                 C.c_void_p]
    _check(L.fcfd_bf16_gemm(M, N, K, _vp(A), sam, sak, _vp(B), sbk, sbn, _vp(C32), ldc, ksplit,
                            stream))


class Surrogate:
    """The chemistry surrogate (56 -> 4 x 256 GELU -> 54, BF16 MFMA, FP32 accumulate) with its
    training step and its surrogate + FP64-fallback inference.  Array arguments are device tensors
    (torch) unless stated otherwise."""

    NORM_FIELDS = ("in_mean", "in_std", "lo", "hi", "out_mean", "out_std")

    def __init__(self, mech: Mechanism):
        self.mech = mech
        self.handle = C.c_void_p()
        _check(lib().fcfd_sur_create(mech.handle, C.byref(self.handle)))
        self.nparams = lib().fcfd_sur_nparams()

    def __del__(self):
        try:
            lib().fcfd_sur_free(self.handle)
        except Exception:
            pass

    # parameters (host numpy float32, flat, in the documented order)
    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == self.nparams
        _check(lib().fcfd_sur_set_params(self.handle, flat.ctypes.data_as(C.POINTER(C.c_float))))

    def get_params(self):
        out = np.zeros(self.nparams, np.float32)
        _check(lib().fcfd_sur_get_params(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def get_grads(self):
        out = np.zeros(self.nparams, np.float32)
        _check(lib().fcfd_sur_get_grads(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_norm(self, norm: dict):
        arr = np.ascontiguousarray(np.stack([np.asarray(norm[f], np.float64)
                                             for f in self.NORM_FIELDS]))
        assert arr.shape == (6, 64)
        _check(lib().fcfd_sur_set_norm(self.handle, _p(arr)))

    def set_thresholds(self, resid_tol=1e-5, t_tol=2e-3, verify_tol=1.0):
        _check(lib().fcfd_sur_set_thresholds(self.handle, C.c_double(resid_tol), C.c_double(t_tol),
                                             C.c_double(verify_tol)))

    def features(self, rho, dt, T, Y, X, in_range=None, stream=None):
        n = int(T.numel())
        _check(lib().fcfd_sur_features(self.handle, C.c_int(n), _vp(rho), _vp(dt), _vp(T), _vp(Y),
                                       _vp(X), _vp(in_range), stream))

    def targets(self, T0, Y0, T1, Y1, Tg, stream=None):
        n = int(T0.numel())
        _check(lib().fcfd_sur_targets(self.handle, C.c_int(n), _vp(T0), _vp(Y0), _vp(T1), _vp(Y1),
                                      _vp(Tg), stream))

    def forward(self, X, Yh, stream=None):
        _check(lib().fcfd_sur_forward(self.handle, C.c_int(int(X.shape[0])), _vp(X), _vp(Yh),
                                      stream))

    def train_step(self, X_all, Tg_all, idx, lr, want_loss=False, stream=None):
        loss = C.c_float()
        _check(lib().fcfd_sur_train_step(self.handle, _vp(X_all), _vp(Tg_all), _vp(idx),
                                         C.c_int(int(idx.numel())), C.c_float(lr),
                                         C.byref(loss) if want_loss else None, stream))
        return loss.value if want_loss else None

    def infer_batch(self, rho, dt, T, Y, accepted=None, stats=None, rtol=1e-5, atol=1e-13,
                    stream=None):
        """Surrogate + FP64 fallback in place on (T, Y); returns the number of fallback cells."""
        o = opts(atol=atol, rtol=rtol, do_reaction=1)
        nf = C.c_int()
        _check(lib().fcfd_sur_infer_batch(self.handle, C.byref(o), C.c_int(int(T.numel())),
                                          _vp(rho), _vp(dt), _vp(T), _vp(Y), _vp(accepted),
                                          C.byref(nf), _vp(stats), stream))
        return nf.value
