"""Snapshots in the reference's HDF5 layout (libfcfd_h5.so) and restarts from them.

write_snapshot() is write_hdf5_b (SRC/write_hdf5.f90:48-330): the solver's per-cell output fields
(fcfd_eos_fields) as one group per output step.  restart() is read_hdf5_restart + set_block_params
(SRC/read_hdf5.f90:13-423, SRC/block.f90:394-431) onto a solver of the same layout.
"""
import ctypes as C
import os

import numpy as np

from .lib import LIB_PATH, FcfdError, _check, _p, lib

LIB_H5 = os.path.join(os.path.dirname(LIB_PATH), "libfcfd_h5.so")
FIELDS = ["rho", "u_x", "  T", "  p", " et", " ek", " ei", "HRR", "HRD", "Qrd", " Tw", "QTw", "x_m",
          " dx", "idx"]
_h5 = None


def h5lib():
    global _h5
    if _h5 is None:
        if not os.path.exists(LIB_H5):
            raise FcfdError(f"{LIB_H5} is missing: build it with `python -m fortrancfd_amd.build`")
        L = C.CDLL(LIB_H5)
        L.fcfd_h5_last_error.restype = C.c_char_p
        _h5 = L
    return _h5


def _h5check(rc):
    if rc != 0:
        raise FcfdError(f"hdf5: {h5lib().fcfd_h5_last_error().decode()}")


def _species_buffer(names):
    buf = bytearray(16 * len(names))
    for k, s in enumerate(names):
        b = s.encode()[:15]
        buf[16 * k:16 * k + len(b)] = b
    return bytes(buf)


def eos_fields(solver):
    """(15 + KK, ncells) output rows of the solver's current state (fcfd_eos_fields)."""
    n = solver.nblk * solver.ncell
    out = np.zeros((15 + solver.mech.kk, n))
    _check(lib().fcfd_eos_fields(solver.handle, _p(out)))
    return out


def write_snapshot(path, solver, nout, nt, dt, time_total, time_restart=None, cpu_total=0.0,
                   cpu_restart=0.0, create=True, fields=None):
    """Append output step nout to path (create: start a new file, as init_output does)."""
    f = eos_fields(solver) if fields is None else np.ascontiguousarray(fields, np.float64)
    names = solver.mech.names
    _h5check(h5lib().fcfd_h5_write_step(
        path.encode(), C.c_int(int(create)), C.c_int(nout), C.c_int(nt), C.c_double(dt),
        C.c_double(time_total), C.c_double(time_total if time_restart is None else time_restart),
        C.c_double(cpu_total), C.c_double(cpu_restart), C.c_int(f.shape[1]), C.c_int(len(names)),
        _species_buffer(names), _p(f)))


def read_last(path, names):
    """The last output group of path: (fields (15 + KK, n), attributes dict)."""
    n = C.c_int()
    _h5check(h5lib().fcfd_h5_last_size(path.encode(), C.byref(n)))
    f = np.zeros((15 + len(names), n.value))
    nout, nt, t, cpu = C.c_int(), C.c_int(), C.c_double(), C.c_double()
    _h5check(h5lib().fcfd_h5_read_last(path.encode(), n, C.c_int(len(names)), _species_buffer(names),
                                       _p(f), C.byref(nout), C.byref(nt), C.byref(t), C.byref(cpu)))
    return f, {"number of output": nout.value, "number of time step": nt.value,
               "time from 0s": t.value, "cpu time total": cpu.value}


def restart(solver, path):
    """Start solver from the last group of path as read_hdf5_restart does: the group's cells must
    sit where the solver's are (x_m and dx within 1e-7, the same idx, SRC/read_hdf5.f90:308-345),
    then the state is rebuilt from rho, u, T, e_total and X.  Returns the counters the reference
    continues from: (nout + 1, nt + 1, time, cpu time) (:250-253)."""
    f, attr = read_last(path, solver.mech.names)
    own = eos_fields(solver)
    if f.shape != own.shape:
        raise FcfdError(f"{path}: {f.shape[1]} cells, the solver has {own.shape[1]}")
    for row, name in ((12, "x_m"), (13, "dx")):
        if np.any(np.abs(f[row] - own[row]) > 1e-7):
            raise FcfdError(f"{name} in the input file does not match that in the hdf5 read file")
    if np.any(f[14] != own[14]):
        raise FcfdError("idx in the input file does not match that in the hdf5 read file")
    rows = np.concatenate([f[[0, 1, 2, 4]], f[15:]])
    _check(lib().fcfd_init_restart(solver.handle, _p(np.ascontiguousarray(rows))))
    return (attr["number of output"] + 1, attr["number of time step"] + 1, attr["time from 0s"],
            attr["cpu time total"])
