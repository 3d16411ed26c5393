"""The Fortran binding (fortrancfd_amd/fortran/fcfd.f90): a Fortran program drives the reference's
loop (calc_dt, solve_fluid, refresh) through ISO_C_BINDING calls into libfcfd_hip.so and must
reproduce the reference's golden runs bit for bit, like the Python binding does."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import REPO, golden

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(REPO, "tests", "fortran", "_build", "sod_driver")


@pytest.mark.parametrize("name", ["sod_weno5js_hllc", "sod_rk4_lxf"])
def test_fortran_driver_reproduces_reference_run(tmp_path, name):
    from fortrancfd_amd.lib import GRI30
    assert os.path.exists(DRIVER), "build the Fortran test driver first (__graft_entry__.build())"
    g = golden("run_" + name)
    nblk, ncell, ns, lv, acu, wen, mus, riem, tim, cflf = [int(x) for x in g["params"]]
    ints = np.array([nblk, ncell, ns, tim, riem, acu, mus, wen, cflf], np.int32)
    dbl = np.array([float(g["dx"]), float(g["cfl"]), float(g["dt"]) if "dt" in g else 0.0])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(ints.tobytes())
        f.write(dbl.tobytes())
        f.write(np.ascontiguousarray(g["cells"], np.float64).tobytes())
    print(f"fortran driver {name}", flush=True)
    r = subprocess.run([DRIVER, GRI30, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, np.float64)
    dts = raw[:ns]
    rest = raw[ns:].reshape(nblk, 2, ncell + 2 * lv, -1)
    np.testing.assert_array_equal(dts, g["dts"])
    np.testing.assert_array_equal(rest[:, 0], g["Qc"])
    np.testing.assert_array_equal(rest[:, 1], g["qf"])
