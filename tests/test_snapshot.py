"""The HDF5 snapshot layout of libfcfd_h5.so (no GPU): what write_hdf5_b writes
(SRC/write_hdf5.f90:48-330) as an independent reader, HDF5's own h5dump, sees it -- the group name
with the time in Fortran E11.5, the datasets with the reference's 3-character names (leading blanks
included) and the species names, the seven attributes with their types -- and what the reader
(read_hdf5_restart's selection of the last group, SRC/read_hdf5.f90:118-139) gives back."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import golden

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def names():
    return list(golden("mech")["names"])


@pytest.fixture(scope="module")
def snap():
    from fortrancfd_amd import snapshot
    try:
        snapshot.h5lib()
    except Exception as e:                      # pragma: no cover - only without HDF5
        pytest.skip(f"libfcfd_h5.so unavailable: {e}")
    return snapshot


class FakeSolver:
    """What write_snapshot needs from a solver when the fields are given."""
    class M:
        pass

    def __init__(self):
        self.mech = FakeSolver.M()
        self.mech.names = names()
        self.mech.kk = len(self.mech.names)


def test_fortran_e11_5_group_names_and_layout(snap, tmp_path):
    kk = len(names())
    rng = np.random.default_rng(5)
    n = 37
    f = rng.uniform(-1.0, 1.0, (15 + kk, n))
    path = str(tmp_path / "results_nout000000.h5")
    s = FakeSolver()
    snap.write_snapshot(path, s, 0, 0, 1e-9, 0.0, create=True, fields=f)
    f2 = f * 2.0
    snap.write_snapshot(path, s, 1, 1234, 2.5e-9, 1.23456789e-5, time_restart=1.0e-5, cpu_total=3.5,
                        cpu_restart=1.5, create=False, fields=f2)
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    hdr = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    assert 'GROUP "nout000000_time0.00000E+00"' in hdr
    assert 'GROUP "nout000001_time0.12346E-04"' in hdr
    for name in snap.FIELDS + names():
        assert f'DATASET "{name}"' in hdr, name
    for a in ("dt", "time from 0s", "time from restart", "cpu time total", "cpu time restart"):
        assert f'ATTRIBUTE "{a}"' in hdr
    assert "chem directory name" not in hdr
    assert hdr.count("H5T_IEEE_F64LE") == 2 * (15 + kk + 5)
    assert hdr.count("H5T_STD_I32LE") == 2 * 2
    assert hdr.count(f"DATASPACE  SIMPLE {{ ( {n} ) / ( {n} ) }}") == 2 * (15 + kk)
    val = subprocess.run([H5DUMP, "-y", "-m", "%.17g", "-d", "/nout000001_time0.12346E-04/  T", path],
                         capture_output=True, text=True, check=True).stdout
    body = val[val.index("DATA {") + 6:]
    body = body[:body.index("}")]
    nums = np.array([float(t) for t in body.replace(",", " ").split()])
    np.testing.assert_array_equal(nums, f2[2])          # the dataset "  T" holds row 2
    got, attr = snap.read_last(path, names())
    np.testing.assert_array_equal(got, f2)
    assert attr == {"number of output": 1, "number of time step": 1234, "time from 0s": 1.23456789e-5,
                    "cpu time total": 3.5}


def test_reader_fills_missing_species_with_zero(snap, tmp_path):
    kk = len(names())
    f = np.arange((15 + kk) * 4, dtype=float).reshape(15 + kk, 4)
    path = str(tmp_path / "r.h5")
    s = FakeSolver()
    s.mech.names = names()[:10]
    snap.write_snapshot(path, s, 3, 7, 1e-9, 2e-6, create=True, fields=f[:25])
    got, attr = snap.read_last(path, names())
    np.testing.assert_array_equal(got[:25], f[:25])
    assert not got[25:].any() and attr["number of output"] == 3
