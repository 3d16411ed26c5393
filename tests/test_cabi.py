"""The C-ABI library (no GPU needed): it loads, exports every entry point include/fcfd.h declares,
reads the GRI-3.0 link files and rejects bad arguments with an error code and a message."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tests.conftest import REPO, golden


def declared_functions(header="fcfd.h"):
    src = open(os.path.join(REPO, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fcfd_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    from fortrancfd_amd.lib import lib
    L = lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_h5_library_exports_every_declared_symbol():
    from fortrancfd_amd import snapshot
    L = snapshot.h5lib()
    names = declared_functions("fcfd_h5.h")
    assert len(names) == 4
    assert not [n for n in names if not hasattr(L, n)]


def test_mechanism_load_and_dims():
    from fortrancfd_amd.lib import Mechanism
    m = Mechanism()
    assert (m.kk, m.ii, m.nequ) == (53, 325, 56)


def test_bad_paths_and_options_fail_loudly():
    from fortrancfd_amd.lib import FcfdError, Mechanism, Solver, lib
    with pytest.raises(FcfdError, match="cannot read"):
        Mechanism("/nonexistent")
    m = Mechanism()
    with pytest.raises(FcfdError, match="flag_weno"):
        Solver(m, 1, 16, 1e-3, flag_acu=2, flag_weno=7)
    with pytest.raises(FcfdError, match="select_time"):
        Solver(m, 1, 16, 1e-3, select_time=3)
    with pytest.raises(FcfdError, match="flag_lbound"):
        Solver(m, 1, 16, 1e-3, flag_lbound=2)
    with pytest.raises(FcfdError, match="inflow / outflow"):
        Solver(m, 1, 16, 1e-3, select_time=2, flag_rbound=1)
    h = C.c_void_p()
    assert lib().fcfd_mech_load(None, None, C.byref(h)) == -2


def test_species_names_from_link_file():
    import bench
    from fortrancfd_amd.lib import GRI30
    assert bench.species_names(GRI30) == list(golden("mech")["names"])
