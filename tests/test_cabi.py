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


def test_launch_table_file_and_overrides(tmp_path):
    """The launch-shape table: entries of this host's architecture ("none" without a GPU) are read,
    others ignored; fcfd_set_launch checks its key and range; fcfd_load_tuning(NULL) goes back to
    the installed default table."""
    from fortrancfd_amd.lib import lib
    L = lib()
    get = lambda key, nsp=53: (lambda v: (L.fcfd_get_launch(key.encode(), nsp, C.byref(v)), v.value)[1])(C.c_int())
    f = tmp_path / "t.txt"
    f.write_text("# comment\nnone 53 flux_variant 2\ngfx000 53 stage_variant 3\nnone 11 chem_wg_per_cu 3  # x\n")
    assert L.fcfd_load_tuning(str(f).encode()) == 0
    assert get("flux_variant") == 2 and get("stage_variant") == 0
    assert get("chem_wg_per_cu", 11) == 3 and get("chem_wg_per_cu") == 5 and get("chem_four_wave_factor") == 8
    assert get("chem_lu") == 2
    assert L.fcfd_set_launch(b"stage_variant", 53, 3) == 0 and get("stage_variant") == 3
    assert L.fcfd_set_launch(b"stage_variant", 53, 4) != 0
    assert L.fcfd_set_launch(b"no_such_key", 53, 1) != 0
    assert L.fcfd_load_tuning(str(tmp_path / "missing.txt").encode()) != 0
    L.fcfd_load_tuning(None)
    assert get("stage_variant") in (0, 1, 2, 3)
