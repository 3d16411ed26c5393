"""Snapshot fields and restart on the GPU against the oracle: fcfd_eos_fields (set_eos_params +
write_hdf5_b's rows, SRC/block.f90:433-470, SRC/write_hdf5.f90:138-160) after reacting steps of the
micro-flow-reactor deck on a stretched mesh, written to HDF5 in the reference's layout, read back,
and restarted into a fresh solver (read_hdf5_restart + set_block_params) -- which then steps on in
agreement with the oracle restarted from the same file.

Bar: every output row within 1e-12 of its largest magnitude (x_m, dx and idx exactly); the restarted
state and the step after it as in test_gpu_aux (compare_fields, all halo columns)."""
import numpy as np
import pytest

from tests.test_gpu_aux import compare_all_columns, deck, names, step_pair
from tests.test_gpu_chem import REL
from tests.test_gpu_configs import ch4_air_cells, compare_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mech():
    from fortrancfd_amd.lib import Mechanism
    return Mechanism()


def reactor(mech, orc):
    from fortrancfd_amd import inputs
    from fortrancfd_amd.lib import Solver
    from oracle.pyoracle import Domain
    nm = names()
    smr, nfine, ncell, dx = (3, 1, 2, 1), 2, 16, 2e-5
    nblk = nfine + 2 + 1
    cells = ch4_air_cells(mech.kk, nblk * ncell, [])
    cells[:, 0] = np.linspace(450.0, 1500.0, nblk * ncell)
    cells[:, 2] = 0.3
    kw = dict(flag_acu=2, flag_weno=2, select_riemann=2, select_time=1, do_tran=1)
    fb = inputs.read_flow_boundary(deck("CH4_GRI3.0_MFR", "bound_flow.txt"), nm)
    w = inputs.read_wall_temperature(deck("CH4_GRI3.0_MFR", "wallT_373_1300.txt"))
    s = Solver(mech, nfine, ncell, dx, smr=smr, do_reaction=1, flag_lbound=1, flag_rbound=1, **kw)
    s.set_flow_boundary(**fb)
    s.set_wall_heat(**w)
    s.set_step(strang=False, hrr_each_step=True)
    d = Domain(orc, cells, nfine, ncell, dx, smr=smr, **kw)
    d.set_flow_boundary(1, 1, **fb)
    d.set_wall_heat(**w)
    return s, d, cells


def test_fields_snapshot_and_restart(mech, orc, tmp_path):
    from fortrancfd_amd import snapshot
    s, d, cells = reactor(mech, orc)
    s.init_cells(cells)
    print("reactor on a stretched mesh: 3 steps, then fields, snapshot, restart", flush=True)
    step_pair(s, d, 3, True, False)
    hrr, hrd = s.get_hrr()
    f_g = snapshot.eos_fields(s)
    idx0 = 1 - s.ncell * 2                              # two refinement blocks on the left
    f_o = d.eos_fields(hrr, hrd, idx0=idx0)
    for r in range(f_g.shape[0]):
        if r in (12, 13, 14):
            np.testing.assert_array_equal(f_g[r], f_o[r])
        else:
            sc = max(np.abs(f_o[r]).max(), 1e-300 if r < 15 else 1e-13)
            assert np.abs(f_g[r] - f_o[r]).max() <= REL * sc, (snapshot.FIELDS + names())[r]
    assert f_g[14][0] == idx0 and (np.diff(f_g[14]) == 1).all()
    assert (f_g[10] > 300).all() and np.abs(f_g[11]).max() > 0      # Tw and qTw with the wall heat on
    path = str(tmp_path / "results_nout000003.h5")
    snapshot.write_snapshot(path, s, 3, 3, 1e-9, 3e-9, create=True, fields=f_g)
    # a fresh solver of the same layout, restarted from the file, and the oracle from the same rows
    s2, d2, _ = reactor(mech, orc)
    nout, nt, t, _ = snapshot.restart(s2, path)
    assert (nout, nt, t) == (4, 4, 3e-9)
    rows, _ = snapshot.read_last(path, mech.names)
    d2.restart(np.concatenate([rows[[0, 1, 2, 4]], rows[15:]]))
    compare_all_columns(s2, d2, orc)
    dts = step_pair(s2, d2, 2, True, False)
    compare_fields(s2, d2, dts, dts, orc)
    compare_all_columns(s2, d2, orc)
