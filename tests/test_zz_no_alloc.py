"""fcfd_chem_batch_noalloc (the allocation-free variant of fcfd_chem_batch, scratch in a static
device array) against fcfd_chem_batch through the C-ABI: two calls in a row on the same GPU, the
second with fewer cells, each bitwise equal to the allocating entry point on the same inputs
(T, Y and the per-cell counters); n above the limit fails with a clear error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, mech, o, T, Y, rho, dt, st):
    ptr = lambda t: C.c_void_p(t.data_ptr())
    return fn(mech.handle, C.byref(o), C.c_int(int(T.numel())), ptr(rho), ptr(dt), ptr(T), ptr(Y),
              ptr(st), None)


def test_noalloc_matches_chem_batch():
    import torch
    from fortrancfd_amd import workloads
    from fortrancfd_amd.lib import FcfdError, Mechanism, _check, lib, opts
    mech = Mechanism()
    L = lib()
    _check(L.fcfd_mech_prepare(mech.handle))
    o = opts(do_reaction=1)
    rho, dt, T, Y = workloads.chemistry_batch(mech, 3000, seed=workloads.SEED + 17)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    for n in (3000, 700):                        # the second call is smaller than the first
        print(f"allocation-free batch of {n} cells", flush=True)
        r, d = dev(rho[:n]), dev(dt[:n])
        Ta, Ya = dev(T[:n]), dev(Y[:n].T)
        Tb, Yb = Ta.clone(), Ya.clone()
        sa = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
        sb = torch.full((n, 4), -1, dtype=torch.int32, device="cuda")
        _check(_call(L.fcfd_chem_batch, mech, o, Ta, Ya, r, d, sa))
        _check(_call(L.fcfd_chem_batch_noalloc, mech, o, Tb, Yb, r, d, sb))
        torch.cuda.synchronize()
        assert torch.equal(Ta, Tb) and torch.equal(Ya, Yb) and torch.equal(sa, sb)
        st = sa.cpu().numpy()
        assert (st[:, 2] > 0).any() and (st[:, 3] == 0).all()   # implicit work was done, nothing failed
        assert not torch.equal(Tb, dev(T[:n]))
    big = (1 << 20) + 1
    L.fcfd_chem_batch_noalloc.restype = C.c_int
    rc = L.fcfd_chem_batch_noalloc(mech.handle, C.byref(o), C.c_int(big), C.c_void_p(1), C.c_void_p(1),
                                   C.c_void_p(1), C.c_void_p(1), None, None)
    assert rc != 0 and b"allocation-free limit" in L.fcfd_last_error()
    with pytest.raises(FcfdError):
        _check(rc)
