"""Error paths of the plan API on the device, forced from the host: a tune change whose
device allocations fail part way leaves the plan's tables, shapes and results as they
were (every kernel launched afterwards reads live tables)."""
import pytest

pytestmark = pytest.mark.gpu

GEO = (64, 91, 72, 72, 1.0, 1.0, 250.0)


def test_failed_tune_change_keeps_plan():
    import torch
    from imgrecsrc_amd import _lib
    from imgrecsrc_amd._lib import CtprjError
    from imgrecsrc_amd.projector import Projector
    L = _lib.lib()
    P = Projector([GEO, (48, 71, 60, 120, 1.0, 1.0, 0.0)])
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(P.img_elems, device="cuda", generator=g)
    y = torch.rand(P.sino_elems, device="cuda", generator=g)
    t0 = P.get_tune()
    f0, a0 = P.forward(x), P.adjoint(y)
    # each change rebuilds tables: flags (geometry + both work lists), a general shape
    # (one work list) and splits (partial-sum space)
    changes = [dict(flags=0), dict(flags=t0["flags"] | 2), dict(fwd_block=64, adj_tile=8),
               dict(adj_split=3, fwd_split=2)]
    failed = 0
    for change in changes:
        for k in range(1, 16):
            assert L.ctprj_debug_fail_alloc(k) == 0
            try:
                P.set_tune(**change)
            except CtprjError as e:
                assert e.code == _lib.CTPRJ_ENOMEM, str(e)
                failed += 1
                assert P.get_tune() == t0
            finally:
                assert L.ctprj_debug_fail_alloc(0) == 0
            if P.get_tune() != t0:  # fewer than k allocations: the change went through
                P.set_tune(**t0)
            assert P.get_tune() == t0
            assert torch.equal(P.forward(x), f0) and torch.equal(P.adjoint(y), a0), (change, k)
    assert failed >= 8
