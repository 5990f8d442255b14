"""tofe_compliance_noalloc: the context-free, allocation-free compliance and element
sensitivities, called through the C ABI and checked bit for bit against tofe_compliance_dev
(c) and tofe_sensitivity (g) of a context on the same U, F and design, on two calls in a
row -- the second on a smaller mesh, so it reuses the device scratch the first call left."""
import ctypes as C

import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu


def _reference(tofe, nel, seed):
    """(U, F, dEdz, c, g) from a context: a solved cantilever with a random design."""
    import torch

    pb = configs.cantilever(nel, seed=seed, lo=0.05, hi=1.0, mg_levels=2)
    with tofe.Context(nel, mg_levels=2, fixed_dof=pb.fixed_dof) as ctx:
        ctx.set_design(pb.x)
        ctx.assemble()
        U, its, reason = ctx.solve(pb.F, rtol=1e-10)
        assert reason == tofe.CONVERGED_RTOL
        dEdz = ctx.field(tofe.FIELD_DEDZ)
        Ud = torch.from_numpy(U).cuda()
        Fd = torch.from_numpy(pb.F).cuda()
        c = ctx.compliance_dev(Ud.data_ptr(), Fd.data_ptr())
        g = ctx.sensitivity(U)
    return U, pb.F, dEdz, c, g


def _noalloc(tofe, nel, U, F, dEdz, want_c=True, want_g=True):
    import torch

    Ud = torch.from_numpy(U).cuda()
    Fd = torch.from_numpy(F).cuda()
    Dd = torch.from_numpy(dEdz).cuda()
    gd = torch.full((dEdz.size,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    c = C.c_double(np.nan)
    rc = tofe.lib().tofe_compliance_noalloc(
        nel[0], nel[1], nel[2], 1.0, 1.0, 1.0, 1.0, 0.3, Ud.data_ptr(), Fd.data_ptr(),
        Dd.data_ptr(), C.byref(c) if want_c else None, gd.data_ptr() if want_g else None)
    assert rc == tofe.TOFE_OK, tofe.lib().tofe_last_error(None)
    return c.value, gd.cpu().numpy()


def test_noalloc_matches_context_entry_points_twice(tofe):
    big, small = (12, 6, 6), (8, 4, 4)
    for nel, seed in ((big, 5), (small, 6)):
        print(f"noalloc: mesh {nel} seed {seed}", flush=True)
        U, F, dEdz, c_ref, g_ref = _reference(tofe, nel, seed)
        c, g = _noalloc(tofe, nel, U, F, dEdz)
        assert c == c_ref
        np.testing.assert_array_equal(g, g_ref)
        assert np.all(g <= 0)
    # one output at a time leaves the other untouched
    c_only, g_untouched = _noalloc(tofe, small, U, F, dEdz, want_g=False)
    assert c_only == c_ref and np.all(np.isnan(g_untouched))
    c_unset, g_only = _noalloc(tofe, small, U, F, dEdz, want_c=False)
    assert np.isnan(c_unset)
    np.testing.assert_array_equal(g_only, g_ref)


def test_noalloc_rejects_oversize_and_bad_arguments(tofe):
    L = tofe.lib()
    c = C.c_double()
    # 2000^3 elements: beyond the 2^31 - 1 maximum, rejected before any device work
    rc = L.tofe_compliance_noalloc(2000, 2000, 2000, 1.0, 1.0, 1.0, 1.0, 0.3, 1, 1, 1,
                                   C.byref(c), 1)
    assert rc == tofe.TOFE_ERR_ARG
    assert b"maximum" in L.tofe_last_error(None)
    rc = L.tofe_compliance_noalloc(4, 2, 2, 1.0, 1.0, 1.0, 1.0, 0.3, None, None, None,
                                   C.byref(c), None)
    assert rc == tofe.TOFE_ERR_ARG
    rc = L.tofe_compliance_noalloc(0, 2, 2, 1.0, 1.0, 1.0, 1.0, 0.3, 1, 1, 1, C.byref(c), None)
    assert rc == tofe.TOFE_ERR_ARG
