"""ephmc_cc_correlator_static, the allocation-free variant of ephmc_cc_correlator (transfer
matrices in a static device array): on two calls in a row, the second on a smaller lattice, its
device-side result equals the allocating entry point's bit for bit (the same kernels in the same
order), and sizes beyond the static array are refused with EPHMC_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

from _gpu import ephmc, gpu_context, say
from _oracle import OracleModel

pytestmark = pytest.mark.gpu


def _static(ctx, dx, dm, ld):
    E = ephmc()
    out = ctx.array(ctx.nt * 9)
    rc = E.LIB.ephmc_cc_correlator_static(ctx.h, dx.ptr, dm.ptr, ld, out.ptr)
    return rc, out


@pytest.mark.parametrize("sizes", [((8, 8, 16), (4, 4, 8)), ((6, 5, 12), (2, 3, 4))])
def test_static_correlator_equals_allocating_one(sizes):
    for k, (l1, l2, nt) in enumerate(sizes):
        say("static correlator call %d: %dx%d, Nt=%d" % (k + 1, l1, l2, nt))
        om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
        ctx = gpu_context(om)
        x = np.random.default_rng(l1 * 100 + nt).standard_normal(om.dim)
        ld, _, Minv = ctx.logdet_inv(ctx.M_hat(x))
        dx, dm = ctx.array(om.dim, x), ctx.array(om.ns * om.ns, Minv)
        rc, out = _static(ctx, dx, dm, ld)
        assert rc == 0, ephmc().LIB.ephmc_last_error().decode()
        ctx.sync()
        g_static = out.download().reshape(nt, 3, 3)
        g_ref = ctx.cc_correlator(dx, dm, ld)
        np.testing.assert_array_equal(g_static, g_ref)
        assert np.abs(g_ref).max() > 0
        ctx.close()


def test_static_correlator_refuses_oversized():
    om = OracleModel(16, 16, 64)
    ctx = gpu_context(om)
    dx, dm = ctx.array(om.dim), ctx.array(om.ns * om.ns)
    rc, _ = _static(ctx, dx, dm, 0.0)
    assert rc == -1
    assert b"static scratch" in ephmc().LIB.ephmc_last_error()
    ctx.close()
