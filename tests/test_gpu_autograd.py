"""The fermion action S_F(x) = -flavors log det(1 + B..B) and its gradient through the C-ABI,
wrapped in a torch.autograd.Function (test harness only): torch.autograd.gradcheck on a small
lattice, and directional finite differences (Richardson-extrapolated) of S_F against <grad, v> at
16x16, Nt=32.  The gradient is -p_dot (organic_hmc.c:86-134: p_dot = -dS_F/dx)."""
import numpy as np
import pytest

from _gpu import gpu_context, say
from _oracle import OracleModel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def action_fn(ctx, flavors):
    class FermionAction(torch.autograd.Function):
        @staticmethod
        def forward(fctx, x):
            xn = x.detach().cpu().numpy().astype(np.float64)
            ld, _, minv = ctx.logdet_inv(ctx.M_hat(xn))
            fctx.save_for_backward(x)
            fctx.minv, fctx.ld = minv, ld
            return x.new_tensor(-flavors * ld)

        @staticmethod
        def backward(fctx, g):
            (x,) = fctx.saved_tensors
            pdot = ctx.force(x.detach().cpu().numpy().astype(np.float64), fctx.minv, fctx.ld)
            return g * x.new_tensor(-pdot)

    return FermionAction.apply


def test_gradcheck_small_lattice():
    om = OracleModel(2, 3, 4, beta=4 * 40 / 32)
    ctx = gpu_context(om)
    S = action_fn(ctx, om.flavors)
    x = torch.tensor(np.random.default_rng(1).standard_normal(om.dim) * 0.5, dtype=torch.float64, requires_grad=True)
    say("gradcheck 2x3x4")
    assert torch.autograd.gradcheck(S, (x,), eps=1e-6, atol=1e-7, rtol=1e-6)


@pytest.mark.parametrize("seed", [2, 3])
def test_directional_derivative(seed):
    om = OracleModel(16, 16, 32)
    ctx = gpu_context(om)
    S = action_fn(ctx, om.flavors)
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.standard_normal(om.dim) * 0.5, dtype=torch.float64, requires_grad=True)
    v = torch.tensor(rng.standard_normal(om.dim), dtype=torch.float64)
    v /= v.norm()
    s = S(x)
    (g,) = torch.autograd.grad(s, x)

    def central(eps):
        with torch.no_grad():
            return float((S(x + eps * v) - S(x - eps * v)) / (2 * eps))

    # Richardson extrapolation of two central differences: truncation O(eps^4); a step of 1e-3
    # keeps the rounding noise of S (|S| * 1e-16 / eps) far below the tolerance
    fd = (4 * central(1e-3) - central(2e-3)) / 3
    say("directional derivative 16x16x32", fd, float(g @ v), float(s))
    assert abs(fd - float(g @ v)) <= 1e-6 * abs(fd)
