"""Random lattice shapes against the oracle (fixed seed): l1, l2 in 1..40 (N <= 600), Nt in 2..24,
Taylor orders 2..8, checkpoint intervals 1, 4 and the default, fields of two scales.  M_hat within
1e-13, log det within 1e-10, the FP64 force within 1e-10 and the mixed force within 1e-5 (all
relative), the force fed the oracle's M_hat^{-1}."""
import numpy as np
import pytest

from _gpu import gpu_context, rel, say
from _oracle import OracleModel

pytestmark = pytest.mark.gpu


def shapes(n, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        l1, l2 = int(rng.integers(1, 41)), int(rng.integers(1, 41))
        while l1 * l2 > 600:
            l2 = max(1, l2 // 2)
        out.append((l1, l2, int(rng.choice([2, 3, 4, 6, 8, 12, 16, 24])), int(rng.choice([2, 4, 6, 8])),
                     float(rng.choice([0.3, 1.0])), int(rng.choice([0, 1, 4])), int(rng.integers(1 << 30))))
    return out


@pytest.mark.parametrize("l1,l2,nt,taylor,scale,ck,seed", shapes(12))
def test_random_shape_vs_oracle(l1, l2, nt, taylor, scale, ck, seed):
    say("random shape", l1, l2, nt, taylor, scale, ck)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32, taylor=taylor)
    x = np.random.default_rng(seed).standard_normal(om.dim) * scale
    Mo = om.M_hat(x)
    ldo, sgo, Minv = om.logdet_inv(Mo)
    Fo = om.force(x, Minv, ldo)
    ctx = gpu_context(om, precision=0, checkpoint=ck, taylor_order=taylor)
    Mg = ctx.M_hat(x)
    assert rel(Mg, Mo) < 1e-13
    ld, sg, _ = ctx.logdet_inv(Mg)
    assert sg == sgo and abs(ld - ldo) <= 1e-10 * max(1.0, abs(ldo))
    assert rel(ctx.force(x, Minv, ldo), Fo) < 1e-10
    ctx.close()
    mixed = gpu_context(om, precision=1, checkpoint=ck, taylor_order=taylor)
    assert rel(mixed.force(x, Minv, ldo), Fo) < 1e-5
    mixed.close()
