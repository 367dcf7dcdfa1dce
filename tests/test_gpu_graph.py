"""ephmc_config.graph: the trajectory body captured once as a hipGraph and replayed gives the
same records and chain state as the eager launches (same kernels, same order)."""
import numpy as np
import pytest

from _gpu import gpu_context, say
from _oracle import OracleModel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("l1,l2,nt,precision", [(4, 4, 16, 0), (6, 6, 32, 1), (16, 16, 32, 0)])
def test_graph_replay_matches_eager(l1, l2, nt, precision):
    say("graph vs eager", l1, l2, nt, precision)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    rng = np.random.default_rng(l1 + nt)
    init = rng.standard_normal(om.dim)
    draws = [(rng.standard_normal(om.dim), float(rng.random())) for _ in range(8)]
    out = []
    for graph in (0, 1):
        ctx = gpu_context(om, precision=precision, steps=5, graph=graph)
        ctx.chain_init(init)
        recs = [ctx.trajectory(nv, u) for nv, u in draws]
        out.append(([(r.accepted, r.dH, r.logdet, r.energy) for r in recs], ctx.chain_x(), ctx.chain_minv()))
        ctx.close()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])
