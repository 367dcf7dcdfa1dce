"""ephmc_autotune: the launch shapes it picks keep the results.  FP64 force against the oracle
to 1e-10 and mixed to 1e-5 after tuning; a chain replayed as a hipGraph after tuning keeps the
accept/reject decisions and dH (1e-9) of an untuned eager chain on the same inputs."""
import numpy as np
import pytest

from _gpu import gpu_context, rel, say
from _oracle import OracleModel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("l1,l2,nt,k", [(3, 3, 8, 1), (6, 5, 16, 4), (16, 16, 32, 1), (32, 32, 8, 8)])
def test_autotune_keeps_force(l1, l2, nt, k, precision):
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    ctx = gpu_context(om, precision=precision, checkpoint=k)
    x = np.random.default_rng(l1 + nt).standard_normal(om.dim) * 0.5
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    res = ctx.autotune()
    # the fp64 log det after tuning (the Metropolis chain's shape) against the oracle
    ld_t, _, _ = ctx.logdet_inv(ctx.M_hat(x))
    assert abs(ld_t - ld) <= 1e-10 * max(1.0, abs(ld))
    say("autotune", l1, l2, nt, k, precision, res)
    assert res["chain_cols"] in (1, 2, 4, 8) and res["adjoint_cols"] in (1, 2, 4)
    assert res["chain_strip"] in (0, 2, 4) and res["chain64_strip"] in (0, 2, 4)
    assert res["chain_strip"] == 0 or l1 % res["chain_strip"] == 0
    assert res["chain_ms"] > 0 and res["adjoint_ms"] > 0 and res["chain64_ms"] > 0
    if precision == 0:
        assert (res["chain64_cols"], res["chain64_strip"]) == (res["chain_cols"], res["chain_strip"])
    assert rel(ctx.force(x, Minv, ld), om.force(x, Minv, ld)) < (1e-10 if precision == 0 else 1e-5)


def test_autotuned_graph_chain_matches_eager():
    om = OracleModel(8, 8, 32, beta=40.0)
    rng = np.random.default_rng(5)
    init = rng.standard_normal(om.dim)
    draws = [(rng.standard_normal(om.dim), float(rng.random())) for _ in range(6)]
    ref = gpu_context(om, steps=6)
    tuned = gpu_context(om, steps=6, graph=1)
    ref.chain_init(init)
    tuned.chain_init(init)
    for i, (nrm, u) in enumerate(draws):
        if i == 3:
            say("autotune mid-chain", tuned.autotune())
        a, b = ref.trajectory(nrm, u), tuned.trajectory(nrm, u)
        assert a.accepted == b.accepted
        assert abs(a.dH - b.dH) < 1e-9
