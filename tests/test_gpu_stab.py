"""The slice product with periodic FP64 stabilisation (ephmc_stable_logdet_inv, ephmc_config.stabilize).

* Where the plain product is well conditioned (the reference's regime), the stabilised log det and
  inverse equal the plain FP64 chain's (the reference's construct_M_hat + logDetM_and_inv_M_hat,
  organic_dense_matrix.c:251-265,547-607) to 1e-10, for QR intervals k = 1, 4, 8, 32.
* With FP32 slices and an FP64 QR every 8 slices, the log det and inverse stay within FP32 accuracy
  of FP64.
* For a field uniform in space every slice is diagonal in plane waves, so log det(1 + B..B) and the
  inverse are known in closed form at any conditioning.  At a dynamic range of e^45 (top
  eigenvalue of the product beyond 1/eps of FP64) the plain FP64 chain loses the "1 +" of the small
  modes, while the stabilised product (FP64 or FP32 slices) keeps log det and inverse accurate.
* A chain with stabilize = 8 takes the same accept/reject decisions as the plain FP64 chain."""
import numpy as np
import pytest

import _dense as D
from _gpu import gpu_context, rel, say
from _oracle import OracleModel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 4, 8, 32])
def test_stable_equals_plain_when_well_conditioned(k):
    say("stable vs plain, k =", k)
    om = OracleModel(6, 6, 32)
    ctx = gpu_context(om)
    x = np.random.default_rng(k).standard_normal(om.dim)
    ld0, s0, inv0 = ctx.logdet_inv(ctx.M_hat(x))
    ld1, s1, inv1 = ctx.stable_logdet_inv(x, 0, k)
    assert s1 == s0
    assert ld1 == pytest.approx(ld0, rel=1e-10)
    assert rel(inv1, inv0) < 1e-10
    ldo, _, invo = om.logdet_inv(om.M_hat(x))
    assert ld1 == pytest.approx(ldo, rel=1e-10)


def test_stable_fp32_slices():
    om = OracleModel(8, 8, 64)
    c32 = gpu_context(om, precision=1)
    x = np.random.default_rng(3).standard_normal(om.dim)
    ld0, _, inv0 = c32.logdet_inv(c32.M_hat(x))                 # FP64 plain
    ld1, _, inv1 = c32.stable_logdet_inv(x, 1, 8)               # FP32 slices, FP64 QR every 8
    say("FP32 slices stabilised: log det rel err %.2e, inverse %.2e" % (abs(ld1 - ld0) / abs(ld0), rel(inv1, inv0)))
    assert ld1 == pytest.approx(ld0, rel=1e-5)
    assert rel(inv1, inv0) < 1e-5


def _uniform_case(nt=96):
    """6x6 lattice, couplings in lattice units with per-slice eigenvalues of A in [-1.1, 0.7]:
    the product's eigenvalues reach e^44, where the plain FP64 product loses 26 % of the log det."""
    om = OracleModel(6, 6, nt)
    t = np.array([0.3, -0.09, 0.06])
    om.set_couplings(t=t, kappa=0.3 * t, mu=-0.2)
    u = np.random.default_rng(11).standard_normal((nt, 3)) * 0.3
    return om, u, D.uniform_field(6, 6, nt, u)


def _closed_form(om, u):
    l1, l2, nt = om.l1, om.l2, om.nt
    logp = np.zeros(l1 * l2)
    for tau in range(nt):
        eps, _ = D.plane_wave_eps(l1, l2, om.t, om.kappa, om.mu, u[tau])
        logp += np.log(np.polyval(D.POLY4[::-1], eps))
    logdet = np.logaddexp(0, logp).sum()                    # sum log(1 + prod p)
    # inverse: plane waves of the periodic lattice diagonalise every slice
    m1, m2 = np.meshgrid(np.arange(l1), np.arange(l2), indexing="ij")
    q1, q2 = 2 * np.pi * m1.ravel() / l1, 2 * np.pi * m2.ravel() / l2
    s1, s2 = np.arange(l1 * l2) % l1, np.arange(l1 * l2) // l1
    F = np.exp(1j * (np.outer(s1, q1) + np.outer(s2, q2))) / np.sqrt(l1 * l2)
    g = np.exp(-np.logaddexp(0, logp))                      # 1 / (1 + prod p)
    inv = (F * g) @ F.conj().T
    return logdet, inv.real, logp.max()


def test_stabilisation_where_the_plain_chain_fails():
    om, u, x = _uniform_case()
    ld_ref, inv_ref, top = _closed_form(om, u)
    say("uniform field: top log eigenvalue of the product %.1f, log det %.6g" % (top, ld_ref))
    assert top > 40                                        # beyond 1/eps of FP64
    n = om.ns
    c64 = gpu_context(om)
    c32 = gpu_context(om, precision=1)
    ld_plain, _, inv_plain = c64.logdet_inv(c64.M_hat(x))
    ld_s64, _, inv_s64 = c64.stable_logdet_inv(x, 0, 8)
    ld_s32, _, inv_s32 = c32.stable_logdet_inv(x, 1, 4)
    Minv = lambda a: a.reshape(n, n).T                      # column-major chunks -> numpy
    errs = {"plain64": (abs(ld_plain - ld_ref) / ld_ref, np.abs(Minv(inv_plain) - inv_ref).max()),
            "stab64": (abs(ld_s64 - ld_ref) / ld_ref, np.abs(Minv(inv_s64) - inv_ref).max()),
            "stab32": (abs(ld_s32 - ld_ref) / ld_ref, np.abs(Minv(inv_s32) - inv_ref).max())}
    say("log det / inverse errors", errs)
    assert errs["stab64"][0] < 1e-10 and errs["stab64"][1] < 1e-8
    assert errs["stab32"][0] < 1e-4 and errs["stab32"][1] < 1e-3
    # the plain product has lost what the stabilised one keeps
    assert errs["plain64"][1] > 100 * errs["stab64"][1]


@pytest.mark.parametrize("precision", [0, 1])
def test_chain_with_stabilize_takes_the_plain_decisions(precision):
    """A chain whose Metropolis log det (FP64) and molecular-dynamics inverse (the context's slice
    precision) come from the product stabilised every 8 slices, against the plain FP64 chain with the
    same random inputs: the same decisions; log det within 1e-8 (FP64) or 1e-5 (FP32 slices)."""
    om = OracleModel(6, 6, 32)
    plain = gpu_context(om, steps=10, checkpoint=0)
    stab = gpu_context(om, steps=10, checkpoint=0, precision=precision, stabilize=8)
    rng = np.random.default_rng(5)
    init = rng.standard_normal(om.dim)
    plain.chain_init(init)
    stab.chain_init(init)
    tol = 1e-8 if precision == 0 else 1e-5
    assert stab.chain_logdet() == pytest.approx(plain.chain_logdet(), rel=1e-10)
    for it in range(8):
        nrm, uu = rng.standard_normal(om.dim), float(rng.random())
        a, b = plain.trajectory(nrm, uu), stab.trajectory(nrm, uu)
        say("traj %d: plain dH %.6g acc %d, stabilised dH %.6g acc %d" % (it, a.dH, a.accepted, b.dH, b.accepted))
        assert a.accepted == b.accepted, it
        assert b.logdet == pytest.approx(a.logdet, rel=tol), it
