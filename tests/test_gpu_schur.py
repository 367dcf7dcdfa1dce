"""SCHUR_FORM_INV (organic_dense_matrix.c:556-603, organic_hmc.c:113-114, organic_dense_matrix.c:270-278).

* ephmc_schur is a real Schur decomposition: Z orthogonal, Z T Z^T = B, T quasi-upper-triangular
  with its 2 x 2 blocks in standard form, the eigenvalues those of numpy.linalg.eigvals, on random
  nonsymmetric matrices and on B = M_hat - 1 of the model.
* Where nothing is zeroed (the reference's couplings: eigenvalues of B in (0, ~1]), the Schur-form
  log det and inverse equal the plain path's to 1e-10.
* Where the spectrum is wide (top eigenvalue e^44 of the product for a field uniform in space), the
  threshold max_ev DBL_EPSILON / 1e-3 zeroes every eigenvalue below it: log det = sum of
  log(1 + P_q) over the plane waves with P_q above the threshold, inverse = plane-wave sum with
  1 / (1 + P_q) there and 1 elsewhere (closed form).
* The force clip: projections of the adjoint sweep and of the constant-force branch are clipped to
  [-2, 2], checked against the unclipped force on an input M_inv scaled to make projections large.
* A seeded chain with SCHUR_FORM_INV takes the plain chain's decisions."""
import numpy as np
import pytest

import _dense as D
from _gpu import gpu_context, rel, say
from _oracle import OracleModel

pytestmark = pytest.mark.gpu


def check_schur(ctx, B):
    n = B.shape[0]
    Zf, Tf, wr, wi = ctx.schur(B.T.copy().ravel())
    Z, T = Zf.reshape(n, n).T, Tf.reshape(n, n).T
    np.testing.assert_allclose(Z.T @ Z, np.eye(n), atol=1e-12 * n)
    np.testing.assert_allclose(Z @ T @ Z.T, B, atol=1e-12 * n * max(1.0, np.abs(B).max()))
    assert np.all(np.tril(T, -2) == 0)
    i = 0
    while i < n:
        if i + 1 < n and T[i + 1, i] != 0:
            assert T[i, i] == T[i + 1, i + 1] and T[i, i + 1] * T[i + 1, i] < 0, i
            assert i + 2 >= n or T[i + 2, i + 1] == 0
            i += 2
        else:
            i += 1
    ev = np.sort_complex(wr + 1j * wi)
    ref = np.sort_complex(np.linalg.eigvals(B))
    np.testing.assert_allclose(ev, ref, atol=1e-9 * max(1.0, np.abs(ref).max()))
    return Z, T


@pytest.mark.parametrize("n", [1, 2, 5, 17, 64, 100])
def test_schur_random(n):
    say("schur random n =", n)
    om = OracleModel(n, 1, 1)
    ctx = gpu_context(om)
    B = np.random.default_rng(n).standard_normal((n, n))
    check_schur(ctx, B)


def test_schur_of_the_model_matrix():
    om = OracleModel(6, 6, 32)
    ctx = gpu_context(om)
    x = np.random.default_rng(2).standard_normal(om.dim)
    B = om.M_hat(x).reshape(om.ns, om.ns).T - np.eye(om.ns)
    check_schur(ctx, B)


def test_schur_inverse_equals_plain_without_zeroing():
    om = OracleModel(6, 6, 32)
    plain = gpu_context(om)
    schur = gpu_context(om, schur_form_inv=1)
    x = np.random.default_rng(4).standard_normal(om.dim)
    M = plain.M_hat(x)
    ld0, s0, inv0 = plain.logdet_inv(M)
    ld1, s1, inv1 = schur.logdet_inv(M)
    assert ld1 == pytest.approx(ld0, rel=1e-10) and s1 == s0
    assert rel(inv1, inv0) < 1e-10


def test_schur_threshold_closed_form():
    """The wide spectrum of a uniform field (the stabilisation test's case): every eigenvalue of
    B below eps = max_ev DBL_EPSILON / 1e-3 is zeroed in T."""
    nt = 96
    om = OracleModel(6, 6, nt)
    t = np.array([0.3, -0.09, 0.06])
    om.set_couplings(t=t, kappa=0.3 * t, mu=-0.2)
    u = np.random.default_rng(11).standard_normal((nt, 3)) * 0.3
    x = D.uniform_field(6, 6, nt, u)
    logp = np.zeros(om.ns)
    for tau in range(nt):
        eps_q, _ = D.plane_wave_eps(6, 6, om.t, om.kappa, om.mu, u[tau])
        logp += np.log(np.polyval(D.POLY4[::-1], eps_q))
    thr = np.exp(logp.max()) * 2.220446049250313e-16 / 1e-3
    say("threshold %.3g, log P range %.1f .. %.1f" % (thr, logp.min(), logp.max()))
    assert thr > 1
    margin = np.min(np.abs(logp - np.log(thr)))
    assert margin > 0.5                           # no eigenvalue near the threshold
    keep = logp > np.log(thr)
    ld_ref = np.logaddexp(0, logp[keep]).sum()
    g = np.where(keep, np.exp(-np.logaddexp(0, logp)), 1.0)
    m1, m2 = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")
    q1, q2 = 2 * np.pi * m1.ravel() / 6, 2 * np.pi * m2.ravel() / 6
    s1, s2 = np.arange(36) % 6, np.arange(36) // 6
    F = np.exp(1j * (np.outer(s1, q1) + np.outer(s2, q2))) / 6
    inv_ref = ((F * g) @ F.conj().T).real
    schur = gpu_context(om, schur_form_inv=1)
    M = schur.M_hat(x)
    ld, _, inv = schur.logdet_inv(M)
    say("Schur log det %.10g, closed form %.10g" % (ld, ld_ref))
    assert ld == pytest.approx(ld_ref, rel=1e-5)
    np.testing.assert_allclose(inv.reshape(36, 36).T, inv_ref, atol=1e-5)


@pytest.mark.parametrize("constant", [0, 1])
def test_force_clip(constant):
    """With M_inv scaled up the bond projections exceed 2: the Schur-mode force equals the plain
    force with each projection clipped to [-2, 2] (the plain force is -kappa_k flavors times the
    projection)."""
    om = OracleModel(4, 4, 8, beta=8 * 40 / 32)
    flags = {"constant_force": constant}
    plain = gpu_context(om, **flags)
    schur = gpu_context(om, schur_form_inv=1, **flags)
    x = np.random.default_rng(6).standard_normal(om.dim)
    ld, _, Minv = plain.logdet_inv(plain.M_hat(x))
    n = om.ns
    big = Minv.reshape(n, n) * 40
    if constant:
        # the Schur constant-force branch uses QM1P = 1 - M_inv: projection 2 delta - (M_inv + M_inv^T)
        nb = om.nnt[:6 * n].reshape(n, 6)
        F = schur.force(x, big.ravel(), ld).reshape(om.nt, n, 3)
        ref = np.zeros((n, 3))
        for i in range(n):
            for k in range(3):
                j = nb[i, 2 * k]
                proj = (2.0 if i == j else 0.0) - (big[j, i] + big[i, j])
                ref[i, k] = -om.kappa[k] * om.flavors * np.clip(proj, -2, 2)
        for tau in range(om.nt):
            np.testing.assert_allclose(F[tau], ref, rtol=1e-13, atol=1e-15)
        assert np.abs(ref).max() > 0
    else:
        Fp = plain.force(x, big.ravel(), ld).reshape(om.nt, n, 3)
        Fs = schur.force(x, big.ravel(), ld).reshape(om.nt, n, 3)
        f = -np.asarray(om.kappa) * om.flavors
        proj = Fp / f
        assert np.abs(proj).max() > 2                 # the clip is exercised
        np.testing.assert_allclose(Fs, f * np.clip(proj, -2, 2), rtol=1e-12, atol=1e-14)


def test_chain_with_schur_takes_the_plain_decisions():
    om = OracleModel(6, 6, 32)
    plain = gpu_context(om, steps=10, checkpoint=0)
    schur = gpu_context(om, steps=10, checkpoint=0, schur_form_inv=1)
    rng = np.random.default_rng(8)
    init = rng.standard_normal(om.dim)
    plain.chain_init(init)
    schur.chain_init(init)
    assert schur.chain_logdet() == pytest.approx(plain.chain_logdet(), rel=1e-10)
    for it in range(4):
        nrm, uu = rng.standard_normal(om.dim), float(rng.random())
        a, b = plain.trajectory(nrm, uu), schur.trajectory(nrm, uu)
        assert a.accepted == b.accepted, it
        assert b.logdet == pytest.approx(a.logdet, rel=1e-8), it


def test_organic_main_with_schur_flag():
    """organic_main with the SCHUR_FORM_INV token in input.txt: the same accept/reject column as the
    run without it, every results column within 1e-8 relative (nothing is zeroed at these
    couplings, and no bond projection reaches the clip)."""
    import os
    import subprocess
    import tempfile
    from test_gpu_host import INPUT
    from test_host_layer import MAIN
    base = INPUT.format(l1=6, l2=6, nt=16, beta=20.0, steps=10, therm=2, meas=6, freq=1, seed=99)
    with tempfile.TemporaryDirectory() as d:
        rows = []
        for tag, extra in (("plain", ""), ("schur", "SCHUR_FORM_INV\n")):
            inp, res = os.path.join(d, tag + ".txt"), os.path.join(d, tag + ".csv")
            open(inp, "w").write(base + extra)
            say("organic_main", tag)
            out = subprocess.run([MAIN, inp, res], capture_output=True, text=True, timeout=300)
            assert out.returncode == 0, out.stderr
            rows.append(np.array([[float(v) for v in line.split("\t") if v] for line in open(res).read().splitlines()]))
    plain, schur = rows
    assert plain.shape == schur.shape == (6, 12)
    np.testing.assert_array_equal(schur[:, 6], plain[:, 6])
    np.testing.assert_allclose(schur, plain, rtol=1e-8, atol=1e-12)
