"""Parity of the HIP engine with the oracle (same seeded inputs, sizes the oracle finishes in
seconds) and, at the L=32, Nt=128 size, with closed forms.  Tolerances: FP64 path 1e-10
relative (log-det, force, M_hat), FP32-mixed force 1e-5 relative."""
import ctypes as C

import numpy as np
import pytest

import _dense as D
from _gpu import ephmc, gpu_context, rel, say
from _oracle import MT, OracleModel, oracle, ptr

pytestmark = pytest.mark.gpu


def field(om, seed, scale=1.0):
    return np.random.default_rng(seed).standard_normal(om.dim) * scale


@pytest.mark.parametrize("l1,l2,nt", [(1, 1, 4), (2, 2, 4), (2, 3, 8), (5, 4, 8), (6, 6, 32), (8, 12, 16), (16, 16, 8)])
def test_M_hat(l1, l2, nt):
    say("M_hat", l1, l2, nt)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    om.set_couplings(kappa=om.kappa * 10)
    ctx = gpu_context(om)
    np.testing.assert_array_equal(ctx.lattice(), om.nnt)
    x = field(om, nt)
    Mg = ctx.M_hat(x)
    Mo = om.M_hat(x)
    assert rel(Mg, Mo) < 1e-13


@pytest.mark.parametrize("n", [2, 33, 37, 63, 97, 1023])
def test_gj_trailing_blocks(n):
    """Block steps whose diagonal block is cut short (1, 5, 31, 1 and 31 columns; odd sizes pair
    the last pivot with an identity row): log det and inverse against numpy on the fast path."""
    say("gj trailing block", n)
    om = OracleModel(n, 1, 1)
    ctx = gpu_context(om)
    A = np.random.default_rng(n).standard_normal((n, n)) * 0.3 / np.sqrt(n) + np.eye(n) * 2
    A[0, 0] = -2                               # one negative pivot: sign -1
    ld, sg, inv = ctx.logdet_inv(A.T.copy().ravel())
    s0, l0 = np.linalg.slogdet(A)
    assert ld == pytest.approx(l0, rel=1e-12) and sg == int(s0) == -1
    assert rel(inv.reshape(n, n).T, np.linalg.inv(A)) < 1e-12
    assert ctx.gj_fallbacks() == 0


@pytest.mark.parametrize("strip", ["2", "4"])
@pytest.mark.parametrize("l1,l2,nt", [(2, 3, 4), (6, 6, 8), (16, 16, 8), (32, 32, 4)])
def test_strip_chain_bit_identical(monkeypatch, l1, l2, nt, strip):
    """The strip form of the slice stencil sums in the same order as the site-per-thread form:
    M_hat must agree to the last bit (EPHMC_CHAIN_STRIP=0 selects the site-per-thread form)."""
    say("strip vs site chain", l1, l2, nt, strip)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    ctx = gpu_context(om)
    x = field(om, 5)
    monkeypatch.setenv("EPHMC_CHAIN_STRIP", strip)
    Ms = ctx.M_hat(x)
    monkeypatch.setenv("EPHMC_CHAIN_STRIP", "0")
    Mg = ctx.M_hat(x)
    np.testing.assert_array_equal(Ms, Mg)


@pytest.mark.parametrize("revert,stop", [(0, 2), (0, 1), (0, 0), (1, 0), (1, -1)])
def test_apply_exp_partial(revert, stop):
    say("apply_exp_M", revert, stop)
    om = OracleModel(5, 4, 2)
    om.set_couplings(kappa=om.kappa * 10)
    ctx = gpu_context(om)
    x = field(om, 3)
    n = om.ns
    M = np.random.default_rng(1).standard_normal(n * n)
    g = ctx.apply_exp_M(x[:3 * n], M, revert, stop)
    o = M.copy()
    w = np.zeros(2 * n * n)
    oracle().orc_apply_exp_mu_K_to_M(om.ref, ptr(x), ptr(o), ptr(w), revert, stop)
    assert rel(g, o) < 1e-14


@pytest.mark.parametrize("n", [1, 5, 36, 100, 256, 300, 1024])
def test_logdet_inverse_random(n):
    say("gj random", n)
    om = OracleModel(n, 1, 1)
    ctx = gpu_context(om)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + np.eye(n) * (0 if n < 50 else 3)
    ld, sg, inv = ctx.logdet_inv(A.T.copy().ravel())
    s0, l0 = np.linalg.slogdet(A)
    assert ld == pytest.approx(l0, rel=1e-11, abs=1e-11) and sg == int(s0)
    assert rel(inv.reshape(n, n).T, np.linalg.inv(A)) < 1e-10
    # row exchanges are needed for these matrices exactly when LU without exchanges has a
    # multiplier above 1; then the pivoted fallback must have run
    import scipy.linalg as sl
    lu, piv = sl.lu_factor(A)
    needs = bool(np.any(piv != np.arange(n)))
    say("gj random", n, "row exchanges", needs, "fallbacks", ctx.gj_fallbacks())
    assert (ctx.gj_fallbacks() > 0) == needs


@pytest.mark.parametrize("n,r,j", [(256, 100, 97), (256, 200, 97), (1024, 610, 608), (1024, 900, 608)])
def test_gj_single_exchange_in_a_middle_block(n, r, j):
    """A diagonally dominant matrix with one entry |A[r][j]| > |A[j][j]| below the diagonal in the
    pivot column j of a block that is neither the first nor the last: LU needs one row exchange
    there (its multiplier is about 1.5), inside the diagonal block (r in the block of j) or below
    it.  The check must raise the flag, the pivoted fallback must run, and log det, sign and
    inverse equal numpy's."""
    say("gj exchange in a middle block", n, r, j)
    om = OracleModel(n, 1, 1)
    ctx = gpu_context(om)
    A = np.random.default_rng(n + r).standard_normal((n, n)) * 0.01 / np.sqrt(n) + np.eye(n) * 2
    A[r, j] = 3.0
    ld, sg, inv = ctx.logdet_inv(A.T.copy().ravel())
    s0, l0 = np.linalg.slogdet(A)
    assert ctx.gj_fallbacks() == 1
    assert ld == pytest.approx(l0, rel=1e-12) and sg == int(s0)
    assert rel(inv.reshape(n, n).T, np.linalg.inv(A)) < 1e-12
    A[r, j] = 1.5                              # multiplier 0.75: no exchange, fast path
    ld, sg, inv = ctx.logdet_inv(A.T.copy().ravel())
    assert ctx.gj_fallbacks() == 1
    assert ld == pytest.approx(np.linalg.slogdet(A)[1], rel=1e-12)


def test_logdet_inverse_diagonally_dominant_no_fallback():
    """A matrix whose multipliers are all below 1 (no exchanges) stays on the fast path."""
    n = 300
    om = OracleModel(n, 1, 1)
    ctx = gpu_context(om)
    A = np.random.default_rng(5).standard_normal((n, n)) * 0.01 + np.eye(n) * 2
    ld, sg, inv = ctx.logdet_inv(A.T.copy().ravel())
    s0, l0 = np.linalg.slogdet(A)
    assert ld == pytest.approx(l0, rel=1e-12) and sg == int(s0)
    assert rel(inv.reshape(n, n).T, np.linalg.inv(A)) < 1e-12
    assert ctx.gj_fallbacks() == 0


@pytest.mark.parametrize("l1,l2,nt", [(6, 6, 32), (16, 16, 64)])
def test_logdet_inverse_M_hat(l1, l2, nt):
    say("gj M_hat", l1, l2, nt)
    om = OracleModel(l1, l2, nt)
    ctx = gpu_context(om)
    x = field(om, 5)
    M = om.M_hat(x)
    ldo, so, invo = om.logdet_inv(M)
    ldg, sg, invg = ctx.logdet_inv(M)
    assert ldg == pytest.approx(ldo, rel=1e-12) and sg == so
    assert rel(invg, invo) < 1e-12
    assert ctx.gj_fallbacks() == 0


@pytest.mark.parametrize("l1,l2,nt,k", [(3, 3, 4, 16), (6, 6, 32, 16), (6, 6, 32, 5), (6, 6, 32, 1), (5, 7, 12, 4),
                                        (2, 2, 8, 3), (16, 16, 16, 8), (5, 7, 40, 1), (3, 3, 4, 1)])
def test_force_fp64(l1, l2, nt, k):
    say("force fp64", l1, l2, nt, k)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    om.set_couplings(kappa=om.kappa * 10)
    ctx = gpu_context(om, checkpoint=k)
    x = field(om, l1 + nt)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    Fo = om.force(x, Minv, ld)
    Fg = ctx.force(x, Minv, ld)
    assert rel(Fg, Fo) < 1e-11


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("l1,l2,nt", [(6, 6, 40), (16, 16, 72), (32, 32, 8)])
def test_force_every_slice_matches_recompute(precision, l1, l2, nt):
    """k = 1 stores every partial product and the sweep reads them back (32 slices per launch);
    k = 8 recomputes them from checkpoints with the forward chain's arithmetic: same bits."""
    say("force k=1 vs k=8", precision, l1, l2, nt)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    x = field(om, 7)
    ref = gpu_context(om)
    _, _, Minv = ref.logdet_inv(ref.M_hat(x))
    F = {k: gpu_context(om, precision=precision, checkpoint=k).force(x, Minv, 0.0) for k in (1, 8)}
    np.testing.assert_array_equal(F[1], F[8])


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("l1,l2,nt", [(6, 6, 40), (32, 32, 8)])
def test_force_w0_order_bit_identical(monkeypatch, precision, l1, l2, nt):
    """The force sweep's two orders of the first W phase (its own barrier, or W_0 written at the
    end of the previous slice) do the same arithmetic in the same order: same bits."""
    say("force w0 order", precision, l1, l2, nt)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    x = field(om, 8)
    ref = gpu_context(om)
    _, _, Minv = ref.logdet_inv(ref.M_hat(x))
    F = {}
    for early in ("0", "1"):
        monkeypatch.setenv("EPHMC_ADJOINT_W0EARLY", early)
        F[early] = gpu_context(om, precision=precision, checkpoint=1).force(x, Minv, 0.0)
    np.testing.assert_array_equal(F["0"], F["1"])


@pytest.mark.parametrize("l1,l2", [(33, 32), (43, 43)])
def test_force_fp64_default_width_near_lds_limit(l1, l2):
    """FP64 force with the default column width at N = 1056 and 1849, the sizes where the sweep's
    reduction scratch decides whether C = 4 (or 2) still fits the 160 KiB of LDS: the default must
    pick a width that launches, and the force equals the oracle's."""
    say("force FP64 default width", l1, l2)
    om = OracleModel(l1, l2, 4, beta=4 * 40 / 32)
    ctx = gpu_context(om, checkpoint=1)
    x = field(om, 21)
    ld, _, Minv = ctx.logdet_inv(ctx.M_hat(x))
    Fg = ctx.force(x, Minv, ld)
    Fo = om.force(x, Minv, ld)
    say("rel err", rel(Fg, Fo))
    assert rel(Fg, Fo) < 1e-10


def test_maximum_lattice_size():
    """The largest lattice a context accepts, 64 x 32 = 2048 sites (Nt = 8; the chain and the sweep
    give each thread two sites): M_hat against the oracle (1e-13), log det and inverse against numpy
    (1e-10), and the force against the oracle's with the same inverse, FP64 (1e-10) with every slice
    stored and recomputed from checkpoints every 4 slices, mixed precision (1e-5).  One lattice row
    more (65 x 32 = 2080 sites) is refused as unsupported."""
    say("maximum lattice 64 x 32")
    om = OracleModel(64, 32, 8, beta=8 * 40 / 32)
    x = field(om, 31)
    n = om.ns
    Mo = om.M_hat(x)
    ctx = gpu_context(om, checkpoint=1)
    Mg = ctx.M_hat(x)
    say("M_hat rel err", rel(Mg, Mo))
    assert rel(Mg, Mo) < 1e-13
    ld, sg, Minv = ctx.logdet_inv(Mg)
    A, Ai = Mg.reshape(n, n).T, Minv.reshape(n, n).T
    sg_np, ld_np = np.linalg.slogdet(A)
    say("log det", ld, ld_np, "residual", np.abs(A @ Ai - np.eye(n)).max())
    assert sg == sg_np and abs(ld - ld_np) <= 1e-10 * max(1.0, abs(ld_np))
    assert np.abs(A @ Ai - np.eye(n)).max() < 1e-10
    Fo = om.force(x, Minv, ld)
    for prec, k, tol in ((0, 1, 1e-10), (0, 4, 1e-10), (1, 1, 1e-5)):
        c = ctx if (prec, k) == (0, 1) else gpu_context(om, precision=prec, checkpoint=k)
        Fg = c.force(x, Minv, ld)
        say("force precision %d checkpoint %d rel err %.2e" % (prec, k, rel(Fg, Fo)))
        assert rel(Fg, Fo) < tol
        if c is not ctx:
            c.close()
    ctx.close()
    with pytest.raises(ephmc().EphmcError, match="error -3"):
        gpu_context(OracleModel(65, 32, 8, beta=8 * 40 / 32))


def test_force_mixed():
    say("force mixed 6x6x32")
    om = OracleModel(6, 6, 32)
    ctx = gpu_context(om, precision=1)
    x = field(om, 2)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    Fo = om.force(x, Minv, ld)
    Fg = ctx.force(x, Minv, ld)
    say("mixed force rel err", rel(Fg, Fo))
    assert rel(Fg, Fo) < 1e-5


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-5)])
@pytest.mark.parametrize("taylor", [2, 6, 8])
def test_taylor_orders(taylor, precision, tol):
    """Slice exponentials of Taylor orders 2, 6 and 8 (polynomial degree 2, 6, 8: one force-sweep
    instance per degree): M_hat and the force against the oracle built with the same order."""
    say("taylor order", taylor, "precision", precision)
    om = OracleModel(7, 5, 16, beta=16 * 40 / 32, taylor=taylor)
    ctx = gpu_context(om, precision=precision, taylor_order=taylor)
    x = field(om, 11)
    Mo = om.M_hat(x)
    assert rel(ctx.M_hat(x), Mo) < 1e-13
    ld, _, Minv = om.logdet_inv(Mo)
    Fo = om.force(x, Minv, ld)
    Fg = ctx.force(x, Minv, ld)
    say("rel err", rel(Fg, Fo))
    assert rel(Fg, Fo) < tol


@pytest.mark.parametrize("checkpoint", [1, 4])
@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-5)])
@pytest.mark.parametrize("cols", ["1", "2", "4"])
@pytest.mark.parametrize("l1,l2,nt", [(7, 5, 8), (13, 11, 4), (23, 23, 4)])
def test_force_ragged(monkeypatch, l1, l2, nt, cols, precision, tol, checkpoint):
    """Force where the last workgroup's columns run past N (N = 35, 143, 529 are odd) and most
    lanes of a workgroup lie past the last site (N = 35 on 256 threads, 529 on 1024). In FP32
    those lanes load and gather but store nothing, and the columns past N read column N-1 and
    must add only zeros to the bond sums. Every slice stored (k = 1) and recomputed from every
    4th slice. beta = 40 Nt / 32 (slice width 1.25, as in the other force tests): with beta = 40 at
    Nt = 4 the slices are 10 wide and the FP32 chain keeps only about 3 digits."""
    say("force ragged", l1, l2, nt, "cols", cols, "precision", precision, "checkpoint", checkpoint)
    monkeypatch.setenv("EPHMC_ADJOINT_COLS", cols)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    ctx = gpu_context(om, precision=precision, checkpoint=checkpoint)
    x = field(om, 7)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    Fo = om.force(x, Minv, ld)
    Fg = ctx.force(x, Minv, ld)
    say("rel err", rel(Fg, Fo))
    assert rel(Fg, Fo) < tol


@pytest.mark.parametrize("nt", [1, 2, 8, 12, 32, 48])
def test_boson_kernels(nt):
    say("bosons", nt)
    om = OracleModel(3, 2, nt)
    ctx = gpu_context(om)
    lib = oracle()
    rng = np.random.default_rng(nt)
    x = rng.standard_normal(om.dim)
    p = rng.standard_normal(om.dim)
    xg, pg = ctx.evolve_bosons(x, p, 0.41)
    xo, po = x.copy(), p.copy()
    lib.orc_evolve_bosons(om.ref, ptr(xo), ptr(po), 0.41)
    assert rel(xg, xo) < 1e-13 and rel(pg, po) < 1e-13
    assert ctx.energy_momenta(p) == pytest.approx(lib.orc_energy_momenta(om.ref, ptr(p)), rel=1e-13)
    assert ctx.hamilton(x, 1.25, 3.5) == pytest.approx(lib.orc_hamilton(om.ref, ptr(x), 1.25, 3.5), rel=1e-13)
    for fn_g, fn_o in ((ctx.sample_fourier_momenta, lib.orc_sample_fourier_momenta),
                       (ctx.sample_free_bosons, lib.orc_sample_free_bosons)):
        st, st2 = MT(), MT()
        lib.orc_mt_seed(C.byref(st), 77)
        lib.orc_mt_seed(C.byref(st2), 77)
        normals = np.zeros(om.dim)
        lib.orc_random_vector(C.byref(st), ptr(normals), om.dim)
        out_o = np.zeros(om.dim)
        fn_o(om.ref, C.byref(st2), ptr(out_o))
        assert rel(fn_g(normals), out_o) < 1e-13


def test_hmc_accept_sequence_and_logdet():
    """Seeded FP64 chain at L=6, Nt=32, Nmd=10: identical accept/reject decisions and log-det
    within 1e-10 relative over the first 100 trajectories (oracle with the same RNG stream)."""
    om = OracleModel(6, 6, 32)
    lib = oracle()
    from _oracle import HMC
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), 1, 10, 0.0, None) == 0
    ctx = gpu_context(om, steps=10)
    # same RNG stream: the chain's free-boson draw, then per trajectory dim normals + 1 uniform
    rng = MT()
    lib.orc_mt_seed(C.byref(rng), 1)
    normals = np.zeros(om.dim)
    lib.orc_random_vector(C.byref(rng), ptr(normals), om.dim)
    ctx.chain_init(normals)
    assert ctx.chain_logdet() == pytest.approx(h.logdet, rel=1e-12)
    dH = C.c_double()
    worst = 0.0
    for it in range(100):
        lib.orc_random_vector(C.byref(rng), ptr(normals), om.dim)
        u = lib.orc_mt_real2(C.byref(rng))
        acc_o = lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None)
        rec = ctx.trajectory(normals, u)
        worst = max(worst, abs(rec.logdet - h.logdet) / abs(h.logdet))
        assert rec.accepted == acc_o, "trajectory %d: gpu %d oracle %d (dH gpu %g oracle %g)" % (
            it, rec.accepted, acc_o, rec.dH, dH.value)
        assert abs(rec.dH - dH.value) < 1e-8
    say("100 trajectories, worst logdet rel diff", worst)
    assert worst < 1e-10
    lib.orc_hmc_free(C.byref(h))


# ---- full-size (L = 32, Nt = 128) properties -----------------------------------------------

def test_full_size_logdet_closed_form():
    say("closed-form logdet 32x32x128")
    om = OracleModel(32, 32, 128)
    ctx = gpu_context(om)
    u = np.random.default_rng(3).standard_normal((128, 3))
    x = D.uniform_field(32, 32, 128, u)
    ld, sg, _ = ctx.logdet_inv(ctx.M_hat(x))
    ref = D.closed_form_logdet(32, 32, 128, om.t, om.kappa, om.mu, u)
    say("logdet", ld, "closed form", ref)
    assert ld == pytest.approx(ref, rel=1e-10) and sg == 1


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-5)])
def test_full_size_force_closed_form(precision, tol):
    say("closed-form force 32x32x128 precision", precision)
    om = OracleModel(32, 32, 128)
    ctx = gpu_context(om, precision=precision, checkpoint=16)
    u = np.random.default_rng(4).standard_normal((128, 3)) * 2
    x = D.uniform_field(32, 32, 128, u)
    ld, _, minv = ctx.logdet_inv(ctx.M_hat(x))
    F = ctx.force(x, minv, ld)
    ref = D.closed_form_force(32, 32, 128, om.t, om.kappa, om.mu, u, om.flavors)
    say("force rel err", rel(F, ref))
    assert rel(F, ref) < tol


def test_full_size_force_finite_differences():
    """Adjoint force vs finite differences of the FP64 log det on 64 random components at 32x32,
    Nt = 128.  The log det carries about 1.5e-14 of rounding noise, so a central difference at eps
    has a noise of about 1.5e-14 / (2 eps) (relative 8e-6 at eps = 1e-5 for these forces of order
    1e-5), while its truncation O(eps^2) is far smaller here (the central difference at 4e-2 is
    within 1e-8).  Checked: central differences at eps = 1e-5, within 5e-5 (noise bound), and the
    Richardson extrapolation of the central differences at 2e-2 and 4e-2 (truncation O(eps^4)),
    within 1e-7 (measured 1.3e-8)."""
    om = OracleModel(32, 32, 128)
    ctx = gpu_context(om, checkpoint=16)
    x = field(om, 9)
    ld, _, minv = ctx.logdet_inv(ctx.M_hat(x))
    F = ctx.force(x, minv, ld)
    idx = np.random.default_rng(64).choice(om.dim, 64, replace=False)

    def central(i, eps):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        return om.flavors * (ctx.logdet_inv(ctx.M_hat(xp))[0] - ctx.logdet_inv(ctx.M_hat(xm))[0]) / (2 * eps)

    fd5 = np.array([central(i, 1e-5) for i in idx])
    err5 = rel(F[idx], fd5)
    fd = np.array([(4 * central(i, 2e-2) - central(i, 4e-2)) / 3 for i in idx])
    err = rel(F[idx], fd)
    say("finite-difference rel err: central eps 1e-5 %.2e, Richardson (2e-2, 4e-2) %.2e" % (err5, err))
    assert err5 < 5e-5
    assert err < 1e-7
