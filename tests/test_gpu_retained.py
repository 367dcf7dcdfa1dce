"""The retained functions of the reference's host interface that the engine serves besides the
trajectory path (organic_dense_matrix.h:14-38, organic_hmc.h:9-21, organic_bosons.h:9-20,
organic_trans_loc.h:4-6), called through libephmc_host.so with the reference's signatures and
compared with the oracle (or dense numpy where the oracle has no such entry point).  Also the
failure paths: a singular M_hat (log det = -inf), the pivots of a factorisation that needs row
exchanges, a NaN proposal (rejected), transient-localisation pairs past the last bin.
FP64: 1e-12 relative unless stated."""
import ctypes as C

import numpy as np
import pytest

import _dense as D
from _gpu import gpu_context, rel, say
from _oracle import HMC, MT, Flags, OracleModel, dp, ip, oracle, ptr, up
from test_host_layer import host

pytestmark = pytest.mark.gpu

U, I, DB = C.c_uint, C.c_int, C.c_double


def lib():
    h = host()
    sig = {
        "apply_mu_K_to_v": [dp, dp, dp, dp, dp, DB, up, U, U],
        "apply_mu_K_to_M": [dp, dp, dp, dp, dp, DB, up, U, U],
        "apply_K_plus_1_v": [dp, dp, dp, dp, dp, DB, up, U, U],
        "apply_pair_expansion_mu_K_to_v": [dp, dp, dp, dp, dp, DB, up, U, U, I, I],
        "apply_exp_mu_K_to_v": [dp, dp, dp, dp, dp, DB, up, U, U, I, I],
        "construct_QM1P": [dp, dp, dp, dp, dp, dp, DB, up, U, U, U, I, C.c_void_p],
        "construct_all_QM1P": [dp, dp, dp, dp, ip, dp, dp, DB, up, U, U, U, C.c_void_p],
        "refine_QM1P": [dp, dp, dp, dp, dp, DB, up, U, U, I],
        "refine_QM1P_approx": [dp, dp, dp, dp, dp, DB, up, U, U, I],
        "construct_hamilton": [dp, dp, dp, dp, dp, DB, up, U, U],
        "construct_M_hat": [dp, dp, dp, dp, dp, DB, up, U, U, U, C.c_void_p],
        "logDetM_and_inv_M_hat": [dp, dp, ip, U, U, C.c_void_p],
        "force": [dp, dp, dp, dp, dp, ip, dp, dp, DB, DB, U, up, U, U, U, DB, C.c_void_p],
        "check_force": [dp, dp, dp, dp, dp, ip, dp, dp, DB, DB, U, up, U, U, U, C.c_void_p],
        "hot_start": [dp, DB, U, U, U],
        "boson_force": [dp, dp, DB, U, U, U, I],
        "hist_trans_loc": [dp, dp, dp, DB, dp, up, U, U, U, DB, dp, dp, C.c_void_p],
        "phonon_correlator": [dp, U, U, U, dp],
        "mat_mul": [dp, dp, dp, U],
        "mat_mul_tr": [dp, dp, dp, U],
        "mat_mat_mul": [dp, dp, dp, U],
        "transpose_rect": [dp, dp, U, U],
        "organic_plan_create": [dp, dp, dp, dp, DB, DB, U, U, U, U, U, DB, C.c_void_p],
        "organic_plan_destroy": [C.c_void_p],
        "evolve_bosons_approx_step": [dp, DB, U, U, U, C.c_void_p, DB, C.c_void_p],
    }
    for name, args in sig.items():
        getattr(h, name).argtypes = args
    h.logDetM_and_inv_M_hat.restype = DB
    h.check_force.restype = DB
    h.organic_plan_create.restype = C.c_void_p
    # apply_2nd_order_mu_K_to_v takes a C double complex by value: two doubles in registers
    h.apply_2nd_order_mu_K_to_v.argtypes = [dp, dp, dp, dp, dp, DB, up, U, U, DB, DB]
    return h


def model(l1=3, l2=4, nt=6, kscale=10, **kw):
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32, **kw)
    om.set_couplings(kappa=om.kappa * kscale)
    return om


def nntp(om):
    a = np.ascontiguousarray(om.nnt, np.uint32)
    return a, a.ctypes.data_as(up)


@pytest.mark.parametrize("l1,l2", [(1, 3), (3, 4), (5, 5)])
def test_stencil_vector_and_matrix(l1, l2):
    say("apply_mu_K_to_v/M", l1, l2)
    h, o = lib(), oracle()
    om = model(l1, l2, 1)
    n = om.ns
    keep, nt_ = nntp(om)
    x = np.random.default_rng(1).standard_normal(3 * n)
    v = np.random.default_rng(2).standard_normal(n)
    w1, w2 = np.zeros(2 * n), np.zeros(2 * n)
    h.apply_mu_K_to_v(ptr(x), ptr(v), ptr(w1), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6)
    o.orc_apply_mu_K_to_v(om.ref, ptr(x), ptr(v), ptr(w2), om.mu)
    np.testing.assert_allclose(w1[:n], w2[:n], rtol=1e-14, atol=1e-15)
    M = np.random.default_rng(3).standard_normal(n * n)
    M1, M2 = M.copy(), M.copy()
    h.apply_mu_K_to_M(ptr(x), ptr(M1), ptr(np.zeros(2 * n * n)), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6)
    o.orc_apply_mu_K_to_M(om.ref, ptr(x), ptr(M2), ptr(np.zeros(2 * n * n)))
    np.testing.assert_allclose(M1, M2, rtol=1e-14, atol=1e-14)
    # refine_QM1P_approx is the stencil on the slice `time` (organic_dense_matrix.c:444-449)
    xf = np.concatenate([np.zeros(3 * n), x])
    M3 = M.copy()
    h.refine_QM1P_approx(ptr(xf), ptr(M3), ptr(np.zeros(2 * n * n)), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6, 1)
    np.testing.assert_array_equal(M3, M1)


def test_pair_products_on_vectors():
    say("apply_K_plus_1_v / apply_2nd_order / pair expansion")
    h, o = lib(), oracle()
    om = model(3, 4, 1, kscale=20)
    n = om.ns
    keep, nt_ = nntp(om)
    x = np.random.default_rng(4).standard_normal(3 * n)
    v0 = np.random.default_rng(5).standard_normal(n)
    A = D.hopping_matrix(om.nnt, n, x, om.t, om.kappa, om.mu)
    w = np.zeros(2 * n)
    v = v0.copy()
    h.apply_K_plus_1_v(ptr(x), ptr(v), ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6)
    np.testing.assert_allclose(v, v0 + A @ v0, rtol=1e-13, atol=1e-14)
    for a in (0.3 + 0.7j, -0.45 + 0j):
        v = v0.copy()
        h.apply_2nd_order_mu_K_to_v(ptr(x), ptr(v), ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6, a.real, a.imag)
        ref = v0 + 2 * a.real * (A @ v0) + ((abs(a) ** 2) * (A @ (A @ v0)) if a.imag else 0)
        np.testing.assert_allclose(v, ref, rtol=1e-13, atol=1e-13)
    for fn, ofn in (("apply_pair_expansion_mu_K_to_v", None), ("apply_exp_mu_K_to_v", None)):
        for revert, stop in ((0, 2), (0, 1), (0, 0), (1, -1), (1, 0), (1, 1)):
            v1, v2 = v0.copy(), v0.copy()
            getattr(h, fn)(ptr(x), ptr(v1), ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6, revert, stop)
            o.orc_apply_exp_mu_K_to_v(om.ref, ptr(x), ptr(v2), ptr(np.zeros(2 * n)), revert, stop)
            np.testing.assert_allclose(v1, v2, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("l1,l2,nt", [(2, 3, 5), (3, 4, 8), (4, 4, 7)])
def test_construct_all_QM1P_and_refine(l1, l2, nt):
    say("construct_all_QM1P / refine_QM1P / construct_QM1P", l1, l2, nt)
    h, o = lib(), oracle()
    om = model(l1, l2, nt)
    n = om.ns
    keep, nt_ = nntp(om)
    x = np.random.default_rng(nt).standard_normal(om.dim)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    Q1 = np.zeros(nt * n * n)
    Q2 = np.zeros(nt * n * n)
    perm1 = np.zeros(2 * nt, np.int32)
    perm2 = np.zeros(2 * nt, np.int32)
    w = np.zeros(8 * n * n)
    h.construct_all_QM1P(ptr(x), ptr(Minv), ptr(Q1), ptr(w), ptr(perm1, ip), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6,
                         nt, None)
    o.orc_construct_all_QM1P(om.ref, ptr(x), ptr(Minv), ptr(Q2), ptr(np.zeros(5 * n * n)), ptr(perm2, ip))
    assert rel(Q1, Q2) < 1e-12
    np.testing.assert_array_equal(perm1, perm2)
    for tau in (0, nt // 2, nt - 1):
        q1 = Q1[tau * n * n:(tau + 1) * n * n].copy()
        q2 = Q2[tau * n * n:(tau + 1) * n * n].copy()
        h.refine_QM1P(ptr(x), ptr(q1), ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6, tau)
        o.orc_refine_QM1P(om.ref, ptr(x), ptr(q2), ptr(np.zeros(5 * n * n)), tau)
        assert rel(q1, q2) < 1e-12, tau
        c1, c2 = np.zeros(n * n), np.zeros(n * n)
        h.construct_QM1P(ptr(x), ptr(Minv), ptr(c1), ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nt_, n, 6, nt, tau, None)
        o.orc_construct_QM1P(om.ref, ptr(x), ptr(Minv), ptr(c2), ptr(np.zeros(5 * n * n)), tau)
        assert rel(c1, c2) < 1e-12, tau
        assert rel(c1, q1) < 1e-12, tau       # construct_QM1P = refine of the same slot
    # the projection of the refined slots is the force (organic_hmc.c:100-126)
    F = np.zeros(om.dim)
    perm = np.zeros(max(n, 2 * nt), np.int32)
    h.force(ptr(x), ptr(F), ptr(Minv), ptr(np.zeros(nt * n * n)), ptr(w), ptr(perm, ip), ptr(om.t), ptr(om.kappa),
            om.mu, om.omega, om.flavors, nt_, n, 6, nt, ld, None)
    assert rel(F, om.force(x, Minv, ld)) < 1e-11


@pytest.mark.parametrize("order", [2, 6, 8])
def test_refine_taylor_orders(order):
    say("refine_QM1P order", order)
    om = model(3, 3, 4, taylor=order)
    ctx = gpu_context(om, taylor_order=order)
    n = om.ns
    x = np.random.default_rng(order).standard_normal(om.dim)
    Q = np.random.default_rng(9).standard_normal(n * n)
    dx, dq = ctx.array(3 * n, x[3 * n:6 * n]), ctx.array(n * n, Q)
    import ephmc as E
    E.check(E.LIB.ephmc_refine_QM1P(ctx.h, dx.ptr, dq.ptr))
    q2 = Q.copy()
    oracle().orc_refine_QM1P(om.ref, ptr(x), ptr(q2), ptr(np.zeros(5 * n * n)), 1)
    assert rel(dq.download(), q2) < 1e-12


def test_construct_hamilton():
    say("construct_hamilton")
    h, o = lib(), oracle()
    for l1, l2 in ((1, 3), (3, 4), (6, 6)):
        om = OracleModel(l1, l2, 16, static_disorder=True)
        n = om.ns
        keep, nt_ = nntp(om)
        x = np.random.default_rng(5).standard_normal(3 * n)
        H1, H2, w = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
        h.construct_hamilton(ptr(x), ptr(H1), ptr(w), ptr(om.t), ptr(om.kappa), -om.mu, nt_, n, 6)
        o.orc_construct_hamilton(om.ref, ptr(x), ptr(H2), ptr(w))
        np.testing.assert_allclose(H1, H2, rtol=1e-15, atol=0)
        np.testing.assert_array_equal(H1.reshape(n, n), H1.reshape(n, n).T)


def test_check_force():
    """check_force (organic_hmc.c:20-54): the GPU force against differences of GPU log dets is
    small; a force with one wrong component is not."""
    say("check_force")
    h = lib()
    om = model(2, 2, 4, kscale=10)
    n, nt = om.ns, om.nt
    keep, nt_ = nntp(om)
    x = np.random.default_rng(2).standard_normal(om.dim)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    F = om.force(x, Minv, ld)
    w = np.zeros(8 * n * n)
    perm = np.zeros(n, np.int32)
    M = np.zeros(n * n)
    err = h.check_force(ptr(x), ptr(np.zeros(om.dim)), ptr(F), ptr(M), ptr(w), ptr(perm, ip), ptr(om.t), ptr(om.kappa),
                        om.mu, om.omega, om.flavors, nt_, n, 6, nt, None)
    assert err < 1e-8
    F[3] *= 1.01
    bad = h.check_force(ptr(x), ptr(np.zeros(om.dim)), ptr(F), ptr(M), ptr(w), ptr(perm, ip), ptr(om.t),
                        ptr(om.kappa), om.mu, om.omega, om.flavors, nt_, n, 6, nt, None)
    assert bad > 100 * err and bad > 1e-6


def test_boson_force_and_cfa_drift():
    say("boson_force / evolve_bosons_approx_step")
    h, o = lib(), oracle()
    flags = {"classical_fourier_acc": 1, "cfa_steps": 3, "cfa_mass": 0.6}
    om = model(3, 2, 8, flags=flags)
    n, nt = om.ns, om.nt
    x = np.random.default_rng(6).standard_normal(om.dim)
    F1 = np.random.default_rng(7).standard_normal(om.dim)
    F2 = F1.copy()
    h.boson_force(ptr(x), ptr(F1), om.omega, n, 6, nt, 0)
    o.orc_boson_force(om.ref, ptr(x), ptr(F2), 0)
    np.testing.assert_allclose(F1, F2, rtol=1e-14, atol=1e-14)
    h.boson_force(ptr(x), ptr(F1), om.omega, n, 6, nt, 1)
    o.orc_boson_force(om.ref, ptr(x), ptr(F2), 1)
    np.testing.assert_allclose(F1, F2, rtol=1e-14, atol=1e-14)
    # the drift on the plan-bound buffers x and x + dim
    buf = np.zeros(5 * om.dim + n * n)
    buf[:om.dim] = x
    buf[om.dim:2 * om.dim] = np.random.default_rng(8).standard_normal(om.dim)
    xc = np.zeros(4 * om.dim)
    mode = Flags(classical_fourier_acc=1, cfa_steps=3, cfa_mass=0.6)
    plan = C.c_void_p(h.organic_plan_create(ptr(buf), ptr(xc), ptr(om.t), ptr(om.kappa), om.mu, om.omega, om.flavors,
                                            om.l1, om.l2, nt, 4, 0.0, C.cast(C.pointer(mode), C.c_void_p)))
    x_ref, p = buf[:om.dim].copy(), buf[om.dim:2 * om.dim].copy()
    h.evolve_bosons_approx_step(ptr(xc), om.omega, n, 6, nt, C.cast(C.pointer(plan), C.c_void_p), 0.3,
                                C.cast(C.pointer(mode), C.c_void_p))
    o.orc_evolve_bosons_approx_step(om.ref, ptr(x_ref), ptr(p), 0.3)
    np.testing.assert_allclose(buf[:om.dim], x_ref, rtol=1e-12, atol=1e-13)
    np.testing.assert_array_equal(buf[om.dim:2 * om.dim], p)
    h.organic_plan_destroy(plan)


def test_hot_start_draws():
    """hot_start (organic_bosons.c:16-34): constant in time, one Box-Muller pair per two link
    components, from the host's MT19937-64 stream (restated with the oracle's generator)."""
    h = lib()
    ns, nt = 5, 4
    loc = 3 * ns
    x = np.zeros(nt * loc)
    h.init_genrand64(77)
    omega = 0.15
    h.hot_start(ptr(x), omega, ns, 6, nt)
    m = MT()
    oracle().orc_mt_seed(C.byref(m), 77)
    sd = np.sqrt(0.5 / np.tanh(0.5 * nt * omega))
    ref = np.zeros(nt * loc)
    for k in range(0, loc, 2):
        r = sd * np.sqrt(-2 * np.log(oracle().orc_mt_real2(C.byref(m))))
        phi = 6.283185307179586 * oracle().orc_mt_real2(C.byref(m))
        for tau in range(nt):
            ref[tau * loc + k] = r * np.cos(phi)
            ref[tau * loc + (k + 1) % loc] = r * np.sin(phi)
    np.testing.assert_array_equal(x, ref)


def test_hist_trans_loc_one_sample():
    say("hist_trans_loc")
    h, o = lib(), oracle()
    om = OracleModel(4, 4, 32, static_disorder=True)
    n = om.ns
    keep, nt_ = nntp(om)
    cm = np.array([7.2, 3.6, -3.6, 0.0, 7.15, 7.15])
    width = o.orc_tl_width(om.ref, 32)
    x = np.zeros(3 * n + 2 * n * n + n + 12)
    g1 = np.zeros(64)
    r1 = np.zeros(1)
    h.init_genrand64(5)
    h.hist_trans_loc(ptr(x), ptr(om.t), ptr(om.kappa), om.mu, ptr(cm), nt_, n, 6, 32, width, ptr(r1), ptr(g1), None)
    m = MT()
    o.orc_mt_seed(C.byref(m), 5)
    xs = np.zeros(3 * n)
    o.orc_random_vector(C.byref(m), ptr(xs), 3 * n)
    np.testing.assert_array_equal(x[:3 * n], xs)
    g2 = np.zeros(32)
    z = o.orc_hist_trans_loc(om.ref, ptr(xs), ptr(cm), width, ptr(g2), 32)
    np.testing.assert_allclose(g1[:32], g2, rtol=1e-9, atol=1e-12 * np.abs(g2).max())
    assert not g1[32:].any()
    assert r1[0] == pytest.approx(z, rel=1e-12)


def test_trans_loc_bins_past_the_end():
    """Pairs whose bin index (E_j - E_i)/width + 1/2 reaches nbins (organic_trans_loc.c:58-60)
    are dropped; with a bin width of a tenth of the usual one most pairs are, and the kept bins
    still equal the brute-force sum."""
    say("tl bins past the end")
    om = OracleModel(4, 3, 16, static_disorder=True)
    n = om.ns
    ctx = gpu_context(om)
    cm = np.array([7.2, 3.6, -3.6, 0.0, 7.15, 7.15])
    nbins = 16
    width = oracle().orc_tl_width(om.ref, nbins) / 10
    xs = np.random.default_rng(3).standard_normal((2, 3 * n))
    g, res = ctx.tl_hist(xs, cm, width, nbins)
    ref = np.zeros(nbins)
    dropped = 0
    for s in range(2):
        H = D.hopping_matrix(om.nnt, n, xs[s], om.t, om.kappa, -om.mu)
        E, V = np.linalg.eigh(H)
        nb = om.nnt[:6 * n].reshape(n, 6)
        J = []
        for k in range(3):
            Ik = np.zeros((n, n))
            for i in range(n):
                f = om.t[k] - om.kappa[k] * xs[s][3 * i + k]
                Ik[i, nb[i, 2 * k]] += f
                Ik[nb[i, 2 * k], i] -= f
            J.append(V.T @ Ik @ V)
        Cm = cm.reshape(2, 3)
        for i in range(n):
            for j in range(n):
                if E[j] < E[i]:
                    continue
                b = int((E[j] - E[i]) / width + .5)
                if b >= nbins:
                    dropped += 1
                    continue
                proj = np.array([J[k][j, i] for k in range(3)])
                ref[b] += (np.exp(-E[i]) + np.exp(-E[j])) * ((Cm @ proj) ** 2).sum() / (width * n)
    assert dropped > n * (n + 1) // 2        # most of the pairs of two samples fall past the end
    np.testing.assert_allclose(g, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_dense_helpers_and_phonon_correlator():
    h = lib()
    n = 6
    m = np.random.default_rng(1).standard_normal(n * n)
    xv = np.random.default_rng(2).standard_normal(n)
    y = np.zeros(n)
    h.mat_mul(ptr(m), ptr(xv), ptr(y), n)
    np.testing.assert_allclose(y, m.reshape(n, n) @ xv, rtol=1e-14)
    h.mat_mul_tr(ptr(m), ptr(xv), ptr(y), n)
    np.testing.assert_allclose(y, m.reshape(n, n).T @ xv, rtol=1e-14)
    m2 = np.random.default_rng(3).standard_normal(n * n)
    ref = (m.reshape(n, n) @ m2.reshape(n, n).T).T.ravel()
    h.mat_mat_mul(ptr(m), ptr(m2), ptr(y), n)
    np.testing.assert_allclose(m2, ref, rtol=1e-13)
    r = np.random.default_rng(4).standard_normal(3 * 5)
    rt = np.zeros(15)
    h.transpose_rect(ptr(r), ptr(rt), 3, 5)
    np.testing.assert_array_equal(rt.reshape(5, 3), r.reshape(3, 5).T)
    ns, nt = 4, 5
    x = np.random.default_rng(5).standard_normal(nt * ns * 3)
    g = np.zeros(nt * 9)
    h.phonon_correlator(ptr(x), ns, 6, nt, ptr(g))
    xr = x.reshape(nt, ns, 3)
    np.testing.assert_allclose(g.reshape(nt, 3, 3), np.einsum("ia,tib->tab", xr[0], xr) / ns, rtol=1e-13)


# ---- failure paths ---------------------------------------------------------------------------

def test_pivots_of_a_factorisation_with_exchanges():
    """logDetM_and_inv_M_hat's p (dgetrf's ipiv, organic_dense_matrix.c:580) after the pivoted
    path ran, against scipy's LU pivots; the identity when no exchange is needed."""
    say("pivots")
    from scipy.linalg import lu_factor
    h = lib()
    om = model(4, 4, 4)
    n = om.ns
    B = np.random.default_rng(12).standard_normal((n, n))
    flat = B.T.copy().ravel()                 # column-major
    p = np.zeros(n, np.int32)
    ld = h.logDetM_and_inv_M_hat(ptr(flat), ptr(np.zeros(4 * n * n)), ptr(p, ip), n, 2, None)
    lu, piv = lu_factor(B)
    np.testing.assert_array_equal(p, piv + 1)
    assert ld == pytest.approx(np.linalg.slogdet(B)[1], rel=1e-12)
    np.testing.assert_allclose(flat.reshape(n, n).T, np.linalg.inv(B), rtol=1e-9, atol=1e-11)
    # a diagonally dominant matrix needs no exchange
    A = np.eye(n) * 10 + np.random.default_rng(13).standard_normal((n, n))
    flat = A.T.copy().ravel()
    h.logDetM_and_inv_M_hat(ptr(flat), ptr(np.zeros(4 * n * n)), ptr(p, ip), n, 2, None)
    np.testing.assert_array_equal(p, np.arange(1, n + 1))


@pytest.mark.parametrize("n", [24, 100])
def test_singular_matrix_gives_minus_infinity(n):
    """A zero pivot: log|det| = -inf as in the reference (organic_dense_matrix.c:582-584), the
    engine reports EPHMC_ERR_NUMERIC, the host call returns -inf without exiting."""
    say("singular", n)
    om = OracleModel(n, 1, 1)
    ctx = gpu_context(om)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A[:, 5] = 0                               # a zero column: the pivoted path finds no pivot
    rc, ld, sg, inv = ctx.logdet_inv_status(A.T.copy().ravel())
    assert rc == -5 and ld == -np.inf
    assert np.isnan(inv).all()
    D_ = np.eye(n)
    D_[n - 1, n - 1] = 0                      # the last pivot zero: the fast path ends at -inf
    rc, ld, sg, inv = ctx.logdet_inv_status(D_.ravel())
    assert rc == -5 and ld == -np.inf
    # the matrix after a singular one factorises normally again
    rc, ld, sg, inv = ctx.logdet_inv_status((np.eye(n) * 2).ravel())
    assert rc == 0 and ld == pytest.approx(n * np.log(2), rel=1e-14)
    h = lib()
    flat = A.T.copy().ravel()
    p = np.zeros(n, np.int32)
    got = h.logDetM_and_inv_M_hat(ptr(flat), ptr(np.zeros(4 * n * n)), ptr(p, ip), n, 2, None)
    assert got == -np.inf


def test_nan_proposal_is_rejected():
    """A proposal with a NaN (here in the momenta) has H = NaN, exp(H_old - H_new) = NaN, and
    !(NaN > u) rejects it (organic_hmc.c:233-234): the chain state is kept, and the next
    trajectory equals the one a chain without the NaN step makes."""
    say("nan proposal")
    om = model(3, 3, 8, kscale=5)
    a = gpu_context(om, steps=4)
    b = gpu_context(om, steps=4)
    rng = np.random.default_rng(3)
    init = rng.standard_normal(om.dim)
    a.chain_init(init)
    b.chain_init(init)
    x0, ld0 = a.chain_x(), a.chain_logdet()
    bad = rng.standard_normal(om.dim)
    bad[17] = np.nan
    rec = a.trajectory(bad, 0.0)
    assert rec.accepted == 0 and not np.isfinite(rec.dH)
    np.testing.assert_array_equal(a.chain_x(), x0)
    assert a.chain_logdet() == ld0
    nrm = rng.standard_normal(om.dim)
    ra, rb = a.trajectory(nrm, 0.5), b.trajectory(nrm, 0.5)
    assert ra.accepted == rb.accepted and ra.dH == rb.dH
    np.testing.assert_array_equal(a.chain_x(), b.chain_x())
