"""Pins for the oracle functions that the dense/closed-form tests of test_oracle_math.py leave open:
the Taylor tables of orders 2/6/8 and their application order, the unit scaling, the boson force,
the classical-Fourier-acceleration drift and the NO_FOURIER_ACC / CLASSICAL_FOURIER_ACC position
updates, the leapfrog scheme (second order, exactly reversible), and the current-current
correlator (dense numpy evaluation from explicit transfer matrices, and the closed form of the
free translation-invariant system).  Each check is chosen so that a plausible mistake in the
oracle (a dropped term, a sign, a wrong slice index or pair order, a misplaced half step) fails."""
import ctypes as C
import math

import numpy as np
import pytest

import _dense as D
from _oracle import HMC, MT, OracleModel, oracle, ptr

# Taylor pair coefficients of simulation/taylor-coeff_pairs.h (TAYLOR_ORDER_2/4/6/8)
TABLES = {
    2: [1 + 1j],
    4: [1.8294933739891903 + 0.9404019519941707j, 0.1705066260108097 + 1.5785318126884535j],
    6: [2.2561549994606459 + 0.8008468192975424j, 1.0805822170351833 + 1.8245938269821296j,
        -0.3367372164958292 + 1.5494473211106643j],
    8: [2.5115385528762511 + 0.6852625793459637j, 1.6810015288012661 + 1.7481130845728929j,
        0.4249750401763088 + 2.0321227226190966j, -0.6175151218538260 + 1.4284655522676216j],
}
# application order of set_taylor_pair_order (organic_dense_matrix.c:133-159), worked out by hand
# for each table size: tuples alternate between the high and the low end, every third from the
# low end, an odd middle pair last
PAIR_ORDER = {2: [0], 4: [0, 1], 6: [0, 2, 1], 8: [1, 2, 0, 3]}


def exp_series(order):
    return np.array([1.0 / math.factorial(k) for k in range(order + 1)])


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_taylor_table_products_are_exp_series(order):
    """The full pair product of every table is sum_{k<=order} z^k/k! (the defining property of
    the pair coefficients); evaluated through the oracle's own stencil on a one-site lattice,
    where A_tau = mu exactly."""
    pairs = [0.5 * c / len(TABLES[order]) for c in TABLES[order]]
    np.testing.assert_allclose(D.pair_poly(pairs), exp_series(order), rtol=0, atol=2e-16)
    om = OracleModel(1, 1, 1, taylor=order)
    om.set_couplings(t=[0, 0, 0], kappa=[0, 0, 0])
    for z in np.linspace(-3, 1, 17):
        om.set_couplings(mu=z)
        v = np.array([1.0])
        w = np.zeros(2)
        oracle().orc_apply_exp_mu_K_to_v(om.ref, ptr(np.zeros(3)), ptr(v), ptr(w), 0, len(TABLES[order]))
        assert v[0] == pytest.approx(np.polyval(exp_series(order)[::-1], z), rel=1e-14, abs=1e-15)


@pytest.mark.parametrize("order", [6, 8])
@pytest.mark.parametrize("revert,stop", [(0, 1), (0, 2), (1, 0), (1, 1)])
def test_partial_pair_products_follow_reference_order(order, revert, stop):
    """Partial pair products (the pieces refine_QM1P combines) apply the pairs at the positions
    set_taylor_pair_order gives; a natural-order table would give different partial products."""
    npairs = len(TABLES[order])
    om = OracleModel(3, 3, 1, taylor=order)
    om.set_couplings(kappa=om.kappa * 30, t=om.t * 5)
    x = np.random.default_rng(order + 10 * stop).standard_normal(om.dim)
    n = om.ns
    M = np.random.default_rng(3).standard_normal((n, n))
    flat = M.T.copy().ravel()
    w = np.zeros(2 * n * n)
    oracle().orc_apply_exp_mu_K_to_M(om.ref, ptr(x), ptr(flat), ptr(w), revert, stop)
    A = D.hopping_matrix(om.nnt, n, x[:3 * n], om.t, om.kappa, om.mu)
    pos = [i for i in range(npairs) if (i > stop if revert else i < stop)]
    if revert:
        pos = pos[::-1]
    pairs = [0.5 * TABLES[order][PAIR_ORDER[order][i]] / npairs for i in pos]
    np.testing.assert_allclose(flat.reshape(n, n).T, D.poly_matrix(A, pairs) @ M, rtol=1e-12, atol=1e-13)
    if {PAIR_ORDER[order][i] for i in pos} != set(pos):
        # the factors commute, so only a different set of pairs is distinguishable: here the
        # natural table order selects other pairs, and its product must not also match
        wrong = [0.5 * TABLES[order][i] / npairs for i in pos]
        assert not np.allclose(D.poly_matrix(A, wrong) @ M, flat.reshape(n, n).T, rtol=1e-6)


def test_scale_params_closed_form():
    """organic_main.c:99-116 at the L=6 couplings: beta = 40 /eV, Nt = 32, J = (0.10, -0.03, 0.02),
    lambda = 0.3 relative, mu0 = 0, omega0 = 0.006 eV; x_avg = sqrt(1/2 / tanh(omega0 / 2 kT))
    with kT = 0.025 eV.  Numbers written out from the formulas."""
    om = OracleModel(6, 6, 32)
    dtau = 40.0 / 32
    x_avg = math.sqrt(0.5 / math.tanh(0.5 * 0.006 / 0.025))
    assert om.x_avg == pytest.approx(x_avg, rel=1e-15)
    assert x_avg == pytest.approx(2.046130, rel=1e-6)         # tanh(0.12) = 0.1194273
    t = np.array([0.125, -0.0375, 0.025])
    np.testing.assert_allclose(om.t, t, rtol=1e-15)
    np.testing.assert_allclose(om.kappa, 0.3 * t / x_avg, rtol=1e-15)
    assert om.mu == pytest.approx(-2 * (0.125 + 0.0375 + 0.025), rel=1e-15)   # -0.375
    assert om.omega == pytest.approx(0.006 * dtau, rel=1e-15)
    # transient localisation (STATIC_DISORDER): x_avg = 1, beta -> -beta, no division by Nt
    od = OracleModel(4, 4, 512, static_disorder=True, mu0=0.01)
    np.testing.assert_allclose(od.t, -40 * np.array([0.10, -0.03, 0.02]), rtol=1e-15)
    np.testing.assert_allclose(od.kappa, 0.3 * od.t, rtol=1e-15)
    assert od.mu == pytest.approx(-0.4 - 2 * 40 * 0.15, rel=1e-14)
    assert od.x_avg == 1.0


def test_boson_force_is_minus_gradient_of_action():
    """boson_force (organic_bosons.c:179-194) = -dS_B/dx, S_B(x) = hamilton(x, 0, 0)
    (organic_hmc.c:67-84); S_B is quadratic, so central differences are exact up to rounding."""
    om = OracleModel(2, 3, 6)
    lib = oracle()
    x = np.random.default_rng(4).standard_normal(om.dim)
    F = np.zeros(om.dim)
    lib.orc_boson_force(om.ref, ptr(x), ptr(F), 1)
    eps = 1e-3
    fd = np.zeros(om.dim)
    for i in range(om.dim):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        fd[i] = -(lib.orc_hamilton(om.ref, ptr(xp), 0.0, 0.0) - lib.orc_hamilton(om.ref, ptr(xm), 0.0, 0.0)) / (2 * eps)
    np.testing.assert_allclose(F, fd, rtol=1e-7, atol=1e-8 * np.abs(fd).max())
    # overwrite = 0 accumulates onto the given p_dot
    G = np.ones(om.dim)
    lib.orc_boson_force(om.ref, ptr(x), ptr(G), 0)
    np.testing.assert_allclose(G, F + 1, rtol=1e-14)


def _kinetic_grad(om, p, eps=1e-4):
    lib = oracle()
    g = np.zeros(om.dim)
    for i in range(om.dim):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        g[i] = (lib.orc_energy_momenta(om.ref, ptr(pp)) - lib.orc_energy_momenta(om.ref, ptr(pm))) / (2 * eps)
    return g


@pytest.mark.parametrize("nt", [6, 7])
def test_cfa_drift_is_kinetic_gradient(nt):
    """evolve_bosons_approx_step (organic_bosons.c:152-177) is the drift x += h dE_kin/dp of the
    Fourier-space mass with omega -> sqrt(m^2 + omega^2); energy_momenta uses the same mass
    (organic_hmc.c:56-65).  Both sides from the oracle, so a wrong mass, a missing 1/Nt or a wrong
    weight of the doubled modes fails; a numpy rfft restatement checks the same map."""
    flags = {"classical_fourier_acc": 1, "cfa_steps": 2, "cfa_mass": 0.7}
    om = OracleModel(2, 2, nt, flags=flags)
    lib = oracle()
    rng = np.random.default_rng(nt)
    x, p = rng.standard_normal(om.dim), rng.standard_normal(om.dim)
    x1 = x.copy()
    h = 0.3
    lib.orc_evolve_bosons_approx_step(om.ref, ptr(x1), ptr(p), h)
    np.testing.assert_allclose(x1 - x, h * _kinetic_grad(om, p), rtol=1e-7, atol=1e-9)
    loc = 3 * om.ns
    n = np.arange(nt // 2 + 1)
    om_k = math.sqrt(0.7 ** 2 + om.omega ** 2)
    f2 = om_k + (2 * np.sin(np.pi * n / nt)) ** 2 / om_k
    X = np.fft.rfft(x.reshape(nt, loc), axis=0) + (h / f2)[:, None] * np.fft.rfft(p.reshape(nt, loc), axis=0)
    np.testing.assert_allclose(x1, np.fft.irfft(X, nt, axis=0).ravel(), rtol=1e-12, atol=1e-12)


def test_no_fa_kinetic_energy_and_drift():
    """NO_FOURIER_ACC: E_kin = p.p/2 and update_x is x += |half| h p (organic_hmc.c:56-65,139-141)."""
    om = OracleModel(2, 2, 4, flags={"no_fourier_acc": 1})
    lib = oracle()
    rng = np.random.default_rng(1)
    x, p = rng.standard_normal(om.dim), rng.standard_normal(om.dim)
    assert lib.orc_energy_momenta(om.ref, ptr(p)) == pytest.approx(0.5 * p @ p, rel=1e-14)
    for half in (0.5, 1.0, -0.5):
        x1, p1 = x.copy(), p.copy()
        lib.orc_update_x(om.ref, ptr(x1), ptr(p1), ptr(np.zeros(om.dim)), 0.2, half)
        np.testing.assert_allclose(x1, x + 0.2 * abs(half) * p, rtol=1e-15, atol=1e-16)
        np.testing.assert_array_equal(p1, p)


def _hmc(om, seed, steps, traj_length=0.0):
    h = HMC()
    assert oracle().orc_hmc_init(C.byref(h), C.byref(om.m), seed, steps, traj_length, None) == 0
    return h


def _H(om, x, p):
    lib = oracle()
    return lib.orc_hamilton(om.ref, ptr(x), om.logdet(x), lib.orc_energy_momenta(om.ref, ptr(p)))


def _run_leapfrog(om, h, x0, p0, steps):
    """orc_leap_frog on (x0, p0) with the given number of steps; returns the end point."""
    x = np.ctypeslib.as_array(h.x, shape=(om.dim,))
    p = np.ctypeslib.as_array(h.p, shape=(om.dim,))
    x[:] = x0
    p[:] = p0
    h.steps = steps
    oracle().orc_leap_frog(C.byref(h))
    return x.copy(), p.copy()


# CLASSICAL_FOURIER_ACC with an odd number of sub-steps: update_x runs cfa_steps/2 (integer
# division) inner sub-steps per half, so an even count drifts the bosons for h (1 + 1/cfa_steps)
# per step -- the reference's behaviour, which the oracle keeps; only odd counts integrate H to
# second order, and the even count is covered by the reversibility check
LEAPFROG_MODES = {
    "fa": None,
    "no_fa": {"no_fourier_acc": 1},
    "cfa3": {"classical_fourier_acc": 1, "cfa_steps": 3, "cfa_mass": 0.5},
    "cfa2": {"classical_fourier_acc": 1, "cfa_steps": 2, "cfa_mass": 0.5},
}


@pytest.mark.parametrize("mode", list(LEAPFROG_MODES))
def test_leapfrog_second_order_and_reversible(mode):
    """leap_frog (organic_hmc.c:167-190) with update_x of each momentum scheme
    (organic_hmc.c:136-165): |dH| falls 4x per halving of the step (a misplaced or missing half
    step makes it first order), and p -> -p retraces the trajectory exactly."""
    om = OracleModel(2, 3, 8, beta=8 * 40 / 32, flags=LEAPFROG_MODES[mode])
    om.set_couplings(kappa=om.kappa * 4)
    lib = oracle()
    h = _hmc(om, 7, 4, 0.4 if mode == "no_fa" else 0.0)
    x0 = np.ctypeslib.as_array(h.x, shape=(om.dim,)).copy()
    p0 = np.zeros(om.dim)
    mt = MT()
    lib.orc_mt_seed(C.byref(mt), 11)
    lib.orc_sample_fourier_momenta(om.ref, C.byref(mt), ptr(p0))
    H0 = _H(om, x0, p0)
    if mode != "cfa2":
        dH = []
        for steps in (32, 64, 128):
            x1, p1 = _run_leapfrog(om, h, x0, p0, steps)
            dH.append(abs(_H(om, x1, p1) - H0))
        r1, r2 = dH[0] / dH[1], dH[1] / dH[2]
        assert 3.6 < r1 < 4.6 and 3.8 < r2 < 4.2, (mode, dH)
    x1, p1 = _run_leapfrog(om, h, x0, p0, 8)
    x2, p2 = _run_leapfrog(om, h, x1, -p1, 8)
    scale = np.abs(x0).max()
    np.testing.assert_allclose(x2, x0, atol=1e-10 * scale)
    np.testing.assert_allclose(-p2, p0, atol=1e-10 * np.abs(p0).max())
    lib.orc_hmc_free(C.byref(h))


# ---- current-current correlator (organic_dense_matrix.c:451-524) -------------------------------

def _dense_correlator(om, x, Minv_dense, identity_start=False):
    """g_ab(tau) from explicit matrices: R_t0(tau) = B_tau ... B_{nt-1} G and
    R_0t(tau) = B_{tau-1} ... B_0 G^T with G = M_hat^{-1} (or 1 for the identity start); the storage
    element R[r*N + c] is the matrix entry (c, r).  g = 1/N sum_ij chi_ia(0) chi_jb(tau) (4 terms)."""
    n, nt = om.ns, om.nt
    B = D.slices(om, x)
    G = np.eye(n) if identity_start else Minv_dense
    nb = om.nnt[:6 * n].reshape(n, 6)
    xr = x.reshape(nt, n, 3)
    out = np.zeros((nt, 3, 3))
    for tau in range(nt):
        Rt = G.copy()
        for s in range(nt - 1, tau - 1, -1):
            Rt = B[s] @ Rt
        R0 = (np.eye(n) if identity_start else G.T).copy()
        for s in range(tau):
            R0 = B[s] @ R0
        # S[r, c] = storage element r*N + c = matrix (c, r)
        St, S0 = Rt.T, R0.T
        for a in range(3):
            ia = nb[:, 2 * a]
            chi_a = om.t[a] - om.kappa[a] * xr[0, :, a]
            for b in range(3):
                jb = nb[:, 2 * b]
                chi_b = om.t[b] - om.kappa[b] * xr[tau, :, b]
                i = np.arange(n)[:, None]
                j = np.arange(n)[None, :]
                IA = ia[:, None]
                JB = jb[None, :]
                corr = St[j, i] * S0[JB, IA] + St[JB, IA] * S0[j, i] - St[JB, i] * S0[j, IA] - St[j, IA] * S0[JB, i]
                out[tau, a, b] = (chi_a[:, None] * chi_b[None, :] * corr).sum() / n
    return out


def _oracle_correlator(om, x, Minv_flat, logdet=0.0):
    n, nt = om.ns, om.nt
    Rt0 = np.zeros(nt * n * n)
    R0t = np.zeros(n * n)
    w = np.zeros(2 * n * n)
    g = np.zeros(nt * 9)
    oracle().orc_cc_correlator(om.ref, ptr(x), ptr(Minv_flat), ptr(Rt0), ptr(R0t), ptr(w), ptr(g), logdet)
    return g.reshape(nt, 3, 3)


@pytest.mark.parametrize("l1,l2,nt", [(2, 3, 3), (3, 3, 4), (3, 2, 2)])
def test_cc_correlator_dense(l1, l2, nt):
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    om.set_couplings(kappa=om.kappa * 20, t=om.t * 3)
    x = np.random.default_rng(l1 * 7 + nt).standard_normal(om.dim)
    Md = D.dense_M_hat(om, x)
    Minv = np.linalg.inv(Md)
    g = _oracle_correlator(om, x, Minv.T.copy().ravel())
    ref = _dense_correlator(om, x, Minv)
    np.testing.assert_allclose(g, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_cc_correlator_zero_filling_identity_start():
    """ZERO_FILLING starts both propagators from the identity (organic_dense_matrix.c:465-468)."""
    om = OracleModel(3, 2, 3, beta=3 * 40 / 32, flags={"zero_filling": 1})
    om.set_couplings(kappa=om.kappa * 20)
    x = np.random.default_rng(2).standard_normal(om.dim)
    g = _oracle_correlator(om, x, np.zeros(om.ns * om.ns))
    ref = _dense_correlator(om, x, None, identity_start=True)
    np.testing.assert_allclose(g, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("l1,l2,nt", [(4, 4, 4), (5, 3, 6)])
def test_cc_correlator_uniform_closed_form(l1, l2, nt):
    """For a field constant in space on every slice, each B_tau is diagonal in plane waves with
    eigenvalue b_tau(q), and the four Wick terms sum to
        g_ab(tau) = 4 chi_a(0) chi_b(tau) / N sum_q sin(q.e_a) sin(q.e_b) P(q) / (1 + P(q))^2,
    P = prod_tau b_tau(q): independent of tau except through chi_b(tau) (the current of a
    translation-invariant system is conserved).  Signs, the neighbour directions and the slice of
    chi_jb are all fixed by it."""
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    u = np.random.default_rng(l1 + nt).standard_normal((nt, 3)) * 2
    x = D.uniform_field(l1, l2, nt, u)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    g = _oracle_correlator(om, x, Minv, ld)
    m1, m2 = np.meshgrid(np.arange(l1), np.arange(l2), indexing="ij")
    q1, q2 = 2 * np.pi * m1.ravel() / l1, 2 * np.pi * m2.ravel() / l2
    sins = np.stack([np.sin(q1), np.sin(q2), np.sin(q2 - q1)])
    P = np.ones(l1 * l2)
    for tau in range(nt):
        eps, _ = D.plane_wave_eps(l1, l2, om.t, om.kappa, om.mu, u[tau])
        P *= np.polyval(D.POLY4[::-1], eps)
    kern = P / (1 + P) ** 2
    ref = np.zeros((nt, 3, 3))
    for tau in range(nt):
        chi0 = om.t - om.kappa * u[0]
        chit = om.t - om.kappa * u[tau]
        for a in range(3):
            for b in range(3):
                ref[tau, a, b] = 4 * chi0[a] * chit[b] * (sins[a] * sins[b] * kern).sum() / om.ns
    np.testing.assert_allclose(g, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


# ---- remaining building blocks: transpose, stencil on a matrix, all QM1P, TL bin width, results --

def test_transpose_and_stencil_on_matrix():
    """transpose (organic_dense_matrix.c:57-71) and apply_mu_K_to_M (:123-131): every row-major
    chunk (a LAPACK column) becomes A_tau times itself."""
    om = OracleModel(3, 4, 2)
    om.set_couplings(kappa=om.kappa * 20)
    lib = oracle()
    n = om.ns
    M = np.random.default_rng(5).standard_normal((n, n))
    flat = M.ravel().copy()
    lib.orc_transpose(ptr(flat), n)
    np.testing.assert_array_equal(flat.reshape(n, n), M.T)
    x = np.random.default_rng(6).standard_normal(om.dim)
    flat = M.T.copy().ravel()            # columns of M as chunks
    lib.orc_apply_mu_K_to_M(om.ref, ptr(x), ptr(flat), ptr(np.zeros(2 * n * n)))
    A = D.hopping_matrix(om.nnt, n, x[:3 * n], om.t, om.kappa, om.mu)
    np.testing.assert_allclose(flat.reshape(n, n).T, A @ M, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("nt", [1, 2, 5, 8])
def test_construct_all_QM1P_dense(nt):
    """construct_all_QM1P (organic_dense_matrix.c:311-381): slot tau holds
    B_{tau+1} ... B_{nt-1} G B_0 ... B_{tau-1} (G = M_hat^{-1}), with tau slices taken from the
    right and nt-1-tau from the left (perm)."""
    om = OracleModel(2, 3, nt, beta=nt * 40 / 32)
    om.set_couplings(kappa=om.kappa * 10)
    lib = oracle()
    n = om.ns
    x = np.random.default_rng(nt).standard_normal(om.dim)
    G = np.linalg.inv(D.dense_M_hat(om, x))
    Q = np.zeros(nt * n * n)
    perm = np.zeros(2 * nt, np.int32)
    from _oracle import ip
    lib.orc_construct_all_QM1P(om.ref, ptr(x), ptr(G.T.copy().ravel()), ptr(Q), ptr(np.zeros(5 * n * n)),
                               ptr(perm, ip))
    B = D.slices(om, x)
    for tau in range(nt):
        ref = G.copy()
        for s in range(nt - 1, tau, -1):
            ref = B[s] @ ref
        for s in range(tau):
            ref = ref @ B[s]
        np.testing.assert_allclose(Q[tau * n * n:(tau + 1) * n * n].reshape(n, n).T, ref, rtol=1e-11,
                                   atol=1e-13 * np.abs(ref).max())
    np.testing.assert_array_equal(perm[:nt], nt - 1 - np.arange(nt))
    np.testing.assert_array_equal(perm[nt:], np.arange(nt))


def test_tl_bin_width():
    """Bin width of run_trans_loc (organic_trans_loc.c:99-100): (-2 mu + 8 sum|kappa|) / Nt, at
    beta = 40, J = (0.10, -0.03, 0.02), lambda = 0.3, mu0 = 0: mu = -12, kappa = 0.3 t."""
    om = OracleModel(4, 4, 512, static_disorder=True)
    w = oracle().orc_tl_width(om.ref, 512)
    assert w == pytest.approx((24 + 8 * 0.3 * 40 * 0.15) / 512, rel=1e-14)


def test_trajectory_results_columns():
    """results[0..10] of trajectory (organic_hmc.c:256-279): <x>, <x^2> - 1/2, H, -log det,
    higher Matsubara norm, 1 - tr(M_hat^{-1})/N, acceptance, exp(-dH), link averages; recomputed
    from the state the trajectory leaves."""
    om = OracleModel(2, 2, 4, beta=4 * 40 / 32)
    om.set_couplings(kappa=om.kappa * 5)
    lib = oracle()
    h = _hmc(om, 3, 4)
    res = np.zeros(12)
    dH = C.c_double()
    acc = lib.orc_trajectory(C.byref(h), ptr(res), None, C.byref(dH), None, None)
    x = np.ctypeslib.as_array(h.x, shape=(om.dim,)).copy()
    Minv = np.ctypeslib.as_array(h.M_hat, shape=(om.ns * om.ns,)).reshape(om.ns, om.ns)
    ld = om.logdet(x)
    assert res[0] == pytest.approx(x.mean(), rel=1e-13)
    assert res[1] == pytest.approx((x ** 2).mean() - 0.5, rel=1e-13)
    assert res[3] == pytest.approx(-ld, rel=1e-12)
    assert res[3] == pytest.approx(-h.logdet, rel=1e-15)
    xr = x.reshape(om.nt, om.ns, 3)
    assert res[4] == pytest.approx(((xr - xr.mean(0)) ** 2).sum(), rel=1e-12)
    assert res[5] == pytest.approx(1 - np.trace(Minv) / om.ns, rel=1e-12)
    assert res[6] == acc
    assert res[7] == pytest.approx(np.exp(-dH.value), rel=1e-12)
    np.testing.assert_allclose(res[8:11], xr.sum((0, 1)) / om.ns, rtol=1e-12)
    lib.orc_hmc_free(C.byref(h))
