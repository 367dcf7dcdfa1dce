"""Pins the oracle's hot path against mathematics: the Taylor polynomial identity, dense numpy
brute force on small lattices, closed forms for spatially uniform fields, numpy slogdet/inv,
central finite differences, numpy FFT conventions and exact invariants of the boson flow."""
import ctypes as C

import numpy as np
import pytest

import _dense as D
from _oracle import MT, OracleModel, dp, ip, oracle, ptr

RNG = np.random.default_rng(12345)


def rand_x(om, scale=1.0, seed=0):
    return np.random.default_rng(seed).standard_normal(om.dim) * scale


def test_taylor_pairs_are_exp_series():
    assert np.allclose(D.pair_poly(), D.POLY4, rtol=0, atol=1e-16)
    # through the oracle's stencil on a one-site lattice: A = z exactly when t = kappa = 0
    om = OracleModel(1, 1, 1)
    om.set_couplings(t=[0, 0, 0], kappa=[0, 0, 0])
    for z in np.linspace(-3, 1, 41):
        om.set_couplings(mu=z)
        v = np.array([1.0])
        w = np.zeros(2)
        oracle().orc_apply_exp_mu_K_to_v(om.ref, ptr(np.zeros(3)), ptr(v), ptr(w), 0, 2)
        assert v[0] == pytest.approx(np.polyval(D.POLY4[::-1], z), rel=1e-14, abs=1e-15)


@pytest.mark.parametrize("l1,l2", [(1, 3), (2, 2), (3, 4), (5, 5), (4, 6)])
def test_stencil_matches_dense(l1, l2):
    om = OracleModel(l1, l2, 2, beta=40.0)
    om.set_couplings(kappa=om.kappa * 20)          # make the field matter
    x = rand_x(om, seed=l1 * 10 + l2)
    v = RNG.standard_normal(om.ns)
    w = np.zeros(om.ns)
    oracle().orc_apply_mu_K_to_v(om.ref, ptr(x), ptr(v), ptr(w), om.mu)
    A = D.hopping_matrix(om.nnt, om.ns, x[:3 * om.ns], om.t, om.kappa, om.mu)
    np.testing.assert_allclose(w, A @ v, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("revert,stop", [(0, 2), (0, 1), (0, 0), (1, 0), (1, -1), (1, 1)])
def test_partial_pair_products(revert, stop):
    om = OracleModel(3, 4, 2)
    om.set_couplings(kappa=om.kappa * 30, t=om.t * 5)
    x = rand_x(om, seed=3)
    n = om.ns
    M = RNG.standard_normal((n, n))
    flat = M.T.copy().ravel()          # column-major chunks
    w = np.zeros(2 * n * n)
    oracle().orc_apply_exp_mu_K_to_M(om.ref, ptr(x), ptr(flat), ptr(w), revert, stop)
    A = D.hopping_matrix(om.nnt, n, x[:3 * n], om.t, om.kappa, om.mu)
    idx = [i for i in range(2) if (i > stop if revert else i < stop)]
    B = D.poly_matrix(A, [D.PAIRS4[i] for i in idx])
    np.testing.assert_allclose(flat.reshape(n, n).T, B @ M, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("l1,l2,nt", [(2, 2, 3), (3, 3, 4), (3, 4, 8), (4, 4, 6)])
def test_M_hat_matches_dense(l1, l2, nt):
    om = OracleModel(l1, l2, nt, beta=40.0)
    om.set_couplings(kappa=om.kappa * 20, t=om.t * 4)
    x = rand_x(om, seed=nt)
    M = om.M_hat(x).reshape(om.ns, om.ns).T
    np.testing.assert_allclose(M, D.dense_M_hat(om, x), rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("l1,l2,nt", [(6, 6, 32), (8, 5, 16), (2, 3, 8)])
def test_logdet_uniform_closed_form(l1, l2, nt):
    om = OracleModel(l1, l2, nt)
    u = np.random.default_rng(l1 + nt).standard_normal((nt, 3)) * 2
    x = D.uniform_field(l1, l2, nt, u)
    ld = om.logdet(x)
    ref = D.closed_form_logdet(l1, l2, nt, om.t, om.kappa, om.mu, u)
    assert ld == pytest.approx(ref, rel=1e-12)


def test_logdet_inverse_numpy():
    om = OracleModel(4, 4, 8)
    om.set_couplings(kappa=om.kappa * 3)
    x = rand_x(om, seed=9)
    M = om.M_hat(x)
    ld, sign, Minv = om.logdet_inv(M)
    A = M.reshape(om.ns, om.ns).T
    s, l = np.linalg.slogdet(A)
    assert ld == pytest.approx(l, rel=1e-13) and sign == int(s)
    np.testing.assert_allclose(Minv.reshape(om.ns, om.ns).T, np.linalg.inv(A), rtol=1e-10, atol=1e-13)
    # general matrices with sign changes and pivoting
    for seed in range(4):
        B = np.random.default_rng(seed).standard_normal((om.ns, om.ns))
        ld, sign, Binv = om.logdet_inv(B.T.copy().ravel())
        s, l = np.linalg.slogdet(B)
        assert ld == pytest.approx(l, rel=1e-12) and sign == int(s)
        np.testing.assert_allclose(Binv.reshape(om.ns, om.ns).T, np.linalg.inv(B), rtol=1e-9, atol=1e-11)


def _fd_force(om, x, idx, eps=1e-5):
    out = []
    for i in idx:
        xp = x.copy(); xp[i] += eps
        xm = x.copy(); xm[i] -= eps
        out.append(om.flavors * (om.logdet(xp) - om.logdet(xm)) / (2 * eps))
    return np.array(out)


@pytest.mark.parametrize("l1,l2,nt", [(3, 3, 4), (3, 4, 5), (4, 4, 8)])
def test_force_finite_differences(l1, l2, nt):
    # time step of the L=6, Nt=32 configuration (beta/Nt = 1.25 /eV): the Taylor slices are
    # accurate there; a much larger step makes M_hat too ill-conditioned for differences
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    om.set_couplings(kappa=om.kappa * 20)
    x = rand_x(om, seed=l1 * nt)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    F = om.force(x, Minv, ld)
    idx = np.arange(om.dim) if om.dim <= 200 else np.random.default_rng(1).choice(om.dim, 64, replace=False)
    fd = _fd_force(om, x, idx)
    err = np.linalg.norm(F[idx] - fd) / np.linalg.norm(fd)
    assert err < 1e-8


@pytest.mark.parametrize("l1,l2,nt", [(6, 6, 8), (5, 4, 16), (2, 2, 4)])
def test_force_uniform_closed_form(l1, l2, nt):
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    u = np.random.default_rng(7).standard_normal((nt, 3)) * 3
    x = D.uniform_field(l1, l2, nt, u)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    F = om.force(x, Minv, ld)
    ref = D.closed_form_force(l1, l2, nt, om.t, om.kappa, om.mu, u, om.flavors)
    np.testing.assert_allclose(F, ref, rtol=1e-10, atol=1e-14 * np.abs(ref).max())


def test_constant_force_is_slice_zero_force():
    """CONSTANT_FORCE (organic_hmc.c:92-98) reuses the tau=0 refined matrix for every slice."""
    om = OracleModel(3, 3, 4, beta=4 * 40 / 32)
    om.set_couplings(kappa=om.kappa * 20)
    x = rand_x(om, seed=4)
    ld, _, Minv = om.logdet_inv(om.M_hat(x))
    F = om.force(x, Minv, ld).reshape(om.nt, -1)
    omc = OracleModel(3, 3, 4, beta=4 * 40 / 32, flags={"constant_force": 1})
    omc.set_couplings(kappa=om.kappa, t=om.t)
    Fc = omc.force(x, Minv, ld).reshape(om.nt, -1)
    # the single refined matrix is the tau = 0 one, projected identically for every slice
    np.testing.assert_allclose(Fc[0], F[0], rtol=1e-12)
    for tau in range(1, om.nt):
        np.testing.assert_array_equal(Fc[tau], Fc[0])


@pytest.mark.parametrize("nt", [1, 2, 7, 12, 16, 48, 128])
def test_fft_matches_numpy_conventions(nt):
    lib = oracle()
    how = 5
    x = RNG.standard_normal((nt, how))
    out = np.zeros(2 * (nt // 2 + 1) * how)
    lib.orc_r2c_many(nt, how, ptr(x.ravel().copy()), ptr(out))
    X = out.reshape(nt // 2 + 1, how, 2)
    ref = np.fft.rfft(x, axis=0)
    np.testing.assert_allclose(X[..., 0] + 1j * X[..., 1], ref, rtol=1e-12, atol=1e-12)
    back = np.zeros(nt * how)
    lib.orc_c2r_many(nt, how, ptr(out), ptr(back))
    np.testing.assert_allclose(back.reshape(nt, how), nt * x, rtol=1e-12, atol=1e-12)


def test_evolve_bosons_exact_flow():
    """Each Matsubara mode rotates with unit frequency: the free boson energy is conserved,
    h -> -h inverts the flow, and the result matches an explicit numpy rotation."""
    om = OracleModel(3, 2, 12)
    lib = oracle()
    x = RNG.standard_normal(om.dim)
    p = RNG.standard_normal(om.dim)
    x0, p0 = x.copy(), p.copy()
    E0 = lib.orc_energy_momenta(om.ref, ptr(p)) + lib.orc_hamilton(om.ref, ptr(x), 0.0, 0.0)
    lib.orc_evolve_bosons(om.ref, ptr(x), ptr(p), 0.37)
    E1 = lib.orc_energy_momenta(om.ref, ptr(p)) + lib.orc_hamilton(om.ref, ptr(x), 0.0, 0.0)
    assert E1 == pytest.approx(E0, rel=1e-13)
    nt, loc = om.nt, 3 * om.ns
    n = np.arange(nt // 2 + 1)
    f2 = om.omega + (2 * np.sin(np.pi * n / nt)) ** 2 / om.omega
    X = np.fft.rfft(x0.reshape(nt, loc), axis=0)
    P = np.fft.rfft(p0.reshape(nt, loc), axis=0)
    c, s = np.cos(0.37), np.sin(0.37)
    Xn = c * X + (s / f2)[:, None] * P
    Pn = c * P - (s * f2)[:, None] * X
    np.testing.assert_allclose(x, np.fft.irfft(Xn, nt, axis=0).ravel(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(p, np.fft.irfft(Pn, nt, axis=0).ravel(), rtol=1e-11, atol=1e-12)
    lib.orc_evolve_bosons(om.ref, ptr(x), ptr(p), -0.37)
    np.testing.assert_allclose(x, x0, atol=1e-12)
    np.testing.assert_allclose(p, p0, atol=1e-12)


def test_momentum_refresh_and_free_bosons_colouring():
    om = OracleModel(2, 2, 16)
    lib = oracle()
    nt, loc = om.nt, 3 * om.ns
    n = np.arange(nt // 2 + 1)
    f2 = om.omega + (2 * np.sin(np.pi * n / nt)) ** 2 / om.omega
    for fn, gain in ((lib.orc_sample_fourier_momenta, np.sqrt(f2)), (lib.orc_sample_free_bosons, 1 / np.sqrt(f2))):
        a, b = MT(), MT()
        lib.orc_mt_seed(C.byref(a), 99)
        lib.orc_mt_seed(C.byref(b), 99)
        xi = np.zeros(om.dim)
        lib.orc_random_vector(C.byref(b), ptr(xi), om.dim)
        out = np.zeros(om.dim)
        fn(om.ref, C.byref(a), ptr(out))
        ref = np.fft.irfft(np.fft.rfft(xi.reshape(nt, loc), axis=0) * gain[:, None], nt, axis=0).ravel()
        np.testing.assert_allclose(out, ref, rtol=1e-11, atol=1e-12)
        # energy of the coloured momenta: E = sum_n |P_n|^2 / (2 nt f2_n) over the full spectrum
    p = np.random.default_rng(2).standard_normal(om.dim)
    P = np.fft.fft(p.reshape(nt, loc), axis=0)
    nf = np.arange(nt)
    f2f = om.omega + (2 * np.sin(np.pi * nf / nt)) ** 2 / om.omega
    ref = (np.abs(P) ** 2 / (2 * nt * f2f[:, None])).sum()
    assert lib.orc_energy_momenta(om.ref, ptr(p)) == pytest.approx(ref, rel=1e-12)


def test_hamilton_and_observables():
    om = OracleModel(3, 3, 6)
    lib = oracle()
    x = RNG.standard_normal(om.dim)
    xr = x.reshape(om.nt, om.ns, 3)
    pot = (xr ** 2).sum()
    kin = ((np.roll(xr, -1, axis=0) - xr) ** 2).sum()
    H = lib.orc_hamilton(om.ref, ptr(x), 1.5, 2.0)
    assert H == pytest.approx(-1.5 * om.flavors + 2.0 + (kin / om.omega + pot * om.omega) / 2, rel=1e-13)
    hm = lib.orc_higher_matsubaras_norm(om.ref, ptr(x))
    assert hm == pytest.approx(((xr - xr.mean(0)) ** 2).sum(), rel=1e-12)


def test_hmc_free_bosons_exact():
    """With kappa = 0 the fermion force vanishes and Fourier acceleration integrates the bosons
    exactly: dH = 0 to rounding and every trajectory is accepted."""
    om = OracleModel(3, 3, 8)
    om.set_couplings(kappa=[0, 0, 0])
    from _oracle import HMC
    lib = oracle()
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), 5, 10, 0.0, None) == 0
    dH = C.c_double()
    for _ in range(5):
        assert lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None) == 1
        assert abs(dH.value) < 1e-9
    lib.orc_hmc_free(C.byref(h))


def test_hmc_detailed_balance_statistics():
    """<exp(-dH)> = 1 (organic_hmc.c:280) within statistical error on a small interacting system."""
    om = OracleModel(3, 3, 8)
    om.set_couplings(kappa=om.kappa * 15)
    from _oracle import HMC
    lib = oracle()
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), 17, 4, 0.0, None) == 0
    dH = C.c_double()
    vals = []
    for _ in range(60):
        lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None)
        vals.append(np.exp(-dH.value))
    vals = np.array(vals)
    assert abs(vals.mean() - 1) < 4 * vals.std() / np.sqrt(len(vals)) + 0.02
    lib.orc_hmc_free(C.byref(h))


def test_eigh_matches_numpy():
    lib = oracle()
    for n in (1, 2, 5, 17, 40):
        A = np.random.default_rng(n).standard_normal((n, n))
        A = A + A.T
        a = A.copy().ravel()
        ev = np.zeros(n)
        assert lib.orc_eigh(ptr(a), ptr(ev), n) == 0
        w, V = np.linalg.eigh(A)
        np.testing.assert_allclose(ev, w, rtol=1e-12, atol=1e-12)
        Vo = a.reshape(n, n)          # chunk i = eigenvector i
        np.testing.assert_allclose(np.abs(np.sum(Vo * V.T, axis=1)), 1, atol=1e-10)


def test_trans_loc_bins_brute_force():
    om = OracleModel(4, 3, 16, beta=40.0, static_disorder=True)
    lib = oracle()
    nbins = 16
    cm = np.array([7.2, 3.6, -3.6, 0.0, 7.15, 7.15])
    x = np.random.default_rng(8).standard_normal(3 * om.ns)
    width = lib.orc_tl_width(om.ref, nbins)
    g = np.zeros(nbins)
    z = lib.orc_hist_trans_loc(om.ref, ptr(x), ptr(cm), width, ptr(g), nbins)
    # brute force: dense H, numpy eigh, current matrices, all pairs E_j >= E_i
    n = om.ns
    H = D.hopping_matrix(om.nnt, n, x, om.t, om.kappa, -om.mu)
    E, V = np.linalg.eigh(H)
    nb = om.nnt[:6 * n].reshape(n, 6)
    J = []
    for k in range(3):
        Ik = np.zeros((n, n))
        for i in range(n):
            f = om.t[k] - om.kappa[k] * x[3 * i + k]
            Ik[i, nb[i, 2 * k]] += f
            Ik[nb[i, 2 * k], i] -= f
        J.append(V.T @ Ik @ V)        # proj_k[j, i] = <v_j| I_k |v_i>
    Cm = cm.reshape(2, 3)
    ref = np.zeros(nbins)
    for i in range(n):
        for j in range(n):
            if E[j] < E[i]:
                continue
            proj = np.array([J[k][j, i] for k in range(3)])
            Inm = ((Cm @ proj) ** 2).sum()
            b = int((E[j] - E[i]) / width + .5)
            if b < nbins:
                ref[b] += (np.exp(-E[i]) + np.exp(-E[j])) * Inm / (width * n)
    np.testing.assert_allclose(g, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    assert z == pytest.approx(np.exp(-E).sum() / n, rel=1e-12)
