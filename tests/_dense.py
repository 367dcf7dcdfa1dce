"""Dense numpy building blocks used to pin the oracle and the GPU path independently of both:
the hopping matrix A_tau as an explicit N x N matrix, the Taylor-pair polynomial as matrix
products, and the closed forms for spatially uniform fields (plane waves diagonalise every A_tau).
Matrices here are ordinary numpy (row, col) arrays; the column-major chunk storage of the
reference corresponds to .T of the flat buffer reshaped (N, N)."""
import numpy as np

# Taylor pair coefficients (simulation/taylor-coeff_pairs.h:35-42) applied in the order of
# set_taylor_pair_order (organic_dense_matrix.c:133-159): a_i = .5 c_i / 2
PAIRS4 = [0.25 * (1.8294933739891903 + 0.9404019519941707j), 0.25 * (0.1705066260108097 + 1.5785318126884535j)]
POLY4 = np.array([1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24])   # exp series up to z^4


def pair_poly(pairs=PAIRS4):
    poly = np.array([1.0])
    for a in pairs:
        poly = np.convolve(poly, [1.0, 2 * a.real, abs(a) ** 2] if a.imag else [1.0, 2 * a.real])
    return poly


def hopping_matrix(nnt, ns, x_tau, t, kappa, mu):
    """A = mu + K(x_tau): symmetric, bond (i, nb(i,2k)) carries t_k - kappa_k x[i,k]."""
    A = np.zeros((ns, ns))
    A[np.arange(ns), np.arange(ns)] = mu
    nb = nnt[:6 * ns].reshape(ns, 6)
    for i in range(ns):
        for k in range(3):
            f = t[k] - kappa[k] * x_tau[3 * i + k]
            j = nb[i, 2 * k]
            A[i, j] += f
            A[j, i] += f
    return A


def poly_matrix(A, pairs=PAIRS4):
    B = np.eye(A.shape[0])
    for a in pairs:
        if a.imag:
            B = B @ (np.eye(A.shape[0]) + 2 * a.real * A + abs(a) ** 2 * (A @ A))
        else:
            B = B @ (np.eye(A.shape[0]) + 2 * a.real * A)
    return B


def slices(om, x):
    ns, nt = om.ns, om.nt
    return [poly_matrix(hopping_matrix(om.nnt, ns, x[tau * 3 * ns:(tau + 1) * 3 * ns], om.t, om.kappa, om.mu))
            for tau in range(nt)]


def dense_M_hat(om, x, plus_one=True):
    B = slices(om, x)
    P = np.eye(om.ns)
    for b in B:
        P = P @ b
    return (np.eye(om.ns) + P) if plus_one else P


def uniform_field(l1, l2, nt, u):
    """x[tau, i, k] = u[tau, k] for every site i."""
    u = np.asarray(u, float).reshape(nt, 3)
    return np.repeat(u[:, None, :], l1 * l2, axis=1).ravel()


def plane_wave_eps(l1, l2, t, kappa, mu, u_tau):
    """eps(q) = mu + 2 sum_k (t_k - kappa_k u_k) cos(q.e_k), e = (1,0), (0,1), (-1,1)."""
    m1, m2 = np.meshgrid(np.arange(l1), np.arange(l2), indexing="ij")
    q1 = 2 * np.pi * m1.ravel() / l1
    q2 = 2 * np.pi * m2.ravel() / l2
    cosk = np.stack([np.cos(q1), np.cos(q2), np.cos(q2 - q1)])
    f = np.asarray(t) - np.asarray(kappa) * np.asarray(u_tau)
    return mu + 2 * (f[:, None] * cosk).sum(0), cosk


def closed_form_logdet(l1, l2, nt, t, kappa, mu, u, poly=POLY4):
    u = np.asarray(u).reshape(nt, 3)
    prod = np.ones(l1 * l2)
    for tau in range(nt):
        eps, _ = plane_wave_eps(l1, l2, t, kappa, mu, u[tau])
        prod *= np.polyval(poly[::-1], eps)
    return np.log1p(prod).sum()


def closed_form_force(l1, l2, nt, t, kappa, mu, u, flavors, poly=POLY4):
    """p_dot[tau, i, k] for a uniform field: flavors/N * d logdet / d u[tau, k] (same for every i)."""
    u = np.asarray(u).reshape(nt, 3)
    ns = l1 * l2
    vals, derivs, cos_all = [], [], None
    for tau in range(nt):
        eps, cosk = plane_wave_eps(l1, l2, t, kappa, mu, u[tau])
        vals.append(np.polyval(poly[::-1], eps))
        derivs.append(np.polyval(np.polyder(poly[::-1]), eps))
        cos_all = cosk
    vals = np.array(vals)
    prod = vals.prod(0)
    frac = prod / (1 + prod)
    F = np.zeros((nt, 3))
    for tau in range(nt):
        for k in range(3):
            d = frac * derivs[tau] / vals[tau] * (-2 * kappa[k] * cos_all[k])
            F[tau, k] = flavors * d.sum() / ns
    return np.repeat(F[:, None, :], ns, axis=1).ravel()
