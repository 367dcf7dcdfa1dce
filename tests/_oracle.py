"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference pieces compiled
into oracle/_ref/libref.so.  Test infrastructure only: the product never imports this."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

dp = C.POINTER(C.c_double)
up = C.POINTER(C.c_uint)
ip = C.POINTER(C.c_int)


def ptr(a, t=dp):
    return a.ctypes.data_as(t)


class Flags(C.Structure):
    """organic_flags (simulation/organic_flags.h:4-14)."""
    _fields_ = [("schur_form_inv", C.c_int), ("zero_filling", C.c_int), ("static_disorder", C.c_int),
                ("single_particle", C.c_int), ("constant_force", C.c_int), ("no_fourier_acc", C.c_int),
                ("classical_fourier_acc", C.c_int), ("cfa_steps", C.c_uint), ("cfa_mass", C.c_double)]


class MT(C.Structure):
    _fields_ = [("s", C.c_uint64 * 312), ("i", C.c_int)]


class Model(C.Structure):
    _fields_ = [("ns", C.c_uint), ("nn", C.c_uint), ("nt", C.c_uint), ("nnt", up),
                ("t", C.c_double * 3), ("kappa", C.c_double * 3), ("mu", C.c_double), ("omega", C.c_double),
                ("flavors", C.c_uint), ("np", C.c_int), ("a_re", C.c_double * 26), ("a_im", C.c_double * 26),
                ("mode", Flags)]


class HMC(C.Structure):
    _fields_ = [("m", C.POINTER(Model)), ("rng", MT)] + \
        [(n, dp) for n in ("x", "p", "p_dot", "x_old", "M_hat", "M_hat_old", "QM1P", "R_0t", "w")] + \
        [("perm", ip), ("logdet", C.c_double), ("steps", C.c_uint), ("traj_length", C.c_double),
         ("contr_mat", dp)]


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        lib = C.CDLL(ORACLE_SO)
        sig = {
            "orc_mt_seed": (None, [C.POINTER(MT), C.c_uint64]),
            "orc_mt_u64": (C.c_uint64, [C.POINTER(MT)]),
            "orc_mt_real2": (C.c_double, [C.POINTER(MT)]),
            "orc_random_vector": (None, [C.POINTER(MT), dp, C.c_uint]),
            "orc_construct_triangular": (up, [C.c_uint, C.c_uint, up, up]),
            "orc_construct_contraction_mat": (dp, [dp, dp, C.c_uint]),
            "orc_contract_greens": (C.c_double, [dp, dp, C.c_uint, C.c_uint]),
            "orc_set_taylor": (C.c_int, [C.POINTER(Model), C.c_int]),
            "orc_scale_params": (None, [C.c_double, dp, dp, C.c_double, C.c_double, C.c_uint, C.c_int,
                                        dp, dp, dp, dp, dp]),
            "orc_apply_mu_K_to_v": (None, [C.POINTER(Model), dp, dp, dp, C.c_double]),
            "orc_apply_mu_K_to_M": (None, [C.POINTER(Model), dp, dp, dp]),
            "orc_apply_exp_mu_K_to_v": (None, [C.POINTER(Model), dp, dp, dp, C.c_int, C.c_int]),
            "orc_apply_exp_mu_K_to_M": (None, [C.POINTER(Model), dp, dp, dp, C.c_int, C.c_int]),
            "orc_construct_M_hat": (None, [C.POINTER(Model), dp, dp, dp]),
            "orc_transpose": (None, [dp, C.c_uint]),
            "orc_logdet_inv": (C.c_double, [C.POINTER(Model), dp, ip, ip]),
            "orc_construct_QM1P": (None, [C.POINTER(Model), dp, dp, dp, dp, C.c_int]),
            "orc_construct_all_QM1P": (None, [C.POINTER(Model), dp, dp, dp, dp, ip]),
            "orc_refine_QM1P": (None, [C.POINTER(Model), dp, dp, dp, C.c_int]),
            "orc_force": (None, [C.POINTER(Model), dp, dp, dp, dp, dp, ip, C.c_double]),
            "orc_force_workspace": (C.c_size_t, [C.POINTER(Model)]),
            "orc_r2c_many": (None, [C.c_uint, C.c_uint, dp, dp]),
            "orc_c2r_many": (None, [C.c_uint, C.c_uint, dp, dp]),
            "orc_sample_free_bosons": (None, [C.POINTER(Model), C.POINTER(MT), dp]),
            "orc_sample_fourier_momenta": (None, [C.POINTER(Model), C.POINTER(MT), dp]),
            "orc_energy_momenta": (C.c_double, [C.POINTER(Model), dp]),
            "orc_evolve_bosons": (None, [C.POINTER(Model), dp, dp, C.c_double]),
            "orc_evolve_bosons_approx_step": (None, [C.POINTER(Model), dp, dp, C.c_double]),
            "orc_boson_force": (None, [C.POINTER(Model), dp, dp, C.c_int]),
            "orc_higher_matsubaras_norm": (C.c_double, [C.POINTER(Model), dp]),
            "orc_hamilton": (C.c_double, [C.POINTER(Model), dp, C.c_double, C.c_double]),
            "orc_update_x": (None, [C.POINTER(Model), dp, dp, dp, C.c_double, C.c_double]),
            "orc_hmc_init": (C.c_int, [C.POINTER(HMC), C.POINTER(Model), C.c_uint64, C.c_uint, C.c_double, dp]),
            "orc_hmc_free": (None, [C.POINTER(HMC)]),
            "orc_leap_frog": (None, [C.POINTER(HMC)]),
            "orc_trajectory": (C.c_int, [C.POINTER(HMC), dp, dp, dp, C.c_void_p, C.c_void_p]),
            "orc_cc_correlator": (None, [C.POINTER(Model), dp, dp, dp, dp, dp, dp, C.c_double]),
            "orc_construct_hamilton": (None, [C.POINTER(Model), dp, dp, dp]),
            "orc_apply_current_to_v": (None, [C.POINTER(Model), dp, dp, dp, C.c_uint]),
            "orc_eigh": (C.c_int, [dp, dp, C.c_uint]),
            "orc_hist_trans_loc": (C.c_double, [C.POINTER(Model), dp, dp, C.c_double, dp, C.c_uint]),
            "orc_tl_width": (C.c_double, [C.POINTER(Model), C.c_uint]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _orc = lib
    return _orc


def ref():
    """The reference's own lattice/RNG/aux/flags sources, compiled by oracle/Makefile."""
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.construct_triangular.restype = up
        lib.construct_triangular.argtypes = [C.c_uint, C.c_uint, up, up]
        lib.init_genrand64.argtypes = [C.c_ulong]
        lib.genrand64_int64.restype = C.c_ulong
        lib.genrand64_real2.restype = C.c_double
        lib.random_vector.argtypes = [dp, C.c_uint]
        lib.construct_contraction_mat.restype = dp
        lib.construct_contraction_mat.argtypes = [dp, dp, C.c_uint]
        lib.contract_greens.restype = C.c_double
        lib.contract_greens.argtypes = [dp, dp, C.c_uint, C.c_uint]
        lib.fprint_results.argtypes = [C.c_void_p, dp, C.c_uint, C.c_uint]
        lib.fprint_corr.argtypes = [C.c_void_p, dp, dp, C.c_uint, C.c_uint]
        lib.fwrite_sigma.argtypes = [C.c_char_p, dp, C.c_double, C.c_uint]
        lib.set_flags.argtypes = [C.c_void_p, C.POINTER(Flags)]
        lib.average_links.argtypes = [dp, dp, C.c_uint, C.c_uint, C.c_uint, dp]
        lib.average.restype = C.c_double
        lib.average.argtypes = [dp, C.c_uint]
        lib.average_sq.restype = C.c_double
        lib.average_sq.argtypes = [dp, C.c_uint]
        lib.ceil_log2.restype = C.c_uint
        lib.ceil_log2.argtypes = [C.c_uint]
        _ref = lib
    return _ref


_libc = None


def libc():
    global _libc
    if _libc is None:
        _libc = C.CDLL(None)
        _libc.fopen.restype = C.c_void_p
        _libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        _libc.fclose.argtypes = [C.c_void_p]
        _libc.free.argtypes = [C.c_void_p]
    return _libc


# ---------------------------------------------------------------------------------------
# Convenience: build an oracle model for the reference's couplings

J_DEFAULT = (0.10, -0.03, 0.02)
LAMBDA_DEFAULT = (0.3, 0.3, 0.3)


class OracleModel:
    """Owns the lattice table and a Model struct (lattice units, organic_main.c:99-116)."""

    def __init__(self, l1, l2, nt, beta=40.0, J=J_DEFAULT, lam=LAMBDA_DEFAULT, mu0=0.0, omega0=0.006,
                 flavors=2, taylor=4, flags=None, static_disorder=False):
        lib = oracle()
        ns, nn = C.c_uint(), C.c_uint()
        self._nnt_p = lib.orc_construct_triangular(l1, l2, C.byref(ns), C.byref(nn))
        self.ns, self.nt, self.l1, self.l2 = ns.value, nt, l1, l2
        self.nnt = np.ctypeslib.as_array(self._nnt_p, shape=(12 * self.ns,)).copy()
        t = np.zeros(3)
        kap = np.zeros(3)
        mu = C.c_double()
        om = C.c_double()
        xa = C.c_double()
        lib.orc_scale_params(beta, ptr(np.array(J, float)), ptr(np.array(lam, float)), mu0, omega0, nt,
                             int(static_disorder), ptr(t), ptr(kap), C.byref(mu), C.byref(om), C.byref(xa))
        self.t, self.kappa, self.mu, self.omega, self.x_avg = t, kap, mu.value, om.value, xa.value
        m = Model()
        m.ns, m.nn, m.nt = self.ns, 6, nt
        m.nnt = self._nnt_p
        m.t[:] = list(t)
        m.kappa[:] = list(kap)
        m.mu, m.omega, m.flavors = self.mu, self.omega, flavors
        assert lib.orc_set_taylor(C.byref(m), taylor) == 0
        if flags:
            for k, v in flags.items():
                setattr(m.mode, k, v)
        self.m = m
        self.flavors = flavors
        self.dim = nt * self.ns * 3

    @property
    def ref(self):
        return C.byref(self.m)

    def set_couplings(self, t=None, kappa=None, mu=None):
        if t is not None:
            self.t = np.asarray(t, float)
            self.m.t[:] = list(self.t)
        if kappa is not None:
            self.kappa = np.asarray(kappa, float)
            self.m.kappa[:] = list(self.kappa)
        if mu is not None:
            self.mu = float(mu)
            self.m.mu = self.mu

    # -- hot-path wrappers ------------------------------------------------------------
    def M_hat(self, x):
        n = self.ns
        M = np.zeros(n * n)
        w = np.zeros(2 * n * n)
        oracle().orc_construct_M_hat(self.ref, ptr(x), ptr(M), ptr(w))
        return M

    def logdet_inv(self, M):
        M = M.copy()
        perm = np.zeros(self.ns, np.int32)
        sign = C.c_int()
        ld = oracle().orc_logdet_inv(self.ref, ptr(M), ptr(perm, ip), C.byref(sign))
        return ld, sign.value, M

    def force(self, x, Minv, logdet=0.0):
        lib = oracle()
        n = self.ns
        ws = lib.orc_force_workspace(self.ref)
        Q = np.zeros(ws)
        w = Q[ws - 5 * n * n:]
        perm = np.zeros(max(n, 2 * self.nt), np.int32)
        pdot = np.zeros(self.dim)
        lib.orc_force(self.ref, ptr(x), ptr(pdot), ptr(Minv), ptr(Q), ptr(w), ptr(perm, ip), logdet)
        return pdot

    def logdet(self, x):
        return self.logdet_inv(self.M_hat(x))[0]
