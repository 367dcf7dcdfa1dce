"""ctypes binding of libephmc.so (the C-ABI declared in include/ephmc.h).

Host arrays are numpy float64; device buffers are DeviceArray objects allocated through the
library's own ephmc_malloc.  Importing this module raises if the HIP library is missing: there
is no CPU fallback.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "libephmc.so")

FP64, MIXED = 0, 1
ERRORS = {-1: "argument", -2: "HIP", -3: "unsupported", -4: "allocation", -5: "numeric"}


class Config(C.Structure):
    """ephmc_config (include/ephmc.h)."""
    _fields_ = [("l1", C.c_uint), ("l2", C.c_uint), ("nt", C.c_uint),
                ("t", C.c_double * 3), ("kappa", C.c_double * 3), ("mu", C.c_double), ("omega", C.c_double),
                ("flavors", C.c_uint), ("steps", C.c_uint), ("traj_length", C.c_double),
                ("taylor_order", C.c_int), ("precision", C.c_int), ("checkpoint", C.c_uint),
                ("zero_filling", C.c_int), ("single_particle", C.c_int), ("constant_force", C.c_int),
                ("no_fourier_acc", C.c_int), ("classical_fourier_acc", C.c_int),
                ("cfa_steps", C.c_uint), ("cfa_mass", C.c_double),
                ("device", C.c_int), ("stream", C.c_void_p), ("graph", C.c_int), ("schur_form_inv", C.c_int), ("stabilize", C.c_uint)]


class TuneResult(C.Structure):
    _fields_ = [("chain_cols", C.c_uint), ("adjoint_cols", C.c_uint), ("chain_ms", C.c_double), ("adjoint_ms", C.c_double),
                ("adjoint_w0_early", C.c_int), ("chain_strip", C.c_uint), ("chain64_cols", C.c_uint),
                ("chain64_strip", C.c_uint), ("chain64_ms", C.c_double)]


class Record(C.Structure):
    _fields_ = [("accepted", C.c_int), ("boltzmann", C.c_double), ("energy", C.c_double),
                ("logdet", C.c_double), ("dH", C.c_double)]


class EphmcError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise EphmcError("libephmc.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    sig = {
        "ephmc_create": [C.POINTER(Config), C.POINTER(C.c_void_p)],
        "ephmc_destroy": [vp],
        "ephmc_synchronize": [vp],
        "ephmc_dims": [vp, C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_size_t)],
        "ephmc_lattice": [vp, C.POINTER(C.c_uint)],
        "ephmc_malloc": [C.c_size_t, C.POINTER(C.c_void_p)],
        "ephmc_free": [vp],
        "ephmc_memcpy_h2d": [vp, vp, vp, C.c_size_t],
        "ephmc_memcpy_d2h": [vp, vp, vp, C.c_size_t],
        "ephmc_apply_exp_M": [vp, vp, vp, C.c_int, C.c_int],
        "ephmc_construct_M_hat": [vp, vp, vp],
        "ephmc_logdet_inv": [vp, vp, dp, C.POINTER(C.c_int)],
        "ephmc_force": [vp, vp, vp, C.c_double, vp],
        "ephmc_sample_free_bosons": [vp, vp, vp],
        "ephmc_sample_fourier_momenta": [vp, vp, vp],
        "ephmc_energy_momenta": [vp, vp, dp],
        "ephmc_evolve_bosons": [vp, vp, vp, C.c_double],
        "ephmc_hamilton": [vp, vp, C.c_double, C.c_double, dp],
        "ephmc_leap_frog": [vp, vp, vp],
        "ephmc_update_x": [vp, vp, vp, C.c_double, C.c_double],
        "ephmc_chain_init": [vp, dp],
        "ephmc_trajectory": [vp, dp, C.c_double, C.POINTER(Record)],
        "ephmc_last_record": [vp, C.POINTER(Record)],
        "ephmc_chain_get_x": [vp, dp],
        "ephmc_chain_set_x": [vp, dp],
        "ephmc_chain_get_minv": [vp, dp],
        "ephmc_chain_get_logdet": [vp, dp],
        "ephmc_chain_buffers": [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)],
        "ephmc_cc_correlator": [vp, vp, vp, C.c_double, dp],
        "ephmc_tl_hist": [vp, dp, C.c_uint, dp, C.c_double, C.c_uint, dp, dp],
        "ephmc_gj_fallbacks": [vp, C.POINTER(C.c_ulong)],
        "ephmc_timing": [vp, C.c_int],
        "ephmc_timing_get": [vp, C.c_char_p, dp, C.POINTER(C.c_ulong)],
        "ephmc_autotune": [vp, C.POINTER(TuneResult)],
        "ephmc_last_pivots": [vp, C.POINTER(C.c_int)],
        "ephmc_md_force": [vp, vp, vp],
        "ephmc_schur": [vp, vp, vp, vp, dp, dp],
        "ephmc_stable_logdet_inv": [vp, vp, C.c_int, C.c_uint, vp, dp, C.POINTER(C.c_int)],
        "ephmc_cc_correlator_static": [vp, vp, vp, C.c_double, vp],
        "ephmc_apply_mu_K": [vp, vp, vp, vp, C.c_uint],
        "ephmc_apply_pairs": [vp, vp, vp, C.c_uint, C.c_int, dp, dp],
        "ephmc_apply_exp_cols": [vp, vp, vp, C.c_uint, C.c_int, C.c_int],
        "ephmc_refine_QM1P": [vp, vp, vp],
        "ephmc_construct_QM1P": [vp, vp, vp, vp, C.c_int],
        "ephmc_construct_all_QM1P": [vp, vp, vp, vp],
        "ephmc_boson_force": [vp, vp, vp, C.c_int],
        "ephmc_evolve_bosons_approx_step": [vp, vp, vp, C.c_double],
        "ephmc_batch_create": [C.POINTER(vp), C.c_uint, C.POINTER(vp)],
        "ephmc_batch_trajectory": [vp, C.POINTER(dp), dp],
        "ephmc_batch_sync": [vp, C.POINTER(Record)],
        "ephmc_batch_info": [vp, C.POINTER(C.c_uint), C.POINTER(C.c_uint)],
        "ephmc_batch_destroy": [vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.ephmc_last_error.restype = C.c_char_p
    lib.ephmc_last_error.argtypes = []
    return lib


LIB = _load()


def check(rc):
    if rc != 0:
        raise EphmcError("ephmc error %d (%s): %s" % (rc, ERRORS.get(rc, "?"), LIB.ephmc_last_error().decode()))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class DeviceArray:
    """A float64 buffer in HBM allocated by the library."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, int(n)
        p = C.c_void_p()
        check(LIB.ephmc_malloc(max(self.n, 1) * 8, C.byref(p)))
        self.ptr = p

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.size == self.n
        check(LIB.ephmc_memcpy_h2d(self.ctx.h, self.ptr, host.ctypes.data_as(C.c_void_p), self.n * 8))
        return self

    def download(self):
        out = np.empty(self.n)
        check(LIB.ephmc_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.n * 8))
        return out

    def free(self):
        if self.ptr:
            LIB.ephmc_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def scale_params(beta, J, lam, mu0, omega0, nt, static_disorder=False):
    """Lattice-unit couplings of organic_main.c:99-116 (lambda relative to J)."""
    x_avg = 1.0 if static_disorder else math.sqrt(.5 / math.tanh(.5 * omega0 / 0.025))
    scale = -beta if static_disorder else beta / nt
    mu = mu0 * scale
    t = [j * scale for j in J]
    kappa = [l * tk / x_avg for l, tk in zip(lam, t)]
    mu -= sum(2 * abs(tk) for tk in t)
    return t, kappa, mu, omega0 * scale


class Context:
    """One chain on one GPU (ephmc_ctx)."""

    def __init__(self, l1, l2, nt, t, kappa, mu, omega, flavors=2, steps=10, traj_length=0.0,
                 precision=FP64, checkpoint=0, taylor_order=4, device=0, **flags):
        cfg = Config()
        cfg.l1, cfg.l2, cfg.nt = l1, l2, nt
        cfg.t[:] = list(t)
        cfg.kappa[:] = list(kappa)
        cfg.mu, cfg.omega, cfg.flavors = mu, omega, flavors
        cfg.steps, cfg.traj_length = steps, traj_length
        cfg.taylor_order, cfg.precision, cfg.checkpoint = taylor_order, precision, checkpoint
        for k, v in flags.items():
            setattr(cfg, k, v)
        cfg.device = device
        h = C.c_void_p()
        check(LIB.ephmc_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.ns, self.nt = l1 * l2, nt
        self.dim = nt * 3 * self.ns
        self.mat = self.ns * self.ns

    def close(self):
        if self.h:
            LIB.ephmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, n, host=None):
        a = DeviceArray(self, n)
        if host is not None:
            a.upload(host)
        return a

    def sync(self):
        check(LIB.ephmc_synchronize(self.h))

    def lattice(self):
        out = np.zeros(12 * self.ns, np.uint32)
        check(LIB.ephmc_lattice(self.h, out.ctypes.data_as(C.POINTER(C.c_uint))))
        return out

    # -- host-array conveniences (each uploads, runs, downloads) ------------------------
    def M_hat(self, x):
        dx, dm = self.array(self.dim, x), self.array(self.mat)
        check(LIB.ephmc_construct_M_hat(self.h, dx.ptr, dm.ptr))
        return dm.download()

    def logdet_inv(self, M):
        dm = self.array(self.mat, M)
        ld, sg = C.c_double(), C.c_int()
        check(LIB.ephmc_logdet_inv(self.h, dm.ptr, C.byref(ld), C.byref(sg)))
        return ld.value, sg.value, dm.download()

    def force(self, x, Minv, logdet=0.0):
        dx, dp_ = self.array(self.dim, x), self.array(self.dim)
        dm = self.array(self.mat, Minv) if Minv is not None else None
        check(LIB.ephmc_force(self.h, dx.ptr, dm.ptr if dm is not None else None, logdet, dp_.ptr))
        return dp_.download()

    def stable_logdet_inv(self, x, precision, k):
        """(log det, sign, inverse) of 1 + B..B from the product stabilised by an FP64 QR every k
        slices, slices in `precision`."""
        dx, dm = self.array(self.dim, x), self.array(self.mat)
        ld, sg = C.c_double(), C.c_int()
        check(LIB.ephmc_stable_logdet_inv(self.h, dx.ptr, precision, k, dm.ptr, C.byref(ld), C.byref(sg)))
        return ld.value, sg.value, dm.download()

    def schur(self, B):
        """Real Schur form of the N x N matrix B (column-major flat): (Z, T, wr, wi)."""
        db, dz, dt = self.array(self.mat, B), self.array(self.mat), self.array(self.mat)
        wr, wi = np.zeros(self.ns), np.zeros(self.ns)
        check(LIB.ephmc_schur(self.h, db.ptr, dz.ptr, dt.ptr, _dp(wr), _dp(wi)))
        return dz.download(), dt.download(), wr, wi

    def md_force(self, x):
        """The force of one molecular-dynamics step at x (the context's chain precision builds
        M_hat; FP64 inverse; adjoint sweep), as ephmc_leap_frog uses it."""
        dx, dp_ = self.array(self.dim, x), self.array(self.dim)
        check(LIB.ephmc_md_force(self.h, dx.ptr, dp_.ptr))
        return dp_.download()

    def apply_exp_M(self, x_tau, M, revert, stop):
        dx, dm = self.array(3 * self.ns, x_tau), self.array(self.mat, M)
        check(LIB.ephmc_apply_exp_M(self.h, dx.ptr, dm.ptr, revert, stop))
        return dm.download()

    def sample_free_bosons(self, normals):
        dn, dx = self.array(self.dim, normals), self.array(self.dim)
        check(LIB.ephmc_sample_free_bosons(self.h, dn.ptr, dx.ptr))
        return dx.download()

    def sample_fourier_momenta(self, normals):
        dn, dp_ = self.array(self.dim, normals), self.array(self.dim)
        check(LIB.ephmc_sample_fourier_momenta(self.h, dn.ptr, dp_.ptr))
        return dp_.download()

    def energy_momenta(self, p):
        dp_ = self.array(self.dim, p)
        e = C.c_double()
        check(LIB.ephmc_energy_momenta(self.h, dp_.ptr, C.byref(e)))
        return e.value

    def evolve_bosons(self, x, p, h):
        dx, dp_ = self.array(self.dim, x), self.array(self.dim, p)
        check(LIB.ephmc_evolve_bosons(self.h, dx.ptr, dp_.ptr, h))
        return dx.download(), dp_.download()

    def hamilton(self, x, logdet, energyP):
        dx = self.array(self.dim, x)
        H = C.c_double()
        check(LIB.ephmc_hamilton(self.h, dx.ptr, logdet, energyP, C.byref(H)))
        return H.value

    def leap_frog(self, x, p):
        dx, dp_ = self.array(self.dim, x), self.array(self.dim, p)
        check(LIB.ephmc_leap_frog(self.h, dx.ptr, dp_.ptr))
        return dx.download(), dp_.download()

    def update_x(self, x, p, h, half):
        dx, dp_ = self.array(self.dim, x), self.array(self.dim, p)
        check(LIB.ephmc_update_x(self.h, dx.ptr, dp_.ptr, h, half))
        return dx.download(), dp_.download()

    # -- chain ------------------------------------------------------------------------
    def _host_field(self, a, what):
        """Contiguous float64 copy of a host field of dim values (the C side reads dim of them)."""
        a = np.ascontiguousarray(a, np.float64).ravel()
        if a.size != self.dim:
            raise ValueError("%s: %d values, the context needs dim = %d" % (what, a.size, self.dim))
        return a

    def chain_init(self, normals):
        normals = self._host_field(normals, "normals")
        check(LIB.ephmc_chain_init(self.h, _dp(normals)))

    def trajectory(self, normals, uniform, wait=True):
        normals = self._host_field(normals, "normals")
        rec = Record()
        check(LIB.ephmc_trajectory(self.h, _dp(normals), uniform, C.byref(rec) if wait else None))
        return rec if wait else None

    def last_record(self):
        rec = Record()
        check(LIB.ephmc_last_record(self.h, C.byref(rec)))
        return rec

    def chain_x(self):
        out = np.empty(self.dim)
        check(LIB.ephmc_chain_get_x(self.h, _dp(out)))
        return out

    def chain_set_x(self, x):
        x = self._host_field(x, "x")
        check(LIB.ephmc_chain_set_x(self.h, _dp(x)))

    def chain_minv(self):
        out = np.empty(self.mat)
        check(LIB.ephmc_chain_get_minv(self.h, _dp(out)))
        return out

    def chain_logdet(self):
        v = C.c_double()
        check(LIB.ephmc_chain_get_logdet(self.h, C.byref(v)))
        return v.value

    # -- measurement ---------------------------------------------------------------------
    def cc_correlator(self, x, Minv, logdet=0.0):
        """g_ab(tau), shape (nt, 3, 3), of field x with M_inv = (1 + B..B)^{-1}; x and Minv are
        host arrays or DeviceArrays already in HBM (the state organic_main hands over)."""
        dx = x if isinstance(x, DeviceArray) else self.array(self.dim, x)
        dm = Minv if isinstance(Minv, DeviceArray) else self.array(self.mat, Minv)
        out = np.zeros(self.nt * 9)
        check(LIB.ephmc_cc_correlator(self.h, dx.ptr, dm.ptr, logdet, _dp(out)))
        return out.reshape(self.nt, 3, 3)

    # -- transient localisation ------------------------------------------------------------
    def tl_hist(self, normals, contr_mat, width, nbins):
        """Disorder samples (rows of `normals`, 3N each): bins accumulated over the samples (not
        divided by their number) and sum_i exp(-E_i)/N per sample."""
        normals = np.ascontiguousarray(normals, np.float64).reshape(-1, 3 * self.ns)
        cm = np.ascontiguousarray(contr_mat, np.float64).reshape(6)
        greens = np.zeros(nbins)
        res = np.zeros(normals.shape[0])
        check(LIB.ephmc_tl_hist(self.h, _dp(normals), normals.shape[0], _dp(cm), width, nbins, _dp(greens), _dp(res)))
        return greens, res

    def last_pivots(self):
        """LAPACK-style 1-based row interchanges of the last log-det / inverse."""
        out = np.zeros(self.ns, np.int32)
        check(LIB.ephmc_last_pivots(self.h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def logdet_inv_status(self, M):
        """(status code, log det, sign, inverse): a singular matrix gives EPHMC_ERR_NUMERIC (-5)
        with log det = -inf instead of raising."""
        dm = self.array(self.mat, M)
        ld, sg = C.c_double(), C.c_int()
        rc = LIB.ephmc_logdet_inv(self.h, dm.ptr, C.byref(ld), C.byref(sg))
        if rc not in (0, -5):
            check(rc)
        return rc, ld.value, sg.value, dm.download()

    def construct_all_QM1P(self, x, Minv):
        dx, dq = self.array(self.dim, x), self.array(self.nt * self.mat)
        dm = self.array(self.mat, Minv) if Minv is not None else None
        check(LIB.ephmc_construct_all_QM1P(self.h, dx.ptr, dm.ptr if dm is not None else None, dq.ptr))
        return dq.download()

    def gj_fallbacks(self):
        v = C.c_ulong()
        check(LIB.ephmc_gj_fallbacks(self.h, C.byref(v)))
        return v.value

    def autotune(self):
        """Launch-shape tuning of this context (ephmc_autotune): returns the chosen columns per
        workgroup (and strip width) of the forward chain and the force sweep, the fp64 chain of
        the Metropolis log det, and their per-call times (ms)."""
        r = TuneResult()
        check(LIB.ephmc_autotune(self.h, C.byref(r)))
        return {"chain_cols": r.chain_cols, "chain_strip": r.chain_strip, "adjoint_cols": r.adjoint_cols,
                "chain_ms": r.chain_ms, "adjoint_ms": r.adjoint_ms, "adjoint_w0_early": r.adjoint_w0_early,
                "chain64_cols": r.chain64_cols, "chain64_strip": r.chain64_strip, "chain64_ms": r.chain64_ms}

    def timing(self, enable=True):
        check(LIB.ephmc_timing(self.h, int(enable)))

    def timing_get(self, group):
        ms, calls = C.c_double(), C.c_ulong()
        check(LIB.ephmc_timing_get(self.h, group.encode(), C.byref(ms), C.byref(calls)))
        return ms.value, calls.value


class Batch:
    """Many chains advanced together (ephmc_batch_*): every kernel of a trajectory is launched once
    over all chains of a group.  The contexts keep their state and can be used alone between batch
    trajectories (after collect()); close the batch before its contexts."""

    def __init__(self, contexts):
        self.ctxs = list(contexts)
        arr = (C.c_void_p * len(self.ctxs))(*[c.h.value for c in self.ctxs])
        h = C.c_void_p()
        check(LIB.ephmc_batch_create(arr, len(self.ctxs), C.byref(h)))
        self.h = h

    def trajectory(self, normals, uniforms, wait=True):
        """One trajectory of every chain; normals[i] and uniforms[i] as for Context.trajectory.
        With wait the records are returned, else collect() returns them later."""
        if len(normals) != len(self.ctxs) or len(uniforms) != len(self.ctxs):
            raise ValueError("one normals array and one uniform per chain")
        fields = [c._host_field(n, "normals") for c, n in zip(self.ctxs, normals)]
        ptrs = (C.POINTER(C.c_double) * len(fields))(*[_dp(f) for f in fields])
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        check(LIB.ephmc_batch_trajectory(self.h, ptrs, _dp(u)))
        return self.collect() if wait else None

    def collect(self):
        recs = (Record * len(self.ctxs))()
        check(LIB.ephmc_batch_sync(self.h, recs))
        return list(recs)

    def info(self):
        g, n = C.c_uint(), C.c_uint()
        check(LIB.ephmc_batch_info(self.h, C.byref(g), C.byref(n)))
        return {"groups": g.value, "launches": n.value}

    def close(self):
        if self.h:
            LIB.ephmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
