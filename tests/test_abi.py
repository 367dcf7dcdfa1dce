"""The C-ABI libraries load on a machine without a GPU and export every symbol their headers
declare (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

from _gpu import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "ephmc.h")


def declared(header):
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(ephmc_\w+)\s*\(", text, re.M)))


def test_header_lists_entry_points():
    names = declared(HEADER)
    assert "ephmc_create" in names and "ephmc_force" in names and len(names) >= 25


def test_engine_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(PKG, "libephmc.so"))
    missing = [n for n in declared(HEADER) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_imports():
    import sys
    sys.path.insert(0, PKG)
    import ephmc
    assert ephmc.LIB is not None
    t, kappa, mu, omega = ephmc.scale_params(40.0, (0.10, -0.03, 0.02), (0.3, 0.3, 0.3), 0.0, 0.006, 128)
    assert omega == pytest.approx(0.006 * 40 / 128)
    assert mu == pytest.approx(-2 * 0.15 * 40 / 128)


# every function the reference's headers declare on the replaced path (organic_dense_matrix.h,
# organic_hmc.h, organic_bosons.h, organic_trans_loc.h; names only)
REFERENCE_HOST_API = [
    "mat_mul", "mat_mul_tr", "mat_mat_mul", "construct_id", "transpose", "transpose_rect", "trace",
    "apply_mu_K_to_v", "apply_mu_K_to_M", "set_taylor_pair_order", "apply_K_plus_1_v",
    "apply_2nd_order_mu_K_to_v", "apply_pair_expansion_mu_K_to_v", "apply_exp_mu_K_to_v", "apply_exp_mu_K_to_M",
    "construct_hamilton", "apply_current_to_v", "construct_M_hat", "construct_QM1P", "construct_all_QM1P",
    "refine_QM1P", "refine_QM1P_approx", "cc_correlator", "phonon_correlator", "logDetM_and_inv_M_hat",
    "check_force", "energy_momenta", "hamilton", "force", "update_x", "leap_frog", "trajectory", "run_hmc",
    "hot_start", "higher_Matsubaras_norm", "sample_free_bosons", "sample_fourier_momenta",
    "energy_fourier_momenta", "evolve_bosons", "evolve_bosons_approx_step", "boson_force",
    "hist_trans_loc", "run_trans_loc",
]


def test_host_library_exports_the_reference_interface():
    lib = C.CDLL(os.path.join(PKG, "libephmc_host.so"))
    missing = [n for n in REFERENCE_HOST_API if not hasattr(lib, n)]
    assert not missing, missing
    hdr = open(os.path.join(PKG, "host", "organic.h")).read()
    declared_host = re.findall(r"^\s*(?:unsigned\s*\*|double\s*\*|void|double|short|unsigned|int|const char \*|organic_plan)\s+(\w+)\s*\(", hdr, re.M)
    missing = [n for n in declared_host if not hasattr(lib, n)]
    assert not missing, missing
    assert set(REFERENCE_HOST_API) <= set(declared_host)


def test_create_refuses_bad_configurations_before_any_device_call():
    """ephmc_create validates its configuration before touching the device (so this runs without
    a GPU): an empty lattice, omega zero or NaN (the Fourier masses divide by it; it is negative
    under static disorder), a lattice beyond 2048 sites, also where l1 * l2 wraps around 32 bits,
    are refused and no context is returned."""
    import sys
    sys.path.insert(0, PKG)
    import ephmc as E
    base = dict(l1=4, l2=4, nt=8, omega=0.01, precision=E.FP64)
    cases = [(dict(l1=0), -1), (dict(nt=0), -1), (dict(omega=0.0), -1), (dict(omega=float("nan")), -1), (dict(l1=65, l2=32), -3), (dict(l1=65536, l2=65536), -3),
             (dict(l1=1 << 31, l2=2), -3), (dict(precision=7), -1)]
    for change, code in cases:
        cfg = E.Config()
        for k, v in {**base, **change}.items():
            setattr(cfg, k, v)
        h = C.c_void_p()
        assert E.LIB.ephmc_create(C.byref(cfg), C.byref(h)) == code, change
        assert not h.value, change

