"""The measurement path and the C host layer on the GPU: the current-current correlator against
the oracle, the retained organic_* functions of libephmc_host.so (reference call signatures)
against the oracle, and organic_main end to end (input file -> results.csv / corr.csv) against
the oracle's chain with the same seed.  FP64 path: 1e-10 relative, accept/reject identical."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from _gpu import gpu_context, rel, say
from _oracle import HMC, OracleModel, dp, ip, oracle, ptr, up
from test_host_layer import MAIN, host

pytestmark = pytest.mark.gpu


def minv_of(om, x):
    ld, sign, Minv = om.logdet_inv(om.M_hat(x))
    return ld, Minv


def oracle_greens(om, x, Minv, ld):
    n, nt = om.ns, om.nt
    g = np.zeros(nt * 9)
    Rt0 = np.zeros(nt * n * n)
    R0t = np.zeros(n * n)
    w = np.zeros(2 * n * n)
    oracle().orc_cc_correlator(om.ref, ptr(x), ptr(Minv), ptr(Rt0), ptr(R0t), ptr(w), ptr(g), ld)
    return g.reshape(nt, 3, 3)


@pytest.mark.parametrize("l1,l2,nt", [(1, 4, 4), (2, 3, 8), (4, 4, 8), (6, 6, 16), (5, 7, 12)])
def test_cc_correlator(l1, l2, nt):
    say("cc_correlator", l1, l2, nt)
    om = OracleModel(l1, l2, nt, beta=nt * 40 / 32)
    x = np.random.default_rng(nt + l1).standard_normal(om.dim)
    ld, Minv = minv_of(om, x)
    ctx = gpu_context(om)
    g = ctx.cc_correlator(x, Minv, ld)
    go = oracle_greens(om, x, Minv, ld)
    assert rel(g, go) < 1e-10
    np.testing.assert_allclose(g, go, rtol=1e-9, atol=1e-12 * np.abs(go).max())
    # the same call on inputs already in HBM (as the host's trajectory loop passes them), twice in
    # a row: the transfer matrices are reused and the kernels are deterministic
    dx, dm = ctx.array(om.dim, x), ctx.array(om.ns * om.ns, Minv)
    g1, g2 = ctx.cc_correlator(dx, dm, ld), ctx.cc_correlator(dx, dm, ld)
    np.testing.assert_array_equal(g1, g)
    np.testing.assert_array_equal(g2, g)


def test_retained_functions_through_host_library():
    """construct_M_hat, logDetM_and_inv_M_hat, force, cc_correlator, apply_exp_mu_K_to_M and
    hamilton with the reference's signatures (organic_dense_matrix.h, organic_hmc.h)."""
    h = host()
    h.construct_M_hat.argtypes = [dp, dp, dp, dp, dp, C.c_double, up, C.c_uint, C.c_uint, C.c_uint, C.c_void_p]
    h.logDetM_and_inv_M_hat.restype = C.c_double
    h.logDetM_and_inv_M_hat.argtypes = [dp, dp, ip, C.c_uint, C.c_uint, C.c_void_p]
    h.force.argtypes = [dp, dp, dp, dp, dp, ip, dp, dp, C.c_double, C.c_double, C.c_uint, up, C.c_uint, C.c_uint,
                        C.c_uint, C.c_double, C.c_void_p]
    h.cc_correlator.argtypes = [dp, dp, dp, dp, dp, dp, dp, C.c_double, up, C.c_uint, C.c_uint, C.c_uint, dp,
                                C.c_double, C.c_void_p]
    h.apply_exp_mu_K_to_M.argtypes = [dp, dp, dp, dp, dp, C.c_double, up, C.c_uint, C.c_uint, C.c_int, C.c_int]
    h.hamilton.restype = C.c_double
    h.hamilton.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
    om = OracleModel(4, 6, 16, beta=20.0)
    n, nt = om.ns, om.nt
    nnt = np.ascontiguousarray(om.nnt, np.uint32)
    nntp = nnt.ctypes.data_as(up)
    x = np.random.default_rng(11).standard_normal(om.dim)
    w = np.zeros(8 * n * n)
    say("host construct_M_hat")
    M = np.zeros(n * n)
    h.construct_M_hat(ptr(x), ptr(M), ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nntp, n, 6, nt, None)
    Mo = om.M_hat(x)
    assert rel(M, Mo) < 1e-13
    say("host logDetM_and_inv_M_hat")
    perm = np.zeros(n, np.int32)
    ld = h.logDetM_and_inv_M_hat(ptr(M), ptr(w), ptr(perm, ip), n, 2, None)
    ldo, _, Minv = om.logdet_inv(Mo)
    assert abs(ld - ldo) < 1e-10 * abs(ldo)
    assert rel(M, Minv) < 1e-10
    say("host force")
    pdot = np.zeros(om.dim)
    h.force(ptr(x), ptr(pdot), ptr(M), None, ptr(w), ptr(perm, ip), ptr(om.t), ptr(om.kappa), om.mu, om.omega, 2,
            nntp, n, 6, nt, ld, None)
    assert rel(pdot, om.force(x, Minv, ldo)) < 1e-10
    say("host cc_correlator")
    g = np.zeros(nt * 9)
    h.cc_correlator(ptr(x), ptr(M), None, None, ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nntp, n, 6, nt, ptr(g), ld,
                    None)
    assert rel(g, oracle_greens(om, x, Minv, ldo).ravel()) < 1e-10
    say("host apply_exp_mu_K_to_M")
    A = np.random.default_rng(12).standard_normal(n * n)
    B = A.copy()
    h.apply_exp_mu_K_to_M(ptr(x), ptr(A), ptr(w), ptr(om.t), ptr(om.kappa), om.mu, nntp, n, 6, 0, 2)
    oracle().orc_apply_exp_mu_K_to_M(om.ref, ptr(x), ptr(B), ptr(w), 0, 2)
    assert rel(A, B) < 1e-13
    say("host hamilton")
    H = h.hamilton(ptr(x), None, ld, 3.5, om.omega, 2, n, 6, nt)
    Ho = oracle().orc_hamilton(om.ref, ptr(x), ldo, 3.5)
    assert abs(H - Ho) < 1e-10 * abs(Ho)


INPUT = """{l1} {l2}
{nt}
{beta!r}
0.1 -0.03 0.02
0.3 0.3 0.3
0
0.006
{steps} 0
{therm}
{meas} {freq}
1 0
0.5 0.8660254037844386
2
SEED {seed}
"""


def test_organic_main_end_to_end():
    """organic_main input.txt results.csv corr.csv (FP64) against the oracle's chain seeded the
    same way: identical accept/reject column, every results column and every corr.csv line
    within 1e-9 relative."""
    l1, l2, nt, steps, therm, meas, freq, seed = 6, 6, 16, 10, 3, 8, 2, 4242
    beta = nt * 40 / 32
    om = OracleModel(l1, l2, nt, beta=beta)
    a, b = np.array([1.0, 0.0]), np.array([0.5, 0.8660254037844386])
    cm = np.ctypeslib.as_array(oracle().orc_construct_contraction_mat(ptr(a), ptr(b), 3), (6,)).copy()
    with tempfile.TemporaryDirectory() as d:
        inp, res, corr = (os.path.join(d, f) for f in ("input.txt", "results.csv", "corr.csv"))
        open(inp, "w").write(INPUT.format(l1=l1, l2=l2, nt=nt, beta=beta, steps=steps, therm=therm, meas=meas,
                                          freq=freq, seed=seed))
        say("organic_main", l1, l2, nt)
        env = dict(os.environ)
        env.pop("EPHMC_PRECISION", None)
        out = subprocess.run([MAIN, inp, res, corr], capture_output=True, text=True, timeout=300, env=env)
        say("organic_main rc", out.returncode, out.stdout.strip()[-200:], out.stderr.strip()[-200:])
        assert out.returncode == 0 and "ERROR" not in out.stdout
        rows = [[float(v) for v in line.split("\t") if v] for line in open(res).read().splitlines()]
        corr_rows = [line.split("\t") for line in open(corr).read().splitlines()]
    assert len(rows) == meas // freq and all(len(r) == 12 for r in rows)
    assert len(corr_rows) == nt * (meas // freq)

    lib = oracle()
    contr = np.array(cm)
    hm = HMC()
    assert lib.orc_hmc_init(C.byref(hm), C.byref(om.m), seed, steps, 0.0, ptr(contr)) == 0
    dH = C.c_double()
    expect, expect_corr = [], []
    for i in range(therm):
        lib.orc_trajectory(C.byref(hm), None, None, C.byref(dH), None, None)
    for i in range(meas):
        if i % freq:
            lib.orc_trajectory(C.byref(hm), None, None, C.byref(dH), None, None)
            continue
        r, g = np.zeros(12), np.zeros(nt * 9)
        lib.orc_trajectory(C.byref(hm), ptr(r), ptr(g), C.byref(dH), None, None)
        expect.append(r)
        expect_corr += [lib.orc_contract_greens(ptr(g), ptr(contr), 6, tau) * nt * nt for tau in range(nt)]
    lib.orc_hmc_free(C.byref(hm))
    got = np.array(rows)
    exp = np.array(expect)
    say("accept columns", got[:, 6], exp[:, 6])
    np.testing.assert_array_equal(got[:, 6], exp[:, 6])
    np.testing.assert_allclose(got, exp, rtol=1e-9, atol=1e-12)
    assert [int(c[0]) for c in corr_rows] == list(range(nt)) * (meas // freq)
    np.testing.assert_allclose([float(c[1]) for c in corr_rows], expect_corr, rtol=1e-9, atol=1e-14)


TL_INPUT = """{l1} {l2}
{nbins}
40
0.1 -0.03 0.02
0.3 0.3 0.3
0
0.006
1 0
0
{meas} 1
1 0
0.5 0.8660254037844386
1
STATIC_DISORDER
SEED {seed}
"""


def test_organic_main_transient_localisation():
    """organic_main with STATIC_DISORDER (run_trans_loc, organic_trans_loc.c:85-114): per-sample
    results.csv lines and the greens.bin histogram against the oracle with the same RNG stream."""
    l1, l2, nbins, meas, seed = 6, 6, 40, 12, 99
    om = OracleModel(l1, l2, nbins, beta=40.0, static_disorder=True)
    a, b = np.array([1.0, 0.0]), np.array([0.5, 0.8660254037844386])
    cm = np.ctypeslib.as_array(oracle().orc_construct_contraction_mat(ptr(a), ptr(b), 3), (6,)).copy()
    with tempfile.TemporaryDirectory() as d:
        inp, res, gbin = (os.path.join(d, f) for f in ("input.txt", "results.csv", "greens.bin"))
        open(inp, "w").write(TL_INPUT.format(l1=l1, l2=l2, nbins=nbins, meas=meas, seed=seed))
        say("organic_main STATIC_DISORDER", l1, l2, nbins)
        out = subprocess.run([MAIN, inp, res, gbin], capture_output=True, text=True, timeout=300)
        say("organic_main rc", out.returncode, out.stdout.strip()[-200:], out.stderr.strip()[-200:])
        assert out.returncode == 0 and "ERROR" not in out.stdout
        z = np.array([float(line.split("\t")[0]) for line in open(res).read().splitlines()])
        raw = open(gbin, "rb").read()
    n_read = int(np.frombuffer(raw[:4], np.uint32)[0])
    width_read = float(np.frombuffer(raw[4:12], np.float64)[0])
    g = np.frombuffer(raw[12:], np.float64)
    assert n_read == nbins and g.size == nbins
    lib = oracle()
    width = lib.orc_tl_width(om.ref, nbins)
    assert width_read == width
    from _oracle import MT
    rng = MT()
    lib.orc_mt_seed(C.byref(rng), seed)
    go = np.zeros(nbins)
    zo = []
    x = np.zeros(3 * om.ns)
    for _ in range(meas):
        lib.orc_random_vector(C.byref(rng), ptr(x), 3 * om.ns)
        zo.append(lib.orc_hist_trans_loc(om.ref, ptr(x), ptr(cm), width, ptr(go), nbins))
    go /= meas
    np.testing.assert_allclose(z, zo, rtol=1e-12)
    assert rel(g, go) < 1e-9


def test_organic_main_restart():
    """STATE <file>: 3 + 3 trajectories with a restart in between give the same results.csv lines
    as 6 trajectories in one run (field and random stream are resumed exactly)."""
    base = INPUT.format(l1=4, l2=4, nt=16, beta=20.0, steps=6, therm=0, meas="{meas}", freq=1, seed=77)
    with tempfile.TemporaryDirectory() as d:
        state = os.path.join(d, "chain.state")
        runs = []
        for tag, meas, extra in (("full", 6, ""), ("a", 3, "STATE %s\n" % state), ("b", 3, "STATE %s\n" % state)):
            inp, res = os.path.join(d, tag + ".txt"), os.path.join(d, tag + ".csv")
            open(inp, "w").write(base.format(meas=meas) + extra)
            say("organic_main restart run", tag)
            out = subprocess.run([MAIN, inp, res], capture_output=True, text=True, timeout=300)
            assert out.returncode == 0, out.stderr
            runs.append(open(res).read().splitlines())
    full, a, b = runs
    assert len(full) == 6 and len(a) == 3 and len(b) == 3
    assert a + b == full
