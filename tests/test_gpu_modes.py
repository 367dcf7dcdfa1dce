"""The reference's flag modes through the HIP engine, against the oracle on the same inputs:
CONSTANT_FORCE (organic_hmc.c:92-98), NO_FOURIER_ACC and CLASSICAL_FOURIER_ACC
(organic_hmc.c:56-65,131,136-165, organic_bosons.c:76-100,152-194), SINGLE_PARTICLE
(organic_dense_matrix.c:260-264,550-551, organic_hmc.c:121) and ZERO_FILLING
(organic_hmc.c:246-254).  FP64 path: 1e-10 relative; chains: identical accept/reject."""
import ctypes as C

import numpy as np
import pytest

from _gpu import gpu_context, rel, say
from _oracle import HMC, MT, OracleModel, oracle, ptr

pytestmark = pytest.mark.gpu

CFA = dict(classical_fourier_acc=1, cfa_steps=4, cfa_mass=0.02)
MODES = {
    "constant_force": dict(constant_force=1),
    "no_fourier_acc": dict(no_fourier_acc=1),
    "classical_fourier_acc": CFA,
    "single_particle": dict(single_particle=1),
    "single_particle+constant_force": dict(single_particle=1, constant_force=1),
}


def model(l1, l2, nt, flags):
    return OracleModel(l1, l2, nt, beta=nt * 40 / 32, flags=flags)


def field(om, seed, scale=0.5):
    return np.random.default_rng(seed).standard_normal(om.dim) * scale


@pytest.mark.parametrize("name", list(MODES))
@pytest.mark.parametrize("l1,l2,nt", [(4, 4, 8), (6, 6, 16), (8, 6, 12)])
def test_force_modes(name, l1, l2, nt):
    flags = MODES[name]
    say("force mode", name, l1, l2, nt)
    om = model(l1, l2, nt, flags)
    ctx = gpu_context(om, checkpoint=4, **flags)
    x = field(om, nt + l1)
    M = om.M_hat(x)                           # B..B without the identity in SINGLE_PARTICLE
    if flags.get("single_particle"):
        ld = float(np.log(np.trace(M.reshape(om.ns, om.ns))))
        Mi = M
        Fg = ctx.force(x, None, ld)
    else:
        ld, _, Mi = om.logdet_inv(M)
        Fg = ctx.force(x, Mi, ld)
    Fo = om.force(x, Mi, ld)
    assert rel(Fg, Fo) < 1e-10


@pytest.mark.parametrize("flags", [dict(single_particle=1), dict(zero_filling=1)])
def test_M_hat_and_logdet_without_identity(flags):
    om = model(5, 4, 8, flags)
    ctx = gpu_context(om, **flags)
    x = field(om, 2)
    Mg, Mo = ctx.M_hat(x), om.M_hat(x)
    assert rel(Mg, Mo) < 1e-13
    if flags.get("single_particle"):
        ldg, sg, Mafter = ctx.logdet_inv(Mg)
        assert abs(ldg - np.log(np.trace(Mo.reshape(om.ns, om.ns)))) < 1e-12 * abs(ldg)
        np.testing.assert_array_equal(Mafter, Mg)


@pytest.mark.parametrize("flags", [dict(no_fourier_acc=1), CFA])
@pytest.mark.parametrize("half", [0.5, 1.0, -0.5])
def test_update_x_modes(flags, half):
    om = model(4, 4, 16, flags)
    ctx = gpu_context(om, **flags)
    x, p = field(om, 3), field(om, 4)
    xg, pg = ctx.update_x(x, p, 0.15, half)
    xo, po, pdot = x.copy(), p.copy(), np.zeros(om.dim)
    oracle().orc_update_x(om.ref, ptr(xo), ptr(po), ptr(pdot), 0.15, half)
    assert rel(xg, xo) < 1e-12
    assert rel(pg, po) < 1e-12


@pytest.mark.parametrize("name", ["constant_force", "no_fourier_acc", "classical_fourier_acc", "single_particle"])
def test_chain_accept_sequence_modes(name):
    """40 trajectories at 4x4, Nt=16, Nmd=6 with the oracle's RNG stream: identical decisions and
    log det / dH to 1e-9."""
    flags = MODES[name]
    # without Fourier acceleration the stiffest boson mode (frequency ~ 2/omega) needs a short step
    traj = 0.2 if name == "no_fourier_acc" else 0.0
    say("chain mode", name)
    om = model(4, 4, 16, flags)
    lib = oracle()
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), 7, 6, traj, None) == 0
    ctx = gpu_context(om, steps=6, traj_length=traj, **flags)
    rng = MT()
    lib.orc_mt_seed(C.byref(rng), 7)
    normals = np.zeros(om.dim)
    lib.orc_random_vector(C.byref(rng), ptr(normals), om.dim)
    ctx.chain_init(normals)
    assert ctx.chain_logdet() == pytest.approx(h.logdet, rel=1e-11)
    dH = C.c_double()
    accepted = 0
    for it in range(40):
        lib.orc_random_vector(C.byref(rng), ptr(normals), om.dim)
        u = lib.orc_mt_real2(C.byref(rng))
        acc_o = lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None)
        rec = ctx.trajectory(normals, u)
        assert rec.accepted == acc_o, "trajectory %d: gpu %d oracle %d (dH %g / %g)" % (it, rec.accepted, acc_o, rec.dH,
                                                                                      dH.value)
        assert np.isfinite(rec.dH) and abs(rec.dH - dH.value) < 1e-8
        assert abs(rec.logdet - h.logdet) <= 1e-9 * abs(h.logdet)
        accepted += rec.accepted
    assert accepted >= 10
    lib.orc_hmc_free(C.byref(h))


def test_zero_filling_chain():
    """ZERO_FILLING: a free-boson field per trajectory, energy = S_B(x), no Metropolis draw."""
    flags = dict(zero_filling=1)
    om = model(4, 4, 16, flags)
    lib = oracle()
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), 11, 6, 0.0, None) == 0
    ctx = gpu_context(om, steps=6, **flags)
    rng = MT()
    lib.orc_mt_seed(C.byref(rng), 11)
    normals = np.zeros(om.dim)
    lib.orc_random_vector(C.byref(rng), ptr(normals), om.dim)
    ctx.chain_init(normals)
    dH = C.c_double()
    for it in range(10):
        lib.orc_random_vector(C.byref(rng), ptr(normals), om.dim)
        res = np.zeros(12)
        lib.orc_trajectory(C.byref(h), ptr(res), None, C.byref(dH), None, None)
        rec = ctx.trajectory(normals, 0.0)
        assert rec.accepted == 1 and rec.boltzmann == 1.0
        assert rec.energy == pytest.approx(res[2], rel=1e-12)
        xg = ctx.chain_x()
        xo = np.ctypeslib.as_array(h.x, shape=(om.dim,))
        assert rel(xg, xo) < 1e-13
        Mg = ctx.chain_minv()
        Mo = np.ctypeslib.as_array(h.M_hat, shape=(om.ns * om.ns,))
        assert rel(Mg, Mo) < 1e-12
    lib.orc_hmc_free(C.byref(h))


@pytest.mark.parametrize("flags", [dict(single_particle=1), dict(zero_filling=1)])
def test_correlator_modes(flags):
    om = model(4, 5, 8, flags)
    ctx = gpu_context(om, **flags)
    x = field(om, 9)
    M = om.M_hat(x)
    ld = float(np.log(np.trace(M.reshape(om.ns, om.ns)))) if flags.get("single_particle") else 0.0
    g = ctx.cc_correlator(x, M, ld)
    n, nt = om.ns, om.nt
    go = np.zeros(nt * 9)
    oracle().orc_cc_correlator(om.ref, ptr(x), ptr(M), ptr(np.zeros(nt * n * n)), ptr(np.zeros(n * n)),
                               ptr(np.zeros(2 * n * n)), ptr(go), ld)
    assert rel(g.ravel(), go) < 1e-10
