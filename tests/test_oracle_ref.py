"""Pins the oracle's host-side pieces bit-exactly against the reference's own sources
(simulation/organic_lattice.c, mt19937-64.c, organic_random.c, organic_aux.c, organic_flags.c)
compiled into oracle/_ref/libref.so by oracle/Makefile."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from _oracle import MT, Flags, dp, libc, oracle, ptr, ref, up


@pytest.mark.parametrize("l1,l2", [(1, 1), (2, 2), (2, 3), (3, 2), (6, 6), (5, 7), (16, 16), (32, 32), (8, 12)])
def test_lattice_bit_exact(ref_available, l1, l2):
    r, o = ref(), oracle()
    ns_r, nn_r, ns_o, nn_o = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
    pr = r.construct_triangular(l1, l2, C.byref(ns_r), C.byref(nn_r))
    po = o.orc_construct_triangular(l1, l2, C.byref(ns_o), C.byref(nn_o))
    assert (ns_r.value, nn_r.value) == (ns_o.value, nn_o.value) == (l1 * l2, 6)
    a = np.ctypeslib.as_array(pr, shape=(12 * l1 * l2,)).copy()
    b = np.ctypeslib.as_array(po, shape=(12 * l1 * l2,)).copy()
    libc().free(C.cast(pr, C.c_void_p))
    libc().free(C.cast(po, C.c_void_p))
    np.testing.assert_array_equal(a, b)


def test_lattice_invariants():
    """Neighbour symmetry and one bond per link on a lattice big enough to have no wrap-around aliasing."""
    o = oracle()
    ns, nn = C.c_uint(), C.c_uint()
    p = o.orc_construct_triangular(7, 5, C.byref(ns), C.byref(nn))
    t = np.ctypeslib.as_array(p, shape=(12 * 35,)).copy()
    nb = t[:6 * 35].reshape(35, 6)
    link = t[6 * 35:].reshape(35, 6)
    for i in range(35):
        for d in range(0, 6, 2):
            assert nb[nb[i, d], d + 1] == i
            assert link[i, d] == 3 * i + d // 2
            assert link[nb[i, d], d + 1] == link[i, d]
    assert sorted(link[:, 0::2].ravel()) == list(range(3 * 35))


@pytest.mark.parametrize("seed", [1, 2, 3, 7, 1000, 1063, 2 ** 40 + 5])
def test_mt19937_64_bit_exact(ref_available, seed):
    r, o = ref(), oracle()
    st = MT()
    o.orc_mt_seed(C.byref(st), seed)
    r.init_genrand64(seed)
    a = [r.genrand64_int64() for _ in range(2000)]
    b = [o.orc_mt_u64(C.byref(st)) for _ in range(2000)]
    assert a == b
    for _ in range(100):
        assert r.genrand64_real2() == o.orc_mt_real2(C.byref(st))


def test_mt19937_64_known_answer():
    """C++ standard [rand.predef]: the 10000th output of mt19937_64 with the default seed 5489."""
    o = oracle()
    st = MT()
    o.orc_mt_seed(C.byref(st), 5489)
    for _ in range(9999):
        o.orc_mt_u64(C.byref(st))
    assert o.orc_mt_u64(C.byref(st)) == 9981545732273789042


@pytest.mark.parametrize("n", [1, 2, 7, 64, 3456])
def test_box_muller_bit_exact(ref_available, n):
    r, o = ref(), oracle()
    r.init_genrand64(11)
    st = MT()
    o.orc_mt_seed(C.byref(st), 11)
    a = np.zeros(n)
    b = np.zeros(n)
    r.random_vector(ptr(a), n)
    o.orc_random_vector(C.byref(st), ptr(b), n)
    np.testing.assert_array_equal(a, b)
    # the streams stay in lock step afterwards
    assert r.genrand64_real2() == o.orc_mt_real2(C.byref(st))


def test_contraction_bit_exact(ref_available):
    r, o = ref(), oracle()
    a = np.array([7.2, 0.0])
    b = np.array([0.0, 14.3])
    cr = r.construct_contraction_mat(ptr(a), ptr(b), 3)
    co = o.orc_construct_contraction_mat(ptr(a), ptr(b), 3)
    A = np.ctypeslib.as_array(cr, shape=(6,)).copy()
    B = np.ctypeslib.as_array(co, shape=(6,)).copy()
    np.testing.assert_array_equal(A, B)
    rng = np.random.default_rng(3)
    g = rng.standard_normal(9 * 5)
    for tau in range(5):
        assert r.contract_greens(ptr(g), cr, 6, tau) == o.orc_contract_greens(ptr(g), co, 6, tau)
        assert r.contract_greens(ptr(g), None, 6, tau) == o.orc_contract_greens(ptr(g), None, 6, tau)


def _fopen(path, mode):
    return libc().fopen(path.encode(), mode.encode())


def test_writers_bit_exact(ref_available):
    """results.csv / corr.csv text formats (organic_aux.c:51-70) against the oracle's writers."""
    r = ref()
    o = oracle()
    o.orc_fprint_results.argtypes = [C.c_void_p, dp, C.c_uint, C.c_uint]
    o.orc_fprint_corr.argtypes = [C.c_void_p, dp, dp, C.c_uint, C.c_uint]
    rng = np.random.default_rng(5)
    res = rng.standard_normal(24) * 10.0 ** rng.integers(-8, 8, 24)
    corr = rng.standard_normal(9 * 8)
    cm = np.array([1.0, 0.5, -0.5, 0.0, 0.9, 0.9])
    with tempfile.TemporaryDirectory() as d:
        outs = []
        for lib, fr, fc in ((r, r.fprint_results, r.fprint_corr), (o, o.orc_fprint_results, o.orc_fprint_corr)):
            p = os.path.join(d, "out%d" % len(outs))
            f = _fopen(p, "w")
            fr(f, ptr(res), 2, 12)
            fc(f, ptr(corr), ptr(cm), 6, 8)
            libc().fclose(f)
            outs.append(open(p).read())
        assert outs[0] == outs[1]
        assert outs[0].splitlines()[0].count("\t") == 12


def test_flags_parser(ref_available):
    r = ref()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "flags.txt")
        with open(p, "w") as f:
            f.write("ZERO_FILLING\nUNKNOWN_TOKEN\nCLASSICAL_FOURIER_ACC\n3 0.25\nNO_FOURIER_ACC\nSEED 17\n")
        fl = Flags()
        fh = _fopen(p, "r")
        r.set_flags(fh, C.byref(fl))
        libc().fclose(fh)
    assert (fl.zero_filling, fl.classical_fourier_acc, fl.no_fourier_acc, fl.schur_form_inv) == (1, 1, 1, 0)
    assert fl.cfa_steps == 3 and fl.cfa_mass == 0.25
