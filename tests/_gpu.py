"""Shared helpers of the GPU tests: the HIP library binding and model set-up that mirrors the
oracle's, so GPU and oracle run the same couplings and fields."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "electron-phonon-hmc_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)


def ephmc():
    import ephmc as E   # raises when libephmc.so is missing: no fallback
    return E


def gpu_context(om, precision=0, steps=10, checkpoint=16, **kw):
    """HIP context with exactly the oracle model's lattice-unit couplings."""
    E = ephmc()
    return E.Context(om.l1, om.l2, om.nt, list(om.t), list(om.kappa), om.mu, om.omega, flavors=om.flavors,
                     steps=steps, precision=precision, checkpoint=checkpoint, **kw)


def say(*a):
    print("[gpu-test]", *a, flush=True)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
