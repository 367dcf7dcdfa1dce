"""Writes the oracle's seeded FP64 HMC chains as golden fixtures for the GPU parity tests that are
too slow to run the oracle beside the GPU: the 16x16, Nt=64 chain (Nmd=10, traj=pi/2, 2 flavours,
seed 2, the default couplings), 20 trajectories.  Per trajectory: acceptance, dH and the log det
after the Metropolis decision, plus the log det of the initial field.

Run from the repository root:  python tests/golden/make_chain_fixtures.py
(needs oracle/liboracle.so; the oracle is the CPU restatement of the reference algorithm)."""
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _oracle import HMC, OracleModel, oracle  # noqa: E402

CHAINS = {"c2_fp64_16x16_nt64_seed2": dict(l=16, nt=64, steps=10, seed=2, ntraj=20)}


def run(l, nt, steps, seed, ntraj):
    om = OracleModel(l, l, nt)
    lib = oracle()
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), seed, steps, 0.0, None) == 0
    out = {"config": dict(l1=l, l2=l, nt=nt, steps=steps, seed=seed, beta=40.0, J=[0.10, -0.03, 0.02],
                          lam=[0.3, 0.3, 0.3], mu0=0.0, omega0=0.006, flavors=2, traj_length="pi/2"),
           "logdet0": h.logdet, "accepted": [], "dH": [], "logdet": []}
    dH = C.c_double()
    for it in range(ntraj):
        t0 = time.time()
        acc = lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None)
        out["accepted"].append(int(acc))
        out["dH"].append(dH.value)
        out["logdet"].append(h.logdet)
        print("traj %d: acc %d dH %.6g logdet %.12g (%.1f s)" % (it, acc, dH.value, h.logdet, time.time() - t0),
              flush=True)
    lib.orc_hmc_free(C.byref(h))
    return out


if __name__ == "__main__":
    for name, cfg in CHAINS.items():
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(run(**cfg), fh, indent=1)
