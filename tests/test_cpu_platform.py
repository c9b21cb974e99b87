"""The multi-threaded restatement of the reference CPU platform (oracle/cpu_platform.c), timed as the CPU baseline:
checked against the FP64 oracle (CPU only).

The pair forces are FP32 (CpuInteraxPairForceFvec.h) and must reproduce the oracle's interacting pair set exactly and
its forces and energies to FP32 accuracy; the three-body force is FP64 (CpuInteraxThreeBodyForce.cpp) and must agree
to rounding.
"""
import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads


def _systems():
    system, pos = workloads.config0()
    yield "config0", system, pos, np.array(system._box, dtype=float).reshape(3, 3)
    system, pos = workloads.water_cluster(40, seed=3)
    yield "cluster", system, pos, None
    system, pos = workloads.alchemical_box(0.3, seed=2)
    yield "alchemical", system, pos, np.array(system._box, dtype=float).reshape(3, 3)


@pytest.mark.parametrize("threads", [1, 4])
def test_pair_forces_match_oracle(threads):
    for name, system, pos, box in _systems():
        for lam in (1.0, 0.4):
            for f in system.getForces():
                if not isinstance(f, ih.InteraxPairForce):
                    continue
                if f.NAME == "InteraxDispersionForce":
                    f.setUseLongRangeCorrection(False)
                e0, f0 = oracle.pair_force(f, pos, box, lam, double_precision=False)
                e1, f1, pairs = oracle.cpu_pair_force(f, pos, box, lam, threads)
                assert pairs == oracle.pair_count(f, pos, box), (name, f.NAME)
                assert abs(e1-e0) <= 2e-6*max(abs(e0), 1.0), (name, f.NAME, lam, e1, e0)
                rms = np.linalg.norm(f1-f0)/np.linalg.norm(f0)
                assert rms < 2e-5, (name, f.NAME, lam, rms)


def test_pair_forces_independent_of_threads():
    system, pos = workloads.config0()
    box = np.array(system._box, dtype=float).reshape(3, 3)
    f = [x for x in system.getForces() if isinstance(x, ih.InteraxPairForce)][1]
    e1, f1, _ = oracle.cpu_pair_force(f, pos, box, 1.0, 1)
    e4, f4, _ = oracle.cpu_pair_force(f, pos, box, 1.0, 4)
    # per-thread sums in FP32 differ only by the order of additions
    assert abs(e4-e1) <= 1e-6*abs(e1)
    assert np.linalg.norm(f4-f1)/np.linalg.norm(f1) < 1e-5


def test_threebody_matches_oracle():
    system, pos = workloads.water_cluster(60, seed=5)
    force = [f for f in system.getForces() if isinstance(f, ih.InteraxThreeBodyForce)][0]
    e0, f0 = oracle.threebody_force(force, pos, None)
    e1, f1, triples = oracle.cpu_threebody_force(force, pos, None, 1.0, 4)
    assert triples == oracle.threebody_count(force, pos, None)
    assert abs(e1-e0) <= 1e-12*abs(e0)
    assert np.linalg.norm(f1-f0) <= 1e-12*np.linalg.norm(f0)


def test_threebody_periodic_matches_oracle():
    system, pos = workloads.atm_fluid(num_sites=2000, seed=4)
    box = np.array(system._box, dtype=float).reshape(3, 3)
    force = [f for f in system.getForces() if isinstance(f, ih.InteraxThreeBodyForce)][0]
    e0, f0 = oracle.threebody_force(force, pos, box)
    e1, f1, triples = oracle.cpu_threebody_force(force, pos, box, 1.0, 4)
    assert triples == oracle.threebody_count(force, pos, box)
    assert abs(e1-e0) <= 1e-11*abs(e0)
    assert np.linalg.norm(f1-f0) <= 1e-11*np.linalg.norm(f0)
