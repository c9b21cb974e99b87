"""ixh_pair_execute_noalloc: the pair evaluation that allocates nothing, its partial sums in the library's static
device scratch, against ixh_pair_execute on the same handles and positions (the same kernels and launch: the results
must be equal bit for bit).  Two no-allocation calls in a row, the second on a smaller system (fewer partials than the
first left in the scratch), then an all-pairs handle, and the refusals: a handle without a list, a stale list.
"""
import ctypes

import numpy as np
import pytest

import interax_hip as ih
import workloads
from device_frames import Frames

TERMS = ih.IXH_TERM_DISPERSION | ih.IXH_TERM_EXCHANGE | ih.IXH_TERM_SHORT_RANGE
FLAGS = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY


def pair_forces(system):
    return [f for f in system.getForces() if isinstance(f, ih.InteraxPairForce)]


def systems():
    big, big_pos = workloads.config0()                                # 648 atoms, 1.862 nm, cutoff 0.9 nm
    small, small_pos = workloads.water_box(4, 1.3, 0.6, seed=3)       # 192 atoms, 1.3 nm, cutoff 0.6 nm
    cluster, cluster_pos = workloads.water_cluster(30, seed=4)        # 90 atoms, no cutoff
    cluster_forces = pair_forces(cluster)
    bare = ih.System(cluster.getNumParticles())
    for f in cluster_forces:
        bare.addForce(f)
    return [(big, big_pos), (small, small_pos), (bare, cluster_pos)]


def result(lib, frames, k, group):
    energy, forces = frames.result(k)
    count = ctypes.c_longlong()
    ih._check(lib.ixh_pair_last_pair_count(group.handle, ctypes.byref(count)))
    return energy, forces, count.value


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [ih.IXH_MIXED, ih.IXH_DOUBLE])
def test_noalloc_matches_execute(precision):
    lib = ih.load_library()
    cases = []
    for system, pos in systems():
        box = system.getDefaultPeriodicBoxVectors()
        frames = Frames(lib, system.getNumParticles(), [pos, pos], box, precision)
        group = ih._PairGroup(0, system, pair_forces(system), precision)
        cases.append((system, frames, group))
    try:
        expected = []
        for system, frames, group in cases:
            print(f"ixh_pair_execute: {system.getNumParticles()} atoms", flush=True)
            frames.clear(0)
            ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[0]), TERMS, FLAGS))
            expected.append(result(lib, frames, 0, group))
        # the no-allocation calls one after another, each system smaller than the one before
        for (system, frames, group), (e, f, c) in zip(cases, expected):
            print(f"ixh_pair_execute_noalloc: {system.getNumParticles()} atoms", flush=True)
            frames.clear(1)
            ih._check(lib.ixh_pair_execute_noalloc(group.handle, ctypes.byref(frames.buffers[1]), TERMS, FLAGS))
            e2, f2, c2 = result(lib, frames, 1, group)
            assert c2 == c and c > 0
            assert e2 == e
            assert np.array_equal(f2, f)
    finally:
        for _, frames, group in cases:
            frames.close()
            group.close()


@pytest.mark.gpu
def test_noalloc_refusals():
    lib = ih.load_library()
    system, pos = workloads.config0()
    box = system.getDefaultPeriodicBoxVectors()
    moved = pos.copy()
    moved[7] += [0.06, 0.0, 0.0]          # beyond half the 0.1 nm list skin
    frames = Frames(lib, system.getNumParticles(), [pos, moved, moved], box)
    group = ih._PairGroup(0, system, pair_forces(system), ih.IXH_MIXED)
    try:
        frames.share_stream()
        print("ixh_pair_execute_noalloc: refusals on 648 atoms", flush=True)
        # no list yet: the first evaluation of a handle is made by ixh_pair_execute
        assert lib.ixh_pair_execute_noalloc(group.handle, ctypes.byref(frames.buffers[0]), TERMS, FLAGS) == ih.IXH_ERROR_STATE
        assert "no neighbour list" in lib.ixh_last_error().decode()
        ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[0]), TERMS, FLAGS))
        # a stale list is refused, without touching the buffers
        frames.clear(1)
        assert lib.ixh_pair_execute_noalloc(group.handle, ctypes.byref(frames.buffers[1]), TERMS, FLAGS) == ih.IXH_ERROR_STATE
        assert "no longer holds" in lib.ixh_last_error().decode()
        energy, forces = frames.result(1)
        assert energy == 0.0 and not forces.any()
        # ixh_pair_execute rebuilds it; then the no-allocation call agrees with it again
        ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[1]), TERMS, FLAGS))
        e, f = frames.result(1)
        frames.clear(2)
        ih._check(lib.ixh_pair_execute_noalloc(group.handle, ctypes.byref(frames.buffers[2]), TERMS, FLAGS))
        e2, f2 = frames.result(2)
        assert e2 == e and np.array_equal(f2, f)
    finally:
        frames.close()
        group.close()
