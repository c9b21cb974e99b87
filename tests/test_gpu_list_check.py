"""The neighbour-list check of the cutoff pair path through the C ABI: several coordinate frames evaluated on one
shared stream (ixh_state_set_stream), the displacement check deferred (IXH_DEFER_LIST_CHECK) and read back with
ixh_pair_list_was_valid, a stale list detected and the step evaluated again with a rebuilt list.

Results are compared with the FP64 oracle at the reference's mixed-precision tolerances (TestInteraxPairForce.h:719-730)
and, bitwise, with the same frames evaluated on their own streams by a fresh handle: the list check and the frame
buffers must not change what an evaluation computes.
"""
import ctypes

import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads

TERMS = ih.IXH_TERM_DISPERSION | ih.IXH_TERM_EXCHANGE | ih.IXH_TERM_SHORT_RANGE
FLAGS = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY


class _Frames:
    def __init__(self, lib, n, positions, box):
        self.lib = lib
        self.states, self.buffers = [], []
        for pos in positions:
            s = ctypes.c_void_p()
            ih._check(lib.ixh_state_create(0, n, ih.IXH_MIXED, ctypes.byref(s)))
            p = np.ascontiguousarray(pos.reshape(-1))
            ih._check(lib.ixh_state_set_positions(s, ih._dptr(p)))
            b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(-1))
            ih._check(lib.ixh_state_set_box(s, ih._dptr(b)))
            self.states.append(s)
            self.buffers.append(None)
            self.refresh(len(self.states)-1)

    def refresh(self, k):
        buf = ih.Buffers()
        ih._check(self.lib.ixh_state_get_buffers(self.states[k], ctypes.byref(buf)))
        self.buffers[k] = buf

    def result(self, k, n):
        f = np.zeros((n, 3))
        ih._check(self.lib.ixh_state_get_forces(self.states[k], ih._dptr(f)))
        e = ctypes.c_double()
        ih._check(self.lib.ixh_state_get_energy(self.states[k], ctypes.byref(e)))
        return e.value, f

    def close(self):
        for s in self.states:
            self.lib.ixh_state_destroy(s)


def _reference(forces, pos, box):
    e, f = 0.0, np.zeros_like(pos)
    for force in forces:
        fe, ff = oracle.pair_force(force, pos, box, 1.0, double_precision=True)
        e += fe
        f += ff
    return e, f


@pytest.mark.gpu
def test_deferred_check_on_one_stream():
    lib = ih.load_library()
    system, base = workloads.config0()
    n = system.getNumParticles()
    box = system.getDefaultPeriodicBoxVectors()
    forces = [f for f in system.getForces() if isinstance(f, ih.InteraxPairForce)]
    small = workloads.perturbed(base, 1, sigma=0.002)      # every particle well within half the skin
    moved = small.copy()
    moved[5] += np.array([0.08, -0.05, 0.03])             # 0.1 nm: beyond half the 0.1 nm skin
    after = workloads.perturbed(moved, 7, sigma=0.002)    # close to the moved frame
    frames_pos = [base, small, moved, after]
    print("deferred list check: 648 atoms, frames base/small/moved/after on one stream", flush=True)

    frames = _Frames(lib, n, frames_pos, box)
    group = ih._PairGroup(0, system, forces, ih.IXH_MIXED)
    try:
        stream = frames.buffers[0].stream
        for k, st in enumerate(frames.states):
            ih._check(lib.ixh_state_set_stream(st, stream))
            frames.refresh(k)
        valid = ctypes.c_int()
        verdicts = []
        for k in range(len(frames_pos)):
            ih._check(lib.ixh_state_clear(frames.states[k]))
            ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[k]), TERMS,
                                           FLAGS | ih.IXH_DEFER_LIST_CHECK))
            ih._check(lib.ixh_pair_list_was_valid(group.handle, ctypes.byref(valid)))
            verdicts.append(valid.value)
            if not valid.value:
                ih._check(lib.ixh_state_clear(frames.states[k]))
                ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[k]), TERMS, FLAGS))
        # the first call builds the list (no check), the moved frame is caught; after its rebuild the last frame is
        # again within half the skin of the new list's positions
        assert verdicts == [1, 1, 0, 1]
        shared = [frames.result(k, n) for k in range(len(frames_pos))]
    finally:
        frames.close()
        group.close()

    # each frame alone, on its own stream, with a fresh handle that builds its list from that frame
    for k, pos in enumerate(frames_pos):
        alone = _Frames(lib, n, [pos], box)
        g = ih._PairGroup(0, system, forces, ih.IXH_MIXED)
        try:
            ih._check(lib.ixh_state_clear(alone.states[0]))
            ih._check(lib.ixh_pair_execute(g.handle, ctypes.byref(alone.buffers[0]), TERMS, FLAGS))
            e_alone, f_alone = alone.result(0, n)
        finally:
            alone.close()
            g.close()
        e_shared, f_shared = shared[k]
        e_ref, f_ref = _reference(forces, pos, box)
        assert abs(e_shared-e_ref) <= 2e-5*abs(e_ref)
        assert np.abs(f_shared-f_ref).max() <= 2e-4*np.abs(f_ref).max()+1e-8
        rms = np.linalg.norm(f_shared-f_ref)/np.linalg.norm(f_ref)
        assert rms <= 1e-5, rms
        # the same interacting pairs in the same tiles where both lists were built from the same positions
        if k in (0, 2):
            assert e_shared == e_alone
            assert np.array_equal(f_shared, f_alone)
