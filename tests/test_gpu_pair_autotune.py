"""The tile kernel's tuner (ixh_pair_autotune): it evaluates into the handle's own scratch buffers (the caller's forces
and energy are left as they are), and the shape it keeps changes the forces not at all (64-bit fixed point, exact in
any order) and the energy only by the order of its FP64 sum."""
import ctypes

import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads
from device_frames import Frames

FLAGS = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [ih.IXH_MIXED, ih.IXH_DOUBLE])
def test_autotune_keeps_forces_and_buffers(precision):
    lib = ih.load_library()
    system, pos = workloads.water_box(10, 3.1, 0.9, seed=7)
    forces = list(system.getForces())
    box = system.getDefaultPeriodicBoxVectors()
    n = system.getNumParticles()
    print(f"pair autotune: {n} atoms, precision {precision}", flush=True)
    frames = Frames(lib, n, [pos, pos], box, precision)
    group = ih._PairGroup(0, system, forces, precision)
    try:
        frames.clear(0)
        ih._check(lib.ixh_pair_set_tiles_per_item(group.handle, 1))
        ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[0]), 7, FLAGS))
        e0, f0 = frames.result(0)
        frames.clear(1)
        chosen = ctypes.c_int()
        ih._check(lib.ixh_pair_autotune(group.handle, ctypes.byref(frames.buffers[1]), 7, ctypes.byref(chosen)))
        e_untouched, f_untouched = frames.result(1)
        ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[1]), 7, FLAGS))
        e1, f1 = frames.result(1)
    finally:
        group.close()
        frames.close()
    assert chosen.value in (1, 2, 3, 4, 6)
    assert e_untouched == 0.0 and not f_untouched.any()
    assert np.array_equal(f0, f1)
    assert e1 == pytest.approx(e0, rel=1e-12)
    e_ref, f_ref = oracle.evaluate(system, pos, box)
    tol = 1e-9 if precision == ih.IXH_DOUBLE else 2e-4
    assert np.abs(f1-f_ref).max() <= tol*np.abs(f_ref).max()+1e-8
