"""The neighbour lists of the three-body force's plain handles: lists within cutoff + 0.1 nm, rebuilt when a site has
moved more than half that skin (interaxThreeBody.cc:11-30), checked on the device at every evaluation.

Several coordinate frames on one stream through the C ABI, with the check deferred (IXH_DEFER_LIST_CHECK, read back by
ixh_threebody_list_was_valid) and without it: small moves keep the lists, a 0.06 nm move of one site is caught, and
every frame's energy, forces and triple count match the oracle at the reference's tolerances
(TestInteraxThreeBodyForce.h:608-617), whatever lists it was evaluated with.
"""
import ctypes

import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads
from device_frames import Frames

FLAGS = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY


def fluid(n=3000, seed=2):
    system, pos = workloads.atm_fluid(n, seed=seed)
    return system, system.getForce(0), pos


def check(energy, forces, force, pos, box, precision):
    e_ref, f_ref, scale = oracle.threebody_force(force, pos, box, with_scale=True)
    etol, ftol = (1e-10, 1e-8) if precision == ih.IXH_DOUBLE else (1e-5, 1e-4)
    assert abs(energy-e_ref) <= etol*scale
    assert np.abs(forces-f_ref).max() <= ftol*np.abs(f_ref).max()+1e-8
    if precision != ih.IXH_DOUBLE:
        assert np.linalg.norm(forces-f_ref) <= 1e-5*np.linalg.norm(f_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [ih.IXH_MIXED, ih.IXH_DOUBLE])
@pytest.mark.parametrize("defer", [True, False])
def test_lists_kept_and_rebuilt(precision, defer):
    lib = ih.load_library()
    system, force, base = fluid()
    box = system.getDefaultPeriodicBoxVectors()
    n = system.getNumParticles()
    small = workloads.perturbed(base, 1, sigma=0.004)     # every site well within half the skin of base
    moved = small.copy()
    moved[11] += [0.05, -0.03, 0.0]                        # 0.058 nm from base: beyond half the 0.1 nm skin
    after = workloads.perturbed(moved, 3, sigma=0.004)
    frames_pos = [base, small, moved, after]
    print(f"three-body lists: {n} sites, precision {precision}, deferred {defer}", flush=True)
    frames = Frames(lib, n, frames_pos, box, precision)
    tb = ih._ThreeBody(0, system, force, precision)
    try:
        frames.share_stream()
        verdicts, results = [], []
        for k in range(len(frames_pos)):
            frames.clear(k)
            flags = FLAGS | (ih.IXH_DEFER_LIST_CHECK if defer else 0)
            ih._check(lib.ixh_threebody_execute(tb.handle, ctypes.byref(frames.buffers[k]), flags))
            valid = ctypes.c_int()
            ih._check(lib.ixh_threebody_list_was_valid(tb.handle, ctypes.byref(valid)))
            verdicts.append(valid.value)
            if not valid.value:
                frames.clear(k)
                ih._check(lib.ixh_threebody_execute(tb.handle, ctypes.byref(frames.buffers[k]), FLAGS))
            count = ctypes.c_longlong()
            ih._check(lib.ixh_threebody_last_triple_count(tb.handle, ctypes.byref(count)))
            results.append((frames.result(k), count.value))
        # deferred: the moved frame is caught and evaluated again; synchronous: the call rebuilds by itself
        assert verdicts == ([1, 1, 0, 1] if defer else [1, 1, 1, 1])
    finally:
        frames.close()
        tb.close()
    for pos, ((energy, forces), count) in zip(frames_pos, results):
        assert count == oracle.threebody_count(force, pos, box)
        check(energy, forces, force, pos, box, precision)
