"""Parameter updates from device arrays (ixh_pair_set_parameters_device, ixh_threebody_set_parameters_device) against
the same updates from host arrays (copyParametersToContext, ReferenceInteraxKernels.cpp:98-112, 142-148, 185-191,
215-217).

Two handles are created with the same parameters and evaluated once (so that their neighbour lists exist); one is
updated through host arrays, the other through device arrays on the evaluation's stream; both are then evaluated on the
same positions.  Energies, forces and the long range correction must agree bit for bit, and match the oracle with the
new parameters at the reference's tolerances (TestInteraxPairForce.h, TestInteraxThreeBodyForce.h:608-617).  The
updates cover: every class changing together (the type-table classes hold), one particle leaving its class (the
classes are formed anew), and one exchange width leaving the short-range width (the shared overlap is given up).
"""
import ctypes

import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads
from device_frames import Frames
from hip_memory import DeviceArray

FLAGS = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY
PRECISIONS = {"mixed": ih.IXH_MIXED, "double": ih.IXH_DOUBLE}
TOL = {"mixed": (2e-5, 2e-4), "double": (1e-10, 1e-9)}
RESOLUTION = 1e-8


def rescale_classes(disp, exch, sr):
    # every oxygen and every hydrogen changes alike: the classes hold
    for i in range(disp.getNumParticles()):
        m1, m2, m3, a, w, al = disp.getParticleParameters(i)
        f = 1.1 if i % 3 == 0 else 0.93
        disp.setParticleParameters(i, f*m1, m2, f*m3, a, 0.97*w, al)
        amp, wx, al = exch.getParticleParameters(i)
        exch.setParticleParameters(i, 1.05*amp, wx, al)
        k, l, ws, al = sr.getParticleParameters(i)
        sr.setParticleParameters(i, 0.9*k, l, ws, al)


def one_particle(disp, exch, sr):
    # particle 5 (a hydrogen) leaves its dispersion and short-range classes
    m1, m2, m3, a, w, al = disp.getParticleParameters(5)
    disp.setParticleParameters(5, m1, m2, m3, 1.3*a, w, al)
    k, l, ws, al = sr.getParticleParameters(5)
    sr.setParticleParameters(5, 1.2*k, l, ws, al)


def split_widths(disp, exch, sr):
    # one exchange width differs from its short-range width: one overlap can no longer serve both terms
    amp, wx, al = exch.getParticleParameters(7)
    exch.setParticleParameters(7, amp, 1.1*wx, al)


CHANGES = {"classes_kept": rescale_classes, "one_particle": one_particle, "split_widths": split_widths}


def evaluate(lib, frames, k, handle):
    frames.clear(k)
    ih._check(lib.ixh_pair_execute(handle, ctypes.byref(frames.buffers[k]), 7, FLAGS))
    return frames.result(k)


def lrc(lib, handle):
    c, a = ctypes.c_double(), ctypes.c_double()
    ih._check(lib.ixh_pair_long_range_coefficients(handle, ctypes.byref(c), ctypes.byref(a)))
    return c.value, a.value


@pytest.mark.gpu
@pytest.mark.parametrize("precision", list(PRECISIONS))
@pytest.mark.parametrize("change", list(CHANGES))
def test_pair_device_update_matches_host_update(precision, change):
    lib = ih.load_library()
    system, pos = workloads.water_box(6, 1.862, 0.8, seed=3, use_lrc=True)
    disp, exch, sr = system.getForces()
    box = system.getDefaultPeriodicBoxVectors()
    n = system.getNumParticles()
    print(f"device parameter update: {n} particles, {precision}, {change}", flush=True)
    frames = Frames(lib, n, [pos, pos], box, PRECISIONS[precision])
    frames.share_stream()
    host = ih._PairGroup(0, system, [disp, exch, sr], PRECISIONS[precision])
    dev = ih._PairGroup(0, system, [disp, exch, sr], PRECISIONS[precision])
    arrays = []
    try:
        before = evaluate(lib, frames, 0, host.handle)
        assert evaluate(lib, frames, 1, dev.handle)[0] == before[0]
        CHANGES[change](disp, exch, sr)
        host.update()
        arrays = [DeviceArray(f.parameter_array()) for f in (disp, exch, sr)]
        params = ih.PairDeviceParameters(arrays[0].address, arrays[1].address, arrays[2].address, disp.getDampingScale(),
                                         frames.buffers[1].stream)
        ih._check(lib.ixh_pair_set_parameters_device(dev.handle, ctypes.byref(params)))
        e_host, f_host = evaluate(lib, frames, 0, host.handle)
        e_dev, f_dev = evaluate(lib, frames, 1, dev.handle)
        lrc_host, lrc_dev = lrc(lib, host.handle), lrc(lib, dev.handle)
        # a second evaluation after the update is unchanged
        e_again, f_again = evaluate(lib, frames, 1, dev.handle)
    finally:
        host.close()
        dev.close()
        frames.close()
        for a in arrays:
            a.free()
    assert e_dev == e_host and np.array_equal(f_dev, f_host)
    assert e_again == e_dev and np.array_equal(f_again, f_dev)
    assert lrc_dev == lrc_host
    assert e_host != before[0]
    # the oracle with the new parameters; its energy includes the long range correction (ReferenceInteraxKernels.cpp:93-94)
    e_ref, f_ref, scale = oracle.evaluate(system, pos, box, with_scale=True)
    c_ref, a_ref = oracle.long_range_correction(disp)
    volume = box[0][0]*box[1][1]*box[2][2]
    etol, ftol = TOL[precision]
    assert abs(e_host+lrc_host[0]/volume-e_ref) <= etol*scale+1e-12, (e_host, e_ref)
    assert np.abs(f_host-f_ref).max() <= ftol*np.abs(f_ref).max()+RESOLUTION
    assert lrc_host[0] == pytest.approx(c_ref, rel=1e-10)
    assert lrc_host[1] == pytest.approx(a_ref, rel=1e-10, abs=1e-300)


@pytest.mark.gpu
def test_pair_noalloc_waits_for_a_device_update():
    # type tables: the update is finished by the next ixh_pair_execute; until then the no-alloc path refuses
    lib = ih.load_library()
    system, pos = workloads.water_box(6, 1.862, 0.8, seed=4)
    disp, exch, sr = system.getForces()
    n = system.getNumParticles()
    frames = Frames(lib, n, [pos], system.getDefaultPeriodicBoxVectors(), ih.IXH_MIXED)
    group = ih._PairGroup(0, system, [disp, exch, sr], ih.IXH_MIXED)
    arrays = []
    try:
        evaluate(lib, frames, 0, group.handle)
        rescale_classes(disp, exch, sr)
        arrays = [DeviceArray(f.parameter_array()) for f in (disp, exch, sr)]
        params = ih.PairDeviceParameters(arrays[0].address, arrays[1].address, arrays[2].address, 1.0, frames.buffers[0].stream)
        ih._check(lib.ixh_pair_set_parameters_device(group.handle, ctypes.byref(params)))
        frames.clear(0)
        assert lib.ixh_pair_execute_noalloc(group.handle, ctypes.byref(frames.buffers[0]), 7, FLAGS) == ih.IXH_ERROR_STATE
        e1, f1 = evaluate(lib, frames, 0, group.handle)
        frames.clear(0)
        ih._check(lib.ixh_pair_execute_noalloc(group.handle, ctypes.byref(frames.buffers[0]), 7, FLAGS))
        e2, f2 = frames.result(0)
    finally:
        group.close()
        frames.close()
        for a in arrays:
            a.free()
    assert e1 == e2 and np.array_equal(f1, f2)


def threebody_update(force):
    for i in range(force.getNumSites()):
        parts, a, w, width, mol, al = force.getSiteParameters(i)
        force.setSiteParameters(i, parts, (1.1 if i % 2 else 0.95)*a, 1.02*w, width, mol, al)
    force.setScaleFactor(0.9)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", list(PRECISIONS))
def test_threebody_device_update_matches_host_update(precision):
    lib = ih.load_library()
    system, pos = workloads.atm_fluid(2000, seed=5)
    force = system.getForce(0)
    box = system.getDefaultPeriodicBoxVectors()
    n = system.getNumParticles()
    print(f"three-body device parameter update: {n} sites, {precision}", flush=True)
    frames = Frames(lib, n, [pos, pos], box, PRECISIONS[precision])
    frames.share_stream()
    host = ih._ThreeBody(0, system, force, PRECISIONS[precision])
    dev = ih._ThreeBody(0, system, force, PRECISIONS[precision])
    arrays = []
    try:
        for k, h in enumerate((host, dev)):
            frames.clear(k)
            ih._check(lib.ixh_threebody_execute(h.handle, ctypes.byref(frames.buffers[k]), FLAGS))
        before = frames.result(0)
        threebody_update(force)
        host.update()
        _, _, alpha, freq, width, _, _ = force.site_arrays()
        arrays = [DeviceArray(np.asarray(x, dtype=np.float64)) for x in (alpha, freq, width)]
        params = ih.ThreeBodyDeviceParameters(arrays[0].address, arrays[1].address, arrays[2].address, force.getScaleFactor(),
                                              force.getDampingScale(), force.getCentroidDampingRadius(),
                                              frames.buffers[1].stream)
        ih._check(lib.ixh_threebody_set_parameters_device(dev.handle, ctypes.byref(params)))
        results = []
        for k, h in enumerate((host, dev)):
            frames.clear(k)
            ih._check(lib.ixh_threebody_execute(h.handle, ctypes.byref(frames.buffers[k]), FLAGS))
            results.append(frames.result(k))
    finally:
        host.close()
        dev.close()
        frames.close()
        for a in arrays:
            a.free()
    (e_host, f_host), (e_dev, f_dev) = results
    assert e_dev == e_host and np.array_equal(f_dev, f_host)
    assert e_host != before[0]
    e_ref, f_ref, scale = oracle.threebody_force(force, pos, box, with_scale=True)
    etol, ftol = (1e-10, 1e-8) if precision == "double" else (1e-5, 1e-4)
    assert abs(e_host-e_ref) <= etol*scale
    assert np.abs(f_host-f_ref).max() <= ftol*np.abs(f_ref).max()+RESOLUTION


@pytest.mark.gpu
def test_threebody_device_update_checks_tang_toennies_widths():
    # InteraxThreeBodyForceImpl.cpp: the Tang-Toennies damping needs a positive width; a device update cannot be
    # checked on the host, so the next evaluation reports it
    lib = ih.load_library()
    system, pos = workloads.atm_fluid(500, seed=6)
    force = system.getForce(0)
    for i in range(force.getNumSites()):
        parts, a, w, width, mol, al = force.getSiteParameters(i)
        force.setSiteParameters(i, parts, a, w, 0.03, mol, al)
    force.setDampingMethod(ih.InteraxThreeBodyForce.TangToennies)
    n = system.getNumParticles()
    frames = Frames(lib, n, [pos], system.getDefaultPeriodicBoxVectors(), ih.IXH_MIXED)
    handle = ih._ThreeBody(0, system, force, ih.IXH_MIXED)
    _, _, alpha, freq, width, _, _ = force.site_arrays()
    width = np.asarray(width, dtype=np.float64).copy()
    width[17] = 0.0
    arrays = [DeviceArray(np.asarray(x, dtype=np.float64)) for x in (alpha, freq, width)]
    try:
        params = ih.ThreeBodyDeviceParameters(arrays[0].address, arrays[1].address, arrays[2].address, 1.0, 1.0, 0.0,
                                              frames.buffers[0].stream)
        ih._check(lib.ixh_threebody_set_parameters_device(handle.handle, ctypes.byref(params)))
        frames.clear(0)
        status = lib.ixh_threebody_execute(handle.handle, ctypes.byref(frames.buffers[0]), FLAGS)
    finally:
        handle.close()
        frames.close()
        for a in arrays:
            a.free()
    assert status == ih.IXH_ERROR_INVALID_ARGUMENT
