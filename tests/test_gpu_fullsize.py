"""Parity at the full sizes of the benchmark configurations, against the FP64 oracle.

- The 98 304-atom water-like box (9.94 nm, cutoff 1.0 nm): the headline workload.  Its box is wide enough for most
  tiles to take one periodic image per pair of blocks, while tiles between two elongated blocks take the minimum
  image of every pair: both paths of the benchmarked kernel run (ixh_pair_image_mode).  Two perturbed coordinate
  sets on one stream, the second with the deferred list check, in mixed precision (force RMS error <= 1e-5, largest
  component error <= 5e-5 of the RMS force, energy 2e-6: TestInteraxExportValues.h:195-196) and in double precision (TestInteraxPairForce.h:719-725: energy 1e-10, forces
  1e-9 of the largest force); the number of interacting pairs equals the oracle's.
- Its long range correction (98 304 dispersion classes: per-atom damping widths) against the closed form with the
  damping functions set to 1, which at r >= 1.0 nm (x >= 30) is exact to ~1e-9 (4.5e-10 on a 1 536-atom box).
- The per-atom-class long range correction of a 1 536-atom box against the oracle's FP64 restatement
  (InteraxDispersionForceImpl.cpp:86-149), with and without switching and alchemical particles.
- The 32 768-site ATM fluid (cutoff 0.7 nm) with triple-count equality; a periodic grid of 17^3 > 4096 cells, so that
  the counting sort's scan carries its running total across chunks of 4096 cells.
- The 1 024-cluster batch with the oracle on every 16th cluster.
"""
import ctypes
import math

import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads
from device_frames import Frames

TERMS = ih.IXH_TERM_DISPERSION | ih.IXH_TERM_EXCHANGE | ih.IXH_TERM_SHORT_RANGE
FLAGS = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY


def reference(forces, pos, box):
    e, f = 0.0, np.zeros_like(pos)
    for force in forces:
        fe, ff = oracle.pair_force(force, pos, box, 1.0, double_precision=True)
        e += fe
        f += ff
    return e, f


def check(precision, energy, forces, e_ref, f_ref):
    err = forces-f_ref
    rms_ref = np.sqrt(np.mean(f_ref**2))
    rms = np.linalg.norm(err)/np.linalg.norm(f_ref)
    worst = np.abs(err).max()
    print(f"  {precision}: force RMS error {rms:.2e}, largest {worst/rms_ref:.2e} of the RMS force, "
          f"energy {abs(energy-e_ref)/abs(e_ref):.2e}", flush=True)
    if precision == "double":
        assert abs(energy-e_ref) <= 1e-10*abs(e_ref)
        assert worst <= 1e-9*np.abs(f_ref).max()+1e-8
    else:
        assert rms <= 1e-5
        assert worst <= 5e-5*rms_ref
        assert abs(energy-e_ref) <= 2e-6*abs(e_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_water_box_98304(precision):
    lib = ih.load_library()
    system, base = workloads.config1()
    forces = [f for f in system.getForces() if isinstance(f, ih.InteraxPairForce)]
    forces[0].setUseLongRangeCorrection(False)     # the correction is checked on its own below
    box = system.getDefaultPeriodicBoxVectors()
    n = system.getNumParticles()
    steps = [workloads.perturbed(base, 0), workloads.perturbed(base, 1)]
    print(f"98304-atom box, {precision}: two coordinate sets on one stream", flush=True)
    code = ih.IXH_DOUBLE if precision == "double" else ih.IXH_MIXED
    frames = Frames(lib, n, steps, box, code)
    group = ih._PairGroup(0, system, forces, code)
    try:
        frames.share_stream()
        results = []
        for k in range(len(steps)):
            frames.clear(k)
            flags = FLAGS | (ih.IXH_DEFER_LIST_CHECK if k > 0 else 0)
            ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[k]), TERMS, flags))
            valid = ctypes.c_int()
            ih._check(lib.ixh_pair_list_was_valid(group.handle, ctypes.byref(valid)))
            # the steps are within a quarter of the skin of each other: the list built for the first holds
            assert valid.value == 1
            count = ctypes.c_longlong()
            ih._check(lib.ixh_pair_last_pair_count(group.handle, ctypes.byref(count)))
            results.append((frames.result(k), count.value))
        per_block, tiles = ctypes.c_longlong(), ctypes.c_longlong()
        ih._check(lib.ixh_pair_image_mode(group.handle, ctypes.byref(per_block), ctypes.byref(tiles)))
        print(f"  {per_block.value} of {tiles.value} tiles take one image shift per block pair", flush=True)
        # both image paths of the tile kernel run: per block pair for most tiles, per pair where two large blocks
        # meet (Morton blocks that straddle a cell boundary are elongated)
        assert 0.5*tiles.value < per_block.value < tiles.value
    finally:
        frames.close()
        group.close()
    for pos, ((energy, f), count) in zip(steps, results):
        e_ref, f_ref = reference(forces, pos, box)
        check(precision, energy, f, e_ref, f_ref)
        assert count == oracle.pair_count(forces[1], pos, box)


def dispersion_coefficients(pi, pj):
    """C6, C8, C10 of a pair from (m1, m2, m3, alpha) (InteraxPairTerms.h:271-279)."""
    m1i, m2i, m3i, ai = pi
    m1j, m2j, m3j, aj = pj
    pre = ai*aj/max(ai*m1j+aj*m1i, 1e-30)
    return pre*m1i*m1j, 1.5*pre*(m1i*m2j+m2i*m1j), pre*(2*(m1i*m3j+m3i*m1j)+4.2*m2i*m2j)


@pytest.mark.gpu
def test_long_range_correction_98304_classes():
    lib = ih.load_library()
    system, _ = workloads.config1()
    disp = system.getForce(0)
    rc = disp.getCutoffDistance()
    group = ih._PairGroup(0, system, [disp], ih.IXH_MIXED)
    coefficient, alchemical = ctypes.c_double(), ctypes.c_double()
    try:
        ih._check(lib.ixh_pair_long_range_coefficients(group.handle, ctypes.byref(coefficient), ctypes.byref(alchemical)))
    finally:
        group.close()
    # every particle is a class of its own (per-atom damping widths): every pair i <= j has weight 1; with f_n = 1
    # the integral of r^2 E(r) beyond rc is -(C6/(3 rc^3) + C8/(5 rc^5) + C10/(7 rc^7))
    n = system.getNumParticles()
    counts = {"O": n//3, "H": 2*n//3}
    moments = {el: workloads.dispersion_moments(el) for el in counts}

    def integral(a, b):
        c6, c8, c10 = dispersion_coefficients(moments[a], moments[b])
        return -(c6/(3*rc**3)+c8/(5*rc**5)+c10/(7*rc**7))
    total = (counts["O"]*(counts["O"]+1)/2*integral("O", "O")+counts["H"]*(counts["H"]+1)/2*integral("H", "H") +
             counts["O"]*counts["H"]*integral("O", "H"))
    expected = 2*math.pi*n*n/(0.5*n*(n+1))*total
    print(f"LRC coefficient {coefficient.value:.12g}, closed form {expected:.12g}", flush=True)
    assert alchemical.value == 0.0
    assert coefficient.value == pytest.approx(expected, rel=1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["plain", "switch_alchemical"])
def test_long_range_correction_per_atom_classes(variant):
    lib = ih.load_library()
    system, _ = workloads.water_box(8, 2.48, 0.9, seed=4, use_lrc=True)
    disp = system.getForce(0)
    if variant == "switch_alchemical":
        disp.setUseSwitchingFunction(True)
        disp.setSwitchingDistance(0.75)
        for i in range(6):
            disp.setParticleParameters(i, *disp.getParticleParameters(i)[:5], True)
    coef_ref, alch_ref = oracle.long_range_correction(disp)
    group = ih._PairGroup(0, system, [disp], ih.IXH_MIXED)
    coefficient, alchemical = ctypes.c_double(), ctypes.c_double()
    try:
        ih._check(lib.ixh_pair_long_range_coefficients(group.handle, ctypes.byref(coefficient), ctypes.byref(alchemical)))
    finally:
        group.close()
    assert coefficient.value == pytest.approx(coef_ref, rel=1e-9)
    if variant == "plain":
        assert alchemical.value == 0.0
    else:
        assert alch_ref != 0.0
        assert alchemical.value == pytest.approx(alch_ref, rel=1e-9)


@pytest.mark.gpu
def test_atm_fluid_32768():
    system, pos = workloads.atm_fluid()
    f = system.getForce(0)
    box = system.getDefaultPeriodicBoxVectors()
    print("ATM fluid: 32768 sites, cutoff 0.7 nm, mixed", flush=True)
    context = ih.Context(system, {"Precision": "mixed"})
    try:
        for k in range(2):
            step = workloads.perturbed(pos, k)
            context.setPositions(step)
            state = context.getState()
            e_ref, f_ref = oracle.threebody_force(f, step, box)
            assert context.triple_counts() == [oracle.threebody_count(f, step, box)]
            rms = np.linalg.norm(state.getForces()-f_ref)/np.linalg.norm(f_ref)
            print(f"  step {k}: force RMS error {rms:.2e}", flush=True)
            assert rms <= 1e-5
            assert np.abs(state.getForces()-f_ref).max() <= 1e-4*np.abs(f_ref).max()
            assert state.getPotentialEnergy() == pytest.approx(e_ref, rel=1e-5)
    finally:
        context.close()


@pytest.mark.gpu
def test_periodic_grid_above_4096_cells():
    rng = np.random.default_rng(17)
    L, n = 12.5, 5000
    pos = workloads.rsa_points(rng, n, L, 0.3)
    system = ih.System(n, [L, 0, 0, 0, L, 0, 0, 0, L])
    f = ih.InteraxThreeBodyForce()
    c = rng.uniform(1e-6, 4e-6, size=n)
    for i in range(n):
        f.addSite([i], (16*c[i]/(9*1000.0))**(1/3), 1000.0)
    f.setCutoffDistance(0.7)      # 17 cells of 0.735 nm per axis: 4913 cells
    f.setUsesPeriodicBoundaryConditions(True)
    system.addForce(f)
    box = system.getDefaultPeriodicBoxVectors()
    for precision in ("mixed", "double"):
        context = ih.Context(system, {"Precision": precision})
        try:
            context.setPositions(pos)
            state = context.getState()
            e_ref, f_ref, scale = oracle.threebody_force(f, pos, box, with_scale=True)
            count = oracle.threebody_count(f, pos, box)
            assert count > 1000
            assert context.triple_counts() == [count]
            tol = (1e-10, 1e-8) if precision == "double" else (1e-5, 1e-4)
            assert abs(state.getPotentialEnergy()-e_ref) <= tol[0]*scale
            assert np.abs(state.getForces()-f_ref).max() <= tol[1]*np.abs(f_ref).max()+1e-8
        finally:
            context.close()


@pytest.mark.gpu
def test_batch_1024_clusters():
    sizes = workloads.cluster_sizes(1024)
    built = [workloads.water_cluster(m, seed=k) for k, m in enumerate(sizes)]
    systems = [b[0] for b in built]
    positions = [b[1] for b in built]
    print(f"1024 clusters, {sum(len(p) for p in positions)} atoms, mixed", flush=True)
    batch = ih.Batch(systems, "mixed")
    try:
        energies, forces = batch.evaluate(positions)
    finally:
        batch.close()
    err2 = ref2 = 0.0
    for k in range(0, len(systems), 16):
        e_ref, f_ref, scale = oracle.evaluate(systems[k], positions[k], with_scale=True)
        d = forces[k]-f_ref
        err2 += float((d**2).sum())
        ref2 += float((f_ref**2).sum())
        assert abs(energies[k]-e_ref) <= 2e-6*max(scale, abs(e_ref)), k
        assert np.abs(d).max() <= 5e-5*np.sqrt(np.mean(f_ref**2))+1e-8, k
        assert np.linalg.norm(d) <= 1e-5*np.linalg.norm(f_ref), k
    print(f"  force RMS error over the sampled clusters {math.sqrt(err2/ref2):.2e}", flush=True)
