"""The dispersion force's own tests restated on the HIP platform (TestInteraxDispersionForce.h:89-272), and the update
guards of copyParametersToContext (ReferenceInteraxKernels.cpp:98-102) through the C ABI.

Expected values are independent of the oracle where the reference's are: the combining rule by hand at f_n = 1
(testCoefficients), the regularized incomplete gamma function of scipy for the damping (testDamping, on both sides of
every switch point x = (n+1)/2 of the series), central differences for forces.  The long range correction is compared
with the oracle's FP64 restatement (itself pinned against Simpson quadrature in test_oracle.py) at the reference's
1e-9.  Tolerances are those of the pair tests (TestInteraxPairForce.h:719-730).
"""
import ctypes

import numpy as np
import pytest

import independent as ind
import interax_hip as ih
import oracle
from test_gpu_pair import DISPERSION_TYPES, PRECISIONS, TOL, make_force

DAMPING_SCALE = 0.7155241370201111


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_coefficients(precision):
    # TestInteraxDispersionForce.h:89-111: at r = 2 nm (x = 89) the damping is 1
    system = ih.System(2)
    f = ih.InteraxDispersionForce()
    f.addParticle(2.0, 0.1, 0.005, 0.002, 0.025)
    f.addParticle(0.5, 0.01, 0.0002, 0.0003, 0.02)
    f.setDampingScale(1.0)
    system.addForce(f)
    context = ih.Context(system, {"Precision": precision})
    r = 2.0
    context.setPositions(np.array([[0, 0, 0], [0, 0, r]]))
    den = 0.002*0.5+0.0003*2.0
    c6 = 0.002*0.0003*2.0*0.5/den
    c8 = 1.5*0.002*0.0003*(2.0*0.01+0.1*0.5)/den
    c10 = (2*0.002*0.0003*(2.0*0.0002+0.005*0.5)+4.2*0.002*0.0003*0.1*0.01)/den
    expected = -(c6/r**6+c8/r**8+c10/r**10)
    state = context.getState()
    etol, ftol = TOL[precision]
    assert state.getPotentialEnergy() == pytest.approx(expected, rel=etol)
    dEdR = 6*c6/r**7+8*c8/r**9+10*c10/r**11
    assert np.allclose(state.getForces(), [[0, 0, dEdR], [0, 0, -dEdR]], rtol=ftol, atol=1e-8)
    context.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_damping(precision):
    # TestInteraxDispersionForce.h:117-138: both sides of x = 3.5, 4.5, 5.5 (closed form / series of f6, f8, f10)
    system = ih.System(2)
    f = make_force("dispersion")
    f.addParticle(*DISPERSION_TYPES[0])
    f.addParticle(*DISPERSION_TYPES[3])
    system.addForce(f)
    context = ih.Context(system, {"Precision": precision})
    width_sum = DISPERSION_TYPES[0][4]+DISPERSION_TYPES[3][4]
    etol, ftol = TOL[precision]
    for x in (0.01, 0.5, 2.0, 3.49, 3.51, 4.49, 4.51, 5.49, 5.51, 8.0, 20.0):
        r = x*width_sum/(2*DAMPING_SCALE)
        context.setPositions(np.array([[0, 0, 0], [r, 0, 0]]))
        state = context.getState()

        def energy(s):
            return ind.dispersion_energy(s, DISPERSION_TYPES[0], DISPERSION_TYPES[3], DAMPING_SCALE)
        expected = energy(r)
        h = 1e-3*r
        dEdR = (energy(r-2*h)-8*energy(r-h)+8*energy(r+h)-energy(r+2*h))/(12*h)
        assert state.getPotentialEnergy() == pytest.approx(expected, rel=etol), x
        assert np.abs(state.getForces()-[[dEdR, 0, 0], [-dEdR, 0, 0]]).max() <= ftol*abs(dEdR)+1e-8, x
    context.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_zero_polarizability(precision):
    # TestInteraxDispersionForce.h:143-158: a lone-pair site (all parameters zero) does not interact
    system = ih.System(3)
    f = make_force("dispersion")
    f.addParticle(*DISPERSION_TYPES[0])
    f.addParticle(0.0, 0.0, 0.0, 0.0, 0.0)
    f.addParticle(*DISPERSION_TYPES[1])
    system.addForce(f)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(np.array([[0, 0, 0], [0.05, 0, 0], [0, 0.3, 0]]))
    state = context.getState()
    expected = ind.dispersion_energy(0.3, DISPERSION_TYPES[0], DISPERSION_TYPES[1], DAMPING_SCALE)
    assert state.getPotentialEnergy() == pytest.approx(expected, rel=TOL[precision][0])
    assert np.all(state.getForces()[1] == 0.0)
    context.close()


def lrc_energy(force, volume, lam):
    coef, alch = oracle.long_range_correction(force)
    return (coef+lam*alch)/volume


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("use_switch", [False, True])
def test_long_range_correction(precision, use_switch):
    # TestInteraxDispersionForce.h:200-258: 8 particles more than a cutoff apart, so the energy is the correction
    n = 8
    system = ih.System(n, [2, 0, 0, 0, 2, 0, 0, 0, 2])
    f = make_force("dispersion")
    for i in range(n):
        f.addParticle(*DISPERSION_TYPES[i % 4], i < 2)
    f.setNonbondedMethod(ih.InteraxPairForce.CutoffPeriodic)
    f.setCutoffDistance(0.6)
    f.setUseLongRangeCorrection(True)
    f.setUseSwitchingFunction(use_switch)
    f.setSwitchingDistance(0.45)
    system.addForce(f)
    positions = np.array([[i % 2, (i//2) % 2, i//4] for i in range(n)], dtype=float)+[0.1, 0.2, 0.3]
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(positions)
    assert context.getState().getPotentialEnergy() == pytest.approx(lrc_energy(f, 8.0, 1.0), rel=1e-9)
    # proportional to 1/V
    context.setPeriodicBoxVectors([2.2, 0, 0], [0, 2.1, 0], [0, 0, 2.3])
    assert context.getState().getPotentialEnergy() == pytest.approx(lrc_energy(f, 2.2*2.1*2.3, 1.0), rel=1e-9)
    context.setPeriodicBoxVectors([2, 0, 0], [0, 2, 0], [0, 0, 2])
    # every pair with an alchemical particle, self pairs included, is scaled by lambda
    context.setParameter("lambda_sterics", 0.3)
    assert context.getState().getPotentialEnergy() == pytest.approx(lrc_energy(f, 8.0, 0.3), rel=1e-9)
    context.setParameter("lambda_sterics", 1.0)
    # updateParametersInContext updates it
    old = f.getParticleParameters(3)
    f.setParticleParameters(3, 2.5, 0.06, 0.002, 0.0015, 0.023, False)
    f.setDampingScale(0.8)
    context.updateParametersInContext(f)
    expected = lrc_energy(f, 8.0, 1.0)
    assert context.getState().getPotentialEnergy() == pytest.approx(expected, rel=1e-9)
    f.setParticleParameters(3, *old)
    f.setDampingScale(DAMPING_SCALE)
    assert abs(expected-lrc_energy(f, 8.0, 1.0)) > 1e-3*abs(expected)
    context.close()
    # no correction when it is disabled or without periodic boundary conditions (TestInteraxDispersionForce.h:259-272)
    f.setUseLongRangeCorrection(False)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(positions)
    assert context.getState().getPotentialEnergy() == 0.0
    context.close()
    f.setUseLongRangeCorrection(True)
    f.setNonbondedMethod(ih.InteraxPairForce.CutoffNonPeriodic)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(positions)
    assert context.getState().getPotentialEnergy() == 0.0
    context.close()


@pytest.mark.gpu
def test_update_guards():
    # ReferenceInteraxKernels.cpp:98-102, 142-148: the handle refuses a new particle count or overlap method
    lib = ih.load_library()
    system = ih.System(3)
    f = ih.InteraxExchangeForce()
    for k in range(3):
        f.addParticle(3.0, 0.02+0.001*k)
    system.addForce(f)
    group = ih._PairGroup(0, system, [f], ih.IXH_MIXED)
    try:
        f.addParticle(3.0, 0.02)
        system.addParticle()
        desc, keep = group._describe()
        assert lib.ixh_pair_set_parameters(group.handle, ctypes.byref(desc)) == ih.IXH_ERROR_STATE
        assert "The number of particles has changed" in lib.ixh_last_error().decode()
        f._particles.pop()
        system._masses.pop()
        f.setOverlapMethod(ih.InteraxExchangeForce.InteraxOrder2)
        desc, keep = group._describe()
        assert lib.ixh_pair_set_parameters(group.handle, ctypes.byref(desc)) == ih.IXH_ERROR_STATE
        assert "overlap method has changed" in lib.ixh_last_error().decode()
        f.setOverlapMethod(ih.InteraxExchangeForce.Exact)
        desc, keep = group._describe()
        assert lib.ixh_pair_set_parameters(group.handle, ctypes.byref(desc)) == ih.IXH_OK
    finally:
        group.close()
