"""Energies at many lambda values from one evaluation (ixh_pair_lambda_energies, Context.getEnergiesAtLambdas):
checked against one getState() per value on the HIP path and against the FP64 oracle at every value.
"""
import ctypes

import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads
from test_gpu_pair import PRECISIONS, large_system

VALUES = np.linspace(0.0, 1.0, 16)


def _separate(context, name, values):
    out = []
    for v in values:
        context.setParameter(name, float(v))
        out.append(context.getState(getForces=False).getPotentialEnergy())
    return np.array(out)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", ["dispersion", "exchange", "short_range"])
def test_matches_separate_evaluations(model, precision):
    system, force, pos = large_system(model, seed=3)
    box = np.array(system._box, dtype=float).reshape(3, 3)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    context.setParameter(force.getAlchemicalParameterName(), 0.37)
    fused = context.getEnergiesAtLambdas(VALUES, force.getAlchemicalParameterName())
    separate = _separate(context, force.getAlchemicalParameterName(), VALUES)
    scale = np.abs(separate).max()
    tol = 1e-11 if precision == "double" else 2e-6
    assert np.abs(fused-separate).max() <= tol*scale+1e-9, (fused, separate)
    # and the FP64 oracle at every value
    for v, e in zip(VALUES, fused):
        ref = oracle.evaluate(system, pos, box, parameters={force.getAlchemicalParameterName(): float(v)})[0]
        assert abs(e-ref) <= (1e-10 if precision == "double" else 2e-5)*max(abs(ref), 1.0), (v, e, ref)
    context.close()


@pytest.mark.gpu
@pytest.mark.parametrize("softcore", [0.0, 0.3])
def test_alchemical_box(softcore):
    """The alchemical job of the served configuration: a 12-atom solute in 1000 water-like molecules, all three pair
    forces fused, fused windows against separate evaluations (without the long range correction here; the windows
    with it are checked against the oracle below)."""
    system, pos = workloads.alchemical_box(softcore, seed=5)
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    fused = context.getEnergiesAtLambdas(VALUES)
    separate = _separate(context, "lambda_sterics", VALUES)
    assert np.abs(fused-separate).max() <= 2e-6*np.abs(separate).max(), (fused-separate)
    # the energy changes with lambda (the solute is switched on and off)
    assert abs(fused[-1]-fused[0]) > 1.0
    context.close()


@pytest.mark.gpu
def test_without_alchemical_particles():
    system, pos = workloads.config0()
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    fused = context.getEnergiesAtLambdas(VALUES[:5])
    e = context.getState(getForces=False).getPotentialEnergy()
    assert np.allclose(fused, e, rtol=1e-12, atol=0)
    context.close()


@pytest.mark.gpu
def test_buffers_untouched_and_arguments_checked():
    system, force, pos = large_system("exchange", seed=4)
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    group = context._groups[0]
    b = context._buffers(0.5)
    ih._check(ih._lib.ixh_state_clear(context._state))
    values = np.ascontiguousarray(VALUES[:4])
    out = np.zeros(4)
    ih._check(ih._lib.ixh_pair_lambda_energies(group.handle, ctypes.byref(b), ih.IXH_TERM_EXCHANGE, ih._dptr(values), 4, ih._dptr(out)))
    # the buffers' forces and energy are not touched
    e = ctypes.c_double()
    ih._check(ih._lib.ixh_state_get_energy(context._state, ctypes.byref(e)))
    forces = np.zeros((system.getNumParticles(), 3))
    ih._check(ih._lib.ixh_state_get_forces(context._state, ih._dptr(forces)))
    assert e.value == 0.0 and not forces.any()
    for count in (0, 17):
        big = np.zeros(17)
        status = ih._lib.ixh_pair_lambda_energies(group.handle, ctypes.byref(b), ih.IXH_TERM_EXCHANGE, ih._dptr(big), count, ih._dptr(big))
        assert status == 1   # IXH_ERROR_INVALID_ARGUMENT
    context.close()


@pytest.mark.gpu
@pytest.mark.parametrize("softcore", [0.0, 0.3])
def test_alchemical_box_against_oracle(softcore):
    """The alchemical job at its full size (12-atom solute in 1000 water-like molecules, 3012 atoms, 3.106 nm) with the
    long range correction, whose alchemical part scales with lambda: every fused window energy against the FP64
    oracle at that lambda (pair terms with the correction left out, plus the oracle's correction coefficients)."""
    system, pos = workloads.alchemical_box(softcore, seed=2, use_lrc=True)
    box = system.getDefaultPeriodicBoxVectors()
    disp = system.getForce(0)
    coef, alch = oracle.long_range_correction(disp)
    assert alch != 0.0
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    fused = context.getEnergiesAtLambdas(VALUES)
    context.close()
    disp.setUseLongRangeCorrection(False)
    volume = box[0, 0]*box[1, 1]*box[2, 2]
    for v, e in zip(VALUES, fused):
        ref = oracle.evaluate(system, pos, box, parameters={"lambda_sterics": float(v)})[0]+(coef+v*alch)/volume
        assert abs(e-ref) <= 2e-6*abs(ref), (v, e, ref)
