"""The HIP platform against the golden vectors of the reference (TestInteraxExportValues.h:150-198) and against the
oracle on the same systems.

Tolerances are those of the reference's export test: InteraxOrder2 in double precision 1e-8 (energy, relative) and
1e-7 (every force component, relative to the RMS force of the term); Exact in double precision 1e-6 and 1e-5; Exact
in single or mixed precision 2e-6 and 5e-5.  Against the oracle (the same method) the mixed precision forces must
also have an RMS error of at most 1e-5 of the RMS force.
"""
import numpy as np
import pytest

import interax_hip as ih
import oracle
from golden_data import load_export_values, system_path

SYSTEMS = ["cluster", "methanol_box"]


def run(name, precision, exact):
    values = load_export_values(name)
    system = ih.load_system_xml(system_path(name))
    if exact:
        for f in system.getForces():
            if hasattr(f, "setOverlapMethod"):
                f.setOverlapMethod(0)
    context = ih.Context(system, {"Precision": precision})
    if values.periodic:
        context.setPeriodicBoxVectors(*values.box)
    context.setPositions(values.positions)
    results = []
    for group, term in enumerate(values.terms):
        state = context.getState(groups=1 << group)
        results.append((term, state.getPotentialEnergy(), state.getForces(), values.energies[group], values.forces[group]))
    context.close()
    return values, system, results


def check(results, energy_tol, force_tol):
    for term, energy, forces, ref_energy, ref_forces in results:
        rms = np.sqrt(np.mean(ref_forces**2))
        e_err = abs(energy-ref_energy)/abs(ref_energy)
        f_err = np.abs(forces-ref_forces).max()/rms
        print(f"  {term}: energy error {e_err:.2e}, force error {f_err:.2e} of the RMS force")
        assert e_err <= energy_tol, term
        assert f_err <= force_tol, term


@pytest.mark.gpu
@pytest.mark.parametrize("name", SYSTEMS)
def test_double_order2(name):
    values, system, results = run(name, "double", False)
    check(results, 1e-8, 1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("name", SYSTEMS)
def test_double_exact(name):
    values, system, results = run(name, "double", True)
    check(results, 1e-6, 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["single", "mixed"])
@pytest.mark.parametrize("name", SYSTEMS)
def test_single_exact(name, precision):
    values, system, results = run(name, precision, True)
    check(results, 2e-6, 5e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("name", SYSTEMS)
def test_mixed_against_oracle(name):
    values, system, results = run(name, "mixed", True)
    box = values.box if values.periodic else None
    for group, (term, energy, forces, _, _) in enumerate(results):
        e_ref, f_ref = oracle.evaluate(system, values.positions, box, groups=1 << group, double_precision=False)
        rms_err = np.linalg.norm(forces-f_ref)/np.linalg.norm(f_ref)
        print(f"  {term}: force RMS error {rms_err:.2e}, energy error {abs(energy-e_ref)/abs(e_ref):.2e}")
        assert rms_err <= 1e-5, term
        assert abs(energy-e_ref) <= 2e-6*abs(e_ref), term
