"""Reordered contexts: positions and forces in a platform's current particle order, with the atom index array.

OpenMM's GPU platforms reorder particles for locality, exchanging only whole molecules that are identical in every
parameter (PairForceInfo / ThreeBodyForce ForceInfo, CommonInteraxKernels.cpp:111-137, 386-412).  Parameters,
exclusions and sites given in System order then stay valid in the current order; what changes is the System index of
the particle in each slot, which the InteraxOrder2 overlap rule uses (expand about the width of the particle with the
lower System index: CommonInteraxKernels.cpp:213-216, interaxSlaterOverlap.cc:14-17).  The platform passes that
index array (ComputeContext::getAtomIndexArray) with the buffers; these tests permute whole identical molecules,
pass the permutation, and compare the un-permuted results with the golden vectors and the oracle.

The golden methanol box uses InteraxOrder2, but its widths differ pairwise by at least 5.7e-4 nm, more than the
0.01 bohr (5.3e-4 nm) band in which InteraxOrder2 expands a pair of unequal widths, so for it the rule happens not to
matter.  The constructed water box below has O and H widths 2e-4 nm apart, inside the band: there the expansion point
decides the result at the 1e-6 level, and the test checks both that the atom index array gives the oracle's answer
and that the identity does not.
"""
import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads
from golden_data import load_export_values, system_path
from hip_memory import DeviceArray


def molecules(system):
    """Connected components of the exclusion graph of the pair forces (the molecules)."""
    n = system.getNumParticles()
    parent = list(range(n))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i
    for f in system.getForces():
        if isinstance(f, ih.InteraxPairForce):
            for a, b in f._exclusions:
                ra, rb = find(a), find(b)
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
    groups = {}
    for i in range(n):
        groups.setdefault(find(i), []).append(i)
    return sorted(groups.values())


def signature(system, atoms):
    """Everything a platform compares before exchanging two molecules: per-atom parameters of every pair force, the
    internal exclusions (in molecule-local indices) and the sites over the molecule's atoms."""
    local = {a: k for k, a in enumerate(atoms)}
    sig = [len(atoms)]
    for f in system.getForces():
        if isinstance(f, ih.InteraxPairForce):
            sig.append(tuple(tuple(f.getParticleParameters(a)) for a in atoms))
            sig.append(tuple(sorted((min(local[a], local[b]), max(local[a], local[b]))
                                    for a, b in f._exclusions if a in local)))
        else:
            sites = []
            for particles, alpha, freq, width, mol, alch in f._sites:
                if particles[0] in local:
                    sites.append((tuple(local[p] for p in particles), alpha, freq, width, alch))
            sig.append(tuple(sites))
    return tuple(sig)


def molecule_permutation(system, seed):
    """atom_index[slot] = System index of the particle placed in that slot, exchanging identical molecules only."""
    mols = molecules(system)
    classes = {}
    for m, atoms in enumerate(mols):
        classes.setdefault(signature(system, atoms), []).append(m)
    rng = np.random.default_rng(seed)
    target = list(range(len(mols)))
    moved = 0
    for members in classes.values():
        shuffled = list(rng.permutation(members))
        for m, into in zip(members, shuffled):
            target[m] = into
            moved += (m != into)
    assert moved > len(mols)//2, "too few exchangeable molecules"
    atom_index = np.empty(system.getNumParticles(), dtype=np.int32)
    for m, atoms in enumerate(mols):
        for slot, atom in zip(mols[target[m]], atoms):
            atom_index[slot] = atom
    assert np.array_equal(np.sort(atom_index), np.arange(system.getNumParticles()))
    return atom_index


def evaluate_reordered(system, positions, atom_index, precision, groups, box=None, use_index=True):
    """Energy and forces (System order) of a Context that holds the particles in the order atom_index."""
    context = ih.Context(system, {"Precision": precision})
    index = DeviceArray(atom_index)
    try:
        if box is not None:
            context.setPeriodicBoxVectors(*box)
        context.setPositions(positions[atom_index])
        if use_index:
            context._set_atom_index(index.address)
        out = []
        for g in groups:
            state = context.getState(groups=g)
            forces = np.empty_like(state.getForces())
            forces[atom_index] = state.getForces()
            out.append((state.getPotentialEnergy(), forces))
        return out
    finally:
        context.close()
        index.free()


@pytest.mark.gpu
def test_golden_methanol_box_reordered():
    # TestInteraxExportValues.h:188-198 at the InteraxOrder2 FP64 tolerances, with whole methanol molecules exchanged
    values = load_export_values("methanol_box")
    system = ih.load_system_xml(system_path("methanol_box"))
    atom_index = molecule_permutation(system, seed=3)
    print(f"methanol box: 840 atoms, {int((atom_index != np.arange(840)).sum())} in another slot", flush=True)
    groups = [1 << g for g in range(len(values.terms))]
    results = evaluate_reordered(system, values.positions, atom_index, "double", groups, values.box)
    for (energy, forces), term, ref_e, ref_f in zip(results, values.terms, values.energies, values.forces):
        rms = np.sqrt(np.mean(ref_f**2))
        assert abs(energy-ref_e) <= 1e-8*abs(ref_e), term
        assert np.abs(forces-ref_f).max() <= 1e-7*rms, term


def near_equal_width_box():
    """The 648-atom water-like box with per-element parameters (identical molecules) and O/H Slater widths 0.0200 and
    0.0202 nm: inside the InteraxOrder2 band, so O-H pairs are expanded about the lower-index particle's width."""
    _, positions = workloads.config0()
    L = 1.862
    system = ih.System(len(positions), [L, 0, 0, 0, L, 0, 0, 0, L])
    disp, exch, sr = ih.InteraxDispersionForce(), ih.InteraxExchangeForce(), ih.InteraxShortRangeForce()
    widths = {"O": 0.0200, "H": 0.0202}
    for i in range(len(positions)):
        el = "O" if i % 3 == 0 else "H"
        disp.addParticle(*workloads.dispersion_moments(el), 1/37.0)
        exch.addParticle(1.5*abs(workloads.QV[el]), widths[el])
        sr.addParticle(-170*abs(workloads.QV[el]), 900*abs(workloads.QV[el]), widths[el])
    for f in (disp, exch, sr):
        f.setNonbondedMethod(ih.InteraxPairForce.CutoffPeriodic)
        f.setCutoffDistance(0.9)
        for m in range(len(positions)//3):
            f.addExclusion(3*m, 3*m+1)
            f.addExclusion(3*m, 3*m+2)
            f.addExclusion(3*m+1, 3*m+2)
        system.addForce(f)
    exch.setOverlapMethod(ih.InteraxExchangeForce.InteraxOrder2)
    sr.setOverlapMethod(ih.InteraxShortRangeForce.InteraxOrder2)
    for k, f in enumerate((disp, exch, sr)):
        f.setForceGroup(k)
    return system, positions


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_order2_rule_follows_atom_index(precision):
    system, positions = near_equal_width_box()
    box = system.getDefaultPeriodicBoxVectors()
    atom_index = molecule_permutation(system, seed=5)
    print(f"near-equal widths: 648 atoms, {precision}", flush=True)
    groups = [2, 4]     # exchange, short range: the forces with an overlap
    results = evaluate_reordered(system, positions, atom_index, precision, groups)
    ignored = evaluate_reordered(system, positions, atom_index, "double", groups, use_index=False)
    for (energy, forces), (energy_id, _), g in zip(results, ignored, groups):
        e_ref, f_ref = oracle.evaluate(system, positions, box, groups=g)
        fmax = np.abs(f_ref).max()
        if precision == "double":
            assert abs(energy-e_ref) <= 1e-10*abs(e_ref)
            assert np.abs(forces-f_ref).max() <= 1e-9*fmax+1e-8
        else:
            assert abs(energy-e_ref) <= 2e-6*abs(e_ref)
            assert np.linalg.norm(forces-f_ref)/np.linalg.norm(f_ref) <= 1e-5
        # slot indices in place of System indices give another expansion point for the O-H pairs
        assert abs(energy_id-e_ref) > 1e-8*abs(e_ref), (energy_id, e_ref)
