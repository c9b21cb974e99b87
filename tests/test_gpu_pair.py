"""The pair forces on the HIP platform against the oracle: the cases of the reference's shared pair-force tests
(TestInteraxPairForce.h:167-695, TestInteraxShortRangeForce.h:127-210) restated, in double and mixed precision.

Tolerances are the reference's (TestInteraxPairForce.h:719-730): double precision energy 1e-10 (relative) and forces
1e-9 of the largest force; single/mixed precision 2e-5 and 2e-4; plus the 1e-8 kJ/mol/nm resolution of 64-bit
fixed-point forces (InteraxTestUtilities.h:104-115).  The mixed precision path must additionally keep the force RMS
error at or below 1e-5 of the RMS force on the systems of the benchmark.
"""
import math

import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads

PRECISIONS = ["double", "mixed", "single"]
TOL = {"double": (1e-10, 1e-9), "mixed": (2e-5, 2e-4), "single": (2e-5, 2e-4)}
RESOLUTION = 1e-8

# particle types of the reference's test models (TestInteraxDispersionForce.h:55-62, TestInteraxExchangeForce.h:52-57,
# TestInteraxShortRangeForce.h:51-56)
DISPERSION_TYPES = [
    (2.0412491330659432, 0.07991696226413432, 0.003917538949072952, 0.0020393931939312773, 0.026844933830306183),
    (2.179116829570064, 0.05077216863583387, 0.001421435052885423, 0.0012613225562556841, 0.021090413137365274),
    (0.47385034770767653, 0.0062664158577877644, 0.00015411387998271836, 0.00023498508736915282, 0.020351567997950105),
    (3.9405044694974447, 0.18787610923101286, 0.010735264902054127, 0.004221894603968141, 0.03096091246647922)]
EXCHANGE_TYPES = [(10.326424719450545, 0.0248302193400194), (1.0254403205662432, 0.019763296864870033),
                  (13.231601566169132, 0.02761867008817038), (1.5623921585722147, 0.021706271973167678)]
SHORT_RANGE_TYPES = [(-1102.281886852146, 5953.779066676866, 0.017966632940037853),
                     (1236.5568512540217, 80470.88690318535, 0.026044362132579005),
                     (-1579.591012686484, 70439.46596886561, 0.027056173026181088),
                     (-1702.373363056105, 47202.476121472, 0.01880253891316055)]
MODELS = ["dispersion", "exchange", "short_range"]


def make_force(model):
    if model == "dispersion":
        f = ih.InteraxDispersionForce()
        f.setDampingScale(0.7155241370201111)
    elif model == "exchange":
        f = ih.InteraxExchangeForce()
    else:
        f = ih.InteraxShortRangeForce()
    return f


def add_particle(force, model, t, alchemical=False):
    if model == "dispersion":
        force.addParticle(*DISPERSION_TYPES[t], alchemical)
    elif model == "exchange":
        force.addParticle(*EXCHANGE_TYPES[t], alchemical)
    else:
        force.addParticle(*SHORT_RANGE_TYPES[t], alchemical)


def create_system(model, types, alchemical=None, box=(2.0, 2.0, 2.0)):
    system = ih.System(len(types), [box[0], 0, 0, 0, box[1], 0, 0, 0, box[2]])
    force = make_force(model)
    for i, t in enumerate(types):
        add_particle(force, model, t, bool(alchemical[i]) if alchemical is not None else False)
    system.addForce(force)
    return system, force


def compare(context, system, positions, precision, lam=1.0, groups=0xFFFFFFFF, box=None, scale_energy=None):
    state = context.getState(groups=groups)
    params = {}
    for f in system.getForces():
        params[f.getAlchemicalParameterName()] = lam
    e_ref, f_ref, abs_sum = oracle.evaluate(system, positions, box, groups=groups, parameters=params, with_scale=True)
    etol, ftol = TOL[precision]
    # relative to the sum of |pair energies| before switching, the scale of rounding errors, as
    # TestInteraxPairForce.h:137-143 compares (or to a given larger scale)
    escale = abs_sum if scale_energy is None else max(abs_sum, scale_energy)
    assert abs(state.getPotentialEnergy()-e_ref) <= etol*max(escale, 1e-300)+1e-12, (state.getPotentialEnergy(), e_ref)
    fmax = np.abs(f_ref).max() if f_ref.size else 0.0
    err = np.abs(state.getForces()-f_ref).max() if f_ref.size else 0.0
    assert err <= ftol*fmax+RESOLUTION, (err, fmax)
    return state, e_ref, f_ref


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_pair_energies(model, precision):
    direction = np.array([0.3, -0.5, 0.8])/math.sqrt(0.98)
    for t1 in range(4):
        for t2 in range(t1, 4):
            system, force = create_system(model, [t1, t2])
            context = ih.Context(system, {"Precision": precision})
            for r in (0.06, 0.1, 0.15, 0.2, 0.25, 0.3, 0.4, 0.5, 0.7):
                pos = np.array([[0.1, 0.2, -0.1], [0.1, 0.2, -0.1]])+np.outer([0, 1], direction*r)
                context.setPositions(pos)
                compare(context, system, pos, precision)
            context.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_exclusions(model, precision):
    system, force = create_system(model, [0, 1, 2, 3, 0])
    force.addExclusion(0, 1)
    force.addExclusion(3, 2)
    force.addExclusion(4, 0)
    pos = np.array([[0, 0, 0], [0.25, 0, 0], [0, 0.3, 0], [0.1, 0.1, 0.2], [-0.2, 0.05, 0.1]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, pos, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_cutoff(model, precision):
    system, force = create_system(model, [0, 1, 2])
    force.setNonbondedMethod(ih.InteraxPairForce.CutoffNonPeriodic)
    force.setCutoffDistance(0.35)
    context = ih.Context(system, {"Precision": precision})
    pos = np.array([[0, 0, 0], [0.2, 0, 0], [0.45, 0, 0]])
    context.setPositions(pos)
    compare(context, system, pos, precision)
    pos = np.array([[0, 0, 0], [0.36, 0, 0], [0, 0.36, 0.36]])
    context.setPositions(pos)
    state = context.getState()
    assert state.getPotentialEnergy() == 0.0
    assert not np.any(state.getForces())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_periodic(model, precision):
    system, force = create_system(model, [0, 1, 3], box=(1.2, 1.1, 1.3))
    force.setNonbondedMethod(ih.InteraxPairForce.CutoffPeriodic)
    force.setCutoffDistance(0.5)
    context = ih.Context(system, {"Precision": precision})
    pos = np.array([[0.05, 0.5, 0.6], [1.1, 0.55, 0.6], [0.1, 1.05, 2.6+0.05]])
    context.setPositions(pos)
    compare(context, system, pos, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_triclinic(model, precision):
    vectors = np.array([[2.0, 0, 0], [0.4, 2.1, 0], [-0.5, 0.7, 1.9]])
    n = 50
    rng = np.random.default_rng(5)
    pos = []
    while len(pos) < n:
        # spread over several periodic copies of the box, at least 0.12 nm apart
        p = (3*rng.uniform(size=3)-1) @ vectors
        if all(np.linalg.norm(_min_image(p-q, vectors)) >= 0.12 for q in pos):
            pos.append(p)
    pos = np.array(pos)
    system, force = create_system(model, [i % 4 for i in range(n)])
    system.setDefaultPeriodicBoxVectors(*vectors)
    force.setNonbondedMethod(ih.InteraxPairForce.CutoffPeriodic)
    force.setCutoffDistance(0.75)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, pos, precision, box=vectors)


def _min_image(d, vectors):
    import independent
    return independent.minimum_image(d, vectors)


@pytest.mark.gpu
@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_switching_function(model, precision, method):
    system, force = create_system(model, [0, 2])
    force.setNonbondedMethod(method)
    force.setCutoffDistance(0.5)
    force.setUseSwitchingFunction(True)
    force.setSwitchingDistance(0.3)
    context = ih.Context(system, {"Precision": precision})
    for r in np.arange(0.25, 0.55, 0.02):
        pos = np.array([[0, 0, 0], [0, r, 0]])
        context.setPositions(pos)
        # near the cutoff S(r) is small: errors are relative to the unswitched energy (TestInteraxPairForce.h:331-333)
        f2 = make_force(model)
        add_particle(f2, model, 0)
        add_particle(f2, model, 2)
        e_unswitched, _ = oracle.pair_force(f2, pos)
        compare(context, system, pos, precision, scale_energy=abs(e_unswitched))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_alchemical(model, precision):
    system, force = create_system(model, [0, 1, 2, 3, 1], [True, True, False, False, False])
    force.setAlchemicalParameterName("lambda_test")
    pos = np.array([[0, 0, 0], [0.2, 0, 0], [0, 0.25, 0], [0.15, 0.2, 0.15], [-0.1, 0.1, 0.25]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    assert context.getParameter("lambda_test") == 1.0
    compare(context, system, pos, precision)
    for lam in (0.0, 0.3, 0.8):
        context.setParameter("lambda_test", lam)
        compare(context, system, pos, precision, lam=lam, scale_energy=1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_softcore(model, precision):
    system, force = create_system(model, [0, 1, 2, 3], [True, True, False, False])
    force.setSoftcoreRadius(0.2)
    force.setNonbondedMethod(ih.InteraxPairForce.CutoffNonPeriodic)
    force.setCutoffDistance(0.6)
    force.setUseSwitchingFunction(True)
    force.setSwitchingDistance(0.4)
    pos = np.array([[0, 0, 0], [0.18, 0, 0], [0.02, 0.1, 0], [0.3, 0.25, 0.15]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    for lam in (1.0, 0.7, 0.2, 0.0):
        context.setParameter(force.getAlchemicalParameterName(), lam)
        compare(context, system, pos, precision, lam=lam)


def large_system(model, seed=0):
    n_mol, L = 60, 2.3
    rng = np.random.default_rng(seed)
    types, alch = [], []
    for i in range(n_mol):
        for j in range(3):
            types.append((i+j) % 4)
            alch.append(i < 4)
    system, force = create_system(model, types, alch, box=(L, L, L))
    force.setNonbondedMethod(ih.InteraxPairForce.CutoffPeriodic)
    force.setCutoffDistance(0.9)
    force.setUseSwitchingFunction(True)
    force.setSwitchingDistance(0.75)
    force.setSoftcoreRadius(0.15)
    box = np.diag([L, L, L])
    pos = []
    for i in range(n_mol):
        while True:
            o = rng.uniform(0, L, size=3)
            mol = [o, o+[0.1, 0, 0], o+[0, 0.1, 0]]
            if all(np.linalg.norm(_min_image(p-q, box)) >= 0.2 for p in pos for q in mol):
                break
        pos.extend(mol)
        for j in range(3):
            for k in range(j):
                force.addExclusion(3*i+k, 3*i+j)
    return system, force, np.array(pos)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_large_system(model, precision):
    system, force, pos = large_system(model)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    for lam in (1.0, 0.4):
        context.setParameter(force.getAlchemicalParameterName(), lam)
        state, e_ref, f_ref = compare(context, system, pos, precision, lam=lam)
        rms = np.linalg.norm(state.getForces()-f_ref)/np.linalg.norm(f_ref)
        assert rms <= (1e-10 if precision == "double" else 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", MODELS)
def test_changing_parameters(model, precision):
    system, force, pos = large_system(model, seed=1)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, pos, precision)
    # TestInteraxPairForce.h:484-512 with the models' changeParameters()
    if model == "dispersion":
        force.setParticleParameters(1, 2.5, 0.06, 0.002, 0.0015, 0.023, force.getParticleIsAlchemical(1))
        force.setDampingScale(0.8)
    elif model == "exchange":
        force.setParticleParameters(1, 2.0, 0.023, force.getParticleIsAlchemical(1))
        force.setParticleParameters(3, 5.0, 0.026, force.getParticleIsAlchemical(3))
    else:
        force.setParticleParameters(1, 900.0, 50000.0, 0.024, force.getParticleIsAlchemical(1))
        force.setParticleParameters(3, -1000.0, 30000.0, 0.02, force.getParticleIsAlchemical(3))
    context.updateParametersInContext(force)
    compare(context, system, pos, precision)
    if model != "dispersion":
        force.setOverlapMethod(2)
        with pytest.raises(ih.InteraxError, match="overlap method has changed"):
            context.updateParametersInContext(force)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_combined_pair_forces(precision):
    # TestInteraxShortRangeForce.h:127-210: the three forces in one group equal the sum of separate groups, and each
    # group of the fused handle equals a System with that force alone
    rng = np.random.default_rng(2)
    n_mol, L = 40, 1.8
    box = np.diag([L, L, L])
    pos = []
    while len(pos) < 2*n_mol:
        o = rng.uniform(0, L, size=3)
        if all(np.linalg.norm(_min_image(p-q, box)) > 0.2 for p in pos for q in (o, o+[0.1, 0, 0])):
            pos.extend([o, o+[0.1, 0, 0]])
    pos = np.array(pos)

    def create(same_group, only=None):
        system = ih.System(2*n_mol, [L, 0, 0, 0, L, 0, 0, 0, L])
        forces = [ih.InteraxDispersionForce(), ih.InteraxExchangeForce(), ih.InteraxShortRangeForce()]
        for i in range(n_mol):
            alch = (i == 0)
            forces[0].addParticle(*DISPERSION_TYPES[0], alch)
            forces[0].addParticle(*DISPERSION_TYPES[2], alch)
            forces[1].addParticle(*EXCHANGE_TYPES[0], alch)
            forces[1].addParticle(*EXCHANGE_TYPES[1], alch)
            forces[2].addParticle(*SHORT_RANGE_TYPES[0], alch)
            forces[2].addParticle(*SHORT_RANGE_TYPES[3], alch)
        for k, f in enumerate(forces):
            for i in range(n_mol):
                f.addExclusion(2*i, 2*i+1)
            f.setNonbondedMethod(2)
            f.setCutoffDistance(0.8)
            f.setUseSwitchingFunction(True)
            f.setSwitchingDistance(0.7)
            f.setSoftcoreRadius(0.1)
            f.setForceGroup(0 if same_group else k)
            if only is None or only == k:
                system.addForce(f)
        forces[0].setUseLongRangeCorrection(True)
        return system

    separate, combined = create(False), create(True)
    c1 = ih.Context(separate, {"Precision": precision})
    c2 = ih.Context(combined, {"Precision": precision})
    assert len(c1._groups) == 1 and len(c2._groups) == 1     # fused into one kernel handle
    for c in (c1, c2):
        c.setPositions(pos)
        c.setParameter("lambda_sterics", 0.6)
    s1, s2 = c1.getState(), c2.getState()
    etol, ftol = TOL[precision]
    assert s1.getPotentialEnergy() == pytest.approx(s2.getPotentialEnergy(), rel=etol)
    assert np.abs(s1.getForces()-s2.getForces()).max() <= ftol*np.abs(s1.getForces()).max()+RESOLUTION
    for k in range(3):
        single = create(False, only=k)
        c = ih.Context(single, {"Precision": precision})
        c.setPositions(pos)
        c.setParameter("lambda_sterics", 0.6)
        expected = c.getState()
        found = c1.getState(groups=1 << k)
        assert found.getPotentialEnergy() == pytest.approx(expected.getPotentialEnergy(), rel=etol, abs=1e-9)
        assert np.abs(found.getForces()-expected.getForces()).max() <= ftol*np.abs(expected.getForces()).max()+RESOLUTION
        compare(c1, separate, pos, precision, lam=0.6, groups=1 << k)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("model", ["exchange", "short_range"])
@pytest.mark.parametrize("method", [0, 2])
def test_overlap_methods(model, method, precision):
    # TestInteraxPairForce.h:579-630: nearly equal and different widths, either particle the narrower one
    double = precision == "double"
    width = 0.02
    for difference in (0.0, 2e-4, 1e-3, 0.005, 0.01, 0.05, 0.12, 0.3, 0.5):
        for order in (0, 1):
            widths = [width, width*(1+difference)]
            force = make_force(model)
            for w in (widths[order], widths[1-order]):
                if model == "exchange":
                    force.addParticle(3.0, w)
                else:
                    force.addParticle(-1500.0, 60000.0, w)
            force.setOverlapMethod(method)
            system = ih.System(2)
            system.addForce(force)
            context = ih.Context(system, {"Precision": precision})
            for r in (0.01, 0.03, 0.06, 0.1, 0.15, 0.2, 0.3):
                if not double and (method == 2 or r <= 0.025):
                    continue   # not tested by the reference in single precision (truncation of the export scheme)
                pos = np.array([[0, 0, 0], [0, 0, r]])
                context.setPositions(pos)
                state = context.getState()
                e_ref, f_ref = oracle.pair_force(force, pos, double_precision=double)
                tol = 1e-9 if double else 3e-5
                assert state.getPotentialEnergy() == pytest.approx(e_ref, rel=tol), (difference, order, r)
                assert abs(state.getForces()[1, 2]-f_ref[1, 2]) <= 10*tol*abs(f_ref[1, 2])+RESOLUTION
            context.close()


@pytest.mark.gpu
def test_illegal_settings_rejected():
    system, force = create_system("exchange", [0, 1])
    system.addParticle()
    with pytest.raises(ih.InteraxError, match="exactly as many particles"):
        ih.Context(system)
    system, force = create_system("exchange", [0, 1])
    force.addExclusion(0, 2)
    with pytest.raises(ih.InteraxError, match="Illegal particle index for an exclusion"):
        ih.Context(system)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_water_box_648(precision):
    # the plain-method configuration: 216 rigid water-like molecules, cutoff 0.9 nm, FP64 against the oracle at the
    # double precision tolerances; mixed precision at a force RMS error <= 1e-5
    system, pos = workloads.config0()
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    box = system.getDefaultPeriodicBoxVectors()
    for group in (1,):
        state = context.getState(groups=0xFFFFFFFF)
        e_ref, f_ref = oracle.evaluate(system, pos, box)
        rms = np.linalg.norm(state.getForces()-f_ref)/np.linalg.norm(f_ref)
        if precision == "double":
            assert abs(state.getPotentialEnergy()-e_ref) <= 1e-10*abs(e_ref)
            assert np.abs(state.getForces()-f_ref).max() <= 1e-9*np.abs(f_ref).max()+RESOLUTION
        else:
            assert rms <= 1e-5
            assert abs(state.getPotentialEnergy()-e_ref) <= 2e-6*abs(e_ref)
        counts = context.pair_counts()
        assert counts == [oracle.pair_count(system.getForce(1), pos, box)]


@pytest.mark.gpu
def test_bitwise_reproducible():
    system, pos = workloads.config0()
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    a = context.getState()
    b = context.getState()
    assert a.getPotentialEnergy() == b.getPotentialEnergy()
    assert np.array_equal(a.getForces(), b.getForces())
