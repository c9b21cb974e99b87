"""The three-body force on the HIP platform against the oracle: the cases of the reference's three-body tests
(TestInteraxThreeBodyForce.h:185-617) restated, in double and mixed precision.

Tolerances are the reference's (TestInteraxThreeBodyForce.h:608-617): double precision energy 1e-10 of the sum of
the absolute triple energies and forces 1e-8 of the largest force; single/mixed precision 1e-5 and 1e-4; plus the
1e-8 kJ/mol/nm resolution of fixed-point forces.
"""
import math

import numpy as np
import pytest

import independent as ind
import interax_hip as ih
import oracle
import workloads

PRECISIONS = ["double", "mixed", "single"]
TOL = {"double": (1e-10, 1e-8), "mixed": (1e-5, 1e-4), "single": (1e-5, 1e-4)}
RESOLUTION = 1e-8
# InteraxFF molecular site parameters (TestInteraxThreeBodyForce.h:172-173)
SITE_TYPES = [(0.004087456746801944, 911.0681141195369, 0.03), (0.0016246535729804365, 1262.664320774599, 0.025),
              (0.008817709933796472, 747.5256543931184, 0.035)]


def single_particle_sites(n, box=None):
    system = ih.System(n, None if box is None else [box, 0, 0, 0, box, 0, 0, 0, box])
    f = ih.InteraxThreeBodyForce()
    for i in range(n):
        a, w, s = SITE_TYPES[i % 3]
        f.addSite([i], a, w, s)
    system.addForce(f)
    return system, f


def compare(context, system, force, pos, precision, lam=1.0, box=None):
    state = context.getState()
    e_ref, f_ref, abs_sum = oracle.threebody_force(force, pos, box, lam, with_scale=True)
    etol, ftol = TOL[precision]
    # the scale of the energy: the sum of |triple energies| (TestInteraxThreeBodyForce.h:150-165)
    abs_scale = max(abs_sum, 1e-300)
    assert abs(state.getPotentialEnergy()-e_ref) <= etol*abs_scale+1e-14, (state.getPotentialEnergy(), e_ref)
    fmax = np.abs(f_ref).max()
    assert np.abs(state.getForces()-f_ref).max() <= ftol*fmax+RESOLUTION
    return state, e_ref, f_ref


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_triple_law_of_cosines(precision):
    system, f = single_particle_sites(3)
    pos = np.array([[0.0, 0, 0], [0.4, 0, 0], [0.1, 0.35, 0.05]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    state = context.getState()
    c9 = ind.c9(*SITE_TYPES[0][:2], *SITE_TYPES[1][:2], *SITE_TYPES[2][:2])
    expected = ind.triple_energy(pos[0], pos[1], pos[2], c9)
    assert state.getPotentialEnergy() == pytest.approx(expected, rel=TOL[precision][0])
    fd = ind.finite_difference_forces(lambda p: ind.triple_energy(p[0], p[1], p[2], c9), pos, h=1e-6)
    assert np.abs(state.getForces()-fd).max() <= max(TOL[precision][1], 1e-7)*np.abs(fd).max()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_centroid_sites(precision):
    rng = np.random.default_rng(3)
    f = ih.InteraxThreeBodyForce()
    pos = []
    for i in range(4):
        center = np.array([0.4*math.cos(2*math.pi*i/4), 0.4*math.sin(2*math.pi*i/4), 0.1*i])
        parts = []
        for j in range(i+1):
            parts.append(len(pos))
            pos.append(center+(rng.uniform(size=3)-0.5)*0.15)
        f.addSite(parts, 0.004+0.001*i, 900.0-50*i)
    system = ih.System(len(pos))
    system.addForce(f)
    pos = np.array(pos)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, f, pos, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_molecules(precision):
    system, f = single_particle_sites(5)
    for i, m in enumerate([0, 0, -1, 3, -1]):
        s = f.getSiteParameters(i)
        f.setSiteParameters(i, s[0], s[1], s[2], s[3], m, s[5])
    pos = np.array([[0, 0, 0], [0.15, 0, 0], [0, 0.4, 0], [0.3, 0.3, 0.2], [-0.2, 0.1, 0.35]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    _, e1, _ = compare(context, system, f, pos, precision)
    assert context.triple_counts() == [7]
    s = f.getSiteParameters(1)
    f.setSiteParameters(1, s[0], s[1], s[2], s[3], 1, s[5])
    context2 = ih.Context(system, {"Precision": precision})
    context2.setPositions(pos)
    _, e2, _ = compare(context2, system, f, pos, precision)
    assert context2.triple_counts() == [10]
    assert abs(e2-e1) > 1e-3*abs(e1)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_cutoff(precision):
    system, f = single_particle_sites(4)
    f.setCutoffDistance(0.5)
    pos = np.array([[0, 0, 0], [0.4, 0, 0], [0.2, 0.3, 0], [0.2, 0.75, 0]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    state, _, _ = compare(context, system, f, pos, precision)
    assert np.all(state.getForces()[3] == 0)
    assert context.triple_counts() == [1]


def random_periodic(n, box, cutoff, seed):
    rng = np.random.default_rng(seed)
    system, f = single_particle_sites(n, box)
    for i in range(n):
        s = f.getSiteParameters(i)
        f.setSiteParameters(i, s[0], s[1], s[2], s[3], i//2, s[5])
    f.setCutoffDistance(cutoff)
    f.setUsesPeriodicBoundaryConditions(True)
    vectors = np.diag([box]*3)
    pos = []
    while len(pos) < n:
        # spread over several periodic copies of the box, at least 0.25 nm apart
        p = rng.uniform(size=3)*box*3-box
        if all(np.linalg.norm(ind.minimum_image(p-q, vectors)) > 0.25 for q in pos):
            pos.append(p)
    return system, f, np.array(pos), vectors


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("box", [3.5, 2.2])
def test_periodic(box, precision):
    # a cutoff larger than half the box: the three minimum images need not close into a triangle
    system, f, pos, vectors = random_periodic(40, box, 1.2, 4)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, f, pos, precision, box=vectors)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("method,radius", [(0, 0.0), (0, 0.3), (1, 0.0), (1, 0.3)])
def test_damping(method, radius, precision):
    system, f = single_particle_sites(4)
    f.setDampingMethod(method)
    f.setDampingScale(0.8)
    f.setCentroidDampingRadius(radius)
    pos = np.array([[0, 0, 0], [0.12, 0.02, 0], [0.05, 0.18, 0.03], [0.2, 0.2, 0.25]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, f, pos, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_alchemical(precision):
    system, f = single_particle_sites(5)
    for i in (0, 1):
        s = f.getSiteParameters(i)
        f.setSiteParameters(i, s[0], s[1], s[2], s[3], s[4], True)
    f.setAlchemicalParameterName("lambda_test")
    pos = np.array([[0, 0, 0], [0.35, 0, 0], [0, 0.4, 0], [0.3, 0.3, 0.2], [-0.2, 0.1, 0.35]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    assert context.getParameter("lambda_test") == 1.0
    for lam in (1.0, 0.0, 0.25, 0.9):
        context.setParameter("lambda_test", lam)
        compare(context, system, f, pos, precision, lam=lam)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_alchemical_overlap(precision):
    # coincident alchemical and ordinary sites at lambda = 0: exactly zero, finite forces
    system, f = single_particle_sites(5)
    for i in (0, 1):
        s = f.getSiteParameters(i)
        f.setSiteParameters(i, s[0], s[1], s[2], s[3], s[4], True)
    pos = np.array([[0, 0.4, 1e-9], [0.35, 0, 0], [0, 0.4, 0], [0.3, 0.3, 0.2], [-0.2, 0.1, 0.35]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    context.setParameter("lambda_sterics", 0.0)
    state = context.getState()
    e_ref, _ = oracle.threebody_force(f, pos, None, 0.0)
    assert state.getPotentialEnergy() == pytest.approx(e_ref, rel=TOL[precision][0])
    assert np.all(np.isfinite(state.getForces()))
    assert np.all(state.getForces()[:2] == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_large_system(precision):
    n_mol, L = 50, 2.4
    rng = np.random.default_rng(6)
    vectors = np.diag([L]*3)
    centers = []
    while len(centers) < n_mol:
        p = rng.uniform(size=3)*L
        if all(np.linalg.norm(ind.minimum_image(p-q, vectors)) > 0.3 for q in centers):
            centers.append(p)
    f = ih.InteraxThreeBodyForce()
    f.setCutoffDistance(1.0)
    f.setUsesPeriodicBoundaryConditions(True)
    f.setCentroidDampingRadius(0.25)
    f.setScaleFactor(1.3)
    pos = []
    for i in range(n_mol):
        parts = []
        for j in range(3):
            parts.append(len(pos))
            pos.append(centers[i]+(rng.uniform(size=3)-0.5)*0.1)
        f.addSite(parts, 0.003+0.002*(i % 3), 800.0+150*(i % 4), 0.03, i, i < 3)
    system = ih.System(len(pos), [L, 0, 0, 0, L, 0, 0, 0, L])
    system.addForce(f)
    pos = np.array(pos)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    context.setParameter("lambda_sterics", 0.6)
    compare(context, system, f, pos, precision, lam=0.6, box=vectors)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_changing_parameters(precision):
    system, f = single_particle_sites(4)
    f.setDampingMethod(1)
    f.setCentroidDampingRadius(0.2)
    pos = np.array([[0, 0, 0], [0.25, 0.02, 0], [0.05, 0.28, 0.03], [0.2, 0.2, 0.25]])
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    e1 = context.getState().getPotentialEnergy()
    f.setSiteParameters(1, [1], 0.006, 1000.0, 0.028, -1, False)
    f.setScaleFactor(0.9)
    f.setDampingScale(1.2)
    f.setCentroidDampingRadius(0.25)
    context.updateParametersInContext(f)
    state, e2, _ = compare(context, system, f, pos, precision)
    assert abs(e1-state.getPotentialEnergy()) > 1e-3*abs(e1)


@pytest.mark.gpu
def test_illegal_settings():
    system, f = single_particle_sites(3)
    f.addSite([], 0.004, 900.0)
    with pytest.raises(ih.InteraxError, match="at least one particle"):
        ih.Context(system)
    system, f = single_particle_sites(3)
    f.addSite([3], 0.004, 900.0)
    with pytest.raises(ih.InteraxError, match="Illegal particle index for a site"):
        ih.Context(system)
    system, f = single_particle_sites(3)
    f.setDampingMethod(1)
    f.setSiteParameters(2, [2], 0.004, 900.0, 0.0, -1, False)
    with pytest.raises(ih.InteraxError, match="positive width"):
        ih.Context(system)
    system, f = single_particle_sites(3)
    f.setCutoffDistance(0.0)
    with pytest.raises(ih.InteraxError, match="cutoff distance must be positive"):
        ih.Context(system)


@pytest.mark.gpu
def test_atm_fluid_against_oracle():
    # the triplet configuration at reduced size (4096 sites, same density and cutoff), mixed precision
    system, pos = workloads.atm_fluid(4096, seed=1)
    f = system.getForce(0)
    box = system.getDefaultPeriodicBoxVectors()
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    state = context.getState()
    e_ref, f_ref = oracle.threebody_force(f, pos, box)
    assert context.triple_counts() == [oracle.threebody_count(f, pos, box)]
    assert np.linalg.norm(state.getForces()-f_ref)/np.linalg.norm(f_ref) <= 1e-5
    assert state.getPotentialEnergy() == pytest.approx(e_ref, rel=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_nonperiodic_cell_grid(precision):
    # more than 4096 sites without periodic boundaries: the neighbour search runs on a grid over the bounding box;
    # membership and triples must match the oracle's all-pairs search
    rng = np.random.default_rng(7)
    n = 5000
    pos = workloads.rsa_points(rng, n, 6.0, 0.25, periodic=False)
    system, f = single_particle_sites(n)
    f.setCutoffDistance(0.55)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, f, pos, precision)
    assert context.triple_counts() == [oracle.threebody_count(f, pos)]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", PRECISIONS)
def test_lists_longer_than_the_first_capacity(precision):
    # 150 sites all within the cutoff of each other: the first site's list of later neighbours (149 entries) is
    # longer than the first capacity (64), so the search and the triple kernel, launched before the longest list is
    # known, run again with more room before the site forces are applied; the next evaluation keeps that room
    rng = np.random.default_rng(21)
    n = 150
    pos = workloads.rsa_points(rng, n, 1.6, 0.2, periodic=False)
    system, f = single_particle_sites(n)
    f.setCutoffDistance(5.0)
    context = ih.Context(system, {"Precision": precision})
    context.setPositions(pos)
    compare(context, system, f, pos, precision)
    moved = pos+rng.normal(0.0, 0.01, size=pos.shape)
    context.setPositions(moved)
    compare(context, system, f, moved, precision)


@pytest.mark.gpu
def test_grid_with_one_cell_along_an_axis():
    # a periodic box that is wide along x and y but thinner than three cutoffs along z: one cell along z
    rng = np.random.default_rng(3)
    lx, lz = 5.0, 1.9
    grid = np.stack(np.meshgrid(np.arange(20), np.arange(20), np.arange(7), indexing="ij"), -1).reshape(-1, 3)
    pos = (grid+0.5)*np.array([lx/20, lx/20, lz/7])+rng.uniform(-0.04, 0.04, size=grid.shape)
    n = len(pos)
    system = ih.System(n, [lx, 0, 0, 0, lx, 0, 0, 0, lz])
    f = ih.InteraxThreeBodyForce()
    for i in range(n):
        a, w, s = SITE_TYPES[i % 3]
        f.addSite([i], a, w, s)
    f.setCutoffDistance(0.7)
    f.setUsesPeriodicBoundaryConditions(True)
    system.addForce(f)
    box = system.getDefaultPeriodicBoxVectors()
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    compare(context, system, f, pos, "mixed", box=box)
    assert context.triple_counts() == [oracle.threebody_count(f, pos, box)]
