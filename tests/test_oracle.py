"""Pinning of the FP64 oracle (CPU only).

Every oracle function is checked against something it does not share code with: the reference's golden vectors,
its 50-digit and closed-form test values, library routines (mpmath, scipy, numpy), an independent brute-force
evaluation (tests/independent.py) and finite differences of the energy.
"""
import math
import os
import subprocess
import sys

import mpmath
import numpy as np
import pytest
from scipy.special import gammainc

import independent as ind
import interax_hip as ih
import oracle
from golden_data import load_export_values, system_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------------------------------------------------------
# golden vectors (TestInteraxExportValues.h:150-198)
# ---------------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["cluster", "methanol_box"])
@pytest.mark.parametrize("exact", [False, True])
def test_golden_vectors(name, exact):
    values = load_export_values(name)
    system = ih.load_system_xml(system_path(name))
    if exact:
        for f in system.getForces():
            if hasattr(f, "setOverlapMethod"):
                f.setOverlapMethod(0)
    energy_tol, force_tol = (1e-6, 1e-5) if exact else (1e-8, 1e-7)
    box = values.box if values.periodic else None
    for group, term in enumerate(values.terms):
        e, f = oracle.evaluate(system, values.positions, box, groups=1 << group)
        ref = values.forces[group]
        rms = np.sqrt(np.mean(ref**2))
        assert abs(e-values.energies[group]) <= energy_tol*abs(values.energies[group]), term
        assert np.abs(f-ref).max() <= force_tol*rms, term


# ---------------------------------------------------------------------------------------------------------------
# Slater overlap (InteraxPairTerms.h:167-216)
# ---------------------------------------------------------------------------------------------------------------

def test_overlap_equal_widths_closed_form():
    # S = exp(-x) (1 + x + x^2/3) / (64 pi a^3) for a = b (TestInteraxExchangeForce.h:110-128)
    a = 0.025
    for r in (1e-6, 0.02, 0.1, 0.3, 1.0):
        x = r/a
        expected = math.exp(-x)*(1+x+x*x/3)/(64*math.pi*a**3)
        dexpected = -math.exp(-x)*x*(1+x)/3/(64*math.pi*a**4)
        s, ds = oracle.slater_overlap(r, a, a)
        assert s == pytest.approx(expected, rel=1e-13)
        assert ds == pytest.approx(dexpected, rel=1e-12, abs=1e-30)


def test_overlap_fifty_digit_value():
    # TestInteraxExchangeForce.h:98: the closed form at 50 digits for nearly equal widths
    s, _ = oracle.slater_overlap(0.01, 0.02, 0.02*(1+0.0002))
    assert s == pytest.approx(5371.8136182115909371938804306761/9, rel=1e-13)


def test_overlap_normalization_at_contact():
    # 8 pi (2a)^3 S(0) = 1 (TestInteraxExchangeForce.h:101-104, TestInteraxShortRangeForce.h:87-100)
    a = 0.02
    s, _ = oracle.slater_overlap(1e-12, a, a)
    assert 8*math.pi*(2*a)**3*s == pytest.approx(1.0, rel=1e-10)


@pytest.mark.parametrize("r,a,b", [(0.02, 0.02, 0.02), (0.05, 0.02, 0.02), (0.1, 0.02, 0.025), (0.15, 0.0198, 0.0276),
                                   (0.05, 0.03, 0.03003), (0.2, 0.025, 0.0248), (0.01, 0.018, 0.031)])
def test_overlap_against_fourier_quadrature(r, a, b):
    # S(r) = 1/(2 pi^2 r) int_0^inf k sin(kr) / ((1 + k^2 a^2)^2 (1 + k^2 b^2)^2) dk (InteraxTestUtilities.h:277-290)
    f = lambda k: k*mpmath.sin(k*r)/((1+k*k*a*a)**2*(1+k*k*b*b)**2)
    integral = mpmath.quadosc(f, [0, mpmath.inf], omega=r)
    expected = float(integral/(2*mpmath.pi**2*r))
    for (x, y) in ((a, b), (b, a)):
        s, _ = oracle.slater_overlap(r, x, y)
        assert s == pytest.approx(expected, rel=1e-9)


@pytest.mark.parametrize("t", [1e-6, 1e-4, 1e-3, 0.01, 0.019, 0.021, 0.05, 0.2, 0.5, 1.0])
@pytest.mark.parametrize("r", [0.001, 0.02, 0.1, 0.4, 1.0])
def test_overlap_both_branches_against_mpmath(t, r):
    # the Exact scheme in double precision, across its Taylor switch (|t| = 0.02, |t| x = 0.3).  Just outside the
    # switch the FP64 closed form keeps ~1e-9 (the scheme's design accuracy, InteraxExchangeForce.h:67-71)
    a = 0.025
    b = a*(1+t)
    expected = ind.overlap(r, a, b)
    s, ds = oracle.slater_overlap(r, a, b, method=0, double_precision=True)
    assert s == pytest.approx(expected, rel=2e-9)
    dexpected = ind.overlap_derivative(r, a, b)
    # at r far below the widths the closed-form derivative (r dF/dr - F)/r^2 cancels further
    assert ds == pytest.approx(dexpected, rel=1e-7 if r >= 0.01 else 1e-5)


def test_overlap_taylor_table_rederived():
    # The Taylor coefficients follow from the closed form: derive them with sympy and compare the expansion the
    # oracle evaluates (order 6, t = 0.01, a range of x) with the exact rational polynomial.
    script = os.path.join(ROOT, "tests", "golden", "derive_overlap_taylor.py")
    out = subprocess.run([sys.executable, script, "--json"], check=True, capture_output=True, text=True).stdout
    import json
    table = json.loads(out)
    l = 0.02
    for t in (0.001, 0.01, 0.015):
        for x in (0.05, 0.5, 2.0, 6.0):
            value = sum(t**n*sum(float(mpmath.mpf(table[n][k][0])/table[n][k][1])*x**k for k in range(len(table[n])))
                        for n in range(7))
            expected = math.exp(-x)*value/(math.pi*l**3)
            s, _ = oracle.slater_overlap(x*l, l, l*(1+t))
            assert s == pytest.approx(expected, rel=1e-13)


def test_overlap_order2_scheme():
    # InteraxOrder2: expansion to order 2 about the first width below |b - a| < 0.01 bohr, else the closed form
    bohr = 0.0529177210903
    a = 0.025
    for db in (0.002*bohr, -0.009*bohr):
        b = a+db
        t = db/a
        for r in (0.02, 0.1):
            x = r/a
            p0 = 1/64+x/64+x*x/192
            p1 = -3/128-3*x/128-x*x/192+x**3/384
            p2 = 3/128+3*x/128+x*x/768-5*x**3/768+x**4/1280
            expected = math.exp(-x)*(p0+t*p1+t*t*p2)/(math.pi*a**3)
            s, _ = oracle.slater_overlap(r, a, b, method=2)
            assert s == pytest.approx(expected, rel=1e-13)
    s, _ = oracle.slater_overlap(0.1, a, a+0.011*bohr, method=2)
    assert s == pytest.approx(ind.overlap(0.1, a, a+0.011*bohr), rel=1e-9)


# ---------------------------------------------------------------------------------------------------------------
# Tang-Toennies (InteraxPairTerms.h:223-253) and dispersion (InteraxPairTerms.h:267-293)
# ---------------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("x", [0.01, 0.5, 2.0, 3.49, 3.51, 4.49, 4.51, 5.49, 5.51, 8.0, 20.0, 40.0])
def test_tang_toennies_is_regularized_gamma(x):
    f, df = oracle.tang_toennies(x)
    for i, n in enumerate((6, 8, 10)):
        expected = float(mpmath.gammainc(n+1, 0, x, regularized=True))
        assert f[i] == pytest.approx(expected, rel=1e-12, abs=1e-300)
        assert df[i] == pytest.approx(math.exp(-x)*x**n/math.factorial(n), rel=1e-13)


def test_dispersion_coefficients_by_hand():
    # TestInteraxDispersionForce.h:89-111: at r = 2 nm the damping is 1
    pi = [2.0, 0.1, 0.005, 0.002, 0.025]
    pj = [0.5, 0.01, 0.0002, 0.0003, 0.02]
    r = 2.0
    den = 0.002*0.5+0.0003*2.0
    c6 = 0.002*0.0003*2.0*0.5/den
    c8 = 1.5*0.002*0.0003*(2.0*0.01+0.1*0.5)/den
    c10 = (2*0.002*0.0003*(2.0*0.0002+0.005*0.5)+4.2*0.002*0.0003*0.1*0.01)/den
    e, dedr = oracle.dispersion_pair(r, pi, pj, 1.0)
    assert e == pytest.approx(-(c6/r**6+c8/r**8+c10/r**10), rel=1e-13)
    assert dedr == pytest.approx(6*c6/r**7+8*c8/r**9+10*c10/r**11, rel=1e-13)


def test_dispersion_against_independent():
    rng = np.random.default_rng(3)
    for _ in range(20):
        pi = [rng.uniform(0.2, 4), rng.uniform(0.005, 0.2), rng.uniform(1e-4, 1e-2), rng.uniform(1e-4, 4e-3), rng.uniform(0.015, 0.035)]
        pj = [rng.uniform(0.2, 4), rng.uniform(0.005, 0.2), rng.uniform(1e-4, 1e-2), rng.uniform(1e-4, 4e-3), rng.uniform(0.015, 0.035)]
        r = rng.uniform(0.05, 1.2)
        beta = rng.uniform(0.5, 1.2)
        e, dedr = oracle.dispersion_pair(r, pi, pj, beta)
        assert e == pytest.approx(ind.dispersion_energy(r, pi, pj, beta), rel=1e-11)
        h = 1e-4*r
        fd = (ind.dispersion_energy(r+h, pi, pj, beta)-ind.dispersion_energy(r-h, pi, pj, beta))/(2*h)
        assert dedr == pytest.approx(fd, rel=1e-6)


def test_gauss_legendre_rule():
    x, w = oracle.gauss_legendre(32)
    ref_x, ref_w = np.polynomial.legendre.leggauss(32)
    order = np.argsort(0.5*(1-ref_x))
    assert np.allclose(np.sort(x), 0.5*(1-ref_x)[order], rtol=0, atol=1e-14)
    assert np.allclose(w[np.argsort(x)], 0.5*ref_w[order], rtol=1e-12)


# ---------------------------------------------------------------------------------------------------------------
# pair loop (ReferenceInteraxPairIxn.h:53-99): brute force and finite differences
# ---------------------------------------------------------------------------------------------------------------

def random_pair_system(method, n=24, box=(1.5, 1.6, 1.7), triclinic=False, switch=False, alchemical=False,
                       softcore=0.0, seed=1):
    rng = np.random.default_rng(seed)
    vectors = np.diag(box).astype(float)
    if triclinic:
        vectors[1, 0] = 0.3
        vectors[2, 0] = -0.2
        vectors[2, 1] = 0.25
    pos = rng.uniform(0, 1, size=(n, 3)) @ vectors
    system = ih.System(n, vectors.reshape(-1))
    forces = [ih.InteraxDispersionForce(), ih.InteraxExchangeForce(), ih.InteraxShortRangeForce()]
    alch = rng.uniform(size=n) < 0.3 if alchemical else np.zeros(n, bool)
    for i in range(n):
        forces[0].addParticle(rng.uniform(0.3, 3), rng.uniform(0.01, 0.1), rng.uniform(1e-4, 5e-3), rng.uniform(2e-4, 3e-3),
                              rng.uniform(0.018, 0.03), bool(alch[i]))
        w = rng.choice([0.021, 0.024, 0.0242, 0.028])
        forces[1].addParticle(rng.uniform(1, 10), w, bool(alch[i]))
        forces[2].addParticle(rng.uniform(-1000, -100), rng.uniform(100, 5000), w, bool(alch[i]))
    for f in forces:
        f.setNonbondedMethod(method)
        f.setCutoffDistance(0.7)
        f.setUseSwitchingFunction(switch)
        f.setSwitchingDistance(0.6)
        f.setSoftcoreRadius(softcore)
        for e in range(6):
            f.addExclusion(2*e, 2*e+1)
        system.addForce(f)
    forces[0].setDampingScale(0.8)
    return system, forces, pos


@pytest.mark.parametrize("method,triclinic,switch,alchemical,softcore", [
    (0, False, False, False, 0.0), (1, False, True, False, 0.0), (2, False, False, False, 0.0),
    (2, True, True, False, 0.0), (2, False, False, True, 0.0), (2, False, True, True, 0.3)])
def test_pair_loop_against_brute_force(method, triclinic, switch, alchemical, softcore):
    system, forces, pos = random_pair_system(method, triclinic=triclinic, switch=switch, alchemical=alchemical,
                                             softcore=softcore)
    box = system.getDefaultPeriodicBoxVectors()
    lam = 0.6
    for f in forces:
        e, fr = oracle.pair_force(f, pos, box, lam)
        assert e == pytest.approx(ind.pair_energy_total(f, pos, box, lam), rel=1e-9, abs=1e-9), f.NAME
        fd = ind.finite_difference_forces(lambda p: oracle.pair_force(f, p, box, lam)[0], pos, h=1e-6)
        scale = np.abs(fd).max()
        assert np.abs(fr-fd).max() <= 1e-5*scale+1e-7, f.NAME


def test_cell_list_matches_all_pairs():
    # more than 256 particles selects the cell grid; compare with the same system split so that it falls back
    rng = np.random.default_rng(5)
    n = 400
    L = 3.2
    pos = rng.uniform(0, L, size=(n, 3))
    f = ih.InteraxExchangeForce()
    for i in range(n):
        f.addParticle(2.0, 0.025)
    f.setNonbondedMethod(2)
    f.setCutoffDistance(1.0)
    box = np.diag([L, L, L])
    e, _ = oracle.pair_force(f, pos, box)
    assert e == pytest.approx(ind.pair_energy_total(f, pos, box), rel=1e-10)
    assert oracle.pair_count(f, pos, box) == sum(1 for i in range(n) for j in range(i+1, n)
                                                 if np.linalg.norm(ind.minimum_image(pos[j]-pos[i], box)) < 1.0)


def test_long_range_correction_against_quadrature():
    # TestInteraxDispersionForce.h:163-197: E_lrc = 2 pi N^2 <I>/V over the N(N+1)/2 pairs, by adaptive quadrature
    from scipy.integrate import quad
    f = ih.InteraxDispersionForce()
    types = [(2.0, 0.08, 0.004, 0.002, 0.027), (0.5, 0.006, 0.00015, 0.00023, 0.02)]
    counts = [5, 7]
    for t, c in zip(types, counts):
        for _ in range(c):
            f.addParticle(*t)
    f.setNonbondedMethod(2)
    f.setCutoffDistance(0.9)
    f.setUseSwitchingFunction(True)
    f.setSwitchingDistance(0.8)
    f.setUseLongRangeCorrection(True)
    f.setDampingScale(0.7)
    coef, alch = oracle.long_range_correction(f)
    total = 0.0
    for i in range(2):
        for j in range(i, 2):
            E = lambda r: ind.dispersion_energy(r, types[i], types[j], 0.7)
            tail = quad(lambda r: r*r*E(r), 0.9, np.inf, epsabs=0, epsrel=1e-12, limit=200)[0]
            sw = quad(lambda r: (1-ind.switch(r, 0.8, 0.9))*r*r*E(r), 0.8, 0.9, epsabs=0, epsrel=1e-12)[0]
            count = 0.5*counts[i]*(counts[i]+1) if i == j else counts[i]*counts[j]
            total += count*(tail+sw)
    N = sum(counts)
    expected = 2*math.pi*N*N/(0.5*N*(N+1))*total
    assert coef == pytest.approx(expected, rel=1e-9)
    assert alch == 0


# ---------------------------------------------------------------------------------------------------------------
# three-body (InteraxThreeBodyTerms.h:51-146, ReferenceInteraxThreeBodyIxn.cpp:84-135)
# ---------------------------------------------------------------------------------------------------------------

def test_single_triple_law_of_cosines():
    # TestInteraxThreeBodyForce.h:185-220
    pa, pb, pc = np.array([0.0, 0, 0]), np.array([0.4, 0, 0]), np.array([0.1, 0.35, 0.05])
    c9v = 2.5e-6
    e, fa, fb, fc = oracle.triple(pb-pa, pc-pa, pc-pb, c9v)
    assert e == pytest.approx(ind.triple_energy(pa, pb, pc, c9v), rel=1e-13)
    pos = np.array([pa, pb, pc])
    fd = ind.finite_difference_forces(lambda p: ind.triple_energy(p[0], p[1], p[2], c9v), pos, h=1e-6)
    assert np.allclose(np.array([fa, fb, fc]), fd, rtol=1e-6, atol=1e-9*np.abs(fd).max())


def test_c9_identity():
    # C9 = 3/4 alpha C6 for identical sites, with C6 = 3/4 alpha^2 w (TestInteraxThreeBodyForce.h:206-213)
    a, w = 1.7e-3, 900.0
    assert oracle.c9(a, w, a, w, a, w) == pytest.approx(0.75*a*(0.75*a*a*w), rel=1e-14)


@pytest.mark.parametrize("x", [0.05, 1.0, 1.99, 2.01, 5.0, 30.0])
def test_tang_toennies3(x):
    f, df = oracle.tang_toennies3(x)
    assert f == pytest.approx(float(gammainc(4, x)), rel=1e-12)
    assert df == pytest.approx(math.exp(-x)*x**3/6, rel=1e-13)


def random_threebody(periodic, damping, radius, alchemical, n_mol=7, seed=2, L=1.9):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, L, size=(3*n_mol, 3))
    f = ih.InteraxThreeBodyForce()
    for m in range(n_mol):
        # a centroid of two particles and a single-particle site in each molecule
        f.addSite([3*m, 3*m+1], rng.uniform(1e-3, 3e-3), rng.uniform(600, 1200), rng.uniform(0.02, 0.03), m,
                  alchemical and m == 0)
        f.addSite([3*m+2], rng.uniform(1e-3, 3e-3), rng.uniform(600, 1200), rng.uniform(0.02, 0.03), m, False)
    f.setCutoffDistance(1.2)
    f.setUsesPeriodicBoundaryConditions(periodic)
    f.setDampingMethod(1 if damping else 0)
    f.setDampingScale(0.9)
    f.setCentroidDampingRadius(radius)
    f.setScaleFactor(1.3)
    system = ih.System(3*n_mol, [L, 0, 0, 0, L, 0, 0, 0, L])
    system.addForce(f)
    return system, f, pos


@pytest.mark.parametrize("periodic,damping,radius,alchemical", [
    (False, False, 0.0, False), (True, False, 0.0, False), (True, True, 0.0, False), (False, True, 0.25, True),
    (True, False, 0.3, True)])
def test_threebody_against_brute_force(periodic, damping, radius, alchemical):
    system, f, pos = random_threebody(periodic, damping, radius, alchemical)
    box = system.getDefaultPeriodicBoxVectors()
    lam = 0.4
    e, fr = oracle.threebody_force(f, pos, box, lam)
    assert e == pytest.approx(ind.threebody_energy_total(f, pos, box, lam), rel=1e-11)
    fd = ind.finite_difference_forces(lambda p: oracle.threebody_force(f, p, box, lam)[0], pos, h=1e-6)
    assert np.abs(fr-fd).max() <= 1e-5*np.abs(fd).max()
