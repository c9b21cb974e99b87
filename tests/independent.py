"""Independent evaluations of the InteraxFF energies for pinning the oracle.

Nothing here shares code with the oracle: the overlap integral is evaluated with mpmath at high precision from its
closed form (docs/SPEC.md:41-45 of the reference) or, for nearly equal widths, by Cauchy's integral formula on a
circle in the complex width plane (as InteraxTestUtilities.h:248-270 does); the Tang-Toennies functions are the
regularized lower incomplete gamma function P(n+1, x) of scipy; minimum images are found by searching the 27
neighbouring images; and forces are central finite differences of the energy.
"""
import itertools
import math

import mpmath
import numpy as np
from scipy.special import gammainc

mpmath.mp.dps = 40


def overlap_closed_mp(r, a, b):
    """S(r; a, b) = F/(8 pi r) at 40 digits (complex widths allowed)."""
    r, a, b = mpmath.mpf(r), mpmath.mpmathify(a), mpmath.mpmathify(b)
    d = b*b-a*a
    ea, eb = mpmath.exp(-r/a), mpmath.exp(-r/b)
    f = 4*a*a*b*b/d**3*(ea-eb)+r/d**2*(a*ea+b*eb)
    return f/(8*mpmath.pi*r)


def overlap(r, a, b):
    """S(r; a, b) to ~1e-20, valid for every pair of positive widths."""
    if a == b:
        x = mpmath.mpf(r)/a
        return float(mpmath.exp(-x)*(1+x+x*x/3)/(64*mpmath.pi*mpmath.mpf(a)**3))
    t = (b-a)/a
    if abs(t) > 0.05:
        return float(mpmath.re(overlap_closed_mp(r, a, b)))
    # Cauchy: S(t) = mean over the circle |z| = 1/4 of S(z) z/(z - t)
    n = 64
    total = mpmath.mpf(0)
    for k in range(n):
        z = mpmath.mpf(0.25)*mpmath.expjpi(2*(k+0.5)/n)
        total += overlap_closed_mp(r, a, a*(1+z))*z/(z-t)
    return float(mpmath.re(total)/n)


def overlap_derivative(r, a, b, h=None):
    h = h or 1e-4*r
    f = lambda s: overlap(s, a, b)
    return (f(r-2*h)-8*f(r-h)+8*f(r+h)-f(r+2*h))/(12*h)


def tang_toennies(n, x):
    return float(gammainc(n+1, x))


def dispersion_energy(r, pi, pj, beta):
    m1i, m2i, m3i, ai, wi = pi
    m1j, m2j, m3j, aj, wj = pj
    if ai*aj == 0:
        return 0.0
    den = max(ai*m1j+aj*m1i, 1e-30)
    c6 = ai*aj*m1i*m1j/den
    c8 = 1.5*ai*aj*(m1i*m2j+m2i*m1j)/den
    c10 = (2*ai*aj*(m1i*m3j+m3i*m1j)+4.2*ai*aj*m2i*m2j)/den
    x = 2*beta*r/(wi+wj)
    return -(tang_toennies(6, x)*c6/r**6+tang_toennies(8, x)*c8/r**8+tang_toennies(10, x)*c10/r**10)


def exchange_energy(r, pi, pj):
    if pi[1]*pj[1] == 0:
        return 0.0
    return pi[0]*pj[0]*overlap(r, pi[1], pj[1])


def short_range_energy(r, pi, pj):
    if pi[2]*pj[2] == 0:
        return 0.0
    st = 8*math.pi*(pi[2]+pj[2])**3*overlap(r, pi[2], pj[2])
    return 0.5*(pi[0]+pj[0])*st+0.5*(pi[1]+pj[1])*st*st


def minimum_image(d, box):
    """Shortest lattice image of d (searching the neighbouring images of the reduced vector)."""
    box = np.asarray(box).reshape(3, 3)
    d = np.array(d, dtype=float)
    d -= box[2]*round(d[2]/box[2, 2])
    d -= box[1]*round(d[1]/box[1, 1])
    d -= box[0]*round(d[0]/box[0, 0])
    best = d
    for i, j, k in itertools.product((-1, 0, 1), repeat=3):
        e = d+i*box[0]+j*box[1]+k*box[2]
        if e @ e < best @ best:
            best = e
    return best


def switch(r, rs, rc):
    if r <= rs:
        return 1.0
    if r >= rc:
        return 0.0
    x = (r-rs)/(rc-rs)
    return 1-10*x**3+15*x**4-6*x**5


def pair_energy_total(force, positions, box=None, lam=1.0):
    """Brute-force energy of a pair force (all pairs, minimum image by search)."""
    params = force.parameter_array()
    kind = force.NAME
    n = force.getNumParticles()
    excl = {tuple(sorted(e)) for e in force._exclusions}
    alch = force._alchemical_array().astype(bool)
    lam = lam if alch.any() else 1.0
    method = force.getNonbondedMethod()
    rc = force.getCutoffDistance()
    R = force.getSoftcoreRadius()
    total = 0.0
    for i in range(n):
        for j in range(i+1, n):
            if (i, j) in excl:
                continue
            d = positions[j]-positions[i]
            if method == 2:
                d = minimum_image(d, box)
            r = float(np.sqrt(d @ d))
            if method != 0 and r >= rc:
                continue
            apair = alch[i] != alch[j]
            reff = (r**6+(1-lam)*R**6)**(1/6) if (apair and R > 0) else r
            if kind == "InteraxDispersionForce":
                e = dispersion_energy(reff, params[i], params[j], force.getDampingScale())
            elif kind == "InteraxExchangeForce":
                e = exchange_energy(reff, params[i], params[j])
            else:
                e = short_range_energy(reff, params[i], params[j])
            if method != 0 and force.getUseSwitchingFunction():
                e *= switch(r, force.getSwitchingDistance(), rc)
            total += (lam if apair else 1.0)*e
    return total


def triple_energy(pa, pb, pc, c9, box=None, rates=None, radius=0.0):
    """ATM energy from the three displacement vectors, by the law of cosines (interior angles of the triangle)."""
    a = pb-pa
    b = pc-pa
    c = pc-pb
    if box is not None:
        a, b, c = (minimum_image(v, box) for v in (a, b, c))
    rab, rac, rbc = (float(np.sqrt(v @ v)) for v in (a, b, c))
    cos_a = (a @ b)/(rab*rac)
    cos_b = (-a @ c)/(rab*rbc)
    cos_c = (b @ c)/(rac*rbc)
    e = c9*(1+3*cos_a*cos_b*cos_c)/(rab*rac*rbc)**3
    damp = 1.0
    for k, r in enumerate((rab, rac, rbc)):
        if rates is not None:
            damp *= float(gammainc(4, rates[k]*r))
        if radius > 0:
            damp *= -math.expm1(-(r/radius)**6)
    return e*damp


def c9(a1, w1, a2, w2, a3, w3):
    return 1.5*a1*a2*a3*w1*w2*w3*(w1+w2+w3)/((w1+w2)*(w2+w3)*(w3+w1))


def threebody_energy_total(force, positions, box=None, lam=1.0):
    """Brute-force O(N^3) energy of an InteraxThreeBodyForce."""
    sites = force._sites
    n = len(sites)
    pos = np.array([np.mean([positions[p] for p in s[0]], axis=0) for s in sites])
    mol = [s[4] if s[4] >= 0 else -1-i for i, s in enumerate(sites)]
    alch = [s[5] for s in sites]
    lam = lam if any(alch) else 1.0
    periodic = force.usesPeriodicBoundaryConditions()
    rc = force.getCutoffDistance()
    beta = force.getDampingScale()

    def delta(i, j):
        d = pos[j]-pos[i]
        return minimum_image(d, box) if periodic else d

    total = 0.0
    for i, j, k in itertools.combinations(range(n), 3):
        if len({mol[i], mol[j], mol[k]}) < 3:
            continue
        if any(float(np.sqrt(delta(p, q) @ delta(p, q))) >= rc for p, q in ((i, j), (i, k), (j, k))):
            continue
        cc = force.getScaleFactor()*c9(sites[i][1], sites[i][2], sites[j][1], sites[j][2], sites[k][1], sites[k][2])
        if alch[i] or alch[j] or alch[k]:
            cc *= lam
        rates = None
        if force.getDampingMethod() == 1:
            w = [sites[q][3] for q in (i, j, k)]
            rates = [2*beta/(w[0]+w[1]), 2*beta/(w[0]+w[2]), 2*beta/(w[1]+w[2])]
        a = delta(i, j)
        b = delta(i, k)
        total += triple_energy(np.zeros(3), a, b, cc, None, rates, force.getCentroidDampingRadius()) \
            if not periodic else _triple_from_vectors(a, b, delta(j, k), cc, rates, force.getCentroidDampingRadius())
    return total


def _triple_from_vectors(a, b, c, c9v, rates, radius):
    u, v, w = a @ a, b @ b, c @ c
    p, q, s = a @ b, a @ c, b @ c
    P = u*v*w
    e = c9v*(P**-1.5-3*p*q*s*P**-2.5)
    for k, r in enumerate((math.sqrt(u), math.sqrt(v), math.sqrt(w))):
        if rates is not None:
            e *= float(gammainc(4, rates[k]*r))
        if radius > 0:
            e *= -math.expm1(-(r/radius)**6)
    return e


def finite_difference_forces(energy, positions, h=1e-5):
    """-dE/dx by central differences."""
    f = np.zeros_like(positions)
    for i in range(len(positions)):
        for k in range(3):
            p = positions.copy()
            p[i, k] += h
            ep = energy(p)
            p[i, k] -= 2*h
            em = energy(p)
            f[i, k] = -(ep-em)/(2*h)
    return f
