"""Seeded synthetic systems of the benchmark configurations (water-like boxes and clusters, an ATM fluid, and an
alchemical solvated solute), built with the force classes of interax_hip.

Parameters follow the per-type values of the benchmark definition, mapped onto the force API as follows (the API
takes per-particle moments, polarizabilities and widths, not per-type C6 or damping rates):

  dispersion  m1 = sqrt(2 C6/k), alpha = k m1 with k = 1e-3, so that C6_ij = sqrt(C6_i C6_j);
              m2 = (0.01/3) m1 (C8 = 0.01 C6); m3 = ((1.2e-4 - 4.2 (0.01/3)^2)/4) m1 (C10 = 0.012 C8);
              width 1/b with a per-atom rate b ~ U(30, 45) nm^-1 and damping scale 1
  exchange    amplitude 1.5 |q_v|, Slater width 1/24 nm (O) and 1/32 nm (H)
  short range K = -170 |q_v|, L = 900 |q_v|, same widths
  three-body  equal frequencies w and alpha_i = (16 c_i / (9 w))^(1/3), so that C9 is the geometric mean of c_i

Geometry: rigid 3-site molecules (O-H 0.09572 nm, H-O-H 104.52 degrees) at uniformly random orientations.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

import interax_hip as ih

OH = 0.09572
HOH = math.radians(104.52)
C6 = {"O": 2.5e-3, "H": 1.5e-4}
QV = {"O": -6.6, "H": -0.8}
SLATER = {"O": 1.0/24, "H": 1.0/32}
K_POL = 1e-3


def dispersion_moments(element: str):
    m1 = math.sqrt(2*C6[element]/K_POL)
    alpha = K_POL*m1
    m2 = (0.01/3)*m1
    m3 = ((1.2e-4-4.2*(0.01/3)**2)/4)*m1
    return m1, m2, m3, alpha


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices (from uniform unit quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    return np.stack([
        np.stack([1-2*(y*y+z*z), 2*(x*y-z*w), 2*(x*z+y*w)], -1),
        np.stack([2*(x*y+z*w), 1-2*(x*x+z*z), 2*(y*z-x*w)], -1),
        np.stack([2*(x*z-y*w), 2*(y*z+x*w), 1-2*(x*x+y*y)], -1)], 1)


def water_sites(rng: np.random.Generator, centers: np.ndarray) -> np.ndarray:
    """O, H, H positions of molecules with their oxygen at `centers`, randomly oriented."""
    n = len(centers)
    local = np.array([[0, 0, 0], [OH, 0, 0], [OH*math.cos(HOH), OH*math.sin(HOH), 0]])
    rot = random_rotations(rng, n)
    pos = centers[:, None, :]+np.einsum("nij,kj->nki", rot, local)
    return pos.reshape(-1, 3)


def add_water_forces(system: ih.System, num_molecules: int, rng: np.random.Generator, method: int, cutoff: float,
                     use_lrc: bool = False, alchemical: Optional[np.ndarray] = None,
                     softcore_radius: float = 0.0) -> Tuple[ih.InteraxDispersionForce, ih.InteraxExchangeForce, ih.InteraxShortRangeForce]:
    """The three pair forces of water-like molecules (atoms O, H, H per molecule), intramolecular pairs excluded."""
    disp, exch, sr = ih.InteraxDispersionForce(), ih.InteraxExchangeForce(), ih.InteraxShortRangeForce()
    n = 3*num_molecules
    rates = rng.uniform(30.0, 45.0, size=n)
    alch = np.zeros(n, dtype=bool) if alchemical is None else alchemical
    for i in range(n):
        el = "O" if i % 3 == 0 else "H"
        m1, m2, m3, alpha = dispersion_moments(el)
        disp.addParticle(m1, m2, m3, alpha, 1.0/rates[i], bool(alch[i]))
        exch.addParticle(1.5*abs(QV[el]), SLATER[el], bool(alch[i]))
        sr.addParticle(-170*abs(QV[el]), 900*abs(QV[el]), SLATER[el], bool(alch[i]))
    for f in (disp, exch, sr):
        f.setNonbondedMethod(method)
        f.setCutoffDistance(cutoff)
        f.setSoftcoreRadius(softcore_radius)
        for m in range(num_molecules):
            f.addExclusion(3*m, 3*m+1)
            f.addExclusion(3*m, 3*m+2)
            f.addExclusion(3*m+1, 3*m+2)
        system.addForce(f)
    disp.setDampingScale(1.0)
    disp.setUseLongRangeCorrection(use_lrc)
    return disp, exch, sr


def water_box(molecules_per_side: int, box: float, cutoff: float, seed: int = 0, use_lrc: bool = False,
              jitter: float = 0.02):
    """Water-like molecules on a jittered cubic lattice in a periodic cube (the plain-method configurations)."""
    rng = np.random.default_rng(seed)
    m = molecules_per_side
    spacing = box/m
    grid = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    centers = (grid+0.5)*spacing+rng.uniform(-jitter, jitter, size=grid.shape)
    positions = water_sites(rng, centers)
    system = ih.System(len(positions), [box, 0, 0, 0, box, 0, 0, 0, box])
    add_water_forces(system, len(centers), rng, ih.InteraxPairForce.CutoffPeriodic, cutoff, use_lrc)
    return system, positions


def config0(seed: int = 0):
    """216 molecules (648 atoms) in a 1.862 nm box, cutoff 0.9 nm."""
    return water_box(6, 1.862, 0.9, seed)


def config1(seed: int = 0):
    """32768 molecules (98304 atoms) in a 9.94 nm box, cutoff 1.0 nm, long range correction."""
    return water_box(32, 9.94, 1.0, seed, use_lrc=True)


def perturbed(positions: np.ndarray, evaluation: int, sigma: float = 0.005) -> np.ndarray:
    """The coordinates of evaluation k of the scale configuration: base + N(0, sigma) per component, seed k."""
    rng = np.random.default_rng(1000+evaluation)
    return positions+rng.normal(0.0, sigma, size=positions.shape)


def rsa_points(rng: np.random.Generator, n: int, box: float, min_sep: float, periodic: bool = True,
               sphere_radius: Optional[float] = None) -> np.ndarray:
    """Random sequential addition with a minimum separation, accelerated by a cell grid.  Candidates are drawn in
    blocks from the generator (the same stream as one draw of 3 per attempt) and tested with scalar arithmetic."""
    cells = max(1, int(box/min_sep))
    size = box/cells
    min2 = min_sep*min_sep
    grid = {}
    points = []
    attempts = 0
    lo, hi = (0.0, box) if sphere_radius is None else (-sphere_radius, sphere_radius)
    r2max = None if sphere_radius is None else sphere_radius**2
    shift = 0.0 if sphere_radius is None else box/2
    block = []
    while len(points) < n:
        if not block:
            block = rng.uniform(lo, hi, size=(4096, 3)).tolist()
            block.reverse()
        attempts += 1
        if attempts > 200*n:
            raise RuntimeError("random sequential addition did not converge")
        x, y, z = block.pop()
        if r2max is not None:
            if x*x+y*y+z*z > r2max:
                continue
            x, y, z = x+shift, y+shift, z+shift
        cx, cy, cz = int(x//size), int(y//size), int(z//size)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = (cx+dx, cy+dy, cz+dz)
                    if periodic:
                        key = (key[0] % cells, key[1] % cells, key[2] % cells)
                    for (qx, qy, qz) in grid.get(key, ()):
                        ex, ey, ez = x-qx, y-qy, z-qz
                        if periodic:
                            ex -= box*round(ex/box)
                            ey -= box*round(ey/box)
                            ez -= box*round(ez/box)
                        if ex*ex+ey*ey+ez*ez < min2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        key = (cx % cells, cy % cells, cz % cells) if periodic else (cx, cy, cz)
        grid.setdefault(key, []).append((x, y, z))
        points.append((x, y, z))
    return np.array(points, dtype=np.float64).reshape(n, 3)


def atm_fluid(num_sites: int = 32768, density: float = 21.0, min_sep: float = 0.3, cutoff: float = 0.7, seed: int = 0):
    """Single-site atoms at a number density in a periodic cube with an ATM force (the triplet configuration)."""
    rng = np.random.default_rng(seed)
    box = (num_sites/density)**(1.0/3.0)
    positions = rsa_points(rng, num_sites, box, min_sep)
    system = ih.System(num_sites, [box, 0, 0, 0, box, 0, 0, 0, box])
    f = ih.InteraxThreeBodyForce()
    w = 1000.0
    c = rng.uniform(1e-6, 4e-6, size=num_sites)
    alpha = (16*c/(9*w))**(1.0/3.0)
    for i in range(num_sites):
        f.addSite([i], alpha[i], w)
    f.setCutoffDistance(cutoff)
    f.setUsesPeriodicBoundaryConditions(True)
    system.addForce(f)
    return system, positions


def water_cluster(num_molecules: int, seed: int, density: float = 33.4, min_sep: float = 0.25,
                  threebody_cutoff: float = 50.0):
    """A nonperiodic cluster: all pair forces without cutoff plus a molecular ATM force on the oxygens."""
    rng = np.random.default_rng(seed)
    radius = (3*num_molecules/(4*math.pi*density))**(1.0/3.0)
    span = 2*radius+2.0
    centers = rsa_points(rng, num_molecules, span, min_sep, periodic=False, sphere_radius=radius)
    positions = water_sites(rng, centers)
    system = ih.System(len(positions))
    add_water_forces(system, num_molecules, rng, ih.InteraxPairForce.NoCutoff, 1.0)
    tb = ih.InteraxThreeBodyForce()
    alpha_w = 1.45e-3   # molecular polarizability (nm^3) and frequency (kJ/mol) of a water-like molecule
    freq = 1150.0
    for m in range(num_molecules):
        tb.addSite([3*m], alpha_w, freq, 0.0, m)
    tb.setCutoffDistance(threebody_cutoff)
    system.addForce(tb)
    return system, positions


def cluster_sizes(num_clusters: int = 1024, seed: int = 0, low: int = 2, high: int = 256) -> List[int]:
    """Molecule counts log-uniform in [low, high]."""
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(x))) for x in rng.uniform(math.log(low), math.log(high), size=num_clusters)]


# A benzene-like rigid solute (6 heavy atoms on a 0.139 nm ring, 6 hydrogens at 0.248 nm): heavy-atom parameters
# C6 1.8e-3 kJ mol^-1 nm^6, valence charge -4.0, Slater width 1/26 nm; hydrogens as the water hydrogens.
SOLUTE_C6, SOLUTE_QV, SOLUTE_SLATER = 1.8e-3, -4.0, 1.0/26


def alchemical_box(softcore_radius: float = 0.0, seed: int = 0, num_waters: int = 1000, box: float = 3.106,
                   cutoff: float = 0.9, use_lrc: bool = False):
    """An alchemical job of the served configuration: a rigid 12-atom solute, every atom alchemical (its 66
    intramolecular pairs excluded), at the centre of a periodic cube with water-like molecules (RSA-placed oxygens,
    at least 0.25 nm apart and 0.3 nm from the solute), cutoff 0.9 nm, soft-core radius R, optionally the dispersion
    long range correction (whose alchemical part scales with lambda)."""
    rng = np.random.default_rng(seed)
    angles = np.arange(6)*np.pi/3
    ring = np.stack([np.cos(angles), np.sin(angles), np.zeros(6)], -1)
    solute = np.concatenate([0.139*ring, 0.248*ring])@random_rotations(rng, 1)[0].T+box/2
    centers = []
    while len(centers) < num_waters:
        p = rng.uniform(0, box, size=3)
        d = solute-p
        d -= box*np.round(d/box)
        if (d*d).sum(1).min() < 0.3**2:
            continue
        if centers:
            c = np.array(centers)-p
            c -= box*np.round(c/box)
            if (c*c).sum(1).min() < 0.25**2:
                continue
        centers.append(p)
    waters = water_sites(rng, np.array(centers))
    positions = np.concatenate([solute, waters])
    n = len(positions)
    system = ih.System(n, [box, 0, 0, 0, box, 0, 0, 0, box])
    disp, exch, sr = ih.InteraxDispersionForce(), ih.InteraxExchangeForce(), ih.InteraxShortRangeForce()
    rates = rng.uniform(30.0, 45.0, size=n)
    for i in range(n):
        alch = i < 12
        if i < 6:
            m1 = math.sqrt(2*SOLUTE_C6/K_POL)
            moments, qv, width = (m1, (0.01/3)*m1, ((1.2e-4-4.2*(0.01/3)**2)/4)*m1, K_POL*m1), SOLUTE_QV, SOLUTE_SLATER
        else:
            el = "H" if i < 12 or (i-12) % 3 != 0 else "O"
            moments, qv, width = dispersion_moments(el), QV[el], SLATER[el]
        disp.addParticle(*moments, 1.0/rates[i], alch)
        exch.addParticle(1.5*abs(qv), width, alch)
        sr.addParticle(-170*abs(qv), 900*abs(qv), width, alch)
    for f in (disp, exch, sr):
        f.setNonbondedMethod(ih.InteraxPairForce.CutoffPeriodic)
        f.setCutoffDistance(cutoff)
        f.setSoftcoreRadius(softcore_radius)
        for i in range(12):
            for j in range(i+1, 12):
                f.addExclusion(i, j)
        for m in range(num_waters):
            o = 12+3*m
            f.addExclusion(o, o+1)
            f.addExclusion(o, o+2)
            f.addExclusion(o+1, o+2)
        system.addForce(f)
    disp.setDampingScale(1.0)
    disp.setUseLongRangeCorrection(use_lrc)
    return system, positions
