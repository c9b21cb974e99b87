"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker of the HIP path; the product
never imports it.  The functions take the force objects of the package's Python binding (plain data holders with
the reference's API) and evaluate them as the reference's Reference platform would (ReferenceInteraxKernels.cpp).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_ucp = ctypes.POINTER(ctypes.c_ubyte)


class TaylorScheme(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("about_smaller_width", ctypes.c_int), ("max_delta", ctypes.c_double),
                ("max_t", ctypes.c_double), ("max_tx", ctypes.c_double)]


class PairForce(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("num_particles", ctypes.c_int), ("params", _dp), ("alchemical", _ucp),
                ("method", ctypes.c_int), ("cutoff", ctypes.c_double), ("use_switch", ctypes.c_int),
                ("switching_distance", ctypes.c_double), ("softcore_radius", ctypes.c_double),
                ("num_exclusions", ctypes.c_int), ("exclusions", _ip), ("damping_scale", ctypes.c_double),
                ("use_long_range_correction", ctypes.c_int), ("overlap_method", ctypes.c_int),
                ("double_precision", ctypes.c_int)]


class ThreeBodyForce(ctypes.Structure):
    _fields_ = [("num_sites", ctypes.c_int), ("site_offsets", _ip), ("site_particles", _ip), ("alpha", _dp),
                ("frequency", _dp), ("width", _dp), ("molecule", _ip), ("alchemical", _ucp),
                ("cutoff", ctypes.c_double), ("scale_factor", ctypes.c_double), ("damping_scale", ctypes.c_double),
                ("centroid_damping_radius", ctypes.c_double), ("damping_method", ctypes.c_int),
                ("periodic", ctypes.c_int)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            build()
        L = ctypes.CDLL(LIBRARY_PATH)
        L.orc_scheme_accurate.argtypes = [ctypes.c_int, ctypes.POINTER(TaylorScheme)]
        L.orc_scheme_order2.argtypes = [ctypes.POINTER(TaylorScheme)]
        L.orc_slater_overlap.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(TaylorScheme), _dp, _dp]
        L.orc_tang_toennies.argtypes = [ctypes.c_double, _dp, _dp]
        L.orc_dispersion_pair.argtypes = [ctypes.c_double, _dp, _dp, ctypes.c_double, _dp, _dp]
        L.orc_exchange_pair.argtypes = [ctypes.c_double, _dp, _dp, ctypes.POINTER(TaylorScheme), _dp, _dp]
        L.orc_short_range_pair.argtypes = [ctypes.c_double, _dp, _dp, ctypes.POINTER(TaylorScheme), _dp, _dp]
        L.orc_pair_compute.argtypes = [ctypes.POINTER(PairForce), _dp, _dp, ctypes.c_double, _dp, _dp]
        L.orc_pair_compute.restype = ctypes.c_int
        L.orc_pair_compute_scale.argtypes = [ctypes.POINTER(PairForce), _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.orc_pair_compute_scale.restype = ctypes.c_int
        L.orc_pair_count.argtypes = [ctypes.POINTER(PairForce), _dp, _dp]
        L.orc_pair_count.restype = ctypes.c_longlong
        L.orc_long_range_correction.argtypes = [ctypes.POINTER(PairForce), _dp, _dp]
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_tang_toennies3.argtypes = [ctypes.c_double, _dp, _dp]
        L.orc_triple.argtypes = [_dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_double]
        L.orc_triple.restype = ctypes.c_double
        L.orc_c9.argtypes = [ctypes.c_double]*6
        L.orc_c9.restype = ctypes.c_double
        L.orc_threebody_compute.argtypes = [ctypes.POINTER(ThreeBodyForce), _dp, _dp, ctypes.c_double, _dp, _dp]
        L.orc_threebody_compute.restype = ctypes.c_int
        L.orc_threebody_compute_scale.argtypes = [ctypes.POINTER(ThreeBodyForce), _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.orc_threebody_compute_scale.restype = ctypes.c_int
        L.orc_threebody_count.argtypes = [ctypes.POINTER(ThreeBodyForce), _dp, _dp]
        L.orc_threebody_count.restype = ctypes.c_longlong
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------------------------------------------
# scalar functions
# ---------------------------------------------------------------------------------------------------------------

def scheme(method: int, double_precision: bool = True) -> TaylorScheme:
    s = TaylorScheme()
    if method == 0:
        lib().orc_scheme_accurate(int(double_precision), ctypes.byref(s))
    else:
        lib().orc_scheme_order2(ctypes.byref(s))
    return s


def slater_overlap(r, a, b, method=0, double_precision=True):
    S, dS = ctypes.c_double(), ctypes.c_double()
    lib().orc_slater_overlap(r, a, b, ctypes.byref(scheme(method, double_precision)), ctypes.byref(S), ctypes.byref(dS))
    return S.value, dS.value


def tang_toennies(x):
    f = np.zeros(3)
    df = np.zeros(3)
    lib().orc_tang_toennies(x, _d(f), _d(df))
    return f, df


def tang_toennies3(x):
    f, df = ctypes.c_double(), ctypes.c_double()
    lib().orc_tang_toennies3(x, ctypes.byref(f), ctypes.byref(df))
    return f.value, df.value


def dispersion_pair(r, pi, pj, beta):
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    pj = np.ascontiguousarray(pj, dtype=np.float64)
    e, d = ctypes.c_double(), ctypes.c_double()
    lib().orc_dispersion_pair(r, _d(pi), _d(pj), beta, ctypes.byref(e), ctypes.byref(d))
    return e.value, d.value


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _d(x), _d(w))
    return x, w


def c9(a1, w1, a2, w2, a3, w3):
    return lib().orc_c9(a1, w1, a2, w2, a3, w3)


def triple(a, b, c, c9v, rates=None, damping_radius=0.0):
    a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c))
    fa, fb, fc = np.zeros(3), np.zeros(3), np.zeros(3)
    r = None if rates is None else np.ascontiguousarray(rates, dtype=np.float64)
    e = lib().orc_triple(_d(a), _d(b), _d(c), c9v, None if r is None else _d(r), _d(fa), _d(fb), _d(fc), damping_radius)
    return e, fa, fb, fc


# ---------------------------------------------------------------------------------------------------------------
# whole forces
# ---------------------------------------------------------------------------------------------------------------

_KINDS = {"InteraxDispersionForce": 0, "InteraxExchangeForce": 1, "InteraxShortRangeForce": 2}


def _pair_struct(force, double_precision=True):
    params = np.ascontiguousarray(force.parameter_array().reshape(-1))
    alch = np.ascontiguousarray(force._alchemical_array())
    excl = np.ascontiguousarray(force._exclusion_array().reshape(-1), dtype=np.int32)
    s = PairForce()
    s.kind = _KINDS[force.NAME]
    s.num_particles = force.getNumParticles()
    s.params = _d(params)
    s.alchemical = alch.ctypes.data_as(_ucp)
    s.method = force.getNonbondedMethod()
    s.cutoff = force.getCutoffDistance()
    s.use_switch = int(force.getUseSwitchingFunction())
    s.switching_distance = force.getSwitchingDistance()
    s.softcore_radius = force.getSoftcoreRadius()
    s.num_exclusions = len(excl)//2
    s.exclusions = excl.ctypes.data_as(_ip)
    if s.kind == 0:
        s.damping_scale = force.getDampingScale()
        s.use_long_range_correction = int(force.getUseLongRangeCorrection())
    else:
        s.overlap_method = force.getOverlapMethod()
    s.double_precision = int(double_precision)
    return s, (params, alch, excl)


def pair_force(force, positions, box=None, lam=1.0, double_precision=True, with_scale=False):
    """Energy and forces (N x 3) of one pair force on the Reference platform; with_scale also returns the sum of
    |pair energies| before switching."""
    s, keep = _pair_struct(force, double_precision)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
    bx = np.ascontiguousarray(np.asarray(box if box is not None else np.eye(3)*1e6, dtype=np.float64).reshape(-1))
    forces = np.zeros(pos.size)
    e = ctypes.c_double()
    scale = ctypes.c_double()
    if lib().orc_pair_compute_scale(ctypes.byref(s), _d(pos), _d(bx), lam, _d(forces), ctypes.byref(e), ctypes.byref(scale)) != 0:
        raise ValueError("oracle: invalid pair force")
    if with_scale:
        return e.value, forces.reshape(-1, 3), scale.value
    return e.value, forces.reshape(-1, 3)


def pair_count(force, positions, box=None):
    s, keep = _pair_struct(force)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
    bx = np.ascontiguousarray(np.asarray(box if box is not None else np.eye(3)*1e6, dtype=np.float64).reshape(-1))
    return lib().orc_pair_count(ctypes.byref(s), _d(pos), _d(bx))


def long_range_correction(force):
    s, keep = _pair_struct(force)
    c, a = ctypes.c_double(), ctypes.c_double()
    lib().orc_long_range_correction(ctypes.byref(s), ctypes.byref(c), ctypes.byref(a))
    return c.value, a.value


def _threebody_struct(force):
    offsets, parts, alpha, freq, width, mol, alch = [np.ascontiguousarray(a) for a in force.site_arrays()]
    s = ThreeBodyForce()
    s.num_sites = force.getNumSites()
    s.site_offsets = offsets.ctypes.data_as(_ip)
    s.site_particles = parts.ctypes.data_as(_ip)
    s.alpha = _d(alpha)
    s.frequency = _d(freq)
    s.width = _d(width)
    s.molecule = mol.ctypes.data_as(_ip)
    s.alchemical = alch.ctypes.data_as(_ucp)
    s.cutoff = force.getCutoffDistance()
    s.scale_factor = force.getScaleFactor()
    s.damping_scale = force.getDampingScale()
    s.centroid_damping_radius = force.getCentroidDampingRadius()
    s.damping_method = force.getDampingMethod()
    s.periodic = int(force.usesPeriodicBoundaryConditions())
    return s, (offsets, parts, alpha, freq, width, mol, alch)


def threebody_force(force, positions, box=None, lam=1.0, with_scale=False):
    """Energy and forces of a three-body force; with_scale also returns the sum of |triple energies|."""
    s, keep = _threebody_struct(force)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
    bx = np.ascontiguousarray(np.asarray(box if box is not None else np.eye(3)*1e6, dtype=np.float64).reshape(-1))
    forces = np.zeros(pos.size)
    e = ctypes.c_double()
    scale = ctypes.c_double()
    if lib().orc_threebody_compute_scale(ctypes.byref(s), _d(pos), _d(bx), lam, _d(forces), ctypes.byref(e), ctypes.byref(scale)) != 0:
        raise ValueError("oracle: invalid three-body force")
    if with_scale:
        return e.value, forces.reshape(-1, 3), scale.value
    return e.value, forces.reshape(-1, 3)


def threebody_count(force, positions, box=None):
    s, keep = _threebody_struct(force)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
    bx = np.ascontiguousarray(np.asarray(box if box is not None else np.eye(3)*1e6, dtype=np.float64).reshape(-1))
    return lib().orc_threebody_count(ctypes.byref(s), _d(pos), _d(bx))


# ---------------------------------------------------------------------------------------------------------------
# the multi-threaded restatement of the CPU platform (oracle/libcpuplatform.so): the timed CPU baseline
# ---------------------------------------------------------------------------------------------------------------

CPU_PLATFORM_PATH = os.path.join(_HERE, "libcpuplatform.so")
_cpl = None


def cpu_platform_lib() -> ctypes.CDLL:
    global _cpl
    if _cpl is None:
        if not os.path.exists(CPU_PLATFORM_PATH):
            build()
        L = ctypes.CDLL(CPU_PLATFORM_PATH)
        L.cpl_pair_compute.argtypes = [ctypes.POINTER(PairForce), _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                       ctypes.POINTER(ctypes.c_longlong)]
        L.cpl_pair_compute.restype = ctypes.c_int
        L.cpl_threebody_compute.argtypes = [ctypes.POINTER(ThreeBodyForce), ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                            _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]
        L.cpl_threebody_compute.restype = ctypes.c_int
        L.cpl_num_threads.restype = ctypes.c_int
        _cpl = L
    return _cpl


def cpu_pair_force(force, positions, box=None, lam=1.0, threads=0):
    """Energy, forces (N x 3) and interacting pair count of one pair force by the FP32 multi-threaded restatement of
    the CPU platform (no long range correction)."""
    s, keep = _pair_struct(force, double_precision=False)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
    bx = np.ascontiguousarray(np.asarray(box if box is not None else np.eye(3)*1e6, dtype=np.float64).reshape(-1))
    forces = np.zeros(pos.size)
    e = ctypes.c_double()
    pairs = ctypes.c_longlong()
    if cpu_platform_lib().cpl_pair_compute(ctypes.byref(s), _d(pos), _d(bx), lam, threads, _d(forces), ctypes.byref(e),
                                           ctypes.byref(pairs)) != 0:
        raise ValueError("cpu platform: unsupported pair force (triclinic box)")
    return e.value, forces.reshape(-1, 3), pairs.value


def cpu_threebody_force(force, positions, box=None, lam=1.0, threads=0):
    """Energy, forces and interacting triple count of a three-body force by the multi-threaded restatement of the
    CPU platform."""
    s, keep = _threebody_struct(force)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
    bx = np.ascontiguousarray(np.asarray(box if box is not None else np.eye(3)*1e6, dtype=np.float64).reshape(-1))
    forces = np.zeros(pos.size)
    e = ctypes.c_double()
    triples = ctypes.c_longlong()
    if cpu_platform_lib().cpl_threebody_compute(ctypes.byref(s), pos.size//3, _d(pos), _d(bx), lam, threads, _d(forces),
                                                ctypes.byref(e), ctypes.byref(triples)) != 0:
        raise ValueError("cpu platform: invalid three-body force")
    return e.value, forces.reshape(-1, 3), triples.value


def evaluate(system, positions, box=None, groups=0xFFFFFFFF, parameters=None, double_precision=True, with_scale=False):
    """Energy and forces of the force groups of a System, as Context.getState(groups) on the Reference platform.
    with_scale also returns the sum of |pair energies| of the pair forces (the rounding-error scale)."""
    parameters = parameters or {}
    n = system.getNumParticles()
    energy = 0.0
    scale = 0.0
    forces = np.zeros((n, 3))
    if box is None:
        box = system.getDefaultPeriodicBoxVectors()
    for f in system.getForces():
        if not groups & (1 << f.getForceGroup()):
            continue
        lam = parameters.get(f.getAlchemicalParameterName(), 1.0)
        if f.NAME == "InteraxThreeBodyForce":
            e, fr, sc = threebody_force(f, positions, box, lam, with_scale=True)
            scale += sc
        else:
            e, fr, sc = pair_force(f, positions, box, lam, double_precision, with_scale=True)
            scale += sc
        energy += e
        forces += fr
    if with_scale:
        return energy, forces, scale
    return energy, forces
