"""Python binding of the interax_hip C ABI (include/interax_hip.h).

It mirrors the public API of the reference plugin's force classes (openmmapi/include/InteraxPairForce.h:70-239,
InteraxDispersionForce.h:70-182, InteraxExchangeForce.h:60-160, InteraxShortRangeForce.h:59-162,
InteraxThreeBodyForce.h:75-258): same method names, argument meaning, defaults and error messages, so that a
System can be described as with the reference and evaluated on the GPU.  `Context` plays the role of an OpenMM
Context on this platform: it owns the device buffers (ixh_state) and one kernel handle per fused group of pair
forces and per three-body force, and evaluates force groups on request (the `groups` argument of getState).

Only the native library does any arithmetic; this module moves arrays across the C ABI.  Loading fails loudly if
libinterax_hip.so has not been built.
"""
from __future__ import annotations

import ctypes
import gzip
import math
import os
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libinterax_hip.so")

IXH_OK = 0
IXH_ERROR_INVALID_ARGUMENT, IXH_ERROR_HIP, IXH_ERROR_STATE, IXH_ERROR_UNSUPPORTED, IXH_ERROR_CAPACITY = 1, 2, 3, 4, 5
IXH_SINGLE, IXH_MIXED, IXH_DOUBLE = 0, 1, 2
IXH_TERM_DISPERSION, IXH_TERM_EXCHANGE, IXH_TERM_SHORT_RANGE = 1, 2, 4
IXH_INCLUDE_FORCES, IXH_INCLUDE_ENERGY = 1, 2
IXH_DEFER_LIST_CHECK = 4
PRECISIONS = {"single": IXH_SINGLE, "mixed": IXH_MIXED, "double": IXH_DOUBLE}


class InteraxError(RuntimeError):
    """The error raised where the reference throws OpenMMException."""


# ----------------------------------------------------------------------------------------------------------------
# ctypes declarations of include/interax_hip.h
# ----------------------------------------------------------------------------------------------------------------

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_ucp = ctypes.POINTER(ctypes.c_ubyte)


class Buffers(ctypes.Structure):
    _fields_ = [("num_particles", ctypes.c_int), ("precision", ctypes.c_int), ("posq", ctypes.c_void_p),
                ("posq_correction", ctypes.c_void_p), ("force", ctypes.c_void_p), ("force_stride", ctypes.c_int),
                ("energy", ctypes.c_void_p), ("box", ctypes.c_double*9), ("lambda_", ctypes.c_double),
                ("stream", ctypes.c_void_p), ("atom_index", ctypes.c_void_p)]


class PairDesc(ctypes.Structure):
    _fields_ = [("num_particles", ctypes.c_int), ("terms", ctypes.c_int), ("method", ctypes.c_int),
                ("cutoff", ctypes.c_double), ("use_switching_function", ctypes.c_int),
                ("switching_distance", ctypes.c_double), ("softcore_radius", ctypes.c_double),
                ("alchemical", _ucp), ("num_exclusions", ctypes.c_int), ("exclusions", _ip),
                ("precision", ctypes.c_int), ("dispersion", _dp), ("damping_scale", ctypes.c_double),
                ("use_long_range_correction", ctypes.c_int), ("exchange", _dp),
                ("exchange_overlap_method", ctypes.c_int), ("short_range", _dp),
                ("short_range_overlap_method", ctypes.c_int), ("default_box", _dp)]


class ThreeBodyDesc(ctypes.Structure):
    _fields_ = [("num_particles", ctypes.c_int), ("num_sites", ctypes.c_int), ("site_offsets", _ip),
                ("site_particles", _ip), ("alpha", _dp), ("frequency", _dp), ("width", _dp), ("molecule", _ip),
                ("alchemical", _ucp), ("cutoff", ctypes.c_double), ("scale_factor", ctypes.c_double),
                ("damping_method", ctypes.c_int), ("damping_scale", ctypes.c_double),
                ("centroid_damping_radius", ctypes.c_double), ("periodic", ctypes.c_int),
                ("precision", ctypes.c_int)]


class PairDeviceParameters(ctypes.Structure):
    _fields_ = [("dispersion", ctypes.c_void_p), ("exchange", ctypes.c_void_p), ("short_range", ctypes.c_void_p),
                ("damping_scale", ctypes.c_double), ("stream", ctypes.c_void_p)]


class ThreeBodyDeviceParameters(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_void_p), ("frequency", ctypes.c_void_p), ("width", ctypes.c_void_p),
                ("scale_factor", ctypes.c_double), ("damping_scale", ctypes.c_double),
                ("centroid_damping_radius", ctypes.c_double), ("stream", ctypes.c_void_p)]


class BatchShape(ctypes.Structure):
    _fields_ = [("pair_workgroups", ctypes.c_int), ("triple_waves", ctypes.c_int), ("triple_pairs_per_item", ctypes.c_int)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("num_systems", ctypes.c_int), ("particle_offsets", _ip), ("pair", ctypes.POINTER(PairDesc)),
                ("site_offsets", _ip), ("threebody", ctypes.POINTER(ThreeBodyDesc)), ("precision", ctypes.c_int)]


# Every exported symbol of include/interax_hip.h: (name, argtypes).
SIGNATURES = {
    "ixh_last_error": [],
    "ixh_version": [_ip, _ip],
    "ixh_device_count": [_ip],
    "ixh_state_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    "ixh_state_destroy": [ctypes.c_void_p],
    "ixh_state_set_positions": [ctypes.c_void_p, _dp],
    "ixh_state_set_box": [ctypes.c_void_p, _dp],
    "ixh_state_set_lambda": [ctypes.c_void_p, ctypes.c_double],
    "ixh_state_clear": [ctypes.c_void_p],
    "ixh_state_set_stream": [ctypes.c_void_p, ctypes.c_void_p],
    "ixh_state_get_forces": [ctypes.c_void_p, _dp],
    "ixh_state_get_energy": [ctypes.c_void_p, _dp],
    "ixh_state_synchronize": [ctypes.c_void_p],
    "ixh_state_get_buffers": [ctypes.c_void_p, ctypes.POINTER(Buffers)],
    "ixh_pair_create": [ctypes.c_int, ctypes.POINTER(PairDesc), ctypes.POINTER(ctypes.c_void_p)],
    "ixh_pair_destroy": [ctypes.c_void_p],
    "ixh_pair_set_parameters": [ctypes.c_void_p, ctypes.POINTER(PairDesc)],
    "ixh_pair_set_parameters_device": [ctypes.c_void_p, ctypes.POINTER(PairDeviceParameters)],
    "ixh_pair_execute": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int, ctypes.c_int],
    "ixh_pair_execute_noalloc": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int, ctypes.c_int],
    "ixh_pair_long_range_coefficients": [ctypes.c_void_p, _dp, _dp],
    "ixh_pair_long_range_energy": [ctypes.c_void_p, _dp, ctypes.c_double, _dp],
    "ixh_pair_lambda_energies": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _dp],
    "ixh_pair_last_pair_count": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)],
    "ixh_pair_set_tiles_per_item": [ctypes.c_void_p, ctypes.c_int],
    "ixh_pair_autotune": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int, _ip],
    "ixh_pair_image_mode": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
    "ixh_pair_list_statistics": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
    "ixh_threebody_create": [ctypes.c_int, ctypes.POINTER(ThreeBodyDesc), ctypes.POINTER(ctypes.c_void_p)],
    "ixh_threebody_destroy": [ctypes.c_void_p],
    "ixh_threebody_set_parameters": [ctypes.c_void_p, ctypes.POINTER(ThreeBodyDesc)],
    "ixh_threebody_set_parameters_device": [ctypes.c_void_p, ctypes.POINTER(ThreeBodyDeviceParameters)],
    "ixh_threebody_execute": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int],
    "ixh_threebody_list_was_valid": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    "ixh_threebody_last_triple_count": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)],
    "ixh_set_profiling": [ctypes.c_int],
    "ixh_pair_kernel_time": [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_longlong)],
    "ixh_pair_reset_kernel_time": [ctypes.c_void_p],
    "ixh_threebody_kernel_time": [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_longlong)],
    "ixh_threebody_reset_kernel_time": [ctypes.c_void_p],
    "ixh_pair_create_segmented": [ctypes.c_int, ctypes.POINTER(PairDesc), ctypes.c_int, _ip, ctypes.POINTER(ctypes.c_void_p)],
    "ixh_pair_execute_segments": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    "ixh_threebody_create_segmented": [ctypes.c_int, ctypes.POINTER(ThreeBodyDesc), ctypes.c_int, _ip,
                                       ctypes.POINTER(ctypes.c_void_p)],
    "ixh_threebody_execute_segments": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int, ctypes.c_void_p],
    "ixh_batch_create": [ctypes.c_int, ctypes.POINTER(BatchDesc), ctypes.POINTER(ctypes.c_void_p)],
    "ixh_batch_destroy": [ctypes.c_void_p],
    "ixh_batch_set_positions": [ctypes.c_void_p, _dp],
    "ixh_batch_kernel_time": [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_longlong), _dp, ctypes.POINTER(ctypes.c_longlong)],
    "ixh_batch_reset_kernel_time": [ctypes.c_void_p],
    "ixh_batch_get_buffers": [ctypes.c_void_p, ctypes.POINTER(Buffers)],
    "ixh_batch_execute": [ctypes.c_void_p, ctypes.c_int],
    "ixh_batch_get_energies": [ctypes.c_void_p, _dp],
    "ixh_batch_get_forces": [ctypes.c_void_p, _dp],
    "ixh_batch_set_graph": [ctypes.c_void_p, ctypes.c_int],
    "ixh_batch_autotune": [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(BatchShape)],
    "ixh_batch_get_shape": [ctypes.c_void_p, ctypes.POINTER(BatchShape)],
    "ixh_batch_set_shape": [ctypes.c_void_p, ctypes.POINTER(BatchShape)],
    "ixh_pair_execute_subset": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int, ctypes.c_int, ctypes.c_void_p, _ip,
                                ctypes.c_int],
    "ixh_threebody_execute_subset": [ctypes.c_void_p, ctypes.POINTER(Buffers), ctypes.c_int, ctypes.c_void_p, _ip, ctypes.c_int],
    "ixh_pair_list_was_valid": [ctypes.c_void_p, _ip],
    "ixh_server_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.POINTER(ctypes.c_void_p)],
    "ixh_server_add_cluster": [ctypes.c_void_p, ctypes.POINTER(PairDesc), ctypes.POINTER(ThreeBodyDesc), _ip],
    "ixh_server_add_periodic": [ctypes.c_void_p, ctypes.POINTER(PairDesc), _dp, _dp, _ip],
    "ixh_server_start": [ctypes.c_void_p],
    "ixh_server_submit": [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double, ctypes.POINTER(ctypes.c_longlong)],
    "ixh_server_submit_windows": [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)],
    "ixh_server_wait": [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, _dp],
    "ixh_server_wait_windows": [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp],
    "ixh_server_memory": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                          ctypes.POINTER(ctypes.c_longlong)],
    "ixh_server_statistics": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
    "ixh_server_replay": [ctypes.c_void_p, ctypes.c_int, _dp, _ip, _dp, ctypes.POINTER(ctypes.c_longlong), _dp, _ip, _dp, _dp, _dp,
                          _dp],
    "ixh_server_destroy": [ctypes.c_void_p],
}

_lib = None


def load_library(path: str = LIBRARY_PATH) -> ctypes.CDLL:
    """Load libinterax_hip.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise InteraxError(f"{path} has not been built (run __graft_entry__.build() or make -C openmm-interax-plugin_amd)")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "ixh_last_error" else ctypes.c_int
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != IXH_OK:
        raise InteraxError(_lib.ixh_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def _uptr(a: np.ndarray):
    return a.ctypes.data_as(_ucp)


# ----------------------------------------------------------------------------------------------------------------
# Force classes (the reference's public API)
# ----------------------------------------------------------------------------------------------------------------

class InteraxPairForce:
    """Base of the pair forces (InteraxPairForce.h:70-239; defaults InteraxPairForce.cpp:41-43)."""
    NoCutoff, CutoffNonPeriodic, CutoffPeriodic = 0, 1, 2

    def __init__(self):
        self._method = self.NoCutoff
        self._cutoff = 1.0
        self._switching_distance = -1.0
        self._softcore_radius = 0.0
        self._use_switch = False
        self._alchemical_name = "lambda_sterics"
        self._exclusions: List[Tuple[int, int]] = []
        self._group = 0
        self._name = self.NAME

    # settings
    def getNonbondedMethod(self): return self._method

    def setNonbondedMethod(self, method):
        if method < 0 or method > 2:
            raise InteraxError("InteraxPairForce: Illegal value for nonbonded method")
        self._method = int(method)

    def getCutoffDistance(self): return self._cutoff
    def setCutoffDistance(self, d): self._cutoff = float(d)
    def getUseSwitchingFunction(self): return self._use_switch
    def setUseSwitchingFunction(self, use): self._use_switch = bool(use)
    def getSwitchingDistance(self): return self._switching_distance
    def setSwitchingDistance(self, d): self._switching_distance = float(d)
    def getNumExclusions(self): return len(self._exclusions)

    def addExclusion(self, p1, p2):
        self._exclusions.append((int(p1), int(p2)))
        return len(self._exclusions)-1

    def getExclusionParticles(self, index): return self._exclusions[index]
    def setExclusionParticles(self, index, p1, p2): self._exclusions[index] = (int(p1), int(p2))
    def getAlchemicalParameterName(self): return self._alchemical_name
    def setAlchemicalParameterName(self, name): self._alchemical_name = str(name)
    def getSoftcoreRadius(self): return self._softcore_radius

    def setSoftcoreRadius(self, r):
        if r < 0:
            raise InteraxError("InteraxPairForce: The soft-core radius cannot be negative")
        self._softcore_radius = float(r)

    def usesPeriodicBoundaryConditions(self): return self._method == self.CutoffPeriodic
    def getForceGroup(self): return self._group
    def setForceGroup(self, g): self._group = int(g)
    def getName(self): return self._name
    def setName(self, name): self._name = str(name)

    # common per-particle views
    def _alchemical_array(self) -> np.ndarray:
        return np.array([self.getParticleIsAlchemical(i) for i in range(self.getNumParticles())], dtype=np.uint8)

    def _exclusion_array(self) -> np.ndarray:
        return np.array(self._exclusions, dtype=np.int32).reshape(-1, 2)


class InteraxDispersionForce(InteraxPairForce):
    """InteraxDispersionForce.h:70-182 (defaults InteraxDispersionForce.cpp:39-40)."""
    NAME = "InteraxDispersionForce"
    TERM = IXH_TERM_DISPERSION

    def __init__(self):
        super().__init__()
        self._particles: List[Tuple[float, float, float, float, float, bool]] = []
        self._damping_scale = 1.0
        self._use_lrc = False

    def getNumParticles(self): return len(self._particles)

    def addParticle(self, m1, m2, m3, alpha, width, isAlchemical=False):
        self._particles.append((float(m1), float(m2), float(m3), float(alpha), float(width), bool(isAlchemical)))
        return len(self._particles)-1

    def getParticleParameters(self, index): return self._particles[index]

    def setParticleParameters(self, index, m1, m2, m3, alpha, width, isAlchemical):
        self._particles[index] = (float(m1), float(m2), float(m3), float(alpha), float(width), bool(isAlchemical))

    def getParticleIsAlchemical(self, index): return self._particles[index][5]
    def getDampingScale(self): return self._damping_scale
    def setDampingScale(self, s): self._damping_scale = float(s)
    def getUseLongRangeCorrection(self): return self._use_lrc
    def setUseLongRangeCorrection(self, use): self._use_lrc = bool(use)

    def parameter_array(self) -> np.ndarray:
        return np.array([p[:5] for p in self._particles], dtype=np.float64).reshape(-1, 5)


class _OverlapForce(InteraxPairForce):
    Exact, InteraxOrder2 = 0, 2

    def __init__(self):
        super().__init__()
        self._overlap_method = self.Exact

    def getOverlapMethod(self): return self._overlap_method

    def setOverlapMethod(self, method):
        if method not in (self.Exact, self.InteraxOrder2):
            raise InteraxError(f"{self.NAME}: Illegal value for overlap method")
        self._overlap_method = int(method)


class InteraxExchangeForce(_OverlapForce):
    """InteraxExchangeForce.h:60-160."""
    NAME = "InteraxExchangeForce"
    TERM = IXH_TERM_EXCHANGE

    def __init__(self):
        super().__init__()
        self._particles: List[Tuple[float, float, bool]] = []

    def getNumParticles(self): return len(self._particles)

    def addParticle(self, amplitude, width, isAlchemical=False):
        self._particles.append((float(amplitude), float(width), bool(isAlchemical)))
        return len(self._particles)-1

    def getParticleParameters(self, index): return self._particles[index]

    def setParticleParameters(self, index, amplitude, width, isAlchemical):
        self._particles[index] = (float(amplitude), float(width), bool(isAlchemical))

    def getParticleIsAlchemical(self, index): return self._particles[index][2]

    def parameter_array(self) -> np.ndarray:
        return np.array([p[:2] for p in self._particles], dtype=np.float64).reshape(-1, 2)


class InteraxShortRangeForce(_OverlapForce):
    """InteraxShortRangeForce.h:59-162."""
    NAME = "InteraxShortRangeForce"
    TERM = IXH_TERM_SHORT_RANGE

    def __init__(self):
        super().__init__()
        self._particles: List[Tuple[float, float, float, bool]] = []

    def getNumParticles(self): return len(self._particles)

    def addParticle(self, k, l, width, isAlchemical=False):
        self._particles.append((float(k), float(l), float(width), bool(isAlchemical)))
        return len(self._particles)-1

    def getParticleParameters(self, index): return self._particles[index]

    def setParticleParameters(self, index, k, l, width, isAlchemical):
        self._particles[index] = (float(k), float(l), float(width), bool(isAlchemical))

    def getParticleIsAlchemical(self, index): return self._particles[index][3]

    def parameter_array(self) -> np.ndarray:
        return np.array([p[:3] for p in self._particles], dtype=np.float64).reshape(-1, 3)


class InteraxThreeBodyForce:
    """InteraxThreeBodyForce.h:75-258 (defaults InteraxThreeBodyForce.cpp:39-41)."""
    NAME = "InteraxThreeBodyForce"
    NoDamping, TangToennies = 0, 1

    def __init__(self):
        self._sites: List[Tuple[List[int], float, float, float, int, bool]] = []
        self._cutoff = 1.2
        self._scale = 1.0
        self._damping_scale = 1.0
        self._centroid_radius = 0.0
        self._damping_method = self.NoDamping
        self._periodic = False
        self._alchemical_name = "lambda_sterics"
        self._group = 0
        self._name = self.NAME

    def getNumSites(self): return len(self._sites)

    def addSite(self, particles, alpha, frequency, width=0.0, molecule=-1, isAlchemical=False):
        self._sites.append((list(map(int, particles)), float(alpha), float(frequency), float(width), int(molecule),
                            bool(isAlchemical)))
        return len(self._sites)-1

    def getSiteParameters(self, index): return self._sites[index]

    def setSiteParameters(self, index, particles, alpha, frequency, width, molecule, isAlchemical):
        self._sites[index] = (list(map(int, particles)), float(alpha), float(frequency), float(width), int(molecule),
                              bool(isAlchemical))

    def getCutoffDistance(self): return self._cutoff
    def setCutoffDistance(self, d): self._cutoff = float(d)
    def getScaleFactor(self): return self._scale
    def setScaleFactor(self, s): self._scale = float(s)
    def getDampingMethod(self): return self._damping_method

    def setDampingMethod(self, method):
        if method not in (self.NoDamping, self.TangToennies):
            raise InteraxError("InteraxThreeBodyForce: Illegal value for damping method")
        self._damping_method = int(method)

    def getDampingScale(self): return self._damping_scale
    def setDampingScale(self, s): self._damping_scale = float(s)
    def getCentroidDampingRadius(self): return self._centroid_radius

    def setCentroidDampingRadius(self, r):
        if r < 0:
            raise InteraxError("InteraxThreeBodyForce: The centroid damping radius cannot be negative")
        self._centroid_radius = float(r)

    def setUsesPeriodicBoundaryConditions(self, periodic): self._periodic = bool(periodic)
    def usesPeriodicBoundaryConditions(self): return self._periodic
    def getAlchemicalParameterName(self): return self._alchemical_name
    def setAlchemicalParameterName(self, name): self._alchemical_name = str(name)
    def getForceGroup(self): return self._group
    def setForceGroup(self, g): self._group = int(g)
    def getName(self): return self._name
    def setName(self, name): self._name = str(name)

    def site_arrays(self):
        """(offsets, particles, alpha, frequency, width, molecule, alchemical) as numpy arrays."""
        offsets = np.zeros(len(self._sites)+1, dtype=np.int32)
        parts: List[int] = []
        for i, s in enumerate(self._sites):
            parts.extend(s[0])
            offsets[i+1] = len(parts)
        col = lambda k, dt: np.array([s[k] for s in self._sites], dtype=dt)
        return (offsets, np.array(parts, dtype=np.int32), col(1, np.float64), col(2, np.float64), col(3, np.float64),
                col(4, np.int32), col(5, np.uint8))


class System:
    """The part of OpenMM's System the plugin needs: particle count, default box, forces."""

    def __init__(self, num_particles: int = 0, box: Optional[Sequence[float]] = None):
        self._masses: List[float] = [1.0]*int(num_particles)
        self._forces: List[object] = []
        self._box = np.array(box if box is not None else [2, 0, 0, 0, 2, 0, 0, 0, 2], dtype=np.float64)

    def addParticle(self, mass=1.0):
        self._masses.append(float(mass))
        return len(self._masses)-1

    def getNumParticles(self): return len(self._masses)
    def getParticleMass(self, index): return self._masses[index]
    def setParticleMass(self, index, mass): self._masses[index] = float(mass)

    def addForce(self, force):
        self._forces.append(force)
        return len(self._forces)-1

    def getForces(self): return list(self._forces)
    def getNumForces(self): return len(self._forces)
    def getForce(self, i): return self._forces[i]

    def setDefaultPeriodicBoxVectors(self, a, b, c):
        self._box = np.array(list(a)+list(b)+list(c), dtype=np.float64)

    def getDefaultPeriodicBoxVectors(self): return self._box.reshape(3, 3).copy()


# ----------------------------------------------------------------------------------------------------------------
# XML (the version 1 format of the reference's serialization proxies, serialization/src/*Proxy.cpp)
# ----------------------------------------------------------------------------------------------------------------

def _b(node, name):
    return int(node.get(name)) != 0


def _fmt(v) -> str:
    """A double as text that reads back to the same value (shortest round-trip form)."""
    return repr(float(v))


def _set(node, **attrs):
    for k, v in attrs.items():
        if isinstance(v, bool):
            node.set(k, "1" if v else "0")
        elif isinstance(v, float):
            node.set(k, _fmt(v))
        else:
            node.set(k, str(v))
    return node


def _serialize_force(force, tag: str = "Force") -> ET.Element:
    """The version-1 node of one Interax force: Interax{Dispersion,Exchange,ShortRange,ThreeBody}ForceProxy.cpp
    serialize (same attribute and child names, values as written by those proxies)."""
    node = ET.Element(tag)
    if isinstance(force, InteraxThreeBodyForce):
        _set(node, type=force.NAME, version=1, forceGroup=force.getForceGroup(), name=force.getName(),
             cutoff=force.getCutoffDistance(), scaleFactor=force.getScaleFactor(),
             usesPeriodic=force.usesPeriodicBoundaryConditions(), alchemicalParameter=force.getAlchemicalParameterName(),
             dampingMethod=force.getDampingMethod(), dampingScale=force.getDampingScale(),
             centroidDampingRadius=force.getCentroidDampingRadius())
        sites = ET.SubElement(node, "Sites")
        for particles, alpha, frequency, width, molecule, alchemical in force._sites:
            site = _set(ET.SubElement(sites, "Site"), alpha=alpha, frequency=frequency, width=width, molecule=molecule,
                        alchemical=alchemical)
            for q in particles:
                _set(ET.SubElement(site, "Particle"), index=q)
        return node
    if not isinstance(force, InteraxPairForce):
        raise InteraxError(f"XmlSerializer: no serialization proxy for {type(force).__name__}")
    _set(node, type=force.NAME, version=1, forceGroup=force.getForceGroup(), name=force.getName(),
         method=force.getNonbondedMethod(), cutoff=force.getCutoffDistance())
    if isinstance(force, _OverlapForce):
        _set(node, overlapMethod=force.getOverlapMethod())
    _set(node, useSwitchingFunction=force.getUseSwitchingFunction(), switchingDistance=force.getSwitchingDistance(),
         alchemicalParameter=force.getAlchemicalParameterName(), softcoreRadius=force.getSoftcoreRadius())
    if isinstance(force, InteraxDispersionForce):
        _set(node, dampingScale=force.getDampingScale(), useLongRangeCorrection=force.getUseLongRangeCorrection())
    particles = ET.SubElement(node, "Particles")
    for p in force._particles:
        q = ET.SubElement(particles, "Particle")
        if isinstance(force, InteraxDispersionForce):
            _set(q, m1=p[0], m2=p[1], m3=p[2], alpha=p[3], width=p[4], alchemical=p[5])
        elif isinstance(force, InteraxExchangeForce):
            _set(q, amplitude=p[0], width=p[1], alchemical=p[2])
        else:
            _set(q, k=p[0], l=p[1], width=p[2], alchemical=p[3])
    exclusions = ET.SubElement(node, "Exclusions")
    for p1, p2 in force._exclusions:
        _set(ET.SubElement(exclusions, "Exclusion"), p1=p1, p2=p2)
    return node


_FORCE_TYPES = ("InteraxDispersionForce", "InteraxExchangeForce", "InteraxShortRangeForce", "InteraxThreeBodyForce")


def _deserialize_force(node):
    """Interax{Dispersion,Exchange,ShortRange,ThreeBody}ForceProxy.cpp deserialize: the version check, then every
    attribute the proxy writes ("forceGroup" defaults to 0 and "name" to the class name, as there)."""
    kind = node.get("type")
    if int(node.get("version")) != 1:
        raise InteraxError("Unsupported version number")
    if kind == "InteraxThreeBodyForce":
        f = InteraxThreeBodyForce()
        f.setForceGroup(int(node.get("forceGroup", 0)))
        f.setName(node.get("name", f.getName()))
        f.setCutoffDistance(float(node.get("cutoff")))
        f.setScaleFactor(float(node.get("scaleFactor")))
        f.setUsesPeriodicBoundaryConditions(_b(node, "usesPeriodic"))
        f.setAlchemicalParameterName(node.get("alchemicalParameter"))
        f.setDampingMethod(int(node.get("dampingMethod")))
        f.setDampingScale(float(node.get("dampingScale")))
        f.setCentroidDampingRadius(float(node.get("centroidDampingRadius")))
        for site in node.find("Sites"):
            f.addSite([int(p.get("index")) for p in site], float(site.get("alpha")), float(site.get("frequency")),
                      float(site.get("width")), int(site.get("molecule")), _b(site, "alchemical"))
        return f
    f = {"InteraxDispersionForce": InteraxDispersionForce, "InteraxExchangeForce": InteraxExchangeForce,
         "InteraxShortRangeForce": InteraxShortRangeForce}[kind]()
    f.setForceGroup(int(node.get("forceGroup", 0)))
    f.setName(node.get("name", f.getName()))
    f.setNonbondedMethod(int(node.get("method")))
    f.setCutoffDistance(float(node.get("cutoff")))
    f.setUseSwitchingFunction(_b(node, "useSwitchingFunction"))
    f.setSwitchingDistance(float(node.get("switchingDistance")))
    f.setAlchemicalParameterName(node.get("alchemicalParameter"))
    f.setSoftcoreRadius(float(node.get("softcoreRadius")))
    if kind == "InteraxDispersionForce":
        f.setDampingScale(float(node.get("dampingScale")))
        f.setUseLongRangeCorrection(_b(node, "useLongRangeCorrection"))
        for p in node.find("Particles"):
            f.addParticle(float(p.get("m1")), float(p.get("m2")), float(p.get("m3")), float(p.get("alpha")),
                          float(p.get("width")), _b(p, "alchemical"))
    elif kind == "InteraxExchangeForce":
        f.setOverlapMethod(int(node.get("overlapMethod")))
        for p in node.find("Particles"):
            f.addParticle(float(p.get("amplitude")), float(p.get("width")), _b(p, "alchemical"))
    else:
        f.setOverlapMethod(int(node.get("overlapMethod")))
        for p in node.find("Particles"):
            f.addParticle(float(p.get("k")), float(p.get("l")), float(p.get("width")), _b(p, "alchemical"))
    for e in node.find("Exclusions"):
        f.addExclusion(int(e.get("p1")), int(e.get("p2")))
    return f


def _serialize_system(system: "System") -> ET.Element:
    """OpenMM's System node (version 1): box, particle masses, the Interax forces (other force types are not part of
    this platform and are not written)."""
    root = _set(ET.Element("System"), type="System", version=1)
    box = ET.SubElement(root, "PeriodicBoxVectors")
    b = system.getDefaultPeriodicBoxVectors()
    for k, row in zip("ABC", b):
        _set(ET.SubElement(box, k), x=float(row[0]), y=float(row[1]), z=float(row[2]))
    particles = ET.SubElement(root, "Particles")
    for m in system._masses:
        _set(ET.SubElement(particles, "Particle"), mass=m)
    ET.SubElement(root, "Constraints")
    forces = ET.SubElement(root, "Forces")
    for f in system.getForces():
        forces.append(_serialize_force(f))
    return root


def _deserialize_system(root) -> "System":
    box = root.find("PeriodicBoxVectors")
    vec = [float(box.find(k).get(c)) for k in "ABC" for c in "xyz"] if box is not None else None
    system = System(0, vec)
    for p in root.find("Particles"):
        system.addParticle(float(p.get("mass", 1.0)))
    for node in root.find("Forces"):
        if node.get("type") in _FORCE_TYPES:
            system.addForce(_deserialize_force(node))
    return system


class XmlSerializer:
    """OpenMM's XmlSerializer for the objects of this platform: a System or one of the four Interax forces.
    serialize() writes the version-1 format of the reference's proxies (serialization/src/*Proxy.cpp); deserialize()
    reads it back (and reads Systems written by OpenMM, ignoring other force types)."""

    @staticmethod
    def serialize(obj) -> str:
        node = _serialize_system(obj) if isinstance(obj, System) else _serialize_force(obj)
        ET.indent(node, space="\t")
        return ET.tostring(node, encoding="unicode")+"\n"

    @staticmethod
    def deserialize(text: str):
        root = ET.fromstring(text)
        if root.tag == "System" or root.get("type") == "System":
            return _deserialize_system(root)
        if root.get("type") not in _FORCE_TYPES:
            raise InteraxError(f"XmlSerializer: no serialization proxy for type {root.get('type')}")
        return _deserialize_force(root)


def load_system_xml(source) -> System:
    """Read a System written by OpenMM's XmlSerializer with the four Interax forces (others are ignored).

    source: a path (optionally .gz) or XML text."""
    if isinstance(source, str) and not source.lstrip().startswith("<"):
        opener = gzip.open if source.endswith(".gz") else open
        with opener(source, "rt") as f:
            text = f.read()
    else:
        text = source
    return _deserialize_system(ET.fromstring(text))


def save_system_xml(system: System, path: str) -> None:
    """Write `system` with XmlSerializer.serialize (gzip-compressed when path ends in .gz)."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wt") as f:
        f.write(XmlSerializer.serialize(system))


# ----------------------------------------------------------------------------------------------------------------
# Context
# ----------------------------------------------------------------------------------------------------------------

def check_pair_force(force: InteraxPairForce, system: System) -> None:
    """InteraxPairForce.cpp:113-136 (checkPairForce)."""
    name = force.NAME
    n = force.getNumParticles()
    if n != system.getNumParticles():
        raise InteraxError(f"{name} must have exactly as many particles as the System it belongs to.")
    for p1, p2 in force._exclusions:
        if p1 < 0 or p1 >= n or p2 < 0 or p2 >= n:
            raise InteraxError(f"{name}: Illegal particle index for an exclusion")
    if force.getNonbondedMethod() == InteraxPairForce.NoCutoff:
        return
    if force.getCutoffDistance() <= 0:
        raise InteraxError(f"{name}: The cutoff distance must be positive")
    if force.getUseSwitchingFunction() and (force.getSwitchingDistance() < 0 or
                                            force.getSwitchingDistance() >= force.getCutoffDistance()):
        raise InteraxError(f"{name}: Switching distance must satisfy 0 <= r_switch < r_cutoff")
    if force.getNonbondedMethod() == InteraxPairForce.CutoffPeriodic:
        b = system.getDefaultPeriodicBoxVectors()
        c = force.getCutoffDistance()
        if c > 0.5*b[0, 0] or c > 0.5*b[1, 1] or c > 0.5*b[2, 2]:
            raise InteraxError(f"{name}: The cutoff distance cannot be greater than half the periodic box size.")


def _fusion_key(force: InteraxPairForce, precision: int):
    excl = force._exclusion_array()
    excl = np.sort(np.sort(excl, axis=1), axis=0) if len(excl) else excl
    ex_key = tuple(map(tuple, np.unique(excl, axis=0))) if len(excl) else ()
    return (force.getNonbondedMethod(), force.getCutoffDistance(), force.getUseSwitchingFunction(),
            force.getSwitchingDistance(), force.getSoftcoreRadius(), force.getAlchemicalParameterName(),
            ex_key, force._alchemical_array().tobytes(), precision)


class _PairGroup:
    """Pair forces sharing geometry settings: one ixh_pair handle, one kernel launch per evaluation."""

    def __init__(self, device: int, system: System, forces: List[InteraxPairForce], precision: int):
        self.forces = forces
        self.device = device
        self.system = system
        self.precision = precision
        self.handle = ctypes.c_void_p()
        desc, self._keep = self._describe()
        _check(_lib.ixh_pair_create(device, ctypes.byref(desc), ctypes.byref(self.handle)))

    def _describe(self):
        f0 = self.forces[0]
        n = f0.getNumParticles()
        keep = []
        d = PairDesc()
        d.num_particles = n
        d.terms = 0
        d.method = f0.getNonbondedMethod()
        d.cutoff = f0.getCutoffDistance()
        d.use_switching_function = int(f0.getUseSwitchingFunction())
        d.switching_distance = f0.getSwitchingDistance()
        d.softcore_radius = f0.getSoftcoreRadius()
        alch = f0._alchemical_array()
        excl = np.ascontiguousarray(f0._exclusion_array().reshape(-1), dtype=np.int32)
        box = np.ascontiguousarray(self.system.getDefaultPeriodicBoxVectors().reshape(-1))
        keep += [alch, excl, box]
        d.alchemical = _uptr(alch)
        d.num_exclusions = len(excl)//2
        d.exclusions = _iptr(excl)
        d.precision = self.precision
        d.default_box = _dptr(box)
        for f in self.forces:
            d.terms |= f.TERM
            arr = np.ascontiguousarray(f.parameter_array().reshape(-1))
            keep.append(arr)
            if isinstance(f, InteraxDispersionForce):
                d.dispersion = _dptr(arr)
                d.damping_scale = f.getDampingScale()
                d.use_long_range_correction = int(f.getUseLongRangeCorrection())
            elif isinstance(f, InteraxExchangeForce):
                d.exchange = _dptr(arr)
                d.exchange_overlap_method = f.getOverlapMethod()
            else:
                d.short_range = _dptr(arr)
                d.short_range_overlap_method = f.getOverlapMethod()
        return d, keep

    def update(self):
        desc, keep = self._describe()
        _check(_lib.ixh_pair_set_parameters(self.handle, ctypes.byref(desc)))

    def close(self):
        if self.handle:
            _lib.ixh_pair_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class _ThreeBody:
    def __init__(self, device: int, system: System, force: InteraxThreeBodyForce, precision: int):
        self.force = force
        self.system = system
        self.precision = precision
        self.handle = ctypes.c_void_p()
        desc, self._keep = self._describe()
        _check(_lib.ixh_threebody_create(device, ctypes.byref(desc), ctypes.byref(self.handle)))

    def _describe(self):
        f = self.force
        arrays = [np.ascontiguousarray(a) for a in f.site_arrays()]
        offsets, parts, alpha, freq, width, mol, alch = arrays
        d = ThreeBodyDesc()
        d.num_particles = self.system.getNumParticles()
        d.num_sites = f.getNumSites()
        d.site_offsets = _iptr(offsets)
        d.site_particles = _iptr(parts)
        d.alpha = _dptr(alpha)
        d.frequency = _dptr(freq)
        d.width = _dptr(width)
        d.molecule = _iptr(mol)
        d.alchemical = _uptr(alch)
        d.cutoff = f.getCutoffDistance()
        d.scale_factor = f.getScaleFactor()
        d.damping_method = f.getDampingMethod()
        d.damping_scale = f.getDampingScale()
        d.centroid_damping_radius = f.getCentroidDampingRadius()
        d.periodic = int(f.usesPeriodicBoundaryConditions())
        d.precision = self.precision
        return d, arrays

    def update(self):
        desc, keep = self._describe()
        _check(_lib.ixh_threebody_set_parameters(self.handle, ctypes.byref(desc)))

    def close(self):
        if self.handle:
            _lib.ixh_threebody_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class State:
    def __init__(self, energy: float, forces: Optional[np.ndarray]):
        self._energy = energy
        self._forces = forces

    def getPotentialEnergy(self) -> float: return self._energy
    def getForces(self) -> np.ndarray: return self._forces


class Context:
    """A System on one MI355X: device buffers plus one kernel handle per fused pair group and three-body force.

    properties: {"Precision": "single" | "mixed" | "double", "DeviceIndex": "0"} as OpenMM's GPU platforms."""

    def __init__(self, system: System, properties: Optional[Dict[str, str]] = None):
        load_library()
        properties = properties or {}
        self.system = system
        self.precision = PRECISIONS[properties.get("Precision", "mixed")]
        self.device = int(properties.get("DeviceIndex", "0"))
        n = system.getNumParticles()
        self._parameters: Dict[str, float] = {}
        self._atom_index: Optional[int] = None
        self._groups: List[_PairGroup] = []
        self._threebody: List[_ThreeBody] = []
        self._state = ctypes.c_void_p()
        pair_forces = [f for f in system.getForces() if isinstance(f, InteraxPairForce)]
        for f in pair_forces:
            check_pair_force(f, system)
        for f in system.getForces():
            if isinstance(f, InteraxThreeBodyForce):
                self._validate_threebody(f, n)
        # default context parameters (InteraxPairForce.cpp:138-146, InteraxThreeBodyForceImpl.cpp:70-84)
        for f in system.getForces():
            if isinstance(f, InteraxPairForce) and any(f.getParticleIsAlchemical(i) for i in range(f.getNumParticles())):
                self._parameters[f.getAlchemicalParameterName()] = 1.0
            if isinstance(f, InteraxThreeBodyForce) and any(s[5] for s in f._sites):
                self._parameters[f.getAlchemicalParameterName()] = 1.0
        _check(_lib.ixh_state_create(self.device, n, self.precision, ctypes.byref(self._state)))
        # fuse pair forces with identical geometry settings (one force of each class per group)
        groups: List[Tuple[object, List[InteraxPairForce]]] = []
        for f in pair_forces:
            key = _fusion_key(f, self.precision)
            for k, members in groups:
                if k == key and all(type(m) is not type(f) for m in members):
                    members.append(f)
                    break
            else:
                groups.append((key, [f]))
        for _, members in groups:
            self._groups.append(_PairGroup(self.device, system, members, self.precision))
        for f in system.getForces():
            if isinstance(f, InteraxThreeBodyForce):
                self._threebody.append(_ThreeBody(self.device, system, f, self.precision))
        self._box = system.getDefaultPeriodicBoxVectors().reshape(-1).copy()
        _check(_lib.ixh_state_set_box(self._state, _dptr(self._box)))

    @staticmethod
    def _validate_threebody(f: InteraxThreeBodyForce, n: int) -> None:
        """InteraxThreeBodyForceImpl.cpp:42-59."""
        for particles, alpha, freq, width, mol, alch in f._sites:
            if not particles:
                raise InteraxError("InteraxThreeBodyForce: A site must contain at least one particle")
            if any(p < 0 or p >= n for p in particles):
                raise InteraxError("InteraxThreeBodyForce: Illegal particle index for a site")
            if f.getDampingMethod() == InteraxThreeBodyForce.TangToennies and width <= 0:
                raise InteraxError("InteraxThreeBodyForce: The Tang-Toennies damping requires a positive width for every site")
        if f.getCutoffDistance() <= 0:
            raise InteraxError("InteraxThreeBodyForce: The cutoff distance must be positive")

    def setPositions(self, positions) -> None:
        p = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
        if p.size != 3*self.system.getNumParticles():
            raise InteraxError("Called setPositions() on a Context with the wrong number of positions")
        _check(_lib.ixh_state_set_positions(self._state, _dptr(p)))

    def setPeriodicBoxVectors(self, a, b, c) -> None:
        self._box = np.array(list(a)+list(b)+list(c), dtype=np.float64)
        _check(_lib.ixh_state_set_box(self._state, _dptr(self._box)))

    def setParameter(self, name: str, value: float) -> None:
        if name not in self._parameters:
            raise InteraxError(f"Called setParameter() with invalid parameter name: {name}")
        self._parameters[name] = float(value)

    def getParameter(self, name: str) -> float:
        return self._parameters[name]

    def _buffers(self, lam: float) -> Buffers:
        b = Buffers()
        _check(_lib.ixh_state_get_buffers(self._state, ctypes.byref(b)))
        b.lambda_ = lam
        b.atom_index = self._atom_index
        return b

    def _set_atom_index(self, device_pointer: Optional[int]) -> None:
        """Run as a platform that has reordered its particles: positions and forces are in that platform's current
        order and device_pointer (an int device array, or None for the identity) holds the System index of every slot,
        as OpenMM's ComputeContext::getAtomIndexArray.  For adapters and tests of reordered contexts."""
        self._atom_index = device_pointer

    def getState(self, getEnergy: bool = True, getForces: bool = True, groups: int = 0xFFFFFFFF) -> State:
        _check(_lib.ixh_state_clear(self._state))
        flags = IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY
        host_energy = 0.0
        for g in self._groups:
            terms = 0
            for f in g.forces:
                if groups & (1 << f.getForceGroup()):
                    terms |= f.TERM
            if terms == 0:
                continue
            lam = self._parameters.get(g.forces[0].getAlchemicalParameterName(), 1.0)
            b = self._buffers(lam)
            _check(_lib.ixh_pair_execute(g.handle, ctypes.byref(b), terms, flags))
            if terms & IXH_TERM_DISPERSION:
                e = ctypes.c_double()
                _check(_lib.ixh_pair_long_range_energy(g.handle, _dptr(self._box), lam, ctypes.byref(e)))
                host_energy += e.value
        for t in self._threebody:
            if groups & (1 << t.force.getForceGroup()):
                lam = self._parameters.get(t.force.getAlchemicalParameterName(), 1.0)
                b = self._buffers(lam)
                _check(_lib.ixh_threebody_execute(t.handle, ctypes.byref(b), flags))
        energy = ctypes.c_double()
        _check(_lib.ixh_state_get_energy(self._state, ctypes.byref(energy)))
        forces = None
        if getForces:
            forces = np.zeros((self.system.getNumParticles(), 3))
            _check(_lib.ixh_state_get_forces(self._state, _dptr(forces)))
        return State(energy.value+host_energy, forces)

    def getEnergiesAtLambdas(self, values, parameter: str = "lambda_sterics", groups: int = 0xFFFFFFFF) -> np.ndarray:
        """Potential energy of the current positions with the context parameter `parameter` set to each of `values`
        (lambda-window fusion; an extension without a counterpart in the reference, where each value needs its own
        getState()).  Pair forces: one evaluation plus the pairs of the alchemical particles at every value
        (ixh_pair_lambda_energies), and the long range correction at every value; three-body forces: the energy is
        linear in lambda, evaluated at 0 and 1.  Forces of other parameters are taken at their current values.  The
        Context's own force and energy buffers are used as scratch."""
        values = np.asarray(values, dtype=np.float64).reshape(-1)
        out = np.zeros(values.size)
        for g in self._groups:
            terms = 0
            for f in g.forces:
                if groups & (1 << f.getForceGroup()):
                    terms |= f.TERM
            if terms == 0:
                continue
            name = g.forces[0].getAlchemicalParameterName()
            lam = self._parameters.get(name, 1.0)
            b = self._buffers(lam)
            for start in range(0, values.size, 16):
                chunk = values[start:start+16] if name == parameter else np.full(min(16, values.size-start), lam)
                chunk = np.ascontiguousarray(chunk)
                e = np.zeros(chunk.size)
                _check(_lib.ixh_pair_lambda_energies(g.handle, ctypes.byref(b), terms, _dptr(chunk), chunk.size, _dptr(e)))
                out[start:start+chunk.size] += e
                if terms & IXH_TERM_DISPERSION:
                    for k, v in enumerate(chunk):
                        lrc = ctypes.c_double()
                        _check(_lib.ixh_pair_long_range_energy(g.handle, _dptr(self._box), float(v), ctypes.byref(lrc)))
                        out[start+k] += lrc.value
        for t in self._threebody:
            if not groups & (1 << t.force.getForceGroup()):
                continue
            name = t.force.getAlchemicalParameterName()
            ends = []
            for lam in ((0.0, 1.0) if name == parameter else (self._parameters.get(name, 1.0),)*2):
                _check(_lib.ixh_state_clear(self._state))
                b = self._buffers(lam)
                _check(_lib.ixh_threebody_execute(t.handle, ctypes.byref(b), IXH_INCLUDE_ENERGY))
                e = ctypes.c_double()
                _check(_lib.ixh_state_get_energy(self._state, ctypes.byref(e)))
                ends.append(e.value)
            out += ends[0]+values*(ends[1]-ends[0]) if name == parameter else ends[0]
        return out

    def updateParametersInContext(self, force) -> None:
        """Force.updateParametersInContext (InteraxPairForceImpl.h:97-100, InteraxThreeBodyForceImpl.cpp:92-95)."""
        for g in self._groups:
            if any(f is force for f in g.forces):
                if force.getNumParticles() != g.forces[0].getNumParticles() and len(g.forces) == 1:
                    raise InteraxError("updateParametersInContext: The number of particles has changed")
                g.update()
                return
        for t in self._threebody:
            if t.force is force:
                t.update()
                return
        raise InteraxError("updateParametersInContext: the force is not part of this Context")

    def pair_counts(self) -> List[int]:
        out = []
        for g in self._groups:
            c = ctypes.c_longlong()
            _check(_lib.ixh_pair_last_pair_count(g.handle, ctypes.byref(c)))
            out.append(c.value)
        return out

    def triple_counts(self) -> List[int]:
        out = []
        for t in self._threebody:
            c = ctypes.c_longlong()
            _check(_lib.ixh_threebody_last_triple_count(t.handle, ctypes.byref(c)))
            out.append(c.value)
        return out

    def close(self) -> None:
        for g in self._groups:
            g.close()
        for t in self._threebody:
            t.close()
        if self._state:
            _lib.ixh_state_destroy(self._state)
            self._state = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent nonperiodic Systems evaluated together (one segmented pair handle, one segmented three-body
    handle, one device state, graph replay).  Every System must have the same force layout: pair forces with the
    NoCutoff method and identical settings, and optionally one nonperiodic InteraxThreeBodyForce with identical
    settings.  evaluate() returns the energy of every System and the forces of all particles, System after System."""

    def __init__(self, systems: Sequence["System"], precision: str = "mixed", device: int = 0):
        load_library()
        if not systems:
            raise InteraxError("Batch: no systems")
        self.systems = list(systems)
        self.precision = PRECISIONS[precision]
        counts = [s.getNumParticles() for s in self.systems]
        self.particle_offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.num_particles = int(self.particle_offsets[-1])
        keep = []
        pair_desc = self._pair_desc(keep)
        tb_desc, site_offsets = self._threebody_desc(keep)
        d = BatchDesc()
        d.num_systems = len(self.systems)
        d.particle_offsets = _iptr(self.particle_offsets)
        d.pair = ctypes.pointer(pair_desc) if pair_desc is not None else None
        d.site_offsets = _iptr(site_offsets) if site_offsets is not None else None
        d.threebody = ctypes.pointer(tb_desc) if tb_desc is not None else None
        d.precision = self.precision
        self.handle = ctypes.c_void_p()
        _check(_lib.ixh_batch_create(device, ctypes.byref(d), ctypes.byref(self.handle)))

    @staticmethod
    def _forces(system, kind):
        return [f for f in system.getForces() if isinstance(f, kind)]

    def _pair_desc(self, keep):
        first = self._forces(self.systems[0], InteraxPairForce)
        if not first:
            return None
        kinds = [type(f) for f in first]
        d = PairDesc()
        d.num_particles = self.num_particles
        d.terms = 0
        f0 = first[0]
        if f0.getNonbondedMethod() != InteraxPairForce.NoCutoff:
            raise InteraxError("Batch: the pair forces of a batch must use the NoCutoff method")
        d.method = f0.getNonbondedMethod()
        d.cutoff = f0.getCutoffDistance()
        d.softcore_radius = f0.getSoftcoreRadius()
        alch, excl = [], []
        params = {k: [] for k in kinds}
        for s, system in enumerate(self.systems):
            forces = self._forces(system, InteraxPairForce)
            if [type(f) for f in forces] != kinds:
                raise InteraxError("Batch: every System must have the same pair forces")
            offset = int(self.particle_offsets[s])
            alch.append(forces[0]._alchemical_array())
            ex = forces[0]._exclusion_array()
            for f in forces:
                check_pair_force(f, system)
                if f.getNonbondedMethod() != d.method or f.getSoftcoreRadius() != d.softcore_radius:
                    raise InteraxError("Batch: the pair forces of a batch must have identical settings")
                params[type(f)].append(f.parameter_array())
            excl.append(ex+offset)
        alch = np.ascontiguousarray(np.concatenate(alch), dtype=np.uint8)
        excl = np.ascontiguousarray(np.concatenate(excl).reshape(-1), dtype=np.int32)
        keep += [alch, excl]
        d.alchemical = _uptr(alch)
        d.num_exclusions = len(excl)//2
        d.exclusions = _iptr(excl)
        d.precision = self.precision
        d.default_box = None
        for f in first:
            arr = np.ascontiguousarray(np.concatenate(params[type(f)]).reshape(-1))
            keep.append(arr)
            d.terms |= f.TERM
            if isinstance(f, InteraxDispersionForce):
                d.dispersion = _dptr(arr)
                d.damping_scale = f.getDampingScale()
                d.use_long_range_correction = 0
            elif isinstance(f, InteraxExchangeForce):
                d.exchange = _dptr(arr)
                d.exchange_overlap_method = f.getOverlapMethod()
            else:
                d.short_range = _dptr(arr)
                d.short_range_overlap_method = f.getOverlapMethod()
        return d

    def _threebody_desc(self, keep):
        forces = [self._forces(s, InteraxThreeBodyForce) for s in self.systems]
        if not forces[0]:
            return None, None
        f0 = forces[0][0]
        site_counts, offsets, parts, alpha, freq, width, mol, alch = [], [], [], [], [], [], [], []
        site_base = 0
        part_base = 0
        for s, fl in enumerate(forces):
            if len(fl) != 1:
                raise InteraxError("Batch: every System must have one InteraxThreeBodyForce")
            f = fl[0]
            if (f.usesPeriodicBoundaryConditions() or f.getCutoffDistance() != f0.getCutoffDistance() or
                    f.getScaleFactor() != f0.getScaleFactor() or f.getDampingMethod() != f0.getDampingMethod() or
                    f.getDampingScale() != f0.getDampingScale() or
                    f.getCentroidDampingRadius() != f0.getCentroidDampingRadius()):
                raise InteraxError("Batch: the three-body forces of a batch must be nonperiodic with identical settings")
            o, p, a, w, wd, m, al = f.site_arrays()
            offsets.append(o[:-1]+part_base)
            part_base += len(p)
            parts.append(p+int(self.particle_offsets[s]))
            alpha.append(a)
            freq.append(w)
            width.append(wd)
            # negative molecule indices become -1 - (global site index), as a single System would number them
            m = m.copy()
            neg = m < 0
            m[neg] = -1-(site_base+np.nonzero(neg)[0])
            mol.append(m)
            alch.append(al)
            site_counts.append(f.getNumSites())
            site_base += f.getNumSites()
        site_offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(site_counts)]), dtype=np.int32)
        parts_all = np.ascontiguousarray(np.concatenate(parts), dtype=np.int32)
        offs = np.ascontiguousarray(np.concatenate(offsets+[[len(parts_all)]]), dtype=np.int32)
        arrays = [offs, parts_all] + [np.ascontiguousarray(np.concatenate(x)) for x in (alpha, freq, width)]
        arrays += [np.ascontiguousarray(np.concatenate(mol), dtype=np.int32), np.ascontiguousarray(np.concatenate(alch), dtype=np.uint8)]
        keep += arrays + [site_offsets]
        d = ThreeBodyDesc()
        d.num_particles = self.num_particles
        d.num_sites = int(site_offsets[-1])
        d.site_offsets = _iptr(arrays[0])
        d.site_particles = _iptr(arrays[1])
        d.alpha = _dptr(arrays[2])
        d.frequency = _dptr(arrays[3])
        d.width = _dptr(arrays[4])
        d.molecule = _iptr(arrays[5])
        d.alchemical = _uptr(arrays[6])
        d.cutoff = f0.getCutoffDistance()
        d.scale_factor = f0.getScaleFactor()
        d.damping_method = f0.getDampingMethod()
        d.damping_scale = f0.getDampingScale()
        d.centroid_damping_radius = f0.getCentroidDampingRadius()
        d.periodic = 0
        d.precision = self.precision
        return d, site_offsets

    def setPositions(self, positions: Sequence[np.ndarray]) -> None:
        """positions: one N_s x 3 array (nm) per System, or one array of all particles."""
        if isinstance(positions, np.ndarray) and positions.shape == (self.num_particles, 3):
            pos = positions
        else:
            pos = np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 3) for p in positions])
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        if pos.size != 3*self.num_particles:
            raise InteraxError("Batch.setPositions: wrong number of particles")
        _check(_lib.ixh_batch_set_positions(self.handle, _dptr(pos)))

    def execute(self, forces: bool = True) -> None:
        flags = IXH_INCLUDE_ENERGY | (IXH_INCLUDE_FORCES if forces else 0)
        _check(_lib.ixh_batch_execute(self.handle, flags))

    def energies(self) -> np.ndarray:
        e = np.zeros(len(self.systems))
        _check(_lib.ixh_batch_get_energies(self.handle, _dptr(e)))
        return e

    def forces(self) -> List[np.ndarray]:
        f = np.zeros(3*self.num_particles)
        _check(_lib.ixh_batch_get_forces(self.handle, _dptr(f)))
        f = f.reshape(-1, 3)
        return [f[self.particle_offsets[s]:self.particle_offsets[s+1]] for s in range(len(self.systems))]

    def evaluate(self, positions=None, forces: bool = True):
        if positions is not None:
            self.setPositions(positions)
        self.execute(forces)
        return self.energies(), (self.forces() if forces else None)

    def set_graph(self, enable: bool) -> None:
        _check(_lib.ixh_batch_set_graph(self.handle, int(enable)))

    def autotune(self) -> dict:
        """Time the launch shapes on the current positions and keep the fastest; returns the chosen shape."""
        chosen = BatchShape()
        _check(_lib.ixh_batch_autotune(self.handle, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY, ctypes.byref(chosen)))
        return {name: getattr(chosen, name) for name, _ in BatchShape._fields_}

    def shape(self) -> dict:
        s = BatchShape()
        _check(_lib.ixh_batch_get_shape(self.handle, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in BatchShape._fields_}

    def set_shape(self, pair_workgroups: int, triple_waves: int, triple_pairs_per_item: int) -> None:
        s = BatchShape(pair_workgroups, triple_waves, triple_pairs_per_item)
        _check(_lib.ixh_batch_set_shape(self.handle, ctypes.byref(s)))

    def close(self) -> None:
        if self.handle:
            _lib.ixh_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class Server:
    """The serving loop of the native library (its own thread): register Systems, start, then submit coordinates and
    wait for the results.  Cluster Systems (NoCutoff pair forces and an optional nonperiodic three-body force, all with
    the same settings) are evaluated in batches admitted as execution lanes and the memory budget free up; periodic
    Systems (cutoff methods, e.g. alchemical boxes evaluated at a given lambda) one request at a time."""

    def __init__(self, precision: str = "mixed", max_batch: int = 64, memory_budget: int = 8 << 30, device: int = 0):
        load_library()
        self.handle = ctypes.c_void_p()
        _check(_lib.ixh_server_create(device, PRECISIONS[precision], max_batch, memory_budget, ctypes.byref(self.handle)))
        self.precision = PRECISIONS[precision]
        self.sizes: List[int] = []
        self._keep = []
        self._windows = {}

    def add_system(self, system: "System", positions: Optional[np.ndarray] = None) -> int:
        """Register a System (also while the server runs); returns its entry index.  The positions of a periodic
        System, if given, are evaluated once at registration (neighbour lists built, buffers sized)."""
        pair_forces = [f for f in system.getForces() if isinstance(f, InteraxPairForce)]
        tbs = [f for f in system.getForces() if isinstance(f, InteraxThreeBodyForce)]
        if not pair_forces:
            raise InteraxError("Server.add_system: the System has no pair force")
        group = _PairGroup.__new__(_PairGroup)
        group.forces, group.system, group.precision = pair_forces, system, self.precision
        desc, keep = group._describe()
        self._keep.append(keep)
        entry = ctypes.c_int()
        if pair_forces[0].getNonbondedMethod() == InteraxPairForce.NoCutoff:
            tdesc = None
            if tbs:
                tb = _ThreeBody.__new__(_ThreeBody)
                tb.force, tb.system, tb.precision = tbs[0], system, self.precision
                tdesc, tkeep = tb._describe()
                self._keep.append(tkeep)
            _check(_lib.ixh_server_add_cluster(self.handle, ctypes.byref(desc), ctypes.byref(tdesc) if tdesc is not None else None,
                                               ctypes.byref(entry)))
        else:
            if tbs:
                raise InteraxError("Server.add_system: periodic Systems with a three-body force are not served")
            box = np.ascontiguousarray(system.getDefaultPeriodicBoxVectors().reshape(-1))
            self._keep.append(box)
            pos = None
            if positions is not None:
                pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
                if pos.size != 3*system.getNumParticles():
                    raise InteraxError("Server.add_system: wrong number of positions")
            _check(_lib.ixh_server_add_periodic(self.handle, ctypes.byref(desc), _dptr(box), _dptr(pos) if pos is not None else None,
                                                ctypes.byref(entry)))
        self.sizes.append(system.getNumParticles())
        return entry.value

    def start(self) -> None:
        _check(_lib.ixh_server_start(self.handle))

    def submit(self, entry: int, positions: np.ndarray, lam: float = 1.0) -> int:
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
        if pos.size != 3*self.sizes[entry]:
            raise InteraxError("Server.submit: wrong number of particles")
        ticket = ctypes.c_longlong()
        _check(_lib.ixh_server_submit(self.handle, entry, _dptr(pos), lam, ctypes.byref(ticket)))
        return ticket.value

    def submit_windows(self, entry: int, positions: np.ndarray, lambdas) -> int:
        """The energies of one configuration of a periodic System at up to 16 lambda values (fused evaluation)."""
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
        if pos.size != 3*self.sizes[entry]:
            raise InteraxError("Server.submit_windows: wrong number of particles")
        lam = np.ascontiguousarray(lambdas, dtype=np.float64)
        ticket = ctypes.c_longlong()
        _check(_lib.ixh_server_submit_windows(self.handle, entry, _dptr(pos), _dptr(lam), len(lam), ctypes.byref(ticket)))
        self._windows[ticket.value] = len(lam)
        return ticket.value

    def wait_windows(self, ticket: int):
        """(energies at the requested lambda values, latency in ms) of a window request."""
        energies = np.zeros(self._windows.pop(ticket))
        ms = ctypes.c_double()
        _check(_lib.ixh_server_wait_windows(self.handle, ticket, _dptr(energies), ctypes.byref(ms)))
        return energies, ms.value

    def memory(self) -> dict:
        resident, in_flight, peak = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        _check(_lib.ixh_server_memory(self.handle, ctypes.byref(resident), ctypes.byref(in_flight), ctypes.byref(peak)))
        return {"resident_bytes": resident.value, "in_flight_bytes": in_flight.value, "peak_in_flight_bytes": peak.value}

    def wait(self, ticket: int, entry: int):
        """(energy, forces, latency in ms) of a submitted request."""
        e = ctypes.c_double()
        ms = ctypes.c_double()
        f = np.zeros(3*self.sizes[entry])
        _check(_lib.ixh_server_wait(self.handle, ticket, ctypes.byref(e), _dptr(f), ctypes.byref(ms)))
        return e.value, f.reshape(-1, 3), ms.value

    def statistics(self):
        b, r = ctypes.c_longlong(), ctypes.c_longlong()
        _check(_lib.ixh_server_statistics(self.handle, ctypes.byref(b), ctypes.byref(r)))
        return b.value, r.value

    def replay(self, arrival_s: np.ndarray, entries: np.ndarray, positions: List[np.ndarray], lambdas=None, windows=None):
        """Submit requests at the given arrival times (seconds from now) from a native thread; returns the latency (ms)
        and energy of every request and the time to the last completion (s).  windows, optional: per request None
        (one lambda, forces) or the lambda values of a window request (the energy returned is the first window's)."""
        n = len(entries)
        arrival = np.ascontiguousarray(arrival_s, dtype=np.float64)
        ent = np.ascontiguousarray(entries, dtype=np.int32)
        lam = np.ascontiguousarray(np.ones(n) if lambdas is None else lambdas, dtype=np.float64)
        flat = [np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(-1)) for p in positions]
        offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum([f.size for f in flat])[:-1]]), dtype=np.int64)
        allpos = np.ascontiguousarray(np.concatenate(flat))
        counts = values = None
        if windows is not None:
            counts = np.ascontiguousarray([0 if w is None else len(w) for w in windows], dtype=np.int32)
            values = np.ascontiguousarray(np.concatenate([np.zeros(0)]+[np.asarray(w, dtype=np.float64) for w in windows
                                                                          if w is not None]), dtype=np.float64)
            if values.size == 0:
                values = np.zeros(1)
        latency = np.zeros(n)
        energies = np.zeros(n)
        makespan = ctypes.c_double()
        _check(_lib.ixh_server_replay(self.handle, n, _dptr(arrival), _iptr(ent), _dptr(lam),
                                      offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _dptr(allpos),
                                      _iptr(counts) if counts is not None else None,
                                      _dptr(values) if values is not None else None,
                                      _dptr(latency), _dptr(energies), ctypes.byref(makespan)))
        return latency, energies, makespan.value

    def close(self) -> None:
        if self.handle:
            _lib.ixh_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()
