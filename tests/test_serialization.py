"""XmlSerializer for the four Interax forces and the System (CPU only).

Restates the reference's serialization tests (serialization/tests/TestSerializeInterax*Force.cpp, same values): every
property survives a round trip, serializing the copy gives the same text, a default force keeps its defaults, and
version 2 is refused.  The golden systems (written by OpenMM) load, every attribute of their Interax force nodes is
one the proxies define, and a round trip through our writer keeps every parameter and the oracle's energy.
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import interax_hip as ih
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

PAIR_GETTERS = ("getForceGroup", "getName", "getNonbondedMethod", "getCutoffDistance", "getUseSwitchingFunction",
                "getSwitchingDistance", "getAlchemicalParameterName", "getSoftcoreRadius", "getNumParticles",
                "getNumExclusions")
EXTRA = {"InteraxDispersionForce": ("getDampingScale", "getUseLongRangeCorrection"),
         "InteraxExchangeForce": ("getOverlapMethod",), "InteraxShortRangeForce": ("getOverlapMethod",)}
THREEBODY_GETTERS = ("getForceGroup", "getName", "getCutoffDistance", "getScaleFactor", "getDampingMethod",
                     "getDampingScale", "getCentroidDampingRadius", "usesPeriodicBoundaryConditions",
                     "getAlchemicalParameterName", "getNumSites")


def compare(a, b):
    assert type(a) is type(b)
    if isinstance(a, ih.InteraxThreeBodyForce):
        for g in THREEBODY_GETTERS:
            assert getattr(a, g)() == getattr(b, g)(), g
        for i in range(a.getNumSites()):
            assert a.getSiteParameters(i) == b.getSiteParameters(i)
        return
    for g in PAIR_GETTERS+EXTRA[a.NAME]:
        assert getattr(a, g)() == getattr(b, g)(), g
    for i in range(a.getNumParticles()):
        assert a.getParticleParameters(i) == b.getParticleParameters(i)
    for i in range(a.getNumExclusions()):
        assert a.getExclusionParticles(i) == b.getExclusionParticles(i)


def round_trip(force):
    text = ih.XmlSerializer.serialize(force)
    copy = ih.XmlSerializer.deserialize(text)
    compare(force, copy)
    assert ih.XmlSerializer.serialize(copy) == text
    return copy


def _pair_settings(force, method):
    force.setForceGroup(3)
    force.setName("custom name")
    force.setNonbondedMethod(method)
    force.setCutoffDistance(1.1)
    force.setUseSwitchingFunction(True)
    force.setSwitchingDistance(0.95)
    force.setAlchemicalParameterName("lambda_test")
    force.setSoftcoreRadius(0.12)


def test_dispersion():
    """TestSerializeInteraxDispersionForce.cpp:43-107."""
    f = ih.InteraxDispersionForce()
    _pair_settings(f, ih.InteraxPairForce.CutoffPeriodic)
    f.setDampingScale(0.7155241370201111)
    f.setUseLongRangeCorrection(True)
    f.addParticle(2.0412491330659432, 0.07991696226413432, 0.003917538949072952, 0.0020393931939312773,
                  0.026844933830306183)
    f.addParticle(0.47385034770767653, 0.0062664158577877644, 0.00015411387998271836, 0.00023498508736915282,
                  0.020351567997950105, True)
    f.addParticle(0.0, 0.0, 0.0, 0.0, 0.0, False)
    f.addExclusion(0, 1)
    f.addExclusion(2, 1)
    round_trip(f)


@pytest.mark.parametrize("method", [ih.InteraxExchangeForce.Exact, ih.InteraxExchangeForce.InteraxOrder2])
def test_exchange(method):
    """TestSerializeInteraxExchangeForce.cpp:43-100."""
    f = ih.InteraxExchangeForce()
    _pair_settings(f, ih.InteraxPairForce.CutoffNonPeriodic)
    f.setOverlapMethod(method)
    f.addParticle(10.326424719450545, 0.0248302193400194)
    f.addParticle(1.0254403205662432, 0.019763296864870033, True)
    f.addParticle(0.0, 0.0, False)
    f.addExclusion(0, 1)
    f.addExclusion(2, 1)
    round_trip(f)


@pytest.mark.parametrize("method", [ih.InteraxShortRangeForce.Exact, ih.InteraxShortRangeForce.InteraxOrder2])
def test_short_range(method):
    """TestSerializeInteraxShortRangeForce.cpp:43-101."""
    f = ih.InteraxShortRangeForce()
    _pair_settings(f, ih.InteraxPairForce.CutoffNonPeriodic)
    f.setOverlapMethod(method)
    f.addParticle(-1102.281886852146, 5953.779066676866, 0.017966632940037853)
    f.addParticle(1236.5568512540217, 80470.88690318535, 0.026044362132579005, True)
    f.addParticle(0.0, 0.0, 0.0, False)
    f.addExclusion(0, 1)
    f.addExclusion(2, 1)
    round_trip(f)


@pytest.mark.parametrize("radius", [0.0, 0.25])
def test_threebody(radius):
    """TestSerializeInteraxThreeBodyForce.cpp:74-100."""
    f = ih.InteraxThreeBodyForce()
    f.setForceGroup(3)
    f.setName("custom name")
    f.setCutoffDistance(1.1)
    f.setScaleFactor(0.8)
    f.setDampingMethod(ih.InteraxThreeBodyForce.TangToennies)
    f.setDampingScale(0.7)
    f.setCentroidDampingRadius(radius)
    f.setUsesPeriodicBoundaryConditions(True)
    f.setAlchemicalParameterName("lambda_test")
    f.addSite([0, 1, 2], 0.004087456746801944, 911.0681141195369, 0.03, 0, False)
    f.addSite([3], 0.0016246535729804365, 1262.664320774599, 0.025, -1, True)
    f.addSite([4, 6], 0.008817709933796472, 747.5256543931184, 0.035, 2, False)
    round_trip(f)


def test_defaults():
    """TestSerializeInteraxDispersionForce.cpp:109-124 and TestSerializeInteraxExchangeForce.cpp:102-116."""
    f = ih.InteraxDispersionForce()
    f.addParticle(1.0, 0.1, 0.01, 0.001, 0.02)
    c = ih.XmlSerializer.deserialize(ih.XmlSerializer.serialize(f))
    assert c.getNonbondedMethod() == ih.InteraxPairForce.NoCutoff
    assert c.getCutoffDistance() == f.getCutoffDistance()
    assert c.getUseSwitchingFunction() is False
    assert c.getAlchemicalParameterName() == "lambda_sterics"
    assert c.getSoftcoreRadius() == 0.0 and c.getDampingScale() == 1.0 and c.getUseLongRangeCorrection() is False
    for cls in (ih.InteraxExchangeForce, ih.InteraxShortRangeForce):
        c = ih.XmlSerializer.deserialize(ih.XmlSerializer.serialize(cls()))
        assert c.getNonbondedMethod() == ih.InteraxPairForce.NoCutoff
        assert c.getUseSwitchingFunction() is False and c.getSoftcoreRadius() == 0.0
        assert c.getAlchemicalParameterName() == "lambda_sterics"
        assert c.getOverlapMethod() == cls.Exact
    c = ih.XmlSerializer.deserialize(ih.XmlSerializer.serialize(ih.InteraxThreeBodyForce()))
    compare(ih.InteraxThreeBodyForce(), c)


@pytest.mark.parametrize("cls", [ih.InteraxDispersionForce, ih.InteraxExchangeForce, ih.InteraxShortRangeForce,
                                 ih.InteraxThreeBodyForce])
def test_unsupported_version(cls):
    """testUnsupportedVersion of each TestSerializeInterax*Force.cpp."""
    text = ih.XmlSerializer.serialize(cls())
    assert 'version="1"' in text
    with pytest.raises(ih.InteraxError, match="Unsupported version number"):
        ih.XmlSerializer.deserialize(text.replace('version="1"', 'version="2"', 1))


def test_missing_group_and_name_take_defaults():
    """The proxies read "forceGroup" with default 0 and "name" with the force's own name as default."""
    f = ih.InteraxExchangeForce()
    f.setForceGroup(5)
    f.setName("x")
    node = ET.fromstring(ih.XmlSerializer.serialize(f))
    del node.attrib["forceGroup"]
    del node.attrib["name"]
    c = ih.XmlSerializer.deserialize(ET.tostring(node, encoding="unicode"))
    assert c.getForceGroup() == 0 and c.getName() == ih.InteraxExchangeForce().getName()


# attributes each proxy writes (Interax*ForceProxy.cpp serialize)
_PAIR_ATTRS = {"type", "version", "forceGroup", "name", "method", "cutoff", "useSwitchingFunction", "switchingDistance",
               "alchemicalParameter", "softcoreRadius"}
_ATTRS = {"InteraxDispersionForce": _PAIR_ATTRS | {"dampingScale", "useLongRangeCorrection"},
          "InteraxExchangeForce": _PAIR_ATTRS | {"overlapMethod"},
          "InteraxShortRangeForce": _PAIR_ATTRS | {"overlapMethod"},
          "InteraxThreeBodyForce": {"type", "version", "forceGroup", "name", "cutoff", "scaleFactor", "usesPeriodic",
                                    "alchemicalParameter", "dampingMethod", "dampingScale", "centroidDampingRadius"}}


@pytest.mark.parametrize("name", ["cluster", "methanol_box"])
def test_golden_systems_round_trip(name):
    path = os.path.join(GOLDEN, name, "system.xml.gz")
    import gzip
    with gzip.open(path, "rt") as f:
        text = f.read()
    root = ET.fromstring(text)
    seen = 0
    for node in root.find("Forces"):
        if node.get("type") in _ATTRS:
            assert set(node.attrib) == _ATTRS[node.get("type")], node.get("type")
            seen += 1
    system = ih.load_system_xml(path)
    assert seen == system.getNumForces() > 0
    assert system.getNumParticles() == len(root.find("Particles"))
    masses = [float(p.get("mass")) for p in root.find("Particles")]
    assert [system.getParticleMass(i) for i in range(system.getNumParticles())] == masses
    copy = ih.XmlSerializer.deserialize(ih.XmlSerializer.serialize(system))
    assert np.array_equal(copy.getDefaultPeriodicBoxVectors(), system.getDefaultPeriodicBoxVectors())
    assert copy.getNumForces() == system.getNumForces()
    for a, b in zip(system.getForces(), copy.getForces()):
        compare(a, b)
    # the copy is the same physics: identical oracle energies on the same coordinates
    if name == "cluster":
        rng = np.random.default_rng(1)
        pos = rng.uniform(0, 1.5, size=(system.getNumParticles(), 3))
        box = None
        assert oracle.evaluate(system, pos, box)[0] == oracle.evaluate(copy, pos, box)[0]
