"""Loader of the golden vectors of the reference plugin (tests/data of the reference, copied under tests/golden/).

The format is that of TestInteraxExportValues.h:65-120: a `box` line, `positions N` followed by N rows, and one
`term <name> <group> <energy>` header per force group followed by N rows of forces.
"""
import gzip
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class ExportValues:
    def __init__(self):
        self.periodic = False
        self.box = None
        self.positions = None
        self.terms = []
        self.energies = []
        self.forces = []


def load_export_values(name: str) -> ExportValues:
    values = ExportValues()
    with gzip.open(os.path.join(GOLDEN, name, "reference.txt.gz"), "rt") as f:
        lines = [l for l in f.read().splitlines()]
    i = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        if not line or line.startswith("#"):
            continue
        words = line.split()
        if words[0] == "box":
            values.periodic = words[1] != "none"
            if values.periodic:
                values.box = np.array([float(w) for w in words[1:10]]).reshape(3, 3)
        elif words[0] == "positions":
            n = int(words[1])
            values.positions = np.array([[float(w) for w in lines[i+k].split()] for k in range(n)])
            i += n
        elif words[0] == "term":
            assert int(words[2]) == len(values.terms)
            values.terms.append(words[1])
            values.energies.append(float(words[3]))
            n = len(values.positions)
            values.forces.append(np.array([[float(w) for w in lines[i+k].split()] for k in range(n)]))
            i += n
    return values


def system_path(name: str) -> str:
    return os.path.join(GOLDEN, name, "system.xml.gz")
