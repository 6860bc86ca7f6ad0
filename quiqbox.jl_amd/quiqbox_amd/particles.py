"""Nuclei and electron counts (Particles.jl, Lexicons.jl:10-13)."""
import math

import numpy as np

AtomElementNames = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
                    "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K", "Ca"]
ELEMENTS = {s: i + 1 for i, s in enumerate(AtomElementNames)}


def getCharge(nuc):
    """Nuclear charge of a symbol, a sequence of symbols or a NuclearCluster."""
    if isinstance(nuc, NuclearCluster):
        return sum(ELEMENTS[s] for s in nuc.symbols)
    if isinstance(nuc, str):
        if nuc not in ELEMENTS:
            raise KeyError("unknown element %r" % nuc)
        return ELEMENTS[nuc]
    return sum(getCharge(s) for s in nuc)


class NuclearCluster:
    """Nuclear symbols and coordinates; sorted by (charge, coordinates) unless
    pairwiseSort is false (Particles.jl:19-52)."""

    def __init__(self, nucSyms, nucCoords, pairwiseSort=True):
        syms = [str(s) for s in nucSyms]
        coords = [tuple(float(x) for x in c) for c in nucCoords]
        if len(syms) == 0:
            raise ValueError("`nucSyms` must not be empty")
        if len(syms) != len(coords):
            raise ValueError("`nucSyms` and `nucCoords` should have the same length.")
        order = range(len(syms))
        if pairwiseSort:
            order = sorted(order, key=lambda i: (ELEMENTS[syms[i]], coords[i]))
        self.symbols = [syms[i] for i in order]
        self.coords = [coords[i] for i in order]

    def __iter__(self):
        return iter(zip(self.symbols, self.coords))

    def __len__(self):
        return len(self.symbols)

    @property
    def charges(self):
        return np.array([ELEMENTS[s] for s in self.symbols], dtype=np.float64)

    @property
    def coord_array(self):
        return np.array(self.coords, dtype=np.float64).reshape(-1, 3)


class OccupationState:
    """Occupation numbers (n_alpha, n_beta) (Particles.jl:73-140)."""

    def __init__(self, layout):
        layout = [int(x) for x in layout]
        if len(layout) == 0 or min(layout) < 0:
            raise ValueError("occupations must be a non-empty list of non-negative integers")
        self.layout = layout

    def __iter__(self):
        return iter(self.layout)

    def __getitem__(self, i):
        return self.layout[i]

    def __len__(self):
        return len(self.layout)

    def total(self):
        return sum(self.layout)


def prepareSpinConfiguration(nuc, difference=None, offset=0):
    """(n_alpha, n_beta) of the neutral (plus offset) system (Particles.jl:146-171)."""
    total = getCharge(nuc) + offset
    if total <= 0:
        raise ValueError("the total occupation must be positive")
    if difference is None:
        nbeta = total // 2
        return OccupationState((total - nbeta, nbeta))
    if abs(difference) > total or (difference % 2) != (total % 2):
        raise ValueError("`difference=%d` results in illegal spin occupations." % difference)
    return OccupationState(((total + difference) // 2, (total - difference) // 2))


def nucRepulsion(nuc, nucCoords=None):
    """sum_{i<j} Z_i Z_j / r_ij (Particles.jl:197-210)."""
    if isinstance(nuc, NuclearCluster):
        syms, coords = nuc.symbols, nuc.coords
    else:
        syms, coords = list(nuc), list(nucCoords)
    res = 0.0
    for i in range(len(syms)):
        for j in range(i + 1, len(syms)):
            d = math.sqrt(sum((a - b) ** 2 for a, b in zip(coords[i], coords[j])))
            res += getCharge(syms[i]) * getCharge(syms[j]) / d
    return res
