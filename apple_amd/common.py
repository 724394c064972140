"""Attribute names, Lame conversion and default potential names.

Restates ``common/attr_name.py:4-44``, ``common/_moduli.py:5-10`` and
``common/_potential_name.py:6-16`` of liblaf/apple: the same VTK attribute
names, so meshes prepared for apple carry over unchanged.
"""

from __future__ import annotations

import collections
import enum

import numpy as np


class AttrName(str, enum.Enum):
    ACTIVATION = "activation"
    DISPLACEMENT = "displacement"
    E = "e"
    FIXED_MASK = "fixed_mask"
    FIXED_VALUE = "fixed_value"
    FORCE = "force"
    FRACTION = "fraction"
    GLOBAL_POINT_ID = "global_point_id"
    LAMBDA = "lambda_"
    MASS_DENSITY = "mass_density"
    MU = "mu"
    NU = "nu"
    PRESTRAIN = "prestrain"

    @property
    def vtk(self) -> str:
        special = {"lambda_": "lambda", "mu": "mu", "nu": "nu"}
        if self.value in special:
            return special[self.value]
        return self.value.title().replace("_", "")


ACTIVATION = AttrName.ACTIVATION
DISPLACEMENT = AttrName.DISPLACEMENT
E = AttrName.E
FIXED_MASK = AttrName.FIXED_MASK
FIXED_VALUE = AttrName.FIXED_VALUE
FORCE = AttrName.FORCE
FRACTION = AttrName.FRACTION
GLOBAL_POINT_ID = AttrName.GLOBAL_POINT_ID
LAMBDA = AttrName.LAMBDA
MASS_DENSITY = AttrName.MASS_DENSITY
MU = AttrName.MU
NU = AttrName.NU
PRESTRAIN = AttrName.PRESTRAIN


def lame_converter(E, nu):
    """(lambda, mu) from Young's modulus E and Poisson ratio nu."""
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    la = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return la, mu


_counter: collections.Counter = collections.Counter()


def default_name(obj) -> str:
    """``ClassName{n}`` with a per-class counter (common/_potential_name.py:9-16)."""
    cls = type(obj).__name__
    n = _counter[cls]
    _counter[cls] += 1
    return f"{cls}{n}"
