"""Writes patch_test.json: the known-answer patch test of liblaf/apple
(tests/forward/test_static_simulation.py:16-93), transcribed as data.

Inputs: the 5-vertex mesh (:17-26) and its 4 tets (VTK connectivity :27-52),
mu = 1 per cell (:55), the 4 corner vertices fixed with vertex 3 displaced by
(0.2, -0.1, 0.15) (:69-72); the reference runs ARAP (:75).  Expected
(:84-93): PRIMARY_SUCCESS, n_free == 3, a lower final energy, the interior
vertex displaced by (0.05, -0.025, 0.0375) to 1e-8.  The boundary field is
affine, so the same answer holds for any hyperelastic law (the tests also run
SNH with lambda = 1).  Run: python tests/golden/make_patch_test.py
"""

import json
import os

FIXTURE = {
    "source": "liblaf/apple tests/forward/test_static_simulation.py:16-93",
    "points": [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.25, 0.25, 0.25]],
    "cells": [[0, 1, 2, 4], [0, 1, 4, 3], [0, 4, 2, 3], [4, 1, 2, 3]],
    "mu": 1.0,
    "fixed_vertices": [0, 1, 2, 3],
    "fixed_values": [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.2, -0.1, 0.15]],
    "reference_material": "arap",
    "expected": {
        "result": "PRIMARY_SUCCESS",
        "n_free": 3,
        "interior_vertex": 4,
        "interior_displacement": [0.05, -0.025, 0.0375],
        "atol": 1e-8,
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "patch_test.json")
    with open(out, "w") as f:
        json.dump(FIXTURE, f, indent=1)
    print(out)
