"""Loaders of the committed golden fixtures (data only)."""
import json
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def cold_front():
    """CSR (block 1) of the cold-front matrix with its stored float
    coefficients, the right-hand side and the reference test's options."""
    d = json.loads((HERE / "cold_front.json").read_text())
    n = d["n"]
    rows = [[] for _ in range(n)]
    for r, c, v in d["entries"]:
        rows[r].append((c, float(np.float32(float(v)))))
    offsets = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for i, row in enumerate(rows):
        for c, v in row:  # the fixture's entry order
            cols.append(c)
            vals.append(v)
        offsets[i + 1] = len(cols)
    rhs = np.zeros(n)
    rhs[d["rhs_ones_at"]] = 1.0
    atol = d["absolute_tolerance_in_float_eps"] * float(np.finfo(np.float32).eps)
    return n, offsets, np.array(cols, np.int32), np.array(vals), rhs, d["restart"], d["maximum_iterations"], atol
