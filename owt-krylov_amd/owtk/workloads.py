"""Synthetic workloads of the batched solver: many independent block-3
systems from the triangulated-grid generator, grid sizes drawn log-uniformly,
one seed per system and a per-system coefficient contrast."""
from __future__ import annotations

import numpy as np

import owtk


def batched_systems(count=512, nx_min=32, nx_max=512, block=3, seed0=3000, max_contrast=1.5):
    """List of (nodes, row_offsets, columns, values) in component-diagonal CSR
    form (diagonal included), the right-hand sides, and the grid of each."""
    systems, rhs, grids = [], [], []
    for k in range(count):
        rng = np.random.default_rng(seed0 + k)
        nx = int(round(np.exp(rng.uniform(np.log(nx_min), np.log(nx_max)))))
        nx = max(nx_min, min(nx_max, nx))
        ny = max(2, nx // 2)
        sigma = float(rng.uniform(0.0, max_contrast))
        s = owtk.generate_split(nx, ny, block, seed0 + k, ordering=owtk.ORDER_RCM, law=owtk.LAW_EDGE_CONTRAST,
                                sigma=sigma)
        o, c, v = s.csr()
        systems.append((s.nodes, o, c, v))
        rhs.append(s.rhs)
        grids.append((nx, ny, sigma))
    return systems, np.concatenate(rhs), grids
