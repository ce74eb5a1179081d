"""CPU: host-side chain rule of ``hipfdtd.adjoint`` (no GPU calls).

``ProbeDFTAdjoint._cells_from_edges`` turns the library's per-edge sums G
into dL/dε per design cell.  It must equal the oracle's transposed edge
average (``oracle/adjoint_oracle.edge_average_T`` of −G/εe) on the design
cells, including boxes that touch the grid's first and last planes.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import adjoint_oracle as ao
from oracle import fdtd_oracle as fo


@pytest.mark.parametrize("box", [(3, 7, 2, 6, 4, 9), (0, 5, 0, 9, 0, 11), (4, 9, 5, 9, 6, 11)])
def test_cells_from_edges_matches_oracle_transpose(box):
    from hipfdtd.adjoint import ProbeDFTAdjoint

    shape = (9, 9, 11)
    rng = np.random.default_rng(11)
    eps = rng.uniform(1.0, 5.0, shape)
    pb = ProbeDFTAdjoint.__new__(ProbeDFTAdjoint)
    pb.shape = shape
    pb.design_box = box
    gb = (max(box[0] - 1, 0), box[1], max(box[2] - 1, 0), box[3], max(box[4] - 1, 0), box[5])
    pb.grad_box = gb
    pb._eps = eps
    G_full = [np.zeros(shape) for _ in range(3)]
    for c in range(3):
        G_full[c][gb[0]:gb[1], gb[2]:gb[3], gb[4]:gb[5]] = rng.standard_normal(
            (gb[1] - gb[0], gb[3] - gb[2], gb[5] - gb[4]))  # fmt: skip
    G_box = np.stack([g[gb[0]:gb[1], gb[2]:gb[3], gb[4]:gb[5]] for g in G_full])
    got = pb._cells_from_edges(G_box)
    ref = np.zeros(shape)
    for c, ee in enumerate(fo.edge_average(eps)):
        ref += ao.edge_average_T(-G_full[c] / ee, c)
    want = ref[box[0]:box[1], box[2]:box[3], box[4]:box[5]]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.max(np.abs(want)))
