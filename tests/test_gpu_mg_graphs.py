"""Graph replay of the coarse part of the Poisson V-cycle (levels >= 1)
against eager launches: vidyut.mg_graphs = 1 against 0.  The replayed graph
holds the same kernels with the same arguments, so whole steps must agree
bit for bit, also after repeated solves (the graph is captured on the second
cycle of a signature and replayed from then on) and for the FP32 cycles of
the mixed form."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize(
    "case,n,extra",
    [
        ("air3d", (64, 48, 40), {}),
        ("air3d", (64, 48, 40), {"vidyut.mixed_precision": 1}),
        ("argon2d", (96, 64), {}),
    ],
)
def test_graph_replay_matches_eager(case, n, extra):
    from vidyut3d_amd import Vidyut

    out = []
    for graphs in (1, 0):
        ov = dict(extra)
        ov["vidyut.mg_graphs"] = graphs
        g = Vidyut(case, inputs=small_inputs(case, n=n, extra=ov))
        g.init_state()
        info = g.evolve(4)
        out.append((g.get_state(), info.poisson_iters))
        g.close()
    (a, ia), (b, ib) = out
    assert ia == ib
    for c in range(a.shape[0]):
        assert np.array_equal(a[c], b[c]), (case, c, float(np.max(np.abs(a[c] - b[c]))))
