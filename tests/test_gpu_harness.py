"""GPU: the force harness (benchmarks/test_forces.f90:129-242) and its error
metric (:66-127) against the oracle's restatement."""
import io

import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import harness, ics

pytestmark = pytest.mark.gpu


def test_errors_ma_matches_oracle():
    m, q, _ = ics.plummer(3000, seed=4)
    Fd = O.direct_forces(m, q[:, 0], q[:, 1], q[:, 2], 1.0, 1e-6)
    F = O.Tree(m, 11).init(q[:, 0], q[:, 1], q[:, 2]).bh_forces(0.49, 1.0, 1e-6)
    e = harness.evaluate_error_on_accelerations(m, q.T, F.T, Fd.T)
    eo = O.errors_ma(m, q, F, Fd)
    assert abs(e[0] - eo[0]) <= 1e-12 * eo[0]
    assert np.allclose(e[1:], eo[1:], rtol=1e-9, atol=1e-15)


def test_forces_time_rows():
    buf = io.StringIO()
    rows = harness.test_forces_time([2000, 5000], [0.3, 0.7], 1e-6, 2, buf, method=12)
    assert len(rows) == 2 * 2 * (1 + 2)
    lines = buf.getvalue().splitlines()
    assert len(lines) == len(rows)
    for ln in lines:
        v = ln.split()
        assert len(v) == 12
    direct = [r for r in rows if r[1] == -1.0]
    assert all(r[5] == [0.0] * 7 for r in direct)
    tree = [r for r in rows if r[1] > 0]
    for r in tree:
        assert 0 < r[5][0] < 1e-2            # McMillan-Aarseth error of an octupole walk
        assert r[2] >= 0 and r[3] > 0 and r[4] > 0
    # smaller theta, smaller error, on the same system
    by = {(r[0], r[1]): r[5][0] for r in tree}
    assert by[(2000, 0.3)] < by[(2000, 0.7)]
