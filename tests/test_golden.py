"""The oracle reproduces the committed golden fixtures (tests/golden/)."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-4] for f in os.listdir(HERE) if f.endswith(".npz"))


def load(case):
    z = np.load(os.path.join(HERE, case + ".npz"))
    return json.loads(str(z["params"])), z


def test_fixtures_present():
    assert len(CASES) >= 4


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_fixture(case):
    kw, z = load(case)
    o = po.Oracle(po.default_params(**kw))
    c, p, v, q = o.scalar(), o.scalar(), o.velocity(), z["q"].copy()
    t = 0.0
    for k in range(1, 11):
        st, dt = o.step(c, p, v, q, t)
        assert st == 0
        t += dt
        if k in (1, 10):
            assert np.array_equal(c, z["c%d" % k])
            assert np.array_equal(p, z["p%d" % k])
            assert np.array_equal(v, z["v%d" % k])
            assert dt == float(z["dt%d" % k])
