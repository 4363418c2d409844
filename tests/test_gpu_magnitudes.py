"""Bird::step's hand-computed cases on the device, on both device paths.

The cases and their literal expectations are in test_oracle_magnitudes.py
(worked out by hand from tests/model/flockers/bird.rs:62-143).  The device
must meet them and also agree bit for bit with the oracle.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import test_oracle_magnitudes as M  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("binned", [False, True])
@pytest.mark.parametrize("name", sorted(M.CASES))
def test_bird_step_case_on_device(oracle, name, binned):
    from krabmaga_amd import Field2D, FlockParams
    case = M.CASES[name]
    x, y, hx, hy = M.case_arrays(case)
    n = len(x)
    f = Field2D(M.SIDE, M.SIDE, M.D, True, n)
    f.set_object_locations(np.arange(n), x, y, hx, hy)
    f.schedule_run(M.case_params(case, FlockParams), 0, 1, binned=binned)
    got, ids = f.state_by_id(0)
    assert np.array_equal(ids, np.arange(n))
    M.check_case(case, got)
    fl = oracle.Flocker(M.SIDE, M.SIDE, M.D, True, x, y, hx, hy,
                        M.case_params(case, lambda **kw: oracle.params(**kw)))
    fl.step(1)
    assert np.array_equal(got.astype(np.float32).view(np.uint32),
                          fl.state().astype(np.float32).view(np.uint32))
