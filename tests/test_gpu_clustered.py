"""Clustered density: Gaussian blobs as dense as the 8,388,608-boid
clustered workload's (peak about 16 birds per cell) on a smaller torus, so
the oracle finishes in seconds.  Tiles this full overflow every LDS working
set the device plans, which exercises the fallbacks; the results must still
be the oracle's bits."""
import numpy as np
import pytest

from krabmaga_amd import Field2D, FlockParams, workloads

pytestmark = pytest.mark.gpu

D = 10.0 / 1.5


@pytest.mark.parametrize("binned", [False, True])
def test_dense_blobs_match_oracle(oracle, binned):
    O = oracle
    n, side, steps = 131072, 10240.0, 2
    x, y, hx, hy = workloads.clustered(n, side, blobs=2, sigma=150.0, uniform_frac=0.2, seed=5)
    cx = np.floor(x / np.float32(D)).astype(np.int64)
    cy = np.floor(y / np.float32(D)).astype(np.int64)
    assert np.bincount(cx * 4096 + cy).max() >= 20  # blob cores are dense
    p = FlockParams(seed=5)
    f = Field2D(side, side, D, True, n)
    f.set_object_locations(np.arange(n, dtype=np.uint32), x, y, hx, hy)
    f.schedule_run(p, 0, steps, binned=binned)
    st, ids = f.state_by_id(0)
    f.close()
    fl = O.Flocker(side, side, D, True, x, y, hx, hy, O.params(seed=5))
    fl.step(steps)
    ref = fl.state()
    assert np.array_equal(ids, np.arange(n, dtype=np.uint32))
    bad = np.flatnonzero(np.any(st.view(np.uint32) != ref.view(np.uint32), axis=0))
    assert bad.size == 0, f"{bad.size} birds differ, first {bad[:5]}"
