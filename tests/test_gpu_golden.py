"""The device against the golden fixtures of the 1,000-boid flockers
configuration (tests/golden/flockers_config0.npz, from the oracle): initial
placement, every bird's neighbour vector at step 0, the states after 1 and
10 schedule steps, and the statistics after 500 -- bit for bit."""
import os
import sys

import numpy as np
import pytest

from krabmaga_amd import Field2D, FlockParams, stats

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_flockers_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("binned", [False, True])
def test_device_reproduces_the_golden_arrays(binned):
    with np.load(G.OUT, allow_pickle=False) as z:
        want = {k: z[k] for k in z.files}
    n = G.N
    f = Field2D(G.SIDE, G.SIDE, G.D, True, n)
    f.init_uniform(n, G.SEED, G.HEADING_MODE)
    st, ids = f.state_by_id(1)  # the write buffer, before the first update
    assert np.array_equal(ids, np.arange(n))
    assert np.array_equal(bits(st), bits(want["init"]))
    f.lazy_update()
    storage, row_ptr, nbr = f.neighbor_sets(G.RADIUS)
    by_id = {int(storage[k]): nbr[row_ptr[k]:row_ptr[k + 1]] for k in range(len(storage))}
    got = np.concatenate([by_id[i] for i in range(n)])
    assert np.array_equal(got, want["nbr_ids"])
    f.close()
    p = FlockParams(seed=G.SEED, radius=G.RADIUS)
    f = Field2D(G.SIDE, G.SIDE, G.D, True, n)
    f.init_uniform(n, G.SEED, G.HEADING_MODE)
    f.schedule_run(p, 0, 1, binned=binned)
    assert np.array_equal(bits(f.state_by_id(0)[0]), bits(want["step1"]))
    f.schedule_run(p, 1, 9, binned=binned)
    assert np.array_equal(bits(f.state_by_id(0)[0]), bits(want["step10"]))
    f.schedule_run(p, 10, 490, binned=binned)
    s = f.state_by_id(0)[0]
    f.close()
    got = np.array([stats.polarisation(s[2], s[3]),
                    stats.mean_nn_distance(s[0], s[1], G.SIDE, G.SIDE)])
    assert np.array_equal(got, want["stats500"])
