"""Golden fixtures of the 1,000-boid flockers configuration
(tests/golden/flockers_config0.npz, written by
tests/golden/make_flockers_golden.py from the oracle): the oracle still
produces them bit for bit.  tests/test_gpu_golden.py holds the device to the
same arrays."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_flockers_golden as G  # noqa: E402


def load():
    with np.load(G.OUT, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_the_golden_arrays(oracle):
    want = load()
    got = G.compute()
    assert sorted(got) == sorted(want)
    for k in want:
        assert got[k].dtype == want[k].dtype and got[k].shape == want[k].shape, k
        assert np.array_equal(got[k].view(np.uint8), want[k].view(np.uint8)), k


def test_golden_arrays_hold_the_invariants():
    z = load()
    n = G.N
    assert z["nbr_ptr"][0] == 0 and z["nbr_ptr"][-1] == z["nbr_ids"].size
    rows = np.split(z["nbr_ids"], z["nbr_ptr"][1:-1])
    # every bird finds itself (distance 0), and the neighbour relation is symmetric
    assert all(i in r for i, r in enumerate(rows))
    pairs = {(i, int(j)) for i, r in enumerate(rows) for j in r}
    assert all((j, i) in pairs for i, j in pairs)
    for k in ("step1", "step10"):
        s = z[k]
        assert s.shape == (4, n)
        assert np.all((s[:2] >= 0) & (s[:2] < G.SIDE))
        h = np.hypot(s[2].astype(np.float64), s[3].astype(np.float64))
        assert np.allclose(h, 0.7, atol=1e-6)  # |last_d| = jump
    assert 0.0 <= z["stats500"][0] <= 1.0 and z["stats500"][1] > 0.0
