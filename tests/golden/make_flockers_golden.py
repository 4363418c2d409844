"""Writes tests/golden/flockers_config0.npz from the CPU oracle.

The workload is the 1,000-boid flockers configuration: a 200 x 200 torus,
cells of 10 / 1.5, radius 10, weights 1, jump 0.7, Philox key 0 and unit
headings at uniform angles (tests/model/flockers/state.rs,
tests/model/flockers/bird.rs:12-17).

Arrays:
  init      (4, 1000) f32  [x, y, hx, hy] by bird id, as placed by init
  nbr_ptr   (1001,)   u32  CSR row starts, birds by id
  nbr_ids             u32  each bird's get_neighbors_within_distance at the
                           initial positions, i.e. the neighbour vectors of
                           schedule step 0, in bag-scan order
  step1     (4, 1000) f32  after 1 Schedule::step
  step10    (4, 1000) f32  after 10
  stats500  (2,)      f64  polarisation and mean nearest-neighbour distance
                           after 500

These are regression pins for the oracle and the device.  They are not
independent evidence of parity with the reference; the ported reference
tests, the brute-force checks and the known-answer vectors are.

    python tests/golden/make_flockers_golden.py    (from the repository root)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N, SIDE, D, RADIUS, SEED, HEADING_MODE = 1000, 200.0, 10.0 / 1.5, 10.0, 0, 1
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flockers_config0.npz")


def compute():
    from krabmaga_amd import stats
    from oracle import oracle as O

    x, y, hx, hy = O.init_uniform(N, SIDE, SIDE, SEED, HEADING_MODE)
    init = np.stack([x, y, hx, hy]).astype(np.float32)
    f = O.Field2D(SIDE, SIDE, D, True)
    for i in range(N):
        f.set_object_location(i, float(x[i]), float(y[i]))
    f.lazy_update()
    ptr, ids = [0], []
    for i in range(N):
        ids += [b[0] for b in f.get_neighbors_within_distance(float(x[i]), float(y[i]), RADIUS)]
        ptr.append(len(ids))
    fl = O.Flocker(SIDE, SIDE, D, True, x, y, hx, hy, O.params(seed=SEED, radius=RADIUS))
    fl.step(1)
    step1 = fl.state()
    fl.step(9)
    step10 = fl.state()
    fl.step(490)
    s = fl.state()
    stats500 = np.array([stats.polarisation(s[2], s[3]),
                         stats.mean_nn_distance(s[0], s[1], SIDE, SIDE)], dtype=np.float64)
    return {"init": init, "nbr_ptr": np.array(ptr, dtype=np.uint32),
            "nbr_ids": np.array(ids, dtype=np.uint32), "step1": step1, "step10": step10,
            "stats500": stats500}


if __name__ == "__main__":
    arrays = compute()
    np.savez_compressed(OUT, **arrays)
    print(OUT, {k: v.shape for k, v in arrays.items()}, os.path.getsize(OUT), "bytes")
