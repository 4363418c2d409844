"""Bird::step's term magnitudes, pinned by hand-computed cases.

tests/model/flockers/bird.rs:62-143 gives each term a fixed size.  It uses
`400 * avoid`, `-cohe / 10`, `0.05 * rand / |rand|`, the avoidance denominator
`sq * sq + 1`, the double division of consistency and the momentum weight.
Every expected value below was worked out by hand, in float64, from those
formulas.  The working is in each case's comment and the results are written
as literals.  None comes from the oracle.  A copy of the oracle with any one
of the following changes fails at least one case here:
400 -> 40, /10 -> /100, sq*sq+1 -> sq+1, 0.05 -> 0.5, the momentum term
dropped, or a single division of consistency.

The randomness draws are the oracle's Philox4x32-10 words for (seed, id,
step).  Philox itself is pinned by the Random123 known-answer vectors in
test_oracle.py.  The draws are written out here as float bits so that a
change in the counter layout shows too.

tests/test_gpu_magnitudes.py runs the same CASES on the device.
"""
import numpy as np
import pytest

f32 = np.float32
D = 10 / 1.5
SIDE = 100.0
TOL = 1e-6          # |heading| <= jump < 1, f32 rounding stays near 1e-7
TOL_POS = 1e-5      # positions near 50, one f32 ulp is 3.8e-6

# Each case: birds (x, y, hx, hy) with ids 0.., the seed, jump and weights
# (unset weights are 0), and the expected heading and position after one
# Schedule::step for the birds listed in "want".
CASES = {
    # Bird 0 at (50, 50) with heading (0, 1); bird 1 at (50 + gap, 50).
    # dx = toroidal_distance(50, 50 + gap) = -gap, dy = 0, count = 1.
    # avoidance.x = 400 * dx / (dx^4 + 1), cohesion.x = -dx / 10, momentum (0, 1).
    # gap 7: 400 * -7 / 2402 + 0.7 = -0.4656952539550374;
    #        (-0.46569525, 1) / 1.10309 * 0.7 = (-0.29551354, 0.63456422).
    # The x heading changes sign at dx^4 + 1 = 4000 (gap ~7.95).
    "avoid_vs_cohesion_gap7": dict(
        birds=[(50.0, 50.0, 0.0, 1.0), (57.0, 50.0, 0.0, -1.0)], seed=0, jump=0.7,
        w=dict(cohesion=1.0, avoidance=1.0, momentum=1.0),
        want={0: ((-0.29551354, 0.63456422), (49.70448646, 50.63456422))}),
    # gap 9: 400 * -9 / 6562 + 0.9 = 0.3513867723255105;
    #        (0.35138677, 1) / 1.05994 * 0.7 = (0.23206102, 0.66041478).
    "avoid_vs_cohesion_gap9": dict(
        birds=[(50.0, 50.0, 0.0, 1.0), (59.0, 50.0, 0.0, -1.0)], seed=0, jump=0.7,
        w=dict(cohesion=1.0, avoidance=1.0, momentum=1.0),
        want={0: ((0.23206102, 0.66041478), (50.23206102, 50.66041478))}),
    # Bird 0 at (50, 50) with heading (0, 0).  Bird 1 is at (53, 50) with
    # heading (0.6, 0.8), bird 2 at (50, 54) with heading (-1, 0); both are
    # in range, count = 2.
    # cohesion = -((-3, 0) + (0, -4)) / 2 / 10 = (0.15, 0.2)
    # consistency = ((0.6, 0.8) + (-1, 0)) / 2 / 2 = (-0.1, 0.2)
    # d = (0.05, 0.4), |d| = 0.4031129, heading = d / |d| * 0.7
    #   = (0.08682431, 0.69459451)
    "consistency_cohesion_three": dict(
        birds=[(50.0, 50.0, 0.0, 0.0), (53.0, 50.0, 0.6, 0.8), (50.0, 54.0, -1.0, 0.0)],
        seed=0, jump=0.7, w=dict(cohesion=1.0, consistency=1.0),
        want={0: ((0.08682431, 0.69459451), (50.08682431, 50.69459451))}),
    # A lone bird finds only itself.  The vector is not empty, so randomness
    # is drawn, and count = 0.  Draws (seed 5, id 0, step 0) in R_BITS.
    # (2 r1 - 1, 2 r2 - 1) = (0.43263507, 0.55734956), norm 0.70556040
    # randomness = 0.05 * that / norm = (0.03065909, 0.03949709)
    # d = randomness + momentum (0.03, 0) = (0.06065909, 0.03949709)
    # heading = d / |d| * 0.7 = (0.58660682, 0.38195869)
    "randomness_vs_momentum": dict(
        birds=[(50.0, 50.0, 0.03, 0.0)], seed=5, jump=0.7,
        w=dict(randomness=1.0, momentum=1.0),
        want={0: ((0.58660682, 0.38195869), (50.58660682, 50.38195869))}),
    # Weights cohesion 2, avoidance 0.5, consistency 1.5, randomness 3,
    # momentum 0.25, jump 0.9, seed 5.
    # Birds: 0 at (50,50) h (0.6,0.8); 1 at (53,50) h (0,1); 2 at (50,54) h (-1,0).
    # All pairs are in range (3, 4, 5).  For bird 0:
    #   avoid = 400 * ((-3,0)/82 + (0,-4)/257) / 2 = (-7.3170732, -3.1128405)
    #   cohesion = (0.15, 0.2), consistency = ((0,1) + (-1,0)) / 4 = (-0.25, 0.25)
    #   randomness = (0.03065909, 0.03949709) (draws above), momentum (0.6, 0.8)
    #   d = 2 coh + 0.5 avoid + 1.5 cons + 3 rand + 0.25 mom,
    #   heading = d / |d| * 0.9 = (-0.89219232, -0.11829146)
    # Birds 1 and 2 are worked out the same way, with their own draws
    # (seed 5, ids 1 and 2, step 0).
    "all_terms_weighted_three": dict(
        birds=[(50.0, 50.0, 0.6, 0.8), (53.0, 50.0, 0.0, 1.0), (50.0, 54.0, -1.0, 0.0)],
        seed=5, jump=0.9,
        w=dict(cohesion=2.0, avoidance=0.5, consistency=1.5, randomness=3.0, momentum=0.25),
        want={0: ((-0.89219232, -0.11829146), (49.10780768, 49.88170854)),
              1: ((0.89350430, 0.10793551), (53.89350430, 50.10793551)),
              2: ((-0.12540423, 0.89122039), (49.87459577, 54.89122039))}),
}

# Philox draws (seed, id, step) -> float bits of (r1, r2).
# (5, 0, 0): r1 = 0.7163175344467163, r2 = 0.7786747813224792.
R_BITS = {(5, 0, 0): (1060593814, 1061639995)}


def case_arrays(case):
    b = np.array(case["birds"], dtype=np.float64)
    return [b[:, k].astype(f32) for k in range(4)]


def case_params(case, cls):
    kw = dict(cohesion=0.0, avoidance=0.0, consistency=0.0, randomness=0.0, momentum=0.0)
    kw.update(case["w"])
    return cls(jump=case["jump"], seed=case["seed"], **kw)


def check_case(case, state):
    """state: (4, n) float array [x, y, hx, hy] by id after one step."""
    s = np.asarray(state, dtype=np.float64)
    for i, (h, p) in case["want"].items():
        assert np.allclose(s[2:, i], h, atol=TOL), (i, s[2:, i], h)
        assert np.allclose(s[:2, i], p, atol=TOL_POS), (i, s[:2, i], p)


@pytest.mark.parametrize("name", sorted(CASES))
def test_bird_step_case(oracle, name):
    case = CASES[name]
    x, y, hx, hy = case_arrays(case)
    fl = oracle.Flocker(SIDE, SIDE, D, True, x, y, hx, hy,
                        case_params(case, lambda **kw: oracle.params(**kw)))
    fl.step(1)
    check_case(case, fl.state())


def test_philox_draws_for_the_cases(oracle):
    for (seed, bird, step), (b1, b2) in R_BITS.items():
        r1, r2 = oracle.step_uniforms(seed, bird, step)
        assert f32(r1).view(np.uint32) == b1 and f32(r2).view(np.uint32) == b2
