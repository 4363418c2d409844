"""Closed-form checks of the flock statistics (CPU)."""
import numpy as np

from krabmaga_amd.stats import mean_nn_distance, polarisation


def test_polarisation_closed_forms():
    assert abs(polarisation(np.full(50, 0.7), np.zeros(50)) - 1.0) < 1e-12
    # opposite pairs cancel
    assert polarisation([1.0, -1.0, 0.0, 0.0], [0.0, 0.0, 1.0, -1.0]) < 1e-12
    # zero headings are ignored; lengths do not matter
    assert abs(polarisation([0.0, 3.0, 0.5], [0.0, 0.0, 0.0]) - 1.0) < 1e-12
    rng = np.random.default_rng(0)
    a = rng.uniform(0, 2 * np.pi, 100000)
    assert polarisation(np.cos(a), np.sin(a)) < 0.02


def test_mean_nn_distance_lattice_and_wrap():
    # a 20 x 20 lattice of spacing 5 on a 100 x 100 torus: every nearest
    # neighbour is 5 away, including across the seam
    g = np.arange(20) * 5.0 + 0.5
    x, y = np.meshgrid(g, g)
    assert abs(mean_nn_distance(x.ravel(), y.ravel(), 100.0, 100.0, block=37) - 5.0) < 1e-9
    # two points near opposite edges are close through the seam
    assert abs(mean_nn_distance([0.5, 99.5], [50.0, 50.0], 100.0, 100.0) - 1.0) < 1e-9
