"""Host-side checks of the random-field generator (statistics of the mode sum)."""
import numpy as np

from ptstokes import grf


def test_modes_are_seeded_and_shaped():
    a, b = grf.modes(7, 3), grf.modes(7, 3)
    assert a.shape == (1024, 4) and np.array_equal(a, b)
    assert not np.array_equal(a, grf.modes(8, 3))
    assert np.all(grf.modes(7, 2)[:, 2] == 0.0)


def test_field_statistics():
    # unit variance and Gaussian correlation exp(-r^2/l^2), l = 0.5, averaged
    # over many independent fields
    rng_seeds = range(40)
    x = np.linspace(0, 40, 4001)
    lag = int(round(0.5 / (x[1] - x[0])))
    v, c = [], []
    for s in rng_seeds:
        f = grf.evaluate(grf.modes(s, 2, corr_len=0.5, nmodes=512), x, np.zeros_like(x))
        v.append(np.mean(f * f))
        c.append(np.mean(f[:-lag] * f[lag:]))
    assert abs(np.mean(v) - 1.0) < 0.1
    assert abs(np.mean(c) - np.exp(-1.0)) < 0.08
