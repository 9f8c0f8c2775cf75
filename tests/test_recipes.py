"""Benchmark-recipe initial states (wr_make_recipe_states; the reference has
only make_collapse_state, so the recipe is this build's own definition).

Pinned by: bit identity of the library's host code with the oracle's C
restatement; the SplitMix64 chain (its published stream is pinned in
test_oracle.py); what the definition fixes mathematically -- z = ln(n_H) /
sigma standard normal (moments), ln T uniform on [ln t_min, ln t_max],
charge neutrality, the He/H ratio of Y = 0.24 and D/H = 3e-5."""
import numpy as np
import pytest

MASS = np.array([9.10938188e-28, 1.67262158e-24, 1.67353251819e-24, 1.67444345638e-24,
                 3.34512158e-24, 3.34603251819e-24, 3.34615409819e-24, 3.34694345638e-24,
                 3.34706503638e-24, 5.01865409819e-24, 5.01956503638e-24, 6.69024316e-24,
                 6.69115409819e-24, 6.69206503638e-24])


def _h_total(xn):
    # hydrogen nuclei: H+, H, H-, 2 H2+, 2 H2, HD+, HD
    return xn[:, 1] + xn[:, 2] + xn[:, 3] + 2 * xn[:, 6] + 2 * xn[:, 8] + xn[:, 9] + xn[:, 10]


@pytest.mark.parametrize("recipe", ["uniform16", "lognormal128"])
def test_recipe_matches_oracle_bit_for_bit(wr, oracle, recipe):
    seed, sigma, t0, t1 = wr.RECIPES[recipe]
    n = 2048
    got = wr.make_recipe_states(n, seed, sigma, t0, t1)
    want = oracle.make_recipe_cells(n, wr.CELL_DTYPE, seed, sigma, t0, t1)
    assert got.tobytes() == want.tobytes()


def test_recipe_statistics(wr):
    n = 1 << 15
    seed, sigma, t0, t1 = wr.RECIPES["lognormal128"]
    c = wr.make_recipe_states(n, seed, sigma, t0, t1)
    xn = c["current"]["xn"]
    z = np.log(_h_total(xn)) / sigma
    # standard normal: mean 0, variance 1, symmetric, 4th moment 3
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1.0) < 0.03
    assert abs(np.mean(z ** 3)) < 0.08 and abs(np.mean(z ** 4) - 3.0) < 0.15
    lt = np.log(c["current"]["T"])
    lo, hi = np.log(t0), np.log(t1)
    assert lt.min() >= lo and lt.max() <= hi
    u = (lt - lo) / (hi - lo)
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1.0 / 12.0) < 0.005
    # composition
    he_h = xn[:, 11:14].sum(1) / _h_total(xn)
    assert np.allclose(he_h, 0.24 / (4 * 0.76), rtol=1e-12)
    d_h = (xn[:, 4] + xn[:, 5] + xn[:, 7] + xn[:, 9] + xn[:, 10]) / _h_total(xn)
    assert np.allclose(d_h, 3e-5, rtol=1e-9)
    charge = xn[:, 1] + xn[:, 4] + xn[:, 6] + xn[:, 9] + xn[:, 12] + 2 * xn[:, 11] - xn[:, 3] - xn[:, 7]
    assert np.allclose(xn[:, 0], charge, rtol=1e-12)
    assert np.allclose(c["current"]["rho"], xn @ MASS, rtol=1e-13)
    assert np.array_equal(c["density_driver"], c["current"]["rho"])
    # streams differ by seed and cell
    assert len(np.unique(c["current"]["T"])) == n
    other = wr.make_recipe_states(64, seed + 1, sigma, t0, t1)
    assert not np.array_equal(other["current"]["T"], c["current"]["T"][:64])


def test_recipe_constant_temperature_and_arguments(wr):
    seed, sigma, t0, t1 = wr.RECIPES["uniform16"]
    c = wr.make_recipe_states(4096, seed, sigma, t0, t1)
    assert np.all(c["current"]["T"] == 300.0)
    z = np.log(_h_total(c["current"]["xn"])) / sigma
    assert abs(z.mean()) < 0.08 and abs(z.std() - 1.0) < 0.05
    with pytest.raises(wr.WarpReactorError):
        wr.make_recipe_states(4, 0, -1.0, 300.0, 300.0)
    with pytest.raises(wr.WarpReactorError):
        wr.make_recipe_states(4, 0, 0.1, 300.0, 100.0)
