"""The oracle's scalar mass and Laplacian (the operators of the Helmholtz and
matrix filters), pinned by closed forms on an adaptive forest with hanging
nodes: 1^T M u is the exact integral of a trilinear field, u^T K u the exact
Dirichlet energy of an affine field, K 1 = 0; and the two filters' fixed
points and transposes."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tmr_amd import workloads


@pytest.fixture(scope="module")
def mk():
    forests, rho, L = workloads.adaptive(seed=8, base_level=2, max_level=4, nlevels=1)
    f = forests[0]
    nums, X = f.getNodeNumbers(), f.getPoints()
    P = np.zeros((int(f.getNodeRange()[1]), 3))
    P[nums[nums >= 0]] = X[nums >= 0]
    M, K = po.scalar_mass_laplacian(f)
    return M, K, P


def test_mass_integrates_trilinear_fields(mk):
    M, K, P = mk
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    u = 1.0 + x + y * z + x * y * z          # over [0, 2] x [0, 1] x [0, 1]
    exact = 2.0 + 2.0 + 2.0 * 0.25 + 2.0 * 0.25  # int 1, x, yz, xyz
    assert abs(np.ones(len(u)) @ (M @ u) - exact) < 1e-12


def test_laplacian_energy_and_null_space(mk):
    M, K, P = mk
    n = M.shape[0]
    assert np.abs(K @ np.ones(n)).max() < 1e-12
    u = 0.5 * P[:, 0] - 2.0 * P[:, 1] + 0.25 * P[:, 2]
    assert abs(u @ (K @ u) - 2.0 * (0.25 + 4.0 + 0.0625)) < 1e-11
    assert abs((K - K.T)).max() < 1e-15 and abs((M - M.T)).max() < 1e-15


@pytest.mark.parametrize("which", ["helmholtz", "matrix"])
def test_filters_fixed_point_and_adjoint(mk, which):
    M, K, P = mk
    n = M.shape[0]
    rng = np.random.default_rng(1)
    x, y = rng.random(n), rng.standard_normal(n)
    if which == "helmholtz":
        F = lambda v, t=False: po.helmholtz_filter(M, K, 0.1, v, transpose=t)
    else:
        F = lambda v, t=False: po.matrix_filter(M, 0.1, 5, v, transpose=t)
    assert np.allclose(F(np.full(n, 3.0)), 3.0, rtol=1e-12, atol=0)
    assert abs(F(x) @ y - x @ F(y, True)) < 1e-10 * abs(F(x) @ y)
    # smoothing: the filtered field varies less than the input
    assert np.std(F(x)) < np.std(x)
