"""The crate's associated-Legendre helpers (special_functions.rs:17-116, legendre_diagonalization.rs:48-109),
restated twice (the package's host factory and the C oracle), pinned by scipy's lpmv (Condon-Shortley phase),
the negative-order relation and Gauss-Legendre orthonormality. No reference test asserts these functions
(special_functions.rs:118-131 only prints)."""
import math

import numpy as np
import pytest
from scipy.special import lpmv

from oracle import oracle as O
from split_operator_amd import hamiltonian_factory as H


def test_double_factorial():
    assert [H.double_factorial(n) for n in range(8)] == [1, 1, 2, 3, 8, 15, 48, 105]
    assert [O.double_factorial(n) for n in range(8)] == [1, 1, 2, 3, 8, 15, 48, 105]


@pytest.mark.parametrize("m", [0, 1, 2, 5, -1, -3])
@pytest.mark.parametrize("x", [-0.9, -0.2, 0.0, 0.37, 0.99])
def test_associated_legendre_vs_scipy(m, x):
    j = 12
    p = H.associated_legendre_polynomials(j, m, x)
    assert np.allclose(p, O.associated_legendre_polynomials(j, m, x), rtol=1e-13, atol=1e-300)
    mu = abs(m)
    for l in range(mu, j + 1):
        ref = lpmv(mu, l, x)
        if m < 0:  # P_l^-m = (-1)^m (l - m)! / (l + m)! P_l^m
            ref *= (-1) ** mu * math.factorial(l - mu) / math.factorial(l + mu)
        assert abs(p[l] - ref) <= 1e-12 * max(1.0, abs(ref)), (l, p[l], ref)
    assert np.all(p[:mu] == 0.0)


@pytest.mark.parametrize("m", [0, 1, 3, -2])
def test_normalised_functions_orthonormal_on_gl_grid(m):
    """sum_j w_j N_l N_l' P_l^m(x_j) P_l'^m(x_j) = delta_ll' (exact for GL quadrature of this degree)."""
    n = 24
    x, w = np.polynomial.legendre.leggauss(n)
    theta = np.arccos(x)[::-1].copy()
    wt = w[::-1].copy()
    polar = type("G", (), {"nodes": theta, "weights": wt, "nodes_no": n, "dimension_no": 1})()
    t, ti = H.associated_legendre_matrices(polar, m)
    mu = abs(m)
    prod = (t @ ti)[mu:, mu:]
    assert np.max(np.abs(prod - np.eye(n - mu))) < 1e-12
    assert np.all(t[:mu] == 0) and np.all(ti[:, :mu] == 0)
    og = O.Grid.new_custom("polar", theta, wt, 1)
    to, tio = O.associated_legendre_matrices(og, m)
    assert np.max(np.abs(to - t)) < 1e-13 and np.max(np.abs(tio - ti)) < 1e-13


def test_normalization_closed_form():
    for l in range(0, 9):
        for m in range(-l, l + 1):
            ref = math.sqrt((l + 0.5) * math.factorial(l - m) / math.factorial(l + m))
            assert abs(H.normalization(l, m) - ref) < 1e-13 * ref
            assert abs(O.normalization(l, m) - ref) < 1e-13 * ref
