"""The Gauss-Legendre tables the quadratures take as input (numpy's
`leggauss`, ascending) against an independent computation (scipy's
`roots_legendre`) and the rule's defining property, at the degrees the
reference uses: 30 for the volume-map ball integral
(volume_map_boundary.rs:1705, discretization.rs:42) and 8 for contact cells
(volume_map_boundary.rs:57, 795-803), plus 4 for the scenes' half-space maps.
The reference's own table comes from gauss-quad 0.3.2, which cannot run here;
this pins the nodes and weights to the mathematical rule within a few ulps,
leaving only their last bits unpinned."""
import numpy as np
import pytest
import scipy.special as sp


@pytest.mark.parametrize("n", [4, 8, 30])
def test_tables_agree_and_integrate_polynomials_exactly(n):
    x, w = np.polynomial.legendre.leggauss(n)
    xs, ws = sp.roots_legendre(n)
    assert np.all(np.diff(x) > 0)                          # ascending, as the kernels expect
    np.testing.assert_allclose(x, xs, rtol=0, atol=4e-16)
    np.testing.assert_allclose(w, ws, rtol=0, atol=4e-15)
    # exact for every monomial of degree < 2n: int_{-1}^{1} t^k dt
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-13
