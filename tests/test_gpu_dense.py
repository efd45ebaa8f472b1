"""Blocked Cholesky of the contact system (phd_dense.hip) against numpy."""
import numpy as np
import pytest

from diffphd_amd import _lib

pytestmark = pytest.mark.gpu


def _solve(A, b, precision):
    n = b.size
    a_arr, a_p = _lib.dptr(np.asarray(A, dtype=np.float64).ravel(order="F"))
    b_arr, b_p = _lib.dptr(b)
    x_arr, x_p = _lib.dptr(np.zeros(n))
    rc = _lib.lib().phd_dense_spd_solve(n, a_p, b_p, x_p, precision)
    return rc, x_arr.copy()


# The back substitution takes four 64-row blocks per launch from the top:
# 65, 130 and 700 leave a partial top group of 2, 3 and 3 blocks, 260 and 1030
# a top group of one block, 256 and 1024 only full groups.
@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 256, 260, 700, 1024, 1030])
def test_dense_spd_solve_matches_numpy(n):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    b = rng.standard_normal(n)
    ref = np.linalg.solve(A, b)
    rc, x64 = _solve(A, b, 0)
    print(n, "fp64", np.linalg.norm(x64 - ref) / np.linalg.norm(ref))
    assert rc == 0
    assert np.linalg.norm(x64 - ref) <= 1e-12 * np.linalg.norm(ref)
    rc, x32 = _solve(A, b, 1)
    print(n, "fp32", np.linalg.norm(x32 - ref) / np.linalg.norm(ref))
    assert rc == 0
    assert np.linalg.norm(x32 - ref) <= 1e-4 * np.linalg.norm(ref)


def test_dense_spd_solve_delassus_like():
    """Smooth kernel matrix (Green's-function-like) plus a small diagonal, the
    shape of the contact system Omega W Omega + E."""
    n = 1500
    rng = np.random.default_rng(7)
    p = rng.uniform(0, 1, (n, 2))
    d = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
    W = 1.0 / (1.0 + 30.0 * d)
    om = rng.uniform(0.2, 1.8, n)
    A = om[:, None] * W * om[None, :] + np.diag(rng.uniform(1e-3, 1e-2, n))
    b = rng.standard_normal(n)
    ref = np.linalg.solve(A, b)
    rc, x = _solve(A, b, 0)
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    print("fp64", err, "cond", np.linalg.cond(A))
    assert rc == 0 and err < 1e-9


def test_dense_spd_solve_reports_indefinite():
    n = 70
    A = np.eye(n)
    A[40, 40] = -1.0
    for precision in (0, 1):
        rc, _ = _solve(A, np.ones(n), precision)
        assert rc == 1, precision
