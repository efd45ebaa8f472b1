"""phd_apply_ainv_noalloc (device vectors, intermediates in the library's fixed
device scratch array) against phd_apply_ainv (library-allocated work vectors)
on the same factor and input, through the C ABI.

Two bodies share the GPU's one scratch array: a 1,225-vertex beam, then a
63-vertex beam. Between them the larger body runs once on an all-NaN input,
which leaves NaN in every scratch entry that the smaller body's S-product
input padding occupies; the smaller call must still be finite and equal to
the allocating entry point (the product stages columns >= n as zeros).
"""
import numpy as np
import pytest

from diffphd_amd import scenes
from helpers import oracle_body, product_body, rel_err

pytestmark = pytest.mark.gpu


def _noalloc(torch, body, b, precision, refine):
    bd = torch.tensor(np.asarray(b, dtype=np.float64), device="cuda:0")
    xd = torch.full_like(bd, 7.0)  # stale output contents must not matter
    torch.cuda.synchronize()
    body.apply_ainv_noalloc(bd.data_ptr(), xd.data_ptr(), precision=precision, refine=refine)
    return xd.cpu().numpy()


def test_noalloc_matches_allocating_entry_point():
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    big_spec = scenes.cantilever(cells=(24, 6, 6))
    small_spec = scenes.cantilever(cells=(6, 2, 2))
    big = product_body(big_spec)
    small = product_body(small_spec)
    big.factor_build(big_spec["dt"])
    small.factor_build(small_spec["dt"])
    assert big.dofs() > small.dofs()
    rng = np.random.default_rng(5)
    bb = rng.standard_normal(big.dofs())
    bs = rng.standard_normal(small.dofs())
    xo = oracle_body(small_spec).apply_ainv(small_spec["dt"], bs)
    for precision, refine in ((0, 0), (1, 0), (1, 2), (0, 1)):
        print(f"case precision={precision} refine={refine}", flush=True)
        # call 1: the larger body, finite input
        xa = big.apply_ainv(big_spec["dt"], bb, precision=precision, refine=refine)
        xn = _noalloc(torch, big, bb, precision, refine)
        assert np.array_equal(xa, xn), rel_err(xn, xa)
        # the larger body on an all-NaN input fills the scratch with NaN
        _noalloc(torch, big, np.full(big.dofs(), np.nan), precision, refine)
        # call 2 (smaller input) right after: same numbers as the allocating path
        xs_n = _noalloc(torch, small, bs, precision, refine)
        xs_a = small.apply_ainv(small_spec["dt"], bs, precision=precision, refine=refine)
        assert np.isfinite(xs_n).all()
        assert np.array_equal(xs_a, xs_n), rel_err(xs_n, xs_a)
        tol = 1e-12 if (precision == 0 or refine >= 2) else 1e-5
        assert rel_err(xs_n, xo) < tol, rel_err(xs_n, xo)


def test_noalloc_rejects_bad_arguments():
    torch = pytest.importorskip("torch")
    spec = scenes.cantilever(cells=(6, 2, 2))
    body = product_body(spec)
    bd = torch.zeros(body.dofs(), dtype=torch.float64, device="cuda:0")
    xd = torch.zeros_like(bd)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="factor is not built"):
        body.apply_ainv_noalloc(bd.data_ptr(), xd.data_ptr(), 0, 0)
    body.factor_build(spec["dt"])
    with pytest.raises(RuntimeError, match="must not alias"):
        body.apply_ainv_noalloc(bd.data_ptr(), bd.data_ptr(), 0, 0)
    with pytest.raises(RuntimeError, match="precision"):
        body.apply_ainv_noalloc(bd.data_ptr(), xd.data_ptr(), 2, 0)
