"""Autotuner of the launch shapes (phd_autotune): the chosen shapes come from
the candidate sets, and A^-1 b and a contact frame are unchanged by the
choice up to the summation order of the partial sums."""
import numpy as np
import pytest

from diffphd_amd import scenes
from helpers import oracle_body, product_body, rel_err, rollout

pytestmark = pytest.mark.gpu


def test_autotune_keeps_results():
    spec = scenes.press(cells=(16, 16, 8))
    dt = spec["dt"]
    body = product_body(spec)
    b = np.random.default_rng(0).standard_normal(body.dofs())
    x64 = body.apply_ainv(dt, b, precision=0)
    x32 = body.apply_ainv(dt, b, precision=1)
    before = body.factor_info()
    assert (before["chunk_w"], before["unit_items"], before["local_block"]) == (1024, 512, 128)
    t = body.autotune(dt, reps=5)
    print("autotune", t)
    assert t["chunk_w"] in (512, 1024, 2048) and t["unit_items"] in (256, 512, 1024, 2048)
    assert t["local_block"] in (64, 128, 256)
    assert t["rows_ms"][t["chunk_w"]] == min(t["rows_ms"].values())
    assert t["cols_ms"][t["unit_items"]] == min(t["cols_ms"].values())
    assert t["local_ms"][t["local_block"]] == min(t["local_ms"].values())
    info = body.factor_info()
    assert (info["chunk_w"], info["unit_items"], info["local_block"]) == (t["chunk_w"], t["unit_items"], t["local_block"])
    y64 = body.apply_ainv(dt, b, precision=0)
    y32 = body.apply_ainv(dt, b, precision=1)
    print("A^-1 b after tuning: fp64", rel_err(y64, x64), "fp32", rel_err(y32, x32))
    assert rel_err(y64, x64) < 1e-13 and rel_err(y32, x32) < 1e-6
    # every candidate shape gives the same product
    for cw in (512, 1024, 2048):
        for ui in (256, 512, 1024, 2048):
            body.set_launch_shapes(cw, ui, 64)
            z64 = body.apply_ainv(dt, b, precision=0)
            z32 = body.apply_ainv(dt, b, precision=1)
            assert rel_err(z64, x64) < 1e-13 and rel_err(z32, x32) < 1e-6, (cw, ui)
    with pytest.raises(RuntimeError):
        body.set_launch_shapes(4096, 512, 128)


def test_autotuned_body_matches_oracle_on_a_contact_frame():
    spec = scenes.press(cells=(8, 8, 4))
    opt = dict(spec["options"], rel_tol=0.0, abs_tol=0.0, max_pd_iter=16)
    prod, orc = product_body(spec), oracle_body(spec)
    t = prod.autotune(spec["dt"], reps=3)
    print("autotune", t)
    qp, _, ap = rollout(prod, spec, 1, opt)
    qo, _, ao = rollout(orc, spec, 1, opt)
    print("q vs oracle", rel_err(qp[1], qo[1]))
    assert ap[1] == ao[1] and rel_err(qp[1], qo[1]) < 1e-10
