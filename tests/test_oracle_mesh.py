"""Pins the oracle's mesh frictional boundary (MeshFrictionalBoundary,
mesh_frictional_boundary.cpp; its use in forward_gpu.cpp:456-528, 1112-1135,
1483-1531): the closest point on a triangle against a brute-force search,
and a rigid slab seen through the mesh boundary against the planar boundary
it then equals (plane at the slab's top plus the contact radius)."""
import numpy as np

from diffphd_amd import scenes
from helpers import oracle_body, rel_err, rollout


def test_triangle_closest_point_matches_a_brute_force_search(oracle_mod):
    rng = np.random.default_rng(0)
    # barycentric grid over the triangle, refined around its best sample
    for _ in range(200):
        a, b, c = rng.standard_normal((3, 3))
        p = 2 * rng.standard_normal(3)
        x = oracle_mod.tri_closest(p, a, b, c)
        n = 120
        u, v = np.meshgrid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1))
        m = u + v <= 1 + 1e-15
        pts = a + np.outer(u[m], b - a) + np.outer(v[m], c - a)
        d = np.linalg.norm(pts - p, axis=1)
        best = d.min()
        dx = np.linalg.norm(x - p)
        # the returned point lies on the triangle and is no farther than any sample
        assert dx <= best + 1e-12
        assert dx >= best - 2.0 / n * max(np.linalg.norm(b - a), np.linalg.norm(c - a))
        # on the triangle: barycentric coordinates in [0, 1]
        T = np.stack([b - a, c - a], 1)
        w, *_ = np.linalg.lstsq(T, x - a, rcond=None)
        assert w.min() >= -1e-9 and w.sum() <= 1 + 1e-9
        assert np.linalg.norm(T @ w + a - x) < 1e-9
    # p in the face region: the orthogonal projection
    a, b, c = np.array([0, 0, 0.0]), np.array([1, 0, 0.0]), np.array([0, 1, 0.0])
    assert np.allclose(oracle_mod.tri_closest([0.2, 0.3, 0.7], a, b, c), [0.2, 0.3, 0.0], atol=1e-15)
    # vertex and edge regions
    assert np.allclose(oracle_mod.tri_closest([-1, -1, 0.5], a, b, c), a)
    assert np.allclose(oracle_mod.tri_closest([0.5, -2, 0.0], a, b, c), [0.5, 0, 0])
    assert np.allclose(oracle_mod.tri_closest([1, 1, 0.0], a, b, c), [0.5, 0.5, 0])


def _planar_twin(spec):
    """The same body with the slab's top face replaced by the plane z = radius."""
    twin = dict(spec)
    radius = spec["boundary"][1][0]
    twin["boundary"] = ("planar", [0.0, 0.0, 1.0, -radius], spec["boundary"][2])
    return twin


def test_rigid_slab_through_the_mesh_boundary_equals_the_planar_boundary(oracle_mod):
    spec = scenes.stack(fixed_bottom="all", both_sides=False, fmu=0.5)
    opt = dict(spec["options"], rel_tol=1e-11, abs_tol=1e-14, max_pd_iter=3000)
    qm, vm, am = rollout(oracle_body(spec), spec, 4, opt)
    qp, vp, ap = rollout(oracle_body(_planar_twin(spec)), _planar_twin(spec), 4, opt)
    print("active", [len(a) for a in am], [len(a) for a in ap], "q", rel_err(qm[-1], qp[-1]))
    assert any(len(a) > 0 for a in am)
    assert am == ap
    assert rel_err(qm[-1], qp[-1]) < 1e-9
