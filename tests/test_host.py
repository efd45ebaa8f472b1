"""CPU tests of the product's host side: the C ABI library loads and exports
every entry point of include/phd.h, the host factor construction (nested
dissection, multifrontal Cholesky, index-free S layout, SpMV plans), mesh
formats, scenes and the benchmark's command line."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from diffphd_amd import _lib, scenes
from diffphd_amd.hex_mesh import hex_grid, tet_grid, write_bin

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "phd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(phd_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(syms) == bound


def _laplacian(V, E, shift=1.0):
    n = V.shape[1]
    edim = E.shape[0]
    r, c = [], []
    for a in range(edim):
        for b in range(edim):
            r.append(E[a])
            c.append(E[b])
    r = np.concatenate(r)
    c = np.concatenate(c)
    W = sp.coo_matrix((np.where(r == c, edim - 1.0, -1.0), (r, c)), shape=(n, n)).tocsr()
    A = (W + shift * sp.eye(n)).tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize("kind", ["hex", "tet", "isolated"])
def test_host_factor_inverse_and_plans(kind):
    if kind == "tet":
        V, E = tet_grid(5, 0.1, (0, 0, 0))
    else:
        V, E, _, _ = hex_grid(np.ones((9, 7, 5)), 0.1, (0, 0, 0))
    A = _laplacian(V, E)
    if kind == "isolated":  # Dirichlet vertices become isolated 1x1 blocks
        A = A.tolil()
        for v in range(0, A.shape[0], 7):
            A.rows[v] = [v]
            A.data[v] = [1.0]
        A = A.tocsc().tolil()
        for v in range(0, A.shape[0], 7):
            col = A[:, v].nonzero()[0]
            for r in col:
                if r != v:
                    A[r, v] = 0
        A = A.tocsr()
        A.eliminate_zeros()
        A.sort_indices()
    n = A.shape[0]
    Ap, app = _lib.iptr(A.indptr)
    Ai, aip = _lib.iptr(A.indices)
    Ax, axp = _lib.dptr(A.data)
    X, xp = _lib.dptr(V.T.ravel())
    out, op = _lib.dptr(np.zeros(7))
    assert _lib.lib().phd_host_factor_check(n, app, aip, axp, xp, op) == 0
    assert out[0] < 1e-12, out  # S^T S A = I
    assert out[1] < 1e-12 and out[2] < 1e-12  # both SpMV plans replay the dense products
    assert out[6] == 0  # L^-1 row k is exactly the subtree range [fd(k), k]


def test_bin_mesh_roundtrip_and_vertex_order(tmp_path):
    import oracle
    V, E, _, flag = hex_grid(np.ones((3, 2, 2)), 0.5, (1.0, 2.0, 3.0))
    # generate_hex_mesh numbers vertices i -> j -> k (hex_mesh.py:20-31)
    assert np.allclose(V[:, 1], [1.0, 2.0, 3.5]) and np.allclose(V[:, 3], [1.0, 2.5, 3.0])
    # element node order 000, 001, 010, 011, 100, ... (hex_mesh.py:43-46)
    assert list(E[:, 0]) == [flag[0, 0, 0], flag[0, 0, 1], flag[0, 1, 0], flag[0, 1, 1], flag[1, 0, 0],
                             flag[1, 0, 1], flag[1, 1, 0], flag[1, 1, 1]]
    p = str(tmp_path / "m.bin")
    write_bin(p, V, E)
    vdim, edim, V2, E2 = oracle.read_bin_mesh(p)
    assert (vdim, edim) == (3, 8) and np.array_equal(V, V2) and np.array_equal(E, E2)


def test_press_scene_is_well_formed():
    spec = scenes.press(cells=(8, 8, 4))
    V = spec["V"]
    kind, params, idx = spec["boundary"]
    assert kind == "press"
    nw = int(params[1])
    assert len(params) == 2 + 4 * nw + 1
    # block rests on the highest ground point of its footprint: no initial penetration
    z0, waves = params[0], np.array(params[2:2 + 4 * nw]).reshape(nw, 4)
    bottom = V[:, np.abs(V[2]) < 1e-12]
    h = z0 + sum(a * np.sin(kx * bottom[0] + ky * bottom[1] + ph) for a, kx, ky, ph in waves)
    assert (bottom[2] - h).min() >= -1e-12
    assert spec["plane_schedule"][-1] == pytest.approx(params[-1] - 0.2 * 4 * 0.01)
    E = np.asarray(spec["E_elem"])
    assert E.max() / E.min() == pytest.approx(10.0)


def test_bench_rejects_multiple_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
