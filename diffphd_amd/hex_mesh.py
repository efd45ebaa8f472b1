"""Hex/tet grid generators and the `.bin` mesh format.

`generate_hex_mesh` follows python/py_diff_pd/common/hex_mesh.py:8-68: vertices
are numbered in i -> j -> k order over the nodes touched by a voxel, element
nodes in the order 000, 001, 010, 011, 100, 101, 110, 111. The `.bin` layout is
the one of cpp/core/src/mesh/mesh.cpp:4-32 (int vdim, int edim, then vertices
and elements as column-major matrices with int rows/cols headers).
"""
import struct

import numpy as np


def hex_grid(voxels, dx, origin):
    """Returns (vertices 3xN float64, elements 8xE int32, voxel_indices, vertex_flag)."""
    voxels = np.asarray(voxels).astype(bool)
    origin = np.asarray(origin, dtype=np.float64)
    cx, cy, cz = voxels.shape
    flag = np.zeros((cx + 1, cy + 1, cz + 1), dtype=bool)
    for ii in (0, 1):
        for jj in (0, 1):
            for kk in (0, 1):
                flag[ii:ii + cx, jj:jj + cy, kk:kk + cz] |= voxels
    vertex_flag = np.full(flag.shape, -1, dtype=np.int64)
    idx = np.flatnonzero(flag.ravel())  # C order == i -> j -> k
    vertex_flag.ravel()[idx] = np.arange(idx.size)
    I, J, K = np.unravel_index(idx, flag.shape)
    vertices = np.stack([origin[0] + dx * I, origin[1] + dx * J, origin[2] + dx * K]).astype(np.float64)
    cells = np.flatnonzero(voxels.ravel())
    ci, cj, ck = np.unravel_index(cells, voxels.shape)
    elems = []
    for ii in (0, 1):
        for jj in (0, 1):
            for kk in (0, 1):
                elems.append(vertex_flag[ci + ii, cj + jj, ck + kk])
    elements = np.stack(elems).astype(np.int32)
    voxel_indices = np.full(voxels.shape, -1, dtype=np.int64)
    voxel_indices.ravel()[cells] = np.arange(cells.size)
    return vertices, elements, voxel_indices, vertex_flag


def write_bin(path, vertices, elements):
    V = np.asarray(vertices, dtype=np.float64)
    E = np.asarray(elements, dtype=np.int32)
    with open(path, "wb") as f:
        f.write(struct.pack("ii", V.shape[0], E.shape[0]))
        f.write(struct.pack("ii", V.shape[0], V.shape[1]))
        f.write(np.ascontiguousarray(V.T).tobytes())
        f.write(struct.pack("ii", E.shape[0], E.shape[1]))
        f.write(np.ascontiguousarray(E.T).tobytes())


def generate_hex_mesh(voxels, dx, origin, bin_file_name, write=True):
    V, E, voxel_indices, vertex_flag = hex_grid(voxels, dx, origin)
    if write:
        write_bin(bin_file_name, V, E)
    return voxel_indices, vertex_flag


def tet_grid(n, dx, origin):
    """An n x n x n cube grid split into 5 tets per cube with alternating parity
    (no reference counterpart; the reference loads tetgen meshes)."""
    nx = ny = nz = n
    node = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    origin = np.asarray(origin, dtype=np.float64)
    V = np.stack([origin[0] + dx * I.ravel(), origin[1] + dx * J.ravel(), origin[2] + dx * K.ravel()]).astype(np.float64)
    even = [(0, 3, 5, 6), (0, 1, 3, 5), (0, 2, 3, 6), (0, 4, 5, 6), (3, 5, 6, 7)]
    odd = [(1, 2, 4, 7), (1, 0, 2, 4), (1, 3, 2, 7), (1, 5, 4, 7), (2, 4, 6, 7)]
    tets = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                c = [node(i + a, j + b, k + d) for a in (0, 1) for b in (0, 1) for d in (0, 1)]
                pat = even if (i + j + k) % 2 == 0 else odd
                for t in pat:
                    tets.append([c[t[0]], c[t[1]], c[t[2]], c[t[3]]])
    E = np.asarray(tets, dtype=np.int32).T
    # orient every tet positively
    P = V[:, E]
    vol = np.einsum("ij,ij->j", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0], axis=0))
    flip = vol < 0
    E[1, flip], E[2, flip] = E[2, flip].copy(), E[1, flip].copy()
    return V, np.ascontiguousarray(E)


def read_tetgen_file(node_file, ele_file):
    """Vertices (n x 3) and tets (m x 4, zero-based) from tetgen's `.node` and
    `.ele` text files (python/py_diff_pd/common/tet_mesh.py:312-324): the first
    line of each holds the count; each following line holds an index then the
    3 coordinates / 4 vertex indices. As in the reference, the `.ele` indices
    are taken as one-based; extra columns (attributes, markers) are ignored."""
    with open(node_file, "r") as f:
        rows = [l.split() for l in f if l.strip() and not l.lstrip().startswith("#")]
    nv = int(rows[0][0])
    verts = np.asarray([[float(x) for x in rows[i + 1][1:4]] for i in range(nv)], dtype=np.float64)
    with open(ele_file, "r") as f:
        rows = [l.split() for l in f if l.strip() and not l.lstrip().startswith("#")]
    ne = int(rows[0][0])
    elements = np.asarray([[int(x) for x in rows[i + 1][1:5]] for i in range(ne)], dtype=np.int64) - 1
    return verts, elements

