"""Tetrahedral billet meshes and their topology (host side, numpy).

* ``kuhn_cylinder`` builds the structured, jittered Kuhn-split cylinder used by the upset and closed-die
  workloads: a box of hex cells mapped onto a disc cross-section, each hex split into 6 tets along its
  (0,0,0)-(1,1,1) diagonal, interior nodes jittered from a seed.
* ``TetMesh`` mirrors the reference's immutable topology record
  (src/ImpressionFEA.Core/ThreeD/TetMesh.cs:7-101): node->tet lists, outward boundary faces with owners and
  symmetry flags, surface-node flags.
* ``build_pattern`` mirrors the ANP two-ring sparsity pattern of the 3D solver
  (src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs:469-521): nodes i, j couple when both lie in the patch of
  some node; full symmetric CSC of 3x3 blocks with sorted neighbour lists.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np

# Kuhn split of the unit hex: one tet per axis permutation, a monotone path from corner 000 to corner 111.
_PERMS = list(itertools.permutations(range(3)))


def _perm_sign(p) -> int:
    s = 1
    p = list(p)
    for i in range(3):
        for j in range(i + 1, 3):
            if p[i] > p[j]:
                s = -s
    return s


def _kuhn_corner_paths():
    """Per tet: 4 corner offsets (a, b, c) in {0,1}^3, positively oriented on the unit cube."""
    out = []
    for p in _PERMS:
        v0 = np.zeros(3, dtype=np.int64)
        v1 = v0.copy()
        v1[p[0]] = 1
        v2 = v1.copy()
        v2[p[1]] = 1
        v3 = np.ones(3, dtype=np.int64)
        path = [v0, v1, v2, v3]
        if _perm_sign(p) < 0:  # det[e_p0, e_p1, e_p2] = sign(p): swap to keep det > 0
            path = [v0, v2, v1, v3]
        out.append(path)
    return out


_KUHN = _kuhn_corner_paths()


def kuhn_box(nx: int, ny: int, nz: int):
    """Node grid indices and tets of an nx*ny*nz hex box, each hex split into 6 Kuhn tets (cell-major order)."""
    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # cell order: i fastest, then j, then k
    ci = ii.transpose(2, 1, 0).ravel()
    cj = jj.transpose(2, 1, 0).ravel()
    ck = kk.transpose(2, 1, 0).ravel()
    tets = np.empty((ci.size, 6, 4), dtype=np.int64)
    for t, path in enumerate(_KUHN):
        for a, off in enumerate(path):
            tets[:, t, a] = nid(ci + off[0], cj + off[1], ck + off[2])
    return tets.reshape(-1, 4)


def kuhn_cylinder(diameter: float, height: float, nxy: int, nz: int, seed: int = 0, jitter: float = 0.05,
                  quarter: bool = False, z0: float = 0.0):
    """Structured jittered Kuhn mesh of a vertical cylinder.

    The (u, v) parameter square ([-1, 1]^2, or [0, 1]^2 for a quarter model) maps onto the disc of radius D/2 by
    u' = u*sqrt(1 - v^2/2), v' = v*sqrt(1 - u^2/2); z runs from z0 to z0 + H. Interior nodes move by a uniform
    jitter of +-``jitter`` of the cell size in each parameter direction before the mapping, so boundary nodes stay
    on the surface.

    Returns ``(x [nn,3] float64, tets [ne,4] int32, onX0 [nn] bool, onY0 [nn] bool)``.
    """
    nx = ny = nxy
    lo = 0.0 if quarter else -1.0
    u = lo + (1.0 - lo) * np.arange(nx + 1) / nx
    v = lo + (1.0 - lo) * np.arange(ny + 1) / ny
    w = np.arange(nz + 1) / nz
    U, V, W = np.meshgrid(u, v, w, indexing="ij")
    U = U.transpose(2, 1, 0).ravel().copy()
    V = V.transpose(2, 1, 0).ravel().copy()
    W = W.transpose(2, 1, 0).ravel().copy()
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    I = I.transpose(2, 1, 0).ravel()
    J = J.transpose(2, 1, 0).ravel()
    K = K.transpose(2, 1, 0).ravel()
    interior = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=(int(interior.sum()), 3))
    du, dv, dw = (1.0 - lo) / nx, (1.0 - lo) / ny, 1.0 / nz
    U[interior] += jit[:, 0] * du
    V[interior] += jit[:, 1] * dv
    W[interior] += jit[:, 2] * dw
    R = 0.5 * diameter
    X = R * U * np.sqrt(1.0 - 0.5 * V * V)
    Y = R * V * np.sqrt(1.0 - 0.5 * U * U)
    Z = z0 + height * W
    x = np.stack([X, Y, Z], axis=1)
    onX0 = np.zeros(x.shape[0], dtype=bool)
    onY0 = np.zeros(x.shape[0], dtype=bool)
    if quarter:
        onX0 = I == 0
        onY0 = J == 0
        x[onX0, 0] = 0.0
        x[onY0, 1] = 0.0
    tets = kuhn_box(nx, ny, nz).astype(np.int32)
    return x, tets, onX0, onY0


# faces of tet (a, b, c, d), outward for a positively oriented tet (TetMesh.cs:34)
_FACE_IDX = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])


@dataclass
class TetMesh:
    """Topology of a linear tet mesh (TetMesh.cs:9-64).

    Boundary faces are in order of first occurrence when the 4 faces of each tet are visited in element order,
    which is the order the reference's face dictionary yields them.
    """

    nn: int
    tets: np.ndarray  # [ne, 4] int32
    onX0: np.ndarray
    onY0: np.ndarray

    def __post_init__(self):
        t = np.ascontiguousarray(self.tets, dtype=np.int32)
        self.tets = t
        self.ne = t.shape[0]
        nn = self.nn
        flat = t.ravel()
        cnt = np.bincount(flat, minlength=nn)
        self.node_tet_start = np.zeros(nn + 1, dtype=np.int32)
        np.cumsum(cnt, out=self.node_tet_start[1:])
        # stable sort by node keeps elements ascending within each node's list (TetMesh.cs:18-26)
        order = np.argsort(flat, kind="stable")
        self.node_tet_list = (order // 4).astype(np.int32)

        faces = t[:, _FACE_IDX].reshape(-1, 3)  # [4 ne, 3] in visiting order
        key = np.sort(faces, axis=1).astype(np.int64)
        kid = (key[:, 0] * nn + key[:, 1]) * nn + key[:, 2]
        uniq, first, counts = np.unique(kid, return_index=True, return_counts=True)
        if counts.max(initial=0) > 2:
            raise ValueError("Tetrahedral mesh is not manifold (face shared by more than two tets).")
        bfirst = np.sort(first[counts == 1])
        self.boundary_faces = faces[bfirst].astype(np.int32)  # [nb, 3]
        self.boundary_face_owner = (bfirst // 4).astype(np.int32)
        bf = self.boundary_faces
        ox, oy = self.onX0, self.onY0
        self.face_is_symmetry = (ox[bf].all(axis=1)) | (oy[bf].all(axis=1))
        self.is_surface_node = np.zeros(nn, dtype=bool)
        self.is_surface_node[bf[~self.face_is_symmetry].ravel()] = True

    @property
    def ndof(self) -> int:
        return 3 * self.nn


def patches(mesh: TetMesh):
    """Sorted unique node set of the tets around each node (CSR: start [nn+1], nodes)."""
    t = mesh.tets
    nn = mesh.nn
    e = mesh.node_tet_list
    owner = np.repeat(np.arange(nn), np.diff(mesh.node_tet_start))
    pairs = np.stack([np.repeat(owner, 4), t[e].ravel()], axis=1).astype(np.int64)
    k = np.unique(pairs[:, 0] * nn + pairs[:, 1])
    own = (k // nn).astype(np.int32)
    nod = (k % nn).astype(np.int32)
    start = np.zeros(nn + 1, dtype=np.int32)
    np.cumsum(np.bincount(own, minlength=nn), out=start[1:])
    return start, nod


def build_pattern(mesh: TetMesh):
    """Two-ring pattern (Forging3DSolver.cs:469-521).

    Returns ``(nbr_start [nn+1], nbr [.] sorted per node, colPtr [ndof+1] int32, rowIdx [nnz] int32)``.
    nbr[i] = union of patch(n) over all n whose patch contains i.
    """
    nn = mesh.nn
    pstart, pnode = patches(mesh)
    plen = np.diff(pstart).astype(np.int64)
    keys = []
    chunk = 1 << 15
    for c0 in range(0, nn, chunk):
        c1 = min(nn, c0 + chunk)
        s0, s1 = int(pstart[c0]), int(pstart[c1])
        ln = plen[c0:c1]
        own = np.repeat(np.arange(c1 - c0), ln)              # local owner of each patch entry
        rep = ln[own]                                          # each entry pairs with its whole patch
        a = np.repeat(pnode[s0:s1].astype(np.int64), rep)
        blk = np.zeros(c1 - c0 + 1, dtype=np.int64)
        np.cumsum(ln * ln, out=blk[1:])
        t = np.arange(blk[-1], dtype=np.int64)
        ot = np.repeat(np.arange(c1 - c0), ln * ln)
        bidx = (pstart[c0:c1].astype(np.int64) - s0)[ot] + (t - blk[ot]) % ln[ot]
        b = pnode[s0:s1][bidx].astype(np.int64)
        keys.append(np.unique(a * nn + b))
    k = np.unique(np.concatenate(keys))
    ii = (k // nn).astype(np.int64)
    nbr = (k % nn).astype(np.int32)
    nbr_start = np.zeros(nn + 1, dtype=np.int64)
    np.cumsum(np.bincount(ii, minlength=nn), out=nbr_start[1:])
    deg = np.diff(nbr_start)
    ndof = 3 * nn
    colPtr = np.zeros(ndof + 1, dtype=np.int64)
    np.cumsum(np.repeat(3 * deg, 3), out=colPtr[1:])
    if colPtr[-1] >= 2**31:
        raise ValueError("pattern exceeds int32 indexing")
    # column 3j+c holds rows 3i, 3i+1, 3i+2 for i in nbr[j] (the same block list for c = 0, 1, 2)
    per_node_rows = (3 * nbr[:, None].astype(np.int64) + np.arange(3)[None, :]).ravel()
    col_nodes = np.repeat(np.arange(nn), 9 * deg)
    pos = np.arange(colPtr[-1], dtype=np.int64) - np.repeat(colPtr[3 * np.arange(nn)], 9 * deg)
    within = pos % (3 * deg[col_nodes])
    rowIdx = per_node_rows[3 * nbr_start[col_nodes] + within].astype(np.int32)
    return nbr_start.astype(np.int32), nbr, colPtr.astype(np.int32), rowIdx
