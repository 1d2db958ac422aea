"""Block-CSR structure of the membrane stiffness and the index maps the GPU kernels use.

``SparseInfo`` restates CVFES's node-adjacency structure (``sparse.py:36-64``): row ``a`` of the
3x3-block matrix holds the sorted, de-duplicated set of nodes that share an element with node
``a`` (node ``a`` included).  ``indptr``/``indices`` are int64 as in the reference; the structure
is built with vectorised sorting instead of per-node Python lists, with an identical result.

``BlockMaps`` adds the integer maps that the MI355X kernels need on top of that structure:

* ``cptr``/``celem``/``cab``: for every block ``p`` the list of element contributions
  ``(e, a_local*3 + b_local)`` in increasing element order - the order in which the reference's
  ``assembling`` (``optimizedSolidAssemble.pyx:133-188``) adds them into ``K[p]``;
* ``cpos[e*9 + ab]``: the inverse map (where element ``e``'s block ``ab`` lands);
* ``nptr``/``nent``: node -> incident (element, local index) lists, used to pull lumped mass and
  nodal loads without scatter races;
* ``ublk``/``vblk``: symmetric storage - the upper blocks (col >= row) in block order, and for every
  block the upper entry that holds it (itself, or its mirror (col, row) read transposed).
"""
from __future__ import annotations

import numpy as np

__all__ = ["SparseInfo", "BlockMaps", "block_structure", "bisection_order", "strip_order"]


def block_structure(elements: np.ndarray, n_nodes: int):
    """Return ``(indptr, indices)`` int64 of the node-adjacency block structure."""
    elements = np.asarray(elements, dtype=np.int64)
    rows = np.repeat(elements, 3, axis=1).ravel()
    cols = np.tile(elements, (1, 3)).ravel()
    keys = np.unique(rows * n_nodes + cols)
    row_of = keys // n_nodes
    indices = (keys % n_nodes).astype(np.int64)
    indptr = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(np.bincount(row_of, minlength=n_nodes), out=indptr[1:])
    return indptr, indices


def bisection_order(nodes: np.ndarray, leaf: int = 64) -> np.ndarray:
    """Compact visiting order of the nodes: recursive coordinate bisection (split the longest extent
    at the median) down to ``leaf`` nodes, parts in tree order.  Neighbouring nodes end up a few
    positions apart, so a block read by one row is read again by its mirror row soon after."""
    nodes = np.asarray(nodes, np.float64)
    out = []
    stack = [np.arange(nodes.shape[0])]
    while stack:
        idx = stack.pop()
        if idx.size <= leaf:
            out.append(idx)
            continue
        pts = nodes[idx]
        ax = int(np.argmax(pts.max(0) - pts.min(0)))
        o = idx[np.argsort(pts[:, ax], kind="stable")]
        h = o.size // 2
        stack.append(o[h:])
        stack.append(o[:h])
    return np.concatenate(out).astype(np.int32) if out else np.zeros(0, np.int32)


def strip_order(nodes: np.ndarray, elements: np.ndarray, width: int = 64) -> np.ndarray:
    """Visiting order for tubular meshes: the wall is cut into angular strips about the tube axis,
    each about ``width`` nodes of a ring wide, and every strip is swept along the axis (in bands of
    half an edge length, by angle inside a band).  A node's neighbours in the next and previous
    ring are then about ``width`` positions away instead of a whole ring, so a block read by one
    row is read again by its mirror row soon after.

    Axis: the principal direction of the node cloud.  Ring size: circumference (median distance to
    the axis) over the median edge length."""
    x = np.asarray(nodes, np.float64)
    x = x - x.mean(0)
    n = x.shape[0]
    if n == 0:
        return np.zeros(0, np.int32)
    _, vec = np.linalg.eigh(x.T @ x)
    axis, e1 = vec[:, -1], vec[:, -2]
    e2 = np.cross(axis, e1)
    t = x @ axis
    perp = x - np.outer(t, axis)
    phi = np.mod(np.arctan2(perp @ e2, perp @ e1), 2.0 * np.pi)
    el = np.asarray(elements, np.int64)
    ed = np.concatenate([el[:, [0, 1]], el[:, [1, 2]], el[:, [2, 0]]]) if el.size else np.zeros((0, 2), np.int64)
    h = float(np.median(np.linalg.norm(x[ed[:, 0]] - x[ed[:, 1]], axis=1))) if ed.size else 1.0
    h = h if h > 0.0 else 1.0
    radius = float(np.median(np.linalg.norm(perp, axis=1)))
    n_ring = max(1.0, 2.0 * np.pi * radius / h)
    n_strips = max(1, int(round(n_ring / max(1, width))))
    strip = np.minimum((phi * (n_strips / (2.0 * np.pi))).astype(np.int64), n_strips - 1)
    # band edges a quarter band off the multiples of h, where the rings of structured tubes sit
    band = np.floor((t - t.min()) / (0.5 * h) + 0.25).astype(np.int64)
    return np.lexsort((phi, band, strip)).astype(np.int32)


class SparseInfo:
    """Block-CSR structure with the attributes of ``sparse.py:36-64``.

    ``indptr [nNodes+1]``, ``indices [length]`` (int64), ``length`` (number of 3x3 blocks),
    ``nNodes``, ``dof``, ``ndof``, ``dofArray``.
    """

    def __init__(self, mesh, dof=3):
        elements = np.asarray(mesh.elementNodeIds, dtype=np.int64)
        self.indptr, self.indices = block_structure(elements, mesh.nNodes)
        self.length = int(self.indices.size)
        self.nNodes = int(mesh.nNodes)
        self.dof = int(dof)
        self.ndof = self.nNodes * self.dof
        self.dofArray = np.arange(dof)

    def New(self):
        return np.zeros((self.length, self.dof, self.dof))


class BlockMaps:
    """Integer maps between elements, nodes and blocks (all int32, ready for the GPU)."""

    def __init__(self, elements: np.ndarray, n_nodes: int, indptr: np.ndarray, indices: np.ndarray):
        elements = np.asarray(elements, dtype=np.int64)
        n_elm = elements.shape[0]
        nnzb = indices.size
        if n_nodes * 3 >= 2 ** 31 or 9 * n_elm >= 2 ** 31 or nnzb >= 2 ** 31:
            raise ValueError("mesh too large for 32-bit block maps")
        rows = np.repeat(elements, 3, axis=1).ravel()          # a of (e, a, b), ab-major per element
        cols = np.tile(elements, (1, 3)).ravel()
        # Block index of each (row, col) pair: offset of col inside the row's sorted column list.
        keys = rows * n_nodes + cols
        row_keys = np.repeat(np.arange(n_nodes, dtype=np.int64), np.diff(indptr)) * n_nodes + indices
        pos = np.searchsorted(row_keys, keys)
        if np.any(pos >= nnzb) or np.any(row_keys[np.minimum(pos, nnzb - 1)] != keys):
            raise ValueError("element pair missing from the block structure")
        order = np.argsort(pos, kind="stable")
        self.cptr = np.zeros(nnzb + 1, dtype=np.int64)
        np.cumsum(np.bincount(pos, minlength=nnzb), out=self.cptr[1:])
        self.celem = (order // 9).astype(np.int32)
        self.cab = (order % 9).astype(np.int32)
        self.cpos = np.empty(9 * n_elm, dtype=np.int32)
        self.cpos[order] = np.arange(9 * n_elm, dtype=np.int32)
        self.block_of = pos.astype(np.int32)                   # [E*9] block of (e, ab)
        # Node -> (element, local) incidence.
        flat = elements.ravel()
        norder = np.argsort(flat, kind="stable")
        self.nptr = np.zeros(n_nodes + 1, dtype=np.int64)
        np.cumsum(np.bincount(flat, minlength=n_nodes), out=self.nptr[1:])
        self.nent = norder.astype(np.int32)                    # e*3 + local
        self.n_contrib = 9 * n_elm
        # slot of each contribution's element in its row node's incident-element list
        node_of_q = np.repeat(np.arange(n_nodes, dtype=np.int64), np.diff(self.nptr))
        elem_of_q = norder // 3
        slot_of_q = np.arange(3 * n_elm, dtype=np.int64) - self.nptr[node_of_q]
        qkeys = node_of_q * n_elm + elem_of_q
        qorder = np.argsort(qkeys)
        row_of_block = np.repeat(np.arange(n_nodes, dtype=np.int64), np.diff(indptr))
        ckeys = row_of_block[pos[order]] * n_elm + (order // 9)
        where = np.searchsorted(qkeys[qorder], ckeys)
        self.cslot = slot_of_q[qorder][where].astype(np.int32)
        self.max_node_elems = int(np.diff(self.nptr).max()) if n_nodes else 0
        # symmetric storage: upper blocks and, per block, the upper entry holding it
        up = indices >= row_of_block
        self.ublk = np.nonzero(up)[0].astype(np.int32)
        uindex = np.cumsum(up) - 1                              # upper index of upper blocks
        mirror = np.searchsorted(row_keys, indices * n_nodes + row_of_block)
        if np.any(row_keys[np.minimum(mirror, nnzb - 1)] != indices * n_nodes + row_of_block):
            raise ValueError("block structure is not symmetric")
        self.vblk = np.where(up, uindex, uindex[mirror]).astype(np.int32)
        self.nnzu = int(self.ublk.size)
        self.cptr = self.cptr.astype(np.int32)
        self.nptr = self.nptr.astype(np.int32)
