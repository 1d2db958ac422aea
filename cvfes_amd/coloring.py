"""Greedy element colouring over shared-node adjacency (``meshColoring.py:7-63``).

Elements of one colour share no node, so a kernel that scatters element contributions into nodal
arrays can run a colour at a time without write races.  The matrix-free MI355X path gathers per
node instead and needs no colours; the colouring is kept for the element-push variant and for the
``colorGroups`` attribute that CVFES's solver reads.
"""
from __future__ import annotations

import numpy as np

__all__ = ["mesh_coloring"]


def mesh_coloring(elements: np.ndarray, n_nodes: int, n_colors: int = 23):
    """Return ``(elmColors [E] float, colorGroups list[int64 array])``.

    Element ``i`` takes the smallest colour not used by an already-coloured element sharing a
    node with it, in element order, as ``meshColoring.py:39-53`` does.  Raises ``RuntimeError`` when
    more than ``n_colors`` colours would be needed (the reference returns -1 there).
    """
    elements = np.asarray(elements, dtype=np.int64)
    n_elm = elements.shape[0]
    # Colour bitmask per node: bit c set when an element of colour c touches the node.
    node_mask = np.zeros(n_nodes, dtype=np.int64)
    colors = np.empty(n_elm, dtype=np.int64)
    for i in range(n_elm):
        a, b, c = elements[i]
        used = int(node_mask[a] | node_mask[b] | node_mask[c])
        col = 0
        while used & (1 << col):
            col += 1
        if col >= n_colors:
            raise RuntimeError(f"Colors are used up! {n_colors} colors are not enough.")
        colors[i] = col
        bit = 1 << col
        node_mask[a] |= bit
        node_mask[b] |= bit
        node_mask[c] |= bit
    n_used = int(colors.max()) + 1 if n_elm else 0
    groups = [np.nonzero(colors == k)[0].astype(np.int64) for k in range(n_used)]
    return colors.astype(float), groups
