"""Fill-reducing orderings for symmetric sparsity patterns.

The reference obtains its permutation from the AMD routine bundled with the
``qdldl`` Python package (``warpmpc/jax_qdldl/core.py:276-279``).  That package is
not part of this project, so this module implements the approximate minimum
degree algorithm of Amestoy, Davis and Duff ("An approximate minimum degree
ordering algorithm", SIAM J. Matrix Anal. Appl. 17(4), 1996) from the paper:

* quotient-graph elimination with element absorption,
* the approximate external degree bound of section 3 (eq. 3.4),
* aggressive element absorption,
* supervariable detection by hashing (indistinguishable variables),
* ties broken by the smallest variable index, so the result is deterministic.

The permutation it returns will generally differ from SuiteSparse AMD's in tie
breaking, so L/D are only elementwise comparable with the reference when the
reference's permutation is supplied explicitly (``build_qdldl_plan(..., p=...)``).
"""

from __future__ import annotations

import heapq

import numpy as np


def _symmetric_adjacency(n: int, indptr: np.ndarray, indices: np.ndarray) -> list[set[int]]:
    adj: list[set[int]] = [set() for _ in range(n)]
    for j in range(n):
        for ptr in range(int(indptr[j]), int(indptr[j + 1])):
            i = int(indices[ptr])
            if i != j:
                adj[i].add(j)
                adj[j].add(i)
    return adj


def amd_order(n: int, indptr: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Approximate minimum degree permutation ``p`` (pivot ``k`` is ``p[k]``).

    ``indptr/indices`` describe any CSC pattern whose symmetric closure is the
    graph to order (upper, lower or full storage all work).
    """

    a_adj = _symmetric_adjacency(n, np.asarray(indptr), np.asarray(indices))
    e_adj: list[set[int]] = [set() for _ in range(n)]
    elements: dict[int, set[int]] = {}
    nv = np.ones(n, dtype=np.int64)  # supervariable weights (0 = absorbed)
    members: list[list[int]] = [[i] for i in range(n)]
    alive = np.ones(n, dtype=bool)  # principal, not yet eliminated
    degree = np.array([len(a_adj[i]) for i in range(n)], dtype=np.int64)
    heap = [(int(degree[i]), i) for i in range(n)]
    heapq.heapify(heap)
    remaining = n
    order: list[int] = []

    def weight(items) -> int:
        return int(sum(int(nv[j]) for j in items))

    while remaining > 0:
        while True:
            deg, piv = heapq.heappop(heap)
            if alive[piv] and deg == degree[piv]:
                break
        # Lp = (A_p U (U_{e in E_p} L_e)) \ {p}
        lp_set = set(a_adj[piv])
        for e in e_adj[piv]:
            lp_set |= elements[e]
        lp_set.discard(piv)
        lp_set = {j for j in lp_set if alive[j]}

        order.extend(members[piv])
        remaining -= int(nv[piv])
        alive[piv] = False
        absorbed = e_adj[piv]
        for e in absorbed:
            del elements[e]
        elements[piv] = lp_set
        a_adj[piv] = set()
        e_adj[piv] = set()

        for i in lp_set:
            e_adj[i] -= absorbed
            e_adj[i].add(piv)
            a_adj[i] -= lp_set
            a_adj[i].discard(piv)

        # |L_e \ L_p| (weighted) for every element touching L_p (eq. 3.3 of the paper)
        lp_weight = weight(lp_set)
        w_ext: dict[int, int] = {}
        for i in lp_set:
            for e in e_adj[i]:
                if e == piv:
                    continue
                if e not in w_ext:
                    w_ext[e] = weight(elements[e])
                w_ext[e] -= int(nv[i])

        # aggressive absorption: elements fully inside L_p are absorbed into p
        for e, w in list(w_ext.items()):
            if w == 0:
                for j in elements[e]:
                    e_adj[j].discard(e)
                del elements[e]
                del w_ext[e]

        for i in lp_set:
            ext_lp = lp_weight - int(nv[i])
            a_w = weight(a_adj[i])
            e_w = sum(w_ext[e] for e in e_adj[i] if e != piv)
            bound = min(remaining - int(nv[i]), int(degree[i]) + ext_lp, a_w + ext_lp + e_w)
            degree[i] = max(bound, 0)

        # supervariable detection among the variables of the new element
        buckets: dict[int, list[int]] = {}
        for i in sorted(lp_set):
            key = (sum(a_adj[i]) + sum(e_adj[i])) % 1000003
            buckets.setdefault(key, []).append(i)
        for group in buckets.values():
            if len(group) < 2:
                continue
            for idx, i in enumerate(group):
                if not alive[i]:
                    continue
                for j in group[idx + 1 :]:
                    if not alive[j]:
                        continue
                    a_i = a_adj[i] - {j}
                    a_j = a_adj[j] - {i}
                    if a_i == a_j and e_adj[i] == e_adj[j]:
                        nv[i] += nv[j]
                        members[i].extend(members[j])
                        nv[j] = 0
                        alive[j] = False
                        for e in e_adj[j]:
                            elements[e].discard(j)
                        for k in a_adj[j]:
                            a_adj[k].discard(j)
                        a_adj[j] = set()
                        e_adj[j] = set()
            for i in group:
                if alive[i]:
                    # the merged variables are no longer external to i
                    degree[i] = min(
                        int(degree[i]),
                        weight(a_adj[i] - {i})
                        + sum(weight(elements[e] - {i}) for e in e_adj[i]),
                    )
        for i in lp_set:
            if alive[i]:
                heapq.heappush(heap, (int(degree[i]), i))

    perm = np.asarray(order, dtype=np.int32)
    if perm.size != n or not np.array_equal(np.sort(perm), np.arange(n)):
        raise AssertionError("ordering is not a permutation")
    return perm


def natural_order(n: int) -> np.ndarray:
    return np.arange(n, dtype=np.int32)


def etree_postorder(parent: np.ndarray) -> np.ndarray:
    """Postorder of an elimination forest, larger subtrees first.

    Returns ``post`` with ``post[i]`` = the node placed at position ``i``.  Any
    postorder of the elimination tree is an equivalent ordering (same L pattern
    up to relabelling, same fill).  Visiting the children of a node in order of
    decreasing subtree size puts small subtrees (for a KKT matrix: the leaf
    multipliers of box constraints) immediately before their parent, so a
    column of L stays open - referenced by later rows - only for a short span
    of rows.  This bounds the per-instance working set of the factorization by
    the band of the matrix instead of by its size.  Ties go to the smaller index.
    """
    parent = np.asarray(parent)
    n = parent.size
    size = np.ones(n, dtype=np.int64)
    for v in range(n):  # etree parents have larger indices than their children
        if parent[v] >= 0:
            size[parent[v]] += size[v]
    children: list[list[int]] = [[] for _ in range(n)]
    roots: list[int] = []
    for v in range(n):
        (children[parent[v]] if parent[v] >= 0 else roots).append(v)
    for ch in children:
        ch.sort(key=lambda c: (-int(size[c]), c))
    roots.sort(key=lambda c: (-int(size[c]), c))
    post: list[int] = []
    for r in roots:
        stack = [(r, 0)]
        while stack:
            v, i = stack.pop()
            if i < len(children[v]):
                stack.append((v, i + 1))
                stack.append((children[v][i], 0))
            else:
                post.append(v)
    return np.asarray(post, dtype=np.int32)


__all__ = ["amd_order", "etree_postorder", "natural_order"]
