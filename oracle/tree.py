"""Tree topology and TDVP path finding (oracle, test infrastructure only).

Restates:
* ``core/graph_node.py:340-355``      neighbour order: parent first, then children.
* ``core/tree_structure.py:207-272``   ``get_leaves`` / ``distance_to_node`` (DFS
                                       pre-order dict, first maximum wins).
* ``core/tree_structure.py:609-690``   ``find_path_to_root`` / ``path_from_to`` /
                                       ``linearise`` (children first).
* ``time_evolution/time_evo_util/update_path.py:63-254``  LeafToRoot update path.
* ``time_evolution/time_evo_util/update_path.py:400-420`` orthogonalisation path.
"""
from __future__ import annotations

from typing import Dict, List, Optional


class Topology:
    """Rooted tree with ordered children, nodes kept in insertion order."""

    def __init__(self):
        self.parent: Dict[str, Optional[str]] = {}
        self.children: Dict[str, List[str]] = {}
        self.root: Optional[str] = None

    # -- construction -----------------------------------------------------
    def add_root(self, nid: str):
        assert self.root is None
        self.root = nid
        self.parent[nid] = None
        self.children[nid] = []

    def add_child(self, nid: str, parent: str):
        assert nid not in self.parent and parent in self.parent
        self.parent[nid] = parent
        self.children[nid] = []
        self.children[parent].append(nid)

    @staticmethod
    def from_parent_map(root: str, edges: List[tuple]) -> "Topology":
        """edges: list of (child, parent) in insertion order."""
        t = Topology()
        t.add_root(root)
        for child, parent in edges:
            t.add_child(child, parent)
        return t

    # -- queries ----------------------------------------------------------
    @property
    def node_ids(self) -> List[str]:
        return list(self.parent.keys())

    def neighbours(self, nid: str) -> List[str]:
        p = self.parent[nid]
        return ([] if p is None else [p]) + list(self.children[nid])

    def neighbour_index(self, nid: str, other: str) -> int:
        return self.neighbours(nid).index(other)

    def is_leaf(self, nid: str) -> bool:
        return len(self.children[nid]) == 0

    def get_leaves(self) -> List[str]:
        leaves = [n for n in self.parent if self.is_leaf(n)]
        if len(self.neighbours(self.root)) == 1:
            leaves.append(self.root)
        return leaves

    def distance_to_node(self, center: str) -> Dict[str, int]:
        out = {center: 0}

        def rec(nid, last, dist):
            out[nid] = dist
            for nb in self.neighbours(nid):
                if nb != last:
                    rec(nb, nid, dist + 1)

        for nb in self.neighbours(center):
            rec(nb, center, 1)
        return out

    def linearise(self) -> List[str]:
        order = []

        def rec(nid):
            for c in self.children[nid]:
                rec(c)
            order.append(nid)

        rec(self.root)
        return order

    def find_path_to_root(self, nid: str) -> List[str]:
        path = [nid]
        while self.parent[nid] is not None:
            nid = self.parent[nid]
            path.append(nid)
        return path

    def path_from_to(self, start: str, end: str) -> List[str]:
        if start == end:
            return [start]
        up = self.find_path_to_root(start)
        down = self.find_path_to_root(end)
        common = set(up) & set(down)
        # lowest common ancestor = first element of ``up`` that is shared
        lca = next(n for n in up if n in common)
        first = up[: up.index(lca) + 1]
        second = down[: down.index(lca)]
        return first + list(reversed(second))

    def subtree_size(self, nid: str) -> int:
        return 1 + sum(self.subtree_size(c) for c in self.children[nid])

    # -- TDVP paths (LeafToRoot) -----------------------------------------------
    def _branch_rec(self, nid: str) -> List[str]:
        if self.is_leaf(nid):
            return [nid]
        path = []
        for c in self.children[nid]:
            path.extend(self._branch_rec(c))
        path.append(nid)
        return path

    def update_path(self) -> List[str]:
        dist = self.distance_to_node(self.root)
        start = max(dist, key=dist.get)
        main_path = self.find_path_to_root(start)
        path: List[str] = []
        for origin in main_path:
            if origin != self.root:
                for c in self.children[origin]:
                    if c not in main_path:
                        path.extend(self._branch_rec(c))
                path.append(origin)
            else:
                path.extend(self._path_down_from_root(path, main_path, dist))
        return path

    def _path_down_from_root(self, visited, main_path, dist) -> List[str]:
        if len(self.children[self.root]) == 1:
            return [self.root]
        candidates = [l for l in self.get_leaves() if l not in visited]
        leaf_dist = {n: d for n, d in dist.items() if n in candidates}
        final = max(leaf_dist, key=leaf_dist.get)
        down = list(reversed(self.find_path_to_root(final)))
        out: List[str] = []
        for origin in down:
            if origin == self.root:
                excluded = (main_path[-2], down[1])
            else:
                excluded = tuple(down)
            for c in self.children[origin]:
                if c not in excluded:
                    out.extend(self._branch_rec(c))
            out.append(origin)
        return out

    def orthogonalisation_path(self, update_path: List[str]) -> List[List[str]]:
        return [self.path_from_to(update_path[i], update_path[i + 1])[1:]
                for i in range(len(update_path) - 1)]

    def second_order_one_site_back_orth_path(self, back_path: List[str]) -> List[List[str]]:
        out = []
        for i, nid in enumerate(back_path[1:-1]):
            out.append(self.path_from_to(nid, back_path[i + 2])[:-1])
        out.append([back_path[-1]])
        return out
