"""Rooted trees: Newick parsing, node numbering, traversal and the level plan.

MCPhylo keeps its tree type and tree algorithms in the un-vendored package
MCPhyloTree 1.1 (reference `Project.toml:20`, re-exported at
`src/MCPhylo.jl:17`).  The likelihood path uses only a handful of its
functions (`post_order`, `get_leaves`, `get_mother`,
`get_branchlength_vector`, `set_branchlength_vector!`, the `.num`,
`.children`, `.nchild` fields; call sites `src/distributions/Phylodist.jl:108,
115-116,125-129`, `src/Likelihood/LikelihoodCalculator_Node.jl:22,25-26,36,71`).
This module provides the same semantics for that path and nothing more.

Numbering.  Nodes are numbered 1..N with the root last (num == N); the
gradient vector and the branch-length vector are indexed by num
(`LikelihoodCalculator_Node.jl:12,69`).  Leaves take 1..n in the order they
appear in the Newick string and internal nodes take n+1..N in post-order.
That convention is the one under which the reference's golden gradient
(`test/likelihood/felsenstein.jl:37`) is reproduced (tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "Node",
    "ParseNewick",
    "post_order",
    "pre_order",
    "get_leaves",
    "get_mother",
    "number_nodes",
    "get_branchlength_vector",
    "set_branchlength_vector",
    "to_newick",
    "TreeArrays",
    "tree_arrays",
    "tree_from_parent",
]


class Node:
    """A tree node: `name`, `inc_length` (length of the branch above it),
    `children` (ordered), `mother` and `num` (1-based, root == N)."""

    __slots__ = ("name", "inc_length", "children", "mother", "num")

    def __init__(self, name: str = "", inc_length: float = 0.0):
        self.name = name
        self.inc_length = float(inc_length)
        self.children: list[Node] = []
        self.mother: Node | None = None
        self.num = 0

    @property
    def nchild(self) -> int:
        return len(self.children)

    def add_child(self, child: "Node") -> None:
        child.mother = self
        self.children.append(child)

    def __repr__(self) -> str:
        return f"Node(name={self.name!r}, num={self.num}, nchild={self.nchild})"


# --------------------------------------------------------------------------
# Newick
# --------------------------------------------------------------------------

def _tokenize(s: str):
    i, n = 0, len(s)
    while i < n:
        ch = s[i]
        if ch.isspace():
            i += 1
            continue
        if ch in "(),:;":
            yield ch
            i += 1
            continue
        if ch == "'":
            j = s.index("'", i + 1)
            yield ("LABEL", s[i + 1 : j])
            i = j + 1
            continue
        j = i
        while j < n and s[j] not in "(),:;" and not s[j].isspace():
            j += 1
        yield ("LABEL", s[i:j])
        i = j


def ParseNewick(newick: str) -> Node:
    """Parse one Newick string into a rooted tree and number it.

    Whitespace and newlines between tokens are ignored (the reference's
    fixture strings span several lines, `test/likelihood/felsenstein.jl:8-14`).
    Raises ValueError on malformed input.
    """
    toks = list(_tokenize(newick))
    if not toks or toks[-1] != ";":
        raise ValueError("Newick string must end with ';'")
    pos = 0

    def parse_label_length(node: Node):
        nonlocal pos
        if pos < len(toks) and isinstance(toks[pos], tuple):
            node.name = toks[pos][1]
            pos += 1
        if pos < len(toks) and toks[pos] == ":":
            pos += 1
            if pos >= len(toks) or not isinstance(toks[pos], tuple):
                raise ValueError("branch length expected after ':'")
            node.inc_length = float(toks[pos][1])
            pos += 1

    def parse_subtree() -> Node:
        nonlocal pos
        node = Node()
        if toks[pos] == "(":
            pos += 1
            while True:
                node.add_child(parse_subtree())
                if toks[pos] == ",":
                    pos += 1
                    continue
                if toks[pos] == ")":
                    pos += 1
                    break
                raise ValueError(f"unexpected token {toks[pos]!r}")
        parse_label_length(node)
        return node

    root = parse_subtree()
    if toks[pos] != ";":
        raise ValueError("trailing tokens after tree")
    number_nodes(root)
    return root


def to_newick(root: Node, digits: int = 17) -> str:
    def rec(v: Node) -> str:
        s = ""
        if v.children:
            s = "(" + ",".join(rec(c) for c in v.children) + ")"
        s += v.name
        if v.mother is not None:
            s += ":" + repr(float(v.inc_length)) if digits >= 17 else f":{v.inc_length:.{digits}g}"
        return s

    return rec(root) + ";"


# --------------------------------------------------------------------------
# Traversal and numbering
# --------------------------------------------------------------------------

def post_order(root: Node) -> list[Node]:
    """Children before parents, children visited in their stored order; the
    root is last (MCPhyloTree `post_order`)."""
    out: list[Node] = []
    stack = [(root, False)]
    while stack:
        v, done = stack.pop()
        if done:
            out.append(v)
            continue
        stack.append((v, True))
        for c in reversed(v.children):
            stack.append((c, False))
    return out


def pre_order(root: Node) -> list[Node]:
    out: list[Node] = []
    stack = [root]
    while stack:
        v = stack.pop()
        out.append(v)
        stack.extend(reversed(v.children))
    return out


def get_leaves(root: Node) -> list[Node]:
    return [v for v in post_order(root) if not v.children]


def get_mother(node: Node) -> Node:
    if node.mother is None:
        raise ValueError("the root has no mother")
    return node.mother


def number_nodes(root: Node) -> None:
    """Leaves 1..n in Newick (left-to-right) order, internal nodes n+1..N in
    post-order, so the root is N."""
    po = post_order(root)
    k = 1
    for v in po:
        if not v.children:
            v.num = k
            k += 1
    for v in po:
        if v.children:
            v.num = k
            k += 1


def get_branchlength_vector(root: Node) -> np.ndarray:
    """Branch lengths indexed by num (length N-1; the root has no branch)."""
    po = post_order(root)
    blv = np.zeros(len(po) - 1)
    for v in po:
        if v.mother is not None:
            blv[v.num - 1] = v.inc_length
    return blv


def set_branchlength_vector(root: Node, blv) -> None:
    blv = np.asarray(blv, dtype=np.float64)
    po = post_order(root)
    if blv.shape != (len(po) - 1,):
        raise ValueError(f"branch length vector must have length {len(po) - 1}")
    for v in po:
        if v.mother is not None:
            v.inc_length = float(blv[v.num - 1])


# --------------------------------------------------------------------------
# Flat arrays for the C ABI and the oracle
# --------------------------------------------------------------------------

class TreeArrays:
    """Flat, num-indexed description of a rooted tree.

    parent[i]   parent num of node num i+1 (0 for the root), int32
    postorder   node nums in post-order, root last, int32
    child_ptr   CSR offsets (N+1) into child_idx, int32
    child_idx   child nums in stored child order, int32
    leaf_nums   nums of the leaves in post-order
    """

    def __init__(self, parent, postorder, child_ptr, child_idx):
        self.parent = np.ascontiguousarray(parent, dtype=np.int32)
        self.postorder = np.ascontiguousarray(postorder, dtype=np.int32)
        self.child_ptr = np.ascontiguousarray(child_ptr, dtype=np.int32)
        self.child_idx = np.ascontiguousarray(child_idx, dtype=np.int32)
        self.N = int(self.parent.shape[0])
        nch = np.diff(self.child_ptr)
        self.leaf_nums = np.array([v for v in self.postorder if nch[v - 1] == 0], dtype=np.int32)

    @property
    def n_leaves(self) -> int:
        return int(self.leaf_nums.shape[0])

    def topology_key(self) -> bytes:
        return self.parent.tobytes() + self.child_idx.tobytes()


def tree_arrays(root: Node) -> TreeArrays:
    po = post_order(root)
    N = len(po)
    if root.num != N:
        raise ValueError("root must carry num == N")
    parent = np.zeros(N, dtype=np.int32)
    kids: list[list[int]] = [[] for _ in range(N)]
    for v in po:
        if v.mother is not None:
            parent[v.num - 1] = v.mother.num
        kids[v.num - 1] = [c.num for c in v.children]
    child_ptr = np.zeros(N + 1, dtype=np.int32)
    for i in range(N):
        child_ptr[i + 1] = child_ptr[i] + len(kids[i])
    child_idx = np.array([c for k in kids for c in k], dtype=np.int32)
    return TreeArrays(parent, [v.num for v in po], child_ptr, child_idx)


def tree_from_parent(parent, blv=None, names=None) -> Node:
    """Build a Node tree from a num-indexed parent array (root has parent 0).
    Children keep increasing-num order.  Nums are kept as given."""
    parent = np.asarray(parent)
    N = parent.shape[0]
    nodes = [Node() for _ in range(N)]
    for i, v in enumerate(nodes):
        v.num = i + 1
        v.name = names[i] if names is not None else str(i + 1)
    root = None
    for i in range(N):
        p = int(parent[i])
        if p == 0:
            if root is not None:
                raise ValueError("more than one root")
            root = nodes[i]
        else:
            nodes[p - 1].add_child(nodes[i])
            if blv is not None:
                nodes[i].inc_length = float(blv[i])
    if root is None or root.num != N:
        raise ValueError("root must be num N")
    return root
