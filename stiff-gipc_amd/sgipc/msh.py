"""Gmsh .msh reader (formats 2.2 and 4.1, ASCII and binary) and the
tetrahedral-mesh view the simulator loads.

The reference reads meshes with the vendored MshIO library
(MshHelper/MshIO, mshio::load_msh) and keeps the nodes in file order and
every element of type 4 (4-node tetrahedron) (gipc::io::load_tet_mesh,
MshHelper/msh_io.cpp:12-90): node tags are re-indexed 0 .. n-1 in the order
the node blocks list them, and tet node tags are mapped through that index.

read_msh returns the sections MshIO parses and its tests compare (nodes per
entity block, elements per entity block); load_tet_mesh returns (V, T).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

# nodes per element of the Gmsh element types a tet mesh file carries
_NODES_PER_TYPE = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6, 10: 9, 11: 10,
                   15: 1}


@dataclass
class NodeBlock:
    entity_dim: int
    entity_tag: int
    parametric: int
    tags: np.ndarray          # (n,) int64
    xyz: np.ndarray           # (n, 3) float64


@dataclass
class ElementBlock:
    entity_dim: int
    entity_tag: int
    element_type: int
    tags: np.ndarray          # (n,) int64
    nodes: np.ndarray         # (n, nodes_per_element) int64 node tags


@dataclass
class Msh:
    version: str
    binary: bool
    node_blocks: list = field(default_factory=list)
    element_blocks: list = field(default_factory=list)

    @property
    def num_nodes(self):
        return int(sum(b.tags.size for b in self.node_blocks))

    @property
    def num_elements(self):
        return int(sum(b.tags.size for b in self.element_blocks))


class _Reader:
    def __init__(self, data: bytes):
        self.d = data
        self.p = 0

    def line(self) -> str:
        e = self.d.index(b"\n", self.p)
        s = self.d[self.p:e].decode("ascii", "replace").strip()
        self.p = e + 1
        return s

    def ints(self):
        return [int(t) for t in self.line().split()]

    def raw(self, n: int) -> bytes:
        b = self.d[self.p:self.p + n]
        if len(b) != n:
            raise ValueError("truncated binary section")
        self.p += n
        return b

    def skip_newline(self):
        if self.p < len(self.d) and self.d[self.p:self.p + 1] == b"\n":
            self.p += 1

    def tokens(self, n: int):
        """n whitespace-separated ASCII tokens, across lines"""
        out = []
        while len(out) < n:
            out.extend(self.line().split())
        return out


def _section_end(r: _Reader, name: str):
    """move past the $End<name> line (a skipped section may hold binary data)"""
    tag = f"$End{name}".encode()
    e = r.d.find(tag, r.p)
    if e < 0:
        raise ValueError(f"unterminated ${name} section")
    r.p = e
    r.line()


def read_msh(path) -> Msh:
    with open(path, "rb") as f:
        r = _Reader(f.read())
    if r.line() != "$MeshFormat":
        raise ValueError(f"{path}: not a Gmsh file")
    ver, ftype, dsize = r.line().split()
    binary = ftype == "1"
    if ver not in ("2.2", "4.1"):
        raise ValueError(f"{path}: unsupported MSH version {ver}")
    if binary:
        one = struct.unpack("<i", r.raw(4))[0]
        if one != 1:
            raise ValueError(f"{path}: big-endian binary files are not supported")
        r.skip_newline()
    if int(dsize) != 8:
        raise ValueError(f"{path}: data size must be 8")
    _section_end(r, "MeshFormat")
    m = Msh(version=ver, binary=binary)
    while r.p < len(r.d):
        s = r.line()
        if not s:
            continue
        if s == "$Nodes":
            (_nodes_v41 if ver == "4.1" else _nodes_v22)(r, m, binary)
            _section_end(r, "Nodes")
        elif s == "$Elements":
            (_elements_v41 if ver == "4.1" else _elements_v22)(r, m, binary)
            _section_end(r, "Elements")
        elif s.startswith("$"):
            _section_end(r, s[1:])
    return m


def _nodes_v41(r, m, binary):
    if binary:
        nblk, nn, lo, hi = struct.unpack("<4Q", r.raw(32))
    else:
        nblk, nn, lo, hi = r.ints()
    for _ in range(nblk):
        if binary:
            dim, tag, par = struct.unpack("<3i", r.raw(12))
            n = struct.unpack("<Q", r.raw(8))[0]
        else:
            dim, tag, par, n = r.ints()
        # parametric nodes of curves/surfaces carry dim extra coordinates (u[, v])
        w = 3 + (dim if par and dim in (1, 2) else 0)
        if binary:
            tags = np.frombuffer(r.raw(8 * n), "<u8").astype(np.int64)
            xyz = np.frombuffer(r.raw(8 * w * n), "<f8").reshape(n, w)[:, :3].copy()
        else:
            tags = np.array([int(t) for t in r.tokens(n)], np.int64)
            xyz = np.array([float(t) for t in r.tokens(w * n)], np.float64).reshape(n, w)[:, :3]
        m.node_blocks.append(NodeBlock(dim, tag, par, tags, xyz))
    if binary:
        r.skip_newline()


def _elements_v41(r, m, binary):
    if binary:
        nblk, ne, lo, hi = struct.unpack("<4Q", r.raw(32))
    else:
        nblk, ne, lo, hi = r.ints()
    for _ in range(nblk):
        if binary:
            dim, tag, etype = struct.unpack("<3i", r.raw(12))
            n = struct.unpack("<Q", r.raw(8))[0]
        else:
            dim, tag, etype, n = r.ints()
        k = _NODES_PER_TYPE.get(etype)
        if k is None:
            raise ValueError(f"unsupported element type {etype}")
        if binary:
            a = np.frombuffer(r.raw(8 * n * (k + 1)), "<u8").astype(np.int64).reshape(n, k + 1)
        else:
            a = np.array([int(t) for t in r.tokens(n * (k + 1))], np.int64).reshape(n, k + 1)
        m.element_blocks.append(ElementBlock(dim, tag, etype, a[:, 0].copy(), a[:, 1:].copy()))
    if binary:
        r.skip_newline()


def _nodes_v22(r, m, binary):
    n = int(r.line())
    if binary:
        rec = np.frombuffer(r.raw(28 * n), np.dtype([("t", "<i4"), ("x", "<f8", 3)]))
        tags, xyz = rec["t"].astype(np.int64), rec["x"].copy()
        r.skip_newline()
    else:
        rows = [r.line().split() for _ in range(n)]
        tags = np.array([int(w[0]) for w in rows], np.int64)
        xyz = np.array([[float(v) for v in w[1:4]] for w in rows], np.float64)
    m.node_blocks.append(NodeBlock(3, 0, 0, tags, xyz.reshape(-1, 3)))


def _elements_v22(r, m, binary):
    n = int(r.line())
    blocks = {}
    if binary:
        got = 0
        while got < n:
            etype, cnt, ntag = struct.unpack("<3i", r.raw(12))
            k = _NODES_PER_TYPE[etype]
            a = np.frombuffer(r.raw(4 * cnt * (1 + ntag + k)), "<i4").reshape(cnt, 1 + ntag + k)
            for row in a:
                ent = int(row[2]) if ntag >= 2 else 0
                blocks.setdefault((etype, ent), []).append((int(row[0]), row[1 + ntag:]))
            got += cnt
        r.skip_newline()
    else:
        for _ in range(n):
            w = [int(t) for t in r.line().split()]
            etag, etype, ntag = w[0], w[1], w[2]
            ent = w[4] if ntag >= 2 else 0
            k = _NODES_PER_TYPE[etype]
            blocks.setdefault((etype, ent), []).append((etag, w[3 + ntag:3 + ntag + k]))
    for (etype, ent), rows in blocks.items():
        tags = np.array([t for t, _ in rows], np.int64)
        nodes = np.array([list(v) for _, v in rows], np.int64)
        dim = {1: 1, 2: 2, 3: 2, 4: 3, 5: 3, 6: 3, 7: 3, 15: 0}.get(etype, 3)
        m.element_blocks.append(ElementBlock(dim, ent, etype, tags, nodes))


def load_tet_mesh(path):
    """(V (n,3) float64, T (m,4) int64) in the reference's conventions: nodes
    in block order re-indexed from 0, every type-4 element in block order."""
    m = read_msh(path)
    if m.num_nodes == 0:
        raise ValueError("Empty gmsh file (no verts)")
    if m.num_elements == 0:
        raise ValueError("Empty gmsh file (no elements)")
    tags = np.concatenate([b.tags for b in m.node_blocks])
    V = np.concatenate([b.xyz for b in m.node_blocks])
    tag_to_idx = np.full(int(tags.max()) + 1, -1, np.int64)
    tag_to_idx[tags] = np.arange(tags.size)
    tets = [b.nodes for b in m.element_blocks if b.element_type == 4]
    if not tets:
        raise ValueError("No tetrahedra found in mesh file")
    T = tag_to_idx[np.concatenate(tets)]
    if (T < 0).any():
        raise ValueError("a tetrahedron references an unknown node")
    return V, T
