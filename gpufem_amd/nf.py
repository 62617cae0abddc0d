"""Neutral-file (.nf) meshes of the reference, read into a Problem.

Follows FileIO::ReadNF (src/Utils/fileIO.cc:73-404): the element block comes
first (the first line containing "%ELEMENT."), then the node coordinates
("%NODE.COORD"), the supports ("%NODE.SUPPORT") and the nodal loads
("%LOAD.CASE.NODAL.FORCE"); each block is the first line containing its key
after the previous block, searching once more from the top of the file when
"%END" is met first (FileIO::findParameter, fileIO.cc:761-795).  Element
lines skip 3 leading integers in 3D and 4 in 2D; Brick20 connectivity is
permuted from the file's node order into the element's node order
(fileIO.cc:184-204); coordinates keep dim components; support rows are
[node, fix_x, fix_y(, fix_z)] and load rows [node, Fx, Fy(, Fz)].
Materials are not read (the reference takes E and nu from its
configuration); E and nu are arguments here.
"""
import numpy as np

from . import GF_BRICK8, GF_BRICK20, GF_Q4, GF_Q8
from .problems import E_DEFAULT, NU_DEFAULT, Problem, _fixed_dofs

ELEMENTS = {
    "%ELEMENT.BRICK20": (GF_BRICK20, 20, 3),
    "%ELEMENT.BRICK8": (GF_BRICK8, 8, 3),
    "%ELEMENT.Q4": (GF_Q4, 4, 2),
    "%ELEMENT.Q8": (GF_Q8, 8, 2),
}
# element node i <- file column BRICK20_FROM_FILE[i] (fileIO.cc:184-204)
BRICK20_FROM_FILE = (0, 7, 6, 11, 18, 19, 12, 8, 1, 5, 17, 13, 2, 3, 4, 10, 16, 15, 14, 9)


class NfError(ValueError):
    pass


class MeshShape:
    """Sizes of an unstructured mesh (the role Grid plays for structured ones)."""

    def __init__(self, elem_type, dim, nen, nnodes, nelem):
        self.elem_type, self.dim, self.nen = elem_type, dim, nen
        self.nnodes, self.nelem = nnodes, nelem
        self.ndof = nnodes * dim


class _Lines:
    """Sequential line reader with the reference's one-time rewind at %END."""

    def __init__(self, text):
        self.lines = text.splitlines()
        self.pos = 0

    def next(self):
        if self.pos >= len(self.lines):
            raise NfError("unexpected end of file")
        s = self.lines[self.pos]
        self.pos += 1
        return s

    def find(self, key):
        rewound = False
        while self.pos < len(self.lines):
            s = self.lines[self.pos]
            self.pos += 1
            if key in s:
                return s
            if "%END" in s:
                if rewound:
                    break
                rewound = True
                self.pos = 0
        raise NfError("block %s not found" % key)


def _ints(s, n):
    v = s.split()
    if len(v) < n:
        raise NfError("short line: %r" % s)
    return [int(x) for x in v[:n]]


def _floats(s, n):
    v = s.split()
    if len(v) < n:
        raise NfError("short line: %r" % s)
    return [float(x) for x in v[:n]]


def parse_nf(text, E=E_DEFAULT, nu=NU_DEFAULT, ngauss=None, name="nf"):
    L = _Lines(text)
    key = L.find("%ELEMENT.").strip().split()[0]
    if key not in ELEMENTS:
        raise NfError("element type not supported: %s" % key)
    et, nen, dim = ELEMENTS[key]
    nelem = _ints(L.next(), 1)[0]
    skip = 3 if dim == 3 else 4
    conn = np.empty((nelem, nen), dtype=np.int32)
    for e in range(nelem):
        v = _ints(L.next(), skip + nen)[skip:]
        conn[e] = [v[i] for i in BRICK20_FROM_FILE] if et == GF_BRICK20 else v
    L.find("%NODE.COORD")
    nnodes = _ints(L.next(), 1)[0]
    coords = np.empty((nnodes, dim))
    for i in range(nnodes):
        coords[i] = _floats(L.next(), 1 + dim)[1:]
    L.find("%NODE.SUPPORT")
    nsup = _ints(L.next(), 1)[0]
    supports = np.array([_ints(L.next(), dim + 1) for _ in range(nsup)],
                        dtype=np.int32).reshape(nsup, dim + 1)
    L.find("%LOAD.CASE.NODAL.FORCE")
    nload = _ints(L.next(), 1)[0]
    loads = np.array([_floats(L.next(), dim + 1) for _ in range(nload)]).reshape(nload, dim + 1)
    if conn.size and (conn.min() < 1 or conn.max() > nnodes):
        raise NfError("connectivity refers to a node outside 1..%d" % nnodes)
    # force vector: assignment per load row, nonzero components only, so a
    # later row overwrites an earlier one's nonzero components
    # (src/FEM/FemData.cc:370-385)
    f = np.zeros(nnodes * dim)
    for row in loads:
        node = int(row[0])
        if not 1 <= node <= nnodes:
            raise NfError("load on a node outside 1..%d" % nnodes)
        for c in range(dim):
            if row[1 + c] != 0.0:
                f[dim * (node - 1) + c] = row[1 + c]
    if ngauss is None:
        ngauss = 3 if et in (GF_Q8, GF_BRICK20) else 2
    shape = MeshShape(et, dim, nen, nnodes, nelem)
    return Problem(name, et, ngauss, shape, conn, coords, E, nu, None, supports,
                   _fixed_dofs(supports, dim), loads, f)


def read_nf(path, E=E_DEFAULT, nu=NU_DEFAULT, ngauss=None):
    with open(path) as fh:
        text = fh.read()
    import os
    return parse_nf(text, E, nu, ngauss, name=os.path.splitext(os.path.basename(path))[0])
