"""Synthetic structured problems for the benchmark workloads and tests.

Each builder returns a Problem with the mesh arrays, the supports and loads in
the reference's neutral-file conventions (supports rows ``[node, fix_x,
fix_y(, fix_z)]`` as read by src/Utils/fileIO.cc:362-404 and applied by
src/FEM/fem.cc:112-145; load rows ``[node, Fx, Fy(, Fz)]`` assigned into the
force vector as src/FEM/FemData.cc:370-385 does), and the force vector.

Workloads (unit elements, nu = 0.3):
  * plate:  Q4 / Q8, left edge (x = 0) clamped, uniform traction 1.0 in -y on
            the right edge as consistent nodal loads (linear edge: 1/2, 1/2;
            quadratic edge: 1/6, 2/3, 1/6 per unit length).
  * cube:   Brick8 / Brick20, z = 0 face clamped, uniform pressure 1.0 in -z on
            the top face (bilinear face: 1/4 each; 8-node face: -1/12 corners,
            1/3 mid-edges).
  * E fields: lognormal per element, E_e = exp(ln 210000 + sigma z_e), z_e
            N(0,1) in element order from a seed.
"""
import dataclasses

import numpy as np

from . import GF_BRICK8, GF_BRICK20, GF_Q4, GF_Q8, Grid, rng_normal, rng_uniform_int

E_DEFAULT = 210000.0
NU_DEFAULT = 0.3


@dataclasses.dataclass
class Problem:
    name: str
    elem_type: int
    ngauss: int
    grid: Grid
    conn: np.ndarray
    coords: np.ndarray
    E: float
    nu: float
    elem_E: object  # np.ndarray or None
    supports: np.ndarray  # int32 [nsup, dim+1]
    fixed_dofs: np.ndarray  # int32, 0-based
    loads: np.ndarray  # float64 [nload, dim+1]
    f: np.ndarray  # force vector [ndof]

    @property
    def dim(self):
        return self.grid.dim

    @property
    def ndof(self):
        return self.grid.ndof


def _supports_from_nodes(nodes, dim):
    sup = np.ones((nodes.size, dim + 1), dtype=np.int32)
    sup[:, 0] = nodes
    return sup


def _fixed_dofs(supports, dim):
    out = []
    for row in supports:
        for j in range(1, dim + 1):
            if row[j] == 1:
                out.append(dim * (int(row[0]) - 1) + (j - 1))
    return np.asarray(out, dtype=np.int32)


def _loads_from_f(f, dim):
    F = f.reshape(-1, dim)
    nodes = np.nonzero(np.any(F != 0.0, axis=1))[0]
    loads = np.zeros((nodes.size, dim + 1), dtype=np.float64)
    loads[:, 0] = nodes + 1
    loads[:, 1:] = F[nodes]
    return loads


def lognormal_E(nelem, seed, sigma=0.5, mean=E_DEFAULT):
    z = rng_normal(seed, nelem)
    return np.exp(np.log(mean) + sigma * z)


def structured(elem_type, nx, ny, nz=0, ngauss=None, E=E_DEFAULT, nu=NU_DEFAULT,
               elem_E=None, name=None, q=1.0):
    """Plate (2D) or cube (3D) with the standard supports and load."""
    g = Grid(elem_type, nx, ny, nz)
    conn, coords = g.fill()
    if ngauss is None:
        ngauss = 3 if elem_type in (GF_Q8, GF_BRICK20) else 2
    if g.dim == 2:
        sup = _supports_from_nodes(g.face_nodes(0), 2)
        f = g.face_load(1, 1, -q)
    else:
        sup = _supports_from_nodes(g.face_nodes(4), 3)
        f = g.face_load(5, 2, -q)
    return Problem(name or "%s_%dx%dx%d" % (elem_type, nx, ny, nz), elem_type, ngauss, g, conn,
                   coords, E, nu, elem_E, sup, _fixed_dofs(sup, g.dim), _loads_from_f(f, g.dim),
                   f)


def random_top_loads(problem, nrhs, seed):
    """nrhs force vectors with N(0,1) components on the top-face nodes."""
    g = problem.grid
    top = g.face_nodes(5 if g.dim == 3 else 3)
    z = rng_normal(seed, nrhs * top.size * g.dim).reshape(nrhs, top.size, g.dim)
    F = np.zeros((nrhs, g.nnodes, g.dim))
    F[:, top - 1, :] = z
    return F.reshape(nrhs, g.ndof)


# ---- the benchmark configurations ----------------------------------------------
SEED_BRICK8_E = 0x5EED0003
SEED_BRICK20_RHS = 0x5EED0004
SEED_Q4_BATCH = 0x5EED0005


def config_q4_plate(nx=64, ny=32):
    return structured(GF_Q4, nx, ny, ngauss=2, name="q4_plate_%dx%d" % (nx, ny))


def config_q8_plate(nx=1024, ny=512):
    return structured(GF_Q8, nx, ny, ngauss=3, name="q8_plate_%dx%d" % (nx, ny))


def config_brick8_cube(n=128, seed=SEED_BRICK8_E, sigma=0.5):
    nel = n * n * n
    return structured(GF_BRICK8, n, n, n, ngauss=2, elem_E=lognormal_E(nel, seed, sigma),
                      name="brick8_cube_%d" % n)


def config_brick20_cube(n=48):
    return structured(GF_BRICK20, n, n, n, ngauss=3, name="brick20_cube_%d" % n)


def q4_batch_sizes(count=256, lo=16, hi=256, seed=SEED_Q4_BATCH):
    nx = rng_uniform_int(seed, count, lo, hi)
    return [(int(a), int(a) // 2) for a in nx]
