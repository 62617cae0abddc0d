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


def by_name(name):
    """The benchmark workload of that name (q4_plate_NXxNY, q8_plate_NXxNY,
    brick8_cube_N, brick20_cube_N)."""
    for prefix, make in (("q4_plate_", config_q4_plate), ("q8_plate_", config_q8_plate)):
        if name.startswith(prefix):
            nx, ny = map(int, name[len(prefix):].split("x"))
            return make(nx, ny)
    if name.startswith("brick8_cube_"):
        return config_brick8_cube(int(name[len("brick8_cube_"):]))
    if name.startswith("brick20_cube_"):
        return config_brick20_cube(int(name[len("brick20_cube_"):]))
    raise ValueError("unknown workload %s" % name)


def q4_batch_sizes(count=256, lo=16, hi=256, seed=SEED_Q4_BATCH):
    nx = rng_uniform_int(seed, count, lo, hi)
    return [(int(a), int(a) // 2) for a in nx]


# ---- learned warm-start family: Brick8 cantilevers ------------------------------
SEED_CANTILEVER = 0x5EED0006
CANTILEVER_LENGTHSCALE = 8.0
CANTILEVER_SIGMA = 0.5
CANTILEVER_TRACTION_MEAN = (0.0, 0.0, -1.0)
CANTILEVER_TRACTION_STD = 0.5


def cantilever_seeds(k):
    """Seeds of member k: (E field, traction x, traction y, traction z)."""
    base = SEED_CANTILEVER + 4 * int(k)
    return base, base + 1, base + 2, base + 3


def cantilever_fields(k, nx=64, ny=32, nz=32, lengthscale=CANTILEVER_LENGTHSCALE,
                      sigma=CANTILEVER_SIGMA):
    """Element moduli and top-face tractions of member k.

    E_e = 210000 exp(sigma g_e - sigma^2 / 2), g a unit-variance Gaussian random
    field over the elements (mean of E is 210000); traction component c at
    top-face node (i, j) is mu_c + s g_c(i, j), g_c a unit-variance field over
    the (nx+1) x (ny+1) face nodes.  Returns (elem_E [nel], traction [ny+1, nx+1, 3]).
    """
    from . import grf
    sE, sx, sy, sz = cantilever_seeds(k)
    g = grf(sE, (nx, ny, nz), lengthscale).reshape(-1)
    elem_E = E_DEFAULT * np.exp(sigma * g - 0.5 * sigma * sigma)
    t = np.empty((ny + 1, nx + 1, 3))
    for c, s in enumerate((sx, sy, sz)):
        t[:, :, c] = (CANTILEVER_TRACTION_MEAN[c] +
                      CANTILEVER_TRACTION_STD * grf(s, (nx + 1, ny + 1), lengthscale))
    return elem_E, t


def cantilever_problem(k, nx=64, ny=32, nz=32, template=None):
    """Member k of the Brick8 cantilever family (2x2x2 Gauss, nu = 0.3).

    The x = 0 face is clamped; the top face (z = nz) carries the random
    traction of cantilever_fields lumped to its nodes (nodal force = traction
    times the node's tributary area: 1/4 per adjacent unit face).  `template`
    (a member of the same size) lends its mesh arrays.
    """
    elem_E, t = cantilever_fields(k, nx, ny, nz)
    if template is None:
        g = Grid(GF_BRICK8, nx, ny, nz)
        conn, coords = g.fill()
        sup = _supports_from_nodes(g.face_nodes(0), 3)
        fixed = _fixed_dofs(sup, 3)
    else:
        g, conn, coords = template.grid, template.conn, template.coords
        sup, fixed = template.supports, template.fixed_dofs
    area = np.ones((ny + 1, nx + 1))
    area[0, :] *= 0.5
    area[-1, :] *= 0.5
    area[:, 0] *= 0.5
    area[:, -1] *= 0.5
    F = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    F[nz] = t * area[:, :, None]
    f = F.reshape(-1)
    return Problem("brick8_cantilever_%dx%dx%d_k%d" % (nx, ny, nz, k), GF_BRICK8, 2, g, conn,
                   coords, E_DEFAULT, NU_DEFAULT, elem_E, sup, fixed, _loads_from_f(f, 3), f)
