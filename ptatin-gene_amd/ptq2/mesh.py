"""Synthetic Q2 meshes, boundary conditions, viscosity fields and test vectors.

Host-side set-up code (numpy).  Numbering follows the structured DMDA of
pTatin3D: node n = i + NX*(j + NY*k) with NX = 2*mx+1, interlaced dofs 3n+l,
element e = ex + mx*(ey + my*ez), quadrature point q = a + 3b + 9c
(src/dmda_element_q2p1.c:368-469, src/element_utils_q2.c:276-333).

Random fields use SplitMix64 streams with an explicit seed per array, so every
input of a test or benchmark is reproducible bit for bit.
"""
import numpy as np

GAUSS_X = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])
GAUSS_W = np.array([5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0])


def basis_1d():
    """B[q][n] = l_n(x_q), D[q][n] = l_n'(x_q) on nodes {-1,0,1} (src/stokes_operators_tensor.c:193-206)."""
    x = GAUSS_X
    B = np.stack([0.5 * (x * x - x), 1.0 - x * x, 0.5 * (x * x + x)], axis=1)
    D = np.stack([x - 0.5, -2.0 * x, x + 0.5], axis=1)
    return B, D


def nnodes(mx, my, mz):
    return (2 * mx + 1) * (2 * my + 1) * (2 * mz + 1)


# ---------------------------------------------------------------- random

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """The n outputs of SplitMix64 seeded with `seed`, starting at stream position `offset`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(offset + 1, offset + n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, lo=0.0, hi=1.0):
    """n doubles uniform in [lo, hi) from SplitMix64(seed) (53-bit mantissas)."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


# ------------------------------------------------------------------ mesh

def q2_uniform(mx, my, mz, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    """DMDASetUniformCoordinates on the Q2 node lattice (src/ptatin3d_stokes.c:153)."""
    x = np.linspace(box[0], box[1], 2 * mx + 1)
    y = np.linspace(box[2], box[3], 2 * my + 1)
    z = np.linspace(box[4], box[5], 2 * mz + 1)
    nx, ny, nz = x.size, y.size, z.size
    c = np.empty((nz, ny, nx, 3))
    # PETSc computes x0 + i*h; linspace's last point is exact, interior points equal x0 + i*h
    c[..., 0] = (box[0] + np.arange(nx) * ((box[1] - box[0]) / (nx - 1)))[None, None, :]
    c[..., 1] = (box[2] + np.arange(ny) * ((box[3] - box[2]) / (ny - 1)))[None, :, None]
    c[..., 2] = (box[4] + np.arange(nz) * ((box[5] - box[4]) / (nz - 1)))[:, None, None]
    return c.reshape(-1)


def bump_ymax(mx, my, mz, coords, amp=-0.3, lam=-5.6):
    """MeshDeformation_GaussianBump_YMAX (src/mesh_deformation.c:48-125)."""
    nx, ny, nz = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    c = np.asarray(coords, dtype=np.float64).reshape(nz, ny, nx, 3)
    gmin = c.reshape(-1, 3).min(axis=0)
    gmax = c.reshape(-1, 3).max(axis=0)
    u = q2_uniform(mx, my, mz, (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)).reshape(nz, ny, nx, 3)
    xn, zn = u[..., 0], u[..., 2]
    yh = amp * np.exp(lam * (xn * xn + zn * zn)) + 1.0
    dy = (yh + 1.0) / float(ny - 1)
    j = np.arange(ny)[None, :, None]
    u[..., 1] = -1.0 + j * dy
    out = (gmax - gmin) * (u + 1.0) / 2.0 + gmin
    return out.reshape(-1)


def affine(coords, A, b):
    """x' = A x + b applied to every node."""
    c = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    return (c @ np.asarray(A, dtype=np.float64).T + np.asarray(b, dtype=np.float64)).reshape(-1)


def jitter(mx, my, mz, coords, frac, h, seed):
    """Move every interior node by U[-frac*h, frac*h] per axis (boundary nodes stay put)."""
    nx, ny, nz = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    c = np.array(coords, dtype=np.float64).reshape(nz, ny, nx, 3)
    d = uniform(seed, c.size, -frac * h, frac * h).reshape(c.shape)
    c[1:-1, 1:-1, 1:-1, :] += d[1:-1, 1:-1, 1:-1, :]
    return c.reshape(-1)


def bc_mask(mx, my, mz, kind="freeslip_freesurface"):
    """Dirichlet mask of the viscous_sinker boundary conditions
    (src/models/viscous_sinker/model_viscous_sinker_ops.c:192-270)."""
    nx, ny, nz = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    m = np.zeros((nz, ny, nx, 3), dtype=np.uint8)
    if kind == "freeslip_freesurface":
        m[:, :, 0, 0] = m[:, :, -1, 0] = 1
        m[:, 0, :, 1] = 1
        m[0, :, :, 2] = m[-1, :, :, 2] = 1
    elif kind == "freeslip":
        m[:, :, 0, 0] = m[:, :, -1, 0] = 1
        m[:, 0, :, 1] = m[:, -1, :, 1] = 1
        m[0, :, :, 2] = m[-1, :, :, 2] = 1
    elif kind == "noslip":
        m[:, :, 0, :] = m[:, :, -1, :] = 1
        m[:, 0, :, :] = m[:, -1, :, :] = 1
        m[0, :, :, :] = m[-1, :, :, :] = 1
    elif kind != "none":
        raise ValueError("unknown boundary condition kind %r" % kind)
    return m.reshape(-1)


def qp_coords(mx, my, mz, coords):
    """Physical coordinates of the 27 quadrature points of every element, shape (E, 27, 3)."""
    B, _ = basis_1d()
    nx, ny, nz = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    g = np.asarray(coords, dtype=np.float64).reshape(nz, ny, nx, 3)
    x1 = np.stack([B[q, 0] * g[:, :, 0:-1:2] + B[q, 1] * g[:, :, 1::2] + B[q, 2] * g[:, :, 2::2]
                   for q in range(3)], axis=3)  # (nz, ny, mx, qa, 3)
    x2 = np.stack([B[q, 0] * x1[:, 0:-1:2] + B[q, 1] * x1[:, 1::2] + B[q, 2] * x1[:, 2::2]
                   for q in range(3)], axis=3)  # (nz, my, mx, qb, qa, 3)
    del x1
    x3 = np.stack([B[q, 0] * x2[0:-1:2] + B[q, 1] * x2[1::2] + B[q, 2] * x2[2::2]
                   for q in range(3)], axis=3)  # (mz, my, mx, qc, qb, qa, 3)
    return x3.reshape(mx * my * mz, 27, 3)


def _lattice_to_qp(mx, my, mz, g, ops):
    """Apply 1D operators (3x3: quadrature point x element node) along x, y, z of a nodal
    field on the Q2 lattice g (nz, ny, nx, k); result (mx*my*mz, 27, k), q = a + 3b + 9c."""
    Ox, Oy, Oz = ops
    x1 = np.stack([Ox[q, 0] * g[:, :, 0:-1:2] + Ox[q, 1] * g[:, :, 1::2] + Ox[q, 2] * g[:, :, 2::2]
                   for q in range(3)], axis=3)  # (nz, ny, mx, qa, k)
    x2 = np.stack([Oy[q, 0] * x1[:, 0:-1:2] + Oy[q, 1] * x1[:, 1::2] + Oy[q, 2] * x1[:, 2::2]
                   for q in range(3)], axis=3)  # (nz, my, mx, qb, qa, k)
    del x1
    x3 = np.stack([Oz[q, 0] * x2[0:-1:2] + Oz[q, 1] * x2[1::2] + Oz[q, 2] * x2[2::2]
                   for q in range(3)], axis=3)  # (mz, my, mx, qc, qb, qa, k)
    return x3.reshape(mx * my * mz, 27, -1)


def qp_jacobian_det(mx, my, mz, coords):
    """det(dx/dxi) at the 27 quadrature points of every element, shape (E, 27)."""
    B, D = basis_1d()
    g = np.asarray(coords, dtype=np.float64).reshape(2 * mz + 1, 2 * my + 1, 2 * mx + 1, 3)
    jx = _lattice_to_qp(mx, my, mz, g, (D, B, B))
    jy = _lattice_to_qp(mx, my, mz, g, (B, D, B))
    jz = _lattice_to_qp(mx, my, mz, g, (B, B, D))
    return np.einsum("eqi,eqi->eq", jx, np.cross(jy, jz))


def body_force(mx, my, mz, coords, fu):
    """Consistent nodal body force b[3n+d] = sum_q w_q detJ_q N_n(x_q) fu[e,q,d] of every
    element, summed over elements: the Be term of FormFunctionLocal_U
    (src/stokes_form_function.c:219-224).  fu has shape (E, 27, 3)."""
    B, _ = basis_1d()
    W = (GAUSS_W[None, None, :] * GAUSS_W[None, :, None] * GAUSS_W[:, None, None]).reshape(27)
    f = (W[None, :] * qp_jacobian_det(mx, my, mz, coords))[:, :, None] * np.asarray(fu).reshape(-1, 27, 3)
    f = f.reshape(mz, my, mx, 3, 3, 3, 3)  # (ez, ey, ex, qc, qb, qa, d)
    nx, ny, nz = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    # transpose of the interpolation, one direction at a time, scattering into the lattice
    t = np.einsum("zyxcbad,ai->zyxcbid", f, B)
    del f
    ox = np.zeros((mz, my, 3, 3, nx, 3))
    tt = np.moveaxis(t, 2, 5)  # (ez, ey, qc, qb, i, ex, d)
    del t
    for i in range(3):
        ox[:, :, :, :, i:i + 2 * mx:2, :] += tt[:, :, :, :, i]
    del tt
    t = np.einsum("zycbxd,bj->zycjxd", ox, B)  # (ez, ey, qc, j, nx, d)
    del ox
    oy = np.zeros((mz, 3, ny, nx, 3))
    tt = np.moveaxis(t, 1, 3)  # (ez, qc, j, ey, nx, d)
    del t
    for j in range(3):
        oy[:, :, j:j + 2 * my:2, :, :] += tt[:, :, j]
    del tt
    t = np.einsum("zcyxd,ck->zkyxd", oy, B)  # (ez, k, ny, nx, d)
    del oy
    out = np.zeros((nz, ny, nx, 3))
    for k in range(3):
        out[k:k + 2 * mz:2] += t[:, k]
    return out.reshape(-1)


def point_lattice(mx, my, mz, coords, n=(2, 2, 2)):
    """Material points on the zero-perturbation lattice layout: n[0]*n[1]*n[2] per element,
    element-major and x fastest, local coordinates xi_d = -1 + (2/n_d)(p_d + 1/2), physical
    coordinates by Q2 interpolation of the element's nodes
    (SwarmMPntStd_CoordAssignment_LatticeLayout3d, src/material_point_std_utils.c:100-210, then
    MaterialPointCoordinateSetUp, src/ptatin3d.c:520-562).  Returns xi (P, 3), wil (P,), x (P, 3)."""
    E = mx * my * mz
    g = [-1.0 + (2.0 / n[d]) * (np.arange(n[d]) + 0.5) for d in range(3)]
    loc = np.stack(np.meshgrid(g[0], g[1], g[2], indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    npc = loc.shape[0]

    def l3(x):
        return np.stack([0.5 * x * (x - 1.0), (1.0 + x) * (1.0 - x), 0.5 * (1.0 + x) * x], -1)

    N = np.einsum("pk,pj,pi->pkji", l3(loc[:, 2]), l3(loc[:, 1]), l3(loc[:, 0])).reshape(npc, 27)
    nx, ny = 2 * mx + 1, 2 * my + 1
    e = np.arange(E)
    ex, ey, ez = e % mx, (e // mx) % my, e // (mx * my)
    off = np.array([i + nx * (j + ny * k) for k in range(3) for j in range(3) for i in range(3)])
    nodes = (2 * ex + nx * (2 * ey + ny * 2 * ez))[:, None] + off[None, :]
    c = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    x = np.einsum("pn,end->epd", N, c[nodes]).reshape(-1, 3)
    xi = np.tile(loc, (E, 1))
    wil = np.repeat(e, npc).astype(np.int32)
    return xi, wil, x


def sinker_single_inclusion(x, eta0=1.0, eta1=10.0, rho0=0.0, rho1=1.0, origin=(0.5, 0.5, 0.5),
                            length=(0.5, 0.5, 0.5)):
    """Point properties of the viscous sinker's single spherical inclusion
    (ViscousSinker_ApplyInitialMaterialGeometry_SingleInclusion,
    src/models/viscous_sinker/model_viscous_sinker_ops.c:760-838): inside when
    sum_d ((x_d - o_d) / (l_d / 2))^2 < 1.  Returns eta, rho (rho = -rho1 inside, -rho0 outside)."""
    r = (((np.asarray(x) - np.asarray(origin)) / (0.5 * np.asarray(length))) ** 2).sum(axis=1)
    inside = r < 1.0
    return np.where(inside, eta1, eta0).astype(np.float64), np.where(inside, -rho1, -rho0).astype(np.float64)


def sphere_inclusions(n, rmin, rmax, seed, lo=0.0, hi=1.0):
    """n spheres, radius U[rmin,rmax], centre U[lo+r, hi-r]^3 (SplitMix64 stream `seed`)."""
    u = uniform(seed, 4 * n).reshape(n, 4)
    r = rmin + (rmax - rmin) * u[:, 0]
    cen = lo + r[:, None] + (hi - lo - 2 * r[:, None]) * u[:, 1:]
    return cen, r


def viscosity_inclusions(xq, centres, radii, eta_in, eta_out=1.0, noise_seed=None, noise_decades=0.5):
    """Quadrature-point viscosity: eta_in inside any sphere, eta_out elsewhere, optionally
    times 10^U[-noise_decades, noise_decades] per point."""
    E = xq.shape[0]
    eta = np.full((E, 27), float(eta_out))
    inside = np.zeros((E, 27), dtype=bool)
    for c, r in zip(centres, radii):
        d2 = ((xq - c[None, None, :]) ** 2).sum(axis=2)
        inside |= d2 <= r * r
    eta[inside] = float(eta_in)
    if noise_seed is not None:
        eta *= 10.0 ** uniform(noise_seed, eta.size, -noise_decades, noise_decades).reshape(eta.shape)
    return eta.reshape(-1)


def test_vector_u(mx, my, mz, coords):
    """_GenerateTestVector: u_l = X + Y + Z + l + 3 (src/tests/test_stokes_operators.c:48-92)."""
    c = np.asarray(coords).reshape(-1, 3)
    s = c.sum(axis=1)
    return (s[:, None] + np.arange(3)[None, :] + 3.0).reshape(-1)


def test_vector_p(mx, my, mz, legacy=True):
    """_GenerateTestVectorDAP (src/tests/test_stokes_operators.c:94-130).  legacy=True uses the
    spacing 2/(M-1) with which the reference's expected files were produced."""
    d = 1 if legacy else 0
    with np.errstate(divide="ignore", invalid="ignore"):
        hx, hy, hz = np.float64(2.0) / (mx - d), np.float64(2.0) / (my - d), np.float64(2.0) / (mz - d)
        i = np.arange(mx)[None, None, :]
        j = np.arange(my)[None, :, None]
        k = np.arange(mz)[:, None, None]
        f = (-1.0 + i * hx) + (-1.0 + j * hy) + (-1.0 + k * hz)
    return (f[..., None] + np.arange(4) + 3.0).reshape(-1)


class Problem:
    """A synthetic operator problem: mesh, BC mask and quadrature-point viscosity."""

    def __init__(self, mx, my, mz, coords, mask, eta):
        self.mx, self.my, self.mz = mx, my, mz
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        self.eta = np.ascontiguousarray(eta, dtype=np.float64)

    @property
    def nvel(self):
        return 3 * nnodes(self.mx, self.my, self.mz)

    @property
    def nel(self):
        return self.mx * self.my * self.mz


def config_jitter_iso(m, seed_jitter=1, frac=0.1):
    """Isoviscous jittered unit cube: interior nodes jittered by U[-0.1h,0.1h] with h the node
    spacing 1/(2m); eta = 1 at every quadrature point; no boundary conditions."""
    c = jitter(m, m, m, q2_uniform(m, m, m), frac, 1.0 / (2 * m), seed_jitter)
    return Problem(m, m, m, c, None, np.ones(27 * m ** 3))


SHEAR = np.array([[1.0, 0.15, 0.05], [0.0, 1.0, 0.1], [0.0, 0.0, 1.0]])


def config_sinker(m, n_incl=12, contrast=1e3, rmin=0.05, rmax=0.15, seed_jitter=1, seed_incl=3,
                  seed_noise=4, frac=0.1, noise=True, bc="none"):
    """Affine-plus-jitter mesh with spherical inclusions of viscosity `contrast` in a unit matrix
    and (optionally) log-uniform noise 10^U[-0.5,0.5] per quadrature point."""
    c = q2_uniform(m, m, m)
    c = jitter(m, m, m, c, frac, 1.0 / (2 * m), seed_jitter)
    cen, r = sphere_inclusions(n_incl, rmin, rmax, seed_incl)
    xq = qp_coords(m, m, m, c)
    eta = viscosity_inclusions(xq, cen, r, contrast, 1.0, seed_noise if noise else None)
    del xq
    c = affine(c, SHEAR, [0.0, 0.0, 0.0])
    mask = None if bc == "none" else bc_mask(m, m, m, bc)
    return Problem(m, m, m, c, mask, eta)


def _inclusion_field(xq, mx, my, mz, ez0, mzg, centres, radii, v_in, v_out):
    """v_in at quadrature points inside any sphere, v_out elsewhere, for the mz element
    layers starting at layer ez0 of a (jittered) unit-cube mesh of mx x my x mzg elements.
    Each sphere only tests the elements whose lattice cell range can reach it (node
    jitter below half a cell)."""
    out = np.full((mz, my, mx, 27), float(v_out))
    x = xq.reshape(mz, my, mx, 27, 3)
    for cen, r in zip(centres, radii):
        lo = np.floor((cen - r) * np.array([mx, my, mzg])).astype(int) - 1
        hi = np.ceil((cen + r) * np.array([mx, my, mzg])).astype(int) + 1
        x0, x1 = max(lo[0], 0), min(hi[0], mx)
        y0, y1 = max(lo[1], 0), min(hi[1], my)
        z0, z1 = max(lo[2] - ez0, 0), min(hi[2] - ez0, mz)
        if x0 >= x1 or y0 >= y1 or z0 >= z1:
            continue
        sub = x[z0:z1, y0:y1, x0:x1]
        inside = ((sub - cen) ** 2).sum(axis=-1) <= r * r
        out[z0:z1, y0:y1, x0:x1][inside] = float(v_in)
    return out.reshape(-1, 27)


def config_buoyancy(m, n_spheres=20, contrast=1e4, drho=0.1, rmin=0.05, rmax=0.15, seed_jitter=1,
                    seed_spheres=5, frac=0.1, bc="freeslip", threads=None):
    """Viscous solve set-up: unit cube with jittered interior nodes (walls stay planar for the
    free-slip conditions), `n_spheres` seeded spheres of viscosity `contrast` and density
    1 + drho in a matrix of viscosity 1 and density 1, gravity (0,-1,0).  Returns the Problem
    and the consistent buoyancy right-hand side (zero on Dirichlet dofs).  The element
    z-slabs are processed by a thread pool (numpy releases the GIL)."""
    import concurrent.futures
    import os

    c = jitter(m, m, m, q2_uniform(m, m, m), frac, 1.0 / (2 * m), seed_jitter)
    cen, r = sphere_inclusions(n_spheres, rmin, rmax, seed_spheres)
    n1 = 2 * m + 1
    g = c.reshape(n1, n1, n1, 3)
    eta = np.empty((m, m * m, 27))
    slab = max(1, min(8, m // 4))
    chunks = [(z0, min(z0 + slab, m)) for z0 in range(0, m, slab)]

    def work(ch):
        z0, z1 = ch
        mzc = z1 - z0
        cs = np.ascontiguousarray(g[2 * z0:2 * z1 + 1]).reshape(-1)
        xq = qp_coords(m, m, mzc, cs)
        eta[z0:z1] = _inclusion_field(xq, m, m, mzc, z0, m, cen, r, contrast, 1.0).reshape(mzc, m * m, 27)
        rho = _inclusion_field(xq, m, m, mzc, z0, m, cen, r, 1.0 + drho, 1.0)
        del xq
        fu = np.zeros(rho.shape + (3,))
        fu[:, :, 1] = -rho
        return body_force(m, m, mzc, cs, fu).reshape(2 * mzc + 1, n1, n1, 3)

    nt = threads or min(16, os.cpu_count() or 1)
    b = np.zeros((n1, n1, n1, 3))
    with concurrent.futures.ThreadPoolExecutor(max_workers=nt) as ex:
        for (z0, z1), part in zip(chunks, ex.map(work, chunks)):
            b[2 * z0:2 * z1 + 1] += part
    b = b.reshape(-1)
    mask = bc_mask(m, m, m, bc)
    b[mask > 0] = 0.0
    return Problem(m, m, m, c, mask, eta.reshape(-1)), b
