"""CPU restatement of the multigrid V-cycle used by libptq2's viscous solver.

TEST INFRASTRUCTURE ONLY (the checker of tests/test_gpu_solver.py).  Follows:
  - the ChebychevRN recurrence of KSPSolve_ChebychevRN_lag_norm (src/ksp_chebyrn.c:278-447)
    with Jacobi (the A11 diagonal of MatGetDiagonal_MFStokes_A11);
  - coordinate injection onto the coarse node lattice (src/dmda_project_coords.c:39-71);
  - trilinear interpolation of the Q2 node lattice (PETSc DMDA Q1 interpolation,
    src/ptatin_driver_linear_ts.c:996-1006) and restriction = its transpose, with the
    corrections and restricted residuals zeroed on Dirichlet dofs;
  - coarse quadrature-point viscosity = mean (arithmetic, geometric or harmonic) or maximum of
    the fine quadrature values of the child elements a coarse point touches;
  - an exact (dense) solve on the coarsest level.
The element operators are the C oracle's (oracle/ptq2_oracle.c).
"""
import numpy as np

from . import oracle as O


def eta_coarsen(mx, my, mz, eta, mode="arithmetic"):
    """coarse (mx/2,...) quadrature values from the fine ones (arithmetic, geometric or
    harmonic mean)"""
    if mode == "geometric":
        return np.exp(eta_coarsen(mx, my, mz, np.log(eta), "arithmetic"))
    if mode == "harmonic":
        return 1.0 / eta_coarsen(mx, my, mz, 1.0 / np.asarray(eta), "arithmetic")
    if mode == "max":
        return _eta_coarsen_max(mx, my, mz, eta)
    e = np.asarray(eta).reshape(mz, my, mx, 3, 3, 3)  # (ez,ey,ex, qc,qb,qa)
    mxc, myc, mzc = mx // 2, my // 2, mz // 2
    out = np.zeros((mzc, myc, mxc, 3, 3, 3))
    child_mean = e.reshape(mzc, 2, myc, 2, mxc, 2, 27).mean(axis=6)  # (cz,c, cy,b, cx,a)
    for qc in range(3):
        cs = [0] if qc == 0 else ([1] if qc == 2 else [0, 1])
        for qb in range(3):
            bs = [0] if qb == 0 else ([1] if qb == 2 else [0, 1])
            for qa in range(3):
                as_ = [0] if qa == 0 else ([1] if qa == 2 else [0, 1])
                acc = 0.0
                for c in cs:
                    for b in bs:
                        for a in as_:
                            acc = acc + child_mean[:, c, :, b, :, a]
                out[:, :, :, qc, qb, qa] = acc / (len(cs) * len(bs) * len(as_))
    return out.reshape(-1)


def _eta_coarsen_max(mx, my, mz, eta):
    e = np.asarray(eta).reshape(mz, my, mx, 3, 3, 3)
    mxc, myc, mzc = mx // 2, my // 2, mz // 2
    child_max = e.reshape(mzc, 2, myc, 2, mxc, 2, 27).max(axis=6)
    out = np.zeros((mzc, myc, mxc, 3, 3, 3))
    for qc in range(3):
        cs = [0] if qc == 0 else ([1] if qc == 2 else [0, 1])
        for qb in range(3):
            bs = [0] if qb == 0 else ([1] if qb == 2 else [0, 1])
            for qa in range(3):
                as_ = [0] if qa == 0 else ([1] if qa == 2 else [0, 1])
                out[:, :, :, qc, qb, qa] = np.max([child_max[:, c, :, b, :, a] for c in cs for b in bs for a in as_],
                                                  axis=0)
    return out.reshape(-1)


# ---------------------------------------------------------------------------------------
# The reference's coefficient hierarchy (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_-
# MPntPStokes_Hierarchy, src/material_point_utils.c:1325-1524, default projection type 0),
# ptq2_mg_opts.coarse_eta_mean = PTQ2_ETA_Q1_PROJECTION:
#  1. fine nodal field on the Q2 node lattice (_BuildQ1CoefficientProjection_QuadraturePoints_-
#     MPntPStokes_FineGrid, :1196-1300): per element and quadrature point, the 8 corner nodes
#     get Q1 weights N_i(xi_p), every other node of the element weight 1; node value =
#     sum(weight * eta_p) / sum(weight) over all elements;
#  2. each coarser lattice: the scalar trilinear restriction scaled by 1 / (R 1)
#     (MatRestrict + DMCreateInterpolationScale, :1450-1462);
#  3. coarse quadrature values: Q1 interpolation from the element's corner nodes, clipped to
#     the range of its 27 nodal values (_..._InterpolateToQuadratePoints, :957-1070).
# The fine level keeps its own quadrature values.
# ---------------------------------------------------------------------------------------
_G3 = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])


def _q1_at_gauss():
    """N[p, c] = Q1 shape function of corner c = cx + 2cy + 4cz at Gauss point p = i + 3j + 9k"""
    N = np.zeros((27, 8))
    for k in range(3):
        for j in range(3):
            for i in range(3):
                xi = (_G3[i], _G3[j], _G3[k])
                for c in range(8):
                    s = [1 if (c >> a) & 1 else -1 for a in range(3)]
                    N[i + 3 * j + 9 * k, c] = 0.125 * (1 + s[0] * xi[0]) * (1 + s[1] * xi[1]) * (1 + s[2] * xi[2])
    return N


def q1_nodal_projection(mx, my, mz, eta):
    e = np.asarray(eta, dtype=np.float64).reshape(mz, my, mx, 27)
    N = _q1_at_gauss()
    A = np.zeros((2 * mz + 1, 2 * my + 1, 2 * mx + 1))
    B = np.zeros_like(A)
    corner = {0: 0, 2: 1, 6: 2, 8: 3, 18: 4, 20: 5, 24: 6, 26: 7}
    for n in range(27):
        i, j, k = n % 3, (n // 3) % 3, n // 9
        if n in corner:
            w = N[:, corner[n]]
            a = e @ w
            b = w.sum()
        else:
            a = e.sum(axis=3)
            b = 27.0
        A[k:k + 2 * mz:2, j:j + 2 * my:2, i:i + 2 * mx:2] += a
        B[k:k + 2 * mz:2, j:j + 2 * my:2, i:i + 2 * mx:2] += b
    return A / B


def _restrict_scaled_1d(a, axis):
    a = np.moveaxis(a, axis, 0)
    nc = (a.shape[0] - 1) // 2 + 1
    num = np.zeros((nc,) + a.shape[1:])
    den = np.zeros(nc)
    for f in range(a.shape[0]):
        if f % 2 == 0:
            num[f // 2] += a[f]
            den[f // 2] += 1.0
        else:
            for c in ((f - 1) // 2, (f + 1) // 2):
                num[c] += 0.5 * a[f]
                den[c] += 0.5
    return np.moveaxis(num / den.reshape((-1,) + (1,) * (a.ndim - 1)), 0, axis)


def q1_restrict_nodal(nodal):
    """coarse lattice value = sum_f P_fc v_f / sum_f P_fc (P trilinear, separable)"""
    v = nodal
    for ax in range(3):
        v = _restrict_scaled_1d(v, ax)
    return v


def q1_interpolate_qp(mx, my, mz, nodal):
    N = _q1_at_gauss()
    out = np.zeros((mz, my, mx, 27))
    cor = np.stack([nodal[2 * cz:2 * cz + 2 * mz:2, 2 * cy:2 * cy + 2 * my:2, 2 * cx:2 * cx + 2 * mx:2]
                    for cz in range(2) for cy in range(2) for cx in range(2)], axis=-1)  # (ez, ey, ex, 8)
    out = cor @ N.T
    lo = np.full((mz, my, mx), np.inf)
    hi = np.full((mz, my, mx), -np.inf)
    for k in range(3):
        for j in range(3):
            for i in range(3):
                v = nodal[k:k + 2 * mz:2, j:j + 2 * my:2, i:i + 2 * mx:2]
                lo = np.minimum(lo, v)
                hi = np.maximum(hi, v)
    return np.clip(out, lo[..., None], hi[..., None]).reshape(-1)


def q1_projection_hierarchy(mx, my, mz, eta, nlevels):
    """coarse quadrature viscosities of levels 1..nlevels-1"""
    nodal = q1_nodal_projection(mx, my, mz, eta)
    out = []
    for _ in range(nlevels - 1):
        nodal = q1_restrict_nodal(nodal)
        mx, my, mz = mx // 2, my // 2, mz // 2
        out.append(q1_interpolate_qp(mx, my, mz, nodal))
    return out


def prolong(mxc, myc, mzc, xc):
    """trilinear interpolation of a coarse nodal 3-vector field to the fine lattice"""
    c = np.asarray(xc).reshape(2 * mzc + 1, 2 * myc + 1, 2 * mxc + 1, 3)

    def up(a, axis):
        n = a.shape[axis]
        shape = list(a.shape)
        shape[axis] = 2 * n - 1
        out = np.zeros(shape)
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(0, None, 2)
        out[tuple(sl)] = a
        lo = [slice(None)] * a.ndim
        hi = [slice(None)] * a.ndim
        mid = [slice(None)] * a.ndim
        lo[axis] = slice(0, n - 1)
        hi[axis] = slice(1, n)
        mid[axis] = slice(1, None, 2)
        out[tuple(mid)] = 0.5 * (a[tuple(lo)] + a[tuple(hi)])
        return out

    f = up(up(up(c, 2), 1), 0)
    return f.reshape(-1)


def restrict(mxc, myc, mzc, xf):
    """transpose of prolong"""
    nc = 3 * (2 * mxc + 1) * (2 * myc + 1) * (2 * mzc + 1)
    # explicit transpose through the linear map on unit vectors would be slow; use the adjoint stencil
    f = np.asarray(xf).reshape(4 * mzc + 1, 4 * myc + 1, 4 * mxc + 1, 3)

    def down(a, axis):
        n = a.shape[axis]
        nc_ = (n + 1) // 2
        shape = list(a.shape)
        shape[axis] = nc_
        out = np.zeros(shape)
        ev = [slice(None)] * a.ndim
        ev[axis] = slice(0, None, 2)
        out += a[tuple(ev)]
        od = [slice(None)] * a.ndim
        od[axis] = slice(1, None, 2)
        odd = 0.5 * a[tuple(od)]
        lo = [slice(None)] * a.ndim
        hi = [slice(None)] * a.ndim
        lo[axis] = slice(0, nc_ - 1)
        hi[axis] = slice(1, nc_)
        out[tuple(lo)] += odd
        out[tuple(hi)] += odd
        return out

    c = down(down(down(f, 0), 1), 2)
    assert c.size == nc
    return c.reshape(-1)


# 1D quadratic prolongation of the Q2 node lattice: fine node 4E+t of coarse element E sits at
# xi = -1 + t/2 and takes the coarse Q2 basis values there.
_Q2W = {0: ((0, 1.0),), 1: ((0, 0.375), (1, 0.75), (2, -0.125)), 2: ((1, 1.0),),
        3: ((0, -0.125), (1, 0.75), (2, 0.375))}


def _p1d_q2(nc):
    """dense 1D quadratic prolongation, (2*nc-1) x nc, nc = 2*mc+1 coarse lattice nodes"""
    nf = 2 * nc - 1
    P = np.zeros((nf, nc))
    for f in range(nf):
        if f == nf - 1:
            P[f, nc - 1] = 1.0
            continue
        E, t = divmod(f, 4)
        for off, w in _Q2W[t]:
            P[f, 2 * E + off] += w
    return P


def prolong_q2(mxc, myc, mzc, xc):
    c = np.asarray(xc).reshape(2 * mzc + 1, 2 * myc + 1, 2 * mxc + 1, 3)
    Px, Py, Pz = _p1d_q2(2 * mxc + 1), _p1d_q2(2 * myc + 1), _p1d_q2(2 * mzc + 1)
    f = np.einsum("ai,bj,ck,kjid->cbad", Px, Py, Pz, c, optimize=True)
    return f.reshape(-1)


def restrict_q2(mxc, myc, mzc, xf):
    f = np.asarray(xf).reshape(4 * mzc + 1, 4 * myc + 1, 4 * mxc + 1, 3)
    Px, Py, Pz = _p1d_q2(2 * mxc + 1), _p1d_q2(2 * myc + 1), _p1d_q2(2 * mzc + 1)
    c = np.einsum("ai,bj,ck,cbad->kjid", Px, Py, Pz, f, optimize=True)
    return c.reshape(-1)


class Level:
    def __init__(self, mx, my, mz, coords, mask, eta):
        self.mx, self.my, self.mz = mx, my, mz
        self.coords, self.mask, self.eta = coords, mask, eta
        self.n = coords.size

    def matmult(self, x):
        return O.matmult_a11(self.mx, self.my, self.mz, self.coords, self.eta, self.mask, x)

    def diag(self):
        return O.matgetdiag_a11(self.mx, self.my, self.mz, self.coords, self.eta, self.mask)


class GalerkinLevel:
    """Level C with the Galerkin operator Pi R A_F P Pi + (I - Pi) of the level F above it
    (Pi zeroes C's Dirichlet dofs; A_F with MatMult semantics), assembled column by column
    into a sparse matrix: the restatement of ptq2_mg_opts.galerkin_levels (the reference's
    Galerkin coarse levels, src/ptatin_driver_linear_ts.c:64-385, with PETSc's MatPtAP)."""

    def __init__(self, F, C, P=None, R=None):
        import scipy.sparse as sp

        P = P or prolong
        R = R or restrict
        self.mx, self.my, self.mz = C.mx, C.my, C.mz
        self.coords, self.mask, self.eta = C.coords, C.mask, C.eta
        self.n = n = C.n
        dmask = np.zeros(n, dtype=bool) if C.mask is None else C.mask > 0
        rows, cols, vals = [], [], []
        for j in range(n):
            if dmask[j]:
                rows.append(np.array([j]))
                cols.append(np.array([j]))
                vals.append(np.array([1.0]))
                continue
            e = np.zeros(n)
            e[j] = 1.0
            v = P(C.mx, C.my, C.mz, e)
            if F.mask is not None:
                v[F.mask > 0] = 0.0
            col = R(C.mx, C.my, C.mz, F.matmult(v))
            col[dmask] = 0.0
            nz = np.nonzero(col)[0]
            rows.append(nz)
            cols.append(np.full(nz.size, j))
            vals.append(col[nz])
        self.A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))

    def matmult(self, x):
        return self.A @ x

    def diag(self):
        return self.A.diagonal().copy()


def hierarchy(mx, my, mz, coords, mask, eta, nlevels, mode="arithmetic", galerkin_levels=0):
    """Levels 0 .. nlevels-1; with galerkin_levels = k, levels nlevels-1-k .. nlevels-2 carry the
    Galerkin operator of the level above them (GalerkinLevel), built from the top down."""
    lv = [Level(mx, my, mz, coords, mask, eta)]
    proj = q1_projection_hierarchy(mx, my, mz, eta, nlevels) if mode == "q1proj" else None
    for l in range(nlevels - 1):
        f = lv[-1]
        mask = None if f.mask is None else inject_from_fine(f, f.mask)
        ec = proj[l] if proj is not None else eta_coarsen(f.mx, f.my, f.mz, f.eta, mode)
        c = Level(f.mx // 2, f.my // 2, f.mz // 2, inject_from_fine(f), mask, ec)
        if l + 1 >= nlevels - 1 - galerkin_levels and l + 1 < nlevels - 1:
            c = GalerkinLevel(f, c)
        lv.append(c)
    return lv


def inject_from_fine(f, vals=None):
    v = f.coords if vals is None else vals
    g = np.asarray(v).reshape(2 * f.mz + 1, 2 * f.my + 1, 2 * f.mx + 1, 3)
    return np.ascontiguousarray(g[::2, ::2, ::2]).reshape(-1)


def chebyshev(L, b, x, dinv, emin, emax, its, zero_guess):
    """KSPSolve_ChebychevRN_lag_norm with KSP_NORM_NONE (src/ksp_chebyrn.c:278-447)."""
    scale = 2.0 / (emax + emin)
    alpha = 1.0 - scale * emin
    mu = 1.0 / alpha
    omegaprod = 2.0 / alpha
    p = [x if not zero_guess else np.zeros_like(b), None, None]
    c = [1.0, mu, 0.0]
    km1, k, kp1 = 0, 1, 2
    r = b.copy() if zero_guess else b - L.matmult(x)
    p[k] = scale * dinv * r + p[km1]
    for _ in range(its):
        c[kp1] = 2.0 * mu * c[k] - c[km1]
        omega = omegaprod * c[k] / c[kp1]
        r = b - L.matmult(p[k])
        p[kp1] = omega * scale * dinv * r + (1.0 - omega) * p[km1] + omega * p[k]
        km1, k, kp1 = k, kp1, km1
    return p[k]


def dense_inverse(L):
    n = L.n
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[:] = 0.0
        e[j] = 1.0
        A[:, j] = L.matmult(e)
    A = 0.5 * (A + A.T)
    return np.linalg.inv(A)


def galerkin_inverse(F, C, P=None, R=None):
    """Inverse of the Galerkin coarse operator Pi R A_F P Pi + (I - Pi) on level C (Pi zeroes
    C's Dirichlet dofs; A_F with MatMult semantics), assembled column by column."""
    P = P or prolong
    R = R or restrict
    n = C.n
    A = np.zeros((n, n))
    dmask = np.zeros(n, dtype=bool) if C.mask is None else C.mask > 0
    for j in range(n):
        if dmask[j]:
            A[j, j] = 1.0
            continue
        e = np.zeros(n)
        e[j] = 1.0
        v = P(C.mx, C.my, C.mz, e)
        if F.mask is not None:
            v[F.mask > 0] = 0.0
        col = R(C.mx, C.my, C.mz, F.matmult(v))
        col[dmask] = 0.0
        A[:, j] = col
    A = 0.5 * (A + A.T)
    return np.linalg.inv(A)


class VCycle:
    """V-cycle of the GPU solver's hierarchy; galerkin=True makes the coarsest operator the
    Galerkin product through the level above it (ptq2_mg_opts.galerkin_coarse)."""

    def __init__(self, levels, emin, emax, its, transfer="trilinear", galerkin=False):
        self.lv = levels
        self.P = prolong_q2 if transfer == "q2" else prolong
        self.R = restrict_q2 if transfer == "q2" else restrict
        self.emin, self.emax, self.its = emin, emax, its
        self.dinv = [1.0 / L.diag() for L in levels[:-1]]
        if galerkin and len(levels) >= 2:
            self.ainv = galerkin_inverse(levels[-2], levels[-1], self.P, self.R)
        else:
            self.ainv = dense_inverse(levels[-1])

    def __call__(self, b, l=0):
        L = self.lv[l]
        if l == len(self.lv) - 1:
            return self.ainv @ b
        x = chebyshev(L, b, None, self.dinv[l], self.emin[l], self.emax[l], self.its, True)
        r = b - L.matmult(x)
        C = self.lv[l + 1]
        bc = self.R(C.mx, C.my, C.mz, r)
        if C.mask is not None:
            bc[C.mask > 0] = 0.0
        xc = self(bc, l + 1)
        corr = self.P(C.mx, C.my, C.mz, xc)
        if L.mask is not None:
            corr[L.mask > 0] = 0.0
        x = x + corr
        return chebyshev(L, b, x, self.dinv[l], self.emin[l], self.emax[l], self.its, False)


# ---------------------------------------------------------------------------------------
# Rigid-body deflation of stiff inclusions (ptq2_mg_opts.rigid_deflation).  No counterpart
# in the reference, whose velocity solve is PCMG + FGMRES only; this restates the library's
# two-level correction so that its GPU form can be checked.
#
# Stiff elements: those with a quadrature viscosity above sqrt(min * max) of all quadrature
# values, when max >= min_contrast * min (a rigid motion of an inclusion must cover every
# stiff quadrature point, so an element counts as soon as one of its points does).  Inclusions: their
# connected sets (elements sharing at least one node), numbered by their first element; with
# more than max_incl the ones with the most elements are kept (ties: the earlier one).  Each
# kept inclusion contributes 6 columns to W, on the nodes of its elements: the translations
# e_a and the rotations e_a x (x - c) / L (c the mean of the nodes' coordinates, L their
# RMS distance to it), Dirichlet dofs zeroed.  With E = W^T A W, the correction of a
# preconditioner M^-1 is (Tang et al.'s A-DEF2)
#   z = W E^-1 W^T r + M^-1 (r - A W E^-1 W^T r),
# E^-1 over the columns kept by a Cholesky factorisation that skips a column whose pivot is
# <= rel * E_jj (a fully constrained or dependent motion).
# ---------------------------------------------------------------------------------------
def stiff_inclusions(mx, my, mz, eta, min_contrast=100.0, max_incl=128):
    """element labels (-1 outside) and the number of inclusions, by a breadth-first walk over
    the 26 neighbours of each element."""
    e = np.asarray(eta, dtype=np.float64).reshape(-1, 27)
    ne = mx * my * mz
    lab = np.full(ne, -1, dtype=np.int64)
    lo, hi = e.min(), e.max()
    if not hi >= min_contrast * lo:
        return lab.astype(np.int32), 0
    stiff = e.max(axis=1) > np.sqrt(lo * hi)
    first, size = [], []
    for e0 in range(ne):
        if not stiff[e0] or lab[e0] >= 0:
            continue
        c = len(first)
        first.append(e0)
        lab[e0] = c
        stack, cnt = [e0], 0
        while stack:
            e = stack.pop()
            cnt += 1
            ex, ey, ez = e % mx, (e // mx) % my, e // (mx * my)
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        x, y, z = ex + dx, ey + dy, ez + dz
                        if 0 <= x < mx and 0 <= y < my and 0 <= z < mz:
                            f = x + mx * (y + my * z)
                            if stiff[f] and lab[f] < 0:
                                lab[f] = c
                                stack.append(f)
        size.append(cnt)
    nc = len(first)
    if nc > max_incl:
        keep = sorted(sorted(range(nc), key=lambda c: (-size[c], first[c]))[:max_incl])
        new = np.full(nc, -1, dtype=np.int64)
        new[keep] = np.arange(len(keep))
        lab = np.where(lab >= 0, new[np.maximum(lab, 0)], -1)
        nc = max_incl
    return lab.astype(np.int32), nc


def inclusion_node_labels(mx, my, mz, elem_label):
    el = O.elnidx(mx, my, mz).reshape(-1, 27)
    nl = np.full(O.nnodes(mx, my, mz), -1, dtype=np.int32)
    st = elem_label >= 0
    nl[el[st].reshape(-1)] = np.repeat(elem_label[st], 27)
    return nl


def rigid_basis(coords, mask, node_label, nc):
    """W (3N x 6 nc): per inclusion the translations, then the rotations about x, y, z."""
    X = np.asarray(coords).reshape(-1, 3)
    W = np.zeros((X.size, 6 * nc))
    for c in range(nc):
        idx = np.nonzero(node_label == c)[0]
        d = X[idx] - X[idx].mean(axis=0)
        d = d / np.sqrt((d * d).sum() / idx.size)
        for a in range(3):
            W[3 * idx + a, 6 * c + a] = 1.0
            e = np.zeros(3)
            e[a] = 1.0
            W[(3 * idx[:, None] + np.arange(3)).reshape(-1), 6 * c + 3 + a] = np.cross(e, d).reshape(-1)
    if mask is not None:
        W[np.asarray(mask) > 0, :] = 0.0
    return W


def drop_cholesky_keep(E, rel=1e-12):
    """columns kept by a left-looking Cholesky factorisation of E that skips a column whose
    pivot is not above rel * E_jj"""
    n = E.shape[0]
    keep, L = [], np.zeros((n, n))
    for j in range(n):
        if not E[j, j] > 0.0:
            continue
        k = len(keep)
        l = np.zeros(k)
        for i in range(k):
            l[i] = (E[keep[i], j] - L[i, :i] @ l[:i]) / L[i, i]
        piv = E[j, j] - l @ l
        if not piv > rel * E[j, j]:
            continue
        L[k, :k] = l
        L[k, k] = np.sqrt(piv)
        keep.append(j)
    return keep


class RigidDeflation:
    """A-DEF2 correction around a preconditioner, for level L (MatMult semantics)."""

    def __init__(self, L, min_contrast=100.0, max_incl=128, rel=1e-12):
        self.elem_label, self.nc = stiff_inclusions(L.mx, L.my, L.mz, L.eta, min_contrast, max_incl)
        self.node_label = inclusion_node_labels(L.mx, L.my, L.mz, self.elem_label)
        self.W = rigid_basis(L.coords, L.mask, self.node_label, self.nc)
        self.AW = np.stack([L.matmult(self.W[:, j]) for j in range(self.W.shape[1])], axis=1) \
            if self.nc else np.zeros((L.n, 0))
        E = self.W.T @ self.AW
        E = 0.5 * (E + E.T)
        self.keep = drop_cholesky_keep(E, rel)
        self.Einv = np.zeros_like(E)
        if self.keep:
            k = np.array(self.keep)
            self.Einv[np.ix_(k, k)] = np.linalg.inv(E[np.ix_(k, k)])

    def __call__(self, M, r):
        if self.nc == 0:
            return M(r)
        y = self.Einv @ (self.W.T @ r)
        return self.W @ y + M(r - self.AW @ y)
