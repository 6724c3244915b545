"""Pins of the CPU oracle (oracle/ptq2_oracle.c) -- no GPU.

1. The reference's own goldens (tests/integration/operator/**/expected, extracted into
   tests/golden/operator_goldens.json): A12 and A21 y.y on every mesh the reference tests
   (they do not depend on the viscosity), A11 and diagonal y.y with eta = 1.
2. An independent numpy restatement (dense 81x81 element matrices from per-point shape
   function derivatives, assembled and multiplied) on tiny jittered meshes.
3. Invariants the mathematics fixes: rigid-body null space, the exact energy of a linear
   field, symmetry, A21 = A12^T, the affine patch test, closed-form basis values.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptq2 import mesh as M


def golden_mesh(mx, my, mz):
    # the expected files were produced on the Gaussian-bump mesh of the viscous_sinker model
    return O.mesh_bump(mx, my, mz, O.mesh_uniform(mx, my, mz))


def _rel(a, b):
    return abs(a - b) / abs(b)


def test_goldens_a12_a21(goldens):
    seen = 0
    for g in goldens:
        mx, my, mz = g["mx"], g["my"], g["mz"]
        c = golden_mesh(mx, my, mz)
        mask = O.bc_mask(mx, my, mz, 0)
        x = O.test_vector_u(mx, my, mz, c)
        y21 = O.matmult_a21(mx, my, mz, c, mask, x)
        assert _rel(y21 @ y21, g["yy"]["compare_mf_A21"]) < 5e-9, g["source"]
        p = O.test_vector_p(mx, my, mz, legacy=True)
        with np.errstate(all="ignore"):
            y12 = O.matmult_a12(mx, my, mz, c, mask, p)
        if g["yy"]["compare_mf_A12"] is None:
            # 1x1x3: the legacy spacing 2/(M-1) divides by zero -> nan, as in the expected file
            assert np.isnan(y12 @ y12), g["source"]
        else:
            assert _rel(y12 @ y12, g["yy"]["compare_mf_A12"]) < 5e-9, g["source"]
        seen += 1
    assert seen == 12


def test_goldens_a11_diag_isoviscous(goldens):
    g = [g for g in goldens if g["eta1"] == 1.0][0]
    mx, my, mz = g["mx"], g["my"], g["mz"]
    c = golden_mesh(mx, my, mz)
    mask = O.bc_mask(mx, my, mz, 0)
    eta = np.ones(27 * mx * my * mz)
    x = O.test_vector_u(mx, my, mz, c)
    y = O.matmult_a11(mx, my, mz, c, eta, mask, x)
    assert _rel(y @ y, g["yy"]["compare_mf_A11"]) < 5e-9
    d = O.matgetdiag_a11(mx, my, mz, c, eta, mask)
    assert _rel(d @ d, g["yy"]["compare_mf_diagA11"]) < 5e-9


def _np_project_q1(mx, my, mz, xi, wil, eta_p, vertex_weight="q1", clip=True, interp="q1"):
    """numpy restatement of the Q1 marker projection (src/material_point_utils.c:663-830) with
    switches for plausible misreadings: vertex_weight "q2" weights every node by the Q2 basis
    instead of (Q1 at the vertices, 1 elsewhere), "one" by 1 everywhere (a plain point average);
    clip=False drops the element range clipping (a no-op: a Q1 interpolant stays inside the range
    of its vertex values); interp "q2" interpolates the 27 nodal values with the Q2 basis."""
    NX, NY = 2 * mx + 1, 2 * my + 1
    N = NX * NY * (2 * mz + 1)

    def b1(x, q2):
        if q2:
            return np.stack([0.5 * x * (x - 1), (1 + x) * (1 - x), 0.5 * (1 + x) * x], -1)
        return np.stack([0.5 * (1 - x), np.zeros_like(x), 0.5 * (1 + x)], -1)

    def n27(x, q2):
        bx, by, bz = b1(x[:, 0], q2), b1(x[:, 1], q2), b1(x[:, 2], q2)
        return np.einsum("pk,pj,pi->pkji", bz, by, bx).reshape(-1, 27)

    W = n27(xi, vertex_weight == "q2")
    if vertex_weight == "one":
        W[:] = 1.0
    if vertex_weight == "q1":
        vert = np.zeros(27, bool)
        vert[[0, 2, 6, 8, 18, 20, 24, 26]] = True
        W[:, ~vert] = 1.0
    e = wil.astype(np.int64)
    ex, ey, ez = e % mx, (e // mx) % my, e // (mx * my)
    loc = np.array([i + NX * (j + NY * k) for k in range(3) for j in range(3) for i in range(3)])
    nodes = (2 * ex + NX * (2 * ey + NY * 2 * ez))[:, None] + loc[None, :]
    A = np.bincount(nodes.ravel(), (W * eta_p[:, None]).ravel(), N)
    B = np.bincount(nodes.ravel(), W.ravel(), N)
    A = A / B
    s = 0.774596669241483
    g = np.array([-s, 0.0, s])
    xq = np.stack([np.tile(g, 9), np.tile(np.repeat(g, 3), 3), np.repeat(g, 9)], -1)
    NQ = n27(xq, interp == "q2")
    E = mx * my * mz
    ee = np.arange(E)
    enodes = (2 * (ee % mx) + NX * (2 * ((ee // mx) % my) + NY * 2 * (ee // (mx * my))))[:, None] + loc[None, :]
    a = A[enodes]
    v = a @ NQ.T
    if clip:
        v = np.clip(v, a.min(1, keepdims=True), a.max(1, keepdims=True))
    return v.ravel()


def _sinker_eta(mx, my, mz, c):
    xi, wil, xp = O.marker_lattice(mx, my, mz, c)
    eta_p, _ = O.sinker_point_properties(xp, eta1=10.0)
    return xi, wil, eta_p


def test_goldens_a11_diag_marker_projection(goldens):
    """The eta1 = 10 sinker goldens (A11 and diagonal y.y on 7 meshes): material points on the
    zero-perturbation 2x2x2 lattice of each element of the bump mesh, the single-sphere rule,
    the Q1 projection onto the quadrature points."""
    seen = 0
    for g in goldens:
        if g["eta1"] != 10.0:
            continue
        mx, my, mz = g["mx"], g["my"], g["mz"]
        c = golden_mesh(mx, my, mz)
        xi, wil, eta_p = _sinker_eta(mx, my, mz, c)
        eta = O.project_q1(mx, my, mz, xi, wil, eta_p)
        mask = O.bc_mask(mx, my, mz, 0)
        x = O.test_vector_u(mx, my, mz, c)
        y = O.matmult_a11(mx, my, mz, c, eta, mask, x)
        d = O.matgetdiag_a11(mx, my, mz, c, eta, mask)
        assert _rel(y @ y, g["yy"]["compare_mf_A11"]) < 5e-9, g["source"]
        assert _rel(d @ d, g["yy"]["compare_mf_diagA11"]) < 5e-9, g["source"]
        seen += 1
    assert seen == 11


def test_marker_projection_restatements_and_misreadings(goldens):
    """The C projection equals an independent numpy restatement (with and without the clipping,
    which cannot act on a Q1 interpolant), and three plausible misreadings (Q2 weights at every
    node, a plain point average, Q2 interpolation to the quadrature points) miss the 8x9x11
    diagonal golden."""
    g = [g for g in goldens if g["eta1"] == 10.0 and (g["mx"], g["my"], g["mz"]) == (8, 9, 11)][0]
    mx, my, mz = 8, 9, 11
    c = golden_mesh(mx, my, mz)
    xi, wil, eta_p = _sinker_eta(mx, my, mz, c)
    eta = O.project_q1(mx, my, mz, xi, wil, eta_p)
    assert np.max(np.abs(_np_project_q1(mx, my, mz, xi, wil, eta_p) - eta)) < 1e-13
    assert np.max(np.abs(_np_project_q1(mx, my, mz, xi, wil, eta_p, clip=False) - eta)) < 1e-13
    mask = O.bc_mask(mx, my, mz, 0)
    for kw in ({"vertex_weight": "q2"}, {"vertex_weight": "one"}, {"interp": "q2"}):
        bad = _np_project_q1(mx, my, mz, xi, wil, eta_p, **kw)
        d = O.matgetdiag_a11(mx, my, mz, c, bad, mask)
        assert _rel(d @ d, g["yy"]["compare_mf_diagA11"]) > 1e-6, kw
    # constancy: a uniform point field projects to the same constant at every quadrature point
    assert np.allclose(O.project_q1(mx, my, mz, xi, wil, np.full(len(wil), 3.5)), 3.5, rtol=0, atol=1e-14)


def test_golden_mesh_is_the_bump_mesh(goldens):
    """The flat mesh does NOT reproduce the goldens (guards against a silently wrong mesh)."""
    g = [g for g in goldens if g["eta1"] == 1.0][0]
    mx, my, mz = g["mx"], g["my"], g["mz"]
    c = O.mesh_uniform(mx, my, mz)
    x = O.test_vector_u(mx, my, mz, c)
    y21 = O.matmult_a21(mx, my, mz, c, O.bc_mask(mx, my, mz, 0), x)
    assert _rel(y21 @ y21, g["yy"]["compare_mf_A21"]) > 1e-4


# ---------------------------------------------------------- numpy restatement

def _np_shape_derivs():
    B, D = M.basis_1d()
    GNI = np.zeros((27, 3, 27))
    for q in range(27):
        a, b, c = q % 3, (q // 3) % 3, q // 9
        for n in range(27):
            i, j, k = n % 3, (n // 3) % 3, n // 9
            GNI[q, 0, n] = D[a, i] * B[b, j] * B[c, k]
            GNI[q, 1, n] = B[a, i] * D[b, j] * B[c, k]
            GNI[q, 2, n] = B[a, i] * B[b, j] * D[c, k]
    W = np.array([M.GAUSS_W[q % 3] * M.GAUSS_W[(q // 3) % 3] * M.GAUSS_W[q // 9] for q in range(27)])
    return GNI, W


def _np_element_matrix(X, eta):
    """81x81 A11 element matrix, K = sum_q w detJ B^T diag(2eta,2eta,2eta,eta,eta,eta) B."""
    GNI, W = _np_shape_derivs()
    K = np.zeros((81, 81))
    for q in range(27):
        J = GNI[q] @ X  # J[d][l] = dx_l/dxi_d
        dN = np.linalg.solve(J, GNI[q])  # dN[m][n] = dN_n/dx_m
        Bm = np.zeros((6, 81))
        Bm[0, 0::3] = dN[0]
        Bm[1, 1::3] = dN[1]
        Bm[2, 2::3] = dN[2]
        Bm[3, 0::3], Bm[3, 1::3] = dN[1], dN[0]
        Bm[4, 0::3], Bm[4, 2::3] = dN[2], dN[0]
        Bm[5, 1::3], Bm[5, 2::3] = dN[2], dN[1]
        Dm = np.diag([2 * eta[q]] * 3 + [eta[q]] * 3)
        K += W[q] * np.linalg.det(J) * (Bm.T @ Dm @ Bm)
    return K


def _elnodes(mx, my, ex, ey, ez):
    nx, ny = 2 * mx + 1, 2 * my + 1
    return np.array([(2 * ex + i) + nx * ((2 * ey + j) + ny * (2 * ez + k))
                     for k in range(3) for j in range(3) for i in range(3)])


def _np_assemble(mx, my, mz, coords, eta):
    n = 3 * M.nnodes(mx, my, mz)
    A = np.zeros((n, n))
    c = coords.reshape(-1, 3)
    e = 0
    for ez in range(mz):
        for ey in range(my):
            for ex in range(mx):
                nodes = _elnodes(mx, my, ex, ey, ez)
                K = _np_element_matrix(c[nodes], eta[27 * e:27 * e + 27])
                dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).reshape(-1)
                A[np.ix_(dofs, dofs)] += K
                e += 1
    return A


@pytest.fixture(scope="module")
def tiny():
    mx, my, mz = 2, 3, 2
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 0.25, 11)
    c = M.affine(c, M.SHEAR, [0.1, -0.2, 0.3])
    eta = 10.0 ** M.uniform(12, 27 * mx * my * mz, -1.0, 1.0)
    return mx, my, mz, c, eta


def test_a11_vs_numpy_assembled(tiny):
    mx, my, mz, c, eta = tiny
    A = _np_assemble(mx, my, mz, c, eta)
    u = M.uniform(13, A.shape[0], -1.0, 1.0)
    ref = A @ u
    for y in (O.a11_tensor(mx, my, mz, c, eta, u), O.a11_ref(mx, my, mz, c, eta, u),
              O.a11_tensor(mx, my, mz, c, eta, u, threads=3)):
        assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13
    # the C per-point element matrix agrees with the numpy one
    nodes = _elnodes(mx, my, 1, 2, 1)
    e = 1 + mx * (2 + my * 1)
    Kc = O.a11_element_matrix(c.reshape(-1, 3)[nodes].reshape(-1), eta[27 * e:27 * e + 27])
    Kn = _np_element_matrix(c.reshape(-1, 3)[nodes], eta[27 * e:27 * e + 27])
    assert np.abs(Kc - Kn).max() < 1e-12 * np.abs(Kn).max()
    # symmetric positive semi-definite
    assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()


def test_diag_is_diag_of_assembled(tiny):
    mx, my, mz, c, eta = tiny
    A = _np_assemble(mx, my, mz, c, eta)
    d = O.diag_a11(mx, my, mz, c, eta)
    assert np.abs(d - np.diag(A)).max() < 1e-13 * np.abs(d).max()


def test_rigid_body_null_space():
    m = 3
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 21)
    eta = 10.0 ** M.uniform(22, 27 * m ** 3, -2.0, 2.0)
    x = c.reshape(-1, 3)
    modes = []
    for d in range(3):
        t = np.zeros_like(x)
        t[:, d] = 1.0
        modes.append(t)
    modes.append(np.stack([-x[:, 1], x[:, 0], 0 * x[:, 0]], axis=1))
    modes.append(np.stack([-x[:, 2], 0 * x[:, 0], x[:, 0]], axis=1))
    modes.append(np.stack([0 * x[:, 0], -x[:, 2], x[:, 1]], axis=1))
    scale = np.linalg.norm(O.a11_tensor(m, m, m, c, eta, M.uniform(23, c.size, -1, 1)))
    for u in modes:
        y = O.a11_tensor(m, m, m, c, eta, u.reshape(-1))
        assert np.linalg.norm(y) < 1e-12 * scale


def test_linear_field_energy():
    """u = (x,0,0): u^T A11 u = int 2 eta eps:eps = 2 eta |Omega| = 2 on the jittered unit cube."""
    m = 4
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 31)
    x = c.reshape(-1, 3)
    u = np.stack([x[:, 0], 0 * x[:, 0], 0 * x[:, 0]], axis=1).reshape(-1)
    y = O.a11_tensor(m, m, m, c, np.ones(27 * m ** 3), u)
    assert abs(u @ y - 2.0) < 1e-12
    # with eta = 3 exactly three times
    y3 = O.a11_tensor(m, m, m, c, 3.0 * np.ones(27 * m ** 3), u)
    assert abs(u @ y3 - 6.0) < 1e-12


def test_affine_patch_test():
    """Affine mesh, constant eta, linear u: every interior row of A11 u vanishes."""
    m = 3
    c = M.affine(M.q2_uniform(m, m, m), M.SHEAR, [0.2, 0.0, -0.1])
    x = c.reshape(-1, 3)
    G = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [-0.5, 0.9, 1.1]])
    u = (x @ G.T).reshape(-1)
    y = O.a11_tensor(m, m, m, c, 2.5 * np.ones(27 * m ** 3), u).reshape(2 * m + 1, 2 * m + 1, 2 * m + 1, 3)
    assert np.abs(y[1:-1, 1:-1, 1:-1]).max() < 1e-12 * np.abs(y).max()
    # pressure patch test: constant p gives zero interior momentum rows
    p = np.zeros(4 * m ** 3)
    p[0::4] = 1.7
    y12 = O.a12(m, m, m, c, p).reshape(2 * m + 1, 2 * m + 1, 2 * m + 1, 3)
    assert np.abs(y12[1:-1, 1:-1, 1:-1]).max() < 1e-12 * np.abs(y12).max()


def test_a21_is_a12_transpose(tiny):
    mx, my, mz, c, _ = tiny
    nv, npr = 3 * M.nnodes(mx, my, mz), 4 * mx * my * mz
    u = M.uniform(41, nv, -1, 1)
    p = M.uniform(42, npr, -1, 1)
    lhs = p @ O.a21(mx, my, mz, c, u)
    rhs = O.a12(mx, my, mz, c, p) @ u
    assert abs(lhs - rhs) < 1e-13 * (abs(lhs) + 1)


def test_full_block_is_sum_of_blocks(tiny):
    mx, my, mz, c, eta = tiny
    nv, npr = 3 * M.nnodes(mx, my, mz), 4 * mx * my * mz
    u = M.uniform(51, nv, -1, 1)
    p = M.uniform(52, npr, -1, 1)
    yu, yp = O.stokes_a(mx, my, mz, c, eta, u, p)
    assert np.abs(yu - O.a11_tensor(mx, my, mz, c, eta, u) - O.a12(mx, my, mz, c, p)).max() < 1e-12
    assert np.abs(yp - O.a21(mx, my, mz, c, u)).max() < 1e-12
    mask = O.bc_mask(mx, my, mz, 0)
    yu2, yp2 = O.matmult_a(mx, my, mz, c, eta, mask, u, p)
    ul = np.where(mask > 0, 0.0, u)
    assert np.abs(yp2 - O.a21(mx, my, mz, c, ul)).max() < 1e-12
    assert np.all(yu2[mask > 0] == u[mask > 0])


def test_basis_closed_forms():
    B, D = M.basis_1d()
    assert np.allclose(B.sum(axis=1), 1.0, atol=1e-15)
    assert np.allclose(D.sum(axis=1), 0.0, atol=1e-15)
    assert B[1, 1] == 1.0 and B[1, 0] == 0.0 and B[1, 2] == 0.0 and D[1, 1] == 0.0
    # the Gauss rule integrates x^4 exactly: 2/5
    assert abs((M.GAUSS_W * M.GAUSS_X ** 4).sum() - 0.4) < 1e-15


def test_matmult_bc_semantics(tiny):
    mx, my, mz, c, eta = tiny
    mask = O.bc_mask(mx, my, mz, 1)
    x = M.uniform(61, mask.size, -1, 1)
    y = O.matmult_a11(mx, my, mz, c, eta, mask, x)
    A = _np_assemble(mx, my, mz, c, eta)
    keep = mask == 0
    ref = A[:, keep] @ x[keep]
    assert np.abs(y[keep] - ref[keep]).max() < 1e-12 * np.abs(ref).max()
    assert np.all(y[~keep] == x[~keep])
    d = O.matgetdiag_a11(mx, my, mz, c, eta, mask)
    assert np.all(d[~keep] == 1.0)


@pytest.mark.parametrize("m", [3, 6])
def test_avx_baseline_matches_oracle(m):
    """The CPU baseline (the AVX NEV=4 restatement of MFStokesWrapper_A11_AVX, colour-ordered
    scatter) equals the oracle's A11 to roundoff on a jittered variable-viscosity mesh whose
    element counts leave partial batches; its result does not depend on the thread count."""
    P = M.config_sinker(m, n_incl=3, contrast=1e3)
    u = M.uniform(2, P.nvel, -1, 1)
    ref = O.a11_tensor(m, m, m, P.coords, P.eta, u)
    y1 = O.a11_avx(m, m, m, P.coords, P.eta, u, threads=1)
    y3 = O.a11_avx(m, m, m, P.coords, P.eta, u, threads=3)
    assert np.linalg.norm(y1 - ref) <= 1e-13 * np.linalg.norm(ref)
    assert np.array_equal(y1, y3)
    assert O.avx_lib().ptq2a_flops_per_element() == 17253.0


def _jittered(m, seed=3, shear=True):
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), seed)
    return M.affine(c, M.SHEAR, [0.2, 0.0, -0.1]) if shear else c


def test_residual_without_forces_is_the_block_matmult():
    """FormFunction_Stokes with zero body forces and zero Dirichlet values equals
    MatMult_MFStokes_A of (u, p) (the goldens-pinned block, tensor form) to roundoff: the
    pressure and continuity signs of the per-point residual restatement follow the block's."""
    m = 4
    c = _jittered(m)
    eta = 10.0 ** M.uniform(5, 27 * m ** 3, -1, 1)
    mask = O.bc_mask(m, m, m, 0)
    u = M.uniform(6, 3 * M.nnodes(m, m, m), -1, 1)
    p = M.uniform(7, 4 * m ** 3, -1, 1)
    fu = np.zeros((m ** 3, 27, 3))
    Fu, Fp = O.residual(m, m, m, c, eta, mask, u, p, fu)
    yu, yp = O.matmult_a(m, m, m, c, eta, mask, u, p)
    assert np.abs(Fu - yu).max() <= 1e-13 * np.abs(yu).max()
    assert np.abs(Fp - yp).max() <= 1e-13 * np.abs(yp).max()
    assert np.array_equal(Fu[mask > 0], u[mask > 0])


def test_residual_dirichlet_values_and_body_forces():
    """Dirichlet values phi enter the element sums (u_D := phi_D) and the Dirichlet rows read
    u_D - phi_D; the momentum load enters as -Be (the oracle's body force, itself checked
    against a point-by-point assembly); a constant continuity load fp = 1 gives Fp = -int N^p:
    minus the element volume in the constant mode (total -1 on the unit cube), and zero in the
    linear modes of an undistorted uniform mesh (they are odd about the element centre)."""
    m = 3
    c = _jittered(m, shear=False)
    E = m ** 3
    eta = 10.0 ** M.uniform(8, 27 * E, -1, 1)
    mask = O.bc_mask(m, m, m, 1)
    n = 3 * M.nnodes(m, m, m)
    u = M.uniform(9, n, -1, 1)
    p = M.uniform(10, 4 * E, -1, 1)
    phi = M.uniform(11, n, -1, 1)
    fu = M.uniform(12, 81 * E, -1, 1).reshape(E, 27, 3)
    Fu, Fp = O.residual(m, m, m, c, eta, mask, u, p, fu, phi=phi)
    ul = np.where(mask > 0, phi, u)
    yu, yp = O.stokes_a(m, m, m, c, eta, ul, p)
    bf = O.body_force(m, m, m, c, fu)
    free = mask == 0
    assert np.abs(Fu[free] - (yu - bf)[free]).max() <= 1e-13 * np.abs(yu).max()
    assert np.array_equal(Fu[mask > 0], u[mask > 0] - phi[mask > 0])
    assert np.abs(Fp - yp).max() <= 1e-13 * np.abs(yp).max()
    assert np.abs(bf - M.body_force(m, m, m, c, fu)).max() <= 1e-14 * np.abs(bf).max()
    # continuity load alone
    zero = np.zeros(n)
    _, Fp1 = O.residual(m, m, m, O.mesh_uniform(m, m, m), eta, None, zero, np.zeros(4 * E), np.zeros((E, 27, 3)),
                        fp=np.ones((E, 27)))
    Fp1 = Fp1.reshape(E, 4)
    assert np.allclose(Fp1[:, 0], -1.0 / E, rtol=1e-13)
    assert np.abs(Fp1[:, 1:]).max() <= 1e-15
    _, Fp2 = O.residual(m, m, m, c, eta, None, zero, np.zeros(4 * E), np.zeros((E, 27, 3)), fp=np.ones((E, 27)))
    assert Fp2.reshape(E, 4)[:, 0].sum() == pytest.approx(-1.0, rel=1e-13)


def test_residual_affine_patch():
    """Patch test: on an affine mesh with constant viscosity, a linear velocity with constant
    strain and the constant pressure that makes the stress tensor constant leave every interior
    momentum row at zero (each interior node's test function integrates its gradient to 0)."""
    m = 3
    c = M.affine(O.mesh_uniform(m, m, m), M.SHEAR, [0.1, 0.2, 0.3])
    E = m ** 3
    eta = np.full(27 * E, 2.5)
    X = c.reshape(-1, 3)
    G = np.array([[0.3, -0.2, 0.5], [0.1, 0.4, -0.3], [0.2, 0.6, -0.7]])
    u = (X @ G.T).reshape(-1)
    p = np.zeros(4 * E)
    p[0::4] = 1.7
    Fu, _ = O.residual(m, m, m, c, eta, None, u, p, np.zeros((E, 27, 3)))
    n1 = 2 * m + 1
    F = Fu.reshape(n1, n1, n1, 3)[1:-1, 1:-1, 1:-1]
    assert np.abs(F).max() <= 1e-13 * np.abs(Fu).max()
