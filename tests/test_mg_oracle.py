"""Pins of the CPU multigrid restatement (oracle/mg_oracle.py) -- no GPU."""
import numpy as np

from oracle import mg_oracle as G
from ptq2 import mesh as M


def test_restriction_is_prolongation_transpose():
    mxc, myc, mzc = 2, 3, 2
    nc = 3 * (2 * mxc + 1) * (2 * myc + 1) * (2 * mzc + 1)
    nf = 3 * (4 * mxc + 1) * (4 * myc + 1) * (4 * mzc + 1)
    a = M.uniform(1, nc, -1, 1)
    b = M.uniform(2, nf, -1, 1)
    assert abs(G.prolong(mxc, myc, mzc, a) @ b - a @ G.restrict(mxc, myc, mzc, b)) < 1e-13


def test_prolongation_reproduces_trilinear_fields():
    m = 4
    fine = M.q2_uniform(m, m, m)
    L = G.Level(m, m, m, fine, None, None)
    x = fine.reshape(-1, 3)
    f = np.stack([1 + 2 * x[:, 0] - x[:, 1] * x[:, 2], x[:, 0] * x[:, 1], 3 * x[:, 2]], axis=1).reshape(-1)
    coarse = G.inject_from_fine(L, f)
    assert np.abs(G.prolong(m // 2, m // 2, m // 2, coarse) - f).max() < 1e-14


def test_eta_coarsening_constant_and_mean():
    m = 4
    eta = np.full(27 * m ** 3, 7.0)
    assert np.allclose(G.eta_coarsen(m, m, m, eta), 7.0)
    eta = M.uniform(3, 27 * m ** 3, 1, 2)
    ec = G.eta_coarsen(m, m, m, eta).reshape(-1, 27)
    e = eta.reshape(m, m, m, 27)
    # the centre point of coarse element 0 averages all 8 children
    assert abs(ec[0, 13] - e[:2, :2, :2].mean()) < 1e-14
    # corner point 0 of coarse element 0 sees only child (0,0,0)
    assert abs(ec[0, 0] - e[0, 0, 0].mean()) < 1e-14


def test_eta_coarsening_rules_are_ordered_and_exact():
    """Harmonic <= geometric <= arithmetic <= maximum at every coarse point (the power-mean
    inequality), equality for constant children, and a hand-computed two-value case."""
    m = 4
    eta = 10.0 ** M.uniform(5, 27 * m ** 3, -2, 4)
    h, g, a, x = (G.eta_coarsen(m, m, m, eta, mode) for mode in ("harmonic", "geometric", "arithmetic", "max"))
    assert np.all(h <= g * (1 + 1e-12)) and np.all(g <= a * (1 + 1e-12)) and np.all(a <= x * (1 + 1e-12))
    assert np.all(h > 0)
    for mode in ("harmonic", "geometric", "arithmetic", "max"):
        assert np.allclose(G.eta_coarsen(m, m, m, np.full(27 * m ** 3, 3.5), mode), 3.5, rtol=1e-14)
    # children of coarse element 0 split along x: the x=0 children hold 1, the x=1 children 4;
    # its centre point (qa = qb = qc = 1) sees all 8 children, 4 of each value
    e = np.ones((m, m, m, 27))
    e[:, :, 1::2] = 4.0
    ec = {mode: G.eta_coarsen(m, m, m, e.reshape(-1), mode).reshape(-1, 27)[0, 13]
          for mode in ("harmonic", "geometric", "arithmetic", "max")}
    assert abs(ec["arithmetic"] - 2.5) < 1e-14
    assert abs(ec["geometric"] - 2.0) < 1e-14
    assert abs(ec["harmonic"] - 1.0 / (0.5 * (1.0 + 0.25))) < 1e-14
    assert ec["max"] == 4.0


def test_chebyshev_reduces_error_like_the_polynomial():
    """On a level with exact eigen-bounds the Chebyshev smoother's error reduction obeys the
    Chebyshev bound 1/T_k((emax+emin)/(emax-emin)) on the smoothed part of the spectrum."""
    m = 2
    P = M.config_jitter_iso(m)
    L = G.Level(m, m, m, P.coords, None, P.eta)
    n = P.coords.size
    A = np.stack([L.matmult(np.eye(n)[j]) for j in range(n)], axis=1)
    d = np.diag(A)
    S = A / np.sqrt(d)[:, None] / np.sqrt(d)[None, :]
    lam, V = np.linalg.eigh(0.5 * (S + S.T))
    emax = lam[-1]
    emin = 0.3 * emax
    its = 4
    # error of x = Cheb(b) against the exact solution, in the high part of the spectrum
    xs = M.uniform(4, n, -1, 1)
    b = A @ xs
    x = G.chebyshev(L, b, None, 1.0 / d, emin, emax, its, True)
    e0 = np.sqrt(d) * xs
    e1 = np.sqrt(d) * (xs - x)
    hi = lam >= emin
    c0 = V.T @ e0
    c1 = V.T @ e1
    th = np.cosh((its + 1) * np.arccosh((emax + emin) / (emax - emin)))
    ratio = np.linalg.norm(c1[hi]) / np.linalg.norm(c0[hi])
    assert ratio <= 1.0 / th * 1.0001


def test_vcycle_preconditions_gmres():
    P = M.config_sinker(8, n_incl=2, contrast=1e2, bc="freeslip")
    lv = G.hierarchy(8, 8, 8, P.coords, P.mask, P.eta, 3)
    # eigen-bounds from a dense estimate on the two smoothed levels
    emin, emax = [], []
    for L in lv[:-1]:
        n = L.n
        rng = np.random.default_rng(0)
        v = rng.standard_normal(n)
        d = L.diag()
        for _ in range(60):
            w = L.matmult(v) / d
            lam = np.linalg.norm(w) / np.linalg.norm(v)
            v = w / np.linalg.norm(w)
        emax.append(1.1 * lam)
        emin.append(0.2 * lam)
    V = G.VCycle(lv, emin, emax, 4)
    xs = M.uniform(5, lv[0].n, -1, 1)
    xs[P.mask > 0] = 0.0
    b = lv[0].matmult(xs)
    r0 = np.linalg.norm(b)
    x = V(b)
    assert np.linalg.norm(b - lv[0].matmult(x)) < 0.05 * r0
    # as a GMRES preconditioner it reaches 1e-8 in a few tens of iterations
    import scipy.sparse.linalg as sla

    n = lv[0].n
    A = sla.LinearOperator((n, n), matvec=lambda v: lv[0].matmult(v))
    Mp = sla.LinearOperator((n, n), matvec=lambda v: V(v))
    hist = []
    x, info = sla.gmres(A, b, M=Mp, rtol=1e-8, restart=30, maxiter=100, callback=lambda r: hist.append(r),
                        callback_type="pr_norm")
    assert info == 0 and len(hist) <= 25
    assert np.linalg.norm(b - lv[0].matmult(x)) <= 1e-8 * r0 * 1.01


def test_galerkin_coarse_operator_is_rap():
    """The Galerkin coarsest operator of the restatement equals P^T A_f P assembled from an
    explicit trilinear prolongation matrix (restricted to interior coarse dofs)."""
    m = 4
    P = M.config_sinker(m, n_incl=3, contrast=1e2, bc="freeslip")
    lv = G.hierarchy(m, m, m, P.coords, P.mask, P.eta, 2)
    F, C = lv
    nf, nc = F.n, C.n
    Pm = np.zeros((nf, nc))
    for j in range(nc):
        e = np.zeros(nc)
        e[j] = 1.0
        Pm[:, j] = G.prolong(C.mx, C.my, C.mz, e)
    Pm[F.mask > 0, :] = 0.0
    Pm[:, C.mask > 0] = 0.0
    Af = np.stack([F.matmult(np.eye(nf)[k]) for k in range(nf)], axis=1)
    Ac = Pm.T @ Af @ Pm
    Ac[C.mask > 0, C.mask > 0] = 1.0
    inv = G.galerkin_inverse(F, C)
    assert np.allclose(inv @ Ac, np.eye(nc), atol=1e-8)
    # the restriction is the transpose of the prolongation
    r = np.random.default_rng(0).standard_normal(nf)
    assert np.allclose(G.restrict(C.mx, C.my, C.mz, r), (np.stack([G.prolong(C.mx, C.my, C.mz, np.eye(nc)[j]) for j in range(nc)], axis=1)).T @ r)


def test_galerkin_level_is_ptap_and_a_5x5x5_stencil():
    """GalerkinLevel (the restatement of ptq2_mg_opts.galerkin_levels) equals Pi P^T A_f P Pi +
    (I - Pi) built from an explicit prolongation matrix (Kronecker product of the 1D
    interpolations, independent of prolong/restrict), couples a node only to nodes within two
    nodes per axis, and is symmetric to rounding; through a matrix-free level and through a
    level that is itself Galerkin."""
    import scipy.sparse as sp

    def p1d(nc):
        nf = 2 * nc - 1
        r, c, v = [], [], []
        for f in range(nf):
            if f % 2 == 0:
                r.append(f); c.append(f // 2); v.append(1.0)
            else:
                r += [f, f]; c += [f // 2, f // 2 + 1]; v += [0.5, 0.5]
        return sp.csr_matrix((v, (r, c)), shape=(nf, nc))

    def ptap(F, C, Af):
        Pk = sp.kron(sp.kron(sp.kron(p1d(2 * C.mz + 1), p1d(2 * C.my + 1)), p1d(2 * C.mx + 1)), sp.identity(3))
        Pm = sp.diags((F.mask == 0).astype(float)) @ Pk.tocsr() @ sp.diags((C.mask == 0).astype(float))
        G_ = Pm.T @ Af @ Pm
        return (G_.toarray() if sp.issparse(G_) else np.asarray(G_)) + np.diag((C.mask > 0).astype(float))

    def check(F, C, Af):
        want = ptap(F, C, Af)
        A = C.A.toarray()
        assert np.abs(A - want).max() <= 1e-12 * np.abs(want).max()
        assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
        NX, NY = 2 * C.mx + 1, 2 * C.my + 1
        r, c = C.A.nonzero()
        nr, nc_ = r // 3, c // 3
        for a, b in ((nr % NX, nc_ % NX), ((nr // NX) % NY, (nc_ // NX) % NY), (nr // (NX * NY), nc_ // (NX * NY))):
            assert np.abs(a - b).max() <= 2

    P = M.config_sinker(4, n_incl=3, contrast=1e3, bc="freeslip")
    lv = G.hierarchy(4, 4, 4, P.coords, P.mask, P.eta, 3, galerkin_levels=1)
    assert type(lv[0]) is G.Level and type(lv[1]) is G.GalerkinLevel and type(lv[2]) is G.Level
    Af = np.stack([lv[0].matmult(np.eye(lv[0].n)[k]) for k in range(lv[0].n)], axis=1)
    check(lv[0], lv[1], Af)
    P = M.config_sinker(8, n_incl=3, contrast=1e3, bc="freeslip")
    lv = G.hierarchy(8, 8, 8, P.coords, P.mask, P.eta, 4, galerkin_levels=2)
    assert [type(L) for L in lv] == [G.Level, G.GalerkinLevel, G.GalerkinLevel, G.Level]
    check(lv[1], lv[2], lv[1].A)


def test_stiff_inclusions_match_connected_component_labelling():
    """Inclusions of the rigid-body correction = scipy's 26-connected labelling of the elements
    with a stiff quadrature point (numbered in raster order), and the size cap keeps the largest."""
    ndimage = __import__("pytest").importorskip("scipy.ndimage")
    rng = np.random.default_rng(7)
    for _ in range(12):
        mx, my, mz = (int(v) for v in rng.integers(2, 9, 3))
        eta = np.ones((mx * my * mz, 27))
        hit = rng.random(eta.shape) < 0.01  # single stiff points make an element stiff
        eta[hit] = 1e4
        lab, nc = G.stiff_inclusions(mx, my, mz, eta.reshape(-1))
        ref, nref = ndimage.label(hit.any(axis=1).reshape(mz, my, mx), structure=np.ones((3, 3, 3)))
        assert nc == nref and np.array_equal(lab, ref.reshape(-1) - 1)
        if nc >= 2:
            sizes = np.bincount(lab[lab >= 0])
            lab1, n1 = G.stiff_inclusions(mx, my, mz, eta.reshape(-1), max_incl=1)
            big = int(np.argmax(sizes))  # first of the largest
            assert n1 == 1 and np.array_equal(lab1 >= 0, lab == big)
    # below the contrast threshold there is nothing to deflate
    assert G.stiff_inclusions(2, 2, 2, np.full(8 * 27, 3.0))[1] == 0


def test_rigid_motions_have_no_strain_and_adef2_reproduces_them():
    """With every element stiff, W holds the 6 rigid motions of the whole mesh: A W = 0 without
    BCs (isoparametric Q2 reproduces linear fields; jittered mesh).  With a Dirichlet mask and
    one inclusion, the A-DEF2 correction maps r = A W c back to W c whatever M is, and a
    dependent column is dropped by the factorisation."""
    m = 3
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 1)
    eta = np.full((m ** 3, 27), 1e4)
    eta[0, 0] = 1.0
    L = G.Level(m, m, m, c, None, eta.reshape(-1))
    D = G.RigidDeflation(L)
    assert D.nc == 1 and (D.node_label == 0).all()
    assert np.abs(D.AW).max() <= 1e-9 * np.abs(D.W).max() * 1e4
    # translations and rotations are independent and of unit RMS size
    assert np.linalg.matrix_rank(D.W) == 6
    mask = M.bc_mask(m, m, m, "freeslip")
    eta = np.ones((m ** 3, 27))
    eta[13] = 1e3  # the centre element
    L = G.Level(m, m, m, c, mask, eta.reshape(-1))
    D = G.RigidDeflation(L)
    assert D.nc == 1 and len(D.keep) == 6
    coef = M.uniform(9, 6, -1, 1)
    z = D(lambda r: np.sin(r), D.AW @ coef)  # any M: its input is zero
    assert np.allclose(z, D.W @ coef, rtol=0, atol=1e-12)
    Wd = np.concatenate([D.W, D.W[:, :1]], axis=1)
    E = Wd.T @ np.stack([L.matmult(Wd[:, j]) for j in range(7)], axis=1)
    assert G.drop_cholesky_keep(0.5 * (E + E.T)) == [0, 1, 2, 3, 4, 5]


def _q1proj_brute(mx, my, mz, eta, nlevels):
    """The reference's coefficient hierarchy written as plain loops (element by element, point by
    point, node by node) for tiny meshes."""
    g = [-np.sqrt(0.6), 0.0, np.sqrt(0.6)]
    NX, NY, NZ = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    A = np.zeros(NX * NY * NZ)
    B = np.zeros(NX * NY * NZ)
    e = np.asarray(eta).reshape(-1, 27)
    for ez in range(mz):
        for ey in range(my):
            for ex in range(mx):
                el = ex + mx * (ey + my * ez)
                for p in range(27):
                    xi = (g[p % 3], g[(p // 3) % 3], g[p // 9])
                    for c in range(3):
                        for b in range(3):
                            for a in range(3):
                                n = (2 * ex + a) + NX * ((2 * ey + b) + NY * (2 * ez + c))
                                if a % 2 == 0 and b % 2 == 0 and c % 2 == 0:
                                    w = 0.125
                                    for s, x in zip((a, b, c), xi):
                                        w *= (1 + x) if s == 2 else (1 - x)
                                else:
                                    w = 1.0
                                A[n] += w * e[el, p]
                                B[n] += w
    nodal = (A / B).reshape(NZ, NY, NX)
    out = []
    for _ in range(nlevels - 1):
        mx, my, mz = mx // 2, my // 2, mz // 2
        cx, cy, cz = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
        c = np.zeros((cz, cy, cx))
        for K in range(cz):
            for J in range(cy):
                for I in range(cx):
                    num = den = 0.0
                    for fk in range(nodal.shape[0]):
                        for fj in range(nodal.shape[1]):
                            for fi in range(nodal.shape[2]):
                                w = 1.0
                                for f, C in ((fk, K), (fj, J), (fi, I)):
                                    w *= 1.0 if f == 2 * C else (0.5 if abs(f - 2 * C) == 1 else 0.0)
                                num += w * nodal[fk, fj, fi]
                                den += w
                    c[K, J, I] = num / den
        nodal = c
        q = np.zeros((mx * my * mz, 27))
        for ez in range(mz):
            for ey in range(my):
                for ex in range(mx):
                    vals = nodal[2 * ez:2 * ez + 3, 2 * ey:2 * ey + 3, 2 * ex:2 * ex + 3]
                    for p in range(27):
                        xi = (g[p % 3], g[(p // 3) % 3], g[p // 9])
                        s = 0.0
                        for cc in range(8):
                            w = 0.125
                            for ax in range(3):
                                w *= (1 + xi[ax]) if (cc >> ax) & 1 else (1 - xi[ax])
                            s += w * vals[2 * (cc >> 2), 2 * ((cc >> 1) & 1), 2 * (cc & 1)]
                        q[ex + mx * (ey + my * ez), p] = min(max(s, vals.min()), vals.max())
        out.append(q.reshape(-1))
    return out


def test_q1_projection_hierarchy_reproduces_linear_fields_inside():
    """What the mathematics fixes for the coefficient hierarchy: every step is a symmetric
    weighted average (corner weights of the nodal projection mirror across the 8 elements of an
    interior node, the restriction weights 1/2, 1, 1/2 are symmetric) or the Q1 interpolation,
    which reproduces linear fields. So a field linear in space comes out exact at the coarse
    quadrature points wherever no boundary node enters. In fine Q2-lattice units (node F at
    F), fine element e spans [2e, 2e+2]; level-l element ex spans fine [2^(l+1) ex, 2^(l+1)(ex+1)]."""
    m = 16
    g = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])
    coef = np.array([5.0, 0.3, 0.2, 0.1])  # f = 5 + 0.3 X + 0.2 Y + 0.1 Z
    e = np.arange(m)
    q = np.arange(27)
    pos = lambda el, gq, scale: scale * (2 * el[:, None] + 1 + gq[None, :])  # noqa: E731
    X = pos(e, g[q % 3], 1)            # [ex, q]
    Y = pos(e, g[(q // 3) % 3], 1)
    Z = pos(e, g[q // 9], 1)
    eta = (coef[0] + coef[1] * X[None, None, :, :] + coef[2] * Y[None, :, None, :]
           + coef[3] * Z[:, None, None, :]).reshape(-1)
    levels = G.q1_projection_hierarchy(m, m, m, eta, 3)
    for lvl, vals in enumerate(levels, start=1):
        mc = m >> lvl
        sc = 2 ** lvl
        ec = np.arange(mc)
        Xc, Yc, Zc = pos(ec, g[q % 3], sc), pos(ec, g[(q // 3) % 3], sc), pos(ec, g[q // 9], sc)
        want = (coef[0] + coef[1] * Xc[None, None, :, :] + coef[2] * Yc[None, :, None, :]
                + coef[3] * Zc[:, None, None, :])
        got = vals.reshape(mc, mc, mc, 27)
        # elements whose coarse lattice nodes and restriction stencils stay off the boundary:
        # level 1 nodes C in [1, m-1]; level 2 nodes D with 2D-1..2D+1 in [1, m-1]
        lo, hi = 1, (mc * 2 - 3) // 2 if lvl == 1 else (m // 2 - 1 - 2) // 2
        inner = slice(lo, hi + 1)
        assert hi >= lo
        assert np.allclose(got[inner, inner, inner], want[inner, inner, inner], rtol=1e-13, atol=0)
        # and the boundary layer is not exact (the check above is not vacuous)
        assert not np.allclose(got[0, 0, 0], want[0, 0, 0], rtol=1e-6, atol=0)


def test_q1_projection_hierarchy_matches_loops_and_invariants():
    """The reference's coefficient hierarchy (coarse_eta_mean = q1proj): the vectorised
    restatement against the plain-loop one; constant fields stay exact; coarse values stay
    within the fine range (all weights are non-negative)."""
    rng = np.random.default_rng(5)
    eta = 10.0 ** rng.uniform(-1, 3, 4 * 2 * 4 * 27)
    fast = G.q1_projection_hierarchy(4, 2, 4, eta, 2)
    slow = _q1proj_brute(4, 2, 4, eta, 2)
    assert np.allclose(fast[0], slow[0], rtol=1e-13, atol=0)
    for lv in G.q1_projection_hierarchy(8, 8, 4, np.full(8 * 8 * 4 * 27, 2.5), 3):
        assert np.allclose(lv, 2.5, rtol=1e-15, atol=0)
    big = 10.0 ** rng.uniform(0, 6, 8 * 8 * 8 * 27)
    for lv in G.q1_projection_hierarchy(8, 8, 8, big, 4):
        assert lv.min() >= big.min() and lv.max() <= big.max()
