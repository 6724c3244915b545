"""GPU multigrid / FGMRES viscous solve against the CPU restatement (oracle/mg_oracle.py)."""
import numpy as np
import pytest

from oracle import mg_oracle as G
from oracle import oracle as O
import ptq2
from ptq2 import _lib
from ptq2 import mesh as M
from ptq2.solver import ViscousSolver

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = torch.device("cuda:0")


def say(*a):
    print(*a, flush=True)


def make(m, contrast, bc="freeslip", n_incl=3):
    P = M.config_sinker(m, n_incl=n_incl, contrast=contrast, bc=bc)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(m, m, m, P.coords, P.mask)
    ctx.set_viscosity(P.eta)
    return P, ctx


@pytest.mark.parametrize("galerkin,autotune", [(False, 0), (True, 0), (True, 2)])
def test_vcycle_fp64_matches_restatement(galerkin, autotune):
    """autotune 2 runs the row-split (three waves per brick) smoother kernel on every level."""
    m = 8
    say("case: FP64 V-cycle 8^3 (3 levels, galerkin coarse %s, autotune %d) vs CPU restatement" % (galerkin, autotune))
    P, ctx = make(m, 1e2)
    S = ViscousSolver(ctx, nlevels=3, smooth_its=4, fp32=False, use_graph=True, galerkin_coarse=galerkin,
                      autotune=autotune)
    assert [k for k, _, _, _ in S.tuning()] == ([1, 1] if autotune == 2 else [0, 0])
    assert [tuple(r) for r in S.levels()] == [(8, 8, 8), (4, 4, 4), (2, 2, 2)]
    emin, emax = S.eigs()
    lv = G.hierarchy(m, m, m, P.coords, P.mask, P.eta, 3)
    V = G.VCycle(lv, emin, emax, 4, galerkin=galerkin)
    r = M.uniform(3, P.nvel, -1, 1)
    r[P.mask > 0] = 0.0
    z = S.vcycle(torch.from_numpy(r).to(DEV)).cpu().numpy()
    zr = V(r)
    assert np.linalg.norm(z - zr) / np.linalg.norm(zr) <= 1e-10
    # the eigen-estimate is close to the true lambda_max(D^-1 A) of the fine level
    L = lv[0]
    d = L.diag()
    v = M.uniform(4, L.n, -1, 1)
    for _ in range(80):
        w = L.matmult(v) / d
        lam = np.linalg.norm(w) / np.linalg.norm(v)
        v = w / np.linalg.norm(w)
    assert 0.9 * lam <= emax[0] / 1.1 <= 1.01 * lam


@pytest.mark.parametrize("mean", ["geometric", "harmonic", "max", "q1proj"])
def test_vcycle_coarse_viscosity_rules(mean):
    """The coarse quadrature viscosity rules other than the arithmetic default, through the whole
    V-cycle, against the restatement's hierarchy built with the same rule; then the V-cycle
    applied in place (input and output the same vector) equals the out-of-place result."""
    m = 8
    say("case: FP64 V-cycle 8^3, coarse viscosity %s, vs CPU restatement" % mean)
    P, ctx = make(m, 1e3)
    S = ViscousSolver(ctx, nlevels=3, smooth_its=3, fp32=False, use_graph=True, galerkin_coarse=False, eta_mean=mean)
    emin, emax = S.eigs()
    lv = G.hierarchy(m, m, m, P.coords, P.mask, P.eta, 3, mode=mean)
    V = G.VCycle(lv, emin, emax, 3)
    r = M.uniform(31, P.nvel, -1, 1)
    r[P.mask > 0] = 0.0
    rd = torch.from_numpy(r).to(DEV)
    z = S.vcycle(rd).cpu().numpy()
    zr = V(r)
    assert np.linalg.norm(z - zr) / np.linalg.norm(zr) <= 1e-10
    w = rd.clone()
    S.vcycle(w, w)
    torch.cuda.synchronize()
    assert np.array_equal(w.cpu().numpy(), z)
    S.destroy()


@pytest.mark.parametrize("autotune", [0, 2, 3])
def test_graph_replays_match_eager(autotune):
    m = 8
    say("case: V-cycle graph replays vs eager launches, autotune %d" % autotune)
    P, ctx = make(m, 1e3)
    Sg = ViscousSolver(ctx, nlevels=3, fp32=True, use_graph=True, autotune=autotune)
    P2, ctx2 = make(m, 1e3)
    Se = ViscousSolver(ctx2, nlevels=3, fp32=True, use_graph=False, autotune=autotune)
    for k in range(4):
        r = torch.from_numpy(M.uniform(10 + k, P.nvel, -1, 1)).to(DEV)
        zg = Sg.vcycle(r).clone()
        _ = torch.ones(1000, device=DEV).sum().item()  # eager work and a sync between replays
        ze = Se.vcycle(r)
        torch.cuda.synchronize()
        assert torch.equal(zg, ze), k


def test_fgmres_solve_16():
    """FGMRES(30) + GMG to rtol 1e-8 at viscosity contrast 1e4, FP64 and FP32 smoothers.

    Both reach the FP64 true residual; the FP32 smoother costs at most a few extra
    iterations (its error is far below the Chebyshev smoothing factor)."""
    m = 16
    P, ctx = make(m, 1e4, n_incl=6)
    xs = M.uniform(21, P.nvel, -1, 1)
    xs[P.mask > 0] = 0.0
    b = O.matmult_a11(m, m, m, P.coords, P.eta, P.mask, xs)
    its = {}
    for fp32 in (False, True):
        say("case: FGMRES + GMG solve 16^3, contrast 1e4, fp32 smoother=%s" % fp32)
        S = ViscousSolver(ctx, nlevels=4, smooth_its=4, fp32=fp32)
        x, info = S.solve(torch.from_numpy(b).to(DEV), rtol=1e-8, max_it=400, restart=30)
        x = x.cpu().numpy()
        say("  iterations %d, rel residual %.3e" % (info["iterations"], info["rel"]))
        r = b - O.matmult_a11(m, m, m, P.coords, P.eta, P.mask, x)
        assert np.linalg.norm(r) <= 1.05e-8 * np.linalg.norm(b)
        h = S.history()
        assert h[0] == pytest.approx(np.linalg.norm(b), rel=1e-12)
        its[fp32] = info["iterations"]
        S.destroy()
    assert its[True] <= its[False] + max(3, its[False] // 10)


@pytest.mark.parametrize("restart", [30, 7])
def test_fgmres_fp32_basis(restart):
    """FP32 storage of the finished Arnoldi vectors: the FP64 true residual still reaches rtol
    1e-8 (checked with the restatement's operator), with at most a few more iterations than the
    FP64 basis; restart 7 crosses many cycles (each recomputes the FP64 residual)."""
    m = 16
    P, ctx = make(m, 1e4, n_incl=6)
    xs = M.uniform(22, P.nvel, -1, 1)
    xs[P.mask > 0] = 0.0
    b = O.matmult_a11(m, m, m, P.coords, P.eta, P.mask, xs)
    its = {}
    for b32 in (False, True):
        say("case: FGMRES(%d) 16^3 contrast 1e4, FP32 smoother, fp32 basis=%s" % (restart, b32))
        S = ViscousSolver(ctx, nlevels=4, smooth_its=4, fp32=True, fp32_basis=b32)
        x, info = S.solve(torch.from_numpy(b).to(DEV), rtol=1e-8, max_it=600, restart=restart)
        x = x.cpu().numpy()
        say("  iterations %d, rel residual %.3e" % (info["iterations"], info["rel"]))
        r = b - O.matmult_a11(m, m, m, P.coords, P.eta, P.mask, x)
        assert np.linalg.norm(r) <= 1.05e-8 * np.linalg.norm(b)
        its[b32] = info["iterations"]
        S.destroy()
    assert its[True] <= its[False] + max(3, its[False] // 10)


def test_vcycle_large_coarse_level():
    """Two levels, 12^3 -> 6^3: the coarse dense inverse (6591 dofs, 103 sweep blocks, padded)
    against the restatement's LAPACK inverse."""
    m = 12
    say("case: FP64 V-cycle 12^3 (2 levels, coarse 6591 dofs) vs CPU restatement")
    P, ctx = make(m, 1e3)
    S = ViscousSolver(ctx, nlevels=2, smooth_its=2, fp32=False, use_graph=True, galerkin_coarse=False)
    emin, emax = S.eigs()
    lv = G.hierarchy(m, m, m, P.coords, P.mask, P.eta, 2)
    V = G.VCycle(lv, emin, emax, 2)
    r = M.uniform(14, P.nvel, -1, 1)
    r[P.mask > 0] = 0.0
    z = S.vcycle(torch.from_numpy(r).to(DEV)).cpu().numpy()
    zr = V(r)
    assert np.linalg.norm(z - zr) / np.linalg.norm(zr) <= 1e-9


def test_batched_solves_match_sequential():
    """Independent problems solved concurrently (one stream and host thread each) give the
    same iterates, bit for bit, as the same solves run one after the other."""
    from ptq2.solver import auto_levels, solve_batch

    cases = [(8, 1e2), (10, 1e4), (12, 1e3)]
    say("case: batched solves %s" % (cases,))
    solvers, rhs, seq = [], [], []
    for k, (m, contrast) in enumerate(cases):
        P, ctx = make(m, contrast)
        S = ViscousSolver(ctx, nlevels=auto_levels(m, 2000), smooth_its=4, fp32=True)
        xs = M.uniform(30 + k, P.nvel, -1, 1)
        xs[P.mask > 0] = 0.0
        b = torch.from_numpy(O.matmult_a11(m, m, m, P.coords, P.eta, P.mask, xs)).to(DEV)
        x, info = S.solve(b, rtol=1e-8, max_it=300)
        torch.cuda.synchronize()
        seq.append((x.clone(), info))
        solvers.append(S)
        rhs.append(b)
    out = solve_batch(solvers, rhs, rtol=1e-8, max_it=300)
    torch.cuda.synchronize()
    for (xs_, is_), (xb, ib) in zip(seq, out):
        assert ib["iterations"] == is_["iterations"]
        assert torch.equal(xs_, xb)
        assert ib["rel"] <= 1.05e-8


@pytest.mark.parametrize("fp32", [False, True])
def test_rigid_deflation_vcycle_matches_restatement(fp32):
    """The two-level correction for stiff inclusions: inclusion labels bit-exact against the
    restatement, the corrected V-cycle against it (FP64 1e-10; FP32 smoothing 1e-4), graph and
    direct launches identical, and r = A W c mapped back to W c."""
    m = 8
    say("case: rigid-body deflation, 8^3 sinker contrast 1e4, fp32=%s" % fp32)
    P, ctx = make(m, 1e4, n_incl=4)
    lv = G.hierarchy(m, m, m, P.coords, P.mask, P.eta, 3)
    D = G.RigidDeflation(lv[0])
    assert D.nc >= 2
    S = ViscousSolver(ctx, nlevels=3, smooth_its=3, fp32=fp32, use_graph=True, galerkin_coarse=False,
                      rigid_deflation=True, autotune=0)
    ni, nk, lab = S.rigid_info(labels=True)
    assert ni == D.nc and nk == len(D.keep) and np.array_equal(lab, D.node_label)
    emin, emax = S.eigs()
    V = G.VCycle(lv, emin, emax, 3)
    r = M.uniform(41, P.nvel, -1, 1)
    r[P.mask > 0] = 0.0
    z = S.vcycle(torch.from_numpy(r).to(DEV)).cpu().numpy()
    zr = D(V, r)
    err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    say("  inclusions %d, columns kept %d, rel err %.3e" % (ni, nk, err))
    assert err <= (1e-4 if fp32 else 1e-10)
    S2 = ViscousSolver(ctx, nlevels=3, smooth_its=3, fp32=fp32, use_graph=False, galerkin_coarse=False,
                       rigid_deflation=True, autotune=0)
    z2 = S2.vcycle(torch.from_numpy(r).to(DEV)).cpu().numpy()
    assert np.array_equal(z, z2)
    S2.destroy()
    if not fp32:
        coef = M.uniform(42, D.W.shape[1], -1, 1)
        zc = S.vcycle(torch.from_numpy(D.AW @ coef).to(DEV)).cpu().numpy()
        assert np.linalg.norm(zc - D.W @ coef) <= 1e-9 * np.linalg.norm(D.W @ coef)
    S.destroy()


def test_rigid_deflation_options_and_cap():
    m = 8
    P, ctx = make(m, 1e4, n_incl=4)
    S = ViscousSolver(ctx, nlevels=3, smooth_its=2, fp32=True, rigid_deflation=False)
    assert S.rigid_info() == (0, 0)
    S.destroy()
    lv0 = G.hierarchy(m, m, m, P.coords, P.mask, P.eta, 1)[0]
    D1 = G.RigidDeflation(lv0, max_incl=1)
    S = ViscousSolver(ctx, nlevels=3, smooth_its=2, fp32=True, rigid_deflation=True, rigid_max_inclusions=1)
    ni, nk, lab = S.rigid_info(labels=True)
    assert ni == 1 and np.array_equal(lab, D1.node_label) and nk == len(D1.keep)
    S.destroy()
    S = ViscousSolver(ctx, nlevels=3, smooth_its=2, fp32=True, rigid_deflation=True, rigid_min_contrast=1e9)
    assert S.rigid_info() == (0, 0)
    S.destroy()
    for bad in ({"rigid_max_inclusions": 0}, {"rigid_max_inclusions": 129}, {"rigid_min_contrast": 0.5}):
        with pytest.raises(ptq2.PTQ2Error):
            ViscousSolver(ctx, nlevels=3, fp32=True, rigid_deflation=True, **bad)
    # the other options' ranges: Galerkin levels 0 .. nlevels-2, autotune 0 .. 4
    for bad in ({"galerkin_levels": 2}, {"galerkin_levels": -1}, {"autotune": 5}, {"autotune": -1}):
        with pytest.raises(ptq2.PTQ2Error) as ei:
            ViscousSolver(ctx, nlevels=3, fp32=True, **bad)
        assert ei.value.code == _lib.ERR_INVALID_ARG


def test_rigid_deflation_speeds_up_high_contrast_solve():
    """Buoyancy problem at 56^3 with contrast 1.93e4 (20 spheres): the deflated solve (FP32
    smoothing, FP32 basis) reaches rtol 1e-8 on the restatement's FP64 residual in clearly
    fewer iterations than the plain one."""
    from ptq2.solver import auto_levels

    m = 56
    say("case: 56^3 buoyancy, contrast 1.93e4, deflated vs plain FGMRES(256)")
    P, b = M.config_buoyancy(m, contrast=10 ** (2 + 16 / 7), seed_spheres=104)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(m, m, m, P.coords, P.mask)
    ctx.set_viscosity(P.eta)
    bd = torch.from_numpy(b).to(DEV)
    its = {}
    for rigid in (True, False):
        S = ViscousSolver(ctx, nlevels=auto_levels(m), smooth_its=4, fp32=True, fp32_basis=True, emin_frac=0.05,
                          rigid_deflation=rigid)
        x, info = S.solve(bd, rtol=1e-8, max_it=1000, restart=256)
        say("  deflation %s: inclusions %s, iterations %d, rel %.3e" % (rigid, S.rigid_info(), info["iterations"],
                                                                         info["rel"]))
        its[rigid] = info["iterations"]
        r = b - O.matmult_a11(m, m, m, P.coords, P.eta, P.mask, x.cpu().numpy())
        assert np.linalg.norm(r) <= 1.05e-8 * np.linalg.norm(b)
        S.destroy()
    assert its[True] < 0.8 * its[False]


@pytest.mark.parametrize("fp32", [False, True])
def test_vcycle_non_cubic_partial_bricks(fp32):
    """V-cycle on a 12x8x20 jittered mesh with free-slip walls and log-uniform viscosity: level 1
    (6x4x10) has partial bricks in x and z, so the smoother's fused face pass meets bricks of
    every size; FP64 against the restatement at 1e-10, FP32 smoothing at 1e-4."""
    mx, my, mz = 12, 8, 20
    say("case: V-cycle %dx%dx%d, 3 levels, fp32=%s" % (mx, my, mz, fp32))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * mz), 5)
    mask = M.bc_mask(mx, my, mz, "freeslip")
    eta = 10.0 ** M.uniform(6, 27 * mx * my * mz, -1.0, 2.0)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    ctx.set_viscosity(eta)
    S = ViscousSolver(ctx, nlevels=3, smooth_its=3, fp32=fp32, use_graph=True, galerkin_coarse=False)
    assert [tuple(r) for r in S.levels()] == [(12, 8, 20), (6, 4, 10), (3, 2, 5)]
    emin, emax = S.eigs()
    lv = G.hierarchy(mx, my, mz, c, mask, eta, 3)
    V = G.VCycle(lv, emin, emax, 3)
    r = M.uniform(7, c.size, -1, 1)
    r[mask > 0] = 0.0
    z = S.vcycle(torch.from_numpy(r).to(DEV)).cpu().numpy()
    zr = V(r)
    err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    say("  rel err %.3e" % err)
    assert err <= (1e-4 if fp32 else 1e-10)
    S.destroy()
    ctx.destroy()


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("dims", [(16, 16, 16), (12, 8, 20)])
def test_fused_smoother_bitwise_equal(dims, fp32):
    """The Chebyshev step with the Jacobi scaling and three-term update fused into the apply's
    store (fused_smoother) gives the unfused smoother's V-cycle bit for bit: on a mesh of full
    tiles (16^3) and on one whose coarse level has partial bricks (12x8x20 -> 6x4x10), free-slip
    walls (Dirichlet dofs on every face), FP64 and FP32 smoothing, graph replay; and a solve
    takes the same iterations."""
    mx, my, mz = dims
    say("case: fused smoother %dx%dx%d fp32=%s" % (mx, my, mz, fp32))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 8)
    mask = M.bc_mask(mx, my, mz, "freeslip")
    eta = 10.0 ** M.uniform(9, 27 * mx * my * mz, -1.0, 2.0)
    r = M.uniform(10, c.size, -1, 1)
    r[mask > 0] = 0.0
    out = {}
    for fused in (False, True):
        ctx = ptq2.Context(0)
        ctx.setup_mesh(mx, my, mz, c, mask)
        ctx.set_viscosity(eta)
        S = ViscousSolver(ctx, nlevels=3, smooth_its=4, fp32=fp32, use_graph=True, galerkin_coarse=True,
                          fused_smoother=fused, autotune=0)
        rd = torch.from_numpy(r).to(DEV)
        z1 = S.vcycle(rd).clone()
        z2 = S.vcycle(rd).clone()
        x, info = S.solve(rd, rtol=1e-8, max_it=300, restart=60)
        torch.cuda.synchronize()
        out[fused] = (z1, z2, x.clone(), info["iterations"])
        S.destroy()
        ctx.destroy()
    assert torch.equal(out[True][0], out[False][0]) and torch.equal(out[True][1], out[False][1])
    assert torch.equal(out[True][0], out[True][1])
    assert out[True][3] == out[False][3] and torch.equal(out[True][2], out[False][2])
    say("  bitwise equal, %d iterations" % out[True][3])


@pytest.mark.parametrize("fp32", [False, True])
def test_autotuner_choices_and_level_times(fp32):
    """The set-up's autotuner: every smoothing level reports a valid launch shape with the
    measured time of the chosen and the default shape (chosen <= default); the row-split
    kernel's V-cycle agrees with the one-wave kernels' to rounding (FP64 1e-12, FP32 1e-5) on
    a mesh with partial bricks; the per-level V-cycle times are positive."""
    mx, my, mz = 24, 20, 28
    say("case: autotuner %dx%dx%d fp32=%s" % (mx, my, mz, fp32))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / 56, 12)
    mask = M.bc_mask(mx, my, mz, "freeslip")
    eta = 10.0 ** M.uniform(13, 27 * mx * my * mz, -1.0, 2.0)
    r = M.uniform(14, c.size, -1, 1)
    r[mask > 0] = 0.0
    rd = torch.from_numpy(r).to(DEV)
    zs = {}
    for at in ((0, 2, 3, 1) if fp32 else (0, 2, 1)):
        ctx = ptq2.Context(0)
        ctx.setup_mesh(mx, my, mz, c, mask)
        ctx.set_viscosity(eta)
        S = ViscousSolver(ctx, nlevels=3, smooth_its=4, fp32=fp32, galerkin_coarse=True, autotune=at)
        tune = S.tuning()
        say("  autotune %d: %s" % (at, tune))
        assert len(tune) == 2
        for kind, f, ms, md in tune:
            assert kind in ((0, 1, 2) if fp32 else (0, 1)) and f in (0, 1)
            if at in (2, 3):
                assert kind == at - 1
            if at == 1:
                assert 0 < ms <= md * 1.0000001
        zs[at] = S.vcycle(rd).clone()
        if at == 1:
            t = S.level_times(rd, reps=3)
            say("  level times %s ms" % np.round(t, 4))
            assert np.all(t > 0) and t.sum() < 1000.0
        S.destroy()
        ctx.destroy()
    tol = 1e-5 if fp32 else 1e-12
    for at in zs:
        err = (torch.linalg.norm(zs[at] - zs[0]) / torch.linalg.norm(zs[0])).item()
        say("  autotune %d vs one-wave kernels: rel %.2e" % (at, err))
        assert err <= tol


@pytest.mark.parametrize("dims,nl,gl,fp32,at", [((8, 8, 8), 3, 1, False, 0), ((12, 8, 16), 3, 1, False, 0),
                                                ((8, 8, 8), 4, 2, False, 0), ((8, 8, 8), 3, 1, True, 0),
                                                ((12, 8, 16), 3, 1, False, 4), ((8, 8, 8), 4, 2, True, 4)])
def test_galerkin_levels_vcycle_matches_restatement(dims, nl, gl, fp32, at):
    """Galerkin operators on the levels above the coarsest (ptq2_mg_opts.galerkin_levels),
    assembled on the GPU by coloured probing into 5x5x5 block stencils: the V-cycle against the
    restatement whose Galerkin levels are assembled column by column on the CPU (FP64 1e-10,
    FP32 smoothing 1e-4), on cubic and non-cubic meshes with partial bricks, through a
    matrix-free level and through a Galerkin level; autotune 4 applies the Galerkin level below
    the matrix-free one as R A P instead of its stencil."""
    mx, my, mz = dims
    say("case: Galerkin levels %dx%dx%d, %d levels, %d Galerkin, fp32=%s, autotune %d" % (mx, my, mz, nl, gl, fp32, at))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 21)
    mask = M.bc_mask(mx, my, mz, "freeslip")
    eta = 10.0 ** M.uniform(22, 27 * mx * my * mz, -1.0, 3.0)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    ctx.set_viscosity(eta)
    S = ViscousSolver(ctx, nlevels=nl, smooth_its=3, fp32=fp32, galerkin_coarse=True, galerkin_levels=gl, autotune=at)
    kinds = [k for k, _, _, _ in S.tuning()]
    assert kinds == [0] * (nl - 1 - gl) + ([4] + [3] * (gl - 1) if at == 4 else [3] * gl)
    emin, emax = S.eigs()
    lv = G.hierarchy(mx, my, mz, c, mask, eta, nl, galerkin_levels=gl)
    V = G.VCycle(lv, emin, emax, 3, galerkin=True)
    # the eigen-estimates of the Galerkin levels see their assembled operators
    for l in range(nl - 1 - gl, nl - 1):
        L = lv[l]
        v = M.uniform(23 + l, L.n, -1, 1)
        d = L.diag()
        for _ in range(60):
            w = L.matmult(v) / d
            lam = np.linalg.norm(w) / np.linalg.norm(v)
            v = w / np.linalg.norm(w)
        assert 0.8 * lam <= emax[l] / 1.1 <= 1.05 * lam, (l, lam, emax[l])
    r = M.uniform(24, c.size, -1, 1)
    r[mask > 0] = 0.0
    z = S.vcycle(torch.from_numpy(r).to(DEV)).cpu().numpy()
    zr = V(r)
    err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    say("  rel err %.3e" % err)
    assert err <= (1e-4 if fp32 else 1e-10)
    S.destroy()
    ctx.destroy()
