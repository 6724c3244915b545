"""Material points -> quadrature points on the device (ptq2_project_points_q1) against the oracle's
restatement of SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes, and through the
operators against the reference's eta1 = 10 sinker goldens (A11 and diagonal y.y)."""
import numpy as np
import pytest

from oracle import oracle as O
import ptq2
from ptq2 import _lib
from ptq2 import mesh as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = torch.device("cuda:0")


def say(*a):
    print(*a, flush=True)


def _rel(a, b):
    return abs(a - b) / abs(b)


def test_points_reproduce_eta1_goldens(goldens):
    """Points on the zero-perturbation 2x2x2 lattice of the bump mesh with the single-sphere rule,
    projected on the GPU: the quadrature values equal the oracle's to rounding, and the GPU A11
    MatMult and diagonal of the reference test vector reproduce the y.y of every eta1 = 10
    expected file (7 meshes, 11 files) to the 9 printed digits."""
    seen = 0
    for g in goldens:
        if g["eta1"] != 10.0:
            continue
        mx, my, mz = g["mx"], g["my"], g["mz"]
        say("case: eta1=10 golden %dx%dx%d (%s)" % (mx, my, mz, g["source"]))
        c = O.mesh_bump(mx, my, mz, O.mesh_uniform(mx, my, mz))
        xi, wil, x = M.point_lattice(mx, my, mz, c)
        eta_p, rho_p = M.sinker_single_inclusion(x, eta1=10.0)
        ref_eta, ref_rho = O.project_q1(mx, my, mz, xi, wil, eta_p, rho_p)
        ctx = ptq2.Context(0)
        mask = O.bc_mask(mx, my, mz, 0)
        ctx.setup_mesh(mx, my, mz, c, mask)
        eta, rho = ctx.project_points(xi, wil, eta_p, rho_p)
        assert np.max(np.abs(eta - ref_eta)) <= 1e-14 * np.max(np.abs(ref_eta))
        assert np.max(np.abs(rho - ref_rho)) <= 1e-14 * max(1.0, np.max(np.abs(ref_rho)))
        ctx.set_viscosity(eta)
        u = O.test_vector_u(mx, my, mz, c)
        y = ctx.a11(u, flags=ptq2.MATMULT)
        d = ctx.a11_diag(flags=ptq2.MATMULT)
        say("  A11 y.y rel %.2e, diag y.y rel %.2e" % (_rel(y @ y, g["yy"]["compare_mf_A11"]),
                                                      _rel(d @ d, g["yy"]["compare_mf_diagA11"])))
        assert _rel(y @ y, g["yy"]["compare_mf_A11"]) < 5e-9, g["source"]
        assert _rel(d @ d, g["yy"]["compare_mf_diagA11"]) < 5e-9, g["source"]
        ctx.destroy()
        seen += 1
    assert seen == 11


def test_points_device_path_shuffled_and_reproducible():
    """Points in random order with random local coordinates (3x3x3 per element, about 5% of them
    dropped) on a jittered 12x10x9 mesh, device pointers: equal to the oracle to rounding, two
    runs bitwise equal, host and device paths bitwise equal."""
    mx, my, mz = 12, 10, 9
    say("case: points, shuffled device path %dx%dx%d" % (mx, my, mz))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / 24, 5)
    xi, wil, _ = M.point_lattice(mx, my, mz, c, n=(3, 3, 3))
    xi = np.clip(xi + M.uniform(41, xi.size, -0.3, 0.3).reshape(xi.shape), -1.0, 1.0)
    keep = M.uniform(42, wil.size) > 0.05  # drop ~5% of the points
    perm = np.argsort(M.uniform(43, int(keep.sum())))
    xi, wil = xi[keep][perm], wil[keep][perm]
    eta_p = 10.0 ** M.uniform(44, wil.size, -2, 3)
    rho_p = M.uniform(45, wil.size, -1, 1)
    ref_eta, ref_rho = O.project_q1(mx, my, mz, xi, wil, eta_p, rho_p)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c)
    args = [torch.from_numpy(np.ascontiguousarray(a)).to(DEV) for a in (xi, wil.astype(np.int32), eta_p, rho_p)]
    e1, r1 = ctx.project_points(*args)
    e2, r2 = ctx.project_points(*args)
    torch.cuda.synchronize()
    assert torch.equal(e1, e2) and torch.equal(r1, r2)
    eh, rh = ctx.project_points(xi, wil, eta_p, rho_p)
    assert np.array_equal(eh, e1.cpu().numpy()) and np.array_equal(rh, r1.cpu().numpy())
    assert np.max(np.abs(eh - ref_eta) / ref_eta) <= 1e-13
    assert np.max(np.abs(rh - ref_rho)) <= 1e-14
    ctx.destroy()


def test_points_errors():
    """An element index out of range is PTQ2_ERR_INVALID_ARG, a NaN value PTQ2_ERR_NONFINITE, a
    box with a ghost layer PTQ2_ERR_UNSUPPORTED; nothing is written on error."""
    m = 3
    say("case: points, error paths")
    c = M.q2_uniform(m, m, m)
    xi, wil, _ = M.point_lattice(m, m, m, c)
    eta_p = np.ones(wil.size)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(m, m, m, c)
    bad = wil.copy()
    bad[5] = m ** 3
    with pytest.raises(ptq2.PTQ2Error) as ei:
        ctx.project_points(xi, bad, eta_p)
    assert ei.value.code == _lib.ERR_INVALID_ARG
    nan = eta_p.copy()
    nan[7] = np.nan
    with pytest.raises(ptq2.PTQ2Error) as ei:
        ctx.project_points(xi, wil, nan)
    assert ei.value.code == _lib.ERR_NONFINITE
    ctx.destroy()
    # a rank's ghosted box (one extra node layer on every side)
    NX = 2 * m + 3
    cg = M.q2_uniform(m + 1, m + 1, m + 1)
    ctx = ptq2.Context(0)
    L = _lib.lib()
    _lib.check(L.ptq2_setup_mesh_box(ctx.handle, m, m, m, NX, NX, NX, 1, 1, 1, cg.ctypes.data, None),
               "ptq2_setup_mesh_box", ctx.handle)
    ctx.nel = m ** 3
    with pytest.raises(ptq2.PTQ2Error) as ei:
        ctx.project_points(xi, wil, eta_p)
    assert ei.value.code == _lib.ERR_UNSUPPORTED
    ctx.destroy()
