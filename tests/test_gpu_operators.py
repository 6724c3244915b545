"""GPU parity of libptq2's operators against the oracle (oracle/ptq2_oracle.c) and the
reference's goldens.  Bars: FP64 relative 2-norm <= 1e-12, FP32 <= 1e-5, Dirichlet rows
bit-exact."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
import ptq2
from ptq2 import _lib
from ptq2 import mesh as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = torch.device("cuda:0")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, dtype=np.float64) - b) / np.linalg.norm(b))


def say(*a):
    print(*a, flush=True)


def setup_a11(P):
    A = ptq2.MatA11MF("hip", device=0)
    A.setup(P)
    return A


@pytest.fixture(scope="module")
def c1():
    """16^3 jittered unit cube, eta = 1, no BCs (the isoviscous parity configuration)."""
    return M.config_jitter_iso(16)


@pytest.fixture(scope="module")
def sinker():
    """20^3 affine+jitter sinker, contrast 1e3 + log noise, free-slip/free-surface BCs."""
    return M.config_sinker(20, bc="freeslip_freesurface")


def test_a11_fp64_isoviscous_16(c1):
    say("case: A11 FP64 raw, 16^3 jittered, eta=1")
    A = setup_a11(c1)
    u = M.uniform(2, c1.nvel, -1.0, 1.0)
    ref = O.a11_tensor(16, 16, 16, c1.coords, c1.eta, u, threads=8)
    y = A.ctx.a11(u)
    assert rel(y, ref) <= 1e-12
    # device pointers give the same bits as the staged host path
    yd = A.ctx.a11(torch.from_numpy(u).to(DEV)).cpu().numpy()
    assert np.array_equal(yd, y)


def test_a11_fp64_sinker_matmult(sinker):
    P = sinker
    m = P.mx
    say("case: A11 FP64 MatMult, 20^3 sinker, free-slip/free-surface")
    A = setup_a11(P)
    u = M.uniform(3, P.nvel, -1.0, 1.0)
    ref = O.matmult_a11(m, m, m, P.coords, P.eta, P.mask, u)
    y = A.mult(u)
    assert rel(y, ref) <= 1e-12
    dmask = P.mask > 0
    assert np.array_equal(y[dmask], u[dmask])
    say("case: A11 FP32 MatMult, 20^3 sinker")
    y32 = A.mult(torch.from_numpy(u.astype(np.float32)).to(DEV), fp32=True).cpu().numpy()
    assert rel(y32, ref) <= 1e-5


@pytest.mark.parametrize("dims", [(20, 20, 20), (9, 4, 5), (1, 1, 3)])
def test_a11_fp32_mfma_form(dims):
    """The FP32 operator with the reference gradient as a dense GEMM on the matrix cores
    (PTQ2_MFMA) against the FP64 oracle (1e-5), MatMult and raw ADD, partial bricks included,
    next to the sum-factorized FP32 kernel; two applies bitwise equal."""
    mx, my, mz = dims
    say("case: A11 FP32 MFMA form %s" % (dims,))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 5)
    cen, r = M.sphere_inclusions(6, 0.1, 0.2, 9)
    eta = M.viscosity_inclusions(M.qp_coords(mx, my, mz, c), cen, r, 1e3, 1.0, 10)
    mask = M.bc_mask(mx, my, mz, "freeslip_freesurface")
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    ctx.set_viscosity(eta)
    n = 3 * M.nnodes(mx, my, mz)
    u = M.uniform(17, n, -1.0, 1.0)
    y0 = M.uniform(18, n, -1.0, 1.0)
    ref = O.matmult_a11(mx, my, mz, c, eta, mask, u)
    raw = y0 + O.a11_tensor(mx, my, mz, c, eta, u, threads=8)
    u32, y32 = u.astype(np.float32), y0.astype(np.float32)
    for fl in (0, _lib.MFMA):
        ym = ctx.a11(u32, flags=_lib.MATMULT | _lib.FP32 | fl)
        ya = ctx.a11(u32, y32.copy(), flags=_lib.ADD | _lib.FP32 | fl)
        say("  flags %#x: MatMult rel %.2e, raw ADD rel %.2e" % (fl, rel(ym, ref), rel(ya, raw)))
        assert rel(ym, ref) <= 1e-5 and rel(ya, raw) <= 1e-5
    assert np.array_equal(ctx.a11(u32, flags=_lib.FP32 | _lib.MFMA), ctx.a11(u32, flags=_lib.FP32 | _lib.MFMA))
    with pytest.raises(ptq2.PTQ2Error):
        ctx.a11(u, flags=_lib.MFMA)  # FP64 has no matrix-core form
    ctx.destroy()


@pytest.mark.parametrize("dims", [(20, 20, 20), (9, 4, 5), (1, 1, 3)])
def test_a11_fp64_geometry_variants(dims):
    """The on-the-fly kernel (default, PTQ2_ONTHEFLY) and the resident-geometry kernel
    (PTQ2_STOREDGEO) against the oracle, MatMult and raw ADD, including partial bricks."""
    mx, my, mz = dims
    say("case: A11 FP64 stored vs on-the-fly geometry %s" % (dims,))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 3)
    c = M.affine(c, M.SHEAR, [0.0, 0.0, 0.0])
    cen, r = M.sphere_inclusions(6, 0.1, 0.2, 7)
    eta = M.viscosity_inclusions(M.qp_coords(mx, my, mz, c), cen, r, 1e3, 1.0, 8)
    mask = M.bc_mask(mx, my, mz, "freeslip_freesurface")
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    ctx.set_viscosity(eta)
    n = 3 * M.nnodes(mx, my, mz)
    u = M.uniform(15, n, -1.0, 1.0)
    y0 = M.uniform(16, n, -1.0, 1.0)
    ref = O.matmult_a11(mx, my, mz, c, eta, mask, u)
    raw = y0 + O.a11_tensor(mx, my, mz, c, eta, u, threads=8)
    for fl in (0, _lib.ONTHEFLY, _lib.STOREDGEO):
        assert rel(ctx.a11(u, flags=_lib.MATMULT | fl), ref) <= 1e-12
        assert rel(ctx.a11(u, y0.copy(), flags=_lib.ADD | fl), raw) <= 1e-12


def test_a11_raw_add_mode(sinker):
    P = sinker
    m = P.mx
    say("case: A11 raw ADD (SpMVOp contract), 20^3")
    A = setup_a11(P)
    u = M.uniform(4, P.nvel, -1.0, 1.0)
    y0 = M.uniform(5, P.nvel, -1.0, 1.0)
    ref = y0 + O.a11_tensor(m, m, m, P.coords, P.eta, u, threads=8)
    y = A.spmv(u, y0.copy())
    assert rel(y, ref) <= 1e-12


def test_a11_deterministic(sinker):
    say("case: A11 FP64 bitwise determinism, 20^3")
    A = setup_a11(sinker)
    u = torch.from_numpy(M.uniform(6, sinker.nvel, -1.0, 1.0)).to(DEV)
    y1 = A.mult(u).clone()
    y2 = A.mult(u)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("dims", [(5, 7, 14), (1, 1, 3), (17, 13, 11), (4, 4, 4), (9, 4, 5)])
def test_partial_bricks_and_goldens(dims, goldens):
    """Meshes that do not divide into 4^3 bricks, on the reference's golden meshes; y.y of
    the GPU result against the reference's expected files where eta does not matter."""
    mx, my, mz = dims
    say("case: all operators, golden mesh %dx%dx%d" % dims)
    c = M.bump_ymax(mx, my, mz, M.q2_uniform(mx, my, mz))
    mask = M.bc_mask(mx, my, mz)
    eta = 10.0 ** M.uniform(7, 27 * mx * my * mz, -1.0, 1.0)
    P = M.Problem(mx, my, mz, c, mask, eta)
    A = setup_a11(P)
    S = ptq2.MatStokesMF(device=0)
    S.setup(P)
    x = M.test_vector_u(mx, my, mz, c)
    y = A.mult(x)
    assert rel(y, O.matmult_a11(mx, my, mz, c, eta, mask, x)) <= 1e-12
    y21 = S.mult_a21(x)
    ref21 = O.matmult_a21(mx, my, mz, c, mask, x)
    assert rel(y21, ref21) <= 1e-12
    p = M.test_vector_p(mx, my, mz, legacy=(mx > 1 and my > 1 and mz > 1))
    y12 = S.mult_a12(p)
    ref12 = O.matmult_a12(mx, my, mz, c, mask, p)
    assert rel(y12, ref12) <= 1e-12
    d = A.get_diagonal()
    assert rel(d, O.matgetdiag_a11(mx, my, mz, c, eta, mask)) <= 1e-12
    for g in goldens:
        if (g["mx"], g["my"], g["mz"]) == dims:
            assert abs(y21 @ y21 - g["yy"]["compare_mf_A21"]) <= 5e-9 * g["yy"]["compare_mf_A21"]
            if g["yy"]["compare_mf_A12"] is not None:
                assert abs(y12 @ y12 - g["yy"]["compare_mf_A12"]) <= 5e-9 * g["yy"]["compare_mf_A12"]


def test_golden_a11_diag_isoviscous_gpu(goldens):
    g = [g for g in goldens if g["eta1"] == 1.0][0]
    mx, my, mz = g["mx"], g["my"], g["mz"]
    say("case: isoviscous A11/diag goldens through the GPU, 8x9x11")
    c = M.bump_ymax(mx, my, mz, M.q2_uniform(mx, my, mz))
    P = M.Problem(mx, my, mz, c, M.bc_mask(mx, my, mz), np.ones(27 * mx * my * mz))
    A = setup_a11(P)
    y = A.mult(M.test_vector_u(mx, my, mz, c))
    assert abs(y @ y - g["yy"]["compare_mf_A11"]) <= 5e-9 * g["yy"]["compare_mf_A11"]
    d = A.get_diagonal()
    assert abs(d @ d - g["yy"]["compare_mf_diagA11"]) <= 5e-9 * g["yy"]["compare_mf_diagA11"]


def test_full_block(sinker):
    P = sinker
    m = P.mx
    say("case: full Stokes block MatMult, 20^3")
    S = ptq2.MatStokesMF(device=0)
    S.setup(P)
    u = M.uniform(8, P.nvel, -1.0, 1.0)
    p = M.uniform(9, 4 * P.nel, -1.0, 1.0)
    yu, yp = S.mult_a(u, p)
    ru, rp = O.matmult_a(m, m, m, P.coords, P.eta, P.mask, u, p)
    assert rel(yu, ru) <= 1e-12 and rel(yp, rp) <= 1e-12


@pytest.mark.parametrize("dims", [(12, 12, 12), (9, 6, 5)])
def test_full_block_fp64_one_pass(dims):
    """FP64 full block in one element pass against the oracle (MFStokesWrapper_A_AVX form) on a
    jittered contrast-1e6 sinker with partial bricks: MatMult 1e-12; raw ADD mode adds onto the
    output; the resident-geometry variant (three passes) agrees; repeated calls are bitwise
    identical."""
    mx, my, mz = dims
    say("case: FP64 full block %s contrast 1e6" % (dims,))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 2)
    cen, r = M.sphere_inclusions(8, 0.1, 0.2, 6)
    eta = M.viscosity_inclusions(M.qp_coords(mx, my, mz, c), cen, r, 1e6, 1.0, None)
    mask = M.bc_mask(mx, my, mz, "freeslip")
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    ctx.set_viscosity(eta)
    nvel, npres = 3 * M.nnodes(mx, my, mz), 4 * mx * my * mz
    u = M.uniform(11, nvel, -1.0, 1.0)
    p = M.uniform(12, npres, -1.0, 1.0)
    ru, rp = O.matmult_a(mx, my, mz, c, eta, mask, u, p)
    yu, yp = ctx.stokes(u, p, flags=_lib.MATMULT)
    assert rel(yu, ru) <= 1e-12 and rel(yp, rp) <= 1e-12
    yu2, yp2 = ctx.stokes(u, p, flags=_lib.MATMULT)
    assert np.array_equal(yu, yu2) and np.array_equal(yp, yp2)
    gu, gp = ctx.stokes(u, p, flags=_lib.MATMULT | _lib.STOREDGEO)
    assert rel(gu, ru) <= 1e-12 and rel(gp, rp) <= 1e-12
    zu, zp = O.stokes_a(mx, my, mz, c, eta, u, p)
    au = M.uniform(13, nvel, -1.0, 1.0)
    ap = M.uniform(14, npres, -1.0, 1.0)
    wu, wp = ctx.stokes(u, p, au.copy(), ap.copy(), flags=_lib.ADD)
    assert rel(wu, zu + au) <= 1e-12 and rel(wp, zp + ap) <= 1e-12
    ctx.destroy()


@pytest.mark.parametrize("dims", [(12, 12, 12), (9, 6, 5)])
def test_full_block_fp32(dims):
    """FP32 full block (one fused pass, FP32-stored factors) against the FP64 oracle on a
    sinker with viscosity contrast 1e6, MatMult semantics with free-slip walls; raw ADD mode
    adds onto the output; repeated calls are bitwise identical."""
    mx, my, mz = dims
    say("case: FP32 full block %s contrast 1e6" % (dims,))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 2)
    cen, r = M.sphere_inclusions(8, 0.1, 0.2, 6)
    eta = M.viscosity_inclusions(M.qp_coords(mx, my, mz, c), cen, r, 1e6, 1.0, None)
    mask = M.bc_mask(mx, my, mz, "freeslip")
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    ctx.set_viscosity(eta)
    nvel, npres = 3 * M.nnodes(mx, my, mz), 4 * mx * my * mz
    u = M.uniform(11, nvel, -1.0, 1.0)
    p = M.uniform(12, npres, -1.0, 1.0)
    ru, rp = O.matmult_a(mx, my, mz, c, eta, mask, u, p)
    yu, yp = ctx.stokes(u.astype(np.float32), p.astype(np.float32), flags=_lib.MATMULT | _lib.FP32)
    assert yu.dtype == np.float32
    assert rel(yu.astype(np.float64), ru) <= 1e-5 and rel(yp.astype(np.float64), rp) <= 1e-5
    yu2, yp2 = ctx.stokes(u.astype(np.float32), p.astype(np.float32), flags=_lib.MATMULT | _lib.FP32)
    assert np.array_equal(yu, yu2) and np.array_equal(yp, yp2)
    # raw mode adds the element sums (no BC steps) onto the output
    au = M.uniform(13, nvel, -1.0, 1.0).astype(np.float32)
    ap = M.uniform(14, npres, -1.0, 1.0).astype(np.float32)
    zu, zp = ctx.stokes(u.astype(np.float32), p.astype(np.float32), flags=_lib.FP32)
    wu, wp = ctx.stokes(u.astype(np.float32), p.astype(np.float32), au.copy(), ap.copy(), flags=_lib.ADD | _lib.FP32)
    assert np.abs(wu - (zu + au)).max() <= 1e-5 * np.abs(zu).max()
    assert np.abs(wp - (zp + ap)).max() <= 1e-5 * np.abs(zp).max()


def test_mesh_box_with_ghost_layers():
    """A rank-local ghosted node box: elements placed at an offset inside a larger box give
    the same element sums on their nodes and leave the other nodes untouched."""
    mx, my, mz = 6, 5, 7
    gx, gy, gz = 2, 1, 3  # ghost node offsets
    NX, NY, NZ = 2 * mx + 1 + 4, 2 * my + 1 + 2, 2 * mz + 1 + 5
    say("case: box with ghost layers")
    # build a box of NX x NY x NZ nodes on a uniform lattice, jitter interior
    xs = np.linspace(0, 1, NX)
    ys = np.linspace(0, 1, NY)
    zs = np.linspace(0, 1, NZ)
    box = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"), axis=-1).transpose(2, 1, 0, 3).copy()
    box[1:-1, 1:-1, 1:-1] += M.uniform(77, box[1:-1, 1:-1, 1:-1].size, -0.1, 0.1).reshape(
        box[1:-1, 1:-1, 1:-1].shape) * (xs[1] - xs[0])
    sub = box[gz:gz + 2 * mz + 1, gy:gy + 2 * my + 1, gx:gx + 2 * mx + 1].copy()
    eta = 10.0 ** M.uniform(78, 27 * mx * my * mz, -1, 1)
    u_box = M.uniform(79, box.size, -1, 1).reshape(box.shape)
    ctx = ptq2.Context(0)
    L = _lib.lib()
    bc = np.ascontiguousarray(box.reshape(-1))
    _lib.check(L.ptq2_setup_mesh_box(ctx.handle, mx, my, mz, NX, NY, NZ, gx, gy, gz, bc.ctypes.data, None),
               "setup_mesh_box", ctx.handle)
    ctx.nvel = box.size
    ctx.npres = 4 * mx * my * mz
    ctx.set_viscosity(eta)
    y_box = M.uniform(80, box.size, -1, 1)
    y_in = y_box.copy()
    ctx.a11(u_box.reshape(-1), y_box, _lib.ADD)
    ref_sub = O.a11_tensor(mx, my, mz, sub.reshape(-1), eta,
                           u_box[gz:gz + 2 * mz + 1, gy:gy + 2 * my + 1, gx:gx + 2 * mx + 1].reshape(-1))
    yb = y_box.reshape(box.shape)
    yi = y_in.reshape(box.shape)
    got = yb[gz:gz + 2 * mz + 1, gy:gy + 2 * my + 1, gx:gx + 2 * mx + 1] - \
        yi[gz:gz + 2 * mz + 1, gy:gy + 2 * my + 1, gx:gx + 2 * mx + 1]
    assert rel(got.reshape(-1), ref_sub) <= 1e-12
    outside = np.ones(box.shape, dtype=bool)
    outside[gz:gz + 2 * mz + 1, gy:gy + 2 * my + 1, gx:gx + 2 * mx + 1] = False
    assert np.array_equal(yb[outside], yi[outside])


def test_error_paths():
    say("case: error paths")
    m = 4
    P = M.config_jitter_iso(m)
    ctx = ptq2.Context(0)
    L = _lib.lib()
    u = M.uniform(1, P.nvel)
    y = np.zeros(P.nvel)
    # not set up
    assert L.ptq2_a11_apply(ctx.handle, u.ctypes.data, y.ctypes.data, 0) == _lib.ERR_NOT_SETUP
    ctx.setup_mesh(m, m, m, P.coords)
    assert L.ptq2_a11_apply(ctx.handle, u.ctypes.data, y.ctypes.data, 0) == _lib.ERR_NOT_SETUP
    # eta <= 0
    bad = P.eta.copy()
    bad[100] = -1.0
    assert L.ptq2_set_viscosity(ctx.handle, bad.ctypes.data, 1, 0) == _lib.ERR_VISCOSITY
    bad[100] = np.nan
    assert L.ptq2_set_viscosity(ctx.handle, bad.ctypes.data, 2, 0) == _lib.ERR_VISCOSITY
    ctx.set_viscosity(P.eta, 3)
    # NaN in the input vector (host path)
    un = u.copy()
    un[17] = np.inf
    assert L.ptq2_a11_apply(ctx.handle, un.ctypes.data, y.ctypes.data, 0) == _lib.ERR_NONFINITE
    assert L.ptq2_a11_apply(ctx.handle, u.ctypes.data, y.ctypes.data, 0) == _lib.OK
    # bad flags
    assert L.ptq2_a11_apply(ctx.handle, u.ctypes.data, y.ctypes.data, 0x100) == _lib.ERR_INVALID_ARG
    assert L.ptq2_a11_apply(ctx.handle, u.ctypes.data, y.ctypes.data, _lib.ADD | _lib.MATMULT) == _lib.ERR_INVALID_ARG
    assert L.ptq2_a11_apply(ctx.handle, u.ctypes.data, y.ctypes.data, _lib.ONTHEFLY | _lib.STOREDGEO) == _lib.ERR_INVALID_ARG
    # inverted element: swap two node planes of one element
    c = P.coords.reshape(2 * m + 1, 2 * m + 1, 2 * m + 1, 3).copy()
    c[2, 2, 2, 0], c[2, 2, 4, 0] = c[2, 2, 4, 0], c[2, 2, 2, 0]
    c[2, 2, 3, 0] = c[2, 2, 2, 0] + 0.9 * (c[2, 2, 4, 0] - c[2, 2, 2, 0]) * 3
    cc = np.ascontiguousarray(c.reshape(-1))
    assert L.ptq2_setup_mesh(ctx.handle, m, m, m, cc.ctypes.data, None, None) == _lib.ERR_MESH
    # non-finite coordinate
    cc = P.coords.copy()
    cc[5] = np.nan
    assert L.ptq2_setup_mesh(ctx.handle, m, m, m, cc.ctypes.data, None, None) == _lib.ERR_NONFINITE
    # connectivity that is not the structured map
    el = O.elnidx(m, m, m)
    el[[0, 1]] = el[[1, 0]]
    assert L.ptq2_setup_mesh(ctx.handle, m, m, m, P.coords.ctypes.data, el.ctypes.data, None) == _lib.ERR_UNSUPPORTED
    el = O.elnidx(m, m, m)
    assert L.ptq2_setup_mesh(ctx.handle, m, m, m, P.coords.ctypes.data, el.ctypes.data, None) == _lib.OK
    # a mesh above the supported maximum is refused before any input is read
    big = 2048
    rc = L.ptq2_setup_mesh_box(ctx.handle, big, big, big, 2 * big + 1, 2 * big + 1, 2 * big + 1, 0, 0, 0,
                               P.coords.ctypes.data, None)
    assert rc == _lib.ERR_SIZE


def test_full_size_properties_128():
    """128^3 (51M velocity dofs, sinker contrast 1e6): properties that hold at any size.
    FP32 vs FP64, symmetry, rigid-body null space, exact energy of a linear field."""
    m = 128
    say("case: 128^3 properties, host set-up")
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 1)
    cen, r = M.sphere_inclusions(20, 0.05, 0.15, 5)
    xq = M.qp_coords(m, m, m, c)
    eta = M.viscosity_inclusions(xq, cen, r, 1e6, 1.0)
    del xq
    P = M.Problem(m, m, m, c, None, eta)
    A = setup_a11(P)
    n = P.nvel
    say("case: 128^3 properties, applies")
    g = torch.Generator(device=DEV)
    g.manual_seed(5)
    x = torch.rand(n, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    z = torch.rand(n, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    Ax = A.ctx.a11(x)
    Az = A.ctx.a11(z)
    s1, s2 = float(z @ Ax), float(x @ Az)
    assert abs(s1 - s2) <= 1e-11 * float(torch.linalg.norm(Ax) * torch.linalg.norm(z))
    Ax32 = A.ctx.a11(x.float(), flags=_lib.FP32)
    assert float(torch.linalg.norm(Ax32.double() - Ax) / torch.linalg.norm(Ax)) <= 1e-5
    # full block at full size (contrast 1e6): the fused FP32 pass against the FP64 passes
    say("case: 128^3 full block FP32 vs FP64")
    pv = torch.rand(4 * m ** 3, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    yu, yp = A.ctx.stokes(x, pv)
    yu32, yp32 = A.ctx.stokes(x.float(), pv.float(), flags=_lib.FP32)
    assert float(torch.linalg.norm(yu32.double() - yu) / torch.linalg.norm(yu)) <= 1e-5
    assert float(torch.linalg.norm(yp32.double() - yp) / torch.linalg.norm(yp)) <= 1e-5
    del yu, yp, yu32, yp32, pv
    cx = torch.from_numpy(c).to(DEV).view(-1, 3)
    rot = torch.stack([-cx[:, 1], cx[:, 0], torch.zeros_like(cx[:, 0])], dim=1).reshape(-1).contiguous()
    assert float(torch.linalg.norm(A.ctx.a11(rot))) <= 1e-11 * float(torch.linalg.norm(Ax))
    lin = torch.stack([cx[:, 0], torch.zeros_like(cx[:, 0]), torch.zeros_like(cx[:, 0])], dim=1).reshape(-1).contiguous()
    A.ctx.set_viscosity(np.ones(27 * m ** 3), 99)
    e = float(lin @ A.ctx.a11(lin))
    assert abs(e - 2.0) <= 1e-10


@pytest.mark.parametrize("dims", [(8, 9, 11), (16, 16, 16), (5, 7, 14), (1, 1, 3), (4, 12, 1)])
def test_body_force_matches_oracle(dims):
    """ptq2_body_force against the oracle's Be loop (ptq2o_body_force, per-point form of
    FormFunctionLocal_U; its pins are in tests/test_oracle.py and tests/test_host.py) on jittered
    meshes with partial bricks: FP64 1e-12; host and device pointers bitwise equal; ADD
    accumulates; MATMULT zeroes the Dirichlet rows; a non-finite input is reported."""
    mx, my, mz = dims
    say("case: body force %dx%dx%d" % dims)
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 3)
    mask = M.bc_mask(mx, my, mz, "freeslip")
    fu = M.uniform(31, 81 * mx * my * mz, -1, 1).reshape(-1, 27, 3)
    ref = O.body_force(mx, my, mz, c, fu)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    y = ctx.body_force(fu)
    assert rel(y, ref) <= 1e-12
    yd = ctx.body_force(torch.from_numpy(fu).to(DEV))
    torch.cuda.synchronize()
    assert np.array_equal(yd.cpu().numpy(), y)
    y0 = M.uniform(32, ref.size, -1, 1)
    ya = ctx.body_force(fu, y0.copy(), _lib.ADD)
    assert rel(ya, y0 + ref) <= 1e-12
    ym = ctx.body_force(fu, None, _lib.MATMULT)
    assert np.all(ym[mask > 0] == 0.0) and np.array_equal(ym[mask == 0], y[mask == 0])
    bad = fu.copy()
    bad[min(3, bad.shape[0] - 1), 4, 1] = np.nan
    with pytest.raises(ptq2.PTQ2Error):
        ctx.body_force(bad)
    ctx.destroy()


@pytest.mark.parametrize("dims,with_fp", [((8, 9, 11), True), ((16, 16, 16), False), ((5, 7, 14), True)])
def test_stokes_residual_matches_oracle(dims, with_fp):
    """ptq2_stokes_residual (FormFunctionLocal_U / _P: A11 u_loc + A12 p - Be, A21 u_loc - int N^p
    fp, Dirichlet rows u - phi) against the oracle's residual on jittered meshes with partial
    bricks, nonzero Dirichlet values and a variable viscosity: FP64 1e-12 relative in each part;
    host and device paths bitwise equal."""
    mx, my, mz = dims
    say("case: residual %dx%dx%d fp=%s" % (mx, my, mz, with_fp))
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(dims)), 7)
    mask = M.bc_mask(mx, my, mz, "freeslip_freesurface")
    ne = mx * my * mz
    eta = 10.0 ** M.uniform(60, 27 * ne, -1, 2)
    nv = 3 * M.nnodes(mx, my, mz)
    u = M.uniform(61, nv, -1, 1)
    p = M.uniform(62, 4 * ne, -1, 1)
    fu = M.uniform(63, 81 * ne, -1, 1)
    fp = M.uniform(64, 27 * ne, -1, 1) if with_fp else None
    phi = M.uniform(65, nv, -0.5, 0.5)
    ref_u, ref_p = O.residual(mx, my, mz, c, eta, mask, u, p, fu, fp, phi)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, mask)
    ctx.set_viscosity(eta)
    Fu, Fp = ctx.residual(u, p, fu, fp, phi)
    eu, ep = rel(Fu, ref_u), rel(Fp, ref_p)
    say("  Fu rel err %.3e, Fp rel err %.3e" % (eu, ep))
    assert eu <= 1e-12 and ep <= 1e-12
    d = mask > 0
    assert np.array_equal(Fu[d], u[d] - phi[d])
    t = [None if a is None else torch.from_numpy(a).to(DEV) for a in (u, p, fu, fp, phi)]
    Fud, Fpd = ctx.residual(*t)
    torch.cuda.synchronize()
    assert np.array_equal(Fud.cpu().numpy(), Fu) and np.array_equal(Fpd.cpu().numpy(), Fp)
    # zero Dirichlet values (phi = None) and no sources: the block MatMult on u with zeroed
    # Dirichlet columns, except the Dirichlet rows (u_D here)
    Fu0, Fp0 = ctx.residual(u, p, np.zeros(81 * ne))
    yu, yp = O.matmult_a(mx, my, mz, c, eta, mask, u, p)
    assert rel(Fu0[~d], yu[~d]) <= 1e-12 and rel(Fp0, yp) <= 1e-12
    ctx.destroy()
