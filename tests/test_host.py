"""Host-side checks that need no GPU: the C ABI library loads and exports every entry
point of include/ptq2.h, its error paths return codes (never abort), and the host-side
mesh/BC/test-vector generators agree with the oracle's restatement."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
import ptq2
from ptq2 import _lib
from ptq2 import mesh as M

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(header="ptq2.h"):
    txt = open(os.path.join(REPO, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ptq2_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    from ptq2 import solver as Sv

    L = _lib.lib()
    headers = sorted(f for f in os.listdir(os.path.join(REPO, "include")) if f.endswith(".h"))
    assert headers == ["ptq2.h", "ptq2_solver.h"]
    names = header_functions("ptq2.h")
    snames = header_functions("ptq2_solver.h")
    assert len(names) >= 15 and len(snames) >= 10
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ptq2_\w+)", out))
    for n in names + snames:
        assert hasattr(L, n), n
        assert n in exported, n
    # the ctypes tables cover exactly the headers
    assert sorted(_lib.SIGNATURES) == names
    assert sorted(Sv._SOLVER_SIGS) == snames


def test_strerror_and_null_arguments():
    L = _lib.lib()
    for code in range(0, 10):
        assert L.ptq2_strerror(code)
    assert L.ptq2_create(None, 0) == _lib.ERR_INVALID_ARG
    assert L.ptq2_destroy(None) == _lib.OK
    assert L.ptq2_sync(None) == _lib.ERR_INVALID_ARG
    assert L.ptq2_a11_apply(None, None, None, 0) == _lib.ERR_INVALID_ARG
    assert L.ptq2_set_viscosity(None, None, 0, 0) == _lib.ERR_INVALID_ARG
    assert L.ptq2_setup_mesh(None, 1, 1, 1, None, None, None) == _lib.ERR_INVALID_ARG
    assert L.ptq2_a11_apply_batch(None, 1, None, None, 1) == _lib.ERR_INVALID_ARG


def test_solver_default_options_round_trip():
    """ptq2_mg_default_opts fills the ctypes mirror of ptq2_mg_opts field by field: the
    documented defaults read back in place (a layout mismatch would scramble them)."""
    from ptq2 import solver as Sv

    o = Sv.default_opts()
    assert (o.nlevels, o.smooth_its, o.fp32_smoother, o.esteig_its, o.use_graph) == (5, 4, 1, 20, 1)
    assert (o.emin_frac, o.emax_frac) == (0.2, 1.1)
    assert (o.coarse_max_dofs, o.galerkin_coarse, o.coarse_eta_mean) == (20000, 1, Sv.ETA_MEANS["arithmetic"])
    assert (o.fp32_basis, o.rigid_deflation, o.rigid_min_contrast, o.rigid_max_inclusions) == (0, 0, 100.0, 128)
    assert (o.fused_smoother, o.autotune) == (0, 1)
    assert sorted(Sv.ETA_MEANS.values()) == [0, 1, 2, 3, 4]
    txt = open(os.path.join(REPO, "include", "ptq2_solver.h")).read()
    for name, v in Sv.ETA_MEANS.items():
        assert re.search(r"PTQ2_ETA_MEAN_%s\s*=\s*%d" % (name.upper(), v), txt), name


def test_create_without_gpu_returns_code():
    """On a host without a GPU (this container) creating a context reports an error code."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    h = ctypes.c_void_p()
    rc = _lib.lib().ptq2_create(ctypes.byref(h), 0)
    assert rc == _lib.ERR_DEVICE
    with pytest.raises(ptq2.PTQ2Error):
        ptq2.Context(0)


def test_unknown_a11_op_raises():
    with pytest.raises(ValueError):
        ptq2.MatA11MF("avx")


def test_mesh_generators_match_oracle():
    for mx, my, mz in [(8, 9, 11), (1, 1, 3), (5, 7, 14)]:
        c = M.q2_uniform(mx, my, mz)
        assert np.array_equal(c, O.mesh_uniform(mx, my, mz))
        b = M.bump_ymax(mx, my, mz, c)
        assert np.array_equal(b, O.mesh_bump(mx, my, mz, c))
        for kind, k in (("freeslip_freesurface", 0), ("freeslip", 1), ("noslip", 2)):
            assert np.array_equal(M.bc_mask(mx, my, mz, kind), O.bc_mask(mx, my, mz, k))
        assert np.array_equal(M.test_vector_u(mx, my, mz, b), O.test_vector_u(mx, my, mz, b))
        with np.errstate(all="ignore"):
            assert np.allclose(M.test_vector_p(mx, my, mz), O.test_vector_p(mx, my, mz), equal_nan=True)
        assert np.array_equal(M.test_vector_p(mx, my, mz, legacy=False), O.test_vector_p(mx, my, mz, legacy=False))


def test_splitmix_reference_values():
    # first outputs of SplitMix64 with seed 0 (published test vector of the generator)
    z = M.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = M.uniform(7, 10000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.02


def test_qp_coords_reproduce_affine_map():
    m = 3
    A = M.SHEAR
    c = M.affine(M.q2_uniform(m, m, m), A, [0.5, 0.0, 1.0])
    xq = M.qp_coords(m, m, m, c)
    # on an affine mesh the QP of element (ex,ey,ez) sits at A((2e+1+xi)/(2m)) + b
    ref = M.affine(M.qp_coords(m, m, m, M.q2_uniform(m, m, m)).reshape(-1), A, [0.5, 0.0, 1.0])
    assert np.abs(xq.reshape(-1) - ref).max() < 1e-14
    g = M.GAUSS_X
    assert np.allclose(M.qp_coords(m, m, m, M.q2_uniform(m, m, m))[0, 0], (1 + g[0]) / (2 * m))


def _body_force_brute(m, coords, fu):
    """Element-by-element, point-by-point evaluation of sum_q w_q detJ_q N_n(x_q) fu_q."""
    B, D = M.basis_1d()
    n1 = 2 * m + 1
    g = coords.reshape(n1, n1, n1, 3)
    b = np.zeros((n1, n1, n1, 3))
    for ez in range(m):
        for ey in range(m):
            for ex in range(m):
                e = ex + m * (ey + m * ez)
                X = g[2 * ez:2 * ez + 3, 2 * ey:2 * ey + 3, 2 * ex:2 * ex + 3]  # (k, j, i, 3)
                for c in range(3):
                    for bq in range(3):
                        for a in range(3):
                            N = np.einsum("k,j,i->kji", B[c], B[bq], B[a])
                            J = np.stack([np.einsum("k,j,i,kjid->d", B[c], B[bq], D[a], X),
                                          np.einsum("k,j,i,kjid->d", B[c], D[bq], B[a], X),
                                          np.einsum("k,j,i,kjid->d", D[c], B[bq], B[a], X)])
                            w = M.GAUSS_W[a] * M.GAUSS_W[bq] * M.GAUSS_W[c]
                            q = a + 3 * bq + 9 * c
                            b[2 * ez:2 * ez + 3, 2 * ey:2 * ey + 3, 2 * ex:2 * ex + 3] += \
                                (w * np.linalg.det(J)) * N[..., None] * fu[e, q][None, None, None, :]
    return b.reshape(-1)


def test_body_force_matches_pointwise_assembly():
    m = 2
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 7)
    c = M.affine(c, M.SHEAR, [0.1, 0.0, -0.2])
    fu = M.uniform(8, m ** 3 * 27 * 3, -1, 1).reshape(-1, 27, 3)
    b = M.body_force(m, m, m, c, fu)
    ref = _body_force_brute(m, c, fu)
    assert np.abs(b - ref).max() <= 1e-14 * np.abs(ref).max()


def test_body_force_integrates_linear_functions():
    # sum_n b_n x_n = int fu . x dV exactly, since the Q2 basis reproduces x
    m = 3
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 9)
    E = m ** 3
    fu = M.uniform(10, E * 27 * 3, 0.5, 1.5).reshape(E, 27, 3)
    b = M.body_force(m, m, m, c, fu).reshape(-1, 3)
    W = (M.GAUSS_W[None, None, :] * M.GAUSS_W[None, :, None] * M.GAUSS_W[:, None, None]).reshape(27)
    wdet = W[None, :] * M.qp_jacobian_det(m, m, m, c)
    xq = M.qp_coords(m, m, m, c)
    X = c.reshape(-1, 3)
    for d in range(3):
        lhs = (b[:, d] * X[:, 0]).sum()
        rhs = (wdet * fu[:, :, d] * xq[:, :, 0]).sum()
        assert lhs == pytest.approx(rhs, rel=1e-13)
    assert wdet.sum() == pytest.approx(1.0, rel=1e-13)  # jitter keeps the boundary: volume 1


@pytest.mark.parametrize("m", [6, 10])
def test_buoyancy_config_chunked_matches_direct(m):
    P, b = M.config_buoyancy(m, threads=3)
    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 1)
    cen, r = M.sphere_inclusions(20, 0.05, 0.15, 5)
    xq = M.qp_coords(m, m, m, c)
    eta = M.viscosity_inclusions(xq, cen, r, 1e4, 1.0, None)
    rho = M.viscosity_inclusions(xq, cen, r, 1.1, 1.0, None).reshape(-1, 27)
    fu = np.zeros(rho.shape + (3,))
    fu[:, :, 1] = -rho
    ref = M.body_force(m, m, m, c, fu)
    ref[P.mask > 0] = 0.0
    assert np.array_equal(P.eta, eta)
    assert np.abs(b - ref).max() <= 1e-15 * np.abs(ref).max()
    assert np.all(b[P.mask > 0] == 0.0)
