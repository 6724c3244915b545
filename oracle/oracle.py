"""ctypes binding of the CPU oracle (liboracle built from ptq2_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only there as the checker.  The product
library never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libptq2_oracle.so")
_lib = None

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i = ctypes.c_int
_d = ctypes.c_double


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        sig = {
            "ptq2o_mesh_uniform": [_i, _i, _i, _d, _d, _d, _d, _d, _d, _D],
            "ptq2o_mesh_bump_ymax": [_i, _i, _i, _d, _d, _D],
            "ptq2o_elnidx": [_i, _i, _i, _I],
            "ptq2o_bc_mask": [_i, _i, _i, _i, _U8],
            "ptq2o_test_vector_u": [_i, _i, _i, _D, _D],
            "ptq2o_test_vector_p": [_i, _i, _i, _i, _D],
            "ptq2o_a11_tensor": [_i, _i, _i, _D, _D, _D, _D],
            "ptq2o_a11_tensor_omp": [_i, _i, _i, _D, _D, _D, _D, _i],
            "ptq2o_a11_ref": [_i, _i, _i, _D, _D, _D, _D],
            "ptq2o_a12": [_i, _i, _i, _D, _D, _D],
            "ptq2o_a21": [_i, _i, _i, _D, _D, _D],
            "ptq2o_diag_a11": [_i, _i, _i, _D, _D, _D],
            "ptq2o_stokes_a": [_i, _i, _i, _D, _D, _D, _D, _D, _D],
            "ptq2o_matmult_a11": [_i, _i, _i, _D, _D, _U8, _D, _D],
            "ptq2o_matmult_a12": [_i, _i, _i, _D, _U8, _D, _D],
            "ptq2o_matmult_a21": [_i, _i, _i, _D, _U8, _D, _D],
            "ptq2o_matgetdiag_a11": [_i, _i, _i, _D, _D, _U8, _D],
            "ptq2o_matmult_a": [_i, _i, _i, _D, _D, _U8, _D, _D, _D, _D],
            "ptq2o_a11_element_matrix": [_D, _D, _D],
            "ptq2o_residual": [_i, _i, _i, _D, _D, _U8, ctypes.c_void_p, _D, _D, _D, ctypes.c_void_p, _D, _D],
            "ptq2o_body_force": [_i, _i, _i, _D, _D, _D],
            "ptq2o_marker_lattice": [_i, _i, _i, _D, _i, _i, _i, _D, _I, _D],
            "ptq2o_project_q1": [_i, _i, _i, ctypes.c_int64, _D, _I, _D, ctypes.c_void_p, _D,
                                 ctypes.c_void_p],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        _lib = L
    return _lib


def nnodes(mx, my, mz):
    return (2 * mx + 1) * (2 * my + 1) * (2 * mz + 1)


def mesh_uniform(mx, my, mz, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    c = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_mesh_uniform(mx, my, mz, *box, c)
    return c


def mesh_bump(mx, my, mz, coords, amp=-0.3, lam=-5.6):
    c = np.array(coords, dtype=np.float64, copy=True)
    lib().ptq2o_mesh_bump_ymax(mx, my, mz, amp, lam, c)
    return c


def elnidx(mx, my, mz):
    e = np.zeros(27 * mx * my * mz, dtype=np.int32)
    lib().ptq2o_elnidx(mx, my, mz, e)
    return e


def bc_mask(mx, my, mz, kind=0):
    m = np.zeros(3 * nnodes(mx, my, mz), dtype=np.uint8)
    lib().ptq2o_bc_mask(mx, my, mz, kind, m)
    return m


def test_vector_u(mx, my, mz, coords):
    u = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_test_vector_u(mx, my, mz, np.ascontiguousarray(coords), u)
    return u


def test_vector_p(mx, my, mz, legacy=True):
    p = np.zeros(4 * mx * my * mz)
    lib().ptq2o_test_vector_p(mx, my, mz, 1 if legacy else 0, p)
    return p


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def a11_tensor(mx, my, mz, coords, eta, u, threads=0):
    y = np.zeros(3 * nnodes(mx, my, mz))
    if threads:
        lib().ptq2o_a11_tensor_omp(mx, my, mz, _f64(coords), _f64(eta), _f64(u), y, threads)
    else:
        lib().ptq2o_a11_tensor(mx, my, mz, _f64(coords), _f64(eta), _f64(u), y)
    return y


def a11_ref(mx, my, mz, coords, eta, u):
    y = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_a11_ref(mx, my, mz, _f64(coords), _f64(eta), _f64(u), y)
    return y


def a12(mx, my, mz, coords, p):
    y = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_a12(mx, my, mz, _f64(coords), _f64(p), y)
    return y


def a21(mx, my, mz, coords, u):
    y = np.zeros(4 * mx * my * mz)
    lib().ptq2o_a21(mx, my, mz, _f64(coords), _f64(u), y)
    return y


def diag_a11(mx, my, mz, coords, eta):
    d = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_diag_a11(mx, my, mz, _f64(coords), _f64(eta), d)
    return d


def stokes_a(mx, my, mz, coords, eta, u, p):
    yu = np.zeros(3 * nnodes(mx, my, mz))
    yp = np.zeros(4 * mx * my * mz)
    lib().ptq2o_stokes_a(mx, my, mz, _f64(coords), _f64(eta), _f64(u), _f64(p), yu, yp)
    return yu, yp


def _mask(mask, n):
    if mask is None:
        return np.zeros(n, dtype=np.uint8)
    return np.ascontiguousarray(mask, dtype=np.uint8)


def matmult_a11(mx, my, mz, coords, eta, mask, x):
    y = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_matmult_a11(mx, my, mz, _f64(coords), _f64(eta), _mask(mask, y.size), _f64(x), y)
    return y


def matmult_a12(mx, my, mz, coords, mask, p):
    y = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_matmult_a12(mx, my, mz, _f64(coords), _mask(mask, y.size), _f64(p), y)
    return y


def matmult_a21(mx, my, mz, coords, mask, u):
    y = np.zeros(4 * mx * my * mz)
    lib().ptq2o_matmult_a21(mx, my, mz, _f64(coords), _mask(mask, u.size), _f64(u), y)
    return y


def matgetdiag_a11(mx, my, mz, coords, eta, mask):
    d = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_matgetdiag_a11(mx, my, mz, _f64(coords), _f64(eta), _mask(mask, d.size), d)
    return d


def matmult_a(mx, my, mz, coords, eta, mask, u, p):
    yu = np.zeros(3 * nnodes(mx, my, mz))
    yp = np.zeros(4 * mx * my * mz)
    lib().ptq2o_matmult_a(mx, my, mz, _f64(coords), _f64(eta), _mask(mask, yu.size), _f64(u), _f64(p), yu, yp)
    return yu, yp


def residual(mx, my, mz, coords, eta, mask, u, p, fu, fp=None, phi=None):
    """FormFunction_Stokes on one rank: (Fu, Fp) with Dirichlet values phi (None = 0) inserted
    into u, body forces fu (E x 27 x 3) and fp (E x 27, None = 0), Fu_D = u_D - phi_D."""
    n = 3 * nnodes(mx, my, mz)
    Fu = np.zeros(n)
    Fp = np.zeros(4 * mx * my * mz)
    fpa = None if fp is None else _f64(fp)
    pha = None if phi is None else _f64(phi)
    lib().ptq2o_residual(mx, my, mz, _f64(coords), _f64(eta), _mask(mask, n),
                         None if pha is None else pha.ctypes.data, _f64(u), _f64(p), _f64(fu),
                         None if fpa is None else fpa.ctypes.data, Fu, Fp)
    return Fu, Fp


def body_force(mx, my, mz, coords, fu):
    """sum_e sum_q w detJ N fu (fu: E x 27 x 3), per quadrature point"""
    y = np.zeros(3 * nnodes(mx, my, mz))
    lib().ptq2o_body_force(mx, my, mz, _f64(coords), _f64(fu), y)
    return y


def marker_lattice(mx, my, mz, coords, n=(2, 2, 2)):
    """Zero-perturbation lattice layout of material points (SwarmMPntStd_CoordAssignment_LatticeLayout3d
    followed by MaterialPointCoordinateSetUp): returns (xi [np,3], wil [np], xp [np,3])."""
    npts = mx * my * mz * n[0] * n[1] * n[2]
    xi = np.zeros(3 * npts)
    xp = np.zeros(3 * npts)
    wil = np.zeros(npts, dtype=np.int32)
    lib().ptq2o_marker_lattice(mx, my, mz, _f64(coords), n[0], n[1], n[2], xi, wil, xp)
    return xi.reshape(-1, 3), wil, xp.reshape(-1, 3)


def sinker_point_properties(xp, eta0=1.0, eta1=10.0, rho0=0.0, rho1=1.0, origin=(0.5, 0.5, 0.5),
                            length=(0.5, 0.5, 0.5)):
    """ViscousSinker_ApplyInitialMaterialGeometry_SingleInclusion (src/models/viscous_sinker/
    model_viscous_sinker_ops.c:760-838), sphere branch: a point is in the inclusion when
    sum_d ((x_d - o_d) / (0.5 l_d))^2 < 1; eta = eta1 / eta0, rho = -rho1 / -rho0."""
    r = ((np.asarray(xp) - np.asarray(origin)) / (0.5 * np.asarray(length))) ** 2
    inside = r.sum(axis=1) < 1.0
    eta = np.where(inside, eta1, eta0).astype(np.float64)
    rho = np.where(inside, -rho1, -rho0).astype(np.float64)
    return eta, rho


def project_q1(mx, my, mz, xi, wil, eta_p, rho_p=None):
    """Q1 projection of point properties onto the quadrature points
    (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes): returns eta_qp (and rho_qp)."""
    xi = _f64(np.asarray(xi).reshape(-1))
    wil = np.ascontiguousarray(wil, dtype=np.int32)
    n = wil.size
    eta_qp = np.zeros(27 * mx * my * mz)
    rho_qp = None if rho_p is None else np.zeros(27 * mx * my * mz)
    rp = None if rho_p is None else _f64(rho_p)
    lib().ptq2o_project_q1(mx, my, mz, n, xi, wil, _f64(eta_p), None if rp is None else rp.ctypes.data, eta_qp,
                           None if rho_qp is None else rho_qp.ctypes.data)
    return eta_qp if rho_p is None else (eta_qp, rho_qp)


def a11_element_matrix(elcoords, eta27):
    K = np.zeros(81 * 81)
    lib().ptq2o_a11_element_matrix(_f64(elcoords), _f64(eta27), K)
    return K.reshape(81, 81)


# ---- CPU baseline: the restatement of the reference's default AVX A11 backend (ptq2_avx.c)
_AVX = os.path.join(_HERE, "libptq2_avx.so")
_avx = None


def avx_lib():
    global _avx
    if _avx is None:
        if not os.path.exists(_AVX):
            build()
        L = ctypes.CDLL(_AVX)
        L.ptq2a_a11_avx.argtypes = [_i, _i, _i, _D, _D, _D, _D, _i]
        L.ptq2a_a11_avx.restype = None
        L.ptq2a_flops_per_element.argtypes = []
        L.ptq2a_flops_per_element.restype = ctypes.c_double
        _avx = L
    return _avx


def a11_avx(mx, my, mz, coords, eta, u, threads=1, y=None):
    """y += A11 u by the AVX NEV=4 restatement (MFStokesWrapper_A11_AVX) on `threads` cores."""
    if y is None:
        y = np.zeros(3 * nnodes(mx, my, mz))
    avx_lib().ptq2a_a11_avx(mx, my, mz, _f64(coords), _f64(eta), _f64(u), y, int(threads))
    return y
