"""Host-side mirror of pTatin3D's matrix-free Stokes operator objects, backed by libptq2.

Names, argument meaning and error behaviour follow the reference:
  MatA11MF / MatA11MFCreate / MatA11MFSetup / MatA11MFDestroy   (src/stokes_operators.c:104-340)
  MatMult_MFStokes_A11, MatGetDiagonal_MFStokes_A11             (src/stokes_operators.c:1072-1222)
  MatStokesMF with MatMult_MFStokes_A12 / _A21 / _A             (src/stokes_operators.c:893-1509)
  the SpMVOp_MatMult seam (include/stokes_operators.h:48-64)   -> MatA11MF.spmv
The backend registry mirrors `-a11_op` (src/stokes_operators.c:123-154): unknown names raise.

Vectors are numpy arrays (staged through the device, synchronous) or torch tensors on
the GPU (device pointers, ordered on torch's current stream).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import ADD, DEVICE, FP32, MATMULT, check

try:  # torch is the device-memory and stream plumbing of the device path
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _is_dev(x):
    return torch is not None and isinstance(x, torch.Tensor) and x.is_cuda


class Context:
    """One libptq2 context: a device-resident mesh plus quadrature-point viscosity."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        check(L.ptq2_create(ctypes.byref(h), int(device)), "ptq2_create")
        self._h = h
        self.device = device
        self.mx = self.my = self.mz = None
        self.nvel = self.npres = self.nel = 0

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("context destroyed")
        return self._h

    def destroy(self):
        if getattr(self, "_h", None) is not None:
            _lib.lib().ptq2_destroy(ctypes.byref(self._h))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _stream(self, *tensors):
        """Order the library's work on torch's current stream (the null stream by default)."""
        if any(_is_dev(t) for t in tensors):
            s = torch.cuda.current_stream(device=tensors[0].device).cuda_stream
            _lib.lib().ptq2_set_stream(self.handle, ctypes.c_void_p(s) if s else None)

    def setup_mesh(self, mx, my, mz, coords, mask=None, elnidx=None):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        mptr = None
        if mask is not None:
            mask = np.ascontiguousarray(mask, dtype=np.uint8)
            mptr = mask.ctypes.data
        eptr = None
        if elnidx is not None:
            elnidx = np.ascontiguousarray(elnidx, dtype=np.int32)
            eptr = elnidx.ctypes.data
        check(_lib.lib().ptq2_setup_mesh(self.handle, mx, my, mz, coords.ctypes.data, eptr, mptr),
              "ptq2_setup_mesh", self.handle)
        self.mx, self.my, self.mz = mx, my, mz
        nv, np_, ne = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(_lib.lib().ptq2_sizes(self.handle, ctypes.byref(nv), ctypes.byref(np_), ctypes.byref(ne)),
              "ptq2_sizes", self.handle)
        self.nvel, self.npres, self.nel = nv.value, np_.value, ne.value

    def set_viscosity(self, eta, state=0):
        if _is_dev(eta):
            self._stream(eta)
            check(_lib.lib().ptq2_set_viscosity(self.handle, eta.data_ptr(), state, DEVICE),
                  "ptq2_set_viscosity", self.handle)
        else:
            eta = np.ascontiguousarray(eta, dtype=np.float64)
            check(_lib.lib().ptq2_set_viscosity(self.handle, eta.ctypes.data, state, 0),
                  "ptq2_set_viscosity", self.handle)

    def residual(self, u, p, fu, fp=None, phi=None, Fu=None, Fp=None):
        """FormFunction_Stokes (Fu, Fp): momentum A11 u_loc + A12 p - Be(fu) with Dirichlet rows
        u - phi, continuity A21 u_loc - int N^p fp (u_loc = u with the Dirichlet values phi)."""
        L = _lib.lib()
        if _is_dev(u):
            if Fu is None:
                Fu = torch.empty(self.nvel, dtype=torch.float64, device=u.device)
            if Fp is None:
                Fp = torch.empty(self.npres, dtype=torch.float64, device=u.device)
            self._stream(u)
            ptr = (lambda t: None if t is None else t.data_ptr())
            check(L.ptq2_stokes_residual(self.handle, u.data_ptr(), p.data_ptr(), fu.data_ptr(), ptr(fp), ptr(phi),
                                         Fu.data_ptr(), Fp.data_ptr(), DEVICE), "ptq2_stokes_residual", self.handle)
            return Fu, Fp
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (u, p, fu, fp, phi)]
        Fu = np.zeros(self.nvel) if Fu is None else Fu
        Fp = np.zeros(self.npres) if Fp is None else Fp
        ptrs = [None if a is None else a.ctypes.data for a in arrs]
        check(L.ptq2_stokes_residual(self.handle, *ptrs, Fu.ctypes.data, Fp.ctypes.data, 0), "ptq2_stokes_residual",
              self.handle)
        return Fu, Fp

    def project_points(self, xi, wil, eta_p, rho_p=None, eta_qp=None, rho_qp=None):
        """Q1 projection of material-point values onto the quadrature points
        (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes); returns eta_qp, or
        (eta_qp, rho_qp) when rho_p is given.  numpy arrays in, numpy out; torch device
        tensors in, device tensors out."""
        L = _lib.lib()
        n27 = 27 * self.nel
        if _is_dev(eta_p):
            dev = eta_p.device
            xi = xi.reshape(-1).contiguous()
            wil = wil.to(torch.int32).contiguous()
            if eta_qp is None:
                eta_qp = torch.empty(n27, dtype=torch.float64, device=dev)
            if rho_p is not None and rho_qp is None:
                rho_qp = torch.empty(n27, dtype=torch.float64, device=dev)
            self._stream(eta_p)
            ptr = (lambda t: None if t is None else t.data_ptr())
            check(L.ptq2_project_points_q1(self.handle, wil.numel(), xi.data_ptr(), wil.data_ptr(), eta_p.data_ptr(),
                                           ptr(rho_p), eta_qp.data_ptr(), ptr(rho_qp), DEVICE),
                  "ptq2_project_points_q1", self.handle)
        else:
            xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1)
            wil = np.ascontiguousarray(wil, dtype=np.int32)
            eta_p = np.ascontiguousarray(eta_p, dtype=np.float64)
            eta_qp = np.zeros(n27) if eta_qp is None else eta_qp
            rp = None
            if rho_p is not None:
                rho_p = np.ascontiguousarray(rho_p, dtype=np.float64)
                rp = rho_p.ctypes.data
                rho_qp = np.zeros(n27) if rho_qp is None else rho_qp
            check(L.ptq2_project_points_q1(self.handle, wil.size, xi.ctypes.data, wil.ctypes.data, eta_p.ctypes.data,
                                           rp, eta_qp.ctypes.data, None if rho_qp is None else rho_qp.ctypes.data, 0),
                  "ptq2_project_points_q1", self.handle)
        return eta_qp if rho_p is None else (eta_qp, rho_qp)

    def sync(self):
        check(_lib.lib().ptq2_sync(self.handle), "ptq2_sync", self.handle)

    # generic dispatch of one-input/one-output entry points
    def _call1(self, fn, name, x, y, flags, xdt, ydt, ny):
        if _is_dev(x):
            if y is None:
                y = torch.zeros(ny, dtype=ydt, device=x.device)
            self._stream(x, y)
            check(fn(self.handle, x.data_ptr(), y.data_ptr(), flags | DEVICE), name, self.handle)
            return y
        xn = np.ascontiguousarray(x, dtype=xdt)
        if y is None:
            y = np.zeros(ny, dtype=ydt)
        check(fn(self.handle, xn.ctypes.data, y.ctypes.data, flags), name, self.handle)
        return y

    def a11(self, u, y=None, flags=0):
        fp32 = bool(flags & FP32)
        L = _lib.lib()
        if _is_dev(u):
            dt = torch.float32 if fp32 else torch.float64
            return self._call1(L.ptq2_a11_apply, "ptq2_a11_apply", u, y, flags, None, dt, self.nvel)
        dt = np.float32 if fp32 else np.float64
        return self._call1(L.ptq2_a11_apply, "ptq2_a11_apply", u, y, flags, dt, dt, self.nvel)

    def a11_diag(self, d=None, flags=0, device=None):
        L = _lib.lib()
        if device is not None or _is_dev(d):
            if d is None:
                d = torch.zeros(self.nvel, dtype=torch.float64, device=device)
            self._stream(d)
            check(L.ptq2_a11_diag(self.handle, d.data_ptr(), flags | DEVICE), "ptq2_a11_diag", self.handle)
            return d
        if d is None:
            d = np.zeros(self.nvel)
        check(L.ptq2_a11_diag(self.handle, d.ctypes.data, flags), "ptq2_a11_diag", self.handle)
        return d

    def a12(self, p, y=None, flags=0):
        L = _lib.lib()
        dt = torch.float64 if _is_dev(p) else np.float64
        return self._call1(L.ptq2_a12_apply, "ptq2_a12_apply", p, y, flags, np.float64, dt, self.nvel)

    def body_force(self, fu, y=None, flags=0):
        """Body-force load of fu (E x 27 x 3 quadrature values, element-major): the Be term of
        FormFunctionLocal_U (src/stokes_form_function.c:219-224)."""
        L = _lib.lib()
        dt = torch.float64 if _is_dev(fu) else np.float64
        return self._call1(L.ptq2_body_force, "ptq2_body_force", fu, y, flags, np.float64, dt, self.nvel)

    def a21(self, u, y=None, flags=0):
        L = _lib.lib()
        dt = torch.float64 if _is_dev(u) else np.float64
        return self._call1(L.ptq2_a21_apply, "ptq2_a21_apply", u, y, flags, np.float64, dt, self.npres)

    def stokes(self, u, p, yu=None, yp=None, flags=0):
        """[A11 A12; A21 0] [u; p]; with FP32 in flags all vectors are float32."""
        L = _lib.lib()
        f32 = bool(flags & FP32)
        if _is_dev(u):
            dt = torch.float32 if f32 else torch.float64
            for v in (u, p, yu, yp):
                if v is not None and v.dtype != dt:
                    raise TypeError("ptq2_stokes_apply: vectors must be %s" % dt)
            if yu is None:
                yu = torch.zeros(self.nvel, dtype=dt, device=u.device)
            if yp is None:
                yp = torch.zeros(self.npres, dtype=dt, device=u.device)
            self._stream(u, p, yu, yp)
            check(L.ptq2_stokes_apply(self.handle, u.data_ptr(), p.data_ptr(), yu.data_ptr(), yp.data_ptr(),
                                      flags | DEVICE), "ptq2_stokes_apply", self.handle)
            return yu, yp
        dt = np.float32 if f32 else np.float64
        un = np.ascontiguousarray(u, dtype=dt)
        pn = np.ascontiguousarray(p, dtype=dt)
        if yu is None:
            yu = np.zeros(self.nvel, dtype=dt)
        if yp is None:
            yp = np.zeros(self.npres, dtype=dt)
        check(L.ptq2_stokes_apply(self.handle, un.ctypes.data, pn.ctypes.data, yu.ctypes.data, yp.ctypes.data,
                                  flags), "ptq2_stokes_apply", self.handle)
        return yu, yp


BATCH_MAX = 16  # PTQ2_BATCH_MAX


def a11_batch(ctxs, us, ys=None, flags=0):
    """A11 of several problems (contexts on one GPU, meshes of any sizes) in one launch of the
    apply kernel and one of the face pass (ptq2_a11_apply_batch): ys[i] = A11_i us[i].  Device
    tensors only (float64, or float32 with FP32 in flags); ordered on torch's current stream."""
    n = len(ctxs)
    if not 1 <= n <= BATCH_MAX or len(us) != n:
        raise ValueError("a11_batch takes 1..%d problems and one input each" % BATCH_MAX)
    if not all(_is_dev(u) for u in us):
        raise ValueError("a11_batch takes device tensors")
    dt = torch.float32 if flags & FP32 else torch.float64
    for c, u in zip(ctxs, us):
        if u.dtype != dt or u.numel() != c.nvel or not u.is_contiguous():
            raise ValueError("input of %d %s values expected" % (c.nvel, dt))
    if ys is None:
        ys = [torch.empty_like(u) for u in us]
    for c, y in zip(ctxs, ys):
        if y.dtype != dt or y.numel() != c.nvel or not y.is_contiguous():
            raise ValueError("output of %d %s values expected" % (c.nvel, dt))
    for c in ctxs:
        c._stream(us[0])
    hs = (ctypes.c_void_p * n)(*[c.handle for c in ctxs])
    up = (ctypes.c_void_p * n)(*[u.data_ptr() for u in us])
    yp = (ctypes.c_void_p * n)(*[y.data_ptr() for y in ys])
    L = _lib.lib()
    check(L.ptq2_a11_apply_batch(ctypes.cast(hs, ctypes.c_void_p), n, ctypes.cast(up, ctypes.c_void_p),
                                 ctypes.cast(yp, ctypes.c_void_p), flags | DEVICE),
          "ptq2_a11_apply_batch", ctxs[0].handle)
    return ys


# -------------------------------------------------------------- -a11_op registry
_A11_OPS = {}


def register_a11_op(name, factory):
    """PetscFunctionListAdd analogue for A11 backends (src/stokes_operators.c:123-142)."""
    _A11_OPS[name] = factory


class MatA11MF:
    """The A11 shell context (struct _p_MatA11MF, include/stokes_operators.h:48-64) with the
    HIP backend behind SpMVOp_SetUp / SpMVOp_MatMult / SpMVOp_Destroy."""

    def __init__(self, op="hip", device=0):
        if op not in _A11_OPS:
            raise ValueError("Unknown -a11_op %r (registered: %s)" % (op, ", ".join(sorted(_A11_OPS))))
        self.op = op
        self.ctx = _A11_OPS[op](device)
        self.is_setup = False
        self.state = 0

    def setup(self, problem, state=0):
        """MatA11MFSetup: mesh, BC list and quadrature-point viscosity."""
        self.ctx.setup_mesh(problem.mx, problem.my, problem.mz, problem.coords, problem.mask)
        self.ctx.set_viscosity(problem.eta, state)
        self.state = state
        self.is_setup = True

    def set_viscosity(self, eta, state):
        self.ctx.set_viscosity(eta, state)
        self.state = state

    def spmv(self, u, y, fp32=False):
        """SpMVOp_MatMult contract: Yu += element sum of A11 u (no boundary handling)."""
        self._need_setup()
        return self.ctx.a11(u, y, ADD | (FP32 if fp32 else 0))

    def mult(self, x, y=None, fp32=False):
        """MatMult_MFStokes_A11: zero Dirichlet columns, apply, y_k = x_k on Dirichlet rows."""
        self._need_setup()
        return self.ctx.a11(x, y, MATMULT | (FP32 if fp32 else 0))

    def get_diagonal(self, d=None, device=None):
        """MatGetDiagonal_MFStokes_A11: diagonal with 1 in Dirichlet slots."""
        self._need_setup()
        return self.ctx.a11_diag(d, MATMULT, device)

    def destroy(self):
        self.ctx.destroy()
        self.is_setup = False

    def _need_setup(self):
        if not self.is_setup:
            raise RuntimeError("MF operator not setup")


class MatStokesMF:
    """Shell context of the Stokes blocks (A12, A21, full A)."""

    def __init__(self, device=0):
        self.ctx = Context(device)
        self.is_setup = False

    def setup(self, problem, state=0):
        self.ctx.setup_mesh(problem.mx, problem.my, problem.mz, problem.coords, problem.mask)
        self.ctx.set_viscosity(problem.eta, state)
        self.is_setup = True

    def mult_a12(self, p, y=None):
        """MatMult_MFStokes_A12: y = A12 p with Dirichlet rows of y set to 0."""
        return self.ctx.a12(p, y, MATMULT)

    def mult_a21(self, u, y=None):
        """MatMult_MFStokes_A21: y = A21 u with Dirichlet columns of u zeroed."""
        return self.ctx.a21(u, y, MATMULT)

    def mult_a(self, u, p, yu=None, yp=None):
        """MatMult_MFStokes_A: [A11 A12; A21 0][u; p] with the velocity BCs."""
        return self.ctx.stokes(u, p, yu, yp, MATMULT)


register_a11_op("hip", lambda device: Context(device))
