"""Host wrapper of the GMG-preconditioned FGMRES viscous solve (include/ptq2_solver.h).

Mirrors the PETSc configuration pTatin3D builds for the velocity block
(pTatin3dStokesKSPConfigureFSGMG, src/ptatin_driver_linear_ts.c:820-883, with the
sinker options of tests/application/linear/sinker_mf_8x8x8/opts): PCMG, Chebyshev +
Jacobi smoothers with the (0, 0.2, 0, 1.1) eigen-estimate transform, FGMRES outside.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


class MGOpts(ctypes.Structure):
    _fields_ = [("nlevels", ctypes.c_int), ("smooth_its", ctypes.c_int), ("fp32_smoother", ctypes.c_int),
                ("emin_frac", ctypes.c_double), ("emax_frac", ctypes.c_double), ("esteig_its", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("coarse_max_dofs", ctypes.c_int), ("galerkin_coarse", ctypes.c_int),
                ("coarse_eta_mean", ctypes.c_int), ("fp32_basis", ctypes.c_int), ("rigid_deflation", ctypes.c_int),
                ("rigid_min_contrast", ctypes.c_double), ("rigid_max_inclusions", ctypes.c_int),
                ("fused_smoother", ctypes.c_int), ("autotune", ctypes.c_int), ("galerkin_levels", ctypes.c_int)]

ETA_MEANS = {"arithmetic": 0, "geometric": 1, "harmonic": 2, "max": 3, "q1proj": 4}


_SOLVER_SIGS = {
    "ptq2_mg_default_opts": (None, [ctypes.POINTER(MGOpts)]),
    "ptq2_mg_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.POINTER(MGOpts)]),
    "ptq2_mg_destroy": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "ptq2_mg_levels": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]),
    "ptq2_mg_eigs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "ptq2_mg_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "ptq2_mg_level_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_void_p]),
    "ptq2_mg_rigid_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.c_void_p]),
    "ptq2_mg_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ptq2_mg_vcycle": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "ptq2_mg_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]),
    "ptq2_mg_history": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "ptq2_mg_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ptq2_mg_stagnated": (ctypes.c_int, [ctypes.c_void_p]),
}
_bound = False


def _sl():
    global _bound
    L = _lib.lib()
    if not _bound:
        for name, (res, args) in _SOLVER_SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return L


def _mgcheck(code, where, h):
    """raise PTQ2Error with the hierarchy's own message for a nonzero status"""
    if code != _lib.OK:
        raise _lib.PTQ2Error(code, where, _sl().ptq2_mg_last_error(h).decode())


def default_opts():
    o = MGOpts()
    _sl().ptq2_mg_default_opts(ctypes.byref(o))
    return o


class ViscousSolver:
    """GMG V-cycle (hipGraph) + FGMRES for MatMult_MFStokes_A11 x = b on one GPU."""

    def __init__(self, ctx, nlevels=5, smooth_its=4, fp32=True, esteig_its=20, use_graph=True, coarse_max_dofs=20000,
                 galerkin_coarse=True, eta_mean="arithmetic", emin_frac=None, emax_frac=None, rigid_deflation=False,
                 rigid_min_contrast=None, rigid_max_inclusions=None, fp32_basis=False, fused_smoother=None,
                 autotune=None, galerkin_levels=0):
        L = _sl()
        o = default_opts()
        o.nlevels, o.smooth_its, o.fp32_smoother = nlevels, smooth_its, 1 if fp32 else 0
        o.esteig_its, o.use_graph, o.coarse_max_dofs = esteig_its, 1 if use_graph else 0, coarse_max_dofs
        o.galerkin_coarse = 1 if galerkin_coarse else 0
        if eta_mean not in ETA_MEANS:
            raise ValueError("eta_mean must be one of %s" % sorted(ETA_MEANS))
        o.coarse_eta_mean = ETA_MEANS[eta_mean]
        if emin_frac is not None:
            o.emin_frac = emin_frac
        if emax_frac is not None:
            o.emax_frac = emax_frac
        o.fp32_basis = 1 if fp32_basis else 0
        if fused_smoother is not None:
            o.fused_smoother = 1 if fused_smoother else 0
        o.galerkin_levels = galerkin_levels
        if autotune is not None:
            o.autotune = {False: 0, True: 1}.get(autotune, autotune)  # 2-4: one forced choice (see ptq2_solver.h)
        o.rigid_deflation = 1 if rigid_deflation else 0
        if rigid_min_contrast is not None:
            o.rigid_min_contrast = rigid_min_contrast
        if rigid_max_inclusions is not None:
            o.rigid_max_inclusions = rigid_max_inclusions
        self.ctx = ctx  # keep the fine context alive
        h = ctypes.c_void_p()
        check(L.ptq2_mg_create(ctypes.byref(h), ctx.handle, ctypes.byref(o)), "ptq2_mg_create", ctx.handle)
        self._h = h
        self.nlevels = nlevels
        self.fp32 = fp32

    def levels(self):
        n = ctypes.c_int()
        m = np.zeros(3 * self.nlevels, dtype=np.int32)
        check(_sl().ptq2_mg_levels(self._h, ctypes.byref(n), m.ctypes.data), "ptq2_mg_levels")
        return m.reshape(-1, 3)

    def eigs(self):
        lo = np.zeros(self.nlevels - 1)
        hi = np.zeros(self.nlevels - 1)
        check(_sl().ptq2_mg_eigs(self._h, lo.ctypes.data, hi.ctypes.data), "ptq2_mg_eigs")
        return lo, hi

    KERNELS = ("brick", "split", "mfma", "stencil", "galerkin_mf")  # PTQ2_MG_KERNEL_*

    def tuning(self):
        """Per smoothing level: (kernel index into KERNELS, fast tile addressing, ms per apply
        chosen, ms per apply of the default shape) as chosen by the set-up's autotuner."""
        k = self.nlevels - 1
        w = np.zeros(k, dtype=np.int32)
        f = np.zeros(k, dtype=np.int32)
        ms = np.zeros(k)
        md = np.zeros(k)
        check(_sl().ptq2_mg_tuning(self._h, w.ctypes.data, f.ctypes.data, ms.ctypes.data, md.ctypes.data),
              "ptq2_mg_tuning")
        return [(int(a), int(b), float(c), float(d)) for a, b, c, d in zip(w, f, ms, md)]

    def level_times(self, r, reps=10):
        """ms of each level's own work in a V-cycle of r (eager launches, mean of reps)."""
        import torch

        z = torch.empty_like(r)
        ms = np.zeros(self.nlevels)
        self._stream(r)
        _mgcheck(_sl().ptq2_mg_level_times(self._h, r.data_ptr(), z.data_ptr(), reps, ms.ctypes.data),
                 "ptq2_mg_level_times", self._h)
        return ms

    def rigid_info(self, labels=False):
        """(inclusions, kept rigid-motion columns[, per-node inclusion labels]) of the
        two-level correction for stiff inclusions"""
        ni, nk = ctypes.c_int(), ctypes.c_int()
        lab = None
        if labels:
            mx, my, mz = self.levels()[0]
            lab = np.zeros((2 * mx + 1) * (2 * my + 1) * (2 * mz + 1), dtype=np.int32)
        check(_sl().ptq2_mg_rigid_info(self._h, ctypes.byref(ni), ctypes.byref(nk),
                                       None if lab is None else lab.ctypes.data), "ptq2_mg_rigid_info")
        return (ni.value, nk.value, lab) if labels else (ni.value, nk.value)

    def _stream(self, t):
        """Order the solver on torch's current stream (the null stream by default)."""
        import torch

        s = torch.cuda.current_stream(device=t.device).cuda_stream
        check(_sl().ptq2_mg_set_stream(self._h, ctypes.c_void_p(s) if s else None), "ptq2_mg_set_stream")

    def vcycle(self, r, z=None):
        import torch

        if z is None:
            z = torch.empty_like(r)
        self._stream(r)
        _mgcheck(_sl().ptq2_mg_vcycle(self._h, r.data_ptr(), z.data_ptr()), "ptq2_mg_vcycle", self._h)
        return z

    def solve(self, b, x=None, rtol=1e-8, max_it=200, restart=30):
        import torch

        if x is None:
            x = torch.empty_like(b)
        self._stream(b)
        it = ctypes.c_int()
        fr = ctypes.c_double()
        r0 = ctypes.c_double()
        _mgcheck(_sl().ptq2_mg_solve(self._h, b.data_ptr(), x.data_ptr(), rtol, max_it, restart, ctypes.byref(it),
                                     ctypes.byref(fr), ctypes.byref(r0)), "ptq2_mg_solve", self._h)
        return x, {"iterations": it.value, "final_rnorm": fr.value, "rnorm0": r0.value,
                   "rel": fr.value / r0.value if r0.value else 0.0,
                   "stagnated": bool(_sl().ptq2_mg_stagnated(self._h))}

    def history(self):
        n = ctypes.c_int()
        _sl().ptq2_mg_history(self._h, None, 0, ctypes.byref(n))
        h = np.zeros(n.value)
        _sl().ptq2_mg_history(self._h, h.ctypes.data, n.value, ctypes.byref(n))
        return h

    def destroy(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _sl().ptq2_mg_destroy(ctypes.byref(self._h))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def auto_levels(m, coarse_max_dofs=20000):
    """Number of levels for an m^3 mesh coarsened by 2 until the coarse level has at most
    coarse_max_dofs velocity dofs (or cannot be halved)."""
    levels = 1
    while m % 2 == 0 and 3 * (2 * m + 1) ** 3 > coarse_max_dofs:
        m //= 2
        levels += 1
    return levels


def solve_batch(solvers, rhs, rtol=1e-8, max_it=400, restart=30):
    """Solve independent viscous problems concurrently: one host thread and one stream per
    solver (the C solver releases the GIL inside ctypes), so the small kernels of the
    different problems overlap on the GPU.  Returns [(x, info)] in input order."""
    import concurrent.futures

    import torch

    main = torch.cuda.current_stream()
    streams = [torch.cuda.Stream(device=b.device) for b in rhs]
    for s in streams:
        s.wait_stream(main)  # the right-hand sides were produced on the current stream

    def run(k):
        with torch.cuda.stream(streams[k]):
            x, info = solvers[k].solve(rhs[k], rtol=rtol, max_it=max_it, restart=restart)
        streams[k].synchronize()
        return x, info

    with concurrent.futures.ThreadPoolExecutor(max_workers=len(solvers)) as ex:
        out = list(ex.map(run, range(len(solvers))))
    for s in streams:
        main.wait_stream(s)
    return out
