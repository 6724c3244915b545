"""Benchmark of the matrix-free Q2 A11 viscous apply on one MI355X.

One step = one FP64 A11 apply y = A11 u over the whole 128^3-element Q2 mesh
(50,923,779 velocity dofs): one launch of the k_a11_fp64 brick kernel (geometry recomputed
from the coordinates) and one of k_gather_faces.  Inputs are resident in HBM before the
timed region.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--no-cpu] [--no-fp32]
                  [--solve M (default 128, 0 = skip)] [--solve-levels L] [--no-batch]

Besides the apply line, `extra` times the resident-geometry FP64 variant and the FP32 smoother
operator, `extra.solve` the viscous solve of the buoyancy problem (20 dense spheres,
viscosity contrast 1e4, free-slip walls) with FP64- and FP32-smoothed GMG to the same true
residual (rtol 1e-8), the second half of the benchmark metric, `extra.batch_apply` the A11 of the
8 uneven batched problems in one launch against 8 launches, and `extra.batch` their solves
(24^3 ... 80^3, contrasts 1e2 ... 1e6), one after another and concurrently.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, os.path.join(REPO, "ptatin-gene_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
# the reference's flop model of the sum-factorized A11 element (src/stokes_operators_avx.c:305-307)
A11_FLOP_PER_ELEMENT = 17253


def algorithmic_bytes_a11(m, value_bytes=8):
    """Canonical bytes of one A11 apply: u read, y written, coordinates, eta, and an int32
    connectivity entry per element node (each array counted once)."""
    n3 = 3 * (2 * m + 1) ** 3
    e = m ** 3
    return value_bytes * n3 + value_bytes * n3 + 8 * n3 + value_bytes * 27 * e + 4 * 27 * e


def build_problem(m):
    from ptq2 import mesh as M

    # affine+jitter mesh, 12 sphere inclusions with eta 1e3 and log-uniform noise, m^3 elements
    return M.config_sinker(m, n_incl=12, contrast=1e3, bc="none")


def settle_clocks(fn, seconds=2.0):
    """Untimed: run the workload for a few seconds first.  The GPU's clocks ramp up under
    sustained load; without this the first variant timed reads up to ~15% slow."""
    import torch

    t = time.perf_counter()
    n = 0
    while time.perf_counter() - t < seconds:
        fn()
        n += 1
        if n % 20 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()


def time_applies(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps  # ms per step


def copy_bandwidth(stream, nbytes=1 << 30, reps=10):
    """Device-to-device copy of nbytes (read + write counted), the measured HBM ceiling next to
    the 8 TB/s peak."""
    import torch

    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    ms = time_applies(lambda: dst.copy_(src), reps, 2, stream)
    del src, dst
    return round(2 * nbytes / (ms * 1e-3) / 1e9, 1)


def block_comparison(m, steps, warmup, stream):
    """The full Stokes block [A11 A12; A21 0] on the sinker with viscosity contrast 1e6 (no
    noise), FP64 and FP32 (one fused pass on FP32-stored factors), random velocity and
    pressure: time per apply and the FP32 result's relative error against FP64."""
    import torch

    import ptq2
    from ptq2 import _lib
    from ptq2 import mesh as M

    P = M.config_sinker(m, n_incl=12, contrast=1e6, noise=False, bc="none")
    ctx = ptq2.Context(0)
    ctx.setup_mesh(m, m, m, P.coords, P.mask)
    ctx.set_viscosity(P.eta)
    del P
    g = torch.Generator(device="cuda:0")
    g.manual_seed(7)
    u = torch.rand(ctx.nvel, dtype=torch.float64, device="cuda:0", generator=g) * 2 - 1
    p = torch.rand(ctx.npres, dtype=torch.float64, device="cuda:0", generator=g) * 2 - 1
    yu, yp = torch.empty_like(u), torch.empty_like(p)
    u32, p32 = u.float(), p.float()
    yu32, yp32 = torch.empty_like(u32), torch.empty_like(p32)
    ctx.stokes(u32, p32, yu32, yp32, _lib.FP32)  # builds the FP32 factors
    ms64 = time_applies(lambda: ctx.stokes(u, p, yu, yp, 0), steps, warmup, stream)
    ms32 = time_applies(lambda: ctx.stokes(u32, p32, yu32, yp32, _lib.FP32), steps, warmup, stream)
    ctx.stokes(u, p, yu, yp, 0)
    ctx.stokes(u32, p32, yu32, yp32, _lib.FP32)
    torch.cuda.synchronize()
    err_u = float(torch.linalg.norm(yu32.double() - yu) / torch.linalg.norm(yu))
    err_p = float(torch.linalg.norm(yp32.double() - yp) / torch.linalg.norm(yp))
    ndof = ctx.nvel + ctx.npres
    out = {"workload": "full Stokes block [A11 A12; A21 0], %d^3 Q2-P1disc elements (%d velocity + %d pressure "
                       "dofs), 12 sinker inclusions eta 1e6 in eta 1, random u and p" % (m, ctx.nvel, ctx.npres),
           "fp64_ms": round(ms64, 4), "fp64_gdofs": round(ndof / (ms64 * 1e-3) / 1e9, 3),
           "fp32_ms": round(ms32, 4), "fp32_gdofs": round(ndof / (ms32 * 1e-3) / 1e9, 3),
           "fp32_rel_err_u": err_u, "fp32_rel_err_p": err_p}
    ctx.destroy()
    return out


def operator_timings(ctx, u, P, steps, warmup, stream):
    """The other single-operator entry points on the bench's 128^3 context: A12 (gradient), A21
    (divergence) and the A11 diagonal, each with its algorithmic bytes (every array once, int32
    connectivity per element node, the convention of the A11 count) and the resulting rate."""
    import torch

    import ptq2

    ne, n3 = P.nel, P.nvel
    conn = 4 * 27 * ne
    p = torch.from_numpy(ptq2.mesh.uniform(3, 4 * ne, -1.0, 1.0)).to(u.device)
    yu = torch.empty_like(u)
    yp = torch.empty_like(p)
    d = torch.empty_like(u)
    out = {}
    for name, fn, nbytes in (
            ("a12", lambda: ctx.a12(p, yu), 8 * 4 * ne + 8 * n3 + 8 * n3 + conn),
            ("a21", lambda: ctx.a21(u, yp), 8 * n3 + 8 * n3 + 8 * 4 * ne + conn),
            ("diag", lambda: ctx.a11_diag(d), 8 * n3 + 8 * 27 * ne + 8 * n3 + conn)):
        ms = time_applies(fn, steps, warmup, stream)
        out[name] = {"ms": round(ms, 4), "algorithmic_gb": round(nbytes / 1e9, 3),
                     "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    del p, yu, yp, d
    return out


def body_force_timing(m, steps, warmup, stream):
    """The body-force load (ptq2_body_force) on the 128^3 mesh from uniform random quadrature
    values: time per call and its bandwidth over the algorithmic bytes (fu 81 x 8 B per element,
    coordinates and y 3 x 8 B per node)."""
    import torch

    import ptq2
    from ptq2 import mesh as M

    c = M.jitter(m, m, m, M.q2_uniform(m, m, m), 0.1, 1.0 / (2 * m), 1)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(m, m, m, c)
    del c
    g = torch.Generator(device="cuda:0")
    g.manual_seed(11)
    fu = torch.rand(81 * ctx.nel, dtype=torch.float64, device="cuda:0", generator=g)
    y = torch.empty(ctx.nvel, dtype=torch.float64, device="cuda:0")
    ms = time_applies(lambda: ctx.body_force(fu, y), steps, warmup, stream)
    nbytes = 8 * (81 * ctx.nel + 2 * ctx.nvel)
    out = {"workload": "body-force load, %d^3 Q2 elements, jittered mesh, random quadrature values" % m,
           "ms": round(ms, 4), "algorithmic_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    del fu, y
    ctx.destroy()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds=8.0):
    """The CPU baseline: the oracle's restatement of the reference's default A11 backend,
    MFStokesWrapper_A11_AVX (NEV = 4 elements per __m256d, sum-factorized contractions, AVX2/FMA,
    OpenMP with a race-free coloured scatter), on a bounded 40^3 sample of the same workload
    (affine+jitter mesh, 12 inclusions eta 1e3 with noise), timed for `seconds` on all host
    cores the process may use and again on 1 core."""
    from oracle import oracle as O
    from ptq2 import mesh as M

    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", allowed))
    threads = max(1, min(threads, allowed))
    m = 40
    P = M.config_sinker(m, n_incl=12, contrast=1e3, bc="none")
    u = M.uniform(2, P.nvel, -1.0, 1.0)
    y = np.zeros(P.nvel)

    def rate(nt):
        O.a11_avx(m, m, m, P.coords, P.eta, u, threads=nt, y=y)  # warm
        n, t0 = 0, time.perf_counter()
        while True:
            O.a11_avx(m, m, m, P.coords, P.eta, u, threads=nt, y=y)
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                return n, el, P.nvel * n / el / 1e9

    n_all, el_all, v_all = rate(threads)
    n_one, el_one, v_one = rate(1)
    return {"value": round(v_all, 5), "unit": "GDoF/s", "cores": threads, "kind": "port",
            "value_1core": round(v_one, 5), "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": "FP64 A11 applies on a 40^3-element Q2 mesh (%d velocity dofs) by the AVX2 NEV=4 restatement "
                      "of MFStokesWrapper_A11_AVX: %d applies in %.1f s on %d threads, %d applies in %.1f s on 1 core"
                      % (P.nvel, n_all, el_all, threads, n_one, el_one)}


def solve_comparison(m, nlevels, rtol=1e-8, restart=100, max_it=1000, emin_frac=0.05, rigid=False,
                     variants=("fp64", "fp32", "mixed"), fused_smoother=None, autotune=None, galerkin_levels=0):
    """The viscous solve of the buoyancy problem (free-slip walls, 20 dense spheres with
    viscosity contrast 1e4) with GMG-preconditioned FGMRES: FP64 throughout ("fp64"), FP32
    Chebyshev smoothers ("fp32"), and FP32 smoothers with the finished Arnoldi vectors stored in
    FP32 ("mixed").  Wall time of each solve from x = 0 to rtol on the host clock
    around synchronisations (the FGMRES control flow runs on the host).  Both use the
    Chebyshev interval [emin_frac, 1.1] x the estimated lambda_max (the reference's drivers
    use 0.2; 0.05 takes fewer iterations on this problem for either precision)."""
    import torch

    import ptq2
    from ptq2 import mesh as M
    from ptq2.solver import ViscousSolver

    t = time.perf_counter()
    P, b = M.config_buoyancy(m)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(m, m, m, P.coords, P.mask)
    ctx.set_viscosity(P.eta)
    bd = torch.from_numpy(b).to("cuda:0")
    del b
    print("solve set-up %.1f s" % (time.perf_counter() - t), file=sys.stderr, flush=True)
    out = {"workload": "A11 viscous solve, %d^3 Q2 elements, free-slip walls, 20 spheres eta 1e4 / rho 1.1, "
                       "buoyancy RHS, FGMRES(%d) + GMG V-cycle (%d levels, Chebyshev(4)-Jacobi on "
                       "[%g, 1.1] lambda_max)%s, rtol %g"
                       % (m, restart, nlevels, emin_frac, ", rigid-body deflation of the inclusions" if rigid else "",
                          rtol)}
    for key, fp32, b32 in (("fp64", False, False), ("fp32", True, False), ("mixed", True, True)):
        if key not in variants:
            continue
        t = time.perf_counter()
        S = ViscousSolver(ctx, nlevels=nlevels, smooth_its=4, fp32=fp32, emin_frac=emin_frac, rigid_deflation=rigid,
                          fp32_basis=b32, fused_smoother=fused_smoother, autotune=autotune,
                          galerkin_levels=galerkin_levels)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t
        S.vcycle(bd)  # warm the graph
        torch.cuda.synchronize()
        if key != "mixed":  # (the mixed solve has fp32's hierarchy)
            out[key + "_tuning"] = [{"kernel": ViscousSolver.KERNELS[k], "fast_tiles": f, "ms": round(a, 4),
                                     "ms_default": round(d, 4)} for k, f, a, d in S.tuning()]
            out[key + "_level_ms"] = [round(v, 4) for v in S.level_times(bd, reps=5)]
            t0 = time.perf_counter()
            for _ in range(5):
                S.vcycle(bd)
            torch.cuda.synchronize()
            out[key + "_vcycle_ms"] = round((time.perf_counter() - t0) / 5 * 1e3, 4)
        t = time.perf_counter()
        x, info = S.solve(bd, rtol=rtol, max_it=max_it, restart=restart)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        out[key + "_s"] = round(el, 4)
        out[key + "_iterations"] = info["iterations"]
        out[key + "_rel_residual"] = info["rel"]
        out[key + "_setup_s"] = round(setup, 3)
        if rigid:
            out[key + "_inclusions_columns"] = list(S.rigid_info())
        print("%s smoother: %d its, rel %.3e, %.3f s" % (key, info["iterations"], info["rel"], el),
              file=sys.stderr, flush=True)
        S.destroy()
        del x
    if "fp64_s" in out and "fp32_s" in out:
        out["speedup_fp32_vs_fp64"] = round(out["fp64_s"] / out["fp32_s"], 3)
    if "fp64_s" in out and "mixed_s" in out:
        out["speedup_mixed_vs_fp64"] = round(out["fp64_s"] / out["mixed_s"], 3)
    ctx.destroy()
    return out


def batch_comparison(sizes=(24, 32, 40, 48, 56, 64, 64, 80), rtol=1e-8, max_it=1500, restart=256, rigid=True,
                     emin_frac=0.05, members=None, galerkin_levels=None, hard_contrast=1e5):
    """Independent buoyancy solves of uneven sizes (viscosity contrasts 1e2..1e6, seeded
    sphere placement, FP32 smoothing), run one after another and then concurrently (one
    stream and host thread per problem).  Host wall time of each variant.  Per problem also
    the rounding floor of the FP64 residual: |A x - A' x| / |b|, A' the same operator
    evaluated through the resident-geometry kernel (another rounding of the same sums); a
    relative residual much below it is out of reach in FP64.  Problems with a viscosity
    contrast of at least hard_contrast get Galerkin operators on every level between the fine
    and the coarsest one and Chebyshev(8) (galerkin_levels given: that many Galerkin levels on
    every problem, Chebyshev(4))."""
    import torch

    import ptq2
    from ptq2 import mesh as M
    from ptq2.solver import ViscousSolver, auto_levels, solve_batch

    contrasts = [10.0 ** (2 + 4 * k / max(1, len(sizes) - 1)) for k in range(len(sizes))]
    keep = list(range(len(sizes))) if members is None else list(members)
    solvers, rhs, ctxs, config = [], [], [], []
    t = time.perf_counter()
    for k, (m, c) in enumerate(zip(sizes, contrasts)):
        if k not in keep:
            continue
        P, b = M.config_buoyancy(m, contrast=c, seed_spheres=100 + k)
        ctx = ptq2.Context(0)
        ctx.setup_mesh(m, m, m, P.coords, P.mask)
        ctx.set_viscosity(P.eta)
        ctxs.append(ctx)
        nl = auto_levels(m)
        if galerkin_levels is None:
            gl, its = (nl - 2, 8) if c >= hard_contrast else (0, 4)
        else:
            gl, its = min(galerkin_levels, nl - 2), 4
        solvers.append(ViscousSolver(ctx, nlevels=nl, smooth_its=its, fp32=True, rigid_deflation=rigid,
                                     emin_frac=emin_frac, galerkin_levels=gl))
        config.append({"levels": nl, "galerkin_levels": gl, "smooth_its": its,
                       "kernels": [ViscousSolver.KERNELS[kk] for kk, _, _, _ in solvers[-1].tuning()],
                       "kernel_ms": [[round(a, 4), round(d, 4)] for _, _, a, d in solvers[-1].tuning()]})
        rhs.append(torch.from_numpy(b).to("cuda:0"))
    print("batch set-up %.1f s" % (time.perf_counter() - t), file=sys.stderr, flush=True)
    for S, b in zip(solvers, rhs):
        S.vcycle(b)
    torch.cuda.synchronize()
    t = time.perf_counter()
    seq = [S.solve(b, rtol=rtol, max_it=max_it, restart=restart)[1] for S, b in zip(solvers, rhs)]
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t
    t = time.perf_counter()
    bat = solve_batch(solvers, rhs, rtol=rtol, max_it=max_it, restart=restart)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t
    floor = []
    for ctx, b, (x, _) in zip(ctxs, rhs, bat):
        y1 = ctx.a11(x, flags=ptq2.MATMULT)
        y2 = ctx.a11(x, flags=ptq2.MATMULT | 0x20)  # PTQ2_STOREDGEO
        floor.append(float("%.3g" % (torch.linalg.norm(y1 - y2).item() / torch.linalg.norm(b).item())))
        del y1, y2
    out = {"workload": "8 buoyancy solves (20 spheres, rho 1.1, free-slip walls), FGMRES(%d) + GMG V-cycle "
                       "(Chebyshev-Jacobi FP32 on [%g, 1.1] lambda_max)%s, rtol %g, at most %d iterations; "
                       "per problem: levels, Galerkin levels above the coarsest, Chebyshev iterations in 'config'"
                       % (restart, emin_frac, ", rigid-body deflation of the inclusions" if rigid else "", rtol, max_it),
           "config": config,
           "sizes": [sizes[k] for k in keep], "contrasts": [float("%.3g" % contrasts[k]) for k in keep],
           "iterations": [i["iterations"] for i in seq], "rel_residual": [i["rel"] for i in seq],
           "stopped_at_floor": [i["stagnated"] for i in seq],
           "fp64_residual_floor": floor,
           "batched_iterations_equal": [a["iterations"] == b[1]["iterations"] for a, b in zip(seq, bat)],
           "sequential_s": round(t_seq, 4), "concurrent_s": round(t_bat, 4),
           "speedup": round(t_seq / t_bat, 3)}
    for S in solvers:
        S.destroy()
    for c in ctxs:
        c.destroy()
    return out


def batch_apply_comparison(steps, warmup, stream, sizes=(24, 32, 40, 48, 56, 64, 64, 80)):
    """A11 of the 8 uneven sinker problems of the batched solves (viscosity contrasts 1e2..1e6):
    the 8 single applies one after another on one stream against one batched launch
    (ptq2_a11_apply_batch), FP64 and FP32.  Time per round of 8 applies."""
    import torch

    import ptq2
    from ptq2 import mesh as M

    contrasts = [10.0 ** (2 + 4 * k / max(1, len(sizes) - 1)) for k in range(len(sizes))]
    ctxs, us = [], []
    for k, (m, c) in enumerate(zip(sizes, contrasts)):
        P, _ = M.config_buoyancy(m, contrast=c, seed_spheres=100 + k)
        ctx = ptq2.Context(0)
        ctx.setup_mesh(m, m, m, P.coords, P.mask)
        ctx.set_viscosity(P.eta)
        ctxs.append(ctx)
        us.append(torch.from_numpy(M.uniform(5 + k, ctx.nvel, -1.0, 1.0)).to("cuda:0"))
    ndofs = sum(c.nvel for c in ctxs)
    out = {"workload": "A11 of 8 problems %s^3 (buoyancy sinkers, contrasts 1e2..1e6), velocity dofs %d"
                       % ("/".join(str(m) for m in sizes), ndofs)}
    for name, flags, conv in (("fp64", 0, lambda t: t), ("fp32", ptq2.FP32, lambda t: t.float())):
        uu = [conv(u) for u in us]
        ys = [torch.empty_like(u) for u in uu]
        for c, u, y in zip(ctxs, uu, ys):
            c.a11(u, y, flags | ptq2.DEVICE)  # FP32 factors and face buffers outside the timing
        torch.cuda.synchronize()

        def single():
            for c, u, y in zip(ctxs, uu, ys):
                c.a11(u, y, flags | ptq2.DEVICE)

        ms_seq = time_applies(single, steps, warmup, stream)
        ms_bat = time_applies(lambda: ptq2.a11_batch(ctxs, uu, ys, flags), steps, warmup, stream)
        out[name] = {"sequential_ms": round(ms_seq, 4), "batched_ms": round(ms_bat, 4),
                     "speedup": round(ms_seq / ms_bat, 3),
                     "batched_gdofs": round(ndofs / (ms_bat * 1e-3) / 1e9, 3)}
    for c in ctxs:
        c.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp32", action="store_true", help=argparse.SUPPRESS)  # the default; kept for old commands
    ap.add_argument("--no-fp32", action="store_true", help="skip timing the FP32 smoother operator")
    ap.add_argument("--solve", type=int, default=128, metavar="M",
                    help="also time the FP32- vs FP64-smoothed GMG solve on an M^3 mesh (0 = skip)")
    ap.add_argument("--solve-levels", type=int, default=5)
    ap.add_argument("--solve-max-it", type=int, default=1000,
                    help="iteration cap of the solves (a profiled run uses a few dozen)")
    ap.add_argument("--solve-variants", default="fp64,fp32,mixed",
                    help="comma-separated subset of fp64, fp32, mixed")
    ap.add_argument("--fused-smoother", type=int, default=None, metavar="0|1",
                    help="Chebyshev steps with the update fused into the apply's store (default: the library's)")
    ap.add_argument("--autotune", type=int, default=None, metavar="0|1|2",
                    help="solver launch-shape autotuner (default: the library's, 1); 2 / 3 = row-split / MFMA kernel everywhere")
    ap.add_argument("--galerkin-levels", type=int, default=None, metavar="K",
                    help="Galerkin (assembled stencil) operators on the K levels above the coarsest (solves: "
                         "default 0; batched solves: default on the problems of contrast >= 1e5 only)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched solves (--batch)")
    ap.add_argument("--batch-members", default=None, help="comma-separated subset of the 8 batched problems (0..7)")
    ap.add_argument("--solve-only", action="store_true",
                    help="skip the apply timings and the CPU baseline: run the solve comparison only")
    ap.add_argument("--batch", action="store_true", help=argparse.SUPPRESS)  # the default; kept for old commands
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 8 uneven batched solves (C5), sequential vs concurrent")
    ap.add_argument("--rigid", choices=("off", "on", "both"), default="both",
                    help="rigid-body deflation of the stiff inclusions in the 128^3 solve comparison (both: "
                         "without it as 'solve', with it as 'solve_rigid')")
    ap.add_argument("--batch-rigid", choices=("off", "on", "both"), default="on",
                    help="rigid-body deflation in the batched solves (both: without it as 'batch_plain')")
    ap.add_argument("--batch-apply", type=int, default=1, metavar="0|1",
                    help="time the A11 of the 8 batched-solve problems: 8 single launches vs one batched launch")
    ap.add_argument("--batch-apply-only", action="store_true",
                    help="run only the batched-apply comparison (the command of its profiler runs)")
    ap.add_argument("--operators-only", action="store_true",
                    help="run only the A12 / A21 / A11-diagonal timings (the command of their profiler runs)")
    ap.add_argument("--body-force-only", action="store_true",
                    help="run only the body-force timing (the command of its profiler runs)")
    ap.add_argument("--block", type=int, default=1, metavar="0|1",
                    help="time the full Stokes block (FP64 and FP32) on the contrast-1e6 sinker (default 1)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates HBM byte counters)")
    ap.add_argument("--settle", type=float, default=2.0,
                    help="seconds of untimed applies before the timed ones (clock ramp)")
    ap.add_argument("--apply-only", action="store_true",
                    help="time the FP64 apply only (no variants, comparisons or CPU baseline): the "
                         "command of the profiler runs")
    args = ap.parse_args()
    args.batch = not args.no_batch
    if args.apply_only:
        args.no_cpu = args.no_fp32 = True
        args.solve = args.block = args.batch_apply = 0
        args.batch = False
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library runs on a single MI355X; --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ptq2
    from ptq2 import _lib

    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 8, "copy_probe_bytes_written": dst.numel() * 8}))
        return

    if args.batch_apply_only:
        extra = {"batch_apply": batch_apply_comparison(args.steps, args.warmup, torch.cuda.current_stream())}
        print(json.dumps({"batch_apply_only": True, "extra": extra}), flush=True)
        return

    if args.body_force_only:
        extra = {"body_force": body_force_timing(args.size, args.steps, args.warmup, torch.cuda.current_stream())}
        print(json.dumps({"body_force_only": True, "extra": extra}), flush=True)
        return

    if args.batch_only:
        members = None if args.batch_members is None else [int(v) for v in args.batch_members.split(",")]
        extra = {}
        for rigid in {"off": [False], "on": [True], "both": [False, True]}[args.batch_rigid]:
            key = "batch_plain" if not rigid and args.batch_rigid == "both" else "batch"
            extra[key] = batch_comparison(rigid=rigid, members=members, galerkin_levels=args.galerkin_levels,
                                          max_it=args.solve_max_it if args.solve_max_it != 1000 else 1500)
        print(json.dumps({"batch_only": True, "extra": extra}), flush=True)
        return

    if args.solve_only:
        rigid = {"off": [False], "on": [True], "both": [False, True]}[args.rigid]
        extra = {}
        for rg in rigid:
            key = "solve_rigid" if rg and args.rigid == "both" else "solve"
            extra[key] = solve_comparison(args.solve, args.solve_levels, rigid=rg, max_it=args.solve_max_it,
                                          variants=tuple(args.solve_variants.split(",")),
                                          fused_smoother=args.fused_smoother, autotune=args.autotune,
                                          galerkin_levels=args.galerkin_levels or 0)
        print(json.dumps({"solve_only": True, "extra": extra}), flush=True)
        return

    m = args.size
    t_setup = time.perf_counter()
    P = build_problem(m)
    dev = torch.device("cuda:0")
    A = ptq2.MatA11MF("hip", device=0)
    A.setup(P)
    u = torch.from_numpy(ptq2.mesh.uniform(2, P.nvel, -1.0, 1.0)).to(dev)
    y = torch.empty_like(u)
    stream = torch.cuda.current_stream()
    setup_s = time.perf_counter() - t_setup
    print("setup %.1f s (%d elements, %d dofs)" % (setup_s, P.nel, P.nvel), file=sys.stderr, flush=True)

    if args.settle > 0:
        settle_clocks(lambda: A.ctx.a11(u, y, 0), args.settle)
    if args.operators_only:
        extra = {"operators": operator_timings(A.ctx, u, P, args.steps, args.warmup, stream)}
        print(json.dumps({"operators_only": True, "extra": extra}), flush=True)
        return
    ms = time_applies(lambda: A.ctx.a11(u, y, 0), args.steps, args.warmup, stream)
    gdofs = P.nvel / (ms * 1e-3) / 1e9
    bytes_alg = algorithmic_bytes_a11(m)
    achieved = bytes_alg / (ms * 1e-3) / 1e9
    launches = 2
    extra = {}
    # the same apply against the FP64 arithmetic roof: its algorithmic intensity (17,253 flop per
    # element over 1,901.6 MB at 128^3: ~19 flop/B) exceeds the machine balance (78.6 TFLOP/s /
    # 8 TB/s = 9.8 flop/B), so FP64 arithmetic, not HBM, is the roof that binds
    tflops = A11_FLOP_PER_ELEMENT * P.nel / (ms * 1e-3) / 1e12
    extra["compute_roofline"] = {"flop_per_element": A11_FLOP_PER_ELEMENT, "achieved_tflops": round(tflops, 2),
                                 "peak_tflops": FP64_PEAK_TFLOPS, "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
                                 "intensity_flop_per_byte": round(A11_FLOP_PER_ELEMENT * P.nel / bytes_alg, 2),
                                 "hbm_frac_at_fp64_peak": round(
                                     FP64_PEAK_TFLOPS * 1e12 / (A11_FLOP_PER_ELEMENT * P.nel) * bytes_alg / 1e9
                                     / HBM_PEAK_GBS, 4)}
    if not args.apply_only:
        ms_sg = time_applies(lambda: A.ctx.a11(u, y, _lib.STOREDGEO), args.steps, args.warmup, stream)
        extra["fp64_storedgeo_ms"] = round(ms_sg, 4)
        extra["fp64_storedgeo_gdofs"] = round(P.nvel / (ms_sg * 1e-3) / 1e9, 3)
    if not args.no_fp32:
        u32 = u.float()
        y32 = torch.empty_like(u32)
        A.ctx.a11(u32, y32, _lib.FP32)  # builds the FP32 factors outside the timed region
        ms32 = time_applies(lambda: A.ctx.a11(u32, y32, _lib.FP32), args.steps, args.warmup, stream)
        extra["fp32_gdofs"] = round(P.nvel / (ms32 * 1e-3) / 1e9, 3)
        extra["fp32_ms"] = round(ms32, 4)
        torch.cuda.synchronize()
        extra["fp32_rel_err_vs_fp64"] = float(torch.linalg.norm(y32.double() - y) / torch.linalg.norm(y))

    if not args.apply_only:
        extra["hbm_copy_gbs"] = copy_bandwidth(stream)
        try:
            extra["operators"] = operator_timings(A.ctx, u, P, args.steps, args.warmup, stream)
        except Exception as exc:  # reported, not fatal to the headline line
            extra["operators"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
    # release the apply set-up before the larger comparisons
    del A, u, y
    u32 = y32 = None
    torch.cuda.empty_cache()
    if args.block:
        try:
            extra["block"] = block_comparison(m, args.steps, args.warmup, stream)
        except Exception as exc:
            extra["block"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
    if args.block:
        try:
            extra["body_force"] = body_force_timing(m, args.steps, args.warmup, stream)
        except Exception as exc:
            extra["body_force"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
    if args.batch_apply:
        try:
            extra["batch_apply"] = batch_apply_comparison(args.steps, args.warmup, stream)
        except Exception as exc:
            extra["batch_apply"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
    variants = {"off": [False], "on": [True], "both": [False, True]}
    for rigid in variants[args.rigid] if args.solve else []:
        key = "solve_rigid" if rigid and args.rigid == "both" else "solve"
        try:
            extra[key] = solve_comparison(args.solve, args.solve_levels, rigid=rigid, max_it=args.solve_max_it,
                                          variants=tuple(args.solve_variants.split(",")),
                                          fused_smoother=args.fused_smoother, autotune=args.autotune,
                                          galerkin_levels=args.galerkin_levels or 0)
        except Exception as exc:  # the apply line above stays valid; report the failure
            extra[key] = {"error": "%s: %s" % (type(exc).__name__, exc)}
    for rigid in variants[args.batch_rigid] if args.batch else []:
        key = "batch_plain" if not rigid and args.batch_rigid == "both" else "batch"
        try:
            extra[key] = batch_comparison(rigid=rigid)
        except Exception as exc:
            extra[key] = {"error": "%s: %s" % (type(exc).__name__, exc)}

    traffic = None
    prof = os.path.join(REPO, "profiles", "a11_fp64_128_traffic.json")
    if m == 128 and os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()

    out = {
        "metric": "Q2 A11 apply GDoF/s (%HBM peak); mixed vs FP64 GMG solve s @ rtol 1e-8",
        "value": round(gdofs, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "A11 FP64 matrix-free apply y=A11 u, %d^3 Q2 elements (%d velocity dofs), affine+jitter mesh, "
                        "12 sphere inclusions eta=1e3 with log-uniform noise per quadrature point" % (m, P.nvel),
            "elements": P.nel,
            "velocity_dofs": P.nvel,
            "launches_per_step": launches,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    if extra:
        out["extra"] = extra
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
