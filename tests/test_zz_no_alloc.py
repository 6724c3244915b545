"""ptq2_a11_apply_noalloc (the allocation-free A11 entry point, face slots in a library-owned
static device array) against ptq2_a11_apply through the C ABI: bitwise equal on two calls in a
row, the second on a smaller mesh, in raw, MATMULT and ADD modes; a mesh whose face slots exceed
the array is PTQ2_ERR_SIZE; host pointers are rejected."""
import ctypes

import numpy as np
import pytest

import ptq2
from ptq2 import _lib
from ptq2 import mesh as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = torch.device("cuda:0")


def _noalloc(ctx, u, y, flags):
    L = _lib.lib()
    L.ptq2_a11_apply_noalloc.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint]
    L.ptq2_a11_apply_noalloc.restype = ctypes.c_int
    return L.ptq2_a11_apply_noalloc(ctx.handle, u.data_ptr(), y.data_ptr(), flags)


def _problem(mx, my, mz, seed):
    c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(mx, my, mz)), seed)
    eta = 10.0 ** M.uniform(seed + 1, 27 * mx * my * mz, -1, 2)
    ctx = ptq2.Context(0)
    ctx.setup_mesh(mx, my, mz, c, M.bc_mask(mx, my, mz, "freeslip"))
    ctx.set_viscosity(eta)
    u = torch.from_numpy(M.uniform(seed + 2, ctx.nvel, -1, 1)).to(DEV)
    return ctx, u


@pytest.mark.parametrize("dims", [((20, 18, 22), (9, 5, 7))])
def test_noalloc_matches_a11_apply(dims):
    for k, (mx, my, mz) in enumerate(dims):
        print("case: noalloc A11 %dx%dx%d" % (mx, my, mz), flush=True)
        ctx, u = _problem(mx, my, mz, 10 + k)
        for flags in (0, ptq2.MATMULT, _lib.ADD):
            y0 = torch.from_numpy(M.uniform(99, ctx.nvel, -1, 1)).to(DEV)
            ref = ctx.a11(u, y0.clone(), flags)
            out = y0.clone()
            rc = _noalloc(ctx, u, out, flags | _lib.DEVICE)
            torch.cuda.synchronize()
            assert rc == 0, ptq2._lib.lib().ptq2_last_error(ctx.handle)
            assert torch.equal(out, ref), (dims[k], flags)
        ctx.destroy()


def test_noalloc_limits():
    print("case: noalloc limits", flush=True)
    ctx, u = _problem(72, 72, 72, 20)  # 18^3 = 5832 bricks > 5753
    y = torch.zeros_like(u)
    assert _noalloc(ctx, u, y, _lib.DEVICE) == _lib.ERR_SIZE
    ctx.destroy()
    ctx, u = _problem(4, 4, 4, 21)
    y = torch.zeros_like(u)
    assert _noalloc(ctx, u, y, 0) == _lib.ERR_INVALID_ARG  # host pointers need staging
    assert _noalloc(ctx, u, y, _lib.DEVICE | _lib.FP32) == _lib.ERR_INVALID_ARG
    ctx.destroy()
