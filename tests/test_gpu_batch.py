"""Several problems in one launch (ptq2_a11_apply_batch): every output equals the single-problem
ptq2_a11_apply of its context bit for bit, FP64 and FP32, raw, MATMULT and ADD modes, on
meshes of uneven sizes (full and partial bricks, one brick, one element), and the argument
checks."""
import numpy as np
import pytest

from oracle import oracle as O
import ptq2
from ptq2 import _lib
from ptq2 import mesh as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = torch.device("cuda:0")

# uneven sizes: full 4^3 bricks, partial bricks on every axis, a single brick, one element
SIZES = [(8, 8, 8), (6, 5, 7), (1, 1, 1), (12, 8, 4), (3, 9, 2), (4, 4, 4), (10, 6, 9)]


def say(*a):
    print(*a, flush=True)


def _problems(sizes, seed0=0):
    ctxs, us = [], []
    for i, (mx, my, mz) in enumerate(sizes):
        c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(mx, my, mz)), seed0 + i)
        ctx = ptq2.Context(0)
        ctx.setup_mesh(mx, my, mz, c, O.bc_mask(mx, my, mz, i % 3))
        ctx.set_viscosity(10.0 ** M.uniform(100 + seed0 + i, mx * my * mz * 27, -2, 2))
        ctxs.append(ctx)
        us.append(torch.from_numpy(M.uniform(200 + seed0 + i, ctx.nvel, -1, 1)).to(DEV))
    return ctxs, us


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("mode", ["raw", "matmult", "add"])
def test_batch_equals_single_applies(fp32, mode):
    say("case: batch of %d problems %s, fp32=%s, mode=%s" % (len(SIZES), SIZES, fp32, mode))
    ctxs, us = _problems(SIZES)
    flags = (ptq2.FP32 if fp32 else 0) | {"raw": 0, "matmult": ptq2.MATMULT, "add": ptq2.ADD}[mode]
    if fp32:
        us = [u.float() for u in us]
    y0 = [torch.from_numpy(M.uniform(300 + i, c.nvel, -1, 1)).to(DEV).to(us[0].dtype) for i, c in enumerate(ctxs)]
    ref = [c.a11(u, y.clone(), flags=flags | ptq2.DEVICE) for c, u, y in zip(ctxs, us, y0)]
    ys = ptq2.a11_batch(ctxs, us, [y.clone() for y in y0] if mode == "add" else None, flags=flags)
    ys2 = ptq2.a11_batch(ctxs, us, [y.clone() for y in y0] if mode == "add" else None, flags=flags)
    torch.cuda.synchronize()
    for i, (a, b, b2) in enumerate(zip(ref, ys, ys2)):
        assert torch.equal(a, b), "problem %d %s" % (i, SIZES[i])
        assert torch.equal(b, b2), "problem %d not reproducible" % i
    # the FP64 batch against the oracle on one of the problems (the parity of the single apply)
    if not fp32 and mode == "matmult":
        i = 1
        mx, my, mz = SIZES[i]
        c = M.jitter(mx, my, mz, M.q2_uniform(mx, my, mz), 0.1, 1.0 / (2 * max(mx, my, mz)), i)
        eta = 10.0 ** M.uniform(100 + i, mx * my * mz * 27, -2, 2)
        want = O.matmult_a11(mx, my, mz, c, eta, O.bc_mask(mx, my, mz, i % 3), us[i].cpu().numpy())
        got = ys[i].cpu().numpy()
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    for c in ctxs:
        c.destroy()


def test_batch_one_problem_and_own_streams():
    """A batch of one equals the single apply; a problem whose context runs on another stream is
    ordered after that stream's queued work, and that stream waits for the batch."""
    say("case: batch of one; contexts on their own streams")
    ctxs, us = _problems([(6, 5, 7), (8, 8, 8)], seed0=7)
    ref = [c.a11(u, flags=ptq2.DEVICE | ptq2.MATMULT) for c, u in zip(ctxs, us)]
    (y,) = ptq2.a11_batch(ctxs[:1], us[:1], flags=ptq2.MATMULT)
    torch.cuda.synchronize()
    assert torch.equal(y, ref[0])
    s1 = torch.cuda.Stream()
    L = _lib.lib()
    import ctypes
    ys = [torch.zeros_like(u) for u in us]
    with torch.cuda.stream(s1):
        u1 = us[1].clone()  # queued on s1 before the batch
    torch.cuda.current_stream().wait_stream(s1)
    hs = (ctypes.c_void_p * 2)(*[c.handle for c in ctxs])
    L.ptq2_set_stream(ctxs[0].handle, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    L.ptq2_set_stream(ctxs[1].handle, ctypes.c_void_p(s1.cuda_stream))
    up = (ctypes.c_void_p * 2)(us[0].data_ptr(), u1.data_ptr())
    yp = (ctypes.c_void_p * 2)(*[t.data_ptr() for t in ys])
    _lib.check(L.ptq2_a11_apply_batch(ctypes.cast(hs, ctypes.c_void_p), 2, ctypes.cast(up, ctypes.c_void_p),
                                      ctypes.cast(yp, ctypes.c_void_p), ptq2.DEVICE | ptq2.MATMULT),
               "ptq2_a11_apply_batch", ctxs[0].handle)
    with torch.cuda.stream(s1):
        y1 = ys[1].clone()  # s1 waits for the batch
    s1.synchronize()
    torch.cuda.synchronize()
    assert torch.equal(ys[0], ref[0]) and torch.equal(y1, ref[1])
    for c in ctxs:
        c.destroy()


def test_batch_argument_errors():
    say("case: batch argument checks")
    ctxs, us = _problems([(4, 4, 4), (3, 3, 3)], seed0=3)
    with pytest.raises(ValueError):
        ptq2.a11_batch([], [])
    with pytest.raises(ValueError):
        ptq2.a11_batch(ctxs * 9, us * 9)  # 18 > PTQ2_BATCH_MAX
    with pytest.raises(ptq2.PTQ2Error) as ei:
        ptq2.a11_batch([ctxs[0], ctxs[0]], [us[0], us[0]])  # one context twice
    assert ei.value.code == _lib.ERR_INVALID_ARG
    with pytest.raises(ValueError):
        ptq2.a11_batch(ctxs, [us[1], us[0]])  # sizes swapped
    fresh = ptq2.Context(0)
    fresh.setup_mesh(3, 3, 3, M.q2_uniform(3, 3, 3))
    with pytest.raises(ptq2.PTQ2Error) as ei:
        ptq2.a11_batch([ctxs[0], fresh], [us[0], torch.zeros(fresh.nvel, dtype=torch.float64, device=DEV)])
    assert ei.value.code == _lib.ERR_NOT_SETUP  # no viscosity on problem 1
    for c in ctxs + [fresh]:
        c.destroy()
