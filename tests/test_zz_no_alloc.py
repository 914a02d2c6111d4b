"""gp_eval_residual_noalloc (the allocation-free residual, device arrays in
the library's layout, scratch in a per-GPU static array) against
gp_eval_residual on the same inputs: bit for bit, for two calls in a row on
contexts of different sizes (the second smaller, so stale scratch contents of
the first would show), a third call back on the first context, and a context
whose scratch need exceeds the maximum, which is refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import get_case

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def put(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0   # host to device

    def get(self, src, a):
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0   # device to host

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _setup(case, M, seed):
    from gradpower_hip.lib import HipContext
    ctx = HipContext(case.desc, M)
    ctx.set_state(case.desc.z0, case.desc.p)
    ev = [(m, e.bus, e.rfault, 1, 2) for m in range(M) for e in case.events[m % case.M]]
    ctx.set_events(ev, [10] * M)
    rng = np.random.default_rng(seed)
    n = case.desc.n
    z = case.desc.z0[None, :] * (1 + 1e-3 * rng.standard_normal((M, n)))
    zold = case.desc.z0[None, :] * (1 + 1e-3 * rng.standard_normal((M, n)))
    fo = (np.arange(M) % 2).astype(np.uint8)
    return ctx, z, zold, fo


def _device_call(hip, ctx, z, zold, fo, dt):
    """The allocation-free entry point on device copies of z, zold, fault_on
    in the library's layout; returns f as [M, n]."""
    M, n = z.shape
    ld = ctx.symbolic_info()["ld_scen"]
    assert ld % 64 == 0 and ld >= M
    def lay(a):
        t = np.zeros((n, ld))
        t[:, :M] = a.T
        return t
    dz, dzo, df, dfo = hip.alloc(8 * n * ld), hip.alloc(8 * n * ld), hip.alloc(8 * n * ld), hip.alloc(ld)
    hip.put(dz, lay(z))
    hip.put(dzo, lay(zold))
    fol = np.zeros(ld, np.uint8)
    fol[:M] = fo
    hip.put(dfo, fol)
    hip.put(df, np.full((n, ld), 7.0))
    rc = ctx.lib.gp_eval_residual_noalloc(ctx.h, dz, dzo, float(dt), dfo, df, None)
    assert rc == 0, ctx.lib.gp_last_error()
    out = np.zeros((n, ld))
    hip.get(df, out)
    assert np.all(out[:, M:] == 7.0)   # rows of scenarios >= M untouched
    return out[:, :M].T.copy()


def test_noalloc_residual_matches_eval_residual():
    hip = _Hip()
    a = get_case("case2_60bus")
    b = get_case("case1_9bus")
    ca = _setup(a, 8, 1)
    cb = _setup(b, 3, 2)
    try:
        for (ctx, z, zold, fo), dt in ((ca, a.dt), (cb, 0.0), (ca, b.dt)):
            ref = ctx.eval_residual(z, zold, dt, fo)
            got = _device_call(hip, ctx, z, zold, fo, dt)
            assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    finally:
        hip.free()
        ca[0].close()
        cb[0].close()


def test_noalloc_residual_refuses_contexts_beyond_its_scratch():
    from gradpower_hip.lib import GP_ERR_UNSUPPORTED, HipContext
    a = get_case("case2_60bus")
    M = 16384     # 2 x 18 clusters x 16384 scenarios x 8 bytes of injections > 4 MiB
    ctx = HipContext(a.desc, M)
    ctx.set_state(a.desc.z0, a.desc.p)
    ld = ctx.symbolic_info()["ld_scen"]
    hip = _Hip()
    try:
        z, f = hip.alloc(8 * a.desc.n * ld), hip.alloc(8 * a.desc.n * ld)   # real arrays of the right size
        rc = ctx.lib.gp_eval_residual_noalloc(ctx.h, z, z, 0.0, None, f, None)
        assert rc == GP_ERR_UNSUPPORTED
        assert b"maximum" in ctx.lib.gp_last_error()
    finally:
        hip.free()
        ctx.close()
