"""The allocation-free induced velocity (uvlm_induced_velocity_noalloc)
against the allocating entry point (uvlm_induced_velocity), through the
C-ABI with device pointers: bitwise the same results on two calls in a row,
the second smaller, in every precision; sizes beyond its fixed scratch are
rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from uvlm_amd import _native as N
from uvlm_amd import workloads as W

pytestmark = pytest.mark.gpu


def _lattice(rng, num, rows, cols):
    wing = np.stack([np.stack(np.meshgrid(np.linspace(0, 1, rows + 1),
                                          np.linspace(-2, 2, cols + 1), indexing="ij"), -1)
                     for _ in range(num)])
    nodes = np.concatenate([wing, 0.05 * rng.normal(size=wing.shape[:-1] + (1,))], -1)
    nodes[..., 0] += np.arange(num)[:, None, None] * 1.7
    gamma = rng.normal(size=(num, rows, cols))
    return np.ascontiguousarray(nodes), np.ascontiguousarray(gamma)


class _Dev:
    """Device buffers of the test (hipMalloc of the ROCm runtime the library
    uses; the test allocates, the entry point does not)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        return p

    def get(self, p, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _call_noalloc(ctx, targets, nodes, gamma, flags):
    d = _Dev()
    try:
        t = d.put(targets.reshape(-1, 3))
        nd, g = d.put(nodes), d.put(gamma)
        nt = targets.reshape(-1, 3).shape[0]
        out = d.put(np.full((nt, 3), np.nan))
        num, rows, cols = gamma.shape
        rc = N.gpu_lib().uvlm_induced_velocity_noalloc(ctx._h, t, nt, nd, g, num, rows, cols,
                                                       flags, out, None)
        assert rc == N.OK, ctx._L.uvlm_last_error(ctx._h).decode()
        assert ctx._L.uvlm_sync(ctx._h) == N.OK
        return d.get(out, (nt, 3))
    finally:
        d.free()


@pytest.mark.parametrize("flags", [N.PREC_F64_EXACT | N.SEQUENTIAL, N.PREC_F64, N.PREC_F32C])
def test_noalloc_matches_allocating_entry(ctx, flags):
    rng = np.random.default_rng(11)
    for num, rows, cols, nt in ((2, 16, 64, 20000), (1, 5, 12, 700)):
        print(f"[noalloc] flags {flags:#x} lattice {num}x{rows}x{cols}, {nt} targets", flush=True)
        nodes, gamma = _lattice(rng, num, rows, cols)
        targets = rng.uniform(-2, 4, size=(nt, 3))
        want = ctx.induced_velocity(targets, nodes, gamma, flags)
        got = _call_noalloc(ctx, targets, nodes, gamma, flags)
        np.testing.assert_array_equal(got, want)


def test_noalloc_no_sources_and_limits(ctx):
    rng = np.random.default_rng(5)
    targets = rng.uniform(size=(50, 3))
    nodes = np.zeros((1, 1, 5, 3))
    gamma = np.zeros((1, 0, 4))
    got = _call_noalloc(ctx, targets, nodes, gamma, N.PREC_F64)
    assert np.all(got == 0.0)
    # 2 * 600 * 600 chain elements of 64 bytes exceed the 16 MiB of records:
    # rejected on the host, nothing launched (the pointers are never read)
    rc = N.gpu_lib().uvlm_induced_velocity_noalloc(ctx._h, C.c_void_p(8), 10, C.c_void_p(8),
                                                   C.c_void_p(8), 1, 599, 599, N.PREC_F64,
                                                   C.c_void_p(8), None)
    assert rc == N.ERR_CAPACITY
    assert b"scratch" in ctx._L.uvlm_last_error(ctx._h)
