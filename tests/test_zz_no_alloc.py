"""fs_normal_equations_noalloc (the normal equations of caller-supplied device
arrays, partial sums in the library's static device scratch) against
fs_normal_equations on the same Jacobian and residual.

Two calls in a row on one context, the second on a smaller input (fewer rows
and columns, pointers of another context's buffers): with the same chunking
as the allocating entry point the products are bit-identical. A third input
has more row chunks than the allocating entry point's one per CU fits in the
scratch; its products then sum in another order and agree to 1e-12 of the
largest entry. Then the errors: more than 256 columns, and host memory in
place of device memory.
"""
import ctypes as ct

import numpy as np
import pytest

import forcesmith_amd as fa
from forcesmith_amd import datasets as dsm

pytestmark = pytest.mark.gpu

FS_BUF_RESIDUAL, FS_BUF_JACOBIAN = 0, 1


class _Hip:
    """Device buffers through the HIP runtime the library itself loaded (the
    caller's own allocations; the entry point under test allocates nothing)."""

    def __init__(self):
        fa.load_library()
        self.lib = ct.CDLL("libamdhip64.so.7")  # the loaded runtime, matched by soname
        self.lib.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.lib.hipFree.argtypes = [ct.c_void_p]
        self.lib.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.held = []

    def empty(self, n):
        p = ct.c_void_p()
        assert self.lib.hipMalloc(ct.byref(p), 8 * n) == 0
        self.held.append(p.value)
        return p.value

    def get(self, p, n):
        out = np.zeros(n)
        assert self.lib.hipMemcpy(out.ctypes.data, p, 8 * n, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def _buf(g, which):
    p, n = ct.c_void_p(), ct.c_int64()
    g._check(g._lib.fs_device_buffer(g._h, which, ct.byref(p), ct.byref(n)))
    return p.value


def _noalloc(g_ctx, g_data, hip):
    """fs_normal_equations_noalloc on g_ctx's context over g_data's device J
    and F; outputs in the caller's device buffers."""
    D, R = g_data.param_count, g_data.residual_count
    A, b, c = hip.empty(D * D), hip.empty(D), hip.empty(1)
    rc = g_ctx._lib.fs_normal_equations_noalloc(g_ctx._h, _buf(g_data, FS_BUF_JACOBIAN),
                                                _buf(g_data, FS_BUF_RESIDUAL), R, D, A, b, c)
    g_ctx._check(rc)
    g_ctx.synchronize()
    return hip.get(A, D * D).reshape(D, D), hip.get(b, D), float(hip.get(c, 1)[0])


def test_noalloc_normal_equations_match_the_allocating_entry_point():
    hip = _Hip()
    big = dsm.config2(n_configs=6, nmin=32, nmax=300)
    small = dsm.config1(n_configs=2)
    g1 = fa.GpuCalculator(big.dataset, big.model)
    g1.prepare()
    g2 = fa.GpuCalculator(small.dataset, small.model)
    g2.prepare()
    A1, b1, c1 = g1.normal_equations()  # leaves J and F of the same x in g1's buffers
    A2, b2, c2 = g2.normal_equations()
    print(f"inputs: {g1.residual_count} x {g1.param_count}, then {g2.residual_count} x {g2.param_count}",
          flush=True)
    for _ in range(2):  # the second round reuses the scratch the first left behind
        An, bn, cn = _noalloc(g1, g1, hip)
        assert np.array_equal(An, A1) and np.array_equal(bn, b1) and cn == c1
        An, bn, cn = _noalloc(g1, g2, hip)  # smaller input, same context
        assert np.array_equal(An, A2) and np.array_equal(bn, b2) and cn == c2
    hip.free()


def test_noalloc_more_chunks_than_the_scratch_holds():
    hip = _Hip()
    w = dsm.config2(n_configs=40, nmin=32, nmax=1200)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    A, b, c = g.normal_equations()
    assert g.residual_count > 64 * 256  # one chunk per CU needs more chunks than fit
    An, bn, cn = _noalloc(g, g, hip)
    hip.free()
    print(f"{g.residual_count} rows: max |dA| {np.abs(An - A).max():.3e} of {np.abs(A).max():.3e}",
          flush=True)
    assert np.array_equal(An, An.T)
    assert np.max(np.abs(An - A)) <= 1e-12 * np.max(np.abs(A))
    assert np.max(np.abs(bn - b)) <= 1e-12 * np.max(np.abs(b))
    assert cn == c


def test_noalloc_rejects_wide_and_host_arguments():
    hip = _Hip()
    w = dsm.config1(n_configs=1)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    g.normal_equations()
    J, F = _buf(g, FS_BUF_JACOBIAN), _buf(g, FS_BUF_RESIDUAL)
    p = hip.empty(257 * 257 + 300)
    rc = g._lib.fs_normal_equations_noalloc(g._h, J, F, 10, 257, p, p + 8 * 257 * 257, None)
    assert rc == -5  # FS_ERR_UNSUPPORTED
    host = np.zeros(64)
    rc = g._lib.fs_normal_equations_noalloc(g._h, J, F, 10, 4, host.ctypes.data, p, None)
    assert rc == -1  # FS_ERR_ARG
    assert "device memory" in g._lib.fs_last_error(g._h).decode()
    hip.free()
