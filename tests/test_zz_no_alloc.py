"""The allocation-free batched tile LU (frk_dgetrf_nopiv_batched_noalloc)
against the allocating entry point frk_dgetrf_nopiv_batched, through the C ABI.

Two calls in a row share the library's fixed device scratch; the second is
smaller, so stale flags, info words or inverse blocks of the first would show.
Bar: the factors of both entry points agree to 1e-12 relative to the largest
entry of U (the dense getrf contract of test/dense_getrf_test.cpp compares
L U with A; the two entry points run different launch sequences for batch > 1,
so agreement is to rounding, not bitwise), and L U reproduces A to the same
tolerance.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import frank_oracle as fo

pytestmark = pytest.mark.gpu


def _tiles(torch, n, scales, zero_pivot=None):
    x = fo.get_sorted_random_vector(2 * n)
    a = fo.laplacend([x], 0, n, 0, n) + 1000.0 * np.eye(n)
    out = []
    for s in scales:
        t = s * a
        if zero_pivot is not None:
            t = np.eye(n)
            t[zero_pivot, zero_pivot] = 0.0
        out.append(torch.from_numpy(np.ascontiguousarray(t)).cuda())
    return out


def _call(fa, name, tiles):
    lib = fa.lib()
    fn = getattr(lib, name)
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_void_p]
    fn.restype = C.c_int
    ptrs = (C.c_void_p * len(tiles))(*[t.data_ptr() for t in tiles])
    info = np.zeros(len(tiles), dtype=np.int32)
    ctx = fa.default_context()
    rc = fn(ctx.h, ptrs, len(tiles), tiles[0].shape[0], tiles[0].stride(0),
            info.ctypes.data_as(C.c_void_p))
    return rc, info


def _check_pair(torch, fa, n, scales):
    print(f"no-alloc getrf n={n} batch={len(scales)}", flush=True)
    a_noalloc = _tiles(torch, n, scales)
    a_ref = [t.clone() for t in a_noalloc]
    host_a = [t.cpu().numpy() for t in a_noalloc]
    rc, info = _call(fa, "frk_dgetrf_nopiv_batched_noalloc", a_noalloc)
    assert rc == 0, fa.lib().frk_last_error()
    assert list(info) == [0] * len(scales)
    rc2, info2 = _call(fa, "frk_dgetrf_nopiv_batched", a_ref)
    assert rc2 == 0 and list(info2) == [0] * len(scales)
    torch.cuda.synchronize()
    for t, r, a in zip(a_noalloc, a_ref, host_a):
        lu, lu_ref = t.cpu().numpy(), r.cpu().numpy()
        scale = np.abs(np.triu(lu_ref)).max()
        np.testing.assert_allclose(lu, lu_ref, rtol=0, atol=1e-12 * scale)
        Lf, Uf = np.tril(lu, -1) + np.eye(n), np.triu(lu)
        np.testing.assert_allclose(Lf @ Uf, a, rtol=0, atol=1e-12 * np.abs(a).max())


def test_noalloc_getrf_two_calls_second_smaller(torch_gpu, fa):
    _check_pair(torch_gpu, fa, 512, (1.0, 2.0, 0.5))
    _check_pair(torch_gpu, fa, 200, (3.0, 1.5))


def test_noalloc_getrf_zero_pivot_and_limits(torch_gpu, fa):
    torch = torch_gpu
    print("no-alloc getrf zero pivot", flush=True)
    t = _tiles(torch, 70, (1.0, 1.0), zero_pivot=40)
    t[0] = _tiles(torch, 70, (1.0,))[0]
    rc, info = _call(fa, "frk_dgetrf_nopiv_batched_noalloc", t)
    assert rc == 1 and list(info) == [0, 41]
    print("no-alloc getrf size limit", flush=True)
    big = [torch.zeros((513, 513), dtype=torch.float64, device="cuda")]
    rc, _ = _call(fa, "frk_dgetrf_nopiv_batched_noalloc", big)
    assert rc == -3 and b"scratch" in fa.lib().frk_last_error()
