"""Handle lifetimes through the C ABI: a context, matrix or QR factor may be
destroyed in any order (a garbage collector finalizing a reference cycle
does not order its finalizers)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_context_destroyed_before_its_matrices(torch_gpu, fa):
    from frank_amd import _lib as L
    ctx = fa.Context()
    x = np.sort(np.random.default_rng(0).random(256))
    A = fa.BlrMatrix.from_kernel([x], 4, eps=1e-8, kcap=32, ctx=ctx)
    B = A.clone()
    F = A.qr()
    L.lib().frk_ctx_destroy(ctx.h)
    ctx.h = None
    assert A.ranks().shape == (4, 4)  # still usable: the matrix holds its context
    L.lib().frk_blr_destroy(A.h)
    A.h = None
    from frank_amd.blr import _hptr
    T = np.empty((64, 64))
    tau = np.empty(64)
    L.check(L.lib().frk_qrf_get_t(F.h, 1, 0, _hptr(T), _hptr(tau)))  # the factors hold their matrix
    assert np.isfinite(T).all()
    F.close()
    B.close()
