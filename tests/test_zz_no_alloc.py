"""esdg_reduce_max_noalloc (include/esdg.h): the allocation-free max reduction, called through the C
ABI, against esdg_reduce_max on the same device arrays -- two calls in a row, the second with a
smaller input whose maximum lies below values the first call saw, so a partial left over from the
first call would show.  Max is exact, so the results are equal bit for bit."""
import ctypes as C

import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, RunParams


def noalloc_fn():
    from esdgsem_mpiocca_amd import device_api as D
    L = D.lib()
    fn = L.esdg_reduce_max_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    return L, fn


def test_noalloc_rejects_sizes_beyond_its_scratch():
    """Argument checks come before any device call: sizes above the maximum and negative sizes
    are rejected with a message (runs without a GPU too)."""
    L, fn = noalloc_fn()
    cap = 4194304
    dummy = C.c_void_p(256)     # never dereferenced: the size check rejects the call first
    assert fn(None, cap + 1, dummy, dummy) == -1
    assert b"exceeds the maximum" in L.esdg_last_error(None)
    assert fn(None, -1, dummy, dummy) == -1


@pytest.mark.gpu
def test_noalloc_equals_reduce_max_on_two_calls():
    import esdgsem_mpiocca_amd as E
    dev = E.device()
    L, fn = noalloc_fn()
    case = Case.gaussian_bump(N=2, nel=2)
    ctx = dev.Context(dev.make_params(case, RunParams()))
    try:
        rng = np.random.default_rng(11)
        n1 = 3_000_001                                    # 11719 partials
        x = rng.uniform(0.0, 1.0, n1)
        x[n1 - 5] = 7.5                                   # the first call's max, in its last block
        n2 = 1_000_003                                    # second call: a prefix, max 2.25
        x[12345] = 2.25
        dx = ctx.array(x)
        out_ref, out_na = ctx.empty(1), ctx.empty(1)
        stream = L.esdg_stream(C.c_void_p(ctx.handle))
        for n, expect in ((n1, 7.5), (n2, 2.25)):
            ctx.reduce_max(n, dx, out_ref)
            rc = fn(C.c_void_p(stream), n, C.c_void_p(dx.ptr), C.c_void_p(out_na.ptr))
            assert rc == 0, L.esdg_last_error(None)
            ctx.finish()
            a, b = out_ref.download()[0], out_na.download()[0]
            assert a == b == expect == x[:n].max()
        # the largest size it accepts, and the empty input (max over nothing is 0, as esdg_reduce_max)
        big = rng.uniform(0.0, 1.0, 4194304)
        dbig = ctx.array(big)
        assert fn(C.c_void_p(stream), big.size, C.c_void_p(dbig.ptr), C.c_void_p(out_na.ptr)) == 0
        ctx.finish()
        assert out_na.download()[0] == big.max()
        assert fn(C.c_void_p(stream), 0, C.c_void_p(dbig.ptr), C.c_void_p(out_na.ptr)) == 0
        ctx.reduce_max(0, dbig, out_ref)
        ctx.finish()
        assert out_na.download()[0] == out_ref.download()[0] == 0.0
    finally:
        ctx.close()
