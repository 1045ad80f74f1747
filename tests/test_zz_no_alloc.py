"""cb_dots_noalloc (partial sums in a static device array, nothing allocated)
against cb_dots (partial sums in a buffer it allocates) through the C ABI:
the same kernels over the same grid, so the results are bitwise equal, on two
calls in a row, the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest

import cubens
from helpers import Hip

pytestmark = pytest.mark.gpu


def _noalloc():
    return cubens.lib().cb_dots_noalloc


def test_noalloc_equals_cb_dots_on_two_calls():
    f = _noalloc()
    hip = Hip()
    stream = hip.stream()
    for n, length, seed in ((8, 1_000_003, 7), (3, 4099, 8)):
        print(f"[noalloc] n={n} length={length}", flush=True)
        rng = np.random.default_rng(seed)
        A = [rng.standard_normal(length) for _ in range(n)]
        B = [rng.standard_normal(length) for _ in range(n)]
        pa = [hip.upload(a) for a in A]
        pb = [hip.upload(b) for b in B]
        want = cubens.dots(pa, pb, length)
        out = hip.upload(np.full(n, np.nan))
        rc = f(n, (C.c_void_p * n)(*pa), (C.c_void_p * n)(*pb), length, C.c_void_p(out), C.c_void_p(stream))
        assert rc == 0, cubens.lib().cb_last_error()
        hip.sync(stream)
        got = hip.download(out, n)
        assert np.array_equal(got, want), (got, want)
        ref = np.array([float(a @ b) for a, b in zip(A, B)])
        assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max() + 1e-12
    hip.free_all()


def test_noalloc_rejects_more_than_eight_products():
    f = _noalloc()
    hip = Hip()
    t = hip.upload(np.ones(4))
    out = hip.upload(np.zeros(9))
    ptrs = (C.c_void_p * 9)(*([t] * 9))
    rc = f(9, ptrs, ptrs, 4, C.c_void_p(out), None)
    assert rc != 0
    assert b"1 to 8" in cubens.lib().cb_last_error()
    hip.free_all()
