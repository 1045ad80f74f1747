"""cb_dots, the fused replacement of the reference's dot (misc.c:252-262),
against numpy's dot products of the same FP64 vectors: each result within
a few ulps of sum |a_i b_i| (a two-level reduction reorders the sum)."""
import numpy as np
import pytest

import cubens
from helpers import Hip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,length", [(8, 1_000_003), (3, 1000), (1, 1), (5, 0), (2, 200_000)])
def test_cb_dots_against_numpy(n, length):
    rng = np.random.default_rng(1000 + n + length)
    A = [rng.standard_normal(length) for _ in range(n)]
    B = [rng.standard_normal(length) for _ in range(n)]
    hip = Hip()
    got = cubens.dots([hip.upload(a) for a in A], [hip.upload(b) for b in B], length)
    hip.free_all()
    for k in range(n):
        want = float(A[k] @ B[k])
        bound = 64 * np.finfo(float).eps * float(np.abs(A[k] * B[k]).sum()) + 1e-300
        assert abs(got[k] - want) <= bound, (k, got[k], want)


def test_cb_dots_rejects_more_than_eight_products():
    hip = Hip()
    t = hip.upload(np.ones(4))
    with pytest.raises(cubens.CubensError):
        cubens.dots([t] * 9, [t] * 9, 4)
    hip.free_all()
