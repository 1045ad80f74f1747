"""cb_dots, the fused replacement of the reference's dot (misc.c:252-262),
against numpy's dot products of the same FP64 vectors: each result within
a few ulps of sum |a_i b_i| (a two-level reduction reorders the sum)."""
import numpy as np
import pytest

import cubens

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.mark.parametrize("n,length", [(8, 1_000_003), (3, 1000), (1, 1), (5, 0), (2, 200_000)])
def test_cb_dots_against_numpy(n, length):
    rng = np.random.default_rng(1000 + n + length)
    A = [rng.standard_normal(length) for _ in range(n)]
    B = [rng.standard_normal(length) for _ in range(n)]
    ta, tb = [_dev(a) for a in A], [_dev(b) for b in B]
    got = cubens.dots([t.data_ptr() for t in ta], [t.data_ptr() for t in tb], length)
    for k in range(n):
        want = float(A[k] @ B[k])
        bound = 64 * np.finfo(float).eps * float(np.abs(A[k] * B[k]).sum()) + 1e-300
        assert abs(got[k] - want) <= bound, (k, got[k], want)


def test_cb_dots_rejects_more_than_eight_products():
    t = _dev(np.ones(4))
    with pytest.raises(cubens.CubensError):
        cubens.dots([t.data_ptr()] * 9, [t.data_ptr()] * 9, 4)
