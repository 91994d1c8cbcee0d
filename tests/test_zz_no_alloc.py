"""icx_matvec_noalloc (device staging from the library's per-GPU scratch array) against
icx_matvec (staging allocated per call), through the C ABI: bit-identical results on two
calls in a row, the second on a smaller job; FP64 and FP32 operators; a job over the
scratch limit is refused with ICX_ERR_ARG and leaves the next call correct."""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu

ERR_ARG = 1


def _job3(ctx, nx, ny, nz, seed):
    j = icx.cantilever_3d(ctx, nx, ny, nz)
    j.set_design(O.random_density(j.nelx, seed))
    return j


def _job2(ctx, nx, ny, seed):
    j = icx.cantilever_2d(ctx, nx, ny)
    j.set_design(O.random_density(j.nelx, seed))
    return j


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matvec_equals_allocating_matvec(ctx, prec):
    big = _job3(ctx, 30, 14, 12, 1)
    small = _job2(ctx, 20, 7, 2)
    rng = np.random.default_rng(prec)
    xb = rng.standard_normal(big.numeq)
    xs = rng.standard_normal(small.numeq)
    # two calls in a row, the second with a smaller input
    yb = big.matvec_noalloc(xb, prec)
    ys = small.matvec_noalloc(xs, prec)
    assert np.array_equal(yb, big.matvec(xb, prec))
    assert np.array_equal(ys, small.matvec(xs, prec))
    # and again in the other order: the array keeps the last call's contents, which
    # must not leak into the next result
    assert np.array_equal(small.matvec_noalloc(xs, prec), ys)
    assert np.array_equal(big.matvec_noalloc(xb, prec), yb)
    # the operator itself: FP64 product against the oracle's assembled matrix
    pr = O.cantilever_2d(20, 7)
    pr.assemble(O.random_density(pr.nelx, 2))
    ref = pr.spmv(xs)
    tol = 1e-13 if prec == 0 else 2e-5
    assert np.abs(ys - ref).max() <= tol * np.abs(ref).max()
    big.close()
    small.close()


def test_noalloc_matvec_refuses_jobs_over_the_scratch_limit(ctx):
    # 160 x 96 x 96 hexa8: prec 0 stages 8 numeq = 36 MB (fits 64 MiB); prec 1 adds
    # 8 N_grid + 4 nelx = 42 MB and is refused
    j = _job3(ctx, 160, 96, 96, 3)
    x = np.random.default_rng(5).standard_normal(j.numeq)
    with pytest.raises(icx.IcxError) as e:
        j.matvec_noalloc(x, 1)
    assert e.value.code == ERR_ARG and "scratch" in str(e.value)
    y = j.matvec_noalloc(x, 0)
    assert np.array_equal(y, j.matvec(x, 0))
    j.close()
