"""The allocation-free slice inner-product matrix (gx_slice_inner_product_matrix_noalloc), called
through the C ABI: bitwise the result of gx_slice_inner_product_matrix, on two calls in a row
whose second input is smaller (the fixed workspace keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

import grid_amd as gx
import npref as R

pytestmark = pytest.mark.gpu


def _field(ctx, dims, Ls, subset, seed):
    a = gx.gaussian_fermion(dims, Ls, seed)
    if subset != gx.FULL:
        a = R.pick(a, dims, subset)
    return gx.Field(ctx, dims, Ls, gx.F64, subset).upload(a)


def _noalloc(x, y):
    m = np.zeros((x.Ls, x.Ls), dtype=np.complex128)
    rc = gx.lib().gx_slice_inner_product_matrix_noalloc(x.h, y.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, gx.lib().gx_last_error()
    return m


def test_noalloc_matches_allocating_entry_point():
    ctx = gx.Context(0)
    cases = [((8, 8, 6, 8), 12, gx.FULL, 501), ((4, 4, 2, 4), 5, gx.ODD, 502)]   # the second is smaller
    for dims, Ls, subset, seed in cases:
        print(f'noalloc slice ipm dims {dims} Ls {Ls} subset {subset}', flush=True)
        x, y = _field(ctx, dims, Ls, subset, seed), _field(ctx, dims, Ls, subset, seed + 1)
        got = _noalloc(x, y)
        ref = gx.slice_inner_product_matrix(x, y)
        assert np.array_equal(got, ref)
        assert np.array_equal(_noalloc(x, x), gx.slice_inner_product_matrix(x, x))
        x.close()
        y.close()
    big = gx.Field(ctx, (4, 4, 4, 4), 17, gx.F64, gx.FULL)
    rc = gx.lib().gx_slice_inner_product_matrix_noalloc(big.h, big.h, np.zeros(2 * 17 * 17).ctypes.data_as(
        ctypes.POINTER(ctypes.c_double)))
    assert rc == 1  # Ls beyond the workspace: GX_EINVAL, no device work
    big.close()
    ctx.close()
