"""ts_matrix_multiply_noalloc, the product that stages its vectors in the library's per-GPU scratch array
instead of allocating: bit-identical to ts_matrix_multiply (the same kernels) on two calls in a row, the
second on a smaller matrix, in both directions; an oversized matrix is rejected with a reason."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_csr(rng, rows, cols, per_row):
    counts = rng.integers(1, per_row + 1, rows)
    rs = np.zeros(rows + 1, np.int64)
    np.cumsum(counts, out=rs[1:])
    ci = np.concatenate([np.sort(rng.choice(cols, c, replace=False)) for c in counts]).astype(np.int32)
    va = rng.uniform(-1, 1, rs[-1])
    return rs.astype(np.int32), ci, va


def _noalloc(a, x, transpose):
    from tomostar_amd import _lib
    from tomostar_amd import lsqr as Q

    L = Q._declare()
    x2 = np.ascontiguousarray(np.atleast_2d(np.asarray(x, np.float64)))
    nout = a.cols if transpose else a.rows
    y = np.full((x2.shape[0], nout), np.nan)
    r = _lib.reason_buffer()
    rc = L.ts_matrix_multiply_noalloc(a.handle, x2.shape[0], _lib.ptr(x2), _lib.ptr(y), 1 if transpose else 0, r, len(r))
    return rc, r.value.decode(errors="replace"), y


@pytest.mark.parametrize("keep_fp64", [True, False])
def test_noalloc_product_equals_the_allocating_one(device, keep_fp64):
    from tomostar_amd import lsqr as Q

    rng = np.random.default_rng(11)
    for rows, cols, k in ((120_000, 50_000, 3), (9_000, 2_500, 2)):  # second call: smaller input
        print(f"noalloc product: {rows} x {cols}, k={k}, keep_fp64={keep_fp64}", flush=True)
        rs, ci, va = _random_csr(rng, rows, cols, 24)
        a = Q.Matrix(device, rows, cols, rs, ci, va, keep_fp64=keep_fp64)
        try:
            for transpose in (False, True):
                x = rng.standard_normal((k, rows if transpose else cols))
                want = Q.multiply(a, x, transpose)
                rc, why, got = _noalloc(a, x, transpose)
                assert rc == 0, why
                assert np.array_equal(got, want), f"transpose={transpose}: max diff {np.max(np.abs(got - want))}"
        finally:
            a.release()


def test_noalloc_product_rejects_a_matrix_beyond_its_scratch(device):
    from tomostar_amd import _lib
    from tomostar_amd import lsqr as Q

    rows, cols = 4_200_000, 16  # 2 x (rows + cols) doubles exceed the 8,388,608 of the array
    rs = np.arange(rows + 1, dtype=np.int32)
    ci = (np.arange(rows) % cols).astype(np.int32)
    a = Q.Matrix(device, rows, cols, rs, ci, np.ones(rows), keep_fp64=False)
    try:
        rc, why, _ = _noalloc(a, np.ones(cols), False)
        assert rc == _lib.TS_EINVAL and "limit" in why
    finally:
        a.release()
