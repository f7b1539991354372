"""csd_zp_apply_dim_noalloc, the dense-map entry point that allocates nothing (its scratch is a
fixed device array per GPU), against csd_zp_apply_dim's int8-MFMA form through the C ABI: two
calls in a row, the second with a smaller input, must give the same residues; requests larger than
the scratch are rejected; the entry point reaches the context's MFMA kernel choice."""
import numpy as np
import pytest

from _pkg import load

pytestmark = pytest.mark.gpu

P57 = [144115176243904321, 144115172965982401, 144115188075855859]
SCRATCH = 64 << 20


def _uniform(rng, primes, shape):
    out = np.zeros(shape, dtype=np.int64)
    for l, p in enumerate(primes):
        out[:, l] = rng.integers(-2 * p, 2 * p, size=out[:, l].shape, dtype=np.int64, endpoint=True)
    return out


def _dev(be, arr):
    t = be.hal_allocate_i64(arr.shape)
    be.hal_import(t, np.ascontiguousarray(arr))
    return t


def _out(be, t):
    return be.hal_export(t).reshape(t.shape)


def test_noalloc_matches_allocating_form():
    pkg = load()
    be = pkg.HipBackend(0)
    try:
        rng = np.random.default_rng(2026)
        primes = P57
        k, m = 384, 200
        a = _uniform(rng, primes, (1, 3, m, k))
        mat = be.zp_matrix(_dev(be, a), [be.hal_create_modulus(p) for p in primes])
        # two calls in a row, the second smaller: 3 x 3 x 384 x 150 vectors (dim 2), then 1 x 3 x 40
        # vectors (dim 3); the scratch keeps the first call's planes, which the second must not see
        for shape, dim in [((3, 3, k, 150), 2), ((1, 3, 40, k), 3)]:
            x = _uniform(rng, primes, shape)
            out_shape = (shape[0], 3, m, shape[3]) if dim == 2 else (shape[0], 3, shape[2], m)
            xs = _dev(be, x)
            want, got = be.hal_allocate_i64(out_shape), be.hal_allocate_i64(out_shape)
            be.zp_apply_dim(mat, want, xs, dim, form="mfma")
            be.zp_apply_dim_noalloc(mat, got, xs, dim)
            w, g = _out(be, want), _out(be, got)
            assert (g == w).all(), (shape, dim)
            assert (g >= 0).all() and all((g[:, l] < p).all() for l, p in enumerate(primes))
        # a K-chunked call and another MFMA kernel of the context
        x = _uniform(rng, primes, (2, 3, 17, k))
        xs = _dev(be, x)
        want, got = be.hal_allocate_i64((2, 3, 17, m)), be.hal_allocate_i64((2, 3, 17, m))
        be.zp_apply_dim(mat, want, xs, 3, form="mfma", max_chunk=128)
        be.zp_set_kernel(1)
        try:
            be.zp_apply_dim_noalloc(mat, got, xs, 3, max_chunk=128)
        finally:
            be.zp_set_kernel(3)
        assert (_out(be, got) == _out(be, want)).all()
    finally:
        be.close()


def test_noalloc_rejects_oversize_and_bad_shapes():
    pkg = load()
    be = pkg.HipBackend(0)
    try:
        primes = P57[:2]
        k = 4096
        a = np.zeros((1, 2, 16, k), dtype=np.int64)
        mat = be.zp_matrix(_dev(be, a), [be.hal_create_modulus(p) for p in primes])
        # 2 moduli x 8 digit planes x 1088 vectors (padded to 64) x 4096 columns > 64 MiB
        nvec = 1025
        assert 2 * 8 * ((nvec + 63) // 64 * 64) * k > SCRATCH
        xs = be.hal_allocate_i64((1, 2, nvec, k))
        dst = be.hal_allocate_i64((1, 2, nvec, 16))
        with pytest.raises(pkg.HalError, match="scratch"):
            be.zp_apply_dim_noalloc(mat, dst, xs, 3)
        with pytest.raises(pkg.HalError):
            be.zp_apply_dim_noalloc(mat, be.hal_allocate_i64((1, 2, 4, 16)), be.hal_allocate_i64((1, 2, 4, k - 1)), 3)
        # the largest accepted size runs: 2 x 8 x 1024 x 4096 bytes is exactly the scratch
        xs = be.hal_allocate_i64((1, 2, 1024, k))
        dst = be.hal_allocate_i64((1, 2, 1024, 16))
        be.zp_apply_dim_noalloc(mat, dst, xs, 3)
        assert (_out(be, dst) == 0).all()
    finally:
        be.close()
