"""Exact dense maps over Z_p on the MI355X (csd_zp_* and csd_ring_transform_matrix through the C
ABI). Both forms, the int8-digit MFMA contraction and the plain VALU kernel, must give exactly
sum_k A[i][k] x[k] mod p in [0, p): checked against Python integers on ragged shapes and extreme
entries (the accumulator bounds), with forced K-chunking; against the NTT-based transforms of the
ring (equal mod p) at field_36864 size; the transform matrices' first columns against the
oracle's CRT kernel traces (ring.rs:1337-1415); and C^-1 C = 1 as a round trip."""
import numpy as np
import pytest

from _pkg import load
from oracle import ring as R
from test_hal_gpu import from_dev, to_dev

pytestmark = pytest.mark.gpu

P57 = [144115176243904321, 144115172965982401, 144115188075855859]
SMALL = [65537, 7681]


@pytest.fixture(scope="module")
def gpu():
    be = load().HipBackend(0)
    yield be
    be.close()


def exact(a, x, primes, dim):
    """dst[c, l, ..] = A_l . x[c, l, ..] along dim, with Python integers"""
    out = np.zeros(x.shape[:2] + ((a.shape[2], x.shape[3]) if dim == 2 else (x.shape[2], a.shape[2])), dtype=np.int64)
    for l, p in enumerate(primes):
        A = a[0, l].astype(object)
        for c in range(x.shape[0]):
            X = x[c, l].astype(object)
            Y = (A.dot(X) if dim == 2 else X.dot(A.T)) % p
            out[c, l] = Y.astype(np.int64)
    return out


def entries(rng, primes, shape, mode):
    out = np.zeros(shape, dtype=np.int64)
    for l, p in enumerate(primes):
        sub = out[:, l]
        if mode == "uniform":
            sub[...] = rng.integers(0, p, size=sub.shape, dtype=np.int64)
        elif mode == "wide":
            sub[...] = rng.integers(-2 * p, 2 * p, size=sub.shape, dtype=np.int64, endpoint=True)
        elif mode == "extreme":
            sub[...] = rng.choice(np.array([p - 1, -p, p, 2 * p, -2 * p, (p - 1) // 2, -(p - 1) // 2, 0], dtype=np.int64),
                                  size=sub.shape)
        elif mode == "max":
            sub[...] = (p - 1) // 2  # every centred product at its maximum: the accumulator bound
    return out


def run(gpu, a, x, primes, dim, form, max_chunk=0):
    mods = [gpu.hal_create_modulus(p) for p in primes]
    mat = gpu.zp_matrix(to_dev(gpu, a), mods)
    m = a.shape[2]
    shape = (x.shape[0], x.shape[1], m, x.shape[3]) if dim == 2 else (x.shape[0], x.shape[1], x.shape[2], m)
    dst = gpu.hal_allocate_i64(shape)
    apply(gpu, mat, dst, to_dev(gpu, x), dim, form, max_chunk)
    return from_dev(gpu, dst)


KERNELS = {"mfma_l2": 0, "mfma_lds32": 1, "mfma_wide": 2, "mfma_dense64": 4}


def apply(gpu, mat, dst, src, dim, form, max_chunk=0):
    """form "mfma" runs the default MFMA kernel (64 x 64 tiles staged in LDS for 8 waves), "mfma_l2"
    the one that loads its operands from L2 directly, "mfma_lds32" the 32 x 64 LDS-staged one,
    "mfma_wide" the 64 x 64 direct-to-LDS one on 4 waves"""
    if form not in KERNELS:
        gpu.zp_apply_dim(mat, dst, src, dim, form=form, max_chunk=max_chunk)
        return
    gpu.zp_set_kernel(KERNELS[form])
    try:
        gpu.zp_apply_dim(mat, dst, src, dim, form="mfma", max_chunk=max_chunk)
    finally:
        gpu.zp_set_kernel(3)


@pytest.mark.parametrize("m,k,nvec,dim,mode", [
    (20, 70, 33, 3, "uniform"), (12, 12, 5, 2, "uniform"), (64, 128, 64, 3, "wide"), (3, 1, 2, 3, "extreme"),
    (33, 200, 65, 2, "extreme"), (16, 64, 16, 3, "max"), (100, 37, 7, 3, "wide"),
    # maps of at most 32 x 32 take the one-thread-per-vector plain kernel: its largest size and
    # non-square ones
    (32, 32, 300, 2, "extreme"), (5, 9, 41, 3, "max"), (24, 17, 70, 3, "extreme")])
@pytest.mark.parametrize("form", ["mfma", "mfma_l2", "mfma_wide", "mfma_lds32", "plain"])
def test_zp_exact_small(gpu, m, k, nvec, dim, mode, form):
    rng = np.random.default_rng(m * 1000 + k + nvec)
    primes = P57[:2] + SMALL[:1]
    a = entries(rng, primes, (1, len(primes), m, k), mode)
    xs = (2, len(primes), k, (nvec + 1) // 2) if dim == 2 else (2, len(primes), (nvec + 1) // 2, k)
    x = entries(rng, primes, xs, mode)
    got = run(gpu, a, x, primes, dim, form)
    assert (got == exact(a, x, primes, dim)).all()


@pytest.mark.parametrize("max_chunk", [64, 128, 192])
def test_zp_chunked(gpu, max_chunk):
    """K split into chunks of max_chunk terms, each reduced and summed mod p"""
    rng = np.random.default_rng(max_chunk)
    primes = P57
    a = entries(rng, primes, (1, 3, 40, 300), "extreme")
    x = entries(rng, primes, (1, 3, 50, 300), "wide")
    want = exact(a, x, primes, 3)
    for form in ("mfma", "mfma_wide", "mfma_lds32"):
        assert (run(gpu, a, x, primes, 3, form, max_chunk) == want).all()


def test_zp_long_k_at_the_bound(gpu):
    """K = 3072 terms of maximal centred products: C_s and the i128 sum at their largest"""
    primes = P57[:1]
    a = np.full((1, 1, 16, 3072), (P57[0] - 1) // 2, dtype=np.int64)
    a[0, 0, 1::2] = -((P57[0] - 1) // 2)
    x = np.full((1, 1, 64, 3072), (P57[0] - 1) // 2, dtype=np.int64)
    x[0, 0, :, 5::7] = -((P57[0] - 1) // 2)
    want = exact(a, x, primes, 3)
    for form in ("mfma", "mfma_l2", "mfma_wide", "mfma_lds32", "plain"):
        assert (run(gpu, a, x, primes, 3, form) == want).all()


def test_zp_errors(gpu):
    pkg = load()
    mods = [gpu.hal_create_modulus(p) for p in P57[:2]]
    a = gpu.hal_allocate_i64((1, 2, 8, 8))
    mat = gpu.zp_matrix(a, mods)
    x = gpu.hal_allocate_i64((1, 2, 4, 8))
    with pytest.raises(pkg.HalError):
        gpu.zp_apply_dim(mat, x, x, 3)  # aliasing
    with pytest.raises(pkg.HalError):
        gpu.zp_apply_dim(mat, gpu.hal_allocate_i64((1, 2, 4, 9)), x, 3)  # row count mismatch
    with pytest.raises(pkg.HalError):
        gpu.zp_apply_dim(mat, gpu.hal_allocate_i64((1, 2, 4, 8)), x, 3, max_chunk=32)
    with pytest.raises(pkg.HalError):
        gpu.zp_matrix(gpu.hal_allocate_i64((1, 3, 8, 8)), mods)  # moduli count


# ---- the transform matrices at field_36864 -----------------------------------------------------

@pytest.fixture(scope="module")
def rings(gpu, cpu):
    pkg = load()
    g = pkg.ring.RingContext(gpu, "field_36864", (960, 980), 76)
    c = R.RingContext(cpu, R.FieldDescriptor("field_36864"), (960, 980), 76)
    yield g, c
    g.close()


BASE = [0, 5, 16]


@pytest.mark.parametrize("dim", [2, 3])
def test_transform_matrix_first_column_is_the_crt_kernel(gpu, rings, dim):
    """C[a][0] = k[a]: the oracle's CRT kernel traces, computed from the field without any NTT"""
    g, c = rings
    mat = from_dev(gpu, g.transform_matrix(BASE, dim))
    c.transform_cache(BASE)
    for l, i in enumerate(BASE):
        p = g.primes[i]
        traces = c._kernels[p][4 if dim == 2 else 5]
        assert [int(v) % p for v in mat[0, l, :, 0]] == [t % p for t in traces]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("form", ["mfma", "mfma_l2", "mfma_wide", "mfma_lds32", "plain"])
def test_dense_transform_equals_ntt_transform(gpu, rings, dim, form):
    """the dense map along one dimension equals that dimension's ntt_mul + intt pair mod p"""
    g, c = rings
    primes = [g.primes[i] for i in BASE]
    rng = np.random.default_rng(7 + dim)
    x = entries(rng, primes, (2, 3, 12, 3072), "uniform")
    mat = g.transform_matrix(BASE, dim)
    mods = [gpu.hal_create_modulus(p) for p in primes]
    zm = gpu.zp_matrix(mat, mods)
    dst = gpu.hal_allocate_i64(x.shape)
    apply(gpu, zm, dst, to_dev(gpu, x), dim, form)
    got = from_dev(gpu, dst)
    # the oracle's own sequence for one dimension
    cache = c.transform_cache(BASE)
    y = x.copy()
    if dim == 2:
        c.backend.hal_ntt_mul_mont64_dim(y, cache["ntts_lprime"], 2, cache["crt_kernel_lprime"], cache["mods"])
        c.backend.hal_intt_mont64_dim(y, cache["ntts_lprime"], 2)
    else:
        c.backend.hal_ntt_mul_mont64_dim(y, cache["ntts_k"], 3, cache["crt_kernel_k"], cache["mods"])
        c.backend.hal_intt_mont64_dim(y, cache["ntts_k"], 3)
    pv = np.array(primes, dtype=np.int64).reshape(1, -1, 1, 1)
    assert (got == np.mod(y, pv)).all()


def test_dense_two_sided_transform_equals_pb_to_crt(gpu, rings):
    """C_L' X C_K^T (two dense maps, MFMA) = pb_to_crt_tensor (the device NTT path) mod p"""
    g, c = rings
    primes = [g.primes[i] for i in BASE]
    rng = np.random.default_rng(99)
    x = entries(rng, primes, (1, 3, 12, 3072), "uniform")
    mods = [gpu.hal_create_modulus(p) for p in primes]
    zk = gpu.zp_matrix(g.transform_matrix(BASE, 3), mods)
    zl = gpu.zp_matrix(g.transform_matrix(BASE, 2), mods)
    t1 = gpu.hal_allocate_i64(x.shape)
    t2 = gpu.hal_allocate_i64(x.shape)
    gpu.zp_apply_dim(zk, t1, to_dev(gpu, x), 3)
    gpu.zp_apply_dim(zl, t2, t1, 2)
    d = to_dev(gpu, x)
    g.pb_to_crt_tensor(d, BASE)
    pv = np.array(primes, dtype=np.int64).reshape(1, -1, 1, 1)
    assert (from_dev(gpu, t2) == np.mod(from_dev(gpu, d), pv)).all()


def test_inverse_matrix_round_trip(gpu, rings):
    """crt_to_pb's matrix undoes pb_to_crt's: C_K^-1 (C_K x) = x mod p, at full size"""
    g, c = rings
    primes = [g.primes[i] for i in BASE]
    rng = np.random.default_rng(5)
    x = entries(rng, primes, (1, 3, 40, 3072), "uniform")
    mods = [gpu.hal_create_modulus(p) for p in primes]
    fwd = gpu.zp_matrix(g.transform_matrix(BASE, 3), mods)
    inv = gpu.zp_matrix(g.transform_matrix(BASE, 3, inverse=True), mods)
    t1 = gpu.hal_allocate_i64(x.shape)
    t2 = gpu.hal_allocate_i64(x.shape)
    gpu.zp_apply_dim(fwd, t1, to_dev(gpu, x), 3)
    gpu.zp_apply_dim(inv, t2, t1, 3)
    assert (from_dev(gpu, t2) == x).all()


def test_homcrt_block_contraction(gpu):
    """a |J_p| = 64 block of the homcrt cyclic transform (bootstrap/homcrt.rs:191-214) as a dense
    circulant with entries uniform mod p, on 512 vectors, both forms exact"""
    primes = P57
    rng = np.random.default_rng(64)
    first = entries(rng, primes, (1, 3, 64, 1), "uniform")[..., 0]
    idx = (np.arange(64)[:, None] - np.arange(64)[None, :]) % 64
    a = first[:, :, idx]
    x = entries(rng, primes, (1, 3, 512, 64), "uniform")
    want = exact(a, x, primes, 3)
    for form in ("mfma", "mfma_l2", "mfma_wide", "mfma_lds32", "plain"):
        assert (run(gpu, a, x, primes, 3, form) == want).all()


def _group_circulant(first, orders):
    """[1, r, n, n] with entry [a][b] = first[:, a - b], the difference in C_o0 x ... (row-major)"""
    n = int(np.prod(orders))
    strides = [int(np.prod(orders[f + 1:])) for f in range(len(orders))]
    idx = np.zeros((n, n), dtype=np.int64)
    ar = np.arange(n)[:, None]
    bc = np.arange(n)[None, :]
    for o, st in zip(orders, strides):
        idx += (((ar // st) % o - (bc // st) % o) % o) * st
    return first[:, :, idx]


@pytest.mark.parametrize("orders,nvec,mode", [((64,), 70, "uniform"), ((2, 64), 129, "extreme"), ((3, 2, 64), 64, "max"),
                                              ((2, 3, 64), 200, "uniform")])
def test_circulant_form_matches_dense(gpu, orders, nvec, mode):
    """the group-circulant MFMA form (A operands from the generator) gives the dense kernels' bits
    and the exact sums"""
    rng = np.random.default_rng(sum(orders) + nvec)
    primes = P57[:2] + SMALL[:1]
    n = int(np.prod(orders))
    first = entries(rng, primes, (1, len(primes), n, 1), mode)[..., 0]
    a = _group_circulant(first, list(orders))
    x = entries(rng, primes, (1, len(primes), nvec, n), mode)
    mods = [gpu.hal_create_modulus(p) for p in primes]
    zm = gpu.zp_matrix(to_dev(gpu, a), mods)
    want = exact(a, x, primes, 3)
    dense = gpu.hal_allocate_i64(x.shape)
    apply(gpu, zm, dense, to_dev(gpu, x), 3, "mfma_dense64")
    zm.set_circulant(list(orders))
    circ = gpu.hal_allocate_i64(x.shape)
    gpu.zp_apply_dim(zm, circ, to_dev(gpu, x), 3)
    assert (from_dev(gpu, dense) == want).all()
    assert (from_dev(gpu, circ) == want).all()
    # chunked K (two chunks of 64 terms for n = 128) takes the same path
    if n >= 128:
        gpu.zp_apply_dim(zm, circ, to_dev(gpu, x), 3, max_chunk=64)
        assert (from_dev(gpu, circ) == want).all()


def test_circulant_declaration_is_checked(gpu):
    """a matrix that is not the group-circulant of its first column, or of the wrong shape, is refused"""
    rng = np.random.default_rng(3)
    primes = P57[:1]
    first = entries(rng, primes, (1, 1, 128, 1), "uniform")[..., 0]
    a = _group_circulant(first, [2, 64])
    a[0, 0, 5, 77] = (a[0, 0, 5, 77] + 1) % primes[0]
    mods = [gpu.hal_create_modulus(p) for p in primes]
    zm = gpu.zp_matrix(to_dev(gpu, a), mods)
    with pytest.raises(load().HalError):
        zm.set_circulant([2, 64])
    with pytest.raises(load().HalError):
        zm.set_circulant([4, 32])
    with pytest.raises(load().HalError):
        zm.set_circulant([64])
