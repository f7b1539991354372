"""Parity of the gfx950 HAL (libcsdbfv_hal.so, through its C ABI) with the oracle CpuBackend:
every op, on seeded random inputs, with broadcasting and strided views, compared bit for bit
(i64 and i128 representatives, L0). The oracle itself is pinned by test_oracle_*.py."""
import numpy as np
import pytest

from _pkg import load
from oracle import cpu_backend as cb
from oracle.cpu_backend import I128

pytestmark = pytest.mark.gpu

Q57 = 144115188075835777
Q62_RADIX2 = 4611686018427379009
Q61_RADIX3 = 3074457345618256921
P36864 = [144115176243904321, 144115172965982401]


@pytest.fixture(scope="module")
def gpu():
    pkg = load()
    be = pkg.HipBackend(0)
    yield be
    be.close()


def to_dev(gpu, arr):
    if arr.dtype == I128:
        t = gpu.hal_allocate_i128(arr.shape)
        gpu.import_i128(t, np.ascontiguousarray(arr))
    else:
        t = gpu.hal_allocate_i64(arr.shape)
        gpu.hal_import(t, np.ascontiguousarray(arr))
    return t


def from_dev(gpu, t):
    if t.elem == 16:
        return gpu.export_i128(t).reshape(t.shape)
    return gpu.hal_export(t).reshape(t.shape)


def rand_i64(rng, shape, lo=None, hi=None):
    if lo is None:
        return rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, size=shape, dtype=np.int64, endpoint=True)
    return rng.integers(lo, hi, size=shape, dtype=np.int64, endpoint=True)


def rand_i128(rng, shape, bits=127):
    out = np.zeros(shape, dtype=I128)
    out["lo"] = rng.integers(0, np.iinfo(np.uint64).max, size=shape, dtype=np.uint64, endpoint=True)
    hi = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, size=shape, dtype=np.int64, endpoint=True)
    if bits < 127:
        hi = hi >> (127 - bits)
    out["hi"] = hi
    return out


def same(a, b):
    if a.dtype == I128:
        return (a["lo"] == b["lo"]).all() and (a["hi"] == b["hi"]).all()
    return (a == b).all()


def mods_for(be, primes):
    return [be.hal_create_modulus(p) for p in primes]


SHAPE = (3, 4, 5, 70)


@pytest.mark.parametrize("src_shape", [(3, 4, 5, 70), (1, 4, 1, 1), (3, 1, 5, 1), (1, 1, 1, 70), (3, 4, 5, 1)])
@pytest.mark.parametrize("op", ["hal_assign", "hal_add_i64", "hal_sub_i64"])
def test_binary_i64(gpu, cpu, op, src_shape):
    rng = np.random.default_rng(1)
    d = rand_i64(rng, SHAPE)
    s = rand_i64(rng, src_shape)
    dc = d.copy()
    getattr(cpu, op)(dc, s)
    dg = to_dev(gpu, d)
    getattr(gpu, op)(dg, to_dev(gpu, s))
    assert same(from_dev(gpu, dg), dc)


@pytest.mark.parametrize("op", ["hal_assign_narrow", "hal_assign_i128", "hal_add_i128"])
def test_binary_i128(gpu, cpu, op):
    rng = np.random.default_rng(2)
    for src_shape in [(3, 4, 5, 70), (1, 4, 5, 1), (3, 1, 1, 70)]:
        s = rand_i128(rng, src_shape)
        d = rand_i64(rng, SHAPE) if op == "hal_assign_narrow" else rand_i128(rng, SHAPE)
        dc = d.copy()
        getattr(cpu, op)(dc, s)
        dg = to_dev(gpu, d)
        getattr(gpu, op)(dg, to_dev(gpu, s))
        assert same(from_dev(gpu, dg), dc)


@pytest.mark.parametrize("op", ["hal_fma_i64", "hal_mul_i64"])
def test_fma_mul(gpu, cpu, op):
    rng = np.random.default_rng(3)
    for ls, rs in [((3, 4, 5, 70), (3, 4, 5, 70)), ((1, 4, 5, 70), (3, 4, 1, 1)), ((3, 4, 5, 1), (1, 1, 1, 70))]:
        l, r = rand_i64(rng, ls), rand_i64(rng, rs)
        d = rand_i128(rng, SHAPE)
        dc = d.copy()
        getattr(cpu, op)(dc, l, r)
        dg = to_dev(gpu, d)
        getattr(gpu, op)(dg, to_dev(gpu, l), to_dev(gpu, r))
        assert same(from_dev(gpu, dg), dc)


@pytest.mark.parametrize("terms", [1, 5, 23, 30])
def test_inner_product(gpu, cpu, terms):
    rng = np.random.default_rng(4 + terms)
    shape = (2, 3, 4, 600)
    lhs = [rand_i64(rng, (1, 3, 4, 600) if t % 3 == 0 else shape) for t in range(terms)]
    rhs = [rand_i64(rng, (2, 3, 4, 1) if t % 2 == 0 else shape) for t in range(terms)]
    d = rand_i128(rng, shape)
    dc = d.copy()
    cpu.hal_inner_product_i64(dc, lhs, rhs)
    dg = to_dev(gpu, d)
    gpu.hal_inner_product_i64(dg, [to_dev(gpu, x) for x in lhs], [to_dev(gpu, x) for x in rhs])
    assert same(from_dev(gpu, dg), dc)


def test_inner_product_small_lines(gpu, cpu):
    rng = np.random.default_rng(9)
    shape = (2, 3, 4, 5)
    lhs = [rand_i64(rng, shape) for _ in range(3)]
    rhs = [rand_i64(rng, (1, 3, 1, 5)) for _ in range(3)]
    dc = np.zeros(shape, dtype=I128)
    cpu.hal_inner_product_i64(dc, lhs, rhs)
    dg = gpu.hal_allocate_i128(shape)
    gpu.hal_inner_product_i64(dg, [to_dev(gpu, x) for x in lhs], [to_dev(gpu, x) for x in rhs])
    assert same(from_dev(gpu, dg), dc)


@pytest.mark.parametrize("lshape,rshape,dshape", [
    ((1, 21, 17, 1), (4, 17, 3, 300), (4, 21, 3, 300)),
    ((4, 9, 7, 1), (1, 7, 2, 130), (4, 9, 2, 130)),
    ((1, 3, 5, 1), (2, 5, 2, 4), (2, 3, 2, 4)),
])
def test_matmul(gpu, cpu, lshape, rshape, dshape):
    rng = np.random.default_rng(5)
    l = rand_i64(rng, lshape)
    l[l % 7 == 0] = 0  # the reference skips zero coefficients; results must not depend on it
    r = rand_i64(rng, rshape)
    dc = rand_i128(rng, dshape)
    dg = to_dev(gpu, dc)
    cpu.hal_matmul_i64(dc, l, r)
    gpu.hal_matmul_i64(dg, to_dev(gpu, l), to_dev(gpu, r))
    assert same(from_dev(gpu, dg), dc)


def test_matmul_sliced_operands(gpu, cpu):
    """The bconv chunking slices the matrix along dim 2 and the source along dim 1."""
    rng = np.random.default_rng(6)
    l = rand_i64(rng, (1, 21, 17, 1))
    r = rand_i64(rng, (4, 17, 3, 96))
    dc = np.zeros((4, 21, 3, 96), dtype=I128)
    cpu.hal_matmul_i64(dc, l[:, :, 5:12], r[:, 5:12])
    lg, rg = to_dev(gpu, l), to_dev(gpu, r)
    dg = gpu.hal_allocate_i128((4, 21, 3, 96))
    gpu.hal_matmul_i64(dg, lg[:, :, 5:12], rg[:, 5:12])
    assert same(from_dev(gpu, dg), dc)


PRIMES = [97, 65537, Q57, 144115188075855823]


@pytest.mark.parametrize("op", ["hal_mul_mont64"])
def test_mul_mont(gpu, cpu, op):
    rng = np.random.default_rng(7)
    mods_c, mods_g = mods_for(cpu, PRIMES), mods_for(gpu, PRIMES)
    for src_shape in [(3, 4, 5, 70), (1, 4, 1, 1), (1, 4, 5, 1), (1, 4, 1, 70)]:
        bound = [(1 << 63) // p for p in PRIMES]
        d = rand_i64(rng, SHAPE)
        s = np.zeros(src_shape, dtype=np.int64)
        for j, p in enumerate(PRIMES):
            s[:, j] = rng.integers(-p, p, size=s[:, j].shape, endpoint=True)
        dc = d.copy()
        getattr(cpu, op)(dc, s, mods_c)
        dg = to_dev(gpu, d)
        getattr(gpu, op)(dg, to_dev(gpu, s), mods_g)
        assert same(from_dev(gpu, dg), dc)
        del bound


@pytest.mark.parametrize("op", ["hal_add_red_mont64_double", "hal_red_mont64_double", "hal_add_red_mont64",
                                "hal_red_mont64", "hal_i128_to_mont64"])
def test_mod_reductions(gpu, cpu, op):
    rng = np.random.default_rng(8)
    mods_c, mods_g = mods_for(cpu, PRIMES), mods_for(gpu, PRIMES)
    # the reductions are exact wrapping formulas, so parity holds on any i128 input (in and out
    # of the |x| <= 2^63 N contract); both ranges are exercised
    for src_shape, bits in [(SHAPE, 127), ((1, 4, 5, 1), 127), (SHAPE, 100)]:
        s = rand_i128(rng, src_shape, bits=bits)
        d = rand_i64(rng, SHAPE)
        dc = d.copy()
        getattr(cpu, op)(dc, s, mods_c)
        dg = to_dev(gpu, d)
        getattr(gpu, op)(dg, to_dev(gpu, s), mods_g)
        assert same(from_dev(gpu, dg), dc)


def test_i64_to_mont(gpu, cpu):
    rng = np.random.default_rng(10)
    mods_c, mods_g = mods_for(cpu, PRIMES), mods_for(gpu, PRIMES)
    d = rand_i64(rng, SHAPE)
    dc = d.copy()
    cpu.hal_i64_to_mont64(dc, mods_c)
    dg = to_dev(gpu, d)
    gpu.hal_i64_to_mont64(dg, mods_g)
    assert same(from_dev(gpu, dg), dc)


def test_strided_views(gpu, cpu):
    """Ops on sliced destination/source views touch exactly the viewed elements."""
    rng = np.random.default_rng(11)
    mods_c, mods_g = mods_for(cpu, PRIMES[:2]), mods_for(gpu, PRIMES[:2])
    d = rand_i64(rng, (4, 6, 5, 130))
    s = rand_i64(rng, (4, 6, 5, 130), -97, 97)
    dc = d.copy()
    cpu.hal_mul_mont64(dc[1:3, 2:4, 1:4, 7:120], s[1:3, 0:2, 0:3, 3:116], mods_c)
    cpu.hal_add_i64(dc[:, 4:], s[:, :2])
    dg, sg = to_dev(gpu, d), to_dev(gpu, s)
    gpu.hal_mul_mont64(dg[1:3, 2:4, 1:4, 7:120], sg[1:3, 0:2, 0:3, 3:116], mods_g)
    gpu.hal_add_i64(dg[:, 4:], sg[:, :2])
    assert same(from_dev(gpu, dg), dc)
    # import/export of strided views
    view = dg[1:3, 1:5, 2:4, 10:100]
    assert (gpu.hal_export(view) == dc[1:3, 1:5, 2:4, 10:100].reshape(-1)).all()
    patch = rand_i64(rng, (2, 4, 2, 90))
    gpu.hal_import(view, patch)
    cpu.hal_import(dc[1:3, 1:5, 2:4, 10:100], patch)
    assert same(from_dev(gpu, dg), dc)


def test_bconv_correct_mont(gpu, cpu):
    rng = np.random.default_rng(12)
    mods_c, mods_g = mods_for(cpu, PRIMES), mods_for(gpu, PRIMES)
    for shift in [64, 66, 90]:
        for rshape in [(1, 4, 1, 1), (3, 4, 1, 1)]:
            lhs = rand_i128(rng, (3, 1, 5, 70), bits=120)
            rhs = np.zeros(rshape, dtype=np.int64)
            for j, p in enumerate(PRIMES):
                rhs[:, j] = rng.integers(-p, p, size=rhs[:, j].shape, endpoint=True)
            d = rand_i64(rng, SHAPE, -(1 << 60), 1 << 60)
            dc = d.copy()
            cpu.hal_bconv_correct_mont64(dc, lhs, rhs, mods_c, shift)
            dg = to_dev(gpu, d)
            gpu.hal_bconv_correct_mont64(dg, to_dev(gpu, lhs), to_dev(gpu, rhs), mods_g, shift)
            assert same(from_dev(gpu, dg), dc)


def test_bconv_correct_pow2(gpu, cpu):
    rng = np.random.default_rng(13)
    src = rand_i128(rng, (3, 2, 5, 70), bits=100)
    for shift, k in [(1, 1), (20, 50), (64, 62), (100, 30)]:
        dc = np.zeros((3, 1, 5, 70), dtype=np.int64)
        cpu.hal_bconv_correct_pow2(dc, src, 123456789, -987654321, shift, k)
        dg = gpu.hal_allocate_i64((3, 1, 5, 70))
        gpu.hal_bconv_correct_pow2(dg, to_dev(gpu, src), 123456789, -987654321, shift, k)
        assert same(from_dev(gpu, dg), dc)
    src3 = rand_i128(rng, (3, 3, 5, 70), bits=110)
    for shift, k in [(7, 100), (64, 126), (90, 64)]:
        dc = np.zeros((3, 2, 5, 70), dtype=np.int64)
        cpu.hal_bconv_correct_pow2_wide(dc, src3, (1 << 70) + 5, -(3 << 90) + 7, shift, k)
        dg = gpu.hal_allocate_i64((3, 2, 5, 70))
        gpu.hal_bconv_correct_pow2_wide(dg, to_dev(gpu, src3), (1 << 70) + 5, -(3 << 90) + 7, shift, k)
        assert same(from_dev(gpu, dg), dc)


@pytest.mark.parametrize("inner,shifts", [([3, 4], [1, 2]), ([4, 3], [3, 1]), ([2, 3, 8, 64], [1, 2, 5, 63]), ([12], [7])])
def test_permute(gpu, cpu, inner, shifts):
    rng = np.random.default_rng(14)
    n = int(np.prod(inner))
    pc, pg = cpu.hal_create_permutation(inner, shifts), gpu.hal_create_permutation(inner, shifts)
    for shape, dim in [((2, 3, 4, n), 3), ((2, 3, n, 5), 2)]:
        s = rand_i64(rng, shape)
        dc = np.zeros(shape, dtype=np.int64)
        cpu.hal_permute(dc, s, pc, dim)
        dg = gpu.hal_allocate_i64(shape)
        gpu.hal_permute(dg, to_dev(gpu, s), pg, dim)
        assert same(from_dev(gpu, dg), dc)


NTT_CASES = [
    (Q57, [4, 3]), (Q57, [16, 2]), (Q57, [27, 1]), (Q57, [4, 9]), (Q57, [1, 1]), (Q57, []), (Q57, [2]),
    (Q57, [8, 9, 2]), (Q62_RADIX2, [16, 2]), (Q62_RADIX2, [16]), (Q61_RADIX3, [27, 1]), (Q61_RADIX3, [81]),
    (P36864[0], [2, 3, 8, 64]), (P36864[1], [4, 3]), (Q57, [3, 64, 64]),
]


@pytest.mark.parametrize("q,inner", NTT_CASES)
@pytest.mark.parametrize("dim", [2, 3])
def test_ntt_parity(gpu, cpu, q, inner, dim):
    """Forward (reduced), inverse and fused ntt_mul, bit-exact, with inputs in [-2q, 2q]."""
    n = int(np.prod(inner)) if inner else 1
    other = 5 if n >= 1024 else 37
    shape = (2, 2, n, other) if dim == 2 else (2, 2, other if n < 1024 else 3, n)
    rng = np.random.default_rng(15 + n)
    data = rand_i64(rng, shape, -2 * q, 2 * q)
    data[0, 0].flat[:4] = [2 * q, -2 * q, q, -q]
    mc, mg = cpu.hal_create_modulus(q), gpu.hal_create_modulus(q)
    pc, pg = cpu.hal_create_ntt_plan(mc, inner), gpu.hal_create_ntt_plan(mg, inner)
    kshape = (1, 2, n, 1) if dim == 2 else (1, 2, 1, n)
    kern = rand_i64(rng, kshape, -q, q)
    for mode in ["fwd", "inv", "mul"]:
        dc = data.copy()
        dg = to_dev(gpu, data)
        if mode == "fwd":
            cpu.hal_ntt_mont64_dim(dc, [pc, pc], dim)
            gpu.hal_ntt_mont64_dim(dg, [pg, pg], dim)
        elif mode == "inv":
            cpu.hal_intt_mont64_dim(dc, [pc, pc], dim)
            gpu.hal_intt_mont64_dim(dg, [pg, pg], dim)
        else:
            cpu.hal_ntt_mul_mont64_dim(dc, [pc, pc], dim, kern, [mc, mc])
            gpu.hal_ntt_mul_mont64_dim(dg, [pg, pg], dim, to_dev(gpu, kern), [mg, mg])
        assert same(from_dev(gpu, dg), dc), mode


def test_ntt_per_limb_plans(gpu, cpu):
    """Different plans (primes) per dimension-1 index, on a strided sub-view."""
    primes = [P36864[0], P36864[1], 144115188075849889]
    mc = [cpu.hal_create_modulus(p) for p in primes]
    mg = [gpu.hal_create_modulus(p) for p in primes]
    pcs = [cpu.hal_create_ntt_plan(m, [4, 3]) for m in mc]
    pgs = [gpu.hal_create_ntt_plan(m, [4, 3]) for m in mg]
    rng = np.random.default_rng(16)
    data = rand_i64(rng, (3, 5, 12, 100), -2 * primes[1], 2 * primes[1])
    dc, dg = data.copy(), to_dev(gpu, data)
    cpu.hal_ntt_mont64_dim(dc[:, 1:4, :, 10:90], pcs, 2)
    gpu.hal_ntt_mont64_dim(dg[:, 1:4, :, 10:90], pgs, 2)
    cpu.hal_intt_mont64_dim(dc[1:, 2:5], pcs, 2)
    gpu.hal_intt_mont64_dim(dg[1:, 2:5], pgs, 2)
    assert same(from_dev(gpu, dg), dc)


def test_errors_are_reported(gpu):
    pkg = load()
    t = gpu.hal_allocate_i64((2, 2, 3, 4))
    with pytest.raises(pkg.HalError):
        gpu.hal_add_i64(t, gpu.hal_allocate_i64((2, 3, 3, 4)))
    with pytest.raises(pkg.HalError):
        gpu.hal_mul_mont64(t, t, [gpu.hal_create_modulus(97)])
    with pytest.raises(pkg.HalError):
        gpu.hal_create_modulus(96)
    plan = gpu.hal_create_ntt_plan(gpu.hal_create_modulus(Q57), [4])
    with pytest.raises(pkg.HalError):
        gpu.hal_ntt_mont64_dim(t, [plan, plan], 2)


@pytest.mark.parametrize("q,inner", [(P36864[0], [2, 3, 8, 64]), (P36864[1], [4, 3]), (Q62_RADIX2, [16, 2]), (Q61_RADIX3, [81])])
def test_generic_ntt_kernel_parity(gpu, cpu, q, inner):
    """The one-stage-per-pass LDS kernel (the fallback for plans without a register-blocked
    variant) stays bit-exact when forced for every plan."""
    pkg = load()
    n = int(np.prod(inner))
    mc, mg = cpu.hal_create_modulus(q), gpu.hal_create_modulus(q)
    pc, pg = cpu.hal_create_ntt_plan(mc, inner), gpu.hal_create_ntt_plan(mg, inner)
    rng = np.random.default_rng(31)
    gpu.ntt_force_generic(1)
    try:
        for dim in (2, 3):
            shape = (2, 1, n, 9) if dim == 2 else (2, 1, 3, n)
            data = rand_i64(rng, shape, -2 * q, 2 * q)
            kern = rand_i64(rng, (1, 1, n, 1) if dim == 2 else (1, 1, 1, n), -q, q)
            dc, dg = data.copy(), to_dev(gpu, data)
            cpu.hal_ntt_mul_mont64_dim(dc, [pc], dim, kern, [mc])
            gpu.hal_ntt_mul_mont64_dim(dg, [pg], dim, to_dev(gpu, kern), [mg])
            cpu.hal_intt_mont64_dim(dc, [pc], dim)
            gpu.hal_intt_mont64_dim(dg, [pg], dim)
            assert same(from_dev(gpu, dg), dc)
    finally:
        gpu.ntt_force_generic(0)


@pytest.mark.parametrize("lines,threads", [(0, 0), (2, 512), (1, 128), (3, 192), (4, 512), (1, 64), (8, 256)])
def test_row_launch_shapes_parity(gpu, cpu, lines, threads):
    """csd_ntt_set_row_shape: every dimension-3 launch shape of the compiled field_36864 kernels is
    bit-exact with the oracle, including a ragged last workgroup (5 lines per limb)"""
    pkg = load()
    primes, _ = pkg.ring.sample_primes("field_36864", (960, 980), 76)
    q = primes[5]
    inner = [2, 3, 8, 64]
    n = int(np.prod(inner))
    mc, mg = cpu.hal_create_modulus(q), gpu.hal_create_modulus(q)
    pc, pg = cpu.hal_create_ntt_plan(mc, inner), gpu.hal_create_ntt_plan(mg, inner)
    rng = np.random.default_rng(lines * 1000 + threads)
    gpu.ntt_set_row_shape(lines, threads)
    try:
        data = rand_i64(rng, (2, 1, 5, n), -2 * q, 2 * q)
        kern = rand_i64(rng, (1, 1, 1, n), -q, q)
        dc, dg = data.copy(), to_dev(gpu, data)
        cpu.hal_ntt_mul_mont64_dim(dc, [pc], 3, kern, [mc])
        gpu.hal_ntt_mul_mont64_dim(dg, [pg], 3, to_dev(gpu, kern), [mg])
        assert same(from_dev(gpu, dg), dc)
        cpu.hal_intt_mont64_dim(dc, [pc], 3)
        gpu.hal_intt_mont64_dim(dg, [pg], 3)
        assert same(from_dev(gpu, dg), dc)
        cpu.hal_ntt_mont64_dim(dc, [pc], 3)
        gpu.hal_ntt_mont64_dim(dg, [pg], 3)
        assert same(from_dev(gpu, dg), dc)
    finally:
        gpu.ntt_set_row_shape(0, 0)


def test_row_launch_shape_errors(gpu):
    pkg = load()
    for lines, threads in [(-1, 0), (9, 0), (1, 96), (1, 1024), (0, -64)]:
        with pytest.raises(pkg.HalError):
            gpu.ntt_set_row_shape(lines, threads)


@pytest.mark.parametrize("field,log2_q,inner", [("field_36864", (960, 980), [2, 3, 8, 64]),
                                                ("field_49152", (1280, 1300), [3, 64, 64])])
@pytest.mark.parametrize("mode", [0, 2, 3])
def test_compiled_row_kernels_parity(gpu, cpu, field, log2_q, inner, mode):
    """Dimension-3 transforms of the benchmark fields' K parts: the compiled-shape row kernels
    (mode 0, the default) and the plan-driven row kernel (mode 2) are both bit-exact with the
    oracle in all three directions, on several lines per workgroup and inputs in [-2q, 2q]."""
    pkg = load()
    primes, _ = pkg.ring.sample_primes(field, log2_q, 76)
    q = primes[3]
    n = int(np.prod(inner))
    mc, mg = cpu.hal_create_modulus(q), gpu.hal_create_modulus(q)
    pc, pg = cpu.hal_create_ntt_plan(mc, inner), gpu.hal_create_ntt_plan(mg, inner)
    rng = np.random.default_rng(n + mode)
    gpu.ntt_force_generic(mode)
    try:
        data = rand_i64(rng, (2, 1, 3, n), -2 * q, 2 * q)
        kern = rand_i64(rng, (1, 1, 1, n), -q, q)
        dc, dg = data.copy(), to_dev(gpu, data)
        cpu.hal_ntt_mont64_dim(dc, [pc], 3)
        gpu.hal_ntt_mont64_dim(dg, [pg], 3)
        assert same(from_dev(gpu, dg), dc)
        cpu.hal_intt_mont64_dim(dc, [pc], 3)
        gpu.hal_intt_mont64_dim(dg, [pg], 3)
        assert same(from_dev(gpu, dg), dc)
        cpu.hal_ntt_mul_mont64_dim(dc, [pc], 3, kern, [mc])
        gpu.hal_ntt_mul_mont64_dim(dg, [pg], 3, to_dev(gpu, kern), [mg])
        assert same(from_dev(gpu, dg), dc)
        cpu.hal_intt_mont64_dim(dc, [pc], 3)
        gpu.hal_intt_mont64_dim(dg, [pg], 3)
        assert same(from_dev(gpu, dg), dc)
    finally:
        gpu.ntt_force_generic(0)


@pytest.mark.parametrize("shape,prime", [((4, 3), P36864[0]), ((2, 3, 8, 64), P36864[0]), ((3, 64, 64), None),
                                         ((4,), None), ((8, 9, 2), None)])
def test_plan_schedule_counts_match_oracle(gpu, shape, prime):
    """The device plan's schedule (csd_ntt_plan_mults) performs exactly the Barrett
    multiplications the oracle executes per transform, in each direction; the oracle's schedule
    is pinned against the reference's stated rules in test_oracle_hal."""
    from test_oracle_hal import _largest_57bit_prime

    n = prime or _largest_57bit_prime(int(np.lcm.reduce(shape)))
    plan = gpu.hal_create_ntt_plan(gpu.hal_create_modulus(n), list(shape))
    oplan = cb.CpuNTTPlan(cb.CpuBackend().hal_create_modulus(n), shape)
    rng = np.random.default_rng(3)
    for d, (inverse, reduce_output) in {"fwd": (False, True), "fwd_unreduced": (False, False),
                                        "inv": (True, False)}.items():
        data = rng.integers(-2 * n, 2 * n, size=oplan.len, dtype=np.int64)
        _, mults = cb.plan_apply_traced(oplan, data, 1, 1, inverse, reduce_output)
        assert plan.mults(d) == mults, (shape, d)
