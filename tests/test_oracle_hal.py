"""Pins the oracle HAL (oracle/oracle_hal.cpp) against the reference's own unit tests for the hot
path: hal/ntt.rs:660-941 and hal/cpu.rs:1127-1875, plus closed forms and brute force. These run on
the CPU only."""
import numpy as np
import pytest

from conftest import Lcg
from oracle import cpu_backend as cb
from oracle.cpu_backend import I128, i128_from_ints, i128_to_ints

Q57 = 144115188075835777
Q62_RADIX2 = 4611686018427379009
Q61_RADIX3 = 3074457345618256921
I64_MAX = (1 << 63) - 1
I64_MIN = -(1 << 63)


def flat(shape, idx):
    return ((idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]) * shape[3] + idx[3]


def random_entry(rng, bound):
    """ntt.rs:681-684"""
    r = (rng() << 62) ^ (rng() << 31) ^ rng()
    return (r % (2 * bound + 1)) - bound


def montgomery_repr(x, p):
    return (x % p) * ((1 << 64) % p) % p


# ---------------------------------------------------------------------------------------------
# hal/mont.rs


def test_montgomery_constants():
    """-N^-1 is an inverse mod 2^64, R^2 and R^3 are the residues of 2^128, 2^192 (with the
    reference's `pmod` mapping 0 to N)."""
    be = cb.CpuBackend()
    for n in [97, 65537, Q57, 144115188075855823, (1 << 57) - 13 if False else 144115176243904321]:
        m = be.hal_create_modulus(n)
        assert (n * m.neg_inv_N) % (1 << 64) == (1 << 64) - 1
        assert m.r2 == ((1 << 128) % n or n)
        assert m.r3 == ((1 << 192) % n or n)


def test_red128_contract():
    """red128 maps |x| <= 2^63 N + N into [-N, N], congruent to 2^-64 x; double_red128 works for
    any i128 and is congruent to 2^-128 x (mont.rs:38-66)."""
    be = cb.CpuBackend()
    rng = Lcg(5)
    for n in [97, 65537, Q57]:
        m = be.hal_create_modulus(n)
        rinv = pow(1 << 64, -1, n)
        xs = [0, 1, -1, (1 << 63) * n + n, -((1 << 63) * n + n), (1 << 63) * n, n, -n]
        for _ in range(50):
            xs.append(random_entry(rng, (1 << 63) * n))
        for x in xs:
            out = cb.red128(m, x)
            assert -n <= out <= n
            assert out % n == x * rinv % n
        ys = [(1 << 127) - 1, -(1 << 127), 12345 << 100, -(98765 << 90)]
        for _ in range(50):
            ys.append(((rng() << 96) ^ (rng() << 64) ^ (rng() << 32) ^ rng()) - (1 << 124))
        for y in ys:
            out = cb.double_red128(m, y)
            assert -n <= out <= n
            assert out % n == y * rinv * rinv % n


def test_mont_ops(cpu):
    """hal/cpu.rs:1456-1581"""
    be = cpu
    mods = [be.hal_create_modulus(97), be.hal_create_modulus(65537)]
    shape = [2, 2, 3, 4]
    rng = Lcg(17)
    a_data = [rng() % 200 - 100 for _ in range(48)]
    b_data = [rng() % 200 - 100 for _ in range(48)]

    def to_mont128(m, x):
        # montgomery form via the public ops: red128(red128(x) * R^3)
        return cb.red128(m, cb.red128(m, x) * m.r3)

    dst = be.hal_allocate_i64(shape)
    be.hal_import(dst, [to_mont128(mods[(pos // 12) % 2], a_data[pos]) for pos in range(48)])
    src = be.hal_allocate_i64(shape)
    be.hal_import(src, b_data)
    be.hal_mul_mont64(dst, src, mods)
    out = dst.reshape(-1)
    for pos in range(48):
        p = mods[(pos // 12) % 2].N
        assert abs(int(out[pos])) <= p
        assert int(out[pos]) % p == (a_data[pos] * b_data[pos]) % p

    src128_data = [(x - 24) * 1000000007 * 1000000007 for x in range(48)]
    src128 = i128_from_ints(src128_data, shape)
    acc = np.full(shape, 1000, dtype=np.int64)
    be.hal_add_red_mont64_double(acc, src128, mods)
    accf = acc.reshape(-1)
    for pos in range(48):
        p = mods[(pos // 12) % 2].N
        red = int(accf[pos]) - 1000
        assert abs(red) <= p
        assert montgomery_repr(montgomery_repr(red, p), p) == src128_data[pos] % p

    over = np.full(shape, 1000, dtype=np.int64)
    be.hal_red_mont64_double(over, src128, mods)
    assert (over == acc - 1000).all()

    mont = be.hal_allocate_i64(shape)
    be.hal_i128_to_mont64(mont, src128, mods)
    mf = mont.reshape(-1)
    for pos in range(48):
        p = mods[(pos // 12) % 2].N
        assert abs(int(mf[pos])) <= p
        assert int(mf[pos]) % p == montgomery_repr(src128_data[pos], p)

    acc = np.full(shape, 1000, dtype=np.int64)
    be.hal_add_red_mont64(acc, src128, mods)
    accf = acc.reshape(-1)
    for pos in range(48):
        p = mods[(pos // 12) % 2].N
        red = int(accf[pos]) - 1000
        assert abs(red) <= p
        assert montgomery_repr(red, p) == src128_data[pos] % p
    over = np.full(shape, 1000, dtype=np.int64)
    be.hal_red_mont64(over, src128, mods)
    assert (over == acc - 1000).all()

    plain = [(x - 24) * 187654321987654321 for x in range(48)]
    plain = [((v + (1 << 63)) % (1 << 64)) - (1 << 63) for v in plain]
    inplace = be.hal_allocate_i64(shape)
    be.hal_import(inplace, plain)
    be.hal_i64_to_mont64(inplace, mods)
    f = inplace.reshape(-1)
    for pos in range(48):
        p = mods[(pos // 12) % 2].N
        assert abs(int(f[pos])) <= p
        assert int(f[pos]) % p == montgomery_repr(plain[pos], p)


# ---------------------------------------------------------------------------------------------
# hal/ntt.rs


def test_barrett_mul():
    """ntt.rs:696-717"""
    for n in [97, 65537, Q57, Q62_RADIX2, Q61_RADIX3]:
        rng = Lcg(1)
        values = [0, 1, -1, n - 1, -(n - 1), (n - 1) // 2, -((n - 1) // 2)]
        values += [random_entry(rng, n - 1) for _ in range(20)]
        inputs = [0, 1, -1, I64_MAX, I64_MIN, I64_MIN + 1, n, -n, 2 * min(n, I64_MAX // 2), 123456789123456789 % n]
        for z in values:
            zu = cb.barrett_new(z, n)
            assert abs(zu[0]) <= (n - 1) // 2
            for x in inputs:
                out = cb.barrett_mul(zu, x, n)
                assert abs(out) <= n
                assert (z * x) % n == out % n


def test_barrett_constant_closed_form():
    """u = round(-z 2^64 / N), halves away from zero (ntt.rs:118-121, 142)."""
    for n in [97, Q57]:
        for z in [1, 2, -2, (n - 1) // 2, -((n - 1) // 2), 5]:
            zz, u = cb.barrett_new(z, n)
            num = -(zz << 64)
            exact = abs(num) * 2 + n
            mag = exact // (2 * n)
            assert u == (mag if num >= 0 else -mag)


@pytest.mark.parametrize("q,lens", [(Q57, [2, 4, 16, 128, 3, 9, 81]), (Q62_RADIX2, [2, 4, 16]), (Q61_RADIX3, [3, 9, 27, 81])])
def test_forward_is_permuted_dft(q, lens):
    """ntt.rs:719-773: the forward transform evaluates at all len-th roots of unity."""
    for n in lens:
        root = cb.find_root_of_unity(q, n)
        assert pow(root, n, q) == 1 and pow(root, n // (2 if n & (n - 1) == 0 else 3), q) != 1
        roots = [pow(root, e, q) for e in range(n)]
        x_poly = np.zeros(n, dtype=np.int64)
        x_poly[1] = 1
        cb.factor_forward(q, n, x_poly, in_bound=2 * q)
        leaf = [roots.index(int(v) % q) for v in x_poly]
        assert len(set(leaf)) == n
        rng = Lcg(42 + n)
        data = [random_entry(rng, 2 * q) for _ in range(n)]
        tr = np.array(data, dtype=np.int64)
        ob = cb.factor_forward(q, n, tr, in_bound=2 * q)
        assert ob <= I64_MAX
        for pos, e in enumerate(leaf):
            assert abs(int(tr[pos])) <= ob
            expected = sum(x * roots[(e * j) % n] for j, x in enumerate(data)) % q
            assert int(tr[pos]) % q == expected


@pytest.mark.parametrize("q,lens", [(Q57, [2, 4, 16, 128, 3, 9, 81]), (Q62_RADIX2, [2, 4, 16]), (Q61_RADIX3, [3, 9, 27, 81])])
def test_factor_roundtrip(q, lens):
    """ntt.rs:775-804: backward . forward = len * identity, at the extreme +-2q inputs too."""
    for n in lens:
        rng = Lcg(7 + n)
        extreme = [2 * q, -2 * q] + [[2 * q, -2 * q, 0, q, -q][rng() % 5] for _ in range(2, n)]
        rand = [random_entry(rng, 2 * q) for _ in range(n)]
        for data in (extreme, rand):
            tr = np.array(data, dtype=np.int64)
            b = cb.factor_forward(q, n, tr, in_bound=2 * q)
            b = cb.factor_backward(q, n, tr, in_bound=b)
            assert b <= I64_MAX
            for pos in range(n):
                assert int(tr[pos]) % q == (data[pos] * n) % q


def test_strided_vectorized(cpu):
    """ntt.rs:806-869: strided/vectorized == contiguous, exactly."""
    q = Q57
    for n in [16, 27]:
        es, vl = 7, 3
        rng = Lcg(99)
        strided = np.zeros(n * es, dtype=np.int64)
        cont = [np.zeros(n, dtype=np.int64) for _ in range(vl)]
        for i in range(n):
            for l in range(vl):
                v = random_entry(rng, 2 * q)
                strided[i * es + l] = v
                cont[l][i] = v
        cb.factor_forward(q, n, strided, es=es, vl=vl, in_bound=2 * q)
        for c in cont:
            cb.factor_forward(q, n, c, in_bound=2 * q)
        for i in range(n):
            for l in range(vl):
                assert cont[l][i] == strided[i * es + l]
    plan = cpu.hal_create_ntt_plan(cpu.hal_create_modulus(q), [4, 3, 2])
    n, es, vl = 24, 7, 3
    rng = Lcg(77)
    strided = np.zeros(n * es, dtype=np.int64)
    cont = [np.zeros(n, dtype=np.int64) for _ in range(vl)]
    for i in range(n):
        for l in range(vl):
            v = random_entry(rng, 2 * q)
            strided[i * es + l] = v
            cont[l][i] = v
    for inverse in (False, True):
        cb.plan_apply(plan, strided, es=es, vl=vl, inverse=inverse, reduce_output=True)
        for c in cont:
            cb.plan_apply(plan, c, inverse=inverse, reduce_output=True)
        for i in range(n):
            for l in range(vl):
                assert cont[l][i] == strided[i * es + l]


@pytest.mark.parametrize("q,inner", [(Q57, [16, 2]), (Q57, [27, 1]), (Q57, [4, 9]), (Q57, [1, 1]),
                                     (Q62_RADIX2, [16, 2]), (Q61_RADIX3, [27, 1])])
def test_plan_roundtrip_and_convolution(cpu, q, inner):
    """ntt.rs:871-941, against a brute-force group convolution."""
    n = inner[0] * inner[1]
    plan = cpu.hal_create_ntt_plan(cpu.hal_create_modulus(q), inner)
    rng = Lcg(5)
    a = [random_entry(rng, 2 * q) for _ in range(n)]
    b = [random_entry(rng, q // 4) for _ in range(n)]
    expected = [0] * n
    for i0 in range(inner[0]):
        for i1 in range(inner[1]):
            for j0 in range(inner[0]):
                for j1 in range(inner[1]):
                    d = ((i0 + j0) % inner[0]) * inner[1] + (i1 + j1) % inner[1]
                    expected[d] = (expected[d] + a[i0 * inner[1] + i1] * b[j0 * inner[1] + j1]) % q
    data = np.array(a, dtype=np.int64)
    cb.plan_apply(plan, data, inverse=False)
    assert all(abs(int(x)) <= q for x in data)
    cb.plan_apply(plan, data, inverse=True)
    for pos in range(n):
        assert abs(int(data[pos])) <= q and (int(data[pos]) - a[pos]) % q == 0
    ah, bh = np.array(a, dtype=np.int64), np.array(b, dtype=np.int64)
    cb.plan_apply(plan, ah)
    cb.plan_apply(plan, bh)
    prod = np.array([int(x) * int(y) % q for x, y in zip(ah, bh)], dtype=np.int64)
    cb.plan_apply(plan, prod, inverse=True)
    for pos in range(n):
        assert int(prod[pos]) % q == expected[pos]
    red, unred = np.array(a, dtype=np.int64), np.array(a, dtype=np.int64)
    cb.plan_apply(plan, red, reduce_output=True)
    cb.plan_apply(plan, unred, reduce_output=False)
    assert all((int(x) - int(y)) % q == 0 for x, y in zip(red, unred))


# ---------------------------------------------------------------------------------------------
# hal/cpu.rs


def test_slice_import_export(cpu):
    """cpu.rs:1141-1186"""
    shape = [2, 3, 4, 5]
    t = cpu.hal_allocate_i64(shape)
    data = list(range(120))
    cpu.hal_import(t, data)
    assert (t.reshape(-1) == np.arange(120)).all()
    view = t[1:2, 0:2, 2:, 1:4]
    assert view.shape == (1, 2, 2, 3)
    out = cpu.hal_export(view)
    exp = [data[flat(shape, [i, j, k, l])] for i in range(1, 2) for j in range(2) for k in range(2, 4) for l in range(1, 4)]
    assert list(out) == exp
    t2 = t.copy()
    patch = list(range(1000, 1012))
    cpu.hal_import(t2[1:2, 0:2, 2:4, 1:4], patch)
    for i in range(2):
        for j in range(3):
            for k in range(4):
                for l in range(5):
                    pos = flat(shape, [i, j, k, l])
                    if i == 1 and j < 2 and 2 <= k < 4 and 1 <= l < 4:
                        assert t2.reshape(-1)[pos] == patch[(j * 2 + (k - 2)) * 3 + (l - 1)]
                    else:
                        assert t2.reshape(-1)[pos] == data[pos]


def test_add_sub_broadcast(cpu):
    """cpu.rs:1218-1262"""
    shape = [2, 2, 3, 4]
    dst = cpu.hal_allocate_i64(shape)
    data = list(range(48))
    cpu.hal_import(dst, data)
    sj = cpu.hal_allocate_i64([1, 2, 1, 1])
    cpu.hal_import(sj, [10, 20])
    cpu.hal_add_i64(dst, sj)
    sl = cpu.hal_allocate_i64([2, 2, 3, 1])
    sld = [x * 100 for x in range(12)]
    cpu.hal_import(sl, sld)
    cpu.hal_sub_i64(dst, sl)
    for i in range(2):
        for j in range(2):
            for k in range(3):
                for l in range(4):
                    pos = flat(shape, [i, j, k, l])
                    assert dst.reshape(-1)[pos] == data[pos] + [10, 20][j] - sld[(i * 2 + j) * 3 + k]
    d128 = i128_from_ints([1] * 8, [2, 1, 2, 2])
    s128 = i128_from_ints([5, 6, 7, 8], [1, 1, 2, 2])
    cpu.hal_add_i128(d128, s128)
    assert i128_to_ints(d128) == [6, 7, 8, 9, 6, 7, 8, 9]


def test_fma_and_matmul(cpu):
    """cpu.rs:1264-1350"""
    shape = [2, 2, 3, 4]
    rng = Lcg(42)
    ld = [rng() % 1000 - 500 for _ in range(48)]
    rd = [rng() % 1000 - 500 for _ in range(48)]
    lhs, rhs = cpu.hal_allocate_i64(shape), cpu.hal_allocate_i64(shape)
    cpu.hal_import(lhs, ld)
    cpu.hal_import(rhs, rd)
    acc = cpu.hal_allocate_i128(shape)
    cpu.hal_fma_i64(acc, lhs, rhs)
    cpu.hal_fma_i64(acc, lhs, rhs)
    assert i128_to_ints(acc) == [2 * a * b for a, b in zip(ld, rd)]
    prod = i128_from_ints([-77] * 48, shape)
    cpu.hal_mul_i64(prod, lhs, rhs)
    assert i128_to_ints(prod) == [a * b for a, b in zip(ld, rd)]
    I, J, K, L, U = 2, 3, 2, 4, 5
    dshape, lshape, rshape = [I, J, K, L], [1, J, U, 1], [I, U, K, L]
    mld = [rng() % 100 - 50 for _ in range(J * U)]
    mrd = [rng() % 100 - 50 for _ in range(I * U * K * L)]
    ml, mr = cpu.hal_allocate_i64(lshape), cpu.hal_allocate_i64(rshape)
    cpu.hal_import(ml, mld)
    cpu.hal_import(mr, mrd)
    dst = i128_from_ints([-1] * 48, dshape)
    cpu.hal_matmul_i64(dst, ml, mr)
    cpu.hal_matmul_i64(dst, ml, mr)
    out = i128_to_ints(dst)
    for i in range(I):
        for j in range(J):
            for k in range(K):
                for l in range(L):
                    e = sum(mld[j * U + u] * mrd[flat(rshape, [i, u, k, l])] for u in range(U))
                    assert out[flat(dshape, [i, j, k, l])] == e
    lrs, rbs = [I, J, U, 1], [1, U, K, L]
    lrd = [rng() % 100 - 50 for _ in range(I * J * U)]
    rbd = [rng() % 100 - 50 for _ in range(U * K * L)]
    lr, rb = cpu.hal_allocate_i64(lrs), cpu.hal_allocate_i64(rbs)
    cpu.hal_import(lr, lrd)
    cpu.hal_import(rb, rbd)
    cpu.hal_matmul_i64(dst, lr, rb)
    out = i128_to_ints(dst)
    for i in range(I):
        for j in range(J):
            for k in range(K):
                for l in range(L):
                    e = sum(lrd[(i * J + j) * U + u] * rbd[flat(rbs, [0, u, k, l])] for u in range(U))
                    assert out[flat(dshape, [i, j, k, l])] == e


def test_inner_product(cpu):
    """cpu.rs:1352-1454"""
    shape = [2, 2, 3, 4]
    n = 5
    rng = Lcg(99)
    ld = [[rng() % 1000 - 500 for _ in range(48)] for _ in range(n)]
    rd = [[rng() % 1000 - 500 for _ in range(48)] for _ in range(n)]
    lhs, rhs = [], []
    for t in range(n):
        a, b = cpu.hal_allocate_i64(shape), cpu.hal_allocate_i64(shape)
        cpu.hal_import(a, ld[t])
        cpu.hal_import(b, rd[t])
        lhs.append(a)
        rhs.append(b)
    dst = i128_from_ints([-123] * 48, shape)
    cpu.hal_inner_product_i64(dst, lhs, rhs)
    assert i128_to_ints(dst) == [sum(ld[t][p] * rd[t][p] for t in range(n)) for p in range(48)]
    chain = i128_from_ints([7] * 48, shape)
    cpu.hal_mul_i64(chain, lhs[0], rhs[0])
    for t in range(1, n):
        cpu.hal_fma_i64(chain, lhs[t], rhs[t])
    assert i128_to_ints(chain) == i128_to_ints(dst)
    lbd = [rng() % 200 - 100 for _ in range(24)]
    rbd = [rng() % 200 - 100 for _ in range(12)]
    lb, rb = cpu.hal_allocate_i64([1, 2, 3, 4]), cpu.hal_allocate_i64([2, 2, 3, 1])
    cpu.hal_import(lb, lbd)
    cpu.hal_import(rb, rbd)
    bdst = cpu.hal_allocate_i128(shape)
    cpu.hal_inner_product_i64(bdst, [lb, lhs[0]], [rb, rhs[0]])
    out = i128_to_ints(bdst)
    for i in range(2):
        for j in range(2):
            for k in range(3):
                for l in range(4):
                    pos = flat(shape, [i, j, k, l])
                    e = lbd[flat([1, 2, 3, 4], [0, j, k, l])] * rbd[flat([2, 2, 3, 1], [i, j, k, 0])] + ld[0][pos] * rd[0][pos]
                    assert out[pos] == e


def test_assign_narrow(cpu):
    """cpu.rs:1583-1627"""
    shape = [2, 2, 3, 4]
    sd = [(x - 24) * 123456789123 for x in range(48)]
    src = i128_from_ints(sd, shape)
    dst = cpu.hal_allocate_i64(shape)
    cpu.hal_assign_narrow(dst, src)
    assert list(dst.reshape(-1)) == sd
    bc = i128_from_ints([x + 100 for x in range(6)], [1, 2, 3, 1])
    cpu.hal_assign_narrow(dst, bc)
    for i in range(2):
        for j in range(2):
            for k in range(3):
                for l in range(4):
                    assert dst.reshape(-1)[flat(shape, [i, j, k, l])] == j * 3 + k + 100
    d128 = i128_from_ints([-1] * 48, shape)
    cpu.hal_assign_i128(d128, src)
    assert i128_to_ints(d128) == sd
    bl = i128_from_ints([(x + 100) << 80 for x in range(6)], [1, 2, 3, 1])
    cpu.hal_assign_i128(d128, bl)
    out = i128_to_ints(d128)
    for i in range(2):
        for j in range(2):
            for k in range(3):
                for l in range(4):
                    assert out[flat(shape, [i, j, k, l])] == (j * 3 + k + 100) << 80


def test_bconv_correct_mont(cpu):
    """cpu.rs:1629-1690"""
    mods = [cpu.hal_create_modulus(97), cpu.hal_create_modulus(65537)]
    shape = [2, 2, 2, 3]
    shift = 66
    small = [3, -7, 12, 0, 25, -1, 8, 4, -13, 6, 2, 9]
    lhs = i128_from_ints([(s << shift) + ((s % 5) << 20) + 12345 for s in small], [2, 1, 2, 3])
    for rhs_shape, rd in (([1, 2, 1, 1], [1001, -4242]), ([2, 2, 1, 1], [1001, -4242, 77, -33])):
        rhs = cpu.hal_allocate_i64(rhs_shape)
        cpu.hal_import(rhs, rd)
        dst = np.full(shape, 500, dtype=np.int64)
        cpu.hal_bconv_correct_mont64(dst, lhs, rhs, mods, shift)
        for i in range(2):
            for j in range(2):
                p = mods[j].N
                for k in range(2):
                    for l in range(3):
                        out = int(dst.reshape(-1)[flat(shape, [i, j, k, l])]) - 500
                        assert abs(out) <= p
                        r = rd[j] if rhs_shape[0] == 1 else rd[i * 2 + j]
                        assert montgomery_repr(out, p) == (small[(i * 2 + k) * 3 + l] * r) % p


def test_bconv_correct_pow2_closed_form(cpu):
    """cpu.rs:983-1057 against the closed form cmod((v - rounddiv(c, 2^s) f0) f1, 2^k)."""
    rng = Lcg(3)
    shape = [2, 1, 2, 3]
    vals = [((rng() << 64) ^ (rng() << 32) ^ rng()) - (1 << 80) for _ in range(12)]
    corr = [((rng() << 40) ^ rng()) - (1 << 60) for _ in range(12)]
    src = i128_from_ints([x for i in range(2) for x in (vals[i * 6:(i + 1) * 6] + corr[i * 6:(i + 1) * 6])], [2, 2, 2, 3])
    f0, f1, shift, k = 123456789, -987654321, 20, 50
    dst = cpu.hal_allocate_i64(shape)
    cpu.hal_bconv_correct_pow2(dst, src, f0, f1, shift, k)
    for pos in range(12):
        q = (corr[pos] + (1 << (shift - 1))) >> shift
        red = ((vals[pos] - q * f0) * f1) % (1 << k)
        exp = red - (1 << k) if red > (1 << (k - 1)) else red
        assert int(dst.reshape(-1)[pos]) == exp
    # wide variant: 126-bit modulus split into two limbs
    src3 = i128_from_ints([x for i in range(2) for x in (vals[i * 6:(i + 1) * 6] + [v >> 3 for v in vals[i * 6:(i + 1) * 6]] + corr[i * 6:(i + 1) * 6])], [2, 3, 2, 3])
    dst2 = cpu.hal_allocate_i64([2, 2, 2, 3])
    F0, F1, K = (1 << 70) + 5, -(3 << 90) + 7, 100
    cpu.hal_bconv_correct_pow2_wide(dst2, src3, F0, F1, shift, K)
    d = dst2.reshape(2, 2, 6)
    for i in range(2):
        for pos in range(6):
            v = vals[i * 6 + pos] + ((vals[i * 6 + pos] >> 3) << 64)
            q = (corr[i * 6 + pos] + (1 << (shift - 1))) >> shift
            red = ((v - q * F0) * F1) % (1 << K)
            exp = red - (1 << K) if red > (1 << (K - 1)) else red
            assert int(d[i, 0, pos]) + (int(d[i, 1, pos]) << 64) == exp


def test_ntt_roundtrip(cpu):
    """cpu.rs:1692-1724"""
    q = 144115188075849889
    m = cpu.hal_create_modulus(q)
    plan = cpu.hal_create_ntt_plan(m, [4, 3])
    rng = Lcg(7)
    data = [(rng() % (2 * q)) - q for _ in range(240)]
    for shape, dim in (([2, 2, 12, 5], 2), ([2, 2, 5, 12], 3)):
        t = cpu.hal_allocate_i64(shape)
        cpu.hal_import(t, data)
        cpu.hal_ntt_mont64_dim(t, [plan, plan], dim)
        cpu.hal_intt_mont64_dim(t, [plan, plan], dim)
        out = cpu.hal_export(t)
        assert all((int(a) - b) % q == 0 for a, b in zip(out, data))


def test_ntt_convolution_theorem(cpu):
    """cpu.rs:1726-1772 on C_8 x C_9 x C_2, against a brute-force convolution."""
    q = 144115188075849889
    m = cpu.hal_create_modulus(q)
    inner = [8, 9, 2]
    n = 144
    plan = cpu.hal_create_ntt_plan(m, inner)
    rng = Lcg(42)
    a = [rng() % 1000 - 500 for _ in range(n)]
    b = [rng() % 1000 - 500 for _ in range(n)]
    exp = [0] * n
    for i0 in range(8):
        for i1 in range(9):
            for i2 in range(2):
                for j0 in range(8):
                    for j1 in range(9):
                        for j2 in range(2):
                            d = (((i0 + j0) % 8) * 9 + (i1 + j1) % 9) * 2 + (i2 + j2) % 2
                            exp[d] += a[(i0 * 9 + i1) * 2 + i2] * b[(j0 * 9 + j1) * 2 + j2]
    at, bt = cpu.hal_allocate_i64([1, 1, n, 1]), cpu.hal_allocate_i64([1, 1, n, 1])
    cpu.hal_import(at, [montgomery_repr(x, q) for x in a])
    cpu.hal_import(bt, [montgomery_repr(x, q) for x in b])
    cpu.hal_ntt_mont64_dim(at, [plan], 2)
    cpu.hal_ntt_mont64_dim(bt, [plan], 2)
    cpu.hal_mul_mont64(at, bt, [m])
    cpu.hal_intt_mont64_dim(at, [plan], 2)
    out = cpu.hal_export(at)
    for pos in range(n):
        assert int(out[pos]) % q == montgomery_repr(exp[pos], q)


def test_ntt_mul_mont64_dim(cpu):
    """cpu.rs:1774-1827: fused == ntt + mul_mont exactly, along both dims."""
    q = 144115188075849889
    m = cpu.hal_create_modulus(q)
    plan = cpu.hal_create_ntt_plan(m, [8, 3])
    rng = Lcg(123)
    n = 24
    for dim in (2, 3):
        shape, kshape = ([2, 2, n, 3], [1, 2, n, 1]) if dim == 2 else ([2, 2, 3, n], [1, 2, 1, n])
        data = [(rng() % (4 * q)) - 2 * q for _ in range(int(np.prod(shape)))]
        kern = [montgomery_repr((rng() % (2 * q)) - q, q) for _ in range(int(np.prod(kshape)))]
        kt = cpu.hal_allocate_i64(kshape)
        cpu.hal_import(kt, kern)
        ref = cpu.hal_allocate_i64(shape)
        cpu.hal_import(ref, data)
        cpu.hal_ntt_mont64_dim(ref, [plan, plan], dim)
        cpu.hal_mul_mont64(ref, kt, [m, m])
        fused = cpu.hal_allocate_i64(shape)
        cpu.hal_import(fused, data)
        cpu.hal_ntt_mul_mont64_dim(fused, [plan, plan], dim, kt, [m, m])
        assert (np.abs(fused) <= q).all()
        assert (ref == fused).all()


def test_permute(cpu):
    """cpu.rs:1829-1875"""
    plan = cpu.hal_create_permutation([3, 4], [1, 2])
    data = list(range(48))
    s3 = [1, 2, 2, 12]
    src = cpu.hal_allocate_i64(s3)
    cpu.hal_import(src, data)
    dst = cpu.hal_allocate_i64(s3)
    cpu.hal_permute(dst, src, plan, 3)
    for j in range(2):
        for k in range(2):
            for a0 in range(3):
                for a1 in range(4):
                    fr = flat(s3, [0, j, k, a0 * 4 + a1])
                    to = flat(s3, [0, j, k, ((a0 + 1) % 3) * 4 + (a1 + 2) % 4])
                    assert dst.reshape(-1)[to] == data[fr]
    s2 = [1, 2, 12, 2]
    src = cpu.hal_allocate_i64(s2)
    cpu.hal_import(src, data)
    dst = cpu.hal_allocate_i64(s2)
    cpu.hal_permute(dst, src, plan, 2)
    for j in range(2):
        for a0 in range(3):
            for a1 in range(4):
                for l in range(2):
                    fr = flat(s2, [0, j, a0 * 4 + a1, l])
                    to = flat(s2, [0, j, ((a0 + 1) % 3) * 4 + (a1 + 2) % 4, l])
                    assert dst.reshape(-1)[to] == data[fr]


def test_oracle_errors(cpu):
    """Contract violations raise instead of computing garbage (the reference asserts)."""
    t = cpu.hal_allocate_i64([2, 2, 3, 4])
    with pytest.raises(RuntimeError):
        cpu.hal_add_i64(t, cpu.hal_allocate_i64([2, 3, 3, 4]))
    with pytest.raises(RuntimeError):
        cpu.hal_mul_mont64(t, t, [cpu.hal_create_modulus(97)])
    with pytest.raises(RuntimeError):
        cpu.hal_create_modulus(96)


# ---------------------------------------------------------------------------------------------
# The lazy schedule against the facts the reference states about it (ntt.rs:18-81)

def _is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d, s = d // 2, s + 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _largest_57bit_prime(order):
    """the largest prime below 2^57 that is 1 mod `order` (so NTTs of that order exist)"""
    k = ((1 << 57) - 2) // order
    while not _is_prime(k * order + 1):
        k -= 1
    return k * order + 1


def _schedule_model(shape, n, inverse, reduce_output):
    """The schedule ntt.rs:18-81 describes, stage by stage, as (kind, factor, stage, skip_ok,
    fused) tuples: entries start bounded by 2N; a forward stage of radix R adds (R - 1) N to the
    bound, or multiplies it by R in the one twiddle-1 block, which skips its reduction only
    while R * bound fits an i64; a reduce pass (back to N) comes only when the next stage could
    exceed i64::MAX; the last forward stage reduces its output unless unreduced output is asked
    for. The inverse multiplies the bound by R per stage (the unmultiplied y0 + y1 branch), so it
    reduces when R * bound would not fit, and fuses len^-1 into its last stage."""
    MAX = (1 << 63) - 1
    factors = [(L, 2 if L & (L - 1) == 0 else 3) for L in shape if L > 1]
    ev = []
    bound = 2 * n
    for fi, (L, R) in enumerate(factors):
        stages = 0
        while R ** stages < L:
            stages += 1
        last = fi + 1 == len(factors)
        order = range(stages - 1, -1, -1) if inverse else range(stages)
        for s in order:
            if not inverse:
                if bound + (R - 1) * n > MAX:
                    ev.append(("reduce", fi))
                    bound = n
                fuse = reduce_output and last and s == stages - 1
                ok = R * bound <= MAX
                ev.append(("fwd", fi, s, ok, fuse))
                bound = n if fuse else (R * bound if ok else bound + (R - 1) * n)
            else:
                if R * bound > MAX:
                    ev.append(("reduce", fi))
                    bound = n
                fuse = last and s == 0
                ev.append(("inv", fi, s, True, fuse))
                bound = n if fuse else R * bound
    return ev


def _traced(shape, n, inverse, reduce_output, seed=1):
    be = cb.CpuBackend()
    plan = cb.CpuNTTPlan(be.hal_create_modulus(n), shape)
    rng = np.random.default_rng(seed)
    data = rng.integers(-2 * n, 2 * n, size=plan.len, dtype=np.int64)
    events, mults = cb.plan_apply_traced(plan, data, 1, 1, inverse, reduce_output)
    return events, mults


SCHEDULE_PLANS = [
    # field_36864: dim 2 [4, 3] and dim 3 [2, 3, 8, 64] (field.rs:535-536); field_49152: [4] and
    # [3, 64, 64]; long one-radix plans for the inverse reduce cadence
    ((4, 3), 144115176243904321),
    ((2, 3, 8, 64), 144115176243904321),
    ((4,), None),
    ((3, 64, 64), None),
    ((1 << 14,), None),
    ((3 ** 8,), None),
    ((64, 64, 3), None),
]


@pytest.mark.parametrize("shape,prime", SCHEDULE_PLANS)
def test_ntt_schedule_matches_stated_facts(shape, prime):
    n = prime or _largest_57bit_prime(int(np.lcm.reduce(shape)))
    assert (1 << 56) < n < (1 << 57)
    for inverse, reduce_output in [(False, True), (False, False), (True, False)]:
        events, mults = _traced(shape, n, inverse, reduce_output)
        model = _schedule_model(shape, n, inverse, reduce_output)
        got = [("reduce", e["factor"]) if e["kind"] == "reduce" else
               (e["kind"], e["factor"], e["stage"], e["skip_ok"], e["fused"]) for e in events]
        assert got == model, (shape, inverse, reduce_output)
        # 57-bit moduli: the fields' forward transforms never need a reduce pass (ntt.rs:32-34).
        # Long one-radix plans do: their twiddle-1 skips double the bound for 5 stages, after
        # which additive growth crosses i64::MAX, as the rules above imply
        if not inverse and (prime is not None or shape in [(4,), (3, 64, 64)]):
            assert not any(e["kind"] == "reduce" for e in events)
        # every stage has exactly one twiddle-1 block (ntt.rs:39-41); the forward skips it while
        # allowed, the inverse always (except in the stage that applies len^-1)
        for e in events:
            if e["kind"] == "fwd":
                assert e["skipped"] == (1 if e["skip_ok"] else 0)
            elif e["kind"] == "inv":
                assert e["skipped"] == (0 if e["fused"] else 1)
        # the observed Barrett multiplications equal the count implied by those decisions
        L = int(np.prod(shape))
        want = 0
        for e in events:
            if e["kind"] == "reduce":
                want += L
                continue
            F = [f for f in shape if f > 1][e["factor"]]
            R = 2 if F & (F - 1) == 0 else 3
            nblocks = R ** e["stage"]
            part = F // nblocks // R
            lines = L // F
            if e["kind"] == "fwd":
                per = part * ((nblocks - e["skipped"]) if R == 2 else (6 * nblocks - 2 * e["skipped"]))
                per += F if e["fused"] else 0
            else:
                if R == 2:
                    per = F if e["fused"] else part * (nblocks - e["skipped"])
                else:
                    per = part * nblocks * 7 if e["fused"] else part * (6 * nblocks - 2 * e["skipped"])
            want += per * lines
        assert mults == want, (shape, inverse, reduce_output)


def test_inverse_reduce_every_six_radix2_stages():
    """ntt.rs:73-75: for 57-bit moduli the inverse needs a reduction pass every 6 radix-2 stages"""
    n = _largest_57bit_prime(1 << 14)
    events, _ = _traced((1 << 14,), n, True, False)
    pos, stage_no = [], 0
    for e in events:
        if e["kind"] == "reduce":
            pos.append(stage_no + 1)
        else:
            stage_no += 1
    assert pos == [6, 12]
    # radix 3 (the bound triples, ntt.rs:67-69): from 2N the first pass comes before stage 4
    # (3 * 54 N > 2^63), then from N every 3 stages (3 * 27 N > 2^63 > 27 N)
    n3 = _largest_57bit_prime(3 ** 8)
    events, _ = _traced((3 ** 8,), n3, True, False)
    pos, stage_no = [], 0
    for e in events:
        if e["kind"] == "reduce":
            pos.append(stage_no + 1)
        else:
            stage_no += 1
    assert pos == [4, 7]


def test_schedule_model_rejects_planted_mistakes():
    """the comparison above fails for a schedule with an extra reduce pass, a dropped twiddle-1
    skip or a fused stage moved"""
    n = 144115176243904321
    shape = (2, 3, 8, 64)
    events, _ = _traced(shape, n, False, True)
    got = [("reduce", e["factor"]) if e["kind"] == "reduce" else
           (e["kind"], e["factor"], e["stage"], e["skip_ok"], e["fused"]) for e in events]
    model = _schedule_model(shape, n, False, True)
    assert got == model
    assert got[:3] + [("reduce", 1)] + got[3:] != model
    dropped = list(got)
    dropped[3] = dropped[3][:3] + (False,) + dropped[3][4:]
    assert dropped != model
    moved = list(got)
    moved[-1] = moved[-1][:4] + (False,)
    moved[-2] = moved[-2][:4] + (True,)
    assert moved != model


def test_forward_twiddle_one_representatives():
    """closed forms of the lazy representatives on length 2 (ntt.rs:36-56, 5-16): the only
    stage's twiddle-1 butterfly adds and subtracts b without reducing it; with reduced output the
    stage Barrett-multiplies both outputs by 1, i.e. x + round(u x / 2^64) N with
    u = round(-2^64 / N) (halves away from zero)"""
    n = Q57
    be = cb.CpuBackend()
    u = -((2 * (1 << 64) + n) // (2 * n))
    plan2 = cb.CpuNTTPlan(be.hal_create_modulus(n), (2,))

    def red(x):
        q = (u * x + (1 << 63)) >> 64
        return x + q * n

    for a, b in [(2 * n - 1, 2 * n - 3), (-2 * n, 2 * n), (12345, -(2 * n - 7))]:
        x = np.array([a, b], dtype=np.int64)
        cb.plan_apply(plan2, x, 1, 1, False, False)
        assert x.tolist() == [a + b, a - b]
        y = np.array([a, b], dtype=np.int64)
        cb.plan_apply(plan2, y, 1, 1, False, True)
        assert y.tolist() == [red(a + b), red(a - b)]
        assert all(abs(v) <= n for v in y.tolist())
