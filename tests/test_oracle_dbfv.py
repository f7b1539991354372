"""The oracle's dBFV digit reduction (oracle/dbfv.py) against what the mathematics fixes:

* split_overflow: the reference's own property test (dbfv.rs:1155-1208) at its three moduli,
  x = x_lo + (b Delta) x_hi mod Q with |x_lo| < 2 b Delta, plus x_hi in (-R/2, R/2];
* dbfv_reduce: the carry chain. With y_0 = x_0 and y_i = x_i + Delta c_(i-1), every output digit z_i
  satisfies y_i = z_i + (b Delta) c_i mod Q for a carry c_i in (-R/2, R/2], and |z_i| < 2 b Delta.
  The checker recovers c_i = (y_i - z_i) (b Delta)^-1 mod Q; it rejects outputs with a planted
  mistake (digits split without their carries, a digit off by one or by Delta, swapped digits).
"""
import numpy as np
import pytest

from oracle import cpu_backend
from oracle import dbfv as D
from oracle import ring as R


@pytest.fixture(scope="module")
def cpu():
    cpu_backend.build()
    return cpu_backend.CpuBackend()


def lifts(ring, t, base):
    """centred integer lifts mod Q of every coefficient of every row of a period-basis tensor"""
    a = np.asarray(t)
    q = ring.modulus_of(base)
    out = []
    for row in range(a.shape[0]):
        vals = []
        for idx in np.ndindex(a.shape[2], a.shape[3]):
            x = 0
            for j, i in enumerate(base):
                p = ring.primes[i]
                v = R.from_mont_repr(int(a[(row, j) + idx]), p) % p
                m = q // p
                x += v * m * pow(m % p, -1, p)
            x %= q
            vals.append(x - q if x > q // 2 else x)
        out.append(vals)
    return out


def uniform(ring, count, base, rng):
    n2, n3 = ring.field.deg_lprime, ring.field.deg_k
    t = ring.allocate(count, base)
    for j, i in enumerate(base):
        t[:, j] = rng.integers(0, ring.primes[i], size=(count, n2, n3), dtype=np.int64)
    return t


def overflow_values(ov):
    a = np.asarray(ov)
    limbs = a.shape[1]
    return [[sum(int(a[row, k][idx]) << (64 * k) for k in range(limbs)) for idx in np.ndindex(a.shape[2], a.shape[3])]
            for row in range(a.shape[0])]


@pytest.mark.parametrize("seed,modulus", [(10, D.DBFVModulus.new(8, 3)), (18, D.DBFVModulus.new_incomplete(8, 108)),
                                          (20, D.DBFVModulus.new(27, 4))])
def test_split_overflow_property(cpu, seed, modulus):
    ring = R.RingContext(cpu, R.FieldDescriptor("field_192"), (200, 220), modulus.log2_t)
    base = ring.initial_rns_base()
    rng = np.random.default_rng(seed)
    x = uniform(ring, 2, base, rng)
    before = lifts(ring, x, base)
    data = D.SplitOverflowData(ring, modulus, base)
    assert data.limbs() == (1 if modulus.log2_t - modulus.log2_b <= 62 else 2)
    ov = D.split_overflow(ring, data, x, 1, base)
    after = lifts(ring, x, base)
    q = ring.modulus_of(base)
    scale = D.rounded_div(q, modulus.t()) * modulus.b()
    r_half = 1 << (data.log2_r - 1)
    for row, (b_row, a_row, o_row) in enumerate(zip(before, after, overflow_values(ov))):
        for j, (xb, xa, o) in enumerate(zip(b_row, a_row, o_row)):
            assert (xa + o * scale - xb) % q == 0, (row, j)
            assert abs(xa) < 2 * scale, (row, j)
            assert -r_half < o <= r_half, (row, j)


def carry_chain_failures(ring, modulus, base, x_in, z_out):
    """the positions where dbfv_reduce's output violates the carry chain (empty if it holds)"""
    q = ring.modulus_of(base)
    delta = D.rounded_div(q, modulus.t())
    scale = delta * modulus.b()
    inv = pow(scale % q, -1, q)
    r_half = 1 << (modulus.log2_t - modulus.log2_b - 1)
    nd = modulus.digits()
    bad = []
    for s in range(2):  # the two rows of every digit carry independently
        carry = [0] * len(z_out[0])
        for i in range(nd):
            x = x_in[2 * i + s] if 2 * i + s < len(x_in) else [0] * len(carry)
            z = z_out[2 * i + s]
            for j in range(len(carry)):
                y = x[j] + delta * carry[j]
                c = ((y - z[j]) * inv) % q
                c = c - q if c > q // 2 else c
                if not (-r_half < c <= r_half) or abs(z[j]) >= 2 * scale:
                    bad.append((i, s, j))
                carry[j] = c
    return bad


@pytest.mark.parametrize("log2_q,modulus,d,seed", [((200, 220), D.DBFVModulus.new(16, 4), 4, 12),
                                                   ((200, 220), D.DBFVModulus.new(16, 4), 2, 13),
                                                   ((260, 280), D.DBFVModulus.new(27, 4), 4, 19),
                                                   ((260, 280), D.DBFVModulus.new(27, 4), 1, 21)])
def test_dbfv_reduce_carry_chain(cpu, log2_q, modulus, d, seed):
    ring = R.RingContext(cpu, R.FieldDescriptor("field_192"), log2_q, modulus.log2_t)
    base = ring.initial_rns_base()
    rng = np.random.default_rng(seed)
    ct = uniform(ring, 2 * d, base, rng)
    ct_copy = np.array(ct)
    x_in = lifts(ring, ct, base)
    out, bound = D.dbfv_reduce(ring, modulus, ct, 1, base)
    assert bound == 1 and out.shape[0] == 2 * modulus.digits()
    z = lifts(ring, out, base)
    assert carry_chain_failures(ring, modulus, base, x_in, z) == []
    # planted mistakes are caught: the digits split without their carries, a digit off by one,
    # a digit shifted by Delta, swapped digits
    q = ring.modulus_of(base)
    delta = D.rounded_div(q, modulus.t())
    if d >= 2:
        data = D.SplitOverflowData(ring, modulus, base)
        no_carry = ring.allocate(2 * modulus.digits(), base)
        no_carry[0:2 * d] = np.asarray(ct_copy)
        for i in range(modulus.digits()):
            D.split_overflow(ring, data, no_carry[2 * i:2 * i + 2], 1, base)
        assert carry_chain_failures(ring, modulus, base, x_in, lifts(ring, no_carry, base))
    off = [row[:] for row in z]
    off[2] = [v + 1 for v in off[2]]
    assert carry_chain_failures(ring, modulus, base, x_in, off)
    shifted = [row[:] for row in z]
    shifted[3][5] = (shifted[3][5] + delta) % q
    assert carry_chain_failures(ring, modulus, base, x_in, shifted)
    swapped = [z[2], z[3], z[0], z[1]] + z[4:]
    assert carry_chain_failures(ring, modulus, base, x_in, swapped)


def test_dbfv_reduce_single_digit_is_identity(cpu):
    """dbfv.rs:971-973: nothing to do for one-digit moduli or empty ciphertexts"""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_192"), (200, 220), 16)
    base = ring.initial_rns_base()
    ct = uniform(ring, 2, base, np.random.default_rng(3))
    out, bound = D.dbfv_reduce(ring, D.DBFVModulus.new(16, 1), ct, 2, base)
    assert out is ct and bound == 2


def test_split_i128_limbs():
    for x in [0, 1, -1, (1 << 63), -(1 << 63) - 5, (1 << 125) + 12345, -(1 << 125)]:
        lo, hi = D.split_i128_limbs(x)
        assert -(1 << 63) <= lo < 1 << 63 and hi * (1 << 64) + lo == x


def one_key(ring, base, ndigits=1):
    """the relinearization key of the zero secret key with key part g_i = (Q/D_i) ((Q/D_i)^-1
    mod D_i) for gadget digit i (the digits of select_digits): rows (0, g_i) in CRT basis. The
    gadget digits d_i of c2 recompose to sum_i d_i g_i = c2 mod Q, so relin maps (c0, c1, c2) to
    (c0, c1 + c2) mod Q"""
    digits = R.select_digits(ndigits, len(base))
    q = ring.modulus_of(base)
    keys = []
    for a, b in digits:
        di = ring.modulus_of(base[a:b])
        g = (q // di) * pow((q // di) % di, -1, di) % q
        key = ring.allocate(2, base)
        for j, i in enumerate(base):
            key[1, j] = R.montgomery_repr(g % ring.primes[i], ring.primes[i])
        keys.append(key)
    return keys, digits


def convolution_failures(ring, t, base, ll, rl, lifted, out_pb, K, tol=(1, 2)):
    """The digit convolution identity under `one_key`: with T_k,c the exact integer components
    sum_j ll_j (x) rl_(k-j) (c = 0: l0 r0, 1: l0 r1 + l1 r0, 2: l1 r1) of the lifted CRT-basis
    operands (two rows per digit), formed here and taken to period basis by the oracle's
    transforms, output digit k (period basis) must be (round(t T_k,0 / Q), round(t T_k,1 / Q) +
    round(t T_k,2 / Q)) mod Q up to the rescale's +-1 per term. Returns the offending
    (digit, component) pairs."""
    d1, d2 = ll.shape[0] // 2, rl.shape[0] // 2
    T = ring.allocate(3 * K, lifted)
    for j_pos, i in enumerate(lifted):
        p = ring.primes[i]
        rinv = pow(1 << 64, -1, p)
        lv = [[int(v) for v in ll[row, j_pos].reshape(-1)] for row in range(2 * d1)]
        rv = [[int(v) for v in rl[row, j_pos].reshape(-1)] for row in range(2 * d2)]
        n = len(lv[0])
        for k in range(K):
            comps = [[0] * n for _ in range(3)]
            for j in range(max(0, k - (d2 - 1)), min(k, d1 - 1) + 1):
                a0, a1, b0, b1 = lv[2 * j], lv[2 * j + 1], rv[2 * (k - j)], rv[2 * (k - j) + 1]
                for e in range(n):
                    comps[0][e] += a0[e] * b0[e]
                    comps[1][e] += a0[e] * b1[e] + a1[e] * b0[e]
                    comps[2][e] += a1[e] * b1[e]
            for c in range(3):
                T[3 * k + c, j_pos] = np.array([x * rinv % p for x in comps[c]], dtype=np.int64).reshape(T.shape[2:])
    ring.crt_to_pb_tensor(T, lifted)
    X = lifts(ring, T, lifted)
    q = ring.modulus_of(base)
    z = lifts(ring, out_pb, base)
    bad = []

    def centred(v):
        v %= q
        return v - q if v > q // 2 else v

    for k in range(K):
        for c in range(2):
            for e in range(len(z[0])):
                want = D.rounded_div(t * X[3 * k + c][e], q)
                if c == 1:
                    want += D.rounded_div(t * X[3 * k + 2][e], q)
                if abs(centred(z[2 * k + c][e] - want)) > tol[c]:
                    bad.append((k, c))
                    break
    return bad


def hom_mul_failures(ring, modulus, base, lhs_red, rhs_red, out):
    """convolution_failures for dbfv_hom_mul (dbfv.rs:1078-1143), from the reduced operands"""
    ll, lifted = R.lift_to_extended(ring, lhs_red, base)
    rl, _ = R.lift_to_extended(ring, rhs_red, base)
    K = min(modulus.digits(), ll.shape[0] // 2 + rl.shape[0] // 2 - 1)
    return convolution_failures(ring, modulus.t(), base, ll, rl, lifted, out, K)


@pytest.mark.parametrize("modulus,d1,d2,square", [(D.DBFVModulus.new(16, 4), 4, 4, False),
                                                  (D.DBFVModulus.new(16, 4), 2, 3, False),
                                                  (D.DBFVModulus.new(16, 3), 3, 3, True)])
def test_dbfv_hom_mul_convolution(cpu, modulus, d1, d2, square):
    ring = R.RingContext(cpu, R.FieldDescriptor("field_192"), (200, 220), modulus.log2_t)
    base = ring.initial_rns_base()
    rng = np.random.default_rng(40 + d1 + d2)
    lhs = uniform(ring, 2 * d1, base, rng)
    rhs = lhs if square else uniform(ring, 2 * d2, base, rng)
    keys, digits = one_key(ring, base, 2)
    lhs_red, _ = D.dbfv_reduce(ring, modulus, np.array(lhs), 1, base)
    rhs_red, _ = D.dbfv_reduce(ring, modulus, np.array(rhs), 1, base)
    if square:
        out, bound = D.dbfv_hom_square(ring, modulus, np.array(lhs), 1, base, keys, digits)
    else:
        out, bound = D.dbfv_hom_mul(ring, modulus, np.array(lhs), 1, np.array(rhs), 1, base, keys, digits)
    K = min(modulus.digits(), lhs_red.shape[0] // 2 + rhs_red.shape[0] // 2 - 1)
    assert bound == 2 and out.shape[0] == 2 * K
    assert hom_mul_failures(ring, modulus, base, lhs_red, rhs_red, out) == []
    crt, cb = D.dbfv_hom_mul(ring, modulus, np.array(lhs), 1, np.array(rhs), 1, base, keys, digits, "crt")
    pb = np.array(crt)
    ring.crt_to_pb_tensor(pb, base)
    assert cb == 1 and hom_mul_failures(ring, modulus, base, lhs_red, rhs_red, pb) == []
    # planted mistakes are caught: output digits swapped, components swapped, c2 not switched in
    swapped = np.array(out)
    swapped[0:2], swapped[2:4] = np.array(out[2:4]), np.array(out[0:2])
    assert hom_mul_failures(ring, modulus, base, lhs_red, rhs_red, swapped)
    comps = np.array(out)
    comps[0], comps[1] = np.array(out[1]), np.array(out[0])
    assert hom_mul_failures(ring, modulus, base, lhs_red, rhs_red, comps)
    zero_keys = [ring.allocate(2, base) for _ in digits]
    no_c2, _ = D.dbfv_hom_mul(ring, modulus, np.array(lhs), 1, np.array(rhs), 1, base, zero_keys, digits)
    assert hom_mul_failures(ring, modulus, base, lhs_red, rhs_red, no_c2)


def test_dbfv_hom_mul_empty_operand(cpu):
    """dbfv.rs:1047-1051: an operand without digits is the product"""
    modulus = D.DBFVModulus.new(16, 4)
    ring = R.RingContext(cpu, R.FieldDescriptor("field_192"), (200, 220), modulus.log2_t)
    base = ring.initial_rns_base()
    lhs = uniform(ring, 4, base, np.random.default_rng(5))
    empty = ring.allocate(0, base)
    keys, digits = one_key(ring, base)
    out, bound = D.dbfv_hom_mul(ring, modulus, lhs, 1, empty, 1, base, keys, digits)
    assert out.shape[0] == 0 and bound == 1


@pytest.mark.parametrize("field,log2_q,t_bits", [("field_18", (100, 120), 4), ("field_192", (200, 220), 40)])
def test_bfv_hom_mul_identity(cpu, field, log2_q, t_bits):
    """The oracle's bfv_hom_mul and relin_crt (bfv.rs:699-775), pinned by the one-digit case of
    the same identity: under the zero secret key the product of two CRT-basis ciphertexts is
    (round(t l0 r0 / Q), round(t (l0 r1 + l1 r0) / Q) + round(t l1 r1 / Q)) mod Q up to +-1 per
    term; it rejects swapped components and a c2 that is not switched in."""
    ring = R.RingContext(cpu, R.FieldDescriptor(field), log2_q, t_bits)
    base = ring.initial_rns_base()
    rng = np.random.default_rng(90 + t_bits)
    lhs, rhs = uniform(ring, 2, base, rng), uniform(ring, 2, base, rng)
    keys, digits = one_key(ring, base, 2)
    t = 1 << t_bits
    out = R.bfv_hom_mul(ring, t, np.array(lhs), np.array(rhs), base, keys, digits)

    def lifted(ct):
        pb = np.array(ct)
        ring.crt_to_pb_tensor(pb, base)
        return R.lift_to_extended(ring, pb, base, crt=np.array(ct))

    ll, lb = lifted(lhs)
    rl, _ = lifted(rhs)

    def fails(o):
        pb = np.array(o)
        ring.crt_to_pb_tensor(pb, base)
        return convolution_failures(ring, t, base, ll, rl, lb, pb, 1)

    assert fails(out) == []
    swapped = np.array(out)
    swapped[0], swapped[1] = np.array(out[1]), np.array(out[0])
    assert fails(swapped)
    no_c2 = R.bfv_hom_mul(ring, t, np.array(lhs), np.array(rhs), base, [ring.allocate(2, base) for _ in digits], digits)
    assert fails(no_c2)
    # the key digits paired with the wrong gadget digits
    crossed = R.bfv_hom_mul(ring, t, np.array(lhs), np.array(rhs), base, keys[::-1], digits)
    assert fails(crossed)
