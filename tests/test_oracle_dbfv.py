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
