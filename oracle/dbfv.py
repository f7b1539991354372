"""ORACLE — test infrastructure only (imported by tests/ and by bench.py's cpu_baseline leg).

A restatement of the dBFV digit reduction of cyclosubfield-dbfv/src/dbfv.rs over the oracle ring
(``oracle.ring.RingContext`` on the ``CpuBackend``). It follows:

* dbfv.rs:56-80    — ``DBFVModulus`` (b = 2^log2_b, t = 2^log2_t, digits);
* hal/mod.rs:66-76 — ``split_i128_limbs``;
* dbfv.rs:760-841  — ``SplitOverflowData::new`` (the BEHZ-style constants of the overflow split);
* dbfv.rs:873-948  — ``split_overflow``: x -> (x_lo, x_hi) with x = x_lo + (b Delta) x_hi mod Q;
* dbfv.rs:950-1025 — ``dbfv_reduce``: the sequential carry chain over the digits.
* dbfv.rs:1026-1143 — ``dbfv_hom_mul``, ``dbfv_hom_square``, ``dbfv_mul_convolve``: reduce,
  lift, digit convolution in one wide batch, rescale, relinearization per output digit.

Pins: ``split_overflow`` is checked against the reference's own property test
(dbfv.rs:1155-1208: x = x_lo + b Delta x_hi mod Q, |x_lo| < 2 b Delta) at its three moduli,
including two-limb R; ``dbfv_reduce`` is checked against the carry-chain invariant it must keep
(every digit's carry recovered from the input and output is a small integer in (-R/2, R/2], and
every output digit is below 2 b Delta), which a dropped, unscaled or misplaced carry breaks
(tests/test_oracle_dbfv.py). The reference checks dbfv_reduce only through decryption
(dbfv.rs:1324-1362), which needs key generation and encryption that this oracle does not restate.

``dbfv_hom_mul`` is a composition of the pinned ring steps (lift, inner product, wide reduction,
rescale, relin). The composition is pinned by the convolution identity under the zero secret key
(tests/test_oracle_dbfv.py): with key rows (0, g_i), g_i the CRT recomposition factor of gadget
digit i, output digit k is (round(t T_k0 / Q), round(t T_k1 / Q) + round(t T_k2 / Q)) mod Q up to
+-1 per term, T_kc the exact convolution components of the lifted operands. Keys of a nonzero
secret are not restated (the reference checks those only through decryption).

Big-integer constants use Python ints, as the reference's feanor-math ZZbig does.
"""
from .ring import _log2_ceil


class DBFVModulus:
    """dbfv.rs:56-80"""

    def __init__(self, log2_b, log2_t):
        if not (log2_b >= 1 and 1 <= log2_t <= 126):
            raise ValueError("t = 2^log2_t must be between 2 and 2^126")
        self.log2_b = log2_b
        self.log2_t = log2_t

    @classmethod
    def new(cls, log2_b, digits):
        return cls(log2_b, digits * log2_b)

    @classmethod
    def new_incomplete(cls, log2_b, log2_t):
        return cls(log2_b, log2_t)

    def t(self):
        return 1 << self.log2_t

    def b(self):
        return 1 << self.log2_b

    def digits(self):
        return (self.log2_t - 1) // self.log2_b + 1


def _as_i64(x):
    x &= (1 << 64) - 1
    return x - (1 << 64) if x >> 63 else x


def split_i128_limbs(x):
    """hal/mod.rs:66-76: x = hi 2^64 + lo with lo the low 64 bits of x as an i64"""
    assert abs(x) < 1 << 126
    lo = _as_i64(x)
    hi = (x - lo) >> 64
    return lo, hi


def rounded_div(a, b):
    """feanor-math rounded_div for b > 0: nearest integer, halves away from zero"""
    q, r = divmod(abs(a), b)
    if 2 * r >= b:
        q += 1
    return q if a >= 0 else -q


class SplitOverflowData:
    """SplitOverflowData::new (dbfv.rs:785-841)"""

    def __init__(self, ring, modulus, base):
        self.log2_b = modulus.log2_b
        self.log2_t = modulus.log2_t
        self.rns_base = list(base)
        self.log2_r = modulus.log2_t - modulus.log2_b
        if not 1 <= self.log2_r <= 126:
            raise ValueError("t/b must be between 2 and 2^126")
        limbs = self.limbs()
        q = ring.modulus_of(base)
        r_big = 1 << self.log2_r
        primes = [ring.primes[i] for i in base]
        max_p, min_p = max(primes), min(primes)
        self.log2_gamma = max(64, _log2_ceil(4 * len(base) * max_p + 1))
        gamma = 1 << self.log2_gamma
        assert ((gamma // min_p) + 1) * max_p * len(base) < 1 << 127
        q_over_p_mod_r = [(q // p) % r_big for p in primes]
        gamma_over_p = [rounded_div(gamma, p) for p in primes]
        values = [split_i128_limbs(c)[0] for c in q_over_p_mod_r]
        if limbs == 2:
            values += [split_i128_limbs(c)[1] for c in q_over_p_mod_r]
        values += gamma_over_p
        self.matrix = ring._import(values, [1, limbs + 1, len(base), 1])
        self.q_mod_r = q % r_big
        # inv_mod_2_128 (dbfv.rs:763-770) is the inverse modulo 2^128; its negation is reduced
        # modulo 2^log2_r
        self.neg_q_inv_mod_r = (-pow(self.q_mod_r, -1, 1 << 128)) % r_big
        delta = rounded_div(q, modulus.t())
        neg_b_delta = -(delta * modulus.b())
        self.r_q_over_p_inv = ring.q_over_p_inv_tensor(base, r_big)
        self.neg_b_delta_mont2 = [ring.scalar_tensor_big_mont2(base, neg_b_delta << (64 * k)) for k in range(limbs)]

    def limbs(self):
        return 1 if self.log2_r <= 62 else 2


def assert_mul_mont_safe(ring, bound, base):
    """ring.rs:955-960"""
    max_p = max(ring.primes[i] for i in base)
    assert bound * max_p <= 1 << 63, "representatives too large for a montgomery multiplication"


def split_overflow(ring, data, x, x_bound, base):
    """split_overflow (dbfv.rs:893-948): x (period basis, [count, r, ., .]) becomes x_lo in
    place; returns x_hi as a [count, limbs, ., .] tensor of plain values in (-R/2, R/2]."""
    be = ring.backend
    shape = x.shape
    assert len(base) == shape[1]
    assert_mul_mont_safe(ring, x_bound, base)
    limbs = data.limbs()
    mods = ring.mods_for(base)
    lifts = be.hal_allocate_i64(shape)
    be.hal_assign(lifts, x)
    be.hal_mul_mont64(lifts, data.r_q_over_p_inv, mods)
    sums = be.hal_allocate_i128([shape[0], limbs + 1, shape[2], shape[3]])
    be.hal_matmul_i64(sums, data.matrix, lifts)
    overflow = be.hal_allocate_i64([shape[0], limbs, shape[2], shape[3]])
    if limbs == 1:
        be.hal_bconv_correct_pow2(overflow, sums, data.q_mod_r, data.neg_q_inv_mod_r, data.log2_gamma, data.log2_r)
    else:
        be.hal_bconv_correct_pow2_wide(overflow, sums, data.q_mod_r, data.neg_q_inv_mod_r, data.log2_gamma, data.log2_r)
    acc = ring.allocate_wide(shape[0], base)
    rows = range(shape[0])
    for k in range(limbs):
        acc.fma_unbounded(rows, overflow[:, k:k + 1], data.neg_b_delta_mont2[k])
    acc.add(rows, x, x_bound)
    ring.reduce_wide_into(x, acc)
    return overflow


def dbfv_reduce(ring, modulus, ct, ct_bound, base):
    """dbfv_reduce (dbfv.rs:965-1025) on a period-basis dBFV ciphertext: rows 2i, 2i + 1 are the
    BFV ciphertext of digit i. Returns (result, bound): `modulus.digits()` digits, each reduced
    modulo b Delta with its carry added to the next digit; the last carry is dropped."""
    be = ring.backend
    assert ct.shape[0] % 2 == 0
    d = ct.shape[0] // 2
    assert d <= modulus.digits()
    if modulus.digits() == 1 or d == 0:
        return ct, ct_bound
    data = SplitOverflowData(ring, modulus, base)
    delta = rounded_div(ring.modulus_of(base), modulus.t())
    delta_mont2 = [ring.scalar_tensor_big_mont2(base, delta << (64 * k)) for k in range(data.limbs())]
    result = ring.allocate(2 * modulus.digits(), base)
    be.hal_assign(result[0:2 * d], ct)
    carry = None
    for i in range(modulus.digits()):
        input_row_bound = ct_bound if i < d else 1
        rows = result[2 * i:2 * i + 2]
        if carry is not None:
            acc = ring.allocate_wide(2, base)
            for k, factor in enumerate(delta_mont2):
                acc.fma_unbounded(range(2), carry[:, k:k + 1], factor)
            acc.add(range(2), rows, input_row_bound)
            ring.reduce_wide_into(rows, acc)
        row_bound = input_row_bound if carry is None else 1
        carry = split_overflow(ring, data, rows, row_bound, base)
    return result, 1


def dbfv_mul_convolve(ring, modulus, lhs_l, rhs_l, lifted, base_len, keys, digits, output_basis):
    """dbfv.rs:1078-1143: rows 3k..3k+3 of one wide batch hold the three components of output
    digit k, each an inner product over the digit pairs (j, k - j); one reduction, transform and
    rescale; then relin per output digit."""
    from .ring import relin, rescale_to_base

    d1, d2 = lhs_l.shape[0] // 2, rhs_l.shape[0] // 2
    d_out = min(modulus.digits(), d1 + d2 - 1)
    acc = ring.allocate_wide(3 * d_out, lifted)
    for k in range(d_out):
        t0, t1, t2 = [], [], []
        for j in range(max(0, k - (d2 - 1)), min(k, d1 - 1) + 1):
            i = k - j
            t0.append((lhs_l[2 * j:2 * j + 1], 1, rhs_l[2 * i:2 * i + 1], 1))
            t1.append((lhs_l[2 * j:2 * j + 1], 1, rhs_l[2 * i + 1:2 * i + 2], 1))
            t1.append((lhs_l[2 * j + 1:2 * j + 2], 1, rhs_l[2 * i:2 * i + 1], 1))
            t2.append((lhs_l[2 * j + 1:2 * j + 2], 1, rhs_l[2 * i + 1:2 * i + 2], 1))
        acc.inner_product(range(3 * k, 3 * k + 1), t0)
        acc.inner_product(range(3 * k + 1, 3 * k + 2), t1)
        acc.inner_product(range(3 * k + 2, 3 * k + 3), t2)
    prod = ring.reduce_wide(acc)
    ring.crt_to_pb_tensor(prod, lifted)
    base = lifted[:base_len]
    rescaled = rescale_to_base(ring, prod, lifted, base_len, modulus.t())
    result = ring.allocate(2 * d_out, base)
    bound = 1
    for k in range(d_out):
        rel, bound = relin(ring, rescaled[3 * k:3 * k + 3], base, keys, digits, output_basis)
        ring.backend.hal_assign(result[2 * k:2 * k + 2], rel)
    return result, bound


def dbfv_hom_mul(ring, modulus, lhs, lhs_bound, rhs, rhs_bound, base, keys, digits, output_basis="pb"):
    """dbfv.rs:1033-1056"""
    from .ring import lift_to_extended

    assert lhs.shape[0] // 2 <= modulus.digits() and rhs.shape[0] // 2 <= modulus.digits()
    lhs_red, lb = dbfv_reduce(ring, modulus, lhs, lhs_bound, base)
    rhs_red, rb = dbfv_reduce(ring, modulus, rhs, rhs_bound, base)
    if lhs_red.shape[0] == 0:
        return lhs_red, lb
    if rhs_red.shape[0] == 0:
        return rhs_red, rb
    assert lb <= 2 and rb <= 2
    lhs_l, lifted = lift_to_extended(ring, lhs_red, base)
    rhs_l, _ = lift_to_extended(ring, rhs_red, base)
    return dbfv_mul_convolve(ring, modulus, lhs_l, rhs_l, lifted, len(base), keys, digits, output_basis)


def dbfv_hom_square(ring, modulus, ct, ct_bound, base, keys, digits, output_basis="pb"):
    """dbfv.rs:1058-1076"""
    from .ring import lift_to_extended

    assert ct.shape[0] // 2 <= modulus.digits()
    red, bound = dbfv_reduce(ring, modulus, ct, ct_bound, base)
    if red.shape[0] == 0:
        return red, bound
    assert bound <= 2
    lifted_ct, lifted = lift_to_extended(ring, red, base)
    return dbfv_mul_convolve(ring, modulus, lifted_ct, lifted_ct, lifted, len(base), keys, digits, output_basis)


# ---------------------------------------------------------------------------------------------
# dBFV accumulators and exact division (dbfv.rs:82-198, 300-408, 464-533)

def cmod(a, modulus):
    """cmod (ring.rs:93-104): Rust's truncating remainder moved into (-modulus/2, modulus/2]"""
    from .ring import _trunc_mod

    half = modulus // 2
    r = _trunc_mod(a, modulus)
    if abs(r) >= half:
        if r < 0:
            r += modulus
        elif r > half:
            r -= modulus
    return r


def balanced_digits(x, b, max_digits):
    """balanced_digits (dbfv.rs:82-94): at most max_digits balanced base-b digits, stopping at 0;
    the part of x beyond max_digits digits is dropped"""
    out = []
    for _ in range(max_digits):
        if x == 0:
            break
        rem = cmod(x, b)
        out.append(rem)
        x = (x - rem) // b
    return out


class Ct:
    """A dBFV ciphertext (dbfv.rs:35): rows 2i, 2i + 1 are digit i; with its RNS base, basis
    ("pb" or "crt") and representative bound, as the reference's RingEls carries them."""

    def __init__(self, data, base, basis="pb", bound=1):
        self.data, self.base, self.basis, self.bound = data, list(base), basis, bound

    def digits(self):
        assert self.data.shape[0] % 2 == 0
        return self.data.shape[0] // 2


class DBFVAccumulator:
    """DBFVAccumulator (dbfv.rs:121-147): a wide batch of 2 digits() rows, the modulus, the basis
    and the number of digits used so far"""

    def __init__(self, ring, modulus, base, basis):
        self.acc = ring.allocate_wide(2 * modulus.digits(), base)
        self.modulus, self.basis, self.used_digits = modulus, basis, 0

    def capacity(self):
        return self.modulus.digits()

    def rns_base(self):
        return self.acc.rns_base


def dbfv_allocate_accumulator(ring, modulus, base, basis):
    """dbfv.rs:150-157"""
    return DBFVAccumulator(ring, modulus, base, basis)


def dbfv_acc_add(ring, acc, ct):
    """dbfv.rs:160-171"""
    d = ct.digits()
    if d == 0:
        return
    assert d <= acc.capacity() and acc.basis == ct.basis and acc.rns_base() == ct.base
    acc.acc.add(range(0, 2 * d), ct.data, ct.bound)
    acc.used_digits = max(acc.used_digits, d)


def dbfv_acc_sub(ring, acc, ct):
    """dbfv.rs:185-196"""
    d = ct.digits()
    if d == 0:
        return
    assert d <= acc.capacity() and acc.basis == ct.basis and acc.rns_base() == ct.base
    acc.acc.sub(range(0, 2 * d), ct.data, ct.bound)
    acc.used_digits = max(acc.used_digits, d)


def scalar_tensor_big(ring, base, scalar):
    """scalar_tensor_big (ring.rs:890-903): mont(smallest_lift(scalar mod p)) per prime"""
    from .ring import mont_repr, smallest_lift

    return ring._import([mont_repr(smallest_lift(scalar, ring.primes[i]), ring.primes[i]) for i in base],
                        [1, len(base), 1, 1])


def dbfv_acc_add_constant(ring, acc, scalar):
    """dbfv.rs:174-182: adds scalar Delta (times the period-basis sign) to row 0; the digit count
    is left as it is"""
    delta = rounded_div(ring.modulus_of(acc.rns_base()), acc.modulus.t())
    if acc.basis == "pb":
        scalar *= ring.field.scale_q_to_m_period_basis()
    tensor = scalar_tensor_big(ring, acc.rns_base(), scalar * delta)
    acc.acc.add(range(0, 1), tensor, 1)


def dbfv_acc_inner_product_tensors(ring, acc, terms):
    """dbfv.rs:300-348: for each output digit k one inner product over the convolution diagonal
    i + j = k of all terms (lhs Ct, list of factor tensors); the rows written must be fresh"""
    capacity = acc.capacity()
    max_new = 0
    for lhs, rhs in terms:
        ld = lhs.digits()
        if ld == 0 or not rhs:
            continue
        assert ld <= capacity and acc.basis == lhs.basis and acc.rns_base() == lhs.base
        max_new = max(max_new, min(capacity, ld + len(rhs) - 1))
    if max_new == 0:
        return
    for k in range(max_new):
        k_terms = []
        for lhs, rhs in terms:
            ld = lhs.digits()
            if ld == 0 or not rhs:
                continue
            if k >= min(capacity, ld + len(rhs) - 1):
                continue
            for j in range(len(rhs)):
                if k < j:
                    break
                i = k - j
                if i >= ld:
                    continue
                k_terms.append((lhs.data[2 * i:2 * i + 2], lhs.bound, rhs[j], 1))
        if k_terms:
            acc.acc.inner_product(range(2 * k, 2 * k + 2), k_terms)
    acc.used_digits = max(acc.used_digits, max_new)


def dbfv_acc_inner_product_scalar(ring, acc, terms):
    """dbfv.rs:356-373: terms are (lhs Ct, scalar); each scalar in balanced base-b digits"""
    m = acc.modulus
    prepared = [(lhs, [ring.scalar_tensor(acc.rns_base(), dg) for dg in balanced_digits(s, m.b(), m.digits())])
                for lhs, s in terms]
    dbfv_acc_inner_product_tensors(ring, acc, prepared)


def dbfv_acc_fma_scalar(ring, acc, lhs, rhs):
    """dbfv_acc_fma_scalar (dbfv.rs:236-248) over dbfv_acc_fma_tensors (dbfv.rs:202-232)"""
    m = acc.modulus
    rhs_t = [ring.scalar_tensor(acc.rns_base(), dg) for dg in balanced_digits(rhs, m.b(), m.digits())]
    ld = lhs.digits()
    if ld == 0 or not rhs_t:
        return
    assert ld <= acc.capacity() and acc.basis == lhs.basis and acc.rns_base() == lhs.base
    new_digits = min(acc.capacity(), ld + len(rhs_t) - 1)
    for i in range(ld):
        for j in range(len(rhs_t)):
            k = i + j
            if k < new_digits:
                acc.acc.fma(range(2 * k, 2 * k + 2), lhs.data[2 * i:2 * i + 2], lhs.bound, rhs_t[j], 1)
    acc.used_digits = max(acc.used_digits, new_digits)


def dbfv_finish_accumulator(ring, acc):
    """dbfv.rs:402-406: the used rows reduced (reduce_wide_rows), bound 1"""
    rows = 2 * acc.used_digits
    out = ring.allocate(rows, acc.rns_base())
    if rows:
        ring.reduce_wide_tensor(out, acc.acc, range(rows))
    return Ct(out, acc.rns_base(), acc.basis, 1)


def dbfv_exact_div_pow2(ring, initial_modulus, val, log2_div):
    """dbfv_exact_div_pow2 (dbfv.rs:500-533): digit 0 becomes sum_(i <= min(k, d - 1)) b^i y_i,
    digit j >= 1 becomes b^k y_(j + k), k = ceil(log2_div / log2_b)"""
    be = ring.backend
    d = val.digits()
    assert d <= initial_modulus.digits()
    assert log2_div < initial_modulus.log2_t, "the result modulus t' = t / 2^log2_div must be at least 2"
    if log2_div == 0 or d <= 1:
        return val
    log2_b = initial_modulus.log2_b
    k = (log2_div - 1) // log2_b + 1
    base = val.base
    assert_mul_mont_safe(ring, val.bound, base)
    out_digits = max(d - k, 1) if d > k else 1
    result = ring.allocate(2 * out_digits, base)
    if out_digits > 1:
        b_pow_k = scalar_tensor_big(ring, base, 1 << (k * log2_b))
        be.hal_assign(result[2:], val.data[2 * (k + 1):])
        be.hal_mul_mont64(result[2:], b_pow_k, ring.mods_for(base))
    factors = [scalar_tensor_big(ring, base, 1 << (i * log2_b)) for i in range(min(k, d - 1) + 1)]
    terms = [(val.data[2 * i:2 * i + 2], val.bound, f, 1) for i, f in enumerate(factors)]
    acc = ring.allocate_wide(2, base)
    acc.inner_product(range(0, 2), terms)
    ring.reduce_wide_tensor(result[0:2], acc, range(2))
    return Ct(result, base, val.basis, 1)


def mod_switch_ct(ring, ct, new_len):
    """mod_switch_down (bfv.rs:628-662) of a period-basis dBFV ciphertext: bound 1 over the
    length-new_len prefix of its base (a copy when nothing is dropped, with its bound)"""
    from .ring import mod_switch_down

    assert ct.basis == "pb" and new_len <= len(ct.base)
    if new_len == len(ct.base):
        return Ct(ct.data.copy(), ct.base, "pb", ct.bound)
    assert_mul_mont_safe(ring, ct.bound + 1, ct.base[:new_len])
    return Ct(mod_switch_down(ring, ct.data, ct.base, new_len), ct.base[:new_len], "pb", 1)
