"""The dBFV digit reduction of cyclosubfield-dbfv/src/dbfv.rs as HAL call sequences on the
MI355X: ``split_overflow`` (dbfv.rs:873-948) and the sequential carry chain ``dbfv_reduce``
(dbfv.rs:950-1025). The constants of ``SplitOverflowData::new`` (dbfv.rs:785-841) are computed
on the host with exact integers, as the reference computes them, and uploaded once; every pass
over the ciphertexts runs in the backend (``hal.HipBackend``) through the ring (``ring.py``):
one Montgomery multiply, one i128 matmul, one power-of-two base-conversion correction and one
wide accumulation per digit.

A dBFV ciphertext is a period-basis batch whose rows 2i, 2i + 1 hold the BFV ciphertext of
digit i (dbfv.rs:32-38). ``dbfv_reduce_batch`` runs many ciphertexts through the same chain with
every HAL call covering all of them; its rows are digit-major, so that each digit of all
ciphertexts is one contiguous run of rows.

Above the reduction sits the dBFV multiplication (dbfv.rs:1026-1143): ``dbfv_hom_mul``,
``dbfv_hom_square`` and ``dbfv_hom_mul_batch`` (many pairs, digit-major) reduce both sides, lift
them to the extended base, form the digit convolution in one wide batch, rescale by t and
relinearize every output digit in one batch (``bfv.relin_batch``)."""
import numpy as np


def _checked(ring, t, base, bound, what):
    """the debug range check (RingContext.set_range_check) of a produced tensor against its
    tracked bound; returns t"""
    if ring.range_checks and t.shape[0]:
        ring.check_range(t, base, bound, what)
    return t


def _log2_ceil(x):
    """feanor-math abs_log2_ceil for x >= 1"""
    return (abs(x) - 1).bit_length()


class DBFVModulus:
    """DBFVModulus (dbfv.rs:56-80): digit base b = 2^log2_b, total precision t = 2^log2_t"""

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


def split_i128_limbs(x):
    """split_i128_limbs (hal/mod.rs:66-76): x = hi 2^64 + lo, lo the low 64 bits of x as an i64"""
    if abs(x) >= 1 << 126:
        raise ValueError("split_i128_limbs needs |x| < 2^126")
    lo = x & ((1 << 64) - 1)
    lo = lo - (1 << 64) if lo >> 63 else lo
    return lo, (x - lo) >> 64


def rounded_div(a, b):
    """round(a / b) for b > 0, halves away from zero (feanor-math rounded_div)"""
    q, r = divmod(abs(a), b)
    if 2 * r >= b:
        q += 1
    return q if a >= 0 else -q


class SplitOverflowData:
    """SplitOverflowData::new (dbfv.rs:785-841): the BEHZ-style constants that split a
    coefficient into its remainder modulo b Delta and the overflow modulo R = t / b."""

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
        # neither inner product of the matmul may overflow an i128 (dbfv.rs:799-806)
        if ((gamma // min_p) + 1) * max_p * len(base) >= 1 << 127:
            raise ValueError("split_overflow: base too large for the i128 correction row")
        q_over_p_mod_r = [(q // p) % r_big for p in primes]
        values = [split_i128_limbs(c)[0] for c in q_over_p_mod_r]
        if limbs == 2:
            values += [split_i128_limbs(c)[1] for c in q_over_p_mod_r]
        values += [rounded_div(gamma, p) for p in primes]
        be = ring.backend
        self.matrix = be.hal_allocate_i64([1, limbs + 1, len(base), 1])
        be.hal_import(self.matrix, values)
        self.q_mod_r = q % r_big
        # the inverse modulo 2^128 (dbfv.rs:763-770), negated and reduced modulo R
        self.neg_q_inv_mod_r = (-pow(self.q_mod_r, -1, 1 << 128)) % r_big
        delta = rounded_div(q, modulus.t())
        neg_b_delta = -(delta * modulus.b())
        self.r_q_over_p_inv = ring.q_over_p_inv_tensor(base, r_big)
        self.neg_b_delta_mont2 = [ring.scalar_tensor_big_mont2(base, neg_b_delta << (64 * k)) for k in range(limbs)]
        self.delta_mont2 = [ring.scalar_tensor_big_mont2(base, delta << (64 * k)) for k in range(limbs)]
        # the same constants as one [1, 1, 3 + 3 limbs, r] table for the fused kernel
        # (csd_dbfv_split_overflow): lifts, value limbs, gamma / p, -b Delta, mont(1), Delta
        tables = [self.r_q_over_p_inv] + [None] * (limbs + 1) + self.neg_b_delta_mont2 + [ring.scalar_tensor(base, 1)]
        tables += self.delta_mont2
        host = []
        for idx, t in enumerate(tables):
            if t is None:
                host.append(values[(idx - 1) * len(base):idx * len(base)])
            else:
                host.append([int(v) for v in be.hal_export(t)])
        self.consts = be.hal_allocate_i64([1, 1, len(host), len(base)])
        be.hal_import(self.consts, [v for row in host for v in row])
        self.max_p, self.min_p = max_p, min_p

    def limbs(self):
        """one i64 limb of overflow for R <= 2^62, else two (bconv_correct_pow2_wide)"""
        return 1 if self.log2_r <= 62 else 2


def split_overflow_data(ring, modulus, base):
    """split_overflow_data (dbfv.rs:846-871): cached per (modulus, base) on the ring"""
    key = ("split_overflow", modulus.log2_b, modulus.log2_t, tuple(base))
    hit = ring._consts.get(key)
    if hit is None:
        hit = ring._consts[key] = SplitOverflowData(ring, modulus, base)
    return hit


def _assert_mul_mont_safe(ring, bound, base):
    """assert_mul_mont_safe (ring.rs:955-960)"""
    if bound * max(ring.primes[i] for i in base) > 1 << 63:
        raise ValueError("representatives too large for a montgomery multiplication")


def _needs_full_reduction(data, steps):
    """the bound tracking of a fresh WideRingEls over the accumulation steps (ring.rs:354-368):
    whether reduce_wide takes the double reduction (ring.rs:843-865)"""
    bound = 0
    for s in steps:
        bound += s
        if bound > (1 << 127) - 1:
            raise OverflowError("wide ring element overflow: too many accumulation steps")
    return bound > (data.min_p << 63)


def _split_fused(ring, data, x, x_bound, base, carry=None, carry_bound=1, x_in=None, zero_in=False):
    """split_overflow, preceded by the carry step when `carry` is given, as one kernel; the
    reduction kinds are those the wide accumulators of the call sequence would choose. The digit
    comes from `x_in` (or is zero with `zero_in`) when given, and x_lo goes to `x`."""
    be = ring.backend
    limbs = data.limbs()
    unbounded = [data.max_p << 63] * limbs
    full_carry = False
    if carry is not None:
        _assert_mul_mont_safe(ring, carry_bound, base)
        full_carry = _needs_full_reduction(data, unbounded + [carry_bound * data.max_p * data.max_p])
        x_bound = 1
    _assert_mul_mont_safe(ring, x_bound, base)
    full_split = _needs_full_reduction(data, unbounded + [x_bound * data.max_p * data.max_p])
    overflow = be.allocate_i64_uninit([x.shape[0], limbs, x.shape[2], x.shape[3]])  # written entirely
    be.dbfv_split_overflow(x, overflow, carry, data.consts, ring.mods_for(base), data.q_mod_r, data.neg_q_inv_mod_r,
                           data.log2_gamma, data.log2_r, full_split, full_carry, x_in=x_in, zero_in=zero_in)
    return overflow


def split_overflow(ring, data, x, x_bound, base, fused=True):
    """split_overflow (dbfv.rs:893-948): the period-basis rows `x` ([count, r, ., .]) become
    x_lo in place, with x = x_lo + (b Delta) x_hi mod Q and |x_lo| <= 2 b Delta; returns x_hi as a
    [count, limbs, ., .] tensor of plain values in (-R/2, R/2]. `fused` runs the whole sequence
    as one kernel (csd_dbfv_split_overflow); otherwise it is the reference's HAL call sequence.
    Both give identical residues."""
    be = ring.backend
    shape = x.shape
    if len(base) != shape[1]:
        raise ValueError("split_overflow: the tensor's limbs must match the base")
    if fused:
        return _split_fused(ring, data, x, x_bound, base)
    _assert_mul_mont_safe(ring, x_bound, base)
    limbs = data.limbs()
    mods = ring.mods_for(base)
    # plain lifts of (x R) (Q/p)^-1 mod p, in (-p, p)
    lifts = be.hal_allocate_i64(shape)
    be.hal_assign(lifts, x)
    be.hal_mul_mont64(lifts, data.r_q_over_p_inv, mods)
    # per coefficient: sum_p lift_p (Q/p) mod R in limbs, and the gamma-scaled multiple-of-Q estimate
    sums = be.hal_allocate_i128([shape[0], limbs + 1, shape[2], shape[3]])
    be.hal_matmul_i64(sums, data.matrix, lifts)
    overflow = be.hal_allocate_i64([shape[0], limbs, shape[2], shape[3]])
    if limbs == 1:
        be.hal_bconv_correct_pow2(overflow, sums, data.q_mod_r, data.neg_q_inv_mod_r, data.log2_gamma, data.log2_r)
    else:
        be.hal_bconv_correct_pow2_wide(overflow, sums, data.q_mod_r, data.neg_q_inv_mod_r, data.log2_gamma, data.log2_r)
    # x_lo = x + overflow (-b Delta), the limbs folded into mont^2(-b Delta 2^(64 k))
    acc = ring.allocate_wide(shape[0], base)
    rows = range(shape[0])
    for k in range(limbs):
        acc.fma_unbounded(rows, overflow[:, k:k + 1], data.neg_b_delta_mont2[k])
    acc.add(rows, x, x_bound)
    ring.reduce_wide_into(x, acc)
    return overflow


# whether the fused chain runs as one launch (csd_dbfv_reduce_chain) or one launch per digit
# (csd_dbfv_split_overflow); both give identical residues
_CHAIN_ONE_LAUNCH = [True]


def _chain_one_launch(ring, modulus, data, result, src, d, per_digit, ct_bound, base):
    """the fused carry chain as one kernel: every digit's carry step and split, with the checks
    and reduction kinds _split_fused takes for it digit by digit"""
    limbs = data.limbs()
    unbounded = [data.max_p << 63] * limbs
    full_split, full_carry = [], []
    for i in range(modulus.digits()):
        input_row_bound = ct_bound if i < d else 1
        x_bound = input_row_bound
        if i > 0:
            _assert_mul_mont_safe(ring, input_row_bound, base)
            full_carry.append(_needs_full_reduction(data, unbounded + [input_row_bound * data.max_p * data.max_p]))
            x_bound = 1
        else:
            full_carry.append(False)
        _assert_mul_mont_safe(ring, x_bound, base)
        full_split.append(_needs_full_reduction(data, unbounded + [x_bound * data.max_p * data.max_p]))
    ring.backend.dbfv_reduce_chain(result, src[0:d * per_digit], per_digit, modulus.digits(), data.consts,
                                   ring.mods_for(base), limbs, data.q_mod_r, data.neg_q_inv_mod_r, data.log2_gamma,
                                   data.log2_r, full_split, full_carry)


def _reduce_chain(ring, modulus, result, d, per_digit, ct_bound, base, fused, src=None):
    """the carry chain of dbfv_reduce (dbfv.rs:990-1023) over `result`, whose digit i occupies
    rows [i per_digit, (i + 1) per_digit) and whose first d digits come from the input: already
    copied into `result`, or read from `src` (same layout) by the fused kernels, which then
    take the digits beyond the input as zero instead of reading zeroed rows"""
    data = split_overflow_data(ring, modulus, base)
    if fused and src is not None and _CHAIN_ONE_LAUNCH[0] and modulus.digits() <= 32 and len(base) <= 40:
        _chain_one_launch(ring, modulus, data, result, src, d, per_digit, ct_bound, base)
        return
    delta_mont2 = data.delta_mont2
    carry = None
    rows = range(per_digit)
    for i in range(modulus.digits()):
        # rows copied from the (possibly unreduced) input have its bound; rows beyond it are zero
        input_row_bound = ct_bound if i < d else 1
        digit = result[i * per_digit:(i + 1) * per_digit]
        if fused:
            # the carry step and the split of this digit as one kernel
            x_in = src[i * per_digit:(i + 1) * per_digit] if src is not None and i < d else None
            zero_in = src is not None and i >= d
            carry = _split_fused(ring, data, digit, input_row_bound, base, carry, input_row_bound, x_in, zero_in)
            continue
        if carry is not None:
            # digit_i += Delta carry; the carry limbs are plain integers, hence the unbounded fma
            acc = ring.allocate_wide(per_digit, base)
            for k, factor in enumerate(delta_mont2):
                acc.fma_unbounded(rows, carry[:, k:k + 1], factor)
            acc.add(rows, digit, input_row_bound)
            ring.reduce_wide_into(digit, acc)
        row_bound = input_row_bound if carry is None else 1
        carry = split_overflow(ring, data, digit, row_bound, base, fused=False)


def _reduce_into_new(ring, modulus, src, d, per_digit, ct_bound, base, fused, out=None):
    """the result batch of dbfv_reduce (dbfv.rs:985-988, 1023-1024). The call sequence allocates
    it zeroed and copies the input digits in; the fused kernels read the input digits in place
    and write every row, so the result needs neither the fill nor the copy. With `out` (fused
    only: a [rows, r, ., .] view, any row and limb strides) the fused kernels write the result
    there."""
    rows = modulus.digits() * per_digit
    if out is not None and (not fused or out.shape[0] != rows or out.shape[1] != len(base)):
        raise ValueError("dbfv_reduce: an output view needs the fused form and the result's shape")
    if fused:
        result = out if out is not None else ring.backend.allocate_i64_uninit(
            [rows, len(base), ring.info["deg_lprime"], ring.info["deg_k"]])
        _reduce_chain(ring, modulus, result, d, per_digit, ct_bound, base, True, src=src)
    else:
        result = ring.allocate(rows, base)
        ring.backend.hal_assign(result[0:d * per_digit], src)
        _reduce_chain(ring, modulus, result, d, per_digit, ct_bound, base, False)
    return _checked(ring, result, base, 1, "dbfv_reduce"), 1


def dbfv_reduce(ring, modulus, ct, ct_bound, base, fused=True):
    """dbfv_reduce (dbfv.rs:965-1025) of one period-basis dBFV ciphertext (rows 2i, 2i + 1 =
    digit i): every digit is reduced modulo b Delta and its carry added to the next digit, the
    carries of digits >= modulus.digits() dropped. Returns (result, bound). `fused` runs each
    digit's carry step and split as one kernel; otherwise the reference's HAL call sequence."""
    if ct.shape[0] % 2:
        raise ValueError("a dBFV ciphertext has two rows per digit")
    d = ct.shape[0] // 2
    if d > modulus.digits():
        raise ValueError("ciphertext has more digits than the modulus")
    if modulus.digits() == 1 or d == 0:
        return ct, ct_bound
    return _reduce_into_new(ring, modulus, ct, d, 2, ct_bound, base, fused)


def dbfv_reduce_batch(ring, modulus, cts, count, ct_bound, base, fused=True):
    """dbfv_reduce for `count` ciphertexts of d digits at once. `cts` holds them digit-major:
    row (2 i + s) count + c is row s of digit i of ciphertext c; the result ([2 digits count,
    r, ., .]) has the same layout. Every HAL call covers all ciphertexts; each ciphertext sees
    the calls and operand bounds of dbfv_reduce, so its residues are identical."""
    if count < 1 or cts.shape[0] % (2 * count):
        raise ValueError("dbfv_reduce_batch: count must be positive and the rows 2 x digits x count")
    d = cts.shape[0] // (2 * count)
    if d > modulus.digits():
        raise ValueError("ciphertexts have more digits than the modulus")
    if modulus.digits() == 1 or d == 0:
        return cts, ct_bound
    return _reduce_into_new(ring, modulus, cts, d, 2 * count, ct_bound, base, fused)


def _reduced_lifted(ring, modulus, ct, ct_bound, count, base, fused):
    """dbfv_reduce_batch, then lift_to_extended (bfv.rs:564-595) of the reduced digits: returns
    (reduced, bound, lifted tensor, lifted base), or (reduced, bound, None, None) when the reduced
    batch has no digits. With the fused chain the reduced digits are written straight into the
    base limbs of the lifted tensor, so the lift's copy of them is left out (reduced is then
    None); the residues are the same."""
    from .bfv import lift_to_extended

    if count < 1 or ct.shape[0] % (2 * count):
        raise ValueError("dBFV ciphertexts need two rows per digit and pair")
    d = ct.shape[0] // (2 * count)
    if d > modulus.digits():
        raise ValueError("ciphertext has more digits than the modulus")
    if not fused or modulus.digits() == 1 or d == 0:
        red, bound = _reduced_operand(ring, modulus, ct, ct_bound, count, base, fused)
        if red.shape[0] == 0:
            return red, bound, None, None
        lifted_t, lifted = lift_to_extended(ring, red, base)
        return red, bound, lifted_t, lifted
    ext = ring.extension_rns_base_for(base)
    lifted = list(base) + ext
    bl = len(base)
    out = ring.allocate_uninit(modulus.digits() * 2 * count, lifted)
    _reduce_into_new(ring, modulus, ct, d, 2 * count, ct_bound, base, True, out=out[:, :bl])  # bound 1
    ring.bconv_tensors(out[:, bl:], ext, out[:, :bl], base, preserve_src=True)
    ring.pb_to_crt_tensor(out, lifted)
    return None, 1, out, lifted


def _mul_convolve(ring, modulus, lhs_l, rhs_l, count, lifted, base_len, keys, digits, output_basis):
    """dbfv_mul_convolve (dbfv.rs:1078-1143) of `count` pairs of lifted CRT-basis dBFV
    ciphertexts (digit-major: row (2 i + s) count + c is row s of digit i of pair c): the digit
    products as a convolution, one inner product per output digit and component covering all
    pairs, in one wide batch; one reduction, basis transform and rescale for all of them; then
    one batched relinearization of every output digit of every pair. The inner products and their
    wide reduction run as one kernel (csd_convolve_reduce), which forms each row's exact i128 sum
    and reduces it by the kind its WideRingEls bound selects. The rows are laid out component-major
    (row (c d_out + k) count + pair instead of 3 k + c) so that the relinearizations run as one
    batch; each row gets the same value, so the residues are those of the reference's layout. The
    result is digit-major like the inputs."""
    from .bfv import relin_batch, rescale_to_base

    be = ring.backend
    n = count
    d1, d2 = lhs_l.shape[0] // (2 * n), rhs_l.shape[0] // (2 * n)
    K = min(modulus.digits(), d1 + d2 - 1)

    # the inner products of every output digit and component and reduce_wide, as one kernel
    prod = ring.convolve_reduce(lhs_l, (2 * n, n, 1), rhs_l, (2 * n, n, 1), n, d1, d2, K, lifted)  # CRT basis
    ring.crt_to_pb_tensor(prod, lifted)
    rescaled = rescale_to_base(ring, prod, lifted, base_len, modulus.t())  # bound 1
    out, bound = relin_batch(ring, rescaled, K * n, lifted[:base_len], keys, digits, output_basis, bound=1,
                             interleave=False, add_c01=output_basis == "crt")
    # product (k, pair) component s sits at row s K n + k n + pair; the result is digit-major, row
    # (2 k + s) n + pair. Seen as [K, n r, ., .] tensors (the n rows of one k, limbs included, are
    # consecutive planes), each component moves with one strided copy; in period basis c0, c1 of
    # the rescaled product are added on the way (add_els, ring.rs:990-1023)
    base = lifted[:base_len]
    result = ring.allocate_uninit(2 * K * n, base)
    for s_ in (0, 1):
        dst = digit_rows_view(result, s_ * n, K, n, 2 * n)
        if output_basis == "pb":
            be.sum_i64(dst, digit_rows_view(out, s_ * K * n, K, n, n), digit_rows_view(rescaled, s_ * K * n, K, n, n))
        else:
            be.hal_assign(dst, digit_rows_view(out, s_ * K * n, K, n, n))
    if output_basis == "pb":
        bound = 2
    return _checked(ring, result, base, bound, "dbfv_hom_mul"), bound


def digit_rows_view(t, first, K, n, step):
    """rows first + k step + pair (k < K, pair < n) of a [rows, r, n2, n3] tensor as a [K, n r, n2,
    n3] view: the n rows of one k and their limbs are consecutive planes (no copy)"""
    from .hal import DeviceTensor

    plane = t.shape[2] * t.shape[3]
    if t.strides[1] != plane or t.strides[0] != t.shape[1] * plane:
        raise ValueError("digit rows need a contiguous tensor")
    return DeviceTensor(t.backend, t.ptr + first * t.strides[0] * t.elem, (K, n * t.shape[1], t.shape[2], t.shape[3]),
                        (step * t.strides[0], plane, t.strides[2], t.strides[3]), t.elem, t._owner or t)


def _reduced_operand(ring, modulus, ct, ct_bound, count, base, fused):
    if count < 1 or ct.shape[0] % (2 * count):
        raise ValueError("dBFV ciphertexts need two rows per digit and pair")
    if ct.shape[0] // (2 * count) > modulus.digits():
        raise ValueError("ciphertext has more digits than the modulus")
    red, bound = dbfv_reduce_batch(ring, modulus, ct, count, ct_bound, base, fused=fused)
    if red.shape[0] and bound > 2:
        raise ValueError("lifting requires representatives bounded by 2p (basis transform input)")
    return red, bound


def dbfv_hom_mul_batch(ring, modulus, lhs, lhs_bound, rhs, rhs_bound, count, base, keys, digits, output_basis="pb",
                       fused=True):
    """dbfv_hom_mul (dbfv.rs:1033-1056) of `count` pairs of period-basis dBFV ciphertexts over
    `base`, digit-major (row (2 i + s) count + c is row s of digit i of pair c; bounds lhs_bound,
    rhs_bound): both sides reduced (dbfv_reduce), lifted to the extended base, multiplied as a
    digit convolution and relinearized with the key (keys[i] the [2, r_key, ., .] CRT-basis key
    of digit i, `digits` its gadget digits). Returns (ciphertexts, bound), digit-major;
    output_basis "pb" (bound 2, the reference's choice before a reduction) or "crt" (bound 1).
    Every HAL call covers all pairs, and each pair gets the residues of its own dbfv_hom_mul. A
    side without digits is returned as it is."""
    if output_basis not in ("pb", "crt"):
        raise ValueError("output_basis must be 'crt' or 'pb'")
    lhs_red, lb, lhs_l, lifted = _reduced_lifted(ring, modulus, lhs, lhs_bound, count, base, fused)
    if lhs_l is None:
        return lhs_red, lb
    rhs_red, rb, rhs_l, _ = _reduced_lifted(ring, modulus, rhs, rhs_bound, count, base, fused)
    if rhs_l is None:
        return rhs_red, rb
    return _mul_convolve(ring, modulus, lhs_l, rhs_l, count, lifted, len(base), keys, digits, output_basis)


def dbfv_hom_mul(ring, modulus, lhs, lhs_bound, rhs, rhs_bound, base, keys, digits, output_basis="pb", fused=True):
    """dbfv_hom_mul (dbfv.rs:1033-1056) of one pair: rows 2i, 2i + 1 = digit i; see
    dbfv_hom_mul_batch (one pair is the batch of one)."""
    return dbfv_hom_mul_batch(ring, modulus, lhs, lhs_bound, rhs, rhs_bound, 1, base, keys, digits, output_basis, fused)


def dbfv_hom_square(ring, modulus, ct, ct_bound, base, keys, digits, output_basis="pb", fused=True):
    """dbfv_hom_square (dbfv.rs:1058-1076): dbfv_hom_mul of a ciphertext with itself, reduced
    and lifted once."""
    if output_basis not in ("pb", "crt"):
        raise ValueError("output_basis must be 'crt' or 'pb'")
    red, bound, lifted_ct, lifted = _reduced_lifted(ring, modulus, ct, ct_bound, 1, base, fused)
    if lifted_ct is None:
        return red, bound
    return _mul_convolve(ring, modulus, lifted_ct, lifted_ct, 1, lifted, len(base), keys, digits, output_basis)


def dbfv_hom_square_batch(ring, modulus, ct, ct_bound, count, base, keys, digits, output_basis="pb", fused=True):
    """dbfv_hom_square (dbfv.rs:1058-1076) of `count` period-basis dBFV ciphertexts, digit-major
    (row (2 i + s) count + c is row s of digit i of ciphertext c): one reduction and one lift for
    all of them, then the digit convolution of each with itself and one batched
    relinearization. Each ciphertext gets the residues of its own dbfv_hom_square."""
    if output_basis not in ("pb", "crt"):
        raise ValueError("output_basis must be 'crt' or 'pb'")
    red, bound, lifted_ct, lifted = _reduced_lifted(ring, modulus, ct, ct_bound, count, base, fused)
    if lifted_ct is None:
        return red, bound
    return _mul_convolve(ring, modulus, lifted_ct, lifted_ct, count, lifted, len(base), keys, digits, output_basis)


# ---------------------------------------------------------------------------------------------
# dBFV ciphertext batches, accumulators and exact division (dbfv.rs:82-198, 300-408, 464-533)

def cmod(a, modulus):
    """cmod (ring.rs:93-104): Rust's truncating remainder moved into (-modulus/2, modulus/2]"""
    r = abs(a) % modulus
    r = -r if a < 0 else r
    half = modulus // 2
    if abs(r) >= half:
        if r < 0:
            r += modulus
        elif r > half:
            r -= modulus
    return r


def balanced_digits(x, b, max_digits):
    """balanced_digits (dbfv.rs:82-94): at most max_digits balanced base-b digits of x, stopping
    at zero (what lies beyond max_digits digits is dropped)"""
    out = []
    for _ in range(max_digits):
        if x == 0:
            break
        rem = cmod(x, b)
        out.append(rem)
        x = (x - rem) // b
    return out


def balanced_digits_vec(xs, b, max_digits):
    """balanced_digits_vec (dbfv.rs:97-112): balanced base-b digit vectors of the coefficients, at
    most max_digits, stopping once every coefficient is zero"""
    out, cur = [], [int(x) for x in xs]
    for _ in range(max_digits):
        if all(x == 0 for x in cur):
            break
        rems = [cmod(x, b) for x in cur]
        cur = [(x - y) // b for x, y in zip(cur, rems)]
        out.append(rems)
    return out


def scale_q_to_m_period_basis(ring):
    """field.rs:447: (-1)^(number of distinct prime factors of m)"""
    m, p, k = ring.info["m"], 2, 0
    while p * p <= m:
        if m % p == 0:
            k += 1
            while m % p == 0:
                m //= p
        p += 1
    if m > 1:
        k += 1
    return -1 if k % 2 else 1


class DBFVBatch:
    """`count` dBFV ciphertexts (dbfv.rs:35) over one RNS base, digit-major: row (2 i + s) count + c
    is component s of digit i of ciphertext c, so digit i of every ciphertext is the contiguous
    run of rows [2 i count, (2 i + 2) count). With the basis ("pb" or "crt") and the
    representative bound a RingEls carries."""

    def __init__(self, data, base, count, basis="pb", bound=1):
        if count < 1 or data.shape[0] % (2 * count):
            raise ValueError("a dBFV batch has 2 x digits x count rows")
        self.data, self.base, self.count, self.basis, self.bound = data, list(base), count, basis, bound

    def digits(self):
        return self.data.shape[0] // (2 * self.count)

    def digit_rows(self, i, j=None):
        """rows of digits i .. j - 1 (j = i + 1 by default) of every ciphertext"""
        j = i + 1 if j is None else j
        return self.data[2 * i * self.count:2 * j * self.count]


class DBFVAccumulator:
    """DBFVAccumulator (dbfv.rs:121-147) for a batch of `count` ciphertexts: a wide batch with
    rows for every digit of the modulus, digit-major like DBFVBatch, the basis and the number of
    digits used so far. Its i128 rows start unwritten: finishing reduces only the used rows,
    which every sequence writes first (the fresh rows of the reference are overwritten too)."""

    def __init__(self, ring, modulus, base, basis, count):
        self.ring, self.modulus, self.basis, self.count = ring, modulus, basis, count
        self.acc = ring.allocate_wide(2 * modulus.digits() * count, base, zero=False)
        self.used_digits = 0

    def capacity(self):
        return self.modulus.digits()

    def rns_base(self):
        return self.acc.rns_base

    def rows(self, i, j):
        """the wide rows of digits i .. j - 1"""
        return range(2 * i * self.count, 2 * j * self.count)


def dbfv_allocate_accumulator(ring, modulus, base, basis, count=1):
    """dbfv_allocate_accumulator (dbfv.rs:150-157)"""
    return DBFVAccumulator(ring, modulus, base, basis, count)


def _check_operand(acc, ct):
    if ct.count != acc.count or ct.basis != acc.basis or ct.base != acc.rns_base():
        raise ValueError("accumulator and ciphertexts differ in batch size, basis or RNS base")
    if ct.digits() > acc.capacity():
        raise ValueError("ciphertext has more digits than the accumulator")


def dbfv_acc_add(ring, acc, ct):
    """dbfv_acc_add (dbfv.rs:160-171), digit-wise over the batch"""
    d = ct.digits()
    if d == 0:
        return
    _check_operand(acc, ct)
    acc.acc.add(acc.rows(0, d), ct.data, ct.bound)
    acc.used_digits = max(acc.used_digits, d)


def dbfv_acc_sub(ring, acc, ct):
    """dbfv_acc_sub (dbfv.rs:185-196)"""
    d = ct.digits()
    if d == 0:
        return
    _check_operand(acc, ct)
    acc.acc.sub(acc.rows(0, d), ct.data, ct.bound)
    acc.used_digits = max(acc.used_digits, d)


def dbfv_acc_add_constant(ring, acc, scalar):
    """dbfv_acc_add_constant (dbfv.rs:174-182): scalar Delta (times the period-basis sign) into
    component 0 of digit 0 of every ciphertext; the digit count is left as it is"""
    delta = rounded_div(ring.modulus_of(acc.rns_base()), acc.modulus.t())
    if acc.basis == "pb":
        scalar *= scale_q_to_m_period_basis(ring)
    acc.acc.add(range(0, acc.count), ring.scalar_tensor_big(acc.rns_base(), scalar * delta), 1)


def dbfv_acc_inner_product_tensors(ring, acc, terms):
    """dbfv_acc_inner_product_tensors (dbfv.rs:300-348): for every output digit k one inner
    product over the convolution diagonal i + j = k of all terms (a DBFVBatch, its factor
    tensors), covering every ciphertext of the batch; the rows written must be fresh"""
    capacity = acc.capacity()
    max_new = 0
    for lhs, rhs in terms:
        ld = lhs.digits()
        if ld == 0 or not rhs:
            continue
        _check_operand(acc, lhs)
        max_new = max(max_new, min(capacity, ld + len(rhs) - 1))
    if max_new == 0:
        return
    for k in range(max_new):
        k_terms = []
        for lhs, rhs in terms:
            ld = lhs.digits()
            if ld == 0 or not rhs or k >= min(capacity, ld + len(rhs) - 1):
                continue
            for j in range(len(rhs)):
                if k < j:
                    break
                i = k - j
                if i < ld:
                    k_terms.append((lhs.digit_rows(i), lhs.bound, rhs[j], 1))
        if k_terms:
            acc.acc.inner_product(acc.rows(k, k + 1), k_terms)
    acc.used_digits = max(acc.used_digits, max_new)


def dbfv_acc_inner_product_scalar(ring, acc, terms):
    """dbfv_acc_inner_product_scalar (dbfv.rs:356-373): terms (DBFVBatch, scalar), each scalar in
    balanced base-b digits"""
    m = acc.modulus
    prepared = [(lhs, [ring.scalar_tensor(acc.rns_base(), dg) for dg in balanced_digits(s, m.b(), m.digits())])
                for lhs, s in terms]
    dbfv_acc_inner_product_tensors(ring, acc, prepared)


def dbfv_finish_accumulator(ring, acc):
    """dbfv_finish_accumulator (dbfv.rs:402-406): the used rows reduced, bound 1"""
    rows = 2 * acc.used_digits * acc.count
    out = ring.allocate_uninit(rows, acc.rns_base())
    if rows:
        ring.reduce_wide_rows_into(out, acc.acc, rows)
    return DBFVBatch(_checked(ring, out, acc.rns_base(), 1, "dbfv_finish_accumulator"), acc.rns_base(), acc.count,
                     acc.basis, 1)


def dbfv_exact_div_pow2(ring, initial_modulus, val, log2_div):
    """dbfv_exact_div_pow2 (dbfv.rs:500-533) of a batch: digit 0 becomes
    sum_(i <= min(k, d - 1)) b^i y_i and digit j >= 1 becomes b^k y_(j + k), k = ceil(log2_div /
    log2_b); every call covers all ciphertexts"""
    be = ring.backend
    d = val.digits()
    if d > initial_modulus.digits():
        raise ValueError("ciphertext has more digits than the modulus")
    if log2_div >= initial_modulus.log2_t:
        raise ValueError("the result modulus t' = t / 2^log2_div must be at least 2")
    if log2_div == 0 or d <= 1:
        return val
    log2_b = initial_modulus.log2_b
    k = (log2_div - 1) // log2_b + 1
    base = val.base
    _assert_mul_mont_safe(ring, val.bound, base)
    out_digits = max(d - k, 1)
    n = val.count
    result = ring.allocate_uninit(2 * out_digits * n, base)
    if out_digits > 1:
        be.hal_assign(result[2 * n:], val.digit_rows(k + 1, d))
        be.hal_mul_mont64(result[2 * n:], ring.scalar_tensor_big(base, 1 << (k * log2_b)), ring.mods_for(base))
    terms = [(val.digit_rows(i), val.bound, ring.scalar_tensor_big(base, 1 << (i * log2_b)), 1)
             for i in range(min(k, d - 1) + 1)]
    acc = ring.allocate_wide(2 * n, base, zero=False)
    acc.inner_product(range(0, 2 * n), terms)
    ring.reduce_wide_into(result[0:2 * n], acc)
    return DBFVBatch(_checked(ring, result, base, 1, "dbfv_exact_div_pow2"), base, n, val.basis, 1)


def dbfv_mod_switch_down(ring, ct, new_len):
    """mod_switch_down (bfv.rs:628-662) of a period-basis batch to the length-new_len prefix of
    its base (bound 1), one call for every ciphertext"""
    if ct.basis != "pb" or not 1 <= new_len <= len(ct.base):
        raise ValueError("mod_switch_down needs period basis and a prefix of the base")
    if new_len == len(ct.base):
        out = ring.allocate_uninit(ct.data.shape[0], ct.base)
        ring.backend.hal_assign(out, ct.data)
        return DBFVBatch(out, ct.base, ct.count, "pb", ct.bound)
    _assert_mul_mont_safe(ring, ct.bound + 1, ct.base[:new_len])
    out = ring.allocate_uninit(ct.data.shape[0], ct.base[:new_len])
    ring.mod_switch_down(out, ct.data, ct.base, new_len)
    return DBFVBatch(_checked(ring, out, ct.base[:new_len], 1, "mod_switch_down"), ct.base[:new_len], ct.count, "pb", 1)


def _mont_repr(x, p):
    """mont_repr (ring.rs:109) of (x rem p): (y 2^64) rem p with Rust's remainder, y = x rem p"""
    y = abs(x) % p
    y = -y if x < 0 else y
    v = (abs(y) << 64) % p
    return -v if y < 0 else v


COMBINE_MAX_TERMS = 16


def dbfv_combine(ring, modulus, like, terms, adds=(), constant=None, basis="pb", lprime=False):
    """One combination of dBFV batches, as one kernel (csd_dbfv_combine): the result of
    dbfv_allocate_accumulator, dbfv_acc_inner_product_scalar(terms), dbfv_acc_add for each of
    `adds`, dbfv_acc_add_constant(constant) and dbfv_finish_accumulator (dbfv.rs:143-406). The
    reduction kind of every row is that of the accumulator's bound bookkeeping (ring.rs:354-366),
    so the residues are those of the accumulator sequence. `like` gives the base and batch size.
    With `lprime` the factors are elements of L', given by their period-basis coefficient vectors,
    and the inner product is dbfv_acc_inner_product_element_lprime's (dbfv.rs:378-396): the
    balanced digit vectors of each become CRT-basis tensors (lprime_element_tensor). At most
    COMBINE_MAX_TERMS terms with a non-zero factor, 4 adds and 12 output digits."""
    from .hal import check, lib, CsdView
    import ctypes

    base, n = like.base, like.count
    cap, b = modulus.digits(), modulus.b()
    max_p = max(ring.primes[i] for i in base)
    min_p = min(ring.primes[i] for i in base)
    step = max_p * max_p
    tl = []
    for lhs, s in terms:
        ds = balanced_digits_vec(s, b, cap) if lprime else balanced_digits(s, b, cap)
        ld = lhs.digits()
        if ld == 0 or not ds:
            continue
        if lhs.count != n or lhs.base != base or lhs.basis != basis:
            raise ValueError("combination operands differ in batch size, base or basis")
        if ld > cap:
            raise ValueError("ciphertext has more digits than the accumulator")
        tl.append((lhs, ds, min(cap, ld + len(ds) - 1)))
    al = []
    for a in adds:
        if a.digits() == 0:
            continue
        if a.count != n or a.base != base or a.basis != basis or a.digits() > cap:
            raise ValueError("added ciphertext differs in batch size, base, basis or digits")
        al.append(a)
    dout = max([t[2] for t in tl] + [a.digits() for a in al] + [0])
    if dout == 0:
        return DBFVBatch(like.data[0:0], base, n, basis, 1)
    if len(tl) > COMBINE_MAX_TERMS or len(al) > 4 or dout > 12:
        raise ValueError("combination beyond the fused kernel's operand limits")
    # the rows' bounds as the accumulator would track them
    bounds = [[0, 0] for _ in range(dout)]
    for lhs, ds, new_t in tl:
        for k in range(new_t):
            for jj in range(len(ds)):
                i = k - jj
                if 0 <= i < lhs.digits():
                    for s_ in range(2):
                        bounds[k][s_] += lhs.bound * step
    for a in al:
        for k in range(a.digits()):
            for s_ in range(2):
                bounds[k][s_] += a.bound * step
    cst = None
    if constant is not None:
        delta = rounded_div(ring.modulus_of(base), modulus.t())
        c = constant * (scale_q_to_m_period_basis(ring) if basis == "pb" else 1)
        cst = ring.scalar_tensor_big(base, c * delta)
        bounds[0][0] += step
    full = 0
    for k in range(dout):
        for s_ in range(2):
            if bounds[k][s_] > (1 << 127) - 1:
                raise OverflowError("wide ring element overflow: too many accumulation steps")
            if bounds[k][s_] > (min_p << 63):
                full |= 1 << (2 * k + s_)
    be = ring.backend
    r = len(base)
    sd = None
    if tl and not lprime:
        key = ("combine_sd", tuple(base), tuple(tuple(ds) for _, ds, _ in tl))
        sd = ring._consts.get(key)
        if sd is None:
            vals = [0] * (len(tl) * 12 * r)
            for t, (_, ds, _) in enumerate(tl):
                for jj, d in enumerate(ds):
                    for j, i in enumerate(base):
                        vals[(t * 12 + jj) * r + j] = _mont_repr(d, ring.primes[i])
            sd = be.hal_allocate_i64([1, len(tl), 12, r])
            be.hal_import(sd, vals)
            ring._consts[key] = sd
    elif tl:
        # [terms, 12, r, n2]: digit jj of factor t, from the cached L' tensors ([1, r, n2, 1] each)
        key = ("combine_lprime", tuple(base), tuple(tuple(tuple(v) for v in ds) for _, ds, _ in tl))
        sd = ring._consts.get(key)
        if sd is None:
            n2 = ring.info["deg_lprime"]
            vals = np.zeros((len(tl), 12, r, n2), dtype=np.int64)
            for t, (_, ds, _) in enumerate(tl):
                for jj, dv in enumerate(ds):
                    vals[t, jj] = np.asarray(be.hal_export(ring.lprime_element_tensor(base, dv))).reshape(r, n2)
            sd = be.hal_allocate_i64([len(tl), 12, r, n2])
            be.hal_import(sd, vals.reshape(-1))
            ring._consts[key] = sd
    out = ring.allocate_uninit(2 * dout * n, base)
    tv = (CsdView * max(1, len(tl)))(*[t[0].data.view() for t in tl])
    av = (CsdView * max(1, len(al)))(*[a.data.view() for a in al])
    tdig = (ctypes.c_int * max(1, len(tl)))(*[t[0].digits() for t in tl])
    tsd = (ctypes.c_int * max(1, len(tl)))(*[len(t[1]) for t in tl])
    adig = (ctypes.c_int * max(1, len(al)))(*[a.digits() for a in al])
    mods = ring.mods_for(base)
    marr = (type(mods[0]) * len(mods))(*mods)
    one = ring.scalar_tensor(base, 1)
    check(lib().csd_dbfv_combine(be.ctx, ctypes.byref(out.view()), ctypes.c_size_t(n), ctypes.c_size_t(dout), tv, tdig,
                                 ctypes.byref(sd.view()) if sd is not None else None, tsd, ctypes.c_size_t(len(tl)),
                                 av, adig, ctypes.c_size_t(len(al)), ctypes.byref(one.view()),
                                 ctypes.byref(cst.view()) if cst is not None else None, ctypes.c_uint32(full), marr,
                                 ctypes.c_size_t(len(mods))))
    return DBFVBatch(_checked(ring, out, base, 1, "dbfv_combine"), base, n, basis, 1)
