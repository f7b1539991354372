"""Homomorphic digit extraction and floor division for dBFV (cyclosubfield-dbfv/src/bootstrap/
digitex.rs) on the MI355X, for a batch of ciphertexts at once.

The reference evaluates the digit-extraction polynomial (a straight-line program of squares,
products and scalar inner products with f(x) = lift(x mod 2) mod 2^e) and the floor division built
on it one ciphertext at a time. Here every step covers `count` ciphertexts held digit-major
(``dbfv.DBFVBatch``): one reduction, lift, digit convolution and relinearization per product for
all of them (``dbfv_hom_mul_batch``, ``dbfv_hom_square_batch``), one inner product per output digit
for each accumulator, one mod-switch per level. Each ciphertext sees the reference's call sequence
and operand bounds, so its residues are those of its own hom_floordiv_dbfv."""
from . import dbfv as D

RNS_FACTOR_SIZE_LOG2 = 57.0  # bootstrap/mod.rs:31

# (log2_pb_normalized_pb_op_norm_estimate, log2_sk_pb_op_norm_estimate) of each field descriptor
# (field.rs:537-767); the noise model of hom_floordiv_dbfv reads them
NOISE_ESTIMATES = {
    "field_36864": (8., 6.), "field_32768": (6., 5.), "field_32768_ex": (6., 5.), "field_24576": (5., 4.),
    "field_65536": (6., 5.), "field_49152": (10., 8.), "field_49152_ex": (8., 6.), "field_192": (0., 0.),
    "field_32": (0., 0.), "field_18": (0., 0.), "field_16": (0., 0.), "field_2": (0., 0.),
}


def floordiv_schedule(field_name, initial_modulus, final_modulus, baseline_len):
    """digitex.rs:275-312: (v, mults, rns_base_lens[v][mults + 1]) of hom_floordiv_dbfv. The noise
    of level i after j products is n_ij = n_i(j-1) + log2 min(b, t) (the scalar product) +
    log2 min(b, t) + log2 |sk| + log2 |pb| (the product, field.rs:464-483), n_(i+1)0 = n_i1, and the
    base drops floor(n_ij / 57) primes."""
    import math

    v = initial_modulus.log2_t - final_modulus.log2_t
    e = initial_modulus.log2_t
    mults = (e - 1).bit_length() if e <= 16 else 8
    if final_modulus.log2_b != initial_modulus.log2_b:
        raise ValueError("initial and final modulus must share the digit base")
    if v > 16:
        raise NotImplementedError("the polynomial program only supports eliminating up to 16 digits")
    if initial_modulus.log2_t > 80:
        raise NotImplementedError("the polynomial program only supports precision up to 80 bits")
    pb_norm, sk_norm = NOISE_ESTIMATES[field_name]
    lt = math.log2(float(min(initial_modulus.b(), initial_modulus.t())))
    noises = [[0.0] * (mults + 1) for _ in range(v)]
    for i in range(v):
        if i > 0:
            noises[i][0] = noises[i - 1][1]
        for j in range(1, mults + 1):
            noises[i][j] = lt + (lt + sk_norm + pb_norm) + noises[i][j - 1]
    lens = [[baseline_len - int(n // RNS_FACTOR_SIZE_LOG2) for n in row] for row in noises]
    if any(x < 0 for row in lens for x in row):
        raise ValueError("the RNS base is too short for the floor division's noise growth")
    return v, mults, lens


class _Products:
    """the products of the circuit, period-basis output (digitex.rs passes Basis::PeriodBasis),
    relinearized with the key `keys` of gadget digits `digits`"""

    def __init__(self, ring, keys, digits, fused=True):
        self.ring, self.keys, self.digits, self.fused = ring, keys, digits, fused

    def square(self, modulus, x):
        out, bound = D.dbfv_hom_square_batch(self.ring, modulus, x.data, x.bound, x.count, x.base, self.keys,
                                             self.digits, "pb", self.fused)
        return D.DBFVBatch(out, x.base, x.count, "pb", bound)

    def mul(self, modulus, x, y):
        if x.base != y.base or x.count != y.count:
            raise ValueError("operands differ in base or batch size")
        out, bound = D.dbfv_hom_mul_batch(self.ring, modulus, x.data, x.bound, y.data, y.bound, x.count, x.base,
                                          self.keys, self.digits, "pb", self.fused)
        return D.DBFVBatch(out, x.base, x.count, "pb", bound)


def _modswitch_to(ring, x, lens, j):
    """digitex.rs:35-44"""
    if lens[j] < len(x.base):
        return D.dbfv_mod_switch_down(ring, x, lens[j])
    if lens[j] != len(x.base):
        raise ValueError("mod-switch to a longer base")
    return x


def _switched_for(ring, target, value):
    if len(target.rns_base()) < len(value.base):
        if target.rns_base() != value.base[:len(target.rns_base())]:
            raise ValueError("the accumulator's base must be a prefix of the value's")
        return D.dbfv_mod_switch_down(ring, value, len(target.rns_base()))
    return value


def _sub_from(ring, target, value):
    """digitex.rs:45-53"""
    D.dbfv_acc_sub(ring, target, _switched_for(ring, target, value))


def _add_to(ring, target, value):
    """digitex.rs:316-324"""
    D.dbfv_acc_add(ring, target, _switched_for(ring, target, value))


def _combine(ring, modulus, like, terms, adds=(), constant=None, fused=True):
    """allocate an accumulator, one scalar inner product, additions (and a constant), finish: the
    pattern of every combination in digitex.rs:62-241; `fused` runs it as one kernel
    (dbfv.dbfv_combine, bit-identical), otherwise as the accumulator's HAL calls"""
    if fused:
        return D.dbfv_combine(ring, modulus, like, terms, adds, constant)
    acc = D.dbfv_allocate_accumulator(ring, modulus, like.base, "pb", like.count)
    D.dbfv_acc_inner_product_scalar(ring, acc, terms)
    for a in adds:
        D.dbfv_acc_add(ring, acc, a)
    if constant is not None:
        D.dbfv_acc_add_constant(ring, acc, constant)
    return D.dbfv_finish_accumulator(ring, acc)


def hom_eval_digit_extract_circuit_dbfv(ring, modulus, ct, subtract_intermediate, subtract_final, mults, products,
                                        rns_base_lens):
    """hom_eval_digit_extract_circuit_dbfv (digitex.rs:18-265) of the batch `ct` (period basis):
    the level values are subtracted from the pending accumulators subtract_intermediate (fewer
    than min(16, 2^mults)) as they appear, and the result from subtract_final"""
    if len(rns_base_lens) != mults + 1:
        raise ValueError("need one base length per level")
    if not (len(subtract_intermediate) < (1 << mults) and len(subtract_intermediate) < 16):
        raise ValueError("too many pending accumulators for the circuit")
    pending = list(subtract_intermediate)

    def take(n):
        out = pending[:n]
        del pending[:n]
        return out

    def ms(x, j):
        return _modswitch_to(ring, x, rns_base_lens, j)

    x2 = ms(products.square(modulus, ct), 1)
    for target in take(1):
        _sub_from(ring, target, x2)
    if mults <= 1:
        _sub_from(ring, subtract_final, x2)
        return

    x4 = ms(products.square(modulus, x2), 2)
    for target in take(2):
        _sub_from(ring, target, x4)
    if mults <= 2:
        _sub_from(ring, subtract_final, x4)
        return

    x2 = ms(x2, 2)
    val = _combine(ring, modulus, x2, fused=products.fused, terms= [(x2, 94), (x4, 121)])
    pre_x8 = ms(products.square(modulus, val), 3)
    x2, x4 = ms(x2, 3), ms(x4, 3)

    if pending or mults <= 3:
        x8 = _combine(ring, modulus, x2, fused=products.fused, terms= [(x2, 112)], adds=[pre_x8])
        for target in take(4):
            _sub_from(ring, target, x8)
        if mults <= 3:
            _sub_from(ring, subtract_final, x8)
            return

    lhs = _combine(ring, modulus, x2, fused=products.fused, terms= [(x2, 12656), (x4, 226), (pre_x8, 113)])
    rhs = _combine(ring, modulus, x2, fused=products.fused, terms= [(x2, 33720), (x4, 2), (pre_x8, 301)])
    pre_x16 = ms(products.mul(modulus, lhs, rhs), 4)
    x2, x4, pre_x8 = ms(x2, 4), ms(x4, 4), ms(pre_x8, 4)

    if pending or mults <= 4:
        x16 = _combine(ring, modulus, x2, fused=products.fused, terms= [(x2, 2535424), (x4, 528), (pre_x8, 22620)], adds=[pre_x16])
        for target in take(8):
            _sub_from(ring, target, x16)
        if mults <= 4:
            _sub_from(ring, subtract_final, x16)
            return

    lhs = _combine(ring, modulus, x2, fused=products.fused, terms= [(x2, 163551), (x4, -64628), (pre_x8, 88914), (pre_x16, -168751)])
    rhs = _combine(ring, modulus, x2, fused=products.fused, terms= [(x2, 142065), (x4, 426380), (pre_x8, 456690), (pre_x16, -290191)])
    pre_x23 = ms(products.mul(modulus, lhs, rhs), 5)

    x2, x4, pre_x8, pre_x16 = ms(x2, 5), ms(x4, 5), ms(pre_x8, 5), ms(pre_x16, 5)
    val = _combine(ring, modulus, x2, fused=products.fused, terms= [(x4, 460241), (pre_x8, -520176), (pre_x16, 272152)], adds=[pre_x23])
    s1 = ms(products.square(modulus, val), 6)

    x2, x4, pre_x8, pre_x16, pre_x23 = ms(x2, 6), ms(x4, 6), ms(pre_x8, 6), ms(pre_x16, 6), ms(pre_x23, 6)
    val = _combine(ring, modulus, x2, fused=products.fused, terms=
                   [(x4, -32106558060291480110540), (pre_x8, 64784965186867951517504),
                    (pre_x16, -70466562361570490438944), (pre_x23, 403089028126664267137012)], adds=[s1])
    s2 = ms(products.mul(modulus, val, s1), 7)

    x2, x4, pre_x8, pre_x16 = ms(x2, 7), ms(x4, 7), ms(pre_x8, 7), ms(pre_x16, 7)
    pre_x23, s1 = ms(pre_x23, 7), ms(s1, 7)
    lhs = _combine(ring, modulus, x2, fused=products.fused, terms=
                   [(x4, -306594109877109866124476), (pre_x8, 55340232221151557156928),
                    (pre_x16, 66482065641637241715808), (pre_x23, -201964177491009025911804),
                    (s1, 982289121925033623516)], adds=[s2], constant=152453116397172589318468)
    rhs = _combine(ring, modulus, x2, fused=products.fused, terms=
                   [(x4, -541458055423539766375756), (pre_x8, -321268494787747983014080),
                    (pre_x16, 389373873907872951793376), (pre_x23, -168833825134626671122316),
                    (s1, 982289121925033623520)], adds=[s2])
    s3 = ms(products.mul(modulus, lhs, rhs), 8)

    x2, x4, pre_x8, pre_x16 = ms(x2, 8), ms(x4, 8), ms(pre_x8, 8), ms(pre_x16, 8)
    pre_x23, s1 = ms(pre_x23, 8), ms(s1, 8)
    res = _combine(ring, modulus, x2, fused=products.fused, terms=
                   [(x4, 138276793813981485937712), (pre_x8, -394317601587992658717952),
                    (pre_x16, 277439031009004544298112), (pre_x23, -62423781945432606733008),
                    (s1, 376442706312192718108336)], adds=[s3])
    _sub_from(ring, subtract_final, res)


def hom_floordiv_dbfv(ring, initial_modulus, final_modulus, ct, keys, digits, fused=True, levels=None):
    """hom_floordiv_dbfv (digitex.rs:267-367) of the batch `ct` (period basis): floor(x / 2^v),
    v = log2 t - log2 t', as a period-basis batch w.r.t. final_modulus. keys[i] is the [2, r_key,
    ., .] CRT-basis relinearization key of gadget digit i. `levels` (default: all v) stops after
    that many extraction levels and returns the result accumulator unfinished, for timing one
    level of the circuit."""
    v, mults, lens = floordiv_schedule(ring.field_name, initial_modulus, final_modulus, len(ct.base))
    products = _Products(ring, keys, digits, fused)
    n = ct.count
    result = D.dbfv_allocate_accumulator(ring, initial_modulus, ct.base[:lens[-1][-1]], "pb", n)
    _add_to(ring, result, ct)
    parts = [D.dbfv_allocate_accumulator(ring, initial_modulus, ct.base[:lens[i][0]], "pb", n) for i in range(v)]
    for part in parts:
        _add_to(ring, part, ct)
    for i in range(v if levels is None else min(levels, v)):
        current = D.dbfv_finish_accumulator(ring, parts.pop(0))
        current = D.dbfv_exact_div_pow2(ring, initial_modulus, current, i)
        current_modulus = D.DBFVModulus.new_incomplete(initial_modulus.log2_b, initial_modulus.log2_t - i)
        hom_eval_digit_extract_circuit_dbfv(ring, current_modulus, current, parts, result, mults, products, lens[i])
    if levels is not None and levels < v:
        return result
    out = D.dbfv_finish_accumulator(ring, result)
    return D.dbfv_exact_div_pow2(ring, initial_modulus, out, v)
