"""ORACLE — test infrastructure only (imported by tests/ and by bench.py's cpu_baseline leg).

A restatement of the dBFV digit extraction of cyclosubfield-dbfv/src/bootstrap/digitex.rs over the
oracle ring and the oracle's dBFV layer (``oracle.dbfv``), one ciphertext at a time as the
reference runs it:

* digitex.rs:18-265  — ``hom_eval_digit_extract_circuit_dbfv``: the straight-line program of
  squares, products and scalar inner products whose result f(x) satisfies f(x) = lift(x mod 2)
  mod 2^e, with mod-switches to the per-level base lengths and the subtractions of each level's
  value from the pending accumulators;
* digitex.rs:267-367 — ``hom_floordiv_dbfv``: floor(x / 2^v) by extracting and subtracting the
  lowest bit v times, with exact divisions by powers of two in between and the base lengths
  derived from the field's noise-growth estimates.

Pins: the composition is checked under the zero secret key (tests/test_oracle_digitex.py): with
relinearization key rows (0, g_i), g_i the CRT recomposition factor of gadget digit i, every
homomorphic operation acts on the c0 rows alone as on Delta-scaled plaintexts, so decrypting with
s = 0 (round(t c0 / Q) per digit, recombined as dbfv_dec does) must give floor(m / 2^v) mod t' for
the encrypted m — the reference's own floordiv tests (digitex.rs:372-482) run with the values and
moduli used there. A planted wrong circuit constant or a dropped subtraction fails that check.
"""
from . import dbfv as D

RNS_FACTOR_SIZE_LOG2 = 57.0  # bootstrap/mod.rs:31


def _modswitch_to(ring, x, rns_base_lens, j):
    """digitex.rs:35-44"""
    new_len = rns_base_lens[j]
    if new_len < len(x.base):
        return D.mod_switch_ct(ring, x, new_len)
    assert new_len == len(x.base)
    return x


def _sub_from(ring, target, value):
    """digitex.rs:45-53"""
    if len(target.rns_base()) < len(value.base):
        assert target.rns_base() == value.base[:len(target.rns_base())]
        D.dbfv_acc_sub(ring, target, D.mod_switch_ct(ring, value, len(target.rns_base())))
    else:
        assert target.rns_base() == value.base
        D.dbfv_acc_sub(ring, target, value)


def _add_to(ring, target, value):
    """digitex.rs:316-324"""
    if len(target.rns_base()) < len(value.base):
        assert target.rns_base() == value.base[:len(target.rns_base())]
        D.dbfv_acc_add(ring, target, D.mod_switch_ct(ring, value, len(target.rns_base())))
    else:
        assert target.rns_base() == value.base
        D.dbfv_acc_add(ring, target, value)


class _Ops:
    """The homomorphic products of the circuit (period-basis output, as digitex.rs passes
    Basis::PeriodBasis) over the relinearization key `keys`/`digits`."""

    def __init__(self, ring, keys, digits):
        self.ring, self.keys, self.digits = ring, keys, digits

    def square(self, modulus, x):
        out, bound = D.dbfv_hom_square(self.ring, modulus, x.data, x.bound, x.base, self.keys, self.digits, "pb")
        return D.Ct(out, x.base, "pb", bound)

    def mul(self, modulus, x, y):
        assert x.base == y.base
        out, bound = D.dbfv_hom_mul(self.ring, modulus, x.data, x.bound, y.data, y.bound, x.base, self.keys,
                                    self.digits, "pb")
        return D.Ct(out, x.base, "pb", bound)


def _combine(ring, modulus, base, terms, adds=(), constant=None):
    """dbfv_allocate_accumulator + dbfv_acc_inner_product_scalar(terms) + dbfv_acc_add(adds) (+
    dbfv_acc_add_constant) + dbfv_finish_accumulator, the pattern of digitex.rs:62-241"""
    acc = D.dbfv_allocate_accumulator(ring, modulus, base, "pb")
    D.dbfv_acc_inner_product_scalar(ring, acc, terms)
    for a in adds:
        D.dbfv_acc_add(ring, acc, a)
    if constant is not None:
        D.dbfv_acc_add_constant(ring, acc, constant)
    return D.dbfv_finish_accumulator(ring, acc)


def hom_eval_digit_extract_circuit_dbfv(ring, modulus, ct, subtract_intermediate, subtract_final, mults, ops,
                                        rns_base_lens):
    """digitex.rs:18-265. `ct` is a period-basis Ct; the accumulators in subtract_intermediate
    (fewer than min(16, 2^mults)) and subtract_final receive the subtractions."""
    assert len(rns_base_lens) == mults + 1
    assert len(subtract_intermediate) < (1 << mults) and len(subtract_intermediate) < 16
    pending = list(subtract_intermediate)

    def take(n):
        out = pending[:n]
        del pending[:n]
        return out

    ms = lambda x, j: _modswitch_to(ring, x, rns_base_lens, j)  # noqa: E731
    x2 = ms(ops.square(modulus, ct), 1)
    for target in take(1):
        _sub_from(ring, target, x2)
    if mults <= 1:
        _sub_from(ring, subtract_final, x2)
        return

    x4 = ms(ops.square(modulus, x2), 2)
    for target in take(2):
        _sub_from(ring, target, x4)
    if mults <= 2:
        _sub_from(ring, subtract_final, x4)
        return

    x2 = ms(x2, 2)
    val = _combine(ring, modulus, x2.base, [(x2, 94), (x4, 121)])
    pre_x8 = ms(ops.square(modulus, val), 3)
    x2, x4 = ms(x2, 3), ms(x4, 3)

    if pending or mults <= 3:
        x8 = _combine(ring, modulus, x2.base, [(x2, 112)], adds=[pre_x8])
        for target in take(4):
            _sub_from(ring, target, x8)
        if mults <= 3:
            _sub_from(ring, subtract_final, x8)
            return

    lhs = _combine(ring, modulus, x2.base, [(x2, 12656), (x4, 226), (pre_x8, 113)])
    rhs = _combine(ring, modulus, x2.base, [(x2, 33720), (x4, 2), (pre_x8, 301)])
    pre_x16 = ms(ops.mul(modulus, lhs, rhs), 4)
    x2, x4, pre_x8 = ms(x2, 4), ms(x4, 4), ms(pre_x8, 4)

    if pending or mults <= 4:
        x16 = _combine(ring, modulus, x2.base, [(x2, 2535424), (x4, 528), (pre_x8, 22620)], adds=[pre_x16])
        for target in take(8):
            _sub_from(ring, target, x16)
        if mults <= 4:
            _sub_from(ring, subtract_final, x16)
            return

    lhs = _combine(ring, modulus, x2.base, [(x2, 163551), (x4, -64628), (pre_x8, 88914), (pre_x16, -168751)])
    rhs = _combine(ring, modulus, x2.base, [(x2, 142065), (x4, 426380), (pre_x8, 456690), (pre_x16, -290191)])
    pre_x23 = ms(ops.mul(modulus, lhs, rhs), 5)

    x2, x4, pre_x8, pre_x16 = ms(x2, 5), ms(x4, 5), ms(pre_x8, 5), ms(pre_x16, 5)
    val = _combine(ring, modulus, x2.base, [(x4, 460241), (pre_x8, -520176), (pre_x16, 272152)], adds=[pre_x23])
    s1 = ms(ops.square(modulus, val), 6)

    x2, x4, pre_x8, pre_x16, pre_x23 = ms(x2, 6), ms(x4, 6), ms(pre_x8, 6), ms(pre_x16, 6), ms(pre_x23, 6)
    val = _combine(ring, modulus, x2.base,
                   [(x4, -32106558060291480110540), (pre_x8, 64784965186867951517504),
                    (pre_x16, -70466562361570490438944), (pre_x23, 403089028126664267137012)], adds=[s1])
    s2 = ms(ops.mul(modulus, val, D.Ct(s1.data.copy(), s1.base, s1.basis, s1.bound)), 7)

    x2, x4, pre_x8, pre_x16 = ms(x2, 7), ms(x4, 7), ms(pre_x8, 7), ms(pre_x16, 7)
    pre_x23, s1 = ms(pre_x23, 7), ms(s1, 7)
    lhs = _combine(ring, modulus, x2.base,
                   [(x4, -306594109877109866124476), (pre_x8, 55340232221151557156928),
                    (pre_x16, 66482065641637241715808), (pre_x23, -201964177491009025911804),
                    (s1, 982289121925033623516)], adds=[s2], constant=152453116397172589318468)
    rhs = _combine(ring, modulus, x2.base,
                   [(x4, -541458055423539766375756), (pre_x8, -321268494787747983014080),
                    (pre_x16, 389373873907872951793376), (pre_x23, -168833825134626671122316),
                    (s1, 982289121925033623520)], adds=[s2])
    s3 = ms(ops.mul(modulus, lhs, rhs), 8)

    x2, x4, pre_x8, pre_x16 = ms(x2, 8), ms(x4, 8), ms(pre_x8, 8), ms(pre_x16, 8)
    pre_x23, s1 = ms(pre_x23, 8), ms(s1, 8)
    res = _combine(ring, modulus, x2.base,
                   [(x4, 138276793813981485937712), (pre_x8, -394317601587992658717952),
                    (pre_x16, 277439031009004544298112), (pre_x23, -62423781945432606733008),
                    (s1, 376442706312192718108336)], adds=[s3])
    _sub_from(ring, subtract_final, res)


def floordiv_schedule(field, initial_modulus, final_modulus, baseline_len):
    """the circuit depth and the per-level RNS base lengths of hom_floordiv_dbfv
    (digitex.rs:275-312): (v, mults, rns_base_lens[v][mults + 1])"""
    v = initial_modulus.log2_t - final_modulus.log2_t
    e = initial_modulus.log2_t
    mults = (e - 1).bit_length() if e <= 16 else 8
    assert final_modulus.log2_b == initial_modulus.log2_b
    if v > 16:
        raise NotImplementedError("the polynomial program only supports eliminating up to 16 digits")
    if initial_modulus.log2_t > 80:
        raise NotImplementedError("the polynomial program only supports precision up to 80 bits")
    tb = min(initial_modulus.b(), initial_modulus.t())
    noises = [[0.0] * (mults + 1) for _ in range(v)]
    for i in range(v):
        if i > 0:
            noises[i][0] = noises[i - 1][1]
        for j in range(1, mults + 1):
            noises[i][j] = (field.noise_growth_hom_mul_scalar_lower_estimate(tb)
                            + field.noise_growth_hom_mul_lower_estimate(tb) + noises[i][j - 1])
    lens = []
    for row in noises:
        out = []
        for n in row:
            drop = int(n // RNS_FACTOR_SIZE_LOG2)
            assert baseline_len - drop >= 0
            out.append(baseline_len - drop)
        lens.append(out)
    return v, mults, lens


def hom_floordiv_dbfv(ring, initial_modulus, final_modulus, ct, keys, digits):
    """digitex.rs:267-367: floor(x / 2^v) of the period-basis Ct `ct`, v = log2 t - log2 t';
    the result is a period-basis Ct w.r.t. final_modulus"""
    v, mults, lens = floordiv_schedule(ring.field, initial_modulus, final_modulus, len(ct.base))
    ops = _Ops(ring, keys, digits)
    result = D.dbfv_allocate_accumulator(ring, initial_modulus, ct.base[:lens[-1][-1]], "pb")
    _add_to(ring, result, ct)
    parts = [D.dbfv_allocate_accumulator(ring, initial_modulus, ct.base[:lens[i][0]], "pb") for i in range(v)]
    for part in parts:
        _add_to(ring, part, ct)
    for i in range(v):
        current = D.dbfv_finish_accumulator(ring, parts.pop(0))
        current = D.dbfv_exact_div_pow2(ring, initial_modulus, current, i)
        current_modulus = D.DBFVModulus.new_incomplete(initial_modulus.log2_b, initial_modulus.log2_t - i)
        hom_eval_digit_extract_circuit_dbfv(ring, current_modulus, current, parts, result, mults, ops, lens[i])
    out = D.dbfv_finish_accumulator(ring, result)
    return D.dbfv_exact_div_pow2(ring, initial_modulus, out, v)
