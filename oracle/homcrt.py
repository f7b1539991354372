"""ORACLE — test infrastructure only (imported by tests/ and by bench.py's cpu_baseline leg).

A restatement of the homomorphic evaluation half of cyclosubfield-dbfv/src/bootstrap/homcrt.rs over
the oracle ring (``oracle.ring``) and the oracle's dBFV layer (``oracle.dbfv``), one ciphertext at
a time as the reference runs it:

* ring.rs:1227-1250  — ``lprime_element_tensor``: an element of L' (period-basis coefficients
  w.r.t. the basis induced by M) as a [1, r, deg_L', 1] CRT-basis factor tensor;
* dbfv.rs:193-259    — ``dbfv_acc_fma_element_lprime`` over the fma skeleton;
* dbfv.rs:378-396    — ``dbfv_acc_inner_product_element_lprime``;
* dbfv.rs:431-450    — ``dbfv_hom_fma_element_lprime``;
* bfv.rs:783-839     — ``bfv_hom_gal``: hoisted Galois automorphisms of one BFV ciphertext;
* homcrt.rs:188-211  — ``apply_cyclic_transform_plain`` (the plaintext transform);
* homcrt.rs:320-349  — ``bsgs_split``, ``factor_galois_elements``;
* homcrt.rs:355-429  — ``hom_cyclic_transform_dbfv``: x -> sum_i coeffs[i] sigma^i(x) on a dBFV
  ciphertext with baby steps sharing the gadget decomposition and one unhoisted giant step per
  block.

Pins (tests/test_oracle_homcrt.py): the L'-element tensor equals the CRT transform of the full
ring element whose period-basis coefficients are a_u at every K position (the transforms are
pinned by the reference's KATs); under the zero secret key (Galois key rows (0, g_i), g_i the CRT
recomposition factor of gadget digit i) bfv_hom_gal is sigma_g on both components mod Q, so the
transform of an encryption of m must decrypt to sum_i coeffs[i] sigma^i(m) mod t, formed
independently by the plaintext transform of the ring's own products; planted mistakes (a wrong
giant-step exponent, a shifted coefficient block, a dropped carry-over of the giant steps) fail.
The BSGS split is checked against the reference's cost rule by brute force.
"""
from . import dbfv as D
from .ring import gadget_decompose, hom_gal_one, clip_digits, mont_repr, _trunc_mod

UNHOISTED_OVERHEAD = 3  # homcrt.rs:318


def homcrt_factors(field):
    """field.rs:217-243 for the primes p | m_K (m_k_factors, field.rs:271): (p, i_p_gen, |J_p|) with
    i_p_gen the smallest generator of (Z/pZ)* embedded as 1 mod m/p and |J_p| = (p - 1) / |H_p|,
    H_p the powers of i_p_gen inside Gal(M|L)"""
    from .ring import crt2

    m = field.m
    ml = set(field.gal_ml)
    out = []
    for p in field.m_factors:
        if field.m_prime != 1 and p == field.m_prime:
            continue
        pf, rest, f = [], p - 1, 2
        while rest > 1:
            if rest % f == 0:
                pf.append(f)
                while rest % f == 0:
                    rest //= f
            f += 1
        gen = next(g for g in range(2, p) if all(pow(g, (p - 1) // q, p) != 1 for q in pf))
        ipg = crt2(gen, p, 1, m // p)
        hp = sum(1 for e in range(p - 1) if pow(ipg, e, m) in ml)
        out.append((p, ipg, (p - 1) // hp))
    return out


def bsgs_split(length):
    """homcrt.rs:320-324: the first baby-step count minimising count + 3 ceil((len - 1) / count)"""
    best, best_cost = None, None
    for count in range(1, length + 1):
        cost = count + UNHOISTED_OVERHEAD * ((length - 1) // count + 1)
        if best_cost is None or cost < best_cost:
            best, best_cost = count, cost
    return best, (length - 1) // best + 1


def pow_mod_m(g, e, m):
    """homcrt.rs:326-332"""
    result = 1
    for _ in range(e):
        result = result * g % m
    return result


def factor_galois_elements(m, gen, length):
    """homcrt.rs:334-346: the baby-step exponents and the giant-step exponent of one factor"""
    if length <= 1:
        return []
    baby, giant = bsgs_split(length)
    result = [pow_mod_m(gen, j, m) for j in range(1, baby)]
    if giant > 1:
        result.append(pow_mod_m(gen, baby, m))
    return result


def lprime_element_tensor(ring, base, coeffs):
    """ring.rs:1227-1250: mont((c mod p) as i64) per limb and coefficient, then the L' part of the
    CRT transform along dim 2 (ntt_mul with the L' kernel, intt)"""
    f = ring.field
    assert len(coeffs) == f.deg_lprime
    values = []
    for i in base:
        p = ring.primes[i]
        for c in coeffs:
            values.append(mont_repr(_trunc_mod(c, p), p))
    t = ring._import(values, [1, len(base), f.deg_lprime, 1])
    cache = ring.transform_cache(base)
    ring.backend.hal_ntt_mul_mont64_dim(t, cache["ntts_lprime"], 2, cache["crt_kernel_lprime"], cache["mods"])
    ring.backend.hal_intt_mont64_dim(t, cache["ntts_lprime"], 2)
    return t


def balanced_digits_vec(xs, b, max_digits):
    """dbfv.rs:97-112: balanced base-b digit vectors, at most max_digits, until all are zero"""
    out, cur = [], list(xs)
    for _ in range(max_digits):
        if all(x == 0 for x in cur):
            break
        rems = [D.cmod(x, b) for x in cur]
        cur = [(x - y) // b for x, y in zip(cur, rems)]
        out.append(rems)
    return out


def dbfv_empty(ring, base, basis):
    """dbfv.rs:47-49"""
    return D.Ct(ring.allocate(0, base), base, basis, 1)


def dbfv_acc_fma_tensors(ring, acc, lhs, rhs_tensors):
    """dbfv.rs:193-224"""
    ld = lhs.digits()
    if ld == 0 or not rhs_tensors:
        return
    assert ld <= acc.capacity() and acc.basis == lhs.basis and acc.rns_base() == lhs.base
    new_digits = min(acc.capacity(), ld + len(rhs_tensors) - 1)
    for i in range(ld):
        for j in range(len(rhs_tensors)):
            k = i + j
            if k < new_digits:
                acc.acc.fma(range(2 * k, 2 * k + 2), lhs.data[2 * i:2 * i + 2], lhs.bound, rhs_tensors[j], 1)
    acc.used_digits = max(acc.used_digits, new_digits)


def _lprime_digit_tensors(ring, acc, coeffs):
    m = acc.modulus
    return [lprime_element_tensor(ring, acc.rns_base(), ds) for ds in balanced_digits_vec(coeffs, m.b(), m.digits())]


def dbfv_acc_fma_element_lprime(ring, acc, lhs, rhs):
    """dbfv.rs:244-259"""
    assert acc.basis == "crt"
    dbfv_acc_fma_tensors(ring, acc, lhs, _lprime_digit_tensors(ring, acc, rhs))


def dbfv_acc_inner_product_element_lprime(ring, acc, terms):
    """dbfv.rs:378-396: terms are (lhs Ct in CRT basis, L' coefficients)"""
    assert acc.basis == "crt"
    D.dbfv_acc_inner_product_tensors(ring, acc, [(lhs, _lprime_digit_tensors(ring, acc, c)) for lhs, c in terms])


def dbfv_hom_fma_element_lprime(ring, modulus, lhs, rhs, summand):
    """dbfv.rs:431-450"""
    assert lhs.basis == "crt" and summand.basis == "crt"
    assert lhs.digits() <= modulus.digits() and summand.digits() <= modulus.digits()
    if lhs.digits() == 0 or all(x == 0 for x in rhs):
        return summand
    acc = D.dbfv_allocate_accumulator(ring, modulus, lhs.base, "crt")
    D.dbfv_acc_add(ring, acc, summand)
    dbfv_acc_fma_element_lprime(ring, acc, lhs, rhs)
    return D.dbfv_finish_accumulator(ring, acc)


def bfv_hom_gal(ring, ct, base, gs, gks):
    """bfv.rs:783-839: one CRT-basis BFV ciphertext ([2, r, ., .]) under the rotations gs with the
    gadget decomposition of c1 shared; gks maps g to (keys, digits) with keys[i] the [2, r_key, .,
    .] key of digit i over a base of which `base` is a prefix"""
    import numpy as np

    if not gs:
        return []
    r = len(base)
    c1 = ct[1:2].copy()
    ring.crt_to_pb_tensor(c1, base)
    digits = clip_digits(gks[gs[0]][1], r)
    parts = gadget_decompose(ring, c1, base, digits)
    out = []
    for g in gs:
        keys, kd = gks[g]
        assert clip_digits(kd, r) == digits, "all galois keys must use the same digits"
        ks = [np.ascontiguousarray(k[:, :r]) for k in keys[:len(digits)]]
        out.append(hom_gal_one(ring, ct, base, g, parts, ks, digits))
    return out


def apply_cyclic_transform_plain(ring, data, base, automorphism, coeffs):
    """homcrt.rs:188-211: sum_i coeffs[i] sigma^i(x) of CRT-basis elements (bound 1)"""
    count = data.shape[0]
    acc = ring.allocate_wide(count, base)
    current = data.copy()
    for i, c in enumerate(coeffs):
        acc.fma(range(count), current, 1, lprime_element_tensor(ring, base, c), 1)
        if i + 1 < len(coeffs):
            nxt = ring.allocate(count, base)
            ring.apply_galois(nxt, current, automorphism)
            current = nxt
    return ring.reduce_wide(acc)


def hom_cyclic_transform_dbfv(ring, modulus, automorphism, coeffs, ct, gks):
    """homcrt.rs:355-429 for a CRT-basis dBFV ciphertext `ct` (oracle.dbfv.Ct)"""
    length = len(coeffs)
    base = ct.base
    if length == 0:
        return dbfv_empty(ring, base, "crt")
    if length == 1:
        return dbfv_hom_fma_element_lprime(ring, modulus, ct, coeffs[0], dbfv_empty(ring, base, "crt"))
    baby, giant = bsgs_split(length)
    m = ring.field.m
    d = ct.digits()
    baby_exps = [pow_mod_m(automorphism, j, m) for j in range(1, baby)]
    babys = [ct] + [D.Ct(ring.allocate(2 * d, base), base, "crt", 1) for _ in range(1, baby)]
    for i in range(d):
        rotated = bfv_hom_gal(ring, ct.data[2 * i:2 * i + 2], base, baby_exps, gks)
        for j, rot in enumerate(rotated):
            babys[j + 1].data[2 * i:2 * i + 2] = rot
    giant_exp = pow_mod_m(automorphism, baby, m)
    out = dbfv_empty(ring, base, "crt")
    for gi in reversed(range(giant)):
        if gi != giant - 1:
            od = out.digits()
            rotated = ring.allocate(2 * od, base)
            for i in range(od):
                [rot] = bfv_hom_gal(ring, out.data[2 * i:2 * i + 2], base, [giant_exp], gks)
                rotated[2 * i:2 * i + 2] = rot
            out = D.Ct(rotated, base, "crt", 1)
        acc = D.dbfv_allocate_accumulator(ring, modulus, base, "crt")
        terms = [(babys[j], coeffs[gi * baby + j]) for j in range(min(baby, length - gi * baby))]
        dbfv_acc_inner_product_element_lprime(ring, acc, terms)
        D.dbfv_acc_add(ring, acc, out)
        out = D.dbfv_finish_accumulator(ring, acc)
    return out
