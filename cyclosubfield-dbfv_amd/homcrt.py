"""The homomorphic CRT transform of bootstrapping (cyclosubfield-dbfv/src/bootstrap/homcrt.rs:
355-449) on the MI355X, batched over dBFV ciphertexts.

Each factor u_p of the transform is the J_p-cyclic map x -> sum_i coeffs[i] sigma^i(x) with
coefficients in L'. ``hom_cyclic_transform_dbfv`` evaluates it with the reference's
baby-step/giant-step split (homcrt.rs:320-324): the baby steps sigma^j(ct) share the gadget
decomposition of every digit (bfv_hom_gal, bfv.rs:783-839), and each giant step applies one
unhoisted automorphism to the carried result before one inner product per output digit over the
baby steps (dbfv_acc_inner_product_element_lprime, dbfv.rs:378-396).

The reference runs one ciphertext and, inside it, one bfv_hom_gal call per dBFV digit. Here one
call covers every digit of every ciphertext of a ``dbfv.DBFVBatch``: the c1 rows of all digits go
through one gadget conversion, and each rotation is one key-switch kernel
(``RingContext.galois_keyswitch_rows``) that reads the Galois key once per group of eight
(digit, ciphertext) pairs and writes its output straight into the digit-major rows of the rotated
batch. Each (digit, ciphertext) pair sees the reference's operands and reductions, so the residues
are those of the reference's per-ciphertext sequence."""
from . import dbfv as D
from .bfv import clip_digits
from .ring import homcrt_factors

UNHOISTED_OVERHEAD = 3  # homcrt.rs:318


def bsgs_split(length):
    """bsgs_split (homcrt.rs:320-324): the first baby-step count minimising
    count + 3 ((length - 1) / count + 1), and the giant-step count"""
    best, best_cost = 1, None
    for count in range(1, length + 1):
        cost = count + UNHOISTED_OVERHEAD * ((length - 1) // count + 1)
        if best_cost is None or cost < best_cost:
            best, best_cost = count, cost
    return best, (length - 1) // best + 1


def pow_mod_m(g, e, m):
    """pow_mod_m (homcrt.rs:326-332)"""
    return pow(g, e, m)


def factor_galois_elements(m, gen, length):
    """factor_galois_elements (homcrt.rs:334-346): the baby-step powers of the generator, then the
    giant step when there is more than one block"""
    if length <= 1:
        return []
    baby, giant = bsgs_split(length)
    res = [pow_mod_m(gen, j, m) for j in range(1, baby)]
    if giant > 1:
        res.append(pow_mod_m(gen, baby, m))
    return res


def required_galois_elements(field_name):
    """required_galois_elements (homcrt.rs:336-349) over the field's factors, sorted"""
    from .ring import field_info

    m = field_info(field_name)["m"]
    return sorted({g for _, gen, n in homcrt_factors(field_name) for g in factor_galois_elements(m, gen, n)})


class GaloisKey:
    """A KeySwitchKey for one Galois element (bfv.rs:64-103) on the device: digit i is rows 2i, 2i+1
    of one [2 d, r_key, deg_L', deg_K] CRT-basis tensor, and `digits` the gadget digit ranges."""

    def __init__(self, data, digits):
        if data.shape[0] != 2 * len(digits):
            raise ValueError("a key has two rows per gadget digit")
        self.data, self.digits = data, list(digits)


def balanced_digits_vec(xs, b, max_digits):
    """balanced_digits_vec (dbfv.rs:97-112): balanced base-b digit vectors of the coefficients, at
    most max_digits, stopping once every coefficient is zero"""
    out, cur = [], [int(x) for x in xs]
    for _ in range(max_digits):
        if all(x == 0 for x in cur):
            break
        rems = [D.cmod(x, b) for x in cur]
        cur = [(x - y) // b for x, y in zip(cur, rems)]
        out.append(rems)
    return out


def dbfv_empty(ring, like):
    """dbfv_empty (dbfv.rs:47-49): a batch with no digits, over the base and batch size of `like`"""
    return D.DBFVBatch(like.data[0:0], like.base, like.count, "crt", 1)


def _lprime_digit_tensors(ring, acc, coeffs):
    m = acc.modulus
    return [ring.lprime_element_tensor(acc.rns_base(), ds) for ds in balanced_digits_vec(coeffs, m.b(), m.digits())]


def dbfv_acc_fma_element_lprime(ring, acc, lhs, rhs):
    """dbfv_acc_fma_element_lprime (dbfv.rs:244-259) over the fma skeleton (dbfv.rs:193-224), for
    every ciphertext of the batch"""
    if acc.basis != "crt":
        raise ValueError("multiplication by an L'-element requires CRT basis")
    tensors = _lprime_digit_tensors(ring, acc, rhs)
    ld = lhs.digits()
    if ld == 0 or not tensors:
        return
    D._check_operand(acc, lhs)
    new_digits = min(acc.capacity(), ld + len(tensors) - 1)
    for i in range(ld):
        for j, t in enumerate(tensors):
            k = i + j
            if k < new_digits:
                acc.acc.fma(acc.rows(k, k + 1), lhs.digit_rows(i), lhs.bound, t, 1)
    acc.used_digits = max(acc.used_digits, new_digits)


def dbfv_acc_inner_product_element_lprime(ring, acc, terms):
    """dbfv_acc_inner_product_element_lprime (dbfv.rs:378-396): terms (DBFVBatch in CRT basis, L'
    coefficients); one inner product per output digit over all terms; the rows written must be
    fresh"""
    if acc.basis != "crt":
        raise ValueError("multiplication by an L'-element requires CRT basis")
    D.dbfv_acc_inner_product_tensors(ring, acc, [(lhs, _lprime_digit_tensors(ring, acc, c)) for lhs, c in terms])


def dbfv_hom_fma_element_lprime(ring, modulus, lhs, rhs, summand):
    """dbfv_hom_fma_element_lprime (dbfv.rs:431-450): summand + lhs rhs, for CRT-basis batches"""
    if lhs.basis != "crt" or summand.basis != "crt":
        raise ValueError("dbfv_hom_fma_element_lprime needs CRT-basis ciphertexts")
    if lhs.digits() > modulus.digits() or summand.digits() > modulus.digits():
        raise ValueError("ciphertext has more digits than the modulus")
    if lhs.digits() == 0 or all(int(x) == 0 for x in rhs):
        return summand
    acc = D.dbfv_allocate_accumulator(ring, modulus, lhs.base, "crt", lhs.count)
    D.dbfv_acc_add(ring, acc, summand)
    dbfv_acc_fma_element_lprime(ring, acc, lhs, rhs)
    return D.dbfv_finish_accumulator(ring, acc)


def bfv_hom_gal_batch(ring, ct, gs, gks):
    """bfv_hom_gal (bfv.rs:783-839) of every digit of every ciphertext of the CRT-basis batch `ct`
    under the rotations gs, with the gadget decomposition of each c1 shared by the rotations.
    gks maps each g to its GaloisKey (over a base of which ct's is a prefix, bfv.rs:825).
    Returns one DBFVBatch per rotation (bound 1), digit-major like `ct`."""
    if ct.basis != "crt":
        raise ValueError("bfv_hom_gal needs CRT-basis ciphertexts")
    if not gs or ct.digits() == 0:
        return [dbfv_empty(ring, ct) for _ in gs]
    if ct.bound > 2:
        raise ValueError("basis transforms require representatives bounded by 2p")
    be = ring.backend
    base, n, d = ct.base, ct.count, ct.digits()
    r = len(base)
    try:
        first = gks[gs[0]]
    except KeyError:
        raise KeyError("required galois key not present") from None
    digits = clip_digits(first.digits, r)
    for g in gs:
        if g not in gks:
            raise KeyError("required galois key not present")
        if clip_digits(gks[g].digits, r) != digits:
            raise ValueError("all galois keys must use the same digits")
    dg = len(digits)
    # the c1 rows of all digits (row i n + c), to period basis, then one gadget conversion
    c1 = ring.allocate_uninit(d * n, base)
    for i in range(d):
        be.hal_assign(c1[i * n:(i + 1) * n], ct.data[(2 * i + 1) * n:(2 * i + 2) * n])
    ring.crt_to_pb_tensor(c1, base)
    parts = ring.allocate_uninit(d * n * dg, base)  # element e = i n + c, gadget digit k at row e dg + k
    ring.bconv_gadget_tensors(parts, c1, base, digits)
    ring.pb_to_crt_tensor(parts, base)
    outs = [ring.allocate_uninit(2 * d * n, base) for _ in gs]
    keys = [gks[g].data[0:2 * dg, 0:r] for g in gs]
    # element e = i n + c: c0 at row 2 i n + c, output row s at (2 i + s) n + c
    ring.galois_keyswitch_rows(base, d * n, n, dg, parts, ct.data, (2 * n, 1), gs, keys, outs, (2 * n, 1, n))
    return [D.DBFVBatch(o, base, n, "crt", 1) for o in outs]


def hom_cyclic_transform_dbfv(ring, modulus, automorphism, coeffs, ct, gks):
    """hom_cyclic_transform_dbfv (homcrt.rs:355-429): sum_i coeffs[i] sigma^i(ct) for a CRT-basis
    batch `ct`, coeffs a list of L' coefficient vectors, automorphism the generator sigma"""
    length = len(coeffs)
    if ct.basis != "crt":
        raise ValueError("the cyclic transform needs CRT-basis ciphertexts")
    if length == 0:
        return dbfv_empty(ring, ct)
    if length == 1:
        return dbfv_hom_fma_element_lprime(ring, modulus, ct, coeffs[0], dbfv_empty(ring, ct))
    baby, giant = bsgs_split(length)
    m = ring.info["m"]
    baby_exps = [pow_mod_m(automorphism, j, m) for j in range(1, baby)]
    babys = [ct] + bfv_hom_gal_batch(ring, ct, baby_exps, gks)
    giant_exp = pow_mod_m(automorphism, baby, m)
    out = dbfv_empty(ring, ct)
    for gi in reversed(range(giant)):
        if gi != giant - 1 and out.digits() > 0:
            [out] = bfv_hom_gal_batch(ring, out, [giant_exp], gks)
        acc = D.dbfv_allocate_accumulator(ring, modulus, ct.base, "crt", ct.count)
        terms = [(babys[j], coeffs[gi * baby + j]) for j in range(min(baby, length - gi * baby))]
        dbfv_acc_inner_product_element_lprime(ring, acc, terms)
        D.dbfv_acc_add(ring, acc, out)
        out = D.dbfv_finish_accumulator(ring, acc)
    return out


def slots_to_coeffs_dbfv(ring, factors, modulus, ct, gks):
    """slots_to_coeffs_dbfv (homcrt.rs:434-449): the cyclic transform of every factor in turn;
    factors is a list of (generator, coefficient list), the u_p of the transform (or their
    inverses for coeffs_to_slots_dbfv, homcrt.rs:452-467)"""
    for gen, coeffs in factors:
        ct = hom_cyclic_transform_dbfv(ring, modulus, gen, coeffs, ct, gks)
    return ct


coeffs_to_slots_dbfv = slots_to_coeffs_dbfv
