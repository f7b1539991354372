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


balanced_digits_vec = D.balanced_digits_vec


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
    if _dense(ct):
        # rows (2 i + 1) n + c of ct to rows i n + c, as one strided copy
        be.hal_assign(D.digit_rows_view(c1, 0, d, n, n), D.digit_rows_view(ct.data, n, d, n, 2 * n))
    else:
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


def _dense(batch):
    """whether a batch's rows are one contiguous [rows, r, n2, n3] block (the combination kernel's
    operand layout)"""
    t = batch.data
    plane = t.shape[2] * t.shape[3]
    return t.strides[3] == 1 and t.strides[2] == t.shape[3] and t.strides[1] == plane and t.strides[0] == t.shape[1] * plane


def hom_cyclic_transform_dbfv(ring, modulus, automorphism, coeffs, ct, gks, fused=True):
    """hom_cyclic_transform_dbfv (homcrt.rs:355-429): sum_i coeffs[i] sigma^i(ct) for a CRT-basis
    batch `ct`, coeffs a list of L' coefficient vectors, automorphism the generator sigma. With
    `fused` each giant step's accumulator (the baby-step inner product, the carried sum and the
    reduction) is one kernel (dbfv.dbfv_combine with L' factors); the residues are the same."""
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
        terms = [(babys[j], coeffs[gi * baby + j]) for j in range(min(baby, length - gi * baby))]
        if fused and len(terms) <= D.COMBINE_MAX_TERMS and all(_dense(x) for x, _ in terms) and _dense(out):
            out = D.dbfv_combine(ring, modulus, ct, terms, adds=[out], basis="crt", lprime=True)
            continue
        acc = D.dbfv_allocate_accumulator(ring, modulus, ct.base, "crt", ct.count)
        dbfv_acc_inner_product_element_lprime(ring, acc, terms)
        D.dbfv_acc_add(ring, acc, out)
        out = D.dbfv_finish_accumulator(ring, acc)
    return out


def slots_to_coeffs_dbfv(ring, factors, modulus, ct, gks, fused=True):
    """slots_to_coeffs_dbfv (homcrt.rs:434-449): the cyclic transform of every factor in turn;
    factors is a list of (generator, coefficient list), the u_p of the transform (or their
    inverses for coeffs_to_slots_dbfv, homcrt.rs:452-467)"""
    for gen, coeffs in factors:
        ct = hom_cyclic_transform_dbfv(ring, modulus, gen, coeffs, ct, gks, fused)
    return ct


coeffs_to_slots_dbfv = slots_to_coeffs_dbfv


# ---------------------------------------------------------------------------------------------
# the transform's coefficients (host set-up, homcrt.rs:29-176)

class CRTTransformFactor:
    """CRTTransformFactor::new (homcrt.rs:70-152): the coefficients u_p (slots to coefficients) and
    inv_u_p of the J_p-cyclic factor for the prime p | m_K, as elements of L' modulo t (period-basis
    coefficients w.r.t. the basis induced by M, balanced), found as Gauss sums in GR(2^e, D),
    D = lcm(f_p, ord_m'(2)), and pulled back to L' by solving against the period basis of L' under
    every embedding of L' (one block of D rows per prime of L' above 2)."""

    def __init__(self, t, p, field, u_p=None, inv_u_p=None, gen_jp=None):
        if u_p is not None:  # reduce_t
            self.p, self.t, self.u_p, self.inv_u_p, self.gen_jp = p, t, u_p, inv_u_p, gen_jp
            return
        from .galoisring import GaloisRing, solve_mod_2e
        from .field import inv_mod

        if t < 2 or t & (t - 1):
            raise ValueError("t must be a power of two")
        e = (t - 1).bit_length()
        m = field.m
        info = field.prime_info(p)
        n2 = field.deg_lprime
        fo = field.mprime_frobenius_order()
        import math

        d = info.f_p * fo // math.gcd(info.f_p, fo)
        gr = GaloisRing(e, d)
        rou_mp = gr.root_of_unity(field.m_prime)
        scale = field.scale_mprime_to_m_period_basis()
        embeddings = field.lprime_embedding_transversal()
        rows = len(embeddings) * d
        matrix = [0] * (rows * n2)
        mp_pows = [gr.one()]
        for _ in range(1, max(field.m_prime, 1)):
            mp_pows.append(gr.mul(mp_pows[-1], rou_mp))
        for block, gamma in enumerate(embeddings):
            for j, g in enumerate(field.lprime_q_transversal()):
                period = gr.zero()
                for h in field.gal_mprime_lprime:
                    conj = gamma * g % m
                    power = conj * h % m % field.m_prime
                    period = gr.add(period, mp_pows[power])
                period = gr.mul_scalar(period, scale)
                for i in range(d):
                    matrix[(block * d + i) * n2 + j] = period[i]
        rou_p = gr.root_of_unity(p)
        p_pows = [gr.one()]
        for _ in range(1, p):
            p_pows.append(gr.mul(p_pows[-1], rou_p))
        gen = info.i_p_gen
        gen_inv = inv_mod(gen, m)
        jp = info.j_p_order()
        hp = len(info.h_p)

        def orbit_sum(base, i, conjugate):
            pg = pow(base, i, m)
            step = m - 1 if conjugate else 1
            s = gr.zero()
            for h in info.h_p:
                el = h * pg % m * step % m
                s = gr.add(s, p_pows[el % p])
            return s

        p_inv = gr.inv(gr.from_scalar(p))
        rhs = [0] * (rows * 2 * jp)

        def set_rhs(col, coeff):
            for block in range(len(embeddings)):
                for row in range(d):
                    rhs[(block * d + row) * 2 * jp + col] = coeff[row]

        for i in range(jp):
            set_rhs(i, orbit_sum(gen_inv, i, False))
            c = gr.sub(orbit_sum(gen, i, True), gr.from_scalar(hp))
            set_rhs(jp + i, gr.mul(c, p_inv))
        sol = solve_mod_2e(e, rows, n2, 2 * jp, matrix, rhs)

        def extract(col):
            return [gr.smallest_lift(sol[c * 2 * jp + col]) for c in range(n2)]

        self.p, self.t, self.gen_jp = p, t, gen
        self.u_p = [extract(i) for i in range(jp)]
        self.inv_u_p = [extract(jp + i) for i in range(jp)]

    def reduce_t(self, new_t):
        """homcrt.rs:164-176"""
        if self.t % new_t:
            raise ValueError("new t must divide t")

        def red(cs):
            return [[D.cmod(c, new_t) for c in v] for v in cs]

        return CRTTransformFactor(new_t, self.p, None, red(self.u_p), red(self.inv_u_p), self.gen_jp)

    def hom_cost(self):
        """homcrt.rs:180-185: (hoisted, unhoisted) automorphisms of one homomorphic evaluation"""
        b, g = bsgs_split(len(self.u_p))
        return max(b - 2, 0), g - 1 + (1 if b > 1 else 0)


class UnsplitCRTTransform:
    """UnsplitCRTTransform (homcrt.rs:214-304): one factor per prime p | m_K, in m_k_factors order"""

    def __init__(self, t, field, factors=None):
        self.t = t
        self.factors = factors if factors is not None else [CRTTransformFactor(t, p, field) for p in field.m_k_factors()]

    def reduce_t(self, new_t):
        return UnsplitCRTTransform(new_t, None, [f.reduce_t(new_t) for f in self.factors])

    def forward(self):
        """(generator, coefficients) per factor for slots_to_coeffs_dbfv"""
        return [(f.gen_jp, f.u_p) for f in self.factors]

    def inverse(self):
        """(generator, coefficients) per factor for coeffs_to_slots_dbfv"""
        return [(f.gen_jp, f.inv_u_p) for f in self.factors]

    def hom_cost(self):
        h = u = 0
        for f in self.factors:
            a, b = f.hom_cost()
            h, u = h + a, u + b
        return h, u


def apply_cyclic_transform_plain(ring, data, base, automorphism, coeffs):
    """apply_cyclic_transform_plain (homcrt.rs:188-211) on the device: sum_i coeffs[i] sigma^i(x)
    of the CRT-basis elements `data` ([count, r, ., .], bound 1), reduced (bound 1)"""
    count = data.shape[0]
    acc = ring.allocate_wide(count, base)
    current = data
    for i, c in enumerate(coeffs):
        acc.fma(range(count), current, 1, ring.lprime_element_tensor(base, c), 1)
        if i + 1 < len(coeffs):
            nxt = ring.allocate_uninit(count, base)
            ring.apply_galois(nxt, current, automorphism)
            current = nxt
    return ring.reduce_wide(acc)
