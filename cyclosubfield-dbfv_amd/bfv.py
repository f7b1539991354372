"""The bfv-layer call sequences of the hot path, as host code over the device ring
(``ring.RingContext``) and backend (``hal.HipBackend``): the scheme functions of
cyclosubfield-dbfv/src/bfv.rs that issue the ring's transforms, base conversions and wide
accumulations. Tensors are ``DeviceTensor`` views in the reference layout [count, r, deg_L', deg_K];
representatives are almost reduced (bound 1) unless stated. Each function names the reference
lines it follows; the HAL call sequence is the reference's, so results are bit-identical with any
backend that is bit-identical per call."""


def select_digits(digits, rns_base_len):
    """select_digits (bfv.rs:388-401): `digits` contiguous prime ranges, the first ones longer."""
    if not 0 < digits <= rns_base_len:
        raise ValueError("need 1 <= digits <= RNS base length")
    small, large = divmod(rns_base_len, digits)
    res, cur = [], 0
    for i in range(digits):
        n = small + 1 if i < large else small
        res.append((cur, cur + n))
        cur += n
    return res


def clip_digits(digits, base_len):
    """clip_digits (bfv.rs:120-126)"""
    return [(a, min(b, base_len)) for a, b in digits if a < base_len]


def lift_to_extended(ring, pb, base, crt=None):
    """lift_to_extended (bfv.rs:564-595): the period-basis batch `pb` over `base`, lifted to
    base + extension in CRT basis (the base part taken from `crt` when given). Returns the
    tensor and the lifted base."""
    be = ring.backend
    base_len = len(base)
    ext = ring.extension_rns_base_for(base)
    lifted = list(base) + ext
    result = ring.allocate(pb.shape[0], lifted)
    tmp = ring.allocate(pb.shape[0], base)
    be.hal_assign(tmp, pb)  # bconv destroys its input
    ring.bconv_tensors(result[:, base_len:], ext, tmp, base)
    if crt is not None:
        ring.pb_to_crt_tensor(result[:, base_len:], ext)
        be.hal_assign(result[:, :base_len], crt)
    else:
        be.hal_assign(result[:, :base_len], pb)
        ring.pb_to_crt_tensor(result, lifted)
    return result, lifted


def rescale_to_base(ring, x, lifted, base_len, t):
    """rescale_to_base (bfv.rs:601-622): round(t x / q') over q' = base for the period-basis
    batch `x` over `lifted` (destroyed)."""
    be = ring.backend
    count = x.shape[0]
    base, ext = lifted[:base_len], lifted[base_len:]
    be.hal_mul_mont64(x, ring.scalar_tensor(lifted, t), ring.mods_for(lifted))   # mul_scalar
    tmp = ring.allocate(count, ext)
    ring.bconv_tensors(tmp, ext, x[:, :base_len], base)
    be.hal_sub_i64(x[:, base_len:], tmp)
    be.hal_mul_mont64(x[:, base_len:], ring.inverse_product_tensor(ext, base), ring.mods_for(ext))
    result = ring.allocate(count, base)
    ring.bconv_tensors(result, base, x[:, base_len:], ext)
    return result


def hom_mul_raw_acc(acc, acc_row, lhs, rhs):
    """hom_mul_raw_acc (bfv.rs:667-683): (l0 r0, l0 r1 + l1 r0, l1 r1) into rows acc_row.. of a
    wide batch, for almost-reduced 2-row CRT-basis tensors."""
    l0, l1 = lhs[0:1], lhs[1:2]
    r0, r1 = rhs[0:1], rhs[1:2]
    acc.fma(range(acc_row, acc_row + 1), l0, 1, r0, 1)
    acc.fma(range(acc_row + 1, acc_row + 2), l0, 1, r1, 1)
    acc.fma(range(acc_row + 1, acc_row + 2), l1, 1, r0, 1)
    acc.fma(range(acc_row + 2, acc_row + 3), l1, 1, r1, 1)


def gadget_decompose(ring, c1_pb, base, digits):
    """gadget_decompose (bfv.rs:408-426): the [d, r, ., .] digits of one period-basis element, in
    CRT basis."""
    be = ring.backend
    result = ring.allocate(len(digits), base)
    tmp = ring.allocate(1, base)
    be.hal_assign(tmp, c1_pb)
    ring.bconv_gadget_tensors(result, tmp, base, digits)
    ring.pb_to_crt_tensor(result, base)
    return result


def key_switch_acc(ring, acc, c1_pb, base, keys, digits):
    """key_switch_acc (bfv.rs:498-528): acc[0..2] = sum_i d_i * key_i over the gadget digits of
    c1_pb; keys[i] is the [2, r_key, ., .] CRT-basis key of digit i, restricted to `base` (a prefix
    of its base) as a copy-free view."""
    r = len(base)
    digits = clip_digits(digits, r)
    parts = gadget_decompose(ring, c1_pb, base, digits)
    terms = [(parts[i:i + 1], 1, keys[i][:, :r], 1) for i in range(len(digits))]
    if terms:
        acc.inner_product(range(0, 2), terms)


def relin_crt(ring, prod, base, keys, digits):
    """relin (bfv.rs:699-741) with CRT-basis output: `prod` is a [3, r, ., .] period-basis product
    (destroyed)."""
    acc = ring.allocate_wide(2, base)
    key_switch_acc(ring, acc, prod[2:3], base, keys, digits)
    ring.pb_to_crt_tensor(prod[0:2], base)
    acc.add(range(0, 2), prod[0:2], 1)
    return ring.reduce_wide(acc)


def bfv_hom_mul(ring, t, lhs, rhs, base, keys, digits):
    """bfv_hom_mul (bfv.rs:747-775) with CRT-basis output, for CRT-basis ciphertexts lhs, rhs
    ([2, r, ., .] each) over `base`; keys/digits form the relinearization key."""
    be = ring.backend
    base_len = len(base)

    def lift(ct):
        ct_pb = ring.allocate(2, base)
        be.hal_assign(ct_pb, ct)
        ring.crt_to_pb_tensor(ct_pb, base)
        return lift_to_extended(ring, ct_pb, base, crt=ct)

    lhs_l, lifted = lift(lhs)
    rhs_l, _ = lift(rhs)
    acc = ring.allocate_wide(3, lifted)
    hom_mul_raw_acc(acc, 0, lhs_l, rhs_l)
    prod = ring.reduce_wide(acc)
    ring.crt_to_pb_tensor(prod, lifted)
    rescaled = rescale_to_base(ring, prod, lifted, base_len, t)
    return relin_crt(ring, rescaled, base, keys, digits)
