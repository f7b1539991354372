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
    result = ring.allocate_uninit(pb.shape[0], lifted)  # every part is written below
    # the reference converts a copy of pb (bconv destroys its input); preserve_src leaves pb
    # unchanged in every conversion form, copying only where the form needs it
    ring.bconv_tensors(result[:, base_len:], ext, pb, base, preserve_src=True)
    if crt is not None:
        ring.pb_to_crt_tensor(result[:, base_len:], ext)
        be.hal_assign(result[:, :base_len], crt)
    else:
        be.hal_assign(result[:, :base_len], pb)
        ring.pb_to_crt_tensor(result, lifted)
    return result, lifted


def rescale_to_base(ring, x, lifted, base_len, t, bound=1):
    """rescale_to_base (bfv.rs:601-622): round(t x / q') over q' = base for the period-basis
    batch `x` over `lifted` (destroyed), whose representatives are bounded by `bound` p. Like the
    reference's mul_scalar (ring.rs:923-924, assert_mul_mont_safe) it rejects bounds with
    bound * max p > 2^63."""
    count = x.shape[0]
    base, ext = lifted[:base_len], lifted[base_len:]
    if bound * max(ring.primes[i] for i in lifted) > 1 << 63:
        raise ValueError("representatives too large for a montgomery multiplication")
    s = ring.scalar_tensor(lifted, t)   # mul_scalar, then bconv, sub and the inverse's mul in one step
    ring.rescale_step(x[:, base_len:], ext, x[:, :base_len], base, s[:, :base_len], s[:, base_len:],
                      ring.inverse_product_tensor(ext, base))
    result = ring.allocate_uninit(count, base)
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


def _rows(t, first, step):
    """rows first, first + step, ... of a batch tensor, as a strided view (no copy)"""
    from .hal import DeviceTensor

    n = (t.shape[0] - first + step - 1) // step
    return DeviceTensor(t.backend, t.ptr + first * t.strides[0] * t.elem, (n,) + t.shape[1:],
                        (t.strides[0] * step,) + t.strides[1:], t.elem, t._owner or t)


def bfv_hom_mul_batch(ring, t, lhs, rhs, base, keys, digits):
    """bfv_hom_mul (bfv.rs:747-775) for B ciphertext pairs at once: lhs, rhs are [2B, r, ., .]
    CRT-basis tensors with pair b in rows 2b, 2b + 1; the result has the same layout. Every HAL
    call covers all B pairs, the gadget base conversion (one element in bfv.rs:417-421) included.
    Each element sees the same calls with the same operands, in the same order, as in B calls of
    bfv_hom_mul, and the wide accumulators track the same per-row bounds, so the residues are
    identical."""
    be = ring.backend
    B = lhs.shape[0] // 2

    def lift(ct):
        ct_pb = ring.allocate_uninit(2 * B, base)
        be.hal_assign(ct_pb, ct)
        ring.crt_to_pb_tensor(ct_pb, base)
        return lift_to_extended(ring, ct_pb, base, crt=ct)

    lhs_l, lifted = lift(lhs)
    rhs_l, _ = lift(rhs)
    # hom_mul_raw_acc (bfv.rs:667-683) and reduce_wide as one kernel, rows laid out [3][B]:
    # component k of pair b in row k B + b (operand rows 2b + s)
    prod = ring.convolve_reduce(lhs_l, (0, 1, 2), rhs_l, (0, 1, 2), B, 1, 1, 1, lifted)
    ring.crt_to_pb_tensor(prod, lifted)
    rescaled = rescale_to_base(ring, prod, lifted, len(base), t)  # [3][B] over base
    out, _ = relin_batch(ring, rescaled, B, base, keys, digits, "crt")
    return out


def relin_batch(ring, prod, K, base, keys, digits, output_basis, bound=1, interleave=True, add_c01=True):
    """relin (bfv.rs:699-741) of K products at once: `prod` is a [3K, r, ., .] period-basis
    tensor (destroyed) with component c of product k in row c K + k, of bound `bound`. Returns
    (ciphertexts [2K, r, ., .] with product k in rows 2k, 2k + 1, or with interleave=False in
    rows k, K + k; bound): CRT basis (bound 1), or period basis (reduced key-switch part plus
    c0, c1: bound + 1). One fused gadget conversion covers all K, and one kernel forms the key
    switch of both key rows with its wide reduction (csd_keyswitch_reduce); each product gets the
    values of its own relin, so the residues are identical."""
    be = ring.backend
    r = len(base)
    dg = clip_digits(digits, r)
    d = len(dg)
    if d == 0:
        raise ValueError("relin: the key has no digits")
    parts = ring.allocate_uninit(K * d, base)  # product k, digit i at row k d + i
    ring.bconv_gadget_tensors(parts, prod[2 * K:3 * K], base, dg)  # may destroy rows 2K.. (unused below)
    ring.pb_to_crt_tensor(parts, base)
    # key_switch_acc (bfv.rs:498-528) for both key rows, the c0, c1 add of the CRT branch and
    # reduce_wide as one kernel; the reduction kind follows the WideRingEls bound
    keys = list(keys[:d])
    if any(k.strides != keys[0].strides for k in keys):  # the kernel takes one stride set for all keys
        copies = []
        for k in keys:  # contiguous copies of the limbs the base uses
            c = be.allocate_i64_uninit([2, r] + list(k.shape[2:]))
            be.hal_assign(c, k[:, :r])
            copies.append(c)
        keys = copies
    max_p = max(ring.primes[i] for i in base)
    min_p = min(ring.primes[i] for i in base)
    wide_bound = d * max_p * max_p
    out_rows = (1, 2) if interleave else (K, 1)  # row of (key row s, product k) = s rows[0] + k rows[1]
    out = ring.allocate_uninit(2 * K, base)
    if output_basis == "crt":
        if bound > 2:
            raise ValueError("basis transforms require representatives bounded by 2p")
        ring.pb_to_crt_tensor(prod[0:2 * K], base)
        wide_bound += max_p * max_p  # acc.add(rows, c0 c1, 1)
        be.keyswitch_reduce(out, out_rows, parts, K, d, keys, prod[0:2 * K], ring.scalar_tensor(base, 1),
                            wide_bound > (min_p << 63), ring.mods_for(base))
        return out, 1
    if output_basis != "pb":
        raise ValueError("output_basis must be 'crt' or 'pb'")
    be.keyswitch_reduce(out, (K, 1), parts, K, d, keys, None, None, wide_bound > (min_p << 63), ring.mods_for(base))
    ring.crt_to_pb_tensor(out, base)
    if not add_c01:  # the caller adds c0, c1 itself (add_els, ring.rs:990-1023)
        if interleave:
            raise ValueError("add_c01=False returns the key-switch part in the component-major layout")
        return out, 1
    be.hal_add_i64(out, prod[0:2 * K])  # add_els (ring.rs:990-1023)
    if not interleave:
        return out, 1 + bound
    result = ring.allocate_uninit(2 * K, base)
    be.hal_assign(_rows(result, 0, 2), out[0:K])
    be.hal_assign(_rows(result, 1, 2), out[K:2 * K])
    return result, 1 + bound
