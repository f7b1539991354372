"""BFV/dBFV bootstrapping of cyclosubfield-dbfv/src/bootstrap/mod.rs on the MI355X, for a batch of
ciphertexts: slots-to-coefficients (the homomorphic CRT transform, ``homcrt``), recombination of
the dBFV digits, mod-switch to the sparse-key-encapsulation base and key switch to the sparse key,
the L'-correction and the trace to K, the homomorphic evaluation of the noisy decryption under
the dBFV encryption of the sparse key, coefficients-to-slots and the floor division that removes
the noise (``digitex``).

Every step runs on the device, batched over the ciphertexts where the reference's steps are per
ciphertext. The reference's two host steps are device kernels here too: decoding both components
after the trace (decode_pb_scaled, exact: csd_ring_decode_scaled) and encoding the decoded
plaintexts, the second as its balanced base-b digit vectors (csd_ring_encode_digits), for the
homomorphic decryption; nothing returns to the host inside a bootstrap except the digit count."""
import math

import numpy as np

from . import dbfv as D
from . import homcrt as H
from . import keys as K

RNS_FACTOR_SIZE_LOG2 = 57.0  # bootstrap/mod.rs:31
SPARSE_KEY_ENCAPS_LEN = 3  # ring.rs:639
AMPLIFICATION_SAMPLES = 8  # bootstrap/mod.rs:200


class SparseKeyEncaps:
    """SparseKeyEncaps (bootstrap/mod.rs:35-43): the key switching key from the main key to the
    sparse key over the small encapsulation base, and the dBFV encryption of the sparse key at
    the intermediate modulus over the initial base"""

    def __init__(self, ksk, enc):
        self.ksk, self.enc = ksk, enc


def sparse_key_encaps_rns_base(ring):
    """ring.rs:735"""
    return list(range(min(SPARSE_KEY_ENCAPS_LEN, ring.initial_rns_base_len)))


def gen_sparse_sk_in_k(ring, field, hwt, rng):
    """gen_sparse_sk_in_k (bootstrap/mod.rs:53-71): hwt distinct positions of K (collisions drawn
    again), each +-1, scaled from the M_K- to the M-induced period basis, repeated over the
    deg_L' blocks of L"""
    dk = field.deg_k
    if hwt > dk:
        raise ValueError("hamming weight exceeds deg_K")
    sk = [0] * dk
    for _ in range(hwt):
        while True:
            j = rng.word() % dk
            if sk[j] == 0:
                sk[j] = (rng.word() % 2) * 2 - 1
                break
    s = field.scale_m_k_to_m_period_basis()
    return [x * s for _ in range(field.deg_lprime) for x in sk]


def log2_sparse_key_amplification(ring, field, coeffs, samples, rng):
    """log2_sparse_key_amplification (bootstrap/mod.rs:212-252): the average over `samples`
    rounding errors r (|r| <= 1/2, modelled as integers scaled by 2^20, K-valued) of
    log2 max |r s_sparse|_pbinf - 20, on the shortest prefix of the base that holds the product"""
    scale = 20
    log2_t = 2 * scale + 20
    t = 1 << log2_t
    required = 2 * log2_t + field.log2_pb_expansion_factor_bound()
    base = []
    while (ring.modulus_of(base).bit_length() - 1 if base else 0) < required:
        if len(base) >= ring.initial_rns_base_len:
            raise ValueError("RNS modulus too small to measure the amplification")
        base.append(len(base))
    rhs = K.encode_pb_unscaled(ring, base, t, coeffs)
    ring.pb_to_crt_tensor(rhs, base)
    dk = field.deg_k
    total = 0
    for _ in range(samples):
        words = rng.u64(dk)
        r_k = [int(w % np.uint64(1 << scale)) - (1 << (scale - 1)) for w in words]
        lhs = K.encode_pb_unscaled(ring, base, t, [x for _ in range(field.deg_lprime) for x in r_k])
        ring.pb_to_crt_tensor(lhs, base)
        acc = ring.allocate_wide(1, base)
        acc.fma(range(1), lhs, 1, rhs, 1)
        prod = ring.reduce_wide(acc)
        ring.crt_to_pb_tensor(prod, base)
        total += max(abs(v) for v in K.decode_pb_unscaled(ring, t, prod, base))
    return math.log2(total / samples) - scale


def _finish_sparse_key_encaps(ring, field, intermediate, sk, coeffs, rng, log2_alpha):
    """finish_sparse_key_encaps (bootstrap/mod.rs:153-180)"""
    enc = K.dbfv_enc(ring, intermediate, ring.initial_rns_base(), coeffs, sk, rng, log2_alpha)
    ske = sparse_key_encaps_rns_base(ring)
    sparse = K.element_from_pb_coeffs(ring, ske, coeffs)
    ring.pb_to_crt_tensor(sparse, ske)
    digits = [(i, i + 1) for i in range(len(ske))]
    ksk = K.gen_ksk(ring, ske, sk.data, K.SecretKey(sparse, ske), digits, rng, log2_alpha)
    full = K.element_from_pb_coeffs(ring, ring.initial_rns_base(), coeffs)
    ring.pb_to_crt_tensor(full, ring.initial_rns_base())
    return SparseKeyEncaps(ksk, enc), K.SecretKey(full, ring.initial_rns_base())


def create_sparse_key_encaps(ring, field, intermediate, sk, hwt, rng, log2_alpha, max_log2_amplification=None,
                             max_attempts=128):
    """create_sparse_key_encaps / _with_bound (bootstrap/mod.rs:76-151): with a bound, sparse keys
    whose amplification exceeds it are drawn again. Returns (encapsulation, sparse key)."""
    for _ in range(max_attempts if max_log2_amplification is not None else 1):
        coeffs = gen_sparse_sk_in_k(ring, field, hwt, rng)
        if max_log2_amplification is None or \
                log2_sparse_key_amplification(ring, field, coeffs, AMPLIFICATION_SAMPLES, rng) <= max_log2_amplification:
            return _finish_sparse_key_encaps(ring, field, intermediate, sk, coeffs, rng, log2_alpha)
    raise ValueError(f"no sparse key of hamming weight {hwt} with log2 amplification <= {max_log2_amplification} "
                     f"found in {max_attempts} attempts")


def trace_to_k(ring, field, els, base):
    """trace_to_k (bootstrap/mod.rs:186-196): every CRT-basis element replaced by the sum of its
    Gal(L|K)-conjugates (the L'-transversal), through one wide accumulator"""
    count = els.shape[0]
    acc = ring.allocate_wide(count, base)
    acc.add(range(count), els, 1)
    rotated = ring.allocate_uninit(count, base)
    for g in field.lprime_q_transversal()[1:]:
        ring.apply_galois(rotated, els, g)
        acc.add(range(count), rotated, 1)
    ring.reduce_wide_into(els, acc)


def key_switch(ring, c0_crt, c1_pb, base, ksk):
    """key_switch (bfv.rs:533-562) of one ciphertext (c0 in CRT, c1 in period basis, one row each)
    to the key of `ksk`: the gadget digits of c1, their inner product with the key digits, plus
    c0 on row 0, reduced; CRT basis"""
    from .bfv import clip_digits

    r = len(base)
    digits = clip_digits(ksk.digits, r)
    parts = ring.allocate_uninit(len(digits), base)
    tmp = ring.allocate_uninit(1, base)
    ring.backend.hal_assign(tmp, c1_pb)
    ring.bconv_gadget_tensors(parts, tmp, base, digits)
    ring.pb_to_crt_tensor(parts, base)
    acc = ring.allocate_wide(2, base)
    keys = ksk.keys()
    terms = [(parts[i:i + 1], 1, keys[i][:, 0:r], 1) for i in range(len(digits))]
    if terms:
        acc.inner_product(range(2), terms)
    acc.add(range(1), c0_crt, 1)
    return ring.reduce_wide(acc)


def dbfv_acc_fma_tensors_batch(ring, acc, lhs, tensors):
    """dbfv_acc_fma_tensors (dbfv.rs:193-224) of one dBFV ciphertext `lhs` (a batch of 1, CRT basis)
    with per-ciphertext factor digits: tensors[j] is [count, r, ., .], row c the j-th digit of
    ciphertext c's factor, so acc digit k of ciphertext c += sum_(i + j = k) lhs digit i *
    tensors[j][c]"""
    n = acc.count
    ld = lhs.digits()
    if ld == 0 or not tensors:
        return
    new_digits = min(acc.capacity(), ld + len(tensors) - 1)
    for i in range(ld):
        for j, t in enumerate(tensors):
            k = i + j
            if k < new_digits:
                for s in range(2):
                    rows = range((2 * k + s) * n, (2 * k + s + 1) * n)
                    acc.acc.fma(rows, lhs.data[2 * i + s:2 * i + s + 1], lhs.bound, t, 1)
    acc.used_digits = max(acc.used_digits, new_digits)


def bootstrap_noise_base_len(ring, field, transform, modulus, ct_base, gks):
    """the slots-to-coefficients base length of bootstrap_bfv_dbfv (bootstrap/mod.rs:302-321)"""
    gk_noise = max(gk.log2_key_switch_noise(ring, ct_base) for gk in gks.values())
    s2c_noise = len(transform.factors) * field.noise_growth_hom_mul_plain_lower_estimate(modulus.b()) * 2.0 \
        + gk_noise + math.log2(float(modulus.t()))
    want = math.ceil((math.log2(float(modulus.t())) + s2c_noise) / RNS_FACTOR_SIZE_LOG2)
    return min(len(ct_base), max(want, len(sparse_key_encaps_rns_base(ring))))


def bootstrap_bfv_dbfv(ring, field, crt_transform, modulus, ct, rk, gks, encaps, fused=True, timer=None):
    """bootstrap_bfv_dbfv (bootstrap/mod.rs:280-408) for a batch `ct` of dBFV ciphertexts (a
    DBFVBatch over the initial base, either basis). crt_transform is the UnsplitCRTTransform at the
    intermediate plaintext modulus, rk the relinearization KeySwitchKey, gks the Galois
    KeySwitchKeys by element, encaps the SparseKeyEncaps. Returns a CRT-basis batch encrypting the
    plaintexts with fresh noise. `timer(name)` is called after each phase when given."""
    be = ring.backend
    tick = timer or (lambda name: None)
    base = ring.initial_rns_base()
    inter_t = crt_transform.t
    log2_inter_t = inter_t.bit_length() - 1
    if inter_t != 1 << log2_inter_t:
        raise ValueError("the intermediate plaintext modulus must be a power of two")
    n = ct.count
    reduced_transform = crt_transform.reduce_t(modulus.t())
    if ct.basis == "crt":
        data = ring.allocate_uninit(ct.data.shape[0], ct.base)
        be.hal_assign(data, ct.data)
        ring.crt_to_pb_tensor(data, ct.base)
        ct = D.DBFVBatch(data, ct.base, n, "pb", 1)
    s2c_len = bootstrap_noise_base_len(ring, field, crt_transform, modulus, ct.base, gks)
    down = D.dbfv_mod_switch_down(ring, ct, s2c_len)
    s2c_base = down.base
    ring.pb_to_crt_tensor(down.data, s2c_base)
    down = D.DBFVBatch(down.data, s2c_base, n, "crt", 1)
    in_coeffs = H.slots_to_coeffs_dbfv(ring, reduced_transform.forward(), modulus, down, gks)
    tick("slots_to_coeffs")

    # recombine the digits: sum_i 2^(i log2_b) digit_i, one wide batch of 2 rows per ciphertext
    acc = ring.allocate_wide(2 * n, s2c_base)
    for i in range(in_coeffs.digits()):
        acc.fma(range(2 * n), in_coeffs.digit_rows(i), in_coeffs.bound,
                ring.scalar_tensor(s2c_base, 1 << (i * modulus.log2_b)), 1)
    bfv = ring.reduce_wide(acc)  # row s n + c: component s of ciphertext c, CRT basis
    ring.crt_to_pb_tensor(bfv, s2c_base)
    ske = sparse_key_encaps_rns_base(ring)
    switched = D.dbfv_mod_switch_down(ring, D.DBFVBatch(bfv, s2c_base, n, "pb", 1), len(ske)).data
    ring.pb_to_crt_tensor(switched[0:n], ske)  # c0 to CRT basis, c1 stays in period basis
    tick("recombine_modswitch")

    ks = ring.allocate_uninit(2 * n, ske)  # row s n + c
    for c in range(n):
        one = key_switch(ring, switched[c:c + 1], switched[n + c:n + c + 1], ske, encaps.ksk)
        be.hal_assign(ks[c:c + 1], one[0:1])
        be.hal_assign(ks[n + c:n + c + 1], one[1:2])
    be.hal_mul_mont64(ks, ring.lprime_element_tensor(ske, field.lprime_trace_one_coeffs()), ring.mods_for(ske))
    trace_to_k(ring, field, ks, ske)
    ring.crt_to_pb_tensor(ks, ske)
    # decode_pb_scaled of both components (rows c: c0, n + c: c1) on the device, exact
    plain = ring.decode_scaled(ks, ske, log2_inter_t)
    tick("keyswitch_trace_decode")

    # the noisy decryption c0 + c1 s_sparse, plus the rounding constant; the plaintexts are encoded
    # over the initial base on the device (c1 as its balanced base-b digit vectors)
    inter = D.DBFVModulus.new_incomplete(modulus.log2_b, log2_inter_t)
    planes, nd = ring.encode_digits(plain[n:2 * n], base, modulus.log2_b, inter.digits())
    macc = D.dbfv_allocate_accumulator(ring, inter, base, "crt", n)
    if nd:
        ring.pb_to_crt_tensor(planes[0:nd * n], base)
        dbfv_acc_fma_tensors_batch(ring, macc, encaps.enc, [planes[j * n:(j + 1) * n] for j in range(nd)])
    noisy = D.dbfv_finish_accumulator(ring, macc)
    delta = D.rounded_div(ring.modulus_of(base), inter_t)
    const, _ = ring.encode_digits(plain[0:n], base)
    ring.pb_to_crt_tensor(const, base)
    wide = ring.allocate_wide(n, base)
    wide.fma(range(n), const, 1, ring.scalar_tensor_big(base, delta), 1)
    wide.add(range(n), noisy.digit_rows(0)[0:n], noisy.bound)
    round_to_floor = (inter_t // modulus.t() // 2) * delta
    wide.add(range(n), ring.scalar_tensor_big(base, round_to_floor), 1)
    ring.reduce_wide_into(noisy.data[0:n], wide)
    tick("hom_decrypt")

    in_slots = H.coeffs_to_slots_dbfv(ring, crt_transform.inverse(), inter, noisy, gks)
    tick("coeffs_to_slots")
    ring.crt_to_pb_tensor(in_slots.data, in_slots.base)
    in_slots = D.DBFVBatch(in_slots.data, in_slots.base, n, "pb", 1)
    from . import digitex

    res = digitex.hom_floordiv_dbfv(ring, inter, modulus, in_slots, rk.keys(), rk.digits, fused=fused)
    ring.pb_to_crt_tensor(res.data, res.base)
    tick("floordiv")
    return D.DBFVBatch(res.data, res.base, n, "crt", 1)
