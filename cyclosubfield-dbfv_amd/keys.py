"""Key generation, encryption and decryption of cyclosubfield-dbfv (bfv.rs:41-386, 433-497,
dbfv.rs:602-722) on the MI355X: secret keys, key-switching, relinearization and Galois keys, BFV
and dBFV encryption and decryption, and the host-side encoding and decoding of plaintexts.

These are the set-up and check steps around the hot path (keys are made once; decryption checks
results). The ring arithmetic runs on the device through the ring layer; only the random samples
(the rounded Gaussians of the errors, the sparse ternary keys) are drawn on the host, and the
uniform ring elements are generated on the device (``RingContext.fill_uniform``). The random
streams are not the reference's splitmix64 streams: keys and ciphertexts are fresh random objects
with the reference's distributions, and correctness is checked by decryption.

Decoding lifts every coefficient to Z exactly (the CRT over the ciphertext's base, with Python
integers) as ring.rs:1465-1480 does."""
import math

import numpy as np

from . import dbfv as D
from .bfv import clip_digits, select_digits  # noqa: F401

MIN_SIGMA = 3.2  # bfv.rs:139
SK_HWT = 128  # bfv.rs:41


class Rng:
    """the host random source: 64-bit words from numpy's PCG64, seeded"""

    def __init__(self, seed):
        self.gen = np.random.Generator(np.random.PCG64(seed))

    def u64(self, n):
        return self.gen.integers(0, 1 << 64, size=n, dtype=np.uint64)

    def word(self):
        return int(self.u64(1)[0])


def log2_modulus_of(ring, base):
    """ring.rs:756-758"""
    return sum(math.log2(float(ring.primes[i])) for i in base)


def sigma_for_log2_alpha(ring, base, log2_alpha):
    """bfv.rs:149-163: Q / alpha, floored at MIN_SIGMA"""
    return max(MIN_SIGMA, 2.0 ** (log2_modulus_of(ring, base) - log2_alpha))


def log2_alpha_for_sigma(ring, base, sigma):
    """bfv.rs:165-168"""
    return log2_modulus_of(ring, base) - math.log2(sigma)


def sample_gaussian(rng, sigma, n):
    """sample_gaussian (bfv.rs:170-176) for n samples: Box-Muller from two 53-bit uniforms, times
    sigma, rounded half away from zero"""
    u1 = ((rng.u64(n) >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    u2 = (rng.u64(n) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2) * sigma
    r = np.sign(g) * np.floor(np.abs(g) + 0.5)
    if np.abs(r).max(initial=0) >= 2.0 ** 62:
        raise ValueError("Gaussian samples beyond the i64 range")
    return r.astype(np.int64)


def element_from_pb_coeffs(ring, base, coeffs):
    """element_from_pb_coeffs (ring.rs:1449-1461): a period-basis element over `base` (bound 1) with
    the given integer coefficients, in montgomery form. Coefficients within the i64 range are
    converted on the device with i64_to_mont64; wider ones (below 2^126) as i128 with
    i128_to_mont64. The representatives are congruent to the reference's."""
    be = ring.backend
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    r = len(base)
    mods = ring.mods_for(base)
    arr = np.asarray(coeffs, dtype=object) if not isinstance(coeffs, np.ndarray) else coeffs
    if arr.size != n2 * n3:
        raise ValueError("one coefficient per period-basis position")
    out = ring.allocate_uninit(1, base)
    if arr.dtype != object or all(-(1 << 63) <= int(c) < (1 << 63) for c in arr):
        vals = np.asarray(arr, dtype=np.int64).reshape(n2, n3)
        be.hal_import(out, np.broadcast_to(vals, (1, r, n2, n3)))
        be.hal_i64_to_mont64(out, mods)
        return out
    words = np.zeros((n2 * n3, 2), dtype=np.int64)
    for k, c in enumerate(arr):
        c = int(c)
        if not -(1 << 126) < c < (1 << 126):
            raise ValueError("coefficient beyond the i128 conversion range")
        w = c & ((1 << 128) - 1)
        lo, hi = w & ((1 << 64) - 1), w >> 64
        words[k, 0] = lo - (1 << 64) if lo >> 63 else lo
        words[k, 1] = hi - (1 << 64) if hi >> 63 else hi
    wide = be.allocate_i128_uninit([1, r, n2, n3])
    be.import_i128(wide, np.ascontiguousarray(np.broadcast_to(words.reshape(1, 1, n2 * n3, 2), (1, r, n2 * n3, 2))))
    be.hal_i128_to_mont64(out, wide, mods)
    return out


def encode_pb_unscaled(ring, base, t, ptxt):
    """bfv.rs:226-235"""
    return element_from_pb_coeffs(ring, base, np.array([D.cmod(int(c), t) for c in ptxt], dtype=object))


def mul_scalar_big(ring, els, base, scalar):
    """ring.rs:932-938 (inputs of bound 1): els *= scalar, bound 1"""
    ring.backend.hal_mul_mont64(els, ring.scalar_tensor_big(base, scalar), ring.mods_for(base))


def encode_pb_scaled(ring, base, t, ptxt):
    """bfv.rs:237-248: Delta = round(Q / t) times the unscaled encoding"""
    el = encode_pb_unscaled(ring, base, t, ptxt)
    mul_scalar_big(ring, el, base, D.rounded_div(ring.modulus_of(base), t))
    return el


def encode_scalar(ring, x):
    """bfv.rs:864-868"""
    s = D.scale_q_to_m_period_basis(ring)
    return [x * s] * (ring.info["deg_lprime"] * ring.info["deg_k"])


def _crt_parts(ring, els, base):
    """the residues y_j = x (Q/p_j)^-1 mod p_j in [0, p_j) of every coefficient of a period-basis
    [1, r, ., .] element, as int64 arrays, with the CRT weights M_j = Q/p_j"""
    be = ring.backend
    q = ring.modulus_of(base)
    primes = [ring.primes[i] for i in base]
    ks = [pow((q // p) % p, -1, p) for p in primes]
    key = ("crt_inv_plain", tuple(base))
    kt = ring._consts.get(key)
    if kt is None:
        kt = be.hal_allocate_i64([1, len(base), 1, 1])
        be.hal_import(kt, np.array(ks, dtype=np.int64))
        ring._consts[key] = kt
    tmp = ring.allocate_uninit(1, base)
    be.hal_assign(tmp, els)
    be.hal_mul_mont64(tmp, kt, ring.mods_for(base))  # mont form x R times plain k_j, reduced: x k_j mod p
    ys = be.hal_export(tmp).reshape(len(base), -1)
    ys = [np.mod(ys[j], np.int64(p)) for j, p in enumerate(primes)]
    return ys, [q // p for p in primes], q


def element_lifts(ring, els, base):
    """element_lifts (ring.rs:1465-1480): the smallest lifts mod Q of the coefficients of a
    period-basis [1, r, ., .] element, as Python integers"""
    ys, ms, q = _crt_parts(ring, els, base)
    x = sum(y.astype(object) * m for y, m in zip(ys, ms)) % q
    return [int(v) - q if v > q // 2 else int(v) for v in x]


def decode_pb_unscaled(ring, t, els, base):
    """decode_pb_unscaled (bfv.rs:250-259) of a period-basis [1, r, ., .] element"""
    return [D.cmod(x % t, t) for x in element_lifts(ring, els, base)]


def decode_pb_scaled(ring, t, els, base):
    """decode_pb_scaled (bfv.rs:262-274): round(t x / Q) mod t per coefficient. With x = sum_j
    y_j M_j - k Q the value t x / Q is sum_j y_j t / p_j modulo t, and no coefficient sits on a
    half (Q odd), so round(.) = floor((2 N + Q) / 2 Q) with N = sum_j y_j t M_j"""
    ys, ms, q = _crt_parts(ring, els, base)
    num = sum(y.astype(object) * (t * m) for y, m in zip(ys, ms))
    v = (2 * num + q) // (2 * q)
    return [D.cmod(int(x) % t, t) for x in v]


def decode_scalar(ring, t, data_pb):
    """bfv.rs:870-878 (None if the plaintext is not a scalar)"""
    if any(x != data_pb[0] for x in data_pb):
        return None
    return D.cmod(data_pb[0] * D.scale_q_to_m_period_basis(ring), t)


class SecretKey:
    """a secret key (bfv.rs:47): one CRT-basis element [1, r0, ., .] over the initial base"""

    def __init__(self, data, base):
        self.data, self.base = data, list(base)

    def restrict(self, r):
        return self.data[:, 0:r]


def gen_sk(ring, rng, hwt=SK_HWT):
    """gen_sk (bfv.rs:180-193): ternary, SK_HWT draws of (position, sign) in the period basis
    (a later draw overwrites an earlier one), in CRT basis over the initial base"""
    n = ring.info["deg_lprime"] * ring.info["deg_k"]
    coeffs = np.zeros(n, dtype=np.int64)
    for _ in range(hwt):
        j = rng.word() % n
        coeffs[j] = (rng.word() % 2) * 2 - 1
    base = ring.initial_rns_base()
    sk = element_from_pb_coeffs(ring, base, coeffs)
    ring.pb_to_crt_tensor(sk, base)
    return SecretKey(sk, base)


def add_error(ring, els, base, rng, sigma):
    """add_error (bfv.rs:195-205): a rounded Gaussian per period-basis coefficient, added through a
    wide accumulator and reduced (bound 1)"""
    n = ring.info["deg_lprime"] * ring.info["deg_k"]
    err = element_from_pb_coeffs(ring, base, sample_gaussian(rng, sigma, n))
    acc = ring.allocate_wide(1, base)
    acc.add(range(1), els, 1)
    acc.add(range(1), err, 1)
    ring.reduce_wide_into(els, acc)


def random_element(ring, base, rng):
    """random_element (bfv.rs:207-224): uniform residues in [0, p) per limb, CRT basis"""
    a = ring.allocate_uninit(1, base)
    ring.fill_uniform(a, base, rng.word())
    return a


def _finish_enc(ring, base, sk, b, rng):
    """finish_enc (bfv.rs:310-327): (b + a s, -a)"""
    be = ring.backend
    a = random_element(ring, base, rng)
    ct = ring.allocate_uninit(2, base)
    acc = ring.allocate_wide(1, base)
    acc.fma(range(1), a, 1, sk.restrict(len(base)), 1)
    acc.add(range(1), b, 1)
    ring.reduce_wide_into(ct[0:1], acc)
    be.hal_mul_mont64(a, ring.scalar_tensor(base, -1), ring.mods_for(base))
    be.hal_assign(ct[1:2], a)
    return ct


def bfv_enc_zero(ring, base, sk, rng, log2_alpha):
    """bfv.rs:277-292: a CRT-basis encryption of zero over `base` (a prefix of the key's base)"""
    b = ring.allocate(1, base)
    add_error(ring, b, base, rng, sigma_for_log2_alpha(ring, base, log2_alpha))
    ring.pb_to_crt_tensor(b, base)
    return _finish_enc(ring, base, sk, b, rng)


def bfv_enc(ring, base, t, ptxt, sk, rng, log2_alpha):
    """bfv.rs:294-308"""
    b = encode_pb_scaled(ring, base, t, ptxt)
    add_error(ring, b, base, rng, sigma_for_log2_alpha(ring, base, log2_alpha))
    ring.pb_to_crt_tensor(b, base)
    return _finish_enc(ring, base, sk, b, rng)


def noisy_plaintext(ring, ct, base, sk):
    """noisy_plaintext (bfv.rs:330-345): c0 + c1 s of a CRT-basis [2, r, ., .] ciphertext, in
    period basis"""
    acc = ring.allocate_wide(1, base)
    acc.fma(range(1), ct[1:2], 1, sk.restrict(len(base)), 1)
    acc.add(range(1), ct[0:1], 1)
    noisy = ring.reduce_wide(acc)
    ring.crt_to_pb_tensor(noisy, base)
    return noisy


def bfv_dec(ring, t, ct, base, sk):
    """bfv.rs:347-359"""
    return decode_pb_scaled(ring, t, noisy_plaintext(ring, ct, base, sk), base)


class KeySwitchKey:
    """KeySwitchKey (bfv.rs:56-93): digit i is rows 2i, 2i+1 of one CRT-basis [2 d, r, ., .] tensor
    (an encryption of old_sk times the digit's CRT factor under the new key), with the digits and
    the log2 of the error deviation it was made with. It serves as a GaloisKey too."""

    def __init__(self, data, digits, base, log2_sigma):
        self.data, self.digits, self.base, self.log2_sigma = data, list(digits), list(base), log2_sigma

    def keys(self):
        return [self.data[2 * i:2 * i + 2] for i in range(len(self.digits))]

    def log2_key_switch_noise(self, ring, base):
        """bfv.rs:77-84"""
        return max((log2_modulus_of(ring, base[a:b]) for a, b in clip_digits(self.digits, len(base))), default=0.0) \
            + self.log2_sigma

    def size_bytes(self):
        return self.data.numel() * 8


def gen_ksk(ring, base, old_sk, new_sk, digits, rng, log2_alpha):
    """gen_ksk (bfv.rs:433-460): per digit a fresh encryption of zero under new_sk with old_sk times
    the digit's CRT factor (1 modulo its primes, 0 modulo the others, smallest lift) added to c0"""
    r = len(base)
    q = ring.modulus_of(base)
    log2_sigma = math.log2(sigma_for_log2_alpha(ring, base, log2_alpha))
    data = ring.allocate_uninit(2 * len(digits), base)
    be = ring.backend
    for i, (a, b) in enumerate(digits):
        ct = bfv_enc_zero(ring, base, new_sk, rng, log2_alpha)
        dq = 1
        for j in range(a, b):
            dq *= ring.primes[base[j]]
        factor = (q // dq) * pow((q // dq) % dq, -1, dq) % q
        factor = factor - q if factor > q // 2 else factor
        acc = ring.allocate_wide(1, base)
        acc.fma(range(1), old_sk[:, 0:r], 1, ring.scalar_tensor_big(base, factor), 1)
        acc.add(range(1), ct[0:1], 1)
        ring.reduce_wide_into(ct[0:1], acc)
        be.hal_assign(data[2 * i:2 * i + 2], ct)
    return KeySwitchKey(data, digits, base, log2_sigma)


def gen_rk(ring, sk, digits, rng, log2_alpha):
    """gen_rk (bfv.rs:462-477): the key switch from sk^2 to sk"""
    base = ring.initial_rns_base()
    acc = ring.allocate_wide(1, base)
    acc.fma(range(1), sk.data, 1, sk.data, 1)
    sq = ring.reduce_wide(acc)
    return gen_ksk(ring, base, sq, sk, digits, rng, log2_alpha)


def gen_gk(ring, sk, g, digits, rng, log2_alpha):
    """gen_gk (bfv.rs:479-496): the key switch from g(sk) to sk"""
    base = ring.initial_rns_base()
    gal = ring.allocate_uninit(1, base)
    ring.apply_galois(gal, sk.data, g)
    return gen_ksk(ring, base, gal, sk, digits, rng, log2_alpha)


def dbfv_log2_alpha_per_digit(modulus, log2_alpha):
    """dbfv.rs:602-604"""
    return log2_alpha + modulus.log2_b * (modulus.digits() - 1)


def dbfv_enc(ring, modulus, base, ptxt, sk, rng, log2_alpha):
    """dbfv_enc (dbfv.rs:661-690): digit 0 encrypts the message, the others zero, each with the
    per-digit error; then dbfv_reduce. A CRT-basis batch of one ciphertext."""
    d = modulus.digits()
    la = dbfv_log2_alpha_per_digit(modulus, log2_alpha)
    digit0 = bfv_enc(ring, base, modulus.t(), ptxt, sk, rng, la)
    if d == 1:
        return D.DBFVBatch(digit0, base, 1, "crt", 1)
    ct = ring.allocate_uninit(2 * d, base)
    ring.backend.hal_assign(ct[0:2], digit0)
    for i in range(1, d):
        ring.backend.hal_assign(ct[2 * i:2 * i + 2], bfv_enc_zero(ring, base, sk, rng, la))
    ring.crt_to_pb_tensor(ct, base)
    res, bound = D.dbfv_reduce(ring, modulus, ct, 1, base)
    if bound > 2:
        raise AssertionError("dbfv_reduce returned an unexpected bound")
    ring.pb_to_crt_tensor(res, base)
    return D.DBFVBatch(res, base, 1, "crt", 1)


def dbfv_dec(ring, modulus, ct, sk, c=0):
    """dbfv_dec (dbfv.rs:699-722) of ciphertext c of a CRT-basis batch: sum_i b^i dec(digit i),
    each digit's part taken modulo t / b^i, balanced modulo t"""
    if ct.basis != "crt":
        raise ValueError("dbfv_dec needs CRT basis")
    n = ring.info["deg_lprime"] * ring.info["deg_k"]
    result = [0] * n
    for i in range(ct.digits()):
        rows = ct.digit_rows(i)
        one = ring.allocate_uninit(2, ct.base)
        ring.backend.hal_assign(one[0:1], rows[c:c + 1])
        ring.backend.hal_assign(one[1:2], rows[ct.count + c:ct.count + c + 1])
        part = bfv_dec(ring, modulus.t(), one, ct.base, sk)
        w = 1 << (modulus.log2_t - i * modulus.log2_b)
        result = [a + (D.cmod(p, w) << (i * modulus.log2_b)) for a, p in zip(result, part)]
    t = modulus.t()
    return [D.cmod(x % t, t) for x in result]
