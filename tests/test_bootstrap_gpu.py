"""Bootstrapping (bootstrap/mod.rs) and the key generation, encryption and decryption around it
(bfv.rs:139-562, dbfv.rs:602-722) on the MI355X, checked by decryption (L2) with the reference's
own test values:

* BFV and dBFV encryption decrypt to the plaintext; a Galois key switch with a generated key
  decrypts to the rotated plaintext; a dBFV multiplication with a generated relinearization key
  decrypts to the product (dbfv.rs:1263-1322) -- the relinearization of a nonzero secret's key,
  which the zero-secret identity of the oracle tests leaves open;
* the homomorphic slots-to-coefficients transform of a real encryption decrypts to the plaintext
  transform (homcrt.rs:188-211 on the device);
* test_bootstrap (bootstrap/mod.rs:409-424): field_18, t = 2^15, (5, 3) -> intermediate 2^20,
  sparse weight 4: an encryption of 2025 bootstraps to an encryption of 2025;
* test_bootstrap_field_32 (bootstrap/mod.rs:429-455): two slot values packed, bootstrapped and
  decrypted to the same plaintext (two primes of L' above 2).
"""
import numpy as np
import pytest

from _pkg import load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    be = load().HipBackend(0)
    yield be
    be.close()


def _setup(gpu, field, bits, t_bits, seed):
    pkg = load()
    ring = pkg.ring.RingContext(gpu, field, bits, t_bits)
    f = pkg.field.FieldDescriptor(field)
    rng = pkg.keys.Rng(seed)
    base = ring.initial_rns_base()
    log2_alpha = pkg.keys.log2_alpha_for_sigma(ring, base, pkg.keys.MIN_SIGMA)
    sk = pkg.keys.gen_sk(ring, rng)
    return pkg, ring, f, rng, base, log2_alpha, sk


def test_bfv_and_dbfv_encryption_round_trip(gpu):
    pkg, ring, f, rng, base, la, sk = _setup(gpu, "field_18", (300, 320), 16, 1)
    K, D = pkg.keys, pkg.dbfv
    t = 1 << 16
    n = f.deg_l
    x = [int(v) for v in np.random.default_rng(2).integers(-(t // 2) + 1, t // 2 + 1, n)]
    ct = K.bfv_enc(ring, base, t, x, sk, rng, la)
    assert K.bfv_dec(ring, t, ct, base, sk) == x
    assert K.bfv_dec(ring, t, K.bfv_enc_zero(ring, base, sk, rng, la), base, sk) == [0] * n
    m = D.DBFVModulus.new(4, 4)
    dct = K.dbfv_enc(ring, m, base, x, sk, rng, la)
    assert dct.digits() == 4
    assert K.dbfv_dec(ring, m, dct, sk) == x
    assert K.decode_scalar(ring, t, K.bfv_dec(ring, t, K.bfv_enc(ring, base, t, K.encode_scalar(ring, 2025), sk, rng, la),
                                              base, sk)) == 2025
    ring.close()


def test_galois_key_switch_with_a_generated_key(gpu):
    """bfv_hom_gal (bfv.rs:783-839) with gen_gk keys: decrypts to sigma_g(m) (the Galois action
    permutes the period-basis coefficients, ring.rs:14-16)"""
    pkg, ring, f, rng, base, la, sk = _setup(gpu, "field_18", (300, 320), 16, 3)
    K, D, H = pkg.keys, pkg.dbfv, pkg.homcrt
    t = 1 << 16
    n = f.deg_l
    x = [int(v) for v in np.random.default_rng(4).integers(-(t // 2) + 1, t // 2 + 1, n)]
    digits = pkg.bfv.select_digits(len(base), len(base))
    gens = [g for g, _ in f.lq_cyclic]
    gs = [gens[0], gens[-1]]
    gks = {g: K.gen_gk(ring, sk, g, digits, rng, la) for g in gs}
    ct = D.DBFVBatch(K.bfv_enc(ring, base, t, x, sk, rng, la), base, 1, "crt", 1)
    rotated = H.bfv_hom_gal_batch(ring, ct, gs, gks)
    xe = K.element_from_pb_coeffs(ring, base, np.array(x, dtype=np.int64))
    for g, rot in zip(gs, rotated):
        want = ring.allocate(1, base)
        ring.apply_galois(want, xe, g)
        assert K.bfv_dec(ring, t, rot.data, base, sk) == K.decode_pb_unscaled(ring, t, want, base), g
    ring.close()


def test_dbfv_multiplication_with_a_generated_relinearization_key(gpu):
    """dbfv_hom_mul (dbfv.rs:1034-1060) of two encryptions decrypts to the plaintext product
    (mul_plain_pb, bfv.rs:845-862), as the reference's test_dbfv_hom_mul checks"""
    pkg, ring, f, rng, base, la, sk = _setup(gpu, "field_18", (800, 820), 20, 5)
    K, D = pkg.keys, pkg.dbfv
    m = D.DBFVModulus.new(5, 3)
    t = m.t()
    n = f.deg_l
    gen = np.random.default_rng(6)
    x = [int(v) for v in gen.integers(-(t // 2) + 1, t // 2 + 1, n)]
    y = [int(v) for v in gen.integers(-(t // 2) + 1, t // 2 + 1, n)]
    digits = pkg.bfv.select_digits(len(base), len(base))
    rk = K.gen_rk(ring, sk, digits, rng, la)
    a, b = K.dbfv_enc(ring, m, base, x, sk, rng, la), K.dbfv_enc(ring, m, base, y, sk, rng, la)
    for c in (a, b):
        ring.crt_to_pb_tensor(c.data, base)
    prod, bound = D.dbfv_hom_mul(ring, m, a.data, 1, b.data, 1, base, rk.keys(), rk.digits, output_basis="crt")
    xe = K.encode_pb_unscaled(ring, base, t, x)
    ye = K.encode_pb_unscaled(ring, base, t, y)
    ring.pb_to_crt_tensor(xe, base)
    ring.pb_to_crt_tensor(ye, base)
    acc = ring.allocate_wide(1, base)
    acc.fma(range(1), xe, 1, ye, 1)
    pe = ring.reduce_wide(acc)
    ring.crt_to_pb_tensor(pe, base)
    want = K.decode_pb_unscaled(ring, t, pe, base)
    assert K.dbfv_dec(ring, m, D.DBFVBatch(prod, base, 1, "crt", bound), sk) == want
    ring.close()


def test_slots_to_coeffs_of_a_real_encryption(gpu):
    """slots_to_coeffs_dbfv (homcrt.rs:437-449) with generated Galois keys decrypts to the plaintext
    transform (apply_cyclic_transform_plain, homcrt.rs:188-211, on the device)"""
    pkg, ring, f, rng, base, la, sk = _setup(gpu, "field_18", (800, 820), 20, 7)
    K, D, H = pkg.keys, pkg.dbfv, pkg.homcrt
    m = D.DBFVModulus.new_incomplete(5, 20)
    t = m.t()
    tr = H.UnsplitCRTTransform(t, f)
    digits = pkg.bfv.select_digits(len(base), len(base))
    gks = {g: K.gen_gk(ring, sk, g, digits, rng, la) for g in H.required_galois_elements("field_18")}
    x = [int(v) for v in np.random.default_rng(8).integers(-(t // 2) + 1, t // 2 + 1, f.deg_l)]
    ct = K.dbfv_enc(ring, m, base, x, sk, rng, la)
    out = H.slots_to_coeffs_dbfv(ring, tr.forward(), m, ct, gks)
    el = K.encode_pb_unscaled(ring, base, t, x)
    ring.pb_to_crt_tensor(el, base)
    for gen, cs in tr.forward():
        el = H.apply_cyclic_transform_plain(ring, el, base, gen, cs)
    ring.crt_to_pb_tensor(el, base)
    assert K.dbfv_dec(ring, m, out, sk) == K.decode_pb_unscaled(ring, t, el, base)
    ring.close()


def _bootstrap_case(gpu, field, seed, hwt, slot_values=None):
    pkg, ring, f, rng, base, la, sk = _setup(gpu, field, (800, 820), 20, seed)
    K, D, H, B = pkg.keys, pkg.dbfv, pkg.homcrt, pkg.bootstrap
    t = 1 << 15
    modulus = D.DBFVModulus.new(5, 3)
    inter = D.DBFVModulus.new_incomplete(modulus.log2_b, 20)
    tr = H.UnsplitCRTTransform(inter.t(), f)
    encaps, _ = B.create_sparse_key_encaps(ring, f, inter, sk, hwt, rng, la)
    digits = pkg.bfv.select_digits(len(base), len(base))
    gks = {g: K.gen_gk(ring, sk, g, digits, rng, la) for g in H.required_galois_elements(field)}
    rk = K.gen_rk(ring, sk, digits, rng, la)
    if slot_values is None:
        ptxt = K.encode_scalar(ring, 2025)
    else:
        ptxt = _pack(pkg, ring, f, tr.reduce_t(t), [D.cmod(v, t) for v in slot_values])
    ct = D.DBFVBatch(K.bfv_enc(ring, base, t, ptxt, sk, rng, la), base, 1, "crt", 1)
    res = B.bootstrap_bfv_dbfv(ring, f, tr, modulus, ct, rk, gks, encaps)
    got = K.dbfv_dec(ring, modulus, res, sk)
    ring.close()
    return got, [D.cmod(v, t) for v in ptxt]


def _pack(pkg, ring, f, tr, slots):
    """pack_slot_values (homcrt.rs:276-293) on the device"""
    K, D, H = pkg.keys, pkg.dbfv, pkg.homcrt
    base = ring.initial_rns_base()
    el = K.encode_pb_unscaled(ring, base, tr.t, [x for _ in range(f.deg_lprime) for x in slots])
    ring.pb_to_crt_tensor(el, base)
    for gen, cs in tr.inverse():
        el = H.apply_cyclic_transform_plain(ring, el, base, gen, cs)
    ring.crt_to_pb_tensor(el, base)
    s = f.scale_m_k_to_m_period_basis()
    return [D.cmod(x * s, tr.t) for x in K.decode_pb_unscaled(ring, tr.t, el, base)]


def test_bootstrap_field_18(gpu):
    """bootstrap/mod.rs:409-424: 2025 survives the bootstrap"""
    got, want = _bootstrap_case(gpu, "field_18", 30, 4)
    assert got == want


def test_bootstrap_field_32(gpu):
    """bootstrap/mod.rs:429-455: two packed slot values survive the bootstrap"""
    got, want = _bootstrap_case(gpu, "field_32", 30, 2, slot_values=[2025, -777])
    assert got == want


def test_bootstrap_batch_and_a_wrong_key(gpu):
    """three ciphertexts bootstrapped as one batch decrypt to their own values; with the
    encapsulation of a different main key the decryption is wrong (the check is sensitive to the
    homomorphic decryption step)"""
    pkg, ring, f, rng, base, la, sk = _setup(gpu, "field_18", (800, 820), 20, 40)
    K, D, H, B = pkg.keys, pkg.dbfv, pkg.homcrt, pkg.bootstrap
    t = 1 << 15
    modulus = D.DBFVModulus.new(5, 3)
    inter = D.DBFVModulus.new_incomplete(modulus.log2_b, 20)
    tr = H.UnsplitCRTTransform(inter.t(), f)
    encaps, _ = B.create_sparse_key_encaps(ring, f, inter, sk, 4, rng, la)
    digits = pkg.bfv.select_digits(len(base), len(base))
    gks = {g: K.gen_gk(ring, sk, g, digits, rng, la) for g in H.required_galois_elements("field_18")}
    rk = K.gen_rk(ring, sk, digits, rng, la)
    values = [2025, -1, 77]
    n = len(values)
    cts = [K.bfv_enc(ring, base, t, K.encode_scalar(ring, v), sk, rng, la) for v in values]
    batch = ring.allocate_uninit(2 * n, base)
    for c, x in enumerate(cts):
        ring.backend.hal_assign(batch[c:c + 1], x[0:1])
        ring.backend.hal_assign(batch[n + c:n + c + 1], x[1:2])
    res = B.bootstrap_bfv_dbfv(ring, f, tr, modulus, D.DBFVBatch(batch, base, n, "crt", 1), rk, gks, encaps)
    for c, v in enumerate(values):
        assert K.decode_scalar(ring, t, K.dbfv_dec(ring, modulus, res, sk, c)) == D.cmod(v, t), c
    other = K.gen_sk(ring, rng)
    wrong, _ = B.create_sparse_key_encaps(ring, f, inter, other, 4, rng, la)
    res2 = B.bootstrap_bfv_dbfv(ring, f, tr, modulus, D.DBFVBatch(batch, base, n, "crt", 1), rk, gks, wrong)
    assert K.decode_scalar(ring, t, K.dbfv_dec(ring, modulus, res2, sk, 0)) != D.cmod(2025, t)
    ring.close()


@pytest.mark.parametrize("nprimes,log2_t", [(3, 76), (1, 20), (2, 64), (4, 126)])
def test_device_decode_and_encode_match_the_host(gpu, nprimes, log2_t):
    """csd_ring_decode_scaled equals decode_pb_scaled (bfv.rs:262-274, exact host big integers) on
    uniform residues and on lifts within +-3 of a rounding boundary (t x / Q = k + 1/2 exactly is
    impossible, these are the nearest ones); csd_ring_encode_digits gives element_from_pb_coeffs'
    representatives mont_repr((c rem p) as i64, p) (ring.rs:109, 1449-1461) of the values and of
    their balanced digit vectors (dbfv.rs:97-112), and the digit count balanced_digits_vec keeps"""
    from oracle import ring as R

    pkg, ring, f, rng, base, la, sk = _setup(gpu, "field_18", (300, 320), 16, 11)
    K, D = pkg.keys, pkg.dbfv
    be = ring.backend
    sub = base[:nprimes]
    primes = [ring.primes[i] for i in sub]
    q = ring.modulus_of(sub)
    t = 1 << log2_t
    n = f.deg_l
    gen = np.random.default_rng(nprimes)
    lifts = [int(x) for x in gen.integers(0, 1 << 62, n)]
    lifts = [(x * (q >> 40) + x) % q for x in lifts]
    ks = list(range(0, n, 3))
    for idx, k in enumerate(ks):  # near (k + 1/2) Q / t
        center = ((2 * (k + 1)) * q) // (2 * t) + (q // (2 * t))
        lifts[k] = (center + (idx % 7) - 3) % q
    host = np.zeros((2, nprimes, f.deg_lprime, f.deg_k), dtype=np.int64)
    for j, p in enumerate(primes):
        host[0, j] = np.array([R.mont_repr(x % p, p) for x in lifts], dtype=np.int64).reshape(f.deg_lprime, f.deg_k)
    host[1] = host[0]
    src = ring.allocate_uninit(2, sub)
    be.hal_import(src, host)
    got = be.export_i128(ring.decode_scaled(src, sub, log2_t))
    got = [(int(h) << 64) + (int(lo) & ((1 << 64) - 1)) for lo, h in zip(got["lo"], got["hi"])]
    want = K.decode_pb_scaled(ring, t, src[0:1], sub)
    exact = [D.cmod(D.rounded_div((x - q if x > q // 2 else x) * t, q), t) for x in lifts]
    assert want == exact
    assert got[:n] == want and got[n:] == want
    # encoding over the whole base, of the values and of their balanced digits (base 2^8)
    vals = ring.decode_scaled(src, sub, log2_t)
    if log2_t <= 100:
        enc, kept = ring.encode_digits(vals[0:1], base)
        assert kept == 1
        e = be.hal_export(enc).reshape(1, len(base), f.deg_lprime, f.deg_k)
        for j, i in enumerate(base):
            p = ring.primes[i]
            assert [int(v) for v in e[0, j].reshape(-1)] == [R.mont_repr(R._trunc_mod(c, p), p) for c in want]
    dmax = (log2_t + 7) // 8
    planes, kept = ring.encode_digits(vals, base, 8, dmax)
    digs = pkg.homcrt.balanced_digits_vec(want, 256, dmax)
    assert kept == len(digs)
    e = be.hal_export(planes).reshape(dmax, 2, len(base), f.deg_lprime, f.deg_k)
    for j in range(dmax):
        dv = digs[j] if j < len(digs) else [0] * n
        for jj, i in enumerate(base[:3]):
            p = ring.primes[i]
            assert [int(v) for v in e[j, 1, jj].reshape(-1)] == [R.mont_repr(d, p) for d in dv], (j, jj)
    ring.close()


def test_bootstrap_field_36864(gpu):
    """the reference's full-size test (bootstrap/mod.rs:461-509, ignored there for its CPU time):
    field_36864, 17 primes, (8, 8) -> intermediate 2^76, sparse weight 32 with amplification at
    most v - 2 = 10, 27 Galois keys of 17 digits, a relinearization key of 9 digits: every one of
    the 3072 slots of the bootstrapped encryption of 2025 holds 2025"""
    pkg = load()
    K, D, H, B = pkg.keys, pkg.dbfv, pkg.homcrt, pkg.bootstrap
    ring = pkg.ring.RingContext(gpu, "field_36864", (960, 980), 76)
    f = pkg.field.FieldDescriptor("field_36864")
    rng = K.Rng(31)
    base = ring.initial_rns_base()
    la = 960.0
    modulus = D.DBFVModulus.new(8, 8)
    inter = D.DBFVModulus.new_incomplete(8, 76)
    sk = K.gen_sk(ring, rng)
    tr = H.UnsplitCRTTransform(inter.t(), f)
    encaps, _ = B.create_sparse_key_encaps(ring, f, inter, sk, 32, rng, la, max_log2_amplification=10.0)
    gdig = pkg.bfv.select_digits(len(base), len(base))
    gks = {g: K.gen_gk(ring, sk, g, gdig, rng, la) for g in H.required_galois_elements("field_36864")}
    assert len(gks) == 27
    rk = K.gen_rk(ring, sk, pkg.bfv.select_digits((len(base) - 1) // 2 + 1, len(base)), rng, la)
    ct = D.DBFVBatch(K.bfv_enc(ring, base, modulus.t(), K.encode_scalar(ring, 2025), sk, rng, la), base, 1, "crt", 1)
    res = B.bootstrap_bfv_dbfv(ring, f, tr, modulus, ct, rk, gks, encaps)
    dec = K.dbfv_dec(ring, modulus, res, sk)
    t = modulus.t()
    el = K.encode_pb_unscaled(ring, base, t, dec)
    ring.pb_to_crt_tensor(el, base)
    for gen, cs in tr.reduce_t(t).forward():
        el = H.apply_cyclic_transform_plain(ring, el, base, gen, cs)
    ring.crt_to_pb_tensor(el, base)
    s = f.scale_m_k_to_m_period_basis()
    slots = [D.cmod(v * s, t) for v in K.decode_pb_unscaled(ring, t, el, base)[:f.deg_k]]
    assert slots == [2025] * f.deg_k
    ring.close()
