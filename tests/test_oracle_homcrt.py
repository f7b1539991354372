"""Pins the oracle's homomorphic CRT-transform evaluation (oracle/homcrt.py, bootstrap/homcrt.rs:
188-429, dbfv.rs:193-450) on the CPU.

The reference checks hom_cyclic_transform_dbfv only through decryption of a full bootstrap
(homcrt.rs tests and bootstrap/mod.rs:419-555). Key generation under a nonzero secret is not
restated here, so the check runs under the zero secret key: each Galois key digit has rows
(0, g_i), g_i the CRT recomposition factor of gadget digit i, so bfv_hom_gal maps (c0, c1) to
(sigma_g(c0), sigma_g(c1)) mod Q exactly and the transform acts on the c0 rows as on
Delta-scaled plaintexts. Decrypting with s = 0 (round(t c0 / Q) per digit, recombined with the
digit weights as dbfv_dec does, dbfv.rs:699-722) must then give sum_i coeffs[i] sigma^i(m) mod t,
which the test forms separately with the plaintext transform (homcrt.rs:188-211) over a modulus
far above t. Planted mistakes in the baby-step/giant-step schedule fail that check."""
import numpy as np
import pytest

from oracle import dbfv as D
from oracle import homcrt as H
from oracle import ring as R
from test_oracle_digitex import _lifts, zero_secret_key


def test_bsgs_split_matches_the_reference_rotation_counts():
    """homcrt.rs:316-324: the splits of field_49152's factors are 4 x 1, 13 x 5 and 24 x 8
    (3 + 13 + 24 = 40 Galois keys, the bootstrap's set); a baby-step count is the first minimum
    of count + 3 ceil((len - 1) / count), checked against every alternative"""
    assert H.bsgs_split(4) == (4, 1)
    assert H.bsgs_split(64) == (13, 5)
    assert H.bsgs_split(192) == (24, 8)
    for n in range(2, 200):
        b, g = H.bsgs_split(n)
        assert b * g >= n and (b - 1) * g < n or g == 1
        cost = b + 3 * g
        assert all(c + 3 * ((n - 1) // c + 1) >= cost for c in range(1, n + 1))
        assert all(c + 3 * ((n - 1) // c + 1) > cost for c in range(1, b))


@pytest.mark.parametrize("field,jp", [("field_18", None), ("field_36864", [4, 6, 24, 64]),
                                      ("field_49152", [4, 64, 192])])
def test_factor_galois_elements_give_the_homcrt_key_set(field, jp):
    """required_galois_elements (homcrt.rs:336-349) from the per-prime data: the J_p orders of the
    bootstrap fields and the 40 keys of field_49152; every element is a unit of Z/mZ"""
    f = R.FieldDescriptor(field)
    facs = H.homcrt_factors(f)
    if jp is not None:
        assert [j for _, _, j in facs] == jp
    elems = sorted({g for _, gen, j in facs for g in H.factor_galois_elements(f.m, gen, j)})
    if field == "field_49152":
        assert len(elems) == 40
    for p, gen, j in facs:
        # gen generates (Z/pZ)* and is 1 mod m/p; gen^j lies in Gal(M|L) (J_p = I_p / H_p)
        assert gen % (f.m // p) == 1 and pow(gen, p - 1, f.m) == 1
        assert pow(gen, j, f.m) in set(f.gal_ml)


def _ring18(cpu):
    return R.RingContext(cpu, R.FieldDescriptor("field_18"), (300, 320), 16)


def test_lprime_element_tensor_is_the_crt_transform_of_the_element(cpu):
    """ring.rs:1227-1250: the CRT-basis values of an L'-element a = sum_u a_u eta'_u are those of
    the ring element whose period-basis coefficient (u, v) is a_u for every K position v (eta'_u is
    the sum of the L-periods over Gal(L|L')), constant along dim 3; the full transform is pinned by
    the reference's KATs"""
    ring = _ring18(cpu)
    f = ring.field
    base = ring.initial_rns_base()
    rng = np.random.default_rng(5)
    for coeffs in ([1] + [0] * (f.deg_lprime - 1), [int(x) for x in rng.integers(-1000, 1000, f.deg_lprime)],
                   [int(x) for x in rng.integers(-(1 << 40), 1 << 40, f.deg_lprime)]):
        t = H.lprime_element_tensor(ring, base, coeffs)
        full = ring.allocate(1, base)
        for j, i in enumerate(base):
            p = ring.primes[i]
            for u, c in enumerate(coeffs):
                full[0, j, u, :] = R.mont_repr(R._trunc_mod(c, p), p)
        ring.pb_to_crt_tensor(full, base)
        for j, i in enumerate(base):
            p = ring.primes[i]
            for u in range(f.deg_lprime):
                assert all((int(full[0, j, u, v]) - int(t[0, j, u, 0])) % p == 0 for v in range(f.deg_k))


def _encrypt(ring, base, modulus, x, rng, noise=2):
    """a CRT-basis dBFV encryption of the plaintext x (period-basis coefficients, balanced mod t)
    under s = 0: digit i has c0 = Delta x_i + e, x_i the balanced base-b digits of x, c1 uniform"""
    n2, n3 = ring.field.deg_lprime, ring.field.deg_k
    delta = D.rounded_div(ring.modulus_of(base), modulus.t())
    digits = H.balanced_digits_vec(list(x), modulus.b(), modulus.digits())
    ct = ring.allocate(2 * len(digits), base)
    for i, xi in enumerate(digits):
        errs = rng.integers(-noise, noise + 1, size=n2 * n3)
        for j, idx in enumerate(base):
            p = ring.primes[idx]
            ct[2 * i, j] = np.array([R.mont_repr(R._trunc_mod(delta * v + int(e), p), p) for v, e in zip(xi, errs)],
                                    dtype=np.int64).reshape(n2, n3)
            ct[2 * i + 1, j] = rng.integers(0, p, size=(n2, n3), dtype=np.int64)
    ring.pb_to_crt_tensor(ct, base)
    return D.Ct(ct, base, "crt", 1)


def _decrypt(ring, modulus, ct):
    """dbfv_dec (dbfv.rs:699-722) under s = 0: sum_k b^k round(t c0_k / Q), balanced mod t"""
    q = ring.modulus_of(ct.base)
    t = modulus.t()
    data = ct.data.copy()
    if ct.basis == "crt":
        ring.crt_to_pb_tensor(data, ct.base)
    total = [0] * (ring.field.deg_lprime * ring.field.deg_k)
    for k in range(ct.digits()):
        part = [D.rounded_div(v * t, q) for v in _lifts(ring, data[2 * k], ct.base)]
        total = [a + (b << (k * modulus.log2_b)) for a, b in zip(total, part)]
    return [D.cmod(v, t) for v in total]


def _plain_transform(ring, base, modulus, gen, coeffs, x):
    """sum_i coeffs[i] sigma^i(x) mod t by the plaintext transform (homcrt.rs:188-211) over the
    whole base (Q far above the integer result), lifted and reduced mod t"""
    n2, n3 = ring.field.deg_lprime, ring.field.deg_k
    el = ring.allocate(1, base)
    for j, idx in enumerate(base):
        p = ring.primes[idx]
        el[0, j] = np.array([R.mont_repr(R._trunc_mod(v, p), p) for v in x], dtype=np.int64).reshape(n2, n3)
    ring.pb_to_crt_tensor(el, base)
    out = H.apply_cyclic_transform_plain(ring, el, base, gen, coeffs)
    ring.crt_to_pb_tensor(out, base)
    return [D.cmod(v, modulus.t()) for v in _lifts(ring, out[0], base)]


def _galois_keys(ring, base, gs, ndigits):
    keys, digits = zero_secret_key(ring, base, ndigits)
    return {g: (keys, digits) for g in gs}


def _case(ring, modulus, gen, length, seed, nbase=None):
    f = ring.field
    base = ring.initial_rns_base()[:nbase] if nbase else ring.initial_rns_base()
    rng = np.random.default_rng(seed)
    t = modulus.t()
    x = [int(v) for v in rng.integers(-(t // 2) + 1, t // 2 + 1, f.deg_lprime * f.deg_k)]
    coeffs = [[int(v) for v in rng.integers(-(t // 2) + 1, t // 2 + 1, f.deg_lprime)] for _ in range(length)]
    gs = [H.pow_mod_m(gen, e, f.m) for e in range(1, length + 1)]
    gks = _galois_keys(ring, base, sorted(set(gs)), min(3, len(base)))
    ct = _encrypt(ring, base, modulus, x, rng)
    return base, x, coeffs, gks, ct


@pytest.mark.parametrize("length", [1, 2, 4, 6, 10, 20])
def test_hom_cyclic_transform_decrypts_to_the_plain_transform(cpu, length):
    """homcrt.rs:355-429 under the zero secret key: lengths with one block (1, 2, 4), two blocks
    (6 = 3 x 2, 10 = 5 x 2) and three (20 = 7 x 3), i.e. carried-over giant steps"""
    ring = _ring18(cpu)
    modulus = D.DBFVModulus.new(4, 4)
    gen = ring.field.lq_cyclic[-1][0]
    base, x, coeffs, gks, ct = _case(ring, modulus, gen, length, 100 + length)
    out = H.hom_cyclic_transform_dbfv(ring, modulus, gen, coeffs, ct, gks)
    assert out.basis == "crt" and out.digits() <= modulus.digits()
    assert _decrypt(ring, modulus, out) == _plain_transform(ring, base, modulus, gen, coeffs, x)


def test_hom_cyclic_transform_on_the_homcrt_factors(cpu):
    """the two factors of field_18 (p = 7, 13) with their own generators and lengths, composed as
    slots_to_coeffs_dbfv does (homcrt.rs:437-449), on a shorter base"""
    ring = _ring18(cpu)
    modulus = D.DBFVModulus.new_incomplete(4, 14)
    facs = H.homcrt_factors(ring.field)
    rng = np.random.default_rng(9)
    base = ring.initial_rns_base()
    f = ring.field
    t = modulus.t()
    x = [int(v) for v in rng.integers(-(t // 2) + 1, t // 2 + 1, f.deg_lprime * f.deg_k)]
    gs = sorted({g for _, gen, j in facs for g in H.factor_galois_elements(f.m, gen, j)})
    gks = _galois_keys(ring, base, gs, 4)
    ct = _encrypt(ring, base, modulus, x, rng)
    want = x
    for p, gen, j in facs:
        coeffs = [[int(v) for v in rng.integers(-(t // 2) + 1, t // 2 + 1, f.deg_lprime)] for _ in range(j)]
        ct = H.hom_cyclic_transform_dbfv(ring, modulus, gen, coeffs, ct, gks)
        want = _plain_transform(ring, base, modulus, gen, coeffs, want)
    assert _decrypt(ring, modulus, ct) == want


def test_planted_schedule_mistakes_are_rejected(cpu, monkeypatch):
    """a wrong giant-step exponent, coefficients shifted by one baby step and a dropped carry-over
    of the previous giant step each change the decryption (length 20: 7 baby x 3 giant steps)"""
    ring = _ring18(cpu)
    modulus = D.DBFVModulus.new(4, 4)
    gen = ring.field.lq_cyclic[-1][0]
    base, x, coeffs, gks, ct = _case(ring, modulus, gen, 20, 7)
    want = _plain_transform(ring, base, modulus, gen, coeffs, x)

    def run():
        return _decrypt(ring, modulus, H.hom_cyclic_transform_dbfv(ring, modulus, gen, coeffs, ct, gks))

    assert run() == want
    orig_pow = H.pow_mod_m
    monkeypatch.setattr(H, "pow_mod_m", lambda g, e, m: orig_pow(g, e + 1 if e == 7 else e, m))
    assert run() != want
    monkeypatch.setattr(H, "pow_mod_m", orig_pow)

    orig_ip = H.dbfv_acc_inner_product_element_lprime

    def shifted(ring_, acc, terms):
        cs = [c for _, c in terms]
        return orig_ip(ring_, acc, [(lhs, cs[(k + 1) % len(cs)]) for k, (lhs, _) in enumerate(terms)])

    monkeypatch.setattr(H, "dbfv_acc_inner_product_element_lprime", shifted)
    assert run() != want
    monkeypatch.setattr(H, "dbfv_acc_inner_product_element_lprime", orig_ip)

    orig_add = D.dbfv_acc_add
    monkeypatch.setattr(D, "dbfv_acc_add", lambda ring_, acc, c: None)
    assert run() != want
    monkeypatch.setattr(D, "dbfv_acc_add", orig_add)
    assert run() == want
