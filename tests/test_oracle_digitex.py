"""Pins the oracle's dBFV digit extraction (oracle/digitex.py, bootstrap/digitex.rs) and its dBFV
accumulators (oracle/dbfv.py, dbfv.rs:121-533) on the CPU.

The reference checks hom_floordiv_dbfv only through decryption (digitex.rs:372-482). Key generation
and encryption under a nonzero secret are not restated here, so the check runs under the zero
secret key: the relinearization key has rows (0, g_i), g_i the CRT recomposition factor of gadget
digit i, so relin maps (c0, c1, c2) to (c0, c1 + c2) and every operation acts on the c0 rows
alone, as on Delta-scaled plaintexts. Decrypting with s = 0 (dbfv_dec, dbfv.rs:656-680: round(t c0
/ Q) per digit, recombined) must then give the floor division the reference's tests expect, for
their values and moduli. The same check rejects planted mistakes (a wrong circuit constant, a
dropped subtraction, a wrong exact division)."""
import numpy as np
import pytest

from oracle import dbfv as D
from oracle import digitex as X
from oracle import ring as R


def zero_secret_key(ring, base, ndigits):
    digits = R.select_digits(ndigits, len(base))
    q = ring.modulus_of(base)
    keys = []
    for a, b in digits:
        di = ring.modulus_of(base[a:b])
        g = (q // di) * pow((q // di) % di, -1, di) % q
        key = ring.allocate(2, base)
        for j, i in enumerate(base):
            key[1, j] = R.montgomery_repr(g % ring.primes[i], ring.primes[i])
        keys.append(key)
    return keys, digits


def encrypt_scalar(ring, base, t, m, rng, noise=3):
    """a period-basis BFV ciphertext of the scalar m under s = 0: c0 = Delta encode_scalar(m) + e
    with |e| <= noise per coefficient, c1 uniform (it never reaches the decryption)"""
    n2, n3 = ring.field.deg_lprime, ring.field.deg_k
    q = ring.modulus_of(base)
    delta = D.rounded_div(q, t)
    val = D.cmod(m * ring.field.scale_q_to_m_period_basis(), t) * delta
    errs = rng.integers(-noise, noise + 1, size=n2 * n3)
    ct = ring.allocate(2, base)
    for j, i in enumerate(base):
        p = ring.primes[i]
        ct[0, j] = np.array([R.mont_repr((val + int(e)) % p, p) for e in errs], dtype=np.int64).reshape(n2, n3)
        ct[1, j] = rng.integers(0, p, size=(n2, n3), dtype=np.int64)
    return D.Ct(ct, base, "pb", 1)


def _lifts(ring, rows, base):
    primes = [ring.primes[i] for i in base]
    q = ring.modulus_of(base)
    out = []
    flat = rows.reshape(len(base), -1)
    for k in range(flat.shape[1]):
        x = 0
        for j, p in enumerate(primes):
            v = R.from_mont_repr(int(flat[j, k]), p) % p
            m = q // p
            x += v * m * pow(m % p, -1, p)
        x %= q
        out.append(x - q if x > q // 2 else x)
    return out


def decrypt_scalar(ring, modulus, ct):
    """dbfv_dec (dbfv.rs:656-680) under s = 0 and decode_scalar (bfv.rs:870-878)"""
    q = ring.modulus_of(ct.base)
    t = modulus.t()
    total = None
    for i in range(ct.digits()):
        lifts = _lifts(ring, ct.data[2 * i], ct.base)
        part = [D.cmod(D.rounded_div(x * t, q), t) for x in lifts]
        w = 1 << (modulus.log2_t - i * modulus.log2_b)
        contrib = [D.cmod(p_, w) << (i * modulus.log2_b) for p_ in part]
        total = contrib if total is None else [a + b for a, b in zip(total, contrib)]
    vals = [R.cmod_pow2(x, modulus.log2_t) for x in total]
    if len(set(vals)) != 1:
        return None  # not a scalar
    return D.cmod(vals[0] * ring.field.scale_q_to_m_period_basis(), t)


def floordiv(ring, init, final, m, seed):
    base = ring.initial_rns_base()
    keys, digits = zero_secret_key(ring, base, len(base))
    ct = encrypt_scalar(ring, base, init.t(), m, np.random.default_rng(seed))
    res = X.hom_floordiv_dbfv(ring, init, final, ct, keys, digits)
    return decrypt_scalar(ring, final, res)


@pytest.fixture(scope="module")
def ring192(cpu):
    return R.RingContext(cpu, R.FieldDescriptor("field_192"), (800, 820), 16)


def test_balanced_digits_and_cmod():
    """balanced digits recompose the scalar mod b^digits, each in (-b/2, b/2]"""
    assert D.cmod(7, 16) == 7 and D.cmod(8, 16) == 8 and D.cmod(9, 16) == -7 and D.cmod(-8, 16) == 8
    for x in [0, 1, -1, 2025, -777, 12656, 33720, -32106558060291480110540, 403089028126664267137012]:
        for log2_b, digits in [(4, 4), (8, 5), (6, 8)]:
            b = 1 << log2_b
            ds = D.balanced_digits(x, b, digits)
            assert len(ds) <= digits and all(-b // 2 < d <= b // 2 for d in ds)
            assert (sum(d * b ** i for i, d in enumerate(ds)) - x) % (b ** digits) == 0


def test_floordiv_schedule_matches_reference_noise_model():
    """digitex.rs:286-312 for the bootstrap's field_49152 floor division (8, 76) -> (8, 64) on 23
    primes: 8 multiplications per level, each estimated at 8 + (8 + 8 + 10) bits of noise, so
    the base shrinks by floor(34 j / 57) within a level"""
    f = R.FieldDescriptor("field_49152")
    v, mults, lens = X.floordiv_schedule(f, D.DBFVModulus(8, 76), D.DBFVModulus(8, 64), 23)
    assert (v, mults) == (12, 8)
    assert lens[0] == [23 - (34 * j) // 57 for j in range(9)]
    assert all(lens[i][0] == lens[i - 1][1] for i in range(1, v))
    assert lens[-1][-1] == 23 - (34 * 11 + 34 * 8) // 57


def test_hom_floordiv_dbfv(ring192):
    """digitex.rs:378-397: 2025 with t = 2^16, (4, 4) -> (4, 3): 126"""
    assert floordiv(ring192, D.DBFVModulus.new(4, 4), D.DBFVModulus.new(4, 3), 2025, 1) == 126


@pytest.mark.parametrize("v", [5, 6, 9])
def test_hom_floordiv_dbfv_multiple_digits(cpu, v):
    """digitex.rs:456-481: v beyond one digit (exact divisions dropping several digits)"""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_192"), (900, 920), 16)
    final = D.DBFVModulus.new_incomplete(4, 16 - v)
    got = floordiv(ring, D.DBFVModulus.new(4, 4), final, 40529, 20 + v)
    assert got == D.cmod((40529 % (1 << 16)) >> v, final.t())


@pytest.mark.parametrize("v", [1, 3])
def test_hom_floordiv_dbfv_deep_circuit(cpu, v):
    """digitex.rs:399-423: precision 36 > 16 selects the full 8-multiplication circuit"""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_192"), (900, 920), 36)
    init = D.DBFVModulus.new_incomplete(8, 36)
    final = D.DBFVModulus.new_incomplete(8, 36 - v)
    value = 12345678901
    got = floordiv(ring, init, final, value, 40 + v)
    assert got == D.cmod((value % (1 << 36)) >> v, final.t())


def test_planted_mistakes_are_rejected(ring192, monkeypatch):
    """a wrong circuit constant, a dropped subtraction and a wrong exact division each change the
    decrypted floor division"""
    init, final = D.DBFVModulus.new(4, 4), D.DBFVModulus.new(4, 3)
    good = floordiv(ring192, init, final, 2025, 1)
    assert good == 126

    orig_ip = D.dbfv_acc_inner_product_scalar
    calls = {"n": 0}

    def wrong_constant(ring, acc, terms):
        calls["n"] += 1
        if calls["n"] == 1:  # the first combination of a level: 94 x2 + 121 x4 -> 95 x2 + 121 x4
            terms = [(terms[0][0], terms[0][1] + 1)] + list(terms[1:])
        return orig_ip(ring, acc, terms)

    monkeypatch.setattr(D, "dbfv_acc_inner_product_scalar", wrong_constant)
    assert floordiv(ring192, init, final, 2025, 1) != 126
    monkeypatch.setattr(D, "dbfv_acc_inner_product_scalar", orig_ip)

    orig_sub = X._sub_from
    seen = {"n": 0}

    def drop_first(ring, target, value):
        seen["n"] += 1
        if seen["n"] > 1:
            orig_sub(ring, target, value)

    monkeypatch.setattr(X, "_sub_from", drop_first)
    assert floordiv(ring192, init, final, 2025, 1) != 126
    monkeypatch.setattr(X, "_sub_from", orig_sub)

    # the exact division restructures the digits only for several digits, as for v = 5 (a
    # one-digit ciphertext decrypts to x / 2^v under t / 2^v unchanged)
    ring = R.RingContext(ring192.backend, R.FieldDescriptor("field_192"), (900, 920), 16)
    final5 = D.DBFVModulus.new_incomplete(4, 11)
    want = D.cmod(40529 >> 5, final5.t())
    assert floordiv(ring, init, final5, 40529, 25) == want
    orig_div = D.dbfv_exact_div_pow2
    monkeypatch.setattr(D, "dbfv_exact_div_pow2", lambda ring, m, val, k: orig_div(ring, m, val, max(k - 1, 0)))
    assert floordiv(ring, init, final5, 40529, 25) != want


def test_accumulator_digit_bookkeeping(ring192):
    """dbfv_acc_* (dbfv.rs:160-406): adding a 2-digit and subtracting a 1-digit ciphertext uses 2
    digits; an inner product with a scalar of two balanced digits reaches digit d + 1; the
    constant goes to row 0 only and leaves the digit count; finishing reduces the used rows"""
    ring = ring192
    base = ring.initial_rns_base()
    m = D.DBFVModulus.new(4, 4)
    rng = np.random.default_rng(3)
    a = encrypt_scalar(ring, base, m.t(), 5, rng)
    two = D.Ct(np.concatenate([a.data, a.data]), base, "pb", 1)
    acc = D.dbfv_allocate_accumulator(ring, m, base, "pb")
    D.dbfv_acc_add(ring, acc, two)
    D.dbfv_acc_sub(ring, acc, a)
    assert acc.used_digits == 2
    D.dbfv_acc_add_constant(ring, acc, 3)
    assert acc.used_digits == 2
    out = D.dbfv_finish_accumulator(ring, acc)
    assert out.digits() == 2
    # digit 0: a - a + 3 -> 3; digit 1: a -> 5 (weight 16)
    assert decrypt_scalar(ring, m, D.Ct(out.data[0:2], base, "pb", 1)) == 3
    acc = D.dbfv_allocate_accumulator(ring, m, base, "pb")
    D.dbfv_acc_inner_product_scalar(ring, acc, [(a, 17)])  # 17 = 1 + 1 * 16: two digits
    assert acc.used_digits == 2
    assert decrypt_scalar(ring, m, D.dbfv_finish_accumulator(ring, acc)) == 85
