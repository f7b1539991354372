"""Pins the oracle ring layer (oracle/ring.py) against the reference's ring/field tests
(ring.rs:1854-2212, field.rs:775-970) and the derived constants of the benchmark fields."""
import numpy as np
import pytest

from oracle import ring as R
from oracle.cpu_backend import i128_to_ints

BCONV_PRIMES = [
    144115188075855859, 144115188075855847, 144115188075855823, 144115188075855811, 144115188075855803,
    144115188075855761, 144115188075855677, 144115188075855599, 144115188075855509, 144115188075855449,
    144115188075855439, 144115188075855421, 144115188075855413,
]


@pytest.mark.parametrize("name,m,deg_l,deg_k,deg_lp,gal_ml", [
    ("field_36864", 8536255, 36864, 3072, 12, 144),
    ("field_65536", 5570645, 65536, 16384, 4, 64),
    ("field_49152_ex", 10158235, 49152, 12288, 4, 160),
    ("field_192", 90627, 192, 96, 2, 296),
    ("field_32", 255, 32, 2, 16, 4),
    ("field_32768_ex", 3342387, 32768, 4096, 8, 64),
    ("field_18", 1729, 18, 6, 3, 72),
    ("field_16", 257, 16, 16, 1, None),
    ("field_2", 7, 2, 2, 1, None),
    ("field_49152", 35297665, 49152, 12288, 4, 528),
])
def test_field_descriptors(name, m, deg_l, deg_k, deg_lp, gal_ml):
    """field.rs:775-900"""
    f = R.FieldDescriptor(name)
    assert (f.m, f.deg_l, f.deg_k, f.deg_lprime) == (m, deg_l, deg_k, deg_lp)
    if gal_ml is not None:
        assert len(f.gal_ml) == gal_ml


def test_gal_to_shift_roundtrip():
    """field.rs:932-950"""
    f = R.FieldDescriptor("field_18")
    for i, g in enumerate(f.lq_transversal()):
        assert f.coset_index(g) == i
        idx = 0
        for s, (_, n) in zip(f.gal_to_shift(g), f.lq_cyclic):
            idx = idx * n + s
        assert idx == i
    for h in f.gal_ml:
        assert f.coset_index(h) == 0


def test_crt2():
    """field.rs:964-969"""
    assert R.crt2(3, 5, 2, 7) == 23
    assert R.crt2(1, 5, 1, 7) == 1
    assert R.crt2(1, 1, 2, 7) == 2


def test_rou_orders_and_primes():
    """The derived prime sets of the benchmark configurations: p = 1 mod rou_order, the largest
    such primes below 2^57, in the order the sampler yields them."""
    f2 = R.FieldDescriptor("field_2")
    assert f2.rou_order() == 14
    primes = R.extend_sampled_primes([], 200, 230, 57, 14)
    assert primes == [144115188075855823, 144115188075855599, 144115188075854843, 144115188075854689]
    f = R.FieldDescriptor("field_36864")
    assert f.rou_order() == 1638960960
    base = R.extend_sampled_primes([], 960, 980, 57, f.rou_order())
    assert len(base) == 17
    assert base[:2] == [144115176243904321, 144115172965982401]
    for p in base:
        assert R.is_prime(p) and (p - 1) % f.rou_order() == 0 and p < 1 << 57
    # the sampler yields strictly decreasing primes, each the largest admissible one below the last
    for a, b in zip(base, base[1:]):
        assert b < a
        for c in range(b + f.rou_order(), a, f.rou_order()):
            assert not R.is_prime(c)


def test_bconv_kat(cpu):
    """ring.rs:1859-1917: the fixed 13-prime known-answer test."""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_2"), raw_primes=BCONV_PRIMES)
    frm, to = BCONV_PRIMES[:7], BCONV_PRIMES[7:]
    src = cpu.hal_allocate_i64([2, 7, 8, 1])
    dst = cpu.hal_allocate_i64([2, 6, 8, 1])
    d1 = [0, 1, 2, 3, 4, 5, 6, 7]
    d2 = [1 << 64, 3 << 100, -7 << 100, 9 << 120, -1 << 120, 5 << 120, 0, 7 << 120]
    vec = []
    for data in (d1, d2):
        for p in frm:
            for x in data:
                vec.append(R.montgomery_repr(R._trunc_mod(x, p), p))
    cpu.hal_import(src, vec)
    ring.bconv_tensors(dst, list(range(7, 13)), src, list(range(7)))
    out = cpu.hal_export(dst)
    for b, data in enumerate((d1, d2)):
        for i, p in enumerate(to):
            for j, x in enumerate(data):
                assert int(out[b * 48 + i * 8 + j]) % p == R.montgomery_repr(x, p)


def test_bconv_gadget(cpu):
    """ring.rs:1919-1972: batched gadget conversion == per-digit conversion."""
    primes = BCONV_PRIMES[:7]
    ring = R.RingContext(cpu, R.FieldDescriptor("field_2"), raw_primes=primes)
    base = list(range(7))
    digits = [(0, 3), (3, 5), (5, 7)]
    data = [0, 1, -1 << 100, 3 << 100, -7 << 90, 9 << 120, 12345, 7 << 120]
    vec = [R.montgomery_repr(R._trunc_mod(x, p), p) for p in primes for x in data]
    src = cpu.hal_allocate_i64([1, 7, 8, 1])
    cpu.hal_import(src, vec)
    dst = cpu.hal_allocate_i64([3, 7, 8, 1])
    ring.bconv_gadget_tensors(dst, src, base, digits)
    out = cpu.hal_export(dst)
    for i, (a, b) in enumerate(digits):
        ds = cpu.hal_allocate_i64([1, b - a, 8, 1])
        cpu.hal_import(ds, vec[a * 8:b * 8])
        exp = cpu.hal_allocate_i64([1, 7, 8, 1])
        ring.bconv_tensors(exp, base, ds, base[a:b])
        ev = cpu.hal_export(exp)
        for pos in range(56):
            p = primes[pos // 8]
            assert int(ev[pos]) % p == int(out[i * 56 + pos]) % p


def _ring18(cpu):
    return R.RingContext(cpu, R.FieldDescriptor("field_18"), (100, 120), 4)


def _square(ring, els, base):
    acc = ring.allocate_wide(1, base)
    acc.fma(range(1), els, 1, els, 1)
    return ring.reduce_wide(acc)


def test_pb_crt_period_basis_relation(cpu):
    """ring.rs:1974-2012: the field_18 KAT (computed with sage in the reference)."""
    ring = _ring18(cpu)
    base = [0]
    p = ring.primes[0]
    els = ring.allocate(1, base)
    data = [0] * 18
    for blk in range(3):
        data[blk * 6] = R.mont_repr(1, p)
    cpu.hal_import(els, data)
    ring.pb_to_crt_tensor(els, base)
    sqr = _square(ring, els, base)
    ring.crt_to_pb_tensor(sqr, base)
    out = cpu.hal_export(sqr)
    exp = [3, 4, 2, -2, -4, -2]
    for blk in range(3):
        for j in range(6):
            assert int(out[blk * 6 + j]) % p == R.montgomery_repr(exp[j], p)


def test_pb_crt_roundtrip(cpu):
    """ring.rs:2014-2035"""
    ring = _ring18(cpu)
    base = ring.initial_rns_base()
    els = ring.allocate(2, base)
    state = 123
    data = []
    for _ in range(2 * len(base) * 18):
        state = (state * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
        data.append(R._trunc_mod(state >> 12, 1 << 63) % 1000000)
    cpu.hal_import(els, data)
    ring.pb_to_crt_tensor(els, base)
    ring.crt_to_pb_tensor(els, base)
    out = cpu.hal_export(els)
    for pos, (e, a) in enumerate(zip(data, out)):
        p = ring.primes[base[(pos // 18) % len(base)]]
        assert (e - int(a)) % p == 0


def test_crt_transform_multiply_small(cpu):
    """ring.rs:2037-2067: the square of a period is bounded by m."""
    ring = _ring18(cpu)
    base = [0]
    p = ring.primes[0]
    els = ring.allocate(1, base)
    data = [0] * 18
    data[0] = R.mont_repr(1, p)
    cpu.hal_import(els, data)
    ring.pb_to_crt_tensor(els, base)
    sqr = _square(ring, els, base)
    ring.crt_to_pb_tensor(sqr, base)
    for x in cpu.hal_export(sqr):
        plain = R.from_mont_repr(int(x), p)
        assert min(plain, p - plain) <= ring.field.m


def test_wide_freshness(cpu):
    """ring.rs:2117-2169"""
    ring = _ring18(cpu)
    base = ring.initial_rns_base()
    r = len(base)
    els = ring.allocate(4, base)
    values = [(pos % 50) + 1 for pos in range(4 * r * 18)]
    cpu.hal_import(els, [R.mont_repr(v, ring.primes[base[(pos // 18) % r]]) for pos, v in enumerate(values)])
    acc = ring.allocate_wide(4, base)
    acc.add(range(1, 3), els[1:3], 1)
    acc.add(range(0, 4), els, 1)
    for _ in range(100):
        acc.add(range(0, 1), els[0:1], 1)
    assert acc.needs_full_reduction(0) and not acc.needs_full_reduction(1)
    fin = ring.reduce_wide(acc)
    out = cpu.hal_export(fin)
    for pos in range(4 * r * 18):
        p = ring.primes[base[(pos // 18) % r]]
        assert abs(int(out[pos])) <= p
        mult = {0: 101, 1: 2, 2: 2}.get(pos // (r * 18), 1)
        assert (mult * values[pos]) % p == R.from_mont_repr(int(out[pos]), p)
    acc = ring.allocate_wide(2, base)
    assert (ring.reduce_wide(acc) == 0).all()


def test_apply_galois(cpu):
    """ring.rs:2171-2212"""
    ring = _ring18(cpu)
    base = [0, 1]
    els = ring.allocate(1, base)
    data = list(range(36))
    cpu.hal_import(els, data)
    g = ring.field.lq_cyclic[1][0]
    dst = ring.allocate(1, base)
    ring.apply_galois(dst, els, g)
    out = cpu.hal_export(dst)
    for j in range(2):
        for a0 in range(3):
            for a1 in range(2):
                for a2 in range(3):
                    assert data[j * 18 + (a0 * 2 + a1) * 3 + a2] == out[j * 18 + (a0 * 2 + (a1 + 1) % 2) * 3 + a2]
    g1, g2 = ring.field.lq_cyclic[0][0], ring.field.lq_cyclic[2][0]
    tmp = ring.allocate(1, base)
    ring.apply_galois(tmp, els, g1)
    ring.apply_galois(dst, tmp, g2)
    exp = ring.allocate(1, base)
    ring.apply_galois(exp, els, g1 * g2 % ring.field.m)
    assert (dst == exp).all()


def test_mod_switch_down_rounds(cpu):
    """bfv.rs:628-662 against the closed form round(Q' x / Q) on small lifts (field_2)."""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_2"), (200, 230), 0)
    base = ring.initial_rns_base()
    P = [ring.primes[i] for i in base]
    Q = ring.modulus_of(base)
    Qk = ring.modulus_of(base[:3])
    xs = [0, 1, -1, Q // 3, -(Q // 5), 12345678901234567890123, Q // 2 - 7, -(Q // 2) + 9]
    els = ring.allocate(4, base)
    vals = []
    for e in range(4):
        for p in P:
            for t in range(2):
                vals.append(R.mont_repr(xs[e * 2 + t] % p, p))
    cpu.hal_import(els, vals)
    out = R.mod_switch_down(ring, els, base, 3)
    o = cpu.hal_export(out)
    for e in range(4):
        for jj, p in enumerate(P[:3]):
            for t in range(2):
                x = xs[e * 2 + t]
                got = R.from_mont_repr(int(o[(e * 3 + jj) * 2 + t]), p)
                # round(Qk x / Q) up to the bconv's +-1 multiple of the dropped modulus
                exact = (Qk * x * 2 + Q) // (2 * Q)
                assert any((exact + d) % p == got for d in (-1, 0, 1))


def test_kernel_values_are_traces(cpu):
    """The CRT kernel of field_2 = Tr_(M|K)(zeta_7^(g^-1)) mod p for the transversal [1, 6]:
    periods of Q(zeta_7)^+ are zeta + zeta^-1 style sums over Gal(M|K) = <2> = {1, 2, 4}."""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_2"), (200, 230), 0)
    p = ring.primes[0]
    plp = cpu.hal_create_ntt_plan(ring.moduli[0], [])
    pk = cpu.hal_create_ntt_plan(ring.moduli[0], [2])
    *_, k_lp, k_k = ring._rns_factor_kernels(p, plp, pk)
    assert k_lp == [1]
    cand = 2
    while True:
        z = pow(cand, (p - 1) // 7, p)
        if z != 1:
            break
        cand += 1
    # transversal [1, 6]: g^-1 = 1 and 6; subgroup {1, 2, 4}
    assert k_k == [(z + z ** 2 + z ** 4) % p, (pow(z, 6, p) + pow(z, 12, p) + pow(z, 24, p)) % p]


def _crt_centred(residues, primes):
    """the integer in (-Q/2, Q/2] with the given residues"""
    Q = 1
    for p in primes:
        Q *= p
    x = 0
    for v, p in zip(residues, primes):
        Qp = Q // p
        x += int(v) % p * Qp * pow(Qp % p, -1, p)
    x %= Q
    return x - Q if x > Q // 2 else x


def _encode(ring, xs, base, shape):
    """integers xs (one per coefficient) as a Montgomery-form period-basis tensor over base"""
    n = shape[2] * shape[3]
    out = np.zeros((1, len(base), shape[2], shape[3]), dtype=np.int64)
    for j, i in enumerate(base):
        p = ring.primes[i]
        out[0, j] = np.array([R.mont_repr(x % p, p) for x in xs], dtype=np.int64).reshape(shape[2], shape[3])
    return out


def _decode(ring, t, base):
    """a Montgomery-form tensor [1, r, ., .] over base -> centred integers per coefficient"""
    primes = [ring.primes[i] for i in base]
    flat = t.reshape(len(base), -1)
    return [_crt_centred([R.from_mont_repr(int(flat[j, k]), primes[j]) for j in range(len(base))], primes)
            for k in range(flat.shape[1])]


def test_lift_to_extended_is_a_lift(cpu):
    """lift_to_extended (bfv.rs:564-595): over base + extension the element represents the
    centred lift of x mod q plus at most one multiple of q (the BEHZ conversion's epsilon)."""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_18"), (100, 120), 4)
    base = ring.initial_rns_base()
    q = ring.modulus_of(base)
    rng = np.random.default_rng(3)
    shape = (1, len(base), 3, 6)
    xs = [int(rng.integers(-(1 << 62), 1 << 62)) * int(rng.integers(1, 1 << 40)) % q for _ in range(18)]
    xs = [x - q if x > q // 2 else x for x in xs]
    lifted, lb = R.lift_to_extended(ring, _encode(ring, xs, base, shape), base)
    ring.crt_to_pb_tensor(lifted, lb)
    got = _decode(ring, lifted, lb)
    for x, y in zip(xs, got):
        assert y - x in (-q, 0, q)


def test_rescale_to_base_rounds(cpu):
    """rescale_to_base (bfv.rs:601-622): for X = A B with |A|, |B| <= q/2 (a tensor-product
    coefficient) over base + extension, the result is round(t X / q) mod q up to +-1."""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_18"), (100, 120), 4)
    base = ring.initial_rns_base()
    ext = ring.extension_rns_base_for(base)
    lifted = base + ext
    q = ring.modulus_of(base)
    t = 16
    rng = np.random.default_rng(4)
    xs = []
    for _ in range(18):
        a = int(rng.integers(0, 1 << 62)) * int(rng.integers(0, 1 << 62)) % q - q // 2
        b = int(rng.integers(0, 1 << 62)) * int(rng.integers(0, 1 << 62)) % q - q // 2
        xs.append(a * b)
    x = _encode(ring, xs, lifted, (1, len(lifted), 3, 6))
    out = R.rescale_to_base(ring, x, lifted, len(base), t)
    got = _decode(ring, out, base)
    for X, y in zip(xs, got):
        want = (2 * t * X + q) // (2 * q)
        d = (y - want) % q
        assert min(d, q - d) <= 1


def test_gadget_decompose_recomposes(cpu):
    """gadget_decompose (bfv.rs:408-426): every digit is x mod D_i up to a multiple of D_i, small,
    and sum_i d_i (D/D_i) ((D/D_i)^-1 mod D_i) = x mod D."""
    ring = R.RingContext(cpu, R.FieldDescriptor("field_18"), (160, 180), 4)
    base = ring.initial_rns_base()
    digits = R.select_digits(2, len(base))
    D = ring.modulus_of(base)
    rng = np.random.default_rng(5)
    xs = [int(rng.integers(0, 1 << 62)) * int(rng.integers(0, 1 << 62)) * int(rng.integers(0, 1 << 62)) % D
          for _ in range(18)]
    parts = R.gadget_decompose(ring, _encode(ring, xs, base, (1, len(base), 3, 6)), base, digits)
    ring.crt_to_pb_tensor(parts, base)
    dec = [_decode(ring, parts[i:i + 1], base) for i in range(len(digits))]
    for k, x in enumerate(xs):
        total = 0
        for i, (a, b) in enumerate(digits):
            Di = ring.modulus_of(base[a:b])
            d = dec[i][k]
            assert (d - x) % Di == 0
            assert abs(d) <= 2 * Di
            g = (D // Di) * pow((D // Di) % Di, -1, Di)
            total += d * g
        assert (total - x) % D == 0


def _uniform_crt(ring, count, base, rng):
    t = ring.allocate(count, base)
    for j, i in enumerate(base):
        t[:, j] = rng.integers(0, ring.primes[i], size=t[:, j].shape, dtype=np.int64)
    return t


def _zero_secret_galois_keys(ring, base, ndigits):
    """Galois key of the zero secret key: for gadget digit i the rows (0, g_i) in CRT basis with
    g_i = (Q/D_i) ((Q/D_i)^-1 mod D_i), the CRT recomposition factor of the digit, so that
    sum_i d_i g_i = c1 mod Q for the gadget digits d_i of c1."""
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


def _galois_identity_failures(ring, ct, base, g, out):
    """Under the zero secret key's Galois key, one rotation of bfv_hom_gal (bfv.rs:816-839) maps
    (c0, c1) to (sigma_g(c0), sigma_g(c1)) mod Q: the key-switch part sum_i sigma_g(d_i) (0, g_i)
    is (0, sigma_g(sum_i d_i g_i)) = (0, sigma_g(c1)), and sigma_g(c0) is added to row 0.
    sigma_g is the oracle's apply_galois, pinned by ring.rs:2171-2212. Returns the rows whose
    residues differ mod p in some limb."""
    want = ring.allocate(2, base)
    ring.apply_galois(want, ct, g)
    bad = []
    for row in range(2):
        for j, i in enumerate(base):
            p = ring.primes[i]
            d = (out[row, j].astype(object) - want[row, j].astype(object)) % p
            if any(int(v) != 0 for v in d.reshape(-1)):
                bad.append(row)
                break
    return bad


@pytest.mark.parametrize("field,bits,ndigits", [("field_18", (160, 180), 2), ("field_192", (200, 240), 3)])
def test_hom_gal_one_zero_secret_identity(cpu, field, bits, ndigits):
    """Pins the oracle's hom_gal_one (the checker of the device Galois key switch): with the zero
    secret key it must give (sigma_g(c0), sigma_g(c1)) mod Q for every rotation of the homcrt
    set. The same check rejects the planted mistakes: c0 added unpermuted, the key rows swapped,
    and gadget digits paired with the wrong key digit."""
    ring = R.RingContext(cpu, R.FieldDescriptor(field), bits, 4)
    base = ring.initial_rns_base()
    assert len(base) >= ndigits
    rng = np.random.default_rng(11)
    keys, digits = _zero_secret_galois_keys(ring, base, ndigits)
    m = ring.field.m
    gens = [g for g, _ in ring.field.lq_cyclic]
    gs = [gens[0], gens[-1], pow(gens[-1], 3, m), gens[0] * gens[-1] % m]
    for trial, g in enumerate(gs):
        ct = _uniform_crt(ring, 2, base, rng)
        c1 = ct[1:2].copy()
        ring.crt_to_pb_tensor(c1, base)
        parts = R.gadget_decompose(ring, c1, base, digits)
        out = R.hom_gal_one(ring, ct, base, g, parts, keys, digits)
        assert _galois_identity_failures(ring, ct, base, g, out) == [], (field, g)
        if trial:
            continue
        # planted: c0 added without the automorphism
        gal_parts = ring.allocate(parts.shape[0], base)
        ring.apply_galois(gal_parts, parts, g)
        acc = ring.allocate_wide(2, base)
        R.key_switch_inner_product(ring, acc, gal_parts, keys, digits)
        acc.add(range(0, 1), ct[0:1], 1)
        assert 0 in _galois_identity_failures(ring, ct, base, g, ring.reduce_wide(acc))
        # planted: key rows swapped
        swapped = [np.ascontiguousarray(k[::-1]) for k in keys]
        out_sw = R.hom_gal_one(ring, ct, base, g, parts, swapped, digits)
        assert _galois_identity_failures(ring, ct, base, g, out_sw) != []
        # planted: digits crossed with the next digit's key
        crossed = keys[1:] + keys[:1]
        out_cr = R.hom_gal_one(ring, ct, base, g, parts, crossed, digits)
        assert 1 in _galois_identity_failures(ring, ct, base, g, out_cr)
