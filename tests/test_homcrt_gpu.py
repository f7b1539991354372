"""The homomorphic CRT transform (bootstrap/homcrt.rs:355-449) on the MI355X, batched over dBFV
ciphertexts and their digits, bit-exact (L0) against the oracle's per-ciphertext restatement
(oracle/homcrt.py, pinned under the zero secret key against the plaintext transform): the L'
element tensors, the row-addressed key switch against the plain batched one, the hoisted
bfv_hom_gal of every digit, and hom_cyclic_transform_dbfv at one-, two- and three-block lengths,
on field_18 and on field_36864's own factors at a reduced base, with uniform keys whose base is
longer than the ciphertexts'. Under the zero secret key the device result also decrypts to the
plaintext transform."""
import numpy as np
import pytest

from _pkg import load
from oracle import dbfv as OD
from oracle import homcrt as OH
from oracle import ring as R
from test_digitex_gpu import batch_of, unbatch
from test_hal_gpu import from_dev, same, to_dev
from test_oracle_homcrt import _decrypt, _encrypt, _plain_transform
from test_oracle_digitex import zero_secret_key
from test_ring_gpu import uniform_residues

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    be = load().HipBackend(0)
    yield be
    be.close()


@pytest.fixture(scope="module")
def rings(gpu, cpu):
    pkg = load()
    made = {}

    def get(field, log2_q, t_bits):
        key = (field, log2_q, t_bits)
        if key not in made:
            made[key] = (pkg.ring.RingContext(gpu, field, log2_q, t_bits),
                         R.RingContext(cpu, R.FieldDescriptor(field), log2_q, t_bits))
        return made[key]

    yield get
    for g, _ in made.values():
        g.close()


def crt_batch_of(gpu, pkg, cts, base):
    b = batch_of(gpu, pkg, cts, base)
    b.basis = "crt"
    return b


@pytest.mark.parametrize("field,log2_q", [("field_18", (300, 320)), ("field_36864", (160, 180))])
def test_lprime_element_tensor(gpu, rings, field, log2_q):
    """ring.rs:1227-1250 on the device, equal to the oracle's residue for residue"""
    g, c = rings(field, log2_q, 16)
    base = g.initial_rns_base()
    rng = np.random.default_rng(4)
    n2 = g.info["deg_lprime"]
    for coeffs in ([1] + [0] * (n2 - 1), [int(x) for x in rng.integers(-(1 << 60), 1 << 60, n2)], [-3] * n2):
        got = from_dev(gpu, g.lprime_element_tensor(base, coeffs))
        assert same(got, OH.lprime_element_tensor(c, base, coeffs)), coeffs
    assert g.lprime_element_tensor(base, coeffs) is g.lprime_element_tensor(base, list(coeffs))  # kept per key


def _uniform_keys(rng, g, gs, key_base, ndigits):
    primes = [g.primes[i] for i in key_base]
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    digits = R.select_digits(ndigits, len(key_base))
    host = {x: uniform_residues(rng, primes, (2 * ndigits, len(key_base), n2, n3)) for x in gs}
    return host, digits


def test_row_addressed_keyswitch_matches_the_batched_form(gpu, rings):
    """csd_ring_galois_keyswitch_rows with the plain batch layout, and with a digit-major dBFV
    layout and keys restricted from a longer base, equal csd_ring_galois_keyswitch row for row"""
    pkg = load()
    g, c = rings("field_36864", (160, 180), 16)
    base = g.initial_rns_base()[:2]
    key_base = g.initial_rns_base()
    r, d, n, digits = len(base), 3, 4, 2
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    rng = np.random.default_rng(8)
    primes = [g.primes[i] for i in base]
    E = d * n
    parts = to_dev(gpu, uniform_residues(rng, primes, (E * digits, r, n2, n3)))
    data = uniform_residues(rng, primes, (2 * d * n, r, n2, n3))
    gens = [x for x, _ in c.field.lq_cyclic]
    gs = [gens[0], gens[-1]]
    keys_full = [uniform_residues(rng, [g.primes[i] for i in key_base], (2 * digits, len(key_base), n2, n3)) for _ in gs]
    keys_r = [to_dev(gpu, np.ascontiguousarray(k[:, :r])) for k in keys_full]
    # the plain form on the c0 rows gathered per element e = i n + c
    c0 = np.stack([data[2 * (e // n) * n + e % n] for e in range(E)])
    want = gpu.hal_allocate_i64((len(gs) * E * 2, r, n2, n3))
    g.galois_keyswitch(base, E, digits, parts, to_dev(gpu, c0), gs, keys_r, want)
    want = from_dev(gpu, want)
    outs = [gpu.hal_allocate_i64((2 * d * n, r, n2, n3)) for _ in gs]
    kviews = [to_dev(gpu, k)[0:2 * digits, 0:r] for k in keys_full]
    g.galois_keyswitch_rows(base, E, n, digits, parts, to_dev(gpu, data), (2 * n, 1), gs, kviews, outs, (2 * n, 1, n))
    for t, o in enumerate(outs):
        got = from_dev(gpu, o)
        for e in range(E):
            i, cc = divmod(e, n)
            for s in range(2):
                assert same(got[(2 * i + s) * n + cc], want[(t * E + e) * 2 + s]), (t, e, s)
    # rows beyond the output tensor are refused on the host
    with pytest.raises(pkg.HalError):
        g.galois_keyswitch_rows(base, E, n, digits, parts, to_dev(gpu, data), (2 * n, 1), gs[:1], kviews[:1],
                                [outs[0][0:2 * d * n - 1]], (2 * n, 1, n))


def _transform_case(gpu, rings, field, log2_q, t_bits, modulus, lengths, n, seed, ndigits, nbase=None, key_extra=0,
                    zero_key=False, fused=True):
    pkg = load()
    g, c = rings(field, log2_q, t_bits)
    init = g.initial_rns_base()
    base = init[:nbase] if nbase else init
    key_base = init[:len(base) + key_extra]
    f = c.field
    rng = np.random.default_rng(seed)
    t = modulus.t()
    gen = f.lq_cyclic[-1][0]
    need = sorted({OH.pow_mod_m(gen, e, f.m) for L in lengths for e in range(1, L + 1)})
    if zero_key:
        keys, digits = zero_secret_key(c, key_base, ndigits)
        host = {x: np.concatenate(keys) for x in need}
    else:
        host, digits = _uniform_keys(rng, g, need, key_base, ndigits)
    ogks = {x: ([host[x][2 * i:2 * i + 2] for i in range(len(digits))], digits) for x in need}
    dgks = {x: pkg.homcrt.GaloisKey(to_dev(gpu, host[x]), digits) for x in need}
    xs = [[int(v) for v in rng.integers(-(t // 2) + 1, t // 2 + 1, f.deg_lprime * f.deg_k)] for _ in range(n)]
    cts = [_encrypt(c, base, modulus, x, rng) for x in xs]
    omod = OD.DBFVModulus(modulus.log2_b, modulus.log2_t)
    for L in lengths:
        coeffs = [[int(v) for v in rng.integers(-(t // 2) + 1, t // 2 + 1, f.deg_lprime)] for _ in range(L)]
        want = [OH.hom_cyclic_transform_dbfv(c, omod, gen, coeffs, ct, ogks) for ct in cts]
        got = pkg.homcrt.hom_cyclic_transform_dbfv(g, modulus, gen, coeffs, crt_batch_of(gpu, pkg, cts, base), dgks, fused)
        assert got.basis == "crt" and got.digits() == want[0].digits() and got.count == n
        for k, (a, b) in enumerate(zip(unbatch(gpu, got), want)):
            assert same(a, b.data), (L, k)
        if zero_key:
            for k, x in enumerate(xs):
                dec = _decrypt(c, omod, OD.Ct(unbatch(gpu, got)[k], base, "crt", 1))
                assert dec == _plain_transform(c, base, omod, gen, coeffs, x), (L, k)


@pytest.mark.parametrize("zero_key,fused", [(False, True), (True, True), (False, False)])
def test_hom_cyclic_transform_field18(gpu, rings, zero_key, fused):
    """lengths 1 (one fma), 4 (baby steps only), 6 (3 x 2) and 20 (7 x 3, two carried giant steps)
    over 3 ciphertexts of 4 digits; keys over a base one prime longer than the ciphertexts'; the
    giant steps as one combination kernel each (fused) and as the accumulator sequence"""
    from _pkg import load as _l

    mod = _l().dbfv.DBFVModulus.new(4, 4)
    _transform_case(gpu, rings, "field_18", (300, 320), 16, mod, [1, 4, 6, 20], 3, 31, 3, nbase=5, key_extra=1,
                    zero_key=zero_key, fused=fused)


def test_hom_cyclic_transform_field36864_factors(gpu, rings):
    """field_36864's generator of the K part at lengths 6 (its J_13 ... factor shape 3 x 2) and 24
    (5 x 5 blocks), 9 ciphertexts (a full group of 8 plus one) of 2 digits, 3 primes of a 5-prime
    key base with 2 gadget digits"""
    from _pkg import load as _l

    mod = _l().dbfv.DBFVModulus.new_incomplete(8, 12)
    _transform_case(gpu, rings, "field_36864", (280, 300), 16, mod, [6, 24], 9, 57, 2, nbase=3, key_extra=2)
