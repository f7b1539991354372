"""The dBFV digit extraction and floor division (bootstrap/digitex.rs) on the MI355X, batched over
several ciphertexts, bit-exact (L0) against the oracle's per-ciphertext restatement
(oracle/digitex.py, pinned by the reference's floor-division tests under the zero secret key):
the batched dBFV accumulators and exact division, the shallow and the full 8-multiplication
circuit, several levels with pending accumulators, the HAL-sequence and fused digit reductions,
uniform keys, and field_49152 at a reduced base. Under the zero secret key the device results
also decrypt to the floor division."""
import numpy as np
import pytest

from _pkg import load
from oracle import dbfv as OD
from oracle import digitex as OX
from oracle import ring as R
from test_hal_gpu import from_dev, same, to_dev
from test_oracle_digitex import decrypt_scalar, encrypt_scalar, zero_secret_key
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


def batch_of(gpu, pkg, cts, base):
    """oracle Ct list -> a digit-major device batch (row (2 i + s) count + c)"""
    n = len(cts)
    d = cts[0].digits()
    arr = np.zeros((2 * d * n,) + cts[0].data.shape[1:], dtype=np.int64)
    for c, ct in enumerate(cts):
        for row in range(2 * d):
            arr[row * n + c] = ct.data[row]
    return pkg.dbfv.DBFVBatch(to_dev(gpu, arr), base, n, "pb", cts[0].bound)


def unbatch(gpu, b):
    a = from_dev(gpu, b.data)
    n = b.count
    return [a[c::n] for c in range(n)]


def run_floordiv(gpu, cpu, rings, field, log2_q, t_bits, init, final, values, seed, key="zero", fused=True, ndigits=None):
    pkg = load()
    g, c = rings(field, log2_q, t_bits)
    base = g.initial_rns_base()
    nd = ndigits or len(base)
    if key == "zero":
        keys, digits = zero_secret_key(c, base, nd)
    else:
        rng = np.random.default_rng(seed + 7)
        keys = [uniform_residues(rng, [g.primes[i] for i in base], (2, len(base), g.info["deg_lprime"], g.info["deg_k"]))
                for _ in range(nd)]
        digits = R.select_digits(nd, len(base))
    rng = np.random.default_rng(seed)
    cts = [encrypt_scalar(c, base, init.t(), m, rng) for m in values]
    want = [OX.hom_floordiv_dbfv(c, OD.DBFVModulus(init.log2_b, init.log2_t),
                                 OD.DBFVModulus(final.log2_b, final.log2_t), ct, keys, digits) for ct in cts]
    got = pkg.digitex.hom_floordiv_dbfv(g, init, final, batch_of(gpu, pkg, cts, base), [to_dev(gpu, k) for k in keys],
                                        digits, fused=fused)
    assert got.base == want[0].base and got.bound == want[0].bound and got.count == len(values)
    for c_, (arr, w) in enumerate(zip(unbatch(gpu, got), want)):
        assert same(arr, np.asarray(w.data)), c_
    return got, want, c


@pytest.mark.parametrize("fused", [True, False])
def test_floordiv_field192(gpu, cpu, rings, fused):
    """digitex.rs:378-397 for three ciphertexts at once: (4, 4) -> (4, 3), v = 4"""
    pkg = load()
    init, final = pkg.dbfv.DBFVModulus.new(4, 4), pkg.dbfv.DBFVModulus.new(4, 3)
    values = [2025, -777, 40000]
    got, want, c = run_floordiv(gpu, cpu, rings, "field_192", (800, 820), 16, init, final, values, 1, fused=fused)
    fm = OD.DBFVModulus.new(4, 3)
    for w, m in zip(want, values):
        assert decrypt_scalar(c, fm, w) == OD.cmod((m % (1 << 16)) >> 4, fm.t())


def test_floordiv_multiple_digits(gpu, cpu, rings):
    """v = 5 > log2 b: four extraction levels with pending accumulators and exact divisions that
    drop digits (digitex.rs:456-481)"""
    pkg = load()
    init, final = pkg.dbfv.DBFVModulus.new(4, 4), pkg.dbfv.DBFVModulus.new_incomplete(4, 11)
    got, want, c = run_floordiv(gpu, cpu, rings, "field_192", (900, 920), 16, init, final, [40529, 123], 5)
    assert decrypt_scalar(c, OD.DBFVModulus.new_incomplete(4, 11), want[0]) == OD.cmod(40529 >> 5, 1 << 11)


def test_floordiv_deep_circuit(gpu, cpu, rings):
    """precision 36: the full 8-multiplication circuit (digitex.rs:399-423), two ciphertexts"""
    pkg = load()
    init = pkg.dbfv.DBFVModulus.new_incomplete(8, 36)
    final = pkg.dbfv.DBFVModulus.new_incomplete(8, 35)
    run_floordiv(gpu, cpu, rings, "field_192", (900, 920), 36, init, final, [12345678901, -5], 9)


def test_floordiv_uniform_keys(gpu, cpu, rings):
    """relinearization with a uniform key of 3 gadget digits (no decryption meaning; residues
    only)"""
    pkg = load()
    init, final = pkg.dbfv.DBFVModulus.new(4, 4), pkg.dbfv.DBFVModulus.new_incomplete(4, 14)
    run_floordiv(gpu, cpu, rings, "field_192", (800, 820), 16, init, final, [11, 22], 13, key="uniform", ndigits=3)


def test_floordiv_field49152_reduced_base(gpu, cpu, rings):
    """field_49152 (the C5 field) on a 13-prime base with 8-bit digits: (8, 8) -> (8, 7), the
    3-multiplication circuit, two ciphertexts"""
    pkg = load()
    init, final = pkg.dbfv.DBFVModulus.new(8, 1), pkg.dbfv.DBFVModulus.new_incomplete(8, 7)
    got, want, c = run_floordiv(gpu, cpu, rings, "field_49152", (700, 720), 8, init, final, [201, -45], 17, ndigits=5)
    assert decrypt_scalar(c, OD.DBFVModulus.new_incomplete(8, 7), want[0]) == OD.cmod(201 >> 1, 1 << 7)


def test_accumulators_and_exact_division(gpu, cpu, rings):
    """the batched accumulator ops and dbfv_exact_div_pow2 against the oracle, per ciphertext"""
    pkg = load()
    D = pkg.dbfv
    g, c = rings("field_192", (800, 820), 16)
    base = g.initial_rns_base()
    m, mo = D.DBFVModulus.new(4, 4), OD.DBFVModulus.new(4, 4)
    rng = np.random.default_rng(21)
    n = 3
    cts3 = [OD.Ct(uniform_residues(rng, [g.primes[i] for i in base], (6,) + (len(base), 2, 96)), base, "pb", 2)
            for _ in range(n)]
    cts1 = [OD.Ct(uniform_residues(rng, [g.primes[i] for i in base], (2, len(base), 2, 96)), base, "pb", 1)
            for _ in range(n)]
    b3, b1 = batch_of(gpu, pkg, cts3, base), batch_of(gpu, pkg, cts1, base)
    acc = D.dbfv_allocate_accumulator(g, m, base, "pb", n)
    D.dbfv_acc_inner_product_scalar(g, acc, [(b3, -32106558060291480110540), (b1, 4711)])
    D.dbfv_acc_add(g, acc, b1)
    D.dbfv_acc_sub(g, acc, b3)
    D.dbfv_acc_add_constant(g, acc, 152453116397172589318468)
    out = D.dbfv_finish_accumulator(g, acc)
    div = D.dbfv_exact_div_pow2(g, m, out, 5)
    for k in range(n):
        oacc = OD.dbfv_allocate_accumulator(c, mo, base, "pb")
        OD.dbfv_acc_inner_product_scalar(c, oacc, [(cts3[k], -32106558060291480110540), (cts1[k], 4711)])
        OD.dbfv_acc_add(c, oacc, cts1[k])
        OD.dbfv_acc_sub(c, oacc, cts3[k])
        OD.dbfv_acc_add_constant(c, oacc, 152453116397172589318468)
        oout = OD.dbfv_finish_accumulator(c, oacc)
        assert oout.digits() == out.digits()
        assert same(unbatch(gpu, out)[k], np.asarray(oout.data))
        odiv = OD.dbfv_exact_div_pow2(c, mo, oout, 5)
        assert same(unbatch(gpu, div)[k], np.asarray(odiv.data))
