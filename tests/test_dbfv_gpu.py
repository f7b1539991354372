"""The dBFV digit reduction on the MI355X (split_overflow and the dbfv_reduce carry chain,
dbfv.rs:873-1025), bit-exact against the oracle's restatement (oracle/dbfv.py, itself pinned by
the reference's property test and the carry-chain invariant) on the same seeded inputs: one-limb
and two-limb overflows (R = t/b up to 2^100), partial digit counts, unreduced inputs, and the
batched form against the per-ciphertext form; the fused chain as one kernel, the fused form with
one kernel per digit, and the reference's HAL call sequence."""
import numpy as np
import pytest

from _pkg import load
from oracle import dbfv as OD
from oracle import ring as R
from oracle.ring import select_digits as OD_select
from test_hal_gpu import from_dev, same, to_dev
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


def moduli(log2_b, log2_t):
    pkg = load()
    return pkg.dbfv.DBFVModulus(log2_b, log2_t), OD.DBFVModulus(log2_b, log2_t)


def inputs(g, base, rows, seed, bound):
    """uniform residues, or for bound 2 unreduced values in (-2p, 2p)"""
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(seed)
    shape = (rows, len(base), g.info["deg_lprime"], g.info["deg_k"])
    x = uniform_residues(rng, primes, shape)
    if bound == 2:
        for j, p in enumerate(primes):
            x[:, j] += rng.integers(-2, 1, size=x[:, j].shape, dtype=np.int64) * p
    return x


SPLIT_CASES = [("field_192", (200, 220), 8, 24), ("field_192", (200, 220), 8, 108), ("field_192", (200, 220), 27, 108),
               ("field_36864", (960, 980), 27, 108), ("field_36864", (960, 980), 16, 64)]


@pytest.mark.parametrize("field,log2_q,log2_b,log2_t", SPLIT_CASES)
@pytest.mark.parametrize("bound", [1, 2])
@pytest.mark.parametrize("fused", [True, False])
def test_split_overflow_parity(gpu, cpu, rings, field, log2_q, log2_b, log2_t, bound, fused):
    pkg = load()
    g, c = rings(field, log2_q, log2_t)
    base = g.initial_rns_base()
    mg, mc = moduli(log2_b, log2_t)
    x = inputs(g, base, 2, 100 + log2_t + bound, bound)
    dc = OD.SplitOverflowData(c, mc, base)
    dg = pkg.dbfv.SplitOverflowData(g, mg, base)
    assert (dg.log2_gamma, dg.q_mod_r, dg.neg_q_inv_mod_r, dg.limbs()) == (dc.log2_gamma, dc.q_mod_r, dc.neg_q_inv_mod_r,
                                                                           dc.limbs())
    assert same(from_dev(gpu, dg.matrix), np.asarray(dc.matrix))
    assert same(from_dev(gpu, dg.r_q_over_p_inv), np.asarray(dc.r_q_over_p_inv))
    for k in range(dc.limbs()):
        assert same(from_dev(gpu, dg.neg_b_delta_mont2[k]), np.asarray(dc.neg_b_delta_mont2[k]))
    xc, xg = x.copy(), to_dev(gpu, x)
    oc = OD.split_overflow(c, dc, xc, bound, base)
    og = pkg.dbfv.split_overflow(g, dg, xg, bound, base, fused=fused)
    assert same(from_dev(gpu, og), np.asarray(oc))
    assert same(from_dev(gpu, xg), xc)


REDUCE_CASES = [("field_192", (200, 220), 16, 64, 4), ("field_192", (200, 220), 16, 64, 2),
                ("field_192", (260, 280), 27, 108, 4), ("field_192", (260, 280), 27, 108, 1),
                ("field_36864", (960, 980), 27, 108, 4)]


@pytest.mark.parametrize("field,log2_q,log2_b,log2_t,d", REDUCE_CASES)
@pytest.mark.parametrize("fused", [True, False])
def test_dbfv_reduce_parity(gpu, cpu, rings, field, log2_q, log2_b, log2_t, d, fused):
    pkg = load()
    g, c = rings(field, log2_q, log2_t)
    base = g.initial_rns_base()
    mg, mc = moduli(log2_b, log2_t)
    x = inputs(g, base, 2 * d, 7 * d + log2_t, 2)
    rc, bc = OD.dbfv_reduce(c, mc, x.copy(), 2, base)
    rg, bg = pkg.dbfv.dbfv_reduce(g, mg, to_dev(gpu, x), 2, base, fused=fused)
    assert bg == bc == 1 and rg.shape[0] == 2 * mg.digits()
    assert same(from_dev(gpu, rg), np.asarray(rc))


@pytest.mark.parametrize("field,log2_q,log2_b,log2_t,d,count", [("field_192", (200, 220), 16, 64, 3, 5),
                                                              ("field_36864", (960, 980), 27, 108, 4, 3)])
def test_dbfv_reduce_batch_matches_single(gpu, rings, field, log2_q, log2_b, log2_t, d, count):
    """the digit-major batch gives every ciphertext the residues of its own dbfv_reduce (the
    fused batch against the unfused call sequence per ciphertext)"""
    pkg = load()
    g, _ = rings(field, log2_q, log2_t)
    base = g.initial_rns_base()
    mg, _ = moduli(log2_b, log2_t)
    x = inputs(g, base, 2 * d * count, 31 + count, 1)  # ciphertext c row k at k count + c
    rb, bb = pkg.dbfv.dbfv_reduce_batch(g, mg, to_dev(gpu, x), count, 1, base)
    assert bb == 1 and rb.shape[0] == 2 * mg.digits() * count
    out = from_dev(gpu, rb)
    for cidx in range(count):
        single, _ = pkg.dbfv.dbfv_reduce(g, mg, to_dev(gpu, x[cidx::count]), 1, base, fused=False)
        assert same(out[cidx::count], from_dev(gpu, single))


@pytest.mark.parametrize("field,log2_q,log2_b,log2_t,d,count,bound", [("field_192", (200, 220), 16, 64, 2, 3, 2),
                                                                    ("field_192", (260, 280), 27, 108, 4, 2, 1),
                                                                    ("field_36864", (960, 980), 8, 64, 3, 2, 2),
                                                                    ("field_36864", (960, 980), 27, 108, 2, 4, 2)])
def test_reduce_chain_one_launch_matches_per_digit(gpu, rings, field, log2_q, log2_b, log2_t, d, count, bound):
    """the whole carry chain in one launch (csd_dbfv_reduce_chain) writes what one fused launch
    per digit writes, and leaves its input untouched: four lanes per position for the small
    batches, one lane per position from 2^18 positions (the last case, 8 rows of 36864)"""
    pkg = load()
    g, _ = rings(field, log2_q, log2_t)
    base = g.initial_rns_base()
    mg, _ = moduli(log2_b, log2_t)
    host = inputs(g, base, 2 * d * count, 53 + d, bound)
    src = to_dev(gpu, host)
    one, _ = pkg.dbfv.dbfv_reduce_batch(g, mg, src, count, bound, base)
    pkg.dbfv._CHAIN_ONE_LAUNCH[0] = False
    try:
        per_digit, _ = pkg.dbfv.dbfv_reduce_batch(g, mg, src, count, bound, base)
    finally:
        pkg.dbfv._CHAIN_ONE_LAUNCH[0] = True
    assert same(from_dev(gpu, one), from_dev(gpu, per_digit))
    assert same(from_dev(gpu, src), host)


def test_reduce_chain_errors(gpu, rings):
    """csd_dbfv_reduce_chain refuses row counts, limbs and overlaps it cannot run"""
    pkg = load()
    g, _ = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    mg, _ = moduli(16, 64)
    data = pkg.dbfv.split_overflow_data(g, mg, base)
    mods = g.mods_for(base)
    src = to_dev(gpu, inputs(g, base, 4, 11, 1))
    out = gpu.allocate_i64_uninit([8] + list(src.shape[1:]))

    def run(x, s, per=2, digits=4, limbs=1):
        gpu.dbfv_reduce_chain(x, s, per, digits, data.consts, mods, limbs, data.q_mod_r, data.neg_q_inv_mod_r,
                              data.log2_gamma, data.log2_r, [False] * digits, [False] * digits)

    run(out, src)  # the valid call
    for args in [dict(digits=3),  # x holds 4 digits
                 dict(per=3),  # rows not a multiple of per
                 dict(limbs=3),
                 dict(digits=0)]:
        with pytest.raises(pkg.HalError):
            run(out, src, **args)
    with pytest.raises(pkg.HalError):
        run(out, out[2:6])  # src inside x, other rows
    run(out, out[0:4])  # src as x's own first rows: read and written by the same thread


def test_dbfv_reduce_edges(gpu, rings):
    pkg = load()
    g, _ = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    x = to_dev(gpu, inputs(g, base, 2, 5, 1))
    one, _ = moduli(16, 16)
    out, bound = pkg.dbfv.dbfv_reduce(g, one, x, 2, base)  # one digit: nothing to reduce
    assert out is x and bound == 2
    m4, _ = moduli(16, 64)
    with pytest.raises(ValueError):
        pkg.dbfv.dbfv_reduce(g, m4, to_dev(gpu, inputs(g, base, 3, 6, 1)), 1, base)
    with pytest.raises(ValueError):
        pkg.dbfv.dbfv_reduce(g, m4, to_dev(gpu, inputs(g, base, 10, 6, 1)), 1, base)
    with pytest.raises(ValueError):
        pkg.dbfv.DBFVModulus(8, 127)
    with pytest.raises(ValueError):
        pkg.dbfv.dbfv_reduce_batch(g, m4, x, 0, 1, base)
    with pytest.raises(ValueError):
        pkg.dbfv.dbfv_reduce_batch(g, m4, to_dev(gpu, inputs(g, base, 6, 7, 1)), 4, 1, base)


def test_fused_split_errors(gpu, rings):
    """csd_dbfv_split_overflow rejects tensors that do not match its constant table"""
    pkg = load()
    g, _ = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    mg, _ = moduli(16, 64)
    data = pkg.dbfv.split_overflow_data(g, mg, base)
    x = to_dev(gpu, inputs(g, base, 2, 9, 1))
    mods = g.mods_for(base)
    ok = gpu.hal_allocate_i64([2, 1, x.shape[2], x.shape[3]])
    for ov, cin, consts in [(gpu.hal_allocate_i64([2, 3, x.shape[2], x.shape[3]]), None, data.consts),  # 3 limbs
                            (gpu.hal_allocate_i64([3, 1, x.shape[2], x.shape[3]]), None, data.consts),  # count
                            (ok, gpu.hal_allocate_i64([2, 2, x.shape[2], x.shape[3]]), data.consts),  # carry limbs
                            (ok, None, gpu.hal_allocate_i64([1, 1, 5, len(base)]))]:  # table rows
        with pytest.raises(pkg.HalError):
            gpu.dbfv_split_overflow(x, ov, cin, consts, mods, data.q_mod_r, data.neg_q_inv_mod_r, data.log2_gamma,
                                    data.log2_r, True, True)


def test_fused_split_out_of_place(gpu, rings):
    """the out-of-place form (digit read from x_in, or zero) writes what the in-place form writes,
    and leaves its input untouched"""
    pkg = load()
    g, _ = rings("field_192", (200, 220), 108)
    base = g.initial_rns_base()
    mg, _ = moduli(27, 108)
    data = pkg.dbfv.split_overflow_data(g, mg, base)
    host = inputs(g, base, 2, 41, 1)
    src = to_dev(gpu, host)
    inplace = to_dev(gpu, host)
    ov_a = pkg.dbfv._split_fused(g, data, inplace, 1, base)
    out = gpu.allocate_i64_uninit(list(src.shape))
    ov_b = pkg.dbfv._split_fused(g, data, out, 1, base, x_in=src)
    assert same(from_dev(gpu, out), from_dev(gpu, inplace)) and same(from_dev(gpu, ov_a), from_dev(gpu, ov_b))
    assert same(from_dev(gpu, src), host)
    # a zero digit after a carry equals the carry step on zeroed rows
    zero_rows = to_dev(gpu, np.zeros_like(host))
    ov_c = pkg.dbfv._split_fused(g, data, zero_rows, 1, base, carry=ov_a, carry_bound=1)
    out_z = gpu.allocate_i64_uninit(list(src.shape))
    ov_d = pkg.dbfv._split_fused(g, data, out_z, 1, base, carry=ov_a, carry_bound=1, x_in=src, zero_in=True)
    assert same(from_dev(gpu, out_z), from_dev(gpu, zero_rows)) and same(from_dev(gpu, ov_c), from_dev(gpu, ov_d))
    with pytest.raises(pkg.HalError):
        pkg.dbfv._split_fused(g, data, out, 1, base, x_in=to_dev(gpu, inputs(g, base, 3, 42, 1)))


# the convolution kernel caches the operand digits in registers with 6 to 12 digits a side (the
# 8 x 7 case); the others, among them 13 digits on one side, take the general kernel
HOM_MUL_CASES = [("field_192", (200, 220), 16, 64, 4, 4, 2, "pb"), ("field_192", (200, 220), 16, 64, 2, 3, 4, "crt"),
                 ("field_192", (260, 280), 27, 108, 4, 1, 3, "pb"), ("field_36864", (960, 980), 27, 108, 4, 4, 9, "pb"),
                 ("field_192", (260, 280), 8, 112, 13, 2, 2, "pb"), ("field_192", (260, 280), 8, 64, 8, 7, 2, "pb")]


@pytest.mark.parametrize("field,log2_q,log2_b,log2_t,d1,d2,ndigits,basis", HOM_MUL_CASES)
@pytest.mark.parametrize("fused", [True, False])
def test_dbfv_hom_mul_parity(gpu, cpu, rings, field, log2_q, log2_b, log2_t, d1, d2, ndigits, basis, fused):
    """dbfv_hom_mul (dbfv.rs:1033-1143): reduction of both operands, lift, digit convolution,
    rescale and the batched relinearization of every output digit, bit-exact against the
    oracle's per-digit sequence; period- and CRT-basis outputs, uneven digit counts"""
    pkg = load()
    g, c = rings(field, log2_q, log2_t)
    base = g.initial_rns_base()
    mg, mc = moduli(log2_b, log2_t)
    lhs = inputs(g, base, 2 * d1, 60 + d1, 2)
    rhs = inputs(g, base, 2 * d2, 61 + d2, 1)
    digits = OD_select(ndigits, len(base))
    rng = np.random.default_rng(62)
    keys = [uniform_residues(rng, [g.primes[i] for i in base], (2, len(base), g.info["deg_lprime"], g.info["deg_k"]))
            for _ in digits]
    oc, bc = OD.dbfv_hom_mul(c, mc, lhs.copy(), 2, rhs.copy(), 1, base, keys, digits, basis)
    og, bg = pkg.dbfv.dbfv_hom_mul(g, mg, to_dev(gpu, lhs), 2, to_dev(gpu, rhs), 1, base,
                                   [to_dev(gpu, k) for k in keys], digits, basis, fused=fused)
    assert bg == bc == (2 if basis == "pb" else 1) and og.shape[0] == oc.shape[0]
    assert same(from_dev(gpu, og), np.asarray(oc))


def test_dbfv_hom_square_parity(gpu, cpu, rings):
    pkg = load()
    g, c = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    mg, mc = moduli(16, 64)
    ct = inputs(g, base, 6, 70, 1)
    digits = OD_select(3, len(base))
    rng = np.random.default_rng(71)
    keys = [uniform_residues(rng, [g.primes[i] for i in base], (2, len(base), g.info["deg_lprime"], g.info["deg_k"]))
            for _ in digits]
    oc, bc = OD.dbfv_hom_square(c, mc, ct.copy(), 1, base, keys, digits)
    og, bg = pkg.dbfv.dbfv_hom_square(g, mg, to_dev(gpu, ct), 1, base, [to_dev(gpu, k) for k in keys], digits)
    assert bg == bc == 2 and same(from_dev(gpu, og), np.asarray(oc))
    # the square equals the product of the ciphertext with itself
    om, _ = pkg.dbfv.dbfv_hom_mul(g, mg, to_dev(gpu, ct), 1, to_dev(gpu, ct), 1, base, [to_dev(gpu, k) for k in keys],
                                  digits)
    assert same(from_dev(gpu, om), from_dev(gpu, og))


def test_dbfv_hom_mul_edges(gpu, rings):
    pkg = load()
    g, _ = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    mg, _ = moduli(16, 64)
    ct = to_dev(gpu, inputs(g, base, 4, 72, 1))
    empty = gpu.hal_allocate_i64([0, len(base), g.info["deg_lprime"], g.info["deg_k"]])
    keys = [to_dev(gpu, inputs(g, base, 2, 73, 1))]
    out, bound = pkg.dbfv.dbfv_hom_mul(g, mg, ct, 1, empty, 1, base, keys, [(0, len(base))])
    assert out.shape[0] == 0 and bound == 1
    with pytest.raises(ValueError):  # more digits than the modulus
        pkg.dbfv.dbfv_hom_mul(g, mg, to_dev(gpu, inputs(g, base, 10, 74, 1)), 1, ct, 1, base, keys, [(0, len(base))])
    with pytest.raises(ValueError):
        pkg.dbfv.dbfv_hom_mul(g, mg, ct, 1, ct, 1, base, keys, [(0, len(base))], output_basis="coeff")


@pytest.mark.parametrize("field,log2_q,log2_b,log2_t,d1,d2,count,basis",
                         [("field_192", (200, 220), 16, 64, 4, 3, 5, "pb"), ("field_36864", (960, 980), 27, 108, 4, 4, 3, "crt")])
def test_dbfv_hom_mul_batch_matches_single(gpu, rings, field, log2_q, log2_b, log2_t, d1, d2, count, basis):
    """the digit-major batch gives every pair the residues of its own dbfv_hom_mul"""
    pkg = load()
    g, _ = rings(field, log2_q, log2_t)
    base = g.initial_rns_base()
    mg, _ = moduli(log2_b, log2_t)
    lhs = inputs(g, base, 2 * d1 * count, 80 + count, 2)  # pair c, row k at k count + c
    rhs = inputs(g, base, 2 * d2 * count, 81 + count, 1)
    digits = OD_select(min(5, len(base)), len(base))
    keys = [to_dev(gpu, inputs(g, base, 2, 82 + i, 1)) for i in range(len(digits))]
    out, bound = pkg.dbfv.dbfv_hom_mul_batch(g, mg, to_dev(gpu, lhs), 2, to_dev(gpu, rhs), 1, count, base, keys, digits,
                                             basis)
    got = from_dev(gpu, out)
    K = out.shape[0] // (2 * count)
    assert K == min(mg.digits(), 2 * mg.digits() - 1) and bound == (2 if basis == "pb" else 1)
    for c in range(count):
        one, _ = pkg.dbfv.dbfv_hom_mul(g, mg, to_dev(gpu, lhs[c::count]), 2, to_dev(gpu, rhs[c::count]), 1, base, keys,
                                       digits, basis, fused=False)
        assert same(got[c::count], from_dev(gpu, one)), c


def test_convolve_reduce_errors(gpu, rings):
    """csd_convolve_reduce rejects destinations and operand row maps that do not fit"""
    pkg = load()
    g, _ = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    shape = (g.info["deg_lprime"], g.info["deg_k"])
    lhs = gpu.hal_allocate_i64((8, len(base)) + shape)
    mods = g.mods_for(base)
    for dst_rows, lrows, d1, K in [(12, (2, 1, 8), 4, 4),     # pair rows beyond the operand
                                   (11, (2, 1, 0), 4, 4),     # dst rows not 3 K count
                                   (24, (2, 1, 0), 4, 8)]:    # more output digits than d1 + d2 - 1
        dst = gpu.hal_allocate_i64((dst_rows, len(base)) + shape)
        with pytest.raises(pkg.HalError):
            gpu.convolve_reduce(dst, lhs, lrows, lhs, lrows, 1 if lrows[2] == 0 else 2, d1, 1, K, [0] * (3 * K), mods)
