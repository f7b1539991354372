"""The bfv-layer sequences of the hot path on the MI355X (the ring multiplication of config C2:
lift to the extended base, tensor product in wide accumulators, rescale, relinearization with
gadget base conversion and the key-switch inner product), bit-exact against the oracle's
sequences (bfv.rs:408-426, 498-528, 564-622, 667-775) on the same seeded inputs."""
import numpy as np
import pytest

from _pkg import load
from oracle import ring as R
from test_hal_gpu import from_dev, same, to_dev
from test_ring_gpu import uniform_residues

pytestmark = pytest.mark.gpu

CASES = [("field_18", (100, 120), 4, 2), ("field_192", (100, 120), 8, 2), ("field_36864", (960, 980), 76, 9)]


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


def setup(rings, field, log2_q, t_bits, ndigits, seed):
    g, c = rings(field, log2_q, t_bits)
    base = g.initial_rns_base()
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(seed)
    shape = (2, len(base), g.info["deg_lprime"], g.info["deg_k"])
    lhs, rhs = uniform_residues(rng, primes, shape), uniform_residues(rng, primes, shape)
    digits = R.select_digits(ndigits, len(base))
    keys = [uniform_residues(rng, primes, shape) for _ in digits]
    return g, c, base, lhs, rhs, digits, keys


@pytest.mark.parametrize("field,log2_q,t_bits,ndigits", CASES)
def test_lift_and_rescale_parity(gpu, cpu, rings, field, log2_q, t_bits, ndigits):
    pkg = load()
    g, c, base, lhs, rhs, digits, keys = setup(rings, field, log2_q, t_bits, ndigits, 11)
    assert g.extension_rns_base_for(base) == c.extension_rns_base_for(base)
    lc, lifted_c = R.lift_to_extended(c, lhs.copy(), base)
    lg, lifted_g = pkg.bfv.lift_to_extended(g, to_dev(gpu, lhs), base)
    assert lifted_g == lifted_c
    assert same(from_dev(gpu, lg), lc)
    # rescale a period-basis tensor over the lifted base
    x = uniform_residues(np.random.default_rng(12), [g.primes[i] for i in lifted_c], (2, len(lifted_c)) + lhs.shape[2:])
    t = 1 << t_bits if t_bits else 2
    rc = R.rescale_to_base(c, x.copy(), lifted_c, len(base), t)
    rg = pkg.bfv.rescale_to_base(g, to_dev(gpu, x), lifted_g, len(base), t)
    assert same(from_dev(gpu, rg), rc)


@pytest.mark.parametrize("field,log2_q,t_bits,ndigits", CASES)
def test_relin_parity(gpu, cpu, rings, field, log2_q, t_bits, ndigits):
    pkg = load()
    g, c, base, lhs, rhs, digits, keys = setup(rings, field, log2_q, t_bits, ndigits, 21)
    prod = np.concatenate([lhs, rhs[0:1]])  # a 3-row period-basis product
    want = R.relin_crt(c, prod.copy(), base, keys, digits)
    got = pkg.bfv.relin_crt(g, to_dev(gpu, prod), base, [to_dev(gpu, k) for k in keys], digits)
    assert same(from_dev(gpu, got), want)


@pytest.mark.parametrize("field,log2_q,t_bits,ndigits", CASES)
def test_hom_mul_parity(gpu, cpu, rings, field, log2_q, t_bits, ndigits):
    """bfv_hom_mul (bfv.rs:747-775): the whole C2 ring multiplication, bit-exact"""
    pkg = load()
    g, c, base, lhs, rhs, digits, keys = setup(rings, field, log2_q, t_bits, ndigits, 31)
    t = 1 << t_bits if t_bits else 2
    want = R.bfv_hom_mul(c, t, lhs, rhs, base, keys, digits)
    got = pkg.bfv.bfv_hom_mul(g, t, to_dev(gpu, lhs), to_dev(gpu, rhs), base, [to_dev(gpu, k) for k in keys], digits)
    assert same(from_dev(gpu, got), want)


def test_select_digits_matches_oracle():
    pkg = load()
    for r in (1, 2, 17, 23, 37):
        for d in range(1, r + 1):
            assert pkg.bfv.select_digits(d, r) == R.select_digits(d, r)


@pytest.mark.parametrize("field,log2_q,t_bits,ndigits,pairs", [("field_18", (100, 120), 4, 2, 3),
                                                               ("field_36864", (960, 980), 76, 9, 3)])
def test_hom_mul_batch_matches_per_pair(gpu, rings, field, log2_q, t_bits, ndigits, pairs):
    """the batched ring multiplication gives, pair by pair, the residues of bfv_hom_mul"""
    pkg = load()
    g, c = rings(field, log2_q, t_bits)
    base = g.initial_rns_base()
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(55)
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    lhs = uniform_residues(rng, primes, (2 * pairs, len(base), n2, n3))
    rhs = uniform_residues(rng, primes, (2 * pairs, len(base), n2, n3))
    digits = R.select_digits(ndigits, len(base))
    keys = [to_dev(gpu, uniform_residues(rng, primes, (2, len(base), n2, n3))) for _ in digits]
    t = 1 << t_bits
    got = from_dev(gpu, pkg.bfv.bfv_hom_mul_batch(g, t, to_dev(gpu, lhs), to_dev(gpu, rhs), base, keys, digits))
    for b in range(pairs):
        one = pkg.bfv.bfv_hom_mul(g, t, to_dev(gpu, lhs[2 * b:2 * b + 2]), to_dev(gpu, rhs[2 * b:2 * b + 2]), base, keys,
                                  digits)
        assert same(got[2 * b:2 * b + 2], from_dev(gpu, one)), b


def test_unzeroed_wide_rows_must_be_written(gpu, rings):
    """an accumulator allocated without the zero fill refuses to reduce rows never written"""
    g, _ = rings("field_18", (100, 120), 4)
    base = g.initial_rns_base()
    acc = g.allocate_wide(2, base, zero=False)
    x = to_dev(gpu, uniform_residues(np.random.default_rng(4), [g.primes[i] for i in base],
                                     (1, len(base), g.info["deg_lprime"], g.info["deg_k"])))
    acc.add(range(0, 1), x, 1)
    with pytest.raises(ValueError):
        g.reduce_wide(acc)
    acc.add(range(1, 2), x, 1)
    out = from_dev(gpu, g.reduce_wide(acc))
    assert same(out[0], out[1])


def test_relin_batch_key_layouts(gpu, rings):
    """relin_batch gives the same residues when the key digits are views with different strides
    (one a limb prefix of a larger key), in both output bases; and csd_keyswitch_reduce rejects
    operands that do not fit"""
    pkg = load()
    g, _ = rings("field_192", (100, 120), 8)
    base = g.initial_rns_base()
    r = len(base)
    primes = [g.primes[i] for i in base]
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    rng = np.random.default_rng(77)
    digits = R.select_digits(2, r)
    K = 3
    prod = uniform_residues(rng, primes, (3 * K, r, n2, n3))
    keys = [uniform_residues(rng, primes, (2, r, n2, n3)) for _ in digits]
    wide = np.zeros((2, r + 2, n2, n3), dtype=np.int64)
    wide[:, :r] = keys[1]
    for basis in ("crt", "pb"):
        same_layout, _ = pkg.bfv.relin_batch(g, to_dev(gpu, prod), K, base, [to_dev(gpu, k) for k in keys], digits, basis)
        mixed = [to_dev(gpu, keys[0]), to_dev(gpu, wide)]  # the second key has r + 2 limbs
        other, _ = pkg.bfv.relin_batch(g, to_dev(gpu, prod), K, base, mixed, digits, basis)
        assert same(from_dev(gpu, same_layout), from_dev(gpu, other))
    parts = gpu.hal_allocate_i64((K * 2, r, n2, n3))
    out = gpu.hal_allocate_i64((2 * K, r, n2, n3))
    kd = [to_dev(gpu, k) for k in keys]
    with pytest.raises(pkg.HalError):  # parts rows are not K x digits
        gpu.keyswitch_reduce(out, (K, 1), parts, K, 3, kd + kd[:1], None, None, False, g.mods_for(base))
    with pytest.raises(pkg.HalError):  # output rows out of range
        gpu.keyswitch_reduce(out, (K, 2), parts, K, 2, kd, None, None, False, g.mods_for(base))


def _oracle_rescale_step(c, x, lifted, base_len, t):
    """Lines 607-616 of rescale_to_base (bfv.rs:601-622) on the oracle: the extension rows of x."""
    be = c.backend
    base, ext = lifted[:base_len], lifted[base_len:]
    be.hal_mul_mont64(x, c.scalar_tensor(lifted, t), c.mods_for(lifted))
    tmp = c.allocate(x.shape[0], ext)
    c.bconv_tensors(tmp, ext, x[:, :base_len], base)
    be.hal_sub_i64(x[:, base_len:], tmp)
    be.hal_mul_mont64(x[:, base_len:], c.inverse_scalar_tensor(ext, c.modulus_of(base)), c.mods_for(ext))
    return x[:, base_len:]


@pytest.mark.parametrize("form", [3, 2, 1, 0])
@pytest.mark.parametrize("layout", ["split", "separate", "ext_first"])
@pytest.mark.parametrize("field,log2_q,t_bits,ndigits", CASES[1:])
def test_rescale_step_parity(gpu, cpu, rings, field, log2_q, t_bits, ndigits, layout, form):
    """csd_ring_rescale_step against the oracle's sequence, bit for bit, in every conversion form:
    source and extension as limb ranges of one tensor (either order) or as separate tensors, the
    extension a strided view of a wider one."""
    pkg = load()
    g, c, base, lhs, rhs, digits, keys = setup(rings, field, log2_q, t_bits, ndigits, 13)
    _, lifted = R.lift_to_extended(c, lhs.copy(), base)
    bl, el = len(base), len(lifted) - len(base)
    x = uniform_residues(np.random.default_rng(14), [g.primes[i] for i in lifted], (3, len(lifted)) + lhs.shape[2:])
    t = 1 << t_bits
    want = _oracle_rescale_step(c, x.copy(), lifted, bl, t)
    s = g.scalar_tensor(lifted, t)
    inv = g.inverse_product_tensor(lifted[bl:], base)
    if layout == "split":
        xd = to_dev(gpu, x)
        src, ext = xd[:, :bl], xd[:, bl:]
    elif layout == "ext_first":  # extension rows first, then the source rows, in one tensor
        xd = to_dev(gpu, np.concatenate([x[:, bl:], x[:, :bl]], axis=1))
        ext, src = xd[:, :el], xd[:, el:]
    else:
        wide = np.zeros((3, el + 2) + x.shape[2:], dtype=np.int64)
        wide[:, 1:el + 1] = x[:, bl:]
        xd = to_dev(gpu, wide)
        src, ext = to_dev(gpu, np.ascontiguousarray(x[:, :bl])), xd[:, 1:el + 1]
    src_before = from_dev(gpu, src)
    g.set_bconv_form(form)
    try:
        g.rescale_step(ext, lifted[bl:], src, base, s[:, :bl], s[:, bl:], inv)
    finally:
        g.set_bconv_form(2)
    assert same(from_dev(gpu, ext), want)
    if form >= 1:
        assert same(from_dev(gpu, src), src_before)  # the fused forms leave the source as it is
    if layout == "separate":  # the rows beside the extension view stay untouched
        assert not from_dev(gpu, xd[:, :1]).any() and not from_dev(gpu, xd[:, el + 1:]).any()


def test_rescale_step_errors(gpu, rings):
    pkg = load()
    g, c, base, lhs, rhs, digits, keys = setup(rings, "field_192", (100, 120), 8, 2, 15)
    _, lifted = R.lift_to_extended(c, lhs.copy(), base)
    bl = len(base)
    x = g.allocate(2, lifted)
    s = g.scalar_tensor(lifted, 1 << 8)
    inv = g.inverse_product_tensor(lifted[bl:], base)
    with pytest.raises(pkg.hal.HalError):  # scalar tensor over the wrong base
        g.rescale_step(x[:, bl:], lifted[bl:], x[:, :bl], base, s, s[:, bl:], inv)
    with pytest.raises(pkg.hal.HalError):  # extension of another element count
        g.rescale_step(x[:1, bl:], lifted[bl:], x[:, :bl], base, s[:, :bl], s[:, bl:], inv)
