"""Parity of the ring layer on the MI355X (csd_ring_* through the C ABI) with the oracle ring
layer, bit for bit: period<->CRT transforms (C1 field_2, field_18, field_36864 at full depth),
CRT kernels, BEHZ base conversion (the reference KAT and random), gadget conversion, mod-switch
and Galois automorphisms; plus the reference's own ring KATs run on the device."""
import numpy as np
import pytest

from _pkg import load
from oracle import ring as R
from test_hal_gpu import from_dev, rand_i64, same, to_dev

pytestmark = pytest.mark.gpu

BCONV_PRIMES = [
    144115188075855859, 144115188075855847, 144115188075855823, 144115188075855811, 144115188075855803,
    144115188075855761, 144115188075855677, 144115188075855599, 144115188075855509, 144115188075855449,
    144115188075855439, 144115188075855421, 144115188075855413,
]


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


def uniform_residues(rng, primes, shape):
    """residues uniform in [0, p) per limb (dimension 1)"""
    out = np.zeros(shape, dtype=np.int64)
    for j, p in enumerate(primes):
        out[:, j] = rng.integers(0, p, size=out[:, j].shape, dtype=np.int64)
    return out


@pytest.mark.parametrize("field,log2_q,t_bits,count", [("field_2", (200, 230), 0, 16), ("field_18", (100, 120), 4, 3),
                                                       ("field_192", (100, 120), 8, 2), ("field_36864", (960, 980), 76, 2),
                                                       ("field_49152", (1280, 1300), 76, 2)])
def test_transform_parity(gpu, cpu, rings, field, log2_q, t_bits, count):
    """Config C1 (field_2, 4 primes, 16 elements) and larger fields: pb->CRT, a pointwise
    Montgomery product, CRT->pb; every residue bit-exact with the oracle."""
    g, c = rings(field, log2_q, t_bits)
    assert g.primes == c.primes
    base = g.initial_rns_base()
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(2024)
    shape = (count, len(base), g.info["deg_lprime"], g.info["deg_k"])
    data = uniform_residues(rng, primes, shape)
    other = uniform_residues(rng, primes, shape)
    dc, oc = data.copy(), other.copy()
    c.pb_to_crt_tensor(dc, base)
    cpu.hal_mul_mont64(dc, oc, c.mods_for(base))
    c.crt_to_pb_tensor(dc, base)
    dg, og = to_dev(gpu, data), to_dev(gpu, other)
    g.pb_to_crt_tensor(dg, base)
    gpu.hal_mul_mont64(dg, og, [gpu.hal_create_modulus(p) for p in primes])
    g.crt_to_pb_tensor(dg, base)
    assert same(from_dev(gpu, dg), dc)


@pytest.mark.parametrize("field,log2_q,t_bits,idxs", [("field_36864", (960, 980), 76, [0, 16, 36]),
                                                      ("field_49152", (1280, 1300), 76, [0, 22, 48]),
                                                      ("field_192", (100, 120), 8, [0, 1]), ("field_18", (100, 120), 4, [0])])
def test_crt_kernels_match_oracle(gpu, cpu, rings, field, log2_q, t_bits, idxs):
    """the device-computed CRT kernels (traces, NTT, inverses) equal the oracle's host ones"""
    g, c = rings(field, log2_q, t_bits)
    for idx in idxs:
        k = g.crt_kernels(idx)
        tc = c.transform_cache([idx])
        exp = [c._kernels[c.primes[idx]][t] for t in range(4)]
        for a, b in zip(k, exp):
            assert list(a) == list(b)
        del tc


def test_field18_square_kat_on_device(gpu, rings):
    """ring.rs:1974-2012 (sage KAT), through the device transforms and wide accumulation."""
    g, _ = rings("field_18", (100, 120), 4)
    base = [0]
    p = g.primes[0]
    els = g.allocate(1, base)
    data = [0] * 18
    for blk in range(3):
        data[blk * 6] = R.mont_repr(1, p)
    gpu.hal_import(els, data)
    g.pb_to_crt_tensor(els, base)
    wide = gpu.hal_allocate_i128([1, 1, 3, 6])
    gpu.hal_mul_i64(wide, els, els)
    sqr = g.allocate(1, base)
    g.reduce_wide_tensor(sqr, wide, base, full=False)
    g.crt_to_pb_tensor(sqr, base)
    out = gpu.hal_export(sqr)
    exp = [3, 4, 2, -2, -4, -2]
    for blk in range(3):
        for j in range(6):
            assert int(out[blk * 6 + j]) % p == R.montgomery_repr(exp[j], p)


@pytest.fixture(params=[3, 1, 0], ids=["mfma", "fused", "hal_sequence"])
def bconv_form(request):
    """runs a test with the matrix-core conversion kernel at every shape (form 3), the fused VALU
    kernel and the reference's HAL sequence: the test calls the yielded function on each ring it
    uses, which is reset to the default after"""
    used = []

    def apply(ring):
        ring.set_bconv_form(request.param)
        used.append(ring)
        return request.param

    yield apply
    for ring in used:
        if ring.handle:  # rings a test closed itself are gone
            ring.set_bconv_form(2)


def test_bconv_kat_on_device(gpu, bconv_form):
    """ring.rs:1859-1917, on the device."""
    ring = load().ring.RingContext(gpu, "field_2", raw_primes=BCONV_PRIMES)
    bconv_form(ring)
    frm, to = BCONV_PRIMES[:7], BCONV_PRIMES[7:]
    d1 = [0, 1, 2, 3, 4, 5, 6, 7]
    d2 = [1 << 64, 3 << 100, -7 << 100, 9 << 120, -1 << 120, 5 << 120, 0, 7 << 120]
    vec = [R.montgomery_repr(R._trunc_mod(x, p), p) for data in (d1, d2) for p in frm for x in data]
    src = gpu.hal_allocate_i64([2, 7, 8, 1])
    gpu.hal_import(src, vec)
    dst = gpu.hal_allocate_i64([2, 6, 8, 1])
    ring.bconv_tensors(dst, list(range(7, 13)), src, list(range(7)))
    out = gpu.hal_export(dst)
    for b, data in enumerate((d1, d2)):
        for i, p in enumerate(to):
            for j, x in enumerate(data):
                assert int(out[b * 48 + i * 8 + j]) % p == R.montgomery_repr(x, p)
    ring.close()


def test_bconv_parity(gpu, cpu, rings, bconv_form):
    """The lift q -> extension of field_36864 (17 -> 20 primes) and its reverse, bit-exact; also
    from a limb slice of a larger tensor into a limb slice (strided views), and with unreduced
    (-2p, 2p) inputs. The fused form leaves its source unchanged."""
    g, c = rings("field_36864", (960, 980), 76)
    form = bconv_form(g)
    base = g.initial_rns_base()
    ext = g.extension_rns_base_for(base)
    assert ext == c.extension_rns_base_for(base)
    rng = np.random.default_rng(7)
    for sb, db in [(base, ext), (ext, base), (base[:5], base[5:9])]:
        primes = [g.primes[i] for i in sb]
        src = uniform_residues(rng, primes, (2, len(sb), 12, 3072))
        src[0, :, 0, :8] = 0  # zero and small values too
        src[1, :, 1, :64] -= np.array(primes, dtype=np.int64).reshape(-1, 1) * rng.integers(0, 2, (len(sb), 64))
        dc = np.zeros((2, len(db), 12, 3072), dtype=np.int64)
        c.bconv_tensors(dc, db, src.copy(), sb)
        dg = gpu.hal_allocate_i64(dc.shape)
        sg = to_dev(gpu, src)
        g.bconv_tensors(dg, db, sg, sb)
        assert same(from_dev(gpu, dg), dc)
        if form >= 1:
            assert same(from_dev(gpu, sg), src)
    # strided views: limbs 1.. of a [2, 1 + r, ., .] tensor into limbs 2.. of a [2, 2 + r', ., .] one
    sb, db = base, ext
    big = np.zeros((2, 1 + len(sb), 12, 3072), dtype=np.int64)
    big[:, 1:] = uniform_residues(rng, [g.primes[i] for i in sb], (2, len(sb), 12, 3072))
    dc = np.zeros((2, len(db), 12, 3072), dtype=np.int64)
    c.bconv_tensors(dc, db, big[:, 1:].copy(), sb)
    bg = to_dev(gpu, big)
    out = gpu.hal_allocate_i64((2, 2 + len(db), 12, 3072))
    g.bconv_tensors(out[:, 2:], db, bg[:, 1:], sb)
    assert same(from_dev(gpu, out)[:, 2:], dc)
    # source and destination as disjoint limb ranges of one tensor (the lift's layout)
    one = np.zeros((2, len(sb) + len(db), 12, 3072), dtype=np.int64)
    one[:, :len(sb)] = big[:, 1:]
    og = to_dev(gpu, one)
    g.bconv_tensors(og[:, len(sb):], db, og[:, :len(sb)], sb, preserve_src=True)
    got = from_dev(gpu, og)
    assert same(got[:, len(sb):], dc) and same(got[:, :len(sb)], big[:, 1:])


def test_bconv_matrix_core_path_is_taken(gpu, rings):
    """The default conversion form runs the matrix-core kernel (its own timer name) for the lift of
    field_36864 and field_49152 to their extensions, the reverse conversions and the rescale step;
    form 1 runs the vector-ALU kernel for the same calls. A conversion from 5 primes stays on the
    vector ALUs in the default form and takes the matrix cores in form 3."""
    for field, log2_q in (("field_36864", (960, 980)), ("field_49152", (1280, 1300))):
        g, _ = rings(field, log2_q, 76)
        assert g.bconv_form() == 2
        base = g.initial_rns_base()
        ext = g.extension_rns_base_for(base)
        n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
        lifted = base + ext
        x = g.allocate(1, lifted)
        g.fill_uniform(x, lifted, 3)
        d = gpu.hal_allocate_i64((1, len(base), n2, n3))
        s = g.scalar_tensor(lifted, 1 << 20)
        inv = g.inverse_product_tensor(ext, base)
        for form, names in ((2, {"bconv_mfma", "rescale_step_mfma"}), (1, {"bconv", "rescale_step"})):
            g.set_bconv_form(form)
            gpu.timing_report()  # clears the record
            gpu.set_timing(True)
            try:
                g.bconv_tensors(x[:, len(base):], ext, x[:, :len(base)], base)
                g.bconv_tensors(d, base, x[:, len(base):], ext)
                g.rescale_step(x[:, len(base):], ext, x[:, :len(base)], base, s[:, :len(base)], s[:, len(base):], inv)
                gpu.sync()
            finally:
                gpu.set_timing(False)
                g.set_bconv_form(2)
            rep = gpu.timing_report()
            assert names <= set(rep), (form, sorted(rep))
            assert rep.get("bconv_mfma" if form == 2 else "bconv", (0,))[0] == 2
        # a conversion from 5 primes: the vector ALUs in form 2 (faster there), the matrix cores in form 3
        src5 = g.allocate(1, base[:5])
        g.fill_uniform(src5, base[:5], 4)
        d4 = gpu.hal_allocate_i64((1, 4, n2, n3))
        for form, name in ((2, "bconv"), (3, "bconv_mfma")):
            g.set_bconv_form(form)
            gpu.timing_report()
            gpu.set_timing(True)
            try:
                g.bconv_tensors(d4, base[5:9], src5, base[:5])
                gpu.sync()
            finally:
                gpu.set_timing(False)
                g.set_bconv_form(2)
            assert set(gpu.timing_report()) == {name}, form


def test_bconv_parity_field49152(gpu, cpu, rings, bconv_form):
    """field_49152's bases (23 primes lifted by its extension, and back), bit-exact with the oracle"""
    g, c = rings("field_49152", (1280, 1300), 76)
    bconv_form(g)
    base = g.initial_rns_base()
    ext = g.extension_rns_base_for(base)
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    rng = np.random.default_rng(49)
    for sb, db in [(base, ext), (ext, base)]:
        src = uniform_residues(rng, [g.primes[i] for i in sb], (1, len(sb), n2, n3))
        dc = np.zeros((1, len(db), n2, n3), dtype=np.int64)
        c.bconv_tensors(dc, db, src.copy(), sb)
        dg = gpu.hal_allocate_i64(dc.shape)
        g.bconv_tensors(dg, db, to_dev(gpu, src), sb)
        assert same(from_dev(gpu, dg), dc)


GADGET_CASES = [("field_36864", (960, 980), 76, "select9", 1), ("field_36864", (960, 980), 76, "gaps", 3),
                ("field_192", (200, 220), 8, "whole", 2), ("field_18", (100, 120), 4, "single", 4),
                ("field_49152", (1280, 1300), 76, "select9", 2)]


@pytest.mark.parametrize("form", [1, 0])
@pytest.mark.parametrize("field,log2_q,t_bits,layout,count", GADGET_CASES)
def test_gadget_bconv_parity(gpu, cpu, rings, field, log2_q, t_bits, layout, count, form):
    """GadgetBConvCache::bconv (ring.rs:1809-1851) on `count` elements at once, against the
    oracle's call per element: the fused kernel (form 1, which leaves its source unchanged) and the
    HAL sequence (form 0); digit layouts of select_digits, with limbs in no digit, one digit over
    the whole base, and one-limb digits"""
    pkg = load()
    g, c = rings(field, log2_q, t_bits)
    base = g.initial_rns_base()
    r = len(base)
    digits = {"select9": lambda: R.select_digits(9, r), "gaps": lambda: [(0, 3), (5, 9), (12, r)],
              "whole": lambda: [(0, r)], "single": lambda: [(u, u + 1) for u in range(r)]}[layout]()
    d = len(digits)
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    rng = np.random.default_rng(8 + count)
    src = uniform_residues(rng, [g.primes[i] for i in base], (count, r, n2, n3))
    src[0, :, 0, :8] = 0  # zero and small values too
    src[-1, :, -1, -3:] = 1
    dc = np.zeros((count * d, r, n2, n3), dtype=np.int64)
    for b in range(count):
        c.bconv_gadget_tensors(dc[b * d:(b + 1) * d], src[b:b + 1].copy(), base, digits)
    sg = to_dev(gpu, src)
    dg = gpu.hal_allocate_i64(dc.shape)
    g.set_gadget_form(form)
    try:
        g.bconv_gadget_tensors(dg, sg, base, digits)
    finally:
        g.set_gadget_form(1)
    assert same(from_dev(gpu, dg), dc)
    if form == 1:
        assert same(from_dev(gpu, sg), src)


def test_gadget_bconv_errors(gpu, rings):
    pkg = load()
    g, _ = rings("field_18", (100, 120), 4)
    base = g.initial_rns_base()
    r = len(base)
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    src = gpu.hal_allocate_i64((2, r, n2, n3))
    for shape in [(3, r, n2, n3), (4, r - 1, n2, n3), (4, r, n2, n3 - 1)]:  # rows must be count x digits
        with pytest.raises(pkg.HalError):
            g.bconv_gadget_tensors(gpu.hal_allocate_i64(shape), src, base, [(0, 1), (1, r)])
    with pytest.raises(pkg.HalError):
        g.set_gadget_form(2)


@pytest.mark.parametrize("r,new_len", [(17, 16), (17, 3), (9, 8), (3, 1), (17, 17)])
def test_mod_switch_parity(gpu, cpu, rings, bconv_form, r, new_len):
    """Dynamic scaling (bfv.rs:628-662) at uneven levels, bit-exact, with unreduced (-2p, 2p)
    inputs in part: one kernel in the fused conversion forms, the HAL sequence in form 0."""
    g, c = rings("field_36864", (960, 980), 76)
    bconv_form(g)
    base = list(range(r))
    rng = np.random.default_rng(100 + r)
    primes = [g.primes[i] for i in base]
    src = uniform_residues(rng, primes, (2, r, 12, 3072))
    src[1, :, 3, :256] -= np.array(primes, dtype=np.int64).reshape(-1, 1) * rng.integers(0, 2, (r, 256))
    dc = R.mod_switch_down(c, src.copy(), base, new_len)
    dg = gpu.hal_allocate_i64((2, new_len, 12, 3072))
    sg = to_dev(gpu, src)
    g.mod_switch_down(dg, sg, base, new_len)
    assert same(from_dev(gpu, dg), dc)
    assert same(from_dev(gpu, sg), src)  # the source is left unchanged


def test_galois_parity(gpu, cpu, rings):
    g, c = rings("field_36864", (960, 980), 76)
    base = [0, 1, 2]
    rng = np.random.default_rng(9)
    src = rand_i64(rng, (2, 3, 12, 3072))
    f = c.field
    for gel in [f.lq_cyclic[0][0], f.lq_cyclic[3][0], f.lq_cyclic[5][0] * f.lq_cyclic[1][0] % f.m, 2]:
        if np.gcd(gel, f.m) != 1:
            continue
        dc = np.zeros_like(src)
        c.apply_galois(dc, src, gel)
        dg = gpu.hal_allocate_i64(src.shape)
        g.apply_galois(dg, to_dev(gpu, src), gel)
        assert same(from_dev(gpu, dg), dc)


def test_transform_overlap_halves(gpu, cpu, rings):
    """A batch large enough for the two-stream half-batch overlap (14 elements, 70 MB): the same
    residues as with the overlap off, and as the oracle."""
    pkg = load()
    g, c = rings("field_36864", (960, 980), 76)
    base = g.initial_rns_base()
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(404)
    shape = (14, len(base), 12, 3072)
    data = uniform_residues(rng, primes, shape)
    other = uniform_residues(rng, primes, (1,) + shape[1:])
    mods = [gpu.hal_create_modulus(p) for p in primes]
    outs = []
    for on in (True, False):
        g.set_overlap(on)
        try:
            dg = to_dev(gpu, data)
            g.pb_to_crt_tensor(dg, base)
            gpu.hal_mul_mont64(dg, to_dev(gpu, other), mods)
            g.crt_to_pb_tensor(dg, base)
            outs.append(from_dev(gpu, dg))
        finally:
            g.set_overlap(False)
    assert same(outs[0], outs[1])
    dc = data[[0, 13]].copy()
    c.pb_to_crt_tensor(dc, base)
    cpu.hal_mul_mont64(dc, other, c.mods_for(base))
    c.crt_to_pb_tensor(dc, base)
    assert same(outs[0][[0, 13]], dc)


def test_transform_round_trip_full_batch(gpu, cpu, rings):
    """The benchmark's full batch ([128, 17, 12, 3072], 642 MB): CRT->pb after pb->CRT gives
    every residue back mod p (a size-independent property), and three elements spread over both
    halves of the two-stream overlap equal the oracle's transform bit for bit."""
    g, c = rings("field_36864", (960, 980), 76)
    base = g.initial_rns_base()
    primes = np.array([g.primes[i] for i in base], dtype=np.int64).reshape(1, -1, 1, 1)
    rng = np.random.default_rng(128)
    data = uniform_residues(rng, primes.reshape(-1).tolist(), (128, len(base), 12, 3072))
    dg = to_dev(gpu, data)
    g.pb_to_crt_tensor(dg, base)
    crt = from_dev(gpu, dg)
    picks = [0, 63, 127]
    dc = data[picks].copy()
    c.pb_to_crt_tensor(dc, base)
    assert same(crt[picks], dc)
    g.crt_to_pb_tensor(dg, base)
    back = from_dev(gpu, dg)
    assert (np.mod(back, primes) == data).all()


@pytest.mark.parametrize("form", [3, 2, 1, 0])
def test_bconv_preserve_src(gpu, cpu, rings, form):
    """csd_ring_bconv with preserve_src leaves the source unchanged in every conversion form, with
    the same residues as the destroying call"""
    g, c = rings("field_36864", (960, 980), 76)
    base = g.initial_rns_base()
    ext = g.extension_rns_base_for(base)
    rng = np.random.default_rng(77)
    src = uniform_residues(rng, [g.primes[i] for i in base], (2, len(base), 12, 3072))
    dc = np.zeros((2, len(ext), 12, 3072), dtype=np.int64)
    c.bconv_tensors(dc, ext, src.copy(), base)
    g.set_bconv_form(form)
    try:
        sg = to_dev(gpu, src)
        dg = gpu.hal_allocate_i64(dc.shape)
        g.bconv_tensors(dg, ext, sg, base, preserve_src=True)
        assert same(from_dev(gpu, dg), dc)
        assert same(from_dev(gpu, sg), src)
    finally:
        g.set_bconv_form(2)


def test_forms_are_per_ring_and_context(gpu, cpu):
    """Two contexts (and their rings) in one process hold different execution forms, used
    interleaved: ring A with the HAL-sequence conversions, ring B fused; context A with the
    one-stage-per-pass NTT kernel, context B with the compiled shapes. Every call keeps its own
    context's form (the HAL sequence destroys the source, the fused form does not) and all give the
    oracle's residues."""
    pkg = load()
    be_a, be_b = pkg.HipBackend(0), pkg.HipBackend(0)
    ra = pkg.ring.RingContext(be_a, "field_192", (200, 240), 8)
    rb = pkg.ring.RingContext(be_b, "field_192", (200, 240), 8)
    oc = R.RingContext(cpu, R.FieldDescriptor("field_192"), (200, 240), 8)
    try:
        ra.set_bconv_form(0)
        be_a.ntt_force_generic(1)
        assert ra.bconv_form() == 0 and rb.bconv_form() == 2
        base = ra.initial_rns_base()
        ext = ra.extension_rns_base_for(base)
        n2, n3 = ra.info["deg_lprime"], ra.info["deg_k"]
        rng = np.random.default_rng(5)
        for it in range(3):
            src = uniform_residues(rng, [ra.primes[i] for i in base], (2, len(base), n2, n3))
            want = np.zeros((2, len(ext), n2, n3), dtype=np.int64)
            oc.bconv_tensors(want, ext, src.copy(), base)
            wt = src.copy()
            oc.pb_to_crt_tensor(wt, base)
            for be, ring, destroys in [(be_a, ra, True), (be_b, rb, False)]:
                sg = to_dev(be, src)
                dg = be.hal_allocate_i64(want.shape)
                ring.bconv_tensors(dg, ext, sg, base)
                assert same(from_dev(be, dg), want)
                assert same(from_dev(be, sg), src) != destroys
                tg = to_dev(be, src)
                ring.pb_to_crt_tensor(tg, base)
                assert same(from_dev(be, tg), wt)
    finally:
        ra.close()
        rb.close()
        be_a.close()
        be_b.close()
