"""The batched Galois key switch (csd_ring_galois_keyswitch: gather by sigma_g, inner product with
HBM-resident key digits, + sigma_g(c0), wide reduction, all fused) is bit-exact with the oracle's
unfused reference sequence apply_galois -> inner_product -> add -> reduce_wide (bfv.rs:816-839)."""
import numpy as np
import pytest

from _pkg import load
from oracle import ring as R
from test_hal_gpu import from_dev, same, to_dev
from test_ring_gpu import uniform_residues

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    be = load().HipBackend(0)
    yield be
    be.close()


@pytest.mark.parametrize("field,log2_q,t_bits,r,digits,batch,ngal", [
    ("field_18", (100, 120), 4, 2, 2, 3, 3),
    ("field_49152", (1280, 1300), 76, 3, 3, 10, 2),
    ("field_36864", (960, 980), 76, 17, 9, 2, 1),
    ("field_36864", (960, 980), 76, 17, 9, 8, 1),  # one full group of 8 ciphertexts, 9 digits deep
    # the C5 shape: field_49152's 23 primes with one gadget digit per prime (the bootstrap's Galois
    # keys, bootstrap/mod.rs:531), a full group of 8 ciphertexts plus a partial one; checked on the
    # first and last ciphertext of the full group and on the partial group's
    ("field_49152", (1280, 1300), 76, 23, 23, 9, 1),
])
def test_galois_keyswitch_parity(gpu, cpu, field, log2_q, t_bits, r, digits, batch, ngal):
    pkg = load()
    g = pkg.ring.RingContext(gpu, field, log2_q, t_bits)
    c = R.RingContext(cpu, R.FieldDescriptor(field), log2_q, t_bits)
    assert g.primes == c.primes
    base = list(range(r))
    primes = [g.primes[i] for i in base]
    f = c.field
    n2, n3 = f.deg_lprime, f.deg_k
    rng = np.random.default_rng(77)
    parts = uniform_residues(rng, primes, (batch * digits, r, n2, n3))
    c0 = uniform_residues(rng, primes, (batch, r, n2, n3))
    keys = [uniform_residues(rng, primes, (2 * digits, r, n2, n3)) for _ in range(ngal)]
    # rotations: generators of the cyclic factors and a product of two of them
    gens = [x for x, _ in f.lq_cyclic]
    gs = [gens[0], gens[-1], gens[0] * gens[-1] % f.m][:ngal]
    out = gpu.hal_allocate_i64((ngal * batch * 2, r, n2, n3))
    g.galois_keyswitch(base, batch, digits, to_dev(gpu, parts), to_dev(gpu, c0), gs, [to_dev(gpu, k) for k in keys], out)
    got = from_dev(gpu, out)
    dig = [(i, i + 1) for i in range(digits)]  # only the count matters for the inner product
    for t, gel in enumerate(gs):
        key_list = [keys[t][2 * i:2 * i + 2] for i in range(digits)]
        for b in (range(batch) if batch <= 8 else [0, 7, batch - 1]):
            ct = np.zeros((2, r, n2, n3), dtype=np.int64)
            ct[0] = c0[b]
            exp = R.hom_gal_one(c, ct, base, gel, parts[b * digits:(b + 1) * digits].copy(), key_list, dig)
            assert same(got[(t * batch + b) * 2:(t * batch + b) * 2 + 2], exp), (t, b)
    g.close()


@pytest.mark.parametrize("field,log2_q,t_bits,r,digits,batch,ngal", [
    ("field_18", (100, 120), 4, 2, 2, 3, 3),
    ("field_49152", (1280, 1300), 76, 3, 3, 4, 5),
])
def test_streamed_keys_match_resident(gpu, field, log2_q, t_bits, r, digits, batch, ngal):
    """Keys in pinned host memory, streamed through the two staging buffers, give the same
    outputs as HBM-resident keys; an odd rotation count exercises both buffers and the tail."""
    pkg = load()
    g = pkg.ring.RingContext(gpu, field, log2_q, t_bits)
    base = list(range(r))
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    parts = gpu.hal_allocate_i64((batch * digits, r, n2, n3))
    c0 = gpu.hal_allocate_i64((batch, r, n2, n3))
    g.fill_uniform(parts, base, 1)
    g.fill_uniform(c0, base, 2)
    gs, _ = pkg.ring.homcrt_galois_elements(field)
    gs = gs[:ngal]
    keys = [gpu.hal_allocate_i64((2 * digits, r, n2, n3)) for _ in gs]
    host = []
    for t, k in enumerate(keys):
        g.fill_uniform(k, base, 100 + t)
        h = gpu.host_alloc(k.shape)
        h.fill_from(k)
        host.append(h)
    want = gpu.hal_allocate_i64((len(gs) * batch * 2, r, n2, n3))
    got = gpu.hal_allocate_i64(want.shape)
    g.galois_keyswitch(base, batch, digits, parts, c0, gs, keys, want)
    g.galois_keyswitch_streamed(base, batch, digits, parts, c0, gs, host, got)
    assert same(from_dev(gpu, got), from_dev(gpu, want))
    # host keys must come from csd_host_alloc blocks
    class Fake:
        ptr = host[0].array.ctypes.data + 8 * host[0].array.size  # one past the block
    with pytest.raises(pkg.HalError):
        g.galois_keyswitch_streamed(base, batch, digits, parts, c0, gs[:1], [Fake()], got)
    for h in host:
        h.free()
    g.close()
