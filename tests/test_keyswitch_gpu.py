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
        for b in range(batch):
            ct = np.zeros((2, r, n2, n3), dtype=np.int64)
            ct[0] = c0[b]
            exp = R.hom_gal_one(c, ct, base, gel, parts[b * digits:(b + 1) * digits].copy(), key_list, dig)
            assert same(got[(t * batch + b) * 2:(t * batch + b) * 2 + 2], exp), (t, b)
    g.close()
