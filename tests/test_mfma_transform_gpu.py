"""The int8-MFMA form of the period<->CRT transforms (mfma_transform.MfmaTransforms: the K part as
F_64, F_48, a diagonal and their inverses on the matrix cores, the L' part through the ring's NTT
passes) against the oracle's transforms: equal mod p in every residue (level L1), canonical in
[0, p); also pb->CRT->pb returns the input mod p."""
import numpy as np
import pytest

from _pkg import load
from oracle import ring as R
from test_hal_gpu import from_dev, to_dev
from test_ring_gpu import uniform_residues

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    pkg = load()
    be = pkg.HipBackend(0)
    from oracle import cpu_backend

    cpu = cpu_backend.CpuBackend()
    g = pkg.ring.RingContext(be, "field_36864", (960, 980), 76)
    c = R.RingContext(cpu, R.FieldDescriptor("field_36864"), (960, 980), 76)
    base = g.initial_rns_base()
    mf = pkg.mfma_transform.MfmaTransforms(g, base)
    yield pkg, be, g, c, base, mf
    del mf
    g.close()
    be.close()


def _equal_mod_p(a, b, primes):
    pv = np.array(primes, dtype=object).reshape(1, -1, 1, 1)
    return (((a.astype(object) - b.astype(object)) % pv) == 0).all()


@pytest.mark.parametrize("inverse", [False, True])
def test_mfma_transform_matches_oracle_mod_p(setup, inverse):
    pkg, be, g, c, base, mf = setup
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(11 + inverse)
    x = uniform_residues(rng, primes, (3, len(base), 12, 3072))
    x[0, :, 0, :5] = 0
    want = x.copy()
    (c.crt_to_pb_tensor if inverse else c.pb_to_crt_tensor)(want, base)
    d = to_dev(be, x)
    (mf.crt_to_pb_tensor if inverse else mf.pb_to_crt_tensor)(d)
    got = from_dev(be, d)
    assert _equal_mod_p(got, want, primes)
    pv = np.array(primes, dtype=np.int64).reshape(1, -1, 1, 1)
    assert (got >= 0).all() and (got < pv).all()


def test_mfma_transform_round_trip(setup):
    pkg, be, g, c, base, mf = setup
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(5)
    x = uniform_residues(rng, primes, (2, len(base), 12, 3072))
    d = to_dev(be, x)
    mf.pb_to_crt_tensor(d)
    mf.crt_to_pb_tensor(d)
    assert _equal_mod_p(from_dev(be, d), x, primes)
