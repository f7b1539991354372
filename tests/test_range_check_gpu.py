"""The debug range check of kernel outputs (csd_hal_range_check, RingContext.set_range_check): the
count of entries beyond k N per limb, and the dBFV layer run with every produced batch checked
against its tracked bound (the bookkeeping of RingEls.bound, ring.rs:33-47). The out-of-range
entries are written from the host; every kernel exits normally."""
import numpy as np
import pytest

from _pkg import load
from oracle import dbfv as OD
from oracle.ring import select_digits as OD_select
from test_dbfv_gpu import gpu, inputs, moduli, rings  # noqa: F401 (fixtures)
from test_hal_gpu import from_dev, same, to_dev
from test_ring_gpu import uniform_residues

pytestmark = pytest.mark.gpu


def test_range_check_counts(gpu, rings):  # noqa: F811
    """entries with |x| > k p are counted per limb's own p, both signs, the boundary k p kept"""
    g, _ = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    primes = [g.primes[i] for i in base]
    x = uniform_residues(np.random.default_rng(3), primes, (2, len(base), g.info["deg_lprime"], g.info["deg_k"]))
    t = to_dev(gpu, x)
    mods = g.mods_for(base)
    assert gpu.range_check(t, mods, 1) == 0
    x[0, 0, 0, 0] = primes[0] + 1           # beyond p
    x[1, -1, 1, 5] = -(primes[-1] + 7)      # beyond p, negative
    x[1, 1, 1, 1] = 2 * primes[1]           # beyond p, exactly 2 p
    x[0, 1, 0, 3] = -primes[1]              # exactly p: kept
    t = to_dev(gpu, x)
    assert gpu.range_check(t, mods, 1) == 3
    assert gpu.range_check(t, mods, 2) == 0
    assert gpu.range_check(t[1:2, 1:], mods[1:], 1) == 2  # a strided view
    with pytest.raises(ArithmeticError):
        g.check_range(t, base, 1, "planted")
    with pytest.raises(load().HalError):
        gpu.range_check(t, mods[1:], 1)  # one modulus per limb


def test_dbfv_layer_under_range_checks(gpu, cpu, rings):  # noqa: F811
    """a dBFV multiplication (reduction, lift, convolution, rescale, relinearization) and a
    reduction with the debug mode on: no entry beyond its tracked bound, and the residues those
    of the oracle"""
    pkg = load()
    g, c = rings("field_192", (200, 220), 64)
    base = g.initial_rns_base()
    mg, mc = moduli(16, 64)
    lhs = inputs(g, base, 8, 71, 2)
    rhs = inputs(g, base, 6, 72, 1)
    digits = OD_select(2, len(base))
    rng = np.random.default_rng(73)
    keys = [uniform_residues(rng, [g.primes[i] for i in base], (2, len(base), g.info["deg_lprime"], g.info["deg_k"]))
            for _ in digits]
    oc, _ = OD.dbfv_hom_mul(c, mc, lhs.copy(), 2, rhs.copy(), 1, base, keys, digits, "pb")
    g.set_range_check(True)
    try:
        og, bg = pkg.dbfv.dbfv_hom_mul(g, mg, to_dev(gpu, lhs), 2, to_dev(gpu, rhs), 1, base,
                                       [to_dev(gpu, k) for k in keys], digits, "pb")
        rg, _ = pkg.dbfv.dbfv_reduce(g, mg, to_dev(gpu, lhs), 2, base)
    finally:
        g.set_range_check(False)
    assert bg == 2 and same(from_dev(gpu, og), np.asarray(oc))
    rc, _ = OD.dbfv_reduce(c, mc, lhs.copy(), 2, base)
    assert same(from_dev(gpu, rg), np.asarray(rc))
