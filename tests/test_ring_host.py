"""CPU-only checks of the product library: it loads, exports every symbol its headers declare,
and its host-side ring setup (field descriptors, Galois shifts, RNS prime sampling) agrees with
the oracle. No device work is started here."""
import os
import re

import pytest

from _pkg import ROOT, load
from oracle import ring as R

FIELDS = ["field_2", "field_16", "field_18", "field_32", "field_192", "field_24576", "field_32768", "field_32768_ex",
          "field_36864", "field_49152", "field_49152_ex", "field_65536"]


def declared_symbols():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^(?:int|const char\*)\s+(csd_\w+)\(", text, flags=re.M)
    return names


def test_library_exports_every_declared_symbol():
    pkg = load()
    lib = pkg.lib()
    names = declared_symbols()
    assert len(names) >= 50
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("name", FIELDS)
def test_field_info_matches_oracle(name):
    ring = load().ring
    info = ring.field_info(name)
    f = R.FieldDescriptor(name)
    assert info["m"] == f.m
    assert (info["deg_l"], info["deg_k"], info["deg_lprime"]) == (f.deg_l, f.deg_k, f.deg_lprime)
    assert info["gal_ml"] == len(f.gal_ml)
    assert info["rou_order"] == f.rou_order()
    assert info["orders"] == f.lq_cyclic_orders()
    assert info["lprime_factor_count"] == f.lprime_factor_count


@pytest.mark.parametrize("name", ["field_18", "field_192", "field_36864", "field_49152"])
def test_gal_to_shift_matches_oracle(name):
    ring = load().ring
    f = R.FieldDescriptor(name)
    tr = f.lq_transversal()
    for i in list(range(0, len(tr), max(1, len(tr) // 50))) + [len(tr) - 1]:
        g = tr[i] * f.gal_ml[-1] % f.m  # another representative of the same coset
        assert ring.gal_to_shift(name, g) == f.gal_to_shift(g)


@pytest.mark.parametrize("name,log2_q,t_bits", [("field_2", (200, 230), 0), ("field_18", (100, 120), 4),
                                                ("field_36864", (960, 980), 76), ("field_49152", (1280, 1300), 76)])
def test_prime_sampler_matches_oracle(name, log2_q, t_bits):
    ring = load().ring
    primes, init = ring.sample_primes(name, log2_q, t_bits)
    f = R.FieldDescriptor(name)
    base = R.extend_sampled_primes([], log2_q[0], log2_q[1], 57, f.rou_order())
    ext = R.extend_sampled_primes(base, 2 * log2_q[1] + 60 + t_bits, 2 * log2_q[1] + 90 + t_bits, 57, f.rou_order())
    assert primes == ext
    assert init == len(base)


def test_benchmark_bases():
    """field_36864 at 960..980 bits: 17 initial primes, 37 in all; field_49152 at 1280..1300 bits:
    23 initial primes (the configurations of the full-size reference tests, bootstrap/mod.rs:467-554)."""
    ring = load().ring
    p, init = ring.sample_primes("field_36864", (960, 980), 76)
    assert init == 17 and len(p) == 37
    p, init = ring.sample_primes("field_49152", (1280, 1300), 76)
    assert init == 23


@pytest.mark.parametrize("name,count,jp", [("field_49152", 40, [4, 64, 192]), ("field_36864", 27, [4, 6, 24, 64]),
                                           ("field_18", 5, [6, 3])])
def test_homcrt_rotation_sets(name, count, jp):
    """required_galois_elements (homcrt.rs:336-349): |J_p| = (p-1)/|I_p ∩ Gal(M|L)| per p | m_K and
    the BSGS split (homcrt.rs:320-334); every element is a unit mod m and maps to a nonzero shift."""
    ring = load().ring
    gs, orders = ring.homcrt_galois_elements(name)
    assert orders == jp
    assert len(gs) == count == len(set(gs))
    f = R.FieldDescriptor(name)
    import math
    for g in gs:
        assert math.gcd(g, f.m) == 1
        assert any(s != 0 for s in f.gal_to_shift(g))
    # baby steps of one factor are consecutive powers of one generator: for |J_p| = 64 the split is
    # 13 baby steps x 5 giant steps (12 powers plus the giant step power 13)
    if name == "field_49152":
        def split(n):
            best = min(range(1, n + 1), key=lambda c: c + 3 * ((n - 1) // c + 1))
            return best, (n - 1) // best + 1
        assert [split(n) for n in jp] == [(4, 1), (13, 5), (24, 8)]


def test_host_errors():
    pkg = load()
    with pytest.raises(pkg.HalError):
        pkg.ring.field_info("field_7")
    with pytest.raises(pkg.HalError):
        pkg.ring.sample_primes("field_2", (230, 200), 0)


def test_floordiv_schedule_matches_oracle():
    """the device package's floor-division schedule (host logic) equals the oracle's, which is
    pinned against the reference's noise model"""
    from oracle import dbfv as OD
    from oracle import digitex as OX
    from oracle import ring as R

    pkg = load()
    for field, init, final, r in [("field_49152", (8, 76), (8, 64), 23), ("field_36864", (8, 76), (8, 64), 17),
                                  ("field_192", (4, 16), (4, 12), 14), ("field_192", (8, 36), (8, 24), 16),
                                  ("field_49152", (8, 8), (8, 7), 13)]:
        got = pkg.digitex.floordiv_schedule(field, pkg.dbfv.DBFVModulus(*init), pkg.dbfv.DBFVModulus(*final), r)
        want = OX.floordiv_schedule(R.FieldDescriptor(field), OD.DBFVModulus(*init), OD.DBFVModulus(*final), r)
        assert got == want, field


@pytest.mark.parametrize("name", ["field_18", "field_192", "field_36864", "field_49152"])
def test_homcrt_factors_match_oracle(name):
    """csd_field_homcrt_factors and the key set of the homomorphic CRT transform (homcrt.rs:336-349)
    against the oracle's per-prime data (field.rs:217-243)"""
    from oracle import homcrt as OH

    pkg = load()
    f = R.FieldDescriptor(name)
    assert pkg.ring.homcrt_factors(name) == OH.homcrt_factors(f)
    want = sorted({g for _, gen, j in OH.homcrt_factors(f) for g in OH.factor_galois_elements(f.m, gen, j)})
    assert pkg.homcrt.required_galois_elements(name) == want
    assert pkg.ring.homcrt_galois_elements(name)[0] == want
