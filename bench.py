#!/usr/bin/env python3
"""Benchmark of the RNS-ring hot path on one MI355X: period<->CRT transforms per second at
field_36864 (n = 36864 = 12 x 3072, full-depth RNS base of 17 primes p = 1 mod 1638960960,
p < 2^57), batch of 64 ciphertexts = 128 ring elements, tensor [128, 17, 12, 3072] int64
(641.7 MB) resident in HBM. One step = pb->CRT followed by CRT->pb of the whole batch
(2 x 128 = 256 transforms), the reference's pb_to_crt_tensor / crt_to_pb_tensor
(cyclosubfield-dbfv/src/ring.rs:1117-1161) on the device.

Prints one JSON line. `value` is transforms/s of the fastest bit-exact device form; the plain
one-kernel-per-HAL-call form is reported beside it. `roofline` is the dominant kernel's
algorithmic HBM bytes per launch over its HIP-event-timed average duration; `cpu_baseline` is
the oracle (the C++ restatement of hal::cpu, single thread) timed on a bounded sample.
"""
import argparse
import ctypes
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FIELD = "field_36864"
LOG2_Q = (960, 980)
T_BITS = 76
SEED = 20261203
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent pair recorded on a given stream (the library's context stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.e0, self.e1 = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(self.e0)) == 0
        assert self.hip.hipEventCreate(ctypes.byref(self.e1)) == 0

    def start(self, stream):
        assert self.hip.hipEventRecord(self.e0, ctypes.c_void_p(stream)) == 0

    def stop(self, stream):
        assert self.hip.hipEventRecord(self.e1, ctypes.c_void_p(stream)) == 0
        assert self.hip.hipEventSynchronize(self.e1) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), self.e0, self.e1) == 0
        return ms.value


def uniform_batch(primes, count, deg_lp, deg_k, seed):
    rng = np.random.default_rng(seed)
    out = np.empty((count, len(primes), deg_lp, deg_k), dtype=np.int64)
    for j, p in enumerate(primes):
        out[:, j] = rng.integers(0, p, size=(count, deg_lp, deg_k), dtype=np.int64)
    return out


# the sources the headline transform kernels are built from; the committed counter summaries
# record their hash and are used only while it matches
PROFILED_SOURCES = ("cyclosubfield-dbfv_amd/csrc/csd_ntt.hip", "cyclosubfield-dbfv_amd/csrc/csd_arith.hpp")


def profiled_sources_sha256():
    h = hashlib.sha256()
    for rel in PROFILED_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def profile_summary(name, count):
    """The committed rocprofv3 counter summary profiles/<name>, or None unless it was measured on a
    batch of `count` ring elements and on the kernel sources as they are now (same hash)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    if data.get("count") != count or data.get("source_sha256") != profiled_sources_sha256():
        return None
    return data


def traffic_from_profiles(kernel, count):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if it applies."""
    data = profile_summary("hbm_traffic.json", count)
    entry = None if data is None else data.get(kernel)
    return None if entry is None else entry.get("bytes_per_launch")


def valu_issue_from_profiles(kernel, avg_ms, count):
    """The VALU issue fraction of `kernel` at `avg_ms` per launch: its wave instructions per launch
    (the committed rocprofv3 SQ_INSTS_VALU summary) x 4 issue cycles, over the 1024 SIMDs at 2.4 GHz
    for the launch time; None when the summary is absent or does not apply."""
    data = profile_summary("valu_counters.json", count)
    entry = None if data is None else data.get(kernel)
    if entry is None or not avg_ms:
        return None
    cycles = entry["sq_insts_valu"] * data["cycles_per_wave_inst"] / data["simds"]
    return {"wave_insts_per_launch": entry["sq_insts_valu"],
            "frac": round(cycles / (data["clock_GHz"] * 1e9 * avg_ms / 1000.0), 4),
            "note": "VALU issue cycles per SIMD over the launch time at 2.4 GHz: the bound of this kernel"}


def cpu_threads():
    """host threads for the CPU baseline: the cores this process may use, at most 16"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(primes_base, host_data, sample, threads, min_seconds=10.0):
    """The oracle (the C++ restatement of hal::cpu) on a bounded sample: pb->CRT + CRT->pb of
    `sample` elements of the batch, first on one thread, then on `threads` threads over disjoint
    elements (every op is independent per element; the C++ calls release the GIL), repeated until
    at least `min_seconds` of wall time."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import cpu_backend
    from oracle import ring as oring

    cpu = cpu_backend.CpuBackend()
    ring = oring.RingContext(cpu, oring.FieldDescriptor(FIELD), LOG2_Q, T_BITS)
    base = ring.initial_rns_base()
    assert [ring.primes[i] for i in base] == primes_base
    ring.transform_cache(base)
    x = np.ascontiguousarray(host_data[:1])
    t0 = time.perf_counter()
    ring.pb_to_crt_tensor(x, base)
    ring.crt_to_pb_tensor(x, base)
    one = 2 / (time.perf_counter() - t0)
    chunks = [np.ascontiguousarray(host_data[i:i + 1]) for i in range(sample)]

    def work(c):
        ring.pb_to_crt_tensor(c, base)
        ring.crt_to_pb_tensor(c, base)

    done, t0 = 0, time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        while True:
            list(ex.map(work, chunks))
            done += 2 * len(chunks)
            dt = time.perf_counter() - t0
            if dt >= min_seconds:
                break
    return {"value": round(done / dt, 3), "unit": "transforms/s", "cores": threads, "kind": "port",
            "sample": f"{sample} ring elements of the same batch, pb->CRT + CRT->pb on {threads} threads over disjoint "
                      f"elements, {done} transforms in {dt:.1f} s; one thread: {one:.1f} transforms/s"}


def _zero_secret_rk(pkg, be, ring, base, ndigits):
    """relinearization key of the zero secret key (gadget digit i: rows (0, g_i) in CRT basis, g_i
    the CRT recomposition factor of the digit), so that a Delta-scaled plaintext in c0 decrypts
    as round(t c0 / Q) through every product"""
    digits = pkg.bfv.select_digits(ndigits, len(base))
    q = ring.modulus_of(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    keys = []
    for a, b in digits:
        di = ring.modulus_of(base[a:b])
        g = (q // di) * pow((q // di) % di, -1, di) % q
        host = np.zeros((2, len(base), n2, n3), dtype=np.int64)
        for j, i in enumerate(base):
            p = ring.primes[i]
            host[1, j] = (g % p) * ((1 << 64) % p) % p
        k = be.hal_allocate_i64(host.shape)
        be.hal_import(k, host)
        keys.append(k)
    return keys, digits


def _mont_table(val, p):
    """Montgomery representatives ((val + e) 2^64 mod p, in [0, p)) for e = -3 .. 3"""
    r = (1 << 64) % p
    return np.array([(val + e) % p * r % p for e in range(-3, 4)], dtype=np.int64)


def _decrypt_scalar_coeffs(pkg, be, ring, modulus, ct, c, ncoef):
    """dbfv_dec (dbfv.rs:656-680) under the zero secret key of ciphertext c of a period-basis batch,
    on its first ncoef coefficients (round(t c0 / Q) per digit, recombined)"""
    D = pkg.dbfv
    base = ct.base
    primes = [ring.primes[i] for i in base]
    q = ring.modulus_of(base)
    t = modulus.t()
    total = [0] * ncoef
    for i in range(ct.digits()):
        row = ct.data[2 * i * ct.count + c:2 * i * ct.count + c + 1]
        h = be.hal_export(row).reshape(len(base), -1)[:, :ncoef]
        for k in range(ncoef):
            x = 0
            for j, p in enumerate(primes):
                v = int(h[j, k]) * pow(1 << 64, -1, p) % p
                m = q // p
                x += v * m * pow(m % p, -1, p)
            x %= q
            x = x - q if x > q // 2 else x
            part = D.cmod(D.rounded_div(x * t, q), t)
            total[k] += D.cmod(part, 1 << (modulus.log2_t - i * modulus.log2_b)) << (i * modulus.log2_b)
    vals = set(D.cmod(v, t) for v in total)
    return vals


def bench_bootstrap_step(pkg, be, batch=32, seed=SEED, cpu=True):
    """The dBFV bootstrapping step at field_49152 (m = 35297665, n = 49152, 23-prime base, t_bits
    76) on `batch` ciphertexts, end to end on the device: homcrt's Galois key switching over the
    full rotation set (40 rotations, 23 gadget digits, keys resident in HBM, 16.6 GB;
    bfv.rs:777-841, homcrt.rs:336-429) followed by the digit extraction of the bootstrap,
    hom_floordiv_dbfv from the intermediate modulus (b = 2^8, t = 2^76) to t = 2^64 (v = 12 levels
    of the 8-product circuit, bootstrap/mod.rs:369-382, digitex.rs:267-367) with the relinearization
    key of ceil(23 / 2) = 12 digits (bootstrap/mod.rs:535). The key-switch inputs are uniform
    residues (hoisted digits and c0); the floor division's inputs are dBFV encryptions under the
    zero secret key of random 76-bit scalars, so the result's decryption is checked."""
    D = pkg.dbfv
    field, log2_q, t_bits = "field_49152", (1280, 1300), 76
    ring = pkg.ring.RingContext(be, field, log2_q, t_bits)
    base = ring.initial_rns_base()
    r = len(base)
    d = r
    gs, _ = pkg.ring.homcrt_galois_elements(field)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    stream = be.stream()
    ev = HipEvents()
    # key switching operands
    parts = be.hal_allocate_i64((batch * d, r, n2, n3))
    c0 = be.hal_allocate_i64((batch, r, n2, n3))
    keys = [be.hal_allocate_i64((2 * d, r, n2, n3)) for _ in gs]
    out = be.hal_allocate_i64((len(gs) * batch * 2, r, n2, n3))
    ring.fill_uniform(parts, base, seed + 1)
    ring.fill_uniform(c0, base, seed + 2)
    for t, k in enumerate(keys):
        ring.fill_uniform(k, base, seed + 100 + t)
    # floor-division operands: `batch` dBFV ciphertexts of 10 digits, digit-major
    init, final = D.DBFVModulus(8, 76), D.DBFVModulus(8, 64)
    rk, rk_digits = _zero_secret_rk(pkg, be, ring, base, (r - 1) // 2 + 1)
    rng = np.random.default_rng(seed + 3)
    msgs = [int(rng.integers(0, 1 << 62)) << 14 | int(rng.integers(0, 1 << 14)) for _ in range(batch)]
    q = ring.modulus_of(base)
    delta = D.rounded_div(q, init.t())
    sign = D.scale_q_to_m_period_basis(ring)
    nd = init.digits()
    # the floor division runs on chunks of `chunk` ciphertexts: its v + 1 i128 accumulators take
    # 4.7 GB per ciphertext at the full base and a product's temporaries about as much again, beside
    # the keys and the key switch's outputs
    chunk = min(batch, 8)
    nchunks = (batch + chunk - 1) // chunk
    chunks = []
    for k in range(nchunks):
        cnt = min(chunk, batch - k * chunk)
        cts = be.hal_allocate_i64((2 * nd * cnt, r, n2, n3))
        ring.fill_uniform(cts, base, seed + 4 + k)  # the c1 rows: uniform (a zero-secret decryption never reads them)
        for i in range(nd):  # c0 of digit i: Delta m + e for digit 0, e for the others, |e| <= 3
            host = np.empty((cnt, r, n2, n3), dtype=np.int64)
            for c in range(cnt):
                err = rng.integers(-3, 4, size=(n2, n3)) + 3
                val = D.cmod(msgs[k * chunk + c] * sign, init.t()) * delta if i == 0 else 0
                for j, idx in enumerate(base):
                    host[c, j] = _mont_table(val, ring.primes[idx])[err]
            be.hal_import(cts[2 * i * cnt:(2 * i + 1) * cnt], host)
            del host
        chunks.append(D.DBFVBatch(cts, base, cnt, "pb", 1))
    # warm-up (builds the conversion and transform caches of every level's base) on ciphertext 0
    first = chunks[0]
    one = D.DBFVBatch(be.hal_allocate_i64((2 * nd, r, n2, n3)), base, 1, "pb", 1)
    for row in range(2 * nd):
        be.hal_assign(one.data[row:row + 1], first.data[row * first.count:row * first.count + 1])
    single = pkg.digitex.hom_floordiv_dbfv(ring, init, final, one, rk, rk_digits)
    ring.galois_keyswitch(base, batch, d, parts, c0, gs, keys, out)
    be.sync()
    # the timed step: every rotation of every ciphertext, then the floor division of the batch
    ev_ks = HipEvents()
    ev.start(stream)
    ev_ks.start(stream)
    ring.galois_keyswitch(base, batch, d, parts, c0, gs, keys, out)
    ks_ms = ev_ks.stop(stream)
    t0 = time.perf_counter()
    results = [pkg.digitex.hom_floordiv_dbfv(ring, init, final, ch, rk, rk_digits) for ch in chunks]
    ms = ev.stop(stream)
    fd_wall = time.perf_counter() - t0
    # checks: the batch gives ciphertext 0 the residues of its own floor division, and the results
    # decrypt to floor(m / 2^12) mod 2^64 (first 16 coefficients of the first and last ciphertext)
    res0 = results[0]
    for row in range(single.data.shape[0]):
        if not (be.hal_export(single.data[row:row + 1]) ==
                be.hal_export(res0.data[row * res0.count:row * res0.count + 1])).all():
            raise SystemExit("bench: batched floor division differs from the single-ciphertext one")
    for k, c in [(0, 0), (nchunks - 1, results[-1].count - 1)]:
        got = _decrypt_scalar_coeffs(pkg, be, ring, final, results[k], c, 16)
        want = {D.cmod(D.cmod((msgs[k * chunk + c] % init.t()) >> 12, final.t()) * sign, final.t())}
        if got != want:
            raise SystemExit(f"bench: floor division of ciphertext {k * chunk + c} decrypts to {got}, want {want}")
    del results, res0
    # per-kernel view of the key switch (HIP events around each launch)
    be.set_timing(True)
    ring.galois_keyswitch(base, batch, d, parts, c0, gs, keys, out)
    be.sync()
    be.set_timing(False)
    rep = be.timing_report()["galois_keyswitch"]
    avg_ms = rep[1] / rep[0]
    bytes_per_launch = rep[2] / rep[0]
    v, mults, _ = pkg.digitex.floordiv_schedule(field, init, final, r)
    products = v * mults
    res_json = {
        "workload": f"field_49152 dBFV bootstrapping step on {batch} ciphertexts: Galois key switching over the {len(gs)} "
                    f"homcrt rotations (r = {r} primes, {d} gadget digits, keys "
                    f"{len(gs) * 2 * d * r * n2 * n3 * 8 / 2**30:.1f} GiB resident in HBM), then hom_floordiv_dbfv "
                    f"t = 2^76 -> 2^64 (b = 2^8, 10 digits, v = 12 levels x 8 products, relinearization key of "
                    f"{len(rk_digits)} digits; chunks of {chunk} ciphertexts)",
        "value": round(ms, 1), "unit": "ms per step", "higher_is_better": False,
        "ciphertexts_per_s": round(batch / (ms / 1000.0), 2),
        "ms_per_ciphertext": round(ms / batch, 2),
        "key_switch": {"ms": round(ks_ms, 3), "value": round(len(gs) * batch / (ks_ms / 1000.0), 1),
                       "unit": "key switches/s",
                       "kernel": {"name": "galois_keyswitch", "avg_ms": round(avg_ms, 4),
                                  "bytes_per_launch": bytes_per_launch,
                                  "GB/s": round(bytes_per_launch / (avg_ms / 1000.0) / 1e9, 1),
                                  "frac": round(bytes_per_launch / (avg_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)}},
        "floor_division": {"ms": round(ms - ks_ms, 1), "dbfv_products": products * batch,
                           "ms_per_product_per_ciphertext": round((ms - ks_ms) / (products * batch), 3),
                           "host_wall_s": round(fd_wall, 2)},
        "checks": "batch == single-ciphertext floor division (bit-identical); decryptions of ciphertexts 0 and "
                  f"{batch - 1} under the zero secret key equal floor(m / 2^12) mod 2^64",
    }
    # the fallback for key sets beyond HBM: keys in pinned host memory, streamed per rotation
    nstream = min(8, len(gs))
    host = []
    for k in keys[:nstream]:
        h = be.host_alloc(k.shape)
        h.fill_from(k)
        host.append(h)
    out_s = be.hal_allocate_i64((nstream * batch * 2, r, n2, n3))
    ring.galois_keyswitch_streamed(base, batch, d, parts, c0, gs[:nstream], host, out_s)  # warm-up
    be.sync()
    ev.start(stream)
    ring.galois_keyswitch_streamed(base, batch, d, parts, c0, gs[:nstream], host, out_s)
    ms_s = ev.stop(stream)
    if not (be.hal_export(out_s[0:2]) == be.hal_export(out[0:2])).all():
        raise SystemExit("bench: streamed-key key switch differs from the resident one")
    key_bytes = 2 * d * r * n2 * n3 * 8
    res_json["key_switch"]["streamed_keys"] = {
        "rotations": nstream, "ms": round(ms_s, 3), "value": round(nstream * batch / (ms_s / 1000.0), 1),
        "unit": "key switches/s", "host_to_device_GB/s": round(nstream * key_bytes / (ms_s / 1000.0) / 1e9, 1),
        "note": "keys in pinned host memory, double-buffered over PCIe; the resident form reads them from HBM",
    }
    for h in host:
        h.free()
    del out_s
    if cpu:
        res_json["cpu_baseline"] = _bootstrap_cpu_baseline(pkg, be, ring, base, parts, c0, keys[0], out, gs[0], d,
                                                           first, batch, rk_digits, len(gs), products)
    del parts, c0, keys, out, chunks, first, single, one
    ring.close()
    return res_json


def _bootstrap_cpu_baseline(pkg, be, ring, base, parts, c0, key0, out, g0, d, first, batch, rk_digits, nrot, products):
    """The oracle on the host, one thread: one hoisted Galois key switch with the key read through
    keystore semantics (keystore.rs:50-127: written to an unlinked scratch file and read through a
    shared mapping; the page cache holds it, as for the reference after its first read), checked
    bit-exact against the device; and one dbfv_hom_square of a 10-digit ciphertext of the batch at
    the full base (the product the floor division runs 96 times per ciphertext). The step estimate
    scales both to the device step: batch x rotations key switches plus batch x 96 products."""
    import mmap
    import tempfile

    from oracle import cpu_backend
    from oracle import dbfv as OD
    from oracle import ring as oring

    r = len(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    cb = cpu_backend.CpuBackend()
    oc = oring.RingContext(cb, oring.FieldDescriptor("field_49152"), (1280, 1300), 76)
    hp = be.hal_export(parts[0:d]).reshape(d, r, n2, n3)
    hc = np.zeros((2, r, n2, n3), dtype=np.int64)
    hc[0] = be.hal_export(c0[0:1]).reshape(r, n2, n3)
    hk = be.hal_export(key0).reshape(2 * d, r, n2, n3)
    fd, path = tempfile.mkstemp(prefix="csd_keystore_")
    os.unlink(path)
    try:
        os.write(fd, hk.tobytes())
        os.fsync(fd)
        mm = mmap.mmap(fd, hk.nbytes, mmap.MAP_SHARED, mmap.PROT_READ)
        if hasattr(mm, "madvise") and hasattr(mmap, "MADV_SEQUENTIAL"):
            mm.madvise(mmap.MADV_SEQUENTIAL)
        kv = np.frombuffer(mm, dtype=np.int64).reshape(hk.shape)
        del hk
        t0 = time.perf_counter()
        exp = oring.hom_gal_one(oc, hc, base, g0, hp, [kv[2 * i:2 * i + 2] for i in range(d)],
                                [(i, i + 1) for i in range(d)])
        ks_dt = time.perf_counter() - t0
        del kv
        mm.close()
    finally:
        os.close(fd)
    if not (be.hal_export(out[0:2]).reshape(2, r, n2, n3) == exp).all():
        raise SystemExit("bench: galois key switch differs from the oracle")
    # one dBFV square of ciphertext 0 of the batch at the full base
    nd = first.digits()
    ct0 = np.empty((2 * nd, r, n2, n3), dtype=np.int64)
    for row in range(2 * nd):
        ct0[row] = be.hal_export(first.data[row * first.count:row * first.count + 1]).reshape(r, n2, n3)
    okeys = []
    q = oc.modulus_of(base)
    for a, b in rk_digits:
        di = oc.modulus_of(base[a:b])
        gq = (q // di) * pow((q // di) % di, -1, di) % q
        k = oc.allocate(2, base)
        for j, i in enumerate(base):
            k[1, j] = oring.montgomery_repr(gq % oc.primes[i], oc.primes[i])
        okeys.append(k)
    t0 = time.perf_counter()
    OD.dbfv_hom_square(oc, OD.DBFVModulus(8, 76), ct0, 1, base, okeys, rk_digits, "pb")
    sq_dt = time.perf_counter() - t0
    step_s = batch * nrot * ks_dt + batch * products * sq_dt
    return {"value": round(step_s * 1000.0, 1), "unit": "ms per step (estimate)", "cores": 1, "kind": "port",
            "key_switches/s": round(1.0 / ks_dt, 3), "dbfv_products/s": round(1.0 / sq_dt, 4),
            "sample": f"1 hoisted key switch (23 digits, key mmap'd from an unlinked scratch file: {ks_dt:.2f} s, checked "
                      f"bit-exact) and 1 dbfv_hom_square of a 10-digit ciphertext over 23 primes ({sq_dt:.1f} s), one "
                      f"thread; value = {batch} x {nrot} key switches + {batch} x {products} products at those rates "
                      f"(an extrapolation, the products of later levels run on shorter bases)"}


def bench_bootstrap36864(pkg, be, reps=3, batch=8, seed=31, cpu=True):
    """bootstrap_bfv_dbfv at field_36864 with the parameters of the reference's full-size test
    (bootstrap/mod.rs:461-509): 17-prime base (log2 q 960..980, t_bits 76), log2 alpha 960, dBFV
    modulus (8, 8) (t = 2^64, b = 2^8), intermediate 2^76 (v = 12), sparse key of weight 32 with
    amplification at most 10, Galois keys of 17 gadget digits for the 27 homcrt elements, a
    relinearization key of 9 digits; an encryption of 2025 under a generated secret key. Keys and
    transform coefficients are made beforehand (not timed). Timed: one ciphertext, end to end, with
    the BEHZ conversions' matrix products on the int8 matrix cores (bconv form 2) and on the VALU
    (form 1); the results must be bit-identical and decrypt to 2025 in every slot. Also the time
    per ciphertext of a batch of `batch` ciphertexts."""
    K, D, H, B = pkg.keys, pkg.dbfv, pkg.homcrt, pkg.bootstrap
    t_setup = time.perf_counter()
    ring = pkg.ring.RingContext(be, "field_36864", (960, 980), 76)
    ring.precompute_rns_factor_transforms()
    f = pkg.field.FieldDescriptor("field_36864")
    rng = K.Rng(seed)
    base = ring.initial_rns_base()
    la = 960.0
    modulus = D.DBFVModulus.new(8, 8)
    inter = D.DBFVModulus.new_incomplete(8, 76)
    sk = K.gen_sk(ring, rng)
    tr = H.UnsplitCRTTransform(inter.t(), f)
    encaps, _ = B.create_sparse_key_encaps(ring, f, inter, sk, 32, rng, la, max_log2_amplification=10.0)
    gdig = pkg.bfv.select_digits(len(base), len(base))
    gs = H.required_galois_elements("field_36864")
    gks = {g: K.gen_gk(ring, sk, g, gdig, rng, la) for g in gs}
    rk = K.gen_rk(ring, sk, pkg.bfv.select_digits((len(base) - 1) // 2 + 1, len(base)), rng, la)
    ct = D.DBFVBatch(K.bfv_enc(ring, base, modulus.t(), K.encode_scalar(ring, 2025), sk, rng, la), base, 1, "crt", 1)
    be.sync()
    setup_s = time.perf_counter() - t_setup

    def run(form, n_rep):
        ring.set_bconv_form(form)
        B.bootstrap_bfv_dbfv(ring, f, tr, modulus, ct, rk, gks, encaps)  # warm-up (caches, pool)
        times, phases, res = [], None, None
        for _ in range(n_rep):
            ph, last = [], [None]

            def tick(name):
                be.sync()
                x = time.perf_counter()
                ph.append((name, round((x - last[0]) * 1000.0, 2)))
                last[0] = x

            be.sync()
            t0 = last[0] = time.perf_counter()
            res = B.bootstrap_bfv_dbfv(ring, f, tr, modulus, ct, rk, gks, encaps, timer=tick)
            be.sync()
            times.append((time.perf_counter() - t0) * 1000.0)
            if phases is None or times[-1] == min(times):
                phases = dict(ph)
        return sorted(times)[len(times) // 2], phases, res

    ms_mfma, ph_mfma, res_mfma = run(2, reps)
    ms_plain, ph_plain, res_plain = run(1, reps)
    ms_all, _, res_all = run(3, reps)
    ring.set_bconv_form(2)
    same = all((be.hal_export(a) == be.hal_export(b)).all()
               for a, b in ((res_mfma.data, res_plain.data), (res_all.data, res_plain.data)))
    if not same:
        raise SystemExit("bench: the bootstrap's matrix-core and VALU conversion forms differ")
    dec = K.dbfv_dec(ring, modulus, res_mfma, sk)
    t = modulus.t()
    el = K.encode_pb_unscaled(ring, base, t, dec)
    ring.pb_to_crt_tensor(el, base)
    for gen, cs in tr.reduce_t(t).forward():
        el = H.apply_cyclic_transform_plain(ring, el, base, gen, cs)
    ring.crt_to_pb_tensor(el, base)
    sgn = f.scale_m_k_to_m_period_basis()
    slots = [D.cmod(v * sgn, t) for v in K.decode_pb_unscaled(ring, t, el, base)[:f.deg_k]]
    if slots != [2025] * f.deg_k:
        raise SystemExit("bench: the bootstrapped ciphertext does not decrypt to 2025 in every slot")
    # a batch of ciphertexts: time per ciphertext
    cts = [K.bfv_enc(ring, base, t, K.encode_scalar(ring, 2025 + c), sk, rng, la) for c in range(batch)]
    data = ring.allocate_uninit(2 * batch, base)
    for c, x in enumerate(cts):
        be.hal_assign(data[c:c + 1], x[0:1])
        be.hal_assign(data[batch + c:batch + c + 1], x[1:2])
    bct = D.DBFVBatch(data, base, batch, "crt", 1)
    be.sync()
    t0 = time.perf_counter()
    bres = B.bootstrap_bfv_dbfv(ring, f, tr, modulus, bct, rk, gks, encaps)
    be.sync()
    batch_ms = (time.perf_counter() - t0) * 1000.0
    for c in (0, batch - 1):
        if K.decode_scalar(ring, t, K.dbfv_dec(ring, modulus, bres, sk, c)) != 2025 + c:
            raise SystemExit("bench: a bootstrapped ciphertext of the batch decrypts wrongly")
    del cts, data, bct, bres
    h, u = tr.hom_cost()
    res_json = {
        "workload": "bootstrap_bfv_dbfv @field_36864 (the reference's full-size test, bootstrap/mod.rs:461-509): "
                    "17 primes, dBFV (8, 8) -> intermediate 2^76, v = 12, sparse weight 32, 27 Galois keys x 17 "
                    "digits, relinearization key of 9 digits; one ciphertext encrypting 2025",
        "value": round(ms_mfma, 1), "unit": "ms per bootstrap", "higher_is_better": False,
        "form": "mfma: the BEHZ conversions' matrix products on int8 MFMA where that measured faster (bconv form 2: "
                "conversions from 9 or more source primes, rescale steps from 17 or more; the VALU kernel below), "
                "every other step as in the plain form",
        "plain": {"ms": round(ms_plain, 1), "form": "plain: the conversions as fused VALU kernels (bconv form 1)"},
        "mfma_over_plain": round(ms_plain / ms_mfma, 3),
        "mfma_every_conversion": {"ms": round(ms_all, 1), "over_plain": round(ms_plain / ms_all, 3),
                                  "form": "every conversion's matrix product on int8 MFMA (bconv form 3)"},
        "phases_ms": ph_mfma, "phases_ms_plain": ph_plain,
        "batch": {"ciphertexts": batch, "ms": round(batch_ms, 1), "ms_per_ciphertext": round(batch_ms / batch, 1)},
        "checks": "the three forms bit-identical; the result decrypts to 2025 in all 3072 slots (extract_slot_values, "
                  "homcrt.rs:250-274); batch ciphertexts 0 and last decrypt to their plaintexts",
        "transform_cost": {"hoisted_automorphisms": h, "unhoisted_automorphisms": u},
        "setup_s": round(setup_s, 2),
        "reference_published": {"value": 6.3 * 60 * 1000, "unit": "ms", "note": "dBFV (t = 2^64, b = 2^8) bootstrap, "
                                "n = 36864, single-threaded CpuBackend on unstated hardware, the 'Very old' section of "
                                "cyclosubfield-benchmarks/benchmarks.md:94 (older code; context, not a target)"},
    }
    if cpu:
        res_json["cpu_baseline"] = _bootstrap36864_cpu_baseline(pkg, be, ring, base, rk, ct)
    ring.close()
    return res_json


def _bootstrap36864_cpu_baseline(pkg, be, ring, base, rk, ct):
    """The oracle, one thread: one dbfv_hom_square of a 10-digit ciphertext at the full 17-prime
    base with the relinearization key's digit layout (the product the floor division runs 96 times,
    and 96 x this is most of the bootstrap), scaled by 96; stated as an estimate."""
    from oracle import dbfv as OD
    from oracle import ring as oring
    from oracle import cpu_backend

    cb = cpu_backend.CpuBackend()
    oc = oring.RingContext(cb, oring.FieldDescriptor("field_36864"), (960, 980), 76)
    r = len(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    rng = np.random.default_rng(5)
    nd = 10
    x = np.empty((2 * nd, r, n2, n3), dtype=np.int64)
    for j, i in enumerate(base):
        x[:, j] = rng.integers(0, oc.primes[i], size=(2 * nd, n2, n3), dtype=np.int64)
    keys = [be.hal_export(k).reshape(2, r, n2, n3) for k in rk.keys()]
    t0 = time.perf_counter()
    OD.dbfv_hom_square(oc, OD.DBFVModulus(8, 76), x, 1, base, keys, rk.digits, "pb")
    sq = time.perf_counter() - t0
    return {"value": round(96 * sq * 1000.0, 1), "unit": "ms per bootstrap (extrapolated from the floor division's products)",
            "cores": 1, "kind": "port", "dbfv_products/s": round(1.0 / sq, 4),
            "sample": f"1 dbfv_hom_square of a 10-digit ciphertext over 17 primes ({sq:.1f} s, one thread); value = 96 x "
                      f"that (the 12 x 8 products of hom_floordiv_dbfv, all priced at the full base although the later "
                      f"levels run on shorter ones; the transforms, key switches and accumulator steps are left out): "
                      f"an extrapolation, not a measured CPU bootstrap"}


def bench_batch(pkg, be, steps, count=256, seed=SEED, cpu=True):
    """Uneven batch at field_36864: `count` ciphertexts with limb counts uniform in 3..17 (the
    levels dynamic scaling leaves), residues uniform mod each prime, through pb->CRT, a CRT-basis
    plaintext product, CRT->pb and mod_switch_down by one limb. The plain form issues the
    per-ciphertext call sequence (ring.rs:1118-1161, bfv.rs:628-662); the batched form runs the
    autotuned work-list plan as one hipGraph replay per step. Both must agree bit for bit."""
    ring = pkg.ring.RingContext(be, FIELD, LOG2_Q, T_BITS)
    base = ring.initial_rns_base()
    R = len(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    npos = n2 * n3
    limbs = [int(x) for x in np.random.default_rng(seed + 3).integers(3, R, size=count, endpoint=True)]
    # one arena per side; ciphertext c is the contiguous [2, r_c, n2, n3] block at its offset
    ct_arena = be.hal_allocate_i64((sum(2 * r for r in limbs), 1, n2, n3))
    out_arena = be.hal_allocate_i64((sum(2 * (r - 1) for r in limbs), 1, n2, n3))
    cts, outs = [], []
    oc = oo = 0
    for r in limbs:
        cts.append(pkg.DeviceTensor(be, ct_arena.ptr + oc * 8, (2, r, n2, n3), (r * npos, npos, n3, 1), 8, ct_arena))
        outs.append(pkg.DeviceTensor(be, out_arena.ptr + oo * 8, (2, r - 1, n2, n3), ((r - 1) * npos, npos, n3, 1), 8,
                                     out_arena))
        oc += 2 * r * npos
        oo += 2 * (r - 1) * npos
    pt = be.hal_allocate_i64((1, R, n2, n3))
    ring.fill_uniform(pt, base, seed + 4)
    mods = [be.hal_create_modulus(ring.primes[i]) for i in base]

    def reset():
        for c, (ct, r) in enumerate(zip(cts, limbs)):
            ring.fill_uniform(ct, base[:r], seed + 1000 + c)

    def plain():
        for ct, out, r in zip(cts, outs, limbs):
            b = base[:r]
            ring.pb_to_crt_tensor(ct, b)
            be.hal_mul_mont64(ct, pt[:, 0:r], mods[:r])
            ring.crt_to_pb_tensor(ct, b)
            ring.mod_switch_down(out, ct, b, r - 1)

    stream = be.stream()
    ev = HipEvents()
    # plain form: reference result and time
    reset()
    plain()
    ref = [be.hal_export(o) for o in outs[:8]]
    ref_last = be.hal_export(outs[-1])
    plain_steps = max(1, steps // 2)
    ev.start(stream)
    for _ in range(plain_steps):
        plain()
    plain_ms = ev.stop(stream) / plain_steps
    # batched form: autotune on the data, then check and time graph replays
    plan = ring.batch_plan(base, cts, outs, pt)
    t0 = time.perf_counter()
    tuning, tuned_ms = plan.autotune(reps=2)
    tune_s = time.perf_counter() - t0
    reset()
    plan.run(graph=True)
    for o, e in zip(outs[:8], ref):
        if not (be.hal_export(o) == e).all():
            raise SystemExit("bench: batched chain differs from the plain per-ciphertext form")
    if not (be.hal_export(outs[-1]) == ref_last).all():
        raise SystemExit("bench: batched chain differs from the plain per-ciphertext form")
    plan.run(graph=True)
    be.sync()
    ev.start(stream)
    for _ in range(steps):
        plan.run(graph=True)
    graph_ms = ev.stop(stream) / steps
    # the other form of the chain, for comparison: persistent one-launch if the tuner chose the
    # staged launches, else the staged launches
    chosen = plan.tuning()
    other = dict(chosen, persistent=0, grid=0) if chosen["persistent"] else dict(chosen, persistent=1, rows=1, grid=0, g2=3, g3=2)
    plan.set_tuning(**other)
    plan.run(graph=True)
    be.sync()
    ev.start(stream)
    for _ in range(steps):
        plan.run(graph=True)
    other_ms = ev.stop(stream) / steps
    plan.set_tuning(**chosen)
    # per-kernel split of one eager chain (the graph replays carry no per-kernel events)
    be.set_timing(True)
    plan.run(graph=False)
    be.sync()
    be.set_timing(False)
    kernels = be.timing_report()
    info = plan.info()
    per_kernel = {k: {"launches": n, "avg_ms": round(t / n, 4), "GB/s": round(b / n / (t / n / 1000.0) / 1e9, 1)}
                  for k, (n, t, b) in kernels.items()}
    dom = max(kernels.items(), key=lambda kv: kv[1][1])
    name, (launches, total_ms, total_bytes) = dom
    achieved = total_bytes / launches / (total_ms / launches / 1000.0) / 1e9
    res = {
        "workload": f"field_36864 uneven batch: {count} ciphertexts, limbs uniform in 3..{R} (mean {np.mean(limbs):.2f}), "
                    f"pb->CRT, plaintext product, CRT->pb, mod_switch_down by one limb; {info['slices']} (element, limb) "
                    f"slices, {info['bytes'] / 1e9:.2f} GB algorithmic HBM traffic per chain",
        "value": round(count / (graph_ms / 1000.0), 1), "unit": "ciphertexts/s",
        "ms_per_step": round(graph_ms, 3),
        "plain": {"ms_per_step": round(plain_ms, 3), "value": round(count / (plain_ms / 1000.0), 1),
                  "form": "per-ciphertext HAL call sequence"},
        "speedup_vs_plain": round(plain_ms / graph_ms, 3),
        "chain_GB/s": round(info["bytes"] / (graph_ms / 1000.0) / 1e9, 1),
        "tuning": tuning, "autotune_s": round(tune_s, 2),
        ("staged_ms_per_step" if chosen["persistent"] else "persistent_ms_per_step"): round(other_ms, 3),
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "avg_launch_ms": round(total_ms / launches, 4)},
        "kernels": per_kernel,
    }
    if cpu:
        # oracle: the reference sequence on a sample that follows the batch's limb distribution
        # (16 ciphertexts at evenly spaced ranks of the sorted limb counts), checked against the
        # device; one thread, then the host's threads over disjoint ciphertexts
        from concurrent.futures import ThreadPoolExecutor

        from oracle import cpu_backend
        from oracle import ring as oring

        cb = cpu_backend.CpuBackend()
        orc = oring.RingContext(cb, oring.FieldDescriptor(FIELD), LOG2_Q, T_BITS)
        order = sorted(range(count), key=lambda c: limbs[c])
        nsamp = min(16, count)
        idx = [order[(2 * k + 1) * count // (2 * nsamp)] for k in range(nsamp)]
        reset()
        hosts = [be.hal_export(cts[c]).reshape(cts[c].shape) for c in idx]
        hpt = be.hal_export(pt).reshape(pt.shape)
        plan.run(graph=True)
        for r in sorted(set(limbs[c] for c in idx)):
            orc.transform_cache(list(range(r)))

        def chain(k):
            c, h = idx[k], hosts[k]
            r = limbs[c]
            b = list(range(r))
            x = h.copy()
            orc.pb_to_crt_tensor(x, b)
            cb.hal_mul_mont64(x, np.ascontiguousarray(hpt[:, :r]), orc.mods_for(b))
            orc.crt_to_pb_tensor(x, b)
            return oring.mod_switch_down(orc, x.copy(), b, r - 1)

        mid = nsamp // 2
        t0 = time.perf_counter()
        exp0 = chain(mid)
        one = time.perf_counter() - t0
        threads = cpu_threads()
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            exps = list(ex.map(chain, range(nsamp)))
        dt = time.perf_counter() - t0
        if not (exps[mid] == exp0).all():
            raise SystemExit("bench: oracle chain is not deterministic")
        for k, c in enumerate(idx):
            if not (be.hal_export(outs[c]).reshape(outs[c].shape) == exps[k]).all():
                raise SystemExit("bench: batched chain differs from the oracle")
        mean_s = float(np.mean([limbs[c] for c in idx]))
        res["cpu_baseline"] = {"value": round(nsamp / dt, 3), "unit": "ciphertexts/s", "cores": threads, "kind": "port",
                               "sample": f"{nsamp} ciphertexts at evenly spaced ranks of the batch's limb counts (mean "
                                         f"{mean_s:.2f}, batch mean {np.mean(limbs):.2f}) on {threads} threads over disjoint "
                                         f"ciphertexts ({dt:.2f} s), each checked bit-exact; one thread: "
                                         f"{1.0 / one:.2f} ciphertexts/s at {limbs[idx[mid]]} limbs"}
    plan.close()
    del cts, outs, ct_arena, out_arena, pt
    ring.close()
    return res


def bench_hom_mul(pkg, be, steps, pairs=64, ndigits=9, seed=SEED, cpu=True):
    """The ring multiplication of a BFV ciphertext pair at field_36864 (bfv_hom_mul,
    bfv.rs:747-775): lift to the 37-prime extended base, tensor product in wide accumulators,
    rescale by t = 2^76, relinearization with a 9-digit key (gadget base conversion + key-switch
    inner product), on `pairs` ciphertext pairs, one plain HAL call sequence per pair."""
    ring = pkg.ring.RingContext(be, FIELD, LOG2_Q, T_BITS)
    base = ring.initial_rns_base()
    R = len(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    t = 1 << T_BITS
    digits = pkg.bfv.select_digits(ndigits, R)
    cts = be.hal_allocate_i64((2 * pairs * 2, R, n2, n3))
    ring.fill_uniform(cts, base, seed + 11)
    keys = be.hal_allocate_i64((2 * ndigits, R, n2, n3))
    ring.fill_uniform(keys, base, seed + 12)
    key_list = [keys[2 * i:2 * i + 2] for i in range(ndigits)]
    ring.precompute_rns_factor_transforms()
    stream = be.stream()
    ev = HipEvents()

    def step():
        out = None
        for k in range(pairs):
            out = pkg.bfv.bfv_hom_mul(ring, t, cts[4 * k:4 * k + 2], cts[4 * k + 2:4 * k + 4], base, key_list, digits)
        return out

    ring.set_gadget_form(0)  # the per-pair form runs the reference's HAL sequences throughout
    ring.set_bconv_form(0)
    try:
        first = step()  # warm-up (also builds the bconv and transform caches)
        be.sync()
        ev.start(stream)
        for _ in range(steps):
            step()
        ms = ev.stop(stream) / steps
    finally:
        ring.set_gadget_form(1)
        ring.set_bconv_form(2)
    # the batched form: every HAL call covers all pairs, the fused gadget conversion included
    # pair k sits at rows 4k, 4k + 1 (lhs) and 4k + 2, 4k + 3 (rhs) of `cts`; gather the two sides
    lhs_all = be.hal_allocate_i64((2 * pairs, R, n2, n3))
    rhs_all = be.hal_allocate_i64((2 * pairs, R, n2, n3))
    for k in range(pairs):
        be.hal_assign(lhs_all[2 * k:2 * k + 2], cts[4 * k:4 * k + 2])
        be.hal_assign(rhs_all[2 * k:2 * k + 2], cts[4 * k + 2:4 * k + 4])
    batched = pkg.bfv.bfv_hom_mul_batch(ring, t, lhs_all, rhs_all, base, key_list, digits)  # warm-up
    if not (be.hal_export(batched[2 * (pairs - 1):2 * pairs]) == be.hal_export(first)).all():
        raise SystemExit("bench: batched ring multiplication differs from the per-pair form")
    be.sync()
    ev.start(stream)
    for _ in range(steps):
        pkg.bfv.bfv_hom_mul_batch(ring, t, lhs_all, rhs_all, base, key_list, digits)
    ms_b = ev.stop(stream) / steps
    res = {
        "workload": f"field_36864 bfv_hom_mul: {pairs} ciphertext pairs, r = {R} primes lifted to "
                    f"{R + len(ring.extension_rns_base_for(base))}, t = 2^{T_BITS}, relinearization key of {ndigits} digits",
        "value": round(pairs / (ms_b / 1000.0), 1), "unit": "ciphertext multiplications/s", "ms_per_step": round(ms_b, 3),
        "form": "batched: one HAL call per step for all pairs, the base conversions as one kernel each with the matrix product "
                "on the int8 matrix cores, the gadget conversion as a fused kernel, bit-identical per pair",
        "per_pair": {"value": round(pairs / (ms / 1000.0), 1), "ms_per_step": round(ms, 3),
                     "form": "the reference's call sequence, one bfv_hom_mul per pair"},
    }
    if cpu:
        from oracle import cpu_backend
        from oracle import ring as oring

        cb = cpu_backend.CpuBackend()
        orc = oring.RingContext(cb, oring.FieldDescriptor(FIELD), LOG2_Q, T_BITS)
        k = pairs - 1
        hl = be.hal_export(cts[4 * k:4 * k + 2]).reshape(2, R, n2, n3)
        hr = be.hal_export(cts[4 * k + 2:4 * k + 4]).reshape(2, R, n2, n3)
        hk = be.hal_export(keys).reshape(2 * ndigits, R, n2, n3)
        orc.transform_cache(base)
        t0 = time.perf_counter()
        want = oring.bfv_hom_mul(orc, t, hl, hr, base, [hk[2 * i:2 * i + 2] for i in range(ndigits)], digits)
        dt = time.perf_counter() - t0
        if not (be.hal_export(first).reshape(want.shape) == want).all():
            raise SystemExit("bench: bfv_hom_mul differs from the oracle")
        res["cpu_baseline"] = {"value": round(1.0 / dt, 3), "unit": "ciphertext multiplications/s", "cores": 1, "kind": "port",
                               "sample": f"1 pair ({dt:.2f} s, caches built beforehand), also checked bit-exact"}
    del cts, keys
    ring.close()
    return res


def bench_dbfv(pkg, be, steps, count=16, log2_b=27, ndigits=4, seed=SEED, cpu=True):
    """The dBFV digit reduction (dbfv_reduce, dbfv.rs:965-1025) at field_36864: `count`
    period-basis ciphertexts of `ndigits` digits with b = 2^27, t = 2^108 (R = t/b = 2^81, two-limb
    overflows), every digit split modulo b Delta with its carry added to the next. One step is the
    whole chain for all ciphertexts (digit-major batch); the fused form runs one kernel per digit,
    the other form the reference's HAL call sequence."""
    log2_t = log2_b * ndigits
    ring = pkg.ring.RingContext(be, FIELD, LOG2_Q, log2_t)
    base = ring.initial_rns_base()
    R = len(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    modulus = pkg.dbfv.DBFVModulus.new(log2_b, ndigits)
    cts = be.hal_allocate_i64((2 * ndigits * count, R, n2, n3))
    ring.fill_uniform(cts, base, seed + 21)
    stream = be.stream()
    ev = HipEvents()
    out_f, _ = pkg.dbfv.dbfv_reduce_batch(ring, modulus, cts, count, 1, base)  # warm-up, builds the constants
    out_u, _ = pkg.dbfv.dbfv_reduce_batch(ring, modulus, cts, count, 1, base, fused=False)
    if not (be.hal_export(out_f) == be.hal_export(out_u)).all():
        raise SystemExit("bench: fused dbfv_reduce differs from the HAL call sequence")
    del out_u
    be.sync()
    be.set_timing(True)
    ev.start(stream)
    for _ in range(steps):
        pkg.dbfv.dbfv_reduce_batch(ring, modulus, cts, count, 1, base)
    ms = ev.stop(stream) / steps
    be.set_timing(False)
    rep = be.timing_report().get("dbfv_split_overflow")
    ev.start(stream)
    for _ in range(steps):
        pkg.dbfv.dbfv_reduce_batch(ring, modulus, cts, count, 1, base, fused=False)
    ms_u = ev.stop(stream) / steps
    coeff_bytes = 2 * count * R * n2 * n3 * 8  # one digit of every ciphertext
    res = {
        "workload": f"field_36864 dbfv_reduce: {count} ciphertexts x {ndigits} digits, b = 2^{log2_b}, t = 2^{log2_t} "
                    f"(R = 2^{log2_t - log2_b}, two-limb carries), r = {R} primes, digit-major batch",
        "value": round(count / (ms / 1000.0), 1), "unit": "ciphertext reductions/s", "ms_per_step": round(ms, 3),
        "form": "fused: one kernel per digit (carry step + split_overflow), bit-identical to the HAL call sequence",
        "hal_sequence": {"value": round(count / (ms_u / 1000.0), 1), "ms_per_step": round(ms_u, 3),
                         "form": "the reference's HAL calls (mul_mont64, matmul_i64, bconv_correct_pow2_wide, wide fma, "
                                 "reduce_wide), batched over the ciphertexts"},
        "speedup_vs_hal_sequence": round(ms_u / ms, 2),
    }
    if rep is not None:
        avg_ms = rep[1] / rep[0]
        bpl = rep[2] / rep[0]
        res["roofline"] = {"bound": "hbm", "kernel": "dbfv_split_overflow", "achieved": round(bpl / (avg_ms / 1000.0) / 1e9, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(bpl / (avg_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                           "bytes_per_launch": bpl, "avg_launch_ms": round(avg_ms, 4),
                           "note": f"per launch one digit of all ciphertexts: {coeff_bytes / 1e6:.0f} MB read and written, "
                                   "plus the carries"}
    if cpu:
        from oracle import cpu_backend
        from oracle import dbfv as odbfv
        from oracle import ring as oring

        cb = cpu_backend.CpuBackend()
        orc = oring.RingContext(cb, oring.FieldDescriptor(FIELD), LOG2_Q, log2_t)
        host = be.hal_export(cts).reshape(2 * ndigits, count, R, n2, n3)
        ct0 = np.ascontiguousarray(host[:, 0])
        t0 = time.perf_counter()
        want, _ = odbfv.dbfv_reduce(orc, odbfv.DBFVModulus.new(log2_b, ndigits), ct0, 1, base)
        dt = time.perf_counter() - t0
        got = be.hal_export(out_f).reshape(2 * ndigits, count, R, n2, n3)[:, 0]
        if not (got == np.asarray(want)).all():
            raise SystemExit("bench: dbfv_reduce differs from the oracle")
        res["cpu_baseline"] = {"value": round(1.0 / dt, 3), "unit": "ciphertext reductions/s", "cores": 1, "kind": "port",
                               "sample": f"1 ciphertext of {ndigits} digits ({dt:.2f} s), also checked bit-exact"}
    del cts, out_f
    ring.close()
    return res


def bench_dbfv_mul(pkg, be, steps, count=16, log2_b=27, ndigits=4, kdigits=9, seed=SEED, cpu=True):
    """The dBFV multiplication (dbfv_hom_mul, dbfv.rs:1033-1143) at field_36864: `count` pairs of
    period-basis ciphertexts of `ndigits` digits, b = 2^27, t = 2^108; both sides reduced, lifted
    to the extended base, multiplied as a digit convolution, rescaled by t and relinearized with a
    `kdigits`-digit key, period-basis output. One step multiplies all pairs (digit-major batch)."""
    log2_t = log2_b * ndigits
    ring = pkg.ring.RingContext(be, FIELD, LOG2_Q, log2_t)
    base = ring.initial_rns_base()
    R = len(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    modulus = pkg.dbfv.DBFVModulus.new(log2_b, ndigits)
    digits = pkg.bfv.select_digits(kdigits, R)
    lhs = be.hal_allocate_i64((2 * ndigits * count, R, n2, n3))
    rhs = be.hal_allocate_i64((2 * ndigits * count, R, n2, n3))
    keys = be.hal_allocate_i64((2 * kdigits, R, n2, n3))
    ring.fill_uniform(lhs, base, seed + 31)
    ring.fill_uniform(rhs, base, seed + 32)
    ring.fill_uniform(keys, base, seed + 33)
    key_list = [keys[2 * i:2 * i + 2] for i in range(kdigits)]
    ring.precompute_rns_factor_transforms()
    stream = be.stream()
    ev = HipEvents()

    def step():
        return pkg.dbfv.dbfv_hom_mul_batch(ring, modulus, lhs, 1, rhs, 1, count, base, key_list, digits)

    out, bound = step()  # warm-up, builds the constants
    host = be.hal_export(out).reshape(2 * ndigits, count, R, n2, n3)
    # the last pair on its own, through the reference's HAL sequence for the reduction
    last = count - 1
    lhs_h = be.hal_export(lhs).reshape(2 * ndigits, count, R, n2, n3)
    rhs_h = be.hal_export(rhs).reshape(2 * ndigits, count, R, n2, n3)
    single, _ = pkg.dbfv.dbfv_hom_mul(ring, modulus, be_import(be, lhs_h[:, last]), 1, be_import(be, rhs_h[:, last]), 1,
                                      base, key_list, digits, fused=False)
    if not (be.hal_export(single).reshape(2 * ndigits, R, n2, n3) == host[:, last]).all():
        raise SystemExit("bench: batched dbfv_hom_mul differs from the single-pair form")
    del single, out
    be.sync()
    ev.start(stream)
    for _ in range(steps):
        step()
    ms = ev.stop(stream) / steps
    res = {
        "workload": f"field_36864 dbfv_hom_mul: {count} ciphertext pairs x {ndigits} digits, b = 2^{log2_b}, "
                    f"t = 2^{log2_t}, r = {R} primes lifted to {R + len(ring.extension_rns_base_for(base))}, "
                    f"relinearization key of {kdigits} digits, period-basis output",
        "value": round(count / (ms / 1000.0), 1), "unit": "dBFV multiplications/s", "ms_per_step": round(ms, 3),
        "form": "digit-major batch: every HAL call covers all pairs; one batched relinearization; "
                "bit-identical per pair",
    }
    if cpu:
        from oracle import cpu_backend
        from oracle import dbfv as odbfv
        from oracle import ring as oring

        cb = cpu_backend.CpuBackend()
        orc = oring.RingContext(cb, oring.FieldDescriptor(FIELD), LOG2_Q, log2_t)
        om = odbfv.DBFVModulus.new(log2_b, ndigits)
        kh = be.hal_export(keys).reshape(2 * kdigits, R, n2, n3)
        okeys = [np.ascontiguousarray(kh[2 * i:2 * i + 2]) for i in range(kdigits)]
        l0, r0 = np.ascontiguousarray(lhs_h[:, 0]), np.ascontiguousarray(rhs_h[:, 0])
        odbfv.dbfv_hom_mul(orc, om, l0.copy(), 1, r0.copy(), 1, base, okeys, digits)  # builds the caches
        t0 = time.perf_counter()
        want, _ = odbfv.dbfv_hom_mul(orc, om, l0.copy(), 1, r0.copy(), 1, base, okeys, digits)
        dt = time.perf_counter() - t0
        if not (host[:, 0] == np.asarray(want)).all():
            raise SystemExit("bench: dbfv_hom_mul differs from the oracle")
        res["cpu_baseline"] = {"value": round(1.0 / dt, 3), "unit": "dBFV multiplications/s", "cores": 1, "kind": "port",
                               "sample": f"1 pair ({dt:.2f} s, caches built beforehand), also checked bit-exact"}
    del lhs, rhs, keys
    ring.close()
    return res


def bench_mfma_form(pkg, be, ring, base, data, host, primes, stream, steps):
    """The int8-MFMA form of the same step on the same batch (mfma_transform: the K part as F_64,
    F_48, the kernel's diagonal and their inverses on the matrix cores, the L' part by the NTT
    passes), equal mod p to the plain form (L1) and checked so; per-kernel times from HIP events."""
    mf = pkg.mfma_transform.MfmaTransforms(ring, base)
    ev = HipEvents()
    x = be.allocate_i64_uninit(list(data.shape))
    be.hal_import(x, host)
    mf.pb_to_crt_tensor(x)
    mf.crt_to_pb_tensor(x)
    be.sync()
    ev.start(stream)
    for _ in range(steps):
        mf.pb_to_crt_tensor(x)
        mf.crt_to_pb_tensor(x)
    ms = ev.stop(stream) / steps
    be.set_timing(True)
    mf.pb_to_crt_tensor(x)
    mf.crt_to_pb_tensor(x)
    be.sync()
    be.set_timing(False)
    kern = {k: {"launches": v[0], "ms": round(v[1], 3)} for k, v in be.timing_report().items()}
    # every step is a round trip: the batch is unchanged mod p
    out = be.hal_export(x[0:1]).reshape(host[0:1].shape)
    pvec = np.array(primes, dtype=object).reshape(1, -1, 1, 1)
    if not (((out.astype(object) - host[0:1].astype(object)) % pvec) == 0).all():
        raise SystemExit("bench: MFMA-form round trip failed")
    del x, mf
    transforms = 2 * data.shape[0]
    return {"value": round(transforms / (ms / 1000.0), 1), "unit": "transforms/s", "ms_per_step": round(ms, 3),
            "parity": "L1: canonical residues equal mod p to the plain form",
            "form": "K part as F_64 (64 x 64), F_48 (48 x 48), diag(F k), F_48^-1, F_64^-1 on int8 MFMA "
                    "(csd_zp_apply_dim, 224 Z_p-MACs per coefficient), L' part by the NTT passes",
            "kernels_per_step": kern}


def be_import(be, host):
    """a device tensor holding the host array `host`"""
    t = be.hal_allocate_i64(host.shape)
    be.hal_import(t, host)
    return t


MFMA_INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA


def bench_zp(pkg, be, steps, nelem=512, seed=SEED, cpu=True):
    """C4: the period->CRT map of field_36864 as dense matrices over Z_p, Y = C_L' X C_K^T per
    prime, applied exactly to `nelem` ring elements of the 17-prime base ([nelem, 17, 12, 3072]):
    the 3072 x 3072 group-circulant K-part matrix along dimension 3 and the 12 x 12 L'-part matrix
    along dimension 2, each as the int8-digit MFMA contraction. 113,688,576 Z_p multiply-adds per
    (element, prime); checked equal mod p to pb_to_crt_tensor of the same elements. The other MFMA
    kernels and the 64-bit VALU contraction run the K part for comparison."""
    ring = pkg.ring.RingContext(be, FIELD, LOG2_Q, T_BITS)
    base = ring.initial_rns_base()
    R = len(base)
    n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
    primes = [ring.primes[i] for i in base]
    mods = [be.hal_create_modulus(p) for p in primes]
    t0 = time.perf_counter()
    ak = ring.transform_matrix(base, 3)
    mat_k = be.zp_matrix(ak, mods)
    # the K-part matrix is the group-circulant of the CRT kernel over C_2 x C_3 x C_8 x C_64: the
    # default kernel then builds its A operands from the kernel's digit planes (checked on the device)
    mat_k.set_circulant(ring.info["orders"][ring.info["lprime_factor_count"]:])
    al = ring.transform_matrix(base, 2)
    mat_l = be.zp_matrix(al, mods)
    be.sync()
    setup_s = time.perf_counter() - t0
    del ak, al
    x = be.hal_allocate_i64((nelem, R, n2, n3))
    ring.fill_uniform(x, base, seed + 7)
    t = be.hal_allocate_i64(x.shape)
    y = be.hal_allocate_i64(x.shape)
    stream = be.stream()
    ev = HipEvents()

    # the L' side is 12-term dot products: the plain form's one-thread-per-vector kernel, not the
    # matrix cores (which would pad each 12-term sum to 64)
    def step():
        be.zp_apply_dim(mat_k, t, x, 3, form="mfma")
        be.zp_apply_dim(mat_l, y, t, 2, form="plain")

    step()  # warm-up
    be.sync()
    ev.start(stream)
    for _ in range(steps):
        step()
    mfma_ms = ev.stop(stream) / steps
    # each side alone, with the kernels' own durations
    be.set_timing(True)
    ev.start(stream)
    for _ in range(steps):
        be.zp_apply_dim(mat_k, t, x, 3, form="mfma")
    k_ms = ev.stop(stream) / steps
    kern_k = be.timing_report()
    be.set_timing(True)
    ev.start(stream)
    for _ in range(steps):
        be.zp_apply_dim(mat_l, y, t, 2, form="plain")
    l_ms = ev.stop(stream) / steps
    be.set_timing(False)
    kern_l = be.timing_report()
    ev.start(stream)
    be.zp_apply_dim(mat_l, y, t, 2, form="mfma")
    l_mfma_ms = ev.stop(stream)

    # parity: the first and the last element equal pb_to_crt_tensor's output mod p
    pv = np.array(primes, dtype=object).reshape(1, -1, 1, 1)
    for e in (0, nelem - 1):
        xe = be.hal_allocate_i64((1, R, n2, n3))
        be.hal_assign(xe, x[e:e + 1])
        ring.pb_to_crt_tensor(xe, base)
        want = be.hal_export(xe).reshape(1, R, n2, n3).astype(object) % pv
        got = be.hal_export(y[e:e + 1]).reshape(1, R, n2, n3).astype(object)
        if not (got == want).all():
            raise SystemExit(f"bench: the dense C4 map differs from pb_to_crt mod p (element {e})")
        del xe

    y0 = be.hal_export(y[0:1]).reshape(1, R, n2, n3).astype(object)  # kept for the CPU check below
    # the other MFMA kernels on the K part: 0 (4 waves loading their operands from L2), 1 (32 x 64
    # tiles staged in LDS, one block per wave), 2 (64 x 64 tiles loaded straight into LDS), 4 (the
    # default's 64 x 64 tiles with the matrix's own digit planes staged); each must give the
    # default kernel's bits
    tk = be.hal_export(t[0:1])
    other_ms = {}
    for form in (0, 1, 2, 4):
        be.zp_set_kernel(form)
        try:
            be.zp_apply_dim(mat_k, y, x, 3, form="mfma")
            ev.start(stream)
            be.zp_apply_dim(mat_k, y, x, 3, form="mfma")
            other_ms[form] = ev.stop(stream)
        finally:
            be.zp_set_kernel(3)
        if not (tk == be.hal_export(y[0:1])).all():
            raise SystemExit(f"bench: MFMA kernel form {form} differs from the default")
    # the 64-bit VALU contraction on the K part of 16 elements, per element
    sub = 16
    ev.start(stream)
    be.zp_apply_dim(mat_k, y[0:sub], x[0:sub], 3, form="plain")
    plain_k_ms = ev.stop(stream) * nelem / sub
    if not (tk == be.hal_export(y[0:1])).all():
        raise SystemExit("bench: the plain K-part contraction differs from the MFMA one")
    # homcrt's cyclic transforms (bootstrap/homcrt.rs:191-214) as J_p-block maps: field_36864's
    # factors give blocks of 4, 6, 24 and 64 slots; each block a dense J x J circulant with
    # entries uniform mod p, over the same batch viewed as [elements, primes, 36864 / J, J]
    rng = np.random.default_rng(seed + 8)
    homcrt = {}
    for J in (4, 6, 24, 64):
        first = np.stack([rng.integers(0, p, size=J, dtype=np.int64) for p in primes]).reshape(1, R, J)
        idx = (np.arange(J)[:, None] - np.arange(J)[None, :]) % J
        ah = be.hal_allocate_i64((1, R, J, J))
        be.hal_import(ah, first[:, :, idx])
        mat_h = be.zp_matrix(ah, mods)
        shape = (nelem, R, n2 * n3 // J, J)
        xv = pkg.hal.DeviceTensor(be, x.ptr, shape, (x.strides[0], x.strides[1], J, 1), 8, x)
        yv = pkg.hal.DeviceTensor(be, y.ptr, shape, (y.strides[0], y.strides[1], J, 1), 8, y)
        form = "plain" if J <= 32 else "mfma"
        be.zp_apply_dim(mat_h, yv, xv, 3, form=form)
        ev.start(stream)
        for _ in range(steps):
            be.zp_apply_dim(mat_h, yv, xv, 3, form=form)
        hms = ev.stop(stream) / steps
        ops = 2 * 64 * R * nelem * (n2 * n3 // J) * J * J
        homcrt[f"J={J}"] = {"ms_per_step": round(hms, 3), "form": form, "int8_ops_equivalent": ops,
                            "tops_equivalent": round(ops / (hms / 1000.0) / 1e12, 1)}
        del mat_h, ah
    # the NTT path of the same map: pb_to_crt_tensor on the whole batch
    xn = be.hal_allocate_i64(x.shape)
    be.hal_assign(xn, x)
    ring.pb_to_crt_tensor(xn, base)
    ev.start(stream)
    for _ in range(steps):
        ring.pb_to_crt_tensor(xn, base)
    ntt_ms = ev.stop(stream) / steps

    ops_k = 2 * 64 * R * n2 * n3 * n3 * nelem
    ops_l = 2 * 64 * R * n3 * n2 * n2 * nelem
    mk = kern_k["zp_mfma"]
    mfma_kernel_ms = mk[1] / mk[0]
    achieved = ops_k / (mfma_kernel_ms / 1000.0) / 1e12
    units = R * nelem
    res = {
        "workload": f"C4: field_36864 period->CRT map as dense matrices, Y = C_L' X C_K^T per prime ({n3} x {n3} "
                    f"group-circulant and {n2} x {n2}), {nelem} ring elements x {R} primes ([{nelem}, {R}, {n2}, {n3}]), "
                    "exact over Z_p, int8-digit MFMA",
        "value": round(units / (mfma_ms / 1000.0), 1), "unit": "element-limb maps/s", "ms_per_step": round(mfma_ms, 3),
        "int8_ops_per_step": ops_k + ops_l,
        "tops": round((ops_k + ops_l) / (mfma_ms / 1000.0) / 1e12, 1),
        "k_part": {"ms_per_step": round(k_ms, 3), "int8_ops": ops_k},
        "l_part": {"ms_per_step": round(l_ms, 3), "int8_ops": ops_l, "form": "plain (one thread per 12-element vector)",
                   "mfma_form_ms": round(l_mfma_ms, 3),
                   "kernels": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4)} for k, v in kern_l.items()}},
        "homcrt_blocks": homcrt,
        "parity": "L1: elements 0 and last equal pb_to_crt_tensor mod p; every MFMA kernel and the VALU form bit-identical",
        "k_part_other_kernels_ms": {"mfma_l2_operands": round(other_ms[0], 3), "mfma_32x64_lds": round(other_ms[1], 3),
                                    "mfma_64x64_direct_lds": round(other_ms[2], 3),
                                    "mfma_64x64_dense_planes": round(other_ms[4], 3),
                                    "plain_valu": round(plain_k_ms, 3)},
        "ntt_path": {"ms_per_step": round(ntt_ms, 3), "value": round(units / (ntt_ms / 1000.0), 1),
                     "note": "the O(n log n) transform of the same map (pb_to_crt_tensor), equal mod p"},
        "speedup_vs_plain_valu": round(plain_k_ms / k_ms, 2),
        "roofline": {"bound": "mfma", "kernel": "zp_mfma (K part, group-circulant form)", "achieved": round(achieved, 1), "peak": MFMA_INT8_PEAK_TOPS,
                     "unit": "TOPS int8", "frac": round(achieved / MFMA_INT8_PEAK_TOPS, 4),
                     "ops_per_launch": ops_k, "avg_launch_ms": round(mfma_kernel_ms, 4)},
        "kernels": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4)} for k, v in kern_k.items()},
        "setup_s": round(setup_s, 2),
    }
    if cpu:
        from oracle import cpu_backend
        from oracle import ring as oring

        cb = cpu_backend.CpuBackend()
        orc = oring.RingContext(cb, oring.FieldDescriptor(FIELD), LOG2_Q, T_BITS)
        orc.transform_cache(base)
        hx = be.hal_export(x[0:1]).reshape(1, R, n2, n3)
        t0 = time.perf_counter()
        orc.pb_to_crt_tensor(hx, base)
        dt = time.perf_counter() - t0
        if not ((hx.astype(object) % pv) == y0).all():
            raise SystemExit("bench: the dense C4 map differs from the oracle")
        res["cpu_baseline"] = {"value": round(R / dt, 1), "unit": "element-limb maps/s", "cores": 1, "kind": "port",
                               "sample": f"1 element x {R} primes through the oracle's pb_to_crt_tensor ({dt:.2f} s), "
                                         "also checked equal mod p"}
    del mat_k, mat_l, x, t, y, xn
    ring.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--count", type=int, default=128, help="ring elements in the batch (2 per ciphertext)")
    ap.add_argument("--cpu-sample", type=int, default=64, help="ring elements in the CPU-baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum wall time of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workloads", default="boot36864,hom_mul,batch,zp,bootstrap,dbfv,dbfv_mul",
                    help="comma-separated extra workloads beside the headline transforms: boot36864 (the metric's "
                         "bootstrap ms @field_36864), hom_mul, batch, zp, bootstrap (the field_49152 step), dbfv, "
                         "dbfv_mul, or none")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs the hot path on a single MI355X "
              "(multi-GPU sharding of the batch / RNS limbs is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package

    pkg = load_package()
    be = pkg.HipBackend(0)
    t_setup = time.perf_counter()
    ring = pkg.ring.RingContext(be, FIELD, LOG2_Q, T_BITS)
    base = ring.initial_rns_base()
    primes = [ring.primes[i] for i in base]
    info = ring.info
    # compute_rns_factor_transform for every prime (ring.rs:1337-1415): prime sampling, then the CRT
    # kernels' traces, NTTs and inverses on the device
    ring.precompute_rns_factor_transforms()
    be.sync()
    setup_s = time.perf_counter() - t_setup
    host = uniform_batch(primes, args.count, info["deg_lprime"], info["deg_k"], SEED)
    data = be.hal_allocate_i64(host.shape)
    be.hal_import(data, host)
    stream = be.stream()
    ev = HipEvents()
    transforms_per_step = 2 * args.count

    def step():
        ring.pb_to_crt_tensor(data, base)
        ring.crt_to_pb_tensor(data, base)

    for _ in range(args.warmup):
        step()
    be.sync()
    t0 = time.perf_counter()
    ev.start(stream)
    for _ in range(args.steps):
        step()
    ms = ev.stop(stream)
    wall = time.perf_counter() - t0
    # per-kernel durations for the roofline: a separate pass with HIP events around each launch
    ev.start(stream)
    be.set_timing(True)
    for _ in range(max(2, args.steps // 2)):
        step()
    be.sync()
    be.set_timing(False)
    kernels = be.timing_report()
    # the two-stream half-batch overlap (off by default), timed beside the default serial passes
    ring.set_overlap(True)
    step()  # warm-up: creates the second stream
    be.sync()
    ev.start(stream)
    for _ in range(max(2, args.steps // 2)):
        step()
    two_stream_ms = ev.stop(stream) / max(2, args.steps // 2)
    ring.set_overlap(False)
    mfma = bench_mfma_form(pkg, be, ring, base, data, host, primes, stream, max(2, args.steps // 2))

    # size-independent check: every step is a round trip, so the batch is unchanged mod p
    out = be.hal_export(data[0:1]).reshape(host[0:1].shape)
    pvec = np.array(primes, dtype=object).reshape(1, -1, 1, 1)
    if not (((out.astype(object) - host[0:1].astype(object)) % pvec) == 0).all():
        raise SystemExit("bench: round-trip check failed")

    ms_per_step = ms / args.steps
    value = transforms_per_step / (ms_per_step / 1000.0)
    # the whole transform at the HBM roofline: one read and one write of every residue of the
    # element (16 B x primes x coefficients) at the 8 TB/s peak
    headline_peak = HBM_PEAK_GBS * 1e9 / (16 * len(primes) * info["deg_lprime"] * info["deg_k"])
    dom = max(kernels.items(), key=lambda kv: kv[1][1])
    name, (launches, total_ms, total_bytes) = dom
    avg_ms = total_ms / launches
    bytes_per_launch = total_bytes / launches
    achieved = bytes_per_launch / (avg_ms / 1000.0) / 1e9

    # the integer-VALU side: modular multiplications per launch (from the plans' schedules)
    # against the measured Barrett-multiplication ceiling of the chip
    m0 = be.hal_create_modulus(primes[0])
    plan_lp = be.hal_create_ntt_plan(m0, info["orders"][:info["lprime_factor_count"]])
    plan_k = be.hal_create_ntt_plan(m0, info["orders"][info["lprime_factor_count"]:])
    rows = args.count * len(primes) * info["deg_lprime"]
    cols = args.count * len(primes) * info["deg_k"]
    mults = {"ntt_mul_dim3": rows * (plan_k.mults("fwd_unreduced") + info["deg_k"]),
             "intt_dim3": rows * plan_k.mults("inv"),
             "ntt_mul_dim2": cols * (plan_lp.mults("fwd_unreduced") + info["deg_lprime"]),
             "intt_dim2": cols * plan_lp.mults("inv")}
    copy_gbs = be.calibrate("copy", 1 << 30)
    barrett_gps = be.calibrate("barrett", 1 << 24)
    per_kernel = {}
    for k, (n, t, b) in kernels.items():
        e = {"launches": n, "avg_ms": round(t / n, 4), "GB/s": round(b / n / (t / n / 1000.0) / 1e9, 1)}
        if k in mults:
            e["Gmul/s"] = round(mults[k] / (t / n / 1000.0) / 1e9, 1)
            e["valu_frac"] = round(e["Gmul/s"] / barrett_gps, 3)
        per_kernel[k] = e
    result = {
        "metric": "period<->CRT transforms/s @field_36864",
        "value": round(value, 1),
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": f"synthetic: residues uniform in [0, p) per limb, numpy PCG64 seed {SEED}",
        "config": {"workload": f"field_36864 pb->CRT + CRT->pb, {args.count} ring elements ({args.count // 2} ciphertexts) "
                               f"x 17 primes x 36864 coefficients, [{args.count}, 17, 12, 3072] int64 in HBM",
                   "field": FIELD, "rns_primes": len(primes), "elements": args.count, "form": "plain"},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "headline_frac": round(value / headline_peak, 4),
                     "headline_peak": round(headline_peak, 1),
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(name, args.count),
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_ms, 4),
                     "valu": {"achieved": per_kernel[name].get("Gmul/s"), "ceiling": round(barrett_gps, 1),
                              "unit": "G modmul/s", "frac": per_kernel[name].get("valu_frac")},
                     "valu_issue": valu_issue_from_profiles(name, avg_ms, args.count)},
        "ceilings": {"barrett_Gmul/s": round(barrett_gps, 1)},
        "calibration": {"copy_kernel_GB/s": round(copy_gbs, 1),
                        "note": "a plain 1 GiB int4 copy kernel, run so that rocprofv3 PMC runs of this command can "
                                "scale FETCH_SIZE by a known read volume; not a bandwidth ceiling"},
        "kernels": per_kernel,
        "two_stream_ms_per_step": round(two_stream_ms, 3),
        "mfma_form": mfma,
        "overlap": "the passes run serially (the default); two_stream_ms_per_step is the optional two-stream "
                   "half-batch overlap, one pass apart, which measures 1-3 % slower",
        "wall_s": round(wall, 3),
        "setup_s": round(setup_s, 3),
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(primes, host, min(args.cpu_sample, args.count), cpu_threads(),
                                              args.cpu_seconds)
    ring.close()
    del data
    workloads = {}
    chosen = set(w.strip() for w in args.workloads.split(",") if w.strip())
    if "hom_mul" in chosen:
        workloads["bfv_hom_mul"] = bench_hom_mul(pkg, be, max(1, args.steps // 5), cpu=not args.no_cpu)
    if "batch" in chosen:
        workloads["batch_uneven_chain"] = bench_batch(pkg, be, max(2, args.steps // 2), cpu=not args.no_cpu)
    if "zp" in chosen:
        workloads["dense_zp_mfma"] = bench_zp(pkg, be, max(2, args.steps // 4), cpu=not args.no_cpu)
    if "bootstrap" in chosen:
        workloads["bootstrap_step_c5"] = bench_bootstrap_step(pkg, be, cpu=not args.no_cpu)
    if "boot36864" in chosen:
        result["bootstrap"] = bench_bootstrap36864(pkg, be, cpu=not args.no_cpu)
    if "dbfv" in chosen:
        workloads["dbfv_reduce"] = bench_dbfv(pkg, be, max(2, args.steps // 2), cpu=not args.no_cpu)
    if "dbfv_mul" in chosen:
        workloads["dbfv_hom_mul"] = bench_dbfv_mul(pkg, be, max(1, args.steps // 5), cpu=not args.no_cpu)
    if workloads:
        result["workloads"] = workloads
    be.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
