"""Python front end of the ring-level C ABI (include/csdbfv_ring.h): the RNS ring of
``ring::RingContext`` (cyclosubfield-dbfv/src/ring.rs:594-1852) and the bfv HAL sequences of
the hot path, computed on the MI355X. Method names follow the reference (``pb_to_crt_tensor``,
``bconv_tensors``, ``mod_switch_down`` ...); tensors are ``DeviceTensor`` views."""
import ctypes

import numpy as np

from .hal import check, lib


def _sizes(xs):
    xs = [int(x) for x in xs]
    return (ctypes.c_size_t * max(1, len(xs)))(*xs), ctypes.c_size_t(len(xs))


def field_info(name):
    """m, deg_L, deg_K, deg_L', |Gal(M|L)|, rou_order, cyclic orders, number of L'-factors."""
    m, dl, dk, dlp, gml, rou = (ctypes.c_uint64() for _ in range(6))
    orders = (ctypes.c_size_t * 16)()
    n, lp = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().csd_field_info(name.encode(), ctypes.byref(m), ctypes.byref(dl), ctypes.byref(dk), ctypes.byref(dlp),
                               ctypes.byref(gml), ctypes.byref(rou), orders, ctypes.c_size_t(16), ctypes.byref(n), ctypes.byref(lp)))
    return dict(m=m.value, deg_l=dl.value, deg_k=dk.value, deg_lprime=dlp.value, gal_ml=gml.value, rou_order=rou.value,
                orders=[orders[i] for i in range(n.value)], lprime_factor_count=lp.value)


def gal_to_shift(name, g):
    out = (ctypes.c_int64 * 16)()
    n = ctypes.c_size_t()
    check(lib().csd_field_gal_to_shift(name.encode(), ctypes.c_uint64(g), out, ctypes.c_size_t(16), ctypes.byref(n)))
    return [out[i] for i in range(n.value)]


def homcrt_galois_elements(name):
    """required_galois_elements (bootstrap/homcrt.rs:336-349) and the |J_p| of each p | m_K."""
    out = (ctypes.c_uint64 * 4096)()
    jp = (ctypes.c_uint64 * 16)()
    n, njp = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().csd_field_homcrt_galois_elements(name.encode(), out, ctypes.c_size_t(4096), ctypes.byref(n), jp,
                                                 ctypes.c_size_t(16), ctypes.byref(njp)))
    return [out[i] for i in range(n.value)], [jp[i] for i in range(njp.value)]


def homcrt_factors(name):
    """(p, i_p_gen, |J_p|) per prime p | m_K (field.rs:217-243, 271): the cyclic factors of the
    homomorphic CRT transform (bootstrap/homcrt.rs:29-40)."""
    p, g, n = (ctypes.c_uint64 * 16)(), (ctypes.c_uint64 * 16)(), (ctypes.c_uint64 * 16)()
    cnt = ctypes.c_size_t()
    check(lib().csd_field_homcrt_factors(name.encode(), p, g, n, ctypes.c_size_t(16), ctypes.byref(cnt)))
    return [(p[i], g[i], n[i]) for i in range(cnt.value)]


def sample_primes(name, log2_q, t_bits):
    out = (ctypes.c_int64 * 128)()
    n, il = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().csd_sample_primes(name.encode(), ctypes.c_size_t(log2_q[0]), ctypes.c_size_t(log2_q[1]), ctypes.c_size_t(t_bits),
                                  out, ctypes.c_size_t(128), ctypes.byref(n), ctypes.byref(il)))
    return [out[i] for i in range(n.value)], il.value


class RingContext:
    def __init__(self, backend, field, log2_q=None, t_bits=0, raw_primes=None):
        self.backend = backend
        self.field_name = field
        self.info = field_info(field)
        h = ctypes.c_void_p()
        if raw_primes is not None:
            arr = (ctypes.c_int64 * len(raw_primes))(*raw_primes)
            check(lib().csd_ring_create_raw(backend.ctx, field.encode(), arr, ctypes.c_size_t(len(raw_primes)), ctypes.byref(h)))
        else:
            check(lib().csd_ring_create(backend.ctx, field.encode(), ctypes.c_size_t(log2_q[0]), ctypes.c_size_t(log2_q[1]),
                                        ctypes.c_size_t(t_bits), ctypes.byref(h)))
        self.handle = h
        out = (ctypes.c_int64 * 128)()
        n, il = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().csd_ring_primes(self.handle, out, ctypes.c_size_t(128), ctypes.byref(n), ctypes.byref(il)))
        self.primes = [out[i] for i in range(n.value)]
        self.initial_rns_base_len = il.value
        self._consts = {}
        self.range_checks = False

    def set_range_check(self, on):
        """Debug mode: every dBFV batch the dbfv layer produces (reductions, products,
        accumulators, combinations, mod switches) is checked against its tracked bound, |x| <=
        bound p for the prime p of each limb (the bookkeeping of RingEls.bound, ring.rs:33-47),
        and a violation raises. Each check synchronizes the device; off by default."""
        self.range_checks = bool(on)

    def check_range(self, t, base, bound, what):
        """raises when an entry of the [count, r, ., .] tensor `t` over `base` is beyond bound p"""
        bad = self.backend.range_check(t, self.mods_for(base), bound)
        if bad:
            raise ArithmeticError(f"{what}: {bad} entries beyond {bound} p (range check)")

    def close(self):
        self._consts = {}
        if self.handle is not None:
            lib().csd_ring_destroy(self.handle)
            self.handle = None

    def initial_rns_base(self):
        return list(range(self.initial_rns_base_len))

    def allocate(self, count, base):
        return self.backend.hal_allocate_i64([count, len(base), self.info["deg_lprime"], self.info["deg_k"]])

    def allocate_uninit(self, count, base):
        """`allocate` without the zero fill, for an output that the next call writes entirely"""
        return self.backend.allocate_i64_uninit([count, len(base), self.info["deg_lprime"], self.info["deg_k"]])

    def precompute_rns_factor_transforms(self):
        check(lib().csd_ring_precompute_transforms(self.handle))

    def crt_kernels(self, idx):
        dlp, dk = self.info["deg_lprime"], self.info["deg_k"]
        a, b, c, d = (np.zeros(n, dtype=np.int64) for n in (dlp, dk, dlp, dk))
        check(lib().csd_ring_crt_kernels(self.handle, ctypes.c_size_t(idx), *[x.ctypes.data_as(ctypes.c_void_p) for x in (a, b, c, d)]))
        return a, b, c, d

    def extension_rns_base_for(self, base):
        b, n = _sizes(base)
        out = (ctypes.c_size_t * 128)()
        cnt = ctypes.c_size_t()
        check(lib().csd_ring_extension_base(self.handle, b, n, out, ctypes.c_size_t(128), ctypes.byref(cnt)))
        return [out[i] for i in range(cnt.value)]

    def pb_to_crt_tensor(self, data, base):
        b, n = _sizes(base)
        check(lib().csd_ring_pb_to_crt(self.handle, ctypes.byref(data.view()), b, n))

    def crt_to_pb_tensor(self, data, base):
        b, n = _sizes(base)
        check(lib().csd_ring_crt_to_pb(self.handle, ctypes.byref(data.view()), b, n))

    def bconv_tensors(self, dst, dst_base, src, src_base, preserve_src=False):
        """bconv_tensors (ring.rs:1526-1535). Like the reference it may destroy src; with
        preserve_src it leaves src unchanged in every form (the fused kernel does anyway, the HAL
        sequence then runs on a copy)."""
        db, dn = _sizes(dst_base)
        sb, sn = _sizes(src_base)
        check(lib().csd_ring_bconv(self.handle, ctypes.byref(dst.view()), db, dn, ctypes.byref(src.view()), sb, sn,
                                   ctypes.c_int(1 if preserve_src else 0)))

    def set_overlap(self, on):
        """Two-stream half-batch overlap of this ring's transforms (csd_ring_set_overlap, off by
        default); results are identical either way."""
        check(lib().csd_ring_set_overlap(self.handle, ctypes.c_int(1 if on else 0)))

    def set_bconv_form(self, form):
        """How this ring's bconv_tensors, rescale_step (and the conversions inside its sequences)
        run (csd_ring_set_bconv_form): 2 (default) one kernel with the matrix product on the int8
        matrix cores where they are the faster (from 9 source primes, 17 for the rescale step), 3
        the matrix-core kernel wherever its tables fit, 1 one fused kernel on the vector ALUs, 0 the
        reference's HAL call sequence; residues are identical."""
        check(lib().csd_ring_set_bconv_form(self.handle, ctypes.c_int(int(form))))

    def bconv_form(self):
        """the current csd_ring_set_bconv_form setting of this ring (3 / 2 matrix cores, 1 fused, 0
        HAL sequence)"""
        return lib().csd_ring_get_bconv_form(self.handle)

    def set_gadget_form(self, form):
        """How this ring's bconv_gadget_tensors runs (csd_ring_set_gadget_form): 1 (default) one
        fused kernel, 0 the reference's HAL call sequence per element; residues are identical."""
        check(lib().csd_ring_set_gadget_form(self.handle, ctypes.c_int(int(form))))

    def rescale_step(self, ext, ext_base, src, src_base, s_src, s_ext, inv):
        """The middle of rescale_to_base (bfv.rs:601-622) for a batch: ext <- mont(mont(ext, s_ext)
        - bconv(mont(src, s_src)), inv), that is mul_mont64, bconv, sub_i64 and mul_mont64, as one
        kernel when the fused conversion applies (src is then left as it is; the HAL sequence, the
        other form, destroys it). src and ext may be disjoint limb ranges of one tensor."""
        eb, en = _sizes(ext_base)
        sb, sn = _sizes(src_base)
        check(lib().csd_ring_rescale_step(self.handle, ctypes.byref(ext.view()), eb, en, ctypes.byref(src.view()), sb, sn,
                                          ctypes.byref(s_src.view()), ctypes.byref(s_ext.view()),
                                          ctypes.byref(inv.view())))

    def bconv_gadget_tensors(self, dst, src, base, digits):
        """bconv_gadget_tensors (ring.rs:1548-1557) for src [B, r, ., .] and dst [B d, r, ., .]
        (element b, digit i at row b d + i); the reference takes B = 1."""
        b, n = _sizes(base)
        bounds = [x for d in digits for x in d]
        bb, _ = _sizes(bounds)
        check(lib().csd_ring_bconv_gadget(self.handle, ctypes.byref(dst.view()), ctypes.byref(src.view()), b, n, bb,
                                          ctypes.c_size_t(len(digits))))

    def mod_switch_down(self, dst, src, base, new_len):
        b, n = _sizes(base)
        check(lib().csd_ring_mod_switch_down(self.handle, ctypes.byref(dst.view()), ctypes.byref(src.view()), b, n,
                                             ctypes.c_size_t(new_len)))

    def apply_galois(self, dst, src, g):
        check(lib().csd_ring_apply_galois(self.handle, ctypes.byref(dst.view()), ctypes.byref(src.view()), ctypes.c_uint64(g)))

    def galois_keyswitch(self, base, batch, digits, parts, c0, gs, keys, out):
        """Batched key switching of bfv_hom_gal (bfv.rs:816-839): see csd_ring_galois_keyswitch."""
        b, n = _sizes(base)
        garr = (ctypes.c_uint64 * len(gs))(*gs)
        karr = (type(parts.view()) * len(keys))(*[k.view() for k in keys])
        check(lib().csd_ring_galois_keyswitch(self.handle, b, n, ctypes.c_size_t(batch), ctypes.c_size_t(digits),
                                              ctypes.byref(parts.view()), ctypes.byref(c0.view()), garr,
                                              ctypes.c_size_t(len(gs)), karr, ctypes.byref(out.view())))

    def galois_keyswitch_streamed(self, base, batch, digits, parts, c0, gs, host_keys, out):
        """galois_keyswitch with the keys in pinned host memory (PinnedHostTensor each), streamed
        to the device rotation by rotation; see csd_ring_galois_keyswitch_streamed."""
        b, n = _sizes(base)
        garr = (ctypes.c_uint64 * len(gs))(*gs)
        P = ctypes.POINTER(ctypes.c_int64)
        karr = (P * len(host_keys))(*[ctypes.cast(ctypes.c_void_p(k.ptr), P) for k in host_keys])
        check(lib().csd_ring_galois_keyswitch_streamed(self.handle, b, n, ctypes.c_size_t(batch), ctypes.c_size_t(digits),
                                                       ctypes.byref(parts.view()), ctypes.byref(c0.view()), garr,
                                                       ctypes.c_size_t(len(gs)), karr, ctypes.byref(out.view())))

    def galois_keyswitch_rows(self, base, batch, group, digits, parts, c0, c0_rows, gs, keys, outs, out_rows):
        """csd_ring_galois_keyswitch_rows: the batched key switch with element e = h group + l
        reading c0 row h c0_rows[0] + l c0_rows[1] and writing row s of rotation t to row
        h out_rows[0] + l out_rows[1] + s out_rows[2] of outs[t]; keys may be restricted views."""
        b, n = _sizes(base)
        garr = (ctypes.c_uint64 * len(gs))(*gs)
        karr = (type(parts.view()) * len(keys))(*[k.view() for k in keys])
        oarr = (type(parts.view()) * len(outs))(*[o.view() for o in outs])
        check(lib().csd_ring_galois_keyswitch_rows(self.handle, b, n, ctypes.c_size_t(batch), ctypes.c_size_t(group),
                                                   ctypes.c_size_t(digits), ctypes.byref(parts.view()),
                                                   ctypes.byref(c0.view()), ctypes.c_int64(c0_rows[0]),
                                                   ctypes.c_int64(c0_rows[1]), garr, ctypes.c_size_t(len(gs)), karr, oarr,
                                                   ctypes.c_int64(out_rows[0]), ctypes.c_int64(out_rows[1]),
                                                   ctypes.c_int64(out_rows[2])))

    def lprime_element_tensor(self, base, coeffs):
        """lprime_element_tensor (ring.rs:1227-1250): the element of L' with period-basis
        coefficients `coeffs` (basis induced by M) as a [1, r, deg_L', 1] CRT-basis factor tensor:
        mont((c rem p) as i64) per limb and coefficient, then the L' half of the CRT transform on
        the device. Kept per (base, coefficients), like the other constant tensors."""
        coeffs = tuple(int(c) for c in coeffs)
        if len(coeffs) != self.info["deg_lprime"]:
            raise ValueError("an element of L' has deg_L' coefficients")
        key = ("lprime", tuple(base), coeffs)
        hit = self._consts.get(key)
        if hit is not None:
            return hit
        vals = []
        for i in base:
            p = self.primes[i]
            for c in coeffs:
                x = abs(c) % p
                x = -x if c < 0 else x  # Rust's remainder keeps the sign of c
                m = (abs(x) << 64) % p
                vals.append(-m if x < 0 else m)
        t = self.backend.hal_allocate_i64([1, len(base), self.info["deg_lprime"], 1])
        self.backend.hal_import(t, np.array(vals, dtype=np.int64))
        b, n = _sizes(base)
        check(lib().csd_ring_lprime_transform(self.handle, ctypes.byref(t.view()), b, n))
        self._consts[key] = t
        return t

    def decode_scaled(self, src, base, log2_t, dst=None):
        """decode_pb_scaled (bfv.rs:262-274) of every element of the period-basis tensor `src` over
        `base` (at most 4 primes) on the device: an i128 [count, 1, ., .] tensor of round(t x / Q)
        balanced modulo t = 2^log2_t (csd_ring_decode_scaled)"""
        if dst is None:
            dst = self.backend.allocate_i128_uninit([src.shape[0], 1, src.shape[2], src.shape[3]])
        b, n = _sizes(base)
        check(lib().csd_ring_decode_scaled(self.handle, ctypes.byref(src.view()), b, n, ctypes.c_size_t(log2_t),
                                           ctypes.byref(dst.view())))
        return dst

    def encode_digits(self, values, base, log2_b=0, max_digits=1):
        """encode_pb_unscaled (bfv.rs:226-235) over `base` of i128 values (log2_b = 0) or of their
        balanced base-2^log2_b digit vectors (dbfv.rs:97-112, max_digits planes, plane i of element c
        at row i count + c): (period-basis tensor, planes kept) (csd_ring_encode_digits)"""
        count = values.shape[0]
        planes = 1 if log2_b == 0 else max_digits
        dst = self.allocate_uninit(planes * count, base)
        b, n = _sizes(base)
        kept = ctypes.c_int64()
        check(lib().csd_ring_encode_digits(self.handle, ctypes.byref(values.view()), ctypes.c_size_t(log2_b),
                                           ctypes.c_size_t(max_digits), ctypes.byref(dst.view()), b, n, ctypes.byref(kept)))
        return dst, kept.value

    def fill_uniform(self, dst, base, seed):
        """Synthetic residues uniform in [0, p) per limb (splitmix64 of the flat index)."""
        mods = [self.backend.hal_create_modulus(self.primes[i]) for i in base]
        arr = (type(mods[0]) * len(mods))(*mods)
        check(lib().csd_fill_uniform(self.backend.ctx, ctypes.byref(dst.view()), arr, ctypes.c_size_t(len(mods)),
                                     ctypes.c_uint64(seed)))

    def reduce_wide_tensor(self, dst, wide, base, full):
        b, n = _sizes(base)
        check(lib().csd_ring_reduce_wide(self.handle, ctypes.byref(dst.view()), ctypes.byref(wide.view()), b, n,
                                         ctypes.c_int(1 if full else 0)))

    def transform_matrix(self, base, dim, inverse=False):
        """The dense [1, r, n, n] matrix of pb_to_crt (or crt_to_pb) along dim 2 or 3."""
        n = self.info["deg_lprime"] if dim == 2 else self.info["deg_k"]
        out = self.backend.hal_allocate_i64((1, len(base), n, n))
        b, nb = _sizes(base)
        check(lib().csd_ring_transform_matrix(self.handle, b, nb, ctypes.c_int(dim), ctypes.c_int(1 if inverse else 0),
                                              ctypes.byref(out.view())))
        return out

    def batch_plan(self, base, cts, outs, pt):
        """A BatchPlan for ciphertexts at the levels len(cts[c][0]) over prefixes of `base`."""
        return BatchPlan(self, base, cts, outs, pt)

    def scalar_tensor(self, base, scalar):
        """scalar_tensor (ring.rs:876-888): mont(scalar mod p) per prime; any i128 scalar. The
        tensors are constants, kept per (base, scalar) so repeated sequences upload nothing."""
        key = ("scalar", tuple(base), scalar)
        hit = self._consts.get(key)
        if hit is not None:
            return hit
        t = self.backend.hal_allocate_i64([1, len(base), 1, 1])
        b, n = _sizes(base)
        if -(1 << 63) <= scalar < (1 << 63):
            check(lib().csd_ring_scalar_tensor(self.handle, ctypes.byref(t.view()), b, n, ctypes.c_int64(scalar)))
        else:
            if not -(1 << 127) <= scalar < (1 << 127):
                raise ValueError("scalar does not fit an i128")
            w = scalar & ((1 << 128) - 1)
            lo, hi = w & ((1 << 64) - 1), w >> 64
            check(lib().csd_ring_scalar_tensor_wide(self.handle, ctypes.byref(t.view()), b, n,
                                                    ctypes.c_int64(lo - (1 << 64) if lo >= 1 << 63 else lo),
                                                    ctypes.c_int64(hi - (1 << 64) if hi >= 1 << 63 else hi)))
        self._consts[key] = t
        return t

    def inverse_product_tensor(self, base, of):
        """inverse_scalar_tensor (ring.rs:1165-1177) of the product of the primes at `of`."""
        key = ("inverse", tuple(base), tuple(of))
        hit = self._consts.get(key)
        if hit is not None:
            return hit
        t = self.backend.hal_allocate_i64([1, len(base), 1, 1])
        b, n = _sizes(base)
        o, no = _sizes(of)
        check(lib().csd_ring_inverse_product_tensor(self.handle, ctypes.byref(t.view()), b, n, o, no))
        self._consts[key] = t
        return t

    def modulus_of(self, base):
        """modulus_of (ring.rs): the product of the primes of `base`"""
        q = 1
        for i in base:
            q *= self.primes[i]
        return q

    def _host_constants(self, key, base, values):
        """a [1, r, 1, 1] tensor of host-computed per-prime constants, uploaded once per key"""
        hit = self._consts.get(key)
        if hit is None:
            hit = self.backend.hal_allocate_i64([1, len(base), 1, 1])
            self.backend.hal_import(hit, np.array(values, dtype=np.int64).reshape(1, len(base), 1, 1))
            self._consts[key] = hit
        return hit

    def q_over_p_inv_tensor(self, base, scale):
        """q_over_p_inv_tensor (ring.rs:1023-1043): the plain smallest lifts of
        scale (Q/p)^-1 mod p per prime p of `base`"""
        q = self.modulus_of(base)
        vals = []
        for i in base:
            p = self.primes[i]
            v = (pow((q // p) % p, -1, p) * (scale % p)) % p
            vals.append(v - p if v > p // 2 else v)
        return self._host_constants(("q_over_p_inv", tuple(base), scale), base, vals)

    def scalar_tensor_big(self, base, scalar):
        """scalar_tensor_big (ring.rs:890-903): mont(smallest_lift(scalar mod p)) per prime, for any
        integer scalar, mont(x) = (x 2^64) rem p with the sign of x (Rust's remainder)"""
        vals = []
        for i in base:
            p = self.primes[i]
            v = scalar % p
            v = v - p if v > p // 2 else v
            r = abs(v << 64) % p
            vals.append(-r if v < 0 else r)
        return self._host_constants(("big", tuple(base), scalar), base, vals)

    def scalar_tensor_big_mont2(self, base, scalar):
        """scalar_tensor_big_mont2 (ring.rs:904-919): mont(mont(smallest_lift(scalar mod p)))
        per prime, mont(x) = (x 2^64) rem p with the sign of x (Rust's remainder)"""
        def mont(x, p):
            r = abs(x << 64) % p
            return -r if x < 0 else r

        vals = []
        for i in base:
            p = self.primes[i]
            v = scalar % p
            v = v - p if v > p // 2 else v
            vals.append(mont(mont(v, p), p))
        return self._host_constants(("big_mont2", tuple(base), scalar), base, vals)

    def mods_for(self, base):
        key = ("mods", tuple(base))
        hit = self._consts.get(key)
        if hit is None:
            hit = self._consts[key] = [self.backend.hal_create_modulus(self.primes[i]) for i in base]
        return hit

    def convolve_reduce(self, lhs, lhs_rows, rhs, rhs_rows, count, d1, d2, K, base, lhs_bound=1, rhs_bound=1):
        """The digit convolution of dbfv_mul_convolve (dbfv.rs:1096-1127), or with d1 = d2 = K = 1 the
        tensor product of hom_mul_raw_acc (bfv.rs:667-683), reduced as reduce_wide (ring.rs:846-865)
        would reduce the WideRingEls rows the reference accumulates: one kernel, the same int64
        rows. Operand row of (digit i, component s, pair b) is i rows[0] + s rows[1] + b rows[2];
        the result ([3 K count, r, ., .], CRT basis) holds component c of output digit k of pair b
        in row (c K + k) count + b."""
        max_p = max(self.primes[i] for i in base)
        min_p = min(self.primes[i] for i in base)
        full = []
        for c in range(3):
            for k in range(K):
                pairs = min(k, d1 - 1) - max(0, k - (d2 - 1)) + 1
                bound = (2 if c == 1 else 1) * pairs * lhs_bound * rhs_bound * max_p * max_p  # WideRingEls.record_step
                if bound > (1 << 127) - 1:
                    raise OverflowError("wide ring element overflow: too many accumulation steps")
                full.append(bound > (min_p << 63))
        out = self.allocate_uninit(3 * K * count, base)
        self.backend.convolve_reduce(out, lhs, lhs_rows, rhs, rhs_rows, count, d1, d2, K, full, self.mods_for(base))
        return out

    def allocate_wide(self, count, base, zero=True):
        """WideRingEls over `base`; with zero=False the i128 rows start unwritten, and reducing a
        row that no operation has written since raises (a zeroed row would reduce to 0)"""
        return WideRingEls(self, count, base, zero)

    def reduce_wide(self, wide):
        """reduce_wide (ring.rs:846-865): one reduction kind per run of rows."""
        out = self.allocate_uninit(wide.count(), wide.rns_base)
        self.reduce_wide_into(out, wide)
        return out

    def reduce_wide_into(self, dst, wide):
        """reduce_wide_into (ring.rs:829-838): the reduced rows written into `dst`"""
        if not wide.zeroed and any(b == 0 for b in wide.bound):
            raise ValueError("reduce_wide: an unzeroed wide row was never written")
        start = 0
        while start < wide.count():
            full = wide.needs_full_reduction(start)
            end = start + 1
            while end < wide.count() and wide.needs_full_reduction(end) == full:
                end += 1
            self.reduce_wide_tensor(dst[start:end], wide.data[start:end], wide.rns_base, full)
            start = end


    def reduce_wide_rows_into(self, dst, wide, rows):
        """reduce_wide_rows (ring.rs:820-824) of wide rows 0 .. rows - 1 into dst: one reduction
        kind per run of rows; every reduced row must have been written"""
        if any(wide.bound[i] == 0 for i in range(rows)) and not wide.zeroed:
            raise ValueError("reduce_wide: an unzeroed wide row was never written")
        start = 0
        while start < rows:
            full = wide.needs_full_reduction(start)
            end = start + 1
            while end < rows and wide.needs_full_reduction(end) == full:
                end += 1
            self.reduce_wide_tensor(dst[start:end], wide.data[start:end], wide.rns_base, full)
            start = end


class WideRingEls:
    """WideRingEls (ring.rs:330-532): i128 accumulators in the doubly-Montgomery domain with the
    reference's per-row bound tracking; fresh rows are overwritten (ring.rs:373-413)."""

    def __init__(self, ring, count, base, zero=True):
        self.ring = ring
        self.rns_base = list(base)
        shape = [count, len(base), ring.info["deg_lprime"], ring.info["deg_k"]]
        self.data = ring.backend.hal_allocate_i128(shape) if zero else ring.backend.allocate_i128_uninit(shape)
        self.zeroed = zero
        self._one = None
        self._neg_one = None
        self.max_p = max(ring.primes[i] for i in base)
        self.min_p = min(ring.primes[i] for i in base)
        self.bound = [0] * count

    def count(self):
        return len(self.bound)

    def one(self):
        if self._one is None:
            self._one = self.ring.scalar_tensor(self.rns_base, 1)
        return self._one

    def neg_one(self):
        if self._neg_one is None:
            self._neg_one = self.ring.scalar_tensor(self.rns_base, -1)
        return self._neg_one

    def record_step(self, rows, step):
        for i in rows:
            nb = self.bound[i] + step
            if nb > (1 << 127) - 1:
                raise OverflowError("wide ring element overflow: too many accumulation steps")
            self.bound[i] = nb

    def needs_full_reduction(self, row):
        return self.bound[row] > (self.min_p << 63)

    def _accumulate(self, rows, lhs, rhs):
        be = self.ring.backend
        rows = list(rows)
        start, end_all = rows[0], rows[-1] + 1
        while start < end_all:
            fresh = self.bound[start] == 0
            end = start + 1
            while end < end_all and (self.bound[end] == 0) == fresh:
                end += 1

            def sl(t):
                return t if t.shape[0] == 1 else t[start - rows[0]:end - rows[0]]

            dst = self.data[start:end]
            if fresh:
                be.hal_mul_i64(dst, sl(lhs), sl(rhs))
            else:
                be.hal_fma_i64(dst, sl(lhs), sl(rhs))
            start = end

    def fma(self, rows, lhs, lhs_bound, rhs, rhs_bound):
        self._accumulate(rows, lhs, rhs)
        self.record_step(rows, lhs_bound * rhs_bound * self.max_p * self.max_p)

    def inner_product(self, rows, terms):
        rows = list(rows)
        if any(self.bound[i] != 0 for i in rows):
            raise ValueError("inner_product overwrites: the rows must be fresh")
        self.ring.backend.hal_inner_product_i64(self.data[rows[0]:rows[-1] + 1], [t[0] for t in terms], [t[2] for t in terms])
        self.record_step(rows, sum(lb * rb * self.max_p * self.max_p for (_, lb, _, rb) in terms))

    def fma_unbounded(self, rows, lhs, rhs):
        """fma_unbounded (ring.rs:493-500): lhs arbitrary i64 (plain values), rhs bounded by max_p"""
        self._accumulate(rows, lhs, rhs)
        self.record_step(rows, self.max_p << 63)

    def add(self, rows, src, src_bound):
        self._accumulate(rows, src, self.one())
        self.record_step(rows, src_bound * self.max_p * self.max_p)

    def sub(self, rows, src, src_bound):
        self._accumulate(rows, src, self.neg_one())
        self.record_step(rows, src_bound * self.max_p * self.max_p)


class BatchTuning(ctypes.Structure):
    """Launch shapes of the batched chain (csd_batch_tuning)."""
    _fields_ = [("rows", ctypes.c_int), ("threads", ctypes.c_int), ("col_threads", ctypes.c_int), ("g2", ctypes.c_int),
                ("g3", ctypes.c_int), ("persistent", ctypes.c_int), ("grid", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchPlan:
    """Ciphertexts at uneven levels through pb_to_crt_tensor (ring.rs:1118-1141), a CRT-basis
    plaintext product (hal mul_mont64), crt_to_pb_tensor (ring.rs:1144-1161) and
    mod_switch_down by one limb (bfv.rs:628-662), each phase one work-list launch over all
    (element, limb) slices; see csd_ring_batch_create. ``cts[c]`` is a contiguous
    [2, limbs[c], deg_L', deg_K] tensor over base[:limbs[c]] (overwritten with the crt_to_pb
    result), ``outs[c]`` receives the [2, limbs[c] - 1, ., .] mod-switched result."""

    def __init__(self, ring, base, cts, outs, pt):
        self.ring = ring
        self.handle = None
        self._keep = (list(cts), list(outs), pt)
        limbs = [c.shape[1] for c in cts]
        n2, n3 = ring.info["deg_lprime"], ring.info["deg_k"]
        if len(outs) != len(cts):
            raise ValueError("batch: one output per ciphertext")
        for c, o, r in zip(cts, outs, limbs):
            if c.shape != (2, r, n2, n3) or o.shape != (2, r - 1, n2, n3):
                raise ValueError("batch: ciphertext or output shape")
            if c.strides != (r * n2 * n3, n2 * n3, n3, 1) or o.strides != ((r - 1) * n2 * n3, n2 * n3, n3, 1):
                raise ValueError("batch: ciphertexts and outputs must be contiguous")
        n = len(cts)
        P = ctypes.POINTER(ctypes.c_int64)
        carr = (P * max(n, 1))(*[ctypes.cast(ctypes.c_void_p(c.ptr), P) for c in cts])
        oarr = (P * max(n, 1))(*[ctypes.cast(ctypes.c_void_p(o.ptr), P) for o in outs])
        larr = (ctypes.c_size_t * max(n, 1))(*limbs)
        b, nb = _sizes(base)
        h = ctypes.c_void_p()
        check(lib().csd_ring_batch_create(ring.handle, b, nb, ctypes.c_size_t(n), carr, oarr, larr,
                                          ctypes.byref(pt.view()), ctypes.byref(h)))
        self.handle = h

    def run(self, graph=True):
        check(lib().csd_batch_run(self.handle, ctypes.c_int(1 if graph else 0)))

    def autotune(self, reps=3):
        t = BatchTuning()
        ms = ctypes.c_double()
        check(lib().csd_batch_autotune(self.handle, ctypes.c_int(reps), ctypes.byref(t), ctypes.byref(ms)))
        return t.as_dict(), ms.value

    def set_tuning(self, **kw):
        t = BatchTuning()
        check(lib().csd_batch_get_tuning(self.handle, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib().csd_batch_set_tuning(self.handle, ctypes.byref(t)))

    def tuning(self):
        t = BatchTuning()
        check(lib().csd_batch_get_tuning(self.handle, ctypes.byref(t)))
        return t.as_dict()

    def info(self):
        ns, ne, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        check(lib().csd_batch_info(self.handle, ctypes.byref(ns), ctypes.byref(ne), ctypes.byref(b)))
        return {"slices": ns.value, "elements": ne.value, "bytes": b.value}

    def close(self):
        if self.handle:
            check(lib().csd_batch_free(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
