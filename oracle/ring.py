"""ORACLE — test infrastructure only (imported by tests/ and by bench.py's cpu_baseline leg).

A restatement of the reference's ring layer and the scheme-level HAL call sequences that the hot
path serves, on top of any backend with the ``hal::Backend`` method names (here: the oracle
``CpuBackend``, numpy tensors). It follows, function by function:

* field.rs:32-520, 529-769 — field descriptors (subgroups, transversals, Galois shifts);
* ring.rs:84-183          — cmod/mont_repr helpers and RNS prime sampling;
* ring.rs:303-532, 793-865 — wide accumulators and their reduction;
* ring.rs:870-1043        — constant tensors;
* ring.rs:1092-1415       — period<->CRT transforms, CRT kernels, Galois permutations;
* ring.rs:1482-1852       — BEHZ base conversion (plain and gadget);
* bfv.rs:388-426, 493-683, 747-841 — digits, gadget decomposition, key-switch inner products,
  lift/rescale/mod-switch, hom_mul and hom_gal sequences.

Pins beyond the reference's KATs: lift_to_extended, rescale_to_base and gadget_decompose are
checked against their closed forms by CRT reconstruction (tests/test_oracle_ring.py). The
compositions relin_crt and bfv_hom_mul are pinned by the identity they satisfy under the zero
secret key (tests/test_oracle_dbfv.py::test_bfv_hom_mul_identity): with key rows (0, g_i), g_i the
CRT recomposition factor of gadget digit i, the product is (round(t l0 r0 / Q), round(t (l0 r1 +
l1 r0) / Q) + round(t l1 r1 / Q)) mod Q up to +-1 per term. Keys of a nonzero secret are not
restated (the reference checks those only through decryption).

Big-integer constants use Python ints, which is what the reference's feanor-math BigInt/Zn
arithmetic computes exactly. Primality (feanor-math's Miller-Rabin, ring.rs:139) is decided with
the deterministic Miller-Rabin base set for 64-bit integers.
"""
import ctypes
import math

import numpy as np

from . import cpu_backend as cb

# --------------------------------------------------------------------------------------------
# lib.rs:25-42 and ring.rs:84-130


def pos_repr(x, modulus):
    return x % modulus


def montgomery_repr(x, modulus):
    return ((x % modulus) * ((1 << 64) % modulus)) % modulus


def _trunc_mod(a, b):
    """Rust's `%` (truncating) for b > 0."""
    r = abs(a) % b
    return -r if a < 0 else r


def cmod_pow2(a, log2_modulus):
    """ring.rs:86-91"""
    assert 1 <= log2_modulus <= 126
    a &= (1 << 128) - 1
    a = a - (1 << 128) if a >> 127 else a
    shift = 128 - log2_modulus
    result = ((a << shift) & ((1 << 128) - 1))
    result = result - (1 << 128) if result >> 127 else result
    result >>= shift
    return result + (1 << log2_modulus) if result == -(1 << (log2_modulus - 1)) else result


def mont_repr(x, p):
    """ring.rs:109: (x << 64) % p with Rust's truncating remainder."""
    return _trunc_mod(x << 64, p)


def mul_mod(a, b, p):
    return _trunc_mod(a * b, p)


def pow_mod(b, e, p):
    return pow(b % p, e, p)


def from_mont_repr(x, p):
    """ring.rs:126-129"""
    r_inv = pow_mod(mont_repr(1, p), p - 2, p)
    return mul_mod(x % p, r_inv, p)


def smallest_lift(x, p):
    """feanor-math `Zn::smallest_lift` for odd p: the representative of smallest absolute value."""
    x %= p
    return x - p if x > p // 2 else x


def is_prime(n):
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    for q in small:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def largest_prime_leq_congruent_to_one(leq_than, congruent_to_one_mod):
    """ring.rs:136-146"""
    assert leq_than > congruent_to_one_mod
    current = leq_than - (leq_than - 1) % congruent_to_one_mod
    while not is_prime(current):
        current -= congruent_to_one_mod
        if current <= 0:
            return None
    return current


def extend_sampled_primes(begin_with, min_bits, max_bits, max_bits_each_modulus, congruent_to_one_mod):
    """ring.rs:150-183, including its f64 bit bookkeeping."""
    assert max_bits > min_bits
    result = list(begin_with)
    current_bits = 0.0
    for p in result:
        current_bits += math.log2(float(p))
    assert math.floor(current_bits) < max_bits
    current_upper_bound = 1 << max_bits_each_modulus
    while current_bits < float(min_bits):
        remaining_bits = float(min_bits) - current_bits
        if remaining_bits < float(max_bits_each_modulus):
            allowed_bits = min(float(max_bits) - current_bits, float(max_bits_each_modulus))
        else:
            required = math.ceil(remaining_bits / float(max_bits_each_modulus))
            allowed_bits = min((float(max_bits) - current_bits) / required, float(max_bits_each_modulus))
        current_upper_bound = min(current_upper_bound, 1 << int(math.floor(allowed_bits)))
        prime = largest_prime_leq_congruent_to_one(current_upper_bound, congruent_to_one_mod)
        current_upper_bound = prime - 1
        while prime in begin_with:
            prime = largest_prime_leq_congruent_to_one(current_upper_bound, congruent_to_one_mod)
            current_upper_bound = prime - 1
        current_bits += math.log2(float(prime))
        result.append(prime)
    return result


# --------------------------------------------------------------------------------------------
# field.rs


def subgroup_elements(gens, m):
    """field.rs:51-66"""
    elements = {1 % m if m > 1 else 1}
    elements = {1}
    frontier = [1]
    while frontier:
        x = frontier.pop()
        for g in gens:
            y = x * g % m
            if y not in elements:
                elements.add(y)
                frontier.append(y)
    return sorted(elements)


def inv_mod(a, n):
    """field.rs:462-475"""
    return pow(a % n, -1, n)


def crt2(a1, m1, a2, m2):
    """field.rs:478-487"""
    if m1 == 1:
        return a2 % m2
    if m2 == 1:
        return a1 % m1
    k = ((a2 + m2 - a1 % m2) % m2) * inv_mod(m1 % m2, m2) % m2
    return (a1 + m1 * k) % (m1 * m2)


FIELD_DATA = {
    # name: (m, m_prime, gal_mprime_lprime_gens, gal_mk_k_gens, gal_lprime_q_cyclic, gal_kq_cyclic)
    # field.rs:529-769
    "field_36864": (8536255, 13, [], [3283177], [(3939811, 4), (7879621, 3)],
                    [(7316791, 2), (1169351, 3), (4560466, 8), (8071246, 64)]),
    "field_32768": (4259905, 13, [1310741], [3604537], [(655371, 4)], [(2949168, 8192)]),
    "field_32768_ex": (3342387, 17, [2162722], [2752556], [(1179667, 8)], [(3276853, 4096)]),
    "field_24576": (851981, 13, [], [196613], [(327686, 4), (458760, 3)], [(393225, 2048)]),
    "field_65536": (5570645, 5, [], [1114131, 2293811], [(4456517, 4)], [(5113926, 2), (5505111, 8192)]),
    "field_49152": (35297665, 13, [29867256], [2715207], [(32582461, 4)],
                    [(1291681, 3), (20965738, 64), (30783429, 64)]),
    "field_49152_ex": (10158235, 5, [], [4063296, 9830566], [(6094942, 4)],
                       [(7798901, 3), (9683156, 2), (4127301, 2048)]),
    "field_192": (90627, 3, [], [30211], [(60419, 2)], [(45427, 2), (4183, 3), (47533, 16)]),
    "field_32": (255, 17, [], [137], [(241, 16)], [(52, 2)]),
    "field_18": (1729, 19, [274], [457], [(1548, 3)], [(1483, 2), (932, 3)]),
    "field_16": (257, 1, [], [2], [], [(155, 16)]),
    "field_2": (7, 1, [], [2], [], [(6, 2)]),
}


# (log2_pb_normalized_pb_op_norm_estimate, log2_sk_pb_op_norm_estimate) of each descriptor
# (field.rs:537-767)
NOISE_ESTIMATES = {
    "field_36864": (8., 6.), "field_32768": (6., 5.), "field_32768_ex": (6., 5.), "field_24576": (5., 4.),
    "field_65536": (6., 5.), "field_49152": (10., 8.), "field_49152_ex": (8., 6.), "field_192": (0., 0.),
    "field_32": (0., 0.), "field_18": (0., 0.), "field_16": (0., 0.), "field_2": (0., 0.),
}


class FieldDescriptor:
    """field.rs:128-300 (the parts the hot path uses)."""

    def __init__(self, name):
        m, m_prime, mp_gens, mk_gens, lq_cyc, kq_cyc = FIELD_DATA[name]
        assert m % 2 == 1 and (m_prime == 1 or m % m_prime == 0)
        self.name = name
        self.m = m
        self.m_prime = m_prime
        self.m_factors = [m_prime] if m_prime != 1 else []
        rest, p = m // m_prime, 3
        while rest > 1:
            if rest % p == 0:
                rest //= p
                assert rest % p != 0
                self.m_factors.append(p)
            p += 2
        phi_m = 1
        for f in self.m_factors:
            phi_m *= f - 1
        self.gal_mprime_lprime = subgroup_elements(mp_gens, m)
        self.gal_mk_k = subgroup_elements(mk_gens, m)
        self.gal_ml = subgroup_elements(mp_gens + mk_gens, m)
        self.lq_cyclic = list(lq_cyc) + list(kq_cyc)
        self.lprime_factor_count = len(lq_cyc)
        self.deg_l = int(np.prod([n for _, n in self.lq_cyclic])) if self.lq_cyclic else 1
        self.deg_lprime = int(np.prod([n for _, n in lq_cyc])) if lq_cyc else 1
        self.deg_k = self.deg_l // self.deg_lprime
        assert phi_m == self.deg_l * len(self.gal_ml)
        self.transversal_index = {}
        for i, g in enumerate(self.lq_transversal()):
            assert g not in self.transversal_index
            self.transversal_index[g] = i

    def scale_q_to_m_period_basis(self):
        """field.rs:447: (-1)^(number of prime factors of m)"""
        return -1 if len(self.m_factors) % 2 == 1 else 1

    def noise_growth_hom_mul_lower_estimate(self, t):
        """field.rs:464-467 (for the secret key weight 128 the reference fixes)"""
        pb_norm, sk_norm = NOISE_ESTIMATES[self.name]
        return math.log2(float(t)) + sk_norm + pb_norm

    def noise_growth_hom_mul_scalar_lower_estimate(self, t):
        """field.rs:483"""
        return math.log2(float(t))

    @property
    def m_k(self):
        return self.m // self.m_prime

    def lq_cyclic_orders(self):
        return [n for _, n in self.lq_cyclic]

    def lprime_q_cyclic_decomposition(self):
        return self.lq_cyclic[: self.lprime_factor_count]

    def kq_cyclic_decomposition(self):
        return self.lq_cyclic[self.lprime_factor_count:]

    def _transversal_of(self, factors):
        """field.rs:275-295"""
        result = [1]
        for g, n in factors:
            powers, power = [], 1
            for _ in range(n):
                powers.append(power)
                power = power * g % self.m
            result = [x * p % self.m for x in result for p in powers]
        return result

    def lq_transversal(self):
        return self._transversal_of(self.lq_cyclic)

    def kq_transversal(self):
        return self._transversal_of(self.kq_cyclic_decomposition())

    def lprime_q_transversal(self):
        return self._transversal_of(self.lprime_q_cyclic_decomposition())

    def coset_index(self, g):
        """field.rs:339-343"""
        assert math.gcd(g % self.m, self.m) == 1
        for h in self.gal_ml:
            idx = self.transversal_index.get(g % self.m * h % self.m)
            if idx is not None:
                return idx
        return None

    def gal_to_shift(self, g):
        """field.rs:327-336"""
        index = self.coset_index(g)
        assert index is not None
        result = [0] * len(self.lq_cyclic)
        for i in reversed(range(len(self.lq_cyclic))):
            n = self.lq_cyclic[i][1]
            result[i] = index % n
            index //= n
        return result

    def rou_order(self):
        """field.rs:440"""
        l = 1
        for _, n in self.lq_cyclic:
            l = l * n // math.gcd(l, n)
        return self.m * l


# --------------------------------------------------------------------------------------------
# ring.rs

EXTENDED_Q_EXTRA_BITS = 60


def _log2_ceil(x):
    """feanor-math `abs_log2_ceil` for x >= 1."""
    return (abs(x) - 1).bit_length()


class WideRingEls:
    """ring.rs:330-532"""

    def __init__(self, ring, count, base):
        self.ring = ring
        self.rns_base = list(base)
        self.data = ring.backend.hal_allocate_i128([count, len(base), ring.field.deg_lprime, ring.field.deg_k])
        self.one = ring.scalar_tensor(base, 1)
        self.neg_one = ring.scalar_tensor(base, -1)
        self.max_p = max(ring.primes[i] for i in base)
        self.min_p = min(ring.primes[i] for i in base)
        self.bound = [0] * count

    def count(self):
        return len(self.bound)

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
        start = rows[0]
        end_all = rows[-1] + 1
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
        assert all(self.bound[i] == 0 for i in rows)
        dst = self.data[rows[0]:rows[-1] + 1]
        self.ring.backend.hal_inner_product_i64(dst, [t[0] for t in terms], [t[2] for t in terms])
        step = sum(lb * rb * self.max_p * self.max_p for (_, lb, _, rb) in terms)
        self.record_step(rows, step)

    def fma_unbounded(self, rows, lhs, rhs):
        """ring.rs:493-500: lhs arbitrary i64 (plain values), rhs bounded by max_p"""
        self._accumulate(rows, lhs, rhs)
        self.record_step(rows, self.max_p << 63)

    def add(self, rows, src, src_bound):
        self._accumulate(rows, src, self.one)
        self.record_step(rows, src_bound * self.max_p * self.max_p)

    def sub(self, rows, src, src_bound):
        self._accumulate(rows, src, self.neg_one)
        self.record_step(rows, src_bound * self.max_p * self.max_p)


class RingContext:
    """ring.rs:594-1852, over a backend with the hal::Backend method names and numpy-like views."""

    def __init__(self, backend, field, log2_q=None, t_bits=0, raw_primes=None):
        self.backend = backend
        self.field = field
        if raw_primes is not None:
            # RingContext::new_raw (ring.rs:657-677): no transform support
            primes = list(raw_primes)
            self.initial_rns_base_len = len(primes)
            self.supports_transform = [False] * len(primes)
        else:
            rou = field.rou_order()
            base = extend_sampled_primes([], log2_q[0], log2_q[1], 57, rou)
            ext = extend_sampled_primes(base, 2 * log2_q[1] + EXTENDED_Q_EXTRA_BITS + t_bits,
                                        2 * log2_q[1] + EXTENDED_Q_EXTRA_BITS + t_bits + 30, 57, rou)
            primes = ext
            self.initial_rns_base_len = len(base)
            self.supports_transform = [True] * len(primes)
            for p in primes:
                assert (p - 1) % rou == 0
        for p in primes:
            assert (1 << 40) <= p <= (1 << 57) - 1
        assert len(set(primes)) == len(primes)
        self.t_bits = t_bits
        self.primes = primes
        self.moduli = [backend.hal_create_modulus(p) for p in primes]
        self._kernels = {}
        self._transform_cache = {}
        self._bconv_cache = {}
        self._gadget_cache = {}
        self._perm_cache = {}

    # --- bases and constants ---------------------------------------------------------------------
    def initial_rns_base(self):
        return list(range(self.initial_rns_base_len))

    def mods_for(self, base):
        return [self.moduli[i] for i in base]

    def modulus_of(self, base):
        q = 1
        for i in base:
            q *= self.primes[i]
        return q

    def extension_rns_base_for(self, base):
        """ring.rs:716-731"""
        assert all(i < self.initial_rns_base_len for i in base)
        required = float(_log2_ceil(self.modulus_of(base)) + self.t_bits + EXTENDED_Q_EXTRA_BITS)
        result, bits = [], 0.0
        for idx in range(self.initial_rns_base_len, len(self.primes)):
            if bits >= required:
                break
            result.append(idx)
            bits += math.log2(float(self.primes[idx]))
        assert bits >= required
        return result

    def _import(self, values, shape):
        t = self.backend.hal_allocate_i64(shape)
        self.backend.hal_import(t, values)
        return t

    def scalar_tensor(self, base, scalar):
        """ring.rs:876-888"""
        vals = [mont_repr(_trunc_mod(scalar, self.primes[i]), self.primes[i]) for i in base]
        return self._import(vals, [1, len(base), 1, 1])

    def inverse_scalar_tensor(self, base, scalar):
        """ring.rs:1165-1177"""
        vals = []
        for i in base:
            p = self.primes[i]
            vals.append(mont_repr(smallest_lift(pow(scalar % p, -1, p), p), p))
        return self._import(vals, [1, len(base), 1, 1])

    def scalar_tensor_big_mont2(self, base, scalar):
        """ring.rs:904-919: mont(mont(smallest_lift(scalar mod p))) per prime"""
        vals = []
        for i in base:
            p = self.primes[i]
            vals.append(mont_repr(mont_repr(smallest_lift(scalar, p), p), p))
        return self._import(vals, [1, len(base), 1, 1])

    def q_over_p_inv_tensor(self, base, scale):
        """ring.rs:1029-1043"""
        q = self.modulus_of(base)
        vals = []
        for i in base:
            p = self.primes[i]
            vals.append(smallest_lift(pow((q // p) % p, -1, p) * (scale % p), p))
        return self._import(vals, [1, len(base), 1, 1])

    def allocate(self, count, base):
        return self.backend.hal_allocate_i64([count, len(base), self.field.deg_lprime, self.field.deg_k])

    # --- transforms (ring.rs:1092-1161, 1252-1415) ------------------------------------------------
    def _rns_factor_kernels(self, p, ntt_lprime, ntt_k):
        """ring.rs:1337-1415"""
        if p in self._kernels:
            return self._kernels[p]
        f = self.field
        m = f.m
        cand = 2
        while True:
            zeta = pow_mod(cand, (p - 1) // m, p)
            if all(pow_mod(zeta, m // fac, p) != 1 for fac in f.m_factors):
                break
            cand += 1
        zeta_mprime = pow_mod(zeta, crt2(1, f.m_prime, 0, f.m_k), p)
        zeta_mk = pow_mod(zeta, crt2(0, f.m_prime, 1, f.m_k), p)

        def traces(zeta_part, part_order, subgroup, transversal):
            sub = np.ascontiguousarray(subgroup, dtype=np.uint64)
            tr = np.ascontiguousarray(transversal, dtype=np.uint64)
            out = np.zeros(len(transversal), dtype=np.int64)
            rc = cb.lib().orc_crt_kernel_traces(
                ctypes.c_int64(p), ctypes.c_int64(zeta_part), ctypes.c_uint64(part_order), ctypes.c_uint64(m),
                sub.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(len(sub)),
                tr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(len(tr)), out.ctypes.data_as(ctypes.c_void_p))
            cb._check(rc)
            return [int(x) for x in out]

        k_lp = traces(zeta_mprime, f.m_prime, f.gal_mprime_lprime, f.lprime_q_transversal())
        k_k = traces(zeta_mk, f.m_k, f.gal_mk_k, f.kq_transversal())

        def to_ntt(kernel, plan, shape, dim):
            t = self._import([mont_repr(x, p) for x in kernel], shape)
            self.backend.hal_ntt_mont64_dim(t, [plan], dim)
            return [int(x) % p for x in self.backend.hal_export(t)]

        n_lp = to_ntt(k_lp, ntt_lprime, [1, 1, f.deg_lprime, 1], 2)
        n_k = to_ntt(k_k, ntt_k, [1, 1, 1, f.deg_k], 3)

        def invert(kernel):
            out = []
            for x in kernel:
                plain = from_mont_repr(x, p)
                assert plain != 0, "CRT kernel is not invertible"
                out.append(mont_repr(pow_mod(plain, p - 2, p), p))
            return out

        res = (n_lp, n_k, invert(n_lp), invert(n_k), k_lp, k_k)
        self._kernels[p] = res
        return res

    def transform_cache(self, base):
        """ring.rs:1252-1303"""
        key = tuple(base)
        if key in self._transform_cache:
            return self._transform_cache[key]
        f = self.field
        orders = f.lq_cyclic_orders()
        nl = f.lprime_factor_count
        ntts_lp, ntts_k, klp, kk, iklp, ikk = [], [], [], [], [], []
        for i in base:
            assert self.supports_transform[i], "RNS factor doesn't support the basis transforms"
            p = self.primes[i]
            plp = self.backend.hal_create_ntt_plan(self.moduli[i], orders[:nl])
            pk = self.backend.hal_create_ntt_plan(self.moduli[i], orders[nl:])
            n_lp, n_k, in_lp, in_k, _, _ = self._rns_factor_kernels(p, plp, pk)
            ntts_lp.append(plp)
            ntts_k.append(pk)
            klp += n_lp
            kk += n_k
            iklp += in_lp
            ikk += in_k
        r = len(base)
        cache = dict(
            mods=self.mods_for(base), ntts_lprime=ntts_lp, ntts_k=ntts_k,
            crt_kernel_lprime=self._import(klp, [1, r, f.deg_lprime, 1]),
            crt_kernel_k=self._import(kk, [1, r, 1, f.deg_k]),
            inv_crt_kernel_lprime=self._import(iklp, [1, r, f.deg_lprime, 1]),
            inv_crt_kernel_k=self._import(ikk, [1, r, 1, f.deg_k]),
        )
        self._transform_cache[key] = cache
        return cache

    def pb_to_crt_tensor(self, data, base):
        """ring.rs:1117-1139"""
        c = self.transform_cache(base)
        be = self.backend
        be.hal_ntt_mul_mont64_dim(data, c["ntts_lprime"], 2, c["crt_kernel_lprime"], c["mods"])
        be.hal_ntt_mul_mont64_dim(data, c["ntts_k"], 3, c["crt_kernel_k"], c["mods"])
        be.hal_intt_mont64(data, c["ntts_lprime"], c["ntts_k"])

    def crt_to_pb_tensor(self, data, base):
        """ring.rs:1143-1161"""
        c = self.transform_cache(base)
        be = self.backend
        be.hal_ntt_mul_mont64_dim(data, c["ntts_lprime"], 2, c["inv_crt_kernel_lprime"], c["mods"])
        be.hal_ntt_mul_mont64_dim(data, c["ntts_k"], 3, c["inv_crt_kernel_k"], c["mods"])
        be.hal_intt_mont64(data, c["ntts_lprime"], c["ntts_k"])

    # --- Galois (ring.rs:1179-1222) ----------------------------------------------------------------
    def galois_perms(self, g):
        shift = tuple(self.field.gal_to_shift(g))
        if shift not in self._perm_cache:
            nl = self.field.lprime_factor_count
            orders = self.field.lq_cyclic_orders()
            self._perm_cache[shift] = (self.backend.hal_create_permutation(orders[:nl], list(shift[:nl])),
                                       self.backend.hal_create_permutation(orders[nl:], list(shift[nl:])))
        return self._perm_cache[shift]

    def apply_galois(self, dst, src, g):
        p_lp, p_k = self.galois_perms(g)
        tmp = self.backend.hal_allocate_i64(src.shape)
        self.backend.hal_permute(tmp, src, p_lp, 2)
        self.backend.hal_permute(dst, tmp, p_k, 3)

    # --- wide accumulators (ring.rs:793-865) -------------------------------------------------------
    def allocate_wide(self, count, base):
        return WideRingEls(self, count, base)

    def reduce_wide_tensor(self, dst, wide, rows):
        rows = list(rows)
        mods = self.mods_for(wide.rns_base)
        start, end_all = rows[0], rows[-1] + 1
        while start < end_all:
            full = wide.needs_full_reduction(start)
            end = start + 1
            while end < end_all and wide.needs_full_reduction(end) == full:
                end += 1
            dst_run = dst[start - rows[0]:end - rows[0]]
            wide_run = wide.data[start:end]
            if full:
                self.backend.hal_red_mont64_double(dst_run, wide_run, mods)
                self.backend.hal_i64_to_mont64(dst_run, mods)
            else:
                self.backend.hal_red_mont64(dst_run, wide_run, mods)
            start = end

    def reduce_wide(self, wide):
        out = self.allocate(wide.count(), wide.rns_base)
        self.reduce_wide_tensor(out, wide, range(wide.count()))
        return out

    def reduce_wide_into(self, dst, wide):
        """ring.rs:829-838"""
        self.reduce_wide_tensor(dst, wide, range(wide.count()))

    # --- base conversion (ring.rs:1482-1852) -------------------------------------------------------
    def bconv_cache(self, src_base, dst_base):
        """BConvCache::new (ring.rs:1606-1661)"""
        key = (tuple(src_base), tuple(dst_base))
        if key in self._bconv_cache:
            return self._bconv_cache[key]
        r, r_ = len(src_base), len(dst_base)
        P = self.primes
        max_p = max(P[i] for i in list(src_base) + list(dst_base))
        min_p = min(P[i] for i in src_base)
        q = self.modulus_of(src_base)
        log2_qtilde = max(64, _log2_ceil(4 * r * max_p + 1))
        qtilde = 1 << log2_qtilde
        correct_max = (qtilde - 1) // min_p + 1
        assert correct_max <= (1 << 63) - 1
        assert r * max_p * correct_max <= (1 << 127) - 1
        acc_before_reduce = ((1 << 127) - 1) // (max_p * max_p)
        assert acc_before_reduce >= 16
        assert -(-r // acc_before_reduce) * max_p <= (1 << 63) - 1
        matrix = []
        for i in dst_base:
            pp = P[i]
            r_sqr = pow(1 << 64, 2, pp)
            for j in src_base:
                matrix.append(smallest_lift((q // P[j]) % pp * r_sqr, pp))
        for j in src_base:
            # ZZi128.rounded_div: qtilde / p has no ties (p odd)
            matrix.append((2 * qtilde + P[j]) // (2 * P[j]))
        q_mod = [smallest_lift(q % P[i] * ((-(1 << 64)) % P[i]), P[i]) for i in dst_base]
        cache = dict(
            log2_qtilde=log2_qtilde, acc=acc_before_reduce,
            q_over_p_inv=self.q_over_p_inv_tensor(src_base, 1),
            matrix=self._import(matrix, [1, r_ + 1, r, 1]),
            q_mod_p=self._import(q_mod, [1, r_, 1, 1]),
            matrix_values=matrix, q_mod_values=q_mod,
        )
        self._bconv_cache[key] = cache
        return cache

    def bconv_tensors(self, dst, dst_base, src, src_base):
        """BConvCache::bconv (ring.rs:1663-1701); destroys src."""
        c = self.bconv_cache(src_base, dst_base)
        be = self.backend
        r, r_ = len(src_base), len(dst_base)
        shape = [src.shape[0], 1, src.shape[2], src.shape[3]]
        assert list(dst.shape) == [shape[0], r_, shape[2], shape[3]]
        src_mods = self.mods_for(src_base)
        be.hal_mul_mont64(src, c["q_over_p_inv"], src_mods)
        dst_mods = self.mods_for(dst_base)
        correction = be.hal_allocate_i128(shape)
        tmp = be.hal_allocate_i128([shape[0], r_ + 1, shape[2], shape[3]])
        for chunk, j in enumerate(range(0, r, c["acc"])):
            e = min(j + c["acc"], r)
            be.hal_matmul_i64(tmp, c["matrix"][:, :, j:e, :], src[:, j:e])
            if chunk == 0:
                be.hal_red_mont64_double(dst, tmp[:, :r_], dst_mods)
                be.hal_assign_i128(correction, tmp[:, r_:])
            else:
                be.hal_add_red_mont64_double(dst, tmp[:, :r_], dst_mods)
                be.hal_add_i128(correction, tmp[:, r_:])
        be.hal_bconv_correct_mont64(dst, correction, c["q_mod_p"], dst_mods, c["log2_qtilde"])
        be.hal_i64_to_mont64(dst, dst_mods)

    def gadget_bconv_cache(self, base, digits):
        """GadgetBConvCache::new (ring.rs:1730-1807)"""
        key = (tuple(base), tuple(digits))
        if key in self._gadget_cache:
            return self._gadget_cache[key]
        P = self.primes
        r, d = len(base), len(digits)
        max_p = max(P[i] for i in base)
        min_p = min(P[i] for i in base)
        max_len = max(b - a for a, b in digits)
        log2_qtilde = max(64, _log2_ceil(4 * max_len * max_p + 1))
        qtilde = 1 << log2_qtilde
        correct_max = (qtilde - 1) // min_p + 1
        assert correct_max <= (1 << 63) - 1, "coefficients of correction vector don't fit into an i64"
        assert max_len * max_p * correct_max < 1 << 127, "correction vector inner product overflows"
        acc_before_reduce = (((1 << 127) - 1) // (max_p * max_p)) % (1 << 64)  # `as usize` truncates
        assert acc_before_reduce >= 16, "modular reductions would have to be performed too often"
        assert -(-r // acc_before_reduce) * max_p <= (1 << 63) - 1, "overflow during accumulation"
        dmods = [self.modulus_of(base[a:b]) for a, b in digits]
        d_over_p_inv = [0] * r
        for (a, b), dm in zip(digits, dmods):
            for u in range(a, b):
                pu = P[base[u]]
                d_over_p_inv[u] = smallest_lift(pow((dm // pu) % pu, -1, pu), pu)
        matrix = [0] * (d * (r + 1) * r)
        d_mod_p = []
        for i, ((a, b), dm) in enumerate(zip(digits, dmods)):
            for j in range(r):
                pj = P[base[j]]
                r_sqr = pow(1 << 64, 2, pj)
                for u in range(a, b):
                    pu = P[base[u]]
                    matrix[(i * (r + 1) + j) * r + u] = smallest_lift((dm // pu) % pj * r_sqr, pj)
                d_mod_p.append(smallest_lift(dm % pj * ((-(1 << 64)) % pj), pj))
            for u in range(a, b):
                pu = P[base[u]]
                matrix[(i * (r + 1) + r) * r + u] = (2 * qtilde + pu) // (2 * pu)
        cache = dict(
            log2_qtilde=log2_qtilde, acc=acc_before_reduce,
            d_over_p_inv=self._import(d_over_p_inv, [1, r, 1, 1]),
            matrix=self._import(matrix, [d, r + 1, r, 1]),
            d_mod_p=self._import(d_mod_p, [d, r, 1, 1]),
        )
        self._gadget_cache[key] = cache
        return cache

    def bconv_gadget_tensors(self, dst, src, base, digits):
        """GadgetBConvCache::bconv (ring.rs:1809-1851); destroys src."""
        c = self.gadget_bconv_cache(base, digits)
        be = self.backend
        r, d = len(base), len(digits)
        mods = self.mods_for(base)
        be.hal_mul_mont64(src, c["d_over_p_inv"], mods)
        correction = be.hal_allocate_i128([d, 1, src.shape[2], src.shape[3]])
        tmp = be.hal_allocate_i128([d, r + 1, src.shape[2], src.shape[3]])
        for chunk, u in enumerate(range(0, r, c["acc"])):
            e = min(u + c["acc"], r)
            be.hal_matmul_i64(tmp, c["matrix"][:, :, u:e, :], src[:, u:e])
            if chunk == 0:
                be.hal_red_mont64_double(dst, tmp[:, :r], mods)
                be.hal_assign_i128(correction, tmp[:, r:])
            else:
                be.hal_add_red_mont64_double(dst, tmp[:, :r], mods)
                be.hal_add_i128(correction, tmp[:, r:])
        be.hal_bconv_correct_mont64(dst, correction, c["d_mod_p"], mods, c["log2_qtilde"])
        be.hal_i64_to_mont64(dst, mods)


# --------------------------------------------------------------------------------------------
# bfv.rs sequences on raw tensors


def select_digits(digits, rns_base_len):
    """bfv.rs:388-401"""
    assert digits <= rns_base_len
    small = rns_base_len // digits
    large = rns_base_len % digits
    res, cur = [], 0
    for i in range(digits):
        n = small + 1 if i < large else small
        res.append((cur, cur + n))
        cur += n
    return res


def mod_switch_down(ring, els, base, new_len):
    """bfv.rs:628-662 on a period-basis tensor `els` over `base`; returns the new tensor."""
    be = ring.backend
    kept, dropped = list(base[:new_len]), list(base[new_len:])
    count = els.shape[0]
    if not dropped:
        out = ring.allocate(count, kept)
        be.hal_assign(out, els)
        return out
    dropped_part = ring.allocate(count, dropped)
    be.hal_assign(dropped_part, els[:, new_len:])
    tmp = ring.allocate(count, kept)
    ring.bconv_tensors(tmp, kept, dropped_part, dropped)
    result = ring.allocate(count, kept)
    be.hal_assign(result, els[:, :new_len])
    be.hal_sub_i64(result, tmp)
    inv = ring.inverse_scalar_tensor(kept, ring.modulus_of(dropped))
    be.hal_mul_mont64(result, inv, ring.mods_for(kept))
    return result


def lift_to_extended(ring, pb, base, crt=None):
    """bfv.rs:564-595; returns (tensor over base + extension in CRT basis, lifted base)."""
    be = ring.backend
    base_len = len(base)
    ext = ring.extension_rns_base_for(base)
    lifted = list(base) + ext
    result = ring.allocate(pb.shape[0], lifted)
    tmp = ring.allocate(pb.shape[0], base)
    be.hal_assign(tmp, pb)
    ring.bconv_tensors(result[:, base_len:], ext, tmp, base)
    if crt is not None:
        ring.pb_to_crt_tensor(result[:, base_len:], ext)
        be.hal_assign(result[:, :base_len], crt)
    else:
        be.hal_assign(result[:, :base_len], pb)
        ring.pb_to_crt_tensor(result, lifted)
    return result, lifted


def rescale_to_base(ring, x, lifted, base_len, t):
    """bfv.rs:601-622; `x` is a period-basis tensor over `lifted` (destroyed)."""
    be = ring.backend
    count = x.shape[0]
    base, ext = lifted[:base_len], lifted[base_len:]
    q = ring.modulus_of(base)
    be.hal_mul_mont64(x, ring.scalar_tensor(lifted, t), ring.mods_for(lifted))
    tmp = ring.allocate(count, ext)
    ring.bconv_tensors(tmp, ext, x[:, :base_len], base)
    be.hal_sub_i64(x[:, base_len:], tmp)
    be.hal_mul_mont64(x[:, base_len:], ring.inverse_scalar_tensor(ext, q), ring.mods_for(ext))
    result = ring.allocate(count, base)
    ring.bconv_tensors(result, base, x[:, base_len:], ext)
    return result


def hom_mul_raw_acc(ring, acc, acc_row, lhs, rhs):
    """bfv.rs:667-683"""
    l0, l1 = lhs[0:1], lhs[1:2]
    r0, r1 = rhs[0:1], rhs[1:2]
    acc.fma(range(acc_row, acc_row + 1), l0, 1, r0, 1)
    acc.fma(range(acc_row + 1, acc_row + 2), l0, 1, r1, 1)
    acc.fma(range(acc_row + 1, acc_row + 2), l1, 1, r0, 1)
    acc.fma(range(acc_row + 2, acc_row + 3), l1, 1, r1, 1)


def gadget_decompose(ring, c1_pb, base, digits):
    """bfv.rs:408-426: [d, r, ., .] digits in CRT basis."""
    be = ring.backend
    result = ring.allocate(len(digits), base)
    tmp = ring.allocate(1, base)
    be.hal_assign(tmp, c1_pb)
    ring.bconv_gadget_tensors(result, tmp, base, digits)
    ring.pb_to_crt_tensor(result, base)
    return result


def key_switch_inner_product(ring, acc, parts, keys, digits):
    """The inner product of key_switch_acc / bfv_hom_gal (bfv.rs:509-527, 822-836):
    acc[0..2] = sum_i parts[i] * keys[i] with keys[i] a [2, r, ., .] tensor."""
    terms = [(parts[i:i + 1], 1, keys[i], 1) for i in range(len(digits))]
    acc.inner_product(range(0, 2), terms)


def hom_gal_one(ring, ct, base, g, parts, keys, digits):
    """One rotation of bfv_hom_gal (bfv.rs:816-839) given the hoisted gadget digits `parts`."""
    be = ring.backend
    gal_parts = ring.allocate(parts.shape[0], base)
    gal_c0 = ring.allocate(1, base)
    ring.apply_galois(gal_parts, parts, g)
    ring.apply_galois(gal_c0, ct[0:1], g)
    acc = ring.allocate_wide(2, base)
    key_switch_inner_product(ring, acc, gal_parts, keys, digits)
    acc.add(range(0, 1), gal_c0, 1)
    return ring.reduce_wide(acc)


def clip_digits(digits, base_len):
    """bfv.rs:120-126"""
    return [(a, min(b, base_len)) for a, b in digits if a < base_len]


def relin_crt(ring, prod, base, keys, digits):
    """relin (bfv.rs:699-741) with CRT-basis output: `prod` is a [3, r, ., .] period-basis
    product (destroyed); keys[i] is the [2, r_key, ., .] CRT-basis key of digit i over a base of
    which `base` is a prefix."""
    r = len(base)
    acc = ring.allocate_wide(2, base)
    digits = clip_digits(digits, r)
    parts = gadget_decompose(ring, prod[2:3], base, digits)           # key_switch_acc, bfv.rs:498-528
    key_switch_inner_product(ring, acc, parts, [np.ascontiguousarray(k[:, :r]) for k in keys], digits)
    ring.pb_to_crt_tensor(prod[0:2], base)
    acc.add(range(0, 2), prod[0:2], 1)
    return ring.reduce_wide(acc)


def relin(ring, prod, base, keys, digits, output_basis):
    """relin (bfv.rs:699-741) of a [3, r, ., .] period-basis product (destroyed) of bound 1:
    "crt" is relin_crt (bound 1); "pb" reduces the key-switch part, transforms it to period basis
    and adds c0, c1 unreduced (add_els, ring.rs:990-1023). Returns (ciphertext, bound)."""
    if output_basis == "crt":
        return relin_crt(ring, prod, base, keys, digits), 1
    assert output_basis == "pb"
    r = len(base)
    acc = ring.allocate_wide(2, base)
    digits = clip_digits(digits, r)
    parts = gadget_decompose(ring, prod[2:3], base, digits)
    key_switch_inner_product(ring, acc, parts, [np.ascontiguousarray(k[:, :r]) for k in keys], digits)
    switched = ring.reduce_wide(acc)
    ring.crt_to_pb_tensor(switched, base)
    ring.backend.hal_add_i64(switched, prod[0:2])
    return switched, 2


def bfv_hom_mul(ring, t, lhs, rhs, base, keys, digits):
    """bfv_hom_mul (bfv.rs:747-775) with CRT-basis output, for CRT-basis ciphertexts lhs, rhs
    ([2, r, ., .] each) over `base`."""
    base_len = len(base)

    def lift(ct):
        ct_pb = ct.copy()
        ring.crt_to_pb_tensor(ct_pb, base)
        return lift_to_extended(ring, ct_pb, base, crt=ct)

    lhs_l, lifted = lift(lhs)
    rhs_l, _ = lift(rhs)
    acc = ring.allocate_wide(3, lifted)
    hom_mul_raw_acc(ring, acc, 0, lhs_l, rhs_l)
    prod = ring.reduce_wide(acc)
    ring.crt_to_pb_tensor(prod, lifted)
    rescaled = rescale_to_base(ring, prod, lifted, base_len, t)
    return relin_crt(ring, rescaled, base, keys, digits)
