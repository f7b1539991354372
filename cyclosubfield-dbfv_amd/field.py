"""The field descriptor of cyclosubfield-dbfv/src/field.rs on the host, for the set-up of
bootstrapping: the Galois groups as subgroups of (Z/mZ)*, the transversals that order the period
basis, the per-prime data of the homomorphic CRT transform, the period-basis sign conventions and
the noise estimates. The descriptor table itself lives in the library (csd_field_data, the data of
field.rs:529-769); everything here is derived from it as FieldDescriptor::new derives it
(field.rs:159-253)."""
import ctypes
import math

from .hal import check, lib

# (log2_pb_normalized_pb_op_norm_estimate, log2_sk_pb_op_norm_estimate) per descriptor
# (field.rs:537-767)
NOISE_ESTIMATES = {
    "field_36864": (8., 6.), "field_32768": (6., 5.), "field_32768_ex": (6., 5.), "field_24576": (5., 4.),
    "field_65536": (6., 5.), "field_49152": (10., 8.), "field_49152_ex": (8., 6.), "field_192": (0., 0.),
    "field_32": (0., 0.), "field_18": (0., 0.), "field_16": (0., 0.), "field_2": (0., 0.),
}

SK_HWT = 128  # bfv.rs:41


def subgroup_elements(gens, m):
    """field.rs:52-67: the subgroup of (Z/mZ)* generated by gens, sorted"""
    elements, frontier = {1}, [1]
    while frontier:
        x = frontier.pop()
        for g in gens:
            y = x * g % m
            if y not in elements:
                elements.add(y)
                frontier.append(y)
    return sorted(elements)


def inv_mod(a, n):
    """field.rs:497-508"""
    return pow(a % n, -1, n)


def crt2(a1, m1, a2, m2):
    """field.rs:510-520"""
    if m1 == 1:
        return a2 % m2
    if m2 == 1:
        return a1 % m1
    k = ((a2 - a1) % m2) * inv_mod(m1 % m2, m2) % m2
    return (a1 + m1 * k) % (m1 * m2)


def element_order(x, n):
    """field.rs:486-495"""
    cur, order = x % n, 1
    while cur != 1:
        cur = cur * x % n
        order += 1
    return order


class PrimeInfo:
    """PrimeInfo (field.rs:93-124)"""

    def __init__(self, p, i_p_gen, f_p, h_p, d_p):
        self.p, self.i_p_gen, self.f_p, self.h_p, self.d_p = p, i_p_gen, f_p, h_p, d_p

    def j_p_order(self):
        return (self.p - 1) // len(self.h_p)


class FieldDescriptor:
    """FieldDescriptor (field.rs:128-484), built from the library's descriptor table"""

    def __init__(self, name):
        m, mp = ctypes.c_uint64(), ctypes.c_uint64()
        cap = 16
        mpg, mkg, cg, co = ((ctypes.c_uint64 * cap)() for _ in range(4))
        nmp, nmk, nc, nlp = (ctypes.c_size_t() for _ in range(4))
        check(lib().csd_field_data(name.encode(), ctypes.byref(m), ctypes.byref(mp), mpg, ctypes.byref(nmp), mkg,
                                   ctypes.byref(nmk), cg, co, ctypes.byref(nc), ctypes.byref(nlp), ctypes.c_size_t(cap)))
        self.name = name
        self.m, self.m_prime = m.value, mp.value
        mp_gens = [mpg[i] for i in range(nmp.value)]
        mk_gens = [mkg[i] for i in range(nmk.value)]
        self.lq_cyclic = [(cg[i], co[i]) for i in range(nc.value)]
        self.lprime_factor_count = nlp.value
        # all prime factors of m, m' first (field.rs:159-178)
        self.m_factors = [self.m_prime] if self.m_prime != 1 else []
        rest, p = self.m // self.m_prime, 2
        while rest > 1:
            if rest % p == 0:
                rest //= p
                if rest % p == 0:
                    raise ValueError("m must be squarefree")
                self.m_factors.append(p)
            p += 1
        self.deg_l = math.prod(n for _, n in self.lq_cyclic)
        self.deg_lprime = math.prod(n for _, n in self.lq_cyclic[:self.lprime_factor_count])
        self.deg_k = self.deg_l // self.deg_lprime
        self.gal_mprime_lprime = subgroup_elements(mp_gens, self.m)
        self.gal_mk_k = subgroup_elements(mk_gens, self.m)
        self.gal_ml = subgroup_elements(mp_gens + mk_gens, self.m)
        self.noise_pb, self.noise_sk = NOISE_ESTIMATES[name]
        self._prime_info = self._init_prime_info()

    def _init_prime_info(self):
        """field.rs:217-243"""
        ml = set(self.gal_ml)
        out = []
        for p in self.m_factors:
            pf, rest, f = [], p - 1, 2
            while rest > 1:
                if rest % f == 0:
                    pf.append(f)
                    while rest % f == 0:
                        rest //= f
                f += 1
            gen = next(g for g in range(2, p) if all(pow(g, (p - 1) // r, p) != 1 for r in pf))
            ipg = crt2(gen, p, 1, self.m // p)
            f_p = element_order(2 % p, p)
            h_p = [x for x in (pow(ipg, e, self.m) for e in range(p - 1)) if x in ml]
            d_p = f_p // math.gcd(f_p, len(h_p))
            out.append(PrimeInfo(p, ipg, f_p, h_p, d_p))
        return out

    def prime_info(self, p):
        for info in self._prime_info:
            if info.p == p:
                return info
        raise ValueError("prime doesn't divide m")

    def m_k(self):
        return self.m // self.m_prime

    def m_k_factors(self):
        """field.rs:271"""
        return self.m_factors if self.m_prime == 1 else self.m_factors[1:]

    def _transversal_of(self, factors):
        """field.rs:299-317"""
        result = [1]
        for g, n in factors:
            powers, power = [], 1
            for _ in range(n):
                powers.append(power)
                power = power * g % self.m
            result = [x * q % self.m for x in result for q in powers]
        return result

    def lq_transversal(self):
        return self._transversal_of(self.lq_cyclic)

    def lprime_q_transversal(self):
        return self._transversal_of(self.lq_cyclic[:self.lprime_factor_count])

    def kq_transversal(self):
        return self._transversal_of(self.lq_cyclic[self.lprime_factor_count:])

    def lprime_residue_degree(self):
        """field.rs:372-374"""
        return 1 if self.m_prime == 1 else self.prime_info(self.m_prime).d_p

    def mprime_frobenius_order(self):
        """field.rs:378-380"""
        return 1 if self.m_prime == 1 else self.prime_info(self.m_prime).f_p

    def crt_transform_degree(self):
        """field.rs:364-367"""
        d = 1
        for info in self._prime_info:
            if info.p != self.m_prime:
                d = d * info.d_p // math.gcd(d, info.d_p)
        return d

    def lprime_embedding_transversal(self):
        """field.rs:392-411: one element of the L'-transversal per prime of L' above 2"""
        if self.m_prime == 1:
            return [1]
        mp = self.m_prime
        decomposition = subgroup_elements([2 % mp] + [h % mp for h in self.gal_mprime_lprime], mp)
        seen, result = set(), []
        for g in self.lprime_q_transversal():
            coset = min(g % mp * s % mp for s in decomposition)
            if coset not in seen:
                seen.add(coset)
                result.append(g)
        if len(result) != self.deg_lprime // self.lprime_residue_degree():
            raise AssertionError("embedding transversal has the wrong size")
        return result

    def scale_q_to_m_period_basis(self):
        """field.rs:447"""
        return -1 if len(self.m_factors) % 2 == 1 else 1

    def scale_m_k_to_m_period_basis(self):
        """field.rs:457"""
        return 1 if self.m_prime == 1 else -1

    def scale_mprime_to_m_period_basis(self):
        """field.rs:451-453"""
        return self.scale_q_to_m_period_basis() * self.scale_m_k_to_m_period_basis()

    def lprime_trace_one_coeffs(self):
        """field.rs:429-435: alpha in L' with Tr_(L'|Q)(alpha) = 1"""
        mu = 1 if self.m_prime == 1 else -1
        coeffs = [0] * self.deg_lprime
        coeffs[0] = self.scale_mprime_to_m_period_basis() * mu
        return coeffs

    def noise_growth_hom_mul_lower_estimate(self, t):
        """field.rs:464-467 (secret keys of weight 128)"""
        return math.log2(float(t)) + self.noise_sk + self.noise_pb

    def noise_growth_hom_mul_plain_lower_estimate(self, t):
        """field.rs:474-476"""
        return math.log2(float(t)) + self.noise_pb

    def noise_growth_hom_mul_scalar_lower_estimate(self, t):
        """field.rs:483"""
        return math.log2(float(t))

    def log2_pb_expansion_factor_bound(self):
        """ring.rs:770-772"""
        return (self.m - 1).bit_length() + len(self.m_factors) + 1
