// Ring layer of the MI355X backend: the RNS ring O_L/q of cyclosubfield-dbfv's
// `ring::RingContext` (cyclosubfield-dbfv/src/ring.rs:594-1852) over the HAL of csd_hal.hip.
//
// Host responsibilities (one-time setup, as in the reference): field descriptors, RNS prime
// sampling, CRT-kernel traces, base-conversion constants. Device work: every transform,
// conversion and reduction goes through the HAL entry points, so the results are exactly the
// reference's HAL call sequences (ring.rs:1117-1161, 1663-1701, 1809-1851; bfv.rs:628-662).
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/csdbfv_ring.h"
#include "csd_internal.hpp"

// ---------------------------------------------------------------------------------------------
// field descriptors (field.rs)

namespace {

uint64_t mulmod_u(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a * b) % m); }

uint64_t inv_mod_u(uint64_t a, uint64_t n) {
    i128 old_r = (i128)(a % n), r = (i128)n, old_s = 1, s = 0;
    while (r != 0) {
        i128 q = old_r / r, t = old_r - q * r;
        old_r = r;
        r = t;
        t = old_s - q * s;
        old_s = s;
        s = t;
    }
    if (!(old_r == 1 || old_r == -1)) throw CsdError("elements are not coprime");
    i128 v = (old_s * old_r) % (i128)n;
    if (v < 0) v += n;
    return (uint64_t)v;
}

uint64_t crt2(uint64_t a1, uint64_t m1, uint64_t a2, uint64_t m2) {
    if (m1 == 1) return a2 % m2;
    if (m2 == 1) return a1 % m1;
    uint64_t k = mulmod_u((a2 + m2 - a1 % m2) % m2, inv_mod_u(m1 % m2, m2), m2);
    return (uint64_t)(((u128)a1 + (u128)m1 * k) % ((u128)m1 * m2));
}

std::vector<uint64_t> subgroup(const std::vector<uint64_t>& gens, uint64_t m) {
    std::vector<uint64_t> out{1};
    std::vector<uint64_t> frontier{1};
    std::unordered_map<uint64_t, bool> seen{{1, true}};
    while (!frontier.empty()) {
        uint64_t x = frontier.back();
        frontier.pop_back();
        for (uint64_t g : gens) {
            uint64_t y = mulmod_u(x, g, m);
            if (!seen.count(y)) {
                seen[y] = true;
                out.push_back(y);
                frontier.push_back(y);
            }
        }
    }
    std::sort(out.begin(), out.end());
    return out;
}

struct FieldData {
    const char* name;
    uint64_t m, m_prime;
    std::vector<uint64_t> mp_gens, mk_gens;
    std::vector<std::pair<uint64_t, size_t>> lq_cyc, kq_cyc;
};

// field.rs:529-769
const std::vector<FieldData>& field_table() {
    static const std::vector<FieldData> t = {
        {"field_36864", 8536255, 13, {}, {3283177}, {{3939811, 4}, {7879621, 3}},
         {{7316791, 2}, {1169351, 3}, {4560466, 8}, {8071246, 64}}},
        {"field_32768", 4259905, 13, {1310741}, {3604537}, {{655371, 4}}, {{2949168, 8192}}},
        {"field_32768_ex", 3342387, 17, {2162722}, {2752556}, {{1179667, 8}}, {{3276853, 4096}}},
        {"field_24576", 851981, 13, {}, {196613}, {{327686, 4}, {458760, 3}}, {{393225, 2048}}},
        {"field_65536", 5570645, 5, {}, {1114131, 2293811}, {{4456517, 4}}, {{5113926, 2}, {5505111, 8192}}},
        {"field_49152", 35297665, 13, {29867256}, {2715207}, {{32582461, 4}},
         {{1291681, 3}, {20965738, 64}, {30783429, 64}}},
        {"field_49152_ex", 10158235, 5, {}, {4063296, 9830566}, {{6094942, 4}},
         {{7798901, 3}, {9683156, 2}, {4127301, 2048}}},
        {"field_192", 90627, 3, {}, {30211}, {{60419, 2}}, {{45427, 2}, {4183, 3}, {47533, 16}}},
        {"field_32", 255, 17, {}, {137}, {{241, 16}}, {{52, 2}}},
        {"field_18", 1729, 19, {274}, {457}, {{1548, 3}}, {{1483, 2}, {932, 3}}},
        {"field_16", 257, 1, {}, {2}, {}, {{155, 16}}},
        {"field_2", 7, 1, {}, {2}, {}, {{6, 2}}},
    };
    return t;
}

struct Field {
    std::string name;
    uint64_t m, m_prime;
    std::vector<uint64_t> m_factors;
    std::vector<std::pair<uint64_t, size_t>> lq_cyc;
    size_t lp_count, deg_l, deg_k, deg_lp;
    std::vector<uint64_t> gal_ml, gal_mp_lp, gal_mk_k;
    std::unordered_map<uint64_t, size_t> tindex;

    uint64_t m_k() const { return m / m_prime; }

    std::vector<uint64_t> transversal_of(size_t lo, size_t hi) const {  // field.rs:275-295
        std::vector<uint64_t> res{1};
        for (size_t f = lo; f < hi; f++) {
            uint64_t g = lq_cyc[f].first;
            size_t n = lq_cyc[f].second;
            std::vector<uint64_t> powers;
            uint64_t p = 1;
            for (size_t e = 0; e < n; e++) {
                powers.push_back(p);
                p = mulmod_u(p, g, m);
            }
            std::vector<uint64_t> next;
            for (uint64_t x : res)
                for (uint64_t q : powers) next.push_back(mulmod_u(x, q, m));
            res = next;
        }
        return res;
    }
    std::vector<uint64_t> lq_transversal() const { return transversal_of(0, lq_cyc.size()); }
    std::vector<uint64_t> lp_transversal() const { return transversal_of(0, lp_count); }
    std::vector<uint64_t> kq_transversal() const { return transversal_of(lp_count, lq_cyc.size()); }
    std::vector<size_t> orders() const {
        std::vector<size_t> o;
        for (auto& c : lq_cyc) o.push_back(c.second);
        return o;
    }
    uint64_t rou_order() const {  // field.rs:440
        uint64_t l = 1;
        for (auto& c : lq_cyc) {
            uint64_t a = l, b = c.second;
            while (b) {
                uint64_t t = a % b;
                a = b;
                b = t;
            }
            l = l / a * c.second;
        }
        return m * l;
    }
    std::vector<int64_t> gal_to_shift(uint64_t g) const {  // field.rs:327-343
        size_t idx = SIZE_MAX;
        for (uint64_t h : gal_ml) {
            auto it = tindex.find(mulmod_u(g % m, h, m));
            if (it != tindex.end()) {
                idx = it->second;
                break;
            }
        }
        if (idx == SIZE_MAX) throw CsdError("not an element of (Z/mZ)* or inconsistent subgroup data");
        std::vector<int64_t> res(lq_cyc.size());
        for (size_t i = lq_cyc.size(); i-- > 0;) {
            res[i] = (int64_t)(idx % lq_cyc[i].second);
            idx /= lq_cyc[i].second;
        }
        return res;
    }
};

std::shared_ptr<Field> make_field(const std::string& name) {  // FieldDescriptor::new (field.rs:159-228)
    static std::map<std::string, std::shared_ptr<Field>> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(name);
    if (it != cache.end()) return it->second;
    const FieldData* d = nullptr;
    for (auto& x : field_table())
        if (name == x.name) d = &x;
    if (!d) throw CsdError("unknown field " + name);
    auto f = std::make_shared<Field>();
    f->name = name;
    f->m = d->m;
    f->m_prime = d->m_prime;
    if (d->m_prime != 1) f->m_factors.push_back(d->m_prime);
    uint64_t rest = d->m / d->m_prime, p = 3, phi = 1;
    while (rest > 1) {
        if (rest % p == 0) {
            rest /= p;
            f->m_factors.push_back(p);
        }
        p += 2;
    }
    for (uint64_t q : f->m_factors) phi *= q - 1;
    f->gal_mp_lp = subgroup(d->mp_gens, f->m);
    f->gal_mk_k = subgroup(d->mk_gens, f->m);
    std::vector<uint64_t> all = d->mp_gens;
    all.insert(all.end(), d->mk_gens.begin(), d->mk_gens.end());
    f->gal_ml = subgroup(all, f->m);
    f->lq_cyc = d->lq_cyc;
    f->lq_cyc.insert(f->lq_cyc.end(), d->kq_cyc.begin(), d->kq_cyc.end());
    f->lp_count = d->lq_cyc.size();
    f->deg_l = 1;
    f->deg_lp = 1;
    for (auto& c : f->lq_cyc) f->deg_l *= c.second;
    for (auto& c : d->lq_cyc) f->deg_lp *= c.second;
    f->deg_k = f->deg_l / f->deg_lp;
    if (phi != f->deg_l * f->gal_ml.size()) throw CsdError("transversal size must match [L : Q]");
    auto tr = f->lq_transversal();
    for (size_t i = 0; i < tr.size(); i++) {
        if (f->tindex.count(tr[i])) throw CsdError("generators do not induce an isomorphism");
        f->tindex[tr[i]] = i;
    }
    cache[name] = f;
    return f;
}

// ---------------------------------------------------------------------------------------------
// RNS primes (ring.rs:136-183)

uint64_t powmod_u(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mulmod_u(r, b, m);
        b = mulmod_u(b, b, m);
        e >>= 1;
    }
    return r;
}

bool is_prime_u64(uint64_t n) {  // deterministic Miller-Rabin for 64-bit integers
    if (n < 2) return false;
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t q : bases)
        if (n % q == 0) return n == q;
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        s++;
    }
    for (uint64_t a : bases) {
        uint64_t x = powmod_u(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < s; i++) {
            x = mulmod_u(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}

int64_t largest_prime_leq_congruent_to_one(int64_t leq_than, int64_t mod) {
    if (!(leq_than > mod)) throw CsdError("prime sampling: bound below the congruence modulus");
    int64_t cur = leq_than - (leq_than - 1) % mod;
    while (!is_prime_u64((uint64_t)cur)) {
        cur -= mod;
        if (cur <= 0) throw CsdError("prime sampling: no prime found");
    }
    return cur;
}

std::vector<int64_t> extend_sampled_primes(const std::vector<int64_t>& begin_with, size_t min_bits, size_t max_bits,
                                           size_t max_each, int64_t mod) {
    std::vector<int64_t> res = begin_with;
    double cur_bits = 0.0;
    for (int64_t p : res) cur_bits += std::log2((double)p);
    int64_t upper = (int64_t)1 << max_each;
    while (cur_bits < (double)min_bits) {
        double remaining = (double)min_bits - cur_bits;
        double allowed;
        if (remaining < (double)max_each) {
            allowed = std::min((double)max_bits - cur_bits, (double)max_each);
        } else {
            double req = std::ceil(remaining / (double)max_each);
            allowed = std::min(((double)max_bits - cur_bits) / req, (double)max_each);
        }
        upper = std::min(upper, (int64_t)1 << (int)std::floor(allowed));
        int64_t prime = largest_prime_leq_congruent_to_one(upper, mod);
        upper = prime - 1;
        while (std::find(begin_with.begin(), begin_with.end(), prime) != begin_with.end()) {
            prime = largest_prime_leq_congruent_to_one(upper, mod);
            upper = prime - 1;
        }
        cur_bits += std::log2((double)prime);
        res.push_back(prime);
    }
    return res;
}

// the smallest representative in absolute value of x mod p (odd p)
int64_t smallest_lift(uint64_t x_mod_p, int64_t p) {
    return x_mod_p > (uint64_t)p / 2 ? (int64_t)x_mod_p - p : (int64_t)x_mod_p;
}

// (x << 64) % p with a truncating remainder (ring.rs:109)
int64_t mont_repr(int64_t x, int64_t p) { return (int64_t)(((i128)x << 64) % (i128)p); }

uint64_t mod_of(int64_t x, int64_t p) {
    int64_t r = x % p;
    return (uint64_t)(r < 0 ? r + p : r);
}

// bit length of a product of positive 64-bit integers
size_t product_bit_length(const std::vector<int64_t>& xs) {
    std::vector<uint32_t> a{1};
    for (int64_t x : xs) {
        uint64_t lo = (uint64_t)x & 0xffffffffu, hi = (uint64_t)x >> 32;
        std::vector<uint32_t> r(a.size() + 3, 0);
        for (int half = 0; half < 2; half++) {
            uint64_t m = half ? hi : lo;
            uint64_t carry = 0;
            for (size_t i = 0; i < a.size(); i++) {
                uint64_t t = (uint64_t)a[i] * m + r[i + half] + carry;
                r[i + half] = (uint32_t)t;
                carry = t >> 32;
            }
            size_t k = a.size() + half;
            while (carry) {
                uint64_t t = (uint64_t)r[k] + carry;
                r[k++] = (uint32_t)t;
                carry = t >> 32;
            }
        }
        while (r.size() > 1 && r.back() == 0) r.pop_back();
        a = r;
    }
    size_t top = a.back() ? 32 - __builtin_clz(a.back()) : 0;
    return (a.size() - 1) * 32 + top;
}

struct DevBuf {
    csd_ctx* ctx = nullptr;
    void* p = nullptr;
    DevBuf() {}
    DevBuf(csd_ctx* c, size_t bytes) : ctx(c), p(c->alloc_bytes(bytes)) {}
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) ctx->free_bytes(p);
    }
};

// ---------------------------------------------------------------------------------------------
// CRT kernels on the device (compute_rns_factor_kernels, ring.rs:1337-1415). The trace of
// zeta_p^(g^-1) over the Galois subgroup, for every transversal element g, is a sum of powers
// zeta_p^e whose exponents e = (x g^-1 mod m) mod order do not depend on the prime; the host
// computes them once per ring, and one thread per g sums its powers (square-and-multiply in
// Montgomery form). The kernels' inverses are Fermat powers, one thread per entry. Both give the
// canonical residues the host formulas give.

__device__ __forceinline__ int64_t canon_mod(const CsdMod& m, int64_t v) {
    if (v < 0) v += m.N;
    if (v >= m.N) v -= m.N;
    return v;
}

// base_m^e in Montgomery form (base_m and the result Montgomery representatives in [-N, N])
__device__ int64_t mont_pow(const CsdMod& m, int64_t base_m, uint64_t e) {
    int64_t r = csd_to_mont64(m, 1);
    int64_t b = base_m;
    while (e) {
        if (e & 1) r = csd_mont_mul(m, r, b);
        b = csd_mont_mul(m, b, b);
        e >>= 1;
    }
    return r;
}

// out[t] = mont(sum_s zeta^E[t][s]) mod p, canonical: the Montgomery form of the trace
__global__ void k_crt_traces(const uint32_t* __restrict__ E, int64_t T, int64_t S, CsdMod m, int64_t zeta_m,
                             int64_t* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int64_t acc = 0;
    for (int64_t s = 0; s < S; s++) {
        acc += canon_mod(m, mont_pow(m, zeta_m, E[t * S + s]));
        if (acc >= m.N) acc -= m.N;
    }
    out[t] = acc;
}

// x = mont(from_mont(x)^(p - 2)) mod p, canonical, for Montgomery-form entries of any
// representative (the NTT-domain kernels); zero entries are rejected on the host first
__global__ void k_crt_invert(int64_t* __restrict__ x, int64_t n, CsdMod m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t v = x[i] % m.N;
    if (v < 0) v += m.N;
    x[i] = canon_mod(m, mont_pow(m, v, (uint64_t)m.N - 2));
}

csd_view make_view(void* p, int64_t s0, int64_t s1, int64_t s2, int64_t s3) {
    csd_view v;
    v.ptr = p;
    v.shape[0] = s0;
    v.shape[1] = s1;
    v.shape[2] = s2;
    v.shape[3] = s3;
    v.strides[3] = 1;
    v.strides[2] = s3;
    v.strides[1] = s2 * s3;
    v.strides[0] = s1 * s2 * s3;
    return v;
}

// a sub-view [lo, hi) along one dimension; `elem` is the element size in bytes
csd_view sub(const csd_view& v, int dim, int64_t lo, int64_t hi, int elem) {
    csd_view r = v;
    r.ptr = (char*)v.ptr + lo * v.strides[dim] * elem;
    r.shape[dim] = hi - lo;
    return r;
}

void ok(int rc) {
    if (rc != 0) throw CsdError(csd_last_error());
}

std::unique_ptr<DevBuf> upload(csd_ctx* ctx, const std::vector<int64_t>& v) {
    auto b = std::make_unique<DevBuf>(ctx, std::max<size_t>(8, v.size() * 8));
    if (!v.empty()) {
        HIPCHECK(hipMemcpyAsync(b->p, v.data(), v.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    return b;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// the ring context

struct Kernels {  // RnsFactorKernels (ring.rs:559-568)
    std::vector<int64_t> lp, k, inv_lp, inv_k;
};

struct TransformCache {  // ring.rs:578-592
    std::vector<csd_modulus> mods;
    std::vector<csd_ntt_plan*> plans_lp, plans_k;
    std::unique_ptr<DevBuf> k_lp, k_k, ik_lp, ik_k;
    ~TransformCache() {
        for (auto p : plans_lp) csd_ntt_plan_free(p);
        for (auto p : plans_k) csd_ntt_plan_free(p);
    }
};

struct BConvCache {  // ring.rs:1596-1604
    size_t log2_qtilde, acc;
    std::unique_ptr<DevBuf> q_over_p_inv, matrix, q_mod_p;
    std::unique_ptr<DevBuf> src_mods;  // the source moduli as CsdMod, for the fused kernel
    std::unique_ptr<DevBuf> matrix_t;  // the matrix rows transposed and padded (csd_bconv_transpose)
    int64_t rdp = 0;
    std::vector<int64_t> host_qopi, host_matrix, host_qm;  // the same constants, kept for fused kernels
    // the matrix-core form (csd_bconv_mfma_tables): weight digit operands and target-row constants,
    // null when the bases or weights are outside what that kernel covers
    std::unique_ptr<DevBuf> mfma_atab, mfma_rows;
};

struct GadgetCache {  // ring.rs:1714-1727
    size_t log2_qtilde, acc;
    std::unique_ptr<DevBuf> d_over_p_inv, matrix, d_mod_p;
    std::unique_ptr<DevBuf> matrix_t;  // the matrix rows transposed and padded (csd_bconv_transpose)
    int64_t rdp = 0;
};


// Whether a and b are disjoint limb ranges of one tensor (or do not overlap at all), so that
// a kernel whose thread owns one (element, position) of both may read b and rewrite a in place:
// equal shapes and strides apart from dim 1, the dim-1 rows of both within one element's
// stride, and the two row ranges apart.
static bool views_overlap(const csd_view& a, const csd_view& b, int elem);
static bool limb_disjoint(const csd_view& a, const csd_view& b) {
    if (!views_overlap(a, b, 8)) return true;
    for (int q = 0; q < 4; q++) {
        if (a.strides[q] != b.strides[q] || a.strides[q] <= 0) return false;
        if (q != 1 && a.shape[q] != b.shape[q]) return false;
    }
    const int64_t s1 = a.strides[1];
    if (a.shape[2] > 1 && a.strides[2] < a.strides[3] * a.shape[3]) return false;          // positions alias
    if (a.strides[2] * (a.shape[2] - 1) + a.strides[3] * (a.shape[3] - 1) >= s1) return false;  // rows interleave
    const intptr_t diff = (intptr_t)b.ptr - (intptr_t)a.ptr;
    if (diff % (intptr_t)(8 * s1)) return false;
    const int64_t m = (int64_t)(diff / (intptr_t)(8 * s1));  // b's first row, counted in a's rows
    const int64_t lo = std::min<int64_t>(0, m), hi = std::max<int64_t>(a.shape[1], m + b.shape[1]);
    if ((hi - lo) * s1 > a.strides[0] && a.shape[0] > 1) return false;  // rows reach the next element
    return m >= a.shape[1] || m + b.shape[1] <= 0;
}

// whether the memory spans of two views (first to last element) intersect
static bool views_overlap(const csd_view& a, const csd_view& b, int elem) {
    auto span = [elem](const csd_view& v, uintptr_t& lo, uintptr_t& hi) {
        int64_t last = 0;
        for (int q = 0; q < 4; q++) {
            if (v.shape[q] == 0) {
                lo = hi = 0;
                return false;
            }
            if (v.strides[q] < 0) {  // not produced by this library's views; treat as overlapping
                lo = 0;
                hi = UINTPTR_MAX;
                return true;
            }
            last += (v.shape[q] - 1) * v.strides[q];
        }
        lo = (uintptr_t)v.ptr;
        hi = lo + (uintptr_t)(last + 1) * (uintptr_t)elem;
        return true;
    };
    uintptr_t alo, ahi, blo, bhi;
    if (!span(a, alo, ahi) || !span(b, blo, bhi)) return false;
    return alo < bhi && blo < ahi;
}

struct csd_ring {
    csd_ctx* ctx;
    hipStream_t aux_stream = nullptr;  // second stream of the half-batch overlap
    std::shared_ptr<Field> f;
    std::vector<int64_t> primes;
    std::vector<csd_modulus> mods;
    std::vector<bool> supports_transform;
    size_t initial_len, t_bits;
    std::map<int64_t, Kernels> kernels;
    std::map<std::vector<size_t>, std::unique_ptr<TransformCache>> tcache;
    std::map<std::pair<std::vector<size_t>, std::vector<size_t>>, std::unique_ptr<BConvCache>> bcache;
    std::map<std::pair<std::vector<size_t>, std::vector<size_t>>, std::unique_ptr<GadgetCache>> gcache;
    std::map<std::vector<int64_t>, std::pair<csd_perm_plan*, csd_perm_plan*>> pcache;
    std::map<std::pair<std::vector<size_t>, std::vector<size_t>>, std::unique_ptr<DevBuf>> inv_scalar_cache;
    // the ring's execution-form settings (every form computes identical residues): see
    // csd_ring_set_overlap, csd_ring_set_bconv_form and csd_ring_set_gadget_form
    std::atomic<bool> overlap_halves{false};
    std::atomic<int> bconv_form{2};  // csd_ring_set_bconv_form: 0 HAL sequence, 1 fused, 2 / 3 fused on the matrix cores
    std::atomic<bool> gadget_fused{true};

    ~csd_ring() {
        if (aux_stream) {
            (void)hipStreamSynchronize(aux_stream);
            (void)hipStreamDestroy(aux_stream);
        }
        for (auto& kv : pcache) {
            csd_perm_plan_free(kv.second.first);
            csd_perm_plan_free(kv.second.second);
        }
    }

    std::vector<csd_modulus> mods_for(const std::vector<size_t>& base) const {
        std::vector<csd_modulus> r;
        for (size_t i : base) r.push_back(mods.at(i));
        return r;
    }

    // products of the primes at `base`, reduced mod p
    uint64_t prod_mod(const std::vector<size_t>& base, int64_t p, size_t skip = SIZE_MAX) const {
        uint64_t r = 1 % (uint64_t)p;
        for (size_t t = 0; t < base.size(); t++)
            if (t != skip) r = mulmod_u(r, (uint64_t)primes[base[t]] % (uint64_t)p, (uint64_t)p);
        return r;
    }

    void check_base(const std::vector<size_t>& base) const {
        for (size_t i : base)
            if (i >= primes.size()) throw CsdError("RNS base index out of range");
    }

    // the exponents of the trace sums of one part of the field (K or L'), shared by all primes:
    // e[t][s] = (sub[s] g_t^-1 mod m) mod order for the transversal g_t and the Galois subgroup
    struct TraceTable {
        std::unique_ptr<DevBuf> e;
        size_t T = 0, S = 0;
    };
    TraceTable tt_lp, tt_k;

    const TraceTable& trace_table(bool k_part) {
        TraceTable& tt = k_part ? tt_k : tt_lp;
        if (tt.e) return tt;
        const Field& F = *f;
        const uint64_t m = F.m;
        const uint64_t order = k_part ? F.m_k() : F.m_prime;
        const std::vector<uint64_t>& sub = k_part ? F.gal_mk_k : F.gal_mp_lp;
        const std::vector<uint64_t> tr = k_part ? F.kq_transversal() : F.lp_transversal();
        if (order >= ((uint64_t)1 << 32)) throw CsdError("CRT kernel traces: cyclotomic order too large");
        std::vector<uint32_t> e(tr.size() * sub.size());
        for (size_t t = 0; t < tr.size(); t++) {
            const uint64_t gi = inv_mod_u(tr[t], m);
            for (size_t s = 0; s < sub.size(); s++) e[t * sub.size() + s] = (uint32_t)(mulmod_u(sub[s], gi, m) % order);
        }
        tt.T = tr.size();
        tt.S = sub.size();
        tt.e = std::make_unique<DevBuf>(ctx, std::max<size_t>(4, e.size() * 4));
        HIPCHECK(hipMemcpyAsync(tt.e->p, e.data(), e.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        return tt;
    }

    // compute_rns_factor_kernels (ring.rs:1337-1415): traces, forward NTTs and inverses on the
    // device
    const Kernels& factor_kernels(size_t idx, csd_ntt_plan* plan_lp, csd_ntt_plan* plan_k) {
        int64_t p = primes[idx];
        auto it = kernels.find(p);
        if (it != kernels.end()) return it->second;
        const Field& F = *f;
        uint64_t m = F.m;
        uint64_t zeta = 0;
        for (uint64_t c = 2;; c++) {
            zeta = powmod_u(c, ((uint64_t)p - 1) / m, (uint64_t)p);
            bool good = true;
            for (uint64_t fac : F.m_factors)
                if (powmod_u(zeta, m / fac, (uint64_t)p) == 1) good = false;
            if (good) break;
        }
        uint64_t z_mp = powmod_u(zeta, crt2(1, F.m_prime, 0, F.m_k()), (uint64_t)p);
        uint64_t z_mk = powmod_u(zeta, crt2(0, F.m_prime, 1, F.m_k()), (uint64_t)p);
        const CsdMod md = make_mod_table(&mods[idx], 1, 1).m[0];
        // traces (device), forward NTT (device), inverse (device); the host keeps the canonical
        // NTT-domain values, checks them for zeros and caches them per prime
        auto kernel_of = [&](uint64_t zp, const TraceTable& tt, csd_ntt_plan* plan, int dim, std::vector<int64_t>& fwd,
                             std::vector<int64_t>& inv) {
            const int64_t T = (int64_t)tt.T;
            DevBuf b(ctx, (size_t)T * 8);
            const int64_t zeta_m = mont_repr((int64_t)zp, p);
            hipLaunchKernelGGL(k_crt_traces, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)tt.e->p, T, (int64_t)tt.S, md, zeta_m, (int64_t*)b.p);
            HIPCHECK(hipGetLastError());
            csd_view v = dim == 2 ? make_view(b.p, 1, 1, T, 1) : make_view(b.p, 1, 1, 1, T);
            ok(csd_hal_ntt_mont64_dim(ctx, &v, &plan, 1, dim));
            fwd.resize((size_t)T);
            ok(csd_hal_export(ctx, &v, fwd.data(), fwd.size()));
            for (auto& x : fwd) {
                x = (int64_t)mod_of(x, p);
                if (x == 0) throw CsdError("CRT kernel is not invertible");  // from_mont(x) == 0
            }
            // from_mont_repr then invert then mont_repr (ring.rs:1395-1407)
            hipLaunchKernelGGL(k_crt_invert, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, ctx->stream, (int64_t*)b.p, T, md);
            HIPCHECK(hipGetLastError());
            inv.resize((size_t)T);
            ok(csd_hal_export(ctx, &v, inv.data(), inv.size()));
        };
        Kernels K;
        kernel_of(z_mp, trace_table(false), plan_lp, 2, K.lp, K.inv_lp);
        kernel_of(z_mk, trace_table(true), plan_k, 3, K.k, K.inv_k);
        return kernels[p] = K;
    }

    TransformCache& transform_cache(const std::vector<size_t>& base) {  // ring.rs:1252-1303
        auto it = tcache.find(base);
        if (it != tcache.end()) return *it->second;
        check_base(base);
        auto c = std::make_unique<TransformCache>();
        const Field& F = *f;
        auto orders = F.orders();
        std::vector<size_t> o_lp(orders.begin(), orders.begin() + F.lp_count), o_k(orders.begin() + F.lp_count, orders.end());
        std::vector<int64_t> klp, kk, iklp, ikk;
        for (size_t idx : base) {
            if (!supports_transform[idx]) throw CsdError("RNS factor doesn't support the basis transforms");
            csd_ntt_plan *plp = nullptr, *pk = nullptr;
            ok(csd_hal_create_ntt_plan(ctx, &mods[idx], o_lp.data(), o_lp.size(), &plp));
            c->plans_lp.push_back(plp);
            ok(csd_hal_create_ntt_plan(ctx, &mods[idx], o_k.data(), o_k.size(), &pk));
            c->plans_k.push_back(pk);
            const Kernels& K = factor_kernels(idx, plp, pk);
            klp.insert(klp.end(), K.lp.begin(), K.lp.end());
            kk.insert(kk.end(), K.k.begin(), K.k.end());
            iklp.insert(iklp.end(), K.inv_lp.begin(), K.inv_lp.end());
            ikk.insert(ikk.end(), K.inv_k.begin(), K.inv_k.end());
        }
        c->mods = mods_for(base);
        c->k_lp = upload(ctx, klp);
        c->k_k = upload(ctx, kk);
        c->ik_lp = upload(ctx, iklp);
        c->ik_k = upload(ctx, ikk);
        auto& ref = *c;
        tcache[base] = std::move(c);
        return ref;
    }

    // pb_to_crt_tensor / crt_to_pb_tensor (ring.rs:1117-1161)
    void transform(const csd_view& data, const std::vector<size_t>& base, bool inverse_kernels) {
        TransformCache& c = transform_cache(base);
        int64_t r = (int64_t)base.size();
        if (data.shape[1] != r || data.shape[2] != (int64_t)f->deg_lp || data.shape[3] != (int64_t)f->deg_k)
            throw CsdError("transform: tensor shape doesn't match [count, r, deg_L', deg_K]");
        csd_view klp = make_view(inverse_kernels ? c.ik_lp->p : c.k_lp->p, 1, r, (int64_t)f->deg_lp, 1);
        csd_view kk = make_view(inverse_kernels ? c.ik_k->p : c.k_k->p, 1, r, 1, (int64_t)f->deg_k);
        auto step = [&](const csd_view& d, int q) {
            if (q == 0) ok(csd_hal_ntt_mul_mont64_dim(ctx, &d, c.plans_lp.data(), r, CSD_DIM2, &klp, c.mods.data(), r));
            if (q == 1) ok(csd_hal_ntt_mul_mont64_dim(ctx, &d, c.plans_k.data(), r, CSD_DIM3, &kk, c.mods.data(), r));
            if (q == 2) ok(csd_hal_intt_mont64_dim(ctx, &d, c.plans_lp.data(), r, CSD_DIM2));
            if (q == 3) ok(csd_hal_intt_mont64_dim(ctx, &d, c.plans_k.data(), r, CSD_DIM3));
        };
        const int64_t cnt = data.shape[0];
        const int64_t slice_bytes = r * (int64_t)f->deg_lp * (int64_t)f->deg_k * 8;
        if (!overlap_halves || cnt < 2 || cnt * slice_bytes < ((int64_t)64 << 20)) {
            for (int q = 0; q < 4; q++) step(data, q);
            return;
        }
        // Two halves of the batch on two streams, the second one step behind the first: the
        // memory-bound dim-2 passes of one half run beside the VALU-bound dim-3 passes of the
        // other. Each element still sees the same four passes in the same order.
        if (!aux_stream) HIPCHECK(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
        const csd_view a = sub(data, 0, 0, cnt / 2, 8), b = sub(data, 0, cnt / 2, cnt, 8);
        hipStream_t main_stream = ctx->stream;
        hipEvent_t ev[2];
        HIPCHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        try {
            step(a, 0);
            HIPCHECK(hipEventRecord(ev[0], main_stream));  // b starts once a's first pass is issued
            HIPCHECK(hipStreamWaitEvent(aux_stream, ev[0], 0));
            for (int q = 0; q < 4; q++) {
                ctx->stream = aux_stream;
                step(b, q);
                ctx->stream = main_stream;
                if (q + 1 < 4) step(a, q + 1);
            }
            HIPCHECK(hipEventRecord(ev[1], aux_stream));
            HIPCHECK(hipStreamWaitEvent(main_stream, ev[1], 0));
        } catch (...) {
            ctx->stream = main_stream;
            (void)hipStreamSynchronize(aux_stream);
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
            throw;
        }
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }

    BConvCache& bconv_cache(const std::vector<size_t>& src, const std::vector<size_t>& dst) {  // ring.rs:1606-1661
        auto key = std::make_pair(src, dst);
        auto it = bcache.find(key);
        if (it != bcache.end()) return *it->second;
        check_base(src);
        check_base(dst);
        size_t r = src.size(), rd = dst.size();
        if (r == 0 || rd == 0) throw CsdError("bconv: empty base");
        int64_t max_p = 0, min_p = INT64_MAX;
        for (size_t i : src) {
            max_p = std::max(max_p, primes[i]);
            min_p = std::min(min_p, primes[i]);
        }
        for (size_t i : dst) max_p = std::max(max_p, primes[i]);
        u128 x = (u128)4 * r * (u128)max_p;  // abs_log2_ceil(4 r max_p + 1) = bit_length(4 r max_p)
        size_t bits = 0;
        while (x) {
            bits++;
            x >>= 1;
        }
        auto c = std::make_unique<BConvCache>();
        c->log2_qtilde = std::max<size_t>(64, bits);
        if (c->log2_qtilde >= 126) throw CsdError("bconv: correction precision out of range");
        u128 qtilde = (u128)1 << c->log2_qtilde;
        u128 corr_max = (qtilde - 1) / (u128)min_p + 1;
        if (corr_max > (u128)INT64_MAX) throw CsdError("coefficients of correction vector don't fit into an i64");
        c->acc = (size_t)((((u128)1 << 127) - 1) / ((u128)max_p * (u128)max_p));
        if (c->acc < 16) throw CsdError("modular reductions would be needed too often");
        // q_over_p_inv_tensor (ring.rs:1029-1043): lift((q/p)^-1 mod p)
        std::vector<int64_t> qopi;
        for (size_t t = 0; t < r; t++) {
            int64_t p = primes[src[t]];
            qopi.push_back(smallest_lift(inv_mod_u(prod_mod(src, p, t), (uint64_t)p), p));
        }
        // rows: lift((q/p_j) * 2^128 mod p'_i); correction row round(q~/p_j)
        std::vector<int64_t> mat;
        for (size_t i = 0; i < rd; i++) {
            int64_t pp = primes[dst[i]];
            uint64_t r_sqr = powmod_u(((u128)1 << 64) % (u128)pp, 2, (uint64_t)pp);
            for (size_t j = 0; j < r; j++) mat.push_back(smallest_lift(mulmod_u(prod_mod(src, pp, j), r_sqr, (uint64_t)pp), pp));
        }
        for (size_t j = 0; j < r; j++) {
            u128 p = (u128)primes[src[j]];
            mat.push_back((int64_t)((2 * qtilde + p) / (2 * p)));
        }
        // q_mod_p' (ring.rs:1651-1658): lift(q * (-2^64) mod p')
        std::vector<int64_t> qm;
        for (size_t i = 0; i < rd; i++) {
            int64_t pp = primes[dst[i]];
            uint64_t neg_r = (uint64_t)pp - (uint64_t)(((u128)1 << 64) % (u128)pp);
            qm.push_back(smallest_lift(mulmod_u(prod_mod(src, pp), neg_r % (uint64_t)pp, (uint64_t)pp), pp));
        }
        c->q_over_p_inv = upload(ctx, qopi);
        c->matrix = upload(ctx, mat);
        c->q_mod_p = upload(ctx, qm);
        {
            auto sm = mods_for(src);
            ModTable t = make_mod_table(sm.data(), sm.size(), (int64_t)sm.size());
            std::vector<int64_t> flat;
            for (size_t u = 0; u < r; u++) flat.insert(flat.end(), {t.m[u].N, t.m[u].ninv, t.m[u].r2, t.m[u].r3});
            c->src_mods = upload(ctx, flat);
        }
        c->matrix_t = upload(ctx, csd_bconv_transpose(mat, 1, (int64_t)rd, (int64_t)r, &c->rdp));
        {
            auto dm = mods_for(dst);
            ModTable t = make_mod_table(dm.data(), dm.size(), (int64_t)dm.size());
            std::vector<CsdMod> dmods(t.m, t.m + rd);
            std::vector<int8_t> atab;
            std::vector<int64_t> rows;
            if (csd_bconv_mfma_tables(mat, (int64_t)rd, (int64_t)r, dmods, qm, atab, rows)) {
                c->mfma_atab.reset(new DevBuf(ctx, atab.size()));
                HIPCHECK(hipMemcpy(c->mfma_atab->p, atab.data(), atab.size(), hipMemcpyHostToDevice));
                c->mfma_rows = upload(ctx, rows);
            }
        }
        c->host_qopi = qopi;
        c->host_matrix = mat;
        c->host_qm = qm;
        auto& ref = *c;
        bcache[key] = std::move(c);
        return ref;
    }

    // Whether form 2 takes a conversion from r source limbs to the matrix cores. One MFMA covers a
    // 32-byte K slice, source limb u in byte u: with few limbs most of it is padding, while the
    // lift, carry chain and reduction tails stay. Measured at field_36864 (bootstrap shapes, MI355X),
    // the MFMA kernel against the fused VALU kernel: r <= 4 takes 1.9 x as long, r = 5-8 1.2-1.4 x,
    // r = 9-12 0.96-1.1 x and r >= 13 0.75-0.97 x; in the rescale form, which adds two products per
    // target row to both, r = 9-16 takes 1.01-1.09 x and r >= 17 1.0 x. Below these thresholds form 2
    // runs the VALU kernel, whose residues are the same; form 3 takes every conversion whose tables
    // fit to the matrix cores.
    static bool mfma_pays(int64_t r, bool rescale) { return rescale ? r >= 17 : r >= 9; }

    // BConvCache::bconv (ring.rs:1663-1701). The fused form (default, when src and dst do not
    // overlap or are disjoint limb ranges of one tensor) is one kernel and leaves src as it is; the HAL sequence destroys src, like the
    // reference, unless preserve_src is set, when it runs on a copy. The form is decided once per
    // call, so what happens to src never depends on a later change of the setting.
    void bconv(const csd_view& dst, const std::vector<size_t>& dbase, const csd_view& src_in, const std::vector<size_t>& sbase,
               bool preserve_src = false) {
        BConvCache& c = bconv_cache(sbase, dbase);
        int64_t r = (int64_t)sbase.size(), rd = (int64_t)dbase.size();
        int64_t cnt = src_in.shape[0], n2 = src_in.shape[2], n3 = src_in.shape[3];
        if (src_in.shape[1] != r) throw CsdError("bconv: source doesn't match its base");
        if (dst.shape[0] != cnt || dst.shape[1] != rd || dst.shape[2] != n2 || dst.shape[3] != n3)
            throw CsdError("bconv: destination shape");
        auto smods = mods_for(sbase), dmods = mods_for(dbase);
        const int form = bconv_form;
        // the fused kernels read a position's source limbs and write its target limbs in one
        // thread (or wave), so src and dst may be disjoint limb ranges of one tensor
        const bool disjoint = limb_disjoint(dst, src_in);
        if ((form == 3 || (form == 2 && mfma_pays(r, false))) && c.mfma_atab && c.acc >= (size_t)r && disjoint) {
            csd_bconv_mfma_launch(ctx, dst, src_in, (const int64_t*)c.q_over_p_inv->p, (const int8_t*)c.mfma_atab->p,
                                  (const int64_t*)c.mfma_rows->p, c.log2_qtilde, (const CsdMod*)c.src_mods->p, "bconv_mfma");
            return;
        }
        if (form >= 1 && c.acc >= (size_t)r && disjoint) {
            const std::vector<size_t> whole = {0, (size_t)r};
            csd_bconv_fused_launch(ctx, dst, src_in, 1, (const int64_t*)c.q_over_p_inv->p, (const int64_t*)c.matrix->p,
                                   (const int64_t*)c.matrix_t->p, c.rdp, (const int64_t*)c.q_mod_p->p, whole,
                                   c.log2_qtilde, dmods.data(), (const CsdMod*)c.src_mods->p, "bconv");
            return;
        }
        std::unique_ptr<DevBuf> copy;
        csd_view src = src_in;
        if (preserve_src) {
            copy.reset(new DevBuf(ctx, (size_t)cnt * r * n2 * n3 * 8));
            src = make_view(copy->p, cnt, r, n2, n3);
            ok(csd_hal_assign(ctx, &src, &src_in));
        }
        csd_view qopi = make_view(c.q_over_p_inv->p, 1, r, 1, 1);
        ok(csd_hal_mul_mont64(ctx, &src, &qopi, smods.data(), r));
        DevBuf corr(ctx, (size_t)cnt * n2 * n3 * 16), tmp(ctx, (size_t)cnt * (rd + 1) * n2 * n3 * 16);
        csd_view vcorr = make_view(corr.p, cnt, 1, n2, n3);
        csd_view vtmp = make_view(tmp.p, cnt, rd + 1, n2, n3);
        csd_view mat = make_view(c.matrix->p, 1, rd + 1, r, 1);
        int chunk = 0;
        for (int64_t j = 0; j < r; j += (int64_t)c.acc, chunk++) {
            int64_t e = std::min<int64_t>(j + (int64_t)c.acc, r);
            csd_view ms = sub(mat, 2, j, e, 8), ss = sub(src, 1, j, e, 8);
            ok(csd_hal_matmul_i64(ctx, &vtmp, &ms, &ss));
            csd_view tv = sub(vtmp, 1, 0, rd, 16), tc = sub(vtmp, 1, rd, rd + 1, 16);
            if (chunk == 0) {
                ok(csd_hal_red_mont64_double(ctx, &dst, &tv, dmods.data(), rd));
                ok(csd_hal_assign_i128(ctx, &vcorr, &tc));
            } else {
                ok(csd_hal_add_red_mont64_double(ctx, &dst, &tv, dmods.data(), rd));
                ok(csd_hal_add_i128(ctx, &vcorr, &tc));
            }
        }
        csd_view qm = make_view(c.q_mod_p->p, 1, rd, 1, 1);
        ok(csd_hal_bconv_correct_mont64(ctx, &dst, &vcorr, &qm, dmods.data(), rd, c.log2_qtilde));
        ok(csd_hal_i64_to_mont64(ctx, &dst, dmods.data(), rd));
    }

    // The middle of rescale_to_base (bfv.rs:601-622): with s the scalar t and v the inverse of
    // the source base's product (one Montgomery value per limb),
    //   ext <- mont(mont(ext, s_ext) - bconv(mont(src, s_src)), v)
    // that is mul_mont64 by s, bconv, sub_i64 and mul_mont64 by v. The fused form is one kernel that
    // leaves src as it is; the HAL sequence (when the fused bconv does not apply) destroys src.
    void rescale_step(const csd_view& ext, const std::vector<size_t>& ebase, const csd_view& src,
                      const std::vector<size_t>& sbase, const csd_view& s_src, const csd_view& s_ext, const csd_view& inv) {
        BConvCache& c = bconv_cache(sbase, ebase);
        int64_t r = (int64_t)sbase.size(), re = (int64_t)ebase.size();
        int64_t cnt = src.shape[0], n2 = src.shape[2], n3 = src.shape[3];
        if (src.shape[1] != r) throw CsdError("rescale: source doesn't match its base");
        if (ext.shape[0] != cnt || ext.shape[1] != re || ext.shape[2] != n2 || ext.shape[3] != n3)
            throw CsdError("rescale: extension shape");
        auto scalar_ok = [](const csd_view& v, int64_t n) {
            return v.shape[0] == 1 && v.shape[1] == n && v.shape[2] == 1 && v.shape[3] == 1 && (n < 2 || v.strides[1] == 1);
        };
        if (!scalar_ok(s_src, r) || !scalar_ok(s_ext, re) || !scalar_ok(inv, re)) throw CsdError("rescale: scalar tensor shape");
        auto smods = mods_for(sbase), emods = mods_for(ebase);
        const int form = bconv_form;
        if ((form == 3 || (form == 2 && mfma_pays(r, true))) && c.mfma_atab && c.acc >= (size_t)r && limb_disjoint(ext, src)) {
            csd_bconv_mfma_launch(ctx, ext, src, (const int64_t*)c.q_over_p_inv->p, (const int8_t*)c.mfma_atab->p,
                                  (const int64_t*)c.mfma_rows->p, c.log2_qtilde, (const CsdMod*)c.src_mods->p, "rescale_step_mfma",
                                  (const int64_t*)s_src.ptr, (const int64_t*)s_ext.ptr, (const int64_t*)inv.ptr);
            return;
        }
        if (form >= 1 && c.acc >= (size_t)r && limb_disjoint(ext, src)) {
            const std::vector<size_t> whole = {0, (size_t)r};
            csd_bconv_fused_launch(ctx, ext, src, 1, (const int64_t*)c.q_over_p_inv->p, (const int64_t*)c.matrix->p,
                                   (const int64_t*)c.matrix_t->p, c.rdp, (const int64_t*)c.q_mod_p->p, whole,
                                   c.log2_qtilde, emods.data(), (const CsdMod*)c.src_mods->p, "rescale_step",
                                   (const int64_t*)s_src.ptr, (const int64_t*)s_ext.ptr, (const int64_t*)inv.ptr);
            return;
        }
        ok(csd_hal_mul_mont64(ctx, &src, &s_src, smods.data(), r));
        ok(csd_hal_mul_mont64(ctx, &ext, &s_ext, emods.data(), re));
        DevBuf tmp(ctx, (size_t)cnt * re * n2 * n3 * 8);
        csd_view vt = make_view(tmp.p, cnt, re, n2, n3);
        bconv(vt, ebase, src, sbase);
        ok(csd_hal_sub_i64(ctx, &ext, &vt));
        ok(csd_hal_mul_mont64(ctx, &ext, &inv, emods.data(), re));
    }

    GadgetCache& gadget_cache(const std::vector<size_t>& base, const std::vector<size_t>& bounds) {  // ring.rs:1730-1807
        auto key = std::make_pair(base, bounds);
        auto it = gcache.find(key);
        if (it != gcache.end()) return *it->second;
        check_base(base);
        size_t r = base.size(), d = bounds.size() / 2;
        if (d == 0) throw CsdError("gadget bconv: no digits");
        for (size_t i = 0; i < d; i++) {
            if (!(bounds[2 * i] < bounds[2 * i + 1] && bounds[2 * i + 1] <= r)) throw CsdError("gadget bconv: invalid digit");
            if (i > 0 && bounds[2 * i - 1] > bounds[2 * i]) throw CsdError("digits must be sorted and disjoint");
        }
        int64_t max_p = 0, min_p = INT64_MAX;
        size_t max_len = 0;
        for (size_t i : base) {
            max_p = std::max(max_p, primes[i]);
            min_p = std::min(min_p, primes[i]);
        }
        for (size_t i = 0; i < d; i++) max_len = std::max(max_len, bounds[2 * i + 1] - bounds[2 * i]);
        u128 x = (u128)4 * max_len * (u128)max_p;
        size_t bits = 0;
        while (x) {
            bits++;
            x >>= 1;
        }
        auto c = std::make_unique<GadgetCache>();
        c->log2_qtilde = std::max<size_t>(64, bits);
        u128 qtilde = (u128)1 << c->log2_qtilde;
        // the overflow guards of ring.rs:1741-1759
        u128 coeff = (qtilde - 1) / (u128)min_p + 1;
        if (coeff > (u128)INT64_MAX) throw CsdError("coefficients of correction vector don't fit into an i64");
        if ((u128)max_len * (u128)max_p > (((u128)1 << 127) - 1) / coeff) throw CsdError("correction vector inner product overflows");
        c->acc = (size_t)((((u128)1 << 127) - 1) / ((u128)max_p * (u128)max_p));  // truncated like `as usize`
        if (c->acc < 16) throw CsdError("to avoid overflow, modular reductions have to be performed too often to be efficient");
        if ((u128)((r + c->acc - 1) / c->acc) * (u128)max_p > (u128)INT64_MAX) throw CsdError("overflow during accumulation");
        std::vector<int64_t> dopi(r, 0), mat(d * (r + 1) * r, 0), dmp;
        for (size_t i = 0; i < d; i++) {
            std::vector<size_t> dig(base.begin() + bounds[2 * i], base.begin() + bounds[2 * i + 1]);
            for (size_t u = bounds[2 * i]; u < bounds[2 * i + 1]; u++) {
                int64_t pu = primes[base[u]];
                dopi[u] = smallest_lift(inv_mod_u(prod_mod(dig, pu, u - bounds[2 * i]), (uint64_t)pu), pu);
            }
            for (size_t j = 0; j < r; j++) {
                int64_t pj = primes[base[j]];
                uint64_t r_sqr = powmod_u(((u128)1 << 64) % (u128)pj, 2, (uint64_t)pj);
                for (size_t u = bounds[2 * i]; u < bounds[2 * i + 1]; u++)
                    mat[(i * (r + 1) + j) * r + u] =
                        smallest_lift(mulmod_u(prod_mod(dig, pj, u - bounds[2 * i]), r_sqr, (uint64_t)pj), pj);
                uint64_t neg_r = (uint64_t)pj - (uint64_t)(((u128)1 << 64) % (u128)pj);
                dmp.push_back(smallest_lift(mulmod_u(prod_mod(dig, pj), neg_r % (uint64_t)pj, (uint64_t)pj), pj));
            }
            for (size_t u = bounds[2 * i]; u < bounds[2 * i + 1]; u++) {
                u128 pu = (u128)primes[base[u]];
                mat[(i * (r + 1) + r) * r + u] = (int64_t)((2 * qtilde + pu) / (2 * pu));
            }
        }
        c->d_over_p_inv = upload(ctx, dopi);
        c->matrix = upload(ctx, mat);
        c->d_mod_p = upload(ctx, dmp);
        c->matrix_t = upload(ctx, csd_bconv_transpose(mat, (int64_t)d, (int64_t)r, (int64_t)r, &c->rdp));
        auto& ref = *c;
        gcache[key] = std::move(c);
        return ref;
    }

    // GadgetBConvCache::bconv (ring.rs:1809-1851) for B source elements at once: src [B, r, ., .],
    // dst [B d, r, ., .] with element b, digit i at row b d + i (the reference takes B = 1). The
    // fused form (default) is one kernel and leaves src as it is; the HAL sequence destroys src.
    void bconv_gadget(const csd_view& dst, const csd_view& src, const std::vector<size_t>& base, const std::vector<size_t>& bounds) {
        GadgetCache& c = gadget_cache(base, bounds);
        int64_t r = (int64_t)base.size(), d = (int64_t)bounds.size() / 2;
        int64_t B = src.shape[0], n2 = src.shape[2], n3 = src.shape[3];
        if (B < 1 || src.shape[1] != r) throw CsdError("gadget bconv: source shape");
        if (dst.shape[0] != B * d || dst.shape[1] != r || dst.shape[2] != n2 || dst.shape[3] != n3)
            throw CsdError("gadget bconv: destination shape");
        auto m = mods_for(base);
        if (gadget_fused && c.acc >= (size_t)r && !views_overlap(dst, src, 8)) {
            csd_bconv_fused_launch(ctx, dst, src, d, (const int64_t*)c.d_over_p_inv->p, (const int64_t*)c.matrix->p,
                                   (const int64_t*)c.matrix_t->p, c.rdp, (const int64_t*)c.d_mod_p->p, bounds,
                                   c.log2_qtilde, m.data(), nullptr, "gadget_bconv");
            return;
        }
        csd_view dopi = make_view(c.d_over_p_inv->p, 1, r, 1, 1);
        DevBuf corr(ctx, (size_t)d * n2 * n3 * 16), tmp(ctx, (size_t)d * (r + 1) * n2 * n3 * 16);
        csd_view vcorr = make_view(corr.p, d, 1, n2, n3), vtmp = make_view(tmp.p, d, r + 1, n2, n3);
        csd_view mat = make_view(c.matrix->p, d, r + 1, r, 1);
        csd_view dm = make_view(c.d_mod_p->p, d, r, 1, 1);
        for (int64_t b = 0; b < B; b++) {
            csd_view sb = sub(src, 0, b, b + 1, 8), db = sub(dst, 0, b * d, (b + 1) * d, 8);
            ok(csd_hal_mul_mont64(ctx, &sb, &dopi, m.data(), r));
            int chunk = 0;
            for (int64_t u = 0; u < r; u += (int64_t)c.acc, chunk++) {
                int64_t e = std::min<int64_t>(u + (int64_t)c.acc, r);
                csd_view ms = sub(mat, 2, u, e, 8), ss = sub(sb, 1, u, e, 8);
                ok(csd_hal_matmul_i64(ctx, &vtmp, &ms, &ss));
                csd_view tv = sub(vtmp, 1, 0, r, 16), tc = sub(vtmp, 1, r, r + 1, 16);
                if (chunk == 0) {
                    ok(csd_hal_red_mont64_double(ctx, &db, &tv, m.data(), r));
                    ok(csd_hal_assign_i128(ctx, &vcorr, &tc));
                } else {
                    ok(csd_hal_add_red_mont64_double(ctx, &db, &tv, m.data(), r));
                    ok(csd_hal_add_i128(ctx, &vcorr, &tc));
                }
            }
            ok(csd_hal_bconv_correct_mont64(ctx, &db, &vcorr, &dm, m.data(), r, c.log2_qtilde));
            ok(csd_hal_i64_to_mont64(ctx, &db, m.data(), r));
        }
    }

    // inverse_scalar_tensor (ring.rs:1165-1177) for the scalar prod(primes at `of`)
    std::vector<int64_t> inverse_product_values(const std::vector<size_t>& base, const std::vector<size_t>& of) const {
        std::vector<int64_t> v;
        for (size_t i : base) {
            int64_t p = primes[i];
            uint64_t s = prod_mod(of, p);
            if (s == 0) throw CsdError("scalar is not invertible");
            v.push_back(mont_repr(smallest_lift(inv_mod_u(s, (uint64_t)p), p), p));
        }
        return v;
    }

    DevBuf& inverse_product_tensor(const std::vector<size_t>& base, const std::vector<size_t>& of) {
        auto key = std::make_pair(base, of);
        auto it = inv_scalar_cache.find(key);
        if (it != inv_scalar_cache.end()) return *it->second;
        auto b = upload(ctx, inverse_product_values(base, of));
        auto& ref = *b;
        inv_scalar_cache[key] = std::move(b);
        return ref;
    }

    // mod_switch_down (bfv.rs:628-662)
    void mod_switch_down(const csd_view& dst, const csd_view& src, const std::vector<size_t>& base, size_t new_len) {
        int64_t cnt = src.shape[0], n2 = src.shape[2], n3 = src.shape[3];
        int64_t r = (int64_t)base.size(), nl = (int64_t)new_len;
        if (src.shape[1] != r || nl > r || nl < 1) throw CsdError("mod_switch_down: invalid base length");
        if (dst.shape[0] != cnt || dst.shape[1] != nl || dst.shape[2] != n2 || dst.shape[3] != n3)
            throw CsdError("mod_switch_down: destination shape");
        std::vector<size_t> kept(base.begin(), base.begin() + nl), dropped(base.begin() + nl, base.end());
        csd_view src_kept = sub(src, 1, 0, nl, 8);
        if (dropped.empty()) {
            ok(csd_hal_assign(ctx, &dst, &src_kept));
            return;
        }
        int64_t rdr = (int64_t)dropped.size();
        csd_view src_dropped0 = sub(src, 1, nl, r, 8);
        {
            // the fused forms: the conversion of the dropped limbs, the subtraction from the kept
            // ones and the multiply by the inverse as one kernel, the same integer steps as below
            BConvCache& c = bconv_cache(dropped, kept);
            if (bconv_form >= 1 && c.acc >= (size_t)rdr && !views_overlap(dst, src, 8)) {
                DevBuf& inv = inverse_product_tensor(kept, dropped);
                auto km = mods_for(kept);
                const std::vector<size_t> whole = {0, (size_t)rdr};
                csd_bconv_fused_launch(ctx, dst, src_dropped0, 1, (const int64_t*)c.q_over_p_inv->p, (const int64_t*)c.matrix->p,
                                       (const int64_t*)c.matrix_t->p, c.rdp, (const int64_t*)c.q_mod_p->p, whole, c.log2_qtilde,
                                       km.data(), (const CsdMod*)c.src_mods->p, "mod_switch_down", nullptr, nullptr,
                                       (const int64_t*)inv.p, &src_kept);
                return;
            }
        }
        ok(csd_hal_assign(ctx, &dst, &src_kept));
        DevBuf tmp(ctx, (size_t)cnt * nl * n2 * n3 * 8);
        csd_view vt = make_view(tmp.p, cnt, nl, n2, n3);
        csd_view src_dropped = sub(src, 1, nl, r, 8);
        bconv(vt, kept, src_dropped, dropped, true);  // src is left unchanged
        ok(csd_hal_sub_i64(ctx, &dst, &vt));
        DevBuf& inv = inverse_product_tensor(kept, dropped);
        csd_view vinv = make_view(inv.p, 1, nl, 1, 1);
        auto m = mods_for(kept);
        ok(csd_hal_mul_mont64(ctx, &dst, &vinv, m.data(), nl));
    }

    // the (dim 2, dim 3) permutations of the automorphism g (ring.rs:1187-1210)
    std::pair<csd_perm_plan*, csd_perm_plan*> galois_perms(uint64_t g) {
        auto shift = f->gal_to_shift(g);
        auto it = pcache.find(shift);
        if (it == pcache.end()) {
            auto o = f->orders();
            size_t nl = f->lp_count;
            std::vector<size_t> s_lp, s_k;
            for (size_t i = 0; i < shift.size(); i++) (i < nl ? s_lp : s_k).push_back((size_t)shift[i]);
            csd_perm_plan *plp = nullptr, *pk = nullptr;
            ok(csd_hal_create_permutation(ctx, o.data(), s_lp.data(), nl, &plp));
            ok(csd_hal_create_permutation(ctx, o.data() + nl, s_k.data(), o.size() - nl, &pk));
            it = pcache.emplace(shift, std::make_pair(plp, pk)).first;
        }
        return it->second;
    }

    // apply_galois (ring.rs:1183-1222)
    void apply_galois(const csd_view& dst, const csd_view& src, uint64_t g) {
        auto perms = galois_perms(g);
        DevBuf tmp(ctx, (size_t)numel(&src) * 8);
        csd_view vt = make_view(tmp.p, src.shape[0], src.shape[1], src.shape[2], src.shape[3]);
        ok(csd_hal_permute(ctx, &vt, &src, perms.first, CSD_DIM2));
        ok(csd_hal_permute(ctx, &dst, &vt, perms.second, CSD_DIM3));
    }

    // The key-switch part of bfv_hom_gal (bfv.rs:816-839) for a batch of ciphertexts whose
    // gadget digits were hoisted, over a set of rotations; see csd_fused.hip. Element e reads c0 and
    // writes outs[t] through the row addressing `rows` (GksRows); keys[t] may be a restriction of a
    // key over a longer base (rows 2i + s at any row stride, limbs as contiguous planes).
    void galois_keyswitch_rows(const std::vector<size_t>& base, int64_t batch, int64_t d, const csd_view& parts,
                               const csd_view& c0, const std::vector<uint64_t>& gs, const csd_view* keys, const csd_view* outs,
                               const GksRows& rows) {
        int64_t r = (int64_t)base.size(), n2 = (int64_t)f->deg_lp, n3 = (int64_t)f->deg_k;
        const int64_t plane = n2 * n3;
        auto contiguous = [](const csd_view& v, int64_t s0, int64_t s1, int64_t s2, int64_t s3) {
            csd_view e = make_view(v.ptr, s0, s1, s2, s3);
            for (int t = 0; t < 4; t++)
                if (v.shape[t] != e.shape[t] || (e.shape[t] > 1 && v.strides[t] != e.strides[t])) return false;
            return true;
        };
        if (batch < 1 || d < 1 || rows.grp < 1) throw CsdError("galois key switch: empty batch or digits");
        if (!contiguous(parts, batch * d, r, n2, n3)) throw CsdError("galois key switch: parts must be a contiguous [batch*d, r, n2, n3] tensor");
        // the largest row an element reaches, for the bounds of c0 and the outputs
        const int64_t hmax = (batch - 1) / rows.grp, lmax = std::min<int64_t>(batch, rows.grp) - 1;
        auto max_row = [&](int64_t hi, int64_t lo) {
            if (hi < 0 || lo < 0) throw CsdError("galois key switch: negative row strides");
            return hmax * hi + lmax * lo;
        };
        const int64_t c0_rows = max_row(rows.c0_hi, rows.c0_lo) + 1;
        if (!contiguous(c0, c0.shape[0], r, n2, n3) || c0.shape[0] < c0_rows)
            throw CsdError("galois key switch: c0 must be a contiguous [rows, r, n2, n3] tensor covering every element");
        if (rows.o_s < 0) throw CsdError("galois key switch: negative row strides");
        const int64_t o_rows = max_row(rows.o_hi, rows.o_lo) + rows.o_s + 1;
        auto m = mods_for(base);
        // WideRingEls bounds (ring.rs:354-365, 441-455): d terms of bound-1 operands, plus the
        // bound-1 c0 on row 0; a row needs the full reduction above 2^63 min_p (ring.rs:366)
        u128 max_p = 0, min_p = ~(u128)0;
        for (size_t i : base) {
            max_p = std::max<u128>(max_p, (u128)primes[i]);
            min_p = std::min<u128>(min_p, (u128)primes[i]);
        }
        u128 b1 = (u128)d * max_p * max_p, b0 = b1 + max_p * max_p;
        int full0 = b0 > (min_p << 63), full1 = b1 > (min_p << 63);
        for (size_t gi = 0; gi < gs.size(); gi++) {
            const csd_view& k = keys[gi];
            if (k.shape[0] < 2 * d || k.shape[1] != r || k.shape[2] != n2 || k.shape[3] != n3 || k.strides[3] != 1 ||
                (n2 > 1 && k.strides[2] != n3) || (r > 1 && k.strides[1] != plane) || k.strides[0] < r * plane)
                throw CsdError("galois key switch: keys must be [>= 2*d, r, n2, n3] with contiguous limb planes");
            if (!contiguous(outs[gi], outs[gi].shape[0], r, n2, n3) || outs[gi].shape[0] < o_rows)
                throw CsdError("galois key switch: outputs must be contiguous [rows, r, n2, n3] tensors covering every element");
            auto perms = galois_perms(gs[gi]);
            csd_galois_keyswitch_launch(ctx, batch, d, r, n2, n3, (const int64_t*)parts.ptr, (const int64_t*)c0.ptr,
                                        (const int64_t*)k.ptr, k.strides[0], perms.first->dperm, perms.second->dperm,
                                        (int64_t*)outs[gi].ptr, rows, m.data(), full0, full1);
        }
    }

    void galois_keyswitch(const std::vector<size_t>& base, int64_t batch, int64_t d, const csd_view& parts, const csd_view& c0,
                          const std::vector<uint64_t>& gs, const csd_view* keys, const csd_view& out) {
        int64_t r = (int64_t)base.size(), n2 = (int64_t)f->deg_lp, n3 = (int64_t)f->deg_k;
        csd_view e = make_view(out.ptr, (int64_t)gs.size() * batch * 2, r, n2, n3);
        for (int t = 0; t < 4; t++)
            if (out.shape[t] != e.shape[t] || (e.shape[t] > 1 && out.strides[t] != e.strides[t]))
                throw CsdError("galois key switch: out must be a contiguous [ngal*batch*2, r, n2, n3] tensor");
        if (c0.shape[0] != batch) throw CsdError("galois key switch: c0 must be a contiguous [batch, r, n2, n3] tensor");
        for (size_t gi = 0; gi < gs.size(); gi++) {
            const csd_view& k = keys[gi];
            csd_view ek = make_view(k.ptr, 2 * d, r, n2, n3);
            for (int t = 0; t < 4; t++)
                if (k.shape[t] != ek.shape[t] || (ek.shape[t] > 1 && k.strides[t] != ek.strides[t]))
                    throw CsdError("galois key switch: keys must be contiguous [2*d, r, n2, n3]");
        }
        std::vector<csd_view> outs;
        for (size_t gi = 0; gi < gs.size(); gi++)
            outs.push_back(sub(out, 0, (int64_t)gi * batch * 2, (int64_t)(gi + 1) * batch * 2, 8));
        GksRows rows{batch, 0, 1, 0, 2, 1};
        galois_keyswitch_rows(base, batch, d, parts, c0, gs, keys, outs.data(), rows);
    }

    // The same key switch with the keys in pinned host memory (the role of the CPU build's
    // KeyStore, keystore.rs:50-127): two device staging buffers; the copy of key t + 1 on the
    // copy stream overlaps the kernel of key t on the compute stream, ordered by events.
    void galois_keyswitch_streamed(const std::vector<size_t>& base, int64_t batch, int64_t d, const csd_view& parts,
                                   const csd_view& c0, const std::vector<uint64_t>& gs, const int64_t* const* host_keys,
                                   const csd_view& out) {
        const int64_t r = (int64_t)base.size(), n2 = (int64_t)f->deg_lp, n3 = (int64_t)f->deg_k;
        const size_t key_bytes = (size_t)(2 * d * r * n2 * n3) * 8;
        for (size_t gi = 0; gi < gs.size(); gi++) {
            if (!host_keys[gi]) throw CsdError("galois key switch: null host key");
            std::lock_guard<std::mutex> g(ctx->mu);
            bool found = false;
            for (auto& kv : ctx->pinned) {
                const char* b = (const char*)kv.first;
                if ((const char*)host_keys[gi] >= b && (const char*)host_keys[gi] + key_bytes <= b + kv.second) found = true;
            }
            if (!found) throw CsdError("galois key switch: host keys must lie in blocks from csd_host_alloc");
        }
        if (!ctx->copy_stream) HIPCHECK(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        DevBuf stage0(ctx, key_bytes), stage1(ctx, key_bytes);
        void* stage[2] = {stage0.p, stage1.p};
        hipEvent_t copied[2], consumed[2];
        for (int k = 0; k < 2; k++) {
            HIPCHECK(hipEventCreateWithFlags(&copied[k], hipEventDisableTiming));
            HIPCHECK(hipEventCreateWithFlags(&consumed[k], hipEventDisableTiming));
        }
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(ctx->copy_stream);
            (void)hipStreamSynchronize(ctx->stream);
            for (int k = 0; k < 2; k++) {
                (void)hipEventDestroy(copied[k]);
                (void)hipEventDestroy(consumed[k]);
            }
        };
        try {
            // the staging buffers came from the pool: earlier work on the compute stream may still use them
            HIPCHECK(hipEventRecord(consumed[0], ctx->stream));
            HIPCHECK(hipEventRecord(consumed[1], ctx->stream));
            const int64_t per_g = batch * 2 * r * n2 * n3;
            std::vector<csd_view> kv(gs.size());
            for (size_t gi = 0; gi < gs.size(); gi++) {
                const int k = (int)(gi & 1);
                HIPCHECK(hipStreamWaitEvent(ctx->copy_stream, consumed[k], 0));
                HIPCHECK(hipMemcpyAsync(stage[k], host_keys[gi], key_bytes, hipMemcpyHostToDevice, ctx->copy_stream));
                HIPCHECK(hipEventRecord(copied[k], ctx->copy_stream));
                HIPCHECK(hipStreamWaitEvent(ctx->stream, copied[k], 0));
                csd_view key = make_view(stage[k], 2 * d, r, n2, n3);
                csd_view o = sub(out, 0, (int64_t)gi * batch * 2, (int64_t)(gi + 1) * batch * 2, 8);
                (void)per_g;
                galois_keyswitch(base, batch, d, parts, c0, std::vector<uint64_t>{gs[gi]}, &key, o);
                HIPCHECK(hipEventRecord(consumed[k], ctx->stream));
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    }
};

static std::vector<size_t> to_vec(const size_t* p, size_t n) { return std::vector<size_t>(p, p + n); }

extern "C" int csd_field_info(const char* name, uint64_t* m, uint64_t* deg_l, uint64_t* deg_k, uint64_t* deg_lprime,
                              uint64_t* gal_ml_len, uint64_t* rou_order, size_t* orders, size_t cap, size_t* n_orders,
                              size_t* lp_count) {
    CSD_TRY auto f = make_field(name);
    *m = f->m;
    *deg_l = f->deg_l;
    *deg_k = f->deg_k;
    *deg_lprime = f->deg_lp;
    *gal_ml_len = f->gal_ml.size();
    *rou_order = f->rou_order();
    auto o = f->orders();
    if (o.size() > cap) throw CsdError("orders buffer too small");
    for (size_t i = 0; i < o.size(); i++) orders[i] = o[i];
    *n_orders = o.size();
    *lp_count = f->lp_count;
    CSD_CATCH
}

extern "C" int csd_field_gal_to_shift(const char* name, uint64_t g, int64_t* shifts, size_t cap, size_t* n) {
    CSD_TRY auto f = make_field(name);
    auto s = f->gal_to_shift(g);
    if (s.size() > cap) throw CsdError("shift buffer too small");
    for (size_t i = 0; i < s.size(); i++) shifts[i] = s[i];
    *n = s.size();
    CSD_CATCH
}

// Per prime p | m_K (field.rs:217-243, m_k_factors field.rs:271): the generator i_p_gen of the
// inertia group I_p (the smallest generator of (Z/pZ)*, embedded as 1 mod m/p) and the length
// |J_p| = (p - 1) / |I_p ∩ Gal(M|L)| of the cyclic transform factor (homcrt.rs:79-81)
struct HomcrtFactor {
    uint64_t p, gen, len;
};

static std::vector<HomcrtFactor> homcrt_factors(const Field& f) {
    std::unordered_map<uint64_t, bool> ml;
    for (uint64_t h : f.gal_ml) ml[h] = true;
    std::vector<HomcrtFactor> res;
    for (uint64_t p : f.m_factors) {
        if (p == f.m_prime && f.m_prime != 1) continue;
        std::vector<uint64_t> pf;
        uint64_t rest = p - 1;
        for (uint64_t q = 2; rest > 1; q++)
            if (rest % q == 0) {
                pf.push_back(q);
                while (rest % q == 0) rest /= q;
            }
        uint64_t gen = 2;
        for (;; gen++) {
            bool ok_ = true;
            for (uint64_t q : pf)
                if (powmod_u(gen, (p - 1) / q, p) == 1) ok_ = false;
            if (ok_) break;
        }
        uint64_t ipg = crt2(gen, p, 1, f.m / p);
        uint64_t hp = 0, x = 1;
        for (uint64_t e = 0; e < p - 1; e++) {
            if (ml.count(x)) hp++;
            x = mulmod_u(x, ipg, f.m);
        }
        res.push_back({p, ipg, (p - 1) / hp});
    }
    return res;
}

// required_galois_elements (bootstrap/homcrt.rs:336-349): per factor the baby-step/giant-step
// split of its length (homcrt.rs:320-324), the baby-step powers of i_p_gen and the giant step
static std::vector<uint64_t> homcrt_galois_elements(const Field& f, std::vector<uint64_t>* orders = nullptr) {
    std::vector<uint64_t> res;
    for (const HomcrtFactor& fac : homcrt_factors(f)) {
        const uint64_t len = fac.len, ipg = fac.gen;
        if (orders) orders->push_back(len);
        if (len <= 1) continue;
        uint64_t best = 1, best_cost = UINT64_MAX;
        for (uint64_t c = 1; c <= len; c++) {
            uint64_t cost = c + 3 * ((len - 1) / c + 1);
            if (cost < best_cost) {
                best_cost = cost;
                best = c;
            }
        }
        uint64_t giants = (len - 1) / best + 1;
        for (uint64_t j = 1; j < best; j++) res.push_back(powmod_u(ipg, j, f.m));
        if (giants > 1) res.push_back(powmod_u(ipg, best, f.m));
    }
    std::sort(res.begin(), res.end());
    res.erase(std::unique(res.begin(), res.end()), res.end());
    return res;
}

extern "C" int csd_field_data(const char* name, uint64_t* m, uint64_t* m_prime, uint64_t* mp_gens, size_t* n_mp,
                              uint64_t* mk_gens, size_t* n_mk, uint64_t* cyc_gens, uint64_t* cyc_orders, size_t* n_cyc,
                              size_t* n_lprime, size_t cap) {
    CSD_TRY for (const FieldData& d : field_table()) {
        if (name == nullptr || std::string(name) != d.name) continue;
        if (d.mp_gens.size() > cap || d.mk_gens.size() > cap || d.lq_cyc.size() + d.kq_cyc.size() > cap)
            throw CsdError("buffer too small");
        *m = d.m;
        *m_prime = d.m_prime;
        std::copy(d.mp_gens.begin(), d.mp_gens.end(), mp_gens);
        *n_mp = d.mp_gens.size();
        std::copy(d.mk_gens.begin(), d.mk_gens.end(), mk_gens);
        *n_mk = d.mk_gens.size();
        size_t k = 0;
        for (auto& c : d.lq_cyc) {
            cyc_gens[k] = c.first;
            cyc_orders[k++] = c.second;
        }
        for (auto& c : d.kq_cyc) {
            cyc_gens[k] = c.first;
            cyc_orders[k++] = c.second;
        }
        *n_cyc = k;
        *n_lprime = d.lq_cyc.size();
        return 0;
    }
    throw CsdError("unknown field descriptor");
    CSD_CATCH
}

extern "C" int csd_field_homcrt_factors(const char* name, uint64_t* p, uint64_t* gen, uint64_t* len, size_t cap, size_t* n) {
    CSD_TRY auto f = make_field(name);
    auto facs = homcrt_factors(*f);
    if (facs.size() > cap) throw CsdError("buffer too small");
    for (size_t i = 0; i < facs.size(); i++) {
        p[i] = facs[i].p;
        gen[i] = facs[i].gen;
        len[i] = facs[i].len;
    }
    *n = facs.size();
    CSD_CATCH
}

extern "C" int csd_field_homcrt_galois_elements(const char* name, uint64_t* out, size_t cap, size_t* n, uint64_t* jp_orders,
                                                size_t jp_cap, size_t* n_jp) {
    CSD_TRY auto f = make_field(name);
    std::vector<uint64_t> orders;
    auto g = homcrt_galois_elements(*f, &orders);
    if (g.size() > cap || orders.size() > jp_cap) throw CsdError("buffer too small");
    std::copy(g.begin(), g.end(), out);
    std::copy(orders.begin(), orders.end(), jp_orders);
    *n = g.size();
    *n_jp = orders.size();
    CSD_CATCH
}

static void sample(const Field& f, size_t lo, size_t hi, size_t t_bits, std::vector<int64_t>& all, size_t& init_len) {
    int64_t rou = (int64_t)f.rou_order();
    auto base = extend_sampled_primes({}, lo, hi, 57, rou);
    all = extend_sampled_primes(base, 2 * hi + 60 + t_bits, 2 * hi + 60 + t_bits + 30, 57, rou);  // ring.rs:620-632
    init_len = base.size();
}

extern "C" int csd_sample_primes(const char* name, size_t lo, size_t hi, size_t t_bits, int64_t* out, size_t cap,
                                 size_t* n, size_t* initial_len) {
    CSD_TRY if (hi <= lo) throw CsdError("log2_q range is empty");
    auto f = make_field(name);
    std::vector<int64_t> all;
    size_t il = 0;
    sample(*f, lo, hi, t_bits, all, il);
    if (all.size() > cap) throw CsdError("prime buffer too small");
    std::copy(all.begin(), all.end(), out);
    *n = all.size();
    *initial_len = il;
    CSD_CATCH
}

extern "C" int csd_ring_create(csd_ctx* ctx, const char* name, size_t lo, size_t hi, size_t t_bits, csd_ring** out) {
    CSD_TRY if (hi <= lo) throw CsdError("log2_q range is empty");
    auto r = std::make_unique<csd_ring>();
    r->ctx = ctx;
    r->f = make_field(name);
    size_t il = 0;
    sample(*r->f, lo, hi, t_bits, r->primes, il);
    if (r->primes.size() > CSD_MAX_MODULI) throw CsdError("too many RNS factors");
    r->initial_len = il;
    r->t_bits = t_bits;
    for (int64_t p : r->primes) {
        if (p < ((int64_t)1 << 40) || p > ((int64_t)1 << 57) - 1) throw CsdError("RNS factor out of range");
        csd_modulus m;
        ok(csd_hal_create_modulus(p, &m));
        r->mods.push_back(m);
        r->supports_transform.push_back(true);
    }
    *out = r.release();
    CSD_CATCH
}

extern "C" int csd_ring_create_raw(csd_ctx* ctx, const char* name, const int64_t* primes, size_t n, csd_ring** out) {
    CSD_TRY auto r = std::make_unique<csd_ring>();
    r->ctx = ctx;
    r->f = make_field(name);
    if (n > CSD_MAX_MODULI) throw CsdError("too many RNS factors");
    for (size_t i = 0; i < n; i++) {
        int64_t p = primes[i];
        if (p < ((int64_t)1 << 40) || p > ((int64_t)1 << 57) - 1) throw CsdError("RNS factor out of range");
        for (size_t k = 0; k < i; k++)
            if (primes[k] == p) throw CsdError("duplicate RNS factor");
        csd_modulus m;
        ok(csd_hal_create_modulus(p, &m));
        r->primes.push_back(p);
        r->mods.push_back(m);
        r->supports_transform.push_back(false);
    }
    r->initial_len = n;
    r->t_bits = 0;
    *out = r.release();
    CSD_CATCH
}

extern "C" int csd_ring_destroy(csd_ring* ring) {
    CSD_TRY delete ring;
    CSD_CATCH
}

extern "C" int csd_ring_primes(csd_ring* ring, int64_t* out, size_t cap, size_t* n, size_t* initial_len) {
    CSD_TRY if (ring->primes.size() > cap) throw CsdError("prime buffer too small");
    std::copy(ring->primes.begin(), ring->primes.end(), out);
    *n = ring->primes.size();
    *initial_len = ring->initial_len;
    CSD_CATCH
}

extern "C" int csd_ring_precompute_transforms(csd_ring* ring) {
    CSD_TRY for (size_t i = 0; i < ring->primes.size(); i++) if (ring->supports_transform[i]) ring->transform_cache({i});
    CSD_CATCH
}

extern "C" int csd_ring_crt_kernels(csd_ring* ring, size_t idx, int64_t* lp, int64_t* k, int64_t* ilp, int64_t* ik) {
    CSD_TRY if (idx >= ring->primes.size()) throw CsdError("prime index out of range");
    ring->transform_cache({idx});
    const Kernels& K = ring->kernels.at(ring->primes[idx]);
    std::copy(K.lp.begin(), K.lp.end(), lp);
    std::copy(K.k.begin(), K.k.end(), k);
    std::copy(K.inv_lp.begin(), K.inv_lp.end(), ilp);
    std::copy(K.inv_k.begin(), K.inv_k.end(), ik);
    CSD_CATCH
}

extern "C" int csd_ring_extension_base(csd_ring* ring, const size_t* base, size_t r, size_t* out, size_t cap, size_t* n) {
    CSD_TRY auto b = to_vec(base, r);
    for (size_t i : b)
        if (i >= ring->initial_len) throw CsdError("extension base: base must lie in the initial base");
    std::vector<int64_t> ps;
    for (size_t i : b) ps.push_back(ring->primes[i]);
    double required = (double)(product_bit_length(ps) + ring->t_bits + 60);  // ring.rs:716-731
    double bits = 0.0;
    std::vector<size_t> res;
    for (size_t idx = ring->initial_len; idx < ring->primes.size(); idx++) {
        if (bits >= required) break;
        res.push_back(idx);
        bits += std::log2((double)ring->primes[idx]);
    }
    if (bits < required) throw CsdError("extension pool too small for lifted multiplication over this base");
    if (res.size() > cap) throw CsdError("buffer too small");
    std::copy(res.begin(), res.end(), out);
    *n = res.size();
    CSD_CATCH
}

extern "C" int csd_ring_pb_to_crt(csd_ring* ring, const csd_view* data, const size_t* base, size_t r) {
    CSD_TRY check_view(data, "pb_to_crt");
    ring->transform(*data, to_vec(base, r), false);
    CSD_CATCH
}

// The L' half of pb_to_crt_tensor along dim 2 (ntt_mul with the L' CRT kernel, then intt) on a
// [count, r, deg_L', w] tensor of any width w: what lprime_element_tensor (ring.rs:1227-1250)
// applies to the [1, r, deg_L', 1] montgomery-form coefficients of an element of L'.
extern "C" int csd_ring_lprime_transform(csd_ring* ring, const csd_view* data, const size_t* base, size_t r) {
    CSD_TRY check_view(data, "lprime_transform");
    auto b = to_vec(base, r);
    TransformCache& c = ring->transform_cache(b);
    if (data->shape[1] != (int64_t)r || data->shape[2] != (int64_t)ring->f->deg_lp)
        throw CsdError("lprime_transform: tensor shape doesn't match [count, r, deg_L', .]");
    csd_view klp = make_view(c.k_lp->p, 1, (int64_t)r, (int64_t)ring->f->deg_lp, 1);
    ok(csd_hal_ntt_mul_mont64_dim(ring->ctx, data, c.plans_lp.data(), r, CSD_DIM2, &klp, c.mods.data(), r));
    ok(csd_hal_intt_mont64_dim(ring->ctx, data, c.plans_lp.data(), r, CSD_DIM2));
    CSD_CATCH
}

extern "C" int csd_ring_decode_scaled(csd_ring* ring, const csd_view* src, const size_t* base, size_t r, size_t log2_t,
                                      const csd_view* dst) {
    CSD_TRY auto b = to_vec(base, r);
    std::vector<int64_t> ps;
    for (size_t i : b) {
        if (i >= ring->primes.size()) throw CsdError("decode_scaled: base index out of range");
        ps.push_back(ring->primes[i]);
    }
    auto m = ring->mods_for(b);
    csd_decode_scaled_launch(ring->ctx, src, ps.data(), m.data(), r, log2_t, dst);
    CSD_CATCH
}

extern "C" int csd_ring_encode_digits(csd_ring* ring, const csd_view* values, size_t log2_b, size_t max_digits,
                                      const csd_view* dst, const size_t* base, size_t r, int64_t* planes) {
    CSD_TRY auto b = to_vec(base, r);
    auto m = ring->mods_for(b);
    int64_t n = csd_encode_digits_launch(ring->ctx, values, log2_b, max_digits, dst, m.data(), r);
    if (planes) *planes = n;
    CSD_CATCH
}

extern "C" int csd_ring_crt_to_pb(csd_ring* ring, const csd_view* data, const size_t* base, size_t r) {
    CSD_TRY check_view(data, "crt_to_pb");
    ring->transform(*data, to_vec(base, r), true);
    CSD_CATCH
}

extern "C" int csd_ring_bconv(csd_ring* ring, const csd_view* dst, const size_t* dst_base, size_t rd, const csd_view* src,
                              const size_t* src_base, size_t rs, int preserve_src) {
    CSD_TRY check_view(dst, "bconv dst");
    check_view(src, "bconv src");
    ring->bconv(*dst, to_vec(dst_base, rd), *src, to_vec(src_base, rs), preserve_src != 0);
    CSD_CATCH
}

extern "C" int csd_ring_rescale_step(csd_ring* ring, const csd_view* ext, const size_t* ext_base, size_t re,
                                     const csd_view* src, const size_t* src_base, size_t rs, const csd_view* s_src,
                                     const csd_view* s_ext, const csd_view* inv) {
    CSD_TRY check_view(ext, "rescale ext");
    check_view(src, "rescale src");
    check_view(s_src, "rescale scalar");
    check_view(s_ext, "rescale scalar");
    check_view(inv, "rescale inverse");
    ring->rescale_step(*ext, to_vec(ext_base, re), *src, to_vec(src_base, rs), *s_src, *s_ext, *inv);
    CSD_CATCH
}

extern "C" int csd_ring_bconv_gadget(csd_ring* ring, const csd_view* dst, const csd_view* src, const size_t* base, size_t r,
                                     const size_t* bounds, size_t nd) {
    CSD_TRY check_view(dst, "gadget bconv dst");
    check_view(src, "gadget bconv src");
    ring->bconv_gadget(*dst, *src, to_vec(base, r), to_vec(bounds, 2 * nd));
    CSD_CATCH
}

extern "C" int csd_ring_mod_switch_down(csd_ring* ring, const csd_view* dst, const csd_view* src, const size_t* base, size_t r,
                                        size_t new_len) {
    CSD_TRY check_view(dst, "mod_switch_down dst");
    check_view(src, "mod_switch_down src");
    ring->mod_switch_down(*dst, *src, to_vec(base, r), new_len);
    CSD_CATCH
}

extern "C" int csd_ring_apply_galois(csd_ring* ring, const csd_view* dst, const csd_view* src, uint64_t g) {
    CSD_TRY check_view(dst, "apply_galois dst");
    check_view(src, "apply_galois src");
    ring->apply_galois(*dst, *src, g);
    CSD_CATCH
}

extern "C" int csd_ring_galois_keyswitch(csd_ring* ring, const size_t* base, size_t r, size_t batch, size_t digits,
                                         const csd_view* parts, const csd_view* c0, const uint64_t* gs, size_t ngal,
                                         const csd_view* keys, const csd_view* out) {
    CSD_TRY if (batch == 0 || digits == 0 || ngal == 0) throw CsdError("galois key switch: empty batch, digits or rotations");
    ring->galois_keyswitch(to_vec(base, r), (int64_t)batch, (int64_t)digits, *parts, *c0, std::vector<uint64_t>(gs, gs + ngal),
                           keys, *out);
    CSD_CATCH
}

extern "C" int csd_ring_galois_keyswitch_rows(csd_ring* ring, const size_t* base, size_t r, size_t batch, size_t group,
                                              size_t digits, const csd_view* parts, const csd_view* c0, int64_t c0_hi,
                                              int64_t c0_lo, const uint64_t* gs, size_t ngal, const csd_view* keys,
                                              const csd_view* outs, int64_t o_hi, int64_t o_lo, int64_t o_s) {
    CSD_TRY if (batch == 0 || digits == 0 || ngal == 0 || group == 0) throw CsdError("galois key switch: empty batch, digits or rotations");
    GksRows rows{(int64_t)group, c0_hi, c0_lo, o_hi, o_lo, o_s};
    ring->galois_keyswitch_rows(to_vec(base, r), (int64_t)batch, (int64_t)digits, *parts, *c0, std::vector<uint64_t>(gs, gs + ngal),
                                keys, outs, rows);
    CSD_CATCH
}

extern "C" int csd_ring_set_overlap(csd_ring* ring, int on) {
    CSD_TRY if (!ring) throw CsdError("null ring");
    ring->overlap_halves = on != 0;
    CSD_CATCH
}

extern "C" int csd_ring_get_bconv_form(const csd_ring* ring) { return ring ? ring->bconv_form.load() : 0; }

extern "C" int csd_ring_set_bconv_form(csd_ring* ring, int form) {
    CSD_TRY if (!ring) throw CsdError("null ring");
    if (form < 0 || form > 3)
        throw CsdError("bconv form must be 0 (HAL sequence), 1 (fused), 2 (fused, matrix cores where they pay) or 3 (fused, "
                       "matrix cores wherever the tables fit)");
    ring->bconv_form = form;
    CSD_CATCH
}

extern "C" int csd_ring_set_gadget_form(csd_ring* ring, int form) {
    CSD_TRY if (!ring) throw CsdError("null ring");
    if (form != 0 && form != 1) throw CsdError("gadget bconv form must be 0 (HAL sequence) or 1 (fused)");
    ring->gadget_fused = form == 1;
    CSD_CATCH
}

extern "C" int csd_ring_galois_keyswitch_streamed(csd_ring* ring, const size_t* base, size_t r, size_t batch, size_t digits,
                                                  const csd_view* parts, const csd_view* c0, const uint64_t* gs, size_t ngal,
                                                  const int64_t* const* host_keys, const csd_view* out) {
    CSD_TRY if (batch == 0 || digits == 0 || ngal == 0) throw CsdError("galois key switch: empty batch, digits or rotations");
    ring->galois_keyswitch_streamed(to_vec(base, r), (int64_t)batch, (int64_t)digits, *parts, *c0,
                                    std::vector<uint64_t>(gs, gs + ngal), host_keys, *out);
    CSD_CATCH
}

extern "C" int csd_ring_reduce_wide(csd_ring* ring, const csd_view* dst, const csd_view* wide, const size_t* base, size_t r,
                                    int full) {
    CSD_TRY auto m = ring->mods_for(to_vec(base, r));
    if (full) {
        ok(csd_hal_red_mont64_double(ring->ctx, dst, wide, m.data(), r));
        ok(csd_hal_i64_to_mont64(ring->ctx, dst, m.data(), r));
    } else {
        ok(csd_hal_red_mont64(ring->ctx, dst, wide, m.data(), r));
    }
    CSD_CATCH
}

extern "C" int csd_ring_scalar_tensor(csd_ring* ring, const csd_view* dst, const size_t* base, size_t r, int64_t scalar) {
    CSD_TRY if (dst->shape[0] != 1 || dst->shape[1] != (int64_t)r || dst->shape[2] != 1 || dst->shape[3] != 1)
        throw CsdError("scalar tensor must have shape [1, r, 1, 1]");
    std::vector<int64_t> v;
    for (size_t t = 0; t < r; t++) {
        int64_t p = ring->primes.at(base[t]);
        v.push_back(mont_repr(scalar % p, p));  // ring.rs:876-888
    }
    ok(csd_hal_import(ring->ctx, dst, v.data(), v.size()));
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// batched uneven chain: pb_to_crt -> mul_mont64 by a CRT-basis plaintext -> crt_to_pb ->
// mod_switch_down by one limb, for ciphertexts at different levels. Every (element, limb) slice
// of every ciphertext is one entry of a work list; each of the eight transform phases is one
// launch over the whole list, the plaintext product rides on the store of the fourth, and the
// mod-switch is one launch over all elements. The phases can be replayed as one hipGraph.

struct csd_batch_plan {
    csd_ring* ring = nullptr;
    std::vector<size_t> base;
    std::vector<size_t> limbs;
    TransformCache* tc = nullptr;
    csd_view pt;
    int64_t n2 = 0, n3 = 0, nslices = 0, nelems = 0;
    std::unique_ptr<DevBuf> tables;
    int64_t** seg_base = nullptr;
    int32_t* seg_limb = nullptr;
    int64_t** ms_in = nullptr;
    int64_t** ms_out = nullptr;
    int32_t* ms_level = nullptr;
    void* levels = nullptr;
    csd_batch_tuning tune{0, 0, 256, 3, 2, 0, 0};
    int num_cus = 0;
    hipGraphExec_t graph = nullptr;

    ~csd_batch_plan() {
        if (graph) (void)hipGraphExecDestroy(graph);
    }

    double chain_bytes() const {
        // eight transform passes over every slice (read + write each value once), the plaintext
        // read of the fourth, and the mod-switch (read r limbs, write r - 1) per element
        const double slice = 8.0 * n2 * n3;
        double b = 8 * 2 * slice * nslices + slice * nslices;
        for (size_t c = 0; c < limbs.size(); c++) b += 2.0 * slice * (2 * limbs[c] - 1);
        return b;
    }

    CsdSegTune seg_tune() const {
        CsdSegTune t;
        t.rows = tune.rows;
        t.threads = tune.threads;
        t.col_threads = tune.col_threads;
        return t;
    }

    // the chain on the stream; ev (optional) receives an event before each phase and one after
    void enqueue(hipEvent_t* ev = nullptr) {
        csd_ctx* ctx = ring->ctx;
        const int64_t R = (int64_t)base.size();
        const CsdSegTune t = seg_tune();
        csd_view klp = make_view(tc->k_lp->p, 1, R, n2, 1), kk = make_view(tc->k_k->p, 1, R, 1, n3);
        csd_view iklp = make_view(tc->ik_lp->p, 1, R, n2, 1), ikk = make_view(tc->ik_k->p, 1, R, 1, n3);
        csd_view ptv = pt;
        ptv.shape[1] = R;
        const csd_modulus* m = tc->mods.data();
        auto phase = [&](int q) {
            if (ev) HIPCHECK(hipEventRecord(ev[q], ctx->stream));
        };
        if (tune.persistent) {
            // the eight transform passes as one persistent launch (csd_batch_chain_launch)
            phase(0);
            const int rows = tune.rows > 0 ? tune.rows : 1;
            const int grid = tune.grid > 0 ? tune.grid : default_grid(rows);
            double cb = 8 * 2 * 8.0 * n2 * n3 * nslices + 8.0 * n2 * n3 * nslices;
            csd_batch_chain_launch(ctx, seg_base, seg_limb, nslices, tc->plans_lp.data(), tc->plans_k.data(), (size_t)R,
                                   (const int64_t*)tc->k_lp->p, (const int64_t*)tc->k_k->p, (const int64_t*)tc->ik_lp->p,
                                   (const int64_t*)tc->ik_k->p, (const int64_t*)pt.ptr, m, rows, grid, cb);
            for (int q = 1; q <= 8; q++) phase(q);
            double msb = 0;
            for (size_t c = 0; c < limbs.size(); c++) msb += 2.0 * 8.0 * n2 * n3 * (2 * limbs[c] - 1);
            csd_mod_switch_one_launch(ctx, ms_in, ms_out, ms_level, levels, nelems, n2 * n3, m, (size_t)R, msb);
            phase(9);
            return;
        }
        // pb_to_crt (ring.rs:1118-1141)
        phase(0);
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_lp.data(), R, 2, 1, &klp, m, "batch_ntt_mul_dim2", t);
        phase(1);
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_k.data(), R, 3, 1, &kk, m, "batch_ntt_mul_dim3", t);
        phase(2);
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_lp.data(), R, 2, 2, nullptr, m, "batch_intt_dim2", t);
        phase(3);
        // intt along dim 3, then the plaintext product (hal mul_mont64) on the way out
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_k.data(), R, 3, 2, &ptv, m, "batch_intt_mul_dim3", t);
        // crt_to_pb (ring.rs:1144-1161)
        phase(4);
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_lp.data(), R, 2, 1, &iklp, m, "batch_ntt_mul_dim2", t);
        phase(5);
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_k.data(), R, 3, 1, &ikk, m, "batch_ntt_mul_dim3", t);
        phase(6);
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_lp.data(), R, 2, 2, nullptr, m, "batch_intt_dim2", t);
        phase(7);
        csd_ntt_segments(ctx, seg_base, seg_limb, nslices, n2, n3, n3, tc->plans_k.data(), R, 3, 2, nullptr, m, "batch_intt_dim3", t);
        // mod_switch_down by one limb (bfv.rs:628-662)
        phase(8);
        double msb = 0;
        for (size_t c = 0; c < limbs.size(); c++) msb += 2.0 * 8.0 * n2 * n3 * (2 * limbs[c] - 1);
        csd_mod_switch_one_launch(ctx, ms_in, ms_out, ms_level, levels, nelems, n2 * n3, m, (size_t)R, msb);
        phase(9);
    }

    // workgroups resident at once: CUs x the persistent chain's occupancy with `rows` lines of LDS
    int default_grid(int rows) {
        if (!num_cus) {
            hipDeviceProp_t prop;
            HIPCHECK(hipGetDeviceProperties(&prop, ring->ctx->device));
            num_cus = prop.multiProcessorCount;
        }
        return num_cus * csd_batch_chain_occupancy(rows);
    }

    bool persistent_supported() const {
        return csd_batch_chain_supported(tc->plans_lp.data(), tc->plans_k.data(), base.size());
    }

    void drop_graph() {
        if (graph) {
            HIPCHECK(hipStreamSynchronize(ring->ctx->stream));
            HIPCHECK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
    }

    void capture() {
        csd_ctx* ctx = ring->ctx;
        const bool timing = ctx->timing;
        ctx->timing = false;  // no event records inside the captured region
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        try {
            enqueue();
        } catch (...) {
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            ctx->timing = timing;
            throw;
        }
        ctx->timing = timing;
        HIPCHECK(hipStreamEndCapture(ctx->stream, &g));
        hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(e);
    }

    void run(bool use_graph) {
        if (nslices == 0) return;
        if (!use_graph) {
            enqueue();
            return;
        }
        if (!graph) capture();
        TimedScope ts(ring->ctx, "batch_chain_graph", chain_bytes());
        HIPCHECK(hipGraphLaunch(graph, ring->ctx->stream));
    }

    void set_tuning(const csd_batch_tuning& t) {
        if (t.g2 < 1 || t.g2 > 3 || t.g3 < 1 || t.g3 > 2) throw CsdError("tuning: 1 <= g2 <= 3 and 1 <= g3 <= 2");
        if (t.col_threads < 64 || t.col_threads > 256 || t.col_threads % 64) throw CsdError("tuning: col_threads");
        if (t.threads < 0 || t.threads > 512 || t.threads % 64) throw CsdError("tuning: threads");
        if (t.rows < 0 || t.rows > 16) throw CsdError("tuning: rows");
        if (t.rows > 0 && (size_t)t.rows * (n3 + n3 / 8 + 1) * 8 > 160 * 1024) throw CsdError("tuning: rows exceed LDS");
        if (t.persistent != 0 && t.persistent != 1) throw CsdError("tuning: persistent is 0 or 1");
        if (t.grid < 0) throw CsdError("tuning: grid");
        if (t.persistent && (t.rows > 4 || t.g2 != 3 || t.g3 != 2)) throw CsdError("tuning: the persistent chain runs 1-4 lines, default grouping");
        if (t.persistent && !persistent_supported()) throw CsdError("tuning: no persistent chain for this field");
        drop_graph();
        if (t.g2 != tune.g2 || t.g3 != tune.g3) {
            for (auto p : tc->plans_lp) ok(csd_ntt_plan_set_grouping(p, t.g2, t.g3));
            for (auto p : tc->plans_k) ok(csd_ntt_plan_set_grouping(p, t.g2, t.g3));
        }
        tune = t;
    }

    // chain time of the current tuning (ms, best of reps) and the per-phase split
    double time_chain(int reps, double* phase_ms) {
        csd_ctx* ctx = ring->ctx;
        hipEvent_t ev[10];
        for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
        double best = 1e300;
        std::vector<double> best_phase(9, 0.0);
        try {
            enqueue();  // warm-up
            for (int q = 0; q < reps; q++) {
                enqueue(ev);
                HIPCHECK(hipStreamSynchronize(ctx->stream));
                float tot = 0;
                HIPCHECK(hipEventElapsedTime(&tot, ev[0], ev[9]));
                if (tot < best) {
                    best = tot;
                    for (int k = 0; k < 9; k++) {
                        float f = 0;
                        HIPCHECK(hipEventElapsedTime(&f, ev[k], ev[k + 1]));
                        best_phase[k] = f;
                    }
                }
            }
        } catch (...) {
            for (auto& e : ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (phase_ms)
            for (int k = 0; k < 9; k++) phase_ms[k] = best_phase[k];
        return best;
    }

    // coordinate search: dim-3 launch shape, then pass grouping, then the dim-2 block size.
    // Every candidate computes the same values; only the time differs.
    double autotune(int reps) {
        const bool timing = ring->ctx->timing;
        ring->ctx->timing = false;
        auto dim3_ms = [](const double* ph) { return ph[1] + ph[3] + ph[5] + ph[7]; };
        auto dim2_ms = [](const double* ph) { return ph[0] + ph[2] + ph[4] + ph[6]; };
        double ph[9];
        csd_batch_tuning best = tune;
        best.persistent = 0;  // the staged search first; the persistent chain is compared after it
        best.grid = 0;
        try {
            set_tuning(best);
            time_chain(reps, ph);
            double best3 = dim3_ms(ph);
            const size_t line = (size_t)(n3 + n3 / 8 + 1) * 8;
            for (int rows : {1, 2, 3, 4, 6, 8})
                for (int threads : {128, 256, 384, 512}) {
                    if ((size_t)rows * line > 160 * 1024 || rows > n2) continue;
                    if (threads > rows * n3) continue;
                    csd_batch_tuning c = best;
                    c.rows = rows;
                    c.threads = threads;
                    set_tuning(c);
                    time_chain(reps, ph);
                    if (dim3_ms(ph) < best3) {
                        best3 = dim3_ms(ph);
                        best = c;
                    }
                }
            for (int g2 = 1; g2 <= 3; g2++)
                for (int g3 = 1; g3 <= 2; g3++) {
                    csd_batch_tuning c = best;
                    c.g2 = g2;
                    c.g3 = g3;
                    set_tuning(c);
                    time_chain(reps, ph);
                    if (dim3_ms(ph) < best3) {
                        best3 = dim3_ms(ph);
                        best = c;
                    }
                }
            set_tuning(best);
            time_chain(reps, ph);
            double best2 = dim2_ms(ph);
            for (int ct : {64, 128, 192, 256}) {
                csd_batch_tuning c = best;
                c.col_threads = ct;
                set_tuning(c);
                time_chain(reps, ph);
                if (dim2_ms(ph) < best2) {
                    best2 = dim2_ms(ph);
                    best = c;
                }
            }
            set_tuning(best);
            // the persistent chain against the best staged form, on the whole chain time
            if (persistent_supported()) {
                double best_total = time_chain(reps, nullptr);
                csd_batch_tuning staged = best;
                for (int rows : {1, 2})
                    for (int mult : {1, 2, 0}) {
                        csd_batch_tuning c = staged;
                        c.persistent = 1;
                        c.rows = rows;
                        c.g2 = 3;
                        c.g3 = 2;
                        c.grid = mult ? num_cus_or_query() * mult : 0;
                        set_tuning(c);
                        const double tot = time_chain(reps, nullptr);
                        if (tot < best_total) {
                            best_total = tot;
                            best = c;
                        }
                    }
                set_tuning(best);
            }
        } catch (...) {
            ring->ctx->timing = timing;
            throw;
        }
        ring->ctx->timing = timing;
        return time_chain(reps, nullptr);
    }

    int num_cus_or_query() {
        if (!num_cus) (void)default_grid(1);
        return num_cus;
    }
};

static csd_batch_plan* batch_create(csd_ring* ring, const std::vector<size_t>& base, size_t count, int64_t* const* cts,
                                    int64_t* const* outs, const size_t* limbs, const csd_view& pt) {
    const Field& F = *ring->f;
    const int64_t R = (int64_t)base.size();
    if (R < 2) throw CsdError("batch: the base needs at least two primes");
    if (R > CSD_MAX_MODULI) throw CsdError("batch: base too long");
    auto plan = std::make_unique<csd_batch_plan>();
    plan->ring = ring;
    plan->base = base;
    plan->n2 = (int64_t)F.deg_lp;
    plan->n3 = (int64_t)F.deg_k;
    check_view(&pt, "batch plaintext");
    if (pt.shape[0] != 1 || pt.shape[1] < R || pt.shape[2] != plan->n2 || pt.shape[3] != plan->n3)
        throw CsdError("batch: plaintext must be [1, >= r, deg_L', deg_K]");
    if (pt.strides[3] != 1 || pt.strides[2] != plan->n3 || pt.strides[1] != plan->n2 * plan->n3)
        throw CsdError("batch: plaintext must be contiguous per limb");
    plan->pt = pt;
    plan->tc = &ring->transform_cache(base);
    const int64_t npos = plan->n2 * plan->n3;
    std::vector<int64_t*> sb, msi, mso;
    std::vector<int32_t> sl, lvl;
    for (size_t c = 0; c < count; c++) {
        const size_t r = limbs[c];
        if (r < 2 || (int64_t)r > R) throw CsdError("batch: limb counts must lie in [2, r]");
        if (!cts[c] || !outs[c]) throw CsdError("batch: null ciphertext pointer");
        plan->limbs.push_back(r);
        for (int64_t e = 0; e < 2; e++) {
            for (size_t j = 0; j < r; j++) {
                sb.push_back(cts[c] + (e * (int64_t)r + (int64_t)j) * npos);
                sl.push_back((int32_t)j);
            }
            msi.push_back(cts[c] + e * (int64_t)r * npos);
            mso.push_back(outs[c] + e * (int64_t)(r - 1) * npos);
            lvl.push_back((int32_t)r);
        }
    }
    plan->nslices = (int64_t)sb.size();
    plan->nelems = (int64_t)msi.size();
    // per-level constants of the one-limb BConvCache (ring.rs:1606-1661) and of
    // inverse_scalar_tensor (ring.rs:1165-1177), taken from the ring's own caches
    const size_t lvb = csd_mod_switch_level_bytes();
    std::vector<char> lv(lvb * (size_t)(R + 1), 0);
    for (int64_t r = 2; r <= R; r++) {
        std::vector<size_t> kept(base.begin(), base.begin() + r - 1), dropped{base[(size_t)r - 1]};
        BConvCache& bc = ring->bconv_cache(dropped, kept);
        if (bc.acc < 1) throw CsdError("batch: bconv accumulation bound");
        std::vector<int64_t> M(bc.host_matrix.begin(), bc.host_matrix.begin() + (r - 1));
        const int64_t C = bc.host_matrix[(size_t)r - 1];
        std::vector<int64_t> inv = ring->inverse_product_values(kept, dropped);
        csd_mod_switch_fill_level(lv.data() + lvb * (size_t)r, (int)r, M.data(), bc.host_qm.data(), inv.data(), C,
                                  (int)bc.log2_qtilde, bc.host_qopi[0]);
    }
    // one device table block: slice pointers and limbs, element pointers and levels, constants
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_sb = 0, o_sl = al(o_sb + sb.size() * 8), o_mi = al(o_sl + sl.size() * 4), o_mo = al(o_mi + msi.size() * 8),
                 o_lv = al(o_mo + mso.size() * 8), o_c = al(o_lv + lvl.size() * 4), total = al(o_c + lv.size());
    std::vector<char> blob(total, 0);
    if (!sb.empty()) memcpy(blob.data() + o_sb, sb.data(), sb.size() * 8);
    if (!sl.empty()) memcpy(blob.data() + o_sl, sl.data(), sl.size() * 4);
    if (!msi.empty()) memcpy(blob.data() + o_mi, msi.data(), msi.size() * 8);
    if (!mso.empty()) memcpy(blob.data() + o_mo, mso.data(), mso.size() * 8);
    if (!lvl.empty()) memcpy(blob.data() + o_lv, lvl.data(), lvl.size() * 4);
    memcpy(blob.data() + o_c, lv.data(), lv.size());
    csd_ctx* ctx = ring->ctx;
    plan->tables = std::make_unique<DevBuf>(ctx, total);
    char* d = (char*)plan->tables->p;
    HIPCHECK(hipMemcpyAsync(d, blob.data(), total, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    plan->seg_base = (int64_t**)(d + o_sb);
    plan->seg_limb = (int32_t*)(d + o_sl);
    plan->ms_in = (int64_t**)(d + o_mi);
    plan->ms_out = (int64_t**)(d + o_mo);
    plan->ms_level = (int32_t*)(d + o_lv);
    plan->levels = d + o_c;
    // the plans of the transform cache may carry a grouping from an earlier batch plan
    for (auto p : plan->tc->plans_lp) ok(csd_ntt_plan_set_grouping(p, plan->tune.g2, plan->tune.g3));
    for (auto p : plan->tc->plans_k) ok(csd_ntt_plan_set_grouping(p, plan->tune.g2, plan->tune.g3));
    return plan.release();
}

extern "C" int csd_ring_batch_create(csd_ring* ring, const size_t* base, size_t r, size_t count, int64_t* const* cts,
                                     int64_t* const* outs, const size_t* limbs, const csd_view* pt, csd_batch_plan** out) {
    CSD_TRY if (!out || !pt || (count && (!cts || !outs || !limbs))) throw CsdError("batch: null argument");
    *out = batch_create(ring, to_vec(base, r), count, cts, outs, limbs, *pt);
    CSD_CATCH
}

extern "C" int csd_batch_run(csd_batch_plan* plan, int use_graph) {
    CSD_TRY plan->run(use_graph != 0);
    CSD_CATCH
}

extern "C" int csd_batch_autotune(csd_batch_plan* plan, int reps, csd_batch_tuning* chosen, double* ms) {
    CSD_TRY if (reps < 1) throw CsdError("autotune: reps must be positive");
    double t = plan->autotune(reps);
    if (chosen) *chosen = plan->tune;
    if (ms) *ms = t;
    CSD_CATCH
}

extern "C" int csd_batch_set_tuning(csd_batch_plan* plan, const csd_batch_tuning* t) {
    CSD_TRY plan->set_tuning(*t);
    CSD_CATCH
}

extern "C" int csd_batch_get_tuning(const csd_batch_plan* plan, csd_batch_tuning* t) {
    CSD_TRY *t = plan->tune;
    CSD_CATCH
}

extern "C" int csd_batch_info(const csd_batch_plan* plan, int64_t* nslices, int64_t* nelems, double* bytes) {
    CSD_TRY if (nslices) *nslices = plan->nslices;
    if (nelems) *nelems = plan->nelems;
    if (bytes) *bytes = plan->chain_bytes();
    CSD_CATCH
}

extern "C" int csd_batch_free(csd_batch_plan* plan) {
    CSD_TRY delete plan;
    CSD_CATCH
}

// scalar_tensor (ring.rs:876-888) for an i128 scalar given as (low, high) words
extern "C" int csd_ring_scalar_tensor_wide(csd_ring* ring, const csd_view* dst, const size_t* base, size_t r, int64_t lo,
                                           int64_t hi) {
    CSD_TRY if (dst->shape[0] != 1 || dst->shape[1] != (int64_t)r || dst->shape[2] != 1 || dst->shape[3] != 1)
        throw CsdError("scalar tensor must have shape [1, r, 1, 1]");
    const i128 scalar = (i128)(((u128)(uint64_t)hi << 64) | (uint64_t)lo);
    std::vector<int64_t> v;
    for (size_t t = 0; t < r; t++) {
        int64_t p = ring->primes.at(base[t]);
        v.push_back(mont_repr((int64_t)(scalar % (i128)p), p));  // truncating remainder, as Rust's %
    }
    ok(csd_hal_import(ring->ctx, dst, v.data(), v.size()));
    CSD_CATCH
}

// inverse_scalar_tensor (ring.rs:1165-1177) for the scalar prod(primes at `of`)
extern "C" int csd_ring_inverse_product_tensor(csd_ring* ring, const csd_view* dst, const size_t* base, size_t r,
                                               const size_t* of, size_t nof) {
    CSD_TRY if (dst->shape[0] != 1 || dst->shape[1] != (int64_t)r || dst->shape[2] != 1 || dst->shape[3] != 1)
        throw CsdError("scalar tensor must have shape [1, r, 1, 1]");
    std::vector<size_t> b = to_vec(base, r), o = to_vec(of, nof);
    ring->check_base(b);
    ring->check_base(o);
    std::vector<int64_t> v = ring->inverse_product_values(b, o);
    ok(csd_hal_import(ring->ctx, dst, v.data(), v.size()));
    CSD_CATCH
}

// The dense matrix of the period->CRT transform (inverse = 0) or of CRT->period (inverse = 1)
// along one dimension, per limb of `base`: pb_to_crt_tensor (ring.rs:1118-1141) is the
// convolution over the cyclic group of each dimension with the CRT kernel, so the matrix is the
// group-circulant C[a][b] = k[a - b]. k is read off the device transform itself, as the image of
// the unit vector e_0 under that dimension's ntt_mul + intt pair.
extern "C" int csd_ring_transform_matrix(csd_ring* ring, const size_t* base, size_t r, int dim, int inverse, const csd_view* out) {
    CSD_TRY check_view(out, "transform matrix");
    if (dim != 2 && dim != 3) throw CsdError("transform matrix: dim must be 2 or 3");
    std::vector<size_t> b = to_vec(base, r);
    TransformCache& c = ring->transform_cache(b);
    const Field& F = *ring->f;
    auto orders = F.orders();
    std::vector<size_t> o(dim == 2 ? orders.begin() : orders.begin() + F.lp_count,
                          dim == 2 ? orders.begin() + F.lp_count : orders.end());
    const int64_t n = dim == 2 ? (int64_t)F.deg_lp : (int64_t)F.deg_k, R = (int64_t)r;
    if (out->shape[0] != 1 || out->shape[1] != R || out->shape[2] != n || out->shape[3] != n)
        throw CsdError("transform matrix: output must be [1, r, n, n]");
    if (out->strides[2] != n || out->strides[1] != n * n) throw CsdError("transform matrix: output must be contiguous");
    csd_ctx* ctx = ring->ctx;
    DevBuf kv(ctx, (size_t)R * n * 8);
    std::vector<int64_t> delta((size_t)(R * n), 0);
    for (int64_t l = 0; l < R; l++) delta[(size_t)(l * n)] = 1;
    csd_view v = dim == 2 ? make_view(kv.p, 1, R, n, 1) : make_view(kv.p, 1, R, 1, n);
    ok(csd_hal_import(ctx, &v, delta.data(), delta.size()));
    if (dim == 2) {
        csd_view kl = make_view(inverse ? c.ik_lp->p : c.k_lp->p, 1, R, n, 1);
        ok(csd_hal_ntt_mul_mont64_dim(ctx, &v, c.plans_lp.data(), R, CSD_DIM2, &kl, c.mods.data(), R));
        ok(csd_hal_intt_mont64_dim(ctx, &v, c.plans_lp.data(), R, CSD_DIM2));
    } else {
        csd_view kk = make_view(inverse ? c.ik_k->p : c.k_k->p, 1, R, 1, n);
        ok(csd_hal_ntt_mul_mont64_dim(ctx, &v, c.plans_k.data(), R, CSD_DIM3, &kk, c.mods.data(), R));
        ok(csd_hal_intt_mont64_dim(ctx, &v, c.plans_k.data(), R, CSD_DIM3));
    }
    csd_circulant_launch(ctx, (const int64_t*)kv.p, (int64_t*)out->ptr, R, o);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    CSD_CATCH
}
