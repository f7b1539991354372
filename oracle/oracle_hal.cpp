// ORACLE — test infrastructure only. A CPU restatement of the reference's `hal::cpu` backend
// (cyclosubfield-dbfv/src/hal/{mont,ntt,cpu}.rs), used by tests/ as the parity checker and by
// bench.py as the CPU baseline. Nothing in the product links or calls this file.
//
// Every function below follows the cited reference lines formula for formula, including the
// wrapping 64/128-bit arithmetic of Rust release builds (Cargo.toml:26-31 has no
// overflow-checks), so that outputs are representative-exact (bit-exact i64/i128) with the
// reference CpuBackend on the same inputs.
//
// Parity status (what pins each part; the reference Rust cannot be built here, so L0 agreement
// with the Rust original rests on this transcription plus the pins listed):
//   * Montgomery (mont.rs)           — tests/test_oracle_hal.py::test_mont_ops (reference test
//                                      cpu.rs:1456-1581 closed forms) + N*(-N^-1) = -1 mod 2^64.
//   * BarrettMul (ntt.rs:123-158)     — test_barrett_mul (ntt.rs:696-717 inputs and checks).
//   * LazyFactorNTT / plan            — test_forward_is_permuted_dft (naive DFT), roundtrips at
//                                      extreme +-2q inputs, plan convolution theorem vs brute-force
//                                      group convolution, strided == contiguous (exact),
//                                      fused ntt_mul == ntt + mul_mont (exact).
//   * element-wise / contraction ops — exact-integer tests mirroring cpu.rs:1141-1875.
//   * permutation                    — cpu.rs:1829-1875 closed-form shifts.
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <stdexcept>
#include <string>

typedef __int128 i128;
typedef unsigned __int128 u128;

extern "C" {
// A strided view over host memory, strides counted in elements of the tensor's type (i64 or
// i128); the last stride must be 1 (cpu.rs:3-6).
typedef struct {
    void* ptr;
    int64_t shape[4];
    int64_t strides[4];
} orc_view;

// MontgomeryModulus (mont.rs:4-10)
typedef struct {
    int64_t N;
    int64_t neg_inv_N;
    int64_t r2;
    int64_t r3;
} orc_modulus;
}

static thread_local std::string g_err;

#define ORC_TRY try {
#define ORC_CATCH                                  \
    }                                              \
    catch (const std::exception& e) {              \
        g_err = e.what();                          \
        return 1;                                  \
    }                                              \
    return 0;

static void require(bool cond, const char* msg) {
    if (!cond) throw std::runtime_error(msg);
}

// ---------------------------------------------------------------------------------------------
// wrapping helpers (Rust release semantics)
static inline i128 wadd128(i128 a, i128 b) { return (i128)((u128)a + (u128)b); }
static inline i128 wsub128(i128 a, i128 b) { return (i128)((u128)a - (u128)b); }
static inline i128 wmul128(i128 a, i128 b) { return (i128)((u128)a * (u128)b); }
static inline int64_t wadd64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
static inline int64_t wsub64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
static inline int64_t wmul64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }

// ---------------------------------------------------------------------------------------------
// hal/mont.rs

// MontgomeryModulus::new (mont.rs:13-33). -N^-1 mod 2^64 is what `extended_ideal_gen` yields
// after truncation to 64 bits (mont.rs:29); it is computed here by Newton iteration, which
// gives the same residue mod 2^64. `pmod` keeps the reference's quirk of mapping 0 to N.
static orc_modulus mont_new(int64_t N) {
    require(N > 1, "modulus must be > 1");
    require(N % 2 == 1, "modulus must be odd");
    uint64_t inv = (uint64_t)N;  // correct to 3 bits for odd N
    for (int i = 0; i < 6; i++) inv *= 2 - (uint64_t)N * inv;
    auto pmod = [N](i128 x) -> int64_t {
        i128 res = (x % (i128)N) + (i128)N;
        if (res > (i128)N) res -= N;
        return (int64_t)res;
    };
    orc_modulus m;
    m.N = N;
    m.neg_inv_N = (int64_t)(0 - inv);
    i128 r_mod_n = pmod((i128)1 << 64);
    m.r2 = pmod(r_mod_n * r_mod_n);
    m.r3 = pmod((i128)m.r2 * (i128)pmod((i128)1 << 64));
    return m;
}

// red128 (mont.rs:38-45)
static inline int64_t red128(const orc_modulus& m, i128 x) {
    int64_t mm = wmul64((int64_t)x, m.neg_inv_N);
    i128 t = (i128)mm * (i128)m.N;
    return (int64_t)(wadd128(x, t) >> 64);
}

// double_red128 (mont.rs:57-66)
static inline int64_t double_red128(const orc_modulus& m, i128 x) {
    int64_t mm = wmul64((int64_t)x, m.neg_inv_N);
    i128 t = (i128)mm * (i128)m.N;
    i128 carry = (i128)((((u128)(uint64_t)x) + ((u128)(uint64_t)t)) >> 64);
    i128 r1 = wadd128(wadd128(x >> 64, t >> 64), carry);
    return red128(m, r1);
}

// to_mont64 / to_mont128 (mont.rs:68-76)
static inline int64_t to_mont64(const orc_modulus& m, int64_t x) { return red128(m, (i128)x * (i128)m.r2); }
static inline int64_t to_mont128(const orc_modulus& m, i128 x) {
    return red128(m, (i128)red128(m, x) * (i128)m.r3);
}

// ---------------------------------------------------------------------------------------------
// hal/ntt.rs

// mul_mod / pow_mod (ntt.rs:87-100); operands are non-negative where they are used
static int64_t mul_mod(int64_t a, int64_t b, int64_t q) { return (int64_t)(((i128)a * (i128)b) % (i128)q); }
static int64_t pow_mod(int64_t b, uint64_t e, int64_t q) {
    int64_t result = 1;
    b %= q;
    while (e > 0) {
        if (e & 1) result = mul_mod(result, b, q);
        b = mul_mod(b, b, q);
        e >>= 1;
    }
    return result;
}

static bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

// find_root_of_unity (ntt.rs:104-115)
static int64_t find_root_of_unity(int64_t q, size_t n) {
    require(n > 1, "root of unity order must be > 1");
    require((q - 1) % (int64_t)n == 0, "modulus doesn't support roots of unity of this order");
    size_t p = is_pow2(n) ? 2 : 3;
    for (int64_t cand = 2;; cand++) {
        int64_t root = pow_mod(cand, (uint64_t)((q - 1) / (int64_t)n), q);
        if (pow_mod(root, (uint64_t)(n / p), q) != 1) return root;
    }
}

// rounded_div, half away from zero (ntt.rs:118-121)
static i128 rounded_div(i128 a, i128 b) { return a >= 0 ? (a + b / 2) / b : -((-a + b / 2) / b); }

// BarrettMul (ntt.rs:126-158)
struct Barrett {
    int64_t z, u;
};
static Barrett barrett_new(int64_t value, int64_t n) {
    require(n > 1 && n % 2 == 1 && n < ((int64_t)1 << 62), "invalid Barrett modulus");
    int64_t z = value % n;
    if (2 * z > n)
        z -= n;
    else if (2 * z < -n)
        z += n;
    i128 u = rounded_div(-((i128)z << 64), (i128)n);
    require(u >= (i128)INT64_MIN && u <= (i128)INT64_MAX, "Barrett constant out of range");
    return Barrett{z, (int64_t)u};
}
// Schedule trace of the lazy NTTs (test instrumentation, off by default): while g_trace is set,
// the first call of each factor in an apply records its reduce passes and per-stage decisions,
// and every Barrett multiplication of the apply is counted.
struct TraceEvent {
    int32_t kind;     // 0 reduce pass, 1 forward stage, 2 inverse stage
    int32_t factor;   // factor index within the plan
    int32_t stage;    // stage index (the tree level; inverse stages count down)
    int32_t skip_ok;  // forward: the twiddle-1 skip was allowed in this stage
    int32_t skipped;  // number of blocks whose Barrett multiplication was skipped
    int32_t fused;    // forward: reduce fused into the stage; inverse: len^-1 fused into the stage
};
static bool g_trace = false;
static bool g_trace_rec = false;
static int32_t g_trace_factor = 0;
static uint64_t g_trace_mults = 0;
static std::vector<TraceEvent> g_trace_events;
static inline void trace_event(int32_t kind, int32_t stage, int32_t skip_ok, int32_t skipped, int32_t fused) {
    if (g_trace_rec) g_trace_events.push_back(TraceEvent{kind, g_trace_factor, stage, skip_ok, skipped, fused});
}

static inline int64_t barrett_mul(Barrett b, int64_t x, int64_t n) {
    if (__builtin_expect(g_trace, 0)) g_trace_mults++;
    int64_t q = (int64_t)(wadd128((i128)b.u * (i128)x, (i128)1 << 63) >> 64);
    return wadd64(wmul64(b.z, x), wmul64(q, n));
}

static const u128 I64MAX = (u128)INT64_MAX;

// reduce_pass (ntt.rs:164-173)
static void reduce_pass(Barrett one, int64_t n, int64_t* data, size_t es, size_t len, size_t vl) {
    trace_event(0, -1, 0, 0, 0);
    for (size_t i = 0; i < len; i++)
        for (size_t l = 0; l < vl; l++) data[i * es + l] = barrett_mul(one, data[i * es + l], n);
}
static inline int64_t reduce_if(bool reduce, Barrett m, int64_t n, int64_t x) {
    return reduce ? barrett_mul(m, x, n) : x;
}

// LazyFactorNTT (ntt.rs:189-538)
struct FactorNTT {
    bool radix2;
    int64_t n;
    Barrett one;
    size_t len;
    std::vector<std::vector<Barrett>> tw, itw;
    Barrett zeta, zeta2;

    FactorNTT(int64_t n_, size_t len_) : n(n_), len(len_) {
        if (is_pow2(len)) {
            radix2 = true;
            require(len >= 2, "NTT length must be >= 2");
            require(2 * (u128)n <= I64MAX, "modulus too large for lazy radix-2 butterflies");
            one = barrett_new(1, n);
            int64_t root = find_root_of_unity(n, len);
            int64_t inv_root = pow_mod(root, len - 1, n);
            std::vector<size_t> exps{0};
            while (exps.size() < len) {
                std::vector<Barrett> s, is;
                for (size_t e : exps) {
                    s.push_back(barrett_new(pow_mod(root, e / 2, n), n));
                    is.push_back(barrett_new(pow_mod(inv_root, e / 2, n), n));
                }
                tw.push_back(s);
                itw.push_back(is);
                std::vector<size_t> next;
                for (size_t e : exps) {
                    next.push_back(e / 2);
                    next.push_back(e / 2 + len / 2);
                }
                exps = next;
            }
        } else {
            radix2 = false;
            size_t l = len;
            while (l > 1) {
                require(l % 3 == 0, "NTT length must be a power of 2 or a power of 3");
                l /= 3;
            }
            require(len >= 3, "NTT length must be >= 3");
            require(3 * (u128)n <= I64MAX, "modulus too large for lazy radix-3 butterflies");
            one = barrett_new(1, n);
            int64_t root = find_root_of_unity(n, len);
            int64_t inv_root = pow_mod(root, len - 1, n);
            int64_t zp = pow_mod(root, len / 3, n);
            int64_t zp2 = mul_mod(zp, zp, n);
            std::vector<size_t> exps{0};
            while (exps.size() < len) {
                std::vector<Barrett> s, is;
                for (size_t e : exps) {
                    int64_t v = pow_mod(root, e / 3, n);
                    int64_t v2 = mul_mod(v, v, n);
                    s.push_back(barrett_new(v, n));
                    s.push_back(barrett_new(mul_mod(zp, v, n), n));
                    s.push_back(barrett_new(mul_mod(zp2, v, n), n));
                    s.push_back(barrett_new(v2, n));
                    s.push_back(barrett_new(mul_mod(zp, v2, n), n));
                    s.push_back(barrett_new(mul_mod(zp2, v2, n), n));
                    int64_t vi = pow_mod(inv_root, e / 3, n);
                    is.push_back(barrett_new(vi, n));
                    is.push_back(barrett_new(mul_mod(vi, vi, n), n));
                }
                tw.push_back(s);
                itw.push_back(is);
                std::vector<size_t> next;
                for (size_t e : exps) {
                    next.push_back(e / 3);
                    next.push_back(e / 3 + len / 3);
                    next.push_back(e / 3 + 2 * (len / 3));
                }
                exps = next;
            }
            zeta = barrett_new(zp, n);
            zeta2 = barrett_new(zp2, n);
        }
    }

    // forward (ntt.rs:301-414)
    u128 forward(int64_t* data, size_t es, size_t vl, u128 in_bound, bool reduce_output) const {
        u128 bound = in_bound;
        size_t stage_len = len;
        size_t nst = tw.size();
        if (radix2) {
            for (size_t s = 0; s < nst; s++) {
                bool fuse = reduce_output && s + 1 == nst;
                if (bound + (u128)n > I64MAX) {
                    reduce_pass(one, n, data, es, len, vl);
                    bound = (u128)n;
                }
                bool skip_ok = 2 * bound <= I64MAX;
                u128 nb = fuse ? (u128)n : (skip_ok ? 2 * bound : bound + (u128)n);
                size_t half = stage_len / 2;
                if (g_trace_rec) {
                    int32_t sk = 0;
                    for (size_t block = 0; block * stage_len < len; block++) sk += (skip_ok && tw[s][block].z == 1);
                    trace_event(1, (int32_t)s, skip_ok, sk, fuse);
                }
                for (size_t block = 0, bs = 0; bs < len; block++, bs += stage_len) {
                    Barrett v = tw[s][block];
                    bool skip = skip_ok && v.z == 1;
                    for (size_t j = 0; j < half; j++) {
                        int64_t* x = data + (bs + j) * es;
                        int64_t* y = data + (bs + j + half) * es;
                        for (size_t l = 0; l < vl; l++) {
                            int64_t a = x[l];
                            int64_t t = skip ? y[l] : barrett_mul(v, y[l], n);
                            x[l] = reduce_if(fuse, one, n, wadd64(a, t));
                            y[l] = reduce_if(fuse, one, n, wsub64(a, t));
                        }
                    }
                }
                bound = nb;
                stage_len = half;
            }
        } else {
            for (size_t s = 0; s < nst; s++) {
                bool fuse = reduce_output && s + 1 == nst;
                if (bound + 2 * (u128)n > I64MAX) {
                    reduce_pass(one, n, data, es, len, vl);
                    bound = (u128)n;
                }
                bool skip_ok = 3 * bound <= I64MAX;
                u128 nb = fuse ? (u128)n : (skip_ok ? 3 * bound : bound + 2 * (u128)n);
                size_t third = stage_len / 3;
                if (g_trace_rec) {
                    int32_t sk = 0;
                    for (size_t block = 0; block * stage_len < len; block++) sk += (skip_ok && tw[s][6 * block].z == 1);
                    trace_event(1, (int32_t)s, skip_ok, sk, fuse);
                }
                for (size_t block = 0, bs = 0; bs < len; block++, bs += stage_len) {
                    const Barrett* w = &tw[s][6 * block];
                    bool skip = skip_ok && w[0].z == 1;
                    for (size_t j = 0; j < third; j++) {
                        int64_t* x = data + (bs + j) * es;
                        int64_t* y = data + (bs + j + third) * es;
                        int64_t* z = data + (bs + j + 2 * third) * es;
                        for (size_t l = 0; l < vl; l++) {
                            int64_t a = x[l], b = y[l], c = z[l];
                            int64_t tb = skip ? b : barrett_mul(w[0], b, n);
                            int64_t tbz = barrett_mul(w[1], b, n);
                            int64_t tbzz = barrett_mul(w[2], b, n);
                            int64_t tc = skip ? c : barrett_mul(w[3], c, n);
                            int64_t tcz = barrett_mul(w[4], c, n);
                            int64_t tczz = barrett_mul(w[5], c, n);
                            x[l] = reduce_if(fuse, one, n, wadd64(wadd64(a, tb), tc));
                            y[l] = reduce_if(fuse, one, n, wadd64(wadd64(a, tbz), tczz));
                            z[l] = reduce_if(fuse, one, n, wadd64(wadd64(a, tbzz), tcz));
                        }
                    }
                }
                bound = nb;
                stage_len = third;
            }
        }
        return bound;
    }

    // backward (ntt.rs:427-537)
    u128 backward(int64_t* data, size_t es, size_t vl, u128 in_bound, const Barrett* final_scale) const {
        u128 bound = in_bound;
        size_t nst = itw.size();
        if (radix2) {
            for (size_t si = nst; si-- > 0;) {
                const Barrett* fs = si == 0 ? final_scale : nullptr;
                if (2 * bound > I64MAX) {
                    reduce_pass(one, n, data, es, len, vl);
                    bound = (u128)n;
                }
                bound = fs ? (u128)n : 2 * bound;
                size_t stage_len = len >> si;
                size_t half = stage_len / 2;
                if (g_trace_rec) {
                    int32_t sk = 0;
                    for (size_t block = 0; !fs && block * stage_len < len; block++) sk += (itw[si][block].z == 1);
                    trace_event(2, (int32_t)si, 1, sk, fs != nullptr);
                }
                for (size_t block = 0, bs = 0; bs < len; block++, bs += stage_len) {
                    bool fuse = fs != nullptr;
                    Barrett ym = fuse ? *fs : itw[si][block];
                    bool skip = !fuse && ym.z == 1;
                    for (size_t j = 0; j < half; j++) {
                        int64_t* x = data + (bs + j) * es;
                        int64_t* y = data + (bs + j + half) * es;
                        for (size_t l = 0; l < vl; l++) {
                            int64_t y0 = x[l], y1 = y[l];
                            x[l] = reduce_if(fuse, ym, n, wadd64(y0, y1));
                            y[l] = skip ? wsub64(y0, y1) : barrett_mul(ym, wsub64(y0, y1), n);
                        }
                    }
                }
            }
        } else {
            for (size_t si = nst; si-- > 0;) {
                const Barrett* fs = si == 0 ? final_scale : nullptr;
                if (3 * bound > I64MAX) {
                    reduce_pass(one, n, data, es, len, vl);
                    bound = (u128)n;
                }
                bound = fs ? (u128)n : 3 * bound;
                size_t stage_len = len;
                for (size_t t = 0; t < si; t++) stage_len /= 3;
                size_t third = stage_len / 3;
                if (g_trace_rec) {
                    int32_t sk = 0;
                    for (size_t block = 0; !fs && block * stage_len < len; block++) sk += (itw[si][2 * block].z == 1);
                    trace_event(2, (int32_t)si, 1, sk, fs != nullptr);
                }
                for (size_t block = 0, bs = 0; bs < len; block++, bs += stage_len) {
                    bool fuse = fs != nullptr;
                    Barrett ym = fuse ? *fs : itw[si][2 * block];
                    Barrett zm = fuse ? *fs : itw[si][2 * block + 1];
                    bool skip = !fuse && ym.z == 1;
                    for (size_t j = 0; j < third; j++) {
                        int64_t* x = data + (bs + j) * es;
                        int64_t* y = data + (bs + j + third) * es;
                        int64_t* z = data + (bs + j + 2 * third) * es;
                        for (size_t l = 0; l < vl; l++) {
                            int64_t y0 = x[l], y1 = y[l], y2 = z[l];
                            int64_t y1z = barrett_mul(zeta, y1, n);
                            int64_t y1zz = barrett_mul(zeta2, y1, n);
                            int64_t y2z = barrett_mul(zeta, y2, n);
                            int64_t y2zz = barrett_mul(zeta2, y2, n);
                            x[l] = reduce_if(fuse, ym, n, wadd64(wadd64(y0, y1), y2));
                            int64_t b = wadd64(wadd64(y0, y1zz), y2z);
                            int64_t c = wadd64(wadd64(y0, y1z), y2zz);
                            y[l] = skip ? b : barrett_mul(ym, b, n);
                            z[l] = skip ? c : barrett_mul(zm, c, n);
                        }
                    }
                }
            }
        }
        return bound;
    }
};

// CpuNTTPlan (ntt.rs:549-658)
struct OrcNTTPlan {
    int64_t n;
    std::vector<FactorNTT> factors;
    size_t len;
    Barrett inv_len, one;

    OrcNTTPlan(int64_t n_, const size_t* shape, size_t nshape) : n(n_) {
        len = 1;
        for (size_t i = 0; i < nshape; i++) {
            len *= shape[i];
            if (shape[i] > 1) factors.emplace_back(n, shape[i]);
        }
        inv_len = barrett_new(pow_mod((int64_t)len, (uint64_t)(n - 2), n), n);
        one = barrett_new(1, n);
    }

    // apply (ntt.rs:581-657)
    void apply(int64_t* data, size_t es, size_t vl, bool inverse, bool reduce_output) const {
        u128 bound = 2 * (u128)n;
        size_t suffix_len = 1;
        for (auto& f : factors) suffix_len *= f.len;
        size_t prefix_len = 1;
        for (size_t fi = 0; fi < factors.size(); fi++) {
            const FactorNTT& f = factors[fi];
            bool last = fi + 1 == factors.size();
            const Barrett* fs = last ? &inv_len : nullptr;
            bool fuse_fwd = last && reduce_output;
            suffix_len /= f.len;
            u128 nb = bound;
            g_trace_factor = (int32_t)fi;
            if (vl == es) {
                size_t inner = suffix_len * es;
                for (size_t p = 0; p < prefix_len; p++) {
                    g_trace_rec = g_trace && p == 0;
                    int64_t* block = data + p * f.len * inner;
                    nb = inverse ? f.backward(block, inner, inner, bound, fs) : f.forward(block, inner, inner, bound, fuse_fwd);
                }
            } else {
                for (size_t p = 0; p < prefix_len; p++)
                    for (size_t s = 0; s < suffix_len; s++) {
                        g_trace_rec = g_trace && p == 0 && s == 0;
                        size_t off = (p * f.len * suffix_len + s) * es;
                        nb = inverse ? f.backward(data + off, suffix_len * es, vl, bound, fs)
                                     : f.forward(data + off, suffix_len * es, vl, bound, fuse_fwd);
                    }
            }
            g_trace_rec = false;
            prefix_len *= f.len;
            bound = nb;
        }
        if (factors.empty() && (inverse || reduce_output)) {
            Barrett fm = inverse ? inv_len : one;
            for (size_t i = 0; i < len; i++)
                for (size_t l = 0; l < vl; l++) data[i * es + l] = barrett_mul(fm, data[i * es + l], n);
        }
    }
};

// ---------------------------------------------------------------------------------------------
// hal/cpu.rs — strided views with broadcasting (cpu.rs:333-474)

template <typename T>
struct View {
    T* p;
    int64_t sh[4];
    int64_t st[4];
    View(const orc_view* v) {
        p = (T*)v->ptr;
        for (int d = 0; d < 4; d++) {
            sh[d] = v->shape[d];
            st[d] = v->strides[d];
        }
        require(sh[3] <= 1 || st[3] == 1, "last stride must be 1");
    }
    inline T& at(int64_t i, int64_t j, int64_t k, int64_t l) const {
        return p[i * st[0] + j * st[1] + k * st[2] + l * st[3]];
    }
};

// broadcast_strides (cpu.rs:333-346)
template <typename T>
static View<T> bcast(const orc_view* v, const int64_t* dst_shape) {
    View<T> r(v);
    for (int d = 0; d < 4; d++) {
        if (r.sh[d] != dst_shape[d]) {
            require(r.sh[d] == 1, "cannot broadcast dimension");
            r.st[d] = 0;
        }
    }
    return r;
}

template <typename D, typename S, typename F>
static void each2(const orc_view* dv, const orc_view* sv, F f) {
    View<D> d(dv);
    View<S> s = bcast<S>(sv, d.sh);
    for (int64_t i = 0; i < d.sh[0]; i++)
        for (int64_t j = 0; j < d.sh[1]; j++)
            for (int64_t k = 0; k < d.sh[2]; k++)
                for (int64_t l = 0; l < d.sh[3]; l++) f(j, d.at(i, j, k, l), s.at(i, j, k, l));
}

template <typename D, typename A, typename B, typename F>
static void each3(const orc_view* dv, const orc_view* av, const orc_view* bv, F f) {
    View<D> d(dv);
    View<A> a = bcast<A>(av, d.sh);
    View<B> b = bcast<B>(bv, d.sh);
    for (int64_t i = 0; i < d.sh[0]; i++)
        for (int64_t j = 0; j < d.sh[1]; j++)
            for (int64_t k = 0; k < d.sh[2]; k++)
                for (int64_t l = 0; l < d.sh[3]; l++) f(d.at(i, j, k, l), a.at(i, j, k, l), b.at(i, j, k, l));
}

static void check_mods(const orc_view* dst, size_t nmods) {
    require((int64_t)nmods == dst->shape[1], "number of moduli must equal dimension 1");
}

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_create_modulus(int64_t prime, orc_modulus* out) {
    ORC_TRY* out = mont_new(prime);
    ORC_CATCH
}

// elementary Montgomery / Barrett entry points, used by the unit tests of a1/a2
int orc_red128(const orc_modulus* m, int64_t lo, int64_t hi, int64_t* out) {
    *out = red128(*m, (i128)(((u128)(uint64_t)hi << 64) | (uint64_t)lo));
    return 0;
}
int orc_double_red128(const orc_modulus* m, int64_t lo, int64_t hi, int64_t* out) {
    *out = double_red128(*m, (i128)(((u128)(uint64_t)hi << 64) | (uint64_t)lo));
    return 0;
}
int orc_barrett_new(int64_t value, int64_t n, int64_t* z, int64_t* u) {
    ORC_TRY Barrett b = barrett_new(value, n);
    *z = b.z;
    *u = b.u;
    ORC_CATCH
}
int orc_barrett_mul(int64_t z, int64_t u, int64_t x, int64_t n, int64_t* out) {
    *out = barrett_mul(Barrett{z, u}, x, n);
    return 0;
}
int orc_find_root_of_unity(int64_t q, size_t n, int64_t* out) {
    ORC_TRY* out = find_root_of_unity(q, n);
    ORC_CATCH
}

// LazyFactorNTT on contiguous/strided data, as the ntt.rs unit tests call it
int orc_factor_forward(int64_t q, size_t len, int64_t* data, size_t es, size_t vl, uint64_t in_bound_lo,
                       int reduce_output, uint64_t* out_bound) {
    ORC_TRY FactorNTT f(q, len);
    *out_bound = (uint64_t)f.forward(data, es, vl, (u128)in_bound_lo, reduce_output != 0);
    ORC_CATCH
}
int orc_factor_backward(int64_t q, size_t len, int64_t* data, size_t es, size_t vl, uint64_t in_bound_lo,
                        uint64_t* out_bound) {
    ORC_TRY FactorNTT f(q, len);
    *out_bound = (uint64_t)f.backward(data, es, vl, (u128)in_bound_lo, nullptr);
    ORC_CATCH
}

int orc_create_ntt_plan(const orc_modulus* m, const size_t* inner_shape, size_t n, void** out) {
    ORC_TRY* out = new OrcNTTPlan(m->N, inner_shape, n);
    ORC_CATCH
}
int orc_free_ntt_plan(void* plan) {
    delete (OrcNTTPlan*)plan;
    return 0;
}
int orc_plan_apply(void* plan, int64_t* data, size_t es, size_t vl, int inverse, int reduce_output) {
    ORC_TRY((OrcNTTPlan*)plan)->apply(data, es, vl, inverse != 0, reduce_output != 0);
    ORC_CATCH
}

// Schedule trace (test instrumentation): orc_ntt_trace(1) clears and enables it, orc_ntt_trace(0)
// disables it; orc_ntt_trace_get copies the recorded events (6 int32 each) and the number of
// Barrett multiplications since it was enabled.
int orc_ntt_trace(int on) {
    ORC_TRY g_trace = on != 0;
    g_trace_rec = false;
    g_trace_events.clear();
    g_trace_mults = 0;
    ORC_CATCH
}

int orc_ntt_trace_get(int32_t* out, size_t cap, size_t* count, uint64_t* mults) {
    ORC_TRY* count = g_trace_events.size();
    *mults = g_trace_mults;
    require(g_trace_events.size() <= cap, "trace buffer too small");
    for (size_t i = 0; i < g_trace_events.size(); i++) {
        const TraceEvent& e = g_trace_events[i];
        int32_t v[6] = {e.kind, e.factor, e.stage, e.skip_ok, e.skipped, e.fused};
        for (int k = 0; k < 6; k++) out[6 * i + k] = v[k];
    }
    ORC_CATCH
}

// hal_import / hal_export (cpu.rs:566-589): row-major order of the view's lines
int orc_hal_import(const orc_view* dst, const int64_t* src, size_t len) {
    ORC_TRY View<int64_t> d(dst);
    require((int64_t)len == d.sh[0] * d.sh[1] * d.sh[2] * d.sh[3], "import length mismatch");
    size_t pos = 0;
    for (int64_t i = 0; i < d.sh[0]; i++)
        for (int64_t j = 0; j < d.sh[1]; j++)
            for (int64_t k = 0; k < d.sh[2]; k++)
                for (int64_t l = 0; l < d.sh[3]; l++) d.at(i, j, k, l) = src[pos++];
    ORC_CATCH
}
int orc_hal_export(const orc_view* src, int64_t* dst, size_t len) {
    ORC_TRY View<int64_t> s(src);
    require((int64_t)len == s.sh[0] * s.sh[1] * s.sh[2] * s.sh[3], "export length mismatch");
    size_t pos = 0;
    for (int64_t i = 0; i < s.sh[0]; i++)
        for (int64_t j = 0; j < s.sh[1]; j++)
            for (int64_t k = 0; k < s.sh[2]; k++)
                for (int64_t l = 0; l < s.sh[3]; l++) dst[pos++] = s.at(i, j, k, l);
    ORC_CATCH
}

// hal_assign / assign_narrow / assign_i128 (cpu.rs:591-631)
int orc_hal_assign(const orc_view* dst, const orc_view* src) {
    ORC_TRY each2<int64_t, int64_t>(dst, src, [](int64_t, int64_t& d, int64_t s) { d = s; });
    ORC_CATCH
}
int orc_hal_assign_narrow(const orc_view* dst, const orc_view* src) {
    ORC_TRY each2<int64_t, i128>(dst, src, [](int64_t, int64_t& d, i128 s) { d = (int64_t)s; });
    ORC_CATCH
}
int orc_hal_assign_i128(const orc_view* dst, const orc_view* src) {
    ORC_TRY each2<i128, i128>(dst, src, [](int64_t, i128& d, i128 s) { d = s; });
    ORC_CATCH
}

// add/sub (cpu.rs:633-670)
int orc_hal_add_i64(const orc_view* dst, const orc_view* src) {
    ORC_TRY each2<int64_t, int64_t>(dst, src, [](int64_t, int64_t& d, int64_t s) { d = wadd64(d, s); });
    ORC_CATCH
}
int orc_hal_sub_i64(const orc_view* dst, const orc_view* src) {
    ORC_TRY each2<int64_t, int64_t>(dst, src, [](int64_t, int64_t& d, int64_t s) { d = wsub64(d, s); });
    ORC_CATCH
}
int orc_hal_add_i128(const orc_view* dst, const orc_view* src) {
    ORC_TRY each2<i128, i128>(dst, src, [](int64_t, i128& d, i128 s) { d = wadd128(d, s); });
    ORC_CATCH
}

// fma / mul (cpu.rs:672-704)
int orc_hal_fma_i64(const orc_view* dst, const orc_view* lhs, const orc_view* rhs) {
    ORC_TRY each3<i128, int64_t, int64_t>(dst, lhs, rhs,
                                          [](i128& d, int64_t a, int64_t b) { d = wadd128(d, (i128)a * (i128)b); });
    ORC_CATCH
}
int orc_hal_mul_i64(const orc_view* dst, const orc_view* lhs, const orc_view* rhs) {
    ORC_TRY each3<i128, int64_t, int64_t>(dst, lhs, rhs, [](i128& d, int64_t a, int64_t b) { d = (i128)a * (i128)b; });
    ORC_CATCH
}

// inner_product (cpu.rs:706-788): term 0 overwrites, later terms wrapping-add; the tiling of
// the reference does not change the result (wrapping addition is associative)
int orc_hal_inner_product_i64(const orc_view* dst, const orc_view* lhs, const orc_view* rhs, size_t terms) {
    ORC_TRY require(terms >= 1, "inner_product_i64 requires at least one term");
    View<i128> d(dst);
    std::vector<View<int64_t>> ls, rs;
    for (size_t t = 0; t < terms; t++) {
        ls.push_back(bcast<int64_t>(&lhs[t], d.sh));
        rs.push_back(bcast<int64_t>(&rhs[t], d.sh));
    }
    for (int64_t i = 0; i < d.sh[0]; i++)
        for (int64_t j = 0; j < d.sh[1]; j++)
            for (int64_t k = 0; k < d.sh[2]; k++)
                for (int64_t l = 0; l < d.sh[3]; l++) {
                    i128 acc = (i128)ls[0].at(i, j, k, l) * (i128)rs[0].at(i, j, k, l);
                    for (size_t t = 1; t < terms; t++)
                        acc = wadd128(acc, (i128)ls[t].at(i, j, k, l) * (i128)rs[t].at(i, j, k, l));
                    d.at(i, j, k, l) = acc;
                }
    ORC_CATCH
}

// matmul (cpu.rs:790-818)
int orc_hal_matmul_i64(const orc_view* dst, const orc_view* lhs, const orc_view* rhs) {
    ORC_TRY View<i128> d(dst);
    View<int64_t> a(lhs), b(rhs);
    int64_t I = d.sh[0], J = d.sh[1], K = d.sh[2], L = d.sh[3], U = a.sh[2];
    require(a.sh[0] == I || a.sh[0] == 1, "matmul lhs dim 0");
    require(a.sh[1] == J, "matmul lhs dim 1");
    require(a.sh[3] == 1, "matmul lhs dim 3 must be 1");
    require(b.sh[0] == I || b.sh[0] == 1, "matmul rhs dim 0");
    require(b.sh[1] == U && b.sh[2] == K && b.sh[3] == L, "matmul rhs shape");
    for (int64_t i = 0; i < I; i++) {
        int64_t li = a.sh[0] == 1 ? 0 : i, ri = b.sh[0] == 1 ? 0 : i;
        for (int64_t j = 0; j < J; j++)
            for (int64_t k = 0; k < K; k++) {
                for (int64_t l = 0; l < L; l++) d.at(i, j, k, l) = 0;
                for (int64_t u = 0; u < U; u++) {
                    i128 c = a.at(li, j, u, 0);
                    if (c == 0) continue;
                    for (int64_t l = 0; l < L; l++) d.at(i, j, k, l) = wadd128(d.at(i, j, k, l), c * (i128)b.at(ri, u, k, l));
                }
            }
    }
    ORC_CATCH
}

// Montgomery element-wise ops (cpu.rs:536-550, 823-916)
int orc_hal_mul_mont64(const orc_view* dst, const orc_view* src, const orc_modulus* mods, size_t nmods) {
    ORC_TRY check_mods(dst, nmods);
    each2<int64_t, int64_t>(dst, src, [mods](int64_t j, int64_t& d, int64_t s) { d = red128(mods[j], (i128)d * (i128)s); });
    ORC_CATCH
}
int orc_hal_add_red_mont64_double(const orc_view* dst, const orc_view* src, const orc_modulus* mods, size_t nmods) {
    ORC_TRY check_mods(dst, nmods);
    each2<int64_t, i128>(dst, src, [mods](int64_t j, int64_t& d, i128 s) { d = wadd64(d, double_red128(mods[j], s)); });
    ORC_CATCH
}
int orc_hal_red_mont64_double(const orc_view* dst, const orc_view* src, const orc_modulus* mods, size_t nmods) {
    ORC_TRY check_mods(dst, nmods);
    each2<int64_t, i128>(dst, src, [mods](int64_t j, int64_t& d, i128 s) { d = double_red128(mods[j], s); });
    ORC_CATCH
}
int orc_hal_add_red_mont64(const orc_view* dst, const orc_view* src, const orc_modulus* mods, size_t nmods) {
    ORC_TRY check_mods(dst, nmods);
    each2<int64_t, i128>(dst, src, [mods](int64_t j, int64_t& d, i128 s) { d = wadd64(d, red128(mods[j], s)); });
    ORC_CATCH
}
int orc_hal_red_mont64(const orc_view* dst, const orc_view* src, const orc_modulus* mods, size_t nmods) {
    ORC_TRY check_mods(dst, nmods);
    each2<int64_t, i128>(dst, src, [mods](int64_t j, int64_t& d, i128 s) { d = red128(mods[j], s); });
    ORC_CATCH
}
int orc_hal_i128_to_mont64(const orc_view* dst, const orc_view* src, const orc_modulus* mods, size_t nmods) {
    ORC_TRY check_mods(dst, nmods);
    each2<int64_t, i128>(dst, src, [mods](int64_t j, int64_t& d, i128 s) { d = to_mont128(mods[j], s); });
    ORC_CATCH
}
int orc_hal_i64_to_mont64(const orc_view* dst, const orc_modulus* mods, size_t nmods) {
    ORC_TRY check_mods(dst, nmods);
    each2<int64_t, int64_t>(dst, dst, [mods](int64_t j, int64_t& d, int64_t) { d = to_mont64(mods[j], d); });
    ORC_CATCH
}

// ntt_dim_impl (cpu.rs:494-534)
static void ntt_dim(const orc_view* dst, void* const* plans, size_t nplans, int dim, bool inverse, bool reduce_output) {
    View<int64_t> d(dst);
    require((int64_t)nplans == d.sh[1], "number of NTT plans must equal dimension 1");
    if (dim == 2) {
        for (size_t j = 0; j < nplans; j++) require((int64_t)((OrcNTTPlan*)plans[j])->len == d.sh[2], "NTT length mismatch");
        for (int64_t i = 0; i < d.sh[0]; i++)
            for (int64_t j = 0; j < d.sh[1]; j++)
                ((OrcNTTPlan*)plans[j])->apply(d.p + i * d.st[0] + j * d.st[1], d.st[2], d.sh[3], inverse, reduce_output);
    } else {
        for (size_t j = 0; j < nplans; j++) require((int64_t)((OrcNTTPlan*)plans[j])->len == d.sh[3], "NTT length mismatch");
        for (int64_t i = 0; i < d.sh[0]; i++)
            for (int64_t j = 0; j < d.sh[1]; j++)
                for (int64_t k = 0; k < d.sh[2]; k++)
                    ((OrcNTTPlan*)plans[j])->apply(d.p + i * d.st[0] + j * d.st[1] + k * d.st[2], 1, 1, inverse, reduce_output);
    }
}

int orc_hal_ntt_mont64_dim(const orc_view* dst, void* const* plans, size_t nplans, int dim) {
    ORC_TRY ntt_dim(dst, plans, nplans, dim, false, true);
    ORC_CATCH
}
int orc_hal_intt_mont64_dim(const orc_view* dst, void* const* plans, size_t nplans, int dim) {
    ORC_TRY ntt_dim(dst, plans, nplans, dim, true, true);
    ORC_CATCH
}
// hal_ntt_mul_mont64_dim (cpu.rs:933-948)
int orc_hal_ntt_mul_mont64_dim(const orc_view* dst, void* const* plans, size_t nplans, int dim, const orc_view* mul,
                               const orc_modulus* mods, size_t nmods) {
    ORC_TRY ntt_dim(dst, plans, nplans, dim, false, false);
    check_mods(dst, nmods);
    each2<int64_t, int64_t>(dst, mul, [mods](int64_t j, int64_t& d, int64_t s) { d = red128(mods[j], (i128)d * (i128)s); });
    ORC_CATCH
}

// bconv_correct_mont64 (cpu.rs:950-980)
int orc_hal_bconv_correct_mont64(const orc_view* dst, const orc_view* lhs, const orc_view* rhs, const orc_modulus* mods,
                                 size_t nmods, size_t shift) {
    ORC_TRY View<int64_t> d(dst);
    View<i128> a(lhs);
    View<int64_t> b(rhs);
    require(shift >= 64 && shift < 128, "shift must be in [64, 128)");
    require(a.sh[0] == d.sh[0] && a.sh[1] == 1 && a.sh[2] == d.sh[2] && a.sh[3] == d.sh[3], "bconv_correct lhs shape");
    require(b.sh[0] == d.sh[0] || b.sh[0] == 1, "bconv_correct rhs dim 0");
    require(b.sh[1] == d.sh[1] && b.sh[2] == 1 && b.sh[3] == 1, "bconv_correct rhs shape");
    check_mods(dst, nmods);
    i128 half = (i128)1 << (shift - 1);
    for (int64_t i = 0; i < d.sh[0]; i++) {
        int64_t ri = b.sh[0] == 1 ? 0 : i;
        for (int64_t j = 0; j < d.sh[1]; j++) {
            i128 r = b.at(ri, j, 0, 0);
            for (int64_t k = 0; k < d.sh[2]; k++)
                for (int64_t l = 0; l < d.sh[3]; l++) {
                    i128 q = wadd128(a.at(i, 0, k, l), half) >> shift;
                    d.at(i, j, k, l) = wadd64(d.at(i, j, k, l), red128(mods[j], wmul128(q, r)));
                }
        }
    }
    ORC_CATCH
}

// rem_euclid by 2^k of a two's-complement value, centred as in cpu.rs:1007-1011
static i128 pow2_centre(i128 x, size_t log2_modulus) {
    u128 mask = ((u128)1 << log2_modulus) - 1;
    i128 reduced = (i128)((u128)x & mask);
    i128 modulus = (i128)1 << log2_modulus;
    return reduced > modulus / 2 ? reduced - modulus : reduced;
}

// bconv_correct_pow2 (cpu.rs:983-1015)
int orc_hal_bconv_correct_pow2(const orc_view* dst, const orc_view* src, int64_t factor0, int64_t factor1, size_t shift,
                               size_t log2_modulus) {
    ORC_TRY View<int64_t> d(dst);
    View<i128> s(src);
    require(d.sh[1] == 1, "bconv_correct_pow2 dst dim 1 must be 1");
    require(s.sh[0] == d.sh[0] && s.sh[1] == 2 && s.sh[2] == d.sh[2] && s.sh[3] == d.sh[3], "bconv_correct_pow2 src shape");
    require(shift >= 1 && shift < 127, "shift out of range");
    require(log2_modulus >= 1 && log2_modulus <= 62, "log2_modulus out of range");
    i128 half = (i128)1 << (shift - 1);
    for (int64_t i = 0; i < d.sh[0]; i++)
        for (int64_t k = 0; k < d.sh[2]; k++)
            for (int64_t l = 0; l < d.sh[3]; l++) {
                i128 quot = wadd128(s.at(i, 1, k, l), half) >> shift;
                i128 v = wmul128(wsub128(s.at(i, 0, k, l), wmul128(quot, (i128)factor0)), (i128)factor1);
                d.at(i, 0, k, l) = (int64_t)pow2_centre(v, log2_modulus);
            }
    ORC_CATCH
}

// bconv_correct_pow2_wide (cpu.rs:1017-1057), with split_i128_limbs (hal/mod.rs:71-76)
int orc_hal_bconv_correct_pow2_wide(const orc_view* dst, const orc_view* src, int64_t f0_lo, int64_t f0_hi, int64_t f1_lo,
                                    int64_t f1_hi, size_t shift, size_t log2_modulus) {
    ORC_TRY View<int64_t> d(dst);
    View<i128> s(src);
    require(d.sh[1] == 2, "bconv_correct_pow2_wide dst dim 1 must be 2");
    require(s.sh[0] == d.sh[0] && s.sh[1] == 3 && s.sh[2] == d.sh[2] && s.sh[3] == d.sh[3], "bconv_correct_pow2_wide src shape");
    require(shift >= 1 && shift < 127, "shift out of range");
    require(log2_modulus >= 1 && log2_modulus <= 126, "log2_modulus out of range");
    i128 f0 = (i128)(((u128)(uint64_t)f0_hi << 64) | (uint64_t)f0_lo);
    i128 f1 = (i128)(((u128)(uint64_t)f1_hi << 64) | (uint64_t)f1_lo);
    i128 half = (i128)1 << (shift - 1);
    for (int64_t i = 0; i < d.sh[0]; i++)
        for (int64_t k = 0; k < d.sh[2]; k++)
            for (int64_t l = 0; l < d.sh[3]; l++) {
                i128 quot = wadd128(s.at(i, 2, k, l), half) >> shift;
                i128 v = wadd128(s.at(i, 0, k, l), (i128)((u128)s.at(i, 1, k, l) << 64));
                i128 c = pow2_centre(wmul128(wsub128(v, wmul128(quot, f0)), f1), log2_modulus);
                int64_t lo = (int64_t)c;
                int64_t hi = (int64_t)(wsub128(c, (i128)lo) >> 64);
                d.at(i, 0, k, l) = lo;
                d.at(i, 1, k, l) = hi;
            }
    ORC_CATCH
}

// create_permutation (cpu.rs:1059-1083)
int orc_hal_create_permutation(const size_t* inner_shape, const size_t* shifts, size_t n, uint32_t* perm_out, size_t len) {
    ORC_TRY size_t total = 1;
    for (size_t t = 0; t < n; t++) total *= inner_shape[t];
    require(total == len, "permutation length mismatch");
    std::vector<size_t> idx(n, 0);
    for (size_t e = 0; e < total; e++) {
        size_t src = 0;
        for (size_t t = 0; t < n; t++) {
            size_t m = inner_shape[t];
            src = src * m + (idx[t] + m - shifts[t] % m) % m;
        }
        perm_out[e] = (uint32_t)src;
        for (size_t t = n; t-- > 0;) {
            if (++idx[t] < inner_shape[t]) break;
            idx[t] = 0;
        }
    }
    ORC_CATCH
}

// permute (cpu.rs:1085-1122)
int orc_hal_permute(const orc_view* dst, const orc_view* src, const uint32_t* perm, size_t len, int dim) {
    ORC_TRY View<int64_t> d(dst), s(src);
    for (int t = 0; t < 4; t++) require(d.sh[t] == s.sh[t], "permute shape mismatch");
    if (dim == 2) {
        require((int64_t)len == d.sh[2], "permutation length mismatch");
        for (int64_t i = 0; i < d.sh[0]; i++)
            for (int64_t j = 0; j < d.sh[1]; j++)
                for (int64_t e = 0; e < d.sh[2]; e++)
                    for (int64_t l = 0; l < d.sh[3]; l++) d.at(i, j, e, l) = s.at(i, j, perm[e], l);
    } else {
        require((int64_t)len == d.sh[3], "permutation length mismatch");
        for (int64_t i = 0; i < d.sh[0]; i++)
            for (int64_t j = 0; j < d.sh[1]; j++)
                for (int64_t k = 0; k < d.sh[2]; k++)
                    for (int64_t l = 0; l < d.sh[3]; l++) d.at(i, j, k, l) = s.at(i, j, k, perm[l]);
    }
    ORC_CATCH
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// ring.rs:1349-1380 — the CRT-kernel traces: for each transversal element g,
// kernel[g] = Σ_(x in subgroup) ζ^((x g^-1 mod m) mod part_order) mod p, with g^-1 taken mod m
// (field.rs:462-475 `inv_mod`). The heavy loop of the ring-level oracle lives here.
static uint64_t inv_mod_u64(uint64_t a, uint64_t n) {
    i128 old_r = (i128)(a % n), r = (i128)n, old_s = 1, s = 0;
    while (r != 0) {
        i128 q = old_r / r;
        i128 t = old_r - q * r;
        old_r = r;
        r = t;
        t = old_s - q * s;
        old_s = s;
        s = t;
    }
    require(old_r == 1 || old_r == -1, "elements are not coprime");
    i128 v = (old_s * old_r) % (i128)n;
    if (v < 0) v += n;
    return (uint64_t)v;
}

extern "C" int orc_crt_kernel_traces(int64_t p, int64_t zeta_part, uint64_t part_order, uint64_t m, const uint64_t* subgroup,
                                     size_t nsub, const uint64_t* transversal, size_t ntrans, int64_t* out) {
    ORC_TRY std::vector<int64_t> powers(part_order);
    int64_t cur = 1;
    for (uint64_t e = 0; e < part_order; e++) {
        powers[e] = cur;
        cur = mul_mod(cur, zeta_part, p);
    }
    for (size_t t = 0; t < ntrans; t++) {
        uint64_t g_inv = inv_mod_u64(transversal[t], m);
        int64_t sum = 0;
        for (size_t i = 0; i < nsub; i++) {
            uint64_t e = (uint64_t)(((u128)subgroup[i] * g_inv) % m) % part_order;
            sum = (sum + powers[e]) % p;
        }
        out[t] = sum;
    }
    ORC_CATCH
}
