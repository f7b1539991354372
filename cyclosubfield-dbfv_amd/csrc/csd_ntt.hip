// Group NTTs for the MI355X HAL: plan construction on the host and the LDS-staged transform
// kernel (hal_ntt_mont64_dim / hal_intt_mont64_dim / hal_ntt_mul_mont64_dim,
// cyclosubfield-dbfv/src/hal/mod.rs:209-238).
//
// A plan for the group C_n0 x ... x C_n(r-1) carries, per cyclic factor, the Barrett twiddles of
// every stage (the residue tree of hal/ntt.rs:183-187, 216-281) and three *schedules*: the
// sequences of butterfly stages and full reduction passes that the reference's lazy transform
// executes for (forward, reduced output), (forward, unreduced output) and (inverse). Those
// decisions depend only on the modulus, the factor lengths and the input bound 2N
// (hal/ntt.rs:316-333, 359-376, 440-447, 482-491, 596-638), so the host replays the reference's
// u128 bound recursion once and the kernel follows the resulting list without any data-dependent
// control flow. The kernel therefore produces the reference's representatives bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "csd_internal.hpp"

// ---------------------------------------------------------------------------------------------
// host modular helpers

int64_t csd_host_mulmod(int64_t a, int64_t b, int64_t q) { return (int64_t)(((i128)a * (i128)b) % (i128)q); }

int64_t csd_host_powmod(int64_t b, uint64_t e, int64_t q) {
    int64_t r = 1 % q;
    b %= q;
    if (b < 0) b += q;
    while (e) {
        if (e & 1) r = csd_host_mulmod(r, b, q);
        b = csd_host_mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}

// BarrettMul::new (hal/ntt.rs:134-144): centred constant, u = round(-z 2^64 / N) with halves
// rounded away from zero
CsdBarrett csd_host_barrett(int64_t value, int64_t n) {
    if (!(n > 1 && n % 2 == 1 && n < ((int64_t)1 << 62))) throw CsdError("invalid modulus for Barrett multiplication");
    int64_t z = value % n;
    if (2 * z > n)
        z -= n;
    else if (2 * z < -n)
        z += n;
    i128 a = -((i128)z << 64);
    i128 b = n;
    i128 u = a >= 0 ? (a + b / 2) / b : -((-a + b / 2) / b);
    return CsdBarrett{z, (int64_t)u};
}

// the smallest candidate power of exact order n (hal/ntt.rs:104-115)
int64_t csd_host_root_of_unity(int64_t q, size_t n) {
    if ((q - 1) % (int64_t)n != 0) throw CsdError("modulus doesn't support roots of unity of the requested order");
    size_t p = (n & (n - 1)) == 0 ? 2 : 3;
    for (int64_t c = 2;; c++) {
        int64_t root = csd_host_powmod(c, (uint64_t)((q - 1) / (int64_t)n), q);
        if (csd_host_powmod(root, n / p, q) != 1) return root;
    }
}

extern "C" int csd_hal_create_modulus(int64_t N, csd_modulus* out) {
    CSD_TRY if (N <= 1 || N % 2 == 0) throw CsdError("modulus must be an odd integer > 1");
    uint64_t inv = (uint64_t)N;
    for (int i = 0; i < 6; i++) inv *= 2 - (uint64_t)N * inv;  // N^-1 mod 2^64 by Newton steps
    auto pmod = [N](i128 x) -> int64_t {                        // mont.rs:17-23, 0 maps to N
        i128 r = x % (i128)N;
        if (r <= 0) r += N;
        return (int64_t)r;
    };
    out->N = N;
    out->neg_inv_N = (int64_t)(0 - inv);
    i128 r1 = pmod((i128)1 << 64);
    out->r2 = pmod(r1 * r1);
    out->r3 = pmod((i128)out->r2 * r1);
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// plans

#define NTT_MAX_FACTORS 8
#define NTT_MAX_STAGES 16
#define NTT_MAX_STEPS 96

enum StepKind : uint8_t {
    STEP_REDUCE = 0,     // all entries *= 1 (Barrett), bound -> N
    STEP_FWD = 1,        // forward stage; flags bit0: twiddle-1 skip allowed, bit1: fused reduction
    STEP_INV = 2,        // inverse stage; flags bit0: fused len^-1 scaling
    STEP_SCALE_ONE = 3,  // identity plan, forward with reduction
    STEP_SCALE_INV = 4,  // identity plan, inverse
};

struct DevStep {
    uint8_t kind, factor, stage, flags;
};

struct DevFactor {
    int32_t len, radix, nstages;
    uint32_t S, magic_S;  // suffix length (product of the later factor lengths)
    uint32_t stage_len[NTT_MAX_STAGES], part[NTT_MAX_STAGES], nblocks[NTT_MAX_STAGES];
    uint32_t magic_part[NTT_MAX_STAGES], magic_nb[NTT_MAX_STAGES];
    int32_t tw_off[NTT_MAX_STAGES], itw_off[NTT_MAX_STAGES];
    CsdBarrett zeta, zeta2;
};

struct DevPlan {
    int64_t n;
    CsdBarrett one, inv_len;
    int32_t len, nfactors;
    DevFactor f[NTT_MAX_FACTORS];
    int32_t nsteps[3];
    DevStep steps[3][NTT_MAX_STEPS];
    // followed in memory by the Barrett twiddle table
};

struct csd_ntt_plan {
    csd_ctx* ctx;
    DevPlan host;
    std::vector<CsdBarrett> tw;
    void* dev;  // DevPlan followed by the twiddles
    const CsdBarrett* dev_tw;
};

// n / d for n * d < 2^32: (n * magic) >> 32 with magic = floor(2^32 / d) + 1; magic 0 means d = 1
static uint32_t magic_for(uint32_t d) { return d <= 1 ? 0u : (uint32_t)((((uint64_t)1) << 32) / d + 1); }

static const u128 I64MAX_U = (u128)INT64_MAX;

// Replays CpuNTTPlan::apply's bound recursion (hal/ntt.rs:581-657) and records what each
// factor's LazyFactorNTT::forward/backward would execute.
static void build_schedule(DevPlan& P, int dir) {
    std::vector<DevStep> st;
    const bool inverse = dir == 2;
    const bool reduce_output = dir == 0;
    const u128 n = (u128)P.n;
    u128 bound = 2 * n;
    for (int fi = 0; fi < P.nfactors; fi++) {
        const DevFactor& F = P.f[fi];
        const bool last = fi + 1 == P.nfactors;
        if (!inverse) {
            const bool fuse_fwd = last && reduce_output;
            for (int s = 0; s < F.nstages; s++) {
                bool fuse = fuse_fwd && s + 1 == F.nstages;
                u128 growth = F.radix == 2 ? n : 2 * n;  // ntt.rs:319, 362
                if (bound + growth > I64MAX_U) {
                    st.push_back(DevStep{STEP_REDUCE, (uint8_t)fi, 0, 0});
                    bound = n;
                }
                u128 mult = (u128)F.radix;
                bool skip_ok = mult * bound <= I64MAX_U;  // ntt.rs:326, 369
                u128 nb = fuse ? n : (skip_ok ? mult * bound : bound + growth);
                st.push_back(DevStep{STEP_FWD, (uint8_t)fi, (uint8_t)s, (uint8_t)((skip_ok ? 1 : 0) | (fuse ? 2 : 0))});
                bound = nb;
            }
        } else {
            for (int s = F.nstages - 1; s >= 0; s--) {
                bool fused_scale = last && s == 0;
                u128 mult = (u128)F.radix;
                if (mult * bound > I64MAX_U) {  // ntt.rs:443, 487
                    st.push_back(DevStep{STEP_REDUCE, (uint8_t)fi, 0, 0});
                    bound = n;
                }
                bound = fused_scale ? n : mult * bound;
                st.push_back(DevStep{STEP_INV, (uint8_t)fi, (uint8_t)s, (uint8_t)(fused_scale ? 1 : 0)});
            }
        }
    }
    if (P.nfactors == 0) {
        if (inverse)
            st.push_back(DevStep{STEP_SCALE_INV, 0, 0, 0});
        else if (reduce_output)
            st.push_back(DevStep{STEP_SCALE_ONE, 0, 0, 0});
    }
    if (st.size() > NTT_MAX_STEPS) throw CsdError("NTT schedule too long");
    P.nsteps[dir] = (int32_t)st.size();
    for (size_t i = 0; i < st.size(); i++) P.steps[dir][i] = st[i];
}

static bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

extern "C" int csd_hal_create_ntt_plan(csd_ctx* ctx, const csd_modulus* modulus, const size_t* inner_shape, size_t nshape,
                                       csd_ntt_plan** out) {
    CSD_TRY csd_ntt_plan* plan = new csd_ntt_plan();
    plan->ctx = ctx;
    plan->dev = nullptr;
    DevPlan& P = plan->host;
    memset(&P, 0, sizeof(P));
    const int64_t n = modulus->N;
    P.n = n;
    P.one = csd_host_barrett(1, n);
    size_t len = 1;
    std::vector<size_t> lens;
    for (size_t i = 0; i < nshape; i++) {
        if (inner_shape[i] == 0) throw CsdError("NTT plan: zero-length factor");
        len *= inner_shape[i];
        if (inner_shape[i] > 1) lens.push_back(inner_shape[i]);
    }
    if (lens.size() > NTT_MAX_FACTORS) throw CsdError("NTT plan: too many factors");
    if (len > (1u << 20)) throw CsdError("NTT plan: transform too long");
    P.len = (int32_t)len;
    P.nfactors = (int32_t)lens.size();
    P.inv_len = csd_host_barrett(csd_host_powmod((int64_t)(len % (size_t)n), (uint64_t)(n - 2), n), n);
    std::vector<CsdBarrett>& tw = plan->tw;
    size_t suffix = len;
    for (size_t fi = 0; fi < lens.size(); fi++) {
        DevFactor& F = P.f[fi];
        size_t L = lens[fi];
        suffix /= L;
        F.len = (int32_t)L;
        F.S = (uint32_t)suffix;
        F.magic_S = magic_for(F.S);
        if (is_pow2(L)) {
            F.radix = 2;
            if (!(2 * (u128)n <= I64MAX_U)) throw CsdError("modulus too large for lazy radix-2 butterflies");
        } else {
            size_t l = L;
            while (l > 1) {
                if (l % 3) throw CsdError("NTT length must be a power of 2 or a power of 3");
                l /= 3;
            }
            F.radix = 3;
            if (!(3 * (u128)n <= I64MAX_U)) throw CsdError("modulus too large for lazy radix-3 butterflies");
        }
        const int R = F.radix;
        int64_t root = csd_host_root_of_unity(n, L);
        int64_t inv_root = csd_host_powmod(root, L - 1, n);
        int64_t zp = 0, zp2 = 0;
        if (R == 3) {
            zp = csd_host_powmod(root, L / 3, n);
            zp2 = csd_host_mulmod(zp, zp, n);
            F.zeta = csd_host_barrett(zp, n);
            F.zeta2 = csd_host_barrett(zp2, n);
        }
        // the residue tree: stage s has R^s blocks with exponents e; children e/R + k L/R
        std::vector<size_t> exps{0};
        int s = 0;
        std::vector<std::vector<CsdBarrett>> fwd, inv;
        while (exps.size() < L) {
            if (s >= NTT_MAX_STAGES) throw CsdError("NTT plan: too many stages");
            std::vector<CsdBarrett> f_st, i_st;
            for (size_t e : exps) {
                int64_t v = csd_host_powmod(root, e / R, n);
                int64_t vi = csd_host_powmod(inv_root, e / R, n);
                if (R == 2) {
                    f_st.push_back(csd_host_barrett(v, n));
                    i_st.push_back(csd_host_barrett(vi, n));
                } else {
                    int64_t v2 = csd_host_mulmod(v, v, n);
                    f_st.push_back(csd_host_barrett(v, n));
                    f_st.push_back(csd_host_barrett(csd_host_mulmod(zp, v, n), n));
                    f_st.push_back(csd_host_barrett(csd_host_mulmod(zp2, v, n), n));
                    f_st.push_back(csd_host_barrett(v2, n));
                    f_st.push_back(csd_host_barrett(csd_host_mulmod(zp, v2, n), n));
                    f_st.push_back(csd_host_barrett(csd_host_mulmod(zp2, v2, n), n));
                    i_st.push_back(csd_host_barrett(vi, n));
                    i_st.push_back(csd_host_barrett(csd_host_mulmod(vi, vi, n), n));
                }
            }
            fwd.push_back(f_st);
            inv.push_back(i_st);
            std::vector<size_t> next;
            for (size_t e : exps)
                for (int k = 0; k < R; k++) next.push_back(e / R + k * (L / R));
            exps = next;
            s++;
        }
        F.nstages = s;
        uint32_t stage_len = (uint32_t)L, nb = 1;
        for (int t = 0; t < s; t++) {
            F.stage_len[t] = stage_len;
            F.part[t] = stage_len / R;
            F.nblocks[t] = nb;
            F.magic_part[t] = magic_for(F.part[t]);
            F.magic_nb[t] = magic_for(nb);
            F.tw_off[t] = (int32_t)tw.size();
            tw.insert(tw.end(), fwd[t].begin(), fwd[t].end());
            F.itw_off[t] = (int32_t)tw.size();
            tw.insert(tw.end(), inv[t].begin(), inv[t].end());
            stage_len /= R;
            nb *= R;
        }
    }
    for (int dir = 0; dir < 3; dir++) build_schedule(P, dir);
    size_t bytes = sizeof(DevPlan) + tw.size() * sizeof(CsdBarrett);
    plan->dev = ctx->alloc_bytes(bytes);
    std::vector<char> blob(bytes);
    memcpy(blob.data(), &P, sizeof(DevPlan));
    if (!tw.empty()) memcpy(blob.data() + sizeof(DevPlan), tw.data(), tw.size() * sizeof(CsdBarrett));
    HIPCHECK(hipMemcpyAsync(plan->dev, blob.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    plan->dev_tw = (const CsdBarrett*)((const char*)plan->dev + sizeof(DevPlan));
    *out = plan;
    CSD_CATCH
}

extern "C" int csd_ntt_plan_free(csd_ntt_plan* plan) {
    CSD_TRY if (!plan) return 0;
    if (plan->dev) plan->ctx->free_bytes(plan->dev);
    delete plan;
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// the LDS-staged transform kernel

struct NttArgs {
    int64_t d0, d1, d2, d3;
    int64_t st[3];
    int64_t* data;
    const DevPlan* plans[CSD_MAX_MODULI];
    int dim;       // 2 or 3
    int dir;       // 0 forward reduced, 1 forward unreduced, 2 inverse
    int V;         // transforms per tile (power of two)
    int logV;
    int64_t chunks;  // tiles per (i, j)
    // optional fused Montgomery multiplication (ntt_mul_mont64_dim)
    const int64_t* mul;
    Str4 smul;
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t magic) {
    return magic ? (uint32_t)(((uint64_t)n * magic) >> 32) : n;
}

#define NTT_THREADS 256

__global__ __launch_bounds__(NTT_THREADS) void k_ntt(NttArgs a, ModTable mt) {
    extern __shared__ int64_t lds[];
    const int64_t chunk = blockIdx.x % a.chunks;
    const int64_t ij = blockIdx.x / a.chunks;
    const int64_t j = ij % a.d1;
    const int64_t i = ij / a.d1;
    const DevPlan* __restrict__ P = a.plans[j];
    const CsdBarrett* __restrict__ TW = (const CsdBarrett*)(P + 1);
    const int L = P->len;
    const int V = a.V;
    const int logV = a.logV;
    const int total = L * V;
    const int64_t n = P->n;
    int64_t* base = a.data + i * a.st[0] + j * a.st[1];

    // load the tile into LDS as [position][transform]
    if (a.dim == 3) {
        // V rows (k0 .. k0+V-1), each a contiguous line of length L
        const int64_t k0 = chunk * V;
        for (int g = threadIdx.x; g < total; g += NTT_THREADS) {
            int v = g / L;
            int p = g - v * L;
            int64_t k = k0 + v;
            lds[p * V + v] = k < a.d2 ? base[k * a.st[2] + p] : 0;
        }
    } else {
        // L rows of a band of V columns starting at l0
        const int64_t l0 = chunk * V;
        for (int g = threadIdx.x; g < total; g += NTT_THREADS) {
            int p = g >> logV;
            int v = g & (V - 1);
            int64_t l = l0 + v;
            lds[g] = l < a.d3 ? base[p * a.st[2] + l] : 0;
        }
    }
    __syncthreads();

    const int ns = P->nsteps[a.dir];
    for (int sidx = 0; sidx < ns; sidx++) {
        const DevStep s = P->steps[a.dir][sidx];
        if (s.kind == STEP_REDUCE || s.kind == STEP_SCALE_ONE || s.kind == STEP_SCALE_INV) {
            const CsdBarrett m = s.kind == STEP_SCALE_INV ? P->inv_len : P->one;
            for (int g = threadIdx.x; g < total; g += NTT_THREADS) lds[g] = csd_barrett(m, lds[g], n);
        } else {
            const DevFactor& F = P->f[s.factor];
            const int st = s.stage;
            const uint32_t S = F.S, part = F.part[st], nbk = F.nblocks[st], slen = F.stage_len[st];
            const uint32_t mS = F.magic_S, mp = F.magic_part[st], mnb = F.magic_nb[st];
            const int nbfly = (L / F.radix) << logV;
            const uint32_t gap = part * S;
            if (F.radix == 2) {
                const CsdBarrett* tw = TW + (s.kind == STEP_FWD ? F.tw_off[st] : F.itw_off[st]);
                for (int w = threadIdx.x; w < nbfly; w += NTT_THREADS) {
                    const int v = w & (V - 1);
                    const uint32_t b = (uint32_t)w >> logV;
                    const uint32_t q1 = fdiv(b, mS);
                    const uint32_t suffix = b - q1 * S;
                    const uint32_t grp = fdiv(q1, mp);
                    const uint32_t jj = q1 - grp * part;
                    const uint32_t blk = grp - fdiv(grp, mnb) * nbk;
                    const uint32_t e0 = (grp * slen + jj) * S + suffix;
                    const int x_i = (int)(e0 << logV) + v;
                    const int y_i = (int)((e0 + gap) << logV) + v;
                    const int64_t x = lds[x_i], y = lds[y_i];
                    const CsdBarrett t = tw[blk];
                    if (s.kind == STEP_FWD) {
                        const bool skip = (s.flags & 1) && t.z == 1;
                        const int64_t ty = skip ? y : csd_barrett(t, y, n);
                        int64_t o0 = (int64_t)((uint64_t)x + (uint64_t)ty);
                        int64_t o1 = (int64_t)((uint64_t)x - (uint64_t)ty);
                        if (s.flags & 2) {
                            o0 = csd_barrett(P->one, o0, n);
                            o1 = csd_barrett(P->one, o1, n);
                        }
                        lds[x_i] = o0;
                        lds[y_i] = o1;
                    } else {
                        const bool fuse = s.flags & 1;
                        const CsdBarrett ym = fuse ? P->inv_len : t;
                        const bool skip = !fuse && ym.z == 1;
                        const int64_t sum = (int64_t)((uint64_t)x + (uint64_t)y);
                        const int64_t dif = (int64_t)((uint64_t)x - (uint64_t)y);
                        lds[x_i] = fuse ? csd_barrett(ym, sum, n) : sum;
                        lds[y_i] = skip ? dif : csd_barrett(ym, dif, n);
                    }
                }
            } else {
                for (int w = threadIdx.x; w < nbfly; w += NTT_THREADS) {
                    const int v = w & (V - 1);
                    const uint32_t b = (uint32_t)w >> logV;
                    const uint32_t q1 = fdiv(b, mS);
                    const uint32_t suffix = b - q1 * S;
                    const uint32_t grp = fdiv(q1, mp);
                    const uint32_t jj = q1 - grp * part;
                    const uint32_t blk = grp - fdiv(grp, mnb) * nbk;
                    const uint32_t e0 = (grp * slen + jj) * S + suffix;
                    const int x_i = (int)(e0 << logV) + v;
                    const int y_i = (int)((e0 + gap) << logV) + v;
                    const int z_i = (int)((e0 + 2 * gap) << logV) + v;
                    const int64_t x = lds[x_i], y = lds[y_i], z = lds[z_i];
                    if (s.kind == STEP_FWD) {
                        const CsdBarrett* w6 = TW + F.tw_off[st] + 6 * blk;
                        const bool skip = (s.flags & 1) && w6[0].z == 1;
                        const int64_t tb = skip ? y : csd_barrett(w6[0], y, n);
                        const int64_t tbz = csd_barrett(w6[1], y, n);
                        const int64_t tbzz = csd_barrett(w6[2], y, n);
                        const int64_t tc = skip ? z : csd_barrett(w6[3], z, n);
                        const int64_t tcz = csd_barrett(w6[4], z, n);
                        const int64_t tczz = csd_barrett(w6[5], z, n);
                        int64_t o0 = (int64_t)((uint64_t)x + (uint64_t)tb + (uint64_t)tc);
                        int64_t o1 = (int64_t)((uint64_t)x + (uint64_t)tbz + (uint64_t)tczz);
                        int64_t o2 = (int64_t)((uint64_t)x + (uint64_t)tbzz + (uint64_t)tcz);
                        if (s.flags & 2) {
                            o0 = csd_barrett(P->one, o0, n);
                            o1 = csd_barrett(P->one, o1, n);
                            o2 = csd_barrett(P->one, o2, n);
                        }
                        lds[x_i] = o0;
                        lds[y_i] = o1;
                        lds[z_i] = o2;
                    } else {
                        const CsdBarrett* w2 = TW + F.itw_off[st] + 2 * blk;
                        const bool fuse = s.flags & 1;
                        const CsdBarrett ym = fuse ? P->inv_len : w2[0];
                        const CsdBarrett zm = fuse ? P->inv_len : w2[1];
                        const bool skip = !fuse && ym.z == 1;
                        const int64_t y1z = csd_barrett(F.zeta, y, n);
                        const int64_t y1zz = csd_barrett(F.zeta2, y, n);
                        const int64_t y2z = csd_barrett(F.zeta, z, n);
                        const int64_t y2zz = csd_barrett(F.zeta2, z, n);
                        const int64_t sum = (int64_t)((uint64_t)x + (uint64_t)y + (uint64_t)z);
                        const int64_t bb = (int64_t)((uint64_t)x + (uint64_t)y1zz + (uint64_t)y2z);
                        const int64_t cc = (int64_t)((uint64_t)x + (uint64_t)y1z + (uint64_t)y2zz);
                        lds[x_i] = fuse ? csd_barrett(ym, sum, n) : sum;
                        lds[y_i] = skip ? bb : csd_barrett(ym, bb, n);
                        lds[z_i] = skip ? cc : csd_barrett(zm, cc, n);
                    }
                }
            }
        }
        __syncthreads();
    }

    // store, optionally fused with the Montgomery multiplication by `mul`
    const CsdMod m = mt.m[j];
    if (a.dim == 3) {
        const int64_t k0 = chunk * V;
        for (int g = threadIdx.x; g < total; g += NTT_THREADS) {
            int v = g / L;
            int p = g - v * L;
            int64_t k = k0 + v;
            if (k < a.d2) {
                int64_t x = lds[p * V + v];
                if (a.mul) x = csd_mont_mul(m, x, a.mul[i * a.smul.s[0] + j * a.smul.s[1] + k * a.smul.s[2] + p * a.smul.s[3]]);
                base[k * a.st[2] + p] = x;
            }
        }
    } else {
        const int64_t l0 = chunk * V;
        for (int g = threadIdx.x; g < total; g += NTT_THREADS) {
            int p = g >> logV;
            int v = g & (V - 1);
            int64_t l = l0 + v;
            if (l < a.d3) {
                int64_t x = lds[g];
                if (a.mul) x = csd_mont_mul(m, x, a.mul[i * a.smul.s[0] + j * a.smul.s[1] + p * a.smul.s[2] + l * a.smul.s[3]]);
                base[p * a.st[2] + l] = x;
            }
        }
    }
}

static void launch_ntt(csd_ctx* ctx, const csd_view* dst, csd_ntt_plan* const* plans, size_t nplans, int dim, int dir,
                       const csd_view* mul, const csd_modulus* mods, size_t nmods) {
    check_view(dst, "ntt");
    if (dim != 2 && dim != 3) throw CsdError("ntt: dim must be 2 or 3");
    if ((int64_t)nplans != dst->shape[1]) throw CsdError("number of NTT plans must equal dimension 1");
    if (nplans > CSD_MAX_MODULI) throw CsdError("too many NTT plans");
    NttArgs a;
    memset(&a, 0, sizeof(a));
    a.d0 = dst->shape[0];
    a.d1 = dst->shape[1];
    a.d2 = dst->shape[2];
    a.d3 = dst->shape[3];
    for (int d = 0; d < 3; d++) a.st[d] = dst->strides[d];
    a.data = (int64_t*)dst->ptr;
    const int64_t L = dim == 2 ? a.d2 : a.d3;
    for (size_t jj = 0; jj < nplans; jj++) {
        if (!plans[jj]) throw CsdError("ntt: null plan");
        if (plans[jj]->host.len != L) throw CsdError("ntt: plan length doesn't match the transformed dimension");
        a.plans[jj] = (const DevPlan*)plans[jj]->dev;
    }
    a.dim = dim;
    a.dir = dir;
    ModTable t;
    memset(&t, 0, sizeof(t));
    if (mul) {
        check_view(mul, "ntt_mul operand");
        a.mul = (const int64_t*)mul->ptr;
        a.smul = bcast_strides(mul, dst, "ntt_mul operand");
        t = make_mod_table(mods, nmods, a.d1);
    }
    if (numel(dst) == 0) return;
    // tile: as many transforms as fit 32 KiB of LDS (at least one)
    int V = 1;
    const int64_t other = dim == 2 ? a.d3 : a.d2;
    while (V * 2 * L <= 4096 && V * 2 <= 256 && V < other) V *= 2;
    a.V = V;
    a.logV = 0;
    while ((1 << a.logV) < V) a.logV++;
    a.chunks = (other + V - 1) / V;
    int64_t blocks = a.d0 * a.d1 * a.chunks;
    if (blocks > (int64_t)INT32_MAX) throw CsdError("ntt: too many tiles");
    size_t lds = (size_t)L * V * 8;
    if (lds > 160 * 1024) throw CsdError("ntt: transform too long for LDS staging");
    static bool attr_set = false;
    if (!attr_set) {
        HIPCHECK(hipFuncSetAttribute((const void*)k_ntt, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    static const char* names[2][3] = {{"ntt_dim2", "ntt_mul_dim2", "intt_dim2"}, {"ntt_dim3", "ntt_mul_dim3", "intt_dim3"}};
    TimedScope ts(ctx, names[dim - 2][dir], 16.0 * (double)numel(dst));
    hipLaunchKernelGGL(k_ntt, dim3((unsigned)blocks), dim3(NTT_THREADS), lds, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
}

extern "C" int csd_hal_ntt_mont64_dim(csd_ctx* ctx, const csd_view* dst, csd_ntt_plan* const* plans, size_t nplans, int dim) {
    CSD_TRY launch_ntt(ctx, dst, plans, nplans, dim, 0, nullptr, nullptr, 0);
    CSD_CATCH
}

extern "C" int csd_hal_intt_mont64_dim(csd_ctx* ctx, const csd_view* dst, csd_ntt_plan* const* plans, size_t nplans, int dim) {
    CSD_TRY launch_ntt(ctx, dst, plans, nplans, dim, 2, nullptr, nullptr, 0);
    CSD_CATCH
}

extern "C" int csd_hal_ntt_mul_mont64_dim(csd_ctx* ctx, const csd_view* dst, csd_ntt_plan* const* plans, size_t nplans, int dim,
                                          const csd_view* mul, const csd_modulus* mods, size_t nmods) {
    CSD_TRY if (!mul) throw CsdError("ntt_mul: null multiplier");
    launch_ntt(ctx, dst, plans, nplans, dim, 1, mul, mods, nmods);
    CSD_CATCH
}
