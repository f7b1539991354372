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

#include <atomic>

#include <algorithm>
#include <cmath>
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
    int32_t len, radix, nstages, index;
    uint32_t S, magic_S;  // suffix length (product of the later factor lengths)
    uint32_t stage_len[NTT_MAX_STAGES], part[NTT_MAX_STAGES], nblocks[NTT_MAX_STAGES];
    uint32_t magic_part[NTT_MAX_STAGES], magic_nb[NTT_MAX_STAGES];
    int32_t tw_off[NTT_MAX_STAGES], itw_off[NTT_MAX_STAGES];
    CsdBarrett zeta, zeta2;
};

// Register-blocked passes: a pass applies stages [s0, s0 + g) of one factor to groups of R^g
// elements held in registers (in decreasing stage order for the inverse).
#define NTT_MAX_PASSES 24
#define INFO_REDUCE 1  // a full reduction pass precedes this stage (in processing order)
#define INFO_SKIP 2    // forward: twiddle-1 blocks may skip their multiplication
#define INFO_FUSE 4    // forward: fused output reduction; inverse: fused len^-1 scaling

struct DevPass {
    uint8_t factor, s0, g, radix;
};

struct DevPlan {
    int64_t n;
    CsdBarrett one, inv_len;
    int32_t len, nfactors;
    DevFactor f[NTT_MAX_FACTORS];
    int32_t nsteps[3];
    DevStep steps[3][NTT_MAX_STEPS];
    // per (direction, factor, stage): INFO_* bits, derived from `steps`
    uint8_t info[3][NTT_MAX_FACTORS][NTT_MAX_STAGES];
    int32_t npasses[3];
    DevPass passes[3][NTT_MAX_PASSES];
    // followed in memory by the Barrett twiddle table
};

// The device copy of a plan header: every Barrett constant in the folded form of
// csd_barrett_fold, which the kernels' csd_barrett_folded expects.
static DevPlan device_header(const DevPlan& P) {
    DevPlan D = P;
    D.one = csd_barrett_fold(D.one, P.n);
    D.inv_len = csd_barrett_fold(D.inv_len, P.n);
    for (int i = 0; i < NTT_MAX_FACTORS; i++) {
        D.f[i].zeta = csd_barrett_fold(D.f[i].zeta, P.n);
        D.f[i].zeta2 = csd_barrett_fold(D.f[i].zeta2, P.n);
    }
    return D;
}

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

static void build_passes(DevPlan& P, int dir, int g2, int g3);

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
    // the same schedule as per-stage flags, for the register-blocked kernels
    bool pending_reduce = false;
    for (const DevStep& s : st) {
        if (s.kind == STEP_REDUCE) {
            pending_reduce = true;
        } else if (s.kind == STEP_FWD || s.kind == STEP_INV) {
            uint8_t v = pending_reduce ? INFO_REDUCE : 0;
            if (s.kind == STEP_FWD) v |= ((s.flags & 1) ? INFO_SKIP : 0) | ((s.flags & 2) ? INFO_FUSE : 0);
            if (s.kind == STEP_INV) v |= (s.flags & 1) ? INFO_FUSE : 0;
            P.info[dir][s.factor][s.stage] = v;
            pending_reduce = false;
        }
    }
    build_passes(P, dir, 3, 2);
}

// Groups each factor's stages into register-blocked LDS passes of at most g2 radix-2 or g3
// radix-3 stages. The grouping changes only the LDS round trips, never the arithmetic: every
// stage still applies its own schedule bits, so all groupings give identical results.
static void build_passes(DevPlan& P, int dir, int g2, int g3) {
    const bool inverse = dir == 2;
    std::vector<DevPass> passes;
    for (int fi = 0; fi < P.nfactors; fi++) {
        const DevFactor& F = P.f[fi];
        const int G = F.radix == 2 ? g2 : g3;
        if (!inverse) {
            for (int s0 = 0; s0 < F.nstages;) {
                int g = std::min(G, F.nstages - s0);
                passes.push_back(DevPass{(uint8_t)fi, (uint8_t)s0, (uint8_t)g, (uint8_t)F.radix});
                s0 += g;
            }
        } else {
            for (int hi = F.nstages; hi > 0;) {
                int g = std::min(G, hi);
                passes.push_back(DevPass{(uint8_t)fi, (uint8_t)(hi - g), (uint8_t)g, (uint8_t)F.radix});
                hi -= g;
            }
        }
    }
    if (passes.size() > NTT_MAX_PASSES) throw CsdError("NTT plan: too many passes");
    P.npasses[dir] = (int32_t)passes.size();
    for (size_t i = 0; i < passes.size(); i++) P.passes[dir][i] = passes[i];
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
        F.index = (int32_t)fi;
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
    const DevPlan D = device_header(P);
    std::vector<CsdBarrett> dtw(tw.size());
    for (size_t i = 0; i < tw.size(); i++) dtw[i] = csd_barrett_fold(tw[i], n);
    memcpy(blob.data(), &D, sizeof(DevPlan));
    if (!dtw.empty()) memcpy(blob.data() + sizeof(DevPlan), dtw.data(), dtw.size() * sizeof(CsdBarrett));
    HIPCHECK(hipMemcpyAsync(plan->dev, blob.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    plan->dev_tw = (const CsdBarrett*)((const char*)plan->dev + sizeof(DevPlan));
    *out = plan;
    CSD_CATCH
}

// Modular multiplications (Barrett or Montgomery) one transform of the plan performs in
// direction dir (0 forward reduced, 1 forward unreduced, 2 inverse), following the schedule and
// the twiddle-1 skips exactly; the algorithmic work unit of the integer-VALU roofline.
extern "C" int csd_ntt_plan_mults(const csd_ntt_plan* plan, int dir, uint64_t* out) {
    CSD_TRY if (dir < 0 || dir > 2) throw CsdError("dir must be 0, 1 or 2");
    const DevPlan& P = plan->host;
    const uint64_t L = (uint64_t)P.len;
    uint64_t cnt = 0;
    for (int i = 0; i < P.nsteps[dir]; i++) {
        const DevStep& s = P.steps[dir][i];
        if (s.kind == STEP_REDUCE || s.kind == STEP_SCALE_ONE || s.kind == STEP_SCALE_INV) {
            cnt += L;
            continue;
        }
        const DevFactor& F = P.f[s.factor];
        const uint64_t prefixes = L / ((uint64_t)F.len * F.S);
        const uint64_t per_block = (uint64_t)F.part[s.stage] * F.S * prefixes;  // butterflies per block index
        for (uint32_t b = 0; b < F.nblocks[s.stage]; b++) {
            if (s.kind == STEP_FWD) {
                const bool fuse = s.flags & 2;
                if (F.radix == 2) {
                    bool skip = (s.flags & 1) && plan->tw[F.tw_off[s.stage] + b].z == 1;
                    cnt += per_block * ((skip ? 0 : 1) + (fuse ? 2 : 0));
                } else {
                    bool skip = (s.flags & 1) && plan->tw[F.tw_off[s.stage] + 6 * b].z == 1;
                    cnt += per_block * ((skip ? 4 : 6) + (fuse ? 3 : 0));
                }
            } else {
                const bool fuse = s.flags & 1;
                if (F.radix == 2) {
                    bool skip = !fuse && plan->tw[F.itw_off[s.stage] + b].z == 1;
                    cnt += per_block * (fuse ? 2 : (skip ? 0 : 1));
                } else {
                    bool skip = !fuse && plan->tw[F.itw_off[s.stage] + 2 * b].z == 1;
                    cnt += per_block * (4 + (fuse ? 3 : (skip ? 0 : 2)));
                }
            }
        }
    }
    *out = cnt;
    CSD_CATCH
}

// Regroups the plan's LDS passes (an autotuning knob, see build_passes) and refreshes the
// device copy of the plan header; results are unchanged.
extern "C" int csd_ntt_plan_set_grouping(csd_ntt_plan* plan, int g2, int g3) {
    CSD_TRY if (g2 < 1 || g2 > 3 || g3 < 1 || g3 > 2) throw CsdError("grouping: 1 <= g2 <= 3 and 1 <= g3 <= 2");
    for (int dir = 0; dir < 3; dir++) build_passes(plan->host, dir, g2, g3);
    const DevPlan D = device_header(plan->host);
    HIPCHECK(hipMemcpyAsync(plan->dev, &D, sizeof(DevPlan), hipMemcpyHostToDevice, plan->ctx->stream));
    HIPCHECK(hipStreamSynchronize(plan->ctx->stream));
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
            for (int g = threadIdx.x; g < total; g += NTT_THREADS) lds[g] = csd_barrett_folded(m, lds[g], n);
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
                        const bool skip = (s.flags & 1) && csd_folded_is_one(t, n);
                        const int64_t ty = skip ? y : csd_barrett_folded(t, y, n);
                        int64_t o0 = (int64_t)((uint64_t)x + (uint64_t)ty);
                        int64_t o1 = (int64_t)((uint64_t)x - (uint64_t)ty);
                        if (s.flags & 2) {
                            o0 = csd_barrett_folded(P->one, o0, n);
                            o1 = csd_barrett_folded(P->one, o1, n);
                        }
                        lds[x_i] = o0;
                        lds[y_i] = o1;
                    } else {
                        const bool fuse = s.flags & 1;
                        const CsdBarrett ym = fuse ? P->inv_len : t;
                        const bool skip = !fuse && csd_folded_is_one(ym, n);
                        const int64_t sum = (int64_t)((uint64_t)x + (uint64_t)y);
                        const int64_t dif = (int64_t)((uint64_t)x - (uint64_t)y);
                        lds[x_i] = fuse ? csd_barrett_folded(ym, sum, n) : sum;
                        lds[y_i] = skip ? dif : csd_barrett_folded(ym, dif, n);
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
                        const bool skip = (s.flags & 1) && csd_folded_is_one(w6[0], n);
                        const int64_t tb = skip ? y : csd_barrett_folded(w6[0], y, n);
                        const int64_t tbz = csd_barrett_folded(w6[1], y, n);
                        const int64_t tbzz = csd_barrett_folded(w6[2], y, n);
                        const int64_t tc = skip ? z : csd_barrett_folded(w6[3], z, n);
                        const int64_t tcz = csd_barrett_folded(w6[4], z, n);
                        const int64_t tczz = csd_barrett_folded(w6[5], z, n);
                        int64_t o0 = (int64_t)((uint64_t)x + (uint64_t)tb + (uint64_t)tc);
                        int64_t o1 = (int64_t)((uint64_t)x + (uint64_t)tbz + (uint64_t)tczz);
                        int64_t o2 = (int64_t)((uint64_t)x + (uint64_t)tbzz + (uint64_t)tcz);
                        if (s.flags & 2) {
                            o0 = csd_barrett_folded(P->one, o0, n);
                            o1 = csd_barrett_folded(P->one, o1, n);
                            o2 = csd_barrett_folded(P->one, o2, n);
                        }
                        lds[x_i] = o0;
                        lds[y_i] = o1;
                        lds[z_i] = o2;
                    } else {
                        const CsdBarrett* w2 = TW + F.itw_off[st] + 2 * blk;
                        const bool fuse = s.flags & 1;
                        const CsdBarrett ym = fuse ? P->inv_len : w2[0];
                        const CsdBarrett zm = fuse ? P->inv_len : w2[1];
                        const bool skip = !fuse && csd_folded_is_one(ym, n);
                        const int64_t y1z = csd_barrett_folded(F.zeta, y, n);
                        const int64_t y1zz = csd_barrett_folded(F.zeta2, y, n);
                        const int64_t y2z = csd_barrett_folded(F.zeta, z, n);
                        const int64_t y2zz = csd_barrett_folded(F.zeta2, z, n);
                        const int64_t sum = (int64_t)((uint64_t)x + (uint64_t)y + (uint64_t)z);
                        const int64_t bb = (int64_t)((uint64_t)x + (uint64_t)y1zz + (uint64_t)y2z);
                        const int64_t cc = (int64_t)((uint64_t)x + (uint64_t)y1z + (uint64_t)y2zz);
                        lds[x_i] = fuse ? csd_barrett_folded(ym, sum, n) : sum;
                        lds[y_i] = skip ? bb : csd_barrett_folded(ym, bb, n);
                        lds[z_i] = skip ? cc : csd_barrett_folded(zm, cc, n);
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

// ---------------------------------------------------------------------------------------------
// register-blocked butterflies (the reference's butterflies, hal/ntt.rs:335-350, 378-407,
// 450-476, 494-531, with the per-stage decisions read from DevPlan::info)
//
// The twiddle-1 fast path (skip) leaves a product out, as the reference does. With LANE set the
// twiddles differ between the lanes of a wave: the product is then formed for every lane and the
// skip is a select, so the twiddle's words are loaded together instead of behind a branch on z.

template <bool LANE>
__device__ __forceinline__ int64_t skip_or_mul(bool skip, const CsdBarrett& t, int64_t x, int64_t n) {
    if constexpr (LANE) {
        const int64_t p = csd_barrett_folded(t, x, n);
        return skip ? x : p;
    } else {
        return skip ? x : csd_barrett_folded(t, x, n);
    }
}

template <bool LANE = false>
__device__ __forceinline__ void bf2_fwd(int64_t& x, int64_t& y, CsdBarrett t, uint8_t info, CsdBarrett one, int64_t n) {
    const bool skip = (info & INFO_SKIP) && csd_folded_is_one(t, n);
    const int64_t ty = skip_or_mul<LANE>(skip, t, y, n);
    int64_t o0 = (int64_t)((uint64_t)x + (uint64_t)ty);
    int64_t o1 = (int64_t)((uint64_t)x - (uint64_t)ty);
    if (info & INFO_FUSE) {
        o0 = csd_barrett_folded(one, o0, n);
        o1 = csd_barrett_folded(one, o1, n);
    }
    x = o0;
    y = o1;
}

template <bool LANE = false>
__device__ __forceinline__ void bf2_inv(int64_t& x, int64_t& y, CsdBarrett t, uint8_t info, CsdBarrett inv_len, int64_t n) {
    const bool fuse = info & INFO_FUSE;
    const CsdBarrett ym = fuse ? inv_len : t;
    const bool skip = !fuse && csd_folded_is_one(ym, n);
    const int64_t sum = (int64_t)((uint64_t)x + (uint64_t)y);
    const int64_t dif = (int64_t)((uint64_t)x - (uint64_t)y);
    x = fuse ? csd_barrett_folded(ym, sum, n) : sum;
    y = skip_or_mul<LANE>(skip, ym, dif, n);
}

template <bool LANE = false>
__device__ __forceinline__ void bf3_fwd(int64_t& x, int64_t& y, int64_t& z, const CsdBarrett* w6, uint8_t info, CsdBarrett one,
                                        int64_t n) {
    const bool skip = (info & INFO_SKIP) && csd_folded_is_one(w6[0], n);
    const int64_t tb = skip_or_mul<LANE>(skip, w6[0], y, n);
    const int64_t tbz = csd_barrett_folded(w6[1], y, n);
    const int64_t tbzz = csd_barrett_folded(w6[2], y, n);
    const int64_t tc = skip_or_mul<LANE>(skip, w6[3], z, n);
    const int64_t tcz = csd_barrett_folded(w6[4], z, n);
    const int64_t tczz = csd_barrett_folded(w6[5], z, n);
    int64_t o0 = (int64_t)((uint64_t)x + (uint64_t)tb + (uint64_t)tc);
    int64_t o1 = (int64_t)((uint64_t)x + (uint64_t)tbz + (uint64_t)tczz);
    int64_t o2 = (int64_t)((uint64_t)x + (uint64_t)tbzz + (uint64_t)tcz);
    if (info & INFO_FUSE) {
        o0 = csd_barrett_folded(one, o0, n);
        o1 = csd_barrett_folded(one, o1, n);
        o2 = csd_barrett_folded(one, o2, n);
    }
    x = o0;
    y = o1;
    z = o2;
}

template <bool LANE = false>
__device__ __forceinline__ void bf3_inv(int64_t& x, int64_t& y, int64_t& z, const CsdBarrett* w2, uint8_t info,
                                        CsdBarrett inv_len, CsdBarrett zeta, CsdBarrett zeta2, int64_t n) {
    const bool fuse = info & INFO_FUSE;
    const CsdBarrett ym = fuse ? inv_len : w2[0];
    const CsdBarrett zm = fuse ? inv_len : w2[1];
    const bool skip = !fuse && csd_folded_is_one(ym, n);
    const int64_t y1z = csd_barrett_folded(zeta, y, n);
    const int64_t y1zz = csd_barrett_folded(zeta2, y, n);
    const int64_t y2z = csd_barrett_folded(zeta, z, n);
    const int64_t y2zz = csd_barrett_folded(zeta2, z, n);
    const int64_t sum = (int64_t)((uint64_t)x + (uint64_t)y + (uint64_t)z);
    const int64_t bb = (int64_t)((uint64_t)x + (uint64_t)y1zz + (uint64_t)y2z);
    const int64_t cc = (int64_t)((uint64_t)x + (uint64_t)y1z + (uint64_t)y2zz);
    x = fuse ? csd_barrett_folded(ym, sum, n) : sum;
    y = skip_or_mul<LANE>(skip, ym, bb, n);
    z = skip_or_mul<LANE>(skip, zm, cc, n);
}

template <int R>
struct Pow {
    static constexpr int of(int e) { return e == 0 ? 1 : R * of(e - 1); }
};

// One level (stage s = s0 + i) of a register block of M = R^G elements: members of a butterfly are
// D = R^(G-1-i) apart, and sub-block `sb` of the block maps to stage-s block (blk0 R^i + sb) mod R^s.
template <int R, int G, int I, bool INV>
__device__ __forceinline__ void block_level(int64_t (&v)[Pow<R>::of(G)], const DevPlan* __restrict__ P,
                                            const CsdBarrett* __restrict__ TW, const DevFactor& F, int dir, int s0,
                                            uint32_t blk0) {
    constexpr int M = Pow<R>::of(G);
    constexpr int D = Pow<R>::of(G - 1 - I);
    constexpr int NSB = Pow<R>::of(I);
    const int s = s0 + I;
    const uint8_t info = P->info[dir][F.index][s];
    const int64_t n = P->n;
    if (info & INFO_REDUCE) {
#pragma unroll
        for (int m = 0; m < M; m++) v[m] = csd_barrett_folded(P->one, v[m], n);
    }
    const uint32_t nbk = F.nblocks[s];
#pragma unroll
    for (int sb = 0; sb < NSB; sb++) {
        uint32_t g = blk0 * (uint32_t)NSB + (uint32_t)sb;
        uint32_t blk = R == 2 ? (g & (nbk - 1)) : (g - fdiv(g, F.magic_nb[s]) * nbk);
#pragma unroll
        for (int jj = 0; jj < D; jj++) {
            const int m0 = sb * D * R + jj;
            if constexpr (R == 2) {
                if constexpr (!INV)
                    bf2_fwd(v[m0], v[m0 + D], TW[F.tw_off[s] + blk], info, P->one, n);
                else
                    bf2_inv(v[m0], v[m0 + D], TW[F.itw_off[s] + blk], info, P->inv_len, n);
            } else {
                if constexpr (!INV)
                    bf3_fwd(v[m0], v[m0 + D], v[m0 + 2 * D], TW + F.tw_off[s] + 6 * blk, info, P->one, n);
                else
                    bf3_inv(v[m0], v[m0 + D], v[m0 + 2 * D], TW + F.itw_off[s] + 2 * blk, info, P->inv_len, F.zeta, F.zeta2, n);
            }
        }
    }
}

template <int R, int G, bool INV>
__device__ __forceinline__ void block_stages(int64_t (&v)[Pow<R>::of(G)], const DevPlan* P, const CsdBarrett* TW,
                                             const DevFactor& F, int dir, int s0, uint32_t blk0) {
    if constexpr (!INV) {
        block_level<R, G, 0, false>(v, P, TW, F, dir, s0, blk0);
        if constexpr (G > 1) block_level<R, G, 1, false>(v, P, TW, F, dir, s0, blk0);
        if constexpr (G > 2) block_level<R, G, 2, false>(v, P, TW, F, dir, s0, blk0);
    } else {
        if constexpr (G > 2) block_level<R, G, 2, true>(v, P, TW, F, dir, s0, blk0);
        if constexpr (G > 1) block_level<R, G, 1, true>(v, P, TW, F, dir, s0, blk0);
        block_level<R, G, 0, true>(v, P, TW, F, dir, s0, blk0);
    }
}

// LDS slot of line element e: one padding slot after every 8 elements, so that the strided
// groups of the late stages (8 consecutive elements per lane) spread over all banks
__device__ __forceinline__ uint32_t lds_slot(uint32_t e) { return e + (e >> 3); }
__host__ __device__ __forceinline__ uint32_t lds_line(uint32_t L) { return L + (L >> 3) + 1; }

// LDS layout of the compiled shapes: unpadded lines. A 3072-line then takes 24 KiB instead of
// 27.7 KiB, so six workgroups fit a CU instead of five; on MI355X that measured 3-4 % faster than
// the padded layout above (and than an XOR swizzle of the same size), the extra bank conflicts
// costing less than the added occupancy gains.
__host__ __device__ __forceinline__ constexpr uint32_t ct_slot(uint32_t e) { return e; }
__host__ __device__ __forceinline__ constexpr uint32_t ct_line(uint32_t L) { return L; }

// Runs one pass over all groups of `rows` LDS-resident lines of length L.
template <int R, int G, bool INV>
__device__ __forceinline__ void run_pass(int64_t* __restrict__ lds, int L, int rows, uint32_t magic_groups,
                                         const DevPlan* __restrict__ P, const CsdBarrett* __restrict__ TW, const DevPass ps,
                                         int dir) {
    constexpr int M = Pow<R>::of(G);
    const DevFactor& F = P->f[ps.factor];
    const uint32_t S = F.S, mS = F.magic_S;
    const uint32_t SL = F.stage_len[ps.s0];
    const uint32_t sub = SL / M;
    const uint32_t msub = sub <= 1 ? 0u : (uint32_t)((((uint64_t)1) << 32) / sub + 1);
    const uint32_t ngroups = (uint32_t)L / M;
    const uint32_t total = ngroups * (uint32_t)rows;
    const uint32_t lline = lds_line((uint32_t)L);
    for (uint32_t w = threadIdx.x; w < total; w += blockDim.x) {
        const uint32_t row = rows == 1 ? 0u : fdiv(w, magic_groups);
        const uint32_t gam = w - row * ngroups;
        const uint32_t t = fdiv(gam, mS);
        const uint32_t suffix = gam - t * S;
        const uint32_t blk0 = fdiv(t, msub);
        const uint32_t jj = t - blk0 * sub;
        const uint32_t base = (blk0 * SL + jj) * S + suffix;
        const uint32_t stride = sub * S;
        int64_t* line = lds + (size_t)row * lline;
        int64_t v[M];
#pragma unroll
        for (int m = 0; m < M; m++) v[m] = line[lds_slot(base + m * stride)];
        block_stages<R, G, INV>(v, P, TW, F, dir, ps.s0, blk0);
#pragma unroll
        for (int m = 0; m < M; m++) line[lds_slot(base + m * stride)] = v[m];
    }
}

// ---------------------------------------------------------------------------------------------
// Compile-time pass shapes for the dimension-3 plans of the benchmark fields. The generic pass
// above derives its index decomposition and its twiddle blocks at run time; here every divisor
// is a constant, and a register block that starts at stage 0 of its factor has block index
// (blk0 R^i + sb) mod R^i = sb, a constant, so its twiddles are wave-uniform scalar loads. The
// arithmetic (stages, schedule bits, twiddles) is the generic pass's, so results are identical.

template <int R, int G, int S0, int I, bool INV>
__device__ __forceinline__ void ct_level(int64_t (&v)[Pow<R>::of(G)], const DevPlan* __restrict__ P,
                                         const CsdBarrett* __restrict__ TW, const DevFactor& F, int dir, uint32_t blk0) {
    constexpr int M = Pow<R>::of(G);
    constexpr int D = Pow<R>::of(G - 1 - I);
    constexpr int NSB = Pow<R>::of(I);
    constexpr int s = S0 + I;
    constexpr uint32_t nbk = (uint32_t)Pow<R>::of(s);
    const uint8_t info = P->info[dir][F.index][s];
    const int64_t n = P->n;
    if (info & INFO_REDUCE) {
#pragma unroll
        for (int m = 0; m < M; m++) v[m] = csd_barrett_folded(P->one, v[m], n);
    }
#pragma unroll
    for (int sb = 0; sb < NSB; sb++) {
        const uint32_t blk = (blk0 * (uint32_t)NSB + (uint32_t)sb) % nbk;
#pragma unroll
        for (int jj = 0; jj < D; jj++) {
            const int m0 = sb * D * R + jj;
            // blocks starting at stage 0 have wave-uniform twiddles; the others vary per lane. The
            // select form (LANE) measured faster in the forward passes and slower in the inverse
            // ones (field_36864, MI355X), so the inverse keeps the branch.
            constexpr bool LANE = S0 > 0 && !INV;
            if constexpr (R == 2) {
                if constexpr (!INV)
                    bf2_fwd<LANE>(v[m0], v[m0 + D], TW[F.tw_off[s] + blk], info, P->one, n);
                else
                    bf2_inv<LANE>(v[m0], v[m0 + D], TW[F.itw_off[s] + blk], info, P->inv_len, n);
            } else {
                if constexpr (!INV)
                    bf3_fwd<LANE>(v[m0], v[m0 + D], v[m0 + 2 * D], TW + F.tw_off[s] + 6 * blk, info, P->one, n);
                else
                    bf3_inv<LANE>(v[m0], v[m0 + D], v[m0 + 2 * D], TW + F.itw_off[s] + 2 * blk, info, P->inv_len, F.zeta,
                                  F.zeta2, n);
            }
        }
    }
}

template <int R, int G, int S0, bool INV>
__device__ __forceinline__ void ct_stages(int64_t (&v)[Pow<R>::of(G)], const DevPlan* P, const CsdBarrett* TW,
                                          const DevFactor& F, int dir, uint32_t blk0) {
    if constexpr (!INV) {
        ct_level<R, G, S0, 0, false>(v, P, TW, F, dir, blk0);
        if constexpr (G > 1) ct_level<R, G, S0, 1, false>(v, P, TW, F, dir, blk0);
        if constexpr (G > 2) ct_level<R, G, S0, 2, false>(v, P, TW, F, dir, blk0);
    } else {
        if constexpr (G > 2) ct_level<R, G, S0, 2, true>(v, P, TW, F, dir, blk0);
        if constexpr (G > 1) ct_level<R, G, S0, 1, true>(v, P, TW, F, dir, blk0);
        ct_level<R, G, S0, 0, true>(v, P, TW, F, dir, blk0);
    }
}

// stages [S0, S0 + G) of factor FI (radix R, suffix S, stage length SL at S0) on `rows` lines of L
template <int L, int FI, int R, int G, int S, int SL, int S0, bool INV>
__device__ __forceinline__ void ct_pass(int64_t* __restrict__ lds, int rows, const DevPlan* __restrict__ P,
                                        const CsdBarrett* __restrict__ TW, int dir) {
    constexpr int M = Pow<R>::of(G);
    constexpr uint32_t sub = SL / M, NG = L / M, stride = sub * S, lline = ct_line(L);
    const DevFactor& F = P->f[FI];
    const uint32_t total = NG * (uint32_t)rows;
    for (uint32_t w = threadIdx.x; w < total; w += blockDim.x) {
        const uint32_t row = w / NG, gam = w % NG;
        const uint32_t t = gam / S, suffix = gam % S;
        const uint32_t blk0 = t / sub, jj = t % sub;
        const uint32_t base = (blk0 * SL + jj) * S + suffix;
        int64_t* line = lds + row * lline;
        int64_t v[M];
#pragma unroll
        for (int m = 0; m < M; m++) v[m] = line[ct_slot(base + m * stride)];
        ct_stages<R, G, S0, INV>(v, P, TW, F, dir, blk0);
#pragma unroll
        for (int m = 0; m < M; m++) line[ct_slot(base + m * stride)] = v[m];
    }
    __syncthreads();
}

// the last pass of a compiled shape, fused with the store: the register block goes straight to
// global memory, through the optional Montgomery multiplier mul[row * ms2 + p * ms3]
template <int L, int FI, int R, int G, int S, int SL, int S0, bool INV>
__device__ __forceinline__ void ct_pass_store(const int64_t* __restrict__ lds, int rows, const DevPlan* __restrict__ P,
                                              const CsdBarrett* __restrict__ TW, int dir, int64_t* __restrict__ out, int64_t st2,
                                              const int64_t* __restrict__ mul, int64_t ms2, int64_t ms3, const CsdMod& md) {
    constexpr int M = Pow<R>::of(G);
    constexpr uint32_t sub = SL / M, NG = L / M, stride = sub * S, lline = ct_line(L);
    const DevFactor& F = P->f[FI];
    const uint32_t total = NG * (uint32_t)rows;
    for (uint32_t w = threadIdx.x; w < total; w += blockDim.x) {
        const uint32_t row = w / NG, gam = w % NG;
        const uint32_t t = gam / S, suffix = gam % S;
        const uint32_t blk0 = t / sub, jj = t % sub;
        const uint32_t base = (blk0 * SL + jj) * S + suffix;
        const int64_t* line = lds + row * lline;
        int64_t v[M];
#pragma unroll
        for (int m = 0; m < M; m++) v[m] = line[ct_slot(base + m * stride)];
        ct_stages<R, G, S0, INV>(v, P, TW, F, dir, blk0);
#pragma unroll
        for (int m = 0; m < M; m++) {
            const uint32_t p = base + m * stride;
            int64_t x = v[m];
            if (mul) x = csd_mont_mul(md, x, mul[row * ms2 + p * ms3]);
            out[row * st2 + p] = x;
        }
    }
}

// The compiled shapes. SHAPE 1: factors [2, 3, 8, 64] (field_36864, deg_K = 3072); SHAPE 2:
// [3, 64, 64] (field_49152, deg_K = 12288); both with the default grouping (3 radix-2 / 1 radix-3
// stages per pass), forward passes in increasing and inverse passes in decreasing stage order.
template <int SHAPE>
struct CtShape;
template <>
struct CtShape<1> {
    static constexpr int L = 3072;
};
template <>
struct CtShape<2> {
    static constexpr int L = 12288;
};

template <int SHAPE, bool INV, bool FUSED_HEAD = false, bool FUSED_TAIL = false>
__device__ __forceinline__ void ct_passes(int64_t* lds, int rows, const DevPlan* P, const CsdBarrett* TW, int dir) {
    if constexpr (SHAPE == 1) {
        constexpr int L = 3072;
        if constexpr (!FUSED_HEAD) {
            ct_pass<L, 0, 2, 1, 1536, 2, 0, INV>(lds, rows, P, TW, dir);
            ct_pass<L, 1, 3, 1, 512, 3, 0, INV>(lds, rows, P, TW, dir);
        }
        ct_pass<L, 2, 2, 3, 64, 8, 0, INV>(lds, rows, P, TW, dir);
        if constexpr (!INV) {
            ct_pass<L, 3, 2, 3, 1, 64, 0, INV>(lds, rows, P, TW, dir);
            if constexpr (!FUSED_TAIL) ct_pass<L, 3, 2, 3, 1, 8, 3, INV>(lds, rows, P, TW, dir);
        } else {
            ct_pass<L, 3, 2, 3, 1, 8, 3, INV>(lds, rows, P, TW, dir);
            if constexpr (!FUSED_TAIL) ct_pass<L, 3, 2, 3, 1, 64, 0, INV>(lds, rows, P, TW, dir);
        }
    } else {
        constexpr int L = 12288;
        if constexpr (!FUSED_HEAD) ct_pass<L, 0, 3, 1, 4096, 3, 0, INV>(lds, rows, P, TW, dir);
        if constexpr (!INV) {
            ct_pass<L, 1, 2, 3, 64, 64, 0, INV>(lds, rows, P, TW, dir);
            ct_pass<L, 1, 2, 3, 64, 8, 3, INV>(lds, rows, P, TW, dir);
            ct_pass<L, 2, 2, 3, 1, 64, 0, INV>(lds, rows, P, TW, dir);
            if constexpr (!FUSED_TAIL) ct_pass<L, 2, 2, 3, 1, 8, 3, INV>(lds, rows, P, TW, dir);
        } else {
            ct_pass<L, 1, 2, 3, 64, 8, 3, INV>(lds, rows, P, TW, dir);
            ct_pass<L, 1, 2, 3, 64, 64, 0, INV>(lds, rows, P, TW, dir);
            ct_pass<L, 2, 2, 3, 1, 8, 3, INV>(lds, rows, P, TW, dir);
            if constexpr (!FUSED_TAIL) ct_pass<L, 2, 2, 3, 1, 64, 0, INV>(lds, rows, P, TW, dir);
        }
    }
}

// the pass ct_passes<..., FUSED_TAIL = true> leaves out, run with the store
template <int SHAPE, bool INV>
__device__ __forceinline__ void ct_store_tail(const int64_t* lds, int rows, const DevPlan* P, const CsdBarrett* TW, int dir,
                                              int64_t* out, int64_t st2, const int64_t* mul, int64_t ms2, int64_t ms3,
                                              const CsdMod& md) {
    if constexpr (SHAPE == 1) {
        if constexpr (!INV)
            ct_pass_store<3072, 3, 2, 3, 1, 8, 3, INV>(lds, rows, P, TW, dir, out, st2, mul, ms2, ms3, md);
        else
            ct_pass_store<3072, 3, 2, 3, 1, 64, 0, INV>(lds, rows, P, TW, dir, out, st2, mul, ms2, ms3, md);
    } else {
        if constexpr (!INV)
            ct_pass_store<12288, 2, 2, 3, 1, 8, 3, INV>(lds, rows, P, TW, dir, out, st2, mul, ms2, ms3, md);
        else
            ct_pass_store<12288, 2, 2, 3, 1, 64, 0, INV>(lds, rows, P, TW, dir, out, st2, mul, ms2, ms3, md);
    }
}

// Global -> LDS load of `rows` lines fused with the leading single-stage passes of a compiled
// shape: SHAPE 1 runs factor 0 (C_2, suffix 1536) and factor 1 (C_3, suffix 512) on the six
// elements e + 512 m (m < 6) of each suffix e < 512 in registers; SHAPE 2 runs factor 0 (C_3,
// suffix 4096) on the three elements e + 4096 m. The groups, twiddles (block 0, uniform) and
// schedule bits are those of the corresponding ct_pass calls.
template <int SHAPE, bool INV>
__device__ __forceinline__ void ct_load_head(int64_t* __restrict__ lds, const int64_t* __restrict__ base, int64_t st2, int rows,
                                             const DevPlan* __restrict__ P, const CsdBarrett* __restrict__ TW, int dir) {
    if constexpr (SHAPE == 1) {
        constexpr uint32_t L = 3072, E = 512, lline = ct_line(L);
        const DevFactor& F0 = P->f[0];
        const DevFactor& F1 = P->f[1];
        for (uint32_t w = threadIdx.x; w < E * (uint32_t)rows; w += blockDim.x) {
            const uint32_t r = w / E, e = w % E;
            int64_t v[6];
#pragma unroll
            for (int m = 0; m < 6; m++) v[m] = base[r * st2 + e + E * m];
#pragma unroll
            for (int m = 0; m < 3; m++) {
                int64_t g[2] = {v[m], v[m + 3]};
                ct_stages<2, 1, 0, INV>(g, P, TW, F0, dir, 0u);
                v[m] = g[0];
                v[m + 3] = g[1];
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                int64_t g[3] = {v[3 * h], v[3 * h + 1], v[3 * h + 2]};
                ct_stages<3, 1, 0, INV>(g, P, TW, F1, dir, 0u);
                v[3 * h] = g[0];
                v[3 * h + 1] = g[1];
                v[3 * h + 2] = g[2];
            }
#pragma unroll
            for (int m = 0; m < 6; m++) lds[r * lline + ct_slot(e + E * m)] = v[m];
        }
    } else {
        constexpr uint32_t L = 12288, E = 4096, lline = ct_line(L);
        const DevFactor& F0 = P->f[0];
        for (uint32_t w = threadIdx.x; w < E * (uint32_t)rows; w += blockDim.x) {
            const uint32_t r = w / E, e = w % E;
            int64_t g[3];
#pragma unroll
            for (int m = 0; m < 3; m++) g[m] = base[r * st2 + e + E * m];
            ct_stages<3, 1, 0, INV>(g, P, TW, F0, dir, 0u);
#pragma unroll
            for (int m = 0; m < 3; m++) lds[r * lline + ct_slot(e + E * m)] = g[m];
        }
    }
    __syncthreads();
}

// the (factor, s0, g) pass list a compiled shape runs, for checking a plan against it
static bool ct_shape_matches(const DevPlan& P, int dir, int shape) {
    struct Pf {
        int len, S;
    };
    std::vector<Pf> fs;
    std::vector<DevPass> ps;
    if (shape == 1) {
        fs = {{2, 1536}, {3, 512}, {8, 64}, {64, 1}};
        ps = {{0, 0, 1, 2}, {1, 0, 1, 3}, {2, 0, 3, 2}};
        if (dir != 2) {
            ps.push_back({3, 0, 3, 2});
            ps.push_back({3, 3, 3, 2});
        } else {
            ps.push_back({3, 3, 3, 2});
            ps.push_back({3, 0, 3, 2});
        }
    } else if (shape == 2) {
        fs = {{3, 4096}, {64, 64}, {64, 1}};
        ps = {{0, 0, 1, 3}};
        for (uint8_t f = 1; f <= 2; f++) {
            if (dir != 2) {
                ps.push_back({f, 0, 3, 2});
                ps.push_back({f, 3, 3, 2});
            } else {
                ps.push_back({f, 3, 3, 2});
                ps.push_back({f, 0, 3, 2});
            }
        }
    } else {
        return false;
    }
    if (P.nfactors != (int)fs.size()) return false;
    for (int f = 0; f < P.nfactors; f++) {
        const DevFactor& F = P.f[f];
        if (F.len != fs[f].len || (int)F.S != fs[f].S || F.index != f) return false;
        uint32_t nb = 1;
        for (int s = 0; s < F.nstages; s++) {
            if (F.nblocks[s] != nb || F.stage_len[s] != (uint32_t)F.len / nb) return false;
            nb *= (uint32_t)F.radix;
        }
    }
    if (P.npasses[dir] != (int)ps.size()) return false;
    for (size_t q = 0; q < ps.size(); q++) {
        const DevPass& a = P.passes[dir][q];
        if (a.factor != ps[q].factor || a.s0 != ps[q].s0 || a.g != ps[q].g || a.radix != ps[q].radix) return false;
    }
    return true;
}

// two consecutive residues, for 16-byte LDS and global accesses
struct __align__(16) I64x2 {
    int64_t a, b;
};

struct RowArgs {
    int64_t d0, d1, d2, d3;
    int64_t st[3];
    int64_t* data;
    const DevPlan* plans[CSD_MAX_MODULI];
    int dir;
    int rows;           // lines per workgroup
    uint32_t magic_L;   // for g / L with g < rows * L
    uint32_t magic_groups[5];  // for w / (L / M), M = 2, 4, 8, 3, 9 (pass kinds (2,1) (2,2) (2,3) (3,1) (3,2))
    int64_t chunks;     // workgroups per (i, j)
    const int64_t* mul;
    Str4 smul;
    // work-list mode (uneven batches): slice s is the [d2, d3] matrix at seg_base[s] with limb
    // index seg_limb[s]; blocks enumerate (slice, chunk) instead of (i, j, chunk)
    int64_t* const* seg_base;
    const int32_t* seg_limb;
};

// dim-3 transforms (contiguous lines): lines staged in LDS, stages run in register blocks;
// SHAPE 0 reads the pass list from the plan, SHAPE 1 and 2 are the compiled shapes above
template <bool INV, int SHAPE = 0>
__global__ __launch_bounds__(512) void k_ntt_rows(RowArgs a, ModTable mt) {
    extern __shared__ int64_t lds[];
    const int64_t chunk = blockIdx.x % a.chunks;
    const int64_t ij = blockIdx.x / a.chunks;
    int64_t i, j;
    int64_t* base;
    const int64_t k0 = chunk * a.rows;
    if (a.seg_base) {
        i = 0;
        j = a.seg_limb[ij];
        base = a.seg_base[ij] + k0 * a.st[2];
    } else {
        j = ij % a.d1;
        i = ij / a.d1;
        base = a.data + i * a.st[0] + j * a.st[1] + k0 * a.st[2];
    }
    const DevPlan* __restrict__ P = a.plans[j];
    const CsdBarrett* __restrict__ TW = (const CsdBarrett*)(P + 1);
    int L;
    if constexpr (SHAPE == 0)
        L = P->len;
    else
        L = CtShape<SHAPE>::L;
    const uint32_t lline = SHAPE ? ct_line((uint32_t)L) : lds_line((uint32_t)L);
    const int rows = (int)min((int64_t)a.rows, a.d2 - k0);
    const int total = rows * L;
    const int dir = a.dir;
    if constexpr (SHAPE != 0) {
        ct_load_head<SHAPE, INV>(lds, base, a.st[2], rows, P, TW, dir);
        if constexpr (INV) {
            // the inverse ends on the stride-8 groups of the last factor's first stages: lanes
            // hold interleaved runs, so its register blocks go to global memory directly, coalesced
            ct_passes<SHAPE, INV, true, true>(lds, rows, P, TW, dir);
            const int64_t* mul = a.mul ? a.mul + i * a.smul.s[0] + j * a.smul.s[1] + k0 * a.smul.s[2] : nullptr;
            ct_store_tail<SHAPE, INV>(lds, rows, P, TW, dir, base, a.st[2], mul, a.smul.s[2], a.smul.s[3], mt.m[j]);
            return;
        } else {
            // the forward ends on 8 consecutive elements per lane; its store goes through LDS below
            ct_passes<SHAPE, INV, true, false>(lds, rows, P, TW, dir);
        }
    } else {
        for (int g = threadIdx.x; g < total; g += blockDim.x) {
            const uint32_t r = a.rows == 1 ? 0u : fdiv((uint32_t)g, a.magic_L);
            const int p = g - (int)r * L;
            lds[r * lline + lds_slot(p)] = base[r * a.st[2] + p];
        }
        __syncthreads();
    }
    if constexpr (SHAPE != 0) {
        // The forward's store. The lines sit in LDS in natural order; when the lines and the
        // multiplier rows are 16-byte aligned with unit stride, pairs of consecutive elements move
        // as 16-byte accesses from row addresses formed once per line (the per-element form below
        // spends more instructions on its 64-bit index arithmetic than on the product).
        const int64_t* mrow = a.mul ? a.mul + i * a.smul.s[0] + j * a.smul.s[1] + k0 * a.smul.s[2] : nullptr;
        const bool pairs = (((uintptr_t)base | (uintptr_t)(a.st[2] * 8)) & 15) == 0 &&
                           (!mrow || ((((uintptr_t)mrow | (uintptr_t)(a.smul.s[2] * 8)) & 15) == 0 && a.smul.s[3] == 1));
        if (pairs) {
            const CsdMod m = mt.m[j];
            for (int r = 0; r < rows; r++) {
                const I64x2* src = (const I64x2*)(lds + r * lline);
                I64x2* dst = (I64x2*)(base + r * a.st[2]);
                const I64x2* mr = mrow ? (const I64x2*)(mrow + r * a.smul.s[2]) : nullptr;
                for (int q = threadIdx.x; q < L / 2; q += blockDim.x) {
                    I64x2 x = src[q];
                    if (mr) {
                        const I64x2 b = mr[q];
                        x.a = csd_mont_mul(m, x.a, b.a);
                        x.b = csd_mont_mul(m, x.b, b.b);
                    }
                    dst[q] = x;
                }
            }
            return;
        }
    } else if (P->nfactors == 0) {
        if (dir != 1) {
            const CsdBarrett m = dir == 2 ? P->inv_len : P->one;
            for (int g = threadIdx.x; g < total; g += blockDim.x) {
                const uint32_t r = a.rows == 1 ? 0u : fdiv((uint32_t)g, a.magic_L);
                const uint32_t s = r * lline + lds_slot(g - r * L);
                lds[s] = csd_barrett_folded(m, lds[s], P->n);
            }
        }
    } else {
        const int np = P->npasses[dir];
        for (int q = 0; q < np; q++) {
            const DevPass ps = P->passes[dir][q];
            if (ps.radix == 2) {
                if (ps.g == 3)
                    run_pass<2, 3, INV>(lds, L, rows, a.magic_groups[2], P, TW, ps, dir);
                else if (ps.g == 2)
                    run_pass<2, 2, INV>(lds, L, rows, a.magic_groups[1], P, TW, ps, dir);
                else
                    run_pass<2, 1, INV>(lds, L, rows, a.magic_groups[0], P, TW, ps, dir);
            } else {
                if (ps.g == 2)
                    run_pass<3, 2, INV>(lds, L, rows, a.magic_groups[4], P, TW, ps, dir);
                else
                    run_pass<3, 1, INV>(lds, L, rows, a.magic_groups[3], P, TW, ps, dir);
            }
            __syncthreads();
        }
    }
    const CsdMod m = mt.m[j];
    for (int g = threadIdx.x; g < total; g += blockDim.x) {
        const uint32_t r = a.rows == 1 ? 0u : (SHAPE ? (uint32_t)g / (uint32_t)L : fdiv((uint32_t)g, a.magic_L));
        const int p = g - (int)r * L;
        int64_t x = lds[r * lline + (SHAPE ? ct_slot((uint32_t)p) : lds_slot((uint32_t)p))];
        if (a.mul)
            x = csd_mont_mul(m, x, a.mul[i * a.smul.s[0] + j * a.smul.s[1] + (k0 + r) * a.smul.s[2] + p * a.smul.s[3]]);
        base[r * a.st[2] + p] = x;
    }
}

// ---------------------------------------------------------------------------------------------
// field_49152's K part ([3, 64, 64], lines of 12288) through 32 KiB of LDS. Factor 0 (C_3, suffix
// 4096) is a single stage whose butterfly links the elements e, e + 4096, e + 8192; after it,
// factors 1 and 2 act inside each 4096-element sub-line (their twiddle blocks are taken mod the
// stage's block count, so the prefix does not enter). The workgroup of a line runs the C_3
// butterflies on the load, keeps sub-line 0 in LDS and writes sub-lines 1 and 2 back in place (the
// line is its own), then transforms the three sub-lines one after another, reloading 1 and 2 from
// L2. A whole line in LDS (96 KiB) leaves one workgroup per CU; sub-lines leave five.
template <bool INV>
__global__ __launch_bounds__(256) void k_ntt_rows_split3(RowArgs a, ModTable mt) {
    extern __shared__ int64_t lds[];
    constexpr uint32_t E = 4096;
    constexpr int L = E;
    const int64_t row = blockIdx.x % a.d2;
    const int64_t ij = blockIdx.x / a.d2;
    const int64_t j = ij % a.d1, i = ij / a.d1;
    int64_t* base = a.data + i * a.st[0] + j * a.st[1] + row * a.st[2];
    const DevPlan* __restrict__ P = a.plans[j];
    const CsdBarrett* __restrict__ TW = (const CsdBarrett*)(P + 1);
    const int dir = a.dir;
    {
        const DevFactor& F0 = P->f[0];
        for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) {
            int64_t g[3] = {base[e], base[E + e], base[2 * E + e]};
            ct_stages<3, 1, 0, INV>(g, P, TW, F0, dir, 0u);
            lds[e] = g[0];
            base[E + e] = g[1];
            base[2 * E + e] = g[2];
        }
        __syncthreads();
    }
    const CsdMod m = mt.m[j];
    for (int k = 0; k < 3; k++) {
        if (k > 0) {
            for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) lds[e] = base[k * E + e];
            __syncthreads();
        }
        int64_t* out = base + (int64_t)k * E;
        const int64_t* mul = a.mul ? a.mul + i * a.smul.s[0] + j * a.smul.s[1] + row * a.smul.s[2] + (int64_t)k * E * a.smul.s[3]
                                   : nullptr;
        if constexpr (!INV) {
            ct_pass<L, 1, 2, 3, 64, 64, 0, INV>(lds, 1, P, TW, dir);
            ct_pass<L, 1, 2, 3, 64, 8, 3, INV>(lds, 1, P, TW, dir);
            ct_pass<L, 2, 2, 3, 1, 64, 0, INV>(lds, 1, P, TW, dir);
            ct_pass<L, 2, 2, 3, 1, 8, 3, INV>(lds, 1, P, TW, dir);
            for (uint32_t p = threadIdx.x; p < E; p += blockDim.x) {
                int64_t x = lds[p];
                if (mul) x = csd_mont_mul(m, x, mul[p * a.smul.s[3]]);
                out[p] = x;
            }
        } else {
            ct_pass<L, 1, 2, 3, 64, 8, 3, INV>(lds, 1, P, TW, dir);
            ct_pass<L, 1, 2, 3, 64, 64, 0, INV>(lds, 1, P, TW, dir);
            ct_pass<L, 2, 2, 3, 1, 8, 3, INV>(lds, 1, P, TW, dir);
            ct_pass_store<L, 2, 2, 3, 1, 64, 0, INV>(lds, 1, P, TW, dir, out, 0, mul, 0, a.smul.s[3], m);
        }
        __syncthreads();  // the next sub-line overwrites the LDS the store read
    }
}

// dim-2 transforms of short length: one thread owns a whole column (F0 * F1 values) in registers.
// F0, F1 are the factor lengths (F1 = 1 for single-factor plans, F0 = F1 = 1 for the identity).
template <int F>
struct ColFactorShape {
    static constexpr int R = (F & (F - 1)) == 0 ? 2 : 3;
    static constexpr int NST = R == 2 ? (F == 2 ? 1 : F == 4 ? 2 : F == 8 ? 3 : F == 16 ? 4 : 0) : (F == 3 ? 1 : F == 9 ? 2 : 0);
};

// stage s of factor f (length F, suffix S, PRE prefixes) on a register-resident column
template <int F, int S, int PRE, bool INV, int s>
__device__ __forceinline__ void col_stage(int64_t (&v)[F * S * PRE], const DevPlan* __restrict__ P,
                                          const CsdBarrett* __restrict__ TW, const DevFactor& Fd, int f, int dir) {
    constexpr int R = ColFactorShape<F>::R;
    constexpr int SL = F / Pow<R>::of(s);
    constexpr int part = SL / R;
    constexpr int d = part * S;
    const int64_t n = P->n;
    const uint8_t info = P->info[dir][f][s];
    if (info & INFO_REDUCE) {
#pragma unroll
        for (int e = 0; e < F * S * PRE; e++) v[e] = csd_barrett_folded(P->one, v[e], n);
    }
#pragma unroll
    for (int b = 0; b < F / SL; b++) {
        CsdBarrett t2;
        const CsdBarrett* t3;
        if constexpr (R == 2)
            t2 = TW[(INV ? Fd.itw_off[s] : Fd.tw_off[s]) + b];
        else
            t3 = TW + (INV ? Fd.itw_off[s] + 2 * b : Fd.tw_off[s] + 6 * b);
#pragma unroll
        for (int pre = 0; pre < PRE; pre++)
#pragma unroll
            for (int jj = 0; jj < part; jj++)
#pragma unroll
                for (int x = 0; x < S; x++) {
                    constexpr int stride_pre = F * S;
                    const int e0 = pre * stride_pre + (b * SL + jj) * S + x;
                    if constexpr (R == 2) {
                        if constexpr (!INV)
                            bf2_fwd(v[e0], v[e0 + d], t2, info, P->one, n);
                        else
                            bf2_inv(v[e0], v[e0 + d], t2, info, P->inv_len, n);
                    } else {
                        if constexpr (!INV)
                            bf3_fwd(v[e0], v[e0 + d], v[e0 + 2 * d], t3, info, P->one, n);
                        else
                            bf3_inv(v[e0], v[e0 + d], v[e0 + 2 * d], t3, info, P->inv_len, Fd.zeta, Fd.zeta2, n);
                    }
                }
    }
}

template <int F, int S, int PRE, bool INV, int SI>
__device__ __forceinline__ void col_stages(int64_t (&v)[F * S * PRE], const DevPlan* P, const CsdBarrett* TW, const DevFactor& Fd,
                                           int f, int dir) {
    constexpr int NST = ColFactorShape<F>::NST;
    if constexpr (SI < NST) {
        col_stage<F, S, PRE, INV, INV ? NST - 1 - SI : SI>(v, P, TW, Fd, f, dir);
        col_stages<F, S, PRE, INV, SI + 1>(v, P, TW, Fd, f, dir);
    }
}

template <int F, int S, int PRE, bool INV>
__device__ __forceinline__ void col_factor(int64_t (&v)[F * S * PRE], const DevPlan* __restrict__ P,
                                           const CsdBarrett* __restrict__ TW, int f, int dir) {
    static_assert(ColFactorShape<F>::NST > 0, "unsupported column factor");
    col_stages<F, S, PRE, INV, 0>(v, P, TW, P->f[f], f, dir);
}

struct ColArgs {
    int64_t d0, d1, d3;
    int64_t st[3];
    int64_t* data;
    const DevPlan* plans[CSD_MAX_MODULI];
    int dir;
    int64_t chunks;
    const int64_t* mul;
    Str4 smul;
    int64_t* const* seg_base;  // work-list mode, as in RowArgs
    const int32_t* seg_limb;
};

template <int F0, int F1, bool INV>
__global__ __launch_bounds__(256) void k_ntt_cols(ColArgs a, ModTable mt) {
    constexpr int L = F0 * F1;
    const int64_t chunk = blockIdx.x % a.chunks;
    const int64_t ij = blockIdx.x / a.chunks;
    const int64_t l = chunk * blockDim.x + threadIdx.x;
    if (l >= a.d3) return;
    int64_t i, j;
    int64_t* base;
    if (a.seg_base) {
        i = 0;
        j = a.seg_limb[ij];
        base = a.seg_base[ij] + l;
    } else {
        j = ij % a.d1;
        i = ij / a.d1;
        base = a.data + i * a.st[0] + j * a.st[1] + l;
    }
    const DevPlan* __restrict__ P = a.plans[j];
    const CsdBarrett* __restrict__ TW = (const CsdBarrett*)(P + 1);
    int64_t v[L];
#pragma unroll
    for (int k = 0; k < L; k++) v[k] = base[k * a.st[2]];
    const int dir = a.dir;
    if constexpr (L == 1) {
        if (dir != 1) v[0] = csd_barrett_folded(dir == 2 ? P->inv_len : P->one, v[0], P->n);
    } else {
        // factor 0 acts on the most significant digit (suffix F1), then factor 1 (suffix 1)
        col_factor<F0, F1, 1, INV>(v, P, TW, 0, dir);
        if constexpr (F1 > 1) col_factor<F1, 1, F0, INV>(v, P, TW, 1, dir);
    }
    const CsdMod m = mt.m[j];
    if (a.mul && a.smul.s[3] == 0) {
        // multipliers constant along dim 3 (the ring's L' kernels): the workgroup's L values are
        // wave-uniform, loaded together before the products instead of one per product
        const int64_t* __restrict__ mrow = a.mul + i * a.smul.s[0] + j * a.smul.s[1];
        int64_t mk[L];
#pragma unroll
        for (int k = 0; k < L; k++) mk[k] = mrow[k * a.smul.s[2]];
#pragma unroll
        for (int k = 0; k < L; k++) base[k * a.st[2]] = csd_mont_mul(m, v[k], mk[k]);
        return;
    }
#pragma unroll
    for (int k = 0; k < L; k++) {
        int64_t x = v[k];
        if (a.mul) x = csd_mont_mul(m, x, a.mul[i * a.smul.s[0] + j * a.smul.s[1] + k * a.smul.s[2] + l * a.smul.s[3]]);
        base[k * a.st[2]] = x;
    }
}

typedef void (*ColKernel)(ColArgs, ModTable);

template <int F0, int F1>
static ColKernel col_kernel_for(bool inv) {
    return inv ? (ColKernel)k_ntt_cols<F0, F1, true> : (ColKernel)k_ntt_cols<F0, F1, false>;
}

// the column kernel for a plan's factor list, or null if it has no register-resident variant
static ColKernel pick_col_kernel(const DevPlan& P, bool inv) {
    int f0 = P.nfactors > 0 ? P.f[0].len : 1, f1 = P.nfactors > 1 ? P.f[1].len : 1;
    if (P.nfactors > 2) return nullptr;
    if (f0 == 4 && f1 == 3) return col_kernel_for<4, 3>(inv);
    if (f0 == 1 && f1 == 1) return col_kernel_for<1, 1>(inv);
    if (f1 != 1) return nullptr;
    switch (f0) {
        case 2: return col_kernel_for<2, 1>(inv);
        case 3: return col_kernel_for<3, 1>(inv);
        case 4: return col_kernel_for<4, 1>(inv);
        case 8: return col_kernel_for<8, 1>(inv);
        case 9: return col_kernel_for<9, 1>(inv);
        case 16: return col_kernel_for<16, 1>(inv);
        default: return nullptr;
    }
}

// ctx->row_lines / row_threads: dim-3 lines and threads per workgroup, 0 = automatic
extern "C" int csd_ntt_set_row_shape(csd_ctx* ctx, int lines, int threads) {
    CSD_TRY if (!ctx) throw CsdError("null context");
    if (lines < 0 || lines > 8 || threads < 0 || threads > 512 || threads % 64 != 0)
        throw CsdError("csd_ntt_set_row_shape: lines must be 0-8, threads 0 or a multiple of 64 up to 512");
    ctx->row_lines = lines;
    ctx->row_threads = threads;
    CSD_CATCH
}

// ctx->ntt_mode: 1 forces the one-stage-per-LDS-pass kernel, 2 keeps dim-3 transforms on the
// plan-driven row kernel, 3 runs field_49152's lines whole instead of as sub-lines (tests)
extern "C" int csd_ntt_force_generic(csd_ctx* ctx, int mode) {
    CSD_TRY if (!ctx) throw CsdError("null context");
    if (mode < 0 || mode > 3) throw CsdError("csd_ntt_force_generic: mode must be 0, 1, 2 or 3");
    ctx->ntt_mode = mode;
    CSD_CATCH
}

typedef void (*RowKernel)(RowArgs, ModTable);

// the dim-3 row kernel for these plans: a compiled shape when every plan matches one, else the
// plan-driven form; the dynamic LDS limit is raised once per kernel
static RowKernel pick_row_kernel(csd_ctx* ctx, csd_ntt_plan* const* plans, size_t nplans, int dir, int* shape_out = nullptr) {
    static const RowKernel ks[2][3] = {{k_ntt_rows<false, 0>, k_ntt_rows<false, 1>, k_ntt_rows<false, 2>},
                                       {k_ntt_rows<true, 0>, k_ntt_rows<true, 1>, k_ntt_rows<true, 2>}};
    for (auto& row : ks)
        for (RowKernel k : row) csd_allow_lds((const void*)k, 160 * 1024);
    int shape = 0;
    if (ctx->ntt_mode != 2 && nplans > 0) {
        for (int c : {1, 2}) {
            bool all = true;
            for (size_t jj = 0; jj < nplans && all; jj++) all = ct_shape_matches(plans[jj]->host, dir, c);
            if (all) {
                shape = c;
                break;
            }
        }
    }
    if (shape_out) *shape_out = shape;
    return ks[dir == 2 ? 1 : 0][shape];
}

// LDS words per line of the row kernel of the given compiled shape (0: the plan-driven kernel)
static uint32_t row_lds_line(int shape, uint32_t L) { return shape ? ct_line(L) : lds_line(L); }

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
    static const char* names[2][3] = {{"ntt_dim2", "ntt_mul_dim2", "intt_dim2"}, {"ntt_dim3", "ntt_mul_dim3", "intt_dim3"}};
    const bool inv = dir == 2;
    if (ctx->ntt_mode != 1 && dim == 2) {
        // register-resident columns, if every plan has the same supported factor list
        ColKernel k = pick_col_kernel(plans[0]->host, inv);
        for (size_t jj = 1; jj < nplans && k; jj++) {
            const DevPlan& p0 = plans[0]->host;
            const DevPlan& pj = plans[jj]->host;
            if (pj.nfactors != p0.nfactors) k = nullptr;
            for (int f = 0; k && f < p0.nfactors; f++)
                if (pj.f[f].len != p0.f[f].len) k = nullptr;
        }
        if (k) {
            ColArgs c;
            memset(&c, 0, sizeof(c));
            c.d0 = a.d0;
            c.d1 = a.d1;
            c.d3 = a.d3;
            for (int d = 0; d < 3; d++) c.st[d] = a.st[d];
            c.data = a.data;
            for (size_t jj = 0; jj < nplans; jj++) c.plans[jj] = a.plans[jj];
            c.dir = dir;
            c.chunks = (a.d3 + 255) / 256;
            c.mul = a.mul;
            c.smul = a.smul;
            int64_t blocks = a.d0 * a.d1 * c.chunks;
            if (blocks > (int64_t)INT32_MAX) throw CsdError("ntt: too many tiles");
            TimedScope ts(ctx, names[0][dir], 16.0 * (double)numel(dst));
            hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, c, t);
            HIPCHECK(hipGetLastError());
            return;
        }
    }
    if (ctx->ntt_mode != 1 && dim == 3 && L <= 16384) {
        RowArgs r;
        memset(&r, 0, sizeof(r));
        r.d0 = a.d0;
        r.d1 = a.d1;
        r.d2 = a.d2;
        r.d3 = a.d3;
        for (int d = 0; d < 3; d++) r.st[d] = a.st[d];
        r.data = a.data;
        for (size_t jj = 0; jj < nplans; jj++) r.plans[jj] = a.plans[jj];
        r.dir = dir;
        // Lines of 3072 (field_36864): one line per 256-thread workgroup, measured fastest on MI355X
        // (24 KiB of LDS, so six workgroups and 24 waves per CU hide the latency of the Barrett
        // chains; with the padded layout, two lines per 512 threads ran 8-14 % slower). Other
        // lengths keep about 6144 coefficients per workgroup.
        const bool measured_3072 = L == 3072;
        int shape = 0;
        RowKernel rk = pick_row_kernel(ctx, plans, nplans, dir, &shape);
        if (shape == 2 && ctx->ntt_mode == 0 && ctx->row_lines == 0 && ctx->row_threads == 0) {
            // field_49152's lines through sub-lines of 32 KiB (k_ntt_rows_split3)
            r.mul = a.mul;
            r.smul = a.smul;
            const int64_t grid = a.d0 * a.d1 * a.d2;
            if (grid > (int64_t)INT32_MAX) throw CsdError("ntt: too many tiles");
            csd_allow_lds((const void*)k_ntt_rows_split3<false>, 32 * 1024);
            csd_allow_lds((const void*)k_ntt_rows_split3<true>, 32 * 1024);
            TimedScope ts(ctx, names[1][dir], 16.0 * (double)numel(dst));
            if (dir == 2)
                hipLaunchKernelGGL(k_ntt_rows_split3<true>, dim3((unsigned)grid), dim3(256), 32 * 1024, ctx->stream, r, t);
            else
                hipLaunchKernelGGL(k_ntt_rows_split3<false>, dim3((unsigned)grid), dim3(256), 32 * 1024, ctx->stream, r, t);
            HIPCHECK(hipGetLastError());
            return;
        }
        const uint32_t lline = row_lds_line(shape, (uint32_t)L);
        int rows = 1;
        while (!measured_3072 && rows * 2 * L <= 6144 && rows * 2 <= a.d2 && rows < 256) rows *= 2;
        const int row_lines = ctx->row_lines, row_threads = ctx->row_threads;
        if (row_lines > 0) {
            // the requested lines, capped by the batch and by what fits the 160 KiB of LDS
            const int64_t fit = std::max<int64_t>(1, (160 * 1024) / ((int64_t)lline * 8));
            rows = (int)std::min<int64_t>(std::min<int64_t>(row_lines, a.d2), fit);
        }
        r.rows = rows;
        // block size: the candidate with the best (resident threads per CU) x (fraction of
        // non-idle iterations over this plan's passes); measured on lines of 3072, two lines per
        // 256-thread block beat one line per 128 or 192 threads
        int threads = 256;
        {
            const DevPlan& P0 = plans[0]->host;
            const size_t lds_block = (size_t)rows * lline * 8;
            double best = -1;
            for (int T : {256, 512}) {
                double busy = 0, slots = 0;
                for (int q = 0; q < P0.npasses[dir]; q++) {
                    const DevPass& ps = P0.passes[dir][q];
                    int M = ps.radix == 2 ? (1 << ps.g) : (ps.g == 1 ? 3 : 9);
                    double groups = (double)rows * (double)(L / M);
                    busy += groups;
                    slots += std::ceil(groups / T) * T;
                }
                int by_lds = (int)std::max<size_t>(1, (160 * 1024) / lds_block);
                int by_threads = 2048 / T;
                double resident = (double)std::min(by_lds, by_threads) * T;
                double score = resident * (slots > 0 ? busy / slots : 1.0);
                if (score > best + 1e-9) {
                    best = score;
                    threads = T;
                }
            }
        }
        if (measured_3072 && row_lines == 0) threads = 256;
        if (row_threads > 0) threads = row_threads;
        r.magic_L = magic_for((uint32_t)L);
        const int Ms[5] = {2, 4, 8, 3, 9};
        for (int q = 0; q < 5; q++) r.magic_groups[q] = L % Ms[q] == 0 ? magic_for((uint32_t)(L / Ms[q])) : 0u;
        r.chunks = (a.d2 + rows - 1) / rows;
        r.mul = a.mul;
        r.smul = a.smul;
        int64_t blocks = a.d0 * a.d1 * r.chunks;
        if (blocks > (int64_t)INT32_MAX) throw CsdError("ntt: too many tiles");
        size_t lds = (size_t)rows * lline * 8;
        if (lds > 160 * 1024) throw CsdError("ntt: line too long for LDS staging");
        TimedScope ts(ctx, names[1][dir], 16.0 * (double)numel(dst));
        hipLaunchKernelGGL(rk, dim3((unsigned)blocks), dim3(threads), lds, ctx->stream, r, t);
        HIPCHECK(hipGetLastError());
        return;
    }
    // generic form: one stage per LDS pass; a tile holds as many transforms as fit 32 KiB
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
    csd_allow_lds((const void*)k_ntt, 160 * 1024);
    static const char* gnames[2][3] = {{"generic_ntt_dim2", "generic_ntt_mul_dim2", "generic_intt_dim2"},
                                       {"generic_ntt_dim3", "generic_ntt_mul_dim3", "generic_intt_dim3"}};
    TimedScope ts(ctx, gnames[dim - 2][dir], 16.0 * (double)numel(dst));
    hipLaunchKernelGGL(k_ntt, dim3((unsigned)blocks), dim3(NTT_THREADS), lds, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
}

// Work-list transforms for uneven batches: `nslices` matrices [d2, d3] (row stride st2) at the
// device addresses seg_base[s], each transformed with plans[seg_limb[s]] along `dim` in direction
// `dir` (0 forward reduced, 1 forward unreduced, 2 inverse), optionally followed by the
// Montgomery multiplication with mul[0, limb, ., .] (broadcast over its length-1 dimensions).
// One launch for the whole work-list; the per-slice results equal the per-tensor HAL calls.
void csd_ntt_segments(csd_ctx* ctx, int64_t* const* seg_base, const int32_t* seg_limb, int64_t nslices, int64_t d2, int64_t d3,
                      int64_t st2, csd_ntt_plan* const* plans, size_t nplans, int dim, int dir, const csd_view* mul,
                      const csd_modulus* mods, const char* timer, const CsdSegTune& tune) {
    if (nplans > CSD_MAX_MODULI) throw CsdError("too many plans");
    const int64_t L = dim == 2 ? d2 : d3;
    for (size_t jj = 0; jj < nplans; jj++)
        if (plans[jj]->host.len != L) throw CsdError("work-list ntt: plan length mismatch");
    ModTable t;
    memset(&t, 0, sizeof(t));
    Str4 smul{{0, 0, 0, 0}};
    if (mul) {
        if (mul->shape[0] != 1 || mul->shape[1] != (int64_t)nplans) throw CsdError("work-list ntt: multiplier shape");
        for (int d = 1; d < 4; d++) smul.s[d] = mul->shape[d] == 1 ? 0 : mul->strides[d];
        if ((mul->shape[2] != 1 && mul->shape[2] != d2) || (mul->shape[3] != 1 && mul->shape[3] != d3))
            throw CsdError("work-list ntt: multiplier shape");
        t = make_mod_table(mods, nplans, (int64_t)nplans);
    }
    if (nslices == 0) return;
    const double bytes = 16.0 * (double)nslices * d2 * d3;
    const bool inv = dir == 2;
    if (dim == 2) {
        ColKernel k = pick_col_kernel(plans[0]->host, inv);
        if (!k) throw CsdError("work-list ntt: no column kernel for this plan");
        for (size_t jj = 1; jj < nplans; jj++) {
            const DevPlan &p0 = plans[0]->host, &pj = plans[jj]->host;
            bool same = pj.nfactors == p0.nfactors;
            for (int f = 0; same && f < p0.nfactors; f++) same = pj.f[f].len == p0.f[f].len;
            if (!same) throw CsdError("work-list ntt: plans must share one factor list");
        }
        const int ct = tune.col_threads;
        if (ct < 64 || ct > 256 || ct % 64) throw CsdError("work-list ntt: column block size must be 64, 128, 192 or 256");
        ColArgs c;
        memset(&c, 0, sizeof(c));
        c.d0 = 1;
        c.d1 = (int64_t)nplans;
        c.d3 = d3;
        c.st[2] = st2;
        for (size_t jj = 0; jj < nplans; jj++) c.plans[jj] = (const DevPlan*)plans[jj]->dev;
        c.dir = dir;
        c.chunks = (d3 + ct - 1) / ct;
        c.mul = mul ? (const int64_t*)mul->ptr : nullptr;
        c.smul = smul;
        c.seg_base = seg_base;
        c.seg_limb = seg_limb;
        if (nslices * c.chunks > (int64_t)INT32_MAX) throw CsdError("work-list ntt: too many tiles");
        TimedScope ts(ctx, timer, bytes);
        hipLaunchKernelGGL(k, dim3((unsigned)(nslices * c.chunks)), dim3(ct), 0, ctx->stream, c, t);
        HIPCHECK(hipGetLastError());
        return;
    }
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.d0 = 1;
    r.d1 = (int64_t)nplans;
    r.d2 = d2;
    r.d3 = d3;
    r.st[2] = st2;
    for (size_t jj = 0; jj < nplans; jj++) r.plans[jj] = (const DevPlan*)plans[jj]->dev;
    r.dir = dir;
    int rows = 1;
    if (tune.rows > 0)
        rows = (int)std::min<int64_t>(tune.rows, d2);
    else
        while (rows * 2 * L <= 6144 && rows * 2 <= d2 && rows < 256) rows *= 2;
    r.rows = rows;
    r.magic_L = magic_for((uint32_t)L);
    const int Ms[5] = {2, 4, 8, 3, 9};
    for (int q = 0; q < 5; q++) r.magic_groups[q] = L % Ms[q] == 0 ? magic_for((uint32_t)(L / Ms[q])) : 0u;
    r.chunks = (d2 + rows - 1) / rows;
    r.mul = mul ? (const int64_t*)mul->ptr : nullptr;
    r.smul = smul;
    r.seg_base = seg_base;
    r.seg_limb = seg_limb;
    int shape = 0;
    RowKernel rk = pick_row_kernel(ctx, plans, nplans, dir, &shape);
    size_t lds = (size_t)rows * row_lds_line(shape, (uint32_t)L) * 8;
    if (lds > 160 * 1024) throw CsdError("work-list ntt: line too long for LDS staging");
    const int threads = tune.threads > 0 ? tune.threads : (L * rows >= 8192 ? 512 : 256);
    if (threads % 64 || threads > 512) throw CsdError("work-list ntt: row block size must be a multiple of 64, at most 512");
    if (nslices * r.chunks > (int64_t)INT32_MAX) throw CsdError("work-list ntt: too many tiles");
    TimedScope ts(ctx, timer, bytes);
    hipLaunchKernelGGL(rk, dim3((unsigned)(nslices * r.chunks)), dim3(threads), lds, ctx->stream, r, t);
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

// ---------------------------------------------------------------------------------------------
// The transform chain of the batch plan as one persistent launch: each workgroup owns whole
// (element, limb) slices, s = blockIdx.x, blockIdx.x + gridDim.x, ..., and runs all eight passes
// of pb_to_crt -> plaintext product -> crt_to_pb on a slice before taking the next one, so a
// slice moves between passes through the L2 / Infinity Cache instead of HBM. The passes are the
// column and compiled-shape row passes above, in the order and with the schedules of the
// separate launches, so the results are identical. Workgroup barriers order the passes (all
// waves of a workgroup share one CU's vector L1, so its global writes are visible after them).

struct ChainArgs {
    int64_t* const* seg_base;
    const int32_t* seg_limb;
    int64_t nslices;
    const DevPlan* plp[CSD_MAX_MODULI];
    const DevPlan* pk[CSD_MAX_MODULI];
    const int64_t *klp, *kk, *iklp, *ikk;  // [R][n2] and [R][n3] CRT-kernel multipliers
    const int64_t* pt;                     // [R][n2][n3] plaintext, CRT basis
    int rows;                              // lines per LDS round of the row passes
};

// one column pass (dim 2) of a slice: thread-owned columns, the whole plan in registers
template <int F0, int F1, bool INV>
__device__ __forceinline__ void chain_cols(int64_t* __restrict__ base, int n3, const DevPlan* __restrict__ P, int dir,
                                           const int64_t* __restrict__ mul, const CsdMod& m) {
    constexpr int L = F0 * F1;
    const CsdBarrett* __restrict__ TW = (const CsdBarrett*)(P + 1);
    for (int l = threadIdx.x; l < n3; l += blockDim.x) {
        int64_t v[L];
#pragma unroll
        for (int k = 0; k < L; k++) v[k] = base[k * n3 + l];
        col_factor<F0, F1, 1, INV>(v, P, TW, 0, dir);
        if constexpr (F1 > 1) col_factor<F1, 1, F0, INV>(v, P, TW, 1, dir);
#pragma unroll
        for (int k = 0; k < L; k++) {
            int64_t x = v[k];
            if (mul) x = csd_mont_mul(m, x, mul[k]);
            base[k * n3 + l] = x;
        }
    }
}

// one row pass (dim 3) of a slice: `rows` lines at a time through LDS; mul_row (length n3) is
// shared by all rows, mul_mat ([n2][n3]) is per row; at most one of them is set
template <int SHAPE, bool INV>
__device__ __forceinline__ void chain_rows(int64_t* __restrict__ lds, int64_t* __restrict__ base, int n2, int rows_per,
                                           const DevPlan* __restrict__ P, int dir, const int64_t* __restrict__ mul_row,
                                           const int64_t* __restrict__ mul_mat, const CsdMod& m) {
    constexpr int L = CtShape<SHAPE>::L;
    constexpr uint32_t lline = ct_line(L);
    const CsdBarrett* __restrict__ TW = (const CsdBarrett*)(P + 1);
    for (int k0 = 0; k0 < n2; k0 += rows_per) {
        const int rows = min(rows_per, n2 - k0);
        const int total = rows * L;
        for (int g = threadIdx.x; g < total; g += blockDim.x) {
            const int r = g / L, p = g - r * L;
            lds[r * lline + ct_slot(p)] = base[(k0 + r) * L + p];
        }
        __syncthreads();
        ct_passes<SHAPE, INV>(lds, rows, P, TW, dir);
        for (int g = threadIdx.x; g < total; g += blockDim.x) {
            const int r = g / L, p = g - r * L;
            int64_t x = lds[r * lline + ct_slot(p)];
            if (mul_row) x = csd_mont_mul(m, x, mul_row[p]);
            if (mul_mat) x = csd_mont_mul(m, x, mul_mat[(k0 + r) * L + p]);
            base[(k0 + r) * L + p] = x;
        }
        __syncthreads();
    }
}

template <int F0, int F1, int SHAPE>
__global__ __launch_bounds__(256) void k_batch_chain(ChainArgs a, ModTable mt) {
    extern __shared__ int64_t lds[];
    constexpr int N2 = F0 * F1, N3 = CtShape<SHAPE>::L;
    for (int64_t s = blockIdx.x; s < a.nslices; s += gridDim.x) {
        int64_t* base = a.seg_base[s];
        const int j = a.seg_limb[s];
        const CsdMod m = mt.m[j];
        const DevPlan* plp = a.plp[j];
        const DevPlan* pk = a.pk[j];
        // pb_to_crt (ring.rs:1118-1141)
        chain_cols<F0, F1, false>(base, N3, plp, 1, a.klp + (int64_t)j * N2, m);
        __syncthreads();
        chain_rows<SHAPE, false>(lds, base, N2, a.rows, pk, 1, a.kk + (int64_t)j * N3, nullptr, m);
        chain_cols<F0, F1, true>(base, N3, plp, 2, nullptr, m);
        __syncthreads();
        // intt along dim 3, then the plaintext product (hal mul_mont64)
        chain_rows<SHAPE, true>(lds, base, N2, a.rows, pk, 2, nullptr, a.pt + (int64_t)j * N2 * N3, m);
        // crt_to_pb (ring.rs:1144-1161)
        chain_cols<F0, F1, false>(base, N3, plp, 1, a.iklp + (int64_t)j * N2, m);
        __syncthreads();
        chain_rows<SHAPE, false>(lds, base, N2, a.rows, pk, 1, a.ikk + (int64_t)j * N3, nullptr, m);
        chain_cols<F0, F1, true>(base, N3, plp, 2, nullptr, m);
        __syncthreads();
        chain_rows<SHAPE, true>(lds, base, N2, a.rows, pk, 2, nullptr, nullptr, m);
    }
}

// Whether the persistent chain has a compiled form for these plans: dim-2 plans of factors
// [4, 3] and dim-3 plans of compiled shape 1 (the field_36864 slices), default pass grouping.
bool csd_batch_chain_supported(csd_ntt_plan* const* plp, csd_ntt_plan* const* pk, size_t R) {
    for (size_t j = 0; j < R; j++) {
        const DevPlan& a = plp[j]->host;
        if (a.nfactors != 2 || a.f[0].len != 4 || a.f[1].len != 3) return false;
        for (int dir : {1, 2})
            if (!ct_shape_matches(pk[j]->host, dir, 1)) return false;
    }
    return R > 0 && R <= CSD_MAX_MODULI;
}

// workgroups of the persistent chain that fit one CU at once with `rows` lines of LDS
int csd_batch_chain_occupancy(int rows) {
    csd_allow_lds((const void*)k_batch_chain<4, 3, 1>, 160 * 1024);
    int n = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batch_chain<4, 3, 1>, 256, (size_t)rows * ct_line(3072) * 8));
    return std::max(n, 1);
}

void csd_batch_chain_launch(csd_ctx* ctx, int64_t* const* seg_base, const int32_t* seg_limb, int64_t nslices,
                            csd_ntt_plan* const* plp, csd_ntt_plan* const* pk, size_t R, const int64_t* klp, const int64_t* kk,
                            const int64_t* iklp, const int64_t* ikk, const int64_t* pt, const csd_modulus* mods, int rows,
                            int grid, double bytes) {
    if (!csd_batch_chain_supported(plp, pk, R)) throw CsdError("persistent chain: no compiled form for these plans");
    if (rows < 1 || rows > 4) throw CsdError("persistent chain: 1 to 4 lines per LDS round");
    if (grid < 1) throw CsdError("persistent chain: empty grid");
    ChainArgs a;
    memset(&a, 0, sizeof(a));
    a.seg_base = seg_base;
    a.seg_limb = seg_limb;
    a.nslices = nslices;
    for (size_t j = 0; j < R; j++) {
        a.plp[j] = (const DevPlan*)plp[j]->dev;
        a.pk[j] = (const DevPlan*)pk[j]->dev;
    }
    a.klp = klp;
    a.kk = kk;
    a.iklp = iklp;
    a.ikk = ikk;
    a.pt = pt;
    a.rows = rows;
    ModTable t = make_mod_table(mods, R, (int64_t)R);
    const size_t lds = (size_t)rows * ct_line(3072) * 8;
    csd_allow_lds((const void*)k_batch_chain<4, 3, 1>, 160 * 1024);
    if (nslices == 0) return;
    TimedScope ts(ctx, "batch_chain_persistent", bytes);
    hipLaunchKernelGGL((k_batch_chain<4, 3, 1>), dim3((unsigned)std::min<int64_t>(grid, nslices)), dim3(256), lds, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
}
