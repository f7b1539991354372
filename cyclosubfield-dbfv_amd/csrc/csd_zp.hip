// Exact dense linear maps over Z_p on the matrix cores.
//
// Y[v][i] = sum_k A[i][k] X[v][k] mod p for r moduli at once (one matrix per modulus), where
// the vectors X[v] run along dimension 2 or 3 of a [count, r, d2, d3] tensor. The dense forms of
// the period<->CRT transform (the group-circulant matrices of the convolutions in
// ring.rs:1118-1161) and of the cyclic transforms of bootstrap/homcrt.rs:191-214 are such maps.
//
// The int8 form: every entry is centred (|x| <= p/2 < 2^56) and split into eight balanced
// base-256 digits x = sum_d x_d 256^d, x_d in [-128, 127]. The 64 digit-pair products are
// grouped by shift s = a + b into 15 int32 accumulators C_s = sum_{a+b=s} A_a X_b, each a chain
// of v_mfma_i32_16x16x64_i8. |C_s| <= K * 8 * 2^14 < 2^31 for K-chunks of at most 16320 terms,
// and sum_s 2^(8s) C_s = sum_k A[i][k] X[v][k] exactly, |.| < K 2^112 < 2^127, so the i128
// recombination is exact; double_red128 and a multiplication by 2^192 mod p (mont.rs:57-66)
// bring it to [0, p). No floating point anywhere: the result is exact.
//
// The plain form computes the same sums with 64 x 64 -> 128-bit products on the VALU.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include <atomic>

#include <algorithm>
#include <cstring>
#include <vector>

#include "csd_internal.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));

static constexpr int ZP_DIGITS = 8;
static constexpr int64_t ZP_MAX_CHUNK = 16320;  // 64 * floor((2^31 - 1) / (8 * 2^14 * 64))
static constexpr int64_t ZP_PLAIN_CHUNK = 8192; // centred products < 2^112: 2^13 terms < 2^125

struct csd_zp_matrix {
    csd_ctx* ctx;
    int64_t r, m, k, mpad, kpad;
    int8_t* planes;   // [r][8][mpad][kpad] digit planes
    int64_t* values;  // [r][m][k] the entries as given
    std::vector<csd_modulus> mods;
    // csd_zp_matrix_set_circulant: the generator's digit planes [r][8][k] and the leading factors
    int8_t* gen = nullptr;
    int ntop = 0;
    int32_t top_orders[7] = {0}, top_strides[7] = {0};
};

// the vectors of a view along `dim`: vector v = (hi, lo) = (v / nlo, v % nlo), element k
struct VecLayout {
    int64_t s_limb, s_hi, s_lo, s_elem, nlo, nvec, nelem;
};

static VecLayout vec_layout(const csd_view* x, int dim) {
    VecLayout L;
    L.s_limb = x->strides[1];
    L.s_hi = x->strides[0];
    if (dim == 3) {
        L.s_lo = x->strides[2];
        L.s_elem = x->strides[3];
        L.nlo = x->shape[2];
        L.nelem = x->shape[3];
    } else if (dim == 2) {
        L.s_lo = x->strides[3];
        L.s_elem = x->strides[2];
        L.nlo = x->shape[3];
        L.nelem = x->shape[2];
    } else {
        throw CsdError("zp: dim must be 2 or 3");
    }
    L.nvec = x->shape[0] * L.nlo;
    return L;
}

__device__ __forceinline__ int64_t zp_centre(int64_t x, int64_t p) {
    // [-2p, 2p] -> [-p/2, p/2]
    const int64_t h = p >> 1;
    if (x > h) x -= p;
    if (x > h) x -= p;
    if (x < -h) x += p;
    if (x < -h) x += p;
    return x;
}

// v = sum_k a_k b_k (exact, any i128) -> v mod p in [0, p)
__device__ __forceinline__ int64_t zp_reduce(const CsdMod& md, i128 v) {
    int64_t t = csd_mont_mul(md, csd_double_red128(md, v), md.r3);  // 2^-128 v 2^192 2^-64 = v
    if (t < 0) t += md.N;
    if (t >= md.N) t -= md.N;
    return t;
}

// ---------------------------------------------------------------------------------------------
// digit planes: thread (l, v, q) splits elements 16q .. 16q + 15 of vector v of limb l

struct DigitArgs {
    const int64_t* src;
    VecLayout L;
    int64_t vpad, kpad;
    int8_t* dst;  // [r][8][vpad][kpad]
};

__global__ __launch_bounds__(256) void k_zp_digits(DigitArgs a, ModTable mt, int64_t total) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t nq = a.kpad / 16;
    const int64_t q = t % nq;
    const int64_t v = (t / nq) % a.vpad;
    const int64_t l = t / (nq * a.vpad);
    const int64_t p = mt.m[l].N;
    uint32_t w[ZP_DIGITS][4];
#pragma unroll
    for (int d = 0; d < ZP_DIGITS; d++)
#pragma unroll
        for (int u = 0; u < 4; u++) w[d][u] = 0;
    if (v < a.L.nvec) {
        const int64_t* base = a.src + l * a.L.s_limb + (v / a.L.nlo) * a.L.s_hi + (v % a.L.nlo) * a.L.s_lo;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int64_t k = q * 16 + e;
            int64_t c = k < a.L.nelem ? zp_centre(base[k * a.L.s_elem], p) : 0;
#pragma unroll
            for (int d = 0; d < ZP_DIGITS; d++) {
                int32_t lo = (int32_t)(c & 255);
                lo = lo >= 128 ? lo - 256 : lo;
                c = (c - lo) >> 8;
                w[d][e >> 2] |= ((uint32_t)lo & 255u) << (8 * (e & 3));
            }
        }
    }
#pragma unroll
    for (int d = 0; d < ZP_DIGITS; d++) {
        uint4 val = make_uint4(w[d][0], w[d][1], w[d][2], w[d][3]);
        *(uint4*)(a.dst + ((l * ZP_DIGITS + d) * a.vpad + v) * a.kpad + q * 16) = val;
    }
}

static void launch_digits(csd_ctx* ctx, const int64_t* src, const VecLayout& L, int64_t r, int64_t vpad, int64_t kpad,
                          int8_t* dst, const ModTable& mt) {
    DigitArgs a;
    a.src = src;
    a.L = L;
    a.vpad = vpad;
    a.kpad = kpad;
    a.dst = dst;
    const int64_t total = r * vpad * (kpad / 16);
    if (total == 0) return;
    if ((total + 255) / 256 > (int64_t)INT32_MAX) throw CsdError("zp: too many digit tiles");
    TimedScope ts(ctx, "zp_digits", 8.0 * r * L.nvec * L.nelem + (double)r * ZP_DIGITS * vpad * kpad);
    hipLaunchKernelGGL(k_zp_digits, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, a, mt, total);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// the int8 contraction. Workgroup = 4 waves = 32 rows x 64 vectors of one limb; wave w owns rows
// i0 + 16 (w & 1) .. +16 and vectors j0 + 32 (w >> 1) .. +32 as two 16 x 16 output blocks.
// MFMA operand layout: lane (x + 16 g) supplies row x of A (resp. vector x of X) and 16
// consecutive k of k-group g; the output of a block sits in lane (v + 16 g) as rows 4 g .. 4 g + 3
// of vector v. A and X use the same k assignment, so every product pairs equal k.

// one K-step's operands of a wave: 8 digit planes of its 16 rows of A, and of its 2 x 16 vectors
__device__ __forceinline__ void zp_load_step(v4i (&ra)[ZP_DIGITS], v4i (&rx)[2][ZP_DIGITS], const int8_t* A, const int8_t* X,
                                             int64_t pa, int64_t px, int64_t kpad, int64_t k0) {
#pragma unroll
    for (int d = 0; d < ZP_DIGITS; d++) ra[d] = *(const v4i*)(A + d * pa + k0);
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int d = 0; d < ZP_DIGITS; d++) rx[u][d] = *(const v4i*)(X + d * px + u * 16 * kpad + k0);
}

// the 2 x 64 digit-pair MFMAs of one K-step, accumulated by shift
__device__ __forceinline__ void zp_mfma_step(v4i (&acc)[2][2 * ZP_DIGITS - 1], const v4i (&ra)[ZP_DIGITS],
                                             const v4i (&rx)[2][ZP_DIGITS]) {
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int da = 0; da < ZP_DIGITS; da++)
#pragma unroll
            for (int dx = 0; dx < ZP_DIGITS; dx++)
                acc[u][da + dx] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ra[da], rx[u][dx], acc[u][da + dx], 0, 0, 0);
}

struct MfmaArgs {
    const int8_t* a;  // [r][8][mpad][kpad]
    const int8_t* x;  // [r][8][npad][kpad]
    int64_t* y;
    VecLayout Ly;     // output vectors (element i = row of A)
    int64_t m, mpad, npad, kpad, chunk;
    int64_t mtiles, ntiles, nblocks;
};

__global__ __launch_bounds__(256) void k_zp_mfma(MfmaArgs a, ModTable mt) {
    // XCD-aware order: the hardware deals consecutive workgroups to the 8 XCDs in turn; logical
    // tile t = (b mod 8) * (nblocks / 8) + b / 8 keeps the tiles sharing A rows on one XCD's L2
    int64_t b = blockIdx.x;
    if (a.nblocks % 8 == 0) b = (b % 8) * (a.nblocks / 8) + b / 8;
    const int64_t nt = b % a.ntiles;
    const int64_t mtile = (b / a.ntiles) % a.mtiles;
    const int64_t l = b / (a.ntiles * a.mtiles);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x16 = lane & 15, g = lane >> 4;
    const int64_t i0 = mtile * 32 + (wave & 1) * 16;
    const int64_t j0 = nt * 64 + (wave >> 1) * 32;
    const int64_t pa = a.mpad * a.kpad, px = a.npad * a.kpad;
    const int8_t* A = a.a + (l * ZP_DIGITS * a.mpad + i0 + x16) * a.kpad + g * 16;
    const int8_t* X = a.x + (l * ZP_DIGITS * a.npad + j0 + x16) * a.kpad + g * 16;
    const CsdMod md = mt.m[l];
    int64_t total[2][4];
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int r = 0; r < 4; r++) total[u][r] = 0;
    for (int64_t kc = 0; kc < a.kpad; kc += a.chunk) {
        const int64_t ke = min(a.kpad, kc + a.chunk);
        v4i acc[2][2 * ZP_DIGITS - 1];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) acc[u][s] = (v4i){0, 0, 0, 0};
        // software pipeline: the operands of a K-step are loaded one step ahead into the other
        // register set, so their memory latency overlaps the 128 MFMAs of the step before
        v4i ca[ZP_DIGITS], cx[2][ZP_DIGITS], na[ZP_DIGITS], nx[2][ZP_DIGITS];
        zp_load_step(ca, cx, A, X, pa, px, a.kpad, kc);
        int64_t k0 = kc;
        for (; k0 + 128 <= ke; k0 += 128) {
            zp_load_step(na, nx, A, X, pa, px, a.kpad, k0 + 64);
            zp_mfma_step(acc, ca, cx);
            if (k0 + 128 < ke) zp_load_step(ca, cx, A, X, pa, px, a.kpad, k0 + 128);
            zp_mfma_step(acc, na, nx);
        }
        if (k0 < ke) zp_mfma_step(acc, ca, cx);  // odd step count: the last step sits in ca
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                u128 v = 0;
#pragma unroll
                for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) v += (u128)(i128)acc[u][s][r] << (8 * s);
                int64_t t = total[u][r] + zp_reduce(md, (i128)v);
                total[u][r] = t >= md.N ? t - md.N : t;
            }
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int64_t v = j0 + u * 16 + x16;
        if (v >= a.Ly.nvec) continue;
        int64_t* out = a.y + l * a.Ly.s_limb + (v / a.Ly.nlo) * a.Ly.s_hi + (v % a.Ly.nlo) * a.Ly.s_lo;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t i = i0 + 4 * g + r;
            if (i < a.m) out[i * a.Ly.s_elem] = total[u][r];
        }
    }
}

// The LDS-staged form: 8 waves share one 32-row x 64-vector tile. Each K-step's digit planes of
// the tile (16 KiB of A, 32 KiB of X) are loaded from global memory once per workgroup, written
// to one of two LDS buffers while the waves run the MFMAs of the other, and every wave keeps one
// 16 x 16 output block (15 accumulators), so two waves per SIMD hide each other's LDS latency.
// The operand layout in LDS is [plane][row][64 k] with the same k-group assignment as above.

static constexpr int ZPL_THREADS = 512;
static constexpr int ZPL_A_BYTES = ZP_DIGITS * 32 * 64;  // 16 KiB
static constexpr int ZPL_X_BYTES = ZP_DIGITS * 64 * 64;  // 32 KiB

__global__ __launch_bounds__(512) void k_zp_mfma_lds(MfmaArgs a, ModTable mt) {
    extern __shared__ __align__(16) int8_t zs[];  // [2][A tile | X tile]
    int64_t b = blockIdx.x;
    if (a.nblocks % 8 == 0) b = (b % 8) * (a.nblocks / 8) + b / 8;
    const int64_t nt = b % a.ntiles;
    const int64_t mtile = (b / a.ntiles) % a.mtiles;
    const int64_t l = b / (a.ntiles * a.mtiles);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, g = lane >> 4;
    const int rb = wave & 1, cb = wave >> 1;  // this wave's 16-row block and 16-vector block
    const int64_t pa = a.mpad * a.kpad, px = a.npad * a.kpad;
    const int8_t* Ag = a.a + (l * ZP_DIGITS * a.mpad + mtile * 32) * a.kpad;
    const int8_t* Xg = a.x + (l * ZP_DIGITS * a.npad + nt * 64) * a.kpad;
    // staging: A has 1024 16-byte chunks (2 per thread), X 2048 (4 per thread); chunk c of a tile
    // is plane c / (rows * 4), row (c / 4) % rows, k-group c % 4
    auto stage_load = [&](v4i (&ra)[2], v4i (&rx)[4], int64_t k0) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int c = tid + q * ZPL_THREADS;
            const int pl = c >> 7, row = (c >> 2) & 31, kg = c & 3;
            ra[q] = *(const v4i*)(Ag + pl * pa + row * a.kpad + k0 + kg * 16);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = tid + q * ZPL_THREADS;
            const int pl = c >> 8, row = (c >> 2) & 63, kg = c & 3;
            rx[q] = *(const v4i*)(Xg + pl * px + row * a.kpad + k0 + kg * 16);
        }
    };
    auto stage_store = [&](const v4i (&ra)[2], const v4i (&rx)[4], int buf) {
        int8_t* base = zs + buf * (ZPL_A_BYTES + ZPL_X_BYTES);
#pragma unroll
        for (int q = 0; q < 2; q++) *(v4i*)(base + (tid + q * ZPL_THREADS) * 16) = ra[q];
#pragma unroll
        for (int q = 0; q < 4; q++) *(v4i*)(base + ZPL_A_BYTES + (tid + q * ZPL_THREADS) * 16) = rx[q];
    };
    const CsdMod md = mt.m[l];
    int64_t total[4] = {0, 0, 0, 0};
    for (int64_t kc = 0; kc < a.kpad; kc += a.chunk) {
        const int64_t ke = min(a.kpad, kc + a.chunk);
        v4i acc[2 * ZP_DIGITS - 1];
#pragma unroll
        for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) acc[s] = (v4i){0, 0, 0, 0};
        v4i ra[2], rx[4];
        stage_load(ra, rx, kc);
        stage_store(ra, rx, 0);
        __syncthreads();
        int buf = 0;
        for (int64_t k0 = kc; k0 < ke; k0 += 64) {
            const bool more = k0 + 64 < ke;
            if (more) stage_load(ra, rx, k0 + 64);
            const int8_t* base = zs + buf * (ZPL_A_BYTES + ZPL_X_BYTES);
            v4i av[ZP_DIGITS], xv[ZP_DIGITS];
#pragma unroll
            for (int d = 0; d < ZP_DIGITS; d++) {
                av[d] = *(const v4i*)(base + ((d * 32 + rb * 16 + x16) * 4 + g) * 16);
                xv[d] = *(const v4i*)(base + ZPL_A_BYTES + ((d * 64 + cb * 16 + x16) * 4 + g) * 16);
            }
#pragma unroll
            for (int da = 0; da < ZP_DIGITS; da++)
#pragma unroll
                for (int dx = 0; dx < ZP_DIGITS; dx++)
                    acc[da + dx] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[da], xv[dx], acc[da + dx], 0, 0, 0);
            if (more) stage_store(ra, rx, buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            u128 v = 0;
#pragma unroll
            for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) v += (u128)(i128)acc[s][r] << (8 * s);
            int64_t t = total[r] + zp_reduce(md, (i128)v);
            total[r] = t >= md.N ? t - md.N : t;
        }
    }
    const int64_t v = nt * 64 + cb * 16 + x16;
    if (v < a.Ly.nvec) {
        int64_t* out = a.y + l * a.Ly.s_limb + (v / a.Ly.nlo) * a.Ly.s_hi + (v % a.Ly.nlo) * a.Ly.s_lo;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t i = mtile * 32 + rb * 16 + 4 * g + r;
            if (i < a.m) out[i * a.Ly.s_elem] = total[r];
        }
    }
}

// The 64 x 64 form on 8 waves: a 64-row x 64-vector tile, each wave one 32-row x 16-vector strip
// (two 16 x 16 output blocks, 30 accumulators). A K-step stages 32 KiB of A and 32 KiB of X
// (double-buffered, 128 KiB): 1.5 x the MFMAs per staged byte of the 32 x 64 form, with two waves
// per SIMD still hiding each other's LDS reads, and half the accumulators per wave of the
// 4-wave 64 x 64 form.

static constexpr int ZP8_THREADS = 512;
static constexpr int ZP8_A_BYTES = ZP_DIGITS * 64 * 64;  // 32 KiB
static constexpr int ZP8_X_BYTES = ZP_DIGITS * 64 * 64;  // 32 KiB

__global__ __launch_bounds__(512) void k_zp_mfma_lds64(MfmaArgs a, ModTable mt) {
    extern __shared__ __align__(16) int8_t zs[];  // [2][A tile | X tile]
    int64_t b = blockIdx.x;
    if (a.nblocks % 8 == 0) b = (b % 8) * (a.nblocks / 8) + b / 8;
    const int64_t nt = b % a.ntiles;
    const int64_t mtile = (b / a.ntiles) % a.mtiles;
    const int64_t l = b / (a.ntiles * a.mtiles);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, g = lane >> 4;
    const int rh = wave & 1, cb = wave >> 1;  // this wave's 32-row half and 16-vector block
    const int64_t pa = a.mpad * a.kpad, px = a.npad * a.kpad;
    const int8_t* Ag = a.a + (l * ZP_DIGITS * a.mpad + mtile * 64) * a.kpad;
    const int8_t* Xg = a.x + (l * ZP_DIGITS * a.npad + nt * 64) * a.kpad;
    // staging: each tile has 2048 16-byte chunks (4 per thread); chunk c is plane c / 256,
    // row (c / 4) % 64, k-group c % 4
    auto stage_load = [&](v4i (&ra)[4], v4i (&rx)[4], int64_t k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = tid + q * ZP8_THREADS;
            const int pl = c >> 8, row = (c >> 2) & 63, kg = c & 3;
            ra[q] = *(const v4i*)(Ag + pl * pa + row * a.kpad + k0 + kg * 16);
            rx[q] = *(const v4i*)(Xg + pl * px + row * a.kpad + k0 + kg * 16);
        }
    };
    auto stage_store = [&](const v4i (&ra)[4], const v4i (&rx)[4], int buf) {
        int8_t* base = zs + buf * (ZP8_A_BYTES + ZP8_X_BYTES);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            *(v4i*)(base + (tid + q * ZP8_THREADS) * 16) = ra[q];
            *(v4i*)(base + ZP8_A_BYTES + (tid + q * ZP8_THREADS) * 16) = rx[q];
        }
    };
    const CsdMod md = mt.m[l];
    int64_t total[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int64_t kc = 0; kc < a.kpad; kc += a.chunk) {
        const int64_t ke = min(a.kpad, kc + a.chunk);
        v4i acc[2][2 * ZP_DIGITS - 1];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) acc[u][s] = (v4i){0, 0, 0, 0};
        v4i ra[4], rx[4];
        stage_load(ra, rx, kc);
        stage_store(ra, rx, 0);
        __syncthreads();
        int buf = 0;
        for (int64_t k0 = kc; k0 < ke; k0 += 64) {
            const bool more = k0 + 64 < ke;
            if (more) stage_load(ra, rx, k0 + 64);
            const int8_t* base = zs + buf * (ZP8_A_BYTES + ZP8_X_BYTES);
            v4i xv[ZP_DIGITS];
#pragma unroll
            for (int d = 0; d < ZP_DIGITS; d++)
                xv[d] = *(const v4i*)(base + ZP8_A_BYTES + ((d * 64 + cb * 16 + x16) * 4 + g) * 16);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                v4i av[ZP_DIGITS];
#pragma unroll
                for (int d = 0; d < ZP_DIGITS; d++)
                    av[d] = *(const v4i*)(base + ((d * 64 + rh * 32 + u * 16 + x16) * 4 + g) * 16);
#pragma unroll
                for (int da = 0; da < ZP_DIGITS; da++)
#pragma unroll
                    for (int dx = 0; dx < ZP_DIGITS; dx++)
                        acc[u][da + dx] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[da], xv[dx], acc[u][da + dx], 0, 0, 0);
            }
            if (more) stage_store(ra, rx, buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                u128 v = 0;
#pragma unroll
                for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) v += (u128)(i128)acc[u][s][r] << (8 * s);
                int64_t t = total[u][r] + zp_reduce(md, (i128)v);
                total[u][r] = t >= md.N ? t - md.N : t;
            }
    }
    const int64_t v = nt * 64 + cb * 16 + x16;
    if (v < a.Ly.nvec) {
        int64_t* out = a.y + l * a.Ly.s_limb + (v / a.Ly.nlo) * a.Ly.s_hi + (v % a.Ly.nlo) * a.Ly.s_lo;
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t i = mtile * 64 + rh * 32 + u * 16 + 4 * g + r;
                if (i < a.m) out[i * a.Ly.s_elem] = total[u][r];
            }
    }
}

// The group-circulant form of k_zp_mfma_lds64. For C[a][b] = g[a - b] (difference in
// C_o0 x ... x C_o(f-1), mixed-radix indices, last order 64), the 64 x 64 tile (row block A,
// column block B) is the circulant of the 64 generator entries g[S 64 .. S 64 + 63], S = A - B in
// the group of the leading factors: C[64 A + i][64 B + j] = g[S 64 + (i - j) mod 64]. A K-step
// stages the tile's X planes as k_zp_mfma_lds64 does, but for A only those 64 entries' 8 digit
// planes, reversed and written twice (h[t] = g[S 64 + (63 - t) mod 64], t < 128, 1 KiB). Lane
// (x16, k-group q) of row i then needs the bytes h[start .. start + 15], start = (63 - i + 16 q)
// mod 64: five aligned dword reads and four byte-aligns. The operand bytes are those of the dense
// digit planes, so the results are k_zp_mfma_lds64's bits.

struct CircMfmaArgs {
    MfmaArgs m;
    const int8_t* gen;  // [r][8][n] digit planes of the generator (column 0 of the matrix)
    int ntop;           // leading factors (all but the last, which is 64)
    int32_t top_orders[7], top_strides[7];  // strides in units of 64-blocks
};

static constexpr int ZPC_H_BYTES = ZP_DIGITS * 128;  // one K-step's reversed, doubled generator planes
#ifndef ZPC_NCB
#define ZPC_NCB 4  // 16-vector blocks per workgroup of the circulant form (2 measured the same)
#endif

__device__ __forceinline__ int32_t zp_block_sub(const CircMfmaArgs& c, int32_t A, int32_t B) {
    int32_t s = 0;
    for (int f = 0; f < c.ntop; f++) {
        const int32_t o = c.top_orders[f], st = c.top_strides[f];
        const int32_t da = (A / st) % o, db = (B / st) % o;
        s += ((da - db + o) % o) * st;
    }
    return s;
}

// NCB 16-vector blocks per workgroup (2 NCB waves, a 64-row x 16 NCB-vector tile)
template <int NCB>
__global__ __launch_bounds__(128 * NCB, 2) void k_zp_mfma_circ(CircMfmaArgs c, ModTable mt) {
    constexpr int THREADS = 128 * NCB, TW = 16 * NCB;  // threads, tile width in vectors
    constexpr int XB = ZP_DIGITS * TW * 64;             // X tile bytes per K-step
    extern __shared__ __align__(16) int8_t zs[];  // [2][X tile | H planes]
    const MfmaArgs& a = c.m;
    int64_t b = blockIdx.x;
    if (a.nblocks % 8 == 0) b = (b % 8) * (a.nblocks / 8) + b / 8;
    const int64_t nt = b % a.ntiles;
    const int64_t mtile = (b / a.ntiles) % a.mtiles;
    const int64_t l = b / (a.ntiles * a.mtiles);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, g = lane >> 4;
    const int rh = wave & 1, cb = wave >> 1;
    const int64_t px = a.npad * a.kpad;
    const int8_t* Xg = a.x + (l * ZP_DIGITS * a.npad + nt * TW) * a.kpad;
    const int8_t* Gg = c.gen + l * ZP_DIGITS * a.kpad;
    constexpr int BUF = XB + ZPC_H_BYTES;
    // staging: the X tile's 16-byte chunks (4 per thread; chunk cc is plane cc / (4 TW), vector
    // (cc / 4) % TW, k-group cc % 4) and, from threads 0-127, one dword of h each (plane tid / 16,
    // bytes 4 (tid % 16) .. + 3, both copies)
    auto stage_load = [&](v4i (&rx)[4], uint32_t& hw, int64_t k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int cc = tid + q * THREADS;
            const int pl = cc / (4 * TW), row = (cc >> 2) % TW, kg = cc & 3;
            rx[q] = *(const v4i*)(Xg + pl * px + row * a.kpad + k0 + kg * 16);
        }
        if (tid < 128) {
            const int d = tid >> 4, q = tid & 15;
            const int32_t S = zp_block_sub(c, (int32_t)mtile, (int32_t)(k0 >> 6));
            const uint32_t w = *(const uint32_t*)(Gg + d * a.kpad + (int64_t)S * 64 + 60 - 4 * q);
            hw = __builtin_amdgcn_perm(w, w, 0x00010203u);  // bytes reversed: g[63 - 4q] first
        }
    };
    auto stage_store = [&](const v4i (&rx)[4], uint32_t hw, int buf) {
        int8_t* base = zs + buf * BUF;
#pragma unroll
        for (int q = 0; q < 4; q++) *(v4i*)(base + (tid + q * THREADS) * 16) = rx[q];
        if (tid < 128) {
            const int d = tid >> 4, q = tid & 15;
            uint32_t* h = (uint32_t*)(base + XB + d * 128);
            h[q] = hw;
            h[16 + q] = hw;
        }
    };
    const CsdMod md = mt.m[l];
    int64_t total[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int64_t kc = 0; kc < a.kpad; kc += a.chunk) {
        const int64_t ke = min(a.kpad, kc + a.chunk);
        v4i acc[2][2 * ZP_DIGITS - 1];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int s2 = 0; s2 < 2 * ZP_DIGITS - 1; s2++) acc[u][s2] = (v4i){0, 0, 0, 0};
        v4i rx[4];
        uint32_t hw = 0;
        stage_load(rx, hw, kc);
        stage_store(rx, hw, 0);
        __syncthreads();
        int buf = 0;
        for (int64_t k0 = kc; k0 < ke; k0 += 64) {
            const bool more = k0 + 64 < ke;
            if (more) stage_load(rx, hw, k0 + 64);
            const int8_t* base = zs + buf * BUF;
            v4i xv[ZP_DIGITS];
#pragma unroll
            for (int d = 0; d < ZP_DIGITS; d++)
                xv[d] = *(const v4i*)(base + ((d * TW + cb * 16 + x16) * 4 + g) * 16);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int row = rh * 32 + u * 16 + x16;
                const int start = (63 - row + 16 * g) & 63;
                const int w0 = start >> 2;
                const uint32_t sh = (uint32_t)(start & 3);
                v4i av[ZP_DIGITS];
#pragma unroll
                for (int d = 0; d < ZP_DIGITS; d++) {
                    const uint32_t* h = (const uint32_t*)(base + XB + d * 128) + w0;
                    uint32_t hv[5];
#pragma unroll
                    for (int e = 0; e < 5; e++) hv[e] = h[e];
                    v4i o;
#pragma unroll
                    for (int e = 0; e < 4; e++) o[e] = (int)__builtin_amdgcn_alignbyte(hv[e + 1], hv[e], sh);
                    av[d] = o;
                }
#pragma unroll
                for (int da = 0; da < ZP_DIGITS; da++)
#pragma unroll
                    for (int dx = 0; dx < ZP_DIGITS; dx++)
                        acc[u][da + dx] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[da], xv[dx], acc[u][da + dx], 0, 0, 0);
            }
            if (more) stage_store(rx, hw, buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                u128 v = 0;
#pragma unroll
                for (int s2 = 0; s2 < 2 * ZP_DIGITS - 1; s2++) v += (u128)(i128)acc[u][s2][r] << (8 * s2);
                int64_t t = total[u][r] + zp_reduce(md, (i128)v);
                total[u][r] = t >= md.N ? t - md.N : t;
            }
    }
    const int64_t v = nt * TW + cb * 16 + x16;
    if (v < a.Ly.nvec) {
        int64_t* out = a.y + l * a.Ly.s_limb + (v / a.Ly.nlo) * a.Ly.s_hi + (v % a.Ly.nlo) * a.Ly.s_lo;
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t i = mtile * 64 + rh * 32 + u * 16 + 4 * g + r;
                if (i < a.m) out[i * a.Ly.s_elem] = total[u][r];
            }
    }
}

// The 64 x 64 form: 4 waves share one 64-row x 64-vector tile, and each wave keeps a 2 x 2
// arrangement of 16 x 16 output blocks (60 accumulators, held in AGPRs). A K-step's operands of a
// wave are 2 x 8 planes of A rows and 2 x 8 planes of X vectors, 32 LDS reads for 256 MFMAs:
// half the LDS bytes per MFMA of the 32 x 64 form. Staging is as above: 32 KiB of A and 32 KiB of
// X per K-step, double-buffered (128 KiB), so one workgroup (one wave per SIMD) runs per CU.

static constexpr int ZPW_THREADS = 256;
static constexpr int ZPW_A_BYTES = ZP_DIGITS * 64 * 64;  // 32 KiB
static constexpr int ZPW_X_BYTES = ZP_DIGITS * 64 * 64;  // 32 KiB

__global__ __launch_bounds__(256) void k_zp_mfma_wide(MfmaArgs a, ModTable mt) {
    extern __shared__ __align__(16) int8_t zs[];  // [2][A tile | X tile]
    int64_t b = blockIdx.x;
    if (a.nblocks % 8 == 0) b = (b % 8) * (a.nblocks / 8) + b / 8;
    const int64_t nt = b % a.ntiles;
    const int64_t mtile = (b / a.ntiles) % a.mtiles;
    const int64_t l = b / (a.ntiles * a.mtiles);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, g = lane >> 4;
    const int rb = wave & 1, cb = wave >> 1;  // this wave's 32-row half and 32-vector half
    const int64_t pa = a.mpad * a.kpad, px = a.npad * a.kpad;
    const int8_t* Ag = a.a + (l * ZP_DIGITS * a.mpad + mtile * 64) * a.kpad;
    const int8_t* Xg = a.x + (l * ZP_DIGITS * a.npad + nt * 64) * a.kpad;
    // staging: A and X have 2048 16-byte chunks each (8 per thread); chunk c of a tile is plane
    // c / 256, row (c / 4) % 64, k-group c % 4, and lands at byte 16 c of the tile. The chunks go
    // from global memory straight into LDS (global_load_lds_dwordx4: lane i of a wave writes its
    // 16 bytes at the wave's base + 16 i), so no registers hold them in between.
    auto stage_lds = [&](int64_t k0, int buf) {
        const int off = buf * (ZPW_A_BYTES + ZPW_X_BYTES);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int c = tid + q * ZPW_THREADS;
            const int pl = c >> 8, row = (c >> 2) & 63, kg = c & 3;
            const int wbase = (q * ZPW_THREADS + wave * 64) * 16;
            __builtin_amdgcn_global_load_lds((const void*)(Ag + pl * pa + row * a.kpad + k0 + kg * 16), zs + off + wbase, 16,
                                             0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(Xg + pl * px + row * a.kpad + k0 + kg * 16),
                                             zs + off + ZPW_A_BYTES + wbase, 16, 0, 0);
        }
    };
    const CsdMod md = mt.m[l];
    int64_t total[2][2][4];
#pragma unroll
    for (int ua = 0; ua < 2; ua++)
#pragma unroll
        for (int ux = 0; ux < 2; ux++)
#pragma unroll
            for (int r = 0; r < 4; r++) total[ua][ux][r] = 0;
    for (int64_t kc = 0; kc < a.kpad; kc += a.chunk) {
        const int64_t ke = min(a.kpad, kc + a.chunk);
        v4i acc[2][2][2 * ZP_DIGITS - 1];
#pragma unroll
        for (int ua = 0; ua < 2; ua++)
#pragma unroll
            for (int ux = 0; ux < 2; ux++)
#pragma unroll
                for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) acc[ua][ux][s] = (v4i){0, 0, 0, 0};
        stage_lds(kc, 0);
        __builtin_amdgcn_s_waitcnt(0);  // the direct loads have landed in LDS
        __syncthreads();
        int buf = 0;
        for (int64_t k0 = kc; k0 < ke; k0 += 64) {
            const bool more = k0 + 64 < ke;
            // the next K-step into the other buffer, whose last reads ended before the barrier
            if (more) stage_lds(k0 + 64, buf ^ 1);
            const int8_t* base = zs + buf * (ZPW_A_BYTES + ZPW_X_BYTES);
            // both vector halves' operands once per K-step, then each row half against them
            v4i xv[2][ZP_DIGITS];
#pragma unroll
            for (int ux = 0; ux < 2; ux++)
#pragma unroll
                for (int d = 0; d < ZP_DIGITS; d++)
                    xv[ux][d] = *(const v4i*)(base + ZPW_A_BYTES + ((d * 64 + cb * 32 + ux * 16 + x16) * 4 + g) * 16);
#pragma unroll
            for (int ua = 0; ua < 2; ua++) {
                v4i av[ZP_DIGITS];
#pragma unroll
                for (int d = 0; d < ZP_DIGITS; d++)
                    av[d] = *(const v4i*)(base + ((d * 64 + rb * 32 + ua * 16 + x16) * 4 + g) * 16);
#pragma unroll
                for (int ux = 0; ux < 2; ux++)
#pragma unroll
                    for (int da = 0; da < ZP_DIGITS; da++)
#pragma unroll
                        for (int dx = 0; dx < ZP_DIGITS; dx++)
                            acc[ua][ux][da + dx] =
                                __builtin_amdgcn_mfma_i32_16x16x64_i8(av[da], xv[ux][dx], acc[ua][ux][da + dx], 0, 0, 0);
            }
            // each wave's LDS reads of this step and direct loads of the next have completed
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            buf ^= 1;
        }
#pragma unroll
        for (int ua = 0; ua < 2; ua++)
#pragma unroll
            for (int ux = 0; ux < 2; ux++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    u128 v = 0;
#pragma unroll
                    for (int s = 0; s < 2 * ZP_DIGITS - 1; s++) v += (u128)(i128)acc[ua][ux][s][r] << (8 * s);
                    int64_t t = total[ua][ux][r] + zp_reduce(md, (i128)v);
                    total[ua][ux][r] = t >= md.N ? t - md.N : t;
                }
    }
#pragma unroll
    for (int ux = 0; ux < 2; ux++) {
        const int64_t v = nt * 64 + cb * 32 + ux * 16 + x16;
        if (v >= a.Ly.nvec) continue;
        int64_t* out = a.y + l * a.Ly.s_limb + (v / a.Ly.nlo) * a.Ly.s_hi + (v % a.Ly.nlo) * a.Ly.s_lo;
#pragma unroll
        for (int ua = 0; ua < 2; ua++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t i = mtile * 64 + rb * 32 + ua * 16 + 4 * g + r;
                if (i < a.m) out[i * a.Ly.s_elem] = total[ua][ux][r];
            }
    }
}

// ctx->zp_kernel: the MFMA kernel form 0 uses: 0 L2 operands, 1 LDS 32 x 64, 2 LDS 64 x 64 on 4
// waves, 3 (default) LDS 64 x 64 on 8 waves
extern "C" int csd_zp_set_kernel(csd_ctx* ctx, int form) {
    CSD_TRY if (!ctx) throw CsdError("null context");
    if (form < 0 || form > 4) throw CsdError("csd_zp_set_kernel: form must be 0, 1, 2, 3 or 4");
    ctx->zp_kernel = form;
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// the plain form: 16 x 16 outputs per workgroup, 64-term slices of A and X staged in LDS

struct PlainArgs {
    const int64_t* a;  // [r][m][k]
    const int64_t* x;
    int64_t* y;
    VecLayout Lx, Ly;
    int64_t m, k, mtiles, ntiles;
};

__global__ __launch_bounds__(256) void k_zp_plain(PlainArgs a, ModTable mt) {
    __shared__ int64_t ta[16][65], tx[16][65];
    const int64_t nt = blockIdx.x % a.ntiles;
    const int64_t mtile = (blockIdx.x / a.ntiles) % a.mtiles;
    const int64_t l = blockIdx.x / (a.ntiles * a.mtiles);
    const int ti = threadIdx.x & 15, tv = threadIdx.x >> 4;
    const int64_t i = mtile * 16 + ti, v = nt * 16 + tv;
    const CsdMod md = mt.m[l];
    const int64_t p = md.N;
    const int64_t* A = a.a + l * a.m * a.k;
    const int64_t* Xl = a.x + l * a.Lx.s_limb;
    int64_t total = 0;
    for (int64_t kc = 0; kc < a.k; kc += ZP_PLAIN_CHUNK) {
        const int64_t ke = min(a.k, kc + ZP_PLAIN_CHUNK);
        u128 acc = 0;
        for (int64_t k0 = kc; k0 < ke; k0 += 64) {
            for (int e = threadIdx.x; e < 16 * 64; e += 256) {
                const int rr = e >> 6, kk = e & 63;
                const int64_t k = k0 + kk;
                const int64_t ia = mtile * 16 + rr, iv = nt * 16 + rr;
                ta[rr][kk] = (ia < a.m && k < ke) ? zp_centre(A[ia * a.k + k], p) : 0;
                tx[rr][kk] = (iv < a.Lx.nvec && k < ke)
                                 ? zp_centre(Xl[(iv / a.Lx.nlo) * a.Lx.s_hi + (iv % a.Lx.nlo) * a.Lx.s_lo + k * a.Lx.s_elem], p)
                                 : 0;
            }
            __syncthreads();
#pragma unroll 8
            for (int kk = 0; kk < 64; kk++) {
                uint64_t lo;
                int64_t hi;
                csd_smul_wide(ta[ti][kk], tx[tv][kk], lo, hi);
                acc += ((u128)(uint64_t)hi << 64) | lo;
            }
            __syncthreads();
        }
        int64_t t = total + zp_reduce(md, (i128)acc);
        total = t >= p ? t - p : t;
    }
    if (i < a.m && v < a.Ly.nvec)
        a.y[l * a.Ly.s_limb + (v / a.Ly.nlo) * a.Ly.s_hi + (v % a.Ly.nlo) * a.Ly.s_lo + i * a.Ly.s_elem] = total;
}

// Maps of at most 32 x 32 (the L' part's 12 x 12, homcrt's small J_p blocks): one thread per
// vector, its K centred entries in registers, the limb's matrix staged in LDS. Centred entries are
// below 2^56 in magnitude, so with a = a1 2^28 + a0 and x = x1 2^28 + x0 (a0, x0 in [0, 2^28)) the
// four partial sums of 32 x 32-bit products over K <= 32 terms stay below 2^62; the exact sum is
// assembled once per output and reduced as k_zp_plain reduces it (one chunk), so the results are
// its bits.
#define ZP_SMALL_MAX 32

template <int K>
__global__ __launch_bounds__(256) void k_zp_small(PlainArgs a, ModTable mt) {
    __shared__ int32_t sa[ZP_SMALL_MAX * ZP_SMALL_MAX][2];  // (a0, a1) of entry [i][k]
    const int64_t chunks = (a.Lx.nvec + 255) / 256;
    const int64_t l = blockIdx.x / chunks;
    const int64_t v = (blockIdx.x % chunks) * 256 + threadIdx.x;
    const CsdMod md = mt.m[l];
    const int64_t p = md.N;
    const int m = (int)a.m;
    for (int e = threadIdx.x; e < m * K; e += 256) {
        const int64_t c = zp_centre(a.a[l * a.m * K + e], p);
        sa[e][0] = (int32_t)(c & ((1 << 28) - 1));
        sa[e][1] = (int32_t)(c >> 28);
    }
    __syncthreads();
    if (v >= a.Lx.nvec) return;
    const int64_t* __restrict__ xv = a.x + l * a.Lx.s_limb + (v / a.Lx.nlo) * a.Lx.s_hi + (v % a.Lx.nlo) * a.Lx.s_lo;
    int64_t* __restrict__ yv = a.y + l * a.Ly.s_limb + (v / a.Ly.nlo) * a.Ly.s_hi + (v % a.Ly.nlo) * a.Ly.s_lo;
    int32_t x0[K], x1[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int64_t c = zp_centre(xv[k * a.Lx.s_elem], p);
        x0[k] = (int32_t)(c & ((1 << 28) - 1));
        x1[k] = (int32_t)(c >> 28);
    }
    for (int i = 0; i < m; i++) {
        int64_t s00 = 0, s01 = 0, s10 = 0, s11 = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int32_t a0 = sa[i * K + k][0], a1 = sa[i * K + k][1];
            s00 += (int64_t)a0 * x0[k];
            s01 += (int64_t)a0 * x1[k];
            s10 += (int64_t)a1 * x0[k];
            s11 += (int64_t)a1 * x1[k];
        }
        const i128 sum = (i128)s00 + ((i128)(s01 + s10) << 28) + ((i128)s11 << 56);
        yv[i * a.Ly.s_elem] = zp_reduce(md, sum);
    }
}

typedef void (*ZpSmallKernel)(PlainArgs, ModTable);
template <int K>
static ZpSmallKernel zp_small_for(int k) {
    if constexpr (K > ZP_SMALL_MAX) {
        return nullptr;
    } else {
        return k == K ? k_zp_small<K> : zp_small_for<K + 1>(k);
    }
}

// ---------------------------------------------------------------------------------------------
// group-circulant matrices: out[l][a][b] = kern[l][a - b], the difference taken in the group
// C_o0 x ... x C_o(f-1) with mixed-radix (row-major) indices

struct CircArgs {
    const int64_t* kern;  // [r][n] contiguous
    int64_t* out;         // [r][n][n] contiguous
    int64_t n, r;
    int nf;
    int64_t orders[8], strides[8];
};

__global__ __launch_bounds__(256) void k_circulant(CircArgs c) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.r * c.n * c.n) return;
    const int64_t bcol = t % c.n, arow = (t / c.n) % c.n, l = t / (c.n * c.n);
    int64_t idx = 0;
    for (int f = 0; f < c.nf; f++) {
        const int64_t da = (arow / c.strides[f]) % c.orders[f], db = (bcol / c.strides[f]) % c.orders[f];
        idx += ((da - db + c.orders[f]) % c.orders[f]) * c.strides[f];
    }
    c.out[t] = c.kern[l * c.n + idx];
}

void csd_circulant_launch(csd_ctx* ctx, const int64_t* kern, int64_t* out, int64_t r, const std::vector<size_t>& orders) {
    CircArgs c;
    memset(&c, 0, sizeof(c));
    if (orders.size() > 8) throw CsdError("circulant: too many cyclic factors");
    c.kern = kern;
    c.out = out;
    c.r = r;
    c.nf = (int)orders.size();
    int64_t n = 1;
    for (size_t f = orders.size(); f-- > 0;) {
        c.orders[f] = (int64_t)orders[f];
        c.strides[f] = n;
        n *= (int64_t)orders[f];
    }
    c.n = n;
    const int64_t total = r * n * n;
    if (total == 0) return;
    hipLaunchKernelGGL(k_circulant, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, c);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

extern "C" int csd_zp_matrix_create(csd_ctx* ctx, const csd_view* a, const csd_modulus* mods, size_t nmods, csd_zp_matrix** out) {
    CSD_TRY check_view(a, "zp matrix");
    if (a->shape[0] != 1) throw CsdError("zp matrix: dimension 0 must be 1");
    const int64_t r = a->shape[1], m = a->shape[2], k = a->shape[3];
    if (r < 1 || m < 1 || k < 1) throw CsdError("zp matrix: empty matrix");
    ModTable mt = make_mod_table(mods, nmods, r);
    for (size_t j = 0; j < nmods; j++)
        if (mods[j].N >= ((int64_t)1 << 57)) throw CsdError("zp matrix: moduli must be below 2^57");
    auto* z = new csd_zp_matrix();
    z->ctx = ctx;
    z->r = r;
    z->m = m;
    z->k = k;
    z->mpad = round_up(m, 64);  // whole 64-row tiles (the 32-row kernels use half tiles)
    z->kpad = round_up(k, 64);
    z->mods.assign(mods, mods + nmods);
    try {
        z->planes = (int8_t*)ctx->alloc_bytes((size_t)r * ZP_DIGITS * z->mpad * z->kpad);
        z->values = (int64_t*)ctx->alloc_bytes((size_t)r * m * k * 8);
        // a contiguous copy of the entries (the plain form reads it)
        csd_view dst;
        dst.ptr = z->values;
        for (int d = 0; d < 4; d++) dst.shape[d] = a->shape[d];
        dst.strides[3] = 1;
        dst.strides[2] = k;
        dst.strides[1] = m * k;
        dst.strides[0] = r * m * k;
        if (csd_hal_assign(ctx, &dst, a) != 0) throw CsdError(csd_last_error());
        VecLayout L = vec_layout(a, 3);
        launch_digits(ctx, (const int64_t*)a->ptr, L, r, z->mpad, z->kpad, z->planes, mt);
    } catch (...) {
        if (z->planes) ctx->free_bytes(z->planes);
        if (z->values) ctx->free_bytes(z->values);
        delete z;
        throw;
    }
    *out = z;
    CSD_CATCH
}

extern "C" int csd_zp_matrix_free(csd_zp_matrix* z) {
    CSD_TRY if (!z) return 0;
    HIPCHECK(hipStreamSynchronize(z->ctx->stream));
    z->ctx->free_bytes(z->planes);
    z->ctx->free_bytes(z->values);
    if (z->gen) z->ctx->free_bytes(z->gen);
    delete z;
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// group-circulant structure of a square matrix (csd_zp_matrix_set_circulant)

struct CircCheckArgs {
    const int64_t* values;  // [r][n][n]
    int64_t n, r;
    int nf;
    int64_t orders[8], strides[8];
    int64_t* bad;           // set to 1 (vector store) where an entry differs from g[a - b]
};

// every entry against column 0 at the group difference of its indices
__global__ __launch_bounds__(256) void k_zp_check_circulant(CircCheckArgs c) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.r * c.n * c.n) return;
    const int64_t bcol = t % c.n, arow = (t / c.n) % c.n, l = t / (c.n * c.n);
    int64_t idx = 0;
    for (int f = 0; f < c.nf; f++) {
        const int64_t da = (arow / c.strides[f]) % c.orders[f], db = (bcol / c.strides[f]) % c.orders[f];
        idx += ((da - db + c.orders[f]) % c.orders[f]) * c.strides[f];
    }
    if (c.values[t] != c.values[(l * c.n + idx) * c.n]) c.bad[0] = 1;
}

// the generator (column 0) as 8 balanced digit planes [r][8][n], k_zp_digits' digits
__global__ __launch_bounds__(256) void k_zp_gen_digits(const int64_t* values, int64_t n, int64_t r, int8_t* gen, ModTable mt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= r * n) return;
    const int64_t a = t % n, l = t / n;
    int64_t c = zp_centre(values[(l * n + a) * n], mt.m[l].N);
#pragma unroll
    for (int d = 0; d < ZP_DIGITS; d++) {
        int32_t lo = (int32_t)(c & 255);
        lo = lo >= 128 ? lo - 256 : lo;
        c = (c - lo) >> 8;
        gen[(l * ZP_DIGITS + d) * n + a] = (int8_t)lo;
    }
}

extern "C" int csd_zp_matrix_set_circulant(csd_zp_matrix* z, const size_t* orders, size_t norders) {
    CSD_TRY if (!z) throw CsdError("null zp matrix");
    if (norders < 1 || norders > 8) throw CsdError("zp circulant: 1 to 8 cyclic factors");
    int64_t n = 1;
    for (size_t f = 0; f < norders; f++) n *= (int64_t)orders[f];
    if (z->m != n || z->k != n) throw CsdError("zp circulant: the matrix must be n x n, n the group order");
    if (orders[norders - 1] != 64)
        throw CsdError("zp circulant: the last cyclic factor must have order 64 (the 64 x 64 tiles are its circulants)");
    csd_ctx* ctx = z->ctx;
    CircCheckArgs c;
    memset(&c, 0, sizeof(c));
    c.values = z->values;
    c.n = n;
    c.r = z->r;
    c.nf = (int)norders;
    int64_t st = 1;
    for (size_t f = norders; f-- > 0;) {
        c.orders[f] = (int64_t)orders[f];
        c.strides[f] = st;
        st *= (int64_t)orders[f];
    }
    int64_t* bad = (int64_t*)ctx->alloc_bytes(sizeof(int64_t));  // csd_zero_fill clears whole words
    int64_t h_bad = 0;
    try {
        csd_zero_fill(ctx, bad, sizeof(int64_t));
        c.bad = bad;
        const int64_t total = z->r * n * n;
        hipLaunchKernelGGL(k_zp_check_circulant, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, c);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(&h_bad, bad, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        ctx->free_bytes(bad);
        throw;
    }
    ctx->free_bytes(bad);
    if (h_bad) throw CsdError("zp circulant: the matrix is not the group-circulant of its first column");
    if (!z->gen) z->gen = (int8_t*)ctx->alloc_bytes((size_t)z->r * ZP_DIGITS * n);
    ModTable mt = make_mod_table(z->mods.data(), z->mods.size(), z->r);
    hipLaunchKernelGGL(k_zp_gen_digits, dim3((unsigned)((z->r * n + 255) / 256)), dim3(256), 0, ctx->stream, z->values, n, z->r,
                       z->gen, mt);
    HIPCHECK(hipGetLastError());
    z->ntop = (int)norders - 1;
    for (int f = 0; f < z->ntop; f++) {
        z->top_orders[f] = (int32_t)orders[f];
        z->top_strides[f] = (int32_t)(c.strides[f] / 64);
    }
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// The scratch of csd_zp_apply_dim_noalloc: the source vectors' digit planes live in this file-scope
// device array instead of a pooled allocation. Every GPU the library runs on holds its own copy
// (CSD_ZP_SCRATCH_BYTES) until the process ends. Host code never passes the array's name to a
// kernel (on the host it is only a stand-in): csd_zp_scratch_setup resolves the device address of
// each GPU's copy once, with that GPU current, and the entry point uses that address.
static __device__ __align__(256) int8_t scratch[CSD_ZP_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, void*> g_scratch_ptr;  // device -> address of that device's copy of `scratch`

void csd_zp_scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return;
    int cur = -1;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != device) throw CsdError("zp scratch setup: the device must be current");
    void* ptr = nullptr;
    HIPCHECK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw CsdError("zp scratch setup: no device address for the scratch array");
    g_scratch_ptr[device] = ptr;
}

static void* zp_scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) throw CsdError("zp scratch: not set up for this device");
    return it->second;
}

// The K-chunk of the int8 contraction: at most ZP_MAX_CHUNK terms, or max_chunk rounded down to
// a multiple of 64 when given.
static int64_t zp_chunk(size_t max_chunk) {
    int64_t chunk = ZP_MAX_CHUNK;
    if (max_chunk) chunk = std::min<int64_t>(chunk, (int64_t)max_chunk / 64 * 64);
    if (chunk < 64) throw CsdError("zp apply: chunk must hold at least 64 terms");
    return chunk;
}

// The int8-MFMA form of a dense map: splits the source vectors into their digit planes in xd
// ([r][8][npad][kpad] bytes, every byte written here, padding included) and runs the context's
// MFMA kernel over them.
static void zp_mfma_run(csd_ctx* ctx, const csd_zp_matrix* z, const csd_view* src, const VecLayout& Lx, const csd_view* dst,
                        const VecLayout& Ly, int64_t nvec, int64_t npad, int64_t chunk, int8_t* xd, const ModTable& mt) {
    launch_digits(ctx, (const int64_t*)src->ptr, Lx, z->r, npad, z->kpad, xd, mt);
    MfmaArgs a;
    a.a = z->planes;
    a.x = xd;
    a.y = (int64_t*)dst->ptr;
    a.Ly = Ly;
    a.m = z->m;
    a.mpad = z->mpad;
    a.npad = npad;
    a.kpad = z->kpad;
    a.chunk = chunk;
    a.mtiles = z->mpad / 32;
    a.ntiles = npad / 64;
    a.nblocks = z->r * a.mtiles * a.ntiles;
    if (a.nblocks > (int64_t)INT32_MAX) throw CsdError("zp apply: too many tiles");
    TimedScope ts(ctx, "zp_mfma",
                  (double)z->r * ZP_DIGITS * (z->mpad * z->kpad + npad * z->kpad) + 8.0 * z->r * nvec * z->m);
    const int zk = ctx->zp_kernel;
    if (zk == 3 && z->gen && z->mpad == z->kpad && z->m == z->k) {
        // the group-circulant form: the A operands from the generator's planes
        CircMfmaArgs c;
        memset(&c, 0, sizeof(c));
        c.m = a;
        c.m.mtiles = z->mpad / 64;
        c.m.nblocks = z->r * c.m.mtiles * c.m.ntiles;
        c.gen = z->gen;
        c.ntop = z->ntop;
        for (int f = 0; f < z->ntop; f++) {
            c.top_orders[f] = z->top_orders[f];
            c.top_strides[f] = z->top_strides[f];
        }
        constexpr int NCB = ZPC_NCB;
        c.m.ntiles = npad / (16 * NCB);
        c.m.nblocks = z->r * c.m.mtiles * c.m.ntiles;
        const size_t lds = 2 * (size_t)(ZP_DIGITS * 16 * NCB * 64 + ZPC_H_BYTES);
        csd_allow_lds((const void*)k_zp_mfma_circ<NCB>, (int)lds);
        hipLaunchKernelGGL(k_zp_mfma_circ<NCB>, dim3((unsigned)c.m.nblocks), dim3(128 * NCB), lds, ctx->stream, c, mt);
    } else if (zk == 3 || zk == 4) {
        a.mtiles = z->mpad / 64;  // mpad is a multiple of 64
        a.nblocks = z->r * a.mtiles * a.ntiles;
        const size_t lds = 2 * (size_t)(ZP8_A_BYTES + ZP8_X_BYTES);
        csd_allow_lds((const void*)k_zp_mfma_lds64, (int)lds);
        hipLaunchKernelGGL(k_zp_mfma_lds64, dim3((unsigned)a.nblocks), dim3(ZP8_THREADS), lds, ctx->stream, a, mt);
    } else if (zk == 2) {
        // 64-row tiles over the same planes (mpad is a multiple of 64)
        a.mtiles = z->mpad / 64;
        a.nblocks = z->r * a.mtiles * a.ntiles;
        const size_t lds = 2 * (size_t)(ZPW_A_BYTES + ZPW_X_BYTES);
        csd_allow_lds((const void*)k_zp_mfma_wide, (int)lds);
        hipLaunchKernelGGL(k_zp_mfma_wide, dim3((unsigned)a.nblocks), dim3(ZPW_THREADS), lds, ctx->stream, a, mt);
    } else if (zk == 1) {
        const size_t lds = 2 * (size_t)(ZPL_A_BYTES + ZPL_X_BYTES);
        csd_allow_lds((const void*)k_zp_mfma_lds, (int)lds);
        hipLaunchKernelGGL(k_zp_mfma_lds, dim3((unsigned)a.nblocks), dim3(ZPL_THREADS), lds, ctx->stream, a, mt);
    } else {
        hipLaunchKernelGGL(k_zp_mfma, dim3((unsigned)a.nblocks), dim3(256), 0, ctx->stream, a, mt);
    }
    HIPCHECK(hipGetLastError());
}

extern "C" int csd_zp_apply_dim(csd_ctx* ctx, const csd_zp_matrix* z, const csd_view* dst, const csd_view* src, int dim, int form,
                                size_t max_chunk) {
    CSD_TRY check_view(dst, "zp apply dst");
    check_view(src, "zp apply src");
    if (form != 0 && form != 1) throw CsdError("zp apply: form must be 0 (int8 MFMA) or 1 (plain)");
    const VecLayout Lx = vec_layout(src, dim), Ly = vec_layout(dst, dim);
    if (src->shape[1] != z->r || dst->shape[1] != z->r) throw CsdError("zp apply: dimension 1 must equal the number of moduli");
    if (Lx.nelem != z->k) throw CsdError("zp apply: source vectors must have the matrix's column count");
    if (Ly.nelem != z->m) throw CsdError("zp apply: destination vectors must have the matrix's row count");
    if (Lx.nvec != Ly.nvec || src->shape[0] != dst->shape[0] || Lx.nlo != Ly.nlo)
        throw CsdError("zp apply: source and destination must hold the same vectors");
    if (dst->ptr == src->ptr) throw CsdError("zp apply: destination must not alias the source");
    ModTable mt = make_mod_table(z->mods.data(), z->mods.size(), z->r);
    const int64_t nvec = Lx.nvec;
    if (nvec == 0) return 0;
    if (form == 1 && z->m <= ZP_SMALL_MAX && z->k <= ZP_SMALL_MAX) {
        PlainArgs a;
        memset(&a, 0, sizeof(a));
        a.a = z->values;
        a.x = (const int64_t*)src->ptr;
        a.y = (int64_t*)dst->ptr;
        a.Lx = Lx;
        a.Ly = Ly;
        a.m = z->m;
        a.k = z->k;
        const int64_t blocks = z->r * ((nvec + 255) / 256);
        if (blocks > (int64_t)INT32_MAX) throw CsdError("zp apply: too many tiles");
        TimedScope ts(ctx, "zp_small", 8.0 * z->r * (z->m * z->k + nvec * (z->k + z->m)));
        hipLaunchKernelGGL(zp_small_for<1>((int)z->k), dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a, mt);
        HIPCHECK(hipGetLastError());
        return 0;
    }
    if (form == 1) {
        PlainArgs a;
        a.a = z->values;
        a.x = (const int64_t*)src->ptr;
        a.y = (int64_t*)dst->ptr;
        a.Lx = Lx;
        a.Ly = Ly;
        a.m = z->m;
        a.k = z->k;
        a.mtiles = (z->m + 15) / 16;
        a.ntiles = (nvec + 15) / 16;
        const int64_t blocks = z->r * a.mtiles * a.ntiles;
        if (blocks > (int64_t)INT32_MAX) throw CsdError("zp apply: too many tiles");
        TimedScope ts(ctx, "zp_plain", 8.0 * z->r * (z->m * z->k + nvec * (z->k + z->m)));
        hipLaunchKernelGGL(k_zp_plain, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a, mt);
        HIPCHECK(hipGetLastError());
        return 0;
    }
    const int64_t chunk = zp_chunk(max_chunk);
    const int64_t npad = round_up(nvec, 64);
    int8_t* xd = (int8_t*)ctx->alloc_bytes((size_t)z->r * ZP_DIGITS * npad * z->kpad);
    try {
        zp_mfma_run(ctx, z, src, Lx, dst, Ly, nvec, npad, chunk, xd, mt);
    } catch (...) {
        ctx->free_bytes(xd);
        throw;
    }
    ctx->free_bytes(xd);
    CSD_CATCH
}

// csd_zp_apply_dim's int8-MFMA form (form 0) without any allocation: the digit planes of the
// source vectors go to the device's scratch array. Calls on one GPU share that array, so calls
// that overlap on one GPU (other streams or contexts of the same device) race; calls on one
// context's stream are ordered and safe. Nothing needs clearing: the digit kernel writes every
// byte the contraction reads, padding included, before the contraction runs.
extern "C" int csd_zp_apply_dim_noalloc(csd_ctx* ctx, const csd_zp_matrix* z, const csd_view* dst, const csd_view* src,
                                        int dim, size_t max_chunk) {
    CSD_TRY if (!ctx || !z) throw CsdError("zp apply: null context or matrix");
    check_view(dst, "zp apply dst");
    check_view(src, "zp apply src");
    const VecLayout Lx = vec_layout(src, dim), Ly = vec_layout(dst, dim);
    if (src->shape[1] != z->r || dst->shape[1] != z->r) throw CsdError("zp apply: dimension 1 must equal the number of moduli");
    if (Lx.nelem != z->k) throw CsdError("zp apply: source vectors must have the matrix's column count");
    if (Ly.nelem != z->m) throw CsdError("zp apply: destination vectors must have the matrix's row count");
    if (Lx.nvec != Ly.nvec || src->shape[0] != dst->shape[0] || Lx.nlo != Ly.nlo)
        throw CsdError("zp apply: source and destination must hold the same vectors");
    if (dst->ptr == src->ptr) throw CsdError("zp apply: destination must not alias the source");
    const int64_t nvec = Lx.nvec;
    if (nvec == 0) return 0;
    const int64_t chunk = zp_chunk(max_chunk);
    const int64_t npad = round_up(nvec, 64);
    const double need = (double)z->r * ZP_DIGITS * (double)npad * (double)z->kpad;
    if (need > (double)CSD_ZP_SCRATCH_BYTES)
        throw CsdError("zp apply (no allocation): the source digit planes need " + std::to_string((long long)need) +
                       " bytes (moduli x 8 x vectors rounded up to 64 x columns rounded up to 64), more than the " +
                       std::to_string((long long)CSD_ZP_SCRATCH_BYTES) + "-byte scratch");
    ModTable mt = make_mod_table(z->mods.data(), z->mods.size(), z->r);
    int8_t* xd = (int8_t*)zp_scratch_of(ctx->device);
    zp_mfma_run(ctx, z, src, Lx, dst, Ly, nvec, npad, chunk, xd, mt);
    CSD_CATCH
}
