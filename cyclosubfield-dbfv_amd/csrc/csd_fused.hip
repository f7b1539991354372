// Fused and batched forms of the ring-level hot path (beyond one-kernel-per-HAL-call).
//
//  * Galois key switching for a batch of ciphertexts and a set of rotations
//    (bfv_hom_gal, cyclosubfield-dbfv/src/bfv.rs:783-841): one kernel per rotation computes,
//    for every ciphertext of the batch, sigma_g of the hoisted gadget digits (a gather through
//    the automorphism's permutation, ring.rs:1183-1222), the inner product with the rotation's
//    key digits (hal/cpu.rs:706-788), the added sigma_g(c0) * mont(1) (ring.rs:496-500) and the
//    wide reduction (ring.rs:846-865). Keys stay resident in HBM and each key digit is read once
//    per group of ciphertexts instead of once per ciphertext. The i128 sums wrap exactly like the
//    reference's, so outputs are bit-identical to the unfused HAL sequence.
//  * Uniform random residues for synthetic inputs and keys (splitmix64 of the element index).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/csdbfv_ring.h"
#include "csd_internal.hpp"
#include "csd_mont_fast.hpp"

// ---------------------------------------------------------------------------------------------
// synthetic residues

__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct FillArgs {
    int64_t d0, d1, d2, d3;
    int64_t st[4];
    int64_t* data;
    uint64_t seed;
};

// data[i, j, k, l] = uniform in [0, p_j): the high word of a 64-bit random times p_j
__global__ void k_fill_uniform(FillArgs a, ModTable mt) {
    int64_t total = a.d0 * a.d1 * a.d2 * a.d3;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = e % a.d3, t = e / a.d3;
        int64_t k = t % a.d2;
        t /= a.d2;
        int64_t j = t % a.d1, i = t / a.d1;
        uint64_t r = splitmix64_at(a.seed, (uint64_t)e);
        a.data[i * a.st[0] + j * a.st[1] + k * a.st[2] + l * a.st[3]] = (int64_t)csd_umulhi(r, (uint64_t)mt.m[j].N);
    }
}

extern "C" int csd_fill_uniform(csd_ctx* ctx, const csd_view* dst, const csd_modulus* mods, size_t nmods, uint64_t seed) {
    CSD_TRY check_view(dst, "fill_uniform");
    ModTable t = make_mod_table(mods, nmods, dst->shape[1]);
    FillArgs a;
    a.d0 = dst->shape[0];
    a.d1 = dst->shape[1];
    a.d2 = dst->shape[2];
    a.d3 = dst->shape[3];
    for (int d = 0; d < 4; d++) a.st[d] = dst->strides[d];
    a.data = (int64_t*)dst->ptr;
    a.seed = seed;
    int64_t total = numel(dst);
    if (total == 0) return 0;
    unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_fill_uniform, dim3(blocks), dim3(256), 0, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// batched Galois key switching

#define GKS_CB 8  // ciphertexts per thread: each key digit load serves GKS_CB ciphertexts

struct GksArgs {
    int64_t r, n2, n3;
    int d, batch;
    const int64_t* parts;  // [batch, d, r, n2, n3]
    const int64_t* c0;     // [batch, r, n2, n3]
    const int64_t* key;    // row 2i + s of digit i at key + (2i + s) krow, limbs contiguous planes
    int64_t krow;
    const uint32_t* perm2;
    const uint32_t* perm3;
    int64_t* out;  // element e = h grp + l: row h o_hi + l o_lo + s o_s
    int64_t grp, c0_hi, c0_lo, o_hi, o_lo, o_s;  // c0 of element e at row h c0_hi + l c0_lo
    int full0, full1;
    int64_t chunks;  // position chunks of 256 along dim 3
    int64_t groups;  // ciphertext groups of GKS_CB
    int64_t nblocks;
    int64_t one[CSD_MAX_MODULI];  // mont(1) per limb
};

__device__ __forceinline__ int64_t reduce_wide_entry(const CsdMod& m, i128 x, int full) {
    // ring.rs:846-865: a single Montgomery reduction, or the double reduction then to_mont64
    return full ? csd_to_mont64(m, csd_double_red128(m, x)) : csd_red128(m, x);
}

__global__ __launch_bounds__(256) void k_galois_keyswitch(GksArgs a, ModTable mt) {
    // block -> (j, k, l chunk, ciphertext group), the group fastest: the groups of one position
    // read the same key entries, so they run back to back, and the XCD-aware order (the
    // hardware deals consecutive workgroups to the 8 XCDs in turn; tile t = (b mod 8) nblocks/8
    // + b/8) keeps them on one XCD, whose L2 then serves the key to all but the first
    int64_t b = blockIdx.x;
    if (a.nblocks % 8 == 0) b = (b % 8) * (a.nblocks / 8) + b / 8;
    const int64_t cg = b % a.groups;
    b /= a.groups;
    const int64_t c = b % a.chunks;
    b /= a.chunks;
    const int64_t k = b % a.n2;
    const int64_t j = b / a.n2;
    const int64_t l = c * 256 + threadIdx.x;
    if (l >= a.n3) return;
    const int64_t plane = a.n2 * a.n3;
    const int64_t limb = a.r * plane;
    const int64_t src_pos = j * plane + (int64_t)a.perm2[k] * a.n3 + a.perm3[l];
    const int64_t dst_pos = j * plane + k * a.n3 + l;
    const int c_first = (int)(cg * GKS_CB);
    const int nc = min(GKS_CB, a.batch - c_first);
    i128 acc0[GKS_CB], acc1[GKS_CB];
#pragma unroll
    for (int q = 0; q < GKS_CB; q++) acc0[q] = acc1[q] = 0;
    if (nc == GKS_CB) {
        // a full group: the loads of a digit carry no per-ciphertext condition, so all ten are in
        // flight together, and the next digit's loads are issued before this digit's products
        const int64_t* P = a.parts + (int64_t)c_first * a.d * limb + src_pos;
        const int64_t cstride = (int64_t)a.d * limb;
        const int64_t* K = a.key + dst_pos;
        const int64_t krow = a.krow;
        int64_t pv[GKS_CB], k0 = 0, k1 = 0;
#pragma unroll
        for (int q = 0; q < GKS_CB; q++) pv[q] = P[q * cstride];
        k0 = K[0];
        k1 = K[krow];
        for (int i = 0; i < a.d; i++) {
            int64_t pn[GKS_CB], k0n = 0, k1n = 0;
            if (i + 1 < a.d) {
#pragma unroll
                for (int q = 0; q < GKS_CB; q++) pn[q] = P[q * cstride + (int64_t)(i + 1) * limb];
                k0n = K[(int64_t)(2 * i + 2) * krow];
                k1n = K[(int64_t)(2 * i + 3) * krow];
            }
#pragma unroll
            for (int q = 0; q < GKS_CB; q++) {
                acc0[q] = (i128)((u128)acc0[q] + (u128)((i128)pv[q] * (i128)k0));
                acc1[q] = (i128)((u128)acc1[q] + (u128)((i128)pv[q] * (i128)k1));
            }
            if (i + 1 < a.d) {
#pragma unroll
                for (int q = 0; q < GKS_CB; q++) pv[q] = pn[q];
                k0 = k0n;
                k1 = k1n;
            }
        }
    } else {
        for (int i = 0; i < a.d; i++) {
            const int64_t k0 = a.key[(int64_t)(2 * i) * a.krow + dst_pos];
            const int64_t k1 = a.key[(int64_t)(2 * i + 1) * a.krow + dst_pos];
#pragma unroll
            for (int q = 0; q < GKS_CB; q++) {
                if (q < nc) {
                    const int64_t p = a.parts[((int64_t)(c_first + q) * a.d + i) * limb + src_pos];
                    acc0[q] = (i128)((u128)acc0[q] + (u128)((i128)p * (i128)k0));
                    acc1[q] = (i128)((u128)acc1[q] + (u128)((i128)p * (i128)k1));
                }
            }
        }
    }
    const CsdMod m = mt.m[j];
    const int64_t one = a.one[j];
#pragma unroll
    for (int q = 0; q < GKS_CB; q++) {
        if (q < nc) {
            const int64_t e = c_first + q, h = e / a.grp, l = e - h * a.grp;
            const int64_t x0 = a.c0[(h * a.c0_hi + l * a.c0_lo) * limb + src_pos];
            const i128 s0 = (i128)((u128)acc0[q] + (u128)((i128)x0 * (i128)one));
            int64_t* o = a.out + (h * a.o_hi + l * a.o_lo) * limb + dst_pos;
            o[0] = reduce_wide_entry(m, s0, a.full0);
            o[a.o_s * limb] = reduce_wide_entry(m, acc1[q], a.full1);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// mod-switch by one limb for a work-list of uneven ciphertext elements (bfv.rs:628-662 with the
// single-prime BConvCache of ring.rs:1606-1701): per element e at level r (limbs 0..r-1, period
// basis) and position pos, with x the dropped limb's value:
//   x1   = mont_mul(x, qopi)                                  qopi = lift((q/p_d)^-1) = 1 here
//   d_i  = double_red(M_i x1) + red(((C x1 + 2^(lq-1)) >> lq) qm_i),   then to_mont64
//   out_i = mont_mul(in_i - d_i, inv_i)                        (i < r - 1)
// M_i = lift(2^128 mod p_i), C = round(2^lq / p_d), qm_i = lift(-p_d 2^64 mod p_i) and
// inv_i = mont(lift(p_d^-1 mod p_i)) are the per-level constants of `MsLevel`.

struct MsLevel {
    int64_t M[CSD_MAX_MODULI], qm[CSD_MAX_MODULI], inv[CSD_MAX_MODULI];
    int64_t C, qopi;
    int32_t lq, r;
};

struct MsArgs {
    int64_t* const* in;     // per element: [r, npos] (period basis)
    int64_t* const* out;    // per element: [r - 1, npos]
    const int32_t* level;   // per element: index into `levels`
    const MsLevel* levels;
    int64_t npos, chunks;
};

__global__ __launch_bounds__(256) void k_mod_switch_one(MsArgs a, ModTable mt) {
    __shared__ int64_t cm[3 * CSD_MAX_MODULI];  // the level's M, qm, inv, staged once per workgroup
    const int64_t e = blockIdx.x / a.chunks;
    const int64_t pos = (blockIdx.x % a.chunks) * 256 + threadIdx.x;
    const MsLevel& lv = a.levels[a.level[e]];
    const int r = lv.r;
    for (int q = threadIdx.x; q < r - 1; q += 256) {
        cm[q] = lv.M[q];
        cm[CSD_MAX_MODULI + q] = lv.qm[q];
        cm[2 * CSD_MAX_MODULI + q] = lv.inv[q];
    }
    __syncthreads();
    if (pos >= a.npos) return;
    const int64_t* in = a.in[e];
    int64_t* out = a.out[e];
    const CsdMod md = mt.m[r - 1];
    const int64_t x1 = csd_mont_mul(md, in[(int64_t)(r - 1) * a.npos + pos], lv.qopi);
    const i128 corr = (i128)lv.C * (i128)x1;
    const i128 q = (i128)((u128)corr + ((u128)1 << (lv.lq - 1))) >> lv.lq;
    for (int i = 0; i < r - 1; i++) {
        const CsdMod m = mt.m[i];
        int64_t d = csd_double_red128(m, (i128)cm[i] * (i128)x1);
        d = (int64_t)((uint64_t)d + (uint64_t)csd_red128(m, (i128)((u128)q * (u128)(i128)cm[CSD_MAX_MODULI + i])));
        d = csd_to_mont64(m, d);
        const int64_t diff = (int64_t)((uint64_t)in[(int64_t)i * a.npos + pos] - (uint64_t)d);
        out[(int64_t)i * a.npos + pos] = csd_mont_mul(m, diff, cm[2 * CSD_MAX_MODULI + i]);
    }
}

void csd_mod_switch_one_launch(csd_ctx* ctx, int64_t* const* in, int64_t* const* out, const int32_t* level, const void* levels,
                               int64_t nelems, int64_t npos, const csd_modulus* mods, size_t nmods, double bytes) {
    MsArgs a;
    a.in = in;
    a.out = out;
    a.level = level;
    a.levels = (const MsLevel*)levels;
    a.npos = npos;
    a.chunks = (npos + 255) / 256;
    ModTable t = make_mod_table(mods, nmods, (int64_t)nmods);
    if (nelems == 0) return;
    TimedScope ts(ctx, "mod_switch_one", bytes);
    hipLaunchKernelGGL(k_mod_switch_one, dim3((unsigned)(nelems * a.chunks)), dim3(256), 0, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
}

size_t csd_mod_switch_level_bytes() { return sizeof(MsLevel); }

void csd_mod_switch_fill_level(void* dst, int r, const int64_t* M, const int64_t* qm, const int64_t* inv, int64_t C, int lq,
                               int64_t qopi) {
    MsLevel* lv = (MsLevel*)dst;
    memset(lv, 0, sizeof(MsLevel));
    for (int i = 0; i < r - 1; i++) {
        lv->M[i] = M[i];
        lv->qm[i] = qm[i];
        lv->inv[i] = inv[i];
    }
    lv->C = C;
    lv->qopi = qopi;
    lv->lq = lq;
    lv->r = r;
}

// exported through the ring (which owns the permutations); see csd_ring.hip
void csd_galois_keyswitch_launch(csd_ctx* ctx, int64_t batch, int64_t d, int64_t r, int64_t n2, int64_t n3, const int64_t* parts,
                                 const int64_t* c0, const int64_t* key, int64_t krow, const uint32_t* perm2,
                                 const uint32_t* perm3, int64_t* out, const GksRows& rows, const csd_modulus* mods, int full0,
                                 int full1) {
    GksArgs a;
    memset(&a, 0, sizeof(a));
    a.r = r;
    a.n2 = n2;
    a.n3 = n3;
    a.d = (int)d;
    a.batch = (int)batch;
    a.parts = parts;
    a.c0 = c0;
    a.key = key;
    a.krow = krow;
    a.perm2 = perm2;
    a.perm3 = perm3;
    a.out = out;
    a.grp = rows.grp;
    a.c0_hi = rows.c0_hi;
    a.c0_lo = rows.c0_lo;
    a.o_hi = rows.o_hi;
    a.o_lo = rows.o_lo;
    a.o_s = rows.o_s;
    a.full0 = full0;
    a.full1 = full1;
    a.chunks = (n3 + 255) / 256;
    for (int64_t jj = 0; jj < r; jj++) a.one[jj] = (int64_t)(((u128)1 << 64) % (u128)mods[jj].N);  // mont(1)
    ModTable t = make_mod_table(mods, (size_t)r, r);
    int64_t groups = (batch + GKS_CB - 1) / GKS_CB;
    int64_t blocks = groups * r * n2 * a.chunks;
    if (blocks > (int64_t)INT32_MAX) throw CsdError("galois key switch: grid too large");
    a.groups = groups;
    a.nblocks = blocks;
    double bytes = 8.0 * (double)(batch * d * r * n2 * n3 + batch * r * n2 * n3 + 2 * batch * r * n2 * n3 + 2 * d * r * n2 * n3);
    TimedScope ts(ctx, "galois_keyswitch", bytes);
    hipLaunchKernelGGL(k_galois_keyswitch, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// base conversions as one kernel: GadgetBConvCache::bconv (ring.rs:1809-1851) and
// BConvCache::bconv (ring.rs:1663-1701)
//
// The HAL sequence multiplies the source by (D_i/p_u)^-1 in Montgomery form, forms the
// [d, r' + 1] i128 matrix product over chunks of `acc` source limbs (each chunk reduced with
// double_red128 and added), adds round(corr / 2^shift) * (-D_i 2^64 mod p_j) through red128 and
// converts to Montgomery form. The BEHZ conversion is the case of one digit over the whole source
// base, with a target base of its own. Row i of the gadget matrix is zero outside digit i, so
// thread (b, pos) only multiplies the digit's own limbs; a zero product adds nothing to a
// wrapping i128 sum, and every chunk, empty or not, is still reduced and added, so the residues
// are those of the sequence. (Ring primes are below 2^57, so a chunk holds at least 2^13
// products and up to 64 limbs always form one chunk.)

struct GbArgs {
    int64_t B, d, r, rd, rdp, n3, P;  // elements, digits, source limbs, target limbs (padded), dim-3 length, positions
    int64_t ss[4], sd[4];             // source [B, r, n2, n3] and destination [B d, rd, n2, n3] strides
    const int64_t* src;
    int64_t* dst;
    const int64_t* dopi;  // [r]: (D_i / p_u)^-1 for u in digit i
    const int64_t* mat;   // [d][rd + 1][r]: row rd holds the correction weights
    const int64_t* matT;  // [d][r][rdp]: (D_i / p_u) R^2 mod p_j, transposed, zero beyond rd
    const int64_t* dmp;   // [d][rd]: -D_i 2^64 mod p_j
    const CsdMod* smods;  // source moduli (device), or null when the source base is the target base
    int shift;            // log2 of q~
    int64_t xs_rows;      // rows of the LDS limb columns (the longest digit)
    // rescale form (all null otherwise): each source limb u is first taken to mont(x, pre[u]), and
    // each target row j of dst, read before it is written, becomes mont(mont(dst, post_s[j]) - v, post_v[j])
    // for the converted value v
    const int64_t* pre;
    const int64_t* post_s;
    const int64_t* post_v;
    // mod-switch form (null otherwise): each target row j of dst becomes mont(keep[j] - v, post_v[j]),
    // keep a [B, rd, n2, n3] view (the kept limbs of the element)
    const int64_t* keep;
    int64_t ks[4];
    int32_t lo[CSD_MAX_MODULI], hi[CSD_MAX_MODULI];
};

// the output forms of k_bconv_fused
#define GB_PLAIN 0
#define GB_RESCALE 1
#define GB_MODSWITCH 2

__device__ __forceinline__ i128 gb_mac(i128 acc, int64_t a, int64_t b) {
    uint64_t lo;
    int64_t hi;
    csd_smul_wide(a, b, lo, hi);
    return (i128)((u128)acc + (((u128)(uint64_t)hi << 64) | lo));
}

#define GB_GROUP 8  // target limbs per pass over the digit's source limbs

// The sums sum_u w_u x_u of the conversions are exact integers: every weight is a smallest lift
// mod a ring prime (|w| < 2^56) and every lifted limb a Montgomery product (|x| <= N < 2^57), so
// with w = w1 2^28 + w0 and x = x1 2^29 + x0 (w0 in [0, 2^28), x0 in [0, 2^29)) each of the four
// partial sums over at most 64 limbs stays below 2^63 in magnitude. They accumulate in int64 with
// one signed 32 x 32 -> 64 multiply-add each, and the exact i128 sum (below 2^119, so equal to
// the HAL's wrapping one) is put together once per target limb.
struct GbSplit {
    int64_t a00, a01, a10, a11;  // sum w0 x0, w0 x1, w1 x0, w1 x1
};

__device__ __forceinline__ void gb_split_mac(GbSplit& s, int64_t w, int32_t x0, int32_t x1) {
    const int32_t w0 = (int32_t)(w & ((1 << 28) - 1)), w1 = (int32_t)(w >> 28);
    s.a00 += (int64_t)w0 * x0;
    s.a01 += (int64_t)w0 * x1;
    s.a10 += (int64_t)w1 * x0;
    s.a11 += (int64_t)w1 * x1;
}

__device__ __forceinline__ i128 gb_split_sum(const GbSplit& s) {
    return (i128)s.a00 + ((i128)s.a01 << 29) + ((i128)s.a10 << 28) + ((i128)s.a11 << 57);
}

// One thread per (element, position). Launched only when every digit's source limbs fit one i128
// chunk, which the ring's primes (below 2^57) guarantee; so each target limb's sum is reduced once.
// The digit's source limbs are loaded eight at a time (so that their loads overlap), lifted, and
// parked in the thread's LDS column for the passes over GB_GROUP target limbs. SPLIT accumulates
// in the four int64 partial sums (worth it for long digits; short ones pay more for the final
// i128 assembly than they save). WV = 1 stages the digit's weights in LDS once per workgroup
// (broadcast reads); WV = 0 reads them with scalar loads, for weight tables beyond the LDS. FM is
// the output form: GB_PLAIN the conversion, GB_RESCALE the rescale step, GB_MODSWITCH the
// subtraction from the kept limbs and the multiply by the inverse of mod_switch_down.
template <bool SPLIT, int WV, int FM>
__global__ __launch_bounds__(256) void k_bconv_fused(GbArgs a, ModTable mt) {
    extern __shared__ int64_t gb_x[];
    int64_t* xs = gb_x + threadIdx.x;
    int64_t* wl = gb_x + a.xs_rows * 256;  // WV == 1: the digit's weights [len][rdp], shared by the block
    const int64_t chunks = (a.P + 255) / 256;
    const int64_t b = blockIdx.x / chunks;
    const int64_t pos = (blockIdx.x % chunks) * 256 + threadIdx.x;
    const bool active = pos < a.P;
    if (WV == 0 && !active) return;  // that form has no barrier
    const int64_t k = active ? pos / a.n3 : 0, l = active ? pos % a.n3 : 0;
    const int64_t* s = a.src + b * a.ss[0] + k * a.ss[2] + l * a.ss[3];
    int64_t* o = a.dst + b * a.d * a.sd[0] + k * a.sd[2] + l * a.sd[3];
    const int r = (int)a.r, rd = (int)a.rd;
    const i128 half = (i128)1 << (a.shift - 1);
    for (int i = 0; i < (int)a.d; i++) {
        const int lo = a.lo[i], hi = a.hi[i];
        const int64_t* corr_row = a.mat + ((int64_t)i * (rd + 1) + rd) * r;
        const int64_t* wi = a.matT + (int64_t)i * r * a.rdp;
        if (WV == 1) {
            __syncthreads();  // the previous digit's weights are no longer read
            for (int e = threadIdx.x; e < (hi - lo) * a.rdp; e += 256) wl[e] = wi[(int64_t)lo * a.rdp + e];
            __syncthreads();
        }
        if (!active) continue;
        i128 corr = 0;
        for (int u0 = lo; u0 < hi; u0 += 8) {
            int64_t raw[8];
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (u0 + t < hi) raw[t] = s[(u0 + t) * a.ss[1]];
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (u0 + t < hi) {
                    const int u = u0 + t;
                    const CsdMod mu = a.smods ? a.smods[u] : mt.m[u];
                    const int64_t xin = FM == GB_RESCALE ? csd_mont_mul(mu, raw[t], a.pre[u]) : raw[t];
                    const int64_t xv = csd_mont_mul(mu, xin, a.dopi[u]);
                    xs[(u - lo) * 256] = xv;
                    corr = gb_mac(corr, corr_row[u], xv);
                }
        }
        const int64_t q = (int64_t)((i128)((u128)corr + (u128)half) >> a.shift);  // |q| <= 2^63
        // the fast product needs |q| < 2^63 - 2^31 (the ring's conversions have |q| < 2^8); the
        // wave takes the general one if any of its positions is beyond 2^62
        const bool qfast = __all(q < ((int64_t)1 << 62) && q > -((int64_t)1 << 62));
        for (int j0 = 0; j0 < rd; j0 += GB_GROUP) {
            GbSplit ssum[GB_GROUP];
            i128 wsum[GB_GROUP];
#pragma unroll
            for (int g = 0; g < GB_GROUP; g++) {
                ssum[g] = GbSplit{0, 0, 0, 0};
                wsum[g] = 0;
            }
#pragma unroll 2
            for (int u = lo; u < hi; u++) {
                const int64_t xv = xs[(u - lo) * 256];
                const int64_t* w = WV == 1 ? wl + (int64_t)(u - lo) * a.rdp + j0 : wi + (int64_t)u * a.rdp + j0;
                if (SPLIT) {
                    const int32_t x0 = (int32_t)(xv & ((1 << 29) - 1)), x1 = (int32_t)(xv >> 29);
#pragma unroll
                    for (int g = 0; g < GB_GROUP; g++)
                        if (j0 + g < rd) gb_split_mac(ssum[g], w[g], x0, x1);
                } else {
#pragma unroll
                    for (int g = 0; g < GB_GROUP; g++)
                        if (j0 + g < rd) wsum[g] = gb_mac(wsum[g], w[g], xv);
                }
            }
#pragma unroll
            for (int g = 0; g < GB_GROUP; g++) {
                const int j = j0 + g;
                if (j < rd) {
                    const i128 sum = SPLIT ? gb_split_sum(ssum[g]) : wsum[g];
                    // the fast Montgomery forms (csd_mont_fast.hpp): |sum| < 2^119, and the values
                    // below stay within [-2N, 2N]
                    const CsdMod& mj = mt.m[j];
                    const CsdFastMod f = csd_fast_mod(mj.N, mj.ninv);
                    const int64_t dm = a.dmp[(int64_t)i * rd + j];
                    int64_t v = csd_double_red_fast(f, (uint64_t)sum, (int64_t)(sum >> 64));
                    const int64_t c = qfast ? csd_mont_mul_fc(f, q, csd_fast_const(dm, mj.ninv)) : csd_mont_mul(mj, q, dm);
                    v = (int64_t)((uint64_t)v + (uint64_t)c);
                    v = csd_mont_mul_fc(f, v, csd_fast_const(mj.r2, mj.ninv));  // to_mont64
                    int64_t* oj = o + i * a.sd[0] + j * a.sd[1];
                    if (FM == GB_RESCALE) {
                        const int64_t e = csd_mont_mul(mj, *oj, a.post_s[j]);  // any i64
                        v = csd_mont_mul_fc(f, (int64_t)((uint64_t)e - (uint64_t)v), csd_fast_const(a.post_v[j], mj.ninv));
                    } else if (FM == GB_MODSWITCH) {
                        // sub_i64 and mul_mont64 of the HAL sequence; the kept limb is any i64
                        const int64_t kv = a.keep[b * a.ks[0] + k * a.ks[2] + l * a.ks[3] + j * a.ks[1]];
                        v = csd_mont_mul(mj, (int64_t)((uint64_t)kv - (uint64_t)v), a.post_v[j]);
                    }
                    *oj = v;
                }
            }
        }
    }
}

void csd_bconv_fused_launch(csd_ctx* ctx, const csd_view& dst, const csd_view& src, int64_t d, const int64_t* dopi,
                            const int64_t* mat, const int64_t* matT, int64_t rdp, const int64_t* dmp,
                            const std::vector<size_t>& bounds, size_t shift, const csd_modulus* dmods,
                            const CsdMod* smods_dev, const char* timer, const int64_t* pre, const int64_t* post_s,
                            const int64_t* post_v, const csd_view* keep) {
    GbArgs a;
    memset(&a, 0, sizeof(a));
    a.B = src.shape[0];
    a.d = d;
    a.r = src.shape[1];
    a.rd = dst.shape[1];
    a.rdp = rdp;
    a.n3 = src.shape[3];
    a.P = src.shape[2] * src.shape[3];
    for (int q = 0; q < 4; q++) {
        a.ss[q] = src.strides[q];
        a.sd[q] = dst.strides[q];
    }
    a.src = (const int64_t*)src.ptr;
    a.dst = (int64_t*)dst.ptr;
    a.dopi = dopi;
    a.mat = mat;
    a.matT = matT;
    a.dmp = dmp;
    a.smods = smods_dev;
    a.shift = (int)shift;
    a.pre = pre;
    a.post_s = post_s;
    a.post_v = post_v;
    // the rescale form takes pre, post_s and post_v together (the kernel reads all three); the
    // mod-switch form (keep) post_v alone
    if (!keep && ((pre == nullptr) != (post_s == nullptr) || (post_s == nullptr) != (post_v == nullptr) || (post_s && d != 1)))
        throw CsdError("fused bconv: the rescale form needs pre, post_s and post_v together, and one digit");
    if (keep) {
        if (pre || post_s || !post_v || d != 1) throw CsdError("fused bconv: the mod-switch form takes post_v only, and one digit");
        for (int q = 0; q < 4; q++)
            if (keep->shape[q] != dst.shape[q]) throw CsdError("fused bconv: kept limbs must have the destination's shape");
        a.keep = (const int64_t*)keep->ptr;
        for (int q = 0; q < 4; q++) a.ks[q] = keep->strides[q];
    }
    if (!smods_dev && a.rd != a.r) throw CsdError("fused bconv: source moduli missing");
    if (rdp < a.rd || rdp % GB_GROUP) throw CsdError("fused bconv: transposed weights not padded");
    for (int64_t i = 0; i < d; i++) {
        a.lo[i] = (int32_t)bounds[2 * i];
        a.hi[i] = (int32_t)bounds[2 * i + 1];
    }
    if (a.B * a.P == 0) return;
    const int64_t blocks = a.B * ((a.P + 255) / 256);
    if (blocks > (int64_t)INT32_MAX) throw CsdError("fused bconv: grid too large");
    ModTable t = make_mod_table(dmods, (size_t)a.rd, a.rd);
    TimedScope ts(ctx, timer, 8.0 * (double)a.B * a.P * ((double)a.r + (double)d * a.rd * (post_s || keep ? 2 : 1)));
    int64_t max_len = 0;
    for (int64_t i = 0; i < d; i++) max_len = std::max<int64_t>(max_len, a.hi[i] - a.lo[i]);
    const size_t lds0 = (size_t)max_len * 256 * sizeof(int64_t);
    a.xs_rows = max_len;
    const size_t lds1 = (size_t)max_len * 256 * sizeof(int64_t) + (size_t)max_len * a.rdp * sizeof(int64_t);
    const bool split = max_len >= 4, wv = lds1 <= 160 * 1024;  // the digit's weights in LDS: 15-20 % faster
    auto launch = [&](auto kern) {
        csd_allow_lds((const void*)kern, 160 * 1024);
        hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), wv ? lds1 : lds0, ctx->stream, a, t);
    };
    if (post_s) {  // the rescale form: one digit of many limbs
        if (wv) launch(k_bconv_fused<true, 1, GB_RESCALE>);
        else launch(k_bconv_fused<true, 0, GB_RESCALE>);
    } else if (keep) {  // the mod-switch form: few dropped limbs
        if (!wv) launch(split ? k_bconv_fused<true, 0, GB_MODSWITCH> : k_bconv_fused<false, 0, GB_MODSWITCH>);
        else if (split) launch(k_bconv_fused<true, 1, GB_MODSWITCH>);
        else launch(k_bconv_fused<false, 1, GB_MODSWITCH>);
    } else if (wv) {
        if (split) launch(k_bconv_fused<true, 1, GB_PLAIN>);
        else launch(k_bconv_fused<false, 1, GB_PLAIN>);
    } else {
        if (split) launch(k_bconv_fused<true, 0, GB_PLAIN>);
        else launch(k_bconv_fused<false, 0, GB_PLAIN>);
    }
    HIPCHECK(hipGetLastError());
}

// the padded, transposed weight layout csd_bconv_fused_launch reads: out[i][u][j] = mat[i][j][u]
std::vector<int64_t> csd_bconv_transpose(const std::vector<int64_t>& mat, int64_t d, int64_t rd, int64_t r, int64_t* rdp) {
    if (r > 64) throw CsdError("fused bconv: at most 64 source limbs");
    *rdp = (rd + GB_GROUP - 1) / GB_GROUP * GB_GROUP;
    std::vector<int64_t> t((size_t)(d * r * *rdp), 0);
    for (int64_t i = 0; i < d; i++)
        for (int64_t j = 0; j < rd; j++)
            for (int64_t u = 0; u < r; u++) {
                const int64_t w = mat[(size_t)((i * (rd + 1) + j) * r + u)];
                if (w >= ((int64_t)1 << 56) || w < -((int64_t)1 << 56)) throw CsdError("fused bconv: weight out of range");
                t[(size_t)((i * r + u) * *rdp + j)] = w;
            }
    return t;
}

// ---------------------------------------------------------------------------------------------
// The BEHZ conversion (one digit over the whole source base; BConvCache::bconv, ring.rs:1663-1701,
// and the rescale form above) with its matrix product on the int8 matrix cores.
//
// The sums S_j = sum_u W[j][u] x_u are exact integers, and so is the correction sum, which rides
// along as matrix row rd. Every weight (|w| < 2^63) and every lifted limb (|x| <= N < 2^57) is
// written in eight balanced base-256 digits, w = sum_a w_a 256^a and x = sum_b x_b 256^b with
// w_a, x_b in [-128, 127]: the bytes of w + B, each minus 128, for B = sum_a 128 256^a. The 64
// digit pairs (a, b) are 64 v_mfma_i32_32x32x32_i8, each over all source limbs at once (the K
// slice of 32 bytes holds limb u in lane half u / 16, byte u % 16), accumulated by shift class
// s = a + b: |C_s| <= 8 * 32 * 2^14 = 2^22 in int32. The classes are then combined exactly,
// c_t = C_2t + 256 C_(2t+1) (|c_t| < 2^31) and R_t = c_t + (R_(t-1) >> 16), which gives
// S = sum_(t<7) (R_t mod 2^16) 2^(16t) + R_7 2^112, |S| < 2^125 and |R_7| <= 2^13, in four
// 32-bit words. The tail (double_red128, the correction q (-D 2^64 mod p_j), to_mont64 and the
// rescale form's post-multiplications) is that of k_bconv_fused, so the residues are identical.
//
// Workgroup: 4 waves, each owning 64 positions of one element. Phase 1 (lane = position) lifts the
// source limbs and writes their digit bytes to the wave's LDS slice as [digit][limb quad][64];
// phase 2, per block of 32 positions, reads the B operands, runs the 64 MFMAs and the carry chain;
// phase 3 runs the tail in the MFMA output layout (lane l, element 4i + t: row 8i + 4 (l / 32) + t,
// position l % 32), where rows 0 .. rd - 1 leave at most a quarter of the lanes idle.

typedef int gm_v4i __attribute__((ext_vector_type(4)));
typedef int gm_v16i __attribute__((ext_vector_type(16)));

#define GM_WAVES 4
#define GM_MAX_SRC 32  // source limbs: one 32-byte K slice
#define GM_MAX_ROWS 32  // target limbs plus the correction row
static constexpr uint64_t GM_BIAS = 0x8080808080808080ull;
// per target row j, from the host: N, ninv, n0 | n1 << 32, n1b (csd_fast_mod), dmp_j balanced
// (lo | hi << 32) and dmp_j ninv, r2 balanced and r2 ninv; in LDS, for the rescale form, also
// post_s[j], post_v[j] balanced and post_v[j] ninv
#define GM_ROW_HOST 8
#define GM_ROW 12

__device__ __forceinline__ CsdFastConst gm_fc(int64_t packed, int64_t cn) {
    return CsdFastConst{{(int32_t)(uint32_t)(uint64_t)packed, (int32_t)(packed >> 32)}, (uint64_t)cn};
}

struct GmArgs {
    int64_t P, n3, chunks;
    int32_t r, rd, nq, shift;
    int64_t ss[4], sd[4];
    const int64_t* src;
    int64_t* dst;
    const int64_t* dopi;   // [r]
    const CsdMod* smods;   // [r] source moduli
    const gm_v4i* atab;    // [8][64]: lane l's 16 bytes of weight digit a (csd_bconv_mfma_tables)
    const int64_t* rows;   // [rd][GM_ROW_HOST]: target limb j's constants (csd_bconv_mfma_tables)
    const int64_t* pre;    // rescale form (null otherwise): as in GbArgs
    const int64_t* post_s;
    const int64_t* post_v;
};

// 4 x 4 byte transpose: out[b] byte e = byte b of in[e]
__device__ __forceinline__ void gm_transpose4(const uint32_t (&in)[4], uint32_t (&out)[4]) {
    const uint32_t lo01 = __builtin_amdgcn_perm(in[1], in[0], 0x05010400u), hi01 = __builtin_amdgcn_perm(in[1], in[0], 0x07030602u);
    const uint32_t lo23 = __builtin_amdgcn_perm(in[3], in[2], 0x05010400u), hi23 = __builtin_amdgcn_perm(in[3], in[2], 0x07030602u);
    out[0] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
    out[1] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
    out[2] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
    out[3] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
}

// ROWG groups of 8 matrix rows hold rows 0 .. rd (8 ROWG >= rd + 1)
template <bool RS, int ROWG>
__global__ __launch_bounds__(64 * GM_WAVES) void k_bconv_mfma(GmArgs a) {
    extern __shared__ uint32_t gm_lds[];
    int64_t* rowc = (int64_t*)gm_lds;  // [GM_MAX_ROWS][GM_ROW]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t* xt = gm_lds + GM_MAX_ROWS * GM_ROW * 2 + wave * (8 * a.nq * 64);  // [digit][quad][64 positions]
    for (int j = threadIdx.x; j < a.rd; j += 64 * GM_WAVES) {
        int64_t* rj = rowc + GM_ROW * j;
#pragma unroll
        for (int f = 0; f < GM_ROW_HOST; f++) rj[f] = a.rows[GM_ROW_HOST * j + f];
        if (RS) {
            const int64_t pv = a.post_v[j];
            const CsdBal bv = csd_bal(pv);
            rj[8] = a.post_s[j];
            rj[9] = (int64_t)(((uint64_t)(uint32_t)bv.hi << 32) | (uint32_t)bv.lo);
            rj[10] = (int64_t)((uint64_t)pv * (uint64_t)rj[1]);
        }
    }
    const int64_t b = blockIdx.x / a.chunks;
    const int64_t p0 = (blockIdx.x % a.chunks) * (64 * GM_WAVES) + wave * 64;
    const int64_t* sb = a.src + b * a.ss[0];
    int64_t* db = a.dst + b * a.sd[0];

    // phase 1: lane = position p0 + lane; lift every source limb, split into digit bytes (lanes
    // past the end read the last position and contribute zeros)
    {
        const int64_t pos = p0 + lane;
        const bool act = pos < a.P;
        const int64_t pc = act ? pos : a.P - 1;
        const int64_t* s = sb + (pc / a.n3) * a.ss[2] + (pc % a.n3) * a.ss[3];
        int64_t raw[GM_MAX_SRC];  // every limb's load in flight at once
#pragma unroll
        for (int u = 0; u < GM_MAX_SRC; u++) raw[u] = u < a.r ? s[u * a.ss[1]] : 0;
#pragma unroll
        for (int q = 0; q < GM_MAX_SRC / 4; q++) {
            if (q >= a.nq) break;
            uint32_t lo[4], hi[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int u = min(4 * q + e, a.r - 1);
                const CsdMod mu = a.smods[u];
                const int64_t xin = RS ? csd_mont_mul(mu, raw[4 * q + e], a.pre[u]) : raw[4 * q + e];
                int64_t xv = csd_mont_mul(mu, xin, a.dopi[u]);
                xv = (act && 4 * q + e < a.r) ? xv : 0;
                const uint64_t d = ((uint64_t)xv + GM_BIAS) ^ GM_BIAS;
                lo[e] = (uint32_t)d;
                hi[e] = (uint32_t)(d >> 32);
            }
            uint32_t dl[4], dh[4];
            gm_transpose4(lo, dl);
            gm_transpose4(hi, dh);
#pragma unroll
            for (int g = 0; g < 4; g++) {
                xt[(g * a.nq + q) * 64 + lane] = dl[g];
                xt[((g + 4) * a.nq + q) * 64 + lane] = dh[g];
            }
        }
    }
    __syncthreads();

    const int h = lane >> 5, n = lane & 31;
    // the correction row rd sits in lane half hc, element ec
    const int hc = (a.rd >> 2) & 1, ec = 4 * (a.rd >> 3) + (a.rd & 3);
    const i128 half = (i128)1 << (a.shift - 1);
#pragma unroll 1
    for (int blk = 0; blk < 2; blk++) {
        // phase 2: B operands of this block's 32 positions, the 64 digit-pair MFMAs, the carry chain
        // (the weight operands are read again per block, so they are not live during the tail)
        gm_v4i A[8], Bv[8];
#pragma unroll
        for (int d = 0; d < 8; d++) A[d] = a.atab[d * 64 + lane];
#pragma unroll
        for (int d = 0; d < 8; d++)
#pragma unroll
            for (int qq = 0; qq < 4; qq++) {
                // quads past the source limbs meet zero weights: any bytes do
                const int q = min(4 * h + qq, a.nq - 1);
                Bv[d][qq] = (int)xt[(d * a.nq + q) * 64 + 32 * blk + n];
            }
        int32_t R[4 * ROWG], low[4 * ROWG];
        uint32_t w[4 * ROWG][4];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            gm_v16i ce = {0}, co = {0};
#pragma unroll
            for (int da = 0; da < 8; da++) {
                const int de = 2 * t - da, dodd = 2 * t + 1 - da;
                if (de >= 0 && de < 8) ce = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[da], Bv[de], ce, 0, 0, 0);
                if (dodd >= 0 && dodd < 8) co = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[da], Bv[dodd], co, 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4 * ROWG; e++) {
                const int32_t c = (int32_t)(((uint32_t)co[e] << 8) + (uint32_t)ce[e]);
                if (t == 0) {
                    R[e] = c;
                } else {
                    const int32_t prev = R[e];
                    R[e] = c + (prev >> 16);
                    if ((t - 1) % 2 == 0) low[e] = prev & 0xffff;
                    else w[e][(t - 1) / 2] = ((uint32_t)prev << 16) | (uint32_t)low[e];
                }
                // pin the chain here: otherwise it is sunk towards the tail and every class
                // accumulator stays live until then
                asm volatile("" : "+v"(R[e]));
            }
            __builtin_amdgcn_sched_barrier(0);  // one class pair's accumulators live at a time
        }
#pragma unroll
        for (int e = 0; e < 4 * ROWG; e++) w[e][3] = ((uint32_t)R[e] << 16) | (uint32_t)low[e];

        // phase 3: the correction q of each position, then the tail of every target row
        int64_t qv = 0;
#pragma unroll
        for (int e = 0; e < 4 * ROWG; e++)
            if (e == ec) {
                const i128 corr = (i128)(((u128)w[e][3] << 96) | ((u128)w[e][2] << 64) | ((u128)w[e][1] << 32) | (u128)w[e][0]);
                qv = (int64_t)((i128)((u128)corr + (u128)half) >> a.shift);
            }
        const int src_lane = n + 32 * hc;
        const int64_t q = (int64_t)(((uint64_t)(uint32_t)__shfl((int)(uint32_t)((uint64_t)qv >> 32), src_lane) << 32) |
                                    (uint32_t)__shfl((int)(uint32_t)qv, src_lane));
        const int64_t pos = p0 + 32 * blk + n;
        const bool pact = pos < a.P;
        const int64_t pc = pact ? pos : a.P - 1;
        int64_t* o = db + (pc / a.n3) * a.sd[2] + (pc % a.n3) * a.sd[3];
#pragma unroll
        for (int e = 0; e < 4 * ROWG; e++) {
            if (8 * (e >> 2) + (e & 3) >= a.rd) continue;  // no target row in either lane half
            const int j = 8 * (e >> 2) + 4 * h + (e & 3);
            const int jc = min(j, a.rd - 1);
            const int64_t* rc = rowc + GM_ROW * jc;
            const uint64_t np = (uint64_t)rc[2];
            const CsdFastMod f = {(uint64_t)rc[1], (uint32_t)np, (uint32_t)(np >> 32), (uint32_t)rc[3]};
            // |S| < 2^125, |q| < 2^61 and the values below within [-2N, 2N]: the fast forms apply
            const uint64_t lo64 = ((uint64_t)w[e][1] << 32) | w[e][0];
            const int64_t hi64 = (int64_t)(((uint64_t)w[e][3] << 32) | w[e][2]);
            int64_t v = csd_double_red_fast(f, lo64, hi64);
            v = (int64_t)((uint64_t)v + (uint64_t)csd_mont_mul_fc(f, q, gm_fc(rc[4], rc[5])));
            v = csd_mont_mul_fc(f, v, gm_fc(rc[6], rc[7]));  // to_mont64
            int64_t* oj = o + jc * a.sd[1];
            if (RS) {
                // the destination holds any i64: the general product for it
                const int64_t ev = csd_mont_mul(CsdMod{rc[0], rc[1], 0, 0}, *oj, rc[8]);
                v = csd_mont_mul_fc(f, (int64_t)((uint64_t)ev - (uint64_t)v), gm_fc(rc[9], rc[10]));
            }
            if (pact && j < a.rd) *oj = v;
        }
    }
}

bool csd_bconv_mfma_tables(const std::vector<int64_t>& mat, int64_t rd, int64_t r, const std::vector<CsdMod>& dmods,
                           const std::vector<int64_t>& dmp, std::vector<int8_t>& atab, std::vector<int64_t>& rows) {
    if (r < 1 || r > GM_MAX_SRC || rd < 1 || rd + 1 > GM_MAX_ROWS) return false;
    if ((int64_t)dmods.size() != rd || (int64_t)dmp.size() != rd || (int64_t)mat.size() != (rd + 1) * r) return false;
    // the balanced 8-digit range is [-B, 2^64 - 1 - B]: below, every i64 is inside (B > 2^63);
    // above, w + B must not wrap
    const int64_t wmax = (int64_t)(~GM_BIAS);
    for (int64_t w : mat)
        if (w > wmax) return false;
    atab.assign(8 * 64 * 16, 0);
    for (int l = 0; l < 64; l++) {
        const int64_t j = l % 32, hh = l / 32;
        for (int e = 0; e < 16; e++) {
            const int64_t u = 16 * hh + e;
            if (j > rd || u >= r) continue;
            const uint64_t d = (uint64_t)mat[(size_t)(j * r + u)] + GM_BIAS;
            for (int dg = 0; dg < 8; dg++) atab[(size_t)((dg * 64 + l) * 16 + e)] = (int8_t)(uint8_t)(((d >> (8 * dg)) & 255) ^ 128);
        }
    }
    rows.assign((size_t)(GM_ROW_HOST * rd), 0);
    auto packed = [](CsdBal b) { return (int64_t)(((uint64_t)(uint32_t)b.hi << 32) | (uint32_t)b.lo); };
    for (int64_t j = 0; j < rd; j++) {
        const CsdMod& m = dmods[(size_t)j];
        if (m.N <= 0 || m.N >= ((int64_t)1 << 62)) return false;
        const CsdFastMod f = csd_fast_mod(m.N, m.ninv);
        int64_t* rj = rows.data() + GM_ROW_HOST * j;
        rj[0] = m.N;
        rj[1] = m.ninv;
        rj[2] = (int64_t)(((uint64_t)f.n1 << 32) | f.n0);
        rj[3] = (int64_t)f.n1b;
        rj[4] = packed(csd_bal(dmp[(size_t)j]));
        rj[5] = (int64_t)((uint64_t)dmp[(size_t)j] * (uint64_t)m.ninv);
        rj[6] = packed(csd_bal(m.r2));
        rj[7] = (int64_t)((uint64_t)m.r2 * (uint64_t)m.ninv);
    }
    return true;
}

void csd_bconv_mfma_launch(csd_ctx* ctx, const csd_view& dst, const csd_view& src, const int64_t* dopi, const int8_t* atab,
                           const int64_t* rows, size_t shift, const CsdMod* smods_dev, const char* timer, const int64_t* pre,
                           const int64_t* post_s, const int64_t* post_v) {
    GmArgs a;
    memset(&a, 0, sizeof(a));
    const int64_t B = src.shape[0];
    a.r = (int32_t)src.shape[1];
    a.rd = (int32_t)dst.shape[1];
    a.n3 = src.shape[3];
    a.P = src.shape[2] * src.shape[3];
    if (a.r < 1 || a.r > GM_MAX_SRC || a.rd < 1 || a.rd + 1 > GM_MAX_ROWS) throw CsdError("mfma bconv: base sizes out of range");
    if (dst.shape[0] != B || dst.shape[2] != src.shape[2] || dst.shape[3] != src.shape[3]) throw CsdError("mfma bconv: shapes");
    if ((pre == nullptr) != (post_s == nullptr) || (post_s == nullptr) != (post_v == nullptr))
        throw CsdError("mfma bconv: the rescale form needs pre, post_s and post_v together");
    if (shift < 1 || shift > 125) throw CsdError("mfma bconv: correction shift out of range");
    for (int q = 0; q < 4; q++) {
        a.ss[q] = src.strides[q];
        a.sd[q] = dst.strides[q];
    }
    a.src = (const int64_t*)src.ptr;
    a.dst = (int64_t*)dst.ptr;
    a.dopi = dopi;
    a.smods = smods_dev;
    a.atab = (const gm_v4i*)atab;
    a.rows = rows;
    a.shift = (int32_t)shift;
    a.nq = (a.r + 3) / 4;
    a.pre = pre;
    a.post_s = post_s;
    a.post_v = post_v;
    if (B * a.P == 0) return;
    a.chunks = (a.P + 64 * GM_WAVES - 1) / (64 * GM_WAVES);
    const int64_t blocks = B * a.chunks;
    if (blocks > (int64_t)INT32_MAX) throw CsdError("mfma bconv: grid too large");
    TimedScope ts(ctx, timer, 8.0 * (double)B * a.P * ((double)a.r + (double)a.rd * (post_s ? 2 : 1)));
    const size_t lds = (size_t)GM_MAX_ROWS * GM_ROW * 8 + (size_t)GM_WAVES * 8 * a.nq * 64 * sizeof(uint32_t);
    const int rowg = (a.rd + 1 + 7) / 8;
    auto launch = [&](auto kern) {
        csd_allow_lds((const void*)kern, (int)lds);
        hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * GM_WAVES), lds, ctx->stream, a);
    };
    if (post_s) {
        if (rowg <= 2) launch(k_bconv_mfma<true, 2>);
        else if (rowg == 3) launch(k_bconv_mfma<true, 3>);
        else launch(k_bconv_mfma<true, 4>);
    } else {
        if (rowg <= 2) launch(k_bconv_mfma<false, 2>);
        else if (rowg == 3) launch(k_bconv_mfma<false, 3>);
        else launch(k_bconv_mfma<false, 4>);
    }
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// digit convolution and wide reduction in one kernel: the inner products of dbfv_mul_convolve
// (dbfv.rs:1096-1127), or of hom_mul_raw_acc (bfv.rs:667-683) as its one-digit case, followed by
// reduce_wide (ring.rs:846-865)
//
// For output digit k and pair b the three components are
//   c = 0: sum_j l_j0 r_(k-j)0,  c = 1: sum_j (l_j0 r_(k-j)1 + l_j1 r_(k-j)0),  c = 2: sum_j l_j1 r_(k-j)1
// over j = max(0, k - d2 + 1) .. min(k, d1 - 1). Each is formed as the HAL forms it, an i128 sum
// with wrapping adds (the order of wrapping adds does not change it), and reduced in registers by
// the kind the wide element's bound picks (red128, or double_red128 and to_mont64), so the
// int64 rows are those of the accumulate-then-reduce sequence and no i128 row reaches HBM.

struct CvArgs {
    int64_t n, r, P, n3;    // pairs, limbs, positions n2 n3, dim-3 length
    int32_t d1, d2, K;
    int64_t lrow[3], rrow[3];  // row of (digit, component, pair) = digit lrow[0] + comp lrow[1] + pair lrow[2]
    int64_t sl[4], sr[4], sd[4];
    const int64_t* lhs;
    const int64_t* rhs;
    int64_t* dst;  // row (c K + k) n + pair
    uint8_t full[3 * CSD_MAX_MODULI];
};

__global__ __launch_bounds__(256) void k_convolve_reduce(CvArgs a, ModTable mt) {
    const int64_t chunks = (a.P + 255) / 256;
    int64_t blk = blockIdx.x;
    const int64_t c0 = blk % chunks;
    blk /= chunks;
    const int64_t jl = blk % a.r, b = blk / a.r;  // limb, pair
    const int64_t pos = c0 * 256 + threadIdx.x;
    if (pos >= a.P) return;
    const int64_t k2 = pos / a.n3, l3 = pos % a.n3;
    const CsdMod m = mt.m[jl];
    const int64_t* L = a.lhs + jl * a.sl[1] + k2 * a.sl[2] + l3 * a.sl[3];
    const int64_t* R = a.rhs + jl * a.sr[1] + k2 * a.sr[2] + l3 * a.sr[3];
    int64_t* D = a.dst + jl * a.sd[1] + k2 * a.sd[2] + l3 * a.sd[3];
    for (int k = 0; k < a.K; k++) {
        i128 s0 = 0, s1 = 0, s2 = 0;
        const int j_lo = k - (a.d2 - 1) > 0 ? k - (a.d2 - 1) : 0, j_hi = k < a.d1 - 1 ? k : a.d1 - 1;
        for (int j = j_lo; j <= j_hi; j++) {
            const int64_t lr = j * a.lrow[0] + b * a.lrow[2], rr = (k - j) * a.rrow[0] + b * a.rrow[2];
            const int64_t l0 = L[lr * a.sl[0]], l1 = L[(lr + a.lrow[1]) * a.sl[0]];
            const int64_t r0 = R[rr * a.sr[0]], r1 = R[(rr + a.rrow[1]) * a.sr[0]];
            s0 = gb_mac(s0, l0, r0);
            s1 = gb_mac(gb_mac(s1, l0, r1), l1, r0);
            s2 = gb_mac(s2, l1, r1);
        }
        const i128 sums[3] = {s0, s1, s2};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int row = c * a.K + k;
            const int64_t v = a.full[row] ? csd_to_mont64(m, csd_double_red128(m, sums[c])) : csd_red128(m, sums[c]);
            D[((int64_t)row * a.n + b) * a.sd[0]] = v;
        }
    }
}

// The same with every operand digit of the position loaded once into registers (d1, d2 <= CV_MAXD):
// each value is read from memory once instead of once per output digit it enters, and all loads
// are in flight together. The sums are formed in the same order, so the rows are identical.
#define CV_MAXD 12
#define CV_MIN_REGS 6
__global__ __launch_bounds__(256) void k_convolve_reduce_regs(CvArgs a, ModTable mt) {
    const int64_t chunks = (a.P + 255) / 256;
    int64_t blk = blockIdx.x;
    const int64_t c0 = blk % chunks;
    blk /= chunks;
    const int64_t jl = blk % a.r, b = blk / a.r;  // limb, pair
    const int64_t pos = c0 * 256 + threadIdx.x;
    if (pos >= a.P) return;
    const int64_t k2 = pos / a.n3, l3 = pos % a.n3;
    const CsdMod m = mt.m[jl];
    const int64_t* L = a.lhs + jl * a.sl[1] + k2 * a.sl[2] + l3 * a.sl[3];
    const int64_t* R = a.rhs + jl * a.sr[1] + k2 * a.sr[2] + l3 * a.sr[3];
    int64_t* D = a.dst + jl * a.sd[1] + k2 * a.sd[2] + l3 * a.sd[3];
    int64_t l0[CV_MAXD], l1[CV_MAXD], r0[CV_MAXD], r1[CV_MAXD];
#pragma unroll
    for (int j = 0; j < CV_MAXD; j++) {
        if (j < a.d1) {
            const int64_t lr = j * a.lrow[0] + b * a.lrow[2];
            l0[j] = L[lr * a.sl[0]];
            l1[j] = L[(lr + a.lrow[1]) * a.sl[0]];
        }
        if (j < a.d2) {
            const int64_t rr = j * a.rrow[0] + b * a.rrow[2];
            r0[j] = R[rr * a.sr[0]];
            r1[j] = R[(rr + a.rrow[1]) * a.sr[0]];
        }
    }
#pragma unroll
    for (int k = 0; k < 2 * CV_MAXD - 1; k++) {
        if (k < a.K) {
            i128 s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
            for (int j = 0; j < CV_MAXD; j++) {
                // j from max(0, k - d2 + 1) to min(k, d1 - 1), as the general kernel runs it
                if (j <= k && k - j < CV_MAXD && j < a.d1 && k - j < a.d2) {
                    s0 = gb_mac(s0, l0[j], r0[k - j]);
                    s1 = gb_mac(gb_mac(s1, l0[j], r1[k - j]), l1[j], r0[k - j]);
                    s2 = gb_mac(s2, l1[j], r1[k - j]);
                }
            }
            const i128 sums[3] = {s0, s1, s2};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int row = c * a.K + k;
                const int64_t v = a.full[row] ? csd_to_mont64(m, csd_double_red128(m, sums[c])) : csd_red128(m, sums[c]);
                D[((int64_t)row * a.n + b) * a.sd[0]] = v;
            }
        }
    }
}

extern "C" int csd_convolve_reduce(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const int64_t* lhs_rows,
                                   const csd_view* rhs, const int64_t* rhs_rows, size_t count, size_t d1, size_t d2,
                                   size_t K, const uint8_t* full, const csd_modulus* mods, size_t nmods) {
    CSD_TRY check_view(dst, "convolve_reduce dst");
    check_view(lhs, "convolve_reduce lhs");
    check_view(rhs, "convolve_reduce rhs");
    if (count == 0 || d1 == 0 || d2 == 0 || K == 0) throw CsdError("convolve_reduce: empty operands");
    if (K > d1 + d2 - 1 || 3 * K > 3 * (size_t)CSD_MAX_MODULI) throw CsdError("convolve_reduce: too many output digits");
    CvArgs a;
    memset(&a, 0, sizeof(a));
    a.n = (int64_t)count;
    a.r = dst->shape[1];
    a.n3 = dst->shape[3];
    a.P = dst->shape[2] * dst->shape[3];
    a.d1 = (int32_t)d1;
    a.d2 = (int32_t)d2;
    a.K = (int32_t)K;
    if (dst->shape[0] != (int64_t)(3 * K * count)) throw CsdError("convolve_reduce: dst rows must be 3 K count");
    auto max_row = [&](const int64_t* rows, size_t d) { return (int64_t)(d - 1) * rows[0] + rows[1] + (int64_t)(count - 1) * rows[2]; };
    for (int q = 0; q < 3; q++) {
        if (lhs_rows[q] < 0 || rhs_rows[q] < 0) throw CsdError("convolve_reduce: negative row strides");
        a.lrow[q] = lhs_rows[q];
        a.rrow[q] = rhs_rows[q];
    }
    if (max_row(lhs_rows, d1) >= lhs->shape[0] || max_row(rhs_rows, d2) >= rhs->shape[0])
        throw CsdError("convolve_reduce: operand rows out of range");
    for (int q = 1; q < 4; q++)
        if (lhs->shape[q] != dst->shape[q] || rhs->shape[q] != dst->shape[q]) throw CsdError("convolve_reduce: shapes differ");
    for (int q = 0; q < 4; q++) {
        a.sl[q] = lhs->strides[q];
        a.sr[q] = rhs->strides[q];
        a.sd[q] = dst->strides[q];
    }
    a.lhs = (const int64_t*)lhs->ptr;
    a.rhs = (const int64_t*)rhs->ptr;
    a.dst = (int64_t*)dst->ptr;
    for (size_t q = 0; q < 3 * K; q++) a.full[q] = full[q] ? 1 : 0;
    ModTable t = make_mod_table(mods, nmods, a.r);
    if (a.P == 0) return 0;
    const int64_t blocks = a.n * a.r * ((a.P + 255) / 256);
    if (blocks > (int64_t)INT32_MAX) throw CsdError("convolve_reduce: grid too large");
    TimedScope ts(ctx, "convolve_reduce", 8.0 * (double)a.n * a.r * a.P * (2.0 * (d1 + d2) + 3.0 * K));
    // the register form pays where operands enter many output digits: measured faster with 8-10
    // digits a side (a bootstrap's products: 22.4 -> 17.4 ms), slower with 4 (16 pairs: 1.26 ->
    // 1.39 ms) and 1 (the tensor product of 64 pairs: 0.99 -> 1.86 ms), its 147 VGPRs halving the
    // waves per SIMD
    if (d1 >= CV_MIN_REGS && d2 >= CV_MIN_REGS && d1 <= CV_MAXD && d2 <= CV_MAXD)
        hipLaunchKernelGGL(k_convolve_reduce_regs, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a, t);
    else
        hipLaunchKernelGGL(k_convolve_reduce, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// relinearization key switch in one kernel: key_switch_acc's inner products (bfv.rs:498-528) for
// both key rows, relin's acc.add of c0, c1 (bfv.rs:722-726, CRT output) and reduce_wide
// (ring.rs:846-865)
//
// For product k, key row s: sum_i parts[k d + i] * key_i[s] (+ prod[s K + k] * one), an i128 with
// wrapping adds as the HAL forms it, reduced in registers by the kind the WideRingEls bound
// selects. Each digit row of `parts` is read once for both key rows.

struct KsArgs {
    int64_t K, d, r, P, n3;
    int64_t sp[4], sk[4], sa[4], so[4];  // parts [K d, r, ...], keys [2, r_key, ...], add [2K, r, ...], out
    int64_t orow_s, orow_k;              // out row of (s, k) = s orow_s + k orow_k
    const int64_t* parts;
    const int64_t* keys[CSD_MAX_MODULI];
    const int64_t* add;  // prod rows s K + k, or null
    const int64_t* one;  // [r] the multiplier of `add` (the WideRingEls one), or null
    int64_t* out;
    int full;
};

__global__ __launch_bounds__(256) void k_keyswitch_reduce(KsArgs a, ModTable mt) {
    // block -> (chunk, limb, product), the product fastest: the products of one position read the
    // same key entries, so they run back to back, and the XCD-aware order (the hardware deals
    // consecutive workgroups to the 8 XCDs in turn; tile t = (b mod 8) nblocks/8 + b/8) keeps them
    // on one XCD, whose L2 then serves the key to all but the first
    const int64_t chunks = (a.P + 255) / 256;
    const int64_t nblocks = chunks * a.r * a.K;
    int64_t blk = blockIdx.x;
    if (nblocks % 8 == 0) blk = (blk % 8) * (nblocks / 8) + blk / 8;
    const int64_t k = blk % a.K;
    blk /= a.K;
    const int64_t j = blk % a.r, c0 = blk / a.r;
    const int64_t pos = c0 * 256 + threadIdx.x;
    if (pos >= a.P) return;
    const int64_t p2 = pos / a.n3, p3 = pos % a.n3;
    const CsdMod m = mt.m[j];
    const int64_t* P = a.parts + k * a.d * a.sp[0] + j * a.sp[1] + p2 * a.sp[2] + p3 * a.sp[3];
    const int64_t koff = j * a.sk[1] + p2 * a.sk[2] + p3 * a.sk[3];
    i128 s0 = 0, s1 = 0;
    int i = 0;
    for (; i + 4 <= (int)a.d; i += 4) {  // four digits' loads in flight together
        int64_t x[4], k0[4], k1[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            x[t] = P[(i + t) * a.sp[0]];
            k0[t] = a.keys[i + t][koff];
            k1[t] = a.keys[i + t][koff + a.sk[0]];
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            s0 = gb_mac(s0, x[t], k0[t]);
            s1 = gb_mac(s1, x[t], k1[t]);
        }
    }
    for (; i < (int)a.d; i++) {
        const int64_t x = P[i * a.sp[0]];
        const int64_t* key = a.keys[i] + koff;
        s0 = gb_mac(s0, x, key[0]);
        s1 = gb_mac(s1, x, key[a.sk[0]]);
    }
    if (a.add) {
        const int64_t* A = a.add + j * a.sa[1] + p2 * a.sa[2] + p3 * a.sa[3];
        const int64_t one = a.one[j];
        s0 = gb_mac(s0, A[k * a.sa[0]], one);
        s1 = gb_mac(s1, A[(a.K + k) * a.sa[0]], one);
    }
    int64_t* O = a.out + j * a.so[1] + p2 * a.so[2] + p3 * a.so[3];
    const i128 sums[2] = {s0, s1};
#pragma unroll
    for (int s = 0; s < 2; s++)
        O[(s * a.orow_s + k * a.orow_k) * a.so[0]] = a.full ? csd_to_mont64(m, csd_double_red128(m, sums[s]))
                                                             : csd_red128(m, sums[s]);
}

extern "C" int csd_keyswitch_reduce(csd_ctx* ctx, const csd_view* out, const int64_t* out_rows, const csd_view* parts,
                                    size_t K, size_t d, const csd_view* keys, const csd_view* add, const csd_view* one,
                                    int full, const csd_modulus* mods, size_t nmods) {
    CSD_TRY check_view(out, "keyswitch_reduce out");
    check_view(parts, "keyswitch_reduce parts");
    if (K == 0 || d == 0 || d > CSD_MAX_MODULI) throw CsdError("keyswitch_reduce: 1 <= digits <= CSD_MAX_MODULI and K >= 1");
    KsArgs a;
    memset(&a, 0, sizeof(a));
    a.K = (int64_t)K;
    a.d = (int64_t)d;
    a.r = out->shape[1];
    a.n3 = out->shape[3];
    a.P = out->shape[2] * out->shape[3];
    if (parts->shape[0] != (int64_t)(K * d)) throw CsdError("keyswitch_reduce: parts rows must be K x digits");
    for (int q = 1; q < 4; q++)
        if (parts->shape[q] != out->shape[q]) throw CsdError("keyswitch_reduce: parts shape");
    if (out_rows[0] < 0 || out_rows[1] < 0 || out_rows[0] + (int64_t)(K - 1) * out_rows[1] >= out->shape[0])
        throw CsdError("keyswitch_reduce: output rows out of range");
    a.orow_s = out_rows[0];
    a.orow_k = out_rows[1];
    for (int q = 0; q < 4; q++) {
        a.sp[q] = parts->strides[q];
        a.so[q] = out->strides[q];
        a.sk[q] = keys[0].strides[q];
    }
    for (size_t i = 0; i < d; i++) {
        check_view(&keys[i], "keyswitch_reduce key");
        if (keys[i].shape[0] != 2 || keys[i].shape[1] < a.r || keys[i].shape[2] != out->shape[2] ||
            keys[i].shape[3] != out->shape[3])
            throw CsdError("keyswitch_reduce: a key is not [2, >= r, ., .]");
        for (int q = 0; q < 4; q++)
            if (keys[i].strides[q] != a.sk[q]) throw CsdError("keyswitch_reduce: keys must share their strides");
        a.keys[i] = (const int64_t*)keys[i].ptr;
    }
    if (add) {
        check_view(add, "keyswitch_reduce add");
        check_view(one, "keyswitch_reduce one");
        if (add->shape[0] != (int64_t)(2 * K) || add->shape[1] != a.r || add->shape[2] != out->shape[2] ||
            add->shape[3] != out->shape[3])
            throw CsdError("keyswitch_reduce: add must be [2 K, r, ., .]");
        if (one->shape[0] != 1 || one->shape[1] != a.r || one->shape[2] != 1 || one->shape[3] != 1 || one->strides[1] != 1)
            throw CsdError("keyswitch_reduce: one must be a contiguous [1, r, 1, 1] tensor");
        for (int q = 0; q < 4; q++) a.sa[q] = add->strides[q];
        a.add = (const int64_t*)add->ptr;
        a.one = (const int64_t*)one->ptr;
    }
    a.parts = (const int64_t*)parts->ptr;
    a.out = (int64_t*)out->ptr;
    a.full = full ? 1 : 0;
    ModTable t = make_mod_table(mods, nmods, a.r);
    if (a.P == 0) return 0;
    const int64_t blocks = a.K * a.r * ((a.P + 255) / 256);
    if (blocks > (int64_t)INT32_MAX) throw CsdError("keyswitch_reduce: grid too large");
    TimedScope ts(ctx, "keyswitch_reduce", 8.0 * (double)a.K * a.r * a.P * ((double)d + 2.0 + (add ? 2.0 : 0.0)));
    hipLaunchKernelGGL(k_keyswitch_reduce, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}
