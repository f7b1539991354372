// The dBFV digit reduction (cyclosubfield-dbfv/src/dbfv.rs:873-1025) fused into one pass over
// the coefficients per digit.
//
// The reference's split_overflow is a sequence of HAL calls over the whole batch:
//   mul_mont64 (lifts of x R (Q/p)^-1), matmul_i64 (Sum_p lift_p (Q/p) mod R in limbs, and the
//   gamma-scaled estimate), bconv_correct_pow2{,_wide} (the overflow), then a wide accumulation
//   (overflow x mont^2(-b Delta 2^(64k)) + x x mont(1)) and reduce_wide;
// dbfv_reduce adds Delta x carry to each digit first (another wide accumulation and reduction).
// Every one of these steps is local to a coefficient position (all r primes of it), so one
// thread per coefficient runs the whole chain in registers. x is read once and x_lo, the
// overflow written once, against about ten passes over HBM for the call sequence. Each step
// uses the same integer formulas as the kernel of its HAL call (csd_mont_mul, wrapping i128 sums,
// the pow2 correction, csd_double_red128 / csd_red128), so the results are bit-identical.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/csdbfv_ring.h"
#include "csd_internal.hpp"

namespace {

struct SoArgs {
    int64_t count, r, n2, n3;
    int64_t sx[3];   // strides of x (dims 0..2; dim 3 contiguous)
    int64_t so[3];   // strides of the overflow
    int64_t sc[3];   // strides of the carry-in (when present)
    int64_t si[3];   // strides of the separate input (when present)
    int64_t* x;
    const int64_t* xin;  // the digit's input when it is not x itself (read-only), or null
    int zero_in;         // the digit's input is zero (digits beyond the ciphertext's)
    int64_t* ov;
    const int64_t* cin;  // carry of the digit below, [count, limbs, n2, n3], or null
    const int64_t* k;    // constant table, 3 + 3 limbs rows of r (row stride r)
    i128 f0, f1;
    int limbs, shift, log2_r;
    int full_split, full_carry;
};

// rows of the constant table
__device__ __forceinline__ int row_lift() { return 0; }                 // r (Q/p)^-1 mod p (plain)
__device__ __forceinline__ int row_value(int k) { return 1 + k; }       // limb k of (Q/p) mod R
__device__ __forceinline__ int row_corr(int limbs) { return 1 + limbs; }  // round(gamma / p)
__device__ __forceinline__ int row_nbd(int limbs, int k) { return 2 + limbs + k; }  // mont^2(-b Delta 2^64k)
__device__ __forceinline__ int row_one(int limbs) { return 2 + 2 * limbs; }           // mont(1)
__device__ __forceinline__ int row_delta(int limbs, int k) { return 3 + 2 * limbs + k; }  // mont^2(Delta 2^64k)

// One digit's step for one coefficient position: the carry step (when the digit has a carry
// in), split_overflow, x_lo written to X; (cl0, cl1) carry the overflow in and out. The r residues
// wait in the thread's own LDS column (slot p at p * stride) between the passes over the primes, so
// the loops keep only the running sums in registers.
struct SoStep {
    int r, limbs, shift, log2_r;
    i128 f0, f1;
};

__device__ __forceinline__ void so_digit(const SoStep& a, const ModTable& mt, const int64_t* __restrict__ K, int64_t* col,
                                         int stride, const int64_t* XI, int64_t sxi, bool zero_in, bool has_carry,
                                         bool full_split, bool full_carry, int64_t* X, int64_t sx1, int64_t& cl0, int64_t& cl1) {
    const int r = a.r, limbs = a.limbs;
    // the digit's residues, eight loads in flight at a time, parked in the LDS column
    if (!zero_in) {
        for (int p0 = 0; p0 < r; p0 += 8) {
            int64_t raw[8];
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (p0 + t < r) raw[t] = XI[(p0 + t) * sxi];
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (p0 + t < r) col[(p0 + t) * stride] = raw[t];
        }
    }
    // pass 1: the carry step (dbfv.rs:995-1003), then the lifts and the matmul rows
    // (dbfv.rs:909-917)
    u128 s0 = 0, s1 = 0, sg = 0;
    for (int p = 0; p < r; p++) {
        const CsdMod m = mt.m[p];
        int64_t xv = zero_in ? 0 : col[p * stride];
        if (has_carry) {
            // wide fma of each carry limb with mont^2(Delta 2^64k) (mul for the first, fma
            // after), add of the digit with mont(1), reduce_wide
            u128 acc = (u128)((i128)cl0 * (i128)K[row_delta(limbs, 0) * r + p]);
            if (limbs == 2) acc += (u128)((i128)cl1 * (i128)K[row_delta(limbs, 1) * r + p]);
            acc += (u128)((i128)xv * (i128)K[row_one(limbs) * r + p]);
            xv = full_carry ? csd_to_mont64(m, csd_double_red128(m, (i128)acc)) : csd_red128(m, (i128)acc);
        }
        col[p * stride] = xv;
        const int64_t lift = csd_mont_mul(m, xv, K[row_lift() * r + p]);
        s0 += (u128)((i128)K[row_value(0) * r + p] * (i128)lift);
        if (limbs == 2) s1 += (u128)((i128)K[row_value(1) * r + p] * (i128)lift);
        sg += (u128)((i128)K[row_corr(limbs) * r + p] * (i128)lift);
    }
    // the overflow (bconv_correct_pow2{,_wide}, dbfv.rs:919-938)
    const u128 half = (u128)1 << (a.shift - 1);
    const i128 quot = (i128)(sg + half) >> a.shift;
    int64_t ov0, ov1 = 0;
    if (limbs == 1) {
        const i128 v = (i128)((s0 - (u128)quot * (u128)a.f0) * (u128)a.f1);
        ov0 = (int64_t)csd_pow2_centre(v, a.log2_r);
    } else {
        const i128 v = (i128)(s0 + (s1 << 64));
        const i128 cc = csd_pow2_centre((i128)(((u128)v - (u128)quot * (u128)a.f0) * (u128)a.f1), a.log2_r);
        ov0 = (int64_t)cc;
        ov1 = (int64_t)((i128)((u128)cc - (u128)(i128)ov0) >> 64);
    }
    cl0 = ov0;
    cl1 = ov1;
    // pass 2: x_lo = reduce(overflow mont^2(-b Delta) + x mont(1)) (dbfv.rs:940-947)
    for (int p = 0; p < r; p++) {
        const CsdMod m = mt.m[p];
        const int64_t xv = col[p * stride];
        u128 acc = (u128)((i128)ov0 * (i128)K[row_nbd(limbs, 0) * r + p]);
        if (limbs == 2) acc += (u128)((i128)ov1 * (i128)K[row_nbd(limbs, 1) * r + p]);
        acc += (u128)((i128)xv * (i128)K[row_one(limbs) * r + p]);
        X[p * sx1] = full_split ? csd_to_mont64(m, csd_double_red128(m, (i128)acc)) : csd_red128(m, (i128)acc);
    }
}

// the constant table staged in LDS once per workgroup, after the columns (broadcast reads,
// faster here than per-prime scalar loads)
__device__ __forceinline__ const int64_t* so_stage_table(int64_t* xs, const int64_t* k, int r, int limbs) {
    int64_t* kl = xs + (int64_t)r * blockDim.x;
    const int rows = 3 + 3 * limbs;
    for (int q = threadIdx.x; q < rows * r; q += blockDim.x) kl[q] = k[q];
    __syncthreads();
    return kl;
}

// One thread per coefficient position: one digit's carry step and split.
__global__ __launch_bounds__(256) void k_split_overflow(SoArgs a, ModTable mt) {
    extern __shared__ int64_t xs[];
    const int64_t per = a.n2 * a.n3;
    const int64_t total = a.count * per;
    const int64_t* K = so_stage_table(xs, a.k, (int)a.r, a.limbs);
    // one coefficient per thread and no grid-stride loop: a loop with a per-lane exit would make
    // the prime loops below divergent
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < total) {
        const int64_t c = e / per, j = e - c * per;
        const int64_t j2 = j / a.n3, j3 = j - j2 * a.n3;
        int64_t* X = a.x + c * a.sx[0] + j2 * a.sx[2] + j3;
        const int64_t* XI = a.xin ? a.xin + c * a.si[0] + j2 * a.si[2] + j3 : X;
        const int64_t sxi = a.xin ? a.si[1] : a.sx[1];
        int64_t cl0 = 0, cl1 = 0;
        if (a.cin) {
            const int64_t* C = a.cin + c * a.sc[0] + j2 * a.sc[2] + j3;
            cl0 = C[0];
            if (a.limbs == 2) cl1 = C[a.sc[1]];
        }
        const SoStep st{(int)a.r, a.limbs, a.shift, a.log2_r, a.f0, a.f1};
        so_digit(st, mt, K, xs + threadIdx.x, (int)blockDim.x, XI, sxi, a.zero_in != 0, a.cin != nullptr, a.full_split != 0,
                 a.full_carry != 0, X, a.sx[1], cl0, cl1);
        int64_t* O = a.ov + c * a.so[0] + j2 * a.so[2] + j3;
        O[0] = cl0;
        if (a.limbs == 2) O[a.so[1]] = cl1;
    }
}

// The whole carry chain of dbfv_reduce (dbfv.rs:990-1023) in one launch: each coefficient
// position of one row of a digit is taken through every digit in turn, the carry of digit i held in
// registers for digit i + 1 (the chain is sequential along the digits only). Digit i of the result
// is rows [i per, (i + 1) per) of x, read from the same rows of src for i < d and zero beyond; the
// carry of the last digit is dropped, as the reference drops it.
//
// Two kernels: for large batches one lane per position runs so_digit digit after digit
// (k_dbfv_reduce_chain1); for small ones four lanes share a position (k_dbfv_reduce_chain4), lane
// quarter q owning primes q, q + 4, ...: a digit's residues stay in registers, the three i128 sums
// of the matmul rows are added across the quarters (wrapping adds, so the order does not change
// them) and every quarter forms the same overflow. With one ciphertext that gives four times the
// waves, which the per-digit latency (loads, then two passes over the primes) needs; a wave holds
// 16 positions. Both give the residues of so_digit.
struct RcArgs {
    int64_t per, n2, n3;
    int64_t sx[3], si[3];
    int64_t* x;
    const int64_t* src;
    const int64_t* k;
    int digits, d;
    uint32_t full_split, full_carry;  // bit i: digit i's reductions
    SoStep st;
};

#define RC_Q 4
#define RC_MAX_R 40  // primes of one chain (the registers of a digit: RC_MAXP per lane)
#define RC_MAXP (RC_MAX_R / RC_Q)
// positions (rows x coefficients) from which one lane per position has waves enough: 2^18 positions
// are 4096 waves, four per SIMD of the 256 CUs
#define RC_ONE_LANE_FROM ((int64_t)1 << 18)

__device__ __forceinline__ u128 rc_sum_quarters(u128 v) {
    uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
#pragma unroll
    for (int m = 16; m <= 32; m *= 2) {
        const uint64_t olo = __shfl_xor(lo, m), ohi = __shfl_xor(hi, m);
        const u128 t = (((u128)hi << 64) | lo) + (((u128)ohi << 64) | olo);
        lo = (uint64_t)t;
        hi = (uint64_t)(t >> 64);
    }
    return ((u128)hi << 64) | lo;
}

__global__ __launch_bounds__(256) void k_dbfv_reduce_chain1(RcArgs a, ModTable mt) {
    extern __shared__ int64_t xs[];
    const int64_t npos = a.n2 * a.n3;
    const int64_t total = a.per * npos;
    const int64_t* K = so_stage_table(xs, a.k, a.st.r, a.st.limbs);
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < total) {
        const int64_t c = e / npos, j = e - c * npos;
        const int64_t j2 = j / a.n3, j3 = j - j2 * a.n3;
        int64_t cl0 = 0, cl1 = 0;
        for (int i = 0; i < a.digits; i++) {
            const int64_t row = (int64_t)i * a.per + c;
            int64_t* X = a.x + row * a.sx[0] + j2 * a.sx[2] + j3;
            const bool zero_in = i >= a.d;
            const int64_t* XI = zero_in ? X : a.src + row * a.si[0] + j2 * a.si[2] + j3;
            so_digit(a.st, mt, K, xs + threadIdx.x, (int)blockDim.x, XI, a.si[1], zero_in, i > 0, (a.full_split >> i) & 1,
                     (a.full_carry >> i) & 1, X, a.sx[1], cl0, cl1);
        }
    }
}

__global__ __launch_bounds__(256) void k_dbfv_reduce_chain4(RcArgs a, ModTable mt) {
    extern __shared__ int64_t xs[];
    const int r = a.st.r, limbs = a.st.limbs;
    const int trows = 3 + 3 * limbs;
    int64_t* K = xs;                          // the constant table [trows][r]
    CsdMod* M = (CsdMod*)(xs + trows * r);    // the moduli [r]
    for (int q = threadIdx.x; q < trows * r; q += blockDim.x) K[q] = a.k[q];
    for (int p = threadIdx.x; p < r; p += blockDim.x) M[p] = mt.m[p];
    __syncthreads();
    const int64_t npos = a.n2 * a.n3;
    const int64_t total = a.per * npos;
    const int lane = threadIdx.x & 63;
    const int qd = lane >> 4;
    const int64_t e = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16 + (lane & 15);
    // lanes past the end work on the last position (they take part in the cross-lane sums) and
    // write nothing
    const bool act = e < total;
    const int64_t ec = act ? e : total - 1;
    const int64_t c = ec / npos, j = ec - c * npos;
    const int64_t j2 = j / a.n3, j3 = j - j2 * a.n3;
    const u128 half = (u128)1 << (a.st.shift - 1);
    int64_t cl0 = 0, cl1 = 0;
    // digit i's residues of this quarter (zero beyond the input's digits); digit i + 1's are
    // loaded while digit i is reduced (in place, src rows of a digit are only written by its own step)
    auto load_digit = [&](int i, int64_t (&v)[RC_MAXP]) {
        if (i < a.d) {
            const int64_t* XI = a.src + ((int64_t)i * a.per + c) * a.si[0] + j2 * a.si[2] + j3;
#pragma unroll
            for (int k = 0; k < RC_MAXP; k++) {
                const int p = qd + RC_Q * k;
                v[k] = p < r ? XI[p * a.si[1]] : 0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < RC_MAXP; k++) v[k] = 0;
        }
    };
    int64_t xn[RC_MAXP];
    load_digit(0, xn);
    for (int i = 0; i < a.digits; i++) {
        const int64_t row = (int64_t)i * a.per + c;
        int64_t* X = a.x + row * a.sx[0] + j2 * a.sx[2] + j3;
        const bool has_carry = i > 0;
        const bool full_carry = (a.full_carry >> i) & 1, full_split = (a.full_split >> i) & 1;
        int64_t xv[RC_MAXP];
#pragma unroll
        for (int k = 0; k < RC_MAXP; k++) xv[k] = xn[k];
        if (i + 1 < a.digits) load_digit(i + 1, xn);
        // pass 1: the carry step (dbfv.rs:995-1003), then the lifts and the matmul rows
        // (dbfv.rs:909-917), over this quarter's primes
        u128 s0 = 0, s1 = 0, sg = 0;
#pragma unroll
        for (int k = 0; k < RC_MAXP; k++) {
            const int p = qd + RC_Q * k;
            if (p < r) {
                const CsdMod m = M[p];
                int64_t v = xv[k];
                if (has_carry) {
                    u128 acc = (u128)((i128)cl0 * (i128)K[row_delta(limbs, 0) * r + p]);
                    if (limbs == 2) acc += (u128)((i128)cl1 * (i128)K[row_delta(limbs, 1) * r + p]);
                    acc += (u128)((i128)v * (i128)K[row_one(limbs) * r + p]);
                    v = full_carry ? csd_to_mont64(m, csd_double_red128(m, (i128)acc)) : csd_red128(m, (i128)acc);
                    xv[k] = v;
                }
                const int64_t lift = csd_mont_mul(m, v, K[row_lift() * r + p]);
                s0 += (u128)((i128)K[row_value(0) * r + p] * (i128)lift);
                if (limbs == 2) s1 += (u128)((i128)K[row_value(1) * r + p] * (i128)lift);
                sg += (u128)((i128)K[row_corr(limbs) * r + p] * (i128)lift);
            }
        }
        s0 = rc_sum_quarters(s0);
        if (limbs == 2) s1 = rc_sum_quarters(s1);
        sg = rc_sum_quarters(sg);
        // the overflow (bconv_correct_pow2{,_wide}, dbfv.rs:919-938), the same in every quarter
        const i128 quot = (i128)(sg + half) >> a.st.shift;
        int64_t ov0, ov1 = 0;
        if (limbs == 1) {
            const i128 v = (i128)((s0 - (u128)quot * (u128)a.st.f0) * (u128)a.st.f1);
            ov0 = (int64_t)csd_pow2_centre(v, a.st.log2_r);
        } else {
            const i128 v = (i128)(s0 + (s1 << 64));
            const i128 cc = csd_pow2_centre((i128)(((u128)v - (u128)quot * (u128)a.st.f0) * (u128)a.st.f1), a.st.log2_r);
            ov0 = (int64_t)cc;
            ov1 = (int64_t)((i128)((u128)cc - (u128)(i128)ov0) >> 64);
        }
        cl0 = ov0;
        cl1 = ov1;
        // pass 2: x_lo = reduce(overflow mont^2(-b Delta) + x mont(1)) (dbfv.rs:940-947)
#pragma unroll
        for (int k = 0; k < RC_MAXP; k++) {
            const int p = qd + RC_Q * k;
            if (p < r) {
                const CsdMod m = M[p];
                u128 acc = (u128)((i128)ov0 * (i128)K[row_nbd(limbs, 0) * r + p]);
                if (limbs == 2) acc += (u128)((i128)ov1 * (i128)K[row_nbd(limbs, 1) * r + p]);
                acc += (u128)((i128)xv[k] * (i128)K[row_one(limbs) * r + p]);
                const int64_t out = full_split ? csd_to_mont64(m, csd_double_red128(m, (i128)acc)) : csd_red128(m, (i128)acc);
                if (act) X[p * a.sx[1]] = out;
            }
        }
    }
}

}  // namespace

extern "C" int csd_dbfv_split_overflow(csd_ctx* ctx, const csd_view* x, const csd_view* x_in, int zero_in,
                                       const csd_view* overflow, const csd_view* carry_in,
                                       const csd_view* consts, const csd_modulus* mods, size_t nmods, int64_t f0_lo,
                                       int64_t f0_hi, int64_t f1_lo, int64_t f1_hi, size_t shift, size_t log2_r,
                                       int full_split, int full_carry) {
    CSD_TRY check_view(x, "dbfv_split_overflow x");
    check_view(overflow, "dbfv_split_overflow overflow");
    check_view(consts, "dbfv_split_overflow constants");
    SoArgs a;
    memset(&a, 0, sizeof(a));
    a.count = x->shape[0];
    a.r = x->shape[1];
    a.n2 = x->shape[2];
    a.n3 = x->shape[3];
    const int limbs = (int)overflow->shape[1];
    if (limbs != 1 && limbs != 2) throw CsdError("dbfv_split_overflow: the overflow has 1 or 2 limbs");
    if (log2_r < 1 || log2_r > (limbs == 1 ? 62u : 126u)) throw CsdError("dbfv_split_overflow: log2_r out of range for the limbs");
    if (shift < 1 || shift >= 127) throw CsdError("dbfv_split_overflow: shift out of range");
    if (a.r < 1 || a.r > CSD_MAX_MODULI || (size_t)a.r != nmods) throw CsdError("dbfv_split_overflow: one modulus per limb of x");
    if (overflow->shape[0] != a.count || overflow->shape[2] != a.n2 || overflow->shape[3] != a.n3)
        throw CsdError("dbfv_split_overflow: overflow shape");
    if (x->strides[3] != 1 || overflow->strides[3] != 1) throw CsdError("dbfv_split_overflow: dimension 3 must be contiguous");
    const int64_t rows = 3 + 3 * (int64_t)limbs;
    if (consts->shape[0] != 1 || consts->shape[1] != 1 || consts->shape[2] != rows || consts->shape[3] != a.r ||
        consts->strides[2] != a.r || consts->strides[3] != 1)
        throw CsdError("dbfv_split_overflow: constants must be a contiguous [1, 1, 3 + 3 limbs, r] table");
    if (carry_in) {
        check_view(carry_in, "dbfv_split_overflow carry");
        if (carry_in->shape[0] != a.count || carry_in->shape[1] != limbs || carry_in->shape[2] != a.n2 ||
            carry_in->shape[3] != a.n3 || carry_in->strides[3] != 1)
            throw CsdError("dbfv_split_overflow: carry shape");
        for (int d = 0; d < 3; d++) a.sc[d] = carry_in->strides[d];
        a.cin = (const int64_t*)carry_in->ptr;
    }
    if (x_in) {
        check_view(x_in, "dbfv_split_overflow input");
        for (int d = 0; d < 4; d++)
            if (x_in->shape[d] != x->shape[d]) throw CsdError("dbfv_split_overflow: input shape differs from x");
        if (x_in->strides[3] != 1) throw CsdError("dbfv_split_overflow: dimension 3 must be contiguous");
        for (int d = 0; d < 3; d++) a.si[d] = x_in->strides[d];
        a.xin = (const int64_t*)x_in->ptr;
    }
    a.zero_in = zero_in ? 1 : 0;
    for (int d = 0; d < 3; d++) {
        a.sx[d] = x->strides[d];
        a.so[d] = overflow->strides[d];
    }
    a.x = (int64_t*)x->ptr;
    a.ov = (int64_t*)overflow->ptr;
    a.k = (const int64_t*)consts->ptr;
    a.f0 = (i128)(((u128)(uint64_t)f0_hi << 64) | (uint64_t)f0_lo);
    a.f1 = (i128)(((u128)(uint64_t)f1_hi << 64) | (uint64_t)f1_lo);
    a.limbs = limbs;
    a.shift = (int)shift;
    a.log2_r = (int)log2_r;
    a.full_split = full_split ? 1 : 0;
    a.full_carry = full_carry ? 1 : 0;
    if (numel(x) == 0) return 0;
    ModTable t = make_mod_table(mods, nmods, a.r);
    TimedScope ts(ctx, "dbfv_split_overflow",
                  8.0 * (double)a.count * a.n2 * a.n3 * ((zero_in ? 1.0 : 2.0) * a.r + limbs + (carry_in ? limbs : 0)));
    const size_t lds = ((size_t)a.r * 256 + (size_t)(3 + 3 * a.limbs) * a.r) * sizeof(int64_t);  // <= 128 + 4.5 KiB
    csd_allow_lds((const void*)k_split_overflow, 160 * 1024);
    const int64_t total = a.count * a.n2 * a.n3;
    if ((total + 255) / 256 > (int64_t)INT32_MAX) throw CsdError("dbfv_split_overflow: too many coefficients for one launch");
    hipLaunchKernelGGL(k_split_overflow, dim3((unsigned)((total + 255) / 256)), dim3(256), lds, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}

// whether the memory spans (first to last element) of two int64 views with non-negative strides meet
static bool rc_spans_meet(const csd_view& a, const csd_view& b) {
    auto span = [](const csd_view& v, uintptr_t& lo, uintptr_t& hi) {
        int64_t last = 0;
        for (int q = 0; q < 4; q++) {
            if (v.strides[q] < 0) throw CsdError("dbfv_reduce_chain: negative strides");
            last += (v.shape[q] - 1) * v.strides[q];
        }
        lo = (uintptr_t)v.ptr;
        hi = lo + (uintptr_t)(last + 1) * 8;
    };
    uintptr_t alo, ahi, blo, bhi;
    span(a, alo, ahi);
    span(b, blo, bhi);
    return alo < bhi && blo < ahi;
}

extern "C" int csd_dbfv_reduce_chain(csd_ctx* ctx, const csd_view* x, const csd_view* src, size_t per, size_t digits,
                                     const csd_view* consts, const csd_modulus* mods, size_t nmods, size_t limbs,
                                     int64_t f0_lo, int64_t f0_hi, int64_t f1_lo, int64_t f1_hi, size_t shift, size_t log2_r,
                                     uint32_t full_split, uint32_t full_carry) {
    CSD_TRY check_view(x, "dbfv_reduce_chain x");
    check_view(src, "dbfv_reduce_chain src");
    check_view(consts, "dbfv_reduce_chain constants");
    RcArgs a;
    memset(&a, 0, sizeof(a));
    if (per < 1) throw CsdError("dbfv_reduce_chain: at least one row per digit");
    if (digits < 1 || digits > 32) throw CsdError("dbfv_reduce_chain: 1 to 32 digits");
    if (limbs != 1 && limbs != 2) throw CsdError("dbfv_reduce_chain: the carries have 1 or 2 limbs");
    if (log2_r < 1 || log2_r > (limbs == 1 ? 62u : 126u)) throw CsdError("dbfv_reduce_chain: log2_r out of range for the limbs");
    if (shift < 1 || shift >= 127) throw CsdError("dbfv_reduce_chain: shift out of range");
    const int64_t r = x->shape[1];
    if (r < 1 || r > RC_MAX_R || (size_t)r != nmods) throw CsdError("dbfv_reduce_chain: one modulus per limb of x, at most 40");
    if (x->shape[0] != (int64_t)(per * digits)) throw CsdError("dbfv_reduce_chain: x must hold digits x per rows");
    if (src->shape[0] % (int64_t)per != 0 || src->shape[0] > x->shape[0]) throw CsdError("dbfv_reduce_chain: src rows");
    for (int d = 1; d < 4; d++)
        if (src->shape[d] != x->shape[d]) throw CsdError("dbfv_reduce_chain: src shape differs from x");
    if (x->strides[3] != 1 || src->strides[3] != 1) throw CsdError("dbfv_reduce_chain: dimension 3 must be contiguous");
    const int64_t rows = 3 + 3 * (int64_t)limbs;
    if (consts->shape[0] != 1 || consts->shape[1] != 1 || consts->shape[2] != rows || consts->shape[3] != r ||
        consts->strides[2] != r || consts->strides[3] != 1)
        throw CsdError("dbfv_reduce_chain: constants must be a contiguous [1, 1, 3 + 3 limbs, r] table");
    // a thread reads its src rows and writes the same rows of x, so src may be x itself (same
    // rows and strides); any other sharing of memory is refused
    bool same = src->ptr == x->ptr;
    for (int d = 0; d < 4; d++) same = same && src->strides[d] == x->strides[d];
    if (numel(x) == 0) return 0;
    if (!same && numel(src) > 0 && rc_spans_meet(*x, *src)) throw CsdError("dbfv_reduce_chain: src overlaps x");
    a.per = (int64_t)per;
    a.n2 = x->shape[2];
    a.n3 = x->shape[3];
    for (int d = 0; d < 3; d++) {
        a.sx[d] = x->strides[d];
        a.si[d] = src->strides[d];
    }
    a.x = (int64_t*)x->ptr;
    a.src = (const int64_t*)src->ptr;
    a.k = (const int64_t*)consts->ptr;
    a.digits = (int)digits;
    a.d = (int)(src->shape[0] / (int64_t)per);
    a.full_split = full_split;
    a.full_carry = full_carry;
    a.st = SoStep{(int)r, (int)limbs, (int)shift, (int)log2_r, (i128)(((u128)(uint64_t)f0_hi << 64) | (uint64_t)f0_lo),
                  (i128)(((u128)(uint64_t)f1_hi << 64) | (uint64_t)f1_lo)};
    ModTable t = make_mod_table(mods, nmods, r);
    TimedScope ts(ctx, "dbfv_reduce_chain", 8.0 * (double)a.per * a.n2 * a.n3 * r * (double)(digits + a.d));
    const int64_t total = a.per * a.n2 * a.n3;
    if (total >= RC_ONE_LANE_FROM) {
        // 256 threads, one position each, its residues parked in its LDS column
        const size_t lds = ((size_t)r * 256 + (size_t)rows * r) * sizeof(int64_t);
        csd_allow_lds((const void*)k_dbfv_reduce_chain1, 160 * 1024);
        const int64_t blocks = (total + 255) / 256;
        if (blocks > (int64_t)INT32_MAX) throw CsdError("dbfv_reduce_chain: too many coefficients for one launch");
        hipLaunchKernelGGL(k_dbfv_reduce_chain1, dim3((unsigned)blocks), dim3(256), lds, ctx->stream, a, t);
    } else {
        // 256 threads: 64 positions of four lanes each
        const size_t lds = (size_t)rows * r * sizeof(int64_t) + (size_t)r * sizeof(CsdMod);
        const int64_t blocks = (total + 63) / 64;
        hipLaunchKernelGGL(k_dbfv_reduce_chain4, dim3((unsigned)blocks), dim3(256), lds, ctx->stream, a, t);
    }
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// Decoding and encoding of plaintexts on the device (the host steps of bootstrap_bfv_dbfv,
// bootstrap/mod.rs:378-395: decode_pb_scaled of both components after the trace, then
// encode_pb_unscaled of the first and of the balanced base-b digits of the second).
//
// decode_pb_scaled (bfv.rs:262-274) is round(t x / Q) mod t of the smallest lift x of every
// coefficient. With t = 2^e and y_j = x (Q/p_j)^-1 mod p_j, x = sum_j y_j (Q/p_j) - k Q, so
//   t x / Q = sum_j y_j 2^e / p_j  (mod t),   y_j 2^e = q_j p_j + r_j,
//   round(t x / Q) = sum_j q_j + round(sum_j r_j / p_j)  (mod t).
// r_j = x c_j mod p_j with c_j = (Q/p_j)^-1 2^e mod p_j (one Montgomery product of the
// montgomery-form residue), q_j = -r_j p_j^-1 mod 2^e (q_j p_j = y_j 2^e - r_j), and
// round(sum_j r_j / p_j) counts the k with 2 sum_j r_j (Q/p_j) >= (2k + 1) Q, compared exactly in
// four 64-bit words (2 sum r_j Q/p_j and (2k + 1) Q are integers of different parity, so no
// coefficient sits on a tie). The result is exact for every input, as the reference's big-integer
// decoding is.

#define DEC_MAX_R 4

struct DecArgs {
    int64_t count, n2, n3;
    int64_t s0, s1, s2;  // strides of the source (dims 0..2; dim 3 contiguous)
    const int64_t* src;
    i128* out;  // [count, 1, n2, n3], contiguous
    int r, e;
    int64_t c[DEC_MAX_R];
    u128 pinv[DEC_MAX_R];
    uint64_t M[DEC_MAX_R][4];  // Q / p_j, little-endian words
    uint64_t T[DEC_MAX_R][4];  // (2k + 1) Q for k < r
    CsdMod m[DEC_MAX_R];
};

__device__ __forceinline__ bool words_geq(const uint64_t* a, const uint64_t* b) {
    for (int w = 3; w >= 0; w--)
        if (a[w] != b[w]) return a[w] > b[w];
    return true;
}

__global__ __launch_bounds__(256) void k_decode_scaled(DecArgs a) {
    const int64_t plane = a.n2 * a.n3;
    const int64_t pos = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pos >= a.count * plane) return;
    const int64_t i = pos / plane, rest = pos - i * plane, k = rest / a.n3, l = rest - k * a.n3;
    const int64_t* x = a.src + i * a.s0 + k * a.s2 + l;
    u128 q = 0;
    uint64_t W[4] = {0, 0, 0, 0};
    for (int j = 0; j < a.r; j++) {
        const CsdMod& m = a.m[j];
        int64_t rj = csd_mont_mul(m, x[j * a.s1], a.c[j]);
        if (rj < 0) rj += m.N;
        if (rj >= m.N) rj -= m.N;
        q -= (u128)(uint64_t)rj * a.pinv[j];
        const uint64_t mult = 2 * (uint64_t)rj;
        u128 carry = 0;
        for (int w = 0; w < 4; w++) {
            u128 t = (u128)a.M[j][w] * mult + W[w] + carry;
            W[w] = (uint64_t)t;
            carry = t >> 64;
        }
    }
    int rounded = 0;
    for (int kk = 0; kk < a.r; kk++) rounded += words_geq(W, a.T[kk]) ? 1 : 0;
    q += (u128)rounded;
    const u128 mask = ((u128)1 << a.e) - 1;
    u128 v = q & mask;
    i128 out = (i128)v;
    if (v > ((u128)1 << (a.e - 1))) out -= (i128)((u128)1 << a.e);  // balanced: (-t/2, t/2]
    a.out[pos] = out;
}

// encode_pb_unscaled (bfv.rs:226-235) with the reference's representatives: mont_repr((c rem p) as
// i64, p) is c R reduced with the sign of c (Rust's remainder), |.| < p.
__device__ __forceinline__ int64_t encode_repr(const CsdMod& m, i128 c) {
    const i128 mag = c < 0 ? -c : c;
    int64_t z = csd_to_mont128(m, mag);
    if (z < 0) z += m.N;
    if (z >= m.N) z -= m.N;
    return c < 0 ? -z : z;
}

struct EncArgs {
    int64_t count, n2, n3, R, D;
    const i128* v;  // [count, 1, n2, n3], contiguous
    int64_t* dst;   // [D count, R, n2, n3], contiguous: plane i of element c at row i count + c
    int log2_b;     // 0: the values themselves (one plane)
    int* ndig;      // the largest digit count (balanced_digits_vec stops when all are zero)
};

__global__ __launch_bounds__(256) void k_encode_digits(EncArgs a, ModTable mt) {
    const int64_t plane = a.n2 * a.n3;
    const int64_t pos = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pos >= a.count * plane) return;
    const int64_t c = pos / plane, p = pos - c * plane;
    i128 cur = a.v[pos];
    if (a.log2_b == 0) {
        for (int64_t j = 0; j < a.R; j++) a.dst[(c * a.R + j) * plane + p] = encode_repr(mt.m[j], cur);
        return;
    }
    const int64_t b = (int64_t)1 << a.log2_b;
    int cnt = 0;
    for (int64_t i = 0; i < a.D; i++) {
        int64_t d = 0;
        if (cur != 0) {
            d = (int64_t)((u128)cur & (u128)(b - 1));  // cmod(cur, b) in (-b/2, b/2] (dbfv.rs:83-95)
            if (d > b / 2) d -= b;
            cur = (cur - d) >> a.log2_b;
            cnt = (int)i + 1;
        }
        int64_t* o = a.dst + ((i * a.count + c) * a.R) * plane + p;
        for (int64_t j = 0; j < a.R; j++) o[j * plane] = encode_repr(mt.m[j], (i128)d);
    }
    if (cnt) atomicMax(a.ndig, cnt);
}

// host multiword helpers (little-endian 64-bit words)
static void words_mul_small(std::vector<uint64_t>& w, uint64_t f) {
    u128 carry = 0;
    for (auto& x : w) {
        u128 t = (u128)x * f + carry;
        x = (uint64_t)t;
        carry = t >> 64;
    }
    if (carry) w.push_back((uint64_t)carry);
}

void csd_decode_scaled_launch(csd_ctx* ctx, const csd_view* src, const int64_t* primes, const csd_modulus* mods, size_t r,
                              size_t log2_t, const csd_view* dst) {
    check_view(src, "decode_scaled source");
    check_view(dst, "decode_scaled output");
    if (r < 1 || r > DEC_MAX_R) throw CsdError("decode_scaled: between 1 and 4 primes");
    if (log2_t < 1 || log2_t > 126) throw CsdError("decode_scaled: t = 2^e with 1 <= e <= 126");
    if (src->shape[1] != (int64_t)r || src->strides[3] != 1) throw CsdError("decode_scaled: source must be [count, r, n2, n3]");
    if (dst->shape[0] != src->shape[0] || dst->shape[1] != 1 || dst->shape[2] != src->shape[2] ||
        dst->shape[3] != src->shape[3] || dst->strides[3] != 1 || dst->strides[2] != dst->shape[3] ||
        dst->strides[0] != dst->shape[2] * dst->shape[3])
        throw CsdError("decode_scaled: output must be a contiguous i128 [count, 1, n2, n3] tensor");
    DecArgs a;
    memset(&a, 0, sizeof(a));
    a.count = src->shape[0];
    a.n2 = src->shape[2];
    a.n3 = src->shape[3];
    a.s0 = src->strides[0];
    a.s1 = src->strides[1];
    a.s2 = src->strides[2];
    a.src = (const int64_t*)src->ptr;
    a.out = (i128*)dst->ptr;
    a.r = (int)r;
    a.e = (int)log2_t;
    std::vector<uint64_t> q{1};
    for (size_t j = 0; j < r; j++) words_mul_small(q, (uint64_t)primes[j]);
    if (q.size() > 4) throw CsdError("decode_scaled: modulus too large");
    q.resize(4, 0);
    for (size_t j = 0; j < r; j++) {
        const int64_t p = primes[j];
        std::vector<uint64_t> mj{1};
        int64_t mj_mod_p = 1;
        for (size_t i = 0; i < r; i++)
            if (i != j) {
                words_mul_small(mj, (uint64_t)primes[i]);
                mj_mod_p = csd_host_mulmod(mj_mod_p, primes[i] % p, p);
            }
        mj.resize(4, 0);
        for (int w = 0; w < 4; w++) a.M[j][w] = mj[w];
        const int64_t kinv = csd_host_powmod(mj_mod_p, (uint64_t)(p - 2), p);
        a.c[j] = csd_host_mulmod(kinv, csd_host_powmod(2, log2_t, p), p);
        u128 inv = (u128)(uint64_t)p;  // p^-1 mod 2^128 by Newton's iteration
        for (int it = 0; it < 7; it++) inv *= (u128)2 - (u128)(uint64_t)p * inv;
        a.pinv[j] = inv;
        a.m[j] = CsdMod{mods[j].N, mods[j].neg_inv_N, mods[j].r2, mods[j].r3};
    }
    for (size_t k = 0; k < r; k++) {
        std::vector<uint64_t> t = q;
        words_mul_small(t, 2 * k + 1);
        if (t.size() > 4 && t[4]) throw CsdError("decode_scaled: modulus too large");
        for (int w = 0; w < 4; w++) a.T[k][w] = t[w];
    }
    const int64_t total = a.count * a.n2 * a.n3;
    if (total == 0) return;
    if ((total + 255) / 256 > (int64_t)INT32_MAX) throw CsdError("decode_scaled: too many coefficients");
    TimedScope ts(ctx, "decode_scaled", (double)total * (8.0 * r + 16.0));
    hipLaunchKernelGGL(k_decode_scaled, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, a);
    HIPCHECK(hipGetLastError());
}

int64_t csd_encode_digits_launch(csd_ctx* ctx, const csd_view* v, size_t log2_b, size_t max_digits, const csd_view* dst,
                                 const csd_modulus* mods, size_t R) {
    check_view(v, "encode_digits values");
    check_view(dst, "encode_digits output");
    const int64_t count = v->shape[0], n2 = v->shape[2], n3 = v->shape[3];
    if (v->shape[1] != 1 || v->strides[3] != 1 || v->strides[2] != n3 || v->strides[0] != n2 * n3)
        throw CsdError("encode_digits: values must be a contiguous i128 [count, 1, n2, n3] tensor");
    const int64_t D = log2_b == 0 ? 1 : (int64_t)max_digits;
    if (log2_b > 62 || D < 1) throw CsdError("encode_digits: digit base 2^1 .. 2^62 and at least one digit");
    if (R < 1 || R > CSD_MAX_MODULI) throw CsdError("encode_digits: target base size");
    if (dst->shape[0] != D * count || dst->shape[1] != (int64_t)R || dst->shape[2] != n2 || dst->shape[3] != n3 ||
        dst->strides[3] != 1 || dst->strides[2] != n3 || dst->strides[1] != n2 * n3 || dst->strides[0] != (int64_t)R * n2 * n3)
        throw CsdError("encode_digits: output must be a contiguous [planes * count, R, n2, n3] tensor");
    EncArgs a;
    memset(&a, 0, sizeof(a));
    a.count = count;
    a.n2 = n2;
    a.n3 = n3;
    a.R = (int64_t)R;
    a.D = D;
    a.v = (const i128*)v->ptr;
    a.dst = (int64_t*)dst->ptr;
    a.log2_b = (int)log2_b;
    const int64_t total = count * n2 * n3;
    if (total == 0) return 0;
    if ((total + 255) / 256 > (int64_t)INT32_MAX) throw CsdError("encode_digits: too many coefficients");
    int* flag = (int*)ctx->alloc_bytes(8);
    int host = 0;
    try {
        csd_zero_fill(ctx, flag, 8);
        a.ndig = flag;
        ModTable t = make_mod_table(mods, R, (int64_t)R);
        {
            TimedScope ts(ctx, "encode_digits", (double)total * (16.0 + 8.0 * (double)(D * R)));
            hipLaunchKernelGGL(k_encode_digits, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, a, t);
            HIPCHECK(hipGetLastError());
        }
        HIPCHECK(hipMemcpyAsync(&host, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        ctx->free_bytes(flag);
        throw;
    }
    ctx->free_bytes(flag);
    return log2_b == 0 ? 1 : host;
}

// ---------------------------------------------------------------------------------------------
// One combination of dBFV ciphertexts with scalars (the pattern of every step of the digit
// extraction circuit, digitex.rs:62-241): dbfv_allocate_accumulator, dbfv_acc_inner_product_scalar
// (each scalar in balanced base-b digits, dbfv.rs:354-373), dbfv_acc_add of further ciphertexts,
// dbfv_acc_add_constant and dbfv_finish_accumulator (dbfv.rs:143-406), as one pass. Every output
// digit k of component s is sum_(t, i + j = k) x_t[i] sd_t[j] + sum_a a[k] mont(1) (+ the constant
// times mont(1) on digit 0, component 0), an exact i128 sum held in registers; the reference
// forms the same sum in the rows of a WideRingEls, so the reduction (red128, or double_red128 and
// to_mont64 where the row's bound exceeds 2^63 min_p, decided on the host from the same bound
// bookkeeping) gives the same residues. Operands are digit-major batches over one base.
//
// The factors may also be elements of L' (dbfv_acc_inner_product_element_lprime, dbfv.rs:378-396,
// the baby-step sums of the homomorphic CRT transform, homcrt.rs:365-429): their digits are then
// [1, r, n2, 1] CRT-basis tensors, constant along dimension 3, and the table holds one value per
// (term, digit, limb, row of dimension 2).

#define CB_MAX_TERMS 16
#define CB_MAX_ADDS 4
#define CB_MAX_DIGITS 12

struct CombArgs {
    int64_t count, r, plane;  // ciphertexts, limbs, n2 * n3
    int nt, na, dout;
    const int64_t* x[CB_MAX_TERMS];
    int ld[CB_MAX_TERMS];
    const int64_t* a[CB_MAX_ADDS];
    int lda[CB_MAX_ADDS];
    const int64_t* sd;     // [nt][CB_MAX_DIGITS][r][ns2]: digit j of factor t per limb (and row of dim 2)
    int nsd[CB_MAX_TERMS];
    int64_t ns2, n3;       // ns2 = 1 (scalar factors) or n2 (L' factors)
    const int64_t* one;    // [r] mont(1)
    const int64_t* cst;    // [r] the constant's tensor, or null
    uint32_t full;         // bit 2 k + s: row kind of output digit k, component s
    int64_t* out;          // [2 dout count, r, plane]
};

__global__ __launch_bounds__(256) void k_dbfv_combine(CombArgs a, ModTable mt) {
    const int64_t pos = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y % a.count, j = blockIdx.y / a.count;
    if (pos >= a.plane) return;
    const int64_t limb = a.r * a.plane;  // one row of the batch
    const int64_t off = j * a.plane + pos;
    i128 acc[CB_MAX_DIGITS][2];
#pragma unroll
    for (int k = 0; k < CB_MAX_DIGITS; k++) acc[k][0] = acc[k][1] = 0;
    const int64_t k2 = a.ns2 > 1 ? pos / a.n3 : 0;
    const int64_t sdstep = a.r * a.ns2;  // from one digit of a factor to the next
    for (int t = 0; t < a.nt; t++) {
        const int64_t* xt = a.x[t];
        const int64_t* sdt = a.sd + ((int64_t)t * CB_MAX_DIGITS * a.r + j) * a.ns2 + k2;
        for (int i = 0; i < a.ld[t]; i++) {
            const int64_t x0 = xt[((2 * i) * a.count + c) * limb + off];
            const int64_t x1 = xt[((2 * i + 1) * a.count + c) * limb + off];
#pragma unroll
            for (int k = 0; k < CB_MAX_DIGITS; k++) {
                const int jj = k - i;
                if (k < a.dout && jj >= 0 && jj < a.nsd[t]) {
                    const int64_t s = sdt[(int64_t)jj * sdstep];
                    acc[k][0] = (i128)((u128)acc[k][0] + (u128)((i128)x0 * (i128)s));
                    acc[k][1] = (i128)((u128)acc[k][1] + (u128)((i128)x1 * (i128)s));
                }
            }
        }
    }
    const int64_t one = a.one[j];
    for (int q = 0; q < a.na; q++) {
        const int64_t* aq = a.a[q];
#pragma unroll
        for (int k = 0; k < CB_MAX_DIGITS; k++) {
            if (k < a.lda[q]) {
                const int64_t y0 = aq[((2 * k) * a.count + c) * limb + off];
                const int64_t y1 = aq[((2 * k + 1) * a.count + c) * limb + off];
                acc[k][0] = (i128)((u128)acc[k][0] + (u128)((i128)y0 * (i128)one));
                acc[k][1] = (i128)((u128)acc[k][1] + (u128)((i128)y1 * (i128)one));
            }
        }
    }
    if (a.cst) acc[0][0] = (i128)((u128)acc[0][0] + (u128)((i128)a.cst[j] * (i128)one));
    const CsdMod m = mt.m[j];
#pragma unroll
    for (int k = 0; k < CB_MAX_DIGITS; k++) {
        if (k < a.dout) {
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const i128 v = acc[k][s];
                const int64_t y = ((a.full >> (2 * k + s)) & 1u) ? csd_to_mont64(m, csd_double_red128(m, v)) : csd_red128(m, v);
                a.out[((2 * k + s) * a.count + c) * limb + off] = y;
            }
        }
    }
}

void csd_dbfv_combine_launch(csd_ctx* ctx, const csd_view* out, int64_t count, int dout, const csd_view* terms,
                             const int* term_digits, const int64_t* sd_dev, int64_t ns2, const int* term_sd, int nt,
                             const csd_view* adds, const int* add_digits, int na, const int64_t* one_dev,
                             const int64_t* cst_dev, uint32_t full, const csd_modulus* mods, size_t r) {
    check_view(out, "dbfv_combine output");
    if (nt < 0 || nt > CB_MAX_TERMS || na < 0 || na > CB_MAX_ADDS) throw CsdError("dbfv_combine: too many operands");
    if (dout < 1 || dout > CB_MAX_DIGITS) throw CsdError("dbfv_combine: output digits out of range");
    if (count < 1 || r < 1 || r > CSD_MAX_MODULI) throw CsdError("dbfv_combine: batch size or base");
    const int64_t n2 = out->shape[2], n3 = out->shape[3], plane = n2 * n3;
    auto contiguous = [&](const csd_view* v, int64_t rows) {
        return v->shape[0] == rows && v->shape[1] == (int64_t)r && v->shape[2] == n2 && v->shape[3] == n3 &&
               v->strides[3] == 1 && (n2 == 1 || v->strides[2] == n3) && v->strides[1] == plane &&
               v->strides[0] == (int64_t)r * plane;
    };
    if (!contiguous(out, 2 * (int64_t)dout * count)) throw CsdError("dbfv_combine: output must be a contiguous [2 dout count, r, n2, n3] batch");
    CombArgs a;
    memset(&a, 0, sizeof(a));
    a.count = count;
    a.r = (int64_t)r;
    a.plane = plane;
    a.nt = nt;
    a.na = na;
    a.dout = dout;
    for (int t = 0; t < nt; t++) {
        if (term_digits[t] < 1 || term_digits[t] > CB_MAX_DIGITS || term_sd[t] < 1 || term_sd[t] > CB_MAX_DIGITS)
            throw CsdError("dbfv_combine: term digits out of range");
        if (!contiguous(&terms[t], 2 * (int64_t)term_digits[t] * count)) throw CsdError("dbfv_combine: terms must be contiguous batches over the base");
        a.x[t] = (const int64_t*)terms[t].ptr;
        a.ld[t] = term_digits[t];
        a.nsd[t] = term_sd[t];
    }
    for (int q = 0; q < na; q++) {
        if (add_digits[q] < 1 || add_digits[q] > dout) throw CsdError("dbfv_combine: added digits out of range");
        if (!contiguous(&adds[q], 2 * (int64_t)add_digits[q] * count)) throw CsdError("dbfv_combine: added ciphertexts must be contiguous batches over the base");
        a.a[q] = (const int64_t*)adds[q].ptr;
        a.lda[q] = add_digits[q];
    }
    if (ns2 != 1 && ns2 != n2) throw CsdError("dbfv_combine: factor table rows");
    a.sd = sd_dev;
    a.ns2 = ns2;
    a.n3 = n3;
    a.one = one_dev;
    a.cst = cst_dev;
    a.full = full;
    a.out = (int64_t*)out->ptr;
    ModTable t = make_mod_table(mods, r, (int64_t)r);
    const int64_t ycount = count * (int64_t)r;
    if (ycount > 65535) throw CsdError("dbfv_combine: too many (ciphertext, limb) pairs for one launch");
    double reads = 0;
    for (int q = 0; q < nt; q++) reads += 2.0 * term_digits[q];
    for (int q = 0; q < na; q++) reads += 2.0 * add_digits[q];
    TimedScope ts(ctx, "dbfv_combine", 8.0 * (double)ycount * (double)plane * (reads + 2.0 * dout));
    hipLaunchKernelGGL(k_dbfv_combine, dim3((unsigned)((plane + 255) / 256), (unsigned)ycount), dim3(256), 0, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
}

extern "C" int csd_dbfv_combine(csd_ctx* ctx, const csd_view* out, size_t count, size_t dout, const csd_view* terms,
                                const int* term_digits, const csd_view* scalars, const int* term_sd, size_t nt,
                                const csd_view* adds, const int* add_digits, size_t na, const csd_view* one,
                                const csd_view* constant, uint32_t full, const csd_modulus* mods, size_t nmods) {
    CSD_TRY if (nt > CB_MAX_TERMS) throw CsdError("dbfv_combine: too many terms");
    if (!one) throw CsdError("dbfv_combine: mont(1) tensor missing");
    check_view(one, "dbfv_combine one");
    if (one->shape[0] * one->shape[1] * one->shape[2] * one->shape[3] != (int64_t)nmods || one->strides[1] != 1)
        throw CsdError("dbfv_combine: mont(1) must be [1, r, 1, 1], contiguous");
    const int64_t* cst = nullptr;
    if (constant) {
        check_view(constant, "dbfv_combine constant");
        if (constant->shape[0] * constant->shape[1] * constant->shape[2] * constant->shape[3] != (int64_t)nmods ||
            constant->strides[1] != 1)
            throw CsdError("dbfv_combine: the constant must be [1, r, 1, 1], contiguous");
        cst = (const int64_t*)constant->ptr;
    }
    const int64_t* sd = nullptr;
    int64_t ns2 = 1;
    if (nt) {
        if (!scalars) throw CsdError("dbfv_combine: factor digit table missing");
        check_view(scalars, "dbfv_combine factors");
        check_view(out, "dbfv_combine output");
        const int64_t r = (int64_t)nmods, n2 = out->shape[2];
        const csd_view& v = *scalars;
        const bool scalar_form = v.shape[0] == 1 && v.shape[1] == (int64_t)nt && v.shape[2] == CB_MAX_DIGITS && v.shape[3] == r &&
                                 v.strides[3] == 1 && v.strides[2] == r && v.strides[1] == CB_MAX_DIGITS * r;
        const bool lprime_form = v.shape[0] == (int64_t)nt && v.shape[1] == CB_MAX_DIGITS && v.shape[2] == r && v.shape[3] == n2 &&
                                 v.strides[3] == 1 && v.strides[2] == n2 && v.strides[1] == r * n2 &&
                                 v.strides[0] == CB_MAX_DIGITS * r * n2;
        if (!scalar_form && !lprime_form)
            throw CsdError("dbfv_combine: factor digits must be a contiguous [1, terms, 12, r] (scalars) or [terms, 12, r, n2] "
                           "(L' elements) table");
        sd = (const int64_t*)v.ptr;
        ns2 = scalar_form ? 1 : n2;
    }
    csd_dbfv_combine_launch(ctx, out, (int64_t)count, (int)dout, terms, term_digits, sd, ns2, term_sd, (int)nt, adds,
                            add_digits, (int)na, (const int64_t*)one->ptr, cst, full, mods, nmods);
    CSD_CATCH
}
