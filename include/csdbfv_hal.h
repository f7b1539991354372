/*
 * csdbfv_hal.h — C ABI of the MI355X (gfx950) backend for the `hal::Backend` trait of
 * cyclosubfield-dbfv (cyclosubfield-dbfv/src/hal/mod.rs:108-350), plus the ring-level fused
 * entry points that the reference's `ring`/`bfv` layers compose from several trait calls.
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero status on failure; `csd_last_error()`
 *    then describes the failure (the reference panics on the same contract violations, e.g.
 *    hal/cpu.rs:537, 568, 959-964, 993-996).
 *  - Tensors live in device memory (HBM). A view is {ptr, shape[4], strides[4]} with strides in
 *    elements of the tensor's type (8 bytes for i64 tensors, 16 bytes for i128 tensors), and the
 *    last stride 1 (hal/cpu.rs:3-6). A source operand may have length 1 in any dimension where the
 *    destination does not; it is then broadcast (hal/mod.rs:86-90).
 *  - Moduli are passed per call as an array indexed by dimension 1 (hal/mod.rs:92-96); at most
 *    CSD_MAX_MODULI entries.
 *  - All work is enqueued in order on the context's HIP stream; `csd_hal_export` and
 *    `csd_ctx_sync` wait for it (the export is the only device->host read, as in the trait).
 */
#ifndef CSDBFV_HAL_H
#define CSDBFV_HAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSD_MAX_MODULI 64

typedef struct csd_ctx csd_ctx;
typedef struct csd_ntt_plan csd_ntt_plan;
typedef struct csd_perm_plan csd_perm_plan;

typedef struct {
    void* ptr;
    int64_t shape[4];
    int64_t strides[4];
} csd_view;

/* MontgomeryModulus (hal/mont.rs:4-10): N, -N^-1 mod 2^64, 2^128 mod N, 2^192 mod N. */
typedef struct {
    int64_t N;
    int64_t neg_inv_N;
    int64_t r2;
    int64_t r3;
} csd_modulus;

enum { CSD_DIM2 = 2, CSD_DIM3 = 3 };

const char* csd_last_error(void);

/* ---- context (no reference counterpart: the reference backend is a stateless CPU struct,
 *      hal/cpu.rs:489-491) ------------------------------------------------------------------- */
int csd_ctx_create(int device, csd_ctx** out);
int csd_ctx_destroy(csd_ctx* ctx);
int csd_ctx_sync(csd_ctx* ctx);
/* the hipStream_t all work of the context is enqueued on */
int csd_ctx_stream(csd_ctx* ctx, void** stream);
/* opt-in kernel timing with HIP events on the context's stream (no reference counterpart; the
 * reference's tracing spans, hal/cpu.rs:560-1086, play this role on the CPU). The report has one
 * line per kernel: "<name> <launches> <total ms> <total algorithmic bytes>", and resets. */
int csd_ctx_set_timing(csd_ctx* ctx, int on);
int csd_ctx_timing_report(csd_ctx* ctx, char* buf, size_t cap);
/* Calibration of the two ceilings the kernels are measured against (no reference counterpart):
 * kind 0: device-to-device copy of n bytes -> achieved GB/s (read + write);
 * kind 1: n independent chains of 64 Barrett multiplications -> G multiplications/s. */
int csd_calibrate(csd_ctx* ctx, int kind, size_t n, double* out);

/* ---- allocation and transfer --------------------------------------------------------------- */
/* hal_allocate_i64 / hal_allocate_i128 (hal/mod.rs:119-122): zero-initialised, row-major;
 * elem_bytes is 8 (i64) or 16 (i128). */
int csd_hal_allocate(csd_ctx* ctx, int elem_bytes, const int64_t shape[4], void** out);
/* The same allocation without the zero fill, for outputs whose every element the next call
 * writes (no counterpart in the reference, whose allocations are always zeroed). */
int csd_hal_allocate_uninit(csd_ctx* ctx, int elem_bytes, const int64_t shape[4], void** out);
int csd_hal_free(csd_ctx* ctx, void* ptr);
/* pinned host memory for host-resident tensors streamed to the device (no trait counterpart; the
 * CPU build's KeyStore::store / CpuTensor::offload, keystore.rs:50-127, cpu.rs:82-185, move key
 * tensors out of RAM the same way). Allocated page-locked from the start. */
int csd_host_alloc(csd_ctx* ctx, size_t bytes, void** out);
int csd_host_free(csd_ctx* ctx, void* ptr);
/* hal_import / hal_export (hal/mod.rs:124-125): host data in row-major order of the view. */
int csd_hal_import(csd_ctx* ctx, const csd_view* dst, const int64_t* src, size_t len);
int csd_hal_export(csd_ctx* ctx, const csd_view* src, int64_t* dst, size_t len);
/* i128 counterparts (no trait method; the reference tests read i128 tensors directly,
 * hal/cpu.rs:1264-1454): two int64 words (low, high) per entry. */
int csd_import_i128(csd_ctx* ctx, const csd_view* dst, const int64_t* src_words, size_t len);
int csd_export_i128(csd_ctx* ctx, const csd_view* src, int64_t* dst_words, size_t len);
/* Debug range check of kernel outputs: *out_of_range receives the number of entries of the int64
 * view with |x| > multiple * N, N the modulus of the entry's limb (dimension 1; one modulus per
 * limb). Synchronizes the context's stream. The host-side counterpart of the reference's bound
 * bookkeeping (RingEls.bound, ring.rs:33-47; assert_mul_mont_safe, ring.rs:955-960), which it
 * checks against the data: no counterpart in the reference's HAL. */
int csd_hal_range_check(csd_ctx* ctx, const csd_view* v, const csd_modulus* mods, size_t nmods, int64_t multiple,
                        uint64_t* out_of_range);

/* ---- copies and wrapping integer arithmetic ------------------------------------------------ */
int csd_hal_assign(csd_ctx* ctx, const csd_view* dst, const csd_view* src);        /* mod.rs:129 */
int csd_hal_assign_narrow(csd_ctx* ctx, const csd_view* dst, const csd_view* src); /* mod.rs:134 */
int csd_hal_assign_i128(csd_ctx* ctx, const csd_view* dst, const csd_view* src);   /* mod.rs:137 */
int csd_hal_add_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* src);       /* mod.rs:139 */
int csd_hal_sub_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* src);       /* mod.rs:140 */
int csd_hal_add_i128(csd_ctx* ctx, const csd_view* dst, const csd_view* src);      /* mod.rs:141 */
int csd_hal_fma_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs); /* mod.rs:144 */
/* dst = lhs + rhs, wrapping int64 (hal_assign then hal_add_i64, mod.rs:129, 139, in one pass; no
 * trait method of its own) */
int csd_hal_sum_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs);
int csd_hal_mul_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs); /* mod.rs:149 */
/* hal_inner_product_i64 (mod.rs:151-163): dst = sum_t lhs[t] * rhs[t], overwriting. */
int csd_hal_inner_product_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs, size_t terms);
/* hal_matmul_i64 (mod.rs:165-173): dst[i,j,k,l] = sum_u lhs[i,j,u,0] * rhs[i,u,k,l]. */
int csd_hal_matmul_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs);

/* ---- Montgomery arithmetic ------------------------------------------------------------------ */
int csd_hal_create_modulus(int64_t prime, csd_modulus* out); /* mod.rs:175, mont.rs:13-33 */
int csd_hal_mul_mont64(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_modulus* mods, size_t nmods); /* mod.rs:179 */
int csd_hal_add_red_mont64_double(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_modulus* mods, size_t nmods); /* mod.rs:185 */
int csd_hal_red_mont64_double(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_modulus* mods, size_t nmods); /* mod.rs:189 */
int csd_hal_add_red_mont64(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_modulus* mods, size_t nmods); /* mod.rs:195 */
int csd_hal_red_mont64(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_modulus* mods, size_t nmods); /* mod.rs:199 */
int csd_hal_i128_to_mont64(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_modulus* mods, size_t nmods); /* mod.rs:203 */
int csd_hal_i64_to_mont64(csd_ctx* ctx, const csd_view* dst, const csd_modulus* mods, size_t nmods); /* mod.rs:207 */

/* ---- NTTs ----------------------------------------------------------------------------------- */
/* hal_create_ntt_plan (mod.rs:209, ntt.rs:561-566): group NTT over C_n0 x ... x C_n(r-1). */
int csd_hal_create_ntt_plan(csd_ctx* ctx, const csd_modulus* modulus, const size_t* inner_shape, size_t n, csd_ntt_plan** out);
int csd_ntt_plan_free(csd_ntt_plan* plan);
/* modular multiplications one transform of the plan performs (dir 0: forward reduced, 1: forward
 * unreduced, 2: inverse), following its reduction schedule and twiddle-1 skips */
int csd_ntt_plan_mults(const csd_ntt_plan* plan, int dir, uint64_t* out);
/* LDS pass grouping of the register-blocked kernels (no reference counterpart; a tuning knob):
 * at most g2 radix-2 (1..3) or g3 radix-3 (1..2) stages per pass. The arithmetic, and so every
 * result, is the same for all groupings; default 3 and 2. */
int csd_ntt_plan_set_grouping(csd_ntt_plan* plan, int g2, int g3);
/* hal_ntt_mont64_dim / hal_intt_mont64_dim / hal_ntt_mul_mont64_dim (mod.rs:216-238);
 * plans[j] is the plan for dimension-1 index j; dim is CSD_DIM2 or CSD_DIM3. */
int csd_hal_ntt_mont64_dim(csd_ctx* ctx, const csd_view* dst, csd_ntt_plan* const* plans, size_t nplans, int dim);
int csd_hal_intt_mont64_dim(csd_ctx* ctx, const csd_view* dst, csd_ntt_plan* const* plans, size_t nplans, int dim);
int csd_hal_ntt_mul_mont64_dim(csd_ctx* ctx, const csd_view* dst, csd_ntt_plan* const* plans, size_t nplans, int dim,
                               const csd_view* mul, const csd_modulus* mods, size_t nmods);
/* Kernel selection for the NTT entry points (no reference counterpart): 0 (default) uses the
 * register-blocked kernels where a plan has one, with the compiled pass shapes of the
 * field_36864 / field_49152 K parts where the plans match them; 1 forces the one-stage-per-LDS-pass
 * kernel for every plan; 2 keeps the register-blocked kernels but reads every pass list from the
 * plan. All produce identical results; the switch exists to test and time each form. */
int csd_ntt_force_generic(csd_ctx* ctx, int mode); /* per context */
/* Launch shape of dimension-3 transforms of lines that fit LDS: lines per workgroup (1-8, capped
 * by the batch and by the 160 KiB of LDS) and threads per workgroup (a multiple of 64 up to 512);
 * 0 keeps the automatic choice. Per context; every shape computes identical results. No
 * counterpart in the reference (a tuning knob). */
int csd_ntt_set_row_shape(csd_ctx* ctx, int lines, int threads);

/* ---- base-conversion corrections -------------------------------------------------------------- */
int csd_hal_bconv_correct_mont64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs,
                                 const csd_modulus* mods, size_t nmods, size_t shift); /* mod.rs:263-270 */
int csd_hal_bconv_correct_pow2(csd_ctx* ctx, const csd_view* dst, const csd_view* src, int64_t factor0, int64_t factor1,
                               size_t shift, size_t log2_modulus); /* mod.rs:289-297 */
/* factor0/factor1 are i128 given as (low, high) words (mod.rs:329-337) */
int csd_hal_bconv_correct_pow2_wide(csd_ctx* ctx, const csd_view* dst, const csd_view* src, int64_t factor0_lo,
                                    int64_t factor0_hi, int64_t factor1_lo, int64_t factor1_hi, size_t shift,
                                    size_t log2_modulus);

/* ---- permutations (Galois automorphisms) -------------------------------------------------------- */
int csd_hal_create_permutation(csd_ctx* ctx, const size_t* inner_shape, const size_t* shifts, size_t n, csd_perm_plan** out); /* mod.rs:341 */
int csd_perm_plan_free(csd_perm_plan* plan);
int csd_hal_permute(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_perm_plan* plan, int dim); /* mod.rs:343-349 */

/* ---- exact dense maps over Z_p (no reference counterpart) ----------------------------------- */
/* The reference applies its Z_p-linear maps (the period<->CRT transforms, ring.rs:1118-1161, and
 * homcrt's cyclic transforms, bootstrap/homcrt.rs:191-214) through NTTs and Galois automorphisms.
 * These entry points apply them as dense matrices instead, for any square or rectangular map:
 *   dst[c, l, ..., i along dim, ...] = sum_k A_l[i][k] src[c, l, ..., k along dim, ...] mod p_l
 * with the result in [0, p_l), equal mod p_l to every other exact evaluation (the NTT path
 * included). Entries of A and src must lie in [-2p, 2p]; moduli below 2^57.
 * form 0: balanced 8-bit digits on int8 MFMA (v_mfma_i32_16x16x64_i8), int32 accumulation over
 *         K-chunks of at most 16320 terms (max_chunk = 0) or max_chunk (a multiple of 64),
 *         exact i128 recombination, Montgomery reduction;
 * form 1: the plain VALU kernel (64 x 64 -> 128-bit products). */
typedef struct csd_zp_matrix csd_zp_matrix;
/* a: [1, r, m, k], one m x k matrix per modulus; digits are split once here */
int csd_zp_matrix_create(csd_ctx* ctx, const csd_view* a, const csd_modulus* mods, size_t nmods, csd_zp_matrix** out);
int csd_zp_matrix_free(csd_zp_matrix* m);
int csd_zp_apply_dim(csd_ctx* ctx, const csd_zp_matrix* a, const csd_view* dst, const csd_view* src, int dim, int form,
                     size_t max_chunk);
/* which int8 MFMA kernel form 0 uses (per context): 3 (default) stages each K-step's digit planes
 * of a 64 x 64 tile in LDS for 8 waves with two 16 x 16 output blocks each, taking the A operands
 * from the generator when the matrix was declared group-circulant (csd_zp_matrix_set_circulant);
 * 4 is form 3 without the circulant path; 1 stages a 32 x 64 tile for 8 waves with one block
 * each; 2 loads them straight into LDS (global_load_lds) for a 64 x 64 tile of 4 waves with 2 x 2
 * blocks each; 0 has 4 waves load their operands from L2 directly, two blocks each. All compute
 * identical results. */
int csd_zp_set_kernel(csd_ctx* ctx, int form);
/* Declares m's matrices group-circulant over C_o0 x ... x C_o(f-1), mixed-radix (row-major)
 * indices, last order 64: A[a][b] = A[a - b][0], the difference taken in the group (the period->
 * CRT matrices of the K part, ring.rs:1118-1141, are such matrices of the CRT kernel). The
 * structure is checked on the device (an error, and m unchanged, if any entry differs); the
 * default MFMA kernel then builds each 64 x 64 tile's A operands from the first column's digit
 * planes instead of streaming the matrix's. Results are unchanged (no reference counterpart). */
int csd_zp_matrix_set_circulant(csd_zp_matrix* m, const size_t* orders, size_t norders);
/* csd_zp_apply_dim's int8-MFMA form (form 0), allocating nothing. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The source vectors' digit planes go to a
 * file-scope device array of CSD_ZP_SCRATCH_BYTES per GPU (held until the process ends) instead
 * of a library allocation; calls needing more (moduli x 8 x vectors rounded up to 64 x columns
 * rounded up to 64 bytes) are rejected with an error. All calls on one GPU share that array:
 * calls that overlap on one GPU (other streams or contexts of the device) race. Results are
 * those of csd_zp_apply_dim(form 0). */
#define CSD_ZP_SCRATCH_BYTES (64u << 20)
int csd_zp_apply_dim_noalloc(csd_ctx* ctx, const csd_zp_matrix* a, const csd_view* dst, const csd_view* src, int dim,
                             size_t max_chunk);

#ifdef __cplusplus
}
#endif

#endif /* CSDBFV_HAL_H */
