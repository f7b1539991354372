/*
 * csdbfv_ring.h — ring-level entry points of the MI355X backend: the RNS ring O_L/q of
 * cyclosubfield-dbfv's `ring::RingContext` (cyclosubfield-dbfv/src/ring.rs:594-1852) and the
 * scheme-level HAL sequences of `bfv` that the hot path serves, computed on the device with the
 * kernels of csdbfv_hal.h. Tensors follow the reference layout [count, r, deg_L', deg_K]
 * (ring.rs:5-12) in Montgomery form. RNS bases are arrays of indices into the context's primes.
 * Conventions (status codes, views, streams) as in csdbfv_hal.h.
 */
#ifndef CSDBFV_RING_H
#define CSDBFV_RING_H

#include "csdbfv_hal.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csd_ring csd_ring;

/* ---- field descriptors (field.rs:128-520, data field.rs:529-769); pure host functions ------- */
/* m, deg_L, deg_K, deg_L', |Gal(M|L)|, rou_order and the cyclic orders (L' factors first). */
int csd_field_info(const char* field_name, uint64_t* m, uint64_t* deg_l, uint64_t* deg_k, uint64_t* deg_lprime,
                   uint64_t* gal_ml_len, uint64_t* rou_order, size_t* orders, size_t orders_cap, size_t* n_orders,
                   size_t* lprime_factor_count);
/* FieldDescriptor::gal_to_shift (field.rs:327-336) */
int csd_field_gal_to_shift(const char* field_name, uint64_t g, int64_t* shifts, size_t cap, size_t* n);
/* required_galois_elements (bootstrap/homcrt.rs:336-349): the Galois elements whose keys the
 * homomorphic CRT transform needs (baby steps and one giant step per prime p | m_K), sorted,
 * and the cyclic transform lengths |J_p| */
int csd_field_homcrt_galois_elements(const char* field_name, uint64_t* out, size_t cap, size_t* n, uint64_t* jp_orders,
                                     size_t jp_cap, size_t* n_jp);
/* The descriptor data of FieldDescriptorData (field.rs:529-769): m, m', the generators of
 * Gal(M'|L') and Gal(M_K|K) in (Z/mZ)*, and the cyclic decomposition of Gal(L|Q) (generator,
 * order), the first n_lprime factors being those of L'. Every list holds at most cap entries. */
int csd_field_data(const char* field_name, uint64_t* m, uint64_t* m_prime, uint64_t* mp_gens, size_t* n_mp,
                   uint64_t* mk_gens, size_t* n_mk, uint64_t* cyc_gens, uint64_t* cyc_orders, size_t* n_cyc,
                   size_t* n_lprime, size_t cap);
/* Per prime p | m_K (FieldDescriptor::m_k_factors, field.rs:271, in that order): p, the
 * generator i_p_gen of the inertia group (PrimeInfo, field.rs:217-243) and the length
 * |J_p| = PrimeInfo::j_p_order (field.rs:123) of the cyclic factor of the homomorphic CRT
 * transform (CRTTransformFactor, bootstrap/homcrt.rs:29-40). */
int csd_field_homcrt_factors(const char* field_name, uint64_t* p, uint64_t* gen, uint64_t* len, size_t cap, size_t* n);
/* the prime sampler of RingContext::new (ring.rs:136-183, 620-650): initial base and the
 * extended pool, no device work. */
int csd_sample_primes(const char* field_name, size_t log2_q_lo, size_t log2_q_hi, size_t t_bits, int64_t* out,
                      size_t cap, size_t* n, size_t* initial_len);

/* ---- contexts ------------------------------------------------------------------------------- */
/* RingContext::new (ring.rs:620-650) */
int csd_ring_create(csd_ctx* ctx, const char* field_name, size_t log2_q_lo, size_t log2_q_hi, size_t t_bits,
                    csd_ring** out);
/* RingContext::new_raw (ring.rs:657-677): arbitrary primes, base conversions only */
int csd_ring_create_raw(csd_ctx* ctx, const char* field_name, const int64_t* primes, size_t n, csd_ring** out);
int csd_ring_destroy(csd_ring* ring);
int csd_ring_primes(csd_ring* ring, int64_t* out, size_t cap, size_t* n, size_t* initial_len);
/* precompute_rns_factor_transforms (ring.rs:1310-1316) */
int csd_ring_precompute_transforms(csd_ring* ring);
/* the NTT-domain CRT kernels of one prime, as RnsFactorKernels (ring.rs:559-568); lengths
 * deg_L' and deg_K each */
int csd_ring_crt_kernels(csd_ring* ring, size_t prime_index, int64_t* lprime, int64_t* k, int64_t* inv_lprime,
                         int64_t* inv_k);
/* extension_rns_base_for (ring.rs:716-731) */
int csd_ring_extension_base(csd_ring* ring, const size_t* base, size_t r, size_t* out, size_t cap, size_t* n);

/* ---- transforms and conversions --------------------------------------------------------------- */
/* pb_to_crt_tensor / crt_to_pb_tensor (ring.rs:1117-1161), in place, plain form (four HAL calls). */
int csd_ring_pb_to_crt(csd_ring* ring, const csd_view* data, const size_t* base, size_t r);
int csd_ring_crt_to_pb(csd_ring* ring, const csd_view* data, const size_t* base, size_t r);
/* 1: batches of at least 2 elements and 64 MiB are transformed as two halves on two streams, the
 * second half one pass behind, so that the dim-2 passes of one half overlap the dim-3 passes of
 * the other. Every element sees the same passes in the same order, so results are identical
 * either way. Default 0 (serial): with both pass kinds VALU-heavy the overlap measures 1-3 %
 * slower (no reference counterpart). */
int csd_ring_set_overlap(csd_ring* ring, int on);
/* bconv_tensors (ring.rs:1526-1535, BConvCache::bconv 1663-1701). The default form is one kernel
 * that leaves src unchanged (used when src and dst do not overlap, or are disjoint limb ranges of
 * one tensor); the HAL-sequence form destroys
 * src, like the reference, when preserve_src is 0, and works on a copy when it is 1. With
 * preserve_src = 1, src is unchanged whatever form the ring's setting selects. */
int csd_ring_bconv(csd_ring* ring, const csd_view* dst, const size_t* dst_base, size_t r_dst, const csd_view* src,
                   const size_t* src_base, size_t r_src, int preserve_src);
/* Lines 607-616 of rescale_to_base (bfv.rs:601-622) for a batch of count elements: mul_scalar by t
 * (mul_mont64 with s_src / s_ext, the [1, r, 1, 1] scalar tensor of t over the source and extension
 * bases), bconv of src into the extension, sub_i64 and mul_mont64 by inv (the inverse scalar
 * tensor of the source base's product over the extension):
 *   ext <- mont(mont(ext, s_ext) - bconv(mont(src, s_src)), inv).
 * src [count, r_src, ., .] and ext [count, r_ext, ., .] may be disjoint limb ranges of one tensor.
 * With the fused conversion form this is one kernel and src is left as it is; otherwise it is the
 * reference's HAL call sequence and src is destroyed. ext's residues are identical either way. */
int csd_ring_rescale_step(csd_ring* ring, const csd_view* ext, const size_t* ext_base, size_t r_ext,
                          const csd_view* src, const size_t* src_base, size_t r_src, const csd_view* s_src,
                          const csd_view* s_ext, const csd_view* inv);
/* 2 (default): csd_ring_bconv, csd_ring_rescale_step and the conversions inside the ring
 * sequences run as one kernel whose matrix product is on the int8 matrix cores where that measured
 * at least as fast as the vector ALUs (conversions from 9 or more source primes, rescale steps from
 * 17 or more), else as form 1; 3: the matrix-core kernel for every conversion whose tables fit (at
 * most 32 source and 31 target primes; others take form 1); 1: one fused kernel on the vector
 * ALUs; 0: the reference's HAL call sequence (ring.rs:1663-1701). Residues are identical in every
 * form (no reference counterpart). */
int csd_ring_set_bconv_form(csd_ring* ring, int form);
/* the current csd_ring_set_bconv_form setting (3 / 2 matrix cores, 1 fused, 0 HAL sequence) */
int csd_ring_get_bconv_form(const csd_ring* ring);
/* bconv_gadget_tensors (ring.rs:1548-1557, GadgetBConvCache::bconv 1809-1851); digits are
 * half-open position ranges [digit_bounds[2i], digit_bounds[2i+1]). src is [B, r, ., .] and dst
 * [B ndigits, r, ., .], element b's digit i in row b ndigits + i; the reference takes B = 1, and
 * each element gets the residues of its own call. The default form is one kernel that leaves src
 * unchanged; the HAL-sequence form destroys src, like the reference. */
int csd_ring_bconv_gadget(csd_ring* ring, const csd_view* dst, const csd_view* src, const size_t* base, size_t r,
                          const size_t* digit_bounds, size_t ndigits);
/* 1 (default): csd_ring_bconv_gadget runs as one fused kernel; 0: as the reference's HAL call
 * sequence (ring.rs:1809-1851), per element. Residues are identical (no reference counterpart). */
int csd_ring_set_gadget_form(csd_ring* ring, int form);
/* The digit convolution of dbfv_mul_convolve (dbfv.rs:1096-1127) and, with d1 = d2 = K = 1, the
 * tensor product of hom_mul_raw_acc (bfv.rs:667-683), each followed by reduce_wide (ring.rs:846-865),
 * as one kernel (no reference counterpart; it replaces 3 K inner products or four wide fma calls
 * plus the reduction, with identical residues and no i128 rows in HBM). Operand row of (digit i,
 * component s, pair b) is i rows[0] + s rows[1] + b rows[2]; dst row (c K + k) count + b holds
 * component c of output digit k of pair b, reduced by red128 (full[c K + k] = 0) or by
 * double_red128 and to_mont64 (full = 1), the kinds the wide rows' bounds select. dst must not
 * overlap the operands. */
int csd_convolve_reduce(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const int64_t* lhs_rows,
                        const csd_view* rhs, const int64_t* rhs_rows, size_t count, size_t d1, size_t d2, size_t K,
                        const uint8_t* full, const csd_modulus* mods, size_t nmods);
/* relin's key switch (bfv.rs:498-528, 699-741) for K products as one kernel (no reference
 * counterpart): for key row s and product k, sum_i parts[k d + i] keys[i][s] (keys[i] a [2, >= r,
 * ., .] CRT-basis key restricted to the first r limbs), plus add[s K + k] one when `add` is given
 * (the CRT-basis c0, c1 with the WideRingEls constant one, [1, r, 1, 1]), reduced as reduce_wide
 * would (full = 1: double_red128 and to_mont64; 0: red128), into out row s out_rows[0] +
 * k out_rows[1]. Residues are those of the inner products, the wide add and the reduction. */
int csd_keyswitch_reduce(csd_ctx* ctx, const csd_view* out, const int64_t* out_rows, const csd_view* parts, size_t K,
                         size_t d, const csd_view* keys, const csd_view* add, const csd_view* one, int full,
                         const csd_modulus* mods, size_t nmods);
/* mod_switch_down (bfv.rs:628-662): dst [count, new_len, ., .] from src [count, r, ., .] (period
 * basis); src is left unchanged */
int csd_ring_mod_switch_down(csd_ring* ring, const csd_view* dst, const csd_view* src, const size_t* base, size_t r,
                             size_t new_len);
/* apply_galois (ring.rs:1183-1222): dst = sigma_g(src), any basis */
int csd_ring_apply_galois(csd_ring* ring, const csd_view* dst, const csd_view* src, uint64_t g);
/* The key-switching inner products of bfv_hom_gal (bfv.rs:816-839), batched: for ciphertexts
 * c < batch with hoisted gadget digits parts[c*digits + i] (CRT basis, [batch*digits, r, ., .])
 * and first components c0[c] ([batch, r, ., .]), and rotations g = gs[t] with keys[t] the
 * [2*digits, r, ., .] tensor of key digit i, row s at index 2i+s (KeySwitchKey.keys, resident
 * in HBM), writes out[(t*batch + c)*2 + s] = row s of
 *   reduce_wide( sum_i sigma_g(parts[c,i]) * key_i + (s == 0 ? sigma_g(c0[c]) * mont(1) : 0) ).
 * All operands are bound-1 (almost reduced); all tensors contiguous. One fused kernel per
 * rotation; bit-identical to apply_galois + inner_product + add + reduce_wide. */
int csd_ring_galois_keyswitch(csd_ring* ring, const size_t* base, size_t r, size_t batch, size_t digits,
                              const csd_view* parts, const csd_view* c0, const uint64_t* gs, size_t ngal,
                              const csd_view* keys, const csd_view* out);
/* The same key switch over a batch laid out in rows (no reference counterpart; the layout of
 * hom_cyclic_transform_dbfv's baby steps, bootstrap/homcrt.rs:380-392, run for all digits and
 * ciphertexts at once): element e = h group + l (0 <= l < group) takes its gadget digits from
 * parts[e*digits + i] (contiguous), its c0 from row h c0_hi + l c0_lo of c0 (contiguous rows) and
 * writes row s of rotation t to row h o_hi + l o_lo + s o_s of outs[t] (contiguous rows). keys[t]
 * is a [>= 2*digits, r, ., .] view whose rows 2i + s may be those of a key over a longer base
 * (KeySwitchKey restricted with restrict_base, bfv.rs:825): any row stride, limbs as contiguous
 * planes. Bit-identical to csd_ring_galois_keyswitch element by element. */
int csd_ring_galois_keyswitch_rows(csd_ring* ring, const size_t* base, size_t r, size_t batch, size_t group,
                                   size_t digits, const csd_view* parts, const csd_view* c0, int64_t c0_hi, int64_t c0_lo,
                                   const uint64_t* gs, size_t ngal, const csd_view* keys, const csd_view* outs, int64_t o_hi,
                                   int64_t o_lo, int64_t o_s);
/* One combination of digit-major dBFV batches over one base as one kernel (no reference
 * counterpart: dbfv_allocate_accumulator, dbfv_acc_inner_product_scalar, dbfv_acc_add,
 * dbfv_acc_add_constant and dbfv_finish_accumulator of dbfv.rs:143-406 in sequence, the pattern of
 * every step of digitex.rs:62-241): out ([2 dout count, r, n2, n3]) digit k, component s =
 * reduce(sum_(t, i + j = k) terms[t] digit i * scalars[0, t, j, :] + sum_q adds[q] digit k * one
 * + [k = s = 0] constant * one). terms[t] has term_digits[t] digits, its scalar term_sd[t]
 * balanced digits (scalars: [1, nt, 12, r], mont form). With an [nt, 12, r, n2] table the factors
 * are elements of L' instead (dbfv_acc_inner_product_element_lprime, dbfv.rs:378-396): digit j of
 * factor t is the CRT-basis tensor scalars[t, j, :, :] broadcast along dimension 3. Bit 2 k + s of
 * `full` selects double_red128 + to_mont64 for that row (the accumulator's bound exceeds 2^63
 * min_p), else red128. At most 16 terms, 4 adds, 12 output digits. Bit-identical to the
 * accumulator sequence. */
int csd_dbfv_combine(csd_ctx* ctx, const csd_view* out, size_t count, size_t dout, const csd_view* terms,
                     const int* term_digits, const csd_view* scalars, const int* term_sd, size_t nt, const csd_view* adds,
                     const int* add_digits, size_t na, const csd_view* one, const csd_view* constant, uint32_t full,
                     const csd_modulus* mods, size_t nmods);
/* decode_pb_scaled (bfv.rs:262-274) on the device, exact: for the period-basis [count, r, n2, n3]
 * tensor src over `base` (r <= 4 primes, montgomery form, bound <= 2) writes round(2^log2_t x / Q)
 * balanced into (-t/2, t/2] to the contiguous i128 [count, 1, n2, n3] tensor dst, x the lift of
 * each coefficient (element_lifts, ring.rs:1465-1480); 1 <= log2_t <= 126. The host decoding of
 * bootstrap_bfv_dbfv (bootstrap/mod.rs:378-380), without the host round trip. */
int csd_ring_decode_scaled(csd_ring* ring, const csd_view* src, const size_t* base, size_t r, size_t log2_t,
                           const csd_view* dst);
/* encode_pb_unscaled (bfv.rs:226-235) over `base` of the i128 values ([count, 1, n2, n3],
 * contiguous, |v| < 2^100) with log2_b = 0 (dst [count, r, n2, n3]), or of their balanced
 * base-2^log2_b digit vectors (balanced_digits_vec, dbfv.rs:97-112: max_digits planes, dst
 * [max_digits * count, r, n2, n3], plane i of element c at row i * count + c) as dbfv_acc_fma_plain
 * encodes them (dbfv.rs:262-289); *planes receives the number of planes balanced_digits_vec keeps
 * (it stops once every coefficient's remaining value is zero). Representatives are those of
 * element_from_pb_coeffs (ring.rs:1449-1461). */
int csd_ring_encode_digits(csd_ring* ring, const csd_view* values, size_t log2_b, size_t max_digits, const csd_view* dst,
                           const size_t* base, size_t r, int64_t* planes);
/* The L' half of pb_to_crt_tensor along dim 2 (ring.rs:1118-1141: ntt_mul with the L' CRT
 * kernel, then intt) on a [count, r, deg_L', w] tensor; with w = 1 and montgomery-form
 * coefficients it is lprime_element_tensor (ring.rs:1227-1250). */
int csd_ring_lprime_transform(csd_ring* ring, const csd_view* data, const size_t* base, size_t r);
/* split_overflow (dbfv.rs:893-948) of the period-basis tensor x ([count, r, n2, n3], Montgomery
 * form, |x| <= x_bound p) in one pass over the coefficients: x becomes x_lo in place and overflow
 * ([count, limbs, n2, n3], limbs 1 for log2_r <= 62, else 2) receives x_hi, with
 * x = x_lo + (b Delta) x_hi mod Q. The digit is read from x_in (same shape) when it is given, or
 * taken as zero when zero_in is set (digits beyond a ciphertext's), and x_lo is written to x
 * either way. When carry_in ([count, limbs, n2, n3]) is given, the digit first
 * becomes reduce_wide(carry_in * mont^2(Delta 2^64k) + x * mont(1)), the carry step of dbfv_reduce
 * (dbfv.rs:995-1003). consts is the contiguous [1, 1, 3 + 3 limbs, r] table of
 * SplitOverflowData (dbfv.rs:741-759), row by row: r (Q/p)^-1 mod p (plain lifts); the limbs of
 * (Q/p) mod R; round(gamma / p); mont^2(-b Delta 2^64k) per limb; mont(1); mont^2(Delta 2^64k)
 * per limb. f0 = Q mod R and f1 = -Q^-1 mod R are i128 (lo, hi words), shift = log2_gamma.
 * full_split / full_carry select the double reduction of reduce_wide (ring.rs:843-865) as the
 * bound tracking of the wide accumulators decides it. Bit-identical to the HAL call sequence. No
 * counterpart in the reference's HAL (a fused form of its dbfv.rs sequence). */
int csd_dbfv_split_overflow(csd_ctx* ctx, const csd_view* x, const csd_view* x_in, int zero_in, const csd_view* overflow,
                            const csd_view* carry_in, const csd_view* consts, const csd_modulus* mods, size_t nmods,
                            int64_t f0_lo, int64_t f0_hi, int64_t f1_lo, int64_t f1_hi, size_t shift, size_t log2_r,
                            int full_split, int full_carry);
/* The carry chain of dbfv_reduce (dbfv.rs:990-1023) in one launch: digit i of the result x
 * ([digits per, r, n2, n3], digit i in rows [i per, (i + 1) per)) is the split of the input digit
 * (the same rows of src for i < src rows / per, zero beyond) after the carry step with the overflow
 * of digit i - 1; the last digit's overflow is dropped. Each (row, position) runs the digits in
 * turn with its carry in registers. consts, f0, f1, shift and log2_r as for
 * csd_dbfv_split_overflow; limbs is the carries' limb count; bit i of full_split / full_carry
 * selects digit i's double reductions; at most 32 digits and 40 primes. src may be x itself (same
 * view), else must not share memory with it. Bit-identical to csd_dbfv_split_overflow digit by digit. No counterpart in the
 * reference's HAL (a fused form of its dbfv.rs sequence). */
int csd_dbfv_reduce_chain(csd_ctx* ctx, const csd_view* x, const csd_view* src, size_t per, size_t digits,
                          const csd_view* consts, const csd_modulus* mods, size_t nmods, size_t limbs, int64_t f0_lo,
                          int64_t f0_hi, int64_t f1_lo, int64_t f1_hi, size_t shift, size_t log2_r, uint32_t full_split,
                          uint32_t full_carry);
/* The same key switch with the keys in pinned host memory instead of HBM, the fallback for key
 * sets larger than the device (the CPU build streams them from its NVMe KeyStore,
 * keystore.rs:50-127): host_keys[t] points to the contiguous [2*digits, r, ., .] key of rotation t
 * inside a block from csd_host_alloc. Keys are copied through two device staging buffers on a
 * copy stream, each copy overlapping the previous rotation's kernel. Bit-identical to
 * csd_ring_galois_keyswitch. */
int csd_ring_galois_keyswitch_streamed(csd_ring* ring, const size_t* base, size_t r, size_t batch, size_t digits,
                                       const csd_view* parts, const csd_view* c0, const uint64_t* gs, size_t ngal,
                                       const int64_t* const* host_keys, const csd_view* out);
/* synthetic inputs: dst[i, j, k, l] uniform in [0, p_j) from splitmix64 of the flat index */
int csd_fill_uniform(csd_ctx* ctx, const csd_view* dst, const csd_modulus* mods, size_t nmods, uint64_t seed);
/* reduce_wide_tensor (ring.rs:846-865) for rows of equal reduction kind: full = 1 takes
 * red_mont64_double + i64_to_mont64, full = 0 a single red_mont64 */
int csd_ring_reduce_wide(csd_ring* ring, const csd_view* dst, const csd_view* wide, const size_t* base, size_t r, int full);
/* the [1, r, 1, 1] constant tensor mont(scalar mod p) of ring.rs:876-888, written into dst */
int csd_ring_scalar_tensor(csd_ring* ring, const csd_view* dst, const size_t* base, size_t r, int64_t scalar);
/* scalar_tensor (ring.rs:876-888) for an i128 scalar (low, high words), e.g. t = 2^76 */
int csd_ring_scalar_tensor_wide(csd_ring* ring, const csd_view* dst, const size_t* base, size_t r, int64_t lo, int64_t hi);
/* inverse_scalar_tensor (ring.rs:1165-1177) of the product of the primes at `of` */
int csd_ring_inverse_product_tensor(csd_ring* ring, const csd_view* dst, const size_t* base, size_t r, const size_t* of,
                                    size_t nof);

/* The dense n x n matrix, per limb of `base`, of pb_to_crt_tensor (ring.rs:1118-1141; inverse =
 * 0) or crt_to_pb_tensor (ring.rs:1144-1161; inverse = 1) restricted to one dimension (dim 2: the
 * L' part, n = deg_L'; dim 3: the K part, n = deg_K): the group-circulant C[a][b] = k[a - b] of
 * the CRT kernel k of that dimension. out: contiguous [1, r, n, n]. */
int csd_ring_transform_matrix(csd_ring* ring, const size_t* base, size_t r, int dim, int inverse, const csd_view* out);

/* ---- batched uneven chain ------------------------------------------------------------------ */
/* The reference runs every ring operation on one RingEls with one RNS base (ring.rs:198-208), so
 * ciphertexts at different levels after dynamic scaling take one call sequence each. A batch plan
 * takes `count` ciphertexts at once: ciphertext c is the contiguous [2, limbs[c], deg_L', deg_K]
 * tensor at cts[c] over the base prefix base[0 .. limbs[c]), 2 <= limbs[c] <= r, and the plan runs
 * per ciphertext exactly
 *   pb_to_crt_tensor (ring.rs:1118-1141), hal mul_mont64 by pt[:, 0 .. limbs[c]) (cpu.rs mul_mont64),
 *   crt_to_pb_tensor (ring.rs:1144-1161), mod_switch_down by one limb (bfv.rs:628-662)
 * writing the [2, limbs[c] - 1, ., .] result to outs[c]; cts[c] is overwritten with the
 * crt_to_pb output. pt is a [1, >= max limbs, deg_L', deg_K] CRT-basis tensor. All slices of all
 * ciphertexts go through each phase in one work-list launch; the results are bit-identical to
 * the per-ciphertext calls. The plan keeps the pointers: cts, outs and pt must stay allocated
 * while it is used. */
typedef struct csd_batch_plan csd_batch_plan;
/* launch shapes: dim-3 lines per workgroup and workgroup size (0 = built-in choice), dim-2
 * columns per workgroup (64..256), NTT pass grouping (radix-2 stages g2 in 1..3, radix-3 stages
 * g3 in 1..2 per LDS pass; applies to the plans of the whole base); persistent = 1 runs the eight
 * transform passes as one persistent launch in which each workgroup owns whole slices (lines per
 * LDS round = rows, 1..4; grid workgroups, 0 = all that fit at once; field_36864 slices, default
 * grouping), 0 runs one launch per pass. */
typedef struct csd_batch_tuning {
    int rows, threads, col_threads, g2, g3, persistent, grid;
} csd_batch_tuning;
int csd_ring_batch_create(csd_ring* ring, const size_t* base, size_t r, size_t count, int64_t* const* cts,
                          int64_t* const* outs, const size_t* limbs, const csd_view* pt, csd_batch_plan** out);
/* runs the chain on the context stream; use_graph = 1 replays a hipGraph of the phases, captured
 * on the first such call (and again after a tuning change) */
int csd_batch_run(csd_batch_plan* plan, int use_graph);
/* times launch-shape candidates on the plan's own data (which it overwrites), keeps the fastest
 * and reports it with its chain time in ms */
int csd_batch_autotune(csd_batch_plan* plan, int reps, csd_batch_tuning* chosen, double* ms);
int csd_batch_set_tuning(csd_batch_plan* plan, const csd_batch_tuning* t);
int csd_batch_get_tuning(const csd_batch_plan* plan, csd_batch_tuning* t);
/* slices (element, limb) and elements in the work list; algorithmic HBM bytes of one chain */
int csd_batch_info(const csd_batch_plan* plan, int64_t* nslices, int64_t* nelems, double* bytes);
int csd_batch_free(csd_batch_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* CSDBFV_RING_H */
