// Internal declarations shared by the translation units of libcsdbfv_hal.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/csdbfv_hal.h"
#include "csd_arith.hpp"

struct CsdError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void csd_set_error(const std::string& s);

#define CSD_TRY try {
#define CSD_CATCH                          \
    }                                      \
    catch (const std::exception& e) {      \
        csd_set_error(e.what());           \
        return 1;                          \
    }                                      \
    return 0;

#define HIPCHECK(x)                                                                                            \
    do {                                                                                                       \
        hipError_t e_ = (x);                                                                                   \
        if (e_ != hipSuccess) throw CsdError(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

// A context: one device, one in-order stream, and a size-bucketed cache of device blocks so
// that the temporaries of composed ring operations do not hit hipMalloc/hipFree (which
// synchronise the device) on every call.
struct csd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // host->device key streaming, created on first use
    std::map<void*, size_t> pinned;     // pinned host blocks handed out by csd_host_alloc
    std::mutex mu;
    std::map<size_t, std::vector<void*>> pool;
    std::map<void*, size_t> live;
    // opt-in per-kernel timing: HIP events recorded on `stream` around each timed launch
    bool timing = false;
    struct TimedLaunch {
        std::string name;
        hipEvent_t start, stop;
        double bytes;
    };
    std::vector<TimedLaunch> timed;
    // launch-form settings of this context (no reference counterpart; every setting computes
    // identical results): csd_ntt_force_generic, csd_ntt_set_row_shape, csd_zp_set_kernel
    std::atomic<int> ntt_mode{0};
    std::atomic<int> row_lines{0};
    std::atomic<int> row_threads{0};
    std::atomic<int> zp_kernel{3};

    void* alloc_bytes(size_t bytes);
    void free_bytes(void* p);
    size_t alloc_size(void* p);
    ~csd_ctx();
};

// Records start/stop events around the launches in its scope when the context's timing is on.
// `bytes` is the launch's algorithmic HBM traffic (read + write of each operand once).
struct TimedScope {
    csd_ctx* ctx;
    size_t idx;
    TimedScope(csd_ctx* c, const char* name, double bytes);
    ~TimedScope();
};

// dst[e] = src[perm[e]] along one dimension (hal/cpu.rs:481-485)
struct csd_perm_plan {
    csd_ctx* ctx;
    uint32_t* dperm;
    size_t len;
    std::vector<uint32_t> host;
};

// Row addressing of the batched Galois key switch: element e = h grp + l (0 <= l < grp) reads c0
// row h c0_hi + l c0_lo and writes output row s at h o_hi + l o_lo + s o_s (rows of r limb planes).
// The plain batch is grp = batch, c0_lo = 1, o_lo = 2, o_s = 1; a digit-major dBFV batch of n
// ciphertexts is grp = n, c0_hi = o_hi = 2 n, c0_lo = o_lo = 1, o_s = n.
struct GksRows {
    int64_t grp, c0_hi, c0_lo, o_hi, o_lo, o_s;
};

// decode_pb_scaled on the device (csd_dbfv.hip): src [count, r, n2, n3] period basis over the
// primes (r <= 4, their moduli), dst i128 [count, 1, n2, n3] = round(2^log2_t x / Q) balanced
void csd_decode_scaled_launch(csd_ctx* ctx, const csd_view* src, const int64_t* primes, const csd_modulus* mods, size_t r,
                              size_t log2_t, const csd_view* dst);
// encode_pb_unscaled of i128 values (log2_b = 0) or of their balanced base-2^log2_b digits
// (max_digits planes) over R moduli; returns the number of digit planes balanced_digits_vec keeps
int64_t csd_encode_digits_launch(csd_ctx* ctx, const csd_view* v, size_t log2_b, size_t max_digits, const csd_view* dst,
                                 const csd_modulus* mods, size_t R);

// key: row 2i + s of key digit i at key + (2i + s) krow (a restricted key keeps its row stride)
void csd_galois_keyswitch_launch(csd_ctx* ctx, int64_t batch, int64_t d, int64_t r, int64_t n2, int64_t n3, const int64_t* parts,
                                 const int64_t* c0, const int64_t* key, int64_t krow, const uint32_t* perm2,
                                 const uint32_t* perm3, int64_t* out, const GksRows& rows, const csd_modulus* mods, int full0,
                                 int full1);

// GadgetBConvCache::bconv / BConvCache::bconv as one kernel over B source elements (see
// csd_fused.hip), for digits whose limbs fit one i128 chunk; smods_dev holds the source moduli,
// or is null when they are the target moduli. matT is csd_bconv_transpose of mat. With pre, post_s
// and post_v (device, one int64 per source / target limb; one digit only) it runs the rescale form:
// source limbs are Montgomery-multiplied by pre first, and each dst row becomes
// mont(mont(dst, post_s) - converted, post_v), dst read in place. With keep (a view of dst's shape)
// and post_v alone it runs the mod-switch form: each dst row becomes mont(keep - converted, post_v).
void csd_bconv_fused_launch(csd_ctx* ctx, const csd_view& dst, const csd_view& src, int64_t d, const int64_t* dopi,
                            const int64_t* mat, const int64_t* matT, int64_t rdp, const int64_t* dmp,
                            const std::vector<size_t>& bounds, size_t shift, const csd_modulus* dmods,
                            const CsdMod* smods_dev, const char* timer, const int64_t* pre = nullptr,
                            const int64_t* post_s = nullptr, const int64_t* post_v = nullptr,
                            const csd_view* keep = nullptr);
std::vector<int64_t> csd_bconv_transpose(const std::vector<int64_t>& mat, int64_t d, int64_t rd, int64_t r, int64_t* rdp);
// The BEHZ conversion (one digit over the whole source base) with its matrix product on the int8
// matrix cores (csd_fused.hip); residues identical to csd_bconv_fused_launch's. Tables from
// csd_bconv_mfma_tables: mat [rd + 1][r] (row rd the correction weights), the target moduli and
// -D 2^64 mod p_j; false when r > 32, rd > 31 or a weight is outside eight balanced digits.
bool csd_bconv_mfma_tables(const std::vector<int64_t>& mat, int64_t rd, int64_t r, const std::vector<CsdMod>& dmods,
                           const std::vector<int64_t>& dmp, std::vector<int8_t>& atab, std::vector<int64_t>& rows);
void csd_bconv_mfma_launch(csd_ctx* ctx, const csd_view& dst, const csd_view& src, const int64_t* dopi, const int8_t* atab,
                           const int64_t* rows, size_t shift, const CsdMod* smods_dev, const char* timer,
                           const int64_t* pre = nullptr, const int64_t* post_s = nullptr, const int64_t* post_v = nullptr);

// launch shape of the work-list NTT kernels (0 = the built-in choice)
struct CsdSegTune {
    int rows = 0;          // dim-3: lines per workgroup
    int threads = 0;       // dim-3: workgroup size
    int col_threads = 256; // dim-2: columns per workgroup
};

// NTT along dim 2 or 3 of `nslices` [d2, d3] matrices (row stride st2) at seg_base[s], each
// transformed with plans[seg_limb[s]]; `mul` (optional, [1, nplans, d2|1, d3|1]) is multiplied in
// Montgomery form after the transform, in any direction.
void csd_ntt_segments(csd_ctx* ctx, int64_t* const* seg_base, const int32_t* seg_limb, int64_t nslices, int64_t d2, int64_t d3,
                      int64_t st2, csd_ntt_plan* const* plans, size_t nplans, int dim, int dir, const csd_view* mul,
                      const csd_modulus* mods, const char* timer, const CsdSegTune& tune);

// the persistent form of the batch plan's transform chain (see csd_ntt.hip)
int csd_batch_chain_occupancy(int rows);
bool csd_batch_chain_supported(csd_ntt_plan* const* plp, csd_ntt_plan* const* pk, size_t R);
void csd_batch_chain_launch(csd_ctx* ctx, int64_t* const* seg_base, const int32_t* seg_limb, int64_t nslices,
                            csd_ntt_plan* const* plp, csd_ntt_plan* const* pk, size_t R, const int64_t* klp, const int64_t* kk,
                            const int64_t* iklp, const int64_t* ikk, const int64_t* pt, const csd_modulus* mods, int rows,
                            int grid, double bytes);

void csd_mod_switch_one_launch(csd_ctx* ctx, int64_t* const* in, int64_t* const* out, const int32_t* level, const void* levels,
                               int64_t nelems, int64_t npos, const csd_modulus* mods, size_t nmods, double bytes);
size_t csd_mod_switch_level_bytes();
void csd_mod_switch_fill_level(void* dst, int r, const int64_t* M, const int64_t* qm, const int64_t* inv, int64_t C, int lq,
                               int64_t qopi);

// out[l][a][b] = kern[l][a - b] over the group C_o0 x ... (row-major mixed-radix indices)
void csd_circulant_launch(csd_ctx* ctx, const int64_t* kern, int64_t* out, int64_t r, const std::vector<size_t>& orders);

struct Shape4 {
    int64_t d[4];
};
struct Str4 {
    int64_t s[4];
};

struct ModTable {
    CsdMod m[CSD_MAX_MODULI];
};

Shape4 shape_of(const csd_view* v);
int64_t numel(const csd_view* v);
void check_view(const csd_view* v, const char* what);
Str4 bcast_strides(const csd_view* src, const csd_view* dst, const char* what);
Str4 own_strides(const csd_view* v);
ModTable make_mod_table(const csd_modulus* mods, size_t nmods, int64_t d1);
// Raises kernel `fn`'s dynamic-LDS limit to `bytes` on the current device, once per (kernel,
// device): the attribute is per device, and this keeps first calls from several threads safe.
void csd_allow_lds(const void* fn, int bytes);
void csd_zero_fill(csd_ctx* ctx, void* p, size_t bytes);
// Resolves the address of this device's copy of the dense maps' scratch array (csd_zp.hip); call
// once per device with that device current, before csd_zp_apply_dim_noalloc runs on it.
void csd_zp_scratch_setup(int device);

// host-side modular helpers (plan and constant construction)
int64_t csd_host_mulmod(int64_t a, int64_t b, int64_t q);
int64_t csd_host_powmod(int64_t b, uint64_t e, int64_t q);
CsdBarrett csd_host_barrett(int64_t value, int64_t n);
int64_t csd_host_root_of_unity(int64_t q, size_t n);
