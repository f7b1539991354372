// MI355X (gfx950) backend for the hal::Backend trait of cyclosubfield-dbfv
// (cyclosubfield-dbfv/src/hal/mod.rs:108-350), behind the C ABI of include/csdbfv_hal.h.
//
// Design (plain form, one kernel per trait call):
//  * element-wise ops: one generic broadcast kernel; rows (i,j,k) of the destination are split
//    into coalesced chunks along the contiguous dimension 3, the modulus index j is uniform per
//    workgroup (scalar loads of the modulus table passed by value);
//  * NTTs: one workgroup stages a tile of transforms in LDS (dim 3: whole rows, dim 2: a band of
//    columns) and runs the plan's exact stage/reduction schedule there, computed on the host
//    from the same u128 bound recursion as hal/ntt.rs:301-537, so outputs are bit-identical to
//    the reference CpuBackend;
//  * contractions (fma/mul/inner product/matmul) accumulate in registers with i128 wrapping
//    arithmetic, which is order-independent, so any accumulation order is bit-exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/csdbfv_hal.h"
#include "csd_arith.hpp"
#include "csd_internal.hpp"

// ---------------------------------------------------------------------------------------------
// errors

static thread_local std::string g_err;

extern "C" const char* csd_last_error(void) { return g_err.c_str(); }

void csd_set_error(const std::string& s) { g_err = s; }

// ---------------------------------------------------------------------------------------------
// context and the caching allocator

csd_ctx::~csd_ctx() {}

void* csd_ctx::alloc_bytes(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    std::lock_guard<std::mutex> g(mu);
    // the smallest cached block of at least `bytes`, if it is at most 1.5 x as large: sequences
    // whose temporaries shrink step by step (bases losing limbs level by level) reuse blocks
    // instead of growing the cache with every size
    auto it = pool.lower_bound(bytes);
    while (it != pool.end() && it->second.empty()) ++it;
    void* p = nullptr;
    if (it != pool.end() && it->first <= bytes + bytes / 2) {
        bytes = it->first;
        p = it->second.back();
        it->second.pop_back();
    } else {
        // keep a reserve free for the runtime (kernel arguments, scratch): when a new block would
        // cut into it, the cached blocks go back to the device first
        size_t free_b = 0, total_b = 0;
        if (!pool.empty() && hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b < bytes + ((size_t)2 << 30)) {
            for (auto& kv : pool)
                for (void* q : kv.second) (void)hipFree(q);
            pool.clear();
        }
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            // release cached blocks and retry once
            for (auto& kv : pool)
                for (void* q : kv.second) (void)hipFree(q);
            pool.clear();
            e = hipMalloc(&p, bytes);
            if (e != hipSuccess) throw CsdError("device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    }
    live[p] = bytes;
    return p;
}

void csd_ctx::free_bytes(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    auto it = live.find(p);
    if (it == live.end()) throw CsdError("free of a pointer not allocated by this context");
    pool[it->second].push_back(p);
    live.erase(it);
}

size_t csd_ctx::alloc_size(void* p) {
    std::lock_guard<std::mutex> g(mu);
    auto it = live.find(p);
    return it == live.end() ? 0 : it->second;
}

TimedScope::TimedScope(csd_ctx* c, const char* name, double bytes) : ctx(c), idx(SIZE_MAX) {
    if (!ctx->timing) return;
    csd_ctx::TimedLaunch t;
    t.name = name;
    t.bytes = bytes;
    HIPCHECK(hipEventCreate(&t.start));
    HIPCHECK(hipEventCreate(&t.stop));
    HIPCHECK(hipEventRecord(t.start, ctx->stream));
    idx = ctx->timed.size();
    ctx->timed.push_back(t);
}

TimedScope::~TimedScope() {
    if (idx != SIZE_MAX) (void)hipEventRecord(ctx->timed[idx].stop, ctx->stream);
}

extern "C" int csd_ctx_set_timing(csd_ctx* ctx, int on) {
    CSD_TRY ctx->timing = on != 0;
    CSD_CATCH
}

// one line per kernel name: "<name> <launches> <total ms> <total algorithmic bytes>"
extern "C" int csd_ctx_timing_report(csd_ctx* ctx, char* buf, size_t cap) {
    CSD_TRY HIPCHECK(hipStreamSynchronize(ctx->stream));
    std::map<std::string, std::tuple<long, double, double>> agg;
    for (auto& t : ctx->timed) {
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, t.start, t.stop));
        auto& a = agg[t.name];
        std::get<0>(a) += 1;
        std::get<1>(a) += ms;
        std::get<2>(a) += t.bytes;
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    ctx->timed.clear();
    std::string out;
    char line[256];
    for (auto& kv : agg) {
        snprintf(line, sizeof line, "%s %ld %.6f %.1f\n", kv.first.c_str(), std::get<0>(kv.second), std::get<1>(kv.second),
                 std::get<2>(kv.second));
        out += line;
    }
    if (out.size() + 1 > cap) throw CsdError("timing report buffer too small");
    memcpy(buf, out.c_str(), out.size() + 1);
    CSD_CATCH
}

__global__ void k_fill_zero(uint64_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}

void csd_zero_fill(csd_ctx* ctx, void* p, size_t bytes) {
    size_t words = bytes / 8;
    if (!words) return;
    unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, ctx->stream, (uint64_t*)p, words);
    HIPCHECK(hipGetLastError());
}

extern "C" int csd_ctx_create(int device, csd_ctx** out) {
    CSD_TRY
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw CsdError("no such HIP device");
    HIPCHECK(hipSetDevice(device));
    csd_zp_scratch_setup(device);  // the no-allocation entry point's scratch on this device
    csd_ctx* c = new csd_ctx();
    c->device = device;
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c;
    CSD_CATCH
}

extern "C" int csd_ctx_destroy(csd_ctx* ctx) {
    CSD_TRY
    if (!ctx) return 0;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (auto& kv : ctx->pool)
        for (void* p : kv.second) (void)hipFree(p);
    for (auto& kv : ctx->live) (void)hipFree(kv.first);
    if (ctx->copy_stream) {
        (void)hipStreamSynchronize(ctx->copy_stream);
        (void)hipStreamDestroy(ctx->copy_stream);
    }
    for (auto& kv : ctx->pinned) (void)hipHostFree(kv.first);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    CSD_CATCH
}

// Pinned (page-locked) host memory, allocated as such (never registered after the fact), for
// tensors that live on the host and are streamed to the device, e.g. Galois keys that do not
// stay resident. Freed with csd_host_free or at context destruction.
extern "C" int csd_host_alloc(csd_ctx* ctx, size_t bytes, void** out) {
    CSD_TRY if (!out) throw CsdError("csd_host_alloc: null output");
    void* p = nullptr;
    HIPCHECK(hipHostMalloc(&p, std::max<size_t>(bytes, 8), hipHostMallocDefault));
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->pinned[p] = bytes;
    *out = p;
    CSD_CATCH
}

extern "C" int csd_host_free(csd_ctx* ctx, void* p) {
    CSD_TRY if (!p) return 0;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->pinned.find(p);
        if (it == ctx->pinned.end()) throw CsdError("csd_host_free: not a block of this context");
        ctx->pinned.erase(it);
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->copy_stream) HIPCHECK(hipStreamSynchronize(ctx->copy_stream));
    HIPCHECK(hipHostFree(p));
    CSD_CATCH
}

extern "C" int csd_ctx_sync(csd_ctx* ctx) {
    CSD_TRY
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    CSD_CATCH
}

extern "C" int csd_ctx_stream(csd_ctx* ctx, void** stream) {
    CSD_TRY
    *stream = (void*)ctx->stream;
    CSD_CATCH
}

static void* allocate_tensor(csd_ctx* ctx, int elem_bytes, const int64_t shape[4], bool zero) {
    if (elem_bytes != 8 && elem_bytes != 16) throw CsdError("elem_bytes must be 8 or 16");
    size_t n = 1;
    for (int d = 0; d < 4; d++) {
        if (shape[d] < 0) throw CsdError("negative shape");
        n *= (size_t)shape[d];
    }
    void* p = ctx->alloc_bytes(n * elem_bytes);
    if (zero) csd_zero_fill(ctx, p, ((n * elem_bytes) + 7) & ~(size_t)7);
    return p;
}

extern "C" int csd_hal_allocate(csd_ctx* ctx, int elem_bytes, const int64_t shape[4], void** out) {
    CSD_TRY *out = allocate_tensor(ctx, elem_bytes, shape, true);
    CSD_CATCH
}

extern "C" int csd_hal_allocate_uninit(csd_ctx* ctx, int elem_bytes, const int64_t shape[4], void** out) {
    CSD_TRY *out = allocate_tensor(ctx, elem_bytes, shape, false);
    CSD_CATCH
}

extern "C" int csd_hal_free(csd_ctx* ctx, void* ptr) {
    CSD_TRY
    ctx->free_bytes(ptr);
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// views and broadcasting (hal/cpu.rs:333-346)

Shape4 shape_of(const csd_view* v) {
    Shape4 s;
    for (int d = 0; d < 4; d++) s.d[d] = v->shape[d];
    return s;
}

int64_t numel(const csd_view* v) { return v->shape[0] * v->shape[1] * v->shape[2] * v->shape[3]; }

void check_view(const csd_view* v, const char* what) {
    if (!v) throw CsdError(std::string(what) + ": null view");
    for (int d = 0; d < 4; d++)
        if (v->shape[d] < 0) throw CsdError(std::string(what) + ": negative shape");
    if (v->shape[3] > 1 && v->strides[3] != 1) throw CsdError(std::string(what) + ": last stride must be 1");
    if (numel(v) > 0 && !v->ptr) throw CsdError(std::string(what) + ": null pointer");
}

Str4 bcast_strides(const csd_view* src, const csd_view* dst, const char* what) {
    Str4 s;
    for (int d = 0; d < 4; d++) {
        if (src->shape[d] != dst->shape[d]) {
            if (src->shape[d] != 1)
                throw CsdError(std::string(what) + ": cannot broadcast dimension " + std::to_string(d) + " of length " +
                               std::to_string(src->shape[d]) + " to length " + std::to_string(dst->shape[d]));
            s.s[d] = 0;
        } else {
            s.s[d] = src->shape[d] == 1 ? 0 : src->strides[d];
        }
    }
    return s;
}

Str4 own_strides(const csd_view* v) {
    Str4 s;
    for (int d = 0; d < 4; d++) s.s[d] = v->shape[d] == 1 ? 0 : v->strides[d];
    return s;
}

void csd_allow_lds(const void* fn, int bytes) {
    static std::mutex mu;
    static std::set<std::pair<const void*, int>> done;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (done.count({fn, dev})) return;
    HIPCHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done.insert({fn, dev});
}

ModTable make_mod_table(const csd_modulus* mods, size_t nmods, int64_t d1) {
    if ((int64_t)nmods != d1) throw CsdError("number of moduli must equal dimension 1 of the destination");
    if (nmods > CSD_MAX_MODULI) throw CsdError("too many moduli (at most CSD_MAX_MODULI)");
    ModTable t;
    memset(&t, 0, sizeof(t));
    for (size_t j = 0; j < nmods; j++) {
        if (mods[j].N <= 1 || (mods[j].N & 1) == 0) throw CsdError("invalid modulus");
        t.m[j] = CsdMod{mods[j].N, mods[j].neg_inv_N, mods[j].r2, mods[j].r3};
    }
    return t;
}

// ---------------------------------------------------------------------------------------------
// element-wise kernels

enum ElemOp {
    OP_ASSIGN,
    OP_ASSIGN_NARROW,
    OP_ASSIGN_I128,
    OP_ADD_I64,
    OP_SUB_I64,
    OP_ADD_I128,
    OP_FMA_I64,
    OP_MUL_I64,
    OP_MUL_MONT,
    OP_ADD_RED_DOUBLE,
    OP_RED_DOUBLE,
    OP_ADD_RED,
    OP_RED,
    OP_I128_TO_MONT,
    OP_I64_TO_MONT,
    OP_SUM_I64,
};

struct ElemArgs {
    int64_t d0, d1, d2, d3;
    int64_t sd[3];  // dst strides of dims 0..2 (dim 3 contiguous)
    Str4 sa, sb;
    void* pd;
    const void* pa;
    const void* pb;
    int64_t chunks;  // chunks per row in line mode
    int line_mode;
};

#define ELEM_BLOCK 256
#define ELEM_ITEMS 4

template <int OP>
__device__ __forceinline__ void elem_apply(const ElemArgs& a, const CsdMod& m, int64_t od, int64_t oa, int64_t ob) {
    int64_t* d64 = (int64_t*)a.pd;
    i128* d128 = (i128*)a.pd;
    const int64_t* a64 = (const int64_t*)a.pa;
    const i128* a128 = (const i128*)a.pa;
    const int64_t* b64 = (const int64_t*)a.pb;
    if constexpr (OP == OP_ASSIGN) d64[od] = a64[oa];
    if constexpr (OP == OP_ASSIGN_NARROW) d64[od] = (int64_t)a128[oa];
    if constexpr (OP == OP_ASSIGN_I128) d128[od] = a128[oa];
    if constexpr (OP == OP_ADD_I64) d64[od] = (int64_t)((uint64_t)d64[od] + (uint64_t)a64[oa]);
    if constexpr (OP == OP_SUB_I64) d64[od] = (int64_t)((uint64_t)d64[od] - (uint64_t)a64[oa]);
    if constexpr (OP == OP_ADD_I128) d128[od] = (i128)((u128)d128[od] + (u128)a128[oa]);
    if constexpr (OP == OP_FMA_I64) d128[od] = (i128)((u128)d128[od] + (u128)((i128)a64[oa] * (i128)b64[ob]));
    if constexpr (OP == OP_MUL_I64) d128[od] = (i128)a64[oa] * (i128)b64[ob];
    if constexpr (OP == OP_MUL_MONT) d64[od] = csd_mont_mul(m, d64[od], a64[oa]);
    if constexpr (OP == OP_ADD_RED_DOUBLE) d64[od] = (int64_t)((uint64_t)d64[od] + (uint64_t)csd_double_red128(m, a128[oa]));
    if constexpr (OP == OP_RED_DOUBLE) d64[od] = csd_double_red128(m, a128[oa]);
    if constexpr (OP == OP_ADD_RED) d64[od] = (int64_t)((uint64_t)d64[od] + (uint64_t)csd_red128(m, a128[oa]));
    if constexpr (OP == OP_RED) d64[od] = csd_red128(m, a128[oa]);
    if constexpr (OP == OP_I128_TO_MONT) d64[od] = csd_to_mont128(m, a128[oa]);
    if constexpr (OP == OP_I64_TO_MONT) d64[od] = csd_to_mont64(m, d64[od]);
    if constexpr (OP == OP_SUM_I64) d64[od] = (int64_t)((uint64_t)a64[oa] + (uint64_t)b64[ob]);
}

template <int OP>
__global__ __launch_bounds__(ELEM_BLOCK) void k_elem(ElemArgs a, ModTable mt) {
    if (a.line_mode) {
        int64_t row = blockIdx.x / a.chunks;
        int64_t c = blockIdx.x - row * a.chunks;
        int64_t k = row % a.d2;
        int64_t t = row / a.d2;
        int64_t j = t % a.d1;
        int64_t i = t / a.d1;
        CsdMod m = mt.m[j];
        int64_t bd = i * a.sd[0] + j * a.sd[1] + k * a.sd[2];
        int64_t ba = i * a.sa.s[0] + j * a.sa.s[1] + k * a.sa.s[2];
        int64_t bb = i * a.sb.s[0] + j * a.sb.s[1] + k * a.sb.s[2];
        int64_t l0 = c * (ELEM_BLOCK * ELEM_ITEMS) + threadIdx.x;
#pragma unroll
        for (int it = 0; it < ELEM_ITEMS; it++) {
            int64_t l = l0 + it * ELEM_BLOCK;
            if (l < a.d3) elem_apply<OP>(a, m, bd + l, ba + l * a.sa.s[3], bb + l * a.sb.s[3]);
        }
    } else {
        int64_t total = a.d0 * a.d1 * a.d2 * a.d3;
        for (int64_t e = (int64_t)blockIdx.x * ELEM_BLOCK + threadIdx.x; e < total; e += (int64_t)gridDim.x * ELEM_BLOCK) {
            int64_t l = e % a.d3;
            int64_t t = e / a.d3;
            int64_t k = t % a.d2;
            t /= a.d2;
            int64_t j = t % a.d1;
            int64_t i = t / a.d1;
            int64_t od = i * a.sd[0] + j * a.sd[1] + k * a.sd[2] + l;
            int64_t oa = i * a.sa.s[0] + j * a.sa.s[1] + k * a.sa.s[2] + l * a.sa.s[3];
            int64_t ob = i * a.sb.s[0] + j * a.sb.s[1] + k * a.sb.s[2] + l * a.sb.s[3];
            elem_apply<OP>(a, mt.m[j], od, oa, ob);
        }
    }
}

template <int OP>
static void launch_elem(csd_ctx* ctx, const csd_view* dst, const csd_view* sa, const csd_view* sb, const ModTable* mt) {
    ElemArgs a;
    memset(&a, 0, sizeof(a));
    a.d0 = dst->shape[0];
    a.d1 = dst->shape[1];
    a.d2 = dst->shape[2];
    a.d3 = dst->shape[3];
    for (int d = 0; d < 3; d++) a.sd[d] = dst->strides[d];
    a.pd = dst->ptr;
    if (sa) {
        a.sa = bcast_strides(sa, dst, "source");
        a.pa = sa->ptr;
    }
    if (sb) {
        a.sb = bcast_strides(sb, dst, "source");
        a.pb = sb->ptr;
    }
    int64_t total = numel(dst);
    if (total == 0) return;
    static ModTable zero_table;
    const ModTable& tab = mt ? *mt : zero_table;
    int64_t blocks;
    if (a.d3 >= 64) {
        a.line_mode = 1;
        a.chunks = (a.d3 + ELEM_BLOCK * ELEM_ITEMS - 1) / (ELEM_BLOCK * ELEM_ITEMS);
        blocks = a.d0 * a.d1 * a.d2 * a.chunks;
        if (blocks > (int64_t)INT32_MAX) throw CsdError("tensor too large for one launch");
    } else {
        a.line_mode = 0;
        blocks = std::min<int64_t>((total + ELEM_BLOCK - 1) / ELEM_BLOCK, 1 << 20);
    }
    hipLaunchKernelGGL(k_elem<OP>, dim3((unsigned)blocks), dim3(ELEM_BLOCK), 0, ctx->stream, a, tab);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// import / export (hal/cpu.rs:566-589): the host buffer is the row-major enumeration of the view

struct CopyArgs {
    int64_t d0, d1, d2, d3;
    int64_t st[4];
    int words;  // 64-bit words per element
};

__global__ void k_scatter(CopyArgs a, int64_t* view_base, const int64_t* packed) {
    int64_t total = a.d0 * a.d1 * a.d2 * a.d3;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = e % a.d3, t = e / a.d3;
        int64_t k = t % a.d2;
        t /= a.d2;
        int64_t j = t % a.d1, i = t / a.d1;
        int64_t off = (i * a.st[0] + j * a.st[1] + k * a.st[2] + l * a.st[3]) * a.words;
        for (int w = 0; w < a.words; w++) view_base[off + w] = packed[e * a.words + w];
    }
}

__global__ void k_gather(CopyArgs a, const int64_t* view_base, int64_t* packed) {
    int64_t total = a.d0 * a.d1 * a.d2 * a.d3;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = e % a.d3, t = e / a.d3;
        int64_t k = t % a.d2;
        t /= a.d2;
        int64_t j = t % a.d1, i = t / a.d1;
        int64_t off = (i * a.st[0] + j * a.st[1] + k * a.st[2] + l * a.st[3]) * a.words;
        for (int w = 0; w < a.words; w++) packed[e * a.words + w] = view_base[off + w];
    }
}

static bool is_contiguous(const csd_view* v) {
    int64_t expect = 1;
    for (int d = 3; d >= 0; d--) {
        if (v->shape[d] != 1 && v->strides[d] != expect) return false;
        expect *= v->shape[d];
    }
    return true;
}

static void transfer(csd_ctx* ctx, const csd_view* v, int words, int64_t* host, size_t len, bool to_device) {
    check_view(v, to_device ? "import" : "export");
    if ((int64_t)len != numel(v)) throw CsdError(to_device ? "import length mismatch" : "export length mismatch");
    if (len == 0) return;
    size_t bytes = len * words * 8;
    if (is_contiguous(v)) {
        if (to_device) {
            HIPCHECK(hipMemcpyAsync(v->ptr, host, bytes, hipMemcpyHostToDevice, ctx->stream));
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        } else {
            HIPCHECK(hipMemcpyAsync(host, v->ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        }
        return;
    }
    int64_t* staging = (int64_t*)ctx->alloc_bytes(bytes);
    CopyArgs a;
    a.d0 = v->shape[0];
    a.d1 = v->shape[1];
    a.d2 = v->shape[2];
    a.d3 = v->shape[3];
    for (int d = 0; d < 4; d++) a.st[d] = v->strides[d];
    a.words = words;
    unsigned blocks = (unsigned)std::min<int64_t>(((int64_t)len + 255) / 256, 65536);
    try {
        if (to_device) {
            HIPCHECK(hipMemcpyAsync(staging, host, bytes, hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, ctx->stream, a, (int64_t*)v->ptr, (const int64_t*)staging);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        } else {
            hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, ctx->stream, a, (const int64_t*)v->ptr, staging);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(host, staging, bytes, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        }
    } catch (...) {
        ctx->free_bytes(staging);
        throw;
    }
    ctx->free_bytes(staging);
}

extern "C" int csd_hal_import(csd_ctx* ctx, const csd_view* dst, const int64_t* src, size_t len) {
    CSD_TRY transfer(ctx, dst, 1, (int64_t*)src, len, true);
    CSD_CATCH
}
extern "C" int csd_hal_export(csd_ctx* ctx, const csd_view* src, int64_t* dst, size_t len) {
    CSD_TRY transfer(ctx, src, 1, dst, len, false);
    CSD_CATCH
}
extern "C" int csd_import_i128(csd_ctx* ctx, const csd_view* dst, const int64_t* src, size_t len) {
    CSD_TRY transfer(ctx, dst, 2, (int64_t*)src, len, true);
    CSD_CATCH
}
extern "C" int csd_export_i128(csd_ctx* ctx, const csd_view* src, int64_t* dst, size_t len) {
    CSD_TRY transfer(ctx, src, 2, dst, len, false);
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// the debug range check of kernel outputs: how many entries of an int64 view have |x| > k N for the
// modulus N of their limb (dimension 1). One thread per entry, the count gathered with one atomic
// add per workgroup into a zeroed device word.

__global__ __launch_bounds__(256) void k_range_check(const int64_t* __restrict__ p, int64_t d1, int64_t d2, int64_t d3,
                                                     int64_t s0, int64_t s1, int64_t s2, int64_t s3, int64_t total,
                                                     const int64_t* __restrict__ lim, unsigned long long* count) {
    __shared__ unsigned int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < total) {
        const int64_t l = e % d3, t = e / d3, k = t % d2, t2 = t / d2, j = t2 % d1, i = t2 / d1;
        const int64_t x = p[i * s0 + j * s1 + k * s2 + l * s3];
        const uint64_t ax = x < 0 ? (uint64_t)0 - (uint64_t)x : (uint64_t)x;
        if (ax > (uint64_t)lim[j]) atomicAdd(&bad, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && bad) atomicAdd(count, (unsigned long long)bad);
}

extern "C" int csd_hal_range_check(csd_ctx* ctx, const csd_view* v, const csd_modulus* mods, size_t nmods, int64_t multiple,
                                   uint64_t* out_of_range) {
    CSD_TRY check_view(v, "range_check");
    if (!out_of_range) throw CsdError("range_check: null result");
    if ((int64_t)nmods != v->shape[1]) throw CsdError("range_check: one modulus per limb");
    if (multiple < 1) throw CsdError("range_check: the bound multiple must be positive");
    std::vector<int64_t> lim(nmods);
    for (size_t j = 0; j < nmods; j++) {
        if (mods[j].N <= 1) throw CsdError("range_check: invalid modulus");
        const u128 b = (u128)(uint64_t)mods[j].N * (u128)(uint64_t)multiple;
        lim[j] = b >= ((u128)1 << 63) ? INT64_MAX : (int64_t)b;  // k N beyond 2^63: every i64 passes
    }
    *out_of_range = 0;
    const int64_t total = numel(v);
    if (total == 0) return 0;
    const size_t bytes = nmods * 8 + 8;
    char* buf = (char*)ctx->alloc_bytes(bytes);
    try {
        int64_t* dlim = (int64_t*)(buf + 8);
        unsigned long long* dcount = (unsigned long long*)buf;
        csd_zero_fill(ctx, dcount, 8);
        HIPCHECK(hipMemcpyAsync(dlim, lim.data(), nmods * 8, hipMemcpyHostToDevice, ctx->stream));
        const int64_t blocks = (total + 255) / 256;
        if (blocks > (int64_t)INT32_MAX) throw CsdError("range_check: tensor too large for one launch");
        hipLaunchKernelGGL(k_range_check, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (const int64_t*)v->ptr,
                           v->shape[1], v->shape[2], v->shape[3], v->strides[0], v->strides[1], v->strides[2], v->strides[3],
                           total, dlim, dcount);
        HIPCHECK(hipGetLastError());
        unsigned long long host = 0;
        HIPCHECK(hipMemcpyAsync(&host, dcount, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        *out_of_range = host;
    } catch (...) {
        ctx->free_bytes(buf);
        throw;
    }
    ctx->free_bytes(buf);
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// element-wise entry points (hal/cpu.rs:591-704, 823-916)

#define BIN_ENTRY(name, OP)                                                                  \
    extern "C" int name(csd_ctx* ctx, const csd_view* dst, const csd_view* src) {           \
        CSD_TRY check_view(dst, #name);                                                      \
        check_view(src, #name);                                                              \
        launch_elem<OP>(ctx, dst, src, nullptr, nullptr);                                    \
        CSD_CATCH                                                                            \
    }
BIN_ENTRY(csd_hal_assign, OP_ASSIGN)
BIN_ENTRY(csd_hal_assign_narrow, OP_ASSIGN_NARROW)
BIN_ENTRY(csd_hal_assign_i128, OP_ASSIGN_I128)
BIN_ENTRY(csd_hal_add_i64, OP_ADD_I64)
BIN_ENTRY(csd_hal_sub_i64, OP_SUB_I64)
BIN_ENTRY(csd_hal_add_i128, OP_ADD_I128)

extern "C" int csd_hal_fma_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs) {
    CSD_TRY check_view(dst, "fma_i64");
    check_view(lhs, "fma_i64");
    check_view(rhs, "fma_i64");
    launch_elem<OP_FMA_I64>(ctx, dst, lhs, rhs, nullptr);
    CSD_CATCH
}
extern "C" int csd_hal_sum_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs) {
    CSD_TRY check_view(dst, "sum_i64");
    check_view(lhs, "sum_i64");
    check_view(rhs, "sum_i64");
    launch_elem<OP_SUM_I64>(ctx, dst, lhs, rhs, nullptr);
    CSD_CATCH
}
extern "C" int csd_hal_mul_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs) {
    CSD_TRY check_view(dst, "mul_i64");
    check_view(lhs, "mul_i64");
    check_view(rhs, "mul_i64");
    launch_elem<OP_MUL_I64>(ctx, dst, lhs, rhs, nullptr);
    CSD_CATCH
}

#define MOD_ENTRY(name, OP)                                                                                    \
    extern "C" int name(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_modulus* mods, size_t n) { \
        CSD_TRY check_view(dst, #name);                                                                        \
        check_view(src, #name);                                                                                \
        ModTable t = make_mod_table(mods, n, dst->shape[1]);                                                   \
        launch_elem<OP>(ctx, dst, src, nullptr, &t);                                                           \
        CSD_CATCH                                                                                              \
    }
MOD_ENTRY(csd_hal_mul_mont64, OP_MUL_MONT)
MOD_ENTRY(csd_hal_add_red_mont64_double, OP_ADD_RED_DOUBLE)
MOD_ENTRY(csd_hal_red_mont64_double, OP_RED_DOUBLE)
MOD_ENTRY(csd_hal_add_red_mont64, OP_ADD_RED)
MOD_ENTRY(csd_hal_red_mont64, OP_RED)
MOD_ENTRY(csd_hal_i128_to_mont64, OP_I128_TO_MONT)

extern "C" int csd_hal_i64_to_mont64(csd_ctx* ctx, const csd_view* dst, const csd_modulus* mods, size_t n) {
    CSD_TRY check_view(dst, "i64_to_mont64");
    ModTable t = make_mod_table(mods, n, dst->shape[1]);
    launch_elem<OP_I64_TO_MONT>(ctx, dst, nullptr, nullptr, &t);
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// inner product (hal/cpu.rs:706-788): the i128 accumulator stays in registers across all terms

#define IP_MAX_TERMS 24  // per launch; keeps the term table inside the kernel-argument segment

struct IpArgs {
    int64_t d0, d1, d2, d3;
    int64_t sd[3];
    i128* pd;
    int terms;
    int accumulate;  // 0: overwrite dst (first chunk of terms), 1: add onto dst
    int line_mode;
    int64_t chunks;
    const int64_t* pl[IP_MAX_TERMS];
    const int64_t* pr[IP_MAX_TERMS];
    Str4 sl[IP_MAX_TERMS];
    Str4 sr[IP_MAX_TERMS];
};

__global__ __launch_bounds__(ELEM_BLOCK) void k_inner_product(IpArgs a) {
    if (a.line_mode) {
        int64_t row = blockIdx.x / a.chunks;
        int64_t c = blockIdx.x - row * a.chunks;
        int64_t k = row % a.d2, t = row / a.d2;
        int64_t j = t % a.d1, i = t / a.d1;
        int64_t l0 = c * (ELEM_BLOCK * ELEM_ITEMS) + threadIdx.x;
        i128* D = a.pd + i * a.sd[0] + j * a.sd[1] + k * a.sd[2];
        i128 acc[ELEM_ITEMS];
#pragma unroll
        for (int it = 0; it < ELEM_ITEMS; it++) {
            int64_t l = l0 + it * ELEM_BLOCK;
            acc[it] = (a.accumulate && l < a.d3) ? D[l] : (i128)0;
        }
        for (int term = 0; term < a.terms; term++) {
            const Str4 sl = a.sl[term], sr = a.sr[term];
            const int64_t* L = a.pl[term] + i * sl.s[0] + j * sl.s[1] + k * sl.s[2];
            const int64_t* R = a.pr[term] + i * sr.s[0] + j * sr.s[1] + k * sr.s[2];
#pragma unroll
            for (int it = 0; it < ELEM_ITEMS; it++) {
                int64_t l = l0 + it * ELEM_BLOCK;
                if (l < a.d3) acc[it] = (i128)((u128)acc[it] + (u128)((i128)L[l * sl.s[3]] * (i128)R[l * sr.s[3]]));
            }
        }
#pragma unroll
        for (int it = 0; it < ELEM_ITEMS; it++) {
            int64_t l = l0 + it * ELEM_BLOCK;
            if (l < a.d3) D[l] = acc[it];
        }
    } else {
        int64_t total = a.d0 * a.d1 * a.d2 * a.d3;
        for (int64_t e = (int64_t)blockIdx.x * ELEM_BLOCK + threadIdx.x; e < total; e += (int64_t)gridDim.x * ELEM_BLOCK) {
            int64_t l = e % a.d3, t = e / a.d3;
            int64_t k = t % a.d2;
            t /= a.d2;
            int64_t j = t % a.d1, i = t / a.d1;
            i128* D = a.pd + i * a.sd[0] + j * a.sd[1] + k * a.sd[2] + l;
            i128 acc = a.accumulate ? *D : (i128)0;
            for (int term = 0; term < a.terms; term++) {
                const Str4 sl = a.sl[term], sr = a.sr[term];
                int64_t x = a.pl[term][i * sl.s[0] + j * sl.s[1] + k * sl.s[2] + l * sl.s[3]];
                int64_t y = a.pr[term][i * sr.s[0] + j * sr.s[1] + k * sr.s[2] + l * sr.s[3]];
                acc = (i128)((u128)acc + (u128)((i128)x * (i128)y));
            }
            *D = acc;
        }
    }
}

extern "C" int csd_hal_inner_product_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs,
                                         size_t terms) {
    CSD_TRY check_view(dst, "inner_product_i64");
    if (terms < 1) throw CsdError("inner_product_i64 requires at least one term");
    for (size_t t = 0; t < terms; t++) {
        check_view(&lhs[t], "inner_product_i64 lhs");
        check_view(&rhs[t], "inner_product_i64 rhs");
        (void)bcast_strides(&lhs[t], dst, "inner_product_i64 lhs");
        (void)bcast_strides(&rhs[t], dst, "inner_product_i64 rhs");
    }
    if (numel(dst) == 0) return 0;
    // more terms than one launch holds: later launches accumulate onto the first (wrapping
    // addition is associative, so the result is unchanged)
    for (size_t t0 = 0; t0 < terms; t0 += IP_MAX_TERMS) {
        size_t nt = std::min<size_t>(IP_MAX_TERMS, terms - t0);
        IpArgs a;
        memset(&a, 0, sizeof(a));
        a.d0 = dst->shape[0];
        a.d1 = dst->shape[1];
        a.d2 = dst->shape[2];
        a.d3 = dst->shape[3];
        for (int d = 0; d < 3; d++) a.sd[d] = dst->strides[d];
        a.pd = (i128*)dst->ptr;
        a.terms = (int)nt;
        a.accumulate = t0 == 0 ? 0 : 1;
        for (size_t t = 0; t < nt; t++) {
            a.pl[t] = (const int64_t*)lhs[t0 + t].ptr;
            a.pr[t] = (const int64_t*)rhs[t0 + t].ptr;
            a.sl[t] = bcast_strides(&lhs[t0 + t], dst, "inner_product_i64 lhs");
            a.sr[t] = bcast_strides(&rhs[t0 + t], dst, "inner_product_i64 rhs");
        }
        int64_t blocks;
        if (a.d3 >= 64) {
            a.line_mode = 1;
            a.chunks = (a.d3 + ELEM_BLOCK * ELEM_ITEMS - 1) / (ELEM_BLOCK * ELEM_ITEMS);
            blocks = a.d0 * a.d1 * a.d2 * a.chunks;
        } else {
            a.line_mode = 0;
            blocks = std::min<int64_t>((numel(dst) + ELEM_BLOCK - 1) / ELEM_BLOCK, 1 << 20);
        }
        hipLaunchKernelGGL(k_inner_product, dim3((unsigned)blocks), dim3(ELEM_BLOCK), 0, ctx->stream, a);
        HIPCHECK(hipGetLastError());
    }
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// matmul (hal/cpu.rs:790-818): dst[i,j,k,l] = sum_u lhs[i,j,u,0] * rhs[i,u,k,l]

struct MmArgs {
    int64_t I, J, K, L, U;
    int64_t sd[3];
    Str4 sl, sr;  // sl over (i, j, u); sr over (i, u, k, l)
    i128* pd;
    const int64_t* pl;
    const int64_t* pr;
    int64_t chunks;
};

#define MM_JT 8  // output rows j computed per thread from one pass over the rhs column

__global__ __launch_bounds__(ELEM_BLOCK) void k_matmul(MmArgs a) {
    // block -> (i, k, j-tile, l-chunk); each thread owns one l and MM_JT rows j
    int64_t jt = (a.J + MM_JT - 1) / MM_JT;
    int64_t b = blockIdx.x;
    int64_t c = b % a.chunks;
    b /= a.chunks;
    int64_t jtile = b % jt;
    b /= jt;
    int64_t k = b % a.K;
    int64_t i = b / a.K;
    int64_t l = c * ELEM_BLOCK + threadIdx.x;
    if (l >= a.L) return;
    i128 acc[MM_JT];
#pragma unroll
    for (int q = 0; q < MM_JT; q++) acc[q] = 0;
    const int64_t* R = a.pr + i * a.sr.s[0] + k * a.sr.s[2] + l * a.sr.s[3];
    const int64_t* Lm = a.pl + i * a.sl.s[0];
    int64_t j0 = jtile * MM_JT;
    for (int64_t u = 0; u < a.U; u++) {
        int64_t r = R[u * a.sr.s[1]];
#pragma unroll
        for (int q = 0; q < MM_JT; q++) {
            int64_t j = j0 + q;
            if (j < a.J) {
                int64_t c0 = Lm[j * a.sl.s[1] + u * a.sl.s[2]];
                acc[q] = (i128)((u128)acc[q] + (u128)((i128)c0 * (i128)r));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < MM_JT; q++) {
        int64_t j = j0 + q;
        if (j < a.J) a.pd[i * a.sd[0] + j * a.sd[1] + k * a.sd[2] + l] = acc[q];
    }
}

extern "C" int csd_hal_matmul_i64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs) {
    CSD_TRY check_view(dst, "matmul_i64");
    check_view(lhs, "matmul_i64 lhs");
    check_view(rhs, "matmul_i64 rhs");
    MmArgs a;
    a.I = dst->shape[0];
    a.J = dst->shape[1];
    a.K = dst->shape[2];
    a.L = dst->shape[3];
    a.U = lhs->shape[2];
    if (!(lhs->shape[0] == a.I || lhs->shape[0] == 1)) throw CsdError("matmul lhs dim 0");
    if (lhs->shape[1] != a.J) throw CsdError("matmul lhs dim 1");
    if (lhs->shape[3] != 1) throw CsdError("matmul requires lhs dim 3 of length 1");
    if (!(rhs->shape[0] == a.I || rhs->shape[0] == 1)) throw CsdError("matmul rhs dim 0");
    if (rhs->shape[1] != a.U || rhs->shape[2] != a.K || rhs->shape[3] != a.L) throw CsdError("matmul rhs shape");
    for (int d = 0; d < 3; d++) a.sd[d] = dst->strides[d];
    a.sl.s[0] = lhs->shape[0] == 1 ? 0 : lhs->strides[0];
    a.sl.s[1] = lhs->strides[1];
    a.sl.s[2] = lhs->strides[2];
    a.sl.s[3] = 0;
    a.sr.s[0] = rhs->shape[0] == 1 ? 0 : rhs->strides[0];
    a.sr.s[1] = rhs->strides[1];
    a.sr.s[2] = rhs->strides[2];
    a.sr.s[3] = rhs->strides[3];
    if (a.L <= 1) a.sr.s[3] = 0;
    a.pd = (i128*)dst->ptr;
    a.pl = (const int64_t*)lhs->ptr;
    a.pr = (const int64_t*)rhs->ptr;
    if (numel(dst) == 0) return 0;
    a.chunks = (a.L + ELEM_BLOCK - 1) / ELEM_BLOCK;
    int64_t blocks = a.I * a.K * ((a.J + MM_JT - 1) / MM_JT) * a.chunks;
    if (blocks > (int64_t)INT32_MAX) throw CsdError("matmul too large for one launch");
    hipLaunchKernelGGL(k_matmul, dim3((unsigned)blocks), dim3(ELEM_BLOCK), 0, ctx->stream, a);
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// bconv corrections (hal/cpu.rs:950-1057)

struct BcArgs {
    int64_t I, J, K, L;
    int64_t sd[3];
    Str4 sl;  // lhs (i, 0, k, l)
    int64_t sr0, sr1;
    int64_t* pd;
    const i128* pl;
    const int64_t* pr;
    int shift;
};

__global__ void k_bconv_correct_mont(BcArgs a, ModTable mt) {
    int64_t total = a.I * a.J * a.K * a.L;
    i128 half = (i128)1 << (a.shift - 1);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = e % a.L, t = e / a.L;
        int64_t k = t % a.K;
        t /= a.K;
        int64_t j = t % a.J, i = t / a.J;
        i128 x = a.pl[i * a.sl.s[0] + k * a.sl.s[2] + l * a.sl.s[3]];
        i128 q = (i128)((u128)x + (u128)half) >> a.shift;
        i128 r = a.pr[i * a.sr0 + j * a.sr1];
        i128 prod = (i128)((u128)q * (u128)r);
        int64_t* d = a.pd + i * a.sd[0] + j * a.sd[1] + k * a.sd[2] + l;
        *d = (int64_t)((uint64_t)*d + (uint64_t)csd_red128(mt.m[j], prod));
    }
}

extern "C" int csd_hal_bconv_correct_mont64(csd_ctx* ctx, const csd_view* dst, const csd_view* lhs, const csd_view* rhs,
                                            const csd_modulus* mods, size_t nmods, size_t shift) {
    CSD_TRY check_view(dst, "bconv_correct_mont64");
    check_view(lhs, "bconv_correct_mont64 lhs");
    check_view(rhs, "bconv_correct_mont64 rhs");
    if (shift < 64 || shift >= 128) throw CsdError("bconv_correct_mont64 requires 64 <= shift < 128");
    BcArgs a;
    a.I = dst->shape[0];
    a.J = dst->shape[1];
    a.K = dst->shape[2];
    a.L = dst->shape[3];
    if (lhs->shape[0] != a.I || lhs->shape[1] != 1 || lhs->shape[2] != a.K || lhs->shape[3] != a.L)
        throw CsdError("bconv_correct_mont64 lhs shape");
    if (!(rhs->shape[0] == a.I || rhs->shape[0] == 1)) throw CsdError("bconv_correct_mont64 rhs dim 0");
    if (rhs->shape[1] != a.J || rhs->shape[2] != 1 || rhs->shape[3] != 1) throw CsdError("bconv_correct_mont64 rhs shape");
    ModTable t = make_mod_table(mods, nmods, a.J);
    for (int d = 0; d < 3; d++) a.sd[d] = dst->strides[d];
    a.sl = own_strides(lhs);
    a.sr0 = rhs->shape[0] == 1 ? 0 : rhs->strides[0];
    a.sr1 = rhs->strides[1];
    a.pd = (int64_t*)dst->ptr;
    a.pl = (const i128*)lhs->ptr;
    a.pr = (const int64_t*)rhs->ptr;
    a.shift = (int)shift;
    int64_t total = numel(dst);
    if (total == 0) return 0;
    unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_bconv_correct_mont, dim3(blocks), dim3(256), 0, ctx->stream, a, t);
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}

struct Pow2Args {
    int64_t I, K, L;
    int64_t sd[4];
    int64_t ss[4];
    int64_t* pd;
    const i128* ps;
    i128 f0, f1;
    int shift, log2_modulus, wide;
};

__global__ void k_bconv_correct_pow2(Pow2Args a) {
    int64_t total = a.I * a.K * a.L;
    i128 half = (i128)1 << (a.shift - 1);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = e % a.L, t = e / a.L;
        int64_t k = t % a.K, i = t / a.K;
        const i128* s = a.ps + i * a.ss[0] + k * a.ss[2] + l * a.ss[3];
        int64_t* d = a.pd + i * a.sd[0] + k * a.sd[2] + l * a.sd[3];
        if (!a.wide) {
            i128 quot = (i128)((u128)s[a.ss[1]] + (u128)half) >> a.shift;
            i128 v = (i128)(((u128)s[0] - (u128)quot * (u128)a.f0) * (u128)a.f1);
            d[0] = (int64_t)csd_pow2_centre(v, a.log2_modulus);
        } else {
            i128 quot = (i128)((u128)s[2 * a.ss[1]] + (u128)half) >> a.shift;
            i128 v = (i128)((u128)s[0] + ((u128)s[a.ss[1]] << 64));
            i128 c = csd_pow2_centre((i128)(((u128)v - (u128)quot * (u128)a.f0) * (u128)a.f1), a.log2_modulus);
            int64_t lo = (int64_t)c;
            int64_t hi = (int64_t)((i128)((u128)c - (u128)(i128)lo) >> 64);
            d[0] = lo;
            d[a.sd[1]] = hi;
        }
    }
}

static int bconv_pow2_common(csd_ctx* ctx, const csd_view* dst, const csd_view* src, i128 f0, i128 f1, size_t shift,
                             size_t log2_modulus, bool wide) {
    check_view(dst, "bconv_correct_pow2");
    check_view(src, "bconv_correct_pow2 src");
    Pow2Args a;
    a.I = dst->shape[0];
    a.K = dst->shape[2];
    a.L = dst->shape[3];
    if (dst->shape[1] != (wide ? 2 : 1)) throw CsdError("bconv_correct_pow2: wrong dst dimension 1");
    if (src->shape[0] != a.I || src->shape[1] != (wide ? 3 : 2) || src->shape[2] != a.K || src->shape[3] != a.L)
        throw CsdError("bconv_correct_pow2: src shape");
    if (shift < 1 || shift >= 127) throw CsdError("bconv_correct_pow2: shift out of range");
    if (log2_modulus < 1 || log2_modulus > (wide ? 126u : 62u)) throw CsdError("bconv_correct_pow2: log2_modulus out of range");
    for (int d = 0; d < 4; d++) {
        a.sd[d] = dst->strides[d];
        a.ss[d] = src->strides[d];
    }
    a.pd = (int64_t*)dst->ptr;
    a.ps = (const i128*)src->ptr;
    a.f0 = f0;
    a.f1 = f1;
    a.shift = (int)shift;
    a.log2_modulus = (int)log2_modulus;
    a.wide = wide ? 1 : 0;
    int64_t total = a.I * a.K * a.L;
    if (total == 0) return 0;
    unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_bconv_correct_pow2, dim3(blocks), dim3(256), 0, ctx->stream, a);
    HIPCHECK(hipGetLastError());
    return 0;
}

extern "C" int csd_hal_bconv_correct_pow2(csd_ctx* ctx, const csd_view* dst, const csd_view* src, int64_t factor0,
                                          int64_t factor1, size_t shift, size_t log2_modulus) {
    CSD_TRY bconv_pow2_common(ctx, dst, src, (i128)factor0, (i128)factor1, shift, log2_modulus, false);
    CSD_CATCH
}

extern "C" int csd_hal_bconv_correct_pow2_wide(csd_ctx* ctx, const csd_view* dst, const csd_view* src, int64_t f0_lo,
                                               int64_t f0_hi, int64_t f1_lo, int64_t f1_hi, size_t shift,
                                               size_t log2_modulus) {
    CSD_TRY i128 f0 = (i128)(((u128)(uint64_t)f0_hi << 64) | (uint64_t)f0_lo);
    i128 f1 = (i128)(((u128)(uint64_t)f1_hi << 64) | (uint64_t)f1_lo);
    bconv_pow2_common(ctx, dst, src, f0, f1, shift, log2_modulus, true);
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// permutations (hal/cpu.rs:1059-1122)


extern "C" int csd_hal_create_permutation(csd_ctx* ctx, const size_t* inner_shape, const size_t* shifts, size_t n,
                                          csd_perm_plan** out) {
    CSD_TRY size_t total = 1;
    for (size_t t = 0; t < n; t++) {
        if (inner_shape[t] == 0) throw CsdError("permutation: zero-length dimension");
        total *= inner_shape[t];
    }
    if (total > UINT32_MAX) throw CsdError("permutation too long");
    std::vector<uint32_t> perm(total);
    // the entry at multi-index a moves to a + shifts, so position idx reads from idx - shifts
    std::vector<size_t> idx(n, 0);
    for (size_t e = 0; e < total; e++) {
        size_t src = 0;
        for (size_t t = 0; t < n; t++) {
            size_t m = inner_shape[t];
            src = src * m + (idx[t] + m - shifts[t] % m) % m;
        }
        perm[e] = (uint32_t)src;
        for (size_t t = n; t-- > 0;) {
            if (++idx[t] < inner_shape[t]) break;
            idx[t] = 0;
        }
    }
    csd_perm_plan* p = new csd_perm_plan();
    p->ctx = ctx;
    p->len = total;
    p->host = perm;
    p->dperm = (uint32_t*)ctx->alloc_bytes(total * 4);
    HIPCHECK(hipMemcpyAsync(p->dperm, perm.data(), total * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    *out = p;
    CSD_CATCH
}

extern "C" int csd_perm_plan_free(csd_perm_plan* plan) {
    CSD_TRY if (!plan) return 0;
    plan->ctx->free_bytes(plan->dperm);
    delete plan;
    CSD_CATCH
}

struct PermArgs {
    int64_t d0, d1, d2, d3;
    int64_t sd[4], ss[4];
    int64_t* pd;
    const int64_t* ps;
    const uint32_t* perm;
    int dim;
    int64_t chunks;
};

__global__ __launch_bounds__(256) void k_permute(PermArgs a) {
    int64_t row = blockIdx.x / a.chunks;
    int64_t c = blockIdx.x - row * a.chunks;
    int64_t k = row % a.d2, t = row / a.d2;
    int64_t j = t % a.d1, i = t / a.d1;
    int64_t l = c * 256 + threadIdx.x;
    if (l >= a.d3) return;
    int64_t sk = a.dim == 2 ? (int64_t)a.perm[k] : k;
    int64_t sl = a.dim == 3 ? (int64_t)a.perm[l] : l;
    a.pd[i * a.sd[0] + j * a.sd[1] + k * a.sd[2] + l * a.sd[3]] = a.ps[i * a.ss[0] + j * a.ss[1] + sk * a.ss[2] + sl * a.ss[3]];
}

extern "C" int csd_hal_permute(csd_ctx* ctx, const csd_view* dst, const csd_view* src, const csd_perm_plan* plan, int dim) {
    CSD_TRY check_view(dst, "permute");
    check_view(src, "permute src");
    for (int d = 0; d < 4; d++)
        if (dst->shape[d] != src->shape[d]) throw CsdError("permute: shape mismatch");
    if (dim != 2 && dim != 3) throw CsdError("permute: dim must be 2 or 3");
    if ((int64_t)plan->len != dst->shape[dim]) throw CsdError("permute: permutation length mismatch");
    PermArgs a;
    a.d0 = dst->shape[0];
    a.d1 = dst->shape[1];
    a.d2 = dst->shape[2];
    a.d3 = dst->shape[3];
    for (int d = 0; d < 4; d++) {
        a.sd[d] = dst->shape[d] == 1 ? 0 : dst->strides[d];
        a.ss[d] = src->shape[d] == 1 ? 0 : src->strides[d];
    }
    a.pd = (int64_t*)dst->ptr;
    a.ps = (const int64_t*)src->ptr;
    a.perm = plan->dperm;
    a.dim = dim;
    if (numel(dst) == 0) return 0;
    a.chunks = (a.d3 + 255) / 256;
    int64_t blocks = a.d0 * a.d1 * a.d2 * a.chunks;
    hipLaunchKernelGGL(k_permute, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a);
    HIPCHECK(hipGetLastError());
    CSD_CATCH
}

// ---------------------------------------------------------------------------------------------
// calibration kernels: the two ceilings the hot-path kernels are measured against

__global__ __launch_bounds__(256) void k_calib_copy(const int4* __restrict__ src, int4* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// `iters` dependent Barrett multiplications per element (the NTT kernels' folded form, `b`
// folded), 4 independent chains per thread
__global__ __launch_bounds__(256) void k_calib_barrett(int64_t* data, CsdBarrett b, int64_t n, int iters) {
    size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    int64_t x0 = data[i], x1 = data[i + 1], x2 = data[i + 2], x3 = data[i + 3];
    for (int k = 0; k < iters; k++) {
        x0 = csd_barrett_folded(b, x0, n);
        x1 = csd_barrett_folded(b, x1, n);
        x2 = csd_barrett_folded(b, x2, n);
        x3 = csd_barrett_folded(b, x3, n);
    }
    data[i] = x0;
    data[i + 1] = x1;
    data[i + 2] = x2;
    data[i + 3] = x3;
}

// kind 0: device-to-device copy of `n` bytes, result in GB/s (read + write bytes / time);
// kind 1: Barrett multiplications, `n` elements x 64 iterations, result in G multiplications/s
extern "C" int csd_calibrate(csd_ctx* ctx, int kind, size_t n, double* out) {
    CSD_TRY hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    float ms = 0;
    if (kind == 0) {
        size_t words = n / 16;
        void* a = ctx->alloc_bytes(words * 16);
        void* b = ctx->alloc_bytes(words * 16);
        csd_zero_fill(ctx, a, words * 16);
        unsigned blocks = 256 * 64;
        hipLaunchKernelGGL(k_calib_copy, dim3(blocks), dim3(256), 0, ctx->stream, (const int4*)a, (int4*)b, words);
        HIPCHECK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < 5; r++)
            hipLaunchKernelGGL(k_calib_copy, dim3(blocks), dim3(256), 0, ctx->stream, (const int4*)a, (int4*)b, words);
        HIPCHECK(hipEventRecord(e1, ctx->stream));
        HIPCHECK(hipEventSynchronize(e1));
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        *out = 5.0 * 2.0 * (double)(words * 16) / (ms / 1000.0) / 1e9;
        ctx->free_bytes(a);
        ctx->free_bytes(b);
    } else if (kind == 1) {
        size_t elems = (n + 1023) / 1024 * 1024;
        void* a = ctx->alloc_bytes(elems * 8);
        csd_zero_fill(ctx, a, elems * 8);
        const int64_t N = 144115176243904321;
        CsdBarrett bm = csd_barrett_fold(csd_host_barrett(123456789123, N), N);
        unsigned blocks = (unsigned)(elems / 4 / 256);
        hipLaunchKernelGGL(k_calib_barrett, dim3(blocks), dim3(256), 0, ctx->stream, (int64_t*)a, bm, N, 8);
        HIPCHECK(hipEventRecord(e0, ctx->stream));
        hipLaunchKernelGGL(k_calib_barrett, dim3(blocks), dim3(256), 0, ctx->stream, (int64_t*)a, bm, N, 64);
        HIPCHECK(hipEventRecord(e1, ctx->stream));
        HIPCHECK(hipEventSynchronize(e1));
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        *out = (double)elems * 64.0 / (ms / 1000.0) / 1e9;
        ctx->free_bytes(a);
    } else {
        throw CsdError("unknown calibration kind");
    }
    HIPCHECK(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CSD_CATCH
}
