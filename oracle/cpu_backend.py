"""ORACLE — test infrastructure only (imported by tests/ and by bench.py's cpu_baseline leg).

A ctypes front end over ``liboracle.so`` (oracle_hal.cpp), the CPU restatement of the reference
``hal::cpu::CpuBackend`` (cyclosubfield-dbfv/src/hal/cpu.rs). Method names and argument meaning
follow the ``hal::Backend`` trait (cyclosubfield-dbfv/src/hal/mod.rs:108-350); tensors are numpy
arrays (``int64`` for i64 tensors, the 16-byte ``I128`` structured dtype for i128 tensors), and
views are ordinary numpy slices, so slicing/splitting follows cpu.rs:237-296.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# little-endian two's-complement i128 laid out as (low u64, high i64)
I128 = np.dtype([("lo", "<u8"), ("hi", "<i8")])


class OrcView(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("shape", ctypes.c_int64 * 4), ("strides", ctypes.c_int64 * 4)]


class OrcModulus(ctypes.Structure):
    """MontgomeryModulus (hal/mont.rs:4-10)."""

    _fields_ = [("N", ctypes.c_int64), ("neg_inv_N", ctypes.c_int64), ("r2", ctypes.c_int64), ("r3", ctypes.c_int64)]

    def modulus(self):
        return self.N


def build(force=False):
    """Compiles liboracle.so with the Makefile next to this file."""
    src = os.path.join(HERE, "oracle_hal.cpp")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-B", "-C", HERE, "liboracle.so"], check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_last_error.restype = ctypes.c_char_p
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def view_of(arr):
    assert arr.ndim == 4, "tensors are 4-dimensional"
    v = OrcView()
    v.ptr = arr.ctypes.data
    for d in range(4):
        v.shape[d] = arr.shape[d]
        v.strides[d] = arr.strides[d] // arr.itemsize
    return v


def i128_from_ints(values, shape):
    """Builds an I128 tensor from Python ints (wrapped mod 2^128)."""
    flat = np.zeros(int(np.prod(shape)), dtype=I128)
    for pos, x in enumerate(values):
        x &= (1 << 128) - 1
        flat[pos] = (x & ((1 << 64) - 1), ((x >> 64) ^ (1 << 63)) - (1 << 63))
    return flat.reshape(shape)


def i128_to_ints(arr):
    """The entries of an I128 tensor as Python ints, in row-major order."""
    a = np.ascontiguousarray(arr).reshape(-1)
    return [(int(h) << 64) | int(lo) for lo, h in zip(a["lo"].tolist(), a["hi"].tolist())]


class CpuNTTPlan:
    def __init__(self, modulus, inner_shape):
        shape = (ctypes.c_size_t * max(1, len(inner_shape)))(*inner_shape)
        handle = ctypes.c_void_p()
        _check(lib().orc_create_ntt_plan(ctypes.byref(modulus), shape, ctypes.c_size_t(len(inner_shape)), ctypes.byref(handle)))
        self.handle = handle
        self.len = int(np.prod(inner_shape)) if len(inner_shape) else 1
        self.modulus = modulus

    def __del__(self):
        if _lib is not None and self.handle:
            _lib.orc_free_ntt_plan(self.handle)
            self.handle = None


class CpuPermutationPlan:
    def __init__(self, perm):
        self.perm = np.ascontiguousarray(perm, dtype=np.uint32)
        self.len = len(self.perm)


def _mods(mods):
    arr = (OrcModulus * len(mods))(*mods)
    return arr, ctypes.c_size_t(len(mods))


def _plans(plans):
    arr = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
    return arr, ctypes.c_size_t(len(plans))


_DIM = {"Dim2": 2, "Dim3": 3, 2: 2, 3: 3}


class CpuBackend:
    """The oracle backend; see the module docstring."""

    # --- allocation and host transfer (hal/mod.rs:115-125) -------------------------------------
    def hal_allocate_i64(self, shape):
        return np.zeros(tuple(shape), dtype=np.int64)

    def hal_allocate_i128(self, shape):
        return np.zeros(tuple(shape), dtype=I128)

    def hal_import(self, dst, src):
        src = np.ascontiguousarray(np.asarray(src, dtype=np.int64).reshape(-1))
        _check(lib().orc_hal_import(ctypes.byref(view_of(dst)), src.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(src.size)))

    def hal_export(self, src, dst=None):
        out = np.zeros(int(np.prod(src.shape)), dtype=np.int64)
        _check(lib().orc_hal_export(ctypes.byref(view_of(src)), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(out.size)))
        if dst is not None:
            dst[:] = out
        return out

    # --- copies and wrapping arithmetic (hal/mod.rs:127-173) -----------------------------------
    def _bin(self, name, dst, src):
        _check(getattr(lib(), name)(ctypes.byref(view_of(dst)), ctypes.byref(view_of(src))))

    def hal_assign(self, dst, src):
        self._bin("orc_hal_assign", dst, src)

    def hal_assign_narrow(self, dst, src):
        self._bin("orc_hal_assign_narrow", dst, src)

    def hal_assign_i128(self, dst, src):
        self._bin("orc_hal_assign_i128", dst, src)

    def hal_add_i64(self, dst, src):
        self._bin("orc_hal_add_i64", dst, src)

    def hal_sub_i64(self, dst, src):
        self._bin("orc_hal_sub_i64", dst, src)

    def hal_add_i128(self, dst, src):
        self._bin("orc_hal_add_i128", dst, src)

    def hal_fma_i64(self, dst, lhs, rhs):
        _check(lib().orc_hal_fma_i64(ctypes.byref(view_of(dst)), ctypes.byref(view_of(lhs)), ctypes.byref(view_of(rhs))))

    def hal_mul_i64(self, dst, lhs, rhs):
        _check(lib().orc_hal_mul_i64(ctypes.byref(view_of(dst)), ctypes.byref(view_of(lhs)), ctypes.byref(view_of(rhs))))

    def hal_inner_product_i64(self, dst, lhs, rhs):
        assert len(lhs) == len(rhs)
        lv = (OrcView * len(lhs))(*[view_of(x) for x in lhs])
        rv = (OrcView * len(rhs))(*[view_of(x) for x in rhs])
        _check(lib().orc_hal_inner_product_i64(ctypes.byref(view_of(dst)), lv, rv, ctypes.c_size_t(len(lhs))))

    def hal_matmul_i64(self, dst, lhs, rhs):
        _check(lib().orc_hal_matmul_i64(ctypes.byref(view_of(dst)), ctypes.byref(view_of(lhs)), ctypes.byref(view_of(rhs))))

    # --- Montgomery arithmetic (hal/mod.rs:175-207) ---------------------------------------------
    def hal_create_modulus(self, prime):
        m = OrcModulus()
        _check(lib().orc_create_modulus(ctypes.c_int64(prime), ctypes.byref(m)))
        return m

    def _modop(self, name, dst, src, mods):
        arr, n = _mods(mods)
        _check(getattr(lib(), name)(ctypes.byref(view_of(dst)), ctypes.byref(view_of(src)), arr, n))

    def hal_mul_mont64(self, dst, src, mods):
        self._modop("orc_hal_mul_mont64", dst, src, mods)

    def hal_add_red_mont64_double(self, dst, src, mods):
        self._modop("orc_hal_add_red_mont64_double", dst, src, mods)

    def hal_red_mont64_double(self, dst, src, mods):
        self._modop("orc_hal_red_mont64_double", dst, src, mods)

    def hal_add_red_mont64(self, dst, src, mods):
        self._modop("orc_hal_add_red_mont64", dst, src, mods)

    def hal_red_mont64(self, dst, src, mods):
        self._modop("orc_hal_red_mont64", dst, src, mods)

    def hal_i128_to_mont64(self, dst, src, mods):
        self._modop("orc_hal_i128_to_mont64", dst, src, mods)

    def hal_i64_to_mont64(self, dst, mods):
        arr, n = _mods(mods)
        _check(lib().orc_hal_i64_to_mont64(ctypes.byref(view_of(dst)), arr, n))

    # --- NTTs (hal/mod.rs:209-250) --------------------------------------------------------------
    def hal_create_ntt_plan(self, modulus, inner_shape):
        return CpuNTTPlan(modulus, list(inner_shape))

    def hal_ntt_mont64_dim(self, dst, plans, dim):
        arr, n = _plans(plans)
        _check(lib().orc_hal_ntt_mont64_dim(ctypes.byref(view_of(dst)), arr, n, _DIM[dim]))

    def hal_intt_mont64_dim(self, dst, plans, dim):
        arr, n = _plans(plans)
        _check(lib().orc_hal_intt_mont64_dim(ctypes.byref(view_of(dst)), arr, n, _DIM[dim]))

    def hal_ntt_mul_mont64_dim(self, dst, plans, dim, mul, mods):
        arr, n = _plans(plans)
        marr, mn = _mods(mods)
        _check(lib().orc_hal_ntt_mul_mont64_dim(ctypes.byref(view_of(dst)), arr, n, _DIM[dim], ctypes.byref(view_of(mul)), marr, mn))

    def hal_ntt_mont64(self, dst, dim2, dim3):
        self.hal_ntt_mont64_dim(dst, dim2, 2)
        self.hal_ntt_mont64_dim(dst, dim3, 3)

    def hal_intt_mont64(self, dst, dim2, dim3):
        self.hal_intt_mont64_dim(dst, dim2, 2)
        self.hal_intt_mont64_dim(dst, dim3, 3)

    # --- base-conversion corrections (hal/mod.rs:252-337) ---------------------------------------
    def hal_bconv_correct_mont64(self, dst, lhs, rhs, mods, shift):
        arr, n = _mods(mods)
        _check(lib().orc_hal_bconv_correct_mont64(ctypes.byref(view_of(dst)), ctypes.byref(view_of(lhs)),
                                                  ctypes.byref(view_of(rhs)), arr, n, ctypes.c_size_t(shift)))

    def hal_bconv_correct_pow2(self, dst, src, factor0, factor1, shift, log2_modulus):
        _check(lib().orc_hal_bconv_correct_pow2(ctypes.byref(view_of(dst)), ctypes.byref(view_of(src)),
                                                ctypes.c_int64(factor0), ctypes.c_int64(factor1),
                                                ctypes.c_size_t(shift), ctypes.c_size_t(log2_modulus)))

    def hal_bconv_correct_pow2_wide(self, dst, src, factor0, factor1, shift, log2_modulus):
        def split(x):
            x &= (1 << 128) - 1
            lo = x & ((1 << 64) - 1)
            hi = x >> 64
            return ctypes.c_int64(lo - (1 << 64) if lo >= 1 << 63 else lo), ctypes.c_int64(hi - (1 << 64) if hi >= 1 << 63 else hi)

        f0l, f0h = split(factor0)
        f1l, f1h = split(factor1)
        _check(lib().orc_hal_bconv_correct_pow2_wide(ctypes.byref(view_of(dst)), ctypes.byref(view_of(src)),
                                                     f0l, f0h, f1l, f1h, ctypes.c_size_t(shift), ctypes.c_size_t(log2_modulus)))

    # --- permutations (hal/mod.rs:339-349) ------------------------------------------------------
    def hal_create_permutation(self, inner_shape, shifts):
        n = len(inner_shape)
        total = int(np.prod(inner_shape)) if n else 1
        perm = np.zeros(total, dtype=np.uint32)
        shp = (ctypes.c_size_t * max(1, n))(*inner_shape)
        sh = (ctypes.c_size_t * max(1, n))(*shifts)
        _check(lib().orc_hal_create_permutation(shp, sh, ctypes.c_size_t(n), perm.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(total)))
        return CpuPermutationPlan(perm)

    def hal_permute(self, dst, src, plan, dim):
        _check(lib().orc_hal_permute(ctypes.byref(view_of(dst)), ctypes.byref(view_of(src)),
                                     plan.perm.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(plan.len), _DIM[dim]))


# --- scalar entry points for the mont.rs / ntt.rs unit tests ------------------------------------
def red128(m, x):
    out = ctypes.c_int64()
    x &= (1 << 128) - 1
    lo, hi = x & ((1 << 64) - 1), x >> 64
    _check(lib().orc_red128(ctypes.byref(m), ctypes.c_int64(lo - (1 << 64) if lo >= 1 << 63 else lo),
                            ctypes.c_int64(hi - (1 << 64) if hi >= 1 << 63 else hi), ctypes.byref(out)))
    return out.value


def double_red128(m, x):
    out = ctypes.c_int64()
    x &= (1 << 128) - 1
    lo, hi = x & ((1 << 64) - 1), x >> 64
    _check(lib().orc_double_red128(ctypes.byref(m), ctypes.c_int64(lo - (1 << 64) if lo >= 1 << 63 else lo),
                                   ctypes.c_int64(hi - (1 << 64) if hi >= 1 << 63 else hi), ctypes.byref(out)))
    return out.value


def barrett_new(value, n):
    z, u = ctypes.c_int64(), ctypes.c_int64()
    _check(lib().orc_barrett_new(ctypes.c_int64(value), ctypes.c_int64(n), ctypes.byref(z), ctypes.byref(u)))
    return z.value, u.value


def barrett_mul(zu, x, n):
    out = ctypes.c_int64()
    _check(lib().orc_barrett_mul(ctypes.c_int64(zu[0]), ctypes.c_int64(zu[1]), ctypes.c_int64(x), ctypes.c_int64(n), ctypes.byref(out)))
    return out.value


def find_root_of_unity(q, n):
    out = ctypes.c_int64()
    _check(lib().orc_find_root_of_unity(ctypes.c_int64(q), ctypes.c_size_t(n), ctypes.byref(out)))
    return out.value


def factor_forward(q, length, data, es=1, vl=1, in_bound=None, reduce_output=False):
    """LazyFactorNTT::forward on an int64 numpy buffer (in place); returns the output bound."""
    assert data.dtype == np.int64 and data.flags["C_CONTIGUOUS"]
    out = ctypes.c_uint64()
    ib = 2 * q if in_bound is None else in_bound
    _check(lib().orc_factor_forward(ctypes.c_int64(q), ctypes.c_size_t(length), data.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_size_t(es), ctypes.c_size_t(vl), ctypes.c_uint64(ib),
                                    ctypes.c_int(1 if reduce_output else 0), ctypes.byref(out)))
    return out.value


def factor_backward(q, length, data, es=1, vl=1, in_bound=None):
    assert data.dtype == np.int64 and data.flags["C_CONTIGUOUS"]
    out = ctypes.c_uint64()
    ib = 2 * q if in_bound is None else in_bound
    _check(lib().orc_factor_backward(ctypes.c_int64(q), ctypes.c_size_t(length), data.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_size_t(es), ctypes.c_size_t(vl), ctypes.c_uint64(ib), ctypes.byref(out)))
    return out.value


def plan_apply(plan, data, es=1, vl=1, inverse=False, reduce_output=True):
    assert data.dtype == np.int64 and data.flags["C_CONTIGUOUS"]
    _check(lib().orc_plan_apply(plan.handle, data.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(es),
                                ctypes.c_size_t(vl), ctypes.c_int(1 if inverse else 0), ctypes.c_int(1 if reduce_output else 0)))


def plan_apply_traced(plan, data, es=1, vl=1, inverse=False, reduce_output=True):
    """plan_apply with the oracle's schedule trace on. Returns (events, barrett_mults): events are
    dicts with kind ("reduce", "fwd", "inv"), factor, stage, skip_ok, skipped (blocks whose
    twiddle multiplication was skipped) and fused, in execution order, recorded once per factor
    (its first line); barrett_mults counts every Barrett multiplication of the whole apply."""
    _check(lib().orc_ntt_trace(ctypes.c_int(1)))
    try:
        plan_apply(plan, data, es, vl, inverse, reduce_output)
        cap = 4096
        buf = (ctypes.c_int32 * (6 * cap))()
        count, mults = ctypes.c_size_t(), ctypes.c_uint64()
        _check(lib().orc_ntt_trace_get(buf, ctypes.c_size_t(cap), ctypes.byref(count), ctypes.byref(mults)))
    finally:
        _check(lib().orc_ntt_trace(ctypes.c_int(0)))
    kinds = {0: "reduce", 1: "fwd", 2: "inv"}
    events = []
    for i in range(count.value):
        k, f, s, ok, sk, fu = buf[6 * i:6 * i + 6]
        events.append(dict(kind=kinds[k], factor=f, stage=s, skip_ok=bool(ok), skipped=sk, fused=bool(fu)))
    return events, mults.value
