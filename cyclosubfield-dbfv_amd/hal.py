"""Python front end of libcsdbfv_hal.so: the MI355X backend for the ``hal::Backend`` trait
(cyclosubfield-dbfv/src/hal/mod.rs:108-350).

``HipBackend`` exposes the trait methods under their reference names (``hal_allocate_i64``,
``hal_ntt_mul_mont64_dim``, ...). Tensors are ``DeviceTensor`` views over HBM with the reference's
view semantics (hal/cpu.rs:187-296): slicing with step-1 ranges gives a strided view, ``split``
gives two disjoint views, the last stride is always 1. Errors of the C ABI (contract violations,
HIP failures) surface as ``HalError`` — the reference panics in the same places.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcsdbfv_hal.so")

I128 = np.dtype([("lo", "<u8"), ("hi", "<i8")])


class HalError(RuntimeError):
    pass


class CsdView(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("shape", ctypes.c_int64 * 4), ("strides", ctypes.c_int64 * 4)]


class CsdModulus(ctypes.Structure):
    """MontgomeryModulus (hal/mont.rs:4-10)."""

    _fields_ = [("N", ctypes.c_int64), ("neg_inv_N", ctypes.c_int64), ("r2", ctypes.c_int64), ("r3", ctypes.c_int64)]

    def modulus(self):
        return self.N


_lib = None


def lib():
    """Loads the native library; there is no fallback: without it nothing in this package runs."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HalError(f"{LIB_PATH} is not built (run cyclosubfield-dbfv_amd/build_lib.py)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.csd_last_error.restype = ctypes.c_char_p
    return _lib


def check(rc):
    if rc != 0:
        raise HalError(lib().csd_last_error().decode())


def _norm_slice(s, n):
    if isinstance(s, int):
        if s < 0:
            s += n
        if not 0 <= s < n:
            raise IndexError("index out of range")
        return s, s + 1
    assert isinstance(s, slice) and s.step in (None, 1), "only step-1 slices are views"
    start, stop, _ = s.indices(n)
    if stop < start:
        stop = start
    return start, stop


class DeviceTensor:
    """A 4-D strided view over device memory (i64 or i128 entries)."""

    def __init__(self, backend, ptr, shape, strides, elem, owner=None):
        self.backend = backend
        self.ptr = ptr
        self.shape = tuple(int(x) for x in shape)
        self.strides = tuple(int(x) for x in strides)
        self.elem = elem
        self._owner = owner  # keeps the allocation alive while views exist

    @property
    def ndim(self):
        return 4

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        idx = tuple(i for i in idx if i is not Ellipsis)
        idx = idx + (slice(None),) * (4 - len(idx))
        off = 0
        shape = []
        for d in range(4):
            a, b = _norm_slice(idx[d], self.shape[d])
            off += a * self.strides[d]
            shape.append(b - a)
        return DeviceTensor(self.backend, self.ptr + off * self.elem, shape, self.strides, self.elem, self._owner or self)

    def slice(self, r0, r1, r2, r3):
        return self[r0, r1, r2, r3]

    def split(self, at, dim):
        """TensorMut::split (hal/cpu.rs:267-284)."""
        assert 0 <= at <= self.shape[dim]
        s1 = list(self.shape)
        s1[dim] = at
        s2 = list(self.shape)
        s2[dim] = self.shape[dim] - at
        first = DeviceTensor(self.backend, self.ptr, s1, self.strides, self.elem, self._owner or self)
        second = DeviceTensor(self.backend, self.ptr + at * self.strides[dim] * self.elem, s2, self.strides, self.elem,
                              self._owner or self)
        return first, second

    def view(self):
        v = CsdView()
        v.ptr = self.ptr
        for d in range(4):
            v.shape[d] = self.shape[d]
            v.strides[d] = self.strides[d]
        return v

    def numel(self):
        return int(np.prod(self.shape))

    def __repr__(self):
        return f"DeviceTensor(shape={self.shape}, strides={self.strides}, {'i64' if self.elem == 8 else 'i128'})"


class _Allocation(DeviceTensor):
    def __del__(self):
        be = self.backend
        if be is not None and be.ctx is not None and _lib is not None:
            _lib.csd_hal_free(be.ctx, ctypes.c_void_p(self.ptr))


class HipNTTPlan:
    def __init__(self, backend, handle, length):
        self.backend, self.handle, self.len = backend, handle, length

    def mults(self, direction):
        """Modular multiplications of one transform: direction 'fwd', 'fwd_unreduced' or 'inv'."""
        out = ctypes.c_uint64()
        d = {"fwd": 0, "fwd_unreduced": 1, "inv": 2}[direction]
        check(lib().csd_ntt_plan_mults(ctypes.c_void_p(self.handle), ctypes.c_int(d), ctypes.byref(out)))
        return out.value

    def __del__(self):
        if self.handle is not None and _lib is not None and self.backend.ctx is not None:
            _lib.csd_ntt_plan_free(ctypes.c_void_p(self.handle))
            self.handle = None


class HipPermutationPlan:
    def __init__(self, backend, handle, length):
        self.backend, self.handle, self.len = backend, handle, length

    def __del__(self):
        if self.handle is not None and _lib is not None and self.backend.ctx is not None:
            _lib.csd_perm_plan_free(ctypes.c_void_p(self.handle))
            self.handle = None


class PinnedHostTensor:
    """A row-major int64 array in pinned host memory (csd_host_alloc), as a numpy view."""

    def __init__(self, backend, shape):
        self.backend = backend
        self.shape = tuple(int(x) for x in shape)
        n = int(np.prod(self.shape))
        p = ctypes.c_void_p()
        check(lib().csd_host_alloc(backend.ctx, ctypes.c_size_t(8 * max(n, 1)), ctypes.byref(p)))
        self.ptr = p.value
        self.array = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_int64)), shape=(max(n, 1),))[:n].reshape(self.shape)

    def fill_from(self, src):
        """device -> pinned host copy of a device tensor with the same number of entries"""
        n = int(np.prod(self.shape))
        check(lib().csd_hal_export(self.backend.ctx, ctypes.byref(src.view()), ctypes.c_void_p(self.ptr), ctypes.c_size_t(n)))

    def free(self):
        if self.ptr is not None and _lib is not None and self.backend.ctx is not None:
            self.array = None
            _lib.csd_host_free(self.backend.ctx, ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ZpMatrix:
    """r dense matrices over Z_p (one per modulus) with their int8 digit planes split once
    (csd_zp_matrix_create)."""

    def __init__(self, backend, handle, shape):
        self.backend, self.handle, self.shape = backend, handle, shape

    def __del__(self):
        if self.handle is not None and _lib is not None and self.backend.ctx is not None:
            _lib.csd_zp_matrix_free(ctypes.c_void_p(self.handle))
            self.handle = None

    def set_circulant(self, orders):
        """Declares the matrices group-circulant over C_o0 x ... x C_o(f-1) (last order 64):
        entry [a][b] = entry [a - b][0]. Checked on the device (an error if any entry differs);
        the default MFMA kernel then builds its A operands from the first column's digit planes
        (csd_zp_matrix_set_circulant). Results are unchanged."""
        arr = (ctypes.c_size_t * len(orders))(*[int(o) for o in orders])
        check(lib().csd_zp_matrix_set_circulant(ctypes.c_void_p(self.handle), arr, ctypes.c_size_t(len(orders))))


def _row_major_strides(shape):
    st = [1, 1, 1, 1]
    for d in (2, 1, 0):
        st[d] = st[d + 1] * shape[d + 1]
    return st


_DIM = {"Dim2": 2, "Dim3": 3, 2: 2, 3: 3}


def _mods(mods):
    return (CsdModulus * len(mods))(*mods), ctypes.c_size_t(len(mods))


def _plans(plans):
    return (ctypes.c_void_p * len(plans))(*[p.handle for p in plans]), ctypes.c_size_t(len(plans))


class HipBackend:
    """The hal::Backend trait on one MI355X, see the module docstring."""

    def __init__(self, device=0):
        self.ctx = None
        ctx = ctypes.c_void_p()
        check(lib().csd_ctx_create(ctypes.c_int(device), ctypes.byref(ctx)))
        self.ctx = ctx
        self.device = device

    def close(self):
        if self.ctx is not None:
            lib().csd_ctx_destroy(self.ctx)
            self.ctx = None

    def sync(self):
        check(lib().csd_ctx_sync(self.ctx))

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().csd_ctx_stream(self.ctx, ctypes.byref(s)))
        return s.value

    def ntt_set_row_shape(self, lines=0, threads=0):
        """launch shape of this context's dimension-3 transforms: lines and threads per workgroup,
        0 for the automatic choice; every shape computes identical results"""
        check(lib().csd_ntt_set_row_shape(self.ctx, ctypes.c_int(lines), ctypes.c_int(threads)))

    def ntt_force_generic(self, mode):
        """NTT kernel selection of this context (csd_ntt_force_generic): 0 the register-blocked
        kernels with compiled shapes, 1 the one-stage-per-pass kernel, 2 the plan-driven row
        kernel; all compute identical results"""
        check(lib().csd_ntt_force_generic(self.ctx, ctypes.c_int(int(mode))))

    def set_timing(self, on):
        """Records HIP events around the library's timed launches on the context stream."""
        check(lib().csd_ctx_set_timing(self.ctx, ctypes.c_int(1 if on else 0)))

    def calibrate(self, kind, n):
        """kind 'copy': device copy of n bytes -> GB/s; kind 'barrett': n elements x 64 dependent
        Barrett multiplications -> G multiplications/s."""
        out = ctypes.c_double()
        check(lib().csd_calibrate(self.ctx, ctypes.c_int({"copy": 0, "barrett": 1}[kind]), ctypes.c_size_t(n), ctypes.byref(out)))
        return out.value

    def timing_report(self):
        """{kernel name: (launches, total ms, total algorithmic bytes)}; resets the record."""
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().csd_ctx_timing_report(self.ctx, buf, ctypes.c_size_t(len(buf))))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms, by = line.split()
            out[name] = (int(n), float(ms), float(by))
        return out

    # --- allocation and transfer (hal/mod.rs:115-125) --------------------------------------------
    def _alloc(self, shape, elem, zero=True):
        shape = [int(x) for x in shape]
        ptr = ctypes.c_void_p()
        fn = lib().csd_hal_allocate if zero else lib().csd_hal_allocate_uninit
        check(fn(self.ctx, ctypes.c_int(elem), (ctypes.c_int64 * 4)(*shape), ctypes.byref(ptr)))
        return _Allocation(self, ptr.value, shape, _row_major_strides(shape), elem)

    def allocate_i64_uninit(self, shape):
        """an i64 tensor without the zero fill, for outputs that the next call writes entirely"""
        return self._alloc(shape, 8, zero=False)

    def allocate_i128_uninit(self, shape):
        """an i128 tensor without the zero fill, for accumulators whose rows are written before
        they are read"""
        return self._alloc(shape, 16, zero=False)

    def hal_allocate_i64(self, shape):
        return self._alloc(shape, 8)

    def hal_allocate_i128(self, shape):
        return self._alloc(shape, 16)

    def hal_import(self, dst, src):
        assert dst.elem == 8
        a = np.ascontiguousarray(np.asarray(src, dtype=np.int64).reshape(-1))
        check(lib().csd_hal_import(self.ctx, ctypes.byref(dst.view()), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.size)))

    def hal_export(self, src, dst=None):
        assert src.elem == 8
        out = np.zeros(src.numel(), dtype=np.int64)
        check(lib().csd_hal_export(self.ctx, ctypes.byref(src.view()), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(out.size)))
        if dst is not None:
            dst[:] = out
        return out

    def import_i128(self, dst, words):
        """words: I128 numpy array (or int64 pairs) in row-major order of the view."""
        assert dst.elem == 16
        a = np.ascontiguousarray(words).view(np.int64).reshape(-1)
        check(lib().csd_import_i128(self.ctx, ctypes.byref(dst.view()), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.size // 2)))

    def export_i128(self, src):
        assert src.elem == 16
        out = np.zeros(src.numel(), dtype=I128)
        check(lib().csd_export_i128(self.ctx, ctypes.byref(src.view()), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(out.size)))
        return out

    # --- copies and wrapping arithmetic (hal/mod.rs:127-173) --------------------------------------
    def _bin(self, fn, dst, src):
        check(getattr(lib(), fn)(self.ctx, ctypes.byref(dst.view()), ctypes.byref(src.view())))

    def hal_assign(self, dst, src):
        self._bin("csd_hal_assign", dst, src)

    def hal_assign_narrow(self, dst, src):
        self._bin("csd_hal_assign_narrow", dst, src)

    def hal_assign_i128(self, dst, src):
        self._bin("csd_hal_assign_i128", dst, src)

    def hal_add_i64(self, dst, src):
        self._bin("csd_hal_add_i64", dst, src)

    def hal_sub_i64(self, dst, src):
        self._bin("csd_hal_sub_i64", dst, src)

    def hal_add_i128(self, dst, src):
        self._bin("csd_hal_add_i128", dst, src)

    def hal_fma_i64(self, dst, lhs, rhs):
        check(lib().csd_hal_fma_i64(self.ctx, ctypes.byref(dst.view()), ctypes.byref(lhs.view()), ctypes.byref(rhs.view())))

    def sum_i64(self, dst, lhs, rhs):
        """csd_hal_sum_i64: dst = lhs + rhs (wrapping), hal_assign and hal_add_i64 in one pass"""
        check(lib().csd_hal_sum_i64(self.ctx, ctypes.byref(dst.view()), ctypes.byref(lhs.view()), ctypes.byref(rhs.view())))

    def hal_mul_i64(self, dst, lhs, rhs):
        check(lib().csd_hal_mul_i64(self.ctx, ctypes.byref(dst.view()), ctypes.byref(lhs.view()), ctypes.byref(rhs.view())))

    def hal_inner_product_i64(self, dst, lhs, rhs):
        assert len(lhs) == len(rhs)
        lv = (CsdView * len(lhs))(*[x.view() for x in lhs])
        rv = (CsdView * len(rhs))(*[x.view() for x in rhs])
        check(lib().csd_hal_inner_product_i64(self.ctx, ctypes.byref(dst.view()), lv, rv, ctypes.c_size_t(len(lhs))))

    def convolve_reduce(self, dst, lhs, lhs_rows, rhs, rhs_rows, count, d1, d2, K, full, mods):
        """csd_convolve_reduce: the digit convolution (or one-digit tensor product) and its wide
        reduction as one kernel; `full` holds the reduction kind of each of the 3 K output rows."""
        lr = (ctypes.c_int64 * 3)(*[int(x) for x in lhs_rows])
        rr = (ctypes.c_int64 * 3)(*[int(x) for x in rhs_rows])
        fl = (ctypes.c_uint8 * len(full))(*[1 if f else 0 for f in full])
        arr, n = _mods(mods)
        check(lib().csd_convolve_reduce(self.ctx, ctypes.byref(dst.view()), ctypes.byref(lhs.view()), lr,
                                        ctypes.byref(rhs.view()), rr, ctypes.c_size_t(count), ctypes.c_size_t(d1),
                                        ctypes.c_size_t(d2), ctypes.c_size_t(K), fl, arr, n))

    def keyswitch_reduce(self, out, out_rows, parts, K, d, keys, add, one, full, mods):
        """csd_keyswitch_reduce: relin's key-switch inner products for both key rows (plus c0, c1
        times `one` when `add` is given) and their wide reduction as one kernel."""
        rows = (ctypes.c_int64 * 2)(*[int(x) for x in out_rows])
        kv = (CsdView * len(keys))(*[k.view() for k in keys])
        arr, n = _mods(mods)
        check(lib().csd_keyswitch_reduce(self.ctx, ctypes.byref(out.view()), rows, ctypes.byref(parts.view()),
                                         ctypes.c_size_t(K), ctypes.c_size_t(d), kv,
                                         ctypes.byref(add.view()) if add is not None else None,
                                         ctypes.byref(one.view()) if one is not None else None,
                                         ctypes.c_int(1 if full else 0), arr, n))

    def hal_matmul_i64(self, dst, lhs, rhs):
        check(lib().csd_hal_matmul_i64(self.ctx, ctypes.byref(dst.view()), ctypes.byref(lhs.view()), ctypes.byref(rhs.view())))

    # --- Montgomery arithmetic (hal/mod.rs:175-207) -----------------------------------------------
    def hal_create_modulus(self, prime):
        m = CsdModulus()
        check(lib().csd_hal_create_modulus(ctypes.c_int64(prime), ctypes.byref(m)))
        return m

    def _modop(self, fn, dst, src, mods):
        arr, n = _mods(mods)
        check(getattr(lib(), fn)(self.ctx, ctypes.byref(dst.view()), ctypes.byref(src.view()), arr, n))

    def hal_mul_mont64(self, dst, src, mods):
        self._modop("csd_hal_mul_mont64", dst, src, mods)

    def hal_add_red_mont64_double(self, dst, src, mods):
        self._modop("csd_hal_add_red_mont64_double", dst, src, mods)

    def hal_red_mont64_double(self, dst, src, mods):
        self._modop("csd_hal_red_mont64_double", dst, src, mods)

    def hal_add_red_mont64(self, dst, src, mods):
        self._modop("csd_hal_add_red_mont64", dst, src, mods)

    def hal_red_mont64(self, dst, src, mods):
        self._modop("csd_hal_red_mont64", dst, src, mods)

    def hal_i128_to_mont64(self, dst, src, mods):
        self._modop("csd_hal_i128_to_mont64", dst, src, mods)

    def hal_i64_to_mont64(self, dst, mods):
        arr, n = _mods(mods)
        check(lib().csd_hal_i64_to_mont64(self.ctx, ctypes.byref(dst.view()), arr, n))

    # --- NTTs (hal/mod.rs:209-250) -----------------------------------------------------------------
    def hal_create_ntt_plan(self, modulus, inner_shape):
        inner_shape = list(inner_shape)
        h = ctypes.c_void_p()
        shp = (ctypes.c_size_t * max(1, len(inner_shape)))(*inner_shape)
        check(lib().csd_hal_create_ntt_plan(self.ctx, ctypes.byref(modulus), shp, ctypes.c_size_t(len(inner_shape)), ctypes.byref(h)))
        return HipNTTPlan(self, h.value, int(np.prod(inner_shape)) if inner_shape else 1)

    def hal_ntt_mont64_dim(self, dst, plans, dim):
        arr, n = _plans(plans)
        check(lib().csd_hal_ntt_mont64_dim(self.ctx, ctypes.byref(dst.view()), arr, n, ctypes.c_int(_DIM[dim])))

    def hal_intt_mont64_dim(self, dst, plans, dim):
        arr, n = _plans(plans)
        check(lib().csd_hal_intt_mont64_dim(self.ctx, ctypes.byref(dst.view()), arr, n, ctypes.c_int(_DIM[dim])))

    def hal_ntt_mul_mont64_dim(self, dst, plans, dim, mul, mods):
        arr, n = _plans(plans)
        marr, mn = _mods(mods)
        check(lib().csd_hal_ntt_mul_mont64_dim(self.ctx, ctypes.byref(dst.view()), arr, n, ctypes.c_int(_DIM[dim]),
                                               ctypes.byref(mul.view()), marr, mn))

    def hal_ntt_mont64(self, dst, dim2, dim3):
        self.hal_ntt_mont64_dim(dst, dim2, 2)
        self.hal_ntt_mont64_dim(dst, dim3, 3)

    def hal_intt_mont64(self, dst, dim2, dim3):
        self.hal_intt_mont64_dim(dst, dim2, 2)
        self.hal_intt_mont64_dim(dst, dim3, 3)

    # --- base-conversion corrections (hal/mod.rs:252-337) ------------------------------------------
    def hal_bconv_correct_mont64(self, dst, lhs, rhs, mods, shift):
        arr, n = _mods(mods)
        check(lib().csd_hal_bconv_correct_mont64(self.ctx, ctypes.byref(dst.view()), ctypes.byref(lhs.view()),
                                                 ctypes.byref(rhs.view()), arr, n, ctypes.c_size_t(shift)))

    def hal_bconv_correct_pow2(self, dst, src, factor0, factor1, shift, log2_modulus):
        check(lib().csd_hal_bconv_correct_pow2(self.ctx, ctypes.byref(dst.view()), ctypes.byref(src.view()),
                                               ctypes.c_int64(factor0), ctypes.c_int64(factor1),
                                               ctypes.c_size_t(shift), ctypes.c_size_t(log2_modulus)))

    def hal_bconv_correct_pow2_wide(self, dst, src, factor0, factor1, shift, log2_modulus):
        def split(x):
            x &= (1 << 128) - 1
            lo, hi = x & ((1 << 64) - 1), x >> 64
            s = lambda w: ctypes.c_int64(w - (1 << 64) if w >= 1 << 63 else w)
            return s(lo), s(hi)

        f0l, f0h = split(factor0)
        f1l, f1h = split(factor1)
        check(lib().csd_hal_bconv_correct_pow2_wide(self.ctx, ctypes.byref(dst.view()), ctypes.byref(src.view()),
                                                    f0l, f0h, f1l, f1h, ctypes.c_size_t(shift), ctypes.c_size_t(log2_modulus)))

    # --- permutations (hal/mod.rs:339-349) ---------------------------------------------------------
    def dbfv_split_overflow(self, x, overflow, carry_in, consts, mods, f0, f1, shift, log2_r, full_split, full_carry,
                            x_in=None, zero_in=False):
        """csd_dbfv_split_overflow: split_overflow (dbfv.rs:893-948) in one pass, optionally after
        the carry step of dbfv_reduce; f0, f1 are i128 constants. The digit is read from `x_in`
        (all zero with `zero_in`) when given, else from `x`; the result goes to `x`."""
        def words(v):
            v &= (1 << 128) - 1
            lo, hi = v & ((1 << 64) - 1), v >> 64
            s = lambda w: ctypes.c_int64(w - (1 << 64) if w >= 1 << 63 else w)
            return s(lo), s(hi)

        arr, n = _mods(mods)
        f0l, f0h = words(f0)
        f1l, f1h = words(f1)
        cin = ctypes.byref(carry_in.view()) if carry_in is not None else None
        xin = ctypes.byref(x_in.view()) if x_in is not None else None
        check(lib().csd_dbfv_split_overflow(self.ctx, ctypes.byref(x.view()), xin, ctypes.c_int(1 if zero_in else 0),
                                            ctypes.byref(overflow.view()), cin,
                                            ctypes.byref(consts.view()), arr, n, f0l, f0h, f1l, f1h,
                                            ctypes.c_size_t(shift), ctypes.c_size_t(log2_r),
                                            ctypes.c_int(1 if full_split else 0), ctypes.c_int(1 if full_carry else 0)))

    def dbfv_reduce_chain(self, x, src, per, digits, consts, mods, limbs, f0, f1, shift, log2_r, full_split, full_carry):
        """csd_dbfv_reduce_chain: the carry chain of dbfv_reduce (dbfv.rs:990-1023) in one launch;
        digit i of x is rows [i per, (i + 1) per), read from the same rows of src (zero beyond its
        rows). full_split / full_carry are per-digit lists of the double-reduction choices."""
        def words(v):
            v &= (1 << 128) - 1
            lo, hi = v & ((1 << 64) - 1), v >> 64
            s = lambda w: ctypes.c_int64(w - (1 << 64) if w >= 1 << 63 else w)
            return s(lo), s(hi)

        def mask(flags):
            return ctypes.c_uint32(sum(1 << i for i, f in enumerate(flags) if f))

        arr, n = _mods(mods)
        f0l, f0h = words(f0)
        f1l, f1h = words(f1)
        check(lib().csd_dbfv_reduce_chain(self.ctx, ctypes.byref(x.view()), ctypes.byref(src.view()), ctypes.c_size_t(per),
                                          ctypes.c_size_t(digits), ctypes.byref(consts.view()), arr, n,
                                          ctypes.c_size_t(limbs), f0l, f0h, f1l, f1h, ctypes.c_size_t(shift),
                                          ctypes.c_size_t(log2_r), mask(full_split), mask(full_carry)))

    def range_check(self, t, mods, multiple):
        """csd_hal_range_check: the number of entries of the int64 tensor `t` with |x| > multiple N,
        N the modulus of the entry's limb (a debug check of kernel outputs; synchronizes)"""
        arr, n = _mods(mods)
        out = ctypes.c_uint64()
        check(lib().csd_hal_range_check(self.ctx, ctypes.byref(t.view()), arr, n, ctypes.c_int64(int(multiple)),
                                        ctypes.byref(out)))
        return out.value

    def host_alloc(self, shape):
        """pinned host memory for a host-resident int64 tensor (PinnedHostTensor)"""
        return PinnedHostTensor(self, shape)

    def zp_matrix(self, a, mods):
        """Dense maps over Z_p: a is a [1, r, m, k] tensor with entries in [-2p, 2p]."""
        h = ctypes.c_void_p()
        arr, n = _mods(mods)
        check(lib().csd_zp_matrix_create(self.ctx, ctypes.byref(a.view()), arr, n, ctypes.byref(h)))
        return ZpMatrix(self, h.value, a.shape)

    def zp_apply_dim(self, mat, dst, src, dim, form="mfma", max_chunk=0):
        """dst = mat . src along dim (2 or 3), exact, result in [0, p): form "mfma" (int8 digits on
        the matrix cores, the kernel chosen by zp_set_kernel) or "plain" (64-bit VALU products)."""
        check(lib().csd_zp_apply_dim(self.ctx, ctypes.c_void_p(mat.handle), ctypes.byref(dst.view()),
                                     ctypes.byref(src.view()), ctypes.c_int(_DIM[dim]),
                                     ctypes.c_int({"mfma": 0, "plain": 1}[form]), ctypes.c_size_t(max_chunk)))

    def zp_apply_dim_noalloc(self, mat, dst, src, dim, max_chunk=0):
        """zp_apply_dim's "mfma" form without any allocation (csd_zp_apply_dim_noalloc): the source
        digit planes go to the device's fixed scratch array; inputs needing more than it holds are
        rejected"""
        check(lib().csd_zp_apply_dim_noalloc(self.ctx, ctypes.c_void_p(mat.handle), ctypes.byref(dst.view()),
                                             ctypes.byref(src.view()), ctypes.c_int(_DIM[dim]),
                                             ctypes.c_size_t(max_chunk)))

    def zp_set_kernel(self, form):
        """which int8 MFMA kernel this context's "mfma" form runs (csd_zp_set_kernel): 3 (the
        default) LDS-staged 64 x 64 tiles for 8 waves with 2 blocks each, the A operands built
        from the generator when the matrix is declared group-circulant; 4 the same without that;
        1 LDS-staged 32 x 64 tiles for 8 waves, 2 direct-to-LDS 64 x 64 tiles for 4 waves with
        2 x 2 blocks each, 0 operands loaded by 4 waves straight from L2"""
        check(lib().csd_zp_set_kernel(self.ctx, ctypes.c_int(int(form))))

    def hal_create_permutation(self, inner_shape, shifts):
        n = len(inner_shape)
        h = ctypes.c_void_p()
        shp = (ctypes.c_size_t * max(1, n))(*inner_shape)
        sh = (ctypes.c_size_t * max(1, n))(*shifts)
        check(lib().csd_hal_create_permutation(self.ctx, shp, sh, ctypes.c_size_t(n), ctypes.byref(h)))
        return HipPermutationPlan(self, h.value, int(np.prod(inner_shape)) if n else 1)

    def hal_permute(self, dst, src, plan, dim):
        check(lib().csd_hal_permute(self.ctx, ctypes.byref(dst.view()), ctypes.byref(src.view()), ctypes.c_void_p(plan.handle),
                                    ctypes.c_int(_DIM[dim])))
