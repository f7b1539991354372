"""The period<->CRT transforms of field_36864-shaped rings on the int8 matrix cores, exact over Z_p.

The K part of pb_to_crt (ring.rs:1117-1137) is the group convolution Y = k * X over G = C_2 x C_3
x C_8 x C_64 of every 3072-line: Y = C_K X with the group circulant C_K[x][y] = k[x - y]. With the
group's characters F = F_48 (x) F_64 (index x = 64 a + b; F_48 = F_2 (x) F_3 (x) F_8 on a = (a0, a1,
a2), F_n[u][v] = w_n^(u v)) the convolution theorem gives C_K = F^-1 diag(F k) F, and F acts as
F_64 on the 64 consecutive elements b of each line and F_48 on the 48 strided ones a. This form
applies those four small dense maps (64 x 64 along b, 48 x 48 along a, their inverses) and the
diagonal on the matrix cores (csd_zp_apply_dim: balanced 8-bit digits, int32 accumulation,
exact recombination and Montgomery reduction), 224 Z_p-MACs per coefficient instead of the 3072
of the dense circulant. The L' part (12 x 12, dim 2) runs the ring's NTT passes (ntt_mul, intt),
which compute C_L' exactly mod p. crt_to_pb is the same with the inverse kernels.

Results are canonical residues in [0, p), equal mod p to the reference's transforms (level L1):
the lazy NTT's representatives are not reproduced by a different algorithm."""
import numpy as np

from .hal import DeviceTensor


def _root(p, order):
    """an element of exact multiplicative order `order` mod p (order | p - 1)"""
    assert (p - 1) % order == 0
    fac, m, f = [], order, 2
    while m > 1:
        if m % f == 0:
            fac.append(f)
            while m % f == 0:
                m //= f
        f += 1
    g = 2
    while True:
        w = pow(g, (p - 1) // order, p)
        if all(pow(w, order // q, p) != 1 for q in fac):
            return w
        g += 1


def _dft(p, n, inverse=False):
    """F_n[u][v] = w^(u v) mod p (w^-1 and the factor n^-1 for the inverse)"""
    w = _root(p, n)
    if inverse:
        w = pow(w, -1, p)
    pw = [pow(w, e, p) for e in range(n)]
    m = np.array([[pw[(u * v) % n] for v in range(n)] for u in range(n)], dtype=object)
    if inverse:
        m = m * pow(n, -1, p) % p
    return m


def _kron(*ms):
    out = np.ones((1, 1), dtype=object)
    for m in ms:
        out = np.kron(out, m)
    return out


class MfmaTransforms:
    """The factor matrices of one ring and base, on the device: F_64, F_48 and their inverses as
    dense Z_p maps (their digit planes split once), and the diagonals F k of the CRT kernel and of
    its inverse, in Montgomery form."""

    def __init__(self, ring, base):
        info = ring.info
        if info["orders"][info["lprime_factor_count"]:] != [2, 3, 8, 64]:
            raise ValueError("the MFMA transform form is built for K parts of shape [2, 3, 8, 64] (field_36864)")
        self.ring, self.base = ring, list(base)
        be = ring.backend
        primes = [ring.primes[i] for i in base]
        self.mods = [be.hal_create_modulus(p) for p in primes]
        r = len(primes)

        def upload(mats):
            host = np.stack([np.asarray(m % p, dtype=np.int64) for m, p in zip(mats, primes)])[None]
            t = be.hal_allocate_i64(host.shape)
            be.hal_import(t, np.ascontiguousarray(host))
            return be.zp_matrix(t, self.mods)

        self.f64 = upload([_dft(p, 64) for p in primes])
        self.f64i = upload([_dft(p, 64, True) for p in primes])
        self.f48 = upload([_kron(_dft(p, 2), _dft(p, 3), _dft(p, 8)) for p in primes])
        self.f48i = upload([_kron(_dft(p, 2, True), _dft(p, 3, True), _dft(p, 8, True)) for p in primes])
        # the diagonals: F applied to the first column of the dense K-part maps
        self.diag = {}
        for inverse in (False, True):
            dense = ring.transform_matrix(base, 3, inverse)  # [1, r, 3072, 3072]
            col = DeviceTensor(be, dense.ptr, (1, r, 3072, 1), dense.strides, 8, dense)
            k = be.hal_allocate_i64((1, r, 48, 64))
            be.hal_assign(DeviceTensor(be, k.ptr, (1, r, 3072, 1), (r * 3072, 3072, 1, 1), 8, k), col)
            del dense, col
            t1 = be.hal_allocate_i64((1, r, 48, 64))
            be.zp_apply_dim(self.f64, t1, k, 3)
            t2 = be.hal_allocate_i64((1, r, 48, 64))
            be.zp_apply_dim(self.f48, t2, t1, 2)
            d = be.hal_export(t2).reshape(r, 3072).astype(object)
            mont = np.stack([(d[j] * ((1 << 64) % p)) % p for j, p in enumerate(primes)]).astype(np.int64)
            dt = be.hal_allocate_i64((1, r, 1, 3072))
            be.hal_import(dt, mont.reshape(1, r, 1, 3072))
            self.diag[inverse] = dt
        # the L' part: the ring's own NTT plans ([4, 3]) and NTT-domain CRT kernels, per prime
        nlp = info["lprime_factor_count"]
        self.plans_lp = [be.hal_create_ntt_plan(m, info["orders"][:nlp]) for m in self.mods]
        kl, kli = np.zeros((1, r, 12, 1), dtype=np.int64), np.zeros((1, r, 12, 1), dtype=np.int64)
        for j, i in enumerate(base):
            lp, _, ilp, _ = ring.crt_kernels(i)
            kl[0, j, :, 0], kli[0, j, :, 0] = lp, ilp
        self.kern_lp = {}
        for inverse, host in ((False, kl), (True, kli)):
            t = be.hal_allocate_i64(host.shape)
            be.hal_import(t, host)
            self.kern_lp[inverse] = t
        self._tmp = {}

    def _scratch(self, shape):
        t = self._tmp.get(shape)
        if t is None:
            t = self._tmp[shape] = self.ring.backend.allocate_i64_uninit(list(shape))
        return t

    def _k_part(self, data, inverse):
        """data [c, r, 12, 3072] <- C_K (or its inverse) applied to every line; canonical output"""
        be = self.ring.backend
        c, r = data.shape[0], data.shape[1]
        tmp = self._scratch((c, r, 12, 3072))

        def lines(t):  # [c, r, 12 * 48, 64]: b along dim 3
            return DeviceTensor(be, t.ptr, (c, r, 12 * 48, 64), (t.strides[0], t.strides[1], 64, 1), 8, t)

        def rows(t, row):  # row `row` of every element as [c, r, 48, 64]: a along dim 2
            return DeviceTensor(be, t.ptr + row * 3072 * 8, (c, r, 48, 64), (t.strides[0], t.strides[1], 64, 1), 8, t)

        be.zp_apply_dim(self.f64, lines(tmp), lines(data), 3)
        for row in range(12):
            be.zp_apply_dim(self.f48, rows(data, row), rows(tmp, row), 2)
        be.hal_mul_mont64(data, self.diag[inverse], self.mods)
        for row in range(12):
            be.zp_apply_dim(self.f48i, rows(tmp, row), rows(data, row), 2)
        be.zp_apply_dim(self.f64i, lines(data), lines(tmp), 3)

    def pb_to_crt_tensor(self, data):
        """pb_to_crt_tensor (ring.rs:1117-1137) of a contiguous [c, r, 12, 3072] batch, L1"""
        self._apply(data, False)

    def crt_to_pb_tensor(self, data):
        """crt_to_pb_tensor (ring.rs:1141-1161) of a contiguous [c, r, 12, 3072] batch, L1"""
        self._apply(data, True)

    def _apply(self, data, inverse):
        if tuple(data.shape[1:]) != (len(self.base), 12, 3072) or data.strides[3] != 1 or data.strides[2] != 3072:
            raise ValueError("the MFMA transform takes [count, r, 12, 3072] batches with contiguous slices")
        # the L' part: NTT, kernel multiply, inverse NTT along dim 2 (exactly C_L' mod p)
        be = self.ring.backend
        be.hal_ntt_mul_mont64_dim(data, self.plans_lp, 2, self.kern_lp[inverse], self.mods)
        be.hal_intt_mont64_dim(data, self.plans_lp, 2)
        self._k_part(data, inverse)
