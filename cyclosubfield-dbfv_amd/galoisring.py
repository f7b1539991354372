"""Arithmetic in Galois rings GR(2^e, D) = Z_(2^e)[y] / (g(y)) and Gaussian elimination modulo 2^e,
on the host, for the set-up of the homomorphic CRT transform (cyclosubfield-dbfv/src/bootstrap/
galoisring.rs). Only the transform's coefficients come from here; they are computed once per
transform, before any ciphertext is touched.

F_2 polynomials are Python integers read as bit vectors (bit i = coefficient of y^i). The
modulus g is the first irreducible y^D + c with odd c in increasing order (galoisring.rs:136-
154), roots of unity are the first ones the reference's deterministic search finds
(galoisring.rs:297-338), so the coefficients are the reference's."""


def _deg(a):
    return a.bit_length() - 1


def _clmul(a, b):
    """the product of two F_2 polynomials"""
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        b >>= 1
    return r


def _rem(a, g):
    dg = _deg(g)
    while a and _deg(a) >= dg:
        a ^= g << (_deg(a) - dg)
    return a


def _mulmod(a, b, g):
    return _rem(_clmul(a, b), g)


def _powmod(base, e, g):
    """base^e mod g for an integer exponent e (galoisring.rs:73-84)"""
    result, base = _rem(1, g), _rem(base, g)
    while e:
        if e & 1:
            result = _mulmod(result, base, g)
        base = _mulmod(base, base, g)
        e >>= 1
    return result


def _gcd(a, b):
    while b:
        a, b = b, _rem(a, b)
    return a


def _is_irreducible(g, d):
    """galoisring.rs:101-134: y^(2^d) = y mod g, and gcd(y^(2^(d/r)) - y, g) = 1 for primes r | d"""
    y = _rem(2, g)

    def frob(k):
        cur = y
        for _ in range(k):
            cur = _mulmod(cur, cur, g)
        return cur

    if frob(d) != y:
        return False
    rest, r = d, 2
    while rest > 1:
        if rest % r == 0:
            while rest % r == 0:
                rest //= r
            if _deg(_gcd(frob(d // r) ^ y, g)) != 0:
                return False
        r += 1
    return True


def find_irreducible(d):
    """galoisring.rs:136-154: y^d + c for the first odd c that makes it irreducible"""
    if d < 1:
        raise ValueError("degree must be positive")
    if d == 1:
        return 2
    c = 1
    while True:
        g = (1 << d) ^ c
        if _is_irreducible(g, d):
            return g
        c += 2


def _inv_f2(a, g):
    """galoisring.rs:156-166: a^(2^d - 2) mod g"""
    d = _deg(g)
    if a == 0:
        raise ValueError("zero is not invertible")
    return _powmod(a, (1 << d) - 2, g)


def inv_mod_2_128(x):
    """galoisring.rs:169-177: the inverse of odd x modulo 2^128 by Newton iteration"""
    if x & 1 != 1:
        raise ValueError("x must be odd")
    mask = (1 << 128) - 1
    r = x & mask
    for _ in range(6):
        r = (r * (2 - x * r)) & mask
    return r


class GaloisRing:
    """GaloisRing (galoisring.rs:180-348): elements are lists of D coefficients in [0, 2^e)"""

    def __init__(self, e, d):
        if not (1 <= e <= 100 and d >= 1):
            raise ValueError("GR(2^e, D) needs 1 <= e <= 100 and D >= 1")
        self.e, self.d = e, d
        self.g = find_irreducible(d)
        self.mask = (1 << e) - 1
        # the monomials y^j (j < d) that occur in g, for the reduction of products
        self._g_bits = [j for j in range(d) if (self.g >> j) & 1]

    def zero(self):
        return [0] * self.d

    def one(self):
        r = self.zero()
        r[0] = 1
        return r

    def from_scalar(self, x):
        r = self.zero()
        r[0] = x & self.mask
        return r

    def add(self, a, b):
        return [(x + y) & self.mask for x, y in zip(a, b)]

    def sub(self, a, b):
        return [(x - y) & self.mask for x, y in zip(a, b)]

    def mul_scalar(self, a, s):
        return [(x * s) & self.mask for x in a]

    def mul(self, a, b):
        """galoisring.rs:228-251: schoolbook product, reduced by the monic g (0/1 coefficients)"""
        d, mask = self.d, self.mask
        prod = [0] * (2 * d - 1)
        for i, x in enumerate(a):
            if x:
                for j, y in enumerate(b):
                    prod[i + j] += x * y
        prod = [c & mask for c in prod]
        for i in range(2 * d - 2, d - 1, -1):
            c = prod[i]
            if c:
                prod[i] = 0
                for j in self._g_bits:
                    prod[i - d + j] = (prod[i - d + j] - c) & mask
        return prod[:d]

    def pow(self, base, e):
        result, base = self.one(), list(base)
        while e > 0:
            if e & 1:
                result = self.mul(result, base)
            base = self.mul(base, base)
            e >>= 1
        return result

    def is_zero(self, x):
        return all(c == 0 for c in x)

    def _to_f2(self, x):
        v = 0
        for i, c in enumerate(x):
            v |= (c & 1) << i
        return v

    def _from_f2(self, v):
        return [(v >> i) & 1 for i in range(self.d)]

    def inv(self, x):
        """galoisring.rs:282-294: invert modulo 2, then Newton r <- r (2 - x r)"""
        x2 = self._to_f2(x)
        if x2 == 0:
            raise ValueError("element is not a unit")
        r = self._from_f2(_inv_f2(x2, self.g))
        for _ in range(7):
            r = self.mul(r, self.sub(self.from_scalar(2), self.mul(x, r)))
        return r

    def root_of_unity(self, n):
        """galoisring.rs:297-338: an element of exact order n (odd, n | 2^d - 1), found in F_(2^d)
        from the starting elements y^k + 1 (k = 2, 3, ...) and Hensel-lifted"""
        if n % 2 != 1:
            raise ValueError("n must be odd")
        if n == 1:
            return self.one()
        if ((1 << self.d) - 1) % n != 0:
            raise ValueError(f"no root of unity of order {n} in F_2^{self.d}")
        cofactor = ((1 << self.d) - 1) // n
        root_f2 = 1
        k = 2
        while True:
            start = (1 << k) ^ 1
            x = _powmod(start, cofactor, self.g)
            if _deg(x) > 0 and _deg(_powmod(x, n, self.g)) == 0:
                root_f2 = x
                break
            k += 1
        root = self._from_f2(root_f2)
        for _ in range(8):
            power = self.pow(root, n - 1)
            num = self.sub(self.mul(power, root), self.one())
            if self.is_zero(num):
                break
            den = self.mul_scalar(power, n)
            root = self.sub(root, self.mul(num, self.inv(den)))
        if self.pow(root, n) != self.one():
            raise AssertionError("Hensel lifting of the root of unity failed")
        return root

    def smallest_lift(self, c):
        """galoisring.rs:341-345"""
        c &= self.mask
        return c - self.mask - 1 if c > self.mask // 2 else c


def solve_mod_2e(e, rows, cols, rhs_cols, matrix, rhs):
    """solve_mod_2e (galoisring.rs:351-400): matrix (rows x cols, row-major) times the solution
    (cols x rhs_cols) = rhs (rows x rhs_cols) modulo 2^e, by elimination with odd pivots; the
    system must be uniquely solvable modulo 2. matrix and rhs are modified."""
    if rows < cols:
        raise ValueError("need at least as many rows as columns")
    mask = (1 << e) - 1
    pivots = []
    for col in range(cols):
        pivot = next((r for r in range(rows) if r not in pivots and matrix[r * cols + col] & 1), None)
        if pivot is None:
            raise ValueError("matrix is not invertible modulo 2")
        inv = inv_mod_2_128(matrix[pivot * cols + col] & ((1 << 128) - 1)) & mask
        for j in range(cols):
            matrix[pivot * cols + j] = (matrix[pivot * cols + j] * inv) & mask
        for j in range(rhs_cols):
            rhs[pivot * rhs_cols + j] = (rhs[pivot * rhs_cols + j] * inv) & mask
        for r in range(rows):
            if r != pivot:
                f = matrix[r * cols + col]
                if f:
                    for j in range(cols):
                        matrix[r * cols + j] = (matrix[r * cols + j] - f * matrix[pivot * cols + j]) & mask
                    for j in range(rhs_cols):
                        rhs[r * rhs_cols + j] = (rhs[r * rhs_cols + j] - f * rhs[pivot * rhs_cols + j]) & mask
        pivots.append(pivot)
    for r in range(rows):
        if r not in pivots:
            for j in range(rhs_cols):
                if rhs[r * rhs_cols + j] & mask:
                    raise ValueError("linear system is not solvable")
    sol = [0] * (cols * rhs_cols)
    for col, pivot in enumerate(pivots):
        for j in range(rhs_cols):
            sol[col * rhs_cols + j] = rhs[pivot * rhs_cols + j] & mask
    return sol
