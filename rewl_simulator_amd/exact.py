"""Exact density of states of the 2D Ising model on the periodic L x L lattice.

Kaufman's partition function (Phys. Rev. 76, 1232 (1949)) written as an
integer polynomial in y = exp(-2 beta J) following Beale (PRL 76, 78 (1996)):

    sum_b g_b y^b = 1/2 (P1 + P2 + P3 + P4),      b = number of broken bonds,

with, for even L, m = L/2, c_k = cos(k pi / L),
    sigma_k = (1 + y^2)^2 - 2 y (1 - y^2) c_k,   q = y^2 (1 - y^2)^2,
    p_j = sigma p_{j-1} - q p_{j-2}  (p_0 = 2, p_1 = sigma)       [A^j + B^j]
    e_j = sigma e_{j-1} - q e_{j-2}  (e_0 = 0, e_1 = 1)            [(A^j - B^j)/(A - B)]
where A, B are the roots of t^2 - sigma t + q (A = y^2 sinh(2K) e^{gamma_k}),
    P1 = prod_{k odd < L} p_m(k)^2
    P2 = prod_{k odd < L} (sigma_k^2 - 4 q) e_m(k)^2
    P3 = p_m(0) p_m(L) prod_{k even, 0 < k < L} p_m(k)^2
    P4 = f(0) f(L) prod_{k even, 0 < k < L} (sigma_k^2 - 4 q) e_m(k)^2
    p_m(0) = (1-y)^L + y^L (1+y)^L,   f(0) = (1-y)^L - y^L (1+y)^L,
    p_m(L) = (1+y)^L + y^L (1-y)^L,   f(L) = (1+y)^L - y^L (1-y)^L.
(The k and 2L-k factors of Kaufman's products coincide, which turns his square
roots into squares; gamma_0 = 2K + ln tanh K keeps its sign through f(0).)

Every coefficient is an integer below 2^N.  They are computed exactly: the
polynomial is evaluated modulo NTT-friendly primes p (p - 1 divisible by the
transform length and by 2L, so that 2 cos(k pi / L) = w^k + w^-k exists in
GF(p)), interpolated by an inverse number-theoretic transform and lifted by the
Chinese remainder theorem.  Energy E = -2N + 2b for J = 1.
"""
import math

import numpy as np


def _is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _primes(modulus, count, below=1 << 30):
    out = []
    c = (below - 1) // modulus
    while len(out) < count:
        p = c * modulus + 1
        if _is_prime(p):
            out.append(p)
        c -= 1
    return out


def _root_of_unity(order, p):
    """A primitive root of unity of the given order modulo prime p."""
    fac = []
    n = p - 1
    f = 2
    while f * f <= n:
        if n % f == 0:
            fac.append(f)
            while n % f == 0:
                n //= f
        f += 1
    if n > 1:
        fac.append(n)
    for g in range(2, p):
        if all(pow(g, (p - 1) // q, p) != 1 for q in fac):
            return pow(g, (p - 1) // order, p)
    raise ValueError("no generator")


def _ntt(a, root, p):
    """Cyclic transform a_hat[t] = sum_b a[b] root^(b t) mod p (len(a) = 2^k)."""
    n = a.size
    a = a.copy()
    j = 0
    idx = np.arange(n)
    bits = n.bit_length() - 1
    rev = np.zeros(n, dtype=np.int64)
    for b in range(bits):
        rev |= ((idx >> b) & 1) << (bits - 1 - b)
    a = a[rev]
    length = 2
    while length <= n:
        w = pow(root, n // length, p)
        half = length // 2
        tw = np.empty(half, dtype=np.int64)
        tw[0] = 1
        for k in range(1, half):
            tw[k] = tw[k - 1] * w % p
        a = a.reshape(-1, length)
        u = a[:, :half].copy()
        v = a[:, half:] * tw % p
        a[:, :half] = (u + v) % p
        a[:, half:] = (u - v) % p
        a = a.reshape(-1)
        length *= 2
        j += 1
    return a


def _powmod_vec(x, e, p):
    r = np.ones_like(x)
    b = x.copy()
    while e:
        if e & 1:
            r = r * b % p
        b = b * b % p
        e >>= 1
    return r


def _g_mod_p(L, M, p):
    """Coefficients of sum_b g_b y^b modulo p (length M)."""
    m = L // 2
    zeta = _root_of_unity(M, p)
    w = _root_of_unity(2 * L, p)
    t = np.arange(M, dtype=np.int64)
    # y_t = zeta^t
    y = np.empty(M, dtype=np.int64)
    y[0] = 1
    for i in range(1, M):
        y[i] = y[i - 1] * zeta % p
    y2 = y * y % p
    one_m_y2 = (1 - y2) % p
    q = y2 * one_m_y2 % p * one_m_y2 % p
    base = (1 + y2) % p
    base = base * base % p
    two_y_1my2 = 2 * y % p * one_m_y2 % p

    def sigma(k):
        ck2 = (pow(w, k, p) + pow(w, (2 * L - k) % (2 * L), p)) % p  # 2 cos(k pi / L)
        # (1+y^2)^2 - y (1-y^2) * (2 cos)
        return (base - y * one_m_y2 % p * ck2) % p

    def pe(k):
        s = sigma(k)
        pp0, pp1 = np.full(M, 2, dtype=np.int64), s.copy()
        e0, e1 = np.zeros(M, dtype=np.int64), np.ones(M, dtype=np.int64)
        for _ in range(2, m + 1):
            pp0, pp1 = pp1, (s * pp1 - q * pp0) % p
            e0, e1 = e1, (s * e1 - q * e0) % p
        if m == 0:
            pp1, e1 = pp0, e0
        disc = (s * s - 4 * q) % p
        return pp1, e1, disc

    P1 = np.ones(M, dtype=np.int64)
    P2 = np.ones(M, dtype=np.int64)
    for k in range(1, L, 2):
        pk, ek, dk = pe(k)
        P1 = P1 * pk % p * pk % p
        P2 = P2 * dk % p * ek % p * ek % p
    ymL = _powmod_vec(y, L, p)
    om = _powmod_vec((1 - y) % p, L, p)
    op = _powmod_vec((1 + y) % p, L, p)
    pm0 = (om + ymL * op) % p
    f0 = (om - ymL * op) % p
    pmL = (op + ymL * om) % p
    fL = (op - ymL * om) % p
    P3 = pm0 * pmL % p
    P4 = f0 * fL % p
    for k in range(2, L, 2):
        pk, ek, dk = pe(k)
        P3 = P3 * pk % p * pk % p
        P4 = P4 * dk % p * ek % p * ek % p
    inv2 = pow(2, p - 2, p)
    G = (P1 + P2 + P3 + P4) % p * inv2 % p
    # coefficients: c_b = M^-1 sum_t G(zeta^t) zeta^(-b t)
    coef = _ntt(G, pow(zeta, p - 2, p), p)
    return coef * pow(M, p - 2, p) % p


def ising_dos(L):
    """Exact g(E) of the L x L periodic Ising model (J = 1), L even >= 2.

    Returns (E, g) with E an int64 array of energies -2N..2N in steps of 2
    (odd numbers of broken bonds never occur, so every other entry is 0) and g
    a list of Python ints."""
    if L < 2 or L % 2:
        raise ValueError("even L >= 2 required")
    N = L * L
    deg = 2 * N
    M = 1
    while M < deg + 1:
        M *= 2
    modulus = M * 2 * L // math.gcd(M, 2 * L)
    nprimes = N // 29 + 3
    primes = _primes(modulus, nprimes)
    residues = [_g_mod_p(L, M, p) for p in primes]
    # Garner / CRT with Python integers
    coeffs = [0] * (deg + 1)
    prod = 1
    for p, r in zip(primes, residues):
        inv = pow(prod % p, p - 2, p)
        for b in range(deg + 1):
            x = coeffs[b]
            d = (int(r[b]) - x) % p * inv % p
            coeffs[b] = x + prod * d
        prod *= p
    E = -2 * N + 2 * np.arange(deg + 1, dtype=np.int64)
    return E, coeffs


def ising_lng(L):
    """(E, ln g) over the levels with g > 0."""
    E, g = ising_dos(L)
    keep = [i for i, v in enumerate(g) if v > 0]
    lng = np.array([_ln_int(g[i]) for i in keep])
    return E[keep], lng


def _ln_int(v):
    """Natural log of a (huge) positive Python int to double precision."""
    if v < (1 << 1000):
        return math.log(v)
    s = v.bit_length() - 900
    return math.log(v >> s) + s * math.log(2.0)


def binned_lng(L, e_min, bin_size, nbins):
    """Exact ln of the number of states per histogram bin [e_min + b*bs, + bs),
    the quantity a Wang-Landau histogram of that binning converges to; -inf for
    empty bins."""
    E, g = ising_dos(L)
    tot = [0] * nbins
    for e, v in zip(E.tolist(), g):
        if v == 0:
            continue
        b = int((e - e_min) // bin_size)
        if 0 <= b < nbins:
            tot[b] += v
    return np.array([_ln_int(t) if t > 0 else -np.inf for t in tot])
