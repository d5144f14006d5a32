"""The mathematics behind the smooth cells, checked in FP64 on the CPU.

The library sums a cell that is small against the core radius through a
Taylor expansion of f(z) = (1 - exp(-|z|^2)) / z in (z, conj z) about the cell
centre (core-radius units).  Its closed forms -- T_m0 = zb^(m+1) phi^(m)(w),
T_mn = E H_{m,n-1} with the Hermite-type polynomials H, the regrouping of the
n >= 1 part into one polynomial P(z, zb), its coefficients in real monomials
x^a y^b (what the kernel stores), and phi^(m) by the upward recurrence or the
series -- are restated here and compared with brute-force sums of the
exact kernel (not with themselves), so a wrong sign, index or factorial in any
of them fails."""
import math

import numpy as np
import pytest

Q = 12  # the degree the library uses (kSmoothQ)
LAM = math.sqrt(math.log(2.0))


def f_exact(z):
    w = (z * z.conjugate()).real
    out = np.zeros_like(z)
    m = w > 1e-30
    out[m] = -np.expm1(-w[m]) / z[m]
    return out


def moments(d, g, q=Q):
    """mu_mn = sum_j g_j d_j^m conj(d_j)^n / (m! n!) for m + n <= q."""
    mu = {}
    for m in range(q + 1):
        for n in range(q + 1 - m):
            mu[(m, n)] = np.sum(g * d ** m * np.conj(d) ** n) / (math.factorial(m) * math.factorial(n))
    return mu


def coefficients(mu, q=Q):
    """nu_m = mu_m0 and c_ij = sum_l mu_{j+l,i+l+1} C(j+l,l) (i+l)!/i! (-1)^(i+j+l)."""
    nu = [mu[(m, 0)] for m in range(q + 1)]
    c = {}
    for j in range(q):
        for i in range(q - j):
            v = 0j
            l = 0
            while i + j + 2 * l + 1 <= q:
                v += (mu[(j + l, i + l + 1)] * math.comb(j + l, l) * math.factorial(i + l)
                      / math.factorial(i) * (-1) ** (i + j + l))
                l += 1
            c[(i, j)] = v
    return nu, c


def phi_derivs(w, q=Q, w0=1.5, nser=14):
    """phi^(m)(w), phi(w) = (1 - e^-w)/w: upward recurrence for w >= w0, else
    the series of phi^(q) and the recurrence downward (as the kernel does)."""
    E = math.exp(-w)
    ph = [0.0] * (q + 1)
    if w >= w0:
        ph[0] = (1 - E) / w
        for m in range(1, q + 1):
            ph[m] = ((E if m & 1 else -E) - m * ph[m - 1]) / w
    else:
        s = sum((-1) ** k * w ** k / (math.factorial(k) * (k + q + 1)) for k in range(nser))
        ph[q] = -s if q & 1 else s
        for m in range(q, 0, -1):
            ph[m - 1] = ((E if m & 1 else -E) - w * ph[m]) / m
    return ph


def evaluate(Z, nu, c, q=Q):
    """F(Z) = zb sum_m phi^(m)(w) nu_m zb^m + E sum_{i,j} c_ij Z^i zb^j."""
    w = abs(Z) ** 2
    ph = phi_derivs(w, q)
    zb = Z.conjugate()
    s0 = sum(ph[m] * nu[m] * zb ** (m + 1) for m in range(q + 1))
    P = sum(v * Z ** i * zb ** j for (i, j), v in c.items())
    return s0 + math.exp(-w) * P


def _cell(seed, s, n=64):
    rng = np.random.default_rng(seed)
    src = rng.uniform(-s, s, n) + 1j * rng.uniform(-s, s, n)  # positions relative to the centre
    g = rng.normal(size=n)
    return src, g


@pytest.mark.parametrize("s", [0.1, 0.3, 0.49, 0.6])
def test_expansion_matches_exact_sum(s):
    src, g = _cell(1, s)
    nu, c = coefficients(moments(-src, g))
    rng = np.random.default_rng(2)
    worst = 0.0
    for r in (0.0, 0.5 * s, s, 0.8, 1.2, 2.0, 3.0, 4.5):
        for ang in rng.uniform(0, 2 * np.pi, 3):
            Z = r * complex(math.cos(ang), math.sin(ang))
            exact = np.sum(g * f_exact(Z - src))
            worst = max(worst, abs(evaluate(Z, nu, c) - exact) / np.sum(np.abs(g)))
    # truncation after total degree 12: |d|^13 / 13!-like terms; half sides up
    # to 0.6 core radii (the largest eta the library accepts)
    assert worst <= {0.1: 1e-14, 0.3: 1e-10, 0.49: 5e-8, 0.6: 6e-7}[s], worst


def test_phi_branches_agree_with_quadrature():
    t = np.linspace(0.0, 1.0, 20001)
    for w in (0.0, 0.3, 1.0, 1.49, 1.5, 2.5, 6.0, 20.0):
        ph = phi_derivs(w)
        for m in range(Q + 1):
            ref = (-1) ** m * np.trapezoid(t ** m * np.exp(-w * t), t)
            assert abs(ph[m] - ref) <= 1e-7 * max(abs(ref), 1e-3), (w, m, ph[m], ref)


def test_closed_forms_of_the_derivatives():
    """T_mn from the closed forms against central differences of f in x, y
    (d/dz = (d/dx - i d/dy)/2, d/dzb = (d/dx + i d/dy)/2)."""
    def T(Z, m, n):
        w = abs(Z) ** 2
        if n == 0:
            return Z.conjugate() ** (m + 1) * phi_derivs(w)[m]
        k = n - 1
        H = sum(math.comb(m, j) * math.factorial(k) / math.factorial(k - j) * (-1) ** (m + k - j)
                * Z ** (k - j) * Z.conjugate() ** (m - j) for j in range(min(m, k) + 1))
        return math.exp(-w) * H

    h = 1e-3
    f = lambda z: f_exact(np.array([z]))[0]
    dz = lambda g: (lambda z: ((g(z + h) - g(z - h)) - 1j * (g(z + 1j * h) - g(z - 1j * h))) / (4 * h))
    dzb = lambda g: (lambda z: ((g(z + h) - g(z - h)) + 1j * (g(z + 1j * h) - g(z - 1j * h))) / (4 * h))
    for Z in (0.7 + 0.4j, -1.3 + 0.9j, 0.2 - 2.1j):
        for m, n in ((1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1)):
            g = f
            for _ in range(m):
                g = dz(g)
            for _ in range(n):
                g = dzb(g)
            assert abs(T(Z, m, n) - g(Z)) <= 2e-4 * max(1.0, abs(T(Z, m, n))), (Z, m, n)


def test_scaled_coordinates():
    """The kernel evaluates at Z' = Z / lam (units where e^{-|Z|^2} = 2^{-|Z'|^2})
    with nu_m lam^(m+1) and c_ij lam^(i+j): the same value."""
    src, g = _cell(3, 0.2)
    nu, c = coefficients(moments(-src, g))
    nus = [v * LAM ** (m + 1) for m, v in enumerate(nu)]
    cs = {(i, j): v * LAM ** (i + j) for (i, j), v in c.items()}
    for Z in (0.1 + 0.05j, 1.1 - 0.7j, 2.5 + 1.5j):
        Zp = Z / LAM
        wp = abs(Zp) ** 2
        ph = phi_derivs(wp * math.log(2.0))
        zb = Zp.conjugate()
        s0 = sum(ph[m] * nus[m] * zb ** (m + 1) for m in range(Q + 1))
        P = sum(v * Zp ** i * zb ** j for (i, j), v in cs.items())
        F = s0 + 2.0 ** (-wp) * P
        assert abs(F - evaluate(Z, nu, c)) <= 1e-12 * np.sum(np.abs(g))


def real_basis(c, q=Q):
    """d_ab = i^b sum_{i+j=a+b} c_ij sum_{k+l=b} (-1)^l C(i,k) C(j,l): P(z, zb)
    rewritten as sum d_ab x^a y^b with z = x + iy."""
    d = {}
    for b in range(q):
        for a in range(q - b):
            n = a + b
            v = 0j
            for i in range(n + 1):
                j = n - i
                w = sum((-1) ** l * math.comb(i, b - l) * math.comb(j, l)
                        for l in range(max(0, b - i), min(j, b) + 1))
                v += w * c[(i, j)]
            d[(a, b)] = v * 1j ** b
    return d


def test_real_monomial_basis():
    """The stored d_ab give the same P as the c_ij at any point (a wrong sign,
    power of i or binomial changes P)."""
    rng = np.random.default_rng(5)
    c = {(i, j): complex(*rng.normal(size=2)) for j in range(Q) for i in range(Q - j)}
    d = real_basis(c)
    for _ in range(20):
        Z = complex(*rng.uniform(-2.5, 2.5, size=2))
        p_z = sum(v * Z ** i * Z.conjugate() ** j for (i, j), v in c.items())
        p_xy = sum(v * Z.real ** a * Z.imag ** b for (a, b), v in d.items())
        assert abs(p_z - p_xy) <= 1e-12 * max(1.0, abs(p_z)) * 10 ** 3, (Z, p_z, p_xy)
    # one monomial alone: z^2 zb = (x^2 + y^2)(x + iy) -> x^3, x y^2, i x^2 y, i y^3
    one = {k: (1.0 if k == (2, 1) else 0.0) for k in c}
    d1 = {k: v for k, v in real_basis(one).items() if abs(v) > 0}
    assert d1 == {(3, 0): 1, (1, 2): 1, (2, 1): 1j, (0, 3): 1j}, d1
