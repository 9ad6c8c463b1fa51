"""ORACLE -- test infrastructure only.

Restatement of the reference's real spherical harmonics, ylm_module.f90:10-72
(Ylm) and :76-121 (Plm, the associated Legendre recurrence), used to pin the
projector tables' angular convention.
"""
import math


def plm(L, M, xi):
    """Associated Legendre P_L^M(x) with the Condon-Shortley sign (ylm_module.f90:76-121)."""
    x = xi
    if abs(abs(x) - 1.0) < 1.0e-13:
        x = math.copysign(1.0, x)
    if M < 0 or M > L or abs(x) > 1.0:
        raise ValueError("Bad arguments (Plm)")
    pmm = 1.0
    if M > 0:
        somx2 = math.sqrt((1.0 - x) * (1.0 + x))
        fact = 1.0
        for _ in range(M):
            pmm = -pmm * fact * somx2
            fact += 2.0
    if L == M:
        return pmm
    pmmp1 = x * (2.0 * M + 1.0) * pmm
    if L == M + 1:
        return pmmp1
    pll = 0.0
    for ll in range(M + 2, L + 1):
        pll = (x * (2.0 * ll - 1.0) * pmmp1 - (ll + M - 1.0) * pmm) / (ll - M)
        pmm, pmmp1 = pmmp1, pll
    return pll


def ylm(x, y, z, L, M_in):
    """Real Y_LM of ylm_module.f90:10-72 (zero at the origin for L > 0)."""
    pi = 3.141592653589793238
    M = abs(M_in)
    pi4 = 4.0 * pi
    if L == 0:
        return 1.0 / math.sqrt(pi4)
    if x == 0.0 and y == 0.0 and z == 0.0:
        return 0.0
    c = 1.0
    for k in range(L - M + 1, L + M + 1):
        c *= k
    clm = math.sqrt((2.0 * L + 1.0) / (c * pi4))
    r = math.sqrt(x * x + y * y + z * z)
    ct = z / r
    if abs(x) < 1.0e-10:
        phi = 0.5 * pi * math.copysign(1.0, y)
    else:
        phi = math.atan(y / x)
        if x < 0:
            phi += pi
    if M_in > 0:
        return math.sqrt(2.0) * clm * plm(L, M, ct) * math.cos(phi * M)
    if M_in == 0:
        return clm * plm(L, M, ct)
    if M % 2 == 0:
        return -math.sqrt(2.0) * clm * plm(L, M, ct) * math.sin(phi * M)
    return math.sqrt(2.0) * clm * plm(L, M, ct) * math.sin(phi * M)
