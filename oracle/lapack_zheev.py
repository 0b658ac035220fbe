"""LAPACK ZHEEV('V', 'U') restated step by step, TEST INFRASTRUCTURE ONLY.

The reference's cdiag (diag.F90:66-128) calls ZHEEV from MKL; the IVO coefficients it writes
(fockivo.F90:232-290) carry that routine's eigenvector signs.  This restatement follows the
published LAPACK sequence: zhetd2 (Householder reduction of the upper triangle from the last
column), zungtr/zung2l (Q from the reflectors), zsteqr (implicit QL/QR with Wilkinson shifts,
2 x 2 blocks by dlaev2, rotations by dlartg, then a selection sort).  dlartg is taken in its
form before LAPACK 3.10 (r > 0 unless |f| > |g| and f < 0), the one that reproduces the
reference's DFPCMONEW files.
"""
import math
import numpy as np

EPS = np.finfo(float).eps / 2   # dlamch('E') (relative machine precision, rounding)
SAFMIN = np.finfo(float).tiny


def zlarfg(n, alpha, x):
    if n <= 0:
        return alpha, 0.0, x
    xnorm = np.linalg.norm(x) if len(x) else 0.0
    alphr, alphi = alpha.real, alpha.imag
    if xnorm == 0.0 and alphi == 0.0:
        return alpha, 0.0 + 0j, x
    beta = -math.copysign(math.sqrt(alphr ** 2 + alphi ** 2 + xnorm ** 2), alphr)
    tau = complex((beta - alphr) / beta, -alphi / beta)
    a = 1.0 / (alpha - beta)
    x = x * a
    return complex(beta, 0.0), tau, x


def zhetd2_upper(A):
    A = np.array(A, dtype=np.complex128)
    n = A.shape[0]
    d = np.zeros(n); e = np.zeros(max(n - 1, 0)); tau = np.zeros(max(n - 1, 0), dtype=complex)
    A[n - 1, n - 1] = A[n - 1, n - 1].real
    for i in range(n - 1, 0, -1):          # 1-based i = n-1..1 ; column i+1 (0-based i)
        alpha = A[i - 1, i]
        beta, taui, x = zlarfg(i, alpha, A[0:i - 1, i])
        A[0:i - 1, i] = x
        e[i - 1] = beta.real
        if taui != 0:
            A[i - 1, i] = 1.0
            v = A[0:i, i].copy()
            # x := taui * A * v (Hermitian, upper triangle of A(1:i,1:i))
            H = np.triu(A[0:i, 0:i]) + np.conj(np.triu(A[0:i, 0:i], 1)).T
            H[np.arange(i), np.arange(i)] = H[np.arange(i), np.arange(i)].real
            w = taui * (H @ v)
            alph = -0.5 * taui * np.vdot(w, v)
            w = w + alph * v
            # A := A - v w^H - w v^H (upper)
            upd = np.outer(v, np.conj(w)) + np.outer(w, np.conj(v))
            iu = np.triu_indices(i)
            A[0:i, 0:i][iu] -= upd[iu]
            A[np.arange(i), np.arange(i)] = A[np.arange(i), np.arange(i)].real
        else:
            A[i - 1, i - 1] = A[i - 1, i - 1].real
        A[i - 1, i] = e[i - 1]
        d[i] = A[i, i].real
        tau[i - 1] = taui
    d[0] = A[0, 0].real
    return A, d, e, tau


def zungtr_upper(A, tau):
    n = A.shape[0]
    Q = np.zeros((n, n), dtype=complex)
    for j in range(n - 1):
        Q[0:j, j] = A[0:j, j + 1]
        Q[n - 1, j] = 0
    Q[n - 1, n - 1] = 1
    # zung2l(m=n-1, n=n-1, k=n-1) on Q(0:n-1, 0:n-1)
    m = k = nn = n - 1
    B = Q[0:m, 0:nn]
    for i in range(1, k + 1):
        ii = nn - k + i                       # 1-based
        rows = m - nn + ii
        B[rows - 1, ii - 1] = 1.0
        v = B[0:rows, ii - 1].copy()
        C = B[0:rows, 0:ii - 1]
        w = np.conj(C).T @ v
        C -= tau[i - 1] * np.outer(v, np.conj(w))
        B[0:rows - 1, ii - 1] = -tau[i - 1] * B[0:rows - 1, ii - 1]
        B[rows - 1, ii - 1] = 1.0 - tau[i - 1]
        B[rows:m, ii - 1] = 0
    Q[0:m, 0:nn] = B
    return Q


def dlartg(f, g, new=True):
    if g == 0:
        return 1.0, 0.0, f
    if f == 0:
        return (0.0, math.copysign(1.0, g), abs(g)) if new else (0.0, 1.0, g)
    if new:
        dd = math.sqrt(f * f + g * g)
        c = abs(f) / dd
        r = math.copysign(dd, f)
        return c, g / r, r
    r = math.sqrt(f * f + g * g)
    c, s = f / r, g / r
    if abs(f) > abs(g) and c < 0:
        c, s, r = -c, -s, -r
    return c, s, r


def dlaev2(a, b, c):
    sm = a + c; df = a - c; adf = abs(df); tb = b + b; ab = abs(tb)
    acmx, acmn = (a, c) if abs(a) > abs(c) else (c, a)
    if adf > ab:
        rt = adf * math.sqrt(1 + (ab / adf) ** 2)
    elif adf < ab:
        rt = ab * math.sqrt(1 + (adf / ab) ** 2)
    else:
        rt = ab * math.sqrt(2.0)
    if sm < 0:
        rt1 = 0.5 * (sm - rt); sgn1 = -1; rt2 = (acmx / rt1) * acmn - (b / rt1) * b
    elif sm > 0:
        rt1 = 0.5 * (sm + rt); sgn1 = 1; rt2 = (acmx / rt1) * acmn - (b / rt1) * b
    else:
        rt1 = 0.5 * rt; rt2 = -0.5 * rt; sgn1 = 1
    if df >= 0:
        cs = df + rt; sgn2 = 1
    else:
        cs = df - rt; sgn2 = -1
    acs = abs(cs)
    if acs > ab:
        ct = -tb / cs; sn1 = 1 / math.sqrt(1 + ct * ct); cs1 = ct * sn1
    else:
        if ab == 0:
            cs1, sn1 = 1.0, 0.0
        else:
            tn = -cs / tb; cs1 = 1 / math.sqrt(1 + tn * tn); sn1 = tn * cs1
    if sgn1 == sgn2:
        tn = cs1; cs1 = -sn1; sn1 = tn
    return rt1, rt2, cs1, sn1


def zlasr_rvb(Z, c, s, j0, nrot):      # columns j0 .. j0+nrot-1 (0-based), backward
    for j in range(nrot - 2, -1, -1):
        ct, st = c[j], s[j]
        t = Z[:, j0 + j + 1].copy()
        Z[:, j0 + j + 1] = ct * t - st * Z[:, j0 + j]
        Z[:, j0 + j] = st * t + ct * Z[:, j0 + j]


def zlasr_rvf(Z, c, s, j0, nrot):
    for j in range(0, nrot - 1):
        ct, st = c[j], s[j]
        t = Z[:, j0 + j + 1].copy()
        Z[:, j0 + j + 1] = ct * t - st * Z[:, j0 + j]
        Z[:, j0 + j] = st * t + ct * Z[:, j0 + j]


def zsteqr(d, e, Z, new_lartg=False):
    d = d.copy(); e = e.copy(); Z = Z.copy()
    n = len(d)
    eps2 = EPS ** 2
    nmaxit = n * 30
    jtot = 0
    l1 = 1
    # 1-based helpers
    D = lambda i: d[i - 1]
    while True:
        if l1 > n:
            break
        if l1 > 1:
            e[l1 - 2] = 0.0
        m = n
        for mm in range(l1, n):
            tst = abs(e[mm - 1])
            if tst == 0.0:
                m = mm; break
            if tst <= (math.sqrt(abs(d[mm - 1])) * math.sqrt(abs(d[mm]))) * EPS:
                e[mm - 1] = 0.0; m = mm; break
        l = l1; lsv = l; lend = m; lendsv = lend; l1 = m + 1
        if lend == l:
            continue
        if abs(d[lend - 1]) < abs(d[l - 1]):
            lend = lsv; l = lendsv
        if lend > l:   # QL
            while True:
                if l != lend:
                    m = lend
                    for mm in range(l, lend):
                        tst = abs(e[mm - 1]) ** 2
                        if tst <= (eps2 * abs(d[mm - 1])) * abs(d[mm]) + SAFMIN:
                            m = mm; break
                else:
                    m = lend
                if m < lend:
                    e[m - 1] = 0.0
                p = d[l - 1]
                if m == l:
                    d[l - 1] = p; l += 1
                    if l <= lend:
                        continue
                    break
                if m == l + 1:
                    rt1, rt2, c, s = dlaev2(d[l - 1], e[l - 1], d[l])
                    zlasr_rvb(Z, [c], [s], l - 1, 2)
                    d[l - 1] = rt1; d[l] = rt2; e[l - 1] = 0.0; l += 2
                    if l <= lend:
                        continue
                    break
                if jtot == nmaxit:
                    break
                jtot += 1
                g = (d[l] - p) / (2.0 * e[l - 1])
                r = math.hypot(g, 1.0)
                g = d[m - 1] - p + (e[l - 1] / (g + math.copysign(r, g)))
                s = 1.0; c = 1.0; p = 0.0
                cw = np.zeros(n); sw = np.zeros(n)
                for i in range(m - 1, l - 1, -1):
                    f = s * e[i - 1]; b = c * e[i - 1]
                    c, s, r = dlartg(g, f, new_lartg)
                    if i != m - 1:
                        e[i] = r
                    g = d[i] - p
                    r = (d[i - 1] - g) * s + 2.0 * c * b
                    p = s * r
                    d[i] = g + p
                    g = c * r - b
                    cw[i - 1] = c; sw[i - 1] = -s
                zlasr_rvb(Z, cw[l - 1:m], sw[l - 1:m], l - 1, m - l + 1)
                d[l - 1] = d[l - 1] - p; e[l - 1] = g
        else:          # QR
            while True:
                if l != lend:
                    m = lend
                    for mm in range(l, lend, -1):
                        tst = abs(e[mm - 2]) ** 2
                        if tst <= (eps2 * abs(d[mm - 1])) * abs(d[mm - 2]) + SAFMIN:
                            m = mm; break
                else:
                    m = lend
                if m > lend:
                    e[m - 2] = 0.0
                p = d[l - 1]
                if m == l:
                    d[l - 1] = p; l -= 1
                    if l >= lend:
                        continue
                    break
                if m == l - 1:
                    rt1, rt2, c, s = dlaev2(d[l - 2], e[l - 2], d[l - 1])
                    zlasr_rvf(Z, [c], [s], l - 2, 2)
                    d[l - 2] = rt1; d[l - 1] = rt2; e[l - 2] = 0.0; l -= 2
                    if l >= lend:
                        continue
                    break
                if jtot == nmaxit:
                    break
                jtot += 1
                g = (d[l - 2] - p) / (2.0 * e[l - 2])
                r = math.hypot(g, 1.0)
                g = d[m - 1] - p + (e[l - 2] / (g + math.copysign(r, g)))
                s = 1.0; c = 1.0; p = 0.0
                cw = np.zeros(n); sw = np.zeros(n)
                for i in range(m, l):
                    f = s * e[i - 1]; b = c * e[i - 1]
                    c, s, r = dlartg(g, f, new_lartg)
                    if i != m:
                        e[i - 2] = r
                    g = d[i - 1] - p
                    r = (d[i] - g) * s + 2.0 * c * b
                    p = s * r
                    d[i - 1] = g + p
                    g = c * r - b
                    cw[i - 1] = c; sw[i - 1] = s
                zlasr_rvf(Z, cw[m - 1:l], sw[m - 1:l], m - 1, l - m + 1)
                d[l - 1] = d[l - 1] - p; e[l - 2] = g
        if jtot >= nmaxit:
            raise RuntimeError("zsteqr did not converge")
    # selection sort
    for ii in range(2, n + 1):
        i = ii - 1; k = i; p = d[i - 1]
        for j in range(ii, n + 1):
            if d[j - 1] < p:
                k = j; p = d[j - 1]
        if k != i:
            d[k - 1] = d[i - 1]; d[i - 1] = p
            Z[:, [i - 1, k - 1]] = Z[:, [k - 1, i - 1]]
    return d, Z


def zheev_vu(A, new_lartg=False):
    n = A.shape[0]
    if n == 1:
        return np.array([A[0, 0].real]), np.ones((1, 1), dtype=complex)
    T, d, e, tau = zhetd2_upper(A)
    Q = zungtr_upper(T, tau)
    return zsteqr(d, e, Q, new_lartg)
