"""The group solver's linear algebra, restated in NumPy lane by lane:
parallel cyclic reduction (one row per lane), one cyclic-reduction level +
PCR (two rows per lane), and the partition method (R rows per lane)
reducing to the two-row case.  Checked against dense LU on symmetric,
strictly diagonally dominant tridiagonals of the kind the kernels build
(widely varying conductances), with identity padding rows."""
import numpy as np
import pytest

def pcr1(a, b, c, d):   # arrays over lanes
    G = a.size
    a, b, c, d = a.copy(), b.copy(), c.copy(), d.copy()
    s = 1
    while s < G:
        ib = 1.0 / b
        idx = np.arange(G)
        am = np.where(idx >= s, np.roll(a, s), 0); cm = np.where(idx >= s, np.roll(c, s), 0)
        dm = np.where(idx >= s, np.roll(d, s), 0); ibm = np.where(idx >= s, np.roll(ib, s), 1)
        ap = np.where(idx + s < G, np.roll(a, -s), 0); cp = np.where(idx + s < G, np.roll(c, -s), 0)
        dp = np.where(idx + s < G, np.roll(d, -s), 0); ibp = np.where(idx + s < G, np.roll(ib, -s), 1)
        al = -a * ibm; ga = -c * ibp
        na = al * am; nc = ga * cp
        b = b + al * cm + ga * ap; d = d + al * dm + ga * dp
        a, c = na, nc
        s <<= 1
    return d / b
def pcr2(a, b, c, d):   # a etc shape (G, 2)
    G = a.shape[0]; idx = np.arange(G)
    a2 = np.where(idx + 1 < G, np.roll(a[:, 0], -1), 0); b2 = np.where(idx + 1 < G, np.roll(b[:, 0], -1), 1)
    c2 = np.where(idx + 1 < G, np.roll(c[:, 0], -1), 0); d2 = np.where(idx + 1 < G, np.roll(d[:, 0], -1), 0)
    al = -a[:, 1] / b[:, 0]; ga = -c[:, 1] / b2
    oa = al * a[:, 0]; oc = ga * c2; ob = b[:, 1] + al * c[:, 0] + ga * a2; od = d[:, 1] + al * d[:, 0] + ga * d2
    xo = pcr1(oa, ob, oc, od)
    xm = np.where(idx > 0, np.roll(xo, 1), 0)
    x = np.zeros((G, 2)); x[:, 1] = xo; x[:, 0] = (d[:, 0] - a[:, 0] * xm - c[:, 0] * xo) / b[:, 0]
    return x
def solve(G, R, a, b, c, d):  # (G, R)
    if R == 1: return pcr1(a[:, 0], b[:, 0], c[:, 0], d[:, 0])[:, None]
    if R == 2: return pcr2(a, b, c, d)
    M = R - 2
    cp = np.zeros((G, M)); al = np.zeros((G, M)); be = np.zeros((G, M)); ga = np.zeros((G, M))
    for j in range(M):
        k = j + 1
        den = b[:, k] - (a[:, k] * cp[:, j-1] if j > 0 else 0)
        inv = 1 / den
        cp[:, j] = c[:, k] * inv
        al[:, j] = (d[:, k] - (a[:, k] * al[:, j-1] if j > 0 else 0)) * inv
        be[:, j] = ((-a[:, 1] if j == 0 else 0) - (a[:, k] * be[:, j-1] if j > 0 else 0)) * inv
        ga[:, j] = ((-c[:, R-2] if j == M-1 else 0) - (a[:, k] * ga[:, j-1] if j > 0 else 0)) * inv
    for j in range(M-2, -1, -1):
        al[:, j] -= cp[:, j] * al[:, j+1]; be[:, j] -= cp[:, j] * be[:, j+1]; ga[:, j] -= cp[:, j] * ga[:, j+1]
    ra = np.stack([a[:, 0], a[:, R-1] * be[:, M-1]], 1)
    rb = np.stack([b[:, 0] + c[:, 0] * be[:, 0], b[:, R-1] + a[:, R-1] * ga[:, M-1]], 1)
    rc = np.stack([c[:, 0] * ga[:, 0], c[:, R-1]], 1)
    rd = np.stack([d[:, 0] - c[:, 0] * al[:, 0], d[:, R-1] - a[:, R-1] * al[:, M-1]], 1)
    rx = pcr2(ra, rb, rc, rd)
    x = np.zeros((G, R)); x[:, 0] = rx[:, 0]; x[:, R-1] = rx[:, 1]
    for j in range(M): x[:, j+1] = al[:, j] + be[:, j] * rx[:, 0] + ga[:, j] * rx[:, 1]
    return x
@pytest.mark.parametrize("G", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("R", [1, 2, 3, 4, 5, 6, 8])
def test_group_solve_matches_dense(G, R):
    rng = np.random.default_rng(G * 100 + R)
    for N in sorted(set([max(2, G * R - 3), G * R, max(2, G * R // 2 + 1)])):
        if True:
            n = G * R
            # symmetric strictly diagonally dominant (as in the kernels), rows >= N identity
            w = rng.uniform(0.1, 5, n + 1) * np.exp(rng.uniform(-8, 8, n + 1))
            a = np.r_[0, w[1:n]]; c = np.r_[w[1:n], 0]
            b = -(a + c) - rng.uniform(1e-3, 2, n)
            d = rng.standard_normal(n)
            a[N:] = 0; c[N:] = 0; b[N:] = 1; d[N:] = 0; c[N-1] = 0
            A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
            xr = np.linalg.solve(A, d)
            x = solve(G, R, a.reshape(G, R), b.reshape(G, R), c.reshape(G, R), d.reshape(G, R)).ravel()
            err = np.abs(x - xr).max() / np.abs(xr).max()
            assert err < 1e-9, (G, R, N, err)
