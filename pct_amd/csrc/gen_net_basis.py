"""Writes pct_net_basis.h: the fixed tables of the learned-path form.

The surrogate path between the entry and exit of a history is, on each of the
five segments sigma in [i/5, (i+1)/5) of the normalised depth sigma = u_1/u_2,
a cubic in tau = 5 sigma - i:

    t(sigma) = t_in + (t_out - t_in) h01(sigma) + (u_2 dtheta) h11(sigma)
               + sum_j d_j phi_j(sigma)

with the cubic Hermite functions h01 = 3 s^2 - 2 s^3, h11 = s^3 - s^2 (value
and slope of the path at both ends: slope 0 at the entry, dtheta at the exit,
the MLP's own boundary conditions) and phi_j the clamped cubic spline through
the knots 0, 1/5, ..., 1 that is 1 at interior knot j+1, 0 at the other knots
and has zero slope at both ends (so the learned offsets d_j never move the
endpoints or their slopes). Every coefficient is computed here in exact
rational arithmetic and rounded once to float.

Run:  python pct_amd/csrc/gen_net_basis.py  (rewrites the header in place)
"""
import os
from fractions import Fraction as F

SEG = 5
H = F(1, SEG)


def poly_mul(a, b):
    out = [F(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def poly_add(a, b):
    n = max(len(a), len(b))
    return [(a[i] if i < len(a) else 0) + (b[i] if i < len(b) else 0) for i in range(n)]


def sigma_poly(i):
    """sigma = (i + tau) / 5 as a polynomial in tau"""
    return [F(i, SEG), F(1, SEG)]


def hermite(i):
    s = sigma_poly(i)
    s2 = poly_mul(s, s)
    s3 = poly_mul(s2, s)
    h01 = poly_add([3 * c for c in s2], [-2 * c for c in s3])
    h11 = poly_add(s3, [-c for c in s2])
    pad = lambda p: (p + [F(0)] * 4)[:4]
    return pad(h01), pad(h11)


def clamped_moments(y):
    """second derivatives M_0..M_5 of the clamped (zero end slope) cubic spline
    through (k/5, y_k) by Gaussian elimination in rationals"""
    n = SEG + 1
    A = [[F(0)] * n for _ in range(n)]
    r = [F(0)] * n
    A[0][0], A[0][1] = F(2), F(1)
    r[0] = 6 * ((y[1] - y[0]) / H - 0) / H
    for k in range(1, n - 1):
        A[k][k - 1], A[k][k], A[k][k + 1] = F(1), F(4), F(1)
        r[k] = 6 * (y[k + 1] - 2 * y[k] + y[k - 1]) / (H * H)
    A[n - 1][n - 2], A[n - 1][n - 1] = F(1), F(2)
    r[n - 1] = 6 * (0 - (y[n - 1] - y[n - 2]) / H) / H
    for c in range(n):
        piv = A[c][c]
        for k in range(c + 1, n):
            f = A[k][c] / piv
            if f:
                for m in range(c, n):
                    A[k][m] -= f * A[c][m]
                r[k] -= f * r[c]
    M = [F(0)] * n
    for c in range(n - 1, -1, -1):
        M[c] = (r[c] - sum(A[c][m] * M[m] for m in range(c + 1, n))) / A[c][c]
    return M


def spline_segments(y):
    M = clamped_moments(y)
    segs = []
    for i in range(SEG):
        q = H * H / 6
        c0 = y[i]
        c1 = (y[i + 1] - y[i]) - q * (2 * M[i] + M[i + 1])
        c2 = q * 3 * M[i]
        c3 = q * (M[i + 1] - M[i])
        segs.append([c0, c1, c2, c3])
    return segs


def tables():
    h01 = [hermite(i)[0] for i in range(SEG)]
    h11 = [hermite(i)[1] for i in range(SEG)]
    phi = []
    for j in range(4):
        y = [F(0)] * (SEG + 1)
        y[j + 1] = F(1)
        phi.append(spline_segments(y))
    return h01, h11, phi


def fl(x):
    s = "%.9g" % float(x)
    if "." not in s and "e" not in s:
        s += ".0"
    return s + "f"


def render():
    h01, h11, phi = tables()
    lines = []
    w = lines.append
    w("/* pct_net_basis.h -- generated by gen_net_basis.py; do not edit.")
    w(" *")
    w(" * Fixed tables of the learned-path form, shared by the HIP kernels and the")
    w(" * CPU oracle so that both step the surrogate path with the same floats.")
    w(" * Segment i of the normalised depth sigma = u_1 / u_2 covers [i/5, (i+1)/5);")
    w(" * tau = 5 sigma - i; each row holds the coefficients c0..c3 of")
    w(" * c0 + c1 tau + c2 tau^2 + c3 tau^3.")
    w(" *   NET_H01[i]    : Hermite 3 s^2 - 2 s^3 (weight of t_out - t_in)")
    w(" *   NET_H11[i]    : Hermite s^3 - s^2     (weight of u_2 * dtheta)")
    w(" *   NET_PHI[j][i] : clamped cubic spline, 1 at knot (j+1)/5, 0 at the other")
    w(" *                   knots, zero slope at sigma = 0 and 1 (weight of offset j)")
    w(" */")
    w("#ifndef PCT_NET_BASIS_H")
    w("#define PCT_NET_BASIS_H")
    w("")
    w("#define PCT_NET_SEGMENTS 5")
    w("#define PCT_NET_KNOTS 4")
    w("")

    def table(name, rows):
        w(f"static constexpr float {name}[PCT_NET_SEGMENTS][4] = {{")
        for r in rows:
            w("    {" + ", ".join(fl(c) for c in r) + "},")
        w("};")

    table("NET_H01", h01)
    table("NET_H11", h11)
    w("static constexpr float NET_PHI[PCT_NET_KNOTS][PCT_NET_SEGMENTS][4] = {")
    for j in range(4):
        w("    {")
        for r in phi[j]:
            w("        {" + ", ".join(fl(c) for c in r) + "},")
        w("    },")
    w("};")
    w("")
    w("#endif")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pct_net_basis.h")
    with open(out, "w") as f:
        f.write(render())
    print(out)
