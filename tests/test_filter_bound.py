"""Soundness of the mixed form's certified plane-side filter (CPU only).

The mixed clip evaluates e . (X, Y, Z, W) in FP32 from FP32-rounded FP64
planes, where (X, Y, Z, W) are the cofactors of the vertex's three planes
(the expansion det4x4(p1, p2, p3, e) of ConvexCell.cl:79-97 / 400-461).  A
result is trusted when |e.v| > kappa * ((|e.x|+|e.y|+|e.z|) MXYZ + |e.w| MW);
otherwise the cell is recomputed in FP64.  Here near-degenerate random
configurations are evaluated the way the kernel does (with and without fused
multiply-adds) and compared with the exact sign of the same expansion over the
FP64 planes (rational arithmetic): every trusted sign must be the exact one.
"""
from fractions import Fraction

import numpy as np

KAPPA = 2.0 ** -18  # kFilterKappa


def _fp32_eval(p1, p2, p3, e, fused):
    f = np.float32

    def mul(a, b):
        return f(a) * f(b)

    def fms(a, b, c, d):  # a*b - c*d
        if fused:
            return f(np.float64(f(a)) * np.float64(f(b)) - np.float64(mul(c, d)))
        return mul(a, b) - mul(c, d)

    m12 = fms(p2[0], p3[1], p2[1], p3[0])
    m13 = fms(p2[0], p3[2], p2[2], p3[0])
    m14 = fms(p2[0], p3[3], p2[3], p3[0])
    m23 = fms(p2[1], p3[2], p2[2], p3[1])
    m24 = fms(p2[1], p3[3], p2[3], p3[1])
    m34 = fms(p2[2], p3[3], p2[3], p3[2])
    a = [f(v) for v in p1]
    X = -a[3] * m23 - a[1] * m34 + a[2] * m24
    Y = a[0] * m34 + a[3] * m13 - a[2] * m14
    Z = -a[0] * m24 + a[1] * m14 - a[3] * m12
    W = a[0] * m23 - a[1] * m13 + a[2] * m12
    ev = [f(v) for v in e]
    sv = ev[0] * X + ev[1] * Y + ev[2] * Z + ev[3] * W
    # the kernel's sums |u_i v_j| + |u_j v_i| as fma(|u_i|, |v_j|, |u_j| |v_i|)
    ab = lambda i, j: f(np.float64(abs(f(p2[i]))) * np.float64(abs(f(p3[j])))  # noqa: E731
                        + np.float64(abs(f(p2[j])) * abs(f(p3[i]))))
    a12, a13, a14, a23, a24, a34 = ab(0, 1), ab(0, 2), ab(0, 3), ab(1, 2), ab(1, 3), ab(2, 3)
    ax, ay, az, aw = (abs(v) for v in a)
    MX = aw * a23 + ay * a34 + az * a24
    MY = ax * a34 + aw * a13 + az * a14
    MZ = ax * a24 + ay * a14 + aw * a12
    MW = ax * a23 + ay * a13 + az * a12
    MXYZ = max(MX, MY, MZ)
    bound = f(KAPPA) * ((abs(ev[0]) + abs(ev[1]) + abs(ev[2])) * MXYZ + abs(ev[3]) * MW)
    return float(sv), float(bound)


def _exact_sign(p1, p2, p3, e):
    P1, P2, P3, E = ([Fraction(float(v)) for v in p] for p in (p1, p2, p3, e))
    m12 = P2[0] * P3[1] - P2[1] * P3[0]
    m13 = P2[0] * P3[2] - P2[2] * P3[0]
    m14 = P2[0] * P3[3] - P2[3] * P3[0]
    m23 = P2[1] * P3[2] - P2[2] * P3[1]
    m24 = P2[1] * P3[3] - P2[3] * P3[1]
    m34 = P2[2] * P3[3] - P2[3] * P3[2]
    X = -P1[3] * m23 - P1[1] * m34 + P1[2] * m24
    Y = P1[0] * m34 + P1[3] * m13 - P1[2] * m14
    Z = -P1[0] * m24 + P1[1] * m14 - P1[3] * m12
    W = P1[0] * m23 - P1[1] * m13 + P1[2] * m12
    s = E[0] * X + E[1] * Y + E[2] * Z + E[3] * W
    return (s > 0) - (s < 0)


def _plane(rng, scale):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    return np.array([n[0], n[1], n[2], rng.uniform(-scale, scale)])


def test_filter_never_trusts_a_wrong_sign():
    rng = np.random.default_rng(2024)
    trusted = flagged = 0
    for it in range(4000):
        scale = 10.0 ** rng.uniform(0, 2.5)   # cell sizes of a [0,1000] frame around its seed
        p1, p2, p3 = (_plane(rng, scale) for _ in range(3))
        # vertex of the three planes, then a test plane through (or very near) it
        A = np.array([p1[:3], p2[:3], p3[:3]])
        if abs(np.linalg.det(A)) < 1e-3:
            continue
        v = np.linalg.solve(A, -np.array([p1[3], p2[3], p3[3]]))
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        off = 10.0 ** rng.uniform(-9, 0) * scale * rng.choice([-1.0, 1.0])
        e = np.array([n[0], n[1], n[2], -(n @ v) + off])
        exact = _exact_sign(p1, p2, p3, e)
        for fused in (False, True):
            sv, bound = _fp32_eval(p1, p2, p3, e, fused)
            if abs(sv) > bound:
                trusted += 1
                assert (sv > 0) - (sv < 0) == exact, (it, fused, sv, bound)
            else:
                flagged += 1
    assert trusted > 1000 and flagged > 100
