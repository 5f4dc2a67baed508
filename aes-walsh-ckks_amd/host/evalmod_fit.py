"""The reference's EvalMod polynomial fit, restated at MPFR's precision.

binboot_han_ki_discrete_multi_interval_approximate (poly_approx.cpp:565-681)
fits 0.5 - 0.5 cos(2 pi K u) on 2K - 1 small intervals around the points
u = c / (2K) (Han and Ki, CT-RSA 2020): per-interval degrees from an error
bound (han_ki_gen_degrees, :196-307), Chebyshev nodes on each interval
(han_ki_gen_nodes, :309-356), Newton divided differences, the interpolant
sampled at Chebyshev-Lobatto points and a Gauss elimination for its
Chebyshev coefficients on [-K s, K s] / 2^r (han_ki_solve_system,
:358-458), all in 256-bit MPFR; the doubles are then made even (odd terms
zeroed) and pruned below a threshold.  The reference links MPFR (absent
from this image as a development library), so the fit is restated here with
mpmath at 256 bits, round-to-nearest, in the reference's operation order:
additions, products and quotients are correctly rounded in both, and every
cosine and pi is computed at 512 bits and rounded once, which is MPFR's
correctly rounded value except for a midpoint at 2^-256.

Running this file regenerates ckks_evalmod_tables.inc (the coefficient
tables the context installs for the reference's presets):

    python3 aes-walsh-ckks_amd/host/evalmod_fit.py
"""
import math
import os
import sys

import mpmath
from mpmath import mp, mpf
from mpmath.libmp import to_float

mp.prec = 256
PI_D = 3.141592653589793238462643383279502884  # pi() (poly_approx.cpp:160-162), a double


def big(v):
    return mpf(v)


def big_pi():
    with mp.workprec(512):
        v = +mp.pi
    return +v


def big_cos(x):
    with mp.workprec(512):
        v = mpmath.cos(x)
    return +v


def big_pow2(e):
    return mpmath.ldexp(mpf(1), e)


def to_double(x):
    """mpfr_get_d(RNDN)"""
    return to_float(x._mpf_, rnd="n")


def gen_degrees(degree, interval_k, dev):
    """han_ki_gen_degrees (poly_approx.cpp:196-307), in doubles."""
    degree_bound = degree + 1
    total = 2 * interval_k - 1
    error = 1.0 / dev
    degs = [1] * interval_k
    bounds = [0.0] * interval_k
    temp = 0.0
    for i in range(1, 2 * interval_k):
        temp -= math.log2(float(i))
    temp += (2.0 * float(interval_k) - 1.0) * math.log2(2.0 * PI_D)
    temp += math.log2(error)
    for i in range(interval_k):
        bounds[i] = temp
        for j in range(1, interval_k - i):
            bounds[i] += math.log2(float(j) + error)
        for j in range(1, interval_k + i):
            bounds[i] += math.log2(float(j) + error)
    min_gap = math.inf
    for _ in range(200):
        if total >= degree_bound:
            break
        mx = max(bounds)
        m = bounds.index(mx)  # std::max_element: the first maximum
        rest = [b for k, b in enumerate(bounds) if k != m]
        if rest:
            min_gap = min(min_gap, mx - max(rest))
        if m != 0:
            if total + 2 > degree_bound:
                break
            for i in range(interval_k):
                bounds[i] -= math.log2(float(total + 1))
                bounds[i] -= math.log2(float(total + 2))
                bounds[i] += 2.0 * math.log2(2.0 * PI_D)
                if i != m:
                    bounds[i] += math.log2(abs(float(i - m)) + error)
                    bounds[i] += math.log2(float(i + m) + error)
                else:
                    bounds[i] += math.log2(error) - 1.0
                    bounds[i] += math.log2(2.0 * float(i) + error)
            total += 2
        else:
            bounds[0] -= math.log2(float(total + 1))
            bounds[0] += math.log2(error) - 1.0
            bounds[0] += math.log2(2.0 * PI_D)
            for i in range(1, interval_k):
                bounds[i] -= math.log2(float(total + 1))
                bounds[i] += math.log2(2.0 * PI_D)
                bounds[i] += math.log2(float(i) + error)
            total += 1
        degs[m] += 1
    return degs, total, min_gap


def gen_nodes(degs, dev, total, interval_k, spacing_d, double_angle, target):
    """han_ki_gen_nodes (poly_approx.cpp:309-356)."""
    interval_size = big(1) / big(dev)
    spacing = big(spacing_d)
    nodes = [mpf(0)] * total
    count = 1 if degs[0] & 1 else 0
    for i in range(interval_k - 1, 0, -1):
        two_degree = big(2 * degs[i])
        for j in range(degs[i]):
            theta = big_pi() * big(2 * j)
            theta /= two_degree
            local = big_cos(theta) * interval_size
            nodes[count] = (big(i) + local) * spacing
            count += 1
            nodes[count] = -nodes[count - 1]
            count += 1
    two_degree_center = big(2 * degs[0])
    for j in range(degs[0] // 2):
        theta = big_pi() * big(2 * j)
        theta /= two_degree_center
        local = big_cos(theta) * interval_size
        nodes[count] = local * spacing
        count += 1
        nodes[count] = -nodes[count - 1]
        count += 1
    return nodes, [target(z, double_angle) for z in nodes]


def solve_system(total, interval_k, double_angle, spacing_d, nodes, y_in):
    """han_ki_solve_system (poly_approx.cpp:358-458)."""
    y = list(y_in)
    for j in range(1, total):
        for i in range(total - j):
            y[i] = y[i + 1] - y[i]
            tmp = nodes[i + j] - nodes[i]
            y[i] /= tmp
    n = total + 1
    scale_factor = big_pow2(double_angle)
    k_over_scale = big(interval_k) * big(spacing_d) / scale_factor
    x = []
    for i in range(n):
        theta = big_pi() * big(i)
        theta /= big(n - 1)
        x.append(k_over_scale * big_cos(theta))
    p = []
    for i in range(n):
        v = y[0]
        for j in range(1, n - 1):
            v *= (x[i] - nodes[j])
            v += y[j]
        p.append(v)
    T = []
    for i in range(n):
        row = [mpf(0)] * n
        row[0] = big(1)
        row[1] = x[i] / k_over_scale
        for j in range(2, n):
            xn = x[i] / k_over_scale
            row[j] = big(2) * xn * row[j - 1] - row[j - 2]
        T.append(row)
    for i in range(n - 1):
        mi, mabs = i, abs(T[i][i])
        for j in range(i + 1, n):
            cand = abs(T[j][i])
            if cand > mabs:
                mabs, mi = cand, j
        if i != mi:
            T[i], T[mi] = T[mi], T[i]
            p[i], p[mi] = p[mi], p[i]
        piv = T[i][i]
        Ti = T[i]
        for j in range(i + 1, n):
            Ti[j] /= piv
        p[i] /= piv
        Ti[i] = big(1)
        for j in range(i + 1, n):
            Tj = T[j]
            f = Tj[i]
            tmp = f * p[i]
            p[j] -= tmp
            for l in range(i + 1, n):
                tmp = f * Ti[l]
                Tj[l] -= tmp
            Tj[i] = mpf(0)
    c = [mpf(0)] * n
    c[n - 1] = p[n - 1]
    for i in range(n - 2, -1, -1):
        v = p[i]
        for j in range(i + 1, n):
            v -= T[i][j] * c[j]
        c[i] = v
    return c[:total]


def chebyshev_series(coeffs, t):
    """evaluate_chebyshev_series (poly_approx.cpp:173-189): Clenshaw in doubles."""
    b1 = b2 = 0.0
    for k in range(len(coeffs), 1, -1):
        b = 2.0 * t * b1 - b2 + coeffs[k - 1]
        b2, b1 = b1, b
    return 0.0 if not coeffs else coeffs[0] + t * b1 - b2


def evalmod_target(u, period):
    """binboot_evalmod_target (poly_approx.cpp:556-561)"""
    return 0.5 * (1.0 - math.cos(2.0 * PI_D * u / period))


def fit(K, degree, double_angle, interval_radius, points, prune):
    """binboot_han_ki_discrete_multi_interval_approximate (poly_approx.cpp:565-681).
    Returns the coefficients (doubles), the interval, the validation max
    error and diagnostics (pruned count, smallest degree-allocation gap)."""
    if not (K > 0) or double_angle > 30 or not (0.0 < interval_radius < 0.25) or points < 2 or prune < 0:
        raise ValueError("invalid EvalMod fit parameters")
    interval_k = int(round(2.0 * K))
    spacing = 0.5 / K
    norm_radius = interval_radius / K
    # han_ki_discrete_cosine_approximate_impl (:460-538)
    local_radius = norm_radius / spacing
    dev = 1.0 / local_radius
    degs, total, gap = gen_degrees(degree, interval_k, dev)

    def target(z, scnum):
        scaled = z / big_pow2(scnum)
        return big(0.5) - big(0.5) * big_cos(big(2) * big_pi() * big(float(K)) * scaled)

    nodes, values = gen_nodes(degs, dev, total, interval_k, spacing, double_angle, target)
    big_coeffs = solve_system(total, interval_k, double_angle, spacing, nodes, values)
    coeffs = [to_double(c) for c in big_coeffs]
    scale = math.pow(2.0, float(double_angle))
    imin = -spacing * float(interval_k) / scale
    imax = spacing * float(interval_k) / scale
    for i in range(1, len(coeffs), 2):
        coeffs[i] = 0.0
    pruned = 0
    for i, c in enumerate(coeffs):
        if c != 0.0 and abs(c) < prune:
            coeffs[i] = 0.0
            pruned += 1
    max_error = 0.0
    half = interval_k - 1
    for ci in range(-half, half + 1):
        cx = 0.5 * float(ci)
        for j in range(points):
            a = float(j) / float(points - 1)
            xv = cx + interval_radius * (2.0 * a - 1.0)
            u = xv / K
            expected = evalmod_target(u, 1.0 / K)
            t = (2.0 * u - imin - imax) / (imax - imin)
            actual = chebyshev_series(coeffs, t)
            for _ in range(double_angle):
                actual = 4.0 * actual * (1.0 - actual)
            max_error = max(max_error, abs(actual - expected))
    return {"coeffs": coeffs, "interval": (imin, imax), "max_error": max_error, "pruned": pruned,
            "total_degree": total, "interval_degrees": degs, "degree_gap": gap,
            "nodes": [to_double(z) for z in nodes], "big_coeffs": big_coeffs}


RADIUS = math.nextafter(0.25, 0.0)
# (K, degree, double angle, interval radius, validation points, prune threshold)
# of the reference's presets (ckks_aes_params.hpp:12-211; walsh15's K from
# its Irwin-Hall bound at failure probability 2^-16, binboot_k_bound.cpp:54-89)
PRESETS = {
    "AESWalsh15": (11.0, 100, 0, RADIUS, 128, 1e-10),
    "XBootParamAes13Reference": (7.0, 60, 0, RADIUS, 64, 0.0),
    "AESXBoot14Secure": (7.0, 60, 0, RADIUS, 64, 0.0),
    "AESXBoot15": (12.0, 120, 0, RADIUS, 64, 0.0),
}


def emit(path):
    seen = {}
    lines = ["// Generated by evalmod_fit.py: the reference's EvalMod polynomial fit",
             "// (binboot_han_ki_discrete_multi_interval_approximate, poly_approx.cpp:565-681)",
             "// at 256-bit precision for the parameters of its presets.  Do not edit.",
             "// {K, degree, double_angle, interval_radius, validation points, prune threshold,",
             "//  interval_min, interval_max, validation max error, coefficients}"]
    for name, prm in PRESETS.items():
        if prm in seen:
            continue
        r = fit(*prm)
        seen[prm] = r
        K, d, da, rad, pts, pr = prm
        cs = ", ".join(float.hex(c) for c in r["coeffs"])
        lines.append(f"// {name}: {r['total_degree']} nodes, pruned {r['pruned']}, "
                     f"validation max error {r['max_error']:.3e}")
        lines.append(f"{{{float.hex(K)}, {d}, {da}, {float.hex(rad)}, {pts}, {float.hex(pr)}, "
                     f"{float.hex(r['interval'][0])}, {float.hex(r['interval'][1])}, {float.hex(r['max_error'])}, "
                     f"{{{cs}}}}},")
        print(f"{name}: max error {r['max_error']:.3e}, pruned {r['pruned']}, "
              f"degree gap {r['degree_gap']:.3e}", file=sys.stderr)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    emit(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ckks_evalmod_tables.inc"))
