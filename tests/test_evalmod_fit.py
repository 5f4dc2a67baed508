"""The reference's EvalMod polynomial (binboot_han_ki_discrete_multi_interval_
approximate, poly_approx.cpp:565-681), restated at 256-bit precision in
aes-walsh-ckks_amd/host/evalmod_fit.py, and the tables the context installs.

What pins it (the reference links MPFR, which this image lacks as a
development library, so its own output cannot be produced here):
* the fitted polynomial passes through the target at every Han-Ki node
  (Chebyshev points of each interval, computed here independently in
  float64), so the node placement, the divided differences, the
  Chebyshev-Lobatto resampling and the elimination are right together;
* on a small case the coefficients equal, double for double, the direct
  solution of the Chebyshev-Vandermonde system at the same nodes (mpmath
  LU at 256 bits), a different algorithm for the same interpolant;
* the doubles do not depend on the working precision (200, 256 and 320
  bits give the same), so MPFR's 256-bit run rounds to the same doubles;
* the validation grid of poly_approx.cpp:643-670, recomputed with numpy's
  Chebyshev evaluation, reproduces the recorded max error;
* with one double-angle step the fit composed with 4p(1-p) still matches
  the target on the grid.
The context's built-in table equals a fresh run of the restatement."""
import math
import os
import sys

import numpy as np
import pytest

from helpers import ROOT, host

sys.path.insert(0, os.path.join(ROOT, "aes-walsh-ckks_amd", "host"))
import evalmod_fit as E  # noqa: E402


def _nodes_float(K, degs, interval_radius):
    """Han-Ki nodes in float64: interval i (spacing s = 1/(2K), radius
    r = interval_radius / K) holds d_i Chebyshev points s (i + (r/s) cos(pi j / d_i))
    and their negatives; the centre interval its d_0 / 2 pairs and 0 when d_0 is odd."""
    s = 0.5 / K
    rl = (interval_radius / K) / s
    out = []
    if degs[0] % 2:
        out.append(0.0)
    for i in range(len(degs) - 1, 0, -1):
        for j in range(degs[i]):
            z = (i + math.cos(math.pi * 2 * j / (2 * degs[i])) * rl) * s
            out += [z, -z]
    for j in range(degs[0] // 2):
        z = math.cos(math.pi * 2 * j / (2 * degs[0])) * rl * s
        out += [z, -z]
    return np.array(out)


@pytest.mark.parametrize("name", ["AESWalsh15", "XBootParamAes13Reference"])
def test_fit_interpolates_target_at_han_ki_nodes(name):
    K, deg, da, rad, pts, prune = E.PRESETS[name]
    r = E.fit(K, deg, da, rad, pts, 0.0)
    nodes = _nodes_float(K, r["interval_degrees"], rad)
    assert len(nodes) == r["total_degree"]
    assert np.allclose(np.sort(nodes), np.sort(r["nodes"]), rtol=0, atol=1e-15)
    lo, hi = r["interval"]
    t = (2 * nodes - lo - hi) / (hi - lo)
    p = np.polynomial.chebyshev.chebval(t, np.array(r["coeffs"]))
    target = 0.5 - 0.5 * np.cos(2 * np.pi * K * nodes)
    assert np.max(np.abs(p - target)) < 1e-11


def _big_nodes(K, deg, rad):
    interval_k = int(round(2 * K))
    spacing = 0.5 / K
    dev = 1.0 / ((rad / K) / spacing)
    degs, total, _ = E.gen_degrees(deg, interval_k, dev)
    nodes, _ = E.gen_nodes(degs, dev, total, interval_k, spacing, 0, lambda z, s: z)
    return nodes


def test_fit_solve_equals_direct_chebyshev_vandermonde_solution():
    from mpmath import mp, mpf, matrix, lu_solve
    K, deg, rad = 2.0, 14, math.nextafter(0.25, 0.0)
    r = E.fit(K, deg, 0, rad, 16, 0.0)
    n = r["total_degree"]
    znodes = _big_nodes(K, deg, rad)
    assert len(znodes) == n
    a = mpf(r["interval"][1])
    A, b = matrix(n, n), matrix(n, 1)
    for i, z in enumerate(znodes):
        t = z / a
        T0, T1 = mpf(1), t
        A[i, 0] = T0
        A[i, 1] = T1
        for j in range(2, n):
            T0, T1 = T1, 2 * t * T1 - T0
            A[i, j] = T1
        b[i] = mpf(0.5) - mpf(0.5) * mp.cos(2 * mp.pi * mpf(K) * z)
    c = lu_solve(A, b)
    direct = [E.to_double(c[j]) for j in range(n)]
    for j in range(1, n, 2):
        direct[j] = 0.0
    assert direct == r["coeffs"]


def test_fit_doubles_do_not_depend_on_precision():
    from mpmath import mp
    prm = E.PRESETS["AESWalsh15"]
    base = E.fit(*prm)["coeffs"]
    try:
        for prec in (200, 320):
            mp.prec = prec
            assert E.fit(*prm)["coeffs"] == base, prec
    finally:
        mp.prec = 256


def test_validation_grid_error_recomputed():
    K, deg, da, rad, pts, prune = E.PRESETS["AESWalsh15"]
    r = E.fit(K, deg, da, rad, pts, prune)
    ik = int(round(2 * K))
    centers = 0.5 * np.arange(-(ik - 1), ik)
    alpha = np.arange(pts) / (pts - 1)
    x = (centers[:, None] + rad * (2 * alpha[None, :] - 1)).ravel()
    u = x / K
    lo, hi = r["interval"]
    p = np.polynomial.chebyshev.chebval((2 * u - lo - hi) / (hi - lo), np.array(r["coeffs"]))
    err = np.max(np.abs(p - (0.5 - 0.5 * np.cos(2 * np.pi * u * K))))
    assert abs(err - r["max_error"]) < 1e-12
    assert err < 3e-9
    # odd terms zero; nothing below the prune threshold survives
    c = np.array(r["coeffs"])
    assert np.all(c[1::2] == 0) and np.all((c == 0) | (np.abs(c) >= prune))


def test_double_angle_fit_composes_to_target():
    K = 3.0
    r = E.fit(K, 30, 1, math.nextafter(0.25, 0.0), 32, 0.0)
    assert r["interval"] == (-0.5, 0.5)
    assert r["max_error"] < 1e-3, r["max_error"]


def test_context_tables_equal_the_restatement():
    H = host()
    P = H.CKKSParams.AESWalsh15()
    assert H.evalmod_polynomial_is_reference_fit(P)
    assert list(H.default_evalmod_polynomial(P)) == E.fit(*E.PRESETS["AESWalsh15"])["coeffs"]
    # double-angle parameters without a built-in fit: no default polynomial
    c = H.aes_walsh15_config()
    bp = c.bootstrap_params
    bp.evalmod_double_angle = 1
    c.bootstrap_params = bp
    P2 = H.CKKSParams(c)
    assert not H.evalmod_polynomial_is_reference_fit(P2)
    assert list(H.default_evalmod_polynomial(P2)) == []
