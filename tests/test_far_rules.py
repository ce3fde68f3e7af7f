"""The far-pair rules of the mixed-precision form (bem_dev.h far_rule): the
element radius is the largest distance from the triangle's centre to a
vertex; pairs farther than 4 radii take 5x5 Gauss-Legendre squares of the
3-part triangle map while |k| x radius <= 1.2, 6x6 up to 3, 12x12 beyond;
pairs farther than 8 radii with |k| x radius <= 0.5 take 4x4.  Checked
here: over general triangles from random vertices (a third of them skewed
slivers), random directions, normals and distances (4 to 32 radii; 8 to 64
for 4x4), the relative error of each group G_i, H_i, dG_i/dv, dH_i/dv (the
integrands of lit_dcoef_v2.c:404-412) against a 20x20 rule stays below
3e-8 at each tier's largest |k| x radius (about 2e-8 is the largest value
seen over thousands of pairs per tier), for real and lossy wavenumbers,
and the 12x12 rule of the FP64 form agrees with the 20x20 rule to
round-off there."""
import numpy as np
import pytest

S3 = np.sqrt(3.0)


def tri_rule(n):
    """the 3-part map of the equilateral reference triangle onto [-1,1]^2
    squares (lit_dcoef_v2.c:223-235)"""
    x, w = np.polynomial.legendre.leggauss(n)
    a, b = np.meshgrid(x, x, indexing="ij")
    wt = ((1 + 0.25 * a + 0.25 * b) * np.outer(w, w)).ravel()
    a, b = a.ravel(), b.ravel()
    z = np.concatenate([0.125 * (3 + 2 * a + 2 * b + a * b), 0.0625 * (-3 + a - 5 * b - a * b),
                        0.0625 * (-3 - 5 * a + b - a * b)])
    e = np.concatenate([0.125 * S3 * (-a + b), 0.0625 * S3 * (3 + 3 * a + b + a * b),
                        0.0625 * S3 * (-3 - a - 3 * b - a * b)])
    return z, e, np.concatenate([wt, wt, wt])


RULES = {n: tri_rule(n) for n in (4, 5, 6, 12, 20)}


def groups(n, c1, c2, W, x, v, k):
    z, e, w = RULES[n]
    d = c1[:, None] * z + c2[:, None] * e - x[:, None]
    R = np.linalg.norm(d, axis=0)
    ikr = 1j * k * R
    ce = np.exp(ikr)
    rdW = (d * W[:, None]).sum(0) / R
    rd = (d * v[:, None]).sum(0) / R
    Nw = np.array([(1 - 2.5 * (z + S3 * e)) / 3, (1 + 5 * z) / 3, (1 - 2.5 * (z - S3 * e)) / 3]) * w
    qq = (k * R) ** 2 + 3 * (ikr - 1)
    return [(Nw * ce / R).sum(1),
            (Nw * (ikr - 1) * ce * rdW / R ** 2).sum(1),
            (Nw * (ikr - 1) * ce / R ** 2 * rd).sum(1),
            (Nw * ce / R ** 3 * (qq * rd * rdW - (ikr - 1) * (W @ v))).sum(1)]


NEAR = (4.0, 5.0, 6.0, 8.0, 12.0, 16.0, 32.0)   # element radii from the centre
FAR = (8.0, 12.0, 16.0, 32.0, 64.0)


def triangle(rng):
    """c1, c2 of r = c0 + c1 z + c2 e for a random triangle (origin at its
    centre c0) and its radius (largest centre-to-vertex distance)"""
    V = rng.standard_normal((3, 3))
    if rng.uniform() < 1.0 / 3.0:   # sliver: third vertex near the first edge
        V[2] = V[0] + (V[1] - V[0]) * rng.uniform(0.1, 0.9) + rng.standard_normal(3) * rng.uniform(0.02, 0.2)
    c0 = V.mean(0)
    c1 = V[1] - c0
    c2 = (V[2] - V[0]) / S3
    rad = max(np.linalg.norm(V[i] - c0) for i in range(3))
    return c1, c2, rad


@pytest.mark.parametrize("n,kR,ratios,bound", [(5, 0.2, NEAR, 3e-8), (5, 1.2, NEAR, 3e-8), (6, 3.0, NEAR, 3e-8),
                                               (12, 3.0, NEAR, 1e-13), (4, 0.2, FAR, 3e-8), (4, 0.5, FAR, 3e-8)])
def test_far_rule_error(n, kR, ratios, bound):
    rng = np.random.default_rng(int(100 * kR) + n)
    worst = 0.0
    for ratio in ratios:
        for _ in range(120):
            c1, c2, rad = triangle(rng)
            W = np.cross(c1, c2)
            u = rng.standard_normal(3)
            x = u / np.linalg.norm(u) * ratio * rad * 1.0001
            v = rng.standard_normal(3)
            v /= np.linalg.norm(v)
            k = kR / rad * np.exp(1j * rng.uniform(0, 0.5 * np.pi))   # lossy media too, |k| fixed
            ref = groups(20, c1, c2, W, x, v, k)
            got = groups(n, c1, c2, W, x, v, k)
            for g, r in zip(got, ref):
                worst = max(worst, np.abs(g - r).max() / np.abs(r).max())
    assert worst < bound, worst


def test_element_radius_is_vertex_distance():
    """c0 + c1, c0 - c1/2 -+ sqrt3/2 c2 are the vertices (the map of the
    reference triangle's corners (1, 0), (-1/2, -+sqrt3/2)); the radius the
    rules use is the largest of their distances from c0"""
    rng = np.random.default_rng(3)
    for _ in range(50):
        c1, c2, rad = triangle(rng)
        verts = [c1, -0.5 * c1 - 0.5 * S3 * c2, -0.5 * c1 + 0.5 * S3 * c2]
        assert abs(max(np.linalg.norm(v) for v in verts) - rad) < 1e-12 * rad
