"""Input generators for the benchmark and test workloads.

Surface meshes are written as Gmsh 2.2 ASCII files in the paired-face form the
solver reads (setup2.c:205-503): every face appears twice, first with the
interior domain id as physical tag and the outward node order, then with the
open-domain tag 99 and the reversed node order.  Also writes the medium file
(setup2.c:153-193) and the plane-wave file (multi_fbeam.c:145-238).
"""
import math
import os

import numpy as np


def icosphere(level, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron with vertices projected to the sphere.
    Level L has 20*4^L triangles (L=2: 320, L=4: 5120, L=5: 20480)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7),
         (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    F = np.array(faces, dtype=np.int64)
    return orient_outward(V, F)


def box(edge, n, center=(0.0, 0.0, 0.0)):
    """Cube of the given edge, each face cut into n x n squares, 2 triangles
    each (12 n^2 triangles)."""
    pts = {}
    verts, faces = [], []

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in pts:
            pts[key] = len(verts)
            verts.append(np.array(p, dtype=np.float64))
        return pts[key]

    h = edge / 2.0
    g = np.linspace(-h, h, n + 1)
    for axis in range(3):
        for side in (-h, h):
            for i in range(n):
                for j in range(n):
                    q = []
                    for (a, b) in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)):
                        p = [0.0, 0.0, 0.0]
                        p[axis] = side
                        p[(axis + 1) % 3] = g[a]
                        p[(axis + 2) % 3] = g[b]
                        q.append(vid(p))
                    faces += [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
    V = np.array(verts) + np.asarray(center, dtype=np.float64)
    return orient_outward(V, np.array(faces, dtype=np.int64))


def cubesphere(radius, n, center=(0.0, 0.0, 0.0)):
    """Sphere from a cube whose faces are cut into n x n squares, vertices
    projected onto the sphere (12 n^2 triangles: any count of that form, where
    an icosphere offers only 20 x 4^level)."""
    V, F = box(2.0, n)
    V = V / np.linalg.norm(V, axis=1)[:, None] * radius + np.asarray(center, dtype=np.float64)
    return orient_outward(V, F)


def cone(radius, height, n_ring, n_side, center=(0.0, 0.0, 0.0)):
    """Right circular cone (apex up) with a flat base: n_ring azimuthal
    segments, n_side rings along the slant and the base."""
    verts, faces = [], []
    apex = len(verts)
    verts.append(np.array([0.0, 0.0, height / 2.0]))
    rings = []
    for k in range(1, n_side + 1):
        fr = k / n_side
        z = height / 2.0 - fr * height
        ring = []
        for a in range(n_ring):
            th = 2 * math.pi * a / n_ring + (0.5 * math.pi / n_ring) * (k % 2)
            ring.append(len(verts))
            verts.append(np.array([fr * radius * math.cos(th), fr * radius * math.sin(th), z]))
        rings.append(ring)
    for a in range(n_ring):
        faces.append((apex, rings[0][a], rings[0][(a + 1) % n_ring]))
    for k in range(n_side - 1):
        r0, r1 = rings[k], rings[k + 1]
        for a in range(n_ring):
            b = (a + 1) % n_ring
            faces += [(r0[a], r1[a], r1[b]), (r0[a], r1[b], r0[b])]
    base = rings[-1]
    zb = -height / 2.0
    inner = [base]
    for k in range(1, n_side):
        fr = 1.0 - k / n_side
        ring = []
        for a in range(n_ring):
            th = 2 * math.pi * a / n_ring + (0.5 * math.pi / n_ring) * ((n_side + k) % 2)
            ring.append(len(verts))
            verts.append(np.array([fr * radius * math.cos(th), fr * radius * math.sin(th), zb]))
        inner.append(ring)
    centre = len(verts)
    verts.append(np.array([0.0, 0.0, zb]))
    for k in range(len(inner) - 1):
        r0, r1 = inner[k], inner[k + 1]
        for a in range(n_ring):
            b = (a + 1) % n_ring
            faces += [(r0[a], r1[a], r1[b]), (r0[a], r1[b], r0[b])]
    last = inner[-1]
    for a in range(n_ring):
        faces.append((centre, last[(a + 1) % n_ring], last[a]))
    V = np.array(verts) + np.asarray(center, dtype=np.float64)
    return orient_outward(V, np.array(faces, dtype=np.int64))


def orient_outward(V, F):
    """Order each triangle so that (r1-r0)x(r2-r0) points away from the
    centroid of the (star-shaped, convex here) body."""
    c = V.mean(axis=0)
    F = F.copy()
    for i, (a, b, d) in enumerate(F):
        n = np.cross(V[b] - V[a], V[d] - V[a])
        if np.dot(n, (V[a] + V[b] + V[d]) / 3.0 - c) < 0:
            F[i] = (a, d, b)
    return V, F


def write_msh(path, V, F, domain=1, group=1):
    """Gmsh 2.2 ASCII in the paired-face form."""
    with open(path, "w") as fp:
        fp.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(V))
        for i, p in enumerate(V):
            fp.write("%d %.17g %.17g %.17g\n" % (i + 1, p[0], p[1], p[2]))
        fp.write("$EndNodes\n$Elements\n%d\n" % (2 * len(F)))
        for i, (a, b, c) in enumerate(F):
            fp.write("%d 2 2 %d %d %d %d %d\n" % (2 * i + 1, domain, group, a + 1, b + 1, c + 1))
            fp.write("%d 2 2 99 %d %d %d %d\n" % (2 * i + 2, group, a + 1, c + 1, b + 1))
        fp.write("$EndElements\n")


def write_medium(path, n_list):
    with open(path, "w") as fp:
        fp.write("## medium data file for bem3_emf_b2\n# number of medium\n%d\n" % len(n_list))
        fp.write("# real(refractive_index) imag(refractive_index)\n")
        for n in n_list:
            fp.write("%.17g %.17g\n" % (complex(n).real, complex(n).imag))


def write_ipw(path, beams):
    """beams: rows of (lambda0, ni, power, re e0x, im e0x, re e0y, im e0y, fx, fy, fz, theta, phi)"""
    with open(path, "w") as fp:
        fp.write("## plane wave\n# number of waves\n%d\n" % len(beams))
        fp.write("#lambda0 ni power real(e0x) imag(e0x) real(e0y) imag(e0y) fx fy fz theta phi\n")
        for b in beams:
            fp.write(" ".join("%.17g" % x for x in b) + "\n")


def xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0):
    """x-polarised plane wave along +z with amplitude E0 (power = n0 E0^2 / 2)."""
    return np.array([[lambda0, n0, 0.5 * n0 * E0 * E0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]])


def paired_arrays(V, F, domain=1):
    """In-memory equivalent of write_msh + the reader's pairing/swap:
    node table [Nn+1,3], element table [Ne+1,4] (open side as main side),
    main/sub domain ids.  Index 0 rows are unused (1-based tables)."""
    Nn, Ne = len(V), len(F)
    rn = np.zeros((Nn + 1, 3))
    rn[1:] = V
    ed = np.zeros((Ne + 1, 4), np.int32)
    ed[1:, 0] = F[:, 0] + 1
    ed[1:, 1] = F[:, 2] + 1
    ed[1:, 2] = F[:, 1] + 1
    md = np.zeros(Ne + 1, np.int32)
    sd = np.zeros(Ne + 1, np.int32)
    sd[1:] = domain
    gd = np.ones(Ne + 1, np.int32)
    gd[0] = 0
    return rn, ed, md, sd, gd


def scene16(seed=2024, box_um=10.0, n_obj=16):
    """Seeded uneven multi-object scene: spheres, cubes and cones of size
    U(0.3,1.5) um with 500..5000 triangles each (the nearest count the
    mesher offers inside that range), refractive index U(1.4,1.7), placed
    without overlap (bounding spheres) in a cube.  'V' is centred at the
    origin; 'center' is the object's position."""
    rng = np.random.default_rng(seed)
    objs = []
    centers = []
    while len(objs) < n_obj:
        kind = rng.choice(["sphere", "cube", "cone"])
        size = rng.uniform(0.3, 1.5)
        ntri = int(rng.integers(500, 5001))
        nidx = rng.uniform(1.4, 1.7)
        rb = 0.5 * size * (math.sqrt(3.0) if kind == "cube" else 1.2)
        for _ in range(1000):
            c = rng.uniform(-box_um / 2 + rb, box_um / 2 - rb, 3)
            if all(np.linalg.norm(c - c2) > rb + r2 + 0.05 for c2, r2 in centers):
                break
        else:
            continue
        centers.append((c, rb))
        # 12 n^2 triangles in [500, 5000]: n in [7, 20]; cones 2m (4 (m // 2) - 2)
        if kind == "sphere":
            V, F = cubesphere(0.5 * size, min(20, max(7, int(round(math.sqrt(ntri / 12.0))))))
        elif kind == "cube":
            V, F = box(size, min(20, max(7, int(round(math.sqrt(ntri / 12.0))))))
        else:
            m = min(35, max(12, int(round(math.sqrt(ntri / 4.0)))))
            V, F = cone(0.5 * size, size, 2 * m, max(1, m // 2))
        objs.append(dict(kind=str(kind), size=size, n=nidx, center=c, V=V, F=F))
    return objs


def scene_beams():
    """Four objects for many incident beams: two spheres of radius 1 um
    (2028 triangles each), a cube of edge 1.5 um (3072) and a cone of base
    radius 0.75 um and height 1.5 um (1520), n = 1.59, centres 3 um from
    the origin along -x, +x, +y and -y.  'V' is centred at the origin;
    'center' is the object's position."""
    parts = [("sphere", cubesphere(1.0, 13), (-3.0, 0.0, 0.0)),
             ("sphere", cubesphere(1.0, 13), (3.0, 0.0, 0.0)),
             ("cube", box(1.5, 16), (0.0, 3.0, 0.0)),
             ("cone", cone(0.75, 1.5, 40, 10), (0.0, -3.0, 0.0))]
    return [dict(kind=k, n=1.59, center=np.array(c), V=VF[0], F=VF[1]) for k, VF, c in parts]


def random_plane_waves(nb, seed, lambda0=1.064, n0=1.33, power=1.0):
    """nb plane waves (ipw.txt rows) with directions uniform on the sphere
    (theta = arccos(1 - 2u), phi = 2 pi v) and polarisation angle psi
    uniform in [0, pi): e0 = (cos psi, sin psi) in the wave's own frame"""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nb):
        th = math.acos(1.0 - 2.0 * rng.uniform())
        ph = 2.0 * math.pi * rng.uniform()
        psi = math.pi * rng.uniform()
        out.append([lambda0, n0, power, math.cos(psi), 0.0, math.sin(psi), 0.0, 0.0, 0.0, 0.0, th, ph])
    return np.array(out)


def ensure_dir(p):
    os.makedirs(p, exist_ok=True)
    return p
