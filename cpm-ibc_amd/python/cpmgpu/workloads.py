"""Synthetic workloads of the benchmark configurations (geometry and data only)."""
import numpy as np


def torus_knot_polyline(R=1.0, r=0.4, segments=2000):
    """Closed polyline sampling the (2,3) torus knot ((R + r cos 3t) cos 2t, (R + r cos 3t) sin 2t, r sin 3t)
    at t_i = 2 pi i / segments; the last vertex repeats the first."""
    t = 2 * np.pi * np.arange(segments + 1) / segments
    V = np.stack([(R + r * np.cos(3 * t)) * np.cos(2 * t), (R + r * np.cos(3 * t)) * np.sin(2 * t), r * np.sin(3 * t)], 1)
    V[-1] = V[0]
    return V


def icosphere(subdivisions=0, radius=1.0):
    """Unit icosahedron refined `subdivisions` times by edge midpoints projected to the sphere:
    20 * 4**subdivisions triangles, outward (counter-clockwise) orientation."""
    g = (1 + 5 ** 0.5) / 2
    V = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
         (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def m(a, b):
            k = (a, b) if a < b else (b, a)
            if k not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[k] = len(V) - 1
            return mid[k]

        F = [f for a, b, c in F for f in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                                          (m(a, b), m(b, c), m(c, a)))]
    return radius * np.array(V), np.array(F, np.int32)


def real_sph_harm(lmax, pts):
    """Real spherical harmonics Y_lm (l <= lmax, |m| <= l) at the directions of pts, orthonormal on the
    unit sphere, from the associated-Legendre recurrence; columns ordered (l, m = -l..l)."""
    P = np.asarray(pts, float)
    r = np.linalg.norm(P, axis=1)
    ct = P[:, 2] / r
    st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
    ph = np.arctan2(P[:, 1], P[:, 0])
    # normalised associated Legendre functions bar P_l^m(ct) (no Condon-Shortley phase)
    Pl = {(0, 0): np.full_like(ct, np.sqrt(1 / (4 * np.pi)))}
    for m in range(1, lmax + 1):
        Pl[(m, m)] = np.sqrt((2 * m + 1) / (2 * m)) * st * Pl[(m - 1, m - 1)]
    for m in range(0, lmax):
        Pl[(m + 1, m)] = np.sqrt(2 * m + 3) * ct * Pl[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            a = np.sqrt((4 * l * l - 1) / (l * l - m * m))
            b = np.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
            Pl[(l, m)] = a * (ct * Pl[(l - 1, m)] - b * Pl[(l - 2, m)])
    cols = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            if m < 0:
                cols.append(np.sqrt(2) * Pl[(l, -m)] * np.sin(-m * ph))
            elif m == 0:
                cols.append(Pl[(l, 0)])
            else:
                cols.append(np.sqrt(2) * Pl[(l, m)] * np.cos(m * ph))
    return np.stack(cols, 0).T      # column-major: stacking rows is contiguous


def displaced_icosphere(subdivisions=6, amplitude=0.1, lmax=6, seed=4):
    """Icosphere with radial displacement 1 + amplitude g / max|g|, g = sum_{l <= lmax} b_lm Y_lm,
    b_lm ~ N(0, 1) from numpy's default_rng(seed): the mesh surface of the serving workload
    (subdivisions 6: 81,920 triangles)."""
    V, F = icosphere(subdivisions)
    Y = real_sph_harm(lmax, V)
    b = np.random.default_rng(seed).standard_normal(Y.shape[1])
    g = Y @ b
    return V * (1 + amplitude * g / np.abs(g).max())[:, None], F


def box_mesh(lo=-1.0, hi=1.0):
    """Axis-aligned cube [lo, hi]^3 as 12 outward triangles."""
    c = np.array([[x, y, z] for x in (lo, hi) for y in (lo, hi) for z in (lo, hi)], float)
    F = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1], [2, 3, 7], [2, 7, 6],
                  [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.int32)
    return c, F


class MT19937_64:
    """The 64-bit Mersenne Twister with the parameters of C++'s std::mt19937_64
    (Matsumoto and Nishimura; w 64, n 312, m 156, r 31), so that seeded streams
    match a C++ caller's."""
    _N, _M = 312, 156
    _MASK = (1 << 64) - 1
    _UPPER, _LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed=5489):
        mt = [seed & self._MASK]
        for i in range(1, self._N):
            mt.append((6364136223846793005 * (mt[-1] ^ (mt[-1] >> 62)) + i) & self._MASK)
        self.mt, self.i = mt, self._N

    def _twist(self):
        mt, N, M = self.mt, self._N, self._M
        for k in range(N):
            y = (mt[k] & self._UPPER) | (mt[(k + 1) % N] & self._LOWER)
            v = mt[(k + M) % N] ^ (y >> 1)
            if y & 1:
                v ^= 0xB5026F5AA96619E9
            mt[k] = v
        self.i = 0

    def __call__(self):
        if self.i >= self._N:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self._MASK

    def uniform(self):
        """Uniform double in [0, 1): the top 53 bits."""
        return (self() >> 11) * (1.0 / (1 << 53))

    def normals(self, n):
        """n standard normals by the Box-Muller transform of pairs (u1, u2), u1 in (0, 1]:
        sqrt(-2 ln u1) cos(2 pi u2) and sqrt(-2 ln u1) sin(2 pi u2)."""
        out = []
        while len(out) < n:
            u1 = 1.0 - self.uniform()
            u2 = self.uniform()
            rad = np.sqrt(-2.0 * np.log(u1))
            out += [rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)]
        return np.array(out[:n])


def multi_surface_set(seed=3):
    """The 8 surfaces of the multi-tube batch: spheres r = 0.5, 1, 1.5, 2; tori (R, r) = (1, 0.3),
    (1.5, 0.5); two ellipsoids with semi-axes U[0.5, 1.5] from MT19937_64(seed).  Entries
    (label, surface type, params, bbox)."""
    g = MT19937_64(seed)
    out = [(f"sphere r={r}", "Sphere", [r], (-r, r)) for r in (0.5, 1.0, 1.5, 2.0)]
    out += [(f"torus R={R} r={r}", "Torus", [R, r], (-(R + r), R + r)) for R, r in ((1.0, 0.3), (1.5, 0.5))]
    for k in range(2):
        ax = [0.5 + g.uniform() for _ in range(3)]
        out.append(("ellipsoid " + "x".join(f"{a:.3f}" for a in ax), "Ellipsoid", ax, (-max(ax), max(ax))))
    return out


def sh_coefficients(instance, lmax=8):
    """a_lm ~ N(0, 1), l <= lmax, |m| <= l, from MT19937_64(1000 + instance)."""
    return MT19937_64(1000 + instance).normals((lmax + 1) ** 2)


def sh_rhs(cp, coeffs, lmax=8, chunk=1 << 17):
    """f(cp) = sum_lm a_lm Y_lm(cp / |cp|) for the columns of coeffs ((lmax + 1)^2 x k): N x k
    (chunks on host threads; numpy releases the interpreter lock inside its array operations)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    coeffs = np.asarray(coeffs, float).reshape((lmax + 1) ** 2, -1)
    out = np.empty((len(cp), coeffs.shape[1]))

    def one(s):
        out[s:s + chunk] = real_sph_harm(lmax, cp[s:s + chunk]) @ coeffs

    starts = range(0, len(cp), chunk)
    if len(starts) <= 1:
        for s in starts:
            one(s)
    else:
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1, len(starts))) as ex:
            list(ex.map(one, starts))
    return out


def mesh_edges(F):
    """Distinct edges of the triangles F and, per face, the indices of its three edges."""
    E = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    uniq, inv = np.unique(np.sort(E, axis=1), axis=0, return_inverse=True)
    return uniq, inv.reshape(-1)


def plane_section_loops(V, F, normal, offset, edges=None):
    """Closed polylines where the plane normal . x = offset cuts the closed triangle mesh (V, F): the
    crossing point of every edge whose ends lie on opposite sides (a vertex on the plane counts as
    above it), joined through the triangles they cut.  Returns a list of (n + 1) x 3 arrays whose last
    vertex repeats the first; the points lie on the mesh's edges.  edges: mesh_edges(F), if known."""
    s = V @ np.asarray(normal, float) - offset
    up = s >= 0
    uniq, inv = edges if edges is not None else mesh_edges(F)
    a, b = uniq[:, 0], uniq[:, 1]
    cross = up[a] != up[b]
    t = np.where(cross, s[a] / np.where(cross, s[a] - s[b], 1.0), 0.0)
    P = V[a] + t[:, None] * (V[b] - V[a])
    fe = inv.reshape(3, len(F)).T
    fc = cross[fe]
    cut = fc.sum(1) == 2
    pairs = fe[cut][fc[cut]].reshape(-1, 2)
    nbr = {}
    for e1, e2 in pairs:
        nbr.setdefault(int(e1), []).append(int(e2))
        nbr.setdefault(int(e2), []).append(int(e1))
    seen, loops = set(), []
    for start in nbr:
        if start in seen:
            continue
        seq, prev, cur = [start], -1, start
        seen.add(start)
        while True:
            nx = [e for e in nbr[cur] if e != prev]
            nxt = nx[0] if nx else nbr[cur][0]
            if nxt == start or nxt in seen:
                break
            seq.append(nxt)
            seen.add(nxt)
            prev, cur = cur, nxt
        if len(seq) >= 3:
            L = P[seq]
            loops.append(np.vstack([L, L[:1]]))
    return loops


def resample_closed(L, n=256):
    """n points spaced evenly by arc length along the closed polyline L (last vertex = first), returned
    as an (n + 1) x 3 closed polyline; points stay on L's segments."""
    seg = np.linalg.norm(np.diff(L, axis=0), axis=1)
    c = np.concatenate([[0.0], np.cumsum(seg)])
    u = c[-1] * np.arange(n) / n
    i = np.clip(np.searchsorted(c, u, side="right") - 1, 0, len(seg) - 1)
    w = np.where(seg[i] > 0, (u - c[i]) / np.where(seg[i] > 0, seg[i], 1.0), 0.0)
    Q = L[i] + w[:, None] * (L[i + 1] - L[i])
    return np.vstack([Q, Q[:1]])


def serving_request(V, F, seed, max_curves=8, samples=256, spacing=0.25):
    """One request of the serving workload: 1..max_curves closed curves where parallel planes (a random
    normal, offsets at least `spacing` apart) cut the mesh, each the longest loop of its plane resampled
    to `samples` points, and a random RGB colour per curve (numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, max_curves + 1))
    nrm = rng.standard_normal(3)
    nrm /= np.linalg.norm(nrm)
    h = V @ nrm
    lo, hi = h.min() + 0.15, h.max() - 0.15
    offs = []
    tries = 0
    while len(offs) < k and tries < 1000:
        o = float(rng.uniform(lo, hi))
        if all(abs(o - p) >= spacing for p in offs):
            offs.append(o)
        tries += 1
    curves = []
    edges = mesh_edges(F)
    for o in sorted(offs):
        loops = plane_section_loops(V, F, nrm, o, edges)
        if loops:
            L = max(loops, key=lambda q: np.linalg.norm(np.diff(q, axis=0), axis=1).sum())
            curves.append(resample_closed(L, samples))
    colours = rng.uniform(0.0, 1.0, (len(curves), 3))
    return curves, colours


def curve_of_points(curves, pts, chunk=4096):
    """Index of the nearest curve (closed polylines) for each point."""
    best = np.full(len(pts), np.inf)
    idx = np.zeros(len(pts), int)
    for c, C in enumerate(curves):
        A, B = C[:-1], C[1:]
        AB = B - A
        L2 = np.maximum((AB * AB).sum(1), 1e-300)
        for s in range(0, len(pts), chunk):
            Pp = pts[s:s + chunk]
            t = np.clip(((Pp[:, None, :] - A[None]) * AB[None]).sum(2) / L2[None], 0.0, 1.0)
            d = np.linalg.norm(A[None] + t[:, :, None] * AB[None] - Pp[:, None, :], axis=2).min(1)
            better = d < best[s:s + chunk]
            best[s:s + chunk] = np.where(better, d, best[s:s + chunk])
            idx[s:s + chunk] = np.where(better, c, idx[s:s + chunk])
    return idx


def plane_section_curves(V, F, k, seed, samples=256, spacing=0.25):
    """k closed curves where parallel planes (a random normal, offsets at least `spacing` apart, numpy
    default_rng(seed)) cut the mesh: the longest loop of each plane resampled to `samples` points."""
    rng = np.random.default_rng(seed)
    nrm = rng.standard_normal(3)
    nrm /= np.linalg.norm(nrm)
    h = V @ nrm
    lo, hi = h.min() + 0.15, h.max() - 0.15
    offs = []
    while len(offs) < k:
        o = float(rng.uniform(lo, hi))
        if all(abs(o - q) >= spacing for q in offs):
            offs.append(o)
    out = []
    edges = mesh_edges(F)
    for o in sorted(offs):
        loops = plane_section_loops(V, F, nrm, o, edges)
        L = max(loops, key=lambda q: np.linalg.norm(np.diff(q, axis=0), axis=1).sum())
        out.append(resample_closed(L, samples))
    return out
