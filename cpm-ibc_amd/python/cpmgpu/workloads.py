"""Synthetic workloads of the benchmark configurations (geometry and data only)."""
import numpy as np


def torus_knot_polyline(R=1.0, r=0.4, segments=2000):
    """Closed polyline sampling the (2,3) torus knot ((R + r cos 3t) cos 2t, (R + r cos 3t) sin 2t, r sin 3t)
    at t_i = 2 pi i / segments; the last vertex repeats the first."""
    t = 2 * np.pi * np.arange(segments + 1) / segments
    V = np.stack([(R + r * np.cos(3 * t)) * np.cos(2 * t), (R + r * np.cos(3 * t)) * np.sin(2 * t), r * np.sin(3 * t)], 1)
    V[-1] = V[0]
    return V


def polyline_parameter(V, pts, t0=0.0, t1=2 * np.pi, chunk=512):
    """Curve parameter of the closest polyline point: vertex i carries t0 + (t1 - t0) i / (len(V) - 1),
    linear along each segment (first minimum wins, exact FP64 segment projection)."""
    A, B = V[:-1], V[1:]
    AB = B - A
    L2 = (AB * AB).sum(1)
    nseg = len(A)
    out = np.empty(len(pts))
    for s in range(0, len(pts), chunk):
        P = pts[s:s + chunk]
        tt = np.clip(np.einsum("pk,sk->ps", P, AB) - (A * AB).sum(1)[None, :], 0, None) / L2[None, :]
        tt = np.minimum(tt, 1.0)
        Q = A[None, :, :] + tt[:, :, None] * AB[None, :, :]
        d = np.linalg.norm(Q - P[:, None, :], axis=2)
        i = np.argmin(d, axis=1)
        out[s:s + chunk] = t0 + (t1 - t0) * (i + tt[np.arange(len(P)), i]) / nseg
    return out
