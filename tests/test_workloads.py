"""Synthetic-workload generators: the seeded 64-bit Mersenne Twister, its
normal deviates and the real spherical harmonics that build right-hand sides
with known solutions on spheres."""
import numpy as np

from cpmgpu import workloads


def test_mt19937_64_known_answer():
    # the C++ standard fixes the 10000th output of a default-constructed
    # std::mt19937_64 (seed 5489) ([rand.predef])
    g = workloads.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_normals_moments_and_determinism():
    z = workloads.MT19937_64(7).normals(40000)
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1) < 0.03
    assert abs((z ** 4).mean() - 3) < 0.15                       # Gaussian kurtosis
    assert np.array_equal(z[:9], workloads.MT19937_64(7).normals(9))
    assert not np.array_equal(z[:9], workloads.MT19937_64(8).normals(9))


def test_real_spherical_harmonics_orthonormal_and_harmonic():
    lmax = 8
    # Gauss-Legendre in cos(theta) x uniform phi integrates degree <= 2 lmax exactly
    ct, w = np.polynomial.legendre.leggauss(lmax + 2)
    ph = 2 * np.pi * np.arange(2 * lmax + 4) / (2 * lmax + 4)
    C, P = np.meshgrid(ct, ph, indexing="ij")
    S = np.sqrt(1 - C * C)
    pts = np.stack([S * np.cos(P), S * np.sin(P), C], -1).reshape(-1, 3)
    W = (w[:, None] * np.full(len(ph), 2 * np.pi / len(ph))[None, :]).reshape(-1)
    Y = workloads.real_sph_harm(lmax, pts)
    G = Y.T @ (W[:, None] * Y)
    assert np.abs(G - np.eye(Y.shape[1])).max() < 1e-12
    # r^l Y_lm(x / r) is a harmonic polynomial: its 3-D Laplacian vanishes, so
    # Y_lm is an eigenfunction of the Laplace-Beltrami operator with -l(l+1)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (20, 3))
    h = 1e-3
    ls = np.concatenate([[l] * (2 * l + 1) for l in range(lmax + 1)])

    def ext(q):
        r = np.linalg.norm(q, axis=1)
        return workloads.real_sph_harm(lmax, q) * r[:, None] ** ls[None, :]

    lap = -6 * ext(x)
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        lap += ext(x + e) + ext(x - e)
    lap /= h * h
    # the stencil's truncation error (h^2 times fourth derivatives of degree-8
    # polynomials) is ~1e-4 here; a non-harmonic extension would leave terms
    # of order l (l + 1) |Y|
    assert np.abs(lap).max() < 2e-3 * max(1.0, np.abs(ext(x)).max())


def test_multi_surface_set_is_fixed():
    s = workloads.multi_surface_set()
    assert [e[1] for e in s] == ["Sphere"] * 4 + ["Torus"] * 2 + ["Ellipsoid"] * 2
    for _, typ, par, bb in s:
        assert bb[0] == -bb[1] and bb[1] >= max(par[:3]) - 1e-15
        if typ == "Ellipsoid":
            assert all(0.5 <= a < 1.5 for a in par)
    assert s == workloads.multi_surface_set()
