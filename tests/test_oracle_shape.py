"""Oracle of halo_shape and sigma_projections (particle_subroutines.f90:160-461), pinned
by point sets whose moments and cell statistics follow by hand."""
import numpy as np

from oracle import shape


def test_jacobi_matches_a_library_eigensolver():
    rng = np.random.default_rng(1)
    for _ in range(50):
        a = rng.normal(size=(3, 3)).astype(np.float32)
        a = (a + a.T).astype(np.float32)
        got = np.sort(shape.jacobi_eigenvalues(a).astype(np.float64))
        want = np.linalg.eigvalsh(a.astype(np.float64))
        # the early stop (off-diagonal sum below 1e-4 of the element sum) leaves errors of
        # order (1e-4 scale)^2 / gap at most
        assert np.allclose(got, want, rtol=0, atol=1e-6 * np.abs(a).sum())


def test_shape_of_axis_points():
    # six unit masses at (+-a, 0, 0), (0, +-b, 0), (0, 0, +-c): the tensor is
    # diag(a^2, b^2, c^2) / 3, the routine returns a, b, c / sqrt(3), largest first
    a, b, c = 3.0, 2.0, 0.5
    x = np.array([a, -a, 0, 0, 0, 0], np.float32)
    y = np.array([0, 0, b, -b, 0, 0], np.float32)
    z = np.array([0, 0, 0, 0, c, -c], np.float32)
    got = shape.halo_shape(x, y, z, np.ones(6, np.float32))
    want = np.sqrt(np.array([a * a, b * b, c * c], np.float32) / np.float32(3))
    assert np.allclose(got, want, rtol=2e-7)
    # rotating the set leaves the axes
    t = 0.3
    rot = np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    p = rot @ np.stack([x, y, z]).astype(np.float64)
    got = shape.halo_shape(*(p.astype(np.float32)), np.ones(6, np.float32))
    assert np.allclose(got, want, rtol=1e-5)


def test_sigma_projections_by_hand():
    # grid nodes -1, 0, 1; four unit-mass particles at the origin with x velocities
    # 1, 3, 1, 3 and no y, z velocity: every projection puts them in its centre cell
    grid = np.array([-1.0, 0.0, 1.0], np.float32)
    n = 4
    x = np.zeros(n, np.float32)
    vx = np.array([1.0, 3.0, 1.0, 3.0], np.float32)
    zero = np.zeros(n, np.float32)
    m = np.ones(n, np.float32)
    out = shape.sigma_projections(grid, np.arange(n), x, x, x, vx, zero, zero, m, 0, 0, 0,
                                  0.5, 0.5, 0.5, 0.1)
    # x projection: all in cell (1, 1), VCM 2, sigma sqrt(4 / 4) = 1 -> SIG_1D_x_05 = 1
    # y and z projections: velocities 0 -> sigma 0
    assert out[0] == np.float32(1.0) and out[1] == 0.0 and out[2] == 0.0
    # planes: only the centre cell has Rbin = 0 < R05 + 2 ll; XY plane (VCM_z = 0, sigma
    # 0) and XZ give 0 / 0 -> 0; YZ: sumV = 4 * 4, sumSigma = 1 * 4 -> V/sigma_x = 2,
    # lambda_x = 0 / 0 -> 0 (Rbin = 0)
    assert out[3] == np.float32(np.float32(2.0) / np.float32(3.0))
    assert out[4] == 0.0
