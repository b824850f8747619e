"""Pins the oracle's welded marching cubes (oracle_mc_weld) against the
pinned soup oracle and against what the geometry fixes: every welded
triangle corner is the soup's vertex (bit for bit when no edge end is
exactly zero), the vertices are exactly the sign-changing grid edges in
(voxel, axis) order, a closed surface welds into a closed 2-manifold (Euler
characteristic 2), and normals are the field's gradient direction (exact for
linear fields, radial for a sphere)."""
import numpy as np
import pytest

import oracle_ffi as O


def _centres(grid, vs, off):
    return [((np.arange(g, dtype=np.float32) + np.float32(0.5)) * np.float32(v)) + np.float32(o)
            for g, v, o in zip(grid, vs, off)]


def _sign_edges(f):
    """Sign-changing grid edges as sorted keys 3*voxel + axis (f is [z][y][x])."""
    neg = f < 0
    gz, gy, gx = f.shape
    keys = []
    for a, (dz, dy, dx) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
        lo = neg[:gz - dz, :gy - dy, :gx - dx]
        hi = neg[dz:, dy:, dx:]
        zz, yy, xx = np.nonzero(lo != hi)
        keys.append(3 * (xx + gx * (yy + gy * zz)).astype(np.int64) + a)
    return np.sort(np.concatenate(keys))


@pytest.mark.parametrize("grid", [(7, 6, 5), (33, 17, 9), (2, 2, 2), (64, 3, 4)])
def test_welded_equals_soup_and_enumerates_sign_edges(grid):
    rng = np.random.default_rng(sum(grid))
    gx, gy, gz = grid
    f = rng.standard_normal((gz, gy, gx)).astype(np.float32)
    vs, off = (0.1, 0.2, 0.3), (-1.0, 0.5, 2.0)
    v, n, t = O.mc_weld(f, grid, vs, off)
    sv, st, _, _ = O.mc_extract(f, grid, vs, off)
    assert t.shape[0] == sv.shape[0]
    assert t.min(initial=0) >= 0 and t.max(initial=-1) < v.shape[0]
    # every corner is the soup's vertex, bit for bit (no exact zeros here)
    assert np.array_equal(v[t].view(np.uint32), sv.view(np.uint32))
    # vertices = sign-changing edges, each used by a triangle
    keys = _sign_edges(f)
    assert v.shape[0] == keys.size
    assert np.unique(t).size == v.shape[0]
    # vertex k lies on edge keys[k]: the two fixed coordinates are centres,
    # the moving one between the end centres
    cx, cy, cz = _centres(grid, vs, off)
    vox, ax = keys // 3, keys % 3
    x, y, z = vox % gx, (vox // gx) % gy, vox // (gx * gy)
    base = np.stack([cx[x], cy[y], cz[z]], 1)
    for a in range(3):
        fixed = ax != a
        assert np.array_equal(v[fixed, a], base[fixed, a])
        lo = base[~fixed, a]
        assert np.all(v[~fixed, a] >= lo) and np.all(v[~fixed, a] <= lo + np.float32(vs[a]) * 1.000001)


def test_exact_zeros_close_to_soup():
    rng = np.random.default_rng(9)
    f = rng.standard_normal((12, 11, 10)).astype(np.float32)
    f[rng.random(f.shape) < 0.15] = 0.0
    f[rng.random(f.shape) < 0.05] = -0.0
    v, _, t = O.mc_weld(f, (10, 11, 12), (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    sv, _, _, _ = O.mc_extract(f, (10, 11, 12), (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    d = np.abs(v[t] - sv)
    assert d.max() <= 2e-7                       # rounding of s + (e - s) only
    assert np.mean(d == 0) > 0.9


def _sphere(grid, c, r, vs):
    gx, gy, gz = grid
    zz, yy, xx = np.meshgrid(*[(np.arange(g) + 0.5) * vs for g in (gz, gy, gx)], indexing="ij")
    return (np.sqrt((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) - r).astype(np.float32)


def test_closed_sphere_is_a_closed_manifold_with_radial_normals():
    grid, vs = (40, 36, 44), 0.05
    c, r = (1.01, 0.93, 1.12), 0.61
    f = _sphere(grid, c, r, vs)
    v, n, t = O.mc_weld(f, grid, (vs, vs, vs), (0.0, 0.0, 0.0))
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()                      # watertight, no boundary
    assert v.shape[0] - cnt.size + t.shape[0] == 2
    radial = (v - np.array(c, np.float32)) / np.linalg.norm(v - np.array(c, np.float32), axis=1)[:, None]
    assert np.abs(np.linalg.norm(n, axis=1) - 1).max() < 1e-6
    assert (np.sum(n * radial, axis=1) > 0.999).all()
    # consistent winding: all triangle normals point the same way relative to n
    fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    s = np.sign(np.sum(fn * n[t[:, 0]], axis=1))
    assert abs(s.mean()) > 0.99


def test_linear_field_normals_exact():
    grid, vs = (9, 8, 7), (0.1, 0.2, 0.15)
    a = np.array([0.3, -0.8, 0.5])
    cx, cy, cz = [c.astype(np.float64) for c in _centres(grid, vs, (0, 0, 0))]
    zz, yy, xx = np.meshgrid(cz, cy, cx, indexing="ij")
    f = (a[0] * xx + a[1] * yy + a[2] * zz - 0.2).astype(np.float32)
    v, n, t = O.mc_weld(f, grid, vs, (0.0, 0.0, 0.0))
    assert v.shape[0] > 20
    np.testing.assert_allclose(n, np.broadcast_to(a / np.linalg.norm(a), n.shape), atol=2e-5)


def test_degenerate_grids_and_constant_fields():
    for grid in [(1, 5, 5), (5, 1, 5), (5, 5, 1)]:
        f = np.full(grid[::-1], -1.0, np.float32)
        f.flat[::2] = 1.0
        v, n, t = O.mc_weld(f, grid, (1, 1, 1), (0, 0, 0))
        assert v.shape[0] == 0 and t.shape[0] == 0
    v, n, t = O.mc_weld(np.ones((4, 4, 4), np.float32), (4, 4, 4), (1, 1, 1), (0, 0, 0))
    assert v.shape[0] == 0 and t.shape[0] == 0
