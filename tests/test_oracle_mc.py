"""Pins the CPU oracle of the reference's marching cubes
(marching_cubes.cu:23-236, layer stride gx*gy) against closed forms and
topology: single-cube cases, a plane's triangle count, and a sphere's mesh,
which must be a closed, consistently oriented genus-0 surface of the right
volume -- a wrong corner offset, edge pair or swap rule breaks one of them."""
import numpy as np
import pytest

import oracle_ffi as O

from test_mc_tables import corner_offsets, edge_corners, tri_table


def _centres(n, vs, off):
    return (np.arange(n) + 0.5) * vs + off


def test_single_cube_cases():
    """All 256 classes on a 2x2x2 grid (where the reference's layer stride
    gx+gy equals gx*gy): the emitted edges are the table row, and each
    vertex is where the linear interpolation of its edge's corner values
    vanishes."""
    t, ec, co = tri_table(), edge_corners(), corner_offsets()
    vs, off = np.array([0.1, 0.2, 0.3]), np.array([-1.0, 0.5, 2.0])
    rng = np.random.default_rng(0)
    for c in range(256):
        vals = np.where([(c >> i) & 1 for i in range(8)], -rng.uniform(0.1, 1, 8),
                        rng.uniform(0.0, 1, 8)).astype(np.float32)
        f = np.empty((2, 2, 2), np.float32)
        for i, (dx, dy, dz) in enumerate(co):
            f[dz, dy, dx] = vals[i]
        n, hist = O.mc_count(f, (2, 2, 2), hist=True)
        assert hist[c] == 1 and hist.sum() == 1
        expect = (t[c] >= 0).sum() // 3 if c not in (0, 255) else 0
        assert n == expect
        if n == 0:
            continue
        verts, tris, edges, cubes = O.mc_extract(f, (2, 2, 2), vs, off)
        assert edges.ravel().tolist() == t[c][: 3 * n].tolist()
        assert tris.ravel().tolist() == list(range(3 * n))
        for k, e in enumerate(edges.ravel()):
            a, b = ec[e]
            pa = (co[a] + 0.5) * vs + off
            pb = (co[b] + 0.5) * vs + off
            p = verts.reshape(-1, 3)[k].astype(np.float64)
            s = np.linalg.norm(p - pa) / np.linalg.norm(pb - pa)
            interp = vals[a] + s * (vals[b] - vals[a])
            assert abs(interp) < 1e-5 * (abs(vals[a]) + abs(vals[b]))
            assert np.linalg.norm(np.cross(p - pa, pb - pa)) < 1e-5  # on the edge


def test_plane_triangle_count():
    """A horizontal plane strictly between two voxel layers cuts every cube of
    that layer with class 0x0F or 0xF0: two triangles per cube."""
    gx, gy, gz = 13, 9, 6
    z = _centres(gz, 0.1, 0.0)
    f = np.broadcast_to((z - 0.27)[:, None, None], (gz, gy, gx)).astype(np.float32)
    n, hist = O.mc_count(f, (gx, gy, gz), hist=True)
    assert n == 2 * (gx - 1) * (gy - 1)
    assert hist.sum() == (gx - 1) * (gy - 1) * (gz - 1)
    verts, *_ = O.mc_extract(f, (gx, gy, gz), (0.1, 0.1, 0.1), (0, 0, 0))
    np.testing.assert_allclose(verts[..., 2], 0.27, atol=1e-6)


def _weld(verts):
    pts = verts.reshape(-1, 3)
    uniq, inv = np.unique(pts.view(np.dtype((np.void, 12))), return_inverse=True)
    return len(uniq), inv.reshape(-1, 3)


@pytest.mark.parametrize("n,r", [(24, 0.31), (41, 0.37)])
def test_sphere_mesh_is_closed_and_has_the_volume(n, r):
    vs = 1.0 / n
    c = _centres(n, vs, -0.5)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    f = (np.sqrt(x * x + y * y + z * z) - r).astype(np.float32)
    verts, tris, edges, cubes = O.mc_extract(f, (n, n, n), [vs] * 3, [-0.5] * 3)
    nv, faces = _weld(verts)
    # every undirected edge is shared by exactly two triangles, with opposite
    # directions (a closed, consistently oriented 2-manifold)
    directed = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    assert len({tuple(e) for e in directed.tolist()}) == len(directed)
    und = np.sort(directed, axis=1)
    _, counts = np.unique(und, axis=0, return_counts=True)
    assert (counts == 2).all()
    nedges = len(counts)
    assert nv - nedges + len(faces) == 2  # Euler characteristic of a sphere
    rad = np.linalg.norm(verts.reshape(-1, 3), axis=1)
    assert np.abs(rad - r).max() < 0.25 * vs
    v = verts.astype(np.float64)
    signed = np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2])).sum() / 6.0
    # with the reference's corner numbering and a negative inside, triangles
    # wind counter-clockwise seen from outside: the signed volume is positive
    assert signed > 0
    assert abs(signed - 4.0 / 3.0 * np.pi * r ** 3) < 0.02 * (4.0 / 3.0 * np.pi * r ** 3)
    # cube order: z, y, x non-decreasing (marching_cubes.cu:163-188, :213-225)
    assert (np.diff(cubes) >= 0).all()


def test_nan_corner_counts_as_outside():
    f = np.full((2, 2, 2), -1.0, np.float32)
    f[0, 0, 0] = np.nan
    n, hist = O.mc_count(f, (2, 2, 2), hist=True)
    assert hist[0xFE] == 1 and n == 1


def test_degenerate_grids_have_no_cubes():
    for g in [(1, 4, 4), (4, 1, 4), (4, 4, 1)]:
        assert O.mc_count(-np.ones(g[::-1], np.float32), g) == 0
