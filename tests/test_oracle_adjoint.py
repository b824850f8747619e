"""Pins the float64 oracle of the differentiable integrate (bilinear depth
sampling, an extension of cu_tsdf_integrate, tsdf_volume.cu:144-190), of the
surface loss on the reference's trilinear sampler (interpolate_trilinearly,
tsdf_volume.cu:46-104) and of the adjoint of both:

  - the trilinear support against affine fields (which trilinear weights
    reproduce exactly) and against scipy.ndimage.map_coordinates;
  - dL/dD against central differences of the loss (quadratic in D, so the
    difference is exact up to rounding);
  - the bilinear forward against a closed form: a depth map linear in pixel
    coordinates is sampled exactly by bilinear interpolation;
  - the weight cap, the +trunc clamp and the -trunc skip of the adjoint
    against a hand-derived single-voxel case;
  - pose and depth gradients against central finite differences of the loss
    (step 1e-4 on smooth depth; 1e-6 on the noisy depth of the small plain
    workload), on the points whose decisions the step cannot change.

Relative error of a gradient is max|fd - g| / max|g| over all its entries."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates
from scipy.spatial.transform import Rotation

import oracle_ffi as O
from recon import scenes

INTR = (525.0, 525.0, 319.5, 239.5)


def _centres(vol):
    g = list(vol.grid)
    vs = [np.float64(np.float32(vol.voxel_size[i])) for i in range(3)]
    off = [np.float64(np.float32(o)) for o in vol.offset]
    z, y, x = np.meshgrid(*[(np.arange(g[i]) + 0.5) * vs[i] + off[i] for i in (2, 1, 0)],
                          indexing="ij")
    return np.stack([x, y, z], -1)


def _unravel(vol, ids):
    gx, gy = vol.grid[0], vol.grid[1]
    return np.stack([ids % gx, (ids // gx) % gy, ids // (gx * gy)], -1)


# neighbour order of interpolate_trilinearly (tsdf_volume.cu:85-92)
REF_NEIGHBOURS = np.array([(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1),
                           (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)])


def test_trilinear_support_reproduces_affine_fields_and_order():
    vol = O.volume((12, 10, 9), [0.05, 0.04, 0.06], (-0.3, -0.2, -0.27), 0.1, 128.0, INTR)
    rng = np.random.default_rng(11)
    c = _centres(vol)
    lo_c, hi_c = c[0, 0, 0], c[-1, -1, -1]
    for _ in range(200):
        a, b = rng.normal(size=3), rng.normal()
        field = (c @ a + b).ravel()
        # interior points, and points in the half voxel below the first
        # centre, where the lower corner clamps to 0 and the weights
        # extrapolate (as the reference's cubic_coord does)
        p = rng.uniform(lo_c - 0.45 * np.array([0.05, 0.04, 0.06]), hi_c - 1e-9)
        ids, coef = O.trilinear_support(vol, p)
        assert abs(coef.sum() - 1.0) < 1e-12
        assert abs(coef @ field[ids] - (p @ a + b)) < 1e-11
        assert (_unravel(vol, ids) - _unravel(vol, ids[:1]) == REF_NEIGHBOURS).all()
        assert (_unravel(vol, ids[:1]) >= 0).all()
    # beyond the last centre the 2x2x2 support would leave the grid
    assert O.trilinear_support(vol, hi_c + 0.01) is None
    # more than 1e-4 below the grid a coordinate moves onto the grid's face
    # (tsdf_volume.cu:53-58); within 1e-4 it extrapolates
    off = np.array([np.float64(np.float32(o)) for o in vol.offset])
    p = np.array([off[0] - 0.03, 0.01, off[2] - 0.00005])
    ids, coef = O.trilinear_support(vol, p)
    ids0, coef0 = O.trilinear_support(vol, np.array([off[0], 0.01, off[2] - 0.00005]))
    assert (ids == ids0).all() and np.allclose(coef, coef0, rtol=0, atol=1e-15)
    assert coef[[0, 1, 4, 5]].sum() > 1.0  # z below the first centre: extrapolated


def test_trilinear_matches_scipy_linear_interpolation():
    vol = O.volume((16, 14, 11), [0.02] * 3, (-0.16, -0.14, -0.11), 0.06, 128.0, INTR)
    rng = np.random.default_rng(12)
    field = rng.normal(size=(11, 14, 16))
    c = _centres(vol)
    pts = rng.uniform(c[0, 0, 0], c[-1, -1, -1] - 1e-9, size=(500, 3))
    vs = np.float64(np.float32(0.02))
    off = np.array([np.float64(np.float32(o)) for o in vol.offset])
    idx = ((pts - off) / vs - 0.5)[:, ::-1].T  # (z, y, x) index coordinates
    want = map_coordinates(field, idx, order=1)
    got = np.empty(len(pts))
    for k, p in enumerate(pts):
        ids, coef = O.trilinear_support(vol, p)
        got[k] = coef @ field.ravel()[ids]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_lambda_is_the_gradient_of_the_loss():
    w = scenes.tiny(grid=(20, 18, 16), voxel_size=0.03)
    vol = O.volume_from_config(w.config())
    rng = np.random.default_rng(13)
    D = rng.normal(0, 0.05, size=O.shape(vol))
    pts = scenes.surface_points(w, 3000, seed=13)
    wts = (rng.uniform(size=len(pts)) < 0.7).astype(np.float64)
    L, lam = O.surface_loss(vol, D, pts, wts)
    # L itself from the supports
    r = np.array([coef @ D.ravel()[ids] for ids, coef in (O.trilinear_support(vol, p) for p in pts)])
    assert abs(L - np.mean(wts * r * r)) < 1e-15 + 1e-12 * L
    active = np.flatnonzero(lam)
    assert len(active) > 100
    for v in rng.choice(active, 40, replace=False):
        h = 1e-3
        Dp, Dm = D.copy(), D.copy()
        Dp.ravel()[v] += h
        Dm.ravel()[v] -= h
        fd = (O.surface_loss(vol, Dp, pts, wts, want_lambda=False) -
              O.surface_loss(vol, Dm, pts, wts, want_lambda=False)) / (2 * h)
        assert abs(fd - lam.ravel()[v]) <= 1e-9 * np.abs(lam).max()
    # voxels outside every weighted support get no gradient
    untouched = np.ones(D.size, bool)
    for p, wk in zip(pts, wts):
        if wk:
            untouched[O.trilinear_support(vol, p)[0]] = False
    assert (lam.ravel()[untouched] == 0).all()


def _pose_from_rotation(rot, t):
    q = rot.as_quat()  # x, y, z, w
    return np.array([q[3], q[0], q[1], q[2], *t])


def test_bilinear_forward_closed_form_on_a_linear_depth_map():
    """depth[j][i] = A + B i + C j: bilinear sampling at (u, v) is exactly
    A + B u + C v, so every updated voxel has D = min(A + B u + C v - z, trunc)
    and W = 1; u, v, z from scipy's rotation."""
    rows, cols = 120, 160
    intr = (150.0, 150.0, 79.5, 59.5)
    vol = O.volume((24, 22, 20), [0.02] * 3, (-0.24, -0.22, -0.2), 0.05, 128.0, intr)
    rng = np.random.default_rng(14)
    jj, ii = np.mgrid[0:rows, 0:cols]
    c = _centres(vol)
    for trial in range(4):
        A, B, Cc = 0.85 + 0.05 * trial, 1e-3 * rng.normal(), 1e-3 * rng.normal()
        depth = (A + B * ii + Cc * jj).astype(np.float32)
        rot = Rotation.from_rotvec(rng.normal(0, 0.45, 3))
        pose = _pose_from_rotation(rot, [*rng.normal(0, 0.03, 2), 0.85])
        D, W = O.diff_forward(vol, depth[None], O.rt_from_poses(pose[None]))
        p = c @ rot.as_matrix().T + pose[4:]
        z = p[..., 2]
        u = p[..., 0] * intr[0] / z + intr[2]
        v = p[..., 1] * intr[1] / z + intr[3]
        inside = (z > 0) & (np.floor(u) >= 0) & (np.floor(u) <= cols - 2) & \
                 (np.floor(v) >= 0) & (np.floor(v) <= rows - 2)
        # the rounded float32 map departs from the plane by <= 1 ulp
        sdf = (A + B * u + Cc * v) - z
        upd = inside & (sdf >= -vol.truncation)
        amb = (np.abs(sdf + vol.truncation) < 1e-6) | \
              (np.minimum(np.abs(u - np.round(u)), np.abs(v - np.round(v))) < 1e-9)
        assert upd.sum() > 1000 and (~upd & ~amb).sum() > 100
        assert ((W == 1) == upd)[~amb].all()
        want = np.minimum(sdf, vol.truncation)
        err = np.abs(D - want)[upd & ~amb]
        assert err.max() < 2e-7, err.max()
        assert (D[~upd & ~amb] == np.float32(vol.truncation)).all()


def test_adjoint_single_voxel_cap_clamp_and_skip():
    """One voxel, identity rotations, constant depth per frame, cap 2.
    Frames 0,1,2,4 update it (frame 2 clamped at +trunc, frame 3 skipped as
    sdf < -trunc), so D = (2/9) t0 + (2/9) t1 + (2/9) trunc + (1/3) t4 and
    dD/drho_z(f) = -k_f with k = (2/9, 2/9, 0, 0, 1/3); dD/domega = p x g."""
    trunc = 0.05
    vol = O.volume((1, 1, 1), [0.02] * 3, (-0.01, -0.01, -0.01), trunc, 2.0, INTR)
    F = 5
    ts = np.array([[0.01, -0.02, 1.0], [-0.03, 0.01, 1.1], [0.02, 0.02, 0.9],
                   [0.0, 0.01, 1.2], [0.015, -0.01, 1.05]])
    sdfs = np.array([0.013, -0.021, 0.2, -0.3, 0.034])
    c = 0.5 * np.float64(np.float32(0.02)) + np.float64(np.float32(-0.01))  # ~0, per axis
    pc = ts + c  # camera points of the voxel centre
    zc = pc[:, 2]
    depth = np.stack([np.full((480, 640), z + s, np.float32) for z, s in zip(zc, sdfs)])
    sdf32 = depth[:, 0, 0].astype(np.float64) - zc
    Rt = np.zeros((F, 12))
    Rt[:, [0, 4, 8]] = 1.0
    Rt[:, 9:] = ts
    D, W = O.diff_forward(vol, depth, Rt)
    k = np.array([2 / 9, 2 / 9, 2 / 9, 0.0, 1 / 3])
    tsdf = np.minimum(sdf32, np.float64(np.float32(trunc)))
    assert W.ravel()[0] == 2.0
    assert abs(D.ravel()[0] - k @ np.where(np.arange(F) == 3, 0.0, tsdf)) < 1e-15
    lam = np.full((1, 1, 1), 0.7)
    gp, gd = O.diff_adjoint(vol, depth, Rt, lam)
    kg = 0.7 * np.array([2 / 9, 2 / 9, 0.0, 0.0, 1 / 3])  # clamped frame: zero
    g = np.zeros((F, 3))
    g[:, 2] = -kg
    np.testing.assert_allclose(gp[:, :3], g, atol=1e-15)
    np.testing.assert_allclose(gp[:, 3:], np.cross(pc, g), atol=1e-15)
    # the depth taps carry the same weights, split bilinearly
    np.testing.assert_allclose(gd.reshape(F, -1).sum(1), kg, atol=1e-15)
    assert (np.count_nonzero(gd.reshape(F, -1), axis=1) <= 4).all()


def _fd_setup(noise, F=4, grid=64, vs=0.016, npts=8000, h=1e-4, kink=1e-4):
    w = scenes.plain_small(num_frames=F, grid=grid, voxel_size=vs)
    depth = w.render(noise=noise)
    vol = O.volume_from_config(w.config())
    Rt = O.rt_from_poses(w.poses)
    pts = scenes.surface_points(w, npts, seed=21)
    wts = O.stable_points(vol, depth, Rt, pts, h=h, kink=kink)
    D, _ = O.diff_forward(vol, depth, Rt)
    L, lam = O.surface_loss(vol, D, pts, wts)
    gp, gd = O.diff_adjoint(vol, depth, Rt, lam)
    return w, depth, vol, Rt, pts, wts, gp, gd


def _pose_fd(vol, depth, Rt, pts, wts, h):
    F = Rt.shape[0]
    fd = np.zeros((F, 6))
    for f in range(F):
        for i in range(6):
            e = np.zeros(6)
            e[i] = h
            Rp, Rm = Rt.copy(), Rt.copy()
            Rp[f] = O.se3_exp_left(Rt[f], e)
            Rm[f] = O.se3_exp_left(Rt[f], -e)
            Lp = O.surface_loss(vol, O.diff_forward(vol, depth, Rp)[0], pts, wts, want_lambda=False)
            Lm = O.surface_loss(vol, O.diff_forward(vol, depth, Rm)[0], pts, wts, want_lambda=False)
            fd[f, i] = (Lp - Lm) / (2 * h)
    return fd


@pytest.mark.parametrize("noise,h,kink,tol", [(False, 1e-4, 2e-5, 1e-4), (True, 1e-6, 1e-4, 1e-5)],
                         ids=["smooth-h1e-4", "noisy-h1e-6"])
def test_pose_gradient_matches_central_differences(noise, h, kink, tol):
    _, depth, vol, Rt, pts, wts, gp, _ = _fd_setup(noise, h=h, kink=kink)
    assert wts.mean() > 0.1
    assert (np.abs(gp).max(axis=1) > 0).all()
    fd = _pose_fd(vol, depth, Rt, pts, wts, h)
    err = np.abs(fd - gp).max() / np.abs(gp).max()
    assert err < tol, (err, gp, fd)


def test_depth_gradient_matches_central_differences():
    _, depth, vol, Rt, pts, wts, gp, gd = _fd_setup(True, h=1e-4)
    rng = np.random.default_rng(22)
    flat = np.abs(gd).ravel()
    top = np.argsort(-flat)[:12]
    valid = np.flatnonzero((depth.ravel() > 1e-3) & (flat > 0))
    picks = np.r_[top, rng.choice(valid, 12, replace=False)]
    for idx in picks:
        f, j, i = np.unravel_index(idx, depth.shape)
        d0 = depth[f, j, i]
        dp, dm = depth.copy(), depth.copy()
        dp[f, j, i] = np.float32(d0 + 1e-4)
        dm[f, j, i] = np.float32(d0 - 1e-4)
        step = np.float64(dp[f, j, i]) - np.float64(dm[f, j, i])
        Lp = O.surface_loss(vol, O.diff_forward(vol, dp, Rt)[0], pts, wts, want_lambda=False)
        Lm = O.surface_loss(vol, O.diff_forward(vol, dm, Rt)[0], pts, wts, want_lambda=False)
        fd = (Lp - Lm) / step
        assert abs(fd - gd[f, j, i]) <= 1e-6 * flat.max(), (f, j, i, fd, gd[f, j, i])


def test_adjoint_is_linear_in_lambda():
    """Gradients of a sum of lambdas equal the sum of the gradients."""
    w = scenes.tiny(num_frames=4)
    depth = w.render()
    vol = O.volume_from_config(w.config())
    Rt = O.rt_from_poses(w.poses)
    rng = np.random.default_rng(23)
    l1 = rng.normal(size=O.shape(vol)) * (rng.uniform(size=O.shape(vol)) < 0.2)
    l2 = rng.normal(size=O.shape(vol)) * (rng.uniform(size=O.shape(vol)) < 0.2)
    g1, d1 = O.diff_adjoint(vol, depth, Rt, l1)
    g2, d2 = O.diff_adjoint(vol, depth, Rt, l2)
    g3, d3 = O.diff_adjoint(vol, depth, Rt, 0.5 * l1 - 2.0 * l2)
    np.testing.assert_allclose(g3, 0.5 * g1 - 2.0 * g2, rtol=1e-10, atol=1e-12 * np.abs(g3).max())
    np.testing.assert_allclose(d3, 0.5 * d1 - 2.0 * d2, rtol=1e-10, atol=1e-12 * np.abs(d3).max())


def test_float32_forward_restatement_agrees_with_float64():
    """The float32 bilinear forward (the GPU kernel's operation order, which
    the GPU tests match bit for bit) against the float64 forward: decisions
    agree on all but a few voxels, D within 1e-6 on almost all updated
    voxels; the rest straddle depth edges, where the bilinear slope turns
    float32 pixel rounding into larger differences (bounded here)."""
    import recon
    w = scenes.plain_small(num_frames=4, grid=64, voxel_size=0.016)
    depth = w.render()
    rt32 = recon.pose_to_matrix(w.poses)
    vol = O.volume_from_config(w.config())
    D32, W32 = O.diff_forward_f32(vol, depth, rt32)
    D64, W64 = O.diff_forward(vol, depth, rt32.astype(np.float64))
    same = W32 == W64
    assert same.mean() > 0.999
    upd = same & (W32 > 0)
    err = np.abs(D32 - D64)[upd]
    assert upd.sum() > 20000
    assert (err < 1e-6).mean() > 0.995 and err.max() < 3e-4
    # untouched voxels keep the initial value
    assert (D32[W32 == 0] == np.float32(vol.truncation)).all()


# ---- model-to-frame loss of pose refinement --------------------------------------

def test_frame_loss_closed_form_on_a_linear_depth_map():
    """A depth map affine in pixel coordinates, d(i, j) = al i + be j + ga, is
    sampled exactly by bilinear interpolation (depth[j][i] at pixel (i, j)),
    so r = d(u, v) - z for the continuous projection (u, v) of each point;
    the gradient is the analytic one of that residual; invalid samples give
    r = 0 and no gradient."""
    rows, cols = 48, 64
    al, be, ga = 0.004, -0.003, 1.2
    jj, ii = np.mgrid[0:rows, 0:cols]
    depth = (al * ii + be * jj + ga).astype(np.float32)
    vol = O.volume((8, 8, 8), [0.1] * 3, (0.0, 0.0, 0.0), 0.3, 128.0, (50.0, 55.0, 31.5, 23.5))
    Rt = O.rt_from_poses(np.array([[1.0, 0, 0, 0, 0, 0, 0]]))
    rng = np.random.default_rng(31)
    pts = np.c_[rng.uniform(-0.4, 0.4, 50), rng.uniform(-0.3, 0.3, 50), rng.uniform(1.0, 1.5, 50)]
    pts = np.r_[pts, [[0.0, 0.0, -1.0], [50.0, 0.0, 1.0]]]       # behind, off the image
    pairs = np.c_[np.zeros(len(pts)), np.arange(len(pts))].astype(np.int32)
    L, g, res = O.frame_loss(vol, depth[None], Rt, pts, pairs)
    u = pts[:50, 0] * 50.0 / pts[:50, 2] + 31.5
    v = pts[:50, 1] * 55.0 / pts[:50, 2] + 23.5
    dd = al * u + be * v + ga
    # the map holds float32 values of the affine function
    np.testing.assert_allclose(res[:50], dd - pts[:50, 2], atol=2e-7)
    assert res[50] == 0.0 and res[51] == 0.0
    assert abs(L - np.sum(res ** 2) / len(pts)) < 1e-15
    # analytic gradient of sum r^2 / K with r = al u + be v + ga - z
    K = len(pts)
    x, y, z = pts[:50].T
    gr = 2 * res[:50] / K
    drdp = np.c_[al * 50.0 / z, be * 55.0 / z, -(al * 50.0 * x + be * 55.0 * y) / z ** 2 - 1]
    want = np.r_[(gr[:, None] * drdp).sum(0), (gr[:, None] * np.cross(pts[:50], drdp)).sum(0)]
    # (slopes from float32 depths: relative error ~ ulp(d) / (0.004 px) = 3e-5)
    np.testing.assert_allclose(g[0], want, rtol=1e-4, atol=1e-12)
    # truncated at tau: residuals of |r| >= tau count tau^2 and drop out of
    # the gradient
    tau = float(np.median(np.abs(res[:50])))
    Lt, gt, _ = O.frame_loss(vol, depth[None], Rt, pts, pairs, tau=tau)
    keep = np.abs(res[:50]) < tau
    assert 10 < keep.sum() < 40
    assert abs(Lt - (np.sum(res[:50][keep] ** 2) + tau ** 2 * (~keep).sum()) / K) < 1e-15
    gk = gr * keep
    want_t = np.r_[(gk[:, None] * drdp).sum(0), (gk[:, None] * np.cross(pts[:50], drdp)).sum(0)]
    np.testing.assert_allclose(gt[0], want_t, rtol=1e-4, atol=1e-12)


def _frame_pair_setup(F=4, npts=3000):
    w = scenes.plain_small(num_frames=F, grid=64, voxel_size=0.016)
    depth = w.render(noise=False)
    vol = O.volume_from_config(w.config())
    Rt = O.rt_from_poses(w.poses)
    # poses a little off, so that residuals are not zero
    rng = np.random.default_rng(41)
    Rt = np.stack([O.se3_exp_left(Rt[f], rng.normal(0, [2e-3] * 3 + [3e-3] * 3)) for f in range(F)])
    pts, nrm = scenes.surface_points(w, npts, seed=42, with_normals=True)
    pairs = scenes.frame_pairs(w, depth, pts, nrm, poses=w.poses)
    return w, depth, vol, Rt, pts, pairs


def _stable_pairs(w, Rt, pts, pairs, margin=0.02):
    """1 for pairs whose projection stays inside one bilinear cell under the
    finite-difference steps (fractional parts of u and v away from 0 and 1)."""
    fx, fy, cx, cy = w.intrinsics
    f, q = pairs[:, 0], pairs[:, 1]
    R = Rt[f, :9].reshape(-1, 3, 3)
    pc = np.einsum("kij,kj->ki", R, pts[q]) + Rt[f, 9:]
    u = pc[:, 0] * fx / pc[:, 2] + cx
    v = pc[:, 1] * fy / pc[:, 2] + cy
    fu, fv = u - np.floor(u), v - np.floor(v)
    return ((fu > margin) & (fu < 1 - margin) & (fv > margin) & (fv < 1 - margin)).astype(np.float64)


def test_frame_loss_gradient_matches_central_differences():
    w, depth, vol, Rt, pts, pairs = _frame_pair_setup()
    assert len(np.unique(pairs[:, 0])) == Rt.shape[0] and len(pairs) > 1000
    wts = _stable_pairs(w, Rt, pts, pairs)
    assert wts.mean() > 0.8
    L, g, res = O.frame_loss(vol, depth, Rt, pts, pairs, wts)
    assert L > 0 and (np.abs(g).max(axis=1) > 0).all()
    h = 1e-6
    fd = np.zeros_like(g)
    for f in range(Rt.shape[0]):
        for i in range(6):
            e = np.zeros(6)
            e[i] = h
            Rp, Rm = Rt.copy(), Rt.copy()
            Rp[f] = O.se3_exp_left(Rt[f], e)
            Rm[f] = O.se3_exp_left(Rt[f], -e)
            fd[f, i] = (O.frame_loss(vol, depth, Rp, pts, pairs, wts, want_grad=False)[0] -
                        O.frame_loss(vol, depth, Rm, pts, pairs, wts, want_grad=False)[0]) / (2 * h)
    err = np.abs(fd - g).max() / np.abs(g).max()
    assert err < 1e-5, (err, g, fd)


def test_frame_pairs_are_grouped_by_frame_and_seen():
    w, depth, vol, Rt, pts, pairs = _frame_pair_setup()
    assert (np.diff(pairs[:, 0]) >= 0).all()
    # at the true poses the noise-free residuals of the pairs are small
    _, _, res = O.frame_loss(vol, depth, O.rt_from_poses(w.poses), pts, pairs, want_grad=False)
    assert np.percentile(np.abs(res), 95) < 2e-3


# ---- Gauss-Newton refinement: normal matrix and step -------------------------------

def test_frame_normal_matrix_slopes_on_a_linear_depth_map():
    """Without normals the normal matrix is sum 2 J J^T / K over the valid
    pairs with |r| < tau, J = (dr/dp, p x dr/dp) from the bilinear slopes; on
    a depth map affine in pixel coordinates dr/dp has the closed form of the
    frame-loss test above."""
    rows, cols = 48, 64
    al, be, ga = 0.004, -0.003, 1.2
    jj, ii = np.mgrid[0:rows, 0:cols]
    depth = (al * ii + be * jj + ga).astype(np.float32)
    vol = O.volume((8, 8, 8), [0.1] * 3, (0.0, 0.0, 0.0), 0.3, 128.0, (50.0, 55.0, 31.5, 23.5))
    Rt = O.rt_from_poses(np.array([[1.0, 0, 0, 0, 0, 0, 0]]))
    rng = np.random.default_rng(32)
    pts = np.c_[rng.uniform(-0.4, 0.4, 60), rng.uniform(-0.3, 0.3, 60), rng.uniform(1.0, 1.5, 60)]
    pts = np.r_[pts, [[0.0, 0.0, -1.0]]]                         # behind: adds nothing
    pairs = np.c_[np.zeros(len(pts)), np.arange(len(pts))].astype(np.int32)
    _, _, res = O.frame_loss(vol, depth[None], Rt, pts, pairs, want_grad=False)
    x, y, z = pts[:60].T
    drdp = np.c_[al * 50.0 / z, be * 55.0 / z, -(al * 50.0 * x + be * 55.0 * y) / z ** 2 - 1]
    J = np.c_[drdp, np.cross(pts[:60], drdp)]
    K = len(pts)
    for tau in (0.0, float(np.median(np.abs(res[:60])))):
        keep = np.ones(60, bool) if tau == 0.0 else np.abs(res[:60]) < tau
        assert keep.sum() >= 20
        want = 2.0 * (J[keep].T @ J[keep]) / K
        H = O.frame_normal_matrix(vol, depth[None], Rt, pts, pairs, tau=tau)
        assert H.shape == (1, 6, 6) and np.array_equal(H[0], H[0].T)
        np.testing.assert_allclose(H[0], want, rtol=1e-4, atol=1e-12 * np.abs(want).max())


def _plane_setup(seed=33, n=400):
    """A plane seen by a rotated, shifted camera: its depth map rendered
    exactly at the pixel centres, points on the plane (world), its normal."""
    rows, cols = 120, 160
    fx, fy, cx, cy = 150.0, 140.0, 79.5, 59.5
    nc = np.array([0.25, -0.15, -1.0])
    nc /= np.linalg.norm(nc)
    c = float(nc @ np.array([0.05, -0.02, 1.3]))          # plane nc . X = c (camera)
    jj, ii = np.mgrid[0:rows, 0:cols]
    ray = np.stack([(ii - cx) / fx, (jj - cy) / fy, np.ones_like(ii, float)], -1)
    depth = (c / (ray @ nc)).astype(np.float32)
    rng = np.random.default_rng(seed)
    u, v = rng.uniform(5, cols - 6, n), rng.uniform(5, rows - 6, n)
    rc = np.c_[(u - cx) / fx, (v - cy) / fy, np.ones(n)]
    Xc = rc * (c / (rc @ nc))[:, None]
    R = Rotation.from_rotvec([0.3, -0.2, 0.5]).as_matrix()
    t = np.array([0.1, -0.05, 0.2])
    Xw = (Xc - t) @ R                                       # camera -> world
    nw = R.T @ nc
    Rt = np.r_[R.ravel(), t][None]
    vol = O.volume((8, 8, 8), [0.1] * 3, (0.0, 0.0, 0.0), 0.3, 128.0, (fx, fy, cx, cy))
    pairs = np.c_[np.zeros(n), np.arange(n)].astype(np.int32)

    def residual(T, i):
        """r of point i under pose T against the exact plane (no sampling)."""
        p = T[:9].reshape(3, 3) @ Xw[i] + T[9:]
        uu, vv = p[0] * fx / p[2] + cx, p[1] * fy / p[2] + cy
        r_ = np.array([(uu - cx) / fx, (vv - cy) / fy, 1.0])
        return c / (r_ @ nc) - p[2]

    return vol, depth, Rt, Xw, np.tile(nw, (n, 1)), pairs, residual


def test_frame_normal_matrix_point_to_plane_on_a_plane():
    """With normals, J = (j, p x j), j = -(z / (n . p)) n: on a plane this is
    the exact derivative of r = d(s) - z for points on it, so the normal
    matrix equals sum 2 J J^T / K with J from central differences of the
    exact plane residual (world normals rotated into the camera)."""
    vol, depth, Rt, Xw, nw, pairs, residual = _plane_setup()
    h = 1e-6
    J = np.zeros((len(Xw), 6))
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        Tp, Tm = O.se3_exp_left(Rt[0], e), O.se3_exp_left(Rt[0], -e)
        J[:, k] = [(residual(Tp, i) - residual(Tm, i)) / (2 * h) for i in range(len(Xw))]
    want = 2.0 * (J.T @ J) / len(Xw)
    H = O.frame_normal_matrix(vol, depth[None], Rt, Xw, pairs, normals=nw)
    np.testing.assert_allclose(H[0], want, rtol=1e-6, atol=1e-8 * np.abs(want).max())
    # its gradient is sum 2 r J / K with the sampled residuals, at a pose off
    # the plane's fit; without normals it is the exact (slope) gradient
    off = O.se3_exp_left(Rt[0], np.array([0.002, -0.001, 0.002, 1e-3, -1.5e-3, 5e-4]))[None]
    _, g_exact, r = O.frame_loss(vol, depth[None], off, Xw, pairs)
    Hn, gn = O.frame_normal_matrix(vol, depth[None], off, Xw, pairs, normals=nw, want_grad=True)
    pc = Xw @ off[0, :9].reshape(3, 3).T + off[0, 9:]
    nc = nw @ off[0, :9].reshape(3, 3).T
    j = -(pc[:, 2] / np.sum(nc * pc, 1))[:, None] * nc
    Jo = np.c_[j, np.cross(pc, j)]
    np.testing.assert_allclose(gn[0], 2.0 * (r[:, None] * Jo).sum(0) / len(Xw), rtol=1e-9,
                               atol=1e-12 * np.abs(gn).max())
    np.testing.assert_allclose(Hn[0], 2.0 * Jo.T @ Jo / len(Xw), rtol=1e-9,
                               atol=1e-12 * np.abs(Hn).max())
    _, gs = O.frame_normal_matrix(vol, depth[None], off, Xw, pairs, want_grad=True)
    np.testing.assert_allclose(gs[0], g_exact[0], rtol=1e-12, atol=1e-15 * np.abs(g_exact).max())
    # on a plane the two linearisations agree to first order in the
    # residual (millimetres against a 1.3 m depth)
    np.testing.assert_allclose(gn[0], g_exact[0], rtol=0, atol=1e-2 * np.abs(g_exact).max())
    # a plane fixes three coordinates (the shift along n, two tilts)
    ev = np.linalg.eigvalsh(H[0])
    assert ev[3] > 1e-3 * ev[5] and ev[2] < 1e-12 * ev[5]
    # a flipped normal gives the same matrix; grazing pairs add nothing
    np.testing.assert_allclose(O.frame_normal_matrix(vol, depth[None], Rt, Xw, pairs, normals=-nw)[0],
                               H[0], rtol=1e-12, atol=0)
    side = np.cross(nw[0], [1.0, 0.0, 0.0])
    side = np.tile(side / np.linalg.norm(side), (len(Xw), 1))
    Hs = O.frame_normal_matrix(vol, depth[None], Rt, Xw, pairs, normals=side)
    pc = Xw @ Rt[0, :9].reshape(3, 3).T + Rt[0, 9:]
    ns = side @ Rt[0, :9].reshape(3, 3).T
    seen = np.abs(np.sum(ns * pc, 1)) >= 0.1 * np.linalg.norm(pc, axis=1)
    j = -(pc[:, 2] / np.sum(ns * pc, 1))[:, None] * ns
    Js = np.c_[j, np.cross(pc, j)][seen]
    np.testing.assert_allclose(Hs[0], 2.0 * Js.T @ Js / len(Xw), rtol=1e-9,
                               atol=1e-12 * np.abs(Hs).max())
    assert 0 < seen.sum() < len(Xw)


def test_gn_update_fits_a_plane_in_one_step():
    """One full Gauss-Newton step (step 1, no damping) from a pose 3 mm and
    0.1 degree off brings the residuals of points on a plane to nearly zero
    (the residual is linear in xi up to second order) without moving the
    pose much further than its observable error (the coordinates a plane
    leaves free take no step), and H = 0 leaves a pose unchanged."""
    vol, depth, Rt, Xw, nw, pairs, _ = _plane_setup(seed=34)
    off = O.se3_exp_left(Rt[0], np.array([0.002, -0.001, 0.002, 1e-3, -1.5e-3, 5e-4]))[None]
    L0, g, r0 = O.frame_loss(vol, depth[None], off, Xw, pairs)
    H = O.frame_normal_matrix(vol, depth[None], off, Xw, pairs, normals=nw)
    new = O.gn_update(off, g, H, step=1.0, damping=0.0)
    L1, _, r1 = O.frame_loss(vol, depth[None], new.astype(np.float64), Xw, pairs)
    assert np.abs(r0).max() > 2e-3
    assert np.abs(r1).max() < 2e-2 * np.abs(r0).max(), (np.abs(r0).max(), np.abs(r1).max())
    # the pose moved by about the size of the observable error, not more
    a, c = scenes.pose_errors(new.astype(np.float64), off)
    assert a[0] < 0.3 and c[0] < 0.02
    # half a step halves the residuals' observable part
    half = O.gn_update(off, g, H, step=0.5, damping=0.0)
    _, _, rh = O.frame_loss(vol, depth[None], half.astype(np.float64), Xw, pairs)
    np.testing.assert_allclose(rh, 0.5 * r0, atol=0.03 * np.abs(r0).max())
    same = O.gn_update(off, g, np.zeros((1, 6, 6)))
    assert np.array_equal(same, off.astype(np.float32))
