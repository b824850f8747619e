"""The differentiable integrate on the GPU (recon_tsdf_diff_integrate,
recon_tsdf_surface_loss, recon_tsdf_diff_adjoint) against the float64
oracle of the same forward, loss and adjoint, and against float64 central
finite differences of the loss.

Both sides use the same float32-narrowed poses.  Decisions (footprint,
validity, the -trunc skip, the +trunc clamp) are taken in float32 on the GPU
and in float64 by the oracle; the gradient checks use the points whose
decisions no step of the check can change (oracle_stable_points), as the
oracle's own finite-difference tests do.  Relative error of a gradient is
max|got - want| / max|want|."""
import numpy as np
import pytest

import oracle_ffi as O
import recon
from diff_autograd import IntegrateSurfaceLoss
from helpers import assert_bitexact
from recon import scenes

pytestmark = pytest.mark.gpu


def _setup(torch, F=4, grid=64, vs=0.016, npts=8000, noise=True, h=1e-4, kink=1e-4, seed=21):
    w = scenes.plain_small(num_frames=F, grid=grid, voxel_size=vs)
    depth = w.render(noise=noise)
    rt32 = recon.pose_to_matrix(w.poses)
    Rt = rt32.astype(np.float64)
    vol_o = O.volume_from_config(w.config())
    pts = scenes.surface_points(w, npts, seed=seed)
    wts = O.stable_points(vol_o, depth, Rt, pts, h=h, kink=kink)
    dev = dict(device="cuda")
    return dict(w=w, depth=depth, rt32=rt32, Rt=Rt, vol_o=vol_o, pts=pts, wts=wts,
                depth_d=torch.from_numpy(depth).cuda(),
                rt_d=torch.from_numpy(rt32).cuda(),
                pts_d=torch.from_numpy(pts.astype(np.float32)).cuda(),
                wts_d=torch.from_numpy(wts.astype(np.float32)).cuda(), dev=dev)


def _rel(got, want):
    return float(np.abs(got - want).max() / np.abs(want).max())


def test_diff_forward_matches_oracles(cuda):
    """Bit-exact against the float32 restatement (the kernel's operation
    order); against float64, decisions agree on all but a few voxels and D
    within 1e-6 except where a sample straddles a depth edge (the bilinear
    slope there, up to ~1 m per pixel, turns float32 pixel rounding of
    ~1e-4 px into ~1e-4 m)."""
    s = _setup(cuda)
    v = recon.TSDFVolume(s["w"].config())
    print("diff forward 64^3 x4", flush=True)
    v.diff_integrate(s["depth_d"], s["rt_d"])
    cuda.cuda.synchronize()
    D, W = v.download()
    D32, W32 = O.diff_forward_f32(s["vol_o"], s["depth"], s["rt32"])
    assert_bitexact(W, W32, "W vs float32 oracle")
    assert_bitexact(D, D32, "D vs float32 oracle")
    Do, Wo = O.diff_forward(s["vol_o"], s["depth"], s["Rt"])
    same = W == Wo
    assert same.mean() > 0.999, (~same).sum()
    upd = same & (W > 0)
    assert upd.sum() > 20000
    err = np.abs(D[upd] - Do[upd])
    assert (err < 1e-6).mean() > 0.995 and err.max() < 3e-4, (np.sort(err)[-5:])
    v.close()


def test_diff_forward_chunks_and_ragged_grid(cuda):
    """70 frames (two chunks of the 64-frame cull mask) on a grid whose x is
    not a multiple of 4: equal to two calls of 35 frames, and to the oracle."""
    w = scenes.tiny(grid=(37, 29, 23), voxel_size=0.6 / 37, image=(48, 36), num_frames=70)
    w.max_weight = 128.0
    depth = w.render()
    rt32 = recon.pose_to_matrix(w.poses)
    dd = cuda.from_numpy(depth).cuda()
    rd = cuda.from_numpy(rt32).cuda()
    a = recon.TSDFVolume(w.config())
    b = recon.TSDFVolume(w.config())
    print("diff forward 70 frames", flush=True)
    a.diff_integrate(dd, rd)
    b.diff_integrate(dd[:35], rd[:35].contiguous())
    b.diff_integrate(dd[35:], rd[35:].contiguous())
    cuda.cuda.synchronize()
    Da, Wa = a.download()
    Db, Wb = b.download()
    assert (Wa == Wb).all() and (Da == Db).all()
    D32, W32 = O.diff_forward_f32(O.volume_from_config(w.config()), depth, rt32)
    assert_bitexact(Wa, W32, "W vs float32 oracle")
    assert_bitexact(Da, D32, "D vs float32 oracle")
    a.close()
    b.close()


def test_surface_loss_matches_oracle(cuda):
    s = _setup(cuda)
    v = recon.TSDFVolume(s["w"].config())
    v.diff_integrate(s["depth_d"], s["rt_d"])
    print("surface loss", flush=True)
    loss, res, ids, lam = v.surface_loss(s["pts_d"], s["wts_d"])
    cuda.cuda.synchronize()
    D, _ = v.download()
    Lo, lam_o = O.surface_loss(s["vol_o"], D.astype(np.float64), s["pts"], s["wts"])
    assert abs(loss.item() - Lo) <= 1e-6 * Lo
    r_o = np.array([c @ D.ravel()[i] for i, c in (O.trilinear_support(s["vol_o"], p) for p in s["pts"])])
    assert np.abs(res.cpu().numpy() - r_o).max() < 1e-6
    dense = np.zeros(D.size)
    ids_h, lam_h = ids.cpu().numpy(), lam.cpu().numpy()
    ok = ids_h >= 0
    np.add.at(dense, ids_h[ok], lam_h[ok].astype(np.float64))
    assert _rel(dense, lam_o.ravel()) < 1e-5
    v.close()


def test_surface_loss_edge_points(cuda):
    """Points outside the grid get no support (residual 0, id -1); no points
    gives a zero loss."""
    w = scenes.tiny()
    v = recon.TSDFVolume(w.config())
    far = cuda.tensor([[1e30, 0, 0], [float("nan"), 0, 0], [-1e30, -1e30, -1e30], [5.0, 5.0, 5.0]],
                      dtype=cuda.float32, device="cuda")
    loss, res, ids, lam = v.surface_loss(far)
    cuda.cuda.synchronize()
    # the -1e30 point moves onto the grid's lower faces (reference clamp)
    ids_h = ids.cpu().numpy().reshape(4, 8)
    assert (ids_h[[0, 1, 3]] == -1).all() and (ids_h[2] >= 0).all()
    assert (res.cpu().numpy()[[0, 1, 3]] == 0).all()
    empty = cuda.zeros((0, 3), dtype=cuda.float32, device="cuda")
    loss0, _, _, _ = v.surface_loss(empty)
    cuda.cuda.synchronize()
    assert loss0.item() == 0.0
    v.close()


def test_adjoint_matches_float64_adjoint(cuda):
    """The same dL/dD on both sides: the oracle's lambda (stable points),
    rounded to float32.  The decision margins only need to cover float32
    rounding here (step 1e-6)."""
    s = _setup(cuda, h=1e-6)
    v = recon.TSDFVolume(s["w"].config())
    D, _ = O.diff_forward(s["vol_o"], s["depth"], s["Rt"])
    _, lam_o = O.surface_loss(s["vol_o"], D, s["pts"], s["wts"])
    nz = np.flatnonzero(lam_o)
    lam32 = lam_o.ravel()[nz].astype(np.float32)
    lam_r = np.zeros(D.size)
    lam_r[nz] = lam32
    gp_o, gd_o = O.diff_adjoint(s["vol_o"], s["depth"], s["Rt"], lam_r.reshape(D.shape))
    print("adjoint vs float64 adjoint", flush=True)
    gp, gd = v.diff_adjoint(s["depth_d"], s["rt_d"], cuda.from_numpy(nz.astype(np.int64)).cuda(),
                            cuda.from_numpy(lam32).cuda())
    cuda.cuda.synchronize()
    gp, gd = gp.cpu().numpy(), gd.cpu().numpy()
    assert (np.abs(gp_o).max(axis=1) > 0).all()
    assert _rel(gp, gp_o) < 1e-4, (gp, gp_o)
    assert _rel(gd, gd_o) < 1e-4
    v.close()


def test_adjoint_pairs_semantics_and_determinism(cuda):
    """Repeated indices add, indices outside the volume are ignored, an empty
    list gives zero gradients, and two runs agree bit for bit."""
    s = _setup(cuda, npts=2000)
    v = recon.TSDFVolume(s["w"].config())
    v.diff_integrate(s["depth_d"], s["rt_d"])
    _, _, ids, lam = v.surface_loss(s["pts_d"], s["wts_d"])
    g1, d1 = v.diff_adjoint(s["depth_d"], s["rt_d"], ids, lam)
    g2, d2 = v.diff_adjoint(s["depth_d"], s["rt_d"], ids, lam)
    # the same pairs twice, each at half weight, plus out-of-range indices
    n = v.num_voxels
    ids2 = cuda.cat([ids, ids, cuda.tensor([-5, n, n + 7], dtype=cuda.int64, device="cuda")])
    lam2 = cuda.cat([lam * 0.5, lam * 0.5, cuda.ones(3, dtype=cuda.float32, device="cuda")])
    g3, _ = v.diff_adjoint(s["depth_d"], s["rt_d"], ids2, lam2, want_depth=False)
    e = cuda.zeros(0, dtype=cuda.int64, device="cuda")
    g4, d4 = v.diff_adjoint(s["depth_d"], s["rt_d"], e, cuda.zeros(0, dtype=cuda.float32, device="cuda"))
    cuda.cuda.synchronize()
    assert cuda.equal(g1, g2)
    np.testing.assert_allclose(g3.cpu().numpy(), g1.cpu().numpy(), rtol=1e-5,
                               atol=1e-6 * g1.abs().max().item())
    assert (g4 == 0).all() and (d4 == 0).all()
    # depth gradients are exact fixed-point sums: identical run to run
    assert cuda.equal(d1, d2) and d1.abs().max().item() > 0
    v.close()


def test_autograd_function_matches_central_differences(cuda):
    """torch.autograd through the C-ABI: gradients in xi (at 0) and in the
    depth maps against float64 central differences of the oracle's loss
    (step 1e-4 on smooth depth), and against the float64 adjoint."""
    s = _setup(cuda, noise=False, h=1e-4, kink=2e-5)
    assert s["wts"].mean() > 0.1
    F = s["depth"].shape[0]
    v = recon.TSDFVolume(s["w"].config())
    xi = cuda.zeros((F, 6), dtype=cuda.float64, device="cuda", requires_grad=True)
    depth = s["depth_d"].clone().requires_grad_(True)
    print("autograd forward/backward", flush=True)
    L = IntegrateSurfaceLoss.apply(xi, depth, v, s["rt_d"], s["pts_d"], s["wts_d"])
    L.backward()
    cuda.cuda.synchronize()
    gp = xi.grad.cpu().numpy()
    gd = depth.grad.cpu().numpy()
    # float64 reference: forward, loss, adjoint, and central differences
    Do, _ = O.diff_forward(s["vol_o"], s["depth"], s["Rt"])
    Lo, lam_o = O.surface_loss(s["vol_o"], Do, s["pts"], s["wts"])
    assert abs(L.item() - Lo) <= 1e-3 * Lo
    gp_o, gd_o = O.diff_adjoint(s["vol_o"], s["depth"], s["Rt"], lam_o)
    assert _rel(gp, gp_o) < 1e-4
    assert _rel(gd, gd_o) < 1e-4
    h = 1e-4
    fd = np.zeros((F, 6))
    for f in range(F):
        for i in range(6):
            e = np.zeros(6)
            e[i] = h
            Rp, Rm = s["Rt"].copy(), s["Rt"].copy()
            Rp[f] = O.se3_exp_left(s["Rt"][f], e)
            Rm[f] = O.se3_exp_left(s["Rt"][f], -e)
            Lp = O.surface_loss(s["vol_o"], O.diff_forward(s["vol_o"], s["depth"], Rp)[0], s["pts"],
                                s["wts"], want_lambda=False)
            Lm = O.surface_loss(s["vol_o"], O.diff_forward(s["vol_o"], s["depth"], Rm)[0], s["pts"],
                                s["wts"], want_lambda=False)
            fd[f, i] = (Lp - Lm) / (2 * h)
    assert _rel(gp, fd) < 1e-4, (gp, fd)
    # depth: the largest entries, by central differences on the float32 map
    flat = np.abs(gd_o).ravel()
    for idx in np.argsort(-flat)[:8]:
        f, j, i = np.unravel_index(idx, s["depth"].shape)
        dp, dm = s["depth"].copy(), s["depth"].copy()
        dp[f, j, i] = np.float32(dp[f, j, i] + 1e-4)
        dm[f, j, i] = np.float32(dm[f, j, i] - 1e-4)
        step = np.float64(dp[f, j, i]) - np.float64(dm[f, j, i])
        Lp = O.surface_loss(s["vol_o"], O.diff_forward(s["vol_o"], dp, s["Rt"])[0], s["pts"], s["wts"],
                            want_lambda=False)
        Lm = O.surface_loss(s["vol_o"], O.diff_forward(s["vol_o"], dm, s["Rt"])[0], s["pts"], s["wts"],
                            want_lambda=False)
        assert abs((Lp - Lm) / step - gd[f, j, i]) <= 1e-4 * flat.max()
    v.close()


def test_differentiable_256_20_frames_against_float64_adjoint(cuda):
    """The differentiable workload at full size: 256^3, voxel 6 mm, 20 frames
    of 640x480, 100k surface points.  GPU forward + loss + adjoint against the
    float64 chain on the stable points."""
    w = scenes.differentiable()
    depth = w.render()
    rt32 = recon.pose_to_matrix(w.poses)
    Rt = rt32.astype(np.float64)
    vol_o = O.volume_from_config(w.config())
    pts = scenes.surface_points(w, 100_000)
    wts = O.stable_points(vol_o, depth, Rt, pts, h=1e-6, kink=1e-4)
    assert wts.sum() > 1000
    v = recon.TSDFVolume(w.config())
    dd = cuda.from_numpy(depth).cuda()
    rd = cuda.from_numpy(rt32).cuda()
    print("differentiable 256^3 x20", flush=True)
    v.diff_integrate(dd, rd)
    loss, _, ids, lam = v.surface_loss(cuda.from_numpy(pts.astype(np.float32)).cuda(),
                                       cuda.from_numpy(wts.astype(np.float32)).cuda())
    gp, _ = v.diff_adjoint(dd, rd, ids, lam, want_depth=False)
    cuda.cuda.synchronize()
    Do, _ = O.diff_forward(vol_o, depth, Rt)
    Lo, lam_o = O.surface_loss(vol_o, Do, pts, wts)
    gp_o, _ = O.diff_adjoint(vol_o, depth, Rt, lam_o, want_depth=False)
    assert abs(loss.item() - Lo) <= 1e-3 * Lo
    assert _rel(gp.cpu().numpy(), gp_o) < 1e-4
    v.close()
