"""Batched pose refinement (recon_diff_batch_*): the batch forward equals the
single-volume differentiable path bit for bit (D, W, loss), its gradients
equal the single-volume adjoint's to float64 reassociation (which the
float64 oracle and central differences pin), and refinement of perturbed poses brings the surface loss
to its ground-truth level and the fused mesh towards the analytic surface."""
import numpy as np
import pytest

import recon
from helpers import assert_bitexact
from recon import scenes

pytestmark = pytest.mark.gpu

# weight of the model-to-frame term against the fused-TSDF surface loss
FRAME_WEIGHT = 1.0


def _fragment_set(torch, n=3, grid=64, frames=6, npts=4000, ragged=True):
    ws = scenes.pose_refinement(num_fragments=n, frames=frames, grid=grid, voxel_size=1.024 / grid)
    if ragged:
        # a volume with a grid that is not a multiple of 4 and a smaller image
        t = scenes.tiny(grid=(37, 29, 23), voxel_size=0.6 / 37, image=(48, 36), num_frames=5)
        t.max_weight = 128.0
        t.meta["init"] = t.poses.copy()
        ws.append(t)
    vols = []
    for i, w in enumerate(ws):
        depth = torch.from_numpy(w.render()).cuda()
        pts = torch.from_numpy(scenes.surface_points(w, npts, seed=100 + i).astype(np.float32)).cuda()
        wts = torch.rand(pts.shape[0], generator=torch.Generator().manual_seed(i)).float().cuda()
        vols.append((w, depth, pts, wts))
    return vols


def _single(torch, w, depth, rt, pts, wts):
    v = recon.TSDFVolume(w.config())
    v.diff_integrate(depth, rt)
    loss, _, ids, lam = v.surface_loss(pts, wts)
    gp, _ = v.diff_adjoint(depth, rt, ids, lam, want_depth=False)
    torch.cuda.synchronize()
    D, W = v.download()
    v.close()
    return D, W, loss.item(), gp.cpu().numpy()


def _check_forward_backward(torch, vols, reps):
    b = recon.DiffBatch([(w.config(), d, p, q) for w, d, p, q in vols])
    rts = [torch.from_numpy(recon.pose_to_matrix(w.meta["init"])).cuda() for w, _, _, _ in vols]
    rt_all = torch.cat(rts).contiguous()
    print(f"diff batch forward/backward, {len(vols)} volumes", flush=True)
    for rep in range(reps):
        loss = b.forward(rt_all)
        grad = b.backward()
        torch.cuda.synchronize()
        f0 = 0
        for i, (w, d, p, q) in enumerate(vols):
            D, W, L, gp = _single(torch, w, d, rts[i], p, q)
            Db, Wb = b.download(i)
            assert_bitexact(Wb, W, f"volume {i} W")
            assert_bitexact(Db, D, f"volume {i} D")
            assert loss[i].item() == L, (i, loss[i].item(), L)
            F = d.shape[0]
            g = grad[f0:f0 + F].cpu().numpy()
            # float64 sums of the same terms, grouped in units of another
            # size: equal to ~1e-11 of the largest component
            np.testing.assert_allclose(g, gp, rtol=1e-9, atol=1e-9 * np.abs(gp).max())
            f0 += F
    # run-to-run identical gradients
    b.forward(rt_all)
    g2 = b.backward()
    torch.cuda.synchronize()
    assert torch.equal(g2, grad)
    b.close()


def test_batch_forward_backward_equal_single_volume(cuda):
    _check_forward_backward(cuda, _fragment_set(cuda), reps=2)


def test_batch_forward_backward_full_size(cuda):
    """All 32 fragments of the refinement workload (128^3 at 8 mm, 10 frames
    of 640x480, 20k points each): losses, D, W and gradients equal the
    single-volume path's."""
    _check_forward_backward(cuda, _fragment_set(cuda, n=32, grid=128, frames=10, npts=20000,
                                                ragged=False), reps=1)


def test_batch_volume_without_points(cuda):
    vols = _fragment_set(cuda, n=2, ragged=False)
    w, d, p, q = vols[1]
    empty = cuda.zeros((0, 3), dtype=cuda.float32, device="cuda")
    b = recon.DiffBatch([(vols[0][0].config(), vols[0][1], vols[0][2], None),
                         (w.config(), d, empty, None)])
    rt = cuda.cat([cuda.from_numpy(recon.pose_to_matrix(x[0].poses)).cuda() for x in vols]).contiguous()
    loss = b.forward(rt)
    g = b.backward()
    cuda.cuda.synchronize()
    F0 = vols[0][1].shape[0]
    assert loss[1].item() == 0.0 and loss[0].item() > 0.0
    assert (g[F0:] == 0).all() and (g[:F0].abs().max() > 0)
    b.close()


def _plain_mesh_rms(torch, w, depth, rts):
    """RMS distance to the analytic surface of the plain integrate's mesh
    (observed vertices) with poses rts (F, 12)."""
    v = recon.TSDFVolume(w.config())
    for f in range(rts.shape[0]):
        R = rts[f, :9].reshape(3, 3).astype(np.float64)
        v.integrate(depth[f], np.r_[scenes.quat_from_matrix(R), rts[f, 9:]])
    verts, _ = v.extract_mesh()
    _, W = v.download()
    v.close()
    vv = verts.cpu().numpy().reshape(-1, 3)
    return scenes.mesh_rms(w, vv[scenes.observed_vertices(w, W, vv)])


def _refinement_spec(torch, ws, frame_weight, npts=20000, min_weight=0.1, frame_trunc=0.0):
    """Batch volumes: the fused-TSDF surface points of each fragment and,
    with frame_weight > 0, its model-to-frame pairs (chosen at the initial
    poses)."""
    spec, depths = [], []
    for i, w in enumerate(ws):
        depth = w.render()
        pts, wts, nrm = scenes.refinement_points(w, depth, npts, seed=1000 + i, with_normals=True)
        assert wts.mean() > min_weight
        d = torch.from_numpy(depth).cuda()
        depths.append(d)
        item = (w.config(), d, torch.from_numpy(pts.astype(np.float32)).cuda(),
                torch.from_numpy(wts).cuda())
        if frame_weight > 0:
            pairs = scenes.frame_pairs(w, depth, pts, nrm)
            item = item + (torch.from_numpy(pairs).cuda(), frame_weight, frame_trunc)
        spec.append(item)
    return spec, depths


def test_frame_term_matches_oracle(cuda):
    """The model-to-frame term: a batch with pairs minus the same batch
    without them gives, per volume, weight * the float64 oracle's loss and
    per frame its pose gradient (the poses, points and depths the GPU reads,
    widened)."""
    import oracle_ffi as O

    ws = scenes.pose_refinement(num_fragments=2, frames=6, grid=64, voxel_size=1.024 / 64)
    lam, tau = 0.7, 0.004
    spec1, _ = _refinement_spec(cuda, ws, lam, npts=4000, min_weight=0.0, frame_trunc=tau)
    spec0 = [s[:4] for s in spec1]
    init = np.concatenate([recon.pose_to_matrix(w.meta["init"]) for w in ws])
    rt = cuda.from_numpy(init).cuda()
    b0, b1 = recon.DiffBatch(spec0), recon.DiffBatch(spec1)
    l0, l1 = b0.forward(rt), b1.forward(rt)
    g0, g1 = b0.backward(), b1.backward()
    cuda.cuda.synchronize()
    dl = (l1 - l0).cpu().numpy()
    dg = (g1 - g0).cpu().numpy()
    b0.close()
    b1.close()
    f0 = 0
    for i, (w, item) in enumerate(zip(ws, spec1)):
        F = w.num_frames
        vol = O.volume_from_config(w.config())
        pts = item[2].cpu().numpy().astype(np.float64)
        pairs = item[4].cpu().numpy()
        Rt = init[f0:f0 + F].astype(np.float64)
        L, g, res = O.frame_loss(vol, item[1].cpu().numpy(), Rt, pts, pairs, tau=tau)
        assert len(pairs) > 1000 and 0.05 < (np.abs(res) >= tau).mean() < 0.95
        assert abs(dl[i] - lam * L) <= 1e-9 * lam * L, (i, dl[i], lam * L)
        np.testing.assert_allclose(dg[f0:f0 + F], lam * g, rtol=1e-7, atol=1e-9 * lam * np.abs(g).max())
        f0 += F


def test_refine_improves_fit(cuda):
    """Four 128^3 fragments, 10 frames, poses perturbed by up to 1 degree and
    1 cm, 50 steps on the fused-TSDF surface loss plus the model-to-frame
    term: the loss falls to the ground-truth level, the plain integrate's
    mesh moves towards the analytic surface, the fragments' rigid
    (common-mode) error shrinks, and so do the frames' own errors -- the mean
    rotation and camera-centre errors at least halve and no frame ends worse
    than the worst frame started.  (The fused TSDF alone is linear in the
    frames' projective distances inside the truncation band, so it observes
    mainly the common mode; the model-to-frame term observes each frame.)"""
    ws = scenes.pose_refinement(num_fragments=4)
    spec, depths = _refinement_spec(cuda, ws, FRAME_WEIGHT)
    init = np.concatenate([recon.pose_to_matrix(w.meta["init"]) for w in ws])
    gt = np.concatenate([recon.pose_to_matrix(w.poses) for w in ws])
    b = recon.DiffBatch(spec)
    loss_gt = b.forward(cuda.from_numpy(gt).cuda()).sum().item()
    rt = cuda.from_numpy(init.copy()).cuda()
    print("refine 50 steps", flush=True)
    hist = b.refine(rt, 50, lr_translation=5e-4, lr_rotation=1e-3, decay=0.95)
    cuda.cuda.synchronize()
    refined = rt.cpu().numpy()
    h = hist.cpu().numpy()
    b.close()
    # common mode: mean over each fragment's frames of the world-frame error
    def common(rts):
        tw = scenes.world_error_twists(rts, gt).reshape(len(ws), -1, 6).mean(axis=1)
        return np.linalg.norm(tw[:, :3], axis=1).mean()
    m0 = np.mean([_plain_mesh_rms(cuda, w, d, init[10 * i:10 * i + 10]) for i, (w, d) in
                  enumerate(zip(ws, depths))])
    m1 = np.mean([_plain_mesh_rms(cuda, w, d, refined[10 * i:10 * i + 10]) for i, (w, d) in
                  enumerate(zip(ws, depths))])
    a0, c0 = scenes.pose_errors(init, gt)
    a1, c1 = scenes.pose_errors(refined, gt)
    print(f"loss {h[0].sum():.3e} -> {h[-1].sum():.3e} (ground truth {loss_gt:.3e}); mesh rms "
          f"{m0 * 1e3:.2f} -> {m1 * 1e3:.2f} mm; common mode {common(init) * 1e3:.2f} -> "
          f"{common(refined) * 1e3:.2f} mm; per frame mean {a0.mean():.3f} -> {a1.mean():.3f} deg, "
          f"{c0.mean() * 1e3:.2f} -> {c1.mean() * 1e3:.2f} mm, max {a0.max():.3f} -> "
          f"{a1.max():.3f} deg, {c0.max() * 1e3:.2f} -> {c1.max() * 1e3:.2f} mm", flush=True)
    assert h[-1].sum() < 1.2 * loss_gt and h[-1].sum() < 0.5 * h[0].sum()
    assert m1 < 0.9 * m0
    assert common(refined) < 0.8 * common(init)
    assert a1.mean() <= 0.5 * a0.mean() and c1.mean() <= 0.5 * c0.mean()
    assert a1.max() <= a0.max() and c1.max() <= c0.max()
    # rotations stay orthonormal
    R = refined[:, :9].reshape(-1, 3, 3).astype(np.float64)
    assert np.abs(R @ R.transpose(0, 2, 1) - np.eye(3)).max() < 1e-5
