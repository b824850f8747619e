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
FRAME_WEIGHT = 10.0


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


def _refinement_spec(torch, ws, frame_weight, npts=20000, min_weight=0.1, frame_trunc=0.0,
                     normals=False):
    """Batch volumes: the fused-TSDF surface points of each fragment and,
    with frame_weight > 0, its model-to-frame pairs (chosen at the initial
    poses) and, with `normals`, the points' normals."""
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
            if normals:
                item = item + (torch.from_numpy(nrm.astype(np.float32)).cuda(),)
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


def _errors_and_mesh(torch, ws, depths, init, refined, gt):
    def common(rts):
        tw = scenes.world_error_twists(rts, gt).reshape(len(ws), -1, 6).mean(axis=1)
        return np.linalg.norm(tw[:, :3], axis=1).mean()
    F = ws[0].num_frames
    m0 = np.mean([_plain_mesh_rms(torch, w, d, init[F * i:F * i + F]) for i, (w, d) in
                  enumerate(zip(ws, depths))])
    m1 = np.mean([_plain_mesh_rms(torch, w, d, refined[F * i:F * i + F]) for i, (w, d) in
                  enumerate(zip(ws, depths))])
    a0, c0 = scenes.pose_errors(init, gt)
    a1, c1 = scenes.pose_errors(refined, gt)
    print(f"mesh rms {m0 * 1e3:.2f} -> {m1 * 1e3:.2f} mm; common mode {common(init) * 1e3:.2f} -> "
          f"{common(refined) * 1e3:.2f} mm; per frame mean {a0.mean():.3f} -> {a1.mean():.3f} deg, "
          f"{c0.mean() * 1e3:.2f} -> {c1.mean() * 1e3:.2f} mm, max {a0.max():.3f} -> "
          f"{a1.max():.3f} deg, {c0.max() * 1e3:.2f} -> {c1.max() * 1e3:.2f} mm", flush=True)
    return m0, m1, common(init), common(refined), a0, c0, a1, c1


def test_refine_improves_fit(cuda):
    """Adam: four 128^3 fragments, 10 frames, poses perturbed by up to 1
    degree and 1 cm, 50 steps on the fused-TSDF surface loss plus the
    model-to-frame term: the loss falls below a tenth of its start, the plain
    integrate's mesh moves towards the analytic surface, the fragments' rigid
    (common-mode) error shrinks and so do the mean per-frame errors.  (A first-
    order step moves the weakly fixed combinations of a frame's rotation and
    shift slowly; refine_gn, below, fixes every frame.)"""
    ws = scenes.pose_refinement(num_fragments=4)
    spec, depths = _refinement_spec(cuda, ws, FRAME_WEIGHT, frame_trunc=0.015)
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
    print(f"loss {h[0].sum():.3e} -> {h[-1].sum():.3e} (ground truth {loss_gt:.3e})")
    m0, m1, cm0, cm1, a0, c0, a1, c1 = _errors_and_mesh(cuda, ws, depths, init, refined, gt)
    assert h[-1].sum() < 0.1 * h[0].sum()
    assert m1 < 0.9 * m0
    assert cm1 < 0.8 * cm0
    assert a1.mean() < a0.mean() and c1.mean() < c0.mean()
    # rotations stay orthonormal
    R = refined[:, :9].reshape(-1, 3, 3).astype(np.float64)
    assert np.abs(R @ R.transpose(0, 2, 1) - np.eye(3)).max() < 1e-5


# Gauss-Newton refinement: the truncation shrinks from 10 mm to 3 mm; the
# normal matrix covers the model-to-frame term only, so that term leads
GN_TRUNC, GN_RATIO, GN_WEIGHT = 0.010, 0.3, 100.0


def test_refine_gn_fixes_every_frame(cuda):
    """Damped Gauss-Newton with the point-to-plane normal matrix: four
    fragments as above, 20 steps.  Every frame's own error falls: the mean
    rotation and camera-centre errors to a quarter or less of their start, no
    frame ends worse than the worst frame started, and the mesh of the plain
    integrate with the refined poses moves towards the analytic surface."""
    ws = scenes.pose_refinement(num_fragments=4)
    spec, depths = _refinement_spec(cuda, ws, GN_WEIGHT, frame_trunc=GN_TRUNC, normals=True)
    init = np.concatenate([recon.pose_to_matrix(w.meta["init"]) for w in ws])
    gt = np.concatenate([recon.pose_to_matrix(w.poses) for w in ws])
    b = recon.DiffBatch(spec)
    rt = cuda.from_numpy(init.copy()).cuda()
    print("refine_gn 20 steps", flush=True)
    hist = b.refine_gn(rt, 20, step=0.7, damping=1e-3, trunc_ratio=GN_RATIO)
    cuda.cuda.synchronize()
    refined = rt.cpu().numpy()
    h = hist.cpu().numpy()
    b.close()
    print(f"loss {h[0].sum():.3e} -> {h[-1].sum():.3e}")
    m0, m1, cm0, cm1, a0, c0, a1, c1 = _errors_and_mesh(cuda, ws, depths, init, refined, gt)
    assert a1.mean() <= 0.25 * a0.mean() and c1.mean() <= 0.25 * c0.mean()
    assert a1.max() <= a0.max() and c1.max() <= c0.max()
    assert m1 < 0.8 * m0
    R = refined[:, :9].reshape(-1, 3, 3).astype(np.float64)
    assert np.abs(R @ R.transpose(0, 2, 1) - np.eye(3)).max() < 1e-5


def test_gn_step_matches_oracle(cuda):
    """One Gauss-Newton step equals the oracle's step (oracle_ffi.gn_update)
    from the batch's surface-loss gradient (a batch without pairs) plus the
    float64 oracle's Gauss-Newton system of the model-to-frame term
    (point-to-plane with normals, the depth slopes without), to the float32
    rounding of the stored pose."""
    import oracle_ffi as O

    ws = scenes.pose_refinement(num_fragments=2, frames=6, grid=64, voxel_size=1.024 / 64)
    lam, step, damping = 3.0, 0.8, 0.05
    init = np.concatenate([recon.pose_to_matrix(w.meta["init"]) for w in ws])
    for normals in (True, False):
        spec, _ = _refinement_spec(cuda, ws, lam, npts=4000, min_weight=0.0,
                                   frame_trunc=GN_TRUNC, normals=normals)
        b0 = recon.DiffBatch([s[:4] for s in spec])
        rt = cuda.from_numpy(init.copy()).cuda()
        b0.forward(rt)
        g = b0.backward().cpu().numpy()
        b0.close()
        H = []
        f0 = 0
        for w, item in zip(ws, spec):
            F = w.num_frames
            vol = O.volume_from_config(w.config())
            nrm = item[7].cpu().numpy().astype(np.float64) if normals else None
            Hf, gf = O.frame_normal_matrix(vol, item[1].cpu().numpy(),
                                           init[f0:f0 + F].astype(np.float64),
                                           item[2].cpu().numpy().astype(np.float64),
                                           item[4].cpu().numpy(), normals=nrm, tau=GN_TRUNC,
                                           want_grad=True)
            H.append(lam * Hf)
            g[f0:f0 + F] += lam * gf
            f0 += F
        want = O.gn_update(init, g, np.concatenate(H), step=step, damping=damping)
        b = recon.DiffBatch(spec)
        print(f"one Gauss-Newton step, normals={normals}", flush=True)
        b.refine_gn(rt, 1, step=step, damping=damping, trunc_ratio=GN_RATIO)
        got = rt.cpu().numpy()
        b.close()
        assert np.abs(got - init).max() > 1e-4
        d = np.abs(got - want)
        assert d.max() <= 2.4e-7 and (d == 0).mean() > 0.9, (normals, d.max(), (d == 0).mean())


def test_gn_truncation_schedule(cuda):
    """Step k of refine_gn(K) scores the model-to-frame term at
    trunc * ratio^min(1, k / max(1, K / 2 - 1)): its recorded loss equals the
    forward loss, at that truncation, of the poses after k steps (runs of k
    and K steps share their first k steps)."""
    ws = scenes.pose_refinement(num_fragments=2, frames=6, grid=64, voxel_size=1.024 / 64)
    init = np.concatenate([recon.pose_to_matrix(w.meta["init"]) for w in ws])
    spec, _ = _refinement_spec(cuda, ws, GN_WEIGHT, npts=4000, min_weight=0.0,
                               frame_trunc=GN_TRUNC, normals=True)
    b = recon.DiffBatch(spec)
    K = 9                                     # ka = 3
    rt9 = cuda.from_numpy(init.copy()).cuda()
    h9 = b.refine_gn(rt9, K, trunc_ratio=GN_RATIO).cpu().numpy()
    for k, scale in ((1, GN_RATIO ** (1.0 / 3.0)), (8, GN_RATIO)):
        rtk = cuda.from_numpy(init.copy()).cuda()
        b.refine_gn(rtk, k, trunc_ratio=GN_RATIO)
        other = recon.DiffBatch([s[:6] + (GN_TRUNC * scale,) + s[7:] for s in spec])
        lk = other.forward(rtk).cpu().numpy()
        other.close()
        np.testing.assert_allclose(h9[k], lk, rtol=1e-12, atol=0)
        if k == 8:
            assert np.array_equal(h9[k], lk)
    b.close()


def test_refine_gn_needs_pairs_in_every_volume(cuda):
    ws = scenes.pose_refinement(num_fragments=2, frames=6, grid=64, voxel_size=1.024 / 64)
    spec, _ = _refinement_spec(cuda, ws, FRAME_WEIGHT, npts=4000, min_weight=0.0,
                               frame_trunc=GN_TRUNC, normals=True)
    b = recon.DiffBatch([spec[0], spec[1][:4]])
    init = np.concatenate([recon.pose_to_matrix(w.meta["init"]) for w in ws])
    rt = cuda.from_numpy(init.copy()).cuda()
    with pytest.raises(Exception, match="model-to-frame pairs"):
        b.refine_gn(rt, 2)
    assert np.array_equal(rt.cpu().numpy(), init.astype(np.float32))
    b.close()
