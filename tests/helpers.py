"""Shared helpers of the parity tests: run a workload through the HIP path
(librecon.so via recon) and through the CPU oracle on the same inputs."""
from __future__ import annotations

import numpy as np

import oracle_ffi as O


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint32)


def assert_bitexact(a: np.ndarray, b: np.ndarray, what: str):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    diff = bits(a) != bits(b)
    if diff.any():
        idx = np.argwhere(diff)[:5]
        raise AssertionError(f"{what}: {int(diff.sum())} of {a.size} differ, first at "
                             f"{idx.tolist()}: {a[tuple(idx[0])]!r} vs {b[tuple(idx[0])]!r}")


def oracle_run(cfg, depth: np.ndarray, poses: np.ndarray, f64=False, nthreads=0):
    """Integrates frames with the float32 oracle (and the float64 twin)."""
    vol = O.volume_from_config(cfg)
    D, W = O.init(vol)
    counts = []
    out = {"vol": vol}
    if f64:
        D64, W64 = D.astype(np.float64), W.astype(np.float64)
        flips = []
    for f in range(depth.shape[0]):
        counts.append(O.integrate_f32(vol, D, W, depth[f], poses[f], nthreads=nthreads))
        if f64:
            n64, fl = O.integrate_f64(vol, D64, W64, depth[f], poses[f], nthreads=nthreads)
            assert n64 == counts[-1]
            flips.append(fl)
    out.update(D=D, W=W, counts=counts)
    if f64:
        out.update(D64=D64, W64=W64, flips=flips)
    return out


def gpu_run(cfg, depth: np.ndarray, poses: np.ndarray, fused=False, pitch_pad=0, count=False):
    import torch

    import recon

    vol = recon.TSDFVolume(cfg)
    F, H, Wd = depth.shape
    if pitch_pad:
        buf = torch.zeros((F, H, Wd + pitch_pad), dtype=torch.float32, device="cuda")
        buf[:, :, :Wd] = torch.from_numpy(depth).cuda()
        dev = buf[:, :, :Wd]
    else:
        dev = torch.from_numpy(np.ascontiguousarray(depth)).cuda()
    counters = torch.zeros(F, dtype=torch.int64, device="cuda") if count else None
    if fused:
        if count:
            vol.set_update_counter(counters[0:1])
        vol.integrate_frames(dev, poses)
    else:
        for f in range(F):
            if count:
                vol.set_update_counter(counters[f:f + 1])
            vol.integrate(dev[f], poses[f])
    torch.cuda.synchronize()
    D, W = vol.download()
    res = {"D": D, "W": W, "vol": vol}
    if count:
        res["counts"] = counters.cpu().numpy()
    return res
