"""recon_mc_extract_noalloc (marching cubes with a static device scratch,
for callers that cannot allocate) against recon_mc_extract through the
C-ABI: identical triangle counts, vertex bits and index lists on two calls
in a row, the second on a smaller field (the static scratch keeps the first
call's contents).  The refusal above its size limit is a host check, tested
without a GPU in test_abi.py."""
import ctypes as C

import numpy as np
import pytest

import recon
from recon import ffi, scenes

pytestmark = pytest.mark.gpu


def _field(torch, w):
    vol = recon.TSDFVolume(w.config())
    dev = torch.from_numpy(w.render()).cuda()
    for f in range(w.num_frames):
        vol.integrate(dev[f], w.poses[f])
    torch.cuda.synchronize()
    D, _ = vol.download()
    info = vol.info
    vol.close()
    return torch.from_numpy(D).cuda(), list(info.grid), list(info.voxel_size), list(info.global_offset)


def _extract(torch, fn, field, grid, vs, off, cap):
    verts = torch.full((max(cap, 1), 3, 3), -7.0, dtype=torch.float32, device="cuda")
    tris = torch.full((max(cap, 1), 3), -7, dtype=torch.int32, device="cuda")
    n = C.c_int64()
    st = fn(field.data_ptr(), ffi.i32x3(grid), ffi.f32x3(vs), ffi.f32x3(off), verts.data_ptr(),
            tris.data_ptr(), cap, C.byref(n), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return st, int(n.value), verts[:n.value].cpu().numpy(), tris[:n.value].cpu().numpy()


def test_noalloc_equals_allocating_extract(cuda):
    lib = ffi.load()
    big = scenes.plain_small()                     # 128^3, 10 frames
    small = scenes.tiny(grid=(40, 36, 28), voxel_size=0.025, image=(64, 48), num_frames=3)
    for name, w in (("128^3", big), ("40x36x28", small)):
        field, grid, vs, off = _field(cuda, w)
        n_ref = C.c_int64()
        assert lib.recon_mc_count(field.data_ptr(), ffi.i32x3(grid), ffi.f32x3(vs), ffi.f32x3(off),
                                  C.byref(n_ref), None) == 0
        cap = int(n_ref.value)
        assert cap > 1000, (name, cap)
        print(f"no-alloc marching cubes on {name}: {cap} triangles", flush=True)
        st0, n0, v0, t0 = _extract(cuda, lib.recon_mc_extract, field, grid, vs, off, cap)
        st1, n1, v1, t1 = _extract(cuda, lib.recon_mc_extract_noalloc, field, grid, vs, off, cap)
        assert st0 == 0 and st1 == 0, (st0, st1, lib.recon_last_error())
        assert n0 == n1 == cap
        assert np.array_equal(v0.view(np.uint32), v1.view(np.uint32)), f"{name}: vertices differ"
        assert np.array_equal(t0, t1), f"{name}: indices differ"
        # too small an output: the count is reported with RECON_ERR_CAPACITY
        st2, n2, _, _ = _extract(cuda, lib.recon_mc_extract_noalloc, field, grid, vs, off, cap - 1)
        assert st2 == ffi.ERR_CAPACITY and n2 == cap
