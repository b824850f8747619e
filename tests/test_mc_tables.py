"""The marching-cubes tables (include/recon/mc_tables.h) against the
reference's own tables and against what the cube geometry fixes.

SHA-256 digests are those of the reference's int32 little-endian
serialisations of marching_cubes_tables.cuh:11-15 (MC_EDGE_VERTICES),
:18-51 (MC_EDGE_TABLE), :93-350 (MC_TRI_TABLE) and the triangle counts.
"""
import hashlib
import re
from pathlib import Path

import numpy as np

HDR = Path(__file__).resolve().parents[1] / "include" / "recon" / "mc_tables.h"

SHA_EDGE_VERTICES = "9d4f01a93f0b8c80e703c494e5ea8bd455cf4410f970976239cccf276fd8e83c"
SHA_EDGE_TABLE = "ffc58719f11be7a8b34988740a15dcd043dc314a3e2fe01e917fd796001815b9"
SHA_TRI_TABLE = "85e6eb7486ad0101a95aaf3b332a15ba6e5d187a24d31c5eb77874d3aa45d996"
SHA_NTRI = "c3ae8bc49cfb9bece3b576fdb62e0da19d1f42c4d719b221bf04c364e02c492a"


def _init(name):
    txt = HDR.read_text()
    m = re.search(r"#define " + name + r" \{(.*?)\}\s*\n\n", txt, re.S)
    assert m, name
    return m.group(1)


def packed():
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]+)ull", _init("RECON_MC_TRI_PACKED_INIT"))]
    assert len(words) == 256
    return words


def tri_table():
    t = -np.ones((256, 16), np.int32)
    for c, w in enumerate(packed()):
        n = w & 0xF
        for k in range(3 * n):
            t[c, k] = (w >> (4 + 4 * k)) & 0xF
    return t


def edge_corners():
    vals = [int(v) for v in re.findall(r"\d+", _init("RECON_MC_EDGE_CORNERS_INIT"))]
    return np.array(vals, np.int32).reshape(12, 2)


def corner_offsets():
    vals = [int(v) for v in re.findall(r"\d+", _init("RECON_MC_CORNER_OFFSET_INIT"))]
    return np.array(vals, np.int32).reshape(8, 3)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<i4").tobytes()).hexdigest()


def edge_table(t):
    e = np.zeros(256, np.int32)
    for c in range(256):
        for v in t[c]:
            if v >= 0:
                e[c] |= 1 << int(v)
    return e


def test_hashes_match_reference_tables():
    t = tri_table()
    assert sha(t) == SHA_TRI_TABLE
    assert sha(edge_corners()) == SHA_EDGE_VERTICES
    assert sha(edge_table(t)) == SHA_EDGE_TABLE
    ntri = (t >= 0).sum(1) // 3
    assert sha(ntri) == SHA_NTRI
    assert ntri.max() == 5 and ntri.sum() == 820


def test_edges_join_adjacent_corners():
    off = corner_offsets()
    assert sorted(map(tuple, off)) == sorted({(x, y, z) for x in (0, 1) for y in (0, 1) for z in (0, 1)})
    for a, b in edge_corners():
        assert np.abs(off[a] - off[b]).sum() == 1  # a cube edge
    # the 12 edges are the 12 distinct cube edges
    assert len({tuple(sorted(p)) for p in edge_corners().tolist()}) == 12


def test_triangles_use_exactly_the_sign_change_edges():
    t, ec = tri_table(), edge_corners()
    for c in range(256):
        inside = [(c >> i) & 1 for i in range(8)]
        crossing = {e for e in range(12) if inside[ec[e, 0]] != inside[ec[e, 1]]}
        used = {int(v) for v in t[c] if v >= 0}
        assert used == crossing, c
        # complementary classes cut the same edges
        used_c = {int(v) for v in t[255 - c] if v >= 0}
        assert used_c == crossing
