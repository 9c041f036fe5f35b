"""The reference's homogeneous particle templates (sc0/sc1/bcc/fcc_crystal.dat,
single-file Gadget-2, read by loadHomogeneousTemplate, ic_basic.hpp:205-330),
kept as fixtures under tests/golden: the header decodes with the library's
gev_gadget2_header (metadata.hpp:159-178), and sc1 tiled N/4 times is the
cell-centred lattice that gev_pcls_init_lattice builds (CPU only; the
façade's loader is run on the GPU in test_host_driver.py)."""
import ctypes
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_template(name):
    """(header, positions [n, 3] in box units) as loadHomogeneousTemplate reads them"""
    from gevpm._lib import gev_gadget2_header
    raw = open(os.path.join(GOLD, name), "rb").read()
    assert ctypes.sizeof(gev_gadget2_header) == 256
    b1 = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    assert b1 == 256 and int(np.frombuffer(raw, np.int32, 1, 260)[0]) == b1
    hdr = gev_gadget2_header.from_buffer_copy(raw[4:260])
    assert hdr.num_files == 1 and hdr.BoxSize > 0 and hdr.npart[1] > 0
    off = 264
    blk = int(np.frombuffer(raw, np.int32, 1, off)[0])
    n0, n1 = hdr.npart[0], hdr.npart[1]
    assert blk == 12 * sum(hdr.npart)
    pos = np.frombuffer(raw, np.float32, 3 * n1, off + 4 + 12 * n0).reshape(n1, 3) / hdr.BoxSize
    assert int(np.frombuffer(raw, np.int32, 1, off + 4 + blk)[0]) == blk
    assert len(raw) == off + 8 + blk
    assert (pos >= 0).all() and (pos <= 1).all()
    return hdr, pos


@pytest.mark.parametrize("name,n", [("sc0_crystal.dat", 64), ("sc1_crystal.dat", 64), ("bcc_crystal.dat", 16),
                                    ("fcc_crystal.dat", 32)])
def test_template_headers_decode(name, n):
    hdr, pos = read_template(name)
    assert list(hdr.npart) == [0, n, 0, 0, 0, 0]
    assert hdr.BoxSize == 1.0
    assert len({tuple(p) for p in pos.tolist()}) == n  # distinct sites


def test_sc1_tiled_is_the_cell_centred_lattice():
    """initializeParticlePositions (ic_basic.hpp:1612-1644) with sc1 and numtile
    = N/4 puts one particle at every cell centre (x + 1/2)/N, which is the
    lattice gev_pcls_init_lattice builds (positions as a set; the reference's ID
    is i + 64 (xtile + numtile (ytile + numtile ztile)), the library's the cell
    index)"""
    _, t = read_template("sc1_crystal.dat")
    assert sorted(set(t.ravel().tolist())) == [0.125, 0.375, 0.625, 0.875]
    N = 16
    numtile = N // 4
    tiles = np.stack(np.meshgrid(np.arange(numtile), np.arange(numtile), np.arange(numtile), indexing="ij"), -1)
    tiles = tiles[..., ::-1].reshape(-1, 3)  # (xtile, ytile, ztile), x fastest as the reference loops
    pos = ((tiles[:, None, :] + t[None, :, :].astype(np.float64)) / numtile).reshape(-1, 3)
    cells = np.floor(pos * N).astype(np.int64)
    assert np.array_equal(pos, (cells + 0.5) / N)  # exactly the cell centres
    key = cells[:, 0] + N * (cells[:, 1] + N * cells[:, 2])
    assert np.array_equal(np.sort(key), np.arange(N ** 3))
