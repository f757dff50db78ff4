"""Gadget format-1/2 reader (readgadget.py / readsnap.py call surface) on snapshots written by
tests/_gadget.py: both byte orders, single and multi-file, mass table and MASS block."""
import math

import numpy as np
import pytest

from pylians_amd import readgadget
from tests._gadget import write_snapshot


def _data(rng):
    n = [5, 7, 4, 0, 3, 0]
    pos = [rng.random((k, 3), dtype=np.float32) * 1000 for k in n]
    vel = [rng.standard_normal((k, 3)).astype(np.float32) * 100 for k in n]
    ids = [np.arange(k, dtype=np.uint32) + 100 * t for t, k in enumerate(n)]
    masses = [rng.random(k, dtype=np.float32) for k in n]
    massarr = [0.0, 0.5, 2.0, 0.0, 0.0, 0.0]
    return n, pos, vel, ids, masses, massarr


@pytest.mark.parametrize("fmt,endian,nfiles", [(1, "<", 1), (2, "<", 1), (1, ">", 2), (2, ">", 3)])
def test_reader_roundtrip(tmp_path, fmt, endian, nfiles):
    rng = np.random.default_rng(fmt * 10 + nfiles)
    n, pos, vel, ids, masses, massarr = _data(rng)
    split = [np.array(n) // nfiles for _ in range(nfiles)]
    split[-1] = np.array(n) - sum(split[:-1])
    base = str(tmp_path / "snap")
    write_snapshot(base, pos, vel, ids, masses, massarr, split, 0.5, 1.0, 1000.0, fmt=fmt, endian=endian)
    h = readgadget.header(base)
    assert h.format == fmt and h.filenum == nfiles and h.boxsize == 1000.0 and h.redshift == 1.0
    assert list(h.nall) == n and np.allclose(h.massarr, massarr)
    assert np.isclose(h.Hubble, 100 * np.sqrt(0.3 * 8 + 0.7))
    for t in (1, 2, 4):
        assert np.array_equal(readgadget.read_block(base, "POS ", [t]), pos[t])
        assert np.array_equal(readgadget.read_block(base, "ID  ", [t]), ids[t])
        v = readgadget.read_block(base, "VEL ", [t])
        assert np.array_equal(v, vel[t] * np.float32(math.sqrt(0.5)))   # peculiar velocities at z != 0
    assert np.array_equal(readgadget.read_block(base, "POS ", [1, 2]), np.concatenate([pos[1], pos[2]]))
    assert np.array_equal(readgadget.read_block(base, "MASS", [1]), np.full(7, 0.5, np.float32))
    assert np.array_equal(readgadget.read_block(base, "MASS", [0]), masses[0])
    m = readgadget.read_block(base, "MASS", [-1])
    assert np.array_equal(m, np.concatenate([masses[0], np.full(7, 0.5, np.float32), np.full(4, 2.0, np.float32),
                                             masses[4]]))


def test_missing_file_and_hdf5():
    with pytest.raises(Exception):
        readgadget.header("/nonexistent/snap")


def test_read_field_per_file(tmp_path):
    """read_field (readgadget.py:63-97) reads one file: the per-file pieces concatenate to read_block;
    a base name resolves to its first file, as fname_format does."""
    rng = np.random.default_rng(5)
    n, pos, vel, ids, masses, massarr = _data(rng)
    split = [np.array(n) // 2, np.array(n) - np.array(n) // 2]
    base = str(tmp_path / "snap")
    write_snapshot(base, pos, vel, ids, masses, massarr, split, 0.25, 3.0, 1000.0, fmt=2)
    for t in (0, 1, 2, 4):
        for block in ("POS ", "VEL ", "MASS", "ID  "):
            parts = [readgadget.read_field("%s.%d" % (base, i), block, t) for i in range(2)]
            assert np.array_equal(np.concatenate(parts), readgadget.read_block(base, block, [t])), (t, block)
        assert np.array_equal(readgadget.read_field(base, "POS ", t), pos[t][:split[0][t]])
    with pytest.raises(Exception):
        readgadget.read_field(base, "RHO ", 0)
