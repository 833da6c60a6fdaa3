"""The cell tile of the neighbour-list build (k_nlist_tile): every shape the
autotuner may pick (atoms staged per LDS chunk, threads per workgroup) gives
the same list, entry for entry, and so the same forces bit for bit; the
tuned shape is one of them.  C2's slab (77,824 atoms, 9 x 9 x 14 cells,
about 70 atoms per cell: one to four chunks per workgroup)."""
import os

import numpy as np
import pytest

from simox_amd import MODE_MIXED, Simox
from simox_amd.systems import config

pytestmark = pytest.mark.gpu

SHAPES = [(1024, 128), (512, 128), (2048, 128), (1024, 64), (1024, 256), (512, 256)]


def with_env(name, value, fn):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def test_tile_shapes_give_the_same_lists_and_forces():
    s = config("c2")
    ref = None
    for k, shape in enumerate(SHAPES):
        print("tile shape", shape, flush=True)
        g = with_env("SIMOX_TILE_SHAPE", str(k), lambda: Simox.from_system(s, mode=MODE_MIXED))
        e = g.force()
        f = g.get_state()[2]
        cnt, cols = g.neighbors()
        assert g.tile_shape() == shape
        g.close()
        if ref is None:
            ref = (cnt, cols, f, e)
            continue
        assert np.array_equal(cnt, ref[0])
        assert np.array_equal(cols, ref[1])
        assert np.array_equal(f, ref[2])
        assert e == ref[3]


def test_tuned_tile_shape_is_a_listed_one():
    s = config("c2")
    g = Simox.from_system(s, mode=MODE_MIXED)
    g.force()
    assert g.tile_shape() in SHAPES
    g.close()
