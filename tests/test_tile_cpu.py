"""The neighbour-list tile kernel's code run on the CPU (tests/tile_harness.hip
calls the workgroup-level functions of simox_amd/csrc/tile.hpp cell by cell,
thread by thread, with every LDS and global index checked, under the host
address sanitizer).

Checks, against the oracle's FORCE lists (oracle/simox_oracle.c, pinned by
the reference build) and against a numpy restatement of FORCE's loop order
(force.f90:43-48,146-155 over cellsort.f90:18-30):
* each atom's list holds the reference's neighbours, duplicates of boxes
  with two cells along an axis (quirk Q8) included;
* the first cnt_fwd entries are exactly the pairs in which the atom is
  FORCE's "i" (own cell with j after i, then the 13 stencil cells), the rest
  those in which it is "j";
* lists with more than cap entries report the overflow;
* cells fuller than one LDS chunk (several chunks per workgroup) give the same
  lists."""
import os
import subprocess

import numpy as np
import pytest

from simox_amd.systems import config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def build_harness(tmp_path_factory, cap=None, nt=None):
    out = str(tmp_path_factory.mktemp("tile") / "tile_harness")
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-g", "-Xarch_host",
           "-fsanitize=address", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           os.path.join(ROOT, "tests", "tile_harness.hip"), "-o", out]
    if cap:
        cmd[1:1] = ["-DHARNESS_CAP=%d" % cap, "-DHARNESS_NT=%d" % nt]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return build_harness(tmp_path_factory)


@pytest.fixture(scope="module")
def harness_small(tmp_path_factory):
    """The 512-atom, 256-thread cell tile (one of the autotuner's shapes)."""
    return build_harness(tmp_path_factory, 512, 256)


def geometry(box, rcut):
    nc = np.array([int(b / rcut) for b in box], dtype=np.int32)
    return nc


def cell_of(x, box, nc):
    """CELL_SORT's cell id (cellsort.f90:69-72), 0-based, x fastest."""
    lc = box / nc
    idx = ((x + 0.5 * box) / lc).astype(np.int64)
    idx = np.mod(idx + nc, nc)
    return idx[:, 0] + nc[0] * (idx[:, 1] + nc[1] * idx[:, 2])


def sort_by_cell(x, box, nc):
    c = cell_of(x, box, nc)
    order = np.argsort(c, kind="stable")  # within a cell ascending index: link order
    ncall = int(np.prod(nc))
    start = np.zeros(ncall + 1, dtype=np.int32)
    start[1:] = np.cumsum(np.bincount(c, minlength=ncall))
    return order, start


def run(harness, tmp_path, x, box, rcut, cap=500):
    box = np.asarray(box, dtype=np.float64)
    nc = geometry(box, rcut)
    order, start = sort_by_cell(x, box, nc)
    xs = np.ascontiguousarray(x[order], dtype=np.float64)
    n = len(x)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([n, *nc, cap], dtype=np.int32).tofile(f)
        np.array([*box, rcut], dtype=np.float64).tofile(f)
        xs.tofile(f)
        start.tofile(f)
    subprocess.run([harness, fin, fout], check=True)
    raw = np.fromfile(fout, dtype=np.int32)
    err = int(raw[0])
    cnt = raw[1:1 + n]
    fwd = raw[1 + n:1 + 2 * n]
    col = raw[1 + 2 * n:].reshape(cap, n)
    return dict(err=err, cnt=cnt, fwd=fwd, col=col, order=order, start=start, nc=nc, xs=xs)


def min_image_r2(xi, xj, box):
    d = xi - xj
    s = np.where(np.abs(d) < 0.25 * box, 0.0, np.trunc(d / box + np.copysign(0.5, d / box)))
    r = d - box * s
    return (r * r).sum(-1)


STENCIL = [(-1, -1, -1), (0, -1, -1), (1, -1, -1), (-1, 0, -1), (0, 0, -1), (1, 0, -1),
           (-1, 1, -1), (0, 1, -1), (1, 1, -1), (-1, -1, 0), (0, -1, 0), (1, -1, 0), (-1, 0, 0)]


def expected_lists(r, box, rcut):
    """FORCE's pairs in sorted indices: fwd[i] = multiset of j with i as "i",
    bwd[i] = those with i as "j"."""
    xs, start, nc = r["xs"], r["start"], r["nc"]
    n = len(xs)
    fwd = [[] for _ in range(n)]
    bwd = [[] for _ in range(n)]
    rc2 = rcut * rcut
    ncall = int(np.prod(nc))
    for m in range(ncall):
        a0, a1 = start[m], start[m + 1]
        if a1 == a0:
            continue
        cx, cy, cz = m % nc[0], (m // nc[0]) % nc[1], m // (nc[0] * nc[1])
        ia = np.arange(a0, a1)
        r2 = min_image_r2(xs[ia][:, None, :], xs[ia][None, :, :], box)
        for p, q in zip(*np.nonzero((r2 < rc2) & (ia[:, None] < ia[None, :]))):
            fwd[ia[p]].append(ia[q])
            bwd[ia[q]].append(ia[p])
        for dx, dy, dz in STENCIL:
            c = ((cx + dx) % nc[0]) + nc[0] * (((cy + dy) % nc[1]) + nc[1] * ((cz + dz) % nc[2]))
            jb = np.arange(start[c], start[c + 1])
            if len(jb) == 0:
                continue
            r2 = min_image_r2(xs[ia][:, None, :], xs[jb][None, :, :], box)
            for p, q in zip(*np.nonzero(r2 < rc2)):
                fwd[ia[p]].append(jb[q])
                bwd[jb[q]].append(ia[p])
    return fwd, bwd


def check_orientation(r, box, rcut):
    fwd, bwd = expected_lists(r, box, rcut)
    cnt, f, col = r["cnt"], r["fwd"], r["col"]
    for i in range(len(cnt)):
        got_f = np.sort(col[:f[i], i])
        got_b = np.sort(col[f[i]:cnt[i], i])
        assert np.array_equal(got_f, np.sort(np.array(fwd[i], dtype=np.int64))), i
        assert np.array_equal(got_b, np.sort(np.array(bwd[i], dtype=np.int64))), i


def check_against_oracle(r, oracle_mod, ids, x, q, box, fft, rcut):
    o = oracle_mod.Oracle(len(x), box, fft, rcut)
    o.cell_sort(x)
    ref = o.force(ids, x, q, want_lists=True)
    order = r["order"]
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    assert np.array_equal(r["cnt"], ref["nlist"][order])
    for p in range(len(order)):
        i = order[p]
        mine = np.sort(order[r["col"][:r["cnt"][p], p]])
        assert np.array_equal(mine, np.sort(ref["list"][i, :ref["nlist"][i]])), p


@pytest.mark.parametrize("case", ["tiny_s1", "tiny_s2", "c1_s1", "alloy_s1"])
def test_tile_lists_match_reference_order(harness, tmp_path, oracle_mod, case):
    d = np.load(os.path.join(GOLDEN, case + ".npz"))
    r = run(harness, tmp_path, d["x"], d["box"], float(d["rcut"]))
    assert r["err"] == 0
    check_against_oracle(r, oracle_mod, d["ids"], d["x"], d["q"], d["box"], d["fft"],
                         float(d["rcut"]))
    check_orientation(r, d["box"], float(d["rcut"]))


def test_tile_lists_c2(harness, tmp_path):
    s = config("c2")
    r = run(harness, tmp_path, s.x, s.box, s.rcut)
    assert r["err"] == 0
    check_orientation(r, s.box, s.rcut)


def test_tile_several_chunks(harness, tmp_path):
    # 3 x 3 x 3 cells of ~100 atoms: the 28 visits stage ~2,800 atoms, more
    # than one LDS chunk holds, so each workgroup runs several chunks; three
    # cells per axis take the FP32 periodic fold
    rng = np.random.default_rng(4)
    box = np.array([18.1, 18.2, 18.3])
    x = (rng.random((2700, 3)) - 0.5) * box
    r = run(harness, tmp_path, x, box, 6.0)
    assert r["err"] == 0
    check_orientation(r, box, 6.0)


def test_tile_overflow_reported(harness, tmp_path):
    rng = np.random.default_rng(5)
    box = np.array([20.0, 20.0, 20.0])
    x = (rng.random((4000, 3)) - 0.5) * box  # ~ 4000 * (4/3 pi 216) / 8000 = 450+ neighbours
    r = run(harness, tmp_path, x, box, 6.5, cap=300)
    assert r["err"] == 2  # SIMOX_ERR_NB_OVERFLOW
    assert r["cnt"].max() == 300


@pytest.mark.parametrize("case", ["tiny_s2", "c1_s1", "alloy_s1"])
def test_tile_shapes_give_the_same_lists(harness, harness_small, tmp_path, case):
    d = np.load(os.path.join(GOLDEN, case + ".npz"))
    r0 = run(harness, tmp_path, d["x"], d["box"], float(d["rcut"]))
    r1 = run(harness_small, tmp_path, d["x"], d["box"], float(d["rcut"]))
    assert np.array_equal(r0["cnt"], r1["cnt"]) and np.array_equal(r0["fwd"], r1["fwd"])
    assert np.array_equal(r0["col"], r1["col"])  # same entries in the same order
