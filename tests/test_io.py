"""The reference's file formats: grid files (init.c:316-436) and the PETSc
binary restart set (main.c:394-420, 799-847).

CPU tests call the library's host I/O directly (no GPU involved) and check it
against hand-written known answers and the numpy restatement in
oracle/vfs_io_oracle.py; the GPU tests run the restart files through device
arrays and a read grid through the metrics.
"""
import os
import struct
import sys

import numpy as np
import pytest

from cases import gpu_case, oracle_case
from conftest import ROOT, rel_err
import vfsgpu as V
from vfsgpu import workloads as W

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import vfs_io_oracle as O  # noqa: E402


def _node(coor, dims, k, j, i):
    IM, JM, KM = dims
    return coor.reshape(KM + 1, JM + 1, IM + 1, 3)[k, j, i]


def test_text_grid_known_answer(tmp_path):
    """IM=3, JM=2, KM=2: x = i + 10 j + 100 k, y = -x, z = x / 8, written by
    hand in the text layout (every x, every y, every z over k, j, i)."""
    IM, JM, KM = 3, 2, 2
    xs = [i + 10 * j + 100 * k for k in range(KM) for j in range(JM) for i in range(IM)]
    body = [str(v) for v in xs] + [str(-v) for v in xs] + ["%.17e" % (v / 8) for v in xs]
    path = tmp_path / "grid.dat"
    path.write_text("1\n%d %d %d\n" % (IM, JM, KM) + "\n".join(body) + "\n")
    dims, coor = V.read_grid(path)
    assert dims == (IM, JM, KM)
    for k in range(KM):
        for j in range(JM):
            for i in range(IM):
                x = i + 10 * j + 100 * k
                assert tuple(_node(coor, dims, k, j, i)) == (x, -x, x / 8)
    # points outside the node range stay zero
    c = coor.reshape(KM + 1, JM + 1, IM + 1, 3)
    assert not c[KM].any() and not c[:, JM].any() and not c[:, :, IM].any()
    # scaling: value / cl * L_dim, in that order (init.c:407)
    _, s = V.read_grid(path, cl=3.0, L_dim=0.7)
    assert np.array_equal(s.reshape(c.shape)[:KM, :JM, :IM], c[:KM, :JM, :IM] / 3.0 * 0.7)


def test_xyz_grid_known_answer(tmp_path):
    """xyz.dat: IM lines 'x _ _', JM lines '_ y _', KM lines '_ _ z'."""
    X, Y, Z = [0.0, 0.5, 1.25], [-1.0, 2.0], [3.0, 4.0, 8.5, 9.0]
    lines = ["%d %d %d" % (3, 2, 4)]
    lines += ["%r 7 7" % x for x in X] + ["7 %r 7" % y for y in Y] + ["7 7 %r" % z for z in Z]
    path = tmp_path / "xyz.dat"
    path.write_text("\n".join(lines) + "\n")
    dims, coor = V.read_grid(path, V.GRID_XYZ)
    assert dims == (3, 2, 4)
    for k, z in enumerate(Z):
        for j, y in enumerate(Y):
            for i, x in enumerate(X):
                assert tuple(_node(coor, dims, k, j, i)) == (x, y, z)


@pytest.mark.parametrize("fmt", [V.GRID_TEXT, V.GRID_BINARY])
def test_grid_formats_against_the_oracle(tmp_path, fmt):
    """The stretched, wavy-wall channel grid of the benchmark generator at
    12x10x8 cells, written in each format, read by the library and by the
    oracle restatement: identical coordinates, and the exact coordinates for
    %.17e text and binary."""
    cfg, spec = W.make_config("C2", (12, 10, 8))
    coor = W.grid(spec)
    IM, JM, KM = cfg["mx"] - 1, cfg["my"] - 1, cfg["mz"] - 1
    path = tmp_path / "grid.dat"
    (O.write_grid_text if fmt == V.GRID_TEXT else O.write_grid_binary)(path, IM, JM, KM, coor)
    dims, got = V.read_grid(path, fmt, cl=2.0, L_dim=0.5)
    odims, want = O.read_grid(path, fmt, cl=2.0, L_dim=0.5)
    assert dims == odims == (IM, JM, KM)
    assert np.array_equal(got, want)
    _, plain = V.read_grid(path, fmt)
    m = np.zeros((KM + 1, JM + 1, IM + 1, 3), bool)
    m[:KM, :JM, :IM] = True
    assert np.array_equal(plain[m.reshape(-1)], coor[m.reshape(-1)])


def test_grid_text_integer_fields_follow_fscanf_i(tmp_path):
    """The header is read with "%i": 0x.. is hexadecimal and a leading 0 octal."""
    n = 2 * 2 * 2
    path = tmp_path / "grid.dat"
    path.write_text("01\n0x2 2 02\n" + "\n".join(["1.5"] * (3 * n)) + "\n")
    dims, coor = V.read_grid(path)
    assert dims == O.read_grid(path)[0] == (2, 2, 2)


def test_grid_errors(tmp_path):
    p = tmp_path / "g.dat"
    p.write_text("2\n3 3 3\n")
    with pytest.raises(V.VfsError, match="multi-block"):
        V.read_grid(p)
    p.write_text("1\n3 3 3\n1.0 2.0\n")
    with pytest.raises(V.VfsError, match="truncated"):
        V.read_grid(p)
    with pytest.raises(V.VfsError, match="cannot open"):
        V.read_grid(tmp_path / "missing.dat")


def test_petsc_vec_bytes_and_round_trip(tmp_path):
    """Header big-endian (1211214, n), values big-endian float64, as the
    oracle writes; the library reads the oracle's file back bit for bit."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal(1000)
    a[:4] = [0.0, -0.0, np.inf, 1e-310]
    p = tmp_path / "v.dat"
    V.write_petsc_vec(p, a)
    raw = p.read_bytes()
    assert raw[:8] == struct.pack(">ii", 1211214, a.size)
    assert raw[8:] == a.astype(">f8").tobytes()
    q = tmp_path / "w.dat"
    O.write_petsc_vec(q, a[::-1])
    b = V.read_petsc_vec(q, a.size)
    assert b.tobytes() == a[::-1].tobytes()
    assert np.array_equal(O.read_petsc_vec(p), a)


def test_petsc_vec_errors(tmp_path):
    p = tmp_path / "v.dat"
    O.write_petsc_vec(p, np.arange(10.0))
    with pytest.raises(V.VfsError, match="length"):
        V.read_petsc_vec(p, 11)
    p.write_bytes(struct.pack(">ii", 1211213, 1) + b"\0" * 8)
    with pytest.raises(V.VfsError, match="not a PETSc"):
        V.read_petsc_vec(p, 1)
    p.write_bytes(struct.pack(">ii", 1211214, 4) + b"\0" * 8)
    with pytest.raises(V.VfsError, match="truncated"):
        V.read_petsc_vec(p, 4)


def test_restart_path():
    assert V.restart_path("out", "vfield", 123, 0) == O.restart_path("out", "vfield", 123, 0) == "out/vfield000123_0.dat"
    assert V.restart_path("/r", "nvfield", 1234567, 3) == "/r/nvfield1234567_3.dat"


@pytest.mark.gpu
def test_restart_round_trip_through_the_device(tmp_path):
    """vfs_restart_write from device arrays -> the four files hold the
    fields (read by the oracle) -> vfs_restart_read into fresh device arrays
    gives them back bit for bit (nvfield into nvert_o)."""
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2")
    ctx = gpu_case(cfg, spec, coor)
    p = W.random_scalar(cfg, 1.0, 61)
    nv = ctx.nvert().host()
    fields = {"vfield": ucont, "ufield": ucat, "pfield": p, "nvfield": nv}
    dev = {k: V.DeviceArray.from_host(v) for k, v in fields.items()}
    ctx.restart_write(tmp_path, 42, 0, dev["vfield"], dev["ufield"], dev["pfield"], dev["nvfield"])
    for stem, v in fields.items():
        assert np.array_equal(O.read_petsc_vec(O.restart_path(tmp_path, stem, 42, 0)), v), stem
    back = {k: (ctx.vec() if v.size == 3 * cfg["mx"] * cfg["my"] * cfg["mz"] else ctx.sca()) for k, v in fields.items()}
    ctx.restart_read(tmp_path, 42, 0, back["vfield"], back["ufield"], back["pfield"], back["nvfield"])
    for stem, v in fields.items():
        assert np.array_equal(back[stem].host(), v), stem
    ctx.close()


@pytest.mark.gpu
def test_grid_file_to_metrics(tmp_path):
    """A %.17e text grid.dat of the channel generator, read by vfs_read_grid,
    gives the metrics of the in-memory grid exactly and matches the oracle's
    FormMetrics to 1e-15."""
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2")
    IM, JM, KM = cfg["mx"] - 1, cfg["my"] - 1, cfg["mz"] - 1
    path = tmp_path / "grid.dat"
    O.write_grid_text(path, IM, JM, KM, coor)
    dims, read = V.read_grid(path)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(read))
    ref = gpu_case(cfg, spec, coor)
    for n in ("csi", "eta", "zet", "aj", "icsi", "jeta", "kzet", "iaj", "cent"):
        g = ctx.metric(n).host()
        assert np.array_equal(g, ref.metric(n).host()), n
        assert rel_err(g, o.g[n]) <= 1e-15, n
    ctx.close()
    ref.close()
