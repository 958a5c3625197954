"""Test infrastructure only (the checker, never the product): a numpy
restatement of the reference's file formats on either side of the hot path.

- Grid files read by MG_Initial (init.c:316-436): text grid.dat (block count,
  "IM JM KM", then every x, every y, every z over k, j, i with i fastest,
  read with fscanf "%le"), binary grid.dat (the same as native int32 and
  float64, -binary 1) and xyz.dat (-xyz 1: "IM JM KM", then IM lines
  "x _ _", JM lines "_ y _", KM lines "_ _ z" of a tensor-product grid).
  Coordinates are value / cl * L_dim (init.c:406-433).
- PETSc binary vector files (VecView / VecLoadIntoVector of the restart set,
  main.c:394-420, 799-847): big-endian int32 class id 1211214, big-endian
  int32 length, big-endian float64 values in the DA's natural order.
"""
import numpy as np

PETSC_VEC_ID = 1211214


def _c_int(tok):
    """fscanf "%i": decimal, 0x hexadecimal or 0 octal."""
    t = tok.lower()
    sign = -1 if t.startswith("-") else 1
    t = t.lstrip("+-")
    if t.startswith("0x"):
        return sign * int(t[2:], 16)
    if len(t) > 1 and t.startswith("0"):
        return sign * int(t[1:], 8)
    return sign * int(t)


def read_grid(path, fmt=0, cl=1.0, L_dim=1.0):
    """((IM, JM, KM), coor) with node (k, j, i) at ((k*(JM+1)+j)*(IM+1)+i)."""
    if fmt == 1:
        raw = open(path, "rb").read()
        nb, IM, JM, KM = np.frombuffer(raw, dtype=np.int32, count=4)
        assert nb == 1
        vals = np.frombuffer(raw, dtype=np.float64, offset=16, count=3 * IM * JM * KM)
        xyz = vals.reshape(3, KM, JM, IM)
    else:
        tok = open(path).read().split()
        if fmt == 0:
            nb, IM, JM, KM = (_c_int(t) for t in tok[:4])
            assert nb == 1
            vals = np.array([float(t) for t in tok[4:4 + 3 * IM * JM * KM]])
            xyz = vals.reshape(3, KM, JM, IM)
        else:
            IM, JM, KM = (_c_int(t) for t in tok[:3])
            v = np.array([float(t) for t in tok[3:3 + 3 * (IM + JM + KM)]]).reshape(-1, 3)
            X, Y, Z = v[:IM, 0], v[IM:IM + JM, 1], v[IM + JM:, 2]
            xyz = np.stack(np.broadcast_arrays(X[None, None, :], Y[None, :, None], Z[:, None, None]))
    coor = np.zeros((KM + 1, JM + 1, IM + 1, 3))
    for q in range(3):
        coor[:KM, :JM, :IM, q] = xyz[q] / cl * L_dim
    return (int(IM), int(JM), int(KM)), coor.reshape(-1)


def write_grid_text(path, IM, JM, KM, coor, fmt="%.17e"):
    """Text grid.dat of the nodes of a padded coordinate array (a writer for
    the tests; the reference itself only reads this format)."""
    c = coor.reshape(KM + 1, JM + 1, IM + 1, 3)[:KM, :JM, :IM]
    with open(path, "w") as f:
        f.write("1\n%d %d %d\n" % (IM, JM, KM))
        for q in range(3):
            np.savetxt(f, c[..., q].reshape(-1), fmt=fmt)


def write_grid_binary(path, IM, JM, KM, coor):
    c = coor.reshape(KM + 1, JM + 1, IM + 1, 3)[:KM, :JM, :IM]
    with open(path, "wb") as f:
        f.write(np.array([1, IM, JM, KM], dtype=np.int32).tobytes())
        for q in range(3):
            f.write(np.ascontiguousarray(c[..., q]).astype(np.float64).tobytes())


def write_petsc_vec(path, a):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    with open(path, "wb") as f:
        f.write(np.array([PETSC_VEC_ID, a.size], dtype=">i4").tobytes())
        f.write(a.astype(">f8").tobytes())


def read_petsc_vec(path):
    raw = open(path, "rb").read()
    cid, n = np.frombuffer(raw, dtype=">i4", count=2)
    assert cid == PETSC_VEC_ID
    return np.frombuffer(raw, dtype=">f8", offset=8, count=n).astype(np.float64)


def restart_path(directory, stem, ti, block):
    """main.c:404, 804: "%s/vfield%06d_%1d.dat" and the like."""
    return "%s/%s%06d_%1d.dat" % (directory, stem, ti, block)
