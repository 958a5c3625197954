"""vfs_ib_search_noalloc against vfs_ib_search through the C-ABI: the same
node list, bit for bit, on two calls in a row where the second has a
smaller input (a coarser body on a smaller grid), so the second call runs
on scratch the first one left behind.  Also the size limit of the static
scratch (8 MiB) and the list-capacity error."""
import sys

import numpy as np
import pytest

from vfsgpu import workloads as W
import vfsgpu as V

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(N, level):
    cfg, spec = W.make_config("C2", N)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(W.grid(spec)))
    x, e, n = W.icosphere(*W.SPHERE, level=level)
    surf = V.IbSurface(x, e, n)
    ctx.nvert_body(surf)
    return ctx, surf


def _nodes(d, n):
    return d.host()[:n * V.IB_NODE_DTYPE.itemsize].copy()


def test_noalloc_search_matches_the_allocating_one():
    for N, level in (((64, 40, 32), 4), ((32, 24, 20), 2)):
        say(f"[no-alloc] C2 N={N} icosphere level {level}")
        ctx, surf = _case(N, level)
        ref, n_ref = ctx.ib_search(surf)
        assert n_ref > 0
        cap = n_ref + 16
        got = V.DeviceArray(cap * V.IB_NODE_DTYPE.itemsize, np.uint8)
        n = ctx.ib_search_noalloc(surf, got, cap)
        assert n == n_ref
        a, b = _nodes(got, n), _nodes(ref, n_ref)
        assert np.array_equal(a, b), "node records differ"
        # a list one entry short is refused with the count filled in
        with pytest.raises(V.VfsError) as e:
            ctx.ib_search_noalloc(surf, got, n_ref - 1)
        assert e.value.rc == -4


def test_noalloc_search_rejects_jobs_beyond_the_static_scratch():
    say("[no-alloc] scratch limit")
    ctx, _ = _case((32, 24, 20), 1)
    # 300000 elements need 9.6 MB of bounding spheres: more than the 8 MiB
    m = 300000
    x = np.random.default_rng(1).uniform(0, 1, (3 * m, 3))
    e = np.arange(3 * m, dtype=np.int32).reshape(m, 3)
    nrm = np.tile([0.0, 0.0, 1.0], (m, 1))
    big = V.IbSurface(x, e, nrm)
    out = V.DeviceArray(16 * V.IB_NODE_DTYPE.itemsize, np.uint8)
    with pytest.raises(V.VfsError) as err:
        ctx.ib_search_noalloc(big, out, 16)
    assert err.value.rc == -5 and "scratch" in str(err.value)
