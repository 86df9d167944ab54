"""pac_tc_noalloc -- the triangle count with its temporaries in a static device array
(no allocation, no host-device copy) -- against pac_tc on the same graphs, called through
the C ABI.  The caller's device buffers come from hipMalloc of the HIP runtime libpac is
linked against, as a caller that cannot allocate during a solve would set them up
beforehand."""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

import partitionandconquer_amd as pac

pytestmark = pytest.mark.gpu

D2H = 2  # hipMemcpyDeviceToHost


@pytest.fixture(scope="module")
def hip():
    pac.lib()  # loads libamdhip64 with libpac; the soname below resolves to that copy
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class DevOut:
    """A caller-owned device buffer for T (1 uint64) and t(u) (V uint64)."""

    def __init__(self, hip, V):
        self.hip, self.V = hip, V
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * (V + 1)) == 0

    def total_ptr(self):
        return self.p

    def pv_ptr(self):
        return C.c_void_p(self.p.value + 8)

    def read(self):
        h = np.zeros(self.V + 1, np.uint64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(h.ctypes.data, self.p, 8 * (self.V + 1), D2H) == 0
        return int(h[0]), h[1:]

    def free(self):
        self.hip.hipFree(self.p)


def noalloc(g, dispatch, out, per_vertex):
    rc = pac.lib().pac_tc_noalloc(g._h, dispatch, out.total_ptr(),
                                  out.pv_ptr() if per_vertex else None)
    if rc:
        raise pac.PacError(rc, pac.lib().pac_last_error().decode())


DISPATCHES = [pac.DISPATCH_FIXED, pac.FORCE_BITMAP, pac.FORCE_MERGE, pac.FORCE_GALLOP]


def test_noalloc_matches_pac_tc_on_two_calls(ctx, hip):
    """A larger graph, then a smaller one in the same buffers: T and t(u) from the
    no-allocation entry point equal pac_tc's for every dispatch, with and without per-vertex
    counts (the second call must not see the first call's queue, descriptor or counts)."""
    big = pac.Graph.rmat(ctx, 16, 16, 0.57, 0.19, 0.19, seed=16)
    big.prepare(64)
    small = pac.Graph.rmat(ctx, 11, 16, 0.57, 0.19, 0.19, seed=11)
    small.prepare(8)
    out = DevOut(hip, big.get_vertex_num())
    try:
        for d in DISPATCHES:
            for g in (big, small):
                T, pv = g.tc(d, pac.LAUNCH_PERSISTENT, per_vertex=True)
                noalloc(g, d, out, per_vertex=True)
                t, got = out.read()
                assert t == T and np.array_equal(got[:g.get_vertex_num()], pv), (d, g)
                noalloc(g, d, out, per_vertex=False)
                assert out.read()[0] == T, (d, g)
    finally:
        out.free()


def test_noalloc_deferred_rows_and_learned_state(ctx, hip):
    """Rows above the tile's LDS list (K_1100 on the 1024-entry tile) go through the
    deferred-row queue in the static array; the learned dispatch is refused until a learned
    solve has computed its choice; a null total is an argument error."""
    n = 1100
    u, v = map(np.array, zip(*itertools.combinations(range(n), 2)))
    g = pac.Graph.from_edges(ctx, u.astype(np.int32), v.astype(np.int32), n)
    out = DevOut(hip, n)
    tile = ctx.get_tile()
    try:
        ctx.set_tile(0)
        g.prepare(4)
        noalloc(g, pac.FORCE_BITMAP, out, per_vertex=True)
        t, pv = out.read()
        assert t == math.comb(n, 3) and np.all(pv == math.comb(n - 1, 2))
        ctx.set_tile(tile)
        ctx.mlp_init(7)
        with pytest.raises(pac.PacError) as e:
            noalloc(g, pac.DISPATCH_LEARNED, out, per_vertex=False)
        assert e.value.code == -6
        T = g.tc(pac.DISPATCH_LEARNED, pac.LAUNCH_PERSISTENT)
        noalloc(g, pac.DISPATCH_LEARNED, out, per_vertex=False)
        assert out.read()[0] == T == math.comb(n, 3)
        rc = pac.lib().pac_tc_noalloc(g._h, pac.DISPATCH_FIXED, None, None)
        assert rc == -1
    finally:
        ctx.set_tile(tile)
        out.free()
