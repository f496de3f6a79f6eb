"""gs3d_sort_pairs_noalloc, the allocation-free entry point: the sort of
gs3d_stage_sort over caller device buffers, with its scratch in a static
device array. Checked against gs3d_stage_sort (and the closed-form stable
order) bit for bit, on consecutive calls whose second input is smaller, so
that what the first call left in the shared scratch cannot leak into the
second result."""
import ctypes as C
import sys

import numpy as np
import pytest

from gs3d import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_keys(n, seed, tiles=8160, dup=0.3):
    rng = np.random.default_rng(seed)
    tile = rng.integers(0, tiles, n, dtype=np.uint64)
    depth = rng.uniform(0.2, 30.0, n).astype(np.float32).view(np.uint32).astype(np.uint64)
    keys = (tile << np.uint64(32)) | depth
    if n:
        idx = rng.random(n) < dup
        keys[idx] = keys[rng.integers(0, n, idx.sum())]
    return keys, rng.permutation(n).astype(np.uint32)


class DeviceBuffers:
    """Device memory for the test's inputs and outputs, from the HIP runtime
    libgs3d.so itself is linked against (a second HIP runtime, such as the one
    bundled with torch, cannot share the process with it)."""

    def __init__(self):
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(nbytes, 16)) == 0
        self.live.append(p.value)
        return p.value

    def upload(self, a):
        d = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return d

    def download(self, d, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        self.hip.hipDeviceSynchronize()
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


@pytest.fixture
def dev(renderer):
    d = DeviceBuffers()
    yield d
    d.close()


def noalloc_sort(renderer, dev, keys, vals, bits, in_place=False):
    dk, dv = dev.upload(keys), dev.upload(vals)
    ok, ov = (dk, dv) if in_place else (dev.alloc(keys.nbytes), dev.alloc(vals.nbytes))
    rc = L.lib().gs3d_sort_pairs_noalloc(renderer._h, dk, dv, keys.size, bits, ok, ov)
    assert rc == 0, L.lib().gs3d_last_error(renderer._h)
    renderer.sync()
    return dev.download(ok, keys), dev.download(ov, vals)


@pytest.mark.parametrize("bits", [8, 4])
def test_noalloc_sort_equals_stage_sort_on_consecutive_calls(renderer, dev, bits):
    for n, seed in ((3_400_000, 1), (100_003, 2), (1, 3)):
        say(f"no-alloc sort bits={bits} n={n}")
        keys, vals = make_keys(n, seed + bits)
        gk, gv = noalloc_sort(renderer, dev, keys, vals, bits)
        sk, sv = renderer.stage_sort(keys, vals, bits)
        assert np.array_equal(gk, sk) and np.array_equal(gv, sv)
        order = np.lexsort((np.arange(n), keys))
        assert np.array_equal(gk, keys[order]) and np.array_equal(gv, vals[order])


def test_noalloc_sort_in_place_and_at_the_limit(renderer, dev):
    n = 1 << 22
    say(f"no-alloc sort in place, n={n} (the maximum)")
    keys, vals = make_keys(n, 11)
    gk, gv = noalloc_sort(renderer, dev, keys, vals, 8, in_place=True)
    order = np.lexsort((np.arange(n), keys))
    assert np.array_equal(gk, keys[order]) and np.array_equal(gv, vals[order])


def test_noalloc_sort_rejects_oversize_and_bad_args(renderer, dev):
    lib = L.lib()
    t = dev.alloc(64)
    assert lib.gs3d_sort_pairs_noalloc(renderer._h, t, t, (1 << 22) + 1, 8, t, t) == -1
    assert b"static scratch" in lib.gs3d_last_error(renderer._h)
    assert lib.gs3d_sort_pairs_noalloc(renderer._h, None, None, 5, 8, None, None) == -1
    assert lib.gs3d_sort_pairs_noalloc(renderer._h, t, t, 4, 6, t, t) == -1
    assert lib.gs3d_sort_pairs_noalloc(renderer._h, None, None, 0, 8, None, None) == 0
    assert lib.gs3d_sort_pairs_noalloc(None, None, None, 0, 8, None, None) == -1
