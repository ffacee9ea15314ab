"""kc_find_components_noalloc through the C ABI: the allocation-free form of the labels and
shortcuts variants, on device CSR buffers the test owns, against kc_find_components on the same
graphs -- two calls in a row, the second on a smaller graph -- and its size limits."""
import numpy as np
import pytest

import components_amd as ca

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory of the test (the caller), from the HIP runtime libkacomp itself links."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes

            ca.lib()  # libkacomp first: its libamdhip64.so.7 is the one the test then binds
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes):
        import ctypes

        self.ptr = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 64)
        assert self.hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert self.hip().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return self

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)


def device_csr(g):
    off = DeviceBuffer(8 * (g.n + 1)).upload(np.asarray(g.offsets, np.uint64))
    tgt = DeviceBuffer(4 * len(g.targets)).upload(np.asarray(g.targets, np.uint32))
    lab = DeviceBuffer(4 * max(g.n, 1)).upload(np.full(max(g.n, 1), 0xFFFFFFFF, np.uint32))
    return off, tgt, lab


def run_noalloc(g, **kw):
    off, tgt, lab = device_csr(g)
    st = ca.find_components_noalloc(g.n, off.ptr.value, tgt.ptr.value, lab.ptr.value, **kw)
    return lab.download(np.uint32, g.n), st


@pytest.mark.parametrize("variant", ["labels", "shortcut"])
def test_noalloc_equals_find_components(variant):
    rng = np.random.default_rng(6)
    graphs = [ca.generate("rmat", n=1 << 16, m=1 << 19, seed=3),  # skewed: every degree bin, hub chunks
              ca.Graph.from_edges(3000, rng.integers(0, 3000, size=(2000, 2)))]  # smaller, many components
    for g in graphs:
        print(f"noalloc {variant}: n={g.n} m={g.m}", flush=True)
        ref, rst = ca.find_components(g, variant=variant)
        lab, st = run_noalloc(g, variant=variant)
        assert np.array_equal(lab, ref)
        # (the sweep count is not compared: in-place sweeps and pointer jumps read keys other
        # threads are updating, so how many sweeps reach the fixed point varies from run to run)
        assert st["components"] == rst["components"] and st["sweeps"] >= 1
        assert st["n"] == g.n and st["m"] == g.m
    # the same entry point with non-default bin thresholds
    g = graphs[0]
    ref, _ = ca.find_components(g, variant=variant, tmax=4, smax=64, wmax=256)
    lab, _ = run_noalloc(g, variant=variant, tmax=4, smax=64, wmax=256)
    assert np.array_equal(lab, ref)


def test_noalloc_limits_and_variants():
    g = ca.Graph.from_edges(4, [(0, 1), (2, 3)])
    off, tgt, lab = device_csr(g)
    with pytest.raises(ca.KcError) as e:
        ca.find_components_noalloc(ca.NOALLOC_MAX_VERTICES + 1, off.ptr.value, tgt.ptr.value, lab.ptr.value,
                                   variant="labels")
    assert e.value.code == ca.KC_ERR_INVALID
    with pytest.raises(ca.KcError) as e:
        ca.find_components_noalloc(g.n, off.ptr.value, tgt.ptr.value, lab.ptr.value, variant="exp")
    assert e.value.code == ca.KC_ERR_UNSUPPORTED
    # more edge slots than the scratch holds: refused from offsets[n] before any row is read
    big = DeviceBuffer(24).upload(np.array([0, ca.NOALLOC_MAX_EDGE_SLOTS + 2, ca.NOALLOC_MAX_EDGE_SLOTS + 2], np.uint64))
    with pytest.raises(ca.KcError) as e:
        ca.find_components_noalloc(2, big.ptr.value, tgt.ptr.value, lab.ptr.value, variant="labels")
    assert e.value.code == ca.KC_ERR_INVALID
    # still usable afterwards
    st = ca.find_components_noalloc(g.n, off.ptr.value, tgt.ptr.value, lab.ptr.value, variant="shortcut")
    assert list(lab.download(np.uint32, 4)) == [0, 0, 2, 2] and st["components"] == 2
