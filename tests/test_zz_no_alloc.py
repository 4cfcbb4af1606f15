"""GPU test of lhip_direct_forces_noalloc, the allocation-free variant of the direct sum (its partial
sums in the library's per-device scratch array): against lhip_compute_forces with method direct on
the same bodies, bit for bit, on calls in a row, the later ones with smaller inputs. The caller's
device arrays are allocated here through the HIP runtime the library itself links (the entry point
allocates nothing)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import lettuce_simulate_physics_amd as L
import oracle as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Runtime:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime liblettuce_hip.so is linked against."""

    def __init__(self):
        L.library()  # loads liblettuce_hip.so and with it its HIP runtime
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.held = []

    def malloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.held.append(p.value)
        return p.value

    def to_device(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        d = self.malloc(a.nbytes)
        assert self.rt.hipMemcpy(d, a.ctypes.data, a.nbytes, H2D) == 0
        return d

    def to_host(self, d: int, n: int) -> np.ndarray:
        out = np.empty(n, np.float32)
        assert self.rt.hipMemcpy(out.ctypes.data, d, out.nbytes, D2H) == 0  # synchronises
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free_all(self):
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


@pytest.fixture
def rt(hip):
    r = Runtime()
    yield r
    r.free_all()


def test_direct_noalloc_equals_compute_forces_on_calls_in_a_row(hip, rt):
    # 16384 bodies split the source range 32 ways (8 MiB of partial sums), then 3000 bodies split 11
    # ways and 1000 bodies (eps = 0, the guarded pairs) 3 ways into the same scratch: a later call
    # must not see an earlier call's partials.
    for n, seed, potential, eps in ((16384, 1, True, 0.01), (3000, 5, False, 0.05), (1000, 6, True, 0.0)):
        print(f"noalloc direct: n={n} potential={potential} eps={eps}", flush=True)
        b = O.plummer(n, seed)
        p = L.ForceParams(G=1.0, eps=eps, method="direct")
        hip.upload(b)
        hip.compute_forces(p, potential)
        ax, ay, az, pot = hip.accelerations()
        o = hip.download()  # the backend's order, which the accelerations follow
        posm = rt.to_device(np.stack([o.x, o.y, o.z, o.m], 1).astype(np.float32))
        nan = np.full(n, np.nan, np.float32)
        out = [rt.to_device(nan) for _ in range(4)]
        L.direct_forces_noalloc(posm, n, 1.0, eps, potential, out[0], out[1], out[2], out[3] if potential else None)
        rt.sync()
        got = [rt.to_host(d, n) for d in out]
        assert np.array_equal(got[0], ax) and np.array_equal(got[1], ay) and np.array_equal(got[2], az), n
        if potential:
            assert np.array_equal(got[3], pot), n
        else:
            assert np.all(np.isnan(got[3]))


def test_direct_noalloc_argument_errors(hip, rt):
    d = rt.malloc(64 * 16)
    with pytest.raises(ValueError):
        L.direct_forces_noalloc(d, 1 << 31, 1.0, 0.01, False, d, d, d)
    with pytest.raises(ValueError):  # the potential without its output
        L.direct_forces_noalloc(d, 16, 1.0, 0.01, True, d, d, d)
    with pytest.raises(ValueError):  # posm not 16-byte aligned
        L.direct_forces_noalloc(d + 4, 16, 1.0, 0.01, False, d, d, d)
    L.direct_forces_noalloc(d, 0, 1.0, 0.01, False, d, d, d)  # nothing to do
