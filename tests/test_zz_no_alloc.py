"""mcpm_energy_momentum_noalloc against mcpm_energy_momentum.

The allocation-free entry point takes caller-owned device velocity arrays
and writes its four sums to a caller-owned device array; its workgroup
partials live in the library's fixed device array.  The test is the
caller: it allocates and copies through the HIP runtime that libmcpm.so
itself loaded.  The result must equal the context entry point's bit for
bit, on two calls in a row, the second with fewer particles (so stale
partials of the first call would show)."""
import ctypes

import numpy as np
import pytest

import mcpm
import workloads as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    mcpm.load_library()  # the runtime libmcpm.so links; dlopen returns the loaded copy
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return hip


class _Dev:
    """One device buffer of n doubles (test plumbing)."""

    def __init__(self, hip, n, host=None):
        self.hip, self.n = hip, n
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(8 * n, 8)) == 0
        if host is not None:
            h = np.ascontiguousarray(host, dtype=np.float64)
            assert hip.hipMemcpy(self.ptr, h.ctypes.data, 8 * n, H2D) == 0

    def get(self):
        h = np.empty(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data, self.ptr, 8 * self.n, D2H) == 0
        return h

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(hip, u1, u2, u3, mass_fn):
    bufs = [_Dev(hip, u1.size, a) for a in (u1, u2, u3)]
    out = _Dev(hip, 4, np.full(4, np.nan))
    mcpm.energy_momentum_noalloc(*[b.ptr.value or 0 for b in bufs], u1.size, mass_fn,
                                 out.ptr.value, 0)
    assert hip.hipDeviceSynchronize() == 0
    got = out.get()
    for b in bufs + [out]:
        b.free()
    return got


def _reference(p, arrs):
    g = mcpm.Simulation(p)
    g.upload(*arrs)
    g.cell_update()
    _, u1, u2, u3 = g.download()
    e = g.energy_momentum()
    g.close()
    return (u1, u2, u3), e


def test_noalloc_energy_matches_the_context_entry_point_twice():
    hip = _hip()
    big_p, big_arrs, _ = W.heat_conduction(precision=mcpm.FP64, seed=301)
    small_p, small_arrs, _ = W.heat_conduction(precision=mcpm.FP64, seed=302, n_cells=37, ppc=11)
    for p, a in ((big_p, big_arrs), (small_p, small_arrs)):
        print(f"noalloc energy N={a[0].size}", flush=True)
        u, ref = _reference(p, a)
        got = _noalloc(hip, *u, p.mass * p.Fn)
        np.testing.assert_array_equal(got, ref)
        assert ref[0] > 0.0


def test_noalloc_energy_of_no_particles_and_bad_arguments():
    hip = _hip()
    out = _Dev(hip, 4, np.full(4, np.nan))
    mcpm.energy_momentum_noalloc(0, 0, 0, 0, 1.0, out.ptr.value, 0)
    assert hip.hipDeviceSynchronize() == 0
    np.testing.assert_array_equal(out.get(), np.zeros(4))
    with pytest.raises(mcpm.McpmError) as e:
        mcpm.energy_momentum_noalloc(0, 0, 0, -1, 1.0, out.ptr.value, 0)
    assert e.value.code == 1
    with pytest.raises(mcpm.McpmError) as e:
        mcpm.energy_momentum_noalloc(0, 0, 0, 10, 1.0, out.ptr.value, 0)
    assert e.value.code == 1
    out.free()
