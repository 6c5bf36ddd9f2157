"""The allocation-free two-point entry point against the allocating one (GPU).

cvchip_field_contract_twopoint_noalloc keeps its block partial sums in a
file-scope device array and adds C(t) to caller-owned device memory; it runs
the same partial-sum kernel as cvchip_field_contract_twopoint and sums the
partials in the same order on the device, so the results agree bit for bit.
Two calls in a row (the second on a smaller input, after the first left its
partials in the shared array), accumulation into the output, and the
rejection of a lattice whose partials exceed the array.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BC = (1.0, 0.0, 0.0, 0.0)


class DeviceBuffer:
    """Device memory of n doubles from the HIP runtime the library itself uses
    (the caller's memory in the entry point's contract)."""

    def __init__(self, chip, n):
        chip.lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")   # the runtime libcvchip.so is linked against
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0
        self.zero()

    def zero(self):
        assert self.hip.hipMemset(self.ptr, 0, ctypes.c_size_t(8 * self.n)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def run_case(chip, dims, n_c, seed, gammas, mom, src):
    print(f"no-alloc two-point dims={dims} n_c={n_c} gammas={gammas} mom={mom}", flush=True)
    ctx = chip.Context(dims, 0.163, 0.01, BC)
    N = 4 * n_c
    chi, phi = ctx.field(N), ctx.field(N)
    chi.fill_random(seed)
    phi.fill_random(seed + 1)
    want = ctx.field_contract_twopoint(gammas[0], gammas[1], chi, phi, n_c, mom, src)
    out = DeviceBuffer(chip, 2 * dims[0])
    ctx.field_contract_twopoint_noalloc(out.ptr.value, gammas[0], gammas[1], chi, phi, n_c, mom, src)
    got = out.get()
    assert np.array_equal(got[0::2] + 1j * got[1::2], want)
    # a second call adds to the output
    ctx.field_contract_twopoint_noalloc(out.ptr.value, gammas[0], gammas[1], chi, phi, n_c, mom, src)
    got2 = out.get()
    out.free()
    assert np.array_equal(got2, 2 * got)


def test_noalloc_twopoint_matches_allocating_entry_point(chip):
    run_case(chip, (16, 16, 16, 16), 3, 7, (5, 5), (1, 0, 2), (3, 1, 4, 1))
    run_case(chip, (8, 4, 4, 6), 1, 11, (2, 7), None, (0, 0, 0, 0))


def test_noalloc_twopoint_rejects_oversized_lattice(chip):
    small = chip.Context((4, 4, 4, 4), 0.163, 0.01, BC)
    chi, phi = small.field(12), small.field(12)
    out = DeviceBuffer(chip, 512)
    big = chip.Context((256, 64, 64, 64), 0.163, 0.01, BC)   # 256 * 1024 blocks > 131072
    with pytest.raises(chip.CvchipError, match="scratch"):
        big.field_contract_twopoint_noalloc(out.ptr.value, 5, 5, chi, phi, 3)
    out.free()
    # the library stays usable
    chi.fill_random(1)
    phi.fill_random(2)
    ok = small.field_contract_twopoint(5, 5, chi, phi, 3)
    assert ok.shape == (4,) and np.all(np.isfinite(ok))
