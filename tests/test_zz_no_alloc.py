"""edus_lindhard_device_noalloc (the entry point that allocates nothing: its
scratch is a fixed per-GPU array) against edus_lindhard_device (workspace
argument), through the C ABI: two calls in a row, the second on a smaller
input, each bit-identical to the workspace entry point (same kernels, same
arithmetic); a shape beyond the fixed scratch is refused with EDUS_ERR_ARG
before anything is launched."""
import sys

import numpy as np
import pytest

import edus
import screened_models as sm

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(nb, nf, N0, N1, seed):
    from edus import hip_rt

    nk = N0 * N1
    U, w = sm.random_bands(nb, N0, N1, seed=seed)
    dU = hip_rt.DeviceBuffer.from_array(U)
    de = hip_rt.DeviceBuffer.from_array(np.ascontiguousarray(w))
    work = hip_rt.DeviceBuffer(U.nbytes)
    dX = hip_rt.DeviceBuffer(U.nbytes)
    dY = hip_rt.DeviceBuffer(U.nbytes)
    edus.lindhard_device(nb, N0, N1, nf, dU.address, de.address, work.address, dX.address, 0)
    edus.lindhard_device_noalloc(nb, N0, N1, nf, dU.address, de.address, dY.address, 0)
    hip_rt.synchronize()
    ref = dX.to_array(np.complex128, (nb * nb, nk))
    got = dY.to_array(np.complex128, (nb * nb, nk))
    return ref, got


def test_noalloc_matches_workspace_entry_twice_second_smaller():
    for nb, nf, N0, N1, seed in ((14, 7, 24, 20, 5), (6, 2, 9, 8, 6)):
        _say("case nb", nb, "nf", nf, "grid", N0, N1)
        ref, got = _both(nb, nf, N0, N1, seed)
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref)


def test_noalloc_refuses_shapes_beyond_its_scratch():
    from edus import hip_rt

    tiny = hip_rt.DeviceBuffer(1024)  # never read: the size check comes first
    with pytest.raises(edus.EdusError, match="exceeds its fixed scratch") as e:
        edus.lindhard_device_noalloc(14, 480, 480, 7, tiny.address, tiny.address, tiny.address, 0)
    assert e.value.code == 1  # EDUS_ERR_ARG
