"""The allocation-free 2D vote, tira_tensorvote2_prebuilt_noalloc, against
the entry point it mirrors, tira_tensorvote2_prebuilt (tensorvote.h:2236-2305),
both called through the C ABI on the same device fields: bit-identical output
(the same kernels) on two calls in a row, the second with a smaller field so
that the shared scratch array still holds the first call's larger payload,
with the neighbour table in host and in device memory. Requests larger than
the scratch array are refused before anything is launched."""
import numpy as np
import pytest

from tiralib_amd import _lib, api, synth
from tiralib_amd._lib import DeviceArray, call


def _mixed_field(shape, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, np.pi, shape)
    q = np.stack([np.cos(th), np.sin(th)], -1)
    t = np.einsum("hwi,hwj->hwij", q, q) * rng.uniform(0.2, 1.0, shape + (1, 1))
    t += 0.3 * np.eye(2)[None, None] * rng.uniform(0.0, 1.0, shape + (1, 1))
    return np.ascontiguousarray(t.reshape(shape + (4,)), np.float32)


CASES = [
    # the 256^2 Bezier field of the 2D benchmark, stick only
    ("bezier 256x256", lambda: synth.bezier_field(seed=synth.SEED_CFG1), 3.0, 0.0, 1, 19, 1, 0, 20),
    # a smaller mixed field, stick + plate, power 2
    ("mixed 96x80", lambda: _mixed_field((96, 80), 5), 2.0, 1.0, 2, 13, 1, 1, 10),
]


@pytest.mark.gpu
@pytest.mark.parametrize("table_on_device", [False, True])
def test_noalloc_equals_prebuilt(hip, table_on_device):
    for name, make, s1, s2, p, w, stick, plate, samples in CASES:
        print(f"noalloc vs prebuilt: {name}, table on device {table_on_device}", flush=True)
        f = make()
        H, W = f.shape[:2]
        nb = api.build_neighbors2d(w, s1, s2, p, samples, stick, plate)
        d_in = DeviceArray.from_host(f)
        o_ref, o_na = DeviceArray(f.shape), DeviceArray(f.shape)
        d_nb = DeviceArray.from_host(nb.view(np.uint8)) if table_on_device else None
        try:
            o_na.zero()
            call("tira_tensorvote2_prebuilt", o_ref.ptr, d_in.ptr, nb.ctypes.data, len(nb), s1, s2, p, w, H, W,
                 stick, plate, samples)
            call("tira_tensorvote2_prebuilt_noalloc", o_na.ptr, d_in.ptr,
                 d_nb.ptr if table_on_device else nb.ctypes.data, len(nb), s1, s2, p, w, H, W, stick, plate, samples)
            ref, got = o_ref.to_host(), o_na.to_host()
        finally:
            for a in (d_in, o_ref, o_na, d_nb):
                if a is not None:
                    a.free()
        assert np.abs(ref).max() > 0
        np.testing.assert_array_equal(got, ref)


def test_noalloc_refuses_oversize():
    """2048 x 2048 pixels need 64 MiB of payload: refused with an argument
    error before any pointer is used or anything is launched (no GPU needed)."""
    nb = api.build_neighbors2d(19, 3.0, 0.0, 1, 20, True, False)
    buf = np.zeros(4, np.float32)
    with pytest.raises(_lib.TiraInvalidArgument, match="scratch"):
        call("tira_tensorvote2_prebuilt_noalloc", buf.ctypes.data, buf.ctypes.data, nb.ctypes.data, len(nb), 3.0, 0.0,
             1, 19, 2048, 2048, 1, 0, 20)
