"""Fixed workload for profiler runs (rocprofv3 --kernel-trace --stats, or one
--pmc counter per run):

    python -m tiralib_amd.profile_run [--steps 5]

Runs a 1 GiB plain copy (the counter calibration reference: 1 GiB read,
1 GiB written) three times, then `steps` iterations of the 3D vote workload of
bench.py (128^3 tubes + sheets, sigma 2, w 13, N 20) and of the batched
eig3 + Gaussian launch. Prints one line per stage.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--skip-batch", action="store_true")
    a = ap.parse_args(argv)
    from tiralib_amd import _lib, api, synth
    from tiralib_amd._lib import DeviceArray, call
    _lib.load()
    nb = 1 << 30
    src, dst = DeviceArray(nb // 4), DeviceArray(nb // 4)
    for _ in range(3):
        call("tira_hip_copy_async", dst.ptr, src.ptr, nb)
    call("tira_hip_synchronize")
    print("copy16: 3 x 1 GiB", flush=True)
    shape = (128, 128, 128)
    field = DeviceArray.from_host(synth.tubes_sheets_volume(shape, seed=synth.SEED_CFG2))
    n = int(np.prod(shape))
    vote, ev = DeviceArray((n, 6)), DeviceArray((n, 3))
    for _ in range(a.steps):
        call("tira_tensorvote3_async", field.ptr, _lib.LAYOUT_MAT9, vote.ptr, _lib.LAYOUT_PACKED6, *shape, 2.0, 0.0,
             13, 1, 1, 1, 20, _lib.PLATE_AUTO)
        call("tira_eig3_async", vote.ptr, n, _lib.LAYOUT_PACKED6, ev.ptr, None, None, _lib.EIG_SCALE_INVARIANT, None)
    call("tira_hip_synchronize")
    print(f"vote3: {a.steps} steps of 128^3 sigma 2 w 13 N 20", flush=True)
    if not a.skip_batch:
        shapes = synth.batch_shapes(48, seed=synth.SEED_CFG3)
        vols = []
        for i, s in enumerate(shapes):
            d = DeviceArray(tuple(s) + (6,))
            call("tira_synth_eig_tensors_async", d.ptr, int(np.prod(s)), 2000 + i, 1e-3, 1e2, 0.10, 1e-7, 1e-4, None)
            vols.append(d)
        b = api.BatchGaussEig3(vols, api.gaussian_taps(11, 1.5), autotune=False)
        evs = [DeviceArray(tuple(s) + (3,)) for s in shapes]
        for _ in range(a.steps):
            b.run()
            for v, e, s in zip(vols, evs, shapes):
                call("tira_eig3_async", v.ptr, int(np.prod(s)), _lib.LAYOUT_PACKED6, e.ptr, None, None,
                     _lib.EIG_SCALE_INVARIANT, None)
        call("tira_hip_synchronize")
        b.close()
        print(f"batch: {a.steps} runs of 48 volumes ({sum(int(np.prod(s)) for s in shapes)} voxels)", flush=True)


if __name__ == "__main__":
    main()
