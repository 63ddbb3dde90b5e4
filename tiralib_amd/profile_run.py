"""Fixed workload for profiler runs (rocprofv3 --kernel-trace --stats, or one
--pmc counter per run):

    python -m tiralib_amd.profile_run [--steps 5]

Runs a 1 GiB plain copy (the counter calibration reference: 1 GiB read,
1 GiB written) three times, then `steps` iterations of: the 3D vote + eig3
step on 128^3 tubes + sheets (sigma 2, w 13, N 20) with the spectral vote
(eigenvalues written by its inverse x pass) and with the direct gather, and
the analytical vote (ATV) of the same volume; the spectral vote of the 256^3 curved-sheet volume (sigma 4, w 25); the batched eig3 +
Gaussian launch and the one-launch eig3 of its 48 volumes; the learned plate
vote on a 96^3 curved-sheet volume (sigma 4, w 25; untrained weights with
the error check off: the kernel's work does not depend on the weights, and
training's ~10^5 small dispatches are left out of counter runs);
and the full pipeline (structure tensor, Gaussian sigma 2, vote sigma 4,
eig3 saliency) on a 192 x 256 x 256 vascular-tree volume, spectral then
gather. Prints one line per
stage.

    python -m tiralib_amd.profile_run --config4 [--steps 3]

runs instead the 1 GiB copy three times, the FP32 FMA loop once (the VALU
counter reference) and the legs of the benchmark's
256^3 curved-sheet workload (12
sheets, 2% noise, seed 1004; sigma 4, w 25): the learned plate as the
tabulated network and per pair (untrained weights, error check off: the
kernels' work does not depend on the weights), the literal 20- and
64-sample integrated plate, the closed form and the spectral vote, `steps`
calls each (the per-pair network and the 64-sample sum once).
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
    ap.add_argument("--skip-mlp", action="store_true")
    ap.add_argument("--skip-pipeline", action="store_true")
    ap.add_argument("--config4", action="store_true")
    a = ap.parse_args(argv)
    from tiralib_amd import _lib, api, synth
    from tiralib_amd._lib import DeviceArray, call
    _lib.load()
    if a.config4:
        return config4(a.steps)
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
    # the benchmark's step: the spectral vote whose inverse x pass also writes
    # the eigenvalues (one call), then the direct gather followed by eig3
    for mode in (_lib.PLATE_AUTO, _lib.PLATE_CLOSED):
        for _ in range(a.steps):
            call("tira_tensorvote3_eig_async", field.ptr, _lib.LAYOUT_MAT9, vote.ptr, _lib.LAYOUT_PACKED6, *shape,
                 2.0, 0.0, 13, 1, 1, 1, 20, mode, ev.ptr, _lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT, None)
        call("tira_hip_synchronize")
    print(f"vote3: {a.steps} steps of 128^3 sigma 2 w 13 N 20 (vote + eig3), spectral then gather", flush=True)
    if not a.skip_mlp:
        sshape = (256, 256, 256)
        svol = synth.curved_sheets_volume_device(sshape, 12, 0.02, seed=synth.SEED_CFG4)
        sout = DeviceArray((int(np.prod(sshape)), 6))
        for _ in range(a.steps):
            call("tira_tensorvote3_async", svol.ptr, _lib.LAYOUT_PACKED6, sout.ptr, _lib.LAYOUT_PACKED6, *sshape,
                 4.0, 0.0, 25, 1, 1, 1, 20, _lib.VOTE_SPECTRAL)
        call("tira_hip_synchronize")
        svol.free()
        sout.free()
        print(f"spectral: {a.steps} votes of the 256^3 curved-sheet volume, sigma 4, w 25", flush=True)
    for _ in range(a.steps):
        call("tira_tensorvote3_atv_async", field.ptr, _lib.LAYOUT_MAT9, vote.ptr, _lib.LAYOUT_PACKED6, *shape, 2.0,
             13)
    call("tira_hip_synchronize")
    print(f"atv3: {a.steps} analytical votes of 128^3 sigma 2 w 13", flush=True)
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
        counts = [int(np.prod(s)) for s in shapes]
        for _ in range(a.steps):
            api.eig3_batch(vols, counts=counts, evals=evs, flags=_lib.EIG_SCALE_INVARIANT)
        call("tira_hip_synchronize")
        b.close()
        print(f"batch: {a.steps} runs of 48 volumes ({sum(counts)} voxels)", flush=True)
    if not a.skip_mlp:
        mshape = (96, 96, 96)
        mvol = synth.curved_sheets_volume_device(mshape, 4, 0.02, seed=synth.SEED_CFG4)
        net = api.PlateMLP(4.0, 0.0, 1, 64, seed=synth.SEED_CFG4)
        mout = DeviceArray((int(np.prod(mshape)), 6))
        for _ in range(a.steps):
            net.vote(mvol, mout, mshape, 25, stick=True, tol=1e30, in_layout=_lib.LAYOUT_PACKED6)
        call("tira_hip_synchronize")
        net.close()
        print(f"mlp: {a.steps} learned-plate votes of 96^3, sigma 4, w 25", flush=True)
    if not a.skip_pipeline:
        pshape = (192, 256, 256)
        pvol, _ = synth.vessel_volume_device(pshape, seed=synth.SEED_CFG5, noise_sigma=0.1)
        sal = DeviceArray(pshape + (3,))
        for mode in (_lib.VOTE_SPECTRAL, _lib.PLATE_CLOSED):
            for _ in range(a.steps):
                api.pipeline3(pvol, smooth_sigma=2.0, vote_sigma=4.0, w=25, plate_mode=mode, saliency=sal)
            call("tira_hip_synchronize")
        print(f"pipeline: {a.steps} runs on 192 x 256 x 256", flush=True)


def config4(steps):
    from tiralib_amd import _lib, api, synth
    from tiralib_amd._lib import DeviceArray, call
    nb = 1 << 30
    src, dst = DeviceArray(nb // 4), DeviceArray(nb // 4)
    for _ in range(3):
        call("tira_hip_copy_async", dst.ptr, src.ptr, nb)
    call("tira_hip_synchronize")
    src.free()
    dst.free()
    print("copy16: 3 x 1 GiB", flush=True)
    # FP32 FMA loop (8 FMA per iteration and thread): the VALU counter reference
    fblocks, fiters = 256 * 64, 4096
    fout = DeviceArray(fblocks)
    call("tira_hip_fma_loop_async", fout.ptr, fblocks, fiters)
    call("tira_hip_synchronize")
    fout.free()
    print(f"fma loop: {fblocks} blocks x 256 threads x {fiters} iterations x 8 FMA", flush=True)
    shape, sigma, w = (256, 256, 256), 4.0, 25
    vol = synth.curved_sheets_volume_device(shape, 12, 0.02, seed=synth.SEED_CFG4)
    out = DeviceArray((int(np.prod(shape)), 6))
    net = api.PlateMLP(sigma, 0.0, 1, 64, seed=synth.SEED_CFG4)

    def trad(mode, samples):
        call("tira_tensorvote3_async", vol.ptr, _lib.LAYOUT_PACKED6, out.ptr, _lib.LAYOUT_PACKED6, *shape, sigma,
             0.0, w, 1, 1, 1, samples, mode)

    legs = [("tabulated learned plate", lambda: net.vote(vol, out, shape, w, stick=True, tol=1e30,
                                                         in_layout=_lib.LAYOUT_PACKED6), steps),
            ("integrated plate, 20 samples", lambda: trad(_lib.PLATE_SUM, 20), steps),
            ("closed form", lambda: trad(_lib.PLATE_CLOSED, 20), steps),
            ("spectral", lambda: trad(_lib.VOTE_SPECTRAL, 20), steps),
            ("integrated plate, 64 samples", lambda: trad(_lib.PLATE_SUM, 64), 1),
            ("learned plate per pair", lambda: net.vote(vol, out, shape, w, stick=True, tol=1e30,
                                                        in_layout=_lib.LAYOUT_PACKED6, method=_lib.MLP_PER_PAIR), 1)]
    for name, fn, reps in legs:
        for _ in range(reps):
            fn()
        call("tira_hip_synchronize")
        print(f"config 4: {reps} x {name}", flush=True)
    net.close()
    out.free()
    vol.free()


if __name__ == "__main__":
    main()
