"""Benchmark: 3D stick + plate tensor voting on one MI355X.

Workload (one step): the 128^3 float32 tubes + sheets volume (20 tubes of
radius 2, 4 planar sheets, 1% isotropic noise, seed 1002) already resident in
HBM -> per-voxel eigen-decomposition -> fused stick + integrated plate vote
(sigma=2, w=13, p=1, N=20 plate samples) -> eig3 of all 2,097,152 voted
tensors. Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mvoxel/s 3D tensor voting, MLP vs integrated plate at fixed error; eig3/s"
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector, published (needs packed FP32)
PEAK_FP32_UNPACKED_TFLOPS = 78.6  # without packed FP32 (this build disables it)
PEAK_HBM_GBS = 8000.0

# Algorithmic FLOP per in-bounds (receiver, voter) pair of the fused kernel
# (FMA = 2), counted from the formulation the kernel evaluates:
STICK_PAIR_FLOP = 31         # t = q.d, eta, r = q - 2 t d, eta * s * r r^T
PLATE_CLOSED_PAIR_FLOP = 65  # s = n.d, 4 polynomials in 1 - s^2, alpha I + beta nn^T + (h d^T + d h^T)
# Reference formulation (SURVEY counting rule): stick 33, N-sample plate 41 N + 12.
ATV3_PAIR_FLOP = 106         # a = e2.v, b = e0.v, six coefficients, 3 symmetric rank-2 updates + trace term


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (scaled by the copy-kernel calibration), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return int(json.load(f)["kernels"][kernel]["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


SPECTRAL_KERNELS = ("spec_fwd_x", "spec_fwd_y", "spec_fwd_z", "spec_mul", "spec_inv_z", "spec_inv_y", "spec_inv_x")


def spectral_bytes(shape, w, in_bytes=36, out_bytes=24, nf=15):
    """Algorithmic HBM bytes per launch of each kernel of the spectral vote:
    every array a pass reads or writes, once (zero padding that the passes
    skip is not counted). Features: nf real channels packed two per complex
    volume of FFT lengths L0 x L1 x L2 (8 B per complex value); kernel spectra
    6 nf octant values (4 B); payload 32 B per voxel (or the input tensors, in_bytes
    per voxel, when the x pass decomposes them itself); output packed-6."""
    from tiralib_amd import api
    n0, n1, n2 = shape
    L0, L1, L2 = (api.spectral_length(n, w) for n in shape)
    h0, h1, h2 = L0 // 2 + 1, L1 // 2 + 1, L2 // 2 + 1
    nz, N, c = (nf + 1) // 2, n0 * n1 * n2, 8
    # tensor input: the x pass eigen-decomposes the input tensors itself (in every channel
    # group) and no payload is written
    fused = True
    return {
        "vote3_prep": 0 if fused else N * (in_bytes + 32),
        "spec_fwd_x": N * (in_bytes if fused else 32) + nz * n0 * n1 * L2 * c,
        "spec_fwd_y": nz * n0 * n1 * L2 * c + nz * n0 * L1 * L2 * c,
        "spec_fwd_z": nz * n0 * L1 * L2 * c + nz * L0 * L1 * L2 * c,
        "spec_mul": nz * L0 * L1 * L2 * c + 6 * nf * h0 * h1 * h2 * 4 + 3 * L0 * L1 * L2 * c,
        "spec_inv_z": 3 * L0 * L1 * L2 * c + 3 * n0 * L1 * L2 * c,
        "spec_inv_y": 3 * n0 * L1 * L2 * c + 3 * n0 * n1 * L2 * c,
        "spec_inv_x": 3 * n0 * n1 * L2 * c + N * out_bytes,
    }


def spectral_roofline(shape, w, stage_ms, in_bytes=36, nf=15, out_bytes=24):
    """Per-kernel achieved HBM bandwidth of the spectral vote and the dominant kernel."""
    b = spectral_bytes(shape, w, in_bytes, out_bytes=out_bytes, nf=nf)
    per = {k: {"ms": round(stage_ms[k], 4), "bytes": int(b[k]),
               "gbs": round(b[k] / (stage_ms[k] * 1e-3) / 1e9, 1),
               "frac": round(b[k] / (stage_ms[k] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
           for k in SPECTRAL_KERNELS if stage_ms.get(k)}
    tot_b = sum(b[k] for k in per)
    tot_ms = sum(stage_ms[k] for k in per)
    dom = max(per, key=lambda k: per[k]["ms"]) if per else None
    return per, dom, {"bytes": int(tot_b), "ms": round(tot_ms, 4),
                      "gbs": round(tot_b / (tot_ms * 1e-3) / 1e9, 1) if tot_ms else None,
                      "frac": round(tot_b / (tot_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4) if tot_ms else None}


def in_bounds_pairs(shape, w):
    r = (w - 1) // 2
    p = 1
    for n in shape:
        p *= n * w - r * (r + 1) if n > r else sum(min(n - 1, i + r) - max(0, i - r) + 1 for i in range(n))
    return p


def cpu_baseline(field, sigma, w, samples, budget_s=12.0):
    """Reference-faithful FP32 CPU port (oracle), 1 thread, bounded sample."""
    from oracle import oracle as O
    s0, s1, s2 = field.shape[:3]
    n = s0 * s1 * s2
    rng = np.random.default_rng(1234)
    t = time.perf_counter()
    O.tensorvote3(field, sigma, 0.0, 1, w, samples=samples, recv=np.array([n // 2]), nthreads=1)
    t_fixed = time.perf_counter() - t  # eigen-decomposition of every voxel + one receiver
    k = 256
    t_recv = None
    while True:
        recv = np.sort(rng.choice(n, k, replace=False))
        t = time.perf_counter()
        O.tensorvote3(field, sigma, 0.0, 1, w, samples=samples, recv=recv, nthreads=1)
        dt = time.perf_counter() - t
        t_recv = max(dt - t_fixed, 1e-9) / k
        if dt > budget_s or k >= n:
            break
        k = min(n, int(k * max(2.0, budget_s / max(dt, 1e-3))))
    t = time.perf_counter()
    O.evals3(field.reshape(-1, 9))
    t_eig_out = time.perf_counter() - t
    total = t_fixed + t_recv * n + t_eig_out
    # the same receiver sample on every host core (OpenMP over receivers; the
    # per-voxel eigen-decomposition stays serial in the restatement)
    ncores = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    O.tensorvote3(field, sigma, 0.0, 1, w, samples=samples, recv=recv, nthreads=ncores)
    dt_all = time.perf_counter() - t
    t_recv_all = max(dt_all - t_fixed, 1e-9) / len(recv)
    total_all = t_fixed + t_recv_all * n + t_eig_out
    return {
        "value": round(n / total / 1e6, 6),
        "unit": "Mvoxel/s",
        "cores": 1,
        "kind": "port",
        "all_cores": {"value": round(n / total_all / 1e6, 6), "cores": ncores,
                      "note": "same sample, receivers spread over all cores; total %.1f s per step" % total_all},
        "sample": (f"eigen-decomposition of all {n} voxels and eig3 of {n} output tensors timed in full; "
                   f"vote (stick + {samples}-sample literal plate) of {k} uniformly sampled receivers timed and "
                   f"extrapolated linearly to {n} receivers; total {total:.1f} s per step"),
    }


def bench_batch(steps, warmup, check=True):
    """Batched eig3 + Gaussian (48 uneven volumes, sigma 1.5, 11 taps) in one
    persistent graph-replayed launch; returns the JSON sub-object."""
    from tiralib_amd import _lib, api, synth
    from tiralib_amd._lib import DeviceArray, Event, Stream, call
    shapes = synth.batch_shapes(48, seed=synth.SEED_CFG3)
    vols = []
    for i, s in enumerate(shapes):
        d = DeviceArray(tuple(s) + (6,))
        call("tira_synth_eig_tensors_async", d.ptr, int(np.prod(s)), 2000 + i, 1e-3, 1e2, 0.10, 1e-7, 1e-4, None)
        vols.append(d)
    call("tira_hip_synchronize")
    taps = api.gaussian_taps(11, 1.5)
    b = api.BatchGaussEig3(vols, taps, autotune=True)
    st = Stream().use()
    for _ in range(warmup):
        b.run()
    st.synchronize()
    e0, e1 = Event(), Event()
    e0.record(st)
    for _ in range(steps):
        b.run()
    e1.record(st)
    st.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    nvox = int(sum(int(np.prod(s)) for s in shapes))
    bytes_launch = nvox * (24 + 24 + 12)
    out = {
        "workload": "48 float32 6-component volumes, edges uniform over {32,48,...,256} (seed 1003), eigenvalues "
                    "log-uniform [1e-3,1e2], 10% near-degenerate (gap 1e-7..1e-4): zero-padded Gaussian sigma=1.5 "
                    "(11 taps) of every field + eig3 of every tensor, one persistent graph-replayed launch",
        "voxels": nvox, "ms": round(ms, 4),
        "eig3_per_s": round(nvox / (ms * 1e-3), 1),
        "plan": b.plan(),
        "roofline": {"kernel": "batch_gauss_eig3", "bound": "hbm", "achieved": round(bytes_launch / (ms * 1e-3) / 1e9, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(bytes_launch / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                     "traffic": profiled_traffic("k_batch_gauss_eig"),
                     "bytes_per_voxel": 60},
    }
    # eig3 alone over all 48 volumes in one launch (packed-6 in, 3 eigenvalues
    # out, 36 B per tensor), FP32 deflation and with the FP64 re-solve
    import ctypes as C
    ev = [DeviceArray(tuple(s) + (3,)) for s in shapes]
    counts = np.ascontiguousarray([int(np.prod(s)) for s in shapes], np.uint64)
    tin = (C.c_void_p * len(vols))(*[v.ptr for v in vols])
    tout = (C.c_void_p * len(ev))(*[e.ptr for e in ev])
    for name, flags in (("eig3_fp32", _lib.EIG_SCALE_INVARIANT),
                        ("eig3_fp64_refine", _lib.EIG_SCALE_INVARIANT | _lib.EIG_REFINE_FP64)):
        def run_eig():
            call("tira_eig3_batch_async", C.cast(tin, C.c_void_p), counts.ctypes.data, len(vols),
                 _lib.LAYOUT_PACKED6, C.cast(tout, C.c_void_p), None, None, flags, None)
        run_eig()
        st.synchronize()
        e0.record(st)
        for _ in range(steps):
            run_eig()
        e1.record(st)
        st.synchronize()
        ems = e0.elapsed_ms(e1) / steps
        out[name] = {"kernel": "eig3_batch (one launch, 48 arrays)", "ms": round(ems, 4),
                     "eig3_per_s": round(nvox / (ems * 1e-3), 1),
                     "gbs": round(nvox * 36 / (ems * 1e-3) / 1e9, 1),
                     "frac_hbm": round(nvox * 36 / (ems * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    if check:
        from oracle import oracle as O
        n = min(200_000, int(np.prod(shapes[0])))
        t6 = np.empty((n, 6), np.float32)
        call("tira_hip_memcpy", t6.ctypes.data, vols[0].ptr, t6.nbytes)
        got = np.empty((n, 3), np.float32)
        call("tira_hip_memcpy", got.ctypes.data, b.evals[0].ptr, got.nbytes)
        ref = O.eig3_packed6_f64(t6)
        err = np.abs(got.astype(np.float64) - ref).max(1) / np.abs(ref).max(1)
        out["eig3_max_normwise_err_vs_fp64"] = float(err.max())
        out["eig3_check_sample"] = f"first {n} tensors of volume 0 vs FP64 Jacobi"
    b.close()
    return out


def _frob6(x):
    """Frobenius norm of packed-6 symmetric tensors (off-diagonals count twice)."""
    x = np.asarray(x, np.float64).reshape(-1, 6)
    return float(np.sqrt((x[:, [0, 3, 5]] ** 2).sum() + 2 * (x[:, [1, 2, 4]] ** 2).sum()))


def bench_mlp(steps, check=True, log=None, keep_net=False):
    """Learned plate surrogate on the 256^3 curved-sheet volume (12 sheets, 2%
    plate noise, seed 1004), sigma 4, w 25: the BF16 MLP trained on the
    device (4M samples from the FP64 64-sample integrator), then stick + plate
    voting with the MLP plate (per-voter error check, integrator fallback)
    against the GPU integrated plate kernel (literal 20- and 64-sample sums)
    and the closed form. Errors against the FP64 64-sample oracle on sampled
    receivers. Returns the JSON sub-object."""
    from tiralib_amd import _lib, api, synth
    from tiralib_amd._lib import DeviceArray, Event, Stream, call
    say = log or (lambda *a: None)
    shape, sigma, w, tol = (256, 256, 256), 4.0, 25, 4e-3
    n = int(np.prod(shape))
    vol = synth.curved_sheets_volume_device(shape, 12, 0.02, seed=synth.SEED_CFG4)
    net = api.PlateMLP(sigma, 0.0, 1, 64, seed=synth.SEED_CFG4)
    t = time.perf_counter()
    tr, va = net.train(nsamples=1 << 22, epochs=100, batch=4096, lr=3e-3, lr_final=1e-6, seed=synth.SEED_CFG4)
    train_s = time.perf_counter() - t
    be = net.bin_errors()
    say(f"mlp: trained in {train_s:.1f} s, validation error {va:.2e}")
    outs = {k: DeviceArray((n, 6)) for k in ("mlp", "closed", "sum20", "sum64", "spectral")}
    nfall = net.vote(vol, outs["mlp"], shape, w, stick=True, tol=tol, in_layout=_lib.LAYOUT_PACKED6,
                     count_fallback=True)
    st = Stream().use()
    e0, e1 = Event(), Event()

    def timed(fn, reps):
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        st.synchronize()
        return e0.elapsed_ms(e1) / reps

    def trad(mode, samples, out):
        call("tira_tensorvote3_async", vol.ptr, _lib.LAYOUT_PACKED6, out.ptr, _lib.LAYOUT_PACKED6, *shape, sigma,
             0.0, w, 1, 1, 1, samples, mode)

    ms = {}
    ms["mlp"] = timed(lambda: net.vote(vol, outs["mlp"], shape, w, stick=True, tol=tol,
                                       in_layout=_lib.LAYOUT_PACKED6), steps)
    say(f"mlp: {ms['mlp']:.0f} ms per vote")
    trad(_lib.PLATE_CLOSED, 20, outs["closed"])
    ms["closed"] = timed(lambda: trad(_lib.PLATE_CLOSED, 20, outs["closed"]), steps)
    trad(_lib.VOTE_SPECTRAL, 20, outs["spectral"])  # builds the kernel spectra once
    ms["spectral"] = timed(lambda: trad(_lib.VOTE_SPECTRAL, 20, outs["spectral"]), steps)
    say(f"spectral (exact integral as an FFT convolution): {ms['spectral']:.1f} ms")
    ms["sum20"] = timed(lambda: trad(_lib.PLATE_SUM, 20, outs["sum20"]), 1)
    say(f"integrated plate N=20: {ms['sum20']:.0f} ms")
    ms["sum64"] = timed(lambda: trad(_lib.PLATE_SUM, 64, outs["sum64"]), 1)
    say(f"integrated plate N=64: {ms['sum64']:.0f} ms")
    pairs = in_bounds_pairs(shape, w)
    res = {
        "workload": "stick + plate voting, 256^3 curved-sheet volume (12 sheets, 2% plate noise, seed 1004), "
                    "sigma=4, w=25, p=1; MLP 5->64->64->64->6 BF16 on MFMA trained on 4M FP64 64-sample "
                    "integrator samples",
        "pairs_per_step": pairs,
        "mlp": {"ms": round(ms["mlp"], 2), "mvoxel_s": round(n / (ms["mlp"] * 1e-3) / 1e6, 3),
                "mfma_tflops_algorithmic": round(pairs * 17792 / (ms["mlp"] * 1e-3) / 1e12, 1),
                "mfma_frac_of_bf16_peak": round(pairs * 17792 / (ms["mlp"] * 1e-3) / 2.5e15, 4),
                "fallback_voters": nfall, "error_check_tol": tol,
                "train_s": round(train_s, 2), "train_val_rel_err": va,
                "bin_err_median": float(np.median(be)), "bin_err_max": float(be.max())},
        "integrated_sum20": {"ms": round(ms["sum20"], 2), "mvoxel_s": round(n / (ms["sum20"] * 1e-3) / 1e6, 3)},
        "integrated_sum64": {"ms": round(ms["sum64"], 2), "mvoxel_s": round(n / (ms["sum64"] * 1e-3) / 1e6, 3)},
        "closed_form": {"ms": round(ms["closed"], 2), "mvoxel_s": round(n / (ms["closed"] * 1e-3) / 1e6, 3)},
        "spectral": {"ms": round(ms["spectral"], 3), "mvoxel_s": round(n / (ms["spectral"] * 1e-3) / 1e6, 1)},
        "speedup_spectral_vs_mlp": round(ms["mlp"] / ms["spectral"], 1),
        "speedup_spectral_vs_sum20": round(ms["sum20"] / ms["spectral"], 1),
        "speedup_mlp_vs_sum20": round(ms["sum20"] / ms["mlp"], 3),
        "speedup_mlp_vs_sum64": round(ms["sum64"] / ms["mlp"], 3),
        "speedup_closed_vs_mlp": round(ms["mlp"] / ms["closed"], 3),
    }
    if check:
        from oracle import oracle as O
        # FP64 oracle on a 48^3 crop whose 24^3 core holds the receivers (their
        # w = 25 windows lie inside the crop)
        full = vol.to_host()
        occ = np.abs(full).sum(-1) > 0
        r = w // 2
        best, c0 = -1, None
        for z in range(r + 12, shape[0] - r - 12, 16):
            for y in range(r + 12, shape[1] - r - 12, 16):
                for x in range(r + 12, shape[2] - r - 12, 16):
                    k = int(occ[z - 12:z + 12, y - 12:y + 12, x - 12:x + 12].sum())
                    if k > best:
                        best, c0 = k, (z, y, x)
        z, y, x = c0
        crop = full[z - 24:z + 24, y - 24:y + 24, x - 24:x + 24]
        rng = np.random.default_rng(7)
        core = np.argwhere(np.abs(crop[12:36, 12:36, 12:36]).sum(-1) > 0) + 12
        pick = core[rng.choice(len(core), min(48, len(core)), replace=False)]
        pick = np.concatenate([pick, rng.integers(12, 36, (16, 3))])
        recv = (pick[:, 0] * 48 + pick[:, 1]) * 48 + pick[:, 2]
        t = time.perf_counter()
        ref9 = O.tensorvote3_f64(synth.mat9_from_packed6(crop).reshape(48, 48, 48, 9), sigma, 0.0, 1, w, samples=64,
                                 recv=recv,
                                 nthreads=min(16, os.cpu_count() or 1))
        ref = ref9[:, [0, 1, 2, 4, 5, 8]]
        gidx = ((pick[:, 0] + z - 24) * shape[1] + (pick[:, 1] + y - 24)) * shape[2] + (pick[:, 2] + x - 24)
        for k in ("mlp", "closed", "sum20", "sum64", "spectral"):
            got = outs[k].to_host().reshape(-1, 6)[gidx]
            key = {"mlp": "mlp", "closed": "closed_form", "sum20": "integrated_sum20",
                   "sum64": "integrated_sum64", "spectral": "spectral"}[k]
            res[key]["rel_frob_err_vs_fp64_oracle"] = _frob6(got - ref) / _frob6(ref)
        res["error_check_sample"] = (f"{len(recv)} receivers (48 on sheets, 16 uniform) in the 24^3 core of a 48^3 "
                                     f"crop; FP64 64-sample oracle, {time.perf_counter() - t:.1f} s")
    if keep_net:
        return res, net
    net.close()
    return res


def bench_pipeline(net=None, log=None):
    """Full pipeline on the 512 x 512 x 384 vascular-tree volume (seed 1005,
    radii 2-8, intensity 1 on 0, N(0, 0.1) noise): structure tensor ->
    Gaussian sigma 2 (13 taps) -> stick + plate vote sigma 4 (w 25) -> eig3
    saliency; the spectral vote (AUTO), the closed-form gather and, with
    `net`, the learned plate with integrator fallback. Per-stage times and
    roofline fractions; the skeleton stage after the spectral run.""" 
    from tiralib_amd import _lib, api, synth
    from tiralib_amd._lib import DeviceArray, Event, Stream, call
    say = log or (lambda *a: None)
    shape = (384, 512, 512)
    n = int(np.prod(shape))
    vol, segs = synth.vessel_volume_device(shape, seed=synth.SEED_CFG5, noise_sigma=0.1)
    st = Stream().use()
    e0, e1 = Event(), Event()
    pairs = in_bounds_pairs(shape, 25)
    prof = _lib.profile_get
    res = {"workload": "512x512x384 vascular tree (seed 1005, %d segments, radii 2-8, N(0,0.1) noise): structure "
                       "tensor -> Gaussian sigma=2 -> stick+plate vote sigma=4 w=25 -> eig3 saliency" % len(segs),
           "voxels": n, "pairs_per_vote": pairs}
    modes = [("spectral", _lib.VOTE_SPECTRAL, None), ("closed", _lib.PLATE_CLOSED, None)]
    if net is not None:
        modes.append(("mlp", _lib.PLATE_MLP, net))
    sal = DeviceArray(shape + (3,))
    hbm = lambda b, t: round(b * n / (t * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
    for name, mode, mlp in modes:
        if mode != _lib.PLATE_MLP:  # warm-up: workspace, offset tables / kernel spectra
            api.pipeline3(vol, smooth_sigma=2.0, vote_sigma=4.0, w=25, plate_mode=mode, saliency=sal)
        st.synchronize()
        call("tira_hip_profile_begin")
        e0.record(st)
        api.pipeline3(vol, smooth_sigma=2.0, vote_sigma=4.0, w=25, plate_mode=mode, mlp=mlp, mlp_tol=4e-3,
                      saliency=sal)
        e1.record(st)
        st.synchronize()
        call("tira_hip_profile_end")
        ms = e0.elapsed_ms(e1)
        if mode == _lib.VOTE_SPECTRAL:
            vote_stages = list(SPECTRAL_KERNELS)
        else:
            vote_stages = ["vote3_mlp" if mode == _lib.PLATE_MLP else "vote3_closed"]
        prep_name = "vote3_prep_mlp" if mode == _lib.PLATE_MLP else "vote3_prep"
        stages = {k: prof(k)[0] for k in ["structure_tensor", "gauss3_tensor", prep_name, *vote_stages, "eig3"]}
        vote_ms = sum(stages[k] for k in vote_stages)
        if mode == _lib.VOTE_SPECTRAL:
            per, dom, tot = spectral_roofline(shape, 25, stages, in_bytes=24)
            vote_roof = {"unit": "GB/s HBM", "bound": "hbm", "dominant_kernel": dom, "achieved": per[dom]["gbs"],
                         "frac": per[dom]["frac"], "all_kernels": tot, "per_kernel": per}
        elif mode == _lib.PLATE_MLP:
            vote_roof = {"unit": "TFLOP/s BF16 MFMA", "achieved": round(pairs * 17792 / (vote_ms * 1e-3) / 1e12, 1),
                         "frac": round(pairs * 17792 / (vote_ms * 1e-3) / 2.5e15, 4)}
        else:
            vote_roof = {"unit": "TFLOP/s FP32", "achieved": round(pairs * 96 / (vote_ms * 1e-3) / 1e12, 1),
                         "frac": round(pairs * 96 / (vote_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4)}
        res[name] = {"ms": round(ms, 2), "mvoxel_s": round(n / (ms * 1e-3) / 1e6, 3), "vote_ms": round(vote_ms, 3),
                     "stages_ms": {k: round(v, 3) for k, v in stages.items()},
                     "roofline": {"structure_tensor_hbm_frac": hbm(28, stages["structure_tensor"]),
                                  "gauss3_tensor_hbm_frac": hbm(48, stages["gauss3_tensor"]),
                                  "eig3_hbm_frac": hbm(36, stages["eig3"]), "vote": vote_roof}}
        say(f"pipeline ({name}): {ms:.0f} ms")
        if mode == _lib.VOTE_SPECTRAL:
            s_spec = sal.to_host()
            res["skeleton"] = bench_skeleton(sal, shape, st)
            say("skeleton (Otsu + Lee thinning) done")
            res["vote_check"] = pipeline_vote_check(vol, shape, st)
            say("FP64 check of the spectral vote done")
        elif mode == _lib.PLATE_CLOSED:
            s_cl = sal.to_host()
            res["spectral_vs_closed_saliency_rel_err"] = float(np.linalg.norm(s_spec - s_cl) / np.linalg.norm(s_cl))
    sal.free()
    vol.free()
    return res


def pipeline_vote_check(vol, shape, st, ncore=48, nuni=16):
    """Vote-level error of the config-5 spectral vote against the FP64 oracle:
    structure tensor -> Gaussian sigma 2 (13 taps) on the device, the spectral
    vote (sigma 4, w 25) of the whole smoothed field, and the FP64 oracle vote
    of receivers in the 24^3 core of a 48^3 crop centred on the densest vessel
    region (their w = 25 windows lie inside the crop): 48 receivers on vessels
    and 16 uniform."""
    from oracle import oracle as O
    from tiralib_amd import _lib, api, synth
    from tiralib_amd._lib import DeviceArray, call
    n = int(np.prod(shape))
    a, b, o = DeviceArray((n, 6)), DeviceArray((n, 6)), DeviceArray((n, 6))
    try:
        call("tira_structure_tensor3_async", vol.ptr, *shape, 1.0, 1.0, 1.0, a.ptr)
        taps = np.ascontiguousarray(api.gaussian_taps(13, 2.0), np.float32)
        call("tira_gaussian_filter3_tensor_async", a.ptr, b.ptr, *shape, 6, taps.ctypes.data, 13, 1)
        call("tira_tensorvote3_async", b.ptr, _lib.LAYOUT_PACKED6, o.ptr, _lib.LAYOUT_PACKED6, *shape, 4.0, 0.0, 25,
             1, 1, 1, 20, _lib.VOTE_SPECTRAL)
        st.synchronize()
        field = b.to_host().reshape(shape + (6,))
        got_all = o.to_host().reshape(shape + (6,))
    finally:
        a.free(); b.free(); o.free()
    mag = np.abs(field).sum(-1)
    best, c0 = -1.0, None
    for z in range(24, shape[0] - 24, 24):
        for y in range(24, shape[1] - 24, 24):
            for x in range(24, shape[2] - 24, 24):
                m = float(mag[z - 12:z + 12, y - 12:y + 12, x - 12:x + 12].sum())
                if m > best:
                    best, c0 = m, (z, y, x)
    z, y, x = c0
    crop = np.ascontiguousarray(field[z - 24:z + 24, y - 24:y + 24, x - 24:x + 24])
    rng = np.random.default_rng(5)
    core = crop[12:36, 12:36, 12:36]
    thr = np.quantile(np.abs(core).sum(-1), 0.9)
    on = np.argwhere(np.abs(core).sum(-1) >= thr) + 12
    pick = np.concatenate([on[rng.choice(len(on), min(ncore, len(on)), replace=False)],
                           rng.integers(12, 36, (nuni, 3))])
    recv = (pick[:, 0] * 48 + pick[:, 1]) * 48 + pick[:, 2]
    t = time.perf_counter()
    ref = O.tensorvote3_f64(synth.mat9_from_packed6(crop.reshape(-1, 6)).reshape(48, 48, 48, 9), 4.0, 0.0, 1, 25,
                            samples=20, recv=recv, nthreads=min(16, os.cpu_count() or 1))[:, [0, 1, 2, 4, 5, 8]]
    got = got_all[pick[:, 0] + z - 24, pick[:, 1] + y - 24, pick[:, 2] + x - 24]
    return {"spectral_vote_rel_frob_err_vs_fp64_oracle": _frob6(got - ref) / _frob6(ref),
            "sample": f"{len(recv)} receivers ({min(ncore, len(on))} on vessels, {nuni} uniform) in the 24^3 core of "
                      f"the 48^3 crop at {c0}; FP64 oracle with 20 plate samples on the smoothed structure tensor "
                      f"({time.perf_counter() - t:.1f} s)"}


def bench_skeleton(sal, shape, st):
    """Consumer end of the config-5 pipeline: Otsu threshold of the stick
    saliency (a PSD voted field, so ranking by value is ranking by magnitude),
    mask = saliency > threshold, Lee 3D thinning of the mask. The thinning loop
    reads counters back per border direction, so the stage is timed as a whole
    with events and split by the library's profile scopes."""
    from tiralib_amd import _lib
    from tiralib_amd._lib import DeviceArray, Event, call
    n = int(np.prod(shape))
    thr, mask = DeviceArray((1,)), DeviceArray(shape, np.uint8)
    e0, e1, e2 = Event(), Event(), Event()
    it = _lib.C.c_uint()
    try:
        call("tira_hip_profile_begin")
        e0.record(st)
        call("tira_otsu_threshold_async", sal.ptr, n, 3, thr.ptr)
        call("tira_threshold_async", sal.ptr, n, 3, thr.ptr, mask.ptr)
        e1.record(st)
        st.synchronize()
        fg = int(mask.to_host().sum())
        e1b = Event()
        e1b.record(st)
        call("tira_lee_thin3_u8", mask.ptr, *shape, _lib.C.byref(it))
        e2.record(st)
        st.synchronize()
        call("tira_hip_profile_end")
        t_mask, t_thin = e0.elapsed_ms(e1), e1b.elapsed_ms(e2)
        otsu_ms, _ = _lib.profile_get("otsu")
        sk = int(mask.to_host().sum())
        t = float(thr.to_host()[0])
    finally:
        thr.free()
        mask.free()
    ms = t_mask + t_thin
    return {"workload": "Otsu threshold of the stick saliency, mask, Lee 3D thinning (512x512x384)",
            "ms": round(ms, 2), "otsu_threshold_mask_ms": round(t_mask, 3), "otsu_ms": round(otsu_ms, 3),
            "lee_thin_ms": round(t_thin, 2), "threshold": t, "mask_voxels": fg, "skeleton_voxels": sk,
            "outer_iterations": int(it.value), "mvoxel_s": round(n / (ms * 1e-3) / 1e6, 1)}


def bench_vote2d(steps=20):
    """Config 1: 256 x 256 field of 8 seeded cubic Bezier curves + 2% noise
    sticks (seed 1001), stick voting sigma 3, w 19, then eigen2 saliency
    l1 - l0. The device-resident step (tira_tensorvote2 + tira_evals2_symmetric
    on device buffers; both calls block, as the reference's do), the
    host-array step through the pybind-module mirror (PCIe copies included),
    and the oracle's CPU restatement on one core; error against the FP64
    oracle."""
    from oracle import oracle as O
    from tiralib_amd import _lib, synth, tensorvote_tk as tvk
    from tiralib_amd._lib import DeviceArray, call
    f = synth.bezier_field(seed=synth.SEED_CFG1)
    H, W = f.shape[:2]
    n = H * W
    d_in, d_out, d_ev = DeviceArray.from_host(f), DeviceArray((H, W, 4)), DeviceArray((n, 2))
    try:
        def step():
            call("tira_tensorvote2", d_out.ptr, d_in.ptr, 3.0, 0.0, 1, 19, H, W, 1, 0, 20, 0)
            call("tira_evals2_symmetric", d_out.ptr, n, d_ev.ptr, 0)
        for _ in range(3):
            step()
        t = time.perf_counter()
        for _ in range(steps):
            step()
        dev_ms = (time.perf_counter() - t) / steps * 1e3
        gpu = d_out.to_host()
    finally:
        d_in.free(); d_out.free(); d_ev.free()
    t = time.perf_counter()
    for _ in range(steps):
        tvk.tensorvote_tk(f, 3.0, 0.0, 1, 19, True, False, 20)
    host_ms = (time.perf_counter() - t) / steps * 1e3
    t = time.perf_counter()
    ref32 = O.tensorvote2(f, 3.0, 0.0, 1, 19, True, False, 20, nthreads=1)
    O.evals2(ref32.reshape(-1, 4))
    cpu_s = time.perf_counter() - t
    ref64 = O.tensorvote2_f64(f, 3.0, 0.0, 1, 19, True, False, 20, nthreads=min(16, os.cpu_count() or 1))
    err = float(np.linalg.norm(gpu.astype(np.float64) - ref64) / np.linalg.norm(ref64))
    pairs = in_bounds_pairs((H, W), 19)
    return {"workload": "2D stick voting, 256x256 Bezier field (8 curves, 2% noise sticks, seed 1001), sigma=3, "
                        "w=19, then eigen2 saliency", "pixels": n, "pairs": pairs,
            "device_ms_per_step": round(dev_ms, 4), "mpx_s": round(n / dev_ms / 1e3, 1),
            "host_arrays_ms_per_step": round(host_ms, 4), "host_arrays_mpx_s": round(n / host_ms / 1e3, 1),
            "cpu_restatement_1core_s": round(cpu_s, 3), "cpu_mpx_s": round(n / cpu_s / 1e6, 3),
            "rel_frob_err_vs_fp64_oracle": err,
            "note": "launch- and synchronisation-bound: 2.28e7 pairs are ~4 us of FP32 VALU at peak"}


def bench_headline_check(d_in, d_vote, field, shape, sigma, w, samples, st, nrecv=600, steps=5):
    """The headline vote's error against the FP64 oracle on sampled receivers,
    and the direct closed-form gather (the per-pair FP32 VALU kernel) timed
    and checked on the same field."""
    from oracle import oracle as O
    from tiralib_amd import _lib
    from tiralib_amd._lib import DeviceArray, Event, call
    n = int(np.prod(shape))
    head = d_vote.to_host()
    d_g = DeviceArray((n, 6))
    e0, e1 = Event(), Event()
    try:
        def gather():
            call("tira_tensorvote3_async", d_in.ptr, _lib.LAYOUT_MAT9, d_g.ptr, _lib.LAYOUT_PACKED6, *shape, sigma,
                 0.0, w, 1, 1, 1, samples, _lib.PLATE_CLOSED)
        gather()
        st.synchronize()
        call("tira_hip_profile_begin")
        e0.record(st)
        for _ in range(steps):
            gather()
        e1.record(st)
        st.synchronize()
        call("tira_hip_profile_end")
        g_ms = _lib.profile_get("vote3_closed")[0]
        g = d_g.to_host()
    finally:
        d_g.free()
    rng = np.random.default_rng(42)
    recv = np.sort(rng.choice(n, nrecv, replace=False))
    t = time.perf_counter()
    ref = O.tensorvote3_f64(field, sigma, 0.0, 1, w, samples=samples, recv=recv,
                            nthreads=min(16, os.cpu_count() or 1))[:, [0, 1, 2, 4, 5, 8]]
    pairs = in_bounds_pairs(shape, w)
    tflops = pairs * (STICK_PAIR_FLOP + PLATE_CLOSED_PAIR_FLOP) / (g_ms * 1e-3) / 1e12
    return {
        "headline_rel_frob_err_vs_fp64_oracle": _frob6(head[recv] - ref) / _frob6(ref),
        "gather_rel_frob_err_vs_fp64_oracle": _frob6(g[recv] - ref) / _frob6(ref),
        "headline_vs_gather_rel_frob_full_field": _frob6(head - g) / _frob6(g),
        "error_sample": f"{nrecv} uniformly sampled receivers, FP64 oracle with {samples} plate samples "
                        f"({time.perf_counter() - t:.1f} s); upper-triangle components",
        "gather": {"kernel": "vote3_closed", "ms": round(g_ms, 4), "mvoxel_s_vote_only": round(n / g_ms / 1e3, 1),
                   "tflops": round(tflops, 2), "frac_of_fp32_peak": round(tflops / PEAK_FP32_TFLOPS, 4),
                   "frac_of_unpacked_peak": round(tflops / PEAK_FP32_UNPACKED_TFLOPS, 4)},
    }


def bench_atv(d_in, field, shape, sigma, w, steps, warmup, cpu_seconds=None):
    """3D analytical tensor voting (tensorvote.h:2635-2739) on the same resident
    volume, and 2D ATV of a 256^2 field (sigma 5, the reference module's window
    rule): Mvoxel/s, FLOP roofline of the gather, optional bounded CPU port."""
    from tiralib_amd import _lib
    from tiralib_amd._lib import DeviceArray, Event, Stream, call
    n = int(np.prod(shape))
    d_out = DeviceArray((n, 6))
    st = Stream().use()

    def run(method):
        def step():
            call("tira_tensorvote3_atv_method_async", d_in.ptr, _lib.LAYOUT_MAT9, d_out.ptr, _lib.LAYOUT_PACKED6,
                 *shape, sigma, w, method)
        for _ in range(warmup):
            step()
        st.synchronize()

        def timed(profiled):
            e0, e1 = Event(), Event()
            if profiled:
                call("tira_hip_profile_begin")
            e0.record(st)
            for _ in range(steps):
                step()
            e1.record(st)
            st.synchronize()
            if profiled:
                call("tira_hip_profile_end")
            return e0.elapsed_ms(e1) / steps
        # the step time without the per-kernel timing events, then a pass that collects them
        ms = timed(False)
        timed(True)
        return ms

    # the default (TIRA_PLATE_AUTO): the spectral form at this size
    ms = run(_lib.PLATE_AUTO)
    stage_ms = {k: _lib.profile_get(k)[0] for k in SPECTRAL_KERNELS}
    per, dom, tot = spectral_roofline(shape, w, stage_ms, in_bytes=36, nf=15)
    spec_out = d_out.to_host()
    # the per-pair gather
    ms_g = run(_lib.PLATE_CLOSED)
    gath_out = d_out.to_host()
    k_ms, k_cnt = _lib.profile_get("vote3_atv")
    p_ms, _ = _lib.profile_get("vote3_atv_prep")
    pairs = in_bounds_pairs(shape, w)
    achieved = pairs * ATV3_PAIR_FLOP / (k_ms * 1e-3) / 1e12
    res = {
        "workload": f"3D ATV {shape[0]}x{shape[1]}x{shape[2]} (tubes+sheets, seed 1002), sigma={sigma}, w={w}, "
                    f"device-resident, packed-6 output",
        "value": round(n / (ms * 1e-3) / 1e6, 3), "unit": "Mvoxel/s", "ms_per_step": round(ms, 4),
        "method": "spectral (31 features: eigen-subspace weights x quartic monomials, and l0)",
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": per[dom]["gbs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": per[dom]["frac"], "all_kernels": tot, "per_kernel": per},
        "spectral_vs_gather_rel_frob": _frob6(spec_out - gath_out) / _frob6(gath_out),
        "gather": {"ms_per_step": round(ms_g, 4), "mvoxel_s": round(n / (ms_g * 1e-3) / 1e6, 3),
                   "stages_ms": {"vote3_atv_prep": round(p_ms, 4), "vote3_atv": round(k_ms, 4)},
                   "roofline": {"kernel": "vote3_atv", "bound": "valu", "achieved": round(achieved, 3),
                                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                                "frac_of_unpacked_peak": round(achieved / PEAK_FP32_UNPACKED_TFLOPS, 4),
                                "traffic": profiled_traffic("k_vote3_atv"), "flop_per_pair": ATV3_PAIR_FLOP,
                                "avg_launch_ms": round(k_ms, 4), "launches": k_cnt}},
    }
    d_out.free()
    # 2D: 256^2 mixed field as the reference's test builds it, window 2 int(3 sigma) + 1
    rng = np.random.default_rng(1001)
    H = W = 256
    ang = rng.uniform(0, np.pi, (H, W))
    l1 = rng.uniform(0.5, 1.0, (H, W))
    l0 = rng.uniform(0.0, 0.5, (H, W))
    f2 = np.zeros((H, W, 4), np.float32)
    f2[..., 0] = l0 + (l1 - l0) * np.cos(ang) ** 2
    f2[..., 1] = f2[..., 2] = (l1 - l0) * np.cos(ang) * np.sin(ang)
    f2[..., 3] = l0 + (l1 - l0) * np.sin(ang) ** 2
    d2, o2 = DeviceArray.from_host(f2), DeviceArray((H, W, 4))
    w2 = 2 * int(3 * 5.0) + 1
    for _ in range(warmup):
        call("tira_tensorvote2", o2.ptr, d2.ptr, 5.0, 0.0, 1, w2, H, W, 1, 1, 10, 1)
    call("tira_hip_synchronize")
    t = time.perf_counter()
    for _ in range(steps):
        call("tira_tensorvote2", o2.ptr, d2.ptr, 5.0, 0.0, 1, w2, H, W, 1, 1, 10, 1)
    call("tira_hip_synchronize")
    ms2 = (time.perf_counter() - t) * 1e3 / steps
    res["atv2d"] = {"workload": f"2D ATV 256x256 mixed stick/ball field, sigma=5, w={w2}, device pointers, "
                                f"synchronous entry", "value": round(H * W / (ms2 * 1e-3) / 1e6, 3),
                    "unit": "Mpixel/s", "ms_per_call": round(ms2, 4)}
    d2.free(); o2.free()
    if cpu_seconds:
        from oracle import oracle as O
        t = time.perf_counter()
        O.tensorvote3_atv(field, sigma, w, recv=np.array([n // 2]), nthreads=1)
        t_fixed = time.perf_counter() - t
        k = 64
        while True:
            recv = np.sort(rng.choice(n, k, replace=False))
            t = time.perf_counter()
            O.tensorvote3_atv(field, sigma, w, recv=recv, nthreads=1)
            dt = time.perf_counter() - t
            if dt > cpu_seconds or k >= n:
                break
            k = min(n, int(k * max(2.0, cpu_seconds / max(dt, 1e-3))))
        per = max(dt - t_fixed, 1e-9) / k
        total = t_fixed + per * n
        res["cpu_baseline"] = {"value": round(n / total / 1e6, 6), "unit": "Mvoxel/s", "cores": 1, "kind": "port",
                               "sample": f"eigen-decomposition of all {n} voxels timed in full; ATV window of {k} "
                                         f"uniformly sampled receivers extrapolated to {n}; {total:.1f} s per step"}
    return res


def calibrate():
    """Achievable rates on this GPU: FP32 FMA loop (unpacked VALU) and a
    16-byte-vector device copy (HBM), each timed with HIP events."""
    from tiralib_amd._lib import DeviceArray, Event, Stream, call
    st = Stream().use()
    e0, e1 = Event(), Event()
    blocks, iters = 256 * 64, 4096
    out = DeviceArray(blocks)
    call("tira_hip_fma_loop_async", out.ptr, blocks, iters)
    e0.record(st)
    call("tira_hip_fma_loop_async", out.ptr, blocks, iters)
    e1.record(st)
    st.synchronize()
    fma_tf = blocks * 256 * 16.0 * iters / (e0.elapsed_ms(e1) * 1e-3) / 1e12
    nb = 1 << 30
    a, b = DeviceArray(nb // 4), DeviceArray(nb // 4)
    call("tira_hip_copy_async", b.ptr, a.ptr, nb)
    e0.record(st)
    for _ in range(5):
        call("tira_hip_copy_async", b.ptr, a.ptr, nb)
    e1.record(st)
    st.synchronize()
    copy_gbs = 5 * 2.0 * nb / (e0.elapsed_ms(e1) * 1e-3) / 1e9
    return {"fp32_fma_loop_tflops": round(fma_tf, 1), "copy_gbs": round(copy_gbs, 1),
            "note": "measured here: FP32 FMA loop without packed FP32, 1 GiB device copy"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 steps (about 35 ms at 128^3): the first step's host-side set-up is a small share
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plate-mode", choices=["auto", "sum", "closed", "spectral"], default="auto")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--no-atv", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the vote shards by z-slabs with an "
              f"r-plane halo exchange, which this build does not run", file=sys.stderr)
        return 2

    from tiralib_amd import _lib, synth
    from tiralib_amd._lib import DeviceArray, Event, Stream, call

    _lib.load()
    assert _lib.device_count() >= 1, "no GPU visible"
    shape, sigma, w, samples, power = (128, 128, 128), 2.0, 13, 20, 1
    mode = {"auto": _lib.PLATE_AUTO, "sum": _lib.PLATE_SUM, "closed": _lib.PLATE_CLOSED,
            "spectral": _lib.VOTE_SPECTRAL}[a.plate_mode]
    field = synth.tubes_sheets_volume(shape, seed=synth.SEED_CFG2)
    n = int(np.prod(shape))
    d_in = DeviceArray.from_host(field)
    d_vote = DeviceArray((n, 6))
    d_ev = DeviceArray((n, 3))
    st = Stream().use()

    def step():
        # the vote and eig3 of the voted field; the spectral vote writes the
        # eigenvalues from its inverse x pass (a separate eig3 otherwise)
        call("tira_tensorvote3_eig_async", d_in.ptr, _lib.LAYOUT_MAT9, d_vote.ptr, _lib.LAYOUT_PACKED6, *shape,
             sigma, 0.0, w, power, 1, 1, samples, mode, d_ev.ptr, _lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT,
             None)

    for _ in range(a.warmup):
        step()
    st.synchronize()
    call("tira_hip_synchronize")

    def timed(profiled):
        """K steps between two events; with `profiled`, the library also brackets every kernel
        with its own pair of timing events (per-kernel durations for the roofline)."""
        e0, e1 = Event(), Event()
        if profiled:
            call("tira_hip_profile_begin")
        e0.record(st)
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        e1.record(st)
        st.synchronize()
        call("tira_hip_synchronize")
        wall_s = time.perf_counter() - t0
        if profiled:
            call("tira_hip_profile_end")
        return e0.elapsed_ms(e1) / a.steps, wall_s

    # the headline clock runs without the per-kernel events; a second pass collects them
    ms, wall = timed(False)
    ms_profiled, _ = timed(True)

    prof = _lib.profile_get
    prep_ms, _ = prof("vote3_prep")
    eig_ms, _ = prof("eig3")
    pairs = in_bounds_pairs(shape, w)
    spectral = prof("spec_mul")[1] > 0
    if spectral:
        stage_ms = {k: prof(k)[0] for k in SPECTRAL_KERNELS}
        # the inverse x pass also writes the 12 B of eigenvalues per voxel
        per, dom, tot = spectral_roofline(shape, w, stage_ms, out_bytes=36)
        vote_ms = sum(stage_ms.values())
        roof = {
            "kernel": dom, "bound": "hbm", "achieved": per[dom]["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": per[dom]["frac"], "traffic": profiled_traffic(dom),
            "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per launch)",
            "algorithmic_bytes": per[dom]["bytes"], "avg_launch_ms": per[dom]["ms"],
            "spectral_vote_all_kernels": tot, "per_kernel": per,
            "pair_equivalent_tflops": round(pairs * (STICK_PAIR_FLOP + PLATE_CLOSED_PAIR_FLOP) / (vote_ms * 1e-3) / 1e12, 1),
        }
        stages = {"vote3_prep": round(prep_ms, 4), **{k: round(v, 4) for k, v in stage_ms.items()},
                  "eig3": round(eig_ms, 4)}
        vote_name = "spectral"
    else:
        vote_name = "vote3_sum" if (mode == _lib.PLATE_SUM) else "vote3_closed"
        vote_ms, vote_cnt = prof(vote_name)
        if vote_name == "vote3_closed":
            pair_flop = STICK_PAIR_FLOP + PLATE_CLOSED_PAIR_FLOP
        else:
            pair_flop = 33 + 41 * samples + 12
        achieved = pairs * pair_flop / (vote_ms * 1e-3) / 1e12
        roof = {
            "kernel": vote_name, "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": profiled_traffic("k_" + vote_name),
            "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per launch)",
            "flop_per_pair": pair_flop, "avg_launch_ms": round(vote_ms, 4), "launches": vote_cnt,
            "frac_of_unpacked_peak": round(achieved / PEAK_FP32_UNPACKED_TFLOPS, 4),
        }
        stages = {"vote3_prep": round(prep_ms, 4), vote_name: round(vote_ms, 4), "eig3": round(eig_ms, 4)}

    out = {
        "metric": METRIC,
        "value": round(n / (ms * 1e-3) / 1e6, 3),
        "unit": "Mvoxel/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3D stick+plate tensor voting 128^3 (tubes+sheets, seed 1002), sigma=2, w=13, p=1, "
                        "N=20 plate samples, then eig3 of the voted field",
            "shape": list(shape), "sigma": sigma, "w": w, "power": power, "samples": samples,
            "plate_mode": {0: "auto", 1: "sum", 2: "closed", 4: "spectral"}[mode],
            "vote_method": "spectral (FFT convolution of quartic voter features, exact plate)" if spectral
                           else vote_name,
            "pairs_per_step": pairs,
        },
        "roofline": roof,
        "stages_ms": stages,
        "stages_note": "per-kernel averages from a second pass of the same steps with the library's per-kernel "
                       "timing events on",
        "ms_per_step_profiled": round(ms_profiled, 4),
        "wall_s_timed": round(wall, 3),
    }
    out["vote2d"] = bench_vote2d()
    out["accuracy_and_gather"] = bench_headline_check(d_in, d_vote, field, shape, sigma, w, samples, st)
    out["calibration"] = calibrate()
    if not a.no_batch:
        out["eig3_batch"] = bench_batch(max(3, a.steps // 2), max(1, a.warmup))
        print("bench: batched eig3 + Gaussian done", file=sys.stderr, flush=True)
    log = lambda m: print("bench: " + m, file=sys.stderr, flush=True)
    if not a.no_atv:
        out["atv"] = bench_atv(d_in, field, shape, sigma, w, max(3, a.steps // 2), max(1, a.warmup),
                               None if a.no_cpu_baseline else min(a.cpu_seconds, 6.0))
        log("analytical tensor voting done")
    net = None
    if not a.no_mlp:
        out["mlp_plate"], net = bench_mlp(2, log=log, keep_net=True)
    if not a.no_pipeline:
        out["pipeline"] = bench_pipeline(net, log=log)
    if net is not None:
        net.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(field, sigma, w, samples, a.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
