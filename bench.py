"""Benchmark of the MI355X multiview deconvolution (plain method, config 2).

Workload (one GPU): 6 views of 512x512x256 float32, PSF 31x31x51
(sxy 1.2, sz 4.0, views every 60 deg about y), 500 seeded beads on a
background of 100, Poisson noise; blending weights normalised on the GPU
(precomputed weights); efficient-Bayesian kernels; blocks of 256^3 with
PSF-sized overlap (18 blocks per view); Tikhonov lambda 0.006.

A step is one deconvolution iteration (every view, every block).  Inputs,
weights, estimate and kernel spectra are resident in HBM before the timed
region.  Prints one JSON line.

Roofline: the dominant kernel's algorithmic bytes per launch over its
average duration (HIP events on the session stream), and the whole path's
algorithmic bytes (8P + 48H + 20E per block, view and iteration) over the
measured step time.  CPU baseline: the oracle (float32 C++ restatement of
the reference CPU path) on one view-iteration of the same workload.
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

HBM_PEAK_GBS = 8000.0
METRIC = "Gvoxel·view-iters/s MV-deconv (batched vs plain) + % HBM roofline at 1e-4"
UNIT = "Gvoxel·view-iter/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(geom_blocks, q, img_n):
    """Algorithmic HBM bytes of each pass class for one view-iteration (all
    blocks; the four transform passes run once in each of the two
    convolutions)."""
    qx, qy, qz = q
    hx = qx // 2 + 1
    H = hx * qy * qz
    P = qx * qy * qz
    b = dict(x_forward=0.0, y_forward=0.0, z_conv=0.0, y_inverse=0.0, x_inverse_quotient=0.0,
             x_inverse_update=0.0, xy_forward=0.0, xy_inverse_quotient=0.0, xy_inverse_update=0.0)
    # with several blocks the z pass of a view reads the kernel spectrum once
    # for all blocks (one line set per workgroup, looping over the blocks)
    if len(geom_blocks) > 1:
        b["z_conv"] += 2 * 8 * H
    for (ex, ey, ez) in geom_blocks:
        E = ex * ey * ez
        b["x_forward"] += 2 * (4 * P + 8 * H)
        b["y_forward"] += 2 * 16 * H
        b["z_conv"] += 2 * (8 * H + 8 * H * ez / qz) + (2 * 8 * H if len(geom_blocks) == 1 else 0)
        b["y_inverse"] += 2 * (8 * H * ez / qz + 8 * H * ez * ey / (qz * qy))
        rows = hx * ey * ez * 8
        b["x_inverse_quotient"] += rows + 8 * E
        b["x_inverse_update"] += rows + 12 * E
        # fused x-y plane kernels (256 x 256 planes): gather + spectrum write;
        # spectrum planes of the output box + epilogue
        b["xy_forward"] += 2 * (4 * P + 8 * H)
        b["xy_inverse_quotient"] += 8 * H * ez / qz + 8 * E
        b["xy_inverse_update"] += 8 * H * ez / qz + 12 * E
    return b


def path_bytes(geom_blocks, q):
    qx, qy, qz = q
    P = qx * qy * qz
    H = (qx // 2 + 1) * qy * qz
    return sum(8 * P + 48 * H + 20 * ex * ey * ez for (ex, ey, ez) in geom_blocks)


def block_effective_sizes(img, ker, bs):
    effs = []
    eff = [b - k + 1 for b, k in zip(bs, ker)]
    nb = [i // e + (1 if i % e else 0) for i, e in zip(img, eff)]
    for k in range(nb[2]):
        for j in range(nb[1]):
            for i in range(nb[0]):
                sz = []
                for d, c in enumerate((i, j, k)):
                    eo = c * eff[d]
                    sz.append(eff[d] if eo + eff[d] <= img[d] else img[d] - eo)
                effs.append(tuple(sz))
    return effs


def host_fields(threads):
    """Host description recorded with every CPU baseline."""
    from oracle import oracle as O
    return dict(nproc=O.nproc(), omp_num_threads=os.environ.get("OMP_NUM_THREADS"), threads=threads,
                cpu_model=O.cpu_model())


def cpu_baseline(ds, w, threads, repeats=3):
    """Oracle float32 path: one view-iteration of the workload (all its blocks),
    median of `repeats` samples; each sample is a 1-iteration run minus a
    0-iteration run (kernel preparation and FFT set-up excluded)."""
    from oracle import oracle as O
    views, psfs = ds.views[:1], ds.psfs[:1]
    blk = w["block"]

    def run(iters):
        t0 = time.perf_counter()
        O.deconvolve(views, psfs, O.EFFICIENT_BAYESIAN, iters, w["lam"], block_xyz=blk, norm_threads=8,
                     threads=threads)
        return time.perf_counter() - t0

    setup = run(0)
    samples = sorted(run(1) - setup for _ in range(repeats))
    it = samples[len(samples) // 2]
    n = int(np.prod(w["size_xyz"]))
    sx, sy, sz = w["size_xyz"]
    if blk:
        nb = int(np.prod([-(-s_ // (b_ - k_ + 1)) for s_, b_, k_ in zip(w["size_xyz"], blk, w["k_xyz"])]))
        where = f"{nb} blocks of {blk[0]}x{blk[1]}x{blk[2]}"
    else:
        where = "one window of img + K - 1"
    return dict(value=n / it / 1e9, unit=UNIT, cores=threads, kind="port",
                sample=f"oracle float32 CPU path, 1 view x 1 iteration of the {sx}x{sy}x{sz} workload "
                       f"({where}), median of {repeats}: {it:.2f} s (samples "
                       f"{', '.join(f'{x:.2f}' for x in samples)} s), setup excluded",
                **host_fields(threads))


def run_c3(args):
    """Config 3: 32 independent stacks, 15 iterations.  Plain = each stack
    alone (mvd_iterate, default FFT shapes); batched = all stacks in one
    batch (one launch per pass and kernel class, hipGraph replay, autotuned
    FFT shapes).  Also reports batching alone, with the default shapes and
    with the autotuned shapes on both sides."""
    import torch
    from spim_registration_amd import FFTPolicy, MVDeconBatch, MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE
    from spim_registration_amd import synthetic as S
    specs = S.c3_stack_specs()
    t = time.perf_counter()
    data = [S.c3_stack(sp) for sp in specs]
    log(f"generated 32 stacks in {time.perf_counter() - t:.1f} s")
    iters = args.steps
    units = sum(int(np.prod(sp["size_xyz"])) * sp["nviews"] for sp in specs) * iters

    def sessions(policy):
        out = []
        for ds in data:
            inp = MVDeconInput()
            for v, p in zip(ds.views, ds.psfs):
                inp.add(MVDeconFFT(v, None, p, [0]))
            out.append(MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 0, 0.0, run=False, fft_policy=int(policy)))
        return out

    res = {}
    # plain per-stack path, default and autotuned shapes
    for name, policy in (("plain_s", FFTPolicy.SMOOTH), ("plain_autotuned_s", FFTPolicy.AUTOTUNE)):
        plain = sessions(policy)
        for d in plain:
            d.iterate(1)  # warm-up (kernel loading)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for d in plain:
            d.iterate(iters)
        torch.cuda.synchronize()
        res[name] = time.perf_counter() - t0
        for d in plain:
            d.close()
    for name, policy in (("batched_same_shapes_s", FFTPolicy.SMOOTH), ("batched_s", FFTPolicy.AUTOTUNE)):
        t = time.perf_counter()
        st = sessions(policy)
        b = MVDeconBatch(st, use_graph=True)
        log(f"{name}: prepared in {time.perf_counter() - t:.1f} s")
        b.iterate(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.iterate(iters)
        torch.cuda.synchronize()
        res[name] = time.perf_counter() - t0
        b.close()
        for d in st:
            d.close()
    out = {
        "metric": "c3 batched speedup over plain per-stack GPU runs",
        "value": round(res["plain_s"] / res["batched_s"], 3),
        "unit": "x",
        "n_gpus": 1,
        "steps": iters,
        "higher_is_better": True,
        "config": {"workload": "c3: 32 stacks, sizes {96..384}^3 step 32, 2-8 views, PSF 19x19x31, "
                               "sigma_z U[3,6], 15 iterations", "units": units},
        "plain_gunits_per_s": round(units / res["plain_s"] / 1e9, 3),
        "batched_same_shapes_gunits_per_s": round(units / res["batched_same_shapes_s"] / 1e9, 3),
        "batched_gunits_per_s": round(units / res["batched_s"] / 1e9, 3),
        "plain_autotuned_gunits_per_s": round(units / res["plain_autotuned_s"] / 1e9, 3),
        "speedup_batching_only": round(res["plain_s"] / res["batched_same_shapes_s"], 3),
        "speedup_batching_only_autotuned": round(res["plain_autotuned_s"] / res["batched_s"], 3),
        "seconds": {k: round(v, 4) for k, v in res.items()},
    }
    print(json.dumps(out), flush=True)


def run_c4(args):
    """Config 4: PSF refinement on 256^3 with 4 views, PSF 31^3, start PSFs
    with sigma scaled by seeded factors in [0.8, 1.2].  Measures the GPU
    likelihood gradient (NLL, d/dpsi and d/dP for all views: forward
    convolution, PSF correlation and adjoint convolution per view), checks the
    PSF gradient against FP64 central differences of the oracle (64 seeded
    voxels per view, h = 1e-3) and runs 50 refinement steps interleaved with
    deconvolution iterations."""
    import torch
    from oracle import oracle as O
    from spim_registration_amd import MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE
    from spim_registration_amd import synthetic as S
    from spim_registration_amd.refinement import refine_psfs

    t = time.perf_counter()
    ds, w = S.workload("c4")
    start = S.perturbed_psfs(w["k_xyz"], w["sxy"], w["sz"], w["angles"], w["seed"])
    log(f"generated c4 in {time.perf_counter() - t:.1f} s")
    inp = MVDeconInput()
    for v, p in zip(ds.views, start):
        inp.add(MVDeconFFT(v, None, p, [0]))
    from spim_registration_amd import FFTPolicy
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 0, 0.0, run=False, fft_policy=int(FFTPolicy.AUTOTUNE))
    log(f"geometry {d.geometry(0)} (autotuned blocks and FFT shape)")
    d.iterate(2)
    nv = w["nviews"]
    n = int(np.prod(w["size_xyz"]))
    # gradient throughput at the kernels the iteration uses (inputs resident;
    # NLL partials and gradients returned to the host)
    kernels = [v.get_kernel1() for v in d.get_data().get_views()]
    for _ in range(args.warmup):
        d.nll_grad(psfs=kernels, psi_on_device=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nll, _, gpsf = d.nll_grad(psfs=kernels, psi_on_device=True)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    value = n * nv * args.steps / sec / 1e9
    log(f"{args.steps} gradient evaluations in {sec:.3f} s -> {value:.3f} Gvoxel·view·grad-eval/s")
    # FD check of the PSF gradient at this point
    psi = d.get_psi(masked=False)
    rng = np.random.default_rng(w["seed"])
    fd_err, fd_err_r = [], []
    t = time.perf_counter()
    for v in range(nv):
        vox = rng.choice(kernels[v].size, 64, replace=False)
        fd = O.nll_psf_fd(ds.views[v], kernels[v], psi, vox, h=1e-3)
        # Richardson combination of h and h/2: removes the O(h^2) truncation of the difference quotient
        fd_r = (4 * O.nll_psf_fd(ds.views[v], kernels[v], psi, vox, h=5e-4) - fd) / 3
        g = gpsf[v].ravel()[vox].astype(np.float64)
        fd_err.append(float(np.abs(g - fd).max() / np.abs(fd).max()))
        fd_err_r.append(float(np.abs(g - fd_r).max() / np.abs(fd_r).max()))
    log(f"FD check (64 voxels x {nv} views, FP64 CPU) in {time.perf_counter() - t:.1f} s: {fd_err}; "
        f"vs Richardson (h, h/2): {fd_err_r}")
    # refinement: 50 steps, one deconvolution iteration each
    truth = [p / p.sum(dtype=np.float64) for p in ds.psfs]

    def psf_err():
        ks = [v.get_kernel1().astype(np.float64) for v in d.get_data().get_views()]
        return float(np.mean([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(ks, truth)]))

    e0 = psf_err()
    # kernel preparation (MVDeconFFT.init for all views: normalisation, the
    # O(V^2) compound kernels, kernel2 and the kernel spectra) on the GPU,
    # beside the oracle's CPU preparation of the same PSFs (spectra excluded)
    ks = [v.get_kernel1() for v in d.get_data().get_views()]
    d.set_kernels(ks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        d.set_kernels(ks)
    torch.cuda.synchronize()
    prep_ms = 1e3 * (time.perf_counter() - t0) / 5
    from oracle.oracle import default_threads
    t0 = time.perf_counter()
    O.init_kernels(ks, O.EFFICIENT_BAYESIAN, 8)
    prep_cpu_ms = 1e3 * (time.perf_counter() - t0)
    log(f"kernel preparation: GPU {prep_ms:.2f} ms, CPU oracle {prep_cpu_ms:.1f} ms ({default_threads()} threads)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = refine_psfs(d, args.refine_steps, iterations=1)
    torch.cuda.synchronize()
    rsec = time.perf_counter() - t0
    e1 = psf_err()
    if hist:
        log(f"{args.refine_steps} refinement steps in {rsec:.2f} s; PSF rel. error {e0:.4f} -> {e1:.4f}; "
            f"NLL {hist[0].nll:.6e} -> {hist[-1].nll:.6e}")
    # the parametric rule (per-view Gaussian widths, common scale kept) on a
    # fresh deconvolution from the same start
    from spim_registration_amd.refinement import gaussian_psf_model, refine_psfs_parametric
    u = S.uniform(w["seed"], 0xC4, 2 * nv)
    theta0 = np.array([[w["sxy"] * (1 + 0.2 * (2 * u[2 * i] - 1)), w["sz"] * (1 + 0.2 * (2 * u[2 * i + 1] - 1))]
                       for i in range(nv)])
    model = gaussian_psf_model(w["k_xyz"], w["angles"])
    inp2 = MVDeconInput()
    for v, p in zip(ds.views, start):
        inp2.add(MVDeconFFT(v, None, p, [0]))
    d2 = MVDeconvolution(inp2, PSFTYPE.EFFICIENT_BAYESIAN, 0, 0.0, run=False, fft_policy=int(FFTPolicy.AUTOTUNE))
    d2.iterate(2)

    def psf_err2():
        ks = [v.get_kernel1().astype(np.float64) for v in d2.get_data().get_views()]
        return float(np.mean([np.linalg.norm(a / a.sum() - b) / np.linalg.norm(b) for a, b in zip(ks, truth)]))

    pe0 = psf_err2()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phist = refine_psfs_parametric(d2, theta0, model, args.refine_steps, iterations=1)
    torch.cuda.synchronize()
    psec = time.perf_counter() - t0
    pe1 = psf_err2()
    pnll = [h.nll for h in phist]
    pmono = all(b <= a + 1e-9 * abs(a) for a, b in zip(pnll, pnll[1:]))
    theta1 = phist[-1].params if phist else theta0
    log(f"parametric refinement: {args.refine_steps} steps in {psec:.2f} s; PSF rel. error {pe0:.4f} -> {pe1:.4f}; "
        f"NLL monotone {pmono}; widths {np.round(theta1, 3).tolist()}")
    d2.close()
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle's float32 gradient (NLL, d/dpsi, d/dP) of all views at the full size
        t0 = time.perf_counter()
        O.nll_grad(list(ds.views), kernels, psi)
        csec = time.perf_counter() - t0
        cpu = dict(value=round(n * nv / csec / 1e9, 5), unit="Gvoxel·view·grad-eval/s", cores=O.default_threads(),
                   kind="port", sample=f"oracle float32 gradient of the 4 views (256^3, PSF 31^3), one evaluation, "
                                       f"{csec:.2f} s")
        log(f"cpu baseline {cpu['value']} ({csec:.1f} s)")
    d.close()
    # path model per view and gradient evaluation: forward convolution, PSF
    # correlation and adjoint convolution over the window img + K - 1, plus the
    # residual and gradient epilogues: 3 (4P + 24H) + 16E
    win = [a + b - 1 for a, b in zip(w["size_xyz"], w["k_xyz"])]
    P = int(np.prod(win))
    H = (win[0] // 2 + 1) * win[1] * win[2]
    model = 3 * (4 * P + 24 * H) + 16 * n
    path_gbs = model * nv * args.steps / sec / 1e9
    out = {
        "metric": "c4 likelihood-gradient throughput (NLL + d/dpsi + d/dPSF, 4 views)",
        "value": round(value, 4),
        "unit": "Gvoxel·view·grad-eval/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * sec / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 reductions)",
        "data": "synthetic",
        "config": {"workload": "c4: 4 views of 256x256x256 f32, PSF 31x31x31, start sigma x U[0.8, 1.2], "
                               "whole image as one block", "views": nv},
        "path_roofline": {"bound": "hbm", "achieved": round(path_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(path_gbs / HBM_PEAK_GBS, 4),
                          "model": "3 (4P + 24H) + 16E bytes per view and gradient evaluation, P = (img + K - 1)^3"},
        "fd_check": {"voxels_per_view": 64, "h": 1e-3, "reference": "FP64 CPU oracle central differences",
                     "max_rel_err_per_view": [round(e, 6) for e in fd_err], "gate": 1e-3,
                     "pass": bool(max(fd_err) <= 1e-3),
                     "max_rel_err_vs_richardson_h_h2": [float(f"{e:.3g}") for e in fd_err_r],
                     "note": "central differences at h = 1e-3 carry an O(h^2) truncation error of about 5e-4 here; "
                             "the Richardson combination (4 D(h/2) - D(h)) / 3 removes it"},
        "kernel_preparation": {"views": nv, "psf": list(w["k_xyz"]), "gpu_ms": round(prep_ms, 3),
                               "cpu_oracle_ms": round(prep_cpu_ms, 1), "cpu_threads": default_threads(),
                               "note": "GPU time includes the 4 + 4 kernel spectra at the FFT shape"},
        "refinement": {"steps": args.refine_steps, "seconds": round(rsec, 3),
                       "ms_per_step": round(1e3 * rsec / max(1, args.refine_steps), 2),
                       "psf_rel_err_start": round(e0, 5), "psf_rel_err_end": round(e1, 5),
                       "nll_first": hist[0].nll if hist else None, "nll_last": hist[-1].nll if hist else None,
                       "rule": "EM (blind Richardson-Lucy) update of every PSF voxel"},
        "refinement_parametric": {
            "steps": args.refine_steps, "seconds": round(psec, 3), "psf_rel_err_start": round(pe0, 5),
            "psf_rel_err_end": round(pe1, 5), "nll_monotone": bool(pmono),
            "widths_start": np.round(theta0, 4).tolist(), "widths_end": np.round(theta1, 4).tolist(),
            "widths_true": [w["sxy"], w["sz"]],
            "rule": "per-view Gaussian widths; gradient in log-widths through the model, its mean over the views "
                    "(the common blur the likelihood cannot identify) removed; backtracked on the NLL"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def resample_workload():
    """Input preparation at the 512x512x256 geometry: 6 views acquired as
    512x512x128 stacks with z spacing 2, view v rotated by 60 v degrees about
    y around the box centre, resampled into the 512x512x256 bounding box at
    offset 0 with PRECOMPUTED blending weights (border -8, -8, -3; range 12)."""
    import math
    rx, ry, rz = 512, 512, 128
    box = (512, 512, 256)
    fwds = []
    for v in range(6):
        a = math.radians(60.0 * v)
        r = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
        m = r @ np.diag([1.0, 1.0, 2.0])
        cr = np.array([(rx - 1) / 2, (ry - 1) / 2, (rz - 1) / 2])
        cb = np.array([(box[0] - 1) / 2, (box[1] - 1) / 2, (box[2] - 1) / 2])
        fwds.append(np.hstack([m, (cb - m @ cr)[:, None]]))
    return (rx, ry, rz), box, fwds


def run_resample(args):
    """Input preparation on the GPU (resampling + blending weights), raw views
    and outputs resident in HBM; a step resamples all 6 views."""
    import torch
    from oracle import oracle as O
    from spim_registration_amd import transform as T

    (rx, ry, rz), box, fwds = resample_workload()
    invs = [T.inverse_affine(f) for f in fwds]
    blend = T.Blending()
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    raws = [(100.0 + 1900.0 * torch.rand((rz, ry, rx), generator=g, device=dev)).contiguous() for _ in fwds]
    nx, ny, nz = box
    outs = [torch.empty((nz, ny, nx), dtype=torch.float32, device=dev) for _ in fwds]
    wts = [torch.empty_like(o) for o in outs]

    def step():
        for raw, inv, o, w in zip(raws, invs, outs, wts):
            T.transform_view_device(raw, inv, (0, 0, 0), box, o, w, blend)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    nv = len(fwds)
    n_out, n_raw = nx * ny * nz, rx * ry * rz
    value = n_out * nv * args.steps / (ms / 1e3) / 1e9
    launch_us = 1e3 * ms / (args.steps * nv)
    bytes_view = 4 * n_raw + 8 * n_out
    gbs = bytes_view / (launch_us * 1e-6) / 1e9
    log(f"{args.steps} steps x {nv} views in {ms:.2f} ms -> {value:.2f} Gvoxel/s, {launch_us:.1f} us per view, "
        f"{gbs:.0f} GB/s algorithmic")
    # parity check: view 1 against the oracle (bit-exact)
    raw1 = raws[1].cpu().numpy()
    img, w = O.transform_view(raw1, invs[1], (0, 0, 0), box, blend.border, blend.range)
    exact = np.array_equal(outs[1].cpu().numpy(), img) and np.array_equal(wts[1].cpu().numpy(), w)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle.oracle import default_threads
        threads = default_threads()
        t0 = time.perf_counter()
        O.transform_view(raw1, invs[1], (0, 0, 0), box, blend.border, blend.range, threads=threads)
        sec = time.perf_counter() - t0
        cpu = dict(value=round(n_out / sec / 1e9, 4), unit="Gvoxel/s", cores=threads, kind="port",
                   sample=f"oracle C++ resampling + blending of 1 view (512x512x256 box), {sec:.2f} s")
    tfile = os.path.join(ROOT, "profiles", "traffic_resample.json")
    traffic = None
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("transform_view")
    out = {
        "metric": "input preparation: resampled voxels per second (image + blending weight)",
        "value": round(value, 3),
        "unit": "Gvoxel·view/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 coordinates and interpolation weights)",
        "data": "synthetic",
        "config": {"workload": "resample: 6 views of 512x512x128 (z spacing 2) rotated by 60 v deg about y into a "
                               "512x512x256 box, trilinear + PRECOMPUTED blending weights", "views": nv},
        "roofline": {"bound": "hbm", "kernel": "transform_view", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_us": round(launch_us, 2),
                     "model": "4 bytes per raw voxel + 8 bytes per output voxel (image and weight) per view"},
        "parity": {"view": 1, "bit_exact_vs_oracle": bool(exact)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_psf(args):
    """PSF extraction (ExtractPSF.extractNextImg) on the GPU: 2000 beads in a
    512x512x128 view, PSF box 41x41x81 (the 1024x1024x512 workload's), model =
    rotation by 45 deg about y with z spacing 2.  A step is one mvd_extract_psf
    call with host arrays (image upload, extraction, normalisation, transform,
    download), so the rate includes PCIe."""
    import math
    import torch
    from oracle import oracle as O
    from spim_registration_amd import psf as P
    from spim_registration_amd.transform import inverse_affine

    rng = np.random.default_rng(11)
    nx, ny, nz = 512, 512, 128
    img = rng.uniform(100.0, 200.0, (nz, ny, nx)).astype(np.float32)
    nb = 2000
    locs = np.stack([rng.uniform(0, nx - 1, nb), rng.uniform(0, ny - 1, nb), rng.uniform(0, nz - 1, nb)], 1)
    size = (41, 41, 81)
    a = math.radians(45.0)
    r = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]]) @ np.diag([1, 1, 2.0])
    model = np.hstack([r, np.zeros((3, 1))]).reshape(12)
    for _ in range(args.warmup):
        P.extract_psf(img, locs, size, model)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        psf, tpsf = P.extract_psf(img, locs, size, model)
    sec = (time.perf_counter() - t0) / args.steps
    m = int(np.prod(size))
    value = m * nb / sec / 1e9
    log(f"{args.steps} extractions, {1e3 * sec:.2f} ms each -> {value:.2f} G interpolations/s (PCIe included)")
    _, rpsf = O.extract_psf(img, locs[:50], size)
    psf50, _ = P.extract_psf(img, locs[:50], size, None)
    exact = bool(np.array_equal(psf50, rpsf))
    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        O.extract_psf(img, locs[:50], size)
        cs = time.perf_counter() - t0
        cpu = dict(value=round(m * 50 / cs / 1e9, 4), unit="G interpolations/s", cores=1, kind="port",
                   sample=f"oracle C++ extraction of 50 beads (41x41x81 box), single thread, {cs:.2f} s")
    out = {
        "metric": "PSF extraction: bead x PSF-voxel interpolations per second (host arrays, PCIe included)",
        "value": round(value, 3), "unit": "G interpolations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * sec, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 positions and weights)", "data": "synthetic",
        "config": {"workload": "psf: 2000 beads in a 512x512x128 view, PSF 41x41x81, model rot 45 deg + z x2",
                   "transformed_size": list(tpsf.shape[::-1])},
        "parity": {"beads": 50, "bit_exact_vs_oracle": exact},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_dog(args):
    """The separable-convolution hook (convolve_N) at the difference-of-Gaussian
    shape: a 512x512x256 float image, Gaussian sigma 2 (padded to 15 taps) and
    sigma 8 (63 taps) on all three axes, border extension.  A step is one
    convolve_N call with host arrays (upload, three axis passes, download);
    the device passes are timed separately with rocprofv3."""
    import torch
    from oracle import sepconv as SC
    from spim_registration_amd.cuda import CUDASeparableConvolution

    w, h, d = 512, 512, 256
    img = np.random.default_rng(13).uniform(0.0, 1000.0, (d, h, w)).astype(np.float32)
    c = CUDASeparableConvolution()
    res = {}
    for sigma, n in ((2.0, 15), (8.0, 63)):
        k = SC.padded_kernel(SC.gaussian_kernel(sigma), n)
        f = getattr(c, f"convolve_{n}")
        for _ in range(args.warmup):
            f(img.copy(), k, k, k, w, h, d, True, True, True, SC.EXTEND_BORDER_PIXELS, 0.0, 0)
        torch.cuda.synchronize()
        bufs = [img.copy() for _ in range(args.steps)]
        t0 = time.perf_counter()
        for b in bufs:
            f(b, k, k, k, w, h, d, True, True, True, SC.EXTEND_BORDER_PIXELS, 0.0, 0)
        sec = (time.perf_counter() - t0) / args.steps
        # parity on a 64-plane slab against the float64 oracle (the slab's own borders)
        sub = img[:64, :128, :128].copy()
        f(sub, k, k, k, 128, 128, 64, True, True, True, SC.EXTEND_BORDER_PIXELS, 0.0, 0)
        ref = SC.separable_convolve(img[:64, :128, :128], k, k, k, SC.EXTEND_BORDER_PIXELS)
        err = float(np.abs(sub - ref).max() / np.abs(ref).max())
        res[n] = dict(ms=round(1e3 * sec, 3), gvox_s=round(w * h * d / sec / 1e9, 3), max_rel_err=err)
        log(f"convolve_{n} (sigma {sigma}): {1e3 * sec:.2f} ms per call -> {w * h * d / sec / 1e9:.2f} Gvoxel/s "
            f"(PCIe included), max rel err {err:.2e}")
    cpu = None
    if not args.no_cpu_baseline:
        slab = img[:128]
        k15 = SC.padded_kernel(SC.gaussian_kernel(2.0), 15)
        t0 = time.perf_counter()
        SC.separable_convolve(slab, k15, k15, k15, SC.EXTEND_BORDER_PIXELS)
        csec = time.perf_counter() - t0
        cpu = dict(value=round(slab.size / csec / 1e9, 4), unit="Gvoxel/s", cores=1, kind="port",
                   sample=f"numpy float64 oracle, 15 taps on x, y, z, 128 of the 256 planes, {csec:.2f} s")
    out = {
        "metric": "separable Gaussian convolution (convolve_N hook), voxels per second, host arrays (PCIe included)",
        "value": res[15]["gvox_s"], "unit": "Gvoxel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res[15]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "dog: 512x512x256 f32, Gaussian on x, y, z, border extension; value at 15 taps"},
        "sizes": {str(k): v for k, v in res.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_export(args):
    """Output path at the 1024x1024x512 box: a float estimate resident in HBM
    is reduced to its min/max (FusionHelper.minMax) and converted to 16 bit
    over the widened range (ExportSpimData2HDF5.convert), big-endian as the
    TIFF pages take it.  A step is both passes (device to device); the TIFF
    file write of the same image (device -> pinned host -> file, pipelined)
    is timed once and reported beside it (storage bound)."""
    import tempfile
    import torch
    from oracle import export as OE
    from spim_registration_amd import export as E

    nx, ny, nz = 1024, 1024, 512
    n = nx * ny * nz
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(17)
    img = (0.01 + torch.rand((nz, ny, nx), generator=g, device=dev)).contiguous()
    out = torch.empty((nz, ny, nx), dtype=torch.uint16, device=dev)

    def step():
        lo, hi = E.u16_range(E.min_max(img))
        E.convert_u16(img, lo, hi, big_endian=True, out=out)
        return lo, hi

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lo, hi = step()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    value = n / sec / 1e9
    # the two kernels on their own, HIP events on the stream they run on
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    reps = 10
    e0.record()
    for _ in range(reps):
        E.convert_u16(img, lo, hi, big_endian=True, out=out)
    e1.record()
    for _ in range(reps):
        E.min_max(img)
    e2.record()
    torch.cuda.synchronize()
    conv_us = 1e3 * e0.elapsed_time(e1) / reps
    mm_us = 1e3 * e1.elapsed_time(e2) / reps
    conv_gbs = 6.0 * n / (conv_us * 1e-6) / 1e9
    mm_gbs = 4.0 * n / (mm_us * 1e-6) / 1e9
    log(f"min/max + 16-bit conversion: {1e3 * sec:.2f} ms per step -> {value:.1f} Gvoxel/s; conversion "
        f"{conv_us:.0f} us ({conv_gbs:.0f} GB/s), min/max {mm_us:.0f} us incl. partials and sync ({mm_gbs:.0f} GB/s)")
    # parity on 8 planes (bit-exact against the oracle)
    sl = img[100:108].cpu().numpy()
    full_mm = E.min_max(img)
    exp = OE.to_u16(sl, *OE.u16_range(full_mm)).byteswap()
    exact = bool(np.array_equal(out[100:108].cpu().numpy().view(np.uint16), exp))
    exact_mm = bool(OE.min_max(img.cpu().numpy()) == full_mm)
    # file write of the same image as a 16-bit TIFF stack
    file_res = None
    try:
        fd, path = tempfile.mkstemp(suffix=".tif")
        os.close(fd)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E.write_tiff(path, img, bits=16)
        fsec = time.perf_counter() - t0
        size = os.path.getsize(path)
        os.remove(path)
        file_res = {"bytes": size, "seconds": round(fsec, 3), "GB_s": round(size / fsec / 1e9, 2)}
        log(f"16-bit TIFF stack of {size / 2**30:.2f} GiB written in {fsec:.2f} s ({size / fsec / 1e9:.2f} GB/s)")
    except Exception as ex:  # storage of the host, not the path under test
        log(f"file write skipped: {ex}")
    cpu = None
    if not args.no_cpu_baseline:
        a = img[:32].cpu().numpy()
        t0 = time.perf_counter()
        r = OE.u16_range(OE.min_max(a))
        OE.to_u16(a, *r)
        csec = time.perf_counter() - t0
        cpu = dict(value=round(a.size / csec / 1e9, 4), unit="Gvoxel/s", cores=1, kind="port",
                   sample=f"numpy oracle min/max + 16-bit conversion of 32 of the 512 planes, {csec:.2f} s")
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_export.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("k_to_u16")
    res = {
        "metric": "output path: min/max + 16-bit conversion of the estimate, voxels per second (inputs in HBM)",
        "value": round(value, 3), "unit": "Gvoxel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * sec, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 -> u16 (f64 arithmetic)", "data": "synthetic",
        "config": {"workload": "export: 1024x1024x512 f32 estimate, minMax, widened range, 16-bit big-endian"},
        "roofline": {"bound": "hbm", "kernel": "k_to_u16", "achieved": round(conv_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(conv_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_us": round(conv_us, 2), "model": "4 bytes read + 2 bytes written per voxel"},
        "min_max": {"avg_us": round(mm_us, 2), "GB_s": round(mm_gbs, 1), "model": "4 bytes read per voxel"},
        "parity": {"min_max_bit_exact": exact_mm, "u16_planes_100_107_bit_exact": exact},
        "tiff_file": file_res,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20 for the 512x512x256 workload, the workload's own count otherwise)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain device copy of 256 MiB (counter calibration for traffic)")
    ap.add_argument("--refine-steps", type=int, default=50, help="c4: PSF refinement steps")
    ap.add_argument("--grad", type=int, default=0,
                    help="c2/c5: also time this many likelihood-gradient evaluations (blocked, checkpointed)")
    ap.add_argument("--batched", action="store_true",
                    help="run the blocks of each view through the batch engine (one launch per pass, hipGraph)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"c3": 15, "c4": 5, "c5": 3, "resample": 10, "psf": 3, "dog": 3, "export": 10}.get(args.workload, 20)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not supported",
              file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "no GPU"
    if args.workload == "c3":
        run_c3(args)
        return
    if args.workload == "c4":
        run_c4(args)
        return
    if args.workload == "resample":
        run_resample(args)
        return
    if args.workload == "psf":
        run_psf(args)
        return
    if args.workload == "dog":
        run_dog(args)
        return
    if args.workload == "export":
        run_export(args)
        return
    from spim_registration_amd import MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE, WeightType
    from spim_registration_amd import synthetic as S

    t = time.perf_counter()
    ds, w = S.workload(args.workload)
    weights = [S.blending_weight(v.shape) for v in ds.views]
    log(f"generated {args.workload} in {time.perf_counter() - t:.1f} s")

    inp = MVDeconInput()
    blk = w["block"]
    for v, p, wt in zip(ds.views, ds.psfs, weights):
        inp.add(MVDeconFFT(v, wt, p, [0], blk is not None, blk))
    t = time.perf_counter()
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 0, w["lam"], 1.0, 0, norm_threads=8,
                        weight_type=WeightType.PRECOMPUTED_WEIGHTS, run=False)
    log(f"prepared in {time.perf_counter() - t:.1f} s; geometry {d.geometry(0)}")

    if args.copy_probe:
        src = torch.empty(64 * 2**20, dtype=torch.float32, device="cuda").fill_(1.0)
        flush = torch.empty(256 * 2**20, dtype=torch.float32, device="cuda").fill_(0.0)  # evicts src from the caches
        dst = src * 1.0  # one elementwise kernel: reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del src, dst, flush
    runner = d
    if args.batched:
        from spim_registration_amd import MVDeconBatch
        runner = MVDeconBatch([d], use_graph=True)
    runner.iterate(args.warmup)
    torch.cuda.synchronize()
    # per-kernel HIP events on the session stream over the timed region itself
    # (the batch engine replays a graph and is timed per kernel in a separate pass)
    d.set_timing(not args.batched)
    t0 = time.perf_counter()
    runner.iterate(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timed_iters = args.steps
    if args.batched:
        runner.close()
        d.set_timing(True)
        d.iterate(1)
        timed_iters = 1
    tim = d.timing()
    d.set_timing(False)
    sec = t1 - t0
    n = int(np.prod(w["size_xyz"]))
    nv = w["nviews"]
    value = n * nv * args.steps / sec / 1e9
    log(f"{args.steps} iterations in {sec:.3f} s -> {value:.2f} {UNIT}")
    geo = d.geometry(0)
    q = geo["fft"]
    effs = block_effective_sizes(w["size_xyz"], w["k_xyz"], blk) if blk else [tuple(w["size_xyz"])]
    kb = kernel_bytes(effs, q, n)
    per_class = {}
    for name, (ms, cnt) in tim.items():
        if cnt and name in kb:
            bytes_total = kb[name] * nv * timed_iters  # one iteration = nv views
            per_class[name] = dict(ms_total=ms, launches=cnt, avg_us=1e3 * ms / cnt,
                                   gbs=bytes_total / (ms * 1e-3) / 1e9)
    dom = max(per_class, key=lambda k: per_class[k]["ms_total"])
    dc = per_class[dom]
    achieved = dc["gbs"]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(dom)
        except Exception:
            traffic = None
    pb = path_bytes(effs, q) * nv * args.steps
    path_gbs = pb / sec / 1e9
    log("per-kernel: " + json.dumps(per_class, indent=None))

    grad = None
    if args.grad > 0:
        kernels = [v.get_kernel1() for v in d.get_data().get_views()]
        d.nll_grad(psfs=kernels, psi_on_device=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.grad):
            d.nll_grad(psfs=kernels, psi_on_device=True)
        torch.cuda.synchronize()
        gs = (time.perf_counter() - t0) / args.grad
        grad = {"seconds_per_evaluation": round(gs, 4), "gvoxel_view_grad_eval_per_s": round(n * nv / gs / 1e9, 4),
                "what": "NLL, d/dpsi and d/dPSF of every view over the deconvolution's blocks, window spectra "
                        "recomputed per block (checkpointing)"}
        log(f"gradient: {gs * 1e3:.1f} ms per evaluation")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.oracle import default_threads
        threads = default_threads()
        t = time.perf_counter()
        cpu = cpu_baseline(ds, w, threads, repeats=1 if args.workload == "c5" else 3)
        log(f"cpu baseline {cpu['value']:.4f} {UNIT} ({time.perf_counter() - t:.1f} s)")
    d.close()

    # a plain device-to-device copy of 1 GiB beside the roofline: the practical
    # HBM rate of a streaming read + write on this device
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    dst.copy_(src)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(5):
        dst.copy_(src)
    c1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * src.numel() * 4 * 5 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src, dst
    log(f"copy probe: {copy_gbs:.0f} GB/s (read + write of 1 GiB)")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * sec / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {nv} views of {w['size_xyz'][0]}x{w['size_xyz'][1]}x"
                               f"{w['size_xyz'][2]} f32, PSF {w['k_xyz'][0]}x{w['k_xyz'][1]}x{w['k_xyz'][2]}, "
                               f"blocks {blk}, lambda {w['lam']}, efficient-Bayesian, precomputed weights",
                   "views": nv, "size_xyz": list(w["size_xyz"]), "block": list(blk) if blk else None,
                   "fft_shape": list(q), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_us": round(dc["avg_us"], 2)},
        "path_roofline": {"achieved": round(path_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(path_gbs / HBM_PEAK_GBS, 4),
                          "model": "8P + 48H + 20E bytes per block, view and iteration"},
        "copy_probe": {"GB/s": round(copy_gbs, 1), "what": "torch device-to-device copy of 1 GiB (read + write)"},
        "kernels": {k: {"avg_us": round(v["avg_us"], 2), "GB/s": round(v["gbs"], 1),
                        "share": round(v["ms_total"] / sum(x["ms_total"] for x in per_class.values()), 3)}
                    for k, v in per_class.items()},
        "cpu_baseline": cpu,
    }
    if grad:
        out["gradient"] = grad
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
