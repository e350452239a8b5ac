"""Benchmark: standard gridding of the C2 workload on one MI355X.

Workload (Briggs robust=0.5 standard gridding, FP32-Kahan vs FP64):
  50-antenna compact array (1225 baselines), log-normal radii <= 3 km,
  2000 integrations x 4 s at 100 GHz, 64 channels x 4 pols (XX,XY,YX,YY)
  complex64, 200 point sources + 5 Gaussian blobs, 10% random flags;
  2048 x 2048 cube at 0.1", padded grid 2457 x 2457 (256 planes), PS kernel
  support 7, oversampling 100.

One step = one ngc_standard_grid call (image mode, chosen FP32-Kahan form)
over all 627.2M visibility samples resident in HBM, including the uv-tile
binning.  Briggs imaging weights are computed once, on the GPU, before the
timed region (also timed and reported).  The dirty-image error of the chosen
form is measured against the FP64 form on the same inputs, per plane.

Further legs, reported in the same JSON line under their own keys (each can
be skipped with --no-<leg> or run alone with --<leg>-only): 'mosaic' (C4,
tiled aperture gridder against the plain FP64 one, and the learned A-term
surrogate), 'batched' (C3, the autotuned hipGraph gridding plan), 'streamed'
(the C5 one-GPU share from pinned host memory) and 'predict' (visibility
prediction at the C2 shapes).

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(n_vis, n_rows, n_planes, npad):
    """Image-pass algorithmic HBM bytes: c64 vis + f32 weight + u8 flag per
    sample, f64 uvw per row, one c64 grid write per cell."""
    return n_vis * 13 + n_rows * 24 + n_planes * npad * npad * 8


def measured_traffic(precision, config):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same
    counter's reading for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"std_grid_tiled_{config}_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        return rec[precision]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(cfg, data, npad, delta, cgk, n_chan_sample, n_chan_cube):
    """The oracle (FP64 C restatement of the reference loop), compiled here
    with -O3 -march=native (no fast-math, no contraction), timed on this host's
    cores in three shapes:

    - cube: the reference's cube-mode chunking by channel, one task per
      channel of the first n_chan_cube channels, each gridding all its
      samples into its own image planes (no private grids, no tree-sum);
    - dask time chunks: one task per time chunk of the first n_chan_sample
      channels with a private grid, then a pairwise tree-sum (the reference's
      graph layer as it chunks continuum data);
    - one thread, one chunk, on the first n_chan_sample channels.

    Returns the three records; each states the CPU model and threads used."""
    import concurrent.futures

    import oracle
    lib = oracle.use_native_build()
    model, avail = oracle.cpu_model()
    n_thr = int(os.environ.get("OMP_NUM_THREADS", avail))
    t0 = time.time()
    n_c = max(n_chan_sample, n_chan_cube)
    uvw = data["uvw"].cpu().numpy()
    vis = data["vis"][:, :, :n_c].cpu().numpy().astype(np.complex128)
    w = data["imaging_weight"][:, :, :n_c].cpu().numpy().astype(np.float64)
    flag = data["flag"][:, :, :n_c].cpu().numpy().astype(bool)
    w[flag] = np.nan
    fr = data["freq"][:n_c]
    log(f"cpu baseline: {model}, {avail} cores available, {n_thr} threads; host copy "
        f"{time.time() - t0:.1f}s; oracle built at {lib}")
    host = f"{model}, {n_thr} of {avail} cores"

    def rec(n_samples, dt, threads, how):
        r = dict(value=round(n_samples / dt / 1e6, 3), unit="Mvis/s", cores=int(threads),
                 kind="port", cpu=host,
                 sample=f"FP64 C oracle (reference loop order, -O3 -march=native), {how}, image "
                        f"mode ({n_samples} unflagged samples), {dt:.1f} s")
        log(f"cpu baseline {r}")
        return r

    # cube: one task per channel (contiguous per-channel inputs, as chunked
    # storage gives them), the GIL released inside each oracle call
    chans = [(np.ascontiguousarray(vis[:, :, c:c + 1]), np.ascontiguousarray(w[:, :, c:c + 1]))
             for c in range(n_chan_cube)]
    n_cube = int((~flag[:, :, :n_chan_cube]).sum())
    t0 = time.time()
    with concurrent.futures.ThreadPoolExecutor(max_workers=n_thr) as ex:
        outs = list(ex.map(lambda c: oracle.standard_grid(
            chans[c][0], uvw, chans[c][1], fr[c:c + 1], cgk, [npad, npad], delta, 7, 100,
            do_psf=False), range(n_chan_cube)))
    dt = time.time() - t0
    del outs, chans
    cube = rec(n_cube, dt, min(n_thr, n_chan_cube),
               f"{n_chan_cube} tasks of one channel each (cube-mode chunking, every task "
               f"writing only its own planes) on {min(n_thr, n_chan_cube)} threads, first "
               f"{n_chan_cube} of {cfg.n_chan} channels")
    vis_s = np.ascontiguousarray(vis[:, :, :n_chan_sample])
    w_s = np.ascontiguousarray(w[:, :, :n_chan_sample])
    n_s = int((~flag[:, :, :n_chan_sample]).sum())
    out = []
    for threads in (0, 1):
        nt = threads or n_thr
        chunk = uvw.shape[0] if threads == 1 else -(-uvw.shape[0] // nt)
        t0 = time.time()
        grid, sw, used = oracle.standard_grid_chunked(vis_s, uvw, w_s, fr[:n_chan_sample], cgk,
                                                      [npad, npad], delta, 7, 100,
                                                      do_psf=False, time_chunk=chunk,
                                                      n_threads=nt)
        dt = time.time() - t0
        del grid
        how = ("1 thread, one chunk" if threads == 1 else
               f"{used} threads, {-(-uvw.shape[0] // chunk)} time chunks with private grids "
               f"and a pairwise tree-sum (the reference's dask graph shape)")
        out.append(rec(n_s, dt, used, f"{how}, first {n_chan_sample} of {cfg.n_chan} channels"))
    return cube, out[0], out[1]


def cpu_baseline_c1(uvw, vis, w, fr, cgk, npad, delta):
    """The oracle on the whole C1 image call (3.37M samples, 16 channels), on
    one thread and in cube mode (one task per channel on up to 16 threads),
    compiled and timed on this host as the C2 CPU baseline is."""
    import concurrent.futures

    import oracle
    oracle.use_native_build()
    model, avail = oracle.cpu_model()
    n_thr = int(os.environ.get("OMP_NUM_THREADS", avail))
    uvw_h = uvw.cpu().numpy()
    vis_h = vis.cpu().numpy().astype(np.complex128)
    w_h = w.cpu().numpy().astype(np.float64)
    n = int(np.isfinite(w_h).sum())
    C = vis_h.shape[2]
    t0 = time.time()
    oracle.standard_grid(vis_h, uvw_h, w_h, fr, cgk, [npad, npad], delta, 7, 100, do_psf=False)
    t1 = time.time() - t0
    chans = [(np.ascontiguousarray(vis_h[:, :, c:c + 1]), np.ascontiguousarray(w_h[:, :, c:c + 1]))
             for c in range(C)]
    nt = min(n_thr, C)
    t0 = time.time()
    with concurrent.futures.ThreadPoolExecutor(max_workers=nt) as ex:
        list(ex.map(lambda c: oracle.standard_grid(chans[c][0], uvw_h, chans[c][1], fr[c:c + 1],
                                                   cgk, [npad, npad], delta, 7, 100,
                                                   do_psf=False), range(C)))
    tc = time.time() - t0
    how = "FP64 C oracle (reference loop order, -O3 -march=native), image mode, whole C1 call"
    return dict(one_thread=dict(value=round(n / t1 / 1e6, 3), unit="Mvis/s", cores=1,
                                kind="port", cpu=f"{model}", sample=f"{how}, 1 thread, {t1:.2f} s"),
                cube=dict(value=round(n / tc / 1e6, 3), unit="Mvis/s", cores=nt, kind="port",
                          cpu=f"{model}, {nt} of {avail} cores",
                          sample=f"{how}, {C} tasks of one channel each on {nt} threads, "
                                 f"{tc:.2f} s"))


def c1_leg(args):
    """C1 at full size: 27 antennas in a 1 km disk, 300 x 10 s integrations at
    1.4 GHz, 16 channels x 2 pols (3.37M samples), 512^2 at 2", natural
    weights, make_psf + make_image.  Times the PSF and image gridding calls
    (inputs resident in HBM) in the chosen form and the FP64 drop-in, states
    the dirty-image error per plane, and times make_psf + make_image end to
    end through the API (FFT, correction, PSF fit) on device-resident inputs."""
    import torch
    from cngi_prototype_amd import _native as nat
    from cngi_prototype_amd import synthetic as syn
    from cngi_prototype_amd.imaging import make_image, make_psf
    from cngi_prototype_amd.imaging._imaging_utils._gridding_convolutional_kernels import \
        cropped_correction_factors, prolate_spheroidal_kernel_1d
    from cngi_prototype_amd.xds import Dataset

    cfg = syn.CONFIGS["C1"]
    N = cfg.image_size
    npad = int(1.2 * N)
    cell = cfg.cell_arcsec * syn.ARCSEC
    delta = np.array([-cell, cell])
    cgk = prolate_spheroidal_kernel_1d(100, 7)
    C, P = cfg.n_chan, cfg.n_pol
    d = syn.device_dataset(cfg)
    torch.cuda.synchronize()
    uvw, vis, w, fr = d["uvw"], d["vis"], d["weight"], d["freq"]
    n_vis = int(cfg.n_vis)
    n_rows = cfg.n_time * cfg.n_baseline
    corr = cropped_correction_factors([npad, npad], [N, N])

    def call(prec, mode):
        return nat.standard_grid(uvw, w, fr, np.arange(C), np.arange(P), cgk, C, [npad, npad],
                                 delta, 7, 100, mode, prec,
                                 vis=vis if mode == nat.GRID_IMAGE else None)

    times, kernel_ms, imgs = {}, {}, {}
    for prec in (args.precision, "fp64"):
        for mode, name in ((nat.GRID_PSF, "psf"), (nat.GRID_IMAGE, "image")):
            for _ in range(max(1, args.warmup)):
                call(prec, mode)
            torch.cuda.synchronize()
            nat.profile_enable(True)
            kms = []
            t0 = time.time()
            for _ in range(args.steps):
                call(prec, mode)
                kms.append(sum(ms for n_, ms in nat.profile_read()
                               if n_ in ("std_grid_tiled", "std_grid_plain")))
            torch.cuda.synchronize()
            times[(prec, name)] = (time.time() - t0) / args.steps * 1e3
            kernel_ms[(prec, name)] = float(np.mean(kms))
            nat.profile_enable(False)
            g, sw = call(prec, mode)
            imgs[(prec, name)] = nat.grid_to_image(g, [N, N], sum_weight=sw, corr=corr)
            del g
    errs = {}
    for name in ("psf", "image"):
        a, b = imgs[(args.precision, name)], imgs[("fp64", name)]
        dd = (a - b).reshape(C * P, -1)
        rr = b.reshape(C * P, -1)
        errs[name] = float((torch.linalg.norm(dd, dim=1) / torch.linalg.norm(rr, dim=1)).max())
    del imgs
    # end to end through the API on device-resident inputs
    dims4 = ["time", "baseline", "chan", "pol"]
    xds = Dataset(data_vars={"DATA": (dims4, vis), "UVW": (["time", "baseline", "uvw_index"], uvw),
                             "DATA_WEIGHT": (dims4, w),
                             "chan_width": (["chan"], np.full(C, cfg.df))},
                  coords={"chan": (["chan"], fr), "pol": (["pol"], np.array([9, 12]))},
                  attrs={"data_groups": [{"0": {"id": "0", "data": "DATA", "uvw": "UVW",
                                                "weight": "DATA_WEIGHT",
                                                "imaging_weight": "DATA_WEIGHT"}}]})
    mxds = Dataset(attrs={"xds0": xds})
    gp = {"image_size": [N, N], "cell_size": [cfg.cell_arcsec] * 2, "fft_padding": 1.2,
          "chan_mode": "cube", "phase_center": syn.phase_center(cfg),
          "precision": args.precision}
    make_image(mxds, make_psf(mxds, Dataset(), gp, {"xds": "xds0"}, {}), gp, {"xds": "xds0"}, {})
    torch.cuda.synchronize()
    t0 = time.time()
    img = make_psf(mxds, Dataset(), gp, {"xds": "xds0"}, {})
    img = make_image(mxds, img, gp, {"xds": "xds0"}, {})
    torch.cuda.synchronize()
    t_api = (time.time() - t0) * 1e3
    del img, mxds, xds
    cpu = None if args.no_cpu_baseline else cpu_baseline_c1(uvw, vis, w, fr, cgk, npad, delta)
    ck = args.precision
    abytes = algorithmic_bytes(n_vis, n_rows, C * P, npad)
    t_k = kernel_ms[(ck, "image")] / 1e3
    pair = times[(ck, "psf")] + times[(ck, "image")]
    pair64 = times[("fp64", "psf")] + times[("fp64", "image")]
    return dict(workload=f"C1: {cfg.n_ant} antennas x {cfg.n_time} integrations, {C} chan x {P} "
                         f"pol at 1.4 GHz, {N}^2 at {cfg.cell_arcsec}\" (padded {npad}^2), "
                         f"natural weights, make_psf + make_image", samples=n_vis,
                value=round(n_vis / (times[(ck, "image")] / 1e3) / 1e6, 1), unit="Mvis/s",
                precision=ck,
                image_ms=round(times[(ck, "image")], 3), psf_ms=round(times[(ck, "psf")], 3),
                psf_plus_image_mvis_s=round(n_vis / (pair / 1e3) / 1e6, 1),
                plain_fp64_image_ms=round(times[("fp64", "image")], 3),
                plain_fp64_psf_ms=round(times[("fp64", "psf")], 3),
                plain_fp64_psf_plus_image_mvis_s=round(n_vis / (pair64 / 1e3) / 1e6, 1),
                speedup_vs_plain_fp64=round(pair64 / pair, 2),
                error_vs_fp64=dict(image_rel_rms_max=errs["image"], psf_rel_rms_max=errs["psf"],
                                   planes=C * P),
                roofline=dict(bound="hbm", achieved=round(abytes / t_k / 1e9, 1),
                              peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=round(abytes / t_k / 1e9 / HBM_PEAK_GBS, 4),
                              algorithmic_bytes=int(abytes),
                              kernel_ms=round(kernel_ms[(ck, "image")], 4)),
                make_psf_plus_make_image_api_ms=round(t_api, 1),
                cpu_baseline=cpu)


def mosaic_surrogate(d, delta, npad, C):
    """Per-channel A-term kernels of C4 from the learned surrogate (trained on
    exact GPU kernels at 38 frequencies x 3 dish pairs x 21^2 sub-cell offsets
    = 50,274 examples), checked tap by tap against the exact per-channel
    kernels (fallback above 1e-4 of the peak).  Returns (kernels, exact
    kernels, supports per channel, report)."""
    import torch
    from cngi_prototype_amd.imaging._imaging_utils import _kernel_surrogate as ks
    from cngi_prototype_amd.imaging._imaging_utils._gcf_maps import cf_chan_map, gcf_kernels
    pairs = np.array([[0, 0], [0, 1], [1, 1]])
    blockage = np.array([0.75, 0.75])
    os_ = np.array([20, 20])
    ms = np.array([18, 18])
    pb_cell = delta * os_
    npad2 = np.array([npad, npad])
    cfc, pb_freq = cf_chan_map(d["freq"], 0.005)
    torch.cuda.synchronize()
    t0 = time.time()
    wck_b, ck_b, sup_b = gcf_kernels(pb_freq, d["dish"], blockage, pairs, True, pb_cell, npad2,
                                     os_, ms, 0.025)
    torch.cuda.synchronize()
    t_bin = time.time() - t0
    half = 0.5 * (d["freq"].max() - d["freq"].min()) + 50e6
    model, tinfo = ks.train_surrogate(d["dish"], blockage, pairs, True, pb_cell, npad2, os_, ms,
                                      sup_b[:, 0, 0, :], pb_freq, half, n_freq=38, steps=10000,
                                      batch=4096, lr=4e-3, log_every=2500)
    ck_c, _, rep = ks.per_channel_kernels(model, d["freq"], cfc, pb_freq, d["dish"], blockage,
                                          pairs, True, pb_cell, npad2, os_, ms, sup_b, ck_b,
                                          wck_b, check=True, keep_exact=True)
    torch.cuda.synchronize()
    sup_c = np.ascontiguousarray(sup_b[:, cfc])
    info = dict(network="4 dense layers, width 256, bf16 matrix-core GEMMs, FP32 "
                        "accumulation, Adam", examples=tinfo["examples"],
                data_s=tinfo["data_s"], train_s=tinfo["train_s"],
                loss_history=tinfo["history"], kernels=int(ck_c.shape[0] * ck_c.shape[1]),
                bins=int(len(pb_freq)), bin_kernels_s=round(t_bin, 3),
                surrogate_kernels_s=rep["surrogate_s"],
                exact_per_channel_s=rep.get("exact_s"),
                max_tap_err_bin_kernel=rep["max_err_bin"],
                max_tap_err_surrogate=rep["max_err_surrogate"],
                fallbacks=len(rep["fallback"]), tolerance=ks.TOL_PEAK)
    del model
    return ck_c, rep.pop("ck_exact"), sup_c, info


def mosaic_leg(args):
    """C4: 25-pointing heterogeneous mosaic (12 x 7 m + 40 x 12 m, 230 GHz, 32 chan x 2
    pol, 1440 x 10 s dumps, 1024^2 at 0.09", GCF oversampling 20, supports 11/15/17).

    The timed chosen form grids with per-channel kernels from the learned
    surrogate (the reference uses one kernel per dish pair for all 32 channels
    of the 0.5% frequency bin).  Beside it: the chosen form with the bin
    kernels, and the plain FP64 drop-in.  Dirty-image errors are stated per
    plane against the FP64 image made with the exact per-channel kernels."""
    import torch
    from cngi_prototype_amd import _native as nat
    from cngi_prototype_amd import synthetic as syn
    from cngi_prototype_amd.imaging import make_gridding_convolution_function
    from cngi_prototype_amd.xds import Dataset

    cfg = syn.CONFIGS["C4"]
    N = cfg.image_size
    npad = int(1.2 * N)
    t0 = time.time()
    d = syn.device_mosaic_dataset(cfg)
    torch.cuda.synchronize()
    log(f"mosaic: generated {cfg.n_vis / 1e6:.1f}M samples in {time.time() - t0:.1f}s")
    pc = syn.phase_center(cfg)
    dims4 = ["time", "baseline", "chan", "pol"]
    xds = Dataset(
        data_vars={"DATA": (dims4, d["vis"]), "UVW": (["time", "baseline", "uvw_index"], d["uvw"]),
                   "DATA_WEIGHT": (dims4, d["weight"]), "FLAG": (dims4, d["flag"]),
                   "ANTENNA1": (["baseline"], d["a1"]), "ANTENNA2": (["baseline"], d["a2"]),
                   "FIELD_ID": (["time", "baseline"], d["field"])},
        coords={"chan": (["chan"], d["freq"]), "pol": (["pol"], np.array([9, 12]))},
        attrs={"data_groups": [{"0": {"id": "0", "data": "DATA", "uvw": "UVW", "flag": "FLAG",
                                      "weight": "DATA_WEIGHT"}}]})
    field = Dataset(data_vars={"PHASE_DIR": (["field_id", "poly_id", "d2"], d["fields"][:, None, :])},
                    coords={"field_id": (["field_id"], np.arange(len(d["fields"])))})
    mxds = Dataset(attrs={"xds0": xds, "FIELD": field})
    gcf_parms = dict(function="casa_airy", list_dish_diameters=list(d["dish"]),
                     list_blockage_diameters=[0.75, 0.75], unique_ant_indx=list(d["types"]),
                     phase_center=pc, oversampling=[20, 20], max_support=[18, 18])
    grid_parms = dict(image_size=[N, N], cell_size=[cfg.cell_arcsec] * 2, fft_padding=1.2,
                      chan_mode="cube", phase_center=pc)
    torch.cuda.synchronize()
    t0 = time.time()
    gcf = make_gridding_convolution_function(mxds, gcf_parms, grid_parms, {"xds": "xds0"})
    torch.cuda.synchronize()
    t_gcf = time.time() - t0
    sup = np.asarray(gcf["SUPPORT"].values)
    log(f"mosaic: GCF in {t_gcf:.2f}s, supports {sup[:, 0, 0, 0].tolist()}")
    cell = cfg.cell_arcsec * syn.ARCSEC
    delta = np.array([-cell, cell])
    C, P = cfg.n_chan, cfg.n_pol
    cfb = np.asarray(gcf["CF_BASELINE_MAP"].values)
    cfp = np.asarray(gcf["CF_POL_MAP"].values)
    pg = gcf["PHASE_GRADIENT"].data
    fid = np.arange(len(d["fields"]))
    # kernel sets: (CONV_KERNEL, SUPPORT, CF_CHAN_MAP)
    bins = (gcf["CONV_KERNEL"].data, sup, np.asarray(gcf["CF_CHAN_MAP"].values))
    chans = exact = None
    surrogate = None
    if not args.mosaic_grid_only:
        try:
            ck_c, ck_x, sup_c, surrogate = mosaic_surrogate(d, delta, npad, C)
            chans = (ck_c, sup_c, np.arange(C))
            exact = (ck_x, sup_c, np.arange(C))
            log(f"mosaic surrogate: {surrogate}")
        except Exception as e:  # report, never fake
            log(f"surrogate failed: {e!r}")
            surrogate = {"error": repr(e)}
    main = chans if chans is not None else bins

    def call(precision, kern, grid=None, sw=None):
        ck, sp, cfc = kern
        return nat.aperture_grid(d["uvw"], d["weight"], d["field"], fid, d["freq"], np.arange(C),
                                 np.arange(P), cfb, cfc, cfp, ck, sp, pg, C,
                                 [npad, npad], delta, [20, 20], nat.GRID_IMAGE, precision,
                                 vis=d["vis"], grid=grid, sum_weight=sw)

    def timed(kern, grid, sw, steps):
        for _ in range(args.warmup):
            call(args.precision, kern, grid, sw)
        torch.cuda.synchronize()
        nat.profile_enable(True)
        phases = {}
        t0 = time.time()
        for _ in range(steps):
            call(args.precision, kern, grid, sw)
            for name, ms in nat.profile_read():
                phases[name] = phases.get(name, 0.0) + ms / steps
        torch.cuda.synchronize()
        ms_step = (time.time() - t0) / steps * 1e3
        nat.profile_enable(False)
        return ms_step, phases

    def plain(kern):
        torch.cuda.synchronize()
        t0 = time.time()
        g, s = call("fp64", kern)
        torch.cuda.synchronize()
        return g, s, (time.time() - t0) * 1e3

    def plane_err(img, ref):
        dd = (img - ref).reshape(C * P, -1)
        rr = ref.reshape(C * P, -1)
        rel = (torch.linalg.norm(dd, dim=1) / torch.linalg.norm(rr, dim=1)).cpu().numpy()
        return dict(rel_rms_max=float(rel.max()), rel_rms_median=float(np.median(rel)),
                    planes=int(C * P))

    grid = torch.empty((C, P, npad, npad), dtype=torch.complex64, device="cuda")
    sw = torch.empty((C, P), dtype=torch.float64, device="cuda")
    ms_step, phases = timed(main, grid, sw, args.steps)
    n_vis = cfg.n_vis
    kern_name = "surrogate per-channel" if chans is not None else "bin"
    log(f"mosaic chosen form {args.precision}, {kern_name} kernels: {ms_step:.2f} ms, "
        f"phases {phases}")
    t_k = phases.get("ap_grid_tiled", float("nan")) / 1e3
    if args.mosaic_grid_only:
        # the timed gridding steps alone (for counter runs)
        return dict(samples=int(n_vis), ms_per_step=round(ms_step, 3),
                    kernel_ms=round(t_k * 1e3, 3), phases_ms=phases)
    # the same call as an aperture gridding plan: every pointing, channel and
    # pol replayed as one captured graph (tile edge autotuned at creation)
    torch.cuda.synchronize()
    t0 = time.time()
    ck_m, sp_m, cfc_m = main
    gp = torch.empty_like(grid)
    swp = torch.empty_like(sw)
    plan = nat.ApertureGridPlan(d["uvw"], d["weight"], d["field"], fid, d["freq"], np.arange(C),
                                np.arange(P), cfb, cfc_m, cfp, ck_m, sp_m, pg, C, [npad, npad],
                                delta, [20, 20], nat.GRID_IMAGE, args.precision, vis=d["vis"],
                                grid=gp, sum_weight=swp, autotune=True)
    torch.cuda.synchronize()
    t_plan_make = time.time() - t0
    for _ in range(args.warmup):
        plan.execute()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(args.steps):
        plan.execute()
    torch.cuda.synchronize()
    ms_plan = (time.time() - t0) / args.steps * 1e3
    plan.status()
    plan_err = float(torch.linalg.norm((gp - grid).ravel()) / torch.linalg.norm(grid.ravel()))
    plan_leg = dict(ms_per_step=round(ms_plan, 3), mvis_s=round(n_vis / (ms_plan / 1e3) / 1e6, 1),
                    create_s=round(t_plan_make, 2), describe=plan.describe(),
                    rel_diff_vs_eager=plan_err)
    plan.close()
    del gp, swp
    log(f"mosaic plan replay: {ms_plan:.2f} ms ({plan_leg['describe']})")
    img_main = nat.grid_to_image(grid, [N, N], sum_weight=sw)
    bin_leg = None
    if chans is not None:
        # the reference's kernels (one per bin), same form
        gb = torch.empty_like(grid)
        swb = torch.empty_like(sw)
        ms_bin, ph_bin = timed(bins, gb, swb, args.steps)
        img_bin = nat.grid_to_image(gb, [N, N], sum_weight=swb)
        del gb
        bin_leg = dict(ms_per_step=round(ms_bin, 3),
                       mvis_s=round(n_vis / (ms_bin / 1e3) / 1e6, 1),
                       kernel_ms=round(ph_bin.get("ap_grid_tiled", float("nan")), 3))
        log(f"mosaic chosen form, bin kernels: {ms_bin:.2f} ms")
    # plain FP64 drop-in with the timed kernel set, once
    g64, sw64, ms_plain = plain(main)
    log(f"mosaic plain fp64: {ms_plain:.1f} ms")
    img64 = nat.grid_to_image(g64, [N, N], sum_weight=sw64)
    del g64
    errors = dict(vs_fp64_same_kernels=plane_err(img_main, img64))
    if exact is not None:
        gx, swx, ms_exact = plain(exact)
        img_x = nat.grid_to_image(gx, [N, N], sum_weight=swx)
        del gx
        log(f"mosaic plain fp64, exact per-channel kernels: {ms_exact:.1f} ms")
        errors["chosen_vs_exact_fp64"] = plane_err(img_main, img_x)
        errors["bin_kernels_vs_exact_fp64"] = plane_err(img_bin, img_x)
        errors["fp64_surrogate_vs_exact_fp64"] = plane_err(img64, img_x)
        del img_x, img_bin
    del img64, img_main
    log(f"mosaic errors: {errors}")
    # A-term prediction: the gridded cube degridded back to every sample with
    # the same kernels, fields and maps (complex64 grid -> complex64 samples)
    def degrid(g, single, out=None):
        ck, sp, cfc = main
        return nat.aperture_degrid(d["uvw"], d["field"], fid, g, d["freq"], np.arange(C),
                                   np.arange(P), cfb, cfc, cfp, ck, sp, pg, delta,
                                   [20, 20], single=single, out=out)

    vis_p = degrid(grid, True)
    torch.cuda.synchronize()
    n_dg = max(1, min(args.steps, 3))
    nat.profile_enable(True)
    dg_phases = {}
    t0 = time.time()
    for _ in range(n_dg):
        degrid(grid, True, vis_p)
        for name, ms in nat.profile_read():
            dg_phases[name] = dg_phases.get(name, 0.0) + ms / n_dg
    torch.cuda.synchronize()
    ms_degrid = (time.time() - t0) / n_dg * 1e3
    nat.profile_enable(False)
    # error of the complex64 path against FP64 throughout, on one grid
    g64 = grid.to(torch.complex128)
    vis64 = degrid(g64, False)
    vis32 = degrid(grid, True)
    dg_err = float(torch.linalg.norm((vis32.to(torch.complex128) - vis64).ravel())
                   / torch.linalg.norm(vis64.ravel()))
    del vis64, vis32, vis_p, g64
    log(f"mosaic aperture degrid: {ms_degrid:.2f} ms, phases {dg_phases}, "
        f"rel error vs fp64 {dg_err:.2e}")
    # algorithmic flops: per tap CK*PG (2) + one complex multiply-add per pol (8)
    taps = np.asarray([sup[cfb[b], 0, 0, 0] ** 2 for b in range(cfg.n_baseline)], dtype=np.float64)
    flops = float(taps.sum()) * cfg.n_time * C * (2 + 8 * P)
    # A-term prediction: per tap PGu*PGv (6), times CK (2), the norm (2) and
    # conj(c) g accumulated per pol (8), in FP64
    t_dg = dg_phases.get("aperture_degrid", float("nan")) / 1e3
    dg_tflops = round(float(taps.sum()) * cfg.n_time * C * (10 + 8 * P) / t_dg / 1e12, 2)
    return dict(workload=f"C4: 25-pointing mosaic, 12x7 m + 40x12 m, {cfg.n_time} x 10 s, "
                         f"{C} chan x {P} pol at 230 GHz, {N}^2 at {cfg.cell_arcsec}\", "
                         f"GCF os 20, supports {sorted(set(sup[:, 0, 0, 0].tolist()))}",
                samples=int(n_vis),
                kernels=kern_name,
                value=round(n_vis / (ms_step / 1e3) / 1e6, 1), unit="Mvis/s",
                ms_per_step=round(ms_step, 3), precision=args.precision,
                plain_fp64_mvis_s=round(n_vis / (ms_plain / 1e3) / 1e6, 2),
                plain_fp64_ms=round(ms_plain, 1),
                speedup_vs_plain_fp64=round(ms_plain / ms_step, 1),
                error_vs_fp64=errors,
                bin_kernels=bin_leg,
                plan=plan_leg,
                kernel_ms=round(t_k * 1e3, 3),
                kernel_tflops=round(flops / t_k / 1e12, 2),
                fp32_peak_tflops=157.3, phases_ms=phases, gcf_s=round(t_gcf, 2),
                aperture_degrid=dict(ms=round(ms_degrid, 3),
                                     mvis_s=round(n_vis / (ms_degrid / 1e3) / 1e6, 1),
                                     phases_ms=dg_phases, kernel_fp64_tflops=dg_tflops,
                                     fp64_peak_tflops=78.6,
                                     rel_err_complex64_vs_fp64=dg_err),
                surrogate=surrogate)


def batched_leg(args):
    """C3: 256 channels x 2 pols with uneven per-plane flags (0-60%), 1000
    integrations, 1024^2 at 0.05", support 7, oversampling 100, natural weights.
    The chosen form is one autotuned gridding plan (hipGraph replay of the
    persistent tiled pipeline) over all 512 planes; it is timed against the
    eager single call and against one call per channel."""
    import torch
    from cngi_prototype_amd import _native as nat
    from cngi_prototype_amd import synthetic as syn
    from cngi_prototype_amd.imaging._imaging_utils._gridding_convolutional_kernels import \
        prolate_spheroidal_kernel_1d

    cfg = syn.CONFIGS["C3"]
    N = cfg.image_size
    npad = int(1.2 * N)
    cell = cfg.cell_arcsec * syn.ARCSEC
    delta = np.array([-cell, cell])
    cgk = prolate_spheroidal_kernel_1d(100, 7)
    C, P = cfg.n_chan, cfg.n_pol
    t0 = time.time()
    d = syn.device_dataset(cfg)
    torch.cuda.synchronize()
    n_unflagged = int((d["flag"] == 0).sum().item())
    log(f"batched: generated {cfg.n_vis / 1e6:.1f}M slots ({n_unflagged / 1e6:.1f}M unflagged) "
        f"in {time.time() - t0:.1f}s")
    uvw, vis, w, flag, fr = d["uvw"], d["vis"], d["weight"], d["flag"], d["freq"]
    args_ = (uvw, w, fr, np.arange(C), np.arange(P), cgk, C, [npad, npad], delta, 7, 100,
             nat.GRID_IMAGE, args.precision)
    torch.cuda.synchronize()
    t0 = time.time()
    plan = nat.GridPlan(*args_, vis=vis, flag=flag, autotune=True)
    torch.cuda.synchronize()
    t_plan = time.time() - t0
    info = plan.describe()
    log(f"batched: plan in {t_plan:.2f}s: {info}")
    for _ in range(args.warmup):
        plan.execute()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(args.steps):
        plan.execute()
    torch.cuda.synchronize()
    ms_plan = (time.time() - t0) / args.steps * 1e3
    plan.status()
    grid, sw = plan.grid, plan.sum_weight
    # eager single call (default tile shape) and one call per channel
    g_e = torch.empty_like(grid)
    sw_e = torch.empty_like(sw)
    nat.standard_grid(*args_, vis=vis, flag=flag, grid=g_e, sum_weight=sw_e)
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(args.steps):
        nat.standard_grid(*args_, vis=vis, flag=flag, grid=g_e, sum_weight=sw_e)
    torch.cuda.synchronize()
    ms_eager = (time.time() - t0) / args.steps * 1e3
    del g_e, sw_e
    gc = torch.empty((1, P, npad, npad), dtype=grid.dtype, device="cuda")
    swc = torch.empty((1, P), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.time()
    for ch in range(C):
        nat.standard_grid(uvw, w[:, :, ch:ch + 1].contiguous(), fr[ch:ch + 1], [0],
                          np.arange(P), cgk, 1, [npad, npad], delta, 7, 100, nat.GRID_IMAGE,
                          args.precision, vis=vis[:, :, ch:ch + 1].contiguous(),
                          flag=flag[:, :, ch:ch + 1].contiguous(), grid=gc, sum_weight=swc)
    torch.cuda.synchronize()
    ms_per_chan = (time.time() - t0) * 1e3
    del gc, swc
    log(f"batched: plan {ms_plan:.2f} ms, eager {ms_eager:.2f} ms, per-channel calls "
        f"{ms_per_chan:.1f} ms (incl. slicing)")
    # the plan's gridding kernel alone: the same call eager, in the tuned tile
    # shape, profiled per phase (a graph replay has no phase marks)
    import re
    m_tu = re.search(r"tile (\d+)x32", info)
    m_ln = re.search(r"(\d+) channel lanes", info)
    if m_tu and m_ln:
        nat.set_tile_rows(int(m_tu.group(1)))
        nat.set_bin_lanes(int(m_ln.group(1)))
    g_k = torch.empty_like(grid)
    sw_k = torch.empty_like(sw)
    nat.standard_grid(*args_, vis=vis, flag=flag, grid=g_k, sum_weight=sw_k)
    torch.cuda.synchronize()
    nat.profile_enable(True)
    kph = {}
    for _ in range(args.steps):
        nat.standard_grid(*args_, vis=vis, flag=flag, grid=g_k, sum_weight=sw_k)
        for name, ms in nat.profile_read():
            kph[name] = kph.get(name, 0.0) + ms / args.steps
    nat.profile_enable(False)
    nat.set_tile_rows(0)   # the library defaults
    nat.set_bin_lanes(2)
    del g_k, sw_k
    abytes = algorithmic_bytes(cfg.n_vis, cfg.n_time * cfg.n_baseline, C * P, npad)
    t_k = kph.get("std_grid_tiled", float("nan")) / 1e3
    roofline = dict(bound="hbm", achieved=round(abytes / t_k / 1e9, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(abytes / t_k / 1e9 / HBM_PEAK_GBS, 4),
                    kernel="std_grid_tiled_kernel (tuned tile shape)",
                    algorithmic_bytes=int(abytes), kernel_ms=round(t_k * 1e3, 3),
                    whole_call_gbs=round(abytes / (ms_plan / 1e3) / 1e9, 1),
                    phases_ms={k: round(v, 3) for k, v in kph.items()})
    log(f"batched roofline {roofline}")
    # error against the plain FP64 drop-in on the same inputs, per plane
    g64, sw64 = nat.standard_grid(uvw, w, fr, np.arange(C), np.arange(P), cgk, C, [npad, npad],
                                  delta, 7, 100, nat.GRID_IMAGE, "fp64", vis=vis, flag=flag)
    img64 = nat.grid_to_image(g64, [N, N], sum_weight=sw64)
    del g64
    img = nat.grid_to_image(grid, [N, N], sum_weight=sw)
    n_planes = C * P
    dd = (img - img64).reshape(n_planes, -1)
    rr = img64.reshape(n_planes, -1)
    rel = (torch.linalg.norm(dd, dim=1) / torch.linalg.norm(rr, dim=1)).cpu().numpy()
    del img, img64
    plan.close()
    return dict(workload=f"C3: 50 antennas x {cfg.n_time} integrations, {C} chan x {P} pol, "
                         f"per-plane flags U(0, 0.6), {N}^2 at {cfg.cell_arcsec}\", support 7, "
                         f"oversampling 100",
                slots=int(cfg.n_vis), unflagged=n_unflagged,
                value=round(n_unflagged / (ms_plan / 1e3) / 1e6, 1), unit="Mvis/s",
                ms_per_step=round(ms_plan, 3), precision=args.precision, plan=info,
                plan_create_s=round(t_plan, 2), eager_call_ms=round(ms_eager, 3),
                roofline=roofline,
                per_channel_calls_ms=round(ms_per_chan, 1),
                error_vs_fp64=dict(rel_rms_max=float(rel.max()),
                                   rel_rms_median=float(np.median(rel)), planes=int(n_planes)))


def streamed_leg(args):
    """C5 in full on one GPU: 64 x 12 m antennas, 3875 integrations over 8 h at
    100 GHz, 64 channels x 2 pols (1.0e9 visibilities, 13.2 GB with weights,
    flags and uvw), 49 pointings, 4096^2 at 0.05" (4915^2 grid, 24.7 GB in
    complex64), GCF oversampling 20.  Inputs in pinned host memory, streamed in
    2 GiB chunks of whole integrations into the resident grid in the chosen
    form (FP32-Kahan tiled aperture gridder) with per-channel kernels of the
    learned surrogate; the first n_chan_check channels are compared with the
    plain FP64 GPU gridder (inputs resident) with the same kernels."""
    import torch
    from cngi_prototype_amd import _native as nat
    from cngi_prototype_amd import synthetic as syn
    from cngi_prototype_amd.imaging import make_gridding_convolution_function
    from cngi_prototype_amd.imaging._imaging_utils import _kernel_surrogate as ks
    from cngi_prototype_amd.imaging._imaging_utils._gcf_maps import cf_chan_map, gcf_kernels
    from cngi_prototype_amd.imaging._imaging_utils._stream import pinned_copy, \
        stream_aperture_grid
    from cngi_prototype_amd.xds import Dataset

    cfg = syn.CONFIGS["C5"] if args.c5_chans >= 64 else syn.scaled(syn.CONFIGS["C5"],
                                                                     n_chan=args.c5_chans)
    n_check = min(8, cfg.n_chan)
    N = cfg.image_size
    npad = int(1.2 * N)
    t0 = time.time()
    d = syn.device_mosaic_dataset(cfg)
    torch.cuda.synchronize()
    host = dict(uvw=pinned_copy(d["uvw"]), vis=pinned_copy(d["vis"]),
                weight=pinned_copy(d["weight"]), field=pinned_copy(d["field"]))
    log(f"streamed: {cfg.n_vis / 1e6:.1f}M visibilities to pinned host memory in "
        f"{time.time() - t0:.1f}s")
    pc = syn.phase_center(cfg)
    dims4 = ["time", "baseline", "chan", "pol"]
    xds = Dataset(
        data_vars={"DATA": (dims4, host["vis"]), "UVW": (["time", "baseline", "uvw_index"],
                                                          host["uvw"]),
                   "DATA_WEIGHT": (dims4, host["weight"]), "FLAG": (dims4, host["weight"]),
                   "ANTENNA1": (["baseline"], d["a1"]), "ANTENNA2": (["baseline"], d["a2"]),
                   "FIELD_ID": (["time", "baseline"], host["field"])},
        coords={"chan": (["chan"], d["freq"]), "pol": (["pol"], np.array([9, 12]))},
        attrs={"data_groups": [{"0": {"id": "0", "data": "DATA", "uvw": "UVW", "flag": "FLAG",
                                      "weight": "DATA_WEIGHT"}}]})
    field = Dataset(data_vars={"PHASE_DIR": (["field_id", "poly_id", "d2"], d["fields"][:, None, :])},
                    coords={"field_id": (["field_id"], np.arange(len(d["fields"])))})
    mxds = Dataset(attrs={"xds0": xds, "FIELD": field})
    del d
    torch.cuda.empty_cache()
    os_ = np.array([20, 20])
    ms = np.array([18, 18])
    cell = cfg.cell_arcsec * syn.ARCSEC
    delta = np.array([-cell, cell])
    dish, blockage = np.array([12.0]), np.array([0.75])
    pairs = np.array([[0, 0]])
    fr = syn.freqs(cfg)
    cfc, pb_freq = cf_chan_map(fr, 0.005)
    pb_cell = delta * os_
    npad2 = np.array([npad, npad])
    t0 = time.time()
    wck_b, ck_b, sup_b = gcf_kernels(pb_freq, dish, blockage, pairs, True, pb_cell, npad2, os_,
                                     ms, 0.025)
    half = 0.5 * (fr.max() - fr.min()) + 20e6
    model, tinfo = ks.train_surrogate(dish, blockage, pairs, True, pb_cell, npad2, os_, ms,
                                      sup_b[:, 0, 0, :], pb_freq, half, n_freq=38, steps=20000,
                                      lr=4e-3, batch=4096)
    gcf = make_gridding_convolution_function(
        mxds, dict(function="casa_airy", list_dish_diameters=list(dish),
                   list_blockage_diameters=list(blockage), unique_ant_indx=[0] * cfg.n_ant,
                   phase_center=pc, oversampling=list(os_), max_support=list(ms),
                   surrogate=model, surrogate_check=True),
        dict(image_size=[N, N], cell_size=[cfg.cell_arcsec] * 2, fft_padding=1.2,
             chan_mode="cube", phase_center=pc), {"xds": "xds0"})
    torch.cuda.synchronize()
    t_gcf = time.time() - t0
    rep = gcf.attrs["surrogate_report"]
    sup = np.asarray(gcf["SUPPORT"].values)
    log(f"streamed: GCF + surrogate in {t_gcf:.1f}s (train {tinfo['train_s']}s), supports "
        f"{sorted(set(sup[..., 0].ravel().tolist()))}, surrogate max err "
        f"{rep['max_err_surrogate']:.2e} vs bin {rep['max_err_bin']:.2e}, "
        f"{len(rep['fallback'])} fallbacks")
    g = dict(field_id=np.asarray(gcf.coords["field_id"].values),
             cf_baseline_map=np.asarray(gcf["CF_BASELINE_MAP"].values),
             cf_chan_map=np.asarray(gcf["CF_CHAN_MAP"].values),
             cf_pol_map=np.asarray(gcf["CF_POL_MAP"].values), support=sup,
             kernel=gcf["CONV_KERNEL"].data, phase_gradient=gcf["PHASE_GRADIENT"].data)
    C, P = cfg.n_chan, cfg.n_pol
    grid = torch.empty((C, P, npad, npad), dtype=torch.complex64, device="cuda")
    sw = torch.empty((C, P), dtype=torch.float64, device="cuda")
    runs = []
    for _ in range(max(1, args.warmup) + 2):
        grid, sw, st = stream_aperture_grid(host, fr, np.arange(C), np.arange(P), g, C,
                                            [npad, npad], delta, os_, nat.GRID_IMAGE,
                                            args.precision, grid=grid, sum_weight=sw,
                                            chunk_bytes=args.c5_chunk_mb << 20)
        runs.append(st)
    st = min(runs[max(1, args.warmup):], key=lambda r: r["wall_s"])
    log(f"streamed chosen form: {st}")
    # plain FP64 on resident inputs, same kernels, on the first n_check channels
    dv = {k: (v[:, :, :n_check] if v.dim() == 4 else v).contiguous().cuda()
          for k, v in host.items()}
    torch.cuda.synchronize()
    t0 = time.time()
    g64, sw64 = nat.aperture_grid(dv["uvw"], dv["weight"], dv["field"], g["field_id"],
                                  fr[:n_check], np.arange(n_check), np.arange(P),
                                  g["cf_baseline_map"], g["cf_chan_map"][:n_check],
                                  g["cf_pol_map"], g["kernel"], sup, g["phase_gradient"],
                                  n_check, [npad, npad], delta, os_, nat.GRID_IMAGE, "fp64",
                                  vis=dv["vis"])
    torch.cuda.synchronize()
    t64 = time.time() - t0
    n_vis_check = int(dv["weight"].numel())
    del dv
    img64 = nat.grid_to_image(g64, [N, N], sum_weight=sw64)
    del g64
    img = nat.grid_to_image(grid[:n_check], [N, N], sum_weight=sw[:n_check])
    n_planes = n_check * P
    dd = (img - img64).reshape(n_planes, -1)
    rr = img64.reshape(n_planes, -1)
    rel = (torch.linalg.norm(dd, dim=1) / torch.linalg.norm(rr, dim=1)).cpu().numpy()
    del img, img64, grid
    t64_per_vis = t64 / n_vis_check
    return dict(workload=f"C5: {cfg.n_chan} of 64 channels, {cfg.n_ant} x 12 m, "
                         f"{cfg.n_time} integrations, 2 pol, 49 pointings, {N}^2 at "
                         f"{cfg.cell_arcsec}\", GCF os 20, supports "
                         f"{sorted(set(sup[..., 0].ravel().tolist()))}",
                visibilities=int(cfg.n_vis),
                value=round(cfg.n_vis / st["wall_s"] / 1e6, 1), unit="Mvis/s",
                note=f"host-resident inputs, PCIe-inclusive (streamed in {args.c5_chunk_mb} MiB "
                     f"chunks)",
                stream=st, plain_fp64_resident_s=round(t64, 2),
                plain_fp64_channels=n_check,
                plain_fp64_mvis_s=round(n_vis_check / t64 / 1e6, 1),
                speedup_vs_plain_fp64=round(t64_per_vis * cfg.n_vis / st["wall_s"], 1),
                precision=args.precision,
                surrogate=dict(max_tap_err=rep["max_err_surrogate"],
                               max_tap_err_bin=rep["max_err_bin"],
                               fallbacks=len(rep["fallback"]), kernels=C,
                               train_s=tinfo["train_s"]),
                error_vs_fp64=dict(rel_rms_max=float(rel.max()),
                                   rel_rms_median=float(np.median(rel)), planes=int(n_planes),
                                   channels=n_check))


def predict_leg(args):
    """Visibility prediction at the C2 shapes: a model cube of 64 channels x
    4 pols at 2048^2 (2000 random point sources per plane, Jy/pixel) is
    transformed to the 2457^2 grid (correction, padding, LDS FFTs) and degridded
    at all 627.2M samples, complex64 grid and visibilities with FP64 taps and
    sums.  Compared with the FP64 path on the first 4 channels.  Roofline of
    the degrid kernel: the grid read once + uvw + the visibilities written."""
    import torch
    from cngi_prototype_amd import _native as nat
    from cngi_prototype_amd import synthetic as syn
    from cngi_prototype_amd.imaging._imaging_utils._gridding_convolutional_kernels import \
        cropped_correction_factors, prolate_spheroidal_kernel_1d

    cfg = syn.CONFIGS["C2"]
    N = cfg.image_size
    npad = int(1.2 * N)
    cell = cfg.cell_arcsec * syn.ARCSEC
    delta = np.array([-cell, cell])
    C, P = cfg.n_chan, cfg.n_pol
    enu = syn.antenna_enu(cfg)
    uvw_h = syn.uvw_tracks(enu, cfg.lat_deg, cfg.dec_deg, cfg.n_time, cfg.dt)
    uvw = torch.as_tensor(uvw_h).cuda()
    T, B = uvw_h.shape[0], uvw_h.shape[1]
    freq = syn.freqs(cfg)
    rng = np.random.default_rng(cfg.seed + 31)
    model = torch.zeros((C, P, N, N), dtype=torch.float64, device="cuda")
    n_src = 2000
    li = torch.as_tensor(rng.integers(N // 8, N - N // 8, size=n_src), device="cuda")
    mi = torch.as_tensor(rng.integers(N // 8, N - N // 8, size=n_src), device="cuda")
    fl = torch.as_tensor(rng.uniform(1e-3, 1.0, size=n_src), device="cuda")
    model[:, :, li, mi] = fl
    cgk = prolate_spheroidal_kernel_1d(100, 7)
    cl, cm = cropped_correction_factors([npad, npad], [N, N])
    grid = torch.empty((C, P, npad, npad), dtype=torch.complex64, device="cuda")
    vis = torch.empty((T, B, C, P), dtype=torch.complex64, device="cuda")

    def step():
        ph = {}
        nat.model_to_grid(model, [npad, npad], cl, cm, single=True, out=grid)
        ph.update(dict(nat.profile_read()))
        nat.standard_degrid(uvw, grid, freq, np.arange(C), np.arange(P), cgk, delta, 7, 100,
                            single=True, out=vis)
        ph.update(dict(nat.profile_read()))
        return ph

    nat.profile_enable(True)
    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    phases = {}
    t0 = time.time()
    for _ in range(args.steps):
        for k, v in step().items():
            phases[k] = phases.get(k, 0.0) + v / args.steps
    torch.cuda.synchronize()
    ms_step = (time.time() - t0) / args.steps * 1e3
    nat.profile_enable(False)
    # FP64 path on the first 4 channels
    nc = 4
    g64 = nat.model_to_grid(model[:nc].contiguous(), [npad, npad], cl, cm, single=False)
    v64 = nat.standard_degrid(uvw, g64, freq[:nc], np.arange(nc), np.arange(P), cgk, delta, 7,
                              100, single=False)
    d = (vis[:, :, :nc].to(torch.complex128) - v64)
    rel = float(torch.linalg.norm(d) / torch.linalg.norm(v64))
    del g64, v64, d, model, grid
    n_samples = T * B * C * P
    t_dg = phases.get("standard_degrid", float("nan")) / 1e3
    # the whole degrid call: binning passes (tiled path) and the kernel
    t_call = sum(v for k, v in phases.items()
                 if k == "standard_degrid" or k.startswith("degrid_bin")) / 1e3
    alg = C * P * npad * npad * 8 + T * B * 24 + n_samples * 8
    gbs = alg / t_dg / 1e9
    log(f"predict: step {ms_step:.2f} ms, phases {phases}")
    return dict(workload=f"C2 shapes: model cube {C} chan x {P} pol at {N}^2 (2000 point "
                         f"sources per plane) -> {npad}^2 grid -> {n_samples / 1e6:.1f}M "
                         f"predicted samples", samples=int(n_samples),
                value=round(n_samples / (ms_step / 1e3) / 1e6, 1), unit="Mvis/s",
                ms_per_step=round(ms_step, 3),
                phases_ms={k: round(v, 3) for k, v in phases.items()},
                degrid_roofline=dict(bound="hbm", achieved=round(gbs, 1), peak=8000.0,
                                     unit="GB/s", frac=round(gbs / 8000.0, 4),
                                     algorithmic_bytes=int(alg),
                                     kernel_ms=round(t_dg * 1e3, 3),
                                     call_ms=round(t_call * 1e3, 3),
                                     call_frac=round(alg / t_call / 1e9 / 8000.0, 4)),
                precision="complex64 grid and visibilities, FP64 taps",
                rel_err_vs_fp64_4chan=rel)


def combined(n_std, ms_std, mosaic):
    """The metric's combined figure: the standard (C2) and mosaic (C4) steps'
    samples over their summed step times."""
    if not mosaic or "ms_per_step" not in mosaic:
        return None
    n_mos = mosaic["samples"]
    ms = ms_std + mosaic["ms_per_step"]
    return dict(value=round((n_std + n_mos) / (ms / 1e3) / 1e6, 1), unit="Mvis/s",
                samples=int(n_std + n_mos), ms=round(ms, 3),
                parts="C2 standard step (unflagged samples) + C4 mosaic step")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-check", action="store_true")
    ap.add_argument("--no-mosaic", action="store_true",
                    help="skip the C4 mosaic leg (reported under 'mosaic')")
    ap.add_argument("--mosaic-only", action="store_true",
                    help="run only the C4 mosaic leg and print its record")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the C3 batched-cube leg (reported under 'batched')")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the C3 batched-cube leg and print its record")
    ap.add_argument("--no-streamed", action="store_true",
                    help="skip the C5 streamed-mosaic leg (reported under 'streamed')")
    ap.add_argument("--streamed-only", action="store_true",
                    help="run only the C5 streamed-mosaic leg and print its record")
    ap.add_argument("--no-predict", action="store_true",
                    help="skip the C2-shape visibility prediction leg (reported under 'predict')")
    ap.add_argument("--predict-only", action="store_true",
                    help="run only the prediction leg and print its record")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the C1 leg (reported under 'c1')")
    ap.add_argument("--c1-only", action="store_true",
                    help="run only the C1 leg and print its record")
    ap.add_argument("--mosaic-grid-only", action="store_true",
                    help="with --mosaic-only: only the timed gridding steps (no plain FP64, "
                         "prediction or surrogate), for profiling")
    ap.add_argument("--c5-chans", type=int, default=64,
                    help="channels of the C5 streamed leg (64: the whole cube)")
    ap.add_argument("--c5-chunk-mb", type=int, default=2048,
                    help="visibility bytes per host->device chunk of the C5 leg (MiB)")
    ap.add_argument("--no-plain-fp64", action="store_true",
                    help="skip the one call of the plain FP64 drop-in on the main workload")
    ap.add_argument("--cpu-chans", type=int, default=6,
                    help="channels of the time-chunked and one-thread CPU baseline samples")
    ap.add_argument("--cpu-cube-chans", type=int, default=16,
                    help="channels (one task each) of the cube-mode CPU baseline sample")
    ap.add_argument("--config", default="C2")
    ap.add_argument("--precision", default="fp32_kahan", choices=["fp32_kahan", "fp32"])
    ap.add_argument("--tile-rows", type=int, default=0, choices=[0, 8, 16, 32],
                    help="uv-tile rows of the gridder (0: library default, 32 rows at this size)")
    ap.add_argument("--tile-waves", type=int, default=0, choices=[0, 1, 2, 4],
                    help="wavefronts per tile of the gridder, splitting its records "
                         "(0: library default)")
    ap.add_argument("--bin-lanes", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="channels per binning wavefront (0: library default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the steps, run the grid's ifftshift (a plain copy kernel of "
                         "known size) so that a counter run sees a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    # the imaging API reports its parameter defaults on stdout, as the reference's
    # does; keep stdout for the one JSON record
    out, sys.stdout = sys.stdout, sys.stderr

    import torch
    from cngi_prototype_amd import _native as nat
    from cngi_prototype_amd import synthetic as syn
    from cngi_prototype_amd.imaging._imaging_utils._gridding_convolutional_kernels import \
        cropped_correction_factors, prolate_spheroidal_kernel_1d

    torch.cuda.set_device(0)
    if args.tile_waves:
        nat.set_tile_waves(args.tile_waves)
    if args.bin_lanes:
        nat.set_bin_lanes(args.bin_lanes)
    if args.mosaic_only:
        print(json.dumps({"mosaic": mosaic_leg(args)}), file=out, flush=True)
        return
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(args)}), file=out, flush=True)
        return
    if args.streamed_only:
        print(json.dumps({"streamed": streamed_leg(args)}), file=out, flush=True)
        return
    if args.predict_only:
        print(json.dumps({"predict": predict_leg(args)}), file=out, flush=True)
        return
    if args.c1_only:
        print(json.dumps({"c1": c1_leg(args)}), file=out, flush=True)
        return
    if args.tile_rows:
        nat.set_tile_rows(args.tile_rows)
    cfg = syn.CONFIGS[args.config]
    N = cfg.image_size
    npad = int(1.2 * N)
    cell = cfg.cell_arcsec * syn.ARCSEC
    delta = np.array([-cell, cell])
    cgk = prolate_spheroidal_kernel_1d(100, 7)
    C, P = cfg.n_chan, cfg.n_pol

    t0 = time.time()
    data = syn.device_dataset(cfg)
    torch.cuda.synchronize()
    log(f"generated {cfg.n_vis / 1e6:.1f}M samples in HBM in {time.time() - t0:.1f}s")
    uvw, vis, w, flag = data["uvw"], data["vis"], data["weight"], data["flag"]
    fr = data["freq"]
    chan_map = np.arange(C)
    pol_map = np.arange(P)

    # Briggs 0.5 imaging weights (density grid + factors + degrid), on the GPU:
    # a first run that allocates the outputs, then a timed one with its phases
    def briggs(dens=None, dsw=None, iw=None):
        dens, dsw = nat.standard_grid(uvw, w, fr, chan_map, pol_map, np.ones(1), C, [N, N],
                                      delta, 1, 0, nat.GRID_DENSITY, "fp64_tiled", flag=flag,
                                      grid=dens, sum_weight=dsw)
        ph = dict(nat.profile_read())
        fac = nat.briggs_factors(dens, dsw, "briggs", 0.5)
        ph.update(dict(nat.profile_read()))
        iw = nat.imaging_weight_degrid(uvw, w, dens, fac, fr, chan_map, pol_map, delta,
                                       flag=flag, out=iw)
        ph.update(dict(nat.profile_read()))
        return dens, dsw, iw, ph

    torch.cuda.synchronize()
    tb = time.time()
    dens, dsw, iw, _ = briggs()
    torch.cuda.synchronize()
    t_briggs_first = time.time() - tb
    nat.profile_enable(True)
    tb = time.time()
    dens, dsw, iw, briggs_ph = briggs(dens, dsw, iw)
    torch.cuda.synchronize()
    t_briggs = time.time() - tb
    nat.profile_enable(False)
    del dens
    data["imaging_weight"] = iw
    log(f"briggs weights: {t_briggs * 1e3:.1f} ms ({t_briggs_first * 1e3:.1f} ms with the "
        f"first allocation), phases {briggs_ph}")

    n_rows = cfg.n_time * cfg.n_baseline
    n_unflagged = int((flag == 0).sum().item())
    n_planes = C * P
    grid = torch.empty((C, P, npad, npad), dtype=torch.complex64, device="cuda")
    sw = torch.empty((C, P), dtype=torch.float64, device="cuda")

    def step():
        nat.standard_grid(uvw, iw, fr, chan_map, pol_map, cgk, C, [npad, npad], delta, 7, 100,
                          nat.GRID_IMAGE, args.precision, vis=vis, flag=flag, grid=grid,
                          sum_weight=sw)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    nat.profile_enable(True)
    phases = {}
    t0 = time.time()
    for _ in range(args.steps):
        step()
        for name, ms in nat.profile_read():
            phases.setdefault(name, []).append(ms)
    torch.cuda.synchronize()
    elapsed = time.time() - t0
    nat.profile_enable(False)
    if args.copy_calibration:
        # ifftshift_kernel<float, float, 2> is a plain permuted copy of the
        # complex64 grid: planes * Npad^2 * 8 bytes in and the same out
        log(f"copy calibration: ifftshift of {n_planes * npad * npad * 8} bytes in and out")
        nat.set_fft_mode(0)   # the rocFFT path runs the ifftshift copy kernel
        nat.grid_to_image(grid, [N, N], sum_weight=sw)
        torch.cuda.synchronize()
        nat.set_fft_mode(1)
    ms_step = elapsed / args.steps * 1e3
    value = n_unflagged / (elapsed / args.steps) / 1e6
    phase_ms = {k: float(np.mean(v)) for k, v in phases.items()}
    log(f"step {ms_step:.2f} ms, phases {phase_ms}")

    abytes = algorithmic_bytes(cfg.n_vis, n_rows, n_planes, npad)
    t_kernel = phase_ms.get("std_grid_tiled", float("nan")) / 1e3
    achieved = abytes / t_kernel / 1e9
    traffic = measured_traffic(args.precision, args.config)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    kernel="std_grid_tiled_kernel<IMAGE,4,float,%s>" % (
                        "true" if args.precision == "fp32_kahan" else "false"),
                    algorithmic_bytes=int(abytes), kernel_ms=round(t_kernel * 1e3, 3),
                    whole_call_gbs=round(abytes / (ms_step / 1e3) / 1e9, 1))

    # error of the chosen form against FP64 on the same inputs, per plane
    err = None
    if not args.no_error_check:
        te = time.time()
        corr = cropped_correction_factors([npad, npad], [N, N])
        step()
        img32 = nat.grid_to_image(grid, [N, N], sum_weight=sw, corr=corr)
        g64, sw64 = nat.standard_grid(uvw, iw, fr, chan_map, pol_map, cgk, C, [npad, npad],
                                      delta, 7, 100, nat.GRID_IMAGE, "fp64_tiled", vis=vis,
                                      flag=flag)
        img64 = nat.grid_to_image(g64, [N, N], sum_weight=sw64, corr=corr)
        # split the error: gridding alone (FP32 grid through the FP64 FFT)
        g32 = torch.empty_like(g64)
        step()
        g32.copy_(grid)
        img32_f64fft = nat.grid_to_image(g32, [N, N], sum_weight=sw, corr=corr)
        del g32
        grid_rel = float((torch.linalg.norm((img32_f64fft - img64).reshape(n_planes, -1), dim=1) /
                          torch.linalg.norm(img64.reshape(n_planes, -1), dim=1)).max().item())
        del img32_f64fft, g64
        d = (img32 - img64).reshape(n_planes, -1)
        r = img64.reshape(n_planes, -1)
        rel = (torch.linalg.norm(d, dim=1) / torch.linalg.norm(r, dim=1)).cpu().numpy()
        pk = (d.abs().max(dim=1).values / r.abs().max(dim=1).values).cpu().numpy()
        err = dict(rel_rms_max=float(rel.max()), rel_rms_median=float(np.median(rel)),
                   max_abs_over_peak=float(pk.max()), planes=int(n_planes),
                   sum_weight_rel=float(((sw - sw64).abs() / sw64.abs()).max().item()),
                   gridding_only_rel_rms_max=grid_rel)
        del img32, img64
        torch.cuda.synchronize()
        log(f"error check {time.time() - te:.1f}s: {err}")

    # the imaging transform of the gridded cube (make_image.py:116-130): the LDS
    # row FFTs (ifftshift folded in), then the columns in two in-place steps
    # whose second writes the cropped, corrected image, profiled per phase; the
    # transform's algorithmic traffic is two passes over the complex64 grid,
    # each reading and writing it once.  rocFFT is timed beside it.
    corr = cropped_correction_factors([npad, npad], [N, N])
    nat.grid_to_image(grid, [N, N], sum_weight=sw, corr=corr)
    torch.cuda.synchronize()
    nat.profile_enable(True)
    n_fft = max(1, min(args.steps, 3))
    fft_ph = {}
    t0 = time.time()
    for _ in range(n_fft):
        img_t = nat.grid_to_image(grid, [N, N], sum_weight=sw, corr=corr)
        for name, ms in nat.profile_read():
            fft_ph[name] = fft_ph.get(name, 0.0) + ms / n_fft
        del img_t
    torch.cuda.synchronize()
    ms_g2i = (time.time() - t0) / n_fft * 1e3
    nat.profile_enable(False)
    grid_bytes = n_planes * npad * npad * 8
    # every phase but a separate crop (the four-step column pass writes the
    # cropped, corrected image itself)
    fft_phases = [k for k in fft_ph if k != "crop_correct"]
    t_fft = sum(fft_ph[k] for k in fft_phases) / 1e3 if fft_phases else float("nan")
    # the same transform through rocFFT (ifftshift kernel + library FFT)
    nat.set_fft_mode(0)
    nat.grid_to_image(grid, [N, N], sum_weight=sw, corr=corr)
    torch.cuda.synchronize()
    nat.profile_enable(True)
    roc_ph = {}
    for _ in range(n_fft):
        img_t = nat.grid_to_image(grid, [N, N], sum_weight=sw, corr=corr)
        for name, ms in nat.profile_read():
            roc_ph[name] = roc_ph.get(name, 0.0) + ms / n_fft
        del img_t
    torch.cuda.synchronize()
    nat.profile_enable(False)
    nat.set_fft_mode(1)
    fft = dict(planes=int(n_planes), length=f"{npad} x {npad} (3^3 7 13 per axis)",
               grid_to_image_ms=round(ms_g2i, 3),
               phases_ms={k: round(v, 3) for k, v in fft_ph.items()},
               fft_ms=round(t_fft * 1e3, 3),
               fft_phases=fft_phases,
               rocfft_phases_ms={k: round(v, 3) for k, v in roc_ph.items()},
               fft_roofline=dict(bound="hbm", achieved=round(4 * grid_bytes / t_fft / 1e9, 1),
                                 peak=HBM_PEAK_GBS, unit="GB/s",
                                 frac=round(4 * grid_bytes / t_fft / 1e9 / HBM_PEAK_GBS, 4),
                                 algorithmic_bytes=int(4 * grid_bytes)),
               make_image_ms=round(ms_step + ms_g2i, 3))
    log(f"imaging transform {fft}")

    # the plain FP64 drop-in (reference loop per sample, FP64 atomics into a
    # complex128 grid) on the same inputs, once
    plain = None
    if not args.no_plain_fp64:
        del grid
        torch.cuda.empty_cache()
        torch.cuda.synchronize()
        t0 = time.time()
        g64, sw64 = nat.standard_grid(uvw, iw, fr, chan_map, pol_map, cgk, C, [npad, npad],
                                      delta, 7, 100, nat.GRID_IMAGE, "fp64", vis=vis, flag=flag)
        torch.cuda.synchronize()
        ms_plain = (time.time() - t0) * 1e3
        del g64, sw64
        torch.cuda.empty_cache()
        plain = dict(ms=round(ms_plain, 1), mvis_s=round(n_unflagged / ms_plain * 1e-3, 1),
                     speedup_of_chosen_form=round(ms_plain / ms_step, 1))
        log(f"plain fp64 drop-in at {cfg.name}: {plain}")
        grid = torch.empty((C, P, npad, npad), dtype=torch.complex64, device="cuda")

    cpu = cpu_dask = cpu1 = None
    if not args.no_cpu_baseline:
        try:
            cpu, cpu_dask, cpu1 = cpu_baseline(cfg, data, npad, delta, cgk, args.cpu_chans,
                                               args.cpu_cube_chans)
        except Exception as e:  # report, never fake
            log(f"cpu baseline failed: {e!r}")

    mosaic = None
    if not args.no_mosaic:
        del data, uvw, vis, w, flag, iw, grid, sw
        torch.cuda.empty_cache()
        try:
            mosaic = mosaic_leg(args)
            log(f"mosaic {mosaic}")
        except Exception as e:  # report, never fake
            log(f"mosaic leg failed: {e!r}")
            mosaic = {"error": repr(e)}

    batched = None
    if not args.no_batched:
        torch.cuda.empty_cache()
        try:
            batched = batched_leg(args)
            log(f"batched {batched}")
        except Exception as e:  # report, never fake
            log(f"batched leg failed: {e!r}")
            batched = {"error": repr(e)}

    streamed = None
    if not args.no_streamed:
        torch.cuda.empty_cache()
        try:
            streamed = streamed_leg(args)
            log(f"streamed {streamed}")
        except Exception as e:  # report, never fake
            log(f"streamed leg failed: {e!r}")
            streamed = {"error": repr(e)}

    c1 = None
    if not args.no_c1:
        torch.cuda.empty_cache()
        try:
            c1 = c1_leg(args)
            log(f"c1 {c1}")
        except Exception as e:  # report, never fake
            log(f"c1 leg failed: {e!r}")
            c1 = {"error": repr(e)}

    predict = None
    if not args.no_predict:
        torch.cuda.empty_cache()
        try:
            predict = predict_leg(args)
            log(f"predict {predict}")
        except Exception as e:  # report, never fake
            log(f"predict leg failed: {e!r}")
            predict = {"error": repr(e)}

    line = {
        "metric": "Mvis/s gridded (std+mosaic) at <=1e-5 rel err vs FP64; % of HBM roofline",
        "value": round(value, 1),
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": args.precision,
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: standard gridding, Briggs 0.5 weights, "
                               f"{cfg.n_ant} antennas x {cfg.n_time} integrations, "
                               f"{C} chan x {P} pol, {N}^2 image (padded {npad}^2), "
                               f"support 7, oversampling 100, 10% flags",
                   "samples": int(cfg.n_vis), "unflagged": n_unflagged,
                   "planes": int(n_planes),
                   "form": f"{args.precision} (uv-tiled bins, a workgroup of "
                           f"{args.tile_waves or 4} wavefront(s) per "
                           f"{args.tile_rows or 32}x32-cell tile splitting its records, "
                           f"lane-owned residue classes, FP32 register runs added to an "
                           f"LDS subgrid (FP64 cells for fp32_kahan))"},
        "roofline": roofline,
        "combined_std_mosaic": combined(n_unflagged, ms_step, mosaic),
        "imaging_transform": fft,
        "plain_fp64": plain,
        "cpu_baseline": cpu,
        "cpu_baseline_dask_time_chunks": cpu_dask,
        "cpu_baseline_1thread": cpu1,
        "error_vs_fp64": err,
        "phases_ms": phase_ms,
        "briggs_ms": round(t_briggs * 1e3, 2),
        "briggs_first_call_ms": round(t_briggs_first * 1e3, 2),
        "briggs_phases_ms": {k: round(v, 3) for k, v in briggs_ph.items()},
        "tile_rows": args.tile_rows or 32,
        "c1": c1,
        "mosaic": mosaic,
        "batched": batched,
        "streamed": streamed,
        "predict": predict,
    }
    print(json.dumps(line), file=out, flush=True)


if __name__ == "__main__":
    main()
