"""Host-streamed aperture gridding.

The visibilities, weights, uvw and field ids of a cube live in pinned host
memory; they are copied to the GPU in chunks of whole integrations (about
`chunk_bytes` of visibilities each, 2 GiB by default: large enough that a
chunk's samples fill the gridder's tiles densely, small against HBM) on a copy
stream while the previous chunk is gridded, and every chunk is gridded into
the same device grid (accumulate).  Two device buffers alternate.  This replaces the reference's dask graph over
time chunks with private grids and a tree-sum (_aperture_grid.py:25-142):
the grid stays resident in HBM and each chunk adds to it in place.
"""
import time

import numpy as np

from ... import _native


def pinned_copy(t):
    """A pinned host copy of a tensor (device or host)."""
    import torch
    out = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    out.copy_(t)
    return out


def stream_aperture_grid(host, freq, chan_map, pol_map, gcf, n_imag_chan, n_uv, delta_lm,
                         oversampling, mode, precision, chunk_bytes=2 << 30, grid=None,
                         sum_weight=None):
    """host: dict of pinned CPU tensors uvw [T, B, 3] float64, weight [T, B, C, P]
    float32, field [T, B] int32 and (image mode) vis [T, B, C, P] complex64.
    gcf: dict with field_id, cf_baseline_map, cf_chan_map, cf_pol_map, support
    (host arrays) and kernel, phase_gradient (device tensors).
    Returns (grid, sum_weight, stats)."""
    import torch
    T, B, C, P = host["weight"].shape
    per_int = B * C * P * (8 if "vis" in host else 4)
    t_chunk = int(max(1, min(T, chunk_bytes // max(1, per_int))))
    bounds = [(t0, min(T, t0 + t_chunk)) for t0 in range(0, T, t_chunk)]
    names = [n for n in ("uvw", "vis", "weight", "field") if n in host]
    dev = [{n: torch.empty((t_chunk,) + tuple(host[n].shape[1:]), dtype=host[n].dtype,
                           device="cuda") for n in names} for _ in range(2)]
    copy_stream = torch.cuda.Stream()
    copied = [torch.cuda.Event(), torch.cuda.Event()]
    compute = torch.cuda.current_stream()
    consumed = [torch.cuda.Event(), torch.cuda.Event()]

    def issue(k):
        t0, t1 = bounds[k]
        b = k % 2
        copy_stream.wait_event(consumed[b])
        with torch.cuda.stream(copy_stream):
            for n in names:
                dev[b][n][:t1 - t0].copy_(host[n][t0:t1], non_blocking=True)
            copied[b].record(copy_stream)

    torch.cuda.synchronize()
    t_start = time.time()
    for b in range(2):
        consumed[b].record(compute)
    issue(0)
    for k, (t0, t1) in enumerate(bounds):
        if k + 1 < len(bounds):
            issue(k + 1)
        b = k % 2
        compute.wait_event(copied[b])
        n = t1 - t0
        d = {name: dev[b][name][:n] for name in names}
        grid, sum_weight = _native.aperture_grid(
            d["uvw"], d["weight"], d["field"], gcf["field_id"], freq, chan_map, pol_map,
            gcf["cf_baseline_map"], gcf["cf_chan_map"], gcf["cf_pol_map"], gcf["kernel"],
            gcf["support"], gcf["phase_gradient"], n_imag_chan, n_uv, delta_lm, oversampling,
            mode, precision, vis=d.get("vis"), grid=grid, sum_weight=sum_weight,
            accumulate=k > 0)
        consumed[b].record(compute)
    torch.cuda.synchronize()
    wall = time.time() - t_start
    moved = sum(host[n].numel() * host[n].element_size() for n in names)
    return grid, sum_weight, dict(chunks=len(bounds), integrations_per_chunk=t_chunk,
                                  wall_s=wall, host_bytes=int(moved),
                                  host_gb_s=round(moved / wall / 1e9, 2))
