"""Learned surrogate of the mosaic A-term convolution kernels.

The reference computes one exact kernel per dish-type pair and per bin of
channels (chan_tolerance_factor, make_gridding_convolution_function.py:
536-564) and uses it for every channel of the bin.  The surrogate gives every
channel its own kernel at the cost of one small network evaluation: an MLP
(4 dense layers of width 256; bf16 matrix-core products with FP32
accumulation, FP32 master weights, Adam) predicts the change of the sampled
kernel taps between the bin's frequency and the channel's frequency.

Inputs of one example (10 features): the two dish diameters / 12 m, their
blockage diameters, 100 (f/f_bin - 1), 10 (f_bin/f_ref - 1), cos and sin of
twice the parallactic angle (symmetric Airy patterns do not depend on it; the
training set draws it at random so the network learns that), and the
sub-cell offset pair (ou/os, ov/os).  Output: the R x R tap block (R = the
largest support) of that offset, as the difference from the bin's exact block
in units of the bin kernel's peak, divided by the frequency offset feature
100 (f/f_bin - 1).  The kernel of a channel is the bin's kernel stretched by
f/f_bin, so the change is to first order proportional to that offset: the
network learns the smooth rate of change (and its slow dependence on the
offset), and the product with the offset is exact where the change vanishes.

The exact kernel is the fallback and the measure of error: training and validation
targets come from the exact GPU kernels, and with check=True every
per-channel kernel is compared with the exact one over the taps the gridder
reads (the support window at every sub-cell offset) and replaced by it when
its largest tap error there exceeds 1e-4 of the peak.  Outside the window,
which no gridder reads, a surrogate kernel keeps the bin kernel's samples.
"""
import math
import time

import numpy as np

from ... import _native
from ._gcf_maps import gcf_kernels

N_IN = 10
TOL_PEAK = 1e-4


class KernelSurrogate:
    """The dense network and its training (all tensors on the device, FP32)."""

    def __init__(self, n_in, n_out, width=256, depth=4, seed=0):
        import torch
        dims = [n_in] + [width] * (depth - 1) + [n_out]
        gen = torch.Generator().manual_seed(seed)
        # parameters, Adam moments and gradients each live in one flat buffer
        # (W_1, b_1, W_2, ...), so one Adam update covers the whole network
        sizes = []
        for i in range(depth):
            sizes += [(dims[i], dims[i + 1]), (dims[i + 1],)]
        total = sum(math.prod(sz) for sz in sizes)
        self.flat = torch.zeros(total, dtype=torch.float32, device="cuda")
        self.m_flat = torch.zeros_like(self.flat)
        self.v_flat = torch.zeros_like(self.flat)
        self.g_flat = torch.zeros_like(self.flat)

        def views(buf):
            out, o = [], 0
            for sz in sizes:
                n = math.prod(sz)
                out.append(buf[o:o + n].view(sz))
                o += n
            return out
        pv = views(self.flat)
        self.W, self.b = pv[0::2], pv[1::2]
        for i in range(depth):
            lim = math.sqrt(6.0 / (dims[i] + dims[i + 1]))
            w = (torch.rand((dims[i], dims[i + 1]), generator=gen, dtype=torch.float32) * 2 - 1)
            self.W[i].copy_((w * lim).cuda())
        self.m = views(self.m_flat)
        self.v = views(self.v_flat)
        self.grads = views(self.g_flat)
        self.steps = 0
        self.dims = dims
        self.meta = {}
        self.x_mean = torch.zeros(n_in, dtype=torch.float32, device="cuda")
        self.x_std = torch.ones(n_in, dtype=torch.float32, device="cuda")
        self.t_scale = 1.0
        # predictions are multiplied by this (unscaled) input feature, if set
        self.scale_feature = None

    def params(self):
        out = []
        for w, b in zip(self.W, self.b):
            out += [w, b]
        return out

    def forward(self, x):
        """Activations [x, a_1, ..., a_L]; hidden layers tanh, last linear."""
        acts = [x]
        for i, (w, b) in enumerate(zip(self.W, self.b)):
            last = i == len(self.W) - 1
            acts.append(_native.bf16_gemm(acts[-1], w, bias=b, act=0 if last else 1))
        return acts

    def loss_and_grads(self, x, t, out=None):
        """Mean squared error (device scalar) and gradients [dW1, db1, ...]
        (written into `out`, a list of tensors of the parameters' shapes, when
        given).  The tanh derivative of each hidden layer is applied in the
        epilogue of the GEMM that propagates the gradient to it."""
        acts = self.forward(x)
        dz, loss = _native.mse_grad(acts[-1], t)
        grads = [None] * (2 * len(self.W)) if out is None else out
        for i in range(len(self.W) - 1, -1, -1):
            grads[2 * i] = _native.bf16_gemm(acts[i], dz, trans_a=True,
                                             out=None if out is None else out[2 * i])
            grads[2 * i + 1] = _native.col_sum(dz, out=None if out is None else out[2 * i + 1])
            if i > 0:
                dz = _native.bf16_gemm(dz, self.W[i], trans_b=True, act=2, aux=acts[i])
        return loss, grads

    def train(self, x, t, steps=3000, batch=4096, lr=2e-3, lr_min=2e-5, seed=1, log_every=0,
              graph=True):
        """Adam with a cosine learning-rate decay over `steps` minibatches drawn
        on the device.  The step (sampling, forward, backward, schedule, Adam)
        keeps its step number in device memory, so with graph=True it runs
        once eagerly and is then captured as a HIP graph and replayed; with
        graph=False every step is launched eagerly (the same kernels and
        arithmetic)."""
        import torch
        n = x.shape[0]
        b = min(batch, n)
        dev = x.device
        counter = torch.zeros(1, dtype=torch.int64, device=dev)
        sched = torch.zeros(3, dtype=torch.float32, device=dev)
        xb = torch.empty((b, x.shape[1]), dtype=torch.float32, device=dev)
        tb = torch.empty((b, t.shape[1]), dtype=torch.float32, device=dev)
        step0 = self.steps
        seed = int(seed) * 0x9E3779B97F4A7C15 + 1

        def one_step():
            _native.mlp_sample_rows(x, t, b, seed, counter, xb, tb)
            loss, _ = self.loss_and_grads(xb, tb, out=self.grads)
            _native.adam_schedule(counter, sched, lr, lr_min, steps, step0)
            _native.adam_step_sched(self.flat, self.g_flat, self.m_flat, self.v_flat, sched)
            return loss

        hist = []

        def log(k, loss):
            if log_every and (k % log_every == 0 or k == steps - 1):
                hist.append((k, float(loss.item())))

        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            loss = one_step()   # eager: sizes the workspaces, binds this stream
            log(0, loss)
            if graph and steps > 1:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
                    gloss = one_step()
                for k in range(1, steps):
                    g.replay()
                    log(k, gloss)
            else:
                for k in range(1, steps):
                    log(k, one_step())
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.steps = step0 + steps
        return hist

    def set_scaling(self, x, t):
        """Standardise the features and scale the targets to unit RMS (kept in
        meta; train() and predict() take and return unscaled values)."""
        self.x_mean = x.mean(dim=0)
        self.x_std = x.std(dim=0).clamp(min=1e-6)
        self.t_scale = float(t.pow(2).mean().sqrt().item()) or 1.0

    def _xs(self, x):
        return ((x - self.x_mean) / self.x_std).contiguous()

    def predict(self, x, batch=1 << 16):
        import torch
        out = []
        for i in range(0, x.shape[0], batch):
            xb = x[i:i + batch]
            y = self.forward(self._xs(xb))[-1] * self.t_scale
            if self.scale_feature is not None:
                y = y * xb[:, self.scale_feature:self.scale_feature + 1]
            out.append(y)
        return torch.cat(out)


# ---------------------------------------------------------------------------
# kernels <-> tap blocks
# ---------------------------------------------------------------------------
def tap_index(support, oversampling, R, K):
    """Kernel index of tap r (0 <= r < R) at sub-cell offset ou (-os/2..os/2):
    os*(r - S//2) + ou + K//2 (the gridder's cf index), -1 beyond the support.
    Returns an int array [os+1, R]."""
    os_ = int(oversampling)
    half = os_ // 2
    ou = np.arange(-half, half + 1)[:, None]
    r = np.arange(R)[None, :]
    k = os_ * (r - support // 2) + ou + K // 2
    return np.where(r < support, k, -1)


def tap_blocks(ck, support, oversampling, R):
    """Tap blocks [os_u+1, os_v+1, R, R] of one K x K kernel (device)."""
    import torch
    K = ck.shape[-1]
    iu = torch.as_tensor(tap_index(int(support[0]), oversampling[0], R, K), device=ck.device)
    iv = torch.as_tensor(tap_index(int(support[1]), oversampling[1], R, K), device=ck.device)
    mu = iu >= 0
    mv = iv >= 0
    blk = ck[iu.clamp(min=0)[:, None, :, None], iv.clamp(min=0)[None, :, None, :]]
    return blk * (mu[:, None, :, None] & mv[None, :, None, :])


def blocks_to_kernel(blocks, support, oversampling, K, base=None):
    """Inverse of tap_blocks: scatter blocks [os_u+1, os_v+1, R, R] back into a
    K x K kernel (zeros, or a copy of `base`, outside the support window, which
    no gridder reads).  Every kernel sample inside the support window is
    written once: with an even oversampling the offset +os/2 of tap r is the
    sample of offset -os/2 of tap r+1, so +os/2 is kept only for the last tap."""
    import torch
    R = blocks.shape[-1]
    dev = blocks.device

    def kept(s, os_):
        idx = tap_index(int(s), int(os_), R, K)
        m = idx >= 0
        if int(os_) % 2 == 0:
            m[-1, :] = False
            m[-1, int(s) - 1] = idx[-1, int(s) - 1] >= 0
        o, r = np.nonzero(m)
        return (torch.as_tensor(o, device=dev), torch.as_tensor(r, device=dev),
                torch.as_tensor(idx[o, r], device=dev))
    ou, ru, iu = kept(support[0], oversampling[0])
    ov, rv, iv = kept(support[1], oversampling[1])
    out = (torch.zeros((K, K), dtype=blocks.dtype, device=dev) if base is None
           else base.clone())
    out[iu[:, None], iv[None, :]] = blocks[ou[:, None], ov[None, :], ru[:, None], rv[None, :]]
    return out


def features(dish_pair, blockage_pair, freq, f_bin, f_ref, pa, ou, ov, oversampling):
    """Feature rows [n, 10] (host float32) for broadcastable inputs."""
    f = np.broadcast_arrays(freq, f_bin, pa, ou, ov)
    n = f[0].size
    x = np.empty((n, N_IN), dtype=np.float32)
    x[:, 0] = dish_pair[0] / 12.0
    x[:, 1] = dish_pair[1] / 12.0
    x[:, 2] = blockage_pair[0]
    x[:, 3] = blockage_pair[1]
    x[:, 4] = 100.0 * (f[0].ravel() / f[1].ravel() - 1.0)
    x[:, 5] = 10.0 * (f[1].ravel() / f_ref - 1.0)
    x[:, 6] = np.cos(2 * f[2].ravel())
    x[:, 7] = np.sin(2 * f[2].ravel())
    x[:, 8] = f[3].ravel() / float(oversampling[0])
    x[:, 9] = f[4].ravel() / float(oversampling[1])
    return x


def _offset_grid(oversampling):
    hu, hv = int(oversampling[0]) // 2, int(oversampling[1]) // 2
    ou, ov = np.meshgrid(np.arange(-hu, hu + 1), np.arange(-hv, hv + 1), indexing="ij")
    return ou.ravel(), ov.ravel()


def exact_blocks(freqs, dish, blockage, pairs, casa, pb_cell, npad, oversampling, max_support,
                 supports, R):
    """Exact tap blocks [n_pairs, n_freq, (os+1)^2, R*R] (device float64) of the
    convolution kernels at `freqs` with the given supports [n_pairs, 2]."""
    import torch
    sup = np.repeat(np.asarray(supports)[:, None, None, :], len(freqs), axis=1)
    _, ck, _ = gcf_kernels(np.asarray(freqs, dtype=np.float64), dish, blockage, pairs, casa,
                           pb_cell, npad, oversampling, max_support, 0.0, supports=sup)
    out = []
    for p in range(len(pairs)):
        rows = []
        for f in range(len(freqs)):
            b = tap_blocks(ck[p, f, 0], supports[p], oversampling, R)
            rows.append(b.reshape(-1, R * R))
        out.append(torch.stack(rows))
    return torch.stack(out), ck


def training_set(dish, blockage, pairs, casa, pb_cell, npad, oversampling, max_support,
                 supports, f_bins, half_width, n_freq, f_ref, seed=0):
    """Examples for every pair, bin frequency, n_freq frequencies spread over
    [f_bin - half_width, f_bin + half_width] and every sub-cell offset pair.
    Returns device (x [n, 10] float32, t [n, R*R] float32, peaks)."""
    import torch
    R = int(np.max(supports))
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    peaks = {}
    for fb in f_bins:
        fr = fb + np.linspace(-half_width, half_width, n_freq)
        blk, _ = exact_blocks(np.concatenate([[fb], fr]), dish, blockage, pairs, casa, pb_cell,
                              npad, oversampling, max_support, supports, R)
        ou, ov = _offset_grid(oversampling)
        for p, (i, j) in enumerate(pairs):
            base = blk[p, 0]
            peak = float(base.abs().max().item())
            peaks[(p, float(fb))] = peak
            for k, f in enumerate(fr):
                pa = rng.uniform(0, np.pi, size=len(ou))
                xs.append(features((dish[i], dish[j]), (blockage[i], blockage[j]), f, fb, f_ref,
                                   pa, ou, ov, oversampling))
                ts.append(((blk[p, k + 1] - base) / peak).float())
    x = torch.as_tensor(np.concatenate(xs), device="cuda")
    t = torch.cat(ts).contiguous()
    return x, t, peaks


def train_surrogate(dish, blockage, pairs, casa, pb_cell, npad, oversampling, max_support,
                    supports, f_bins, half_width, n_freq=24, steps=3000, batch=4096, seed=0,
                    log_every=0, relative=True, lr=2e-3):
    """Build the exact training set on the GPU and train a KernelSurrogate
    (relative: the network predicts the change per unit frequency offset).
    Returns (model, info)."""
    t0 = time.time()
    f_ref = float(np.mean(f_bins))
    x, t, peaks = training_set(dish, blockage, pairs, casa, pb_cell, npad, oversampling,
                               max_support, supports, f_bins, half_width, n_freq, f_ref, seed)
    t_data = time.time() - t0
    import torch
    R = int(np.max(supports))
    model = KernelSurrogate(N_IN, R * R, seed=seed)
    if relative:
        # learn the change per unit of the frequency offset feature (examples at
        # the bin frequency itself carry no change and are left out)
        model.scale_feature = 4
        d = x[:, 4:5]
        keep = (d.abs() > 1e-9).squeeze(1)
        x, t = x[keep].contiguous(), (t[keep] / d[keep]).contiguous()
    model.set_scaling(x, t)
    t0 = time.time()
    hist = model.train(model._xs(x), (t / model.t_scale).contiguous(), steps=steps, batch=batch,
                       lr=lr, seed=seed + 1, log_every=log_every)
    torch.cuda.synchronize()
    model.meta = dict(f_ref=f_ref, R=R, supports=np.asarray(supports).tolist(),
                      oversampling=[int(o) for o in oversampling],
                      half_width=float(half_width), f_bins=[float(f) for f in f_bins],
                      dish=[float(d) for d in dish], blockage=[float(b) for b in blockage],
                      pairs=np.asarray(pairs).tolist(), scale_feature=model.scale_feature)
    info = dict(examples=int(x.shape[0]), data_s=round(t_data, 2),
                train_s=round(time.time() - t0, 2), steps=steps, history=hist)
    return model, info


def predict_blocks(model, pair, dish, blockage, freqs, f_bin, base_blocks, peak, oversampling,
                   pa=0.0):
    """Per-frequency tap blocks [n_freq, (os+1)^2, R*R]: the bin's exact blocks
    plus the predicted change, at parallactic angle `pa` (radians; the exact
    kernels of symmetric Airy patterns do not depend on it)."""
    import torch
    i, j = pair
    ou, ov = _offset_grid(oversampling)
    x = np.concatenate([features((dish[i], dish[j]), (blockage[i], blockage[j]), f, f_bin,
                                 model.meta["f_ref"], pa, ou, ov, oversampling) for f in freqs])
    y = model.predict(torch.as_tensor(x, device="cuda"))
    return base_blocks[None] + peak * y.double().reshape(len(freqs), len(ou), -1)


def per_channel_kernels(model, freq_chan, cf_chan_map, pb_freq, dish, blockage, pairs, casa,
                        pb_cell, npad, oversampling, max_support, supports, ck_bins, wck_bins,
                        check=True, keep_exact=False):
    """Per-channel CONV_KERNEL [n_pairs, n_chan, 1, K, K] from the surrogate,
    the bins' WEIGHT_CONV_KERNEL repeated per channel, and a report (largest
    tap error of the surrogate and of the bin kernel per channel when check,
    and which kernels fell back to the exact one; with keep_exact the exact
    per-channel kernels of the check as report["ck_exact"])."""
    import torch
    freq_chan = np.asarray(freq_chan, dtype=np.float64)
    n_pairs, n_chan = len(pairs), len(freq_chan)
    K = ck_bins.shape[-1]
    R = int(model.meta["R"])
    sup_pair = np.asarray(supports)[:, 0, 0, :]   # supports of the first bin per pair
    ck = torch.empty((n_pairs, n_chan, 1, K, K), dtype=ck_bins.dtype, device=ck_bins.device)
    wck = wck_bins[:, cf_chan_map].contiguous()
    report = dict(fallback=[], max_err_surrogate=0.0, max_err_bin=0.0, checked=bool(check))
    t0 = time.time()
    for p, (i, j) in enumerate(pairs):
        for b, fb in enumerate(pb_freq):
            chans = np.nonzero(np.asarray(cf_chan_map) == b)[0]
            if len(chans) == 0:
                continue
            sup = np.asarray(supports)[p, b, 0]
            base = tap_blocks(ck_bins[p, b, 0], sup, oversampling, R).reshape(-1, R * R)
            peak = float(base.abs().max().item())
            blocks = predict_blocks(model, (i, j), dish, blockage, freq_chan[chans], fb, base,
                                    peak, oversampling)
            for k, ch in enumerate(chans):
                ck[p, ch, 0] = blocks_to_kernel(blocks[k].reshape(
                    oversampling[0] + 1, oversampling[1] + 1, R, R), sup, oversampling, K,
                    base=ck_bins[p, b, 0])
    report["surrogate_s"] = round(time.time() - t0, 3)
    if check:
        t0 = time.time()
        _, ck_exact, _ = gcf_kernels(freq_chan, dish, blockage, pairs, casa, pb_cell, npad,
                                     oversampling, max_support, 0.0,
                                     supports=np.asarray(supports)[:, cf_chan_map])
        report["exact_s"] = round(time.time() - t0, 3)
        for p in range(n_pairs):
            for ch in range(n_chan):
                # over the taps the gridder reads: the support window at every
                # sub-cell offset
                sup = np.asarray(supports)[p, cf_chan_map[ch], 0]
                ex = tap_blocks(ck_exact[p, ch, 0], sup, oversampling, R)
                peak = float(ex.abs().max().item())
                e_s = float((tap_blocks(ck[p, ch, 0], sup, oversampling, R) - ex)
                            .abs().max().item()) / peak
                e_b = float((tap_blocks(ck_bins[p, cf_chan_map[ch], 0], sup, oversampling, R)
                             - ex).abs().max().item()) / peak
                report["max_err_surrogate"] = max(report["max_err_surrogate"], e_s)
                report["max_err_bin"] = max(report["max_err_bin"], e_b)
                if e_s > TOL_PEAK:
                    ck[p, ch, 0] = ck_exact[p, ch, 0]
                    report["fallback"].append((int(p), int(ch), e_s))
        if keep_exact:
            report["ck_exact"] = ck_exact
    return ck, wck, report
