"""Seeded synthetic interferometer data for the gridding workloads.

Antenna layouts in a local East-North-Up frame, uvw tracks by Earth rotation
(Thompson, Moran & Swenson, 3rd ed., eq. 4.1), point-source (and Gaussian)
skies, thermal noise and random flags.  Small cases are built with numpy on
the host; large ones (hundreds of millions of samples) are built directly in
device memory with torch, chunk by chunk, so that the benchmark never holds
them on the host.

Visibility convention: V(u, v, w) = sum_k S_k exp(+2 pi i (u l_k + v m_k +
w (n_k - 1))) with u, v, w in wavelengths.  With the reference gridder's uv
scale (-freq*cell*N/c with cell[0] negated) a source at direction cosines
(l, m) then appears at image coordinates (l, m) of make_empty_sky_image.
"""
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0
OMEGA_EARTH = 7.2921150e-5  # rad/s, sidereal
ARCSEC = np.pi / (180 * 3600)


@dataclass
class Config:
    name: str
    n_ant: int
    layout: str              # 'uniform_disk' | 'lognormal'
    r_max: float             # metres
    lat_deg: float
    dec_deg: float
    n_time: int
    dt: float                # s
    f0: float                # Hz
    df: float                # Hz
    n_chan: int
    n_pol: int
    image_size: int
    cell_arcsec: float
    n_src: int
    flux_range: tuple = (1e-3, 1.0)
    fov_deg: float = 0.5
    noise: float = 0.01      # Jy per real/imag component
    flag_frac: float = 0.0   # fraction of (row, chan) samples flagged
    flag_per_plane_max: float = 0.0  # uneven per-(chan, pol) flag fraction U(0, max)
    n_blob: int = 0
    lognormal_mu: float = np.log(400.0)
    lognormal_sigma: float = 0.8
    spectral_index: bool = False
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def n_baseline(self):
        return self.n_ant * (self.n_ant - 1) // 2

    @property
    def n_vis(self):
        return self.n_time * self.n_baseline * self.n_chan * self.n_pol


# The workloads of the benchmark.  Where their description leaves a value open
# (integration time, pols, antenna layout, site latitude, sky) the value below
# is this package's choice.
CONFIGS = {
    # plain standard gridding, make_psf + make_image, natural weighting
    "C1": Config("C1", 27, "uniform_disk", 1000.0, 34.0784, 30.0, 300, 10.0, 1.4e9, 1e6, 16, 2,
                 512, 2.0, 20, seed=1),
    # Briggs 0.5, 4 pols, 10% flags, 2048^2 at 0.1"
    "C2": Config("C2", 50, "lognormal", 3000.0, -23.0234, -23.0, 2000, 4.0, 100e9, 10e6, 64, 4,
                 2048, 0.1, 200, flag_frac=0.10, n_blob=5, fov_deg=0.02, seed=2),
    # batched cube: 256 chan x 2 pol, uneven per-plane flags 0-60%; cell 0.05"
    # keeps the longest (6 km) baselines on the 1228^2 grid at 100.5 GHz
    "C3": Config("C3", 50, "lognormal", 3000.0, -23.0234, -23.0, 1000, 8.0, 100e9, 2e6, 256, 2,
                 1024, 0.05, 100, flag_per_plane_max=0.6, spectral_index=True, fov_deg=0.01,
                 seed=3),
    # 25-pointing heterogeneous mosaic: 12 x 7 m + 40 x 12 m, 230 GHz, 4 h of 10 s dumps
    # (cell 0.09": the 12 m - 12 m GCF support is 17 at oversampling 20,
    # max_support 18; 7 m - 7 m is 11 and 7 m - 12 m 15)
    "C4": Config("C4", 52, "lognormal", 1500.0, -23.0234, -23.0, 1440, 10.0, 230e9, 15.625e6,
                 32, 2, 1024, 0.09, 60, fov_deg=0.012, lognormal_mu=np.log(150.0),
                 lognormal_sigma=0.7, seed=4,
                 extra=dict(n_field=25, dish=(7.0, 12.0), n_dish=(12, 40))),
    # large streamed mosaic cube: 64 x 12 m antennas, 8 h of 7.43 s dumps at 100 GHz,
    # 64 chan x 2 pol (1.0e9 visibilities), 49 pointings, 4096^2 at 0.05"
    "C5": Config("C5", 64, "lognormal", 3000.0, -23.0234, -23.0, 3875, 7.432, 100e9, 7.8125e6,
                 64, 2, 4096, 0.05, 80, fov_deg=0.03, seed=5,
                 extra=dict(n_field=49, dish=(12.0,), n_dish=(64,))),
}


def scaled(cfg, **kw):
    """Copy of a config with some fields replaced (for small test versions)."""
    import dataclasses
    return dataclasses.replace(cfg, **kw)


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------
def antenna_enu(cfg):
    rng = np.random.default_rng(cfg.seed)
    n = cfg.n_ant
    if cfg.layout == "uniform_disk":
        r = cfg.r_max * np.sqrt(rng.uniform(size=n))
    elif cfg.layout == "lognormal":
        r = np.empty(n)
        k = 0
        while k < n:
            x = rng.lognormal(cfg.lognormal_mu, cfg.lognormal_sigma)
            if x <= cfg.r_max:
                r[k] = x
                k += 1
    else:
        raise ValueError(cfg.layout)
    th = rng.uniform(0, 2 * np.pi, size=n)
    return np.stack([r * np.cos(th), r * np.sin(th), np.zeros(n)], axis=1)


def baseline_pairs(n_ant):
    a1, a2 = np.triu_indices(n_ant, k=1)
    return a1.astype(np.int32), a2.astype(np.int32)


def uvw_tracks(enu, lat_deg, dec_deg, n_time, dt):
    """uvw [n_time, n_baseline, 3] in metres, hour angle centred on transit."""
    lat = np.deg2rad(lat_deg)
    dec = np.deg2rad(dec_deg)
    a1, a2 = baseline_pairs(len(enu))
    b = enu[a2] - enu[a1]
    E, N, U = b[:, 0], b[:, 1], b[:, 2]
    X = -np.sin(lat) * N + np.cos(lat) * U
    Y = E
    Z = np.cos(lat) * N + np.sin(lat) * U
    t = (np.arange(n_time) - (n_time - 1) / 2.0) * dt
    H = OMEGA_EARTH * t
    sH, cH = np.sin(H)[:, None], np.cos(H)[:, None]
    u = sH * X + cH * Y
    v = -np.sin(dec) * cH * X + np.sin(dec) * sH * Y + np.cos(dec) * Z
    w = np.cos(dec) * cH * X - np.cos(dec) * sH * Y + np.sin(dec) * Z
    return np.stack([u, v, w], axis=2)


def sky(cfg):
    rng = np.random.default_rng(cfg.seed + 1000)
    half = np.deg2rad(cfg.fov_deg) / 2
    l = rng.uniform(-half, half, cfg.n_src)
    m = rng.uniform(-half, half, cfg.n_src)
    lo, hi = np.log(cfg.flux_range[0]), np.log(cfg.flux_range[1])
    s = np.exp(rng.uniform(lo, hi, cfg.n_src))
    alpha = rng.normal(-0.7, 0.2, cfg.n_src) if cfg.spectral_index else np.zeros(cfg.n_src)
    blobs = []
    for _ in range(cfg.n_blob):
        blobs.append((rng.uniform(-half, half), rng.uniform(-half, half),
                      float(np.exp(rng.uniform(lo, hi))), float(rng.uniform(0.2, 1.0) * half / 10)))
    return dict(l=l, m=m, flux=s, alpha=alpha, blobs=blobs)


def dish_types(cfg):
    """Dish-type index of every antenna (type-0 antennas first, so that
    type(ant1) <= type(ant2) on every baseline) and the diameters."""
    counts = cfg.extra.get("n_dish", (cfg.n_ant,))
    dish = np.asarray(cfg.extra.get("dish", (12.0,)), dtype=np.float64)
    types = np.concatenate([np.full(n, k, dtype=int) for k, n in enumerate(counts)])
    assert len(types) == cfg.n_ant
    return types, dish


def mosaic_fields(cfg, phase_dir=None):
    """Pointing centres [n_field, 2] (ra, dec radians) on a hexagonal grid
    spaced by lambda/(2 D_max) at the lowest frequency, whose mean is the
    phase centre (rows of 5 for 25 fields)."""
    n_field = int(cfg.extra.get("n_field", 1))
    ra0, dec0 = phase_center(cfg) if phase_dir is None else phase_dir
    _, dish = dish_types(cfg)
    step = C_LIGHT / cfg.f0 / (2 * dish.max())
    side = int(np.ceil(np.sqrt(n_field)))
    pts = []
    for r in range(side):
        for c in range(side):
            pts.append(((c - (side - 1) / 2 + (0.5 if r % 2 else 0.0) - 0.25) * step,
                        (r - (side - 1) / 2) * step * np.sqrt(3) / 2))
    pts = np.array(pts[:n_field])
    pts -= pts.mean(axis=0)
    return np.stack([ra0 + pts[:, 0] / np.cos(dec0), dec0 + pts[:, 1]], 1)


def field_schedule(cfg):
    """FIELD_ID of every integration: the fields are visited in turn, in
    contiguous blocks whose lengths differ by at most one dump."""
    n_field = int(cfg.extra.get("n_field", 1))
    return ((np.arange(cfg.n_time) * n_field) // cfg.n_time).astype(np.int32)


def airy_voltage(r_rad, freq, dish):
    """Unblocked Airy voltage pattern 2 J1(x)/x, x = pi D r / lambda."""
    from scipy.special import j1
    x = np.pi * dish * np.asarray(r_rad)[..., None] * np.asarray(freq) / C_LIGHT
    xs = np.where(x == 0, 1.0, x)
    return np.where(x == 0, 1.0, 2 * j1(xs) / xs)


def freqs(cfg):
    return cfg.f0 + np.arange(cfg.n_chan) * cfg.df


# ---------------------------------------------------------------------------
# host (numpy) construction, small configurations
# ---------------------------------------------------------------------------
def visibilities_numpy(uvw, freq, sk, n_pol, noise, seed, f_ref=None):
    """complex128 [T, B, C, P]; XX = YY = I, cross hands carry noise only."""
    rng = np.random.default_rng(seed + 2000)
    T, B, _ = uvw.shape
    C = len(freq)
    f_ref = freq[0] if f_ref is None else f_ref
    lam = freq / C_LIGHT
    u = uvw[:, :, 0, None] * lam
    v = uvw[:, :, 1, None] * lam
    w = uvw[:, :, 2, None] * lam
    I = np.zeros((T, B, C), dtype=np.complex128)
    for lk, mk, sk_, ak in zip(sk["l"], sk["m"], sk["flux"], sk["alpha"]):
        nk = np.sqrt(1 - lk * lk - mk * mk)
        spec = (freq / f_ref) ** ak
        I += sk_ * spec * np.exp(2j * np.pi * (u * lk + v * mk + w * (nk - 1)))
    for (lk, mk, sk_, sig) in sk["blobs"]:
        I += sk_ * np.exp(-2 * np.pi ** 2 * sig ** 2 * (u * u + v * v)) * np.exp(
            2j * np.pi * (u * lk + v * mk))
    vis = np.zeros((T, B, C, n_pol), dtype=np.complex128)
    if n_pol == 1:
        vis[..., 0] = I
    elif n_pol == 2:
        vis[..., 0] = I
        vis[..., 1] = I
    else:
        vis[..., 0] = I
        vis[..., n_pol - 1] = I
    vis += noise * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))
    return vis


def flags_numpy(cfg, shape, seed):
    rng = np.random.default_rng(seed + 3000)
    T, B, C, P = shape
    flag = np.zeros(shape, dtype=np.uint8)
    if cfg.flag_frac > 0:
        f = rng.uniform(size=(T, B, C)) < cfg.flag_frac
        flag |= f[..., None].astype(np.uint8)
    if cfg.flag_per_plane_max > 0:
        frac = rng.uniform(0, cfg.flag_per_plane_max, size=(C, P))
        flag |= (rng.uniform(size=shape) < frac[None, None]).astype(np.uint8)
    return flag


def make_mxds(cfg, vis=None, uvw=None, weight=None, flag=None, apply_flags=True):
    """Build a multi-dataset in the reference's layout: mxds.attrs['xds0'] with
    DATA, UVW, DATA_WEIGHT, FLAG, ANTENNA1/2, coords time/baseline/chan/pol,
    chan_width, data_groups.  With apply_flags, flagged samples become NaN in
    DATA and DATA_WEIGHT (cngi/vis/apply_flags.py:49-53)."""
    from .xds import DataArray, Dataset
    enu = antenna_enu(cfg)
    if uvw is None:
        uvw = uvw_tracks(enu, cfg.lat_deg, cfg.dec_deg, cfg.n_time, cfg.dt)
    fr = freqs(cfg)
    if vis is None:
        vis = visibilities_numpy(uvw, fr, sky(cfg), cfg.n_pol, cfg.noise, cfg.seed)
    if weight is None:
        weight = np.ones(vis.shape, dtype=np.float64)
    if flag is None:
        flag = flags_numpy(cfg, vis.shape, cfg.seed)
    if apply_flags and flag.any():
        vis = vis.copy()
        weight = weight.copy()
        vis[flag.astype(bool)] = np.nan
        weight[flag.astype(bool)] = np.nan
    a1, a2 = baseline_pairs(cfg.n_ant)
    dims4 = ["time", "baseline", "chan", "pol"]
    xds = Dataset(
        data_vars={
            "DATA": (dims4, vis),
            "UVW": (["time", "baseline", "uvw_index"], uvw),
            "DATA_WEIGHT": (dims4, weight),
            "FLAG": (dims4, flag),
            "ANTENNA1": (["baseline"], a1),
            "ANTENNA2": (["baseline"], a2),
            "chan_width": (["chan"], np.full(cfg.n_chan, cfg.df)),
        },
        coords={"time": (["time"], np.arange(cfg.n_time) * cfg.dt),
                "baseline": (["baseline"], np.arange(cfg.n_baseline)),
                "chan": (["chan"], fr),
                "pol": (["pol"], np.array([9, 12] if cfg.n_pol == 2 else
                                          ([9, 10, 11, 12] if cfg.n_pol == 4 else [9])[:cfg.n_pol]))},
        attrs={"data_groups": [{"0": {"id": "0", "data": "DATA", "uvw": "UVW", "flag": "FLAG",
                                      "weight": "DATA_WEIGHT"}}]})
    mxds = Dataset(attrs={"xds0": xds})
    return mxds


def make_mosaic_mxds(cfg, apply_flags=True):
    """make_mxds for a mosaic: every integration points at one field
    (FIELD_ID [time, baseline]), visibilities are phase-referenced to the
    common phase centre and each source is attenuated by the product of the
    two antennas' Airy voltage patterns at its distance from that field's
    centre.  Adds mxds.attrs['FIELD'] (PHASE_DIR [n_field, 1, 2], coord
    field_id) as read by make_gridding_convolution_function."""
    from .xds import Dataset
    enu = antenna_enu(cfg)
    uvw = uvw_tracks(enu, cfg.lat_deg, cfg.dec_deg, cfg.n_time, cfg.dt)
    fr = freqs(cfg)
    sk = sky(cfg)
    ra0, dec0 = phase_center(cfg)
    fields = mosaic_fields(cfg)
    sched = field_schedule(cfg)
    types, dish = dish_types(cfg)
    a1, a2 = baseline_pairs(cfg.n_ant)
    # source directions from (l, m) about the phase centre (inverse SIN)
    l, m = sk["l"], sk["m"]
    n = np.sqrt(1 - l * l - m * m)
    dec_s = np.arcsin(m * np.cos(dec0) + n * np.sin(dec0))
    ra_s = ra0 + np.arctan2(l, n * np.cos(dec0) - m * np.sin(dec0))
    T, B, C, P = cfg.n_time, cfg.n_baseline, cfg.n_chan, cfg.n_pol
    lam = fr / C_LIGHT
    vis = np.zeros((T, B, C, P), dtype=np.complex128)
    for f in range(len(fields)):
        rows = np.nonzero(sched == f)[0]
        if len(rows) == 0:
            continue
        raf, decf = fields[f]
        cosr = (np.sin(dec_s) * np.sin(decf) +
                np.cos(dec_s) * np.cos(decf) * np.cos(ra_s - raf))
        r = np.arccos(np.clip(cosr, -1, 1))                               # [S]
        vp = np.stack([airy_voltage(r, fr, d) for d in dish])              # [n_dish, S, C]
        g = vp[types[a1]] * vp[types[a2]]                                  # [B, S, C]
        u = uvw[rows, :, 0, None] * lam
        v = uvw[rows, :, 1, None] * lam
        w = uvw[rows, :, 2, None] * lam
        I = np.zeros((len(rows), B, C), dtype=np.complex128)
        for k in range(len(l)):
            I += (sk["flux"][k] * g[None, :, k, :] *
                  np.exp(2j * np.pi * (u * l[k] + v * m[k] + w * (n[k] - 1))))
        vis[rows, :, :, 0] = I
        vis[rows, :, :, P - 1] = I
    rng = np.random.default_rng(cfg.seed + 2000)
    vis += cfg.noise * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))
    mxds = make_mxds(cfg, vis=vis, uvw=uvw, apply_flags=apply_flags)
    xds = mxds.attrs["xds0"]
    from .xds import DataArray
    xds["FIELD_ID"] = DataArray(np.repeat(sched[:, None], B, axis=1), dims=["time", "baseline"])
    field = Dataset(data_vars={"PHASE_DIR": (["field_id", "poly_id", "d2"], fields[:, None, :])},
                    coords={"field_id": (["field_id"], np.arange(len(fields)))})
    mxds.attrs["FIELD"] = field
    return mxds


def phase_center(cfg):
    """(ra, dec) in radians of the pointing: ra = 0, dec = cfg.dec_deg."""
    return np.array([0.0, np.deg2rad(cfg.dec_deg)])


# ---------------------------------------------------------------------------
# device (torch) construction, large configurations
# ---------------------------------------------------------------------------
def device_dataset(cfg, device="cuda", vis_dtype=None, weight_value=1.0, rows_per_chunk=8192,
                   with_vis=True):
    """Build uvw (float64), vis (complex64), weight (float32) and flag (uint8)
    directly in device memory for a large configuration.  Returns a dict of
    tensors [T, B, 3], [T, B, C, P]."""
    import torch
    vis_dtype = vis_dtype or torch.complex64
    enu = antenna_enu(cfg)
    uvw_np = uvw_tracks(enu, cfg.lat_deg, cfg.dec_deg, cfg.n_time, cfg.dt)
    T, B = cfg.n_time, cfg.n_baseline
    C, P = cfg.n_chan, cfg.n_pol
    uvw = torch.from_numpy(uvw_np).to(device)
    fr = torch.tensor(freqs(cfg), dtype=torch.float64, device=device)
    sk = sky(cfg)
    gen = torch.Generator(device=device)
    gen.manual_seed(cfg.seed + 4000)
    weight = torch.full((T, B, C, P), float(weight_value), dtype=torch.float32, device=device)
    flag = torch.zeros((T, B, C, P), dtype=torch.uint8, device=device)
    if cfg.flag_frac > 0:
        f = torch.rand((T, B, C), generator=gen, device=device) < cfg.flag_frac
        flag |= f[..., None].to(torch.uint8)
    if cfg.flag_per_plane_max > 0:
        frac = torch.rand((C, P), generator=gen, device=device, dtype=torch.float32) * \
            cfg.flag_per_plane_max
        for t0 in range(0, T, 64):
            t1 = min(T, t0 + 64)
            r = torch.rand((t1 - t0, B, C, P), generator=gen, device=device)
            flag[t0:t1] |= (r < frac).to(torch.uint8)
    out = dict(uvw=uvw, weight=weight, flag=flag, freq=freqs(cfg))
    if not with_vis:
        return out
    vis = torch.empty((T, B, C, P), dtype=vis_dtype, device=device)
    lk = torch.tensor(sk["l"], dtype=torch.float32, device=device)
    mk = torch.tensor(sk["m"], dtype=torch.float32, device=device)
    nk1 = torch.sqrt(1 - lk * lk - mk * mk) - 1
    flux = torch.tensor(sk["flux"], dtype=torch.float32, device=device)
    alpha = torch.tensor(sk["alpha"], dtype=torch.float32, device=device)
    spec = (fr.float()[:, None] / float(cfg.f0)) ** alpha[None, :]       # [C, S]
    lam = (fr / C_LIGHT).float()                                           # [C]
    uvw_flat = uvw.reshape(T * B, 3).float()
    n_rows = T * B
    vflat = vis.view(n_rows, C, P)
    for r0 in range(0, n_rows, rows_per_chunk):
        r1 = min(n_rows, r0 + rows_per_chunk)
        u = uvw_flat[r0:r1, 0, None, None] * lam[None, :, None]           # [R, C, 1]
        v = uvw_flat[r0:r1, 1, None, None] * lam[None, :, None]
        w = uvw_flat[r0:r1, 2, None, None] * lam[None, :, None]
        ph = 2 * np.pi * (u * lk + v * mk + w * nk1)                        # [R, C, S]
        amp = flux[None, None, :] * spec[None]
        I = torch.complex((amp * torch.cos(ph)).sum(-1), (amp * torch.sin(ph)).sum(-1))
        for (bl, bm, bs, sig) in sk["blobs"]:
            env = bs * torch.exp(-2 * np.pi ** 2 * sig ** 2 * (u * u + v * v))[..., 0]
            bph = 2 * np.pi * (u * bl + v * bm)[..., 0]
            I = I + torch.complex(env * torch.cos(bph), env * torch.sin(bph))
        noise = torch.complex(torch.randn((r1 - r0, C, P), generator=gen, device=device),
                              torch.randn((r1 - r0, C, P), generator=gen, device=device)) * cfg.noise
        chunk = noise
        chunk[:, :, 0] += I
        if P > 1:
            chunk[:, :, P - 1] += I
        vflat[r0:r1] = chunk.to(vis_dtype)
    out["vis"] = vis
    return out


def device_mosaic_dataset(cfg, device="cuda", rows_per_chunk=4096):
    """make_mosaic_mxds's data built in device memory for a large mosaic:
    uvw (float64 [T, B, 3]), FIELD_ID (int32 [T, B]), vis (complex64), weight
    (float32, ones), flag (uint8, none set).  Sources are attenuated by the
    two antennas' Airy voltage patterns at their distance from the field
    centre; phases are referenced to the common phase centre."""
    import torch
    enu = antenna_enu(cfg)
    uvw_np = uvw_tracks(enu, cfg.lat_deg, cfg.dec_deg, cfg.n_time, cfg.dt)
    T, B, C, P = cfg.n_time, cfg.n_baseline, cfg.n_chan, cfg.n_pol
    fr = freqs(cfg)
    sk = sky(cfg)
    ra0, dec0 = phase_center(cfg)
    fields = mosaic_fields(cfg)
    sched = field_schedule(cfg)
    types, dish = dish_types(cfg)
    a1, a2 = baseline_pairs(cfg.n_ant)
    l, m = sk["l"], sk["m"]
    n = np.sqrt(1 - l * l - m * m)
    dec_s = np.arcsin(m * np.cos(dec0) + n * np.sin(dec0))
    ra_s = ra0 + np.arctan2(l, n * np.cos(dec0) - m * np.sin(dec0))
    uvw = torch.from_numpy(uvw_np).to(device)
    field = torch.from_numpy(np.repeat(sched[:, None], B, axis=1).astype(np.int32)).to(device)
    vis = torch.empty((T, B, C, P), dtype=torch.complex64, device=device)
    lam = torch.tensor(fr / C_LIGHT, dtype=torch.float32, device=device)
    lk = torch.tensor(l, dtype=torch.float32, device=device)
    mk = torch.tensor(m, dtype=torch.float32, device=device)
    nk1 = torch.tensor(n - 1, dtype=torch.float32, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(cfg.seed + 4000)
    ta1 = torch.from_numpy(types[a1]).to(device)
    ta2 = torch.from_numpy(types[a2]).to(device)
    for f in range(len(fields)):
        rows_t = np.nonzero(sched == f)[0]
        if len(rows_t) == 0:
            continue
        raf, decf = fields[f]
        cosr = (np.sin(dec_s) * np.sin(decf) +
                np.cos(dec_s) * np.cos(decf) * np.cos(ra_s - raf))
        r = np.arccos(np.clip(cosr, -1, 1))
        vp = torch.tensor(np.stack([airy_voltage(r, fr, d) for d in dish]), dtype=torch.float32,
                          device=device)                                   # [n_dish, S, C]
        amp = (vp[ta1] * vp[ta2]) * torch.tensor(sk["flux"], dtype=torch.float32,
                                                  device=device)[None, :, None]  # [B, S, C]
        t0, t1 = int(rows_t[0]), int(rows_t[-1]) + 1
        flat = uvw[t0:t1].reshape(-1, 3).float()
        vflat = vis[t0:t1].view(-1, C, P)
        nr = flat.shape[0]
        for r0 in range(0, nr, rows_per_chunk):
            r1 = min(nr, r0 + rows_per_chunk)
            bidx = torch.arange(r0, r1, device=device) % B
            u = flat[r0:r1, 0, None, None] * lam[None, :, None]            # [R, C, 1]
            v = flat[r0:r1, 1, None, None] * lam[None, :, None]
            w = flat[r0:r1, 2, None, None] * lam[None, :, None]
            ph = 2 * np.pi * (u * lk + v * mk + w * nk1)                     # [R, C, S]
            a = amp[bidx].transpose(1, 2)                                    # [R, C, S]
            I = torch.complex((a * torch.cos(ph)).sum(-1), (a * torch.sin(ph)).sum(-1))
            noise = torch.complex(torch.randn((r1 - r0, C, P), generator=gen, device=device),
                                  torch.randn((r1 - r0, C, P), generator=gen, device=device))
            chunk = noise * cfg.noise
            chunk[:, :, 0] += I
            chunk[:, :, P - 1] += I
            vflat[r0:r1] = chunk.to(torch.complex64)
    weight = torch.ones((T, B, C, P), dtype=torch.float32, device=device)
    flag = torch.zeros((T, B, C, P), dtype=torch.uint8, device=device)
    return dict(uvw=uvw, field=field, vis=vis, weight=weight, flag=flag, freq=fr,
                fields=fields, types=types, dish=dish, a1=a1, a2=a2)
