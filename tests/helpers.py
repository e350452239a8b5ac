"""Shared test inputs: small seeded versions of the benchmark configurations."""
import numpy as np

from cngi_prototype_amd import synthetic as syn

ARCSEC = np.pi / (180 * 3600)


def small_case(name="C1", n_time=24, n_chan=3, n_pol=2, image_size=128, cell=None, seed=None,
               flag_frac=0.0, noise=0.01, n_src=5, nan_frac=0.0):
    base = syn.CONFIGS[name]
    cfg = syn.scaled(base, n_time=n_time, n_chan=n_chan, n_pol=n_pol, n_src=n_src,
                     image_size=image_size, flag_frac=flag_frac, noise=noise,
                     seed=base.seed if seed is None else seed, flag_per_plane_max=0.0)
    enu = syn.antenna_enu(cfg)
    uvw = syn.uvw_tracks(enu, cfg.lat_deg, cfg.dec_deg, cfg.n_time, cfg.dt)
    fr = syn.freqs(cfg)
    vis = syn.visibilities_numpy(uvw, fr, syn.sky(cfg), n_pol, noise, cfg.seed)
    rng = np.random.default_rng(cfg.seed + 77)
    weight = rng.uniform(0.5, 2.0, size=vis.shape)
    if nan_frac > 0:
        m = rng.uniform(size=vis.shape) < nan_frac
        weight[m] = np.nan
    cell_arcsec = base.cell_arcsec if cell is None else cell
    npad = int(1.2 * image_size)
    # keep the uv coverage on the grid: scale cell so that u_max sits at ~70%
    umax = np.abs(uvw[:, :, :2]).max() * fr.max() / syn.C_LIGHT
    cell_fit = 0.7 * (npad / 2) / (umax * npad)
    cell_rad = min(cell_arcsec * ARCSEC, cell_fit)
    delta = np.array([-cell_rad, cell_rad])
    return dict(cfg=cfg, uvw=uvw, freq=fr, vis=vis, weight=weight, npad=npad,
                image_size=image_size, delta=delta, cell_arcsec=cell_rad / ARCSEC)


def mosaic_case(n_time=40, n_chan=4, n_pol=2, n_field=4, image_size=160, cell_arcsec=0.3,
                seed=4, flag_frac=0.05):
    """Small heterogeneous mosaic: 3 x 7 m + 5 x 12 m antennas within 150 m,
    n_field pointings visited in turn, 230 GHz.  Returns the mxds, the dish
    types and the grid geometry used by the mosaic parity tests."""
    base = syn.CONFIGS["C4"]
    cfg = syn.scaled(base, n_ant=8, r_max=150.0, lognormal_mu=np.log(50.0),
                     extra=dict(n_field=n_field, dish=(7.0, 12.0), n_dish=(3, 5)),
                     n_time=n_time, dt=60.0, n_chan=n_chan, df=50e6, n_pol=n_pol,
                     image_size=image_size, cell_arcsec=cell_arcsec, n_src=6, fov_deg=0.01,
                     flag_frac=flag_frac, seed=seed)
    mxds = syn.make_mosaic_mxds(cfg)
    types, dish = syn.dish_types(cfg)
    npad = int(1.2 * image_size)
    delta = np.array([-cell_arcsec, cell_arcsec]) * ARCSEC
    return dict(cfg=cfg, mxds=mxds, types=types, dish=dish, npad=npad, image_size=image_size,
                delta=delta, phase_center=syn.phase_center(cfg))
