"""Gridding plans (ngc_grid_plan_*): the autotuned, hipGraph-captured tiled
pipeline against the eager call and the oracle, over several replays with
eager work and synchronisation between them."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from helpers import small_case

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t.to(dtype) if dtype is not None else t


def _case():
    c = small_case(n_time=30, n_chan=6, n_pol=2, nan_frac=0.05)
    rng = np.random.default_rng(5)
    # uneven per-plane flags, as in the batched-cube workload
    frac = rng.uniform(0, 0.6, size=(6, 2))
    flag = (rng.uniform(size=c["weight"].shape) < frac[None, None]).astype(np.uint8)
    return c, flag


@pytest.mark.parametrize("precision", ["fp64_tiled", "fp32_kahan", "fp32"])
@pytest.mark.parametrize("mode", ["image", "psf"])
@pytest.mark.parametrize("autotune", [False, True])
def test_plan_replays_match_eager_and_oracle(gpu, precision, mode, autotune):
    import torch
    import cngi_prototype_amd._native as nat
    c, flag = _case()
    cgk = oracle.ps_kernel_1d(100, 7)
    npad = c["npad"]
    single = precision != "fp64_tiled"
    vdt = torch.complex64 if single else torch.complex128
    wdt = torch.float32 if single else torch.float64
    vis = _dev(c["vis"], vdt)
    w = _dev(c["weight"], wdt)
    uvw = _dev(c["uvw"])
    fl = _dev(flag)
    n_chan = w.shape[2]
    gm = nat.GRID_IMAGE if mode == "image" else nat.GRID_PSF
    args = (uvw, w, c["freq"], np.arange(n_chan), np.arange(2), cgk, n_chan, [npad, npad],
            c["delta"], 7, 100, gm, precision)
    g_e, sw_e = nat.standard_grid(*args, vis=vis if mode == "image" else None, flag=fl)
    g_e = g_e.cpu().numpy()
    sw_e = sw_e.cpu().numpy()
    plan = nat.GridPlan(*args, vis=vis if mode == "image" else None, flag=fl, autotune=autotune)
    assert "tile" in plan.describe()
    if autotune:
        assert "tuning ms" in plan.describe()
    # oracle on the same (rounded) inputs with flags as NaN weights
    wr = c["weight"].astype(np.float32 if single else np.float64).astype(np.float64)
    wr = np.where(flag.astype(bool), np.nan, wr)
    visr = c["vis"].astype(np.complex64 if single else np.complex128).astype(np.complex128)
    g_ref, sw_ref = oracle.standard_grid(visr, c["uvw"], wr, c["freq"], cgk, [npad, npad],
                                         c["delta"], 7, 100, do_psf=(mode == "psf"))
    for rep in range(3):
        g, sw = plan.execute()
        torch.cuda.synchronize()
        gh = g.cpu().numpy()
        assert rel_err(sw.cpu().numpy(), sw_ref) < 1e-12, rep
        assert rel_err(gh, g_ref) < (1e-6 if single else 1e-12), rep
        assert rel_err(gh, g_e) < (1e-6 if single else 1e-13), rep
        # eager work between replays, on the same grid memory
        g.zero_()
        sw.zero_()
        torch.cuda.synchronize()
    plan.status()
    plan.close()


def test_plan_follows_inputs_rewritten_in_place(gpu):
    import torch
    import cngi_prototype_amd._native as nat
    c, flag = _case()
    cgk = oracle.ps_kernel_1d(100, 7)
    npad = c["npad"]
    vis = _dev(c["vis"], torch.complex64)
    w = _dev(c["weight"], torch.float32)
    plan = nat.GridPlan(_dev(c["uvw"]), w, c["freq"], np.arange(6), np.arange(2), cgk, 6,
                        [npad, npad], c["delta"], 7, 100, nat.GRID_IMAGE, "fp32", vis=vis,
                        flag=_dev(flag), autotune=False)
    g1 = plan.execute()[0].clone()
    vis.mul_(2.0)
    g2 = plan.execute()[0]
    torch.cuda.synchronize()
    assert rel_err(g2.cpu().numpy(), 2 * g1.cpu().numpy()) < 1e-6
    plan.status()


def test_plan_reports_record_overflow(gpu):
    """A plan made with most samples flagged holds too few records once the
    flags are cleared: the fill pass stops at the capacity and status()
    reports it (the kernels stay inside their buffers)."""
    import torch
    import cngi_prototype_amd._native as nat
    c, _ = _case()
    cgk = oracle.ps_kernel_1d(100, 7)
    npad = c["npad"]
    flag = np.ones(c["weight"].shape, dtype=np.uint8)
    flag[:3] = 0
    fl = _dev(flag)
    plan = nat.GridPlan(_dev(c["uvw"]), _dev(c["weight"], torch.float32), c["freq"],
                        np.arange(6), np.arange(2), cgk, 6, [npad, npad], c["delta"], 7, 100,
                        nat.GRID_PSF, "fp32", flag=fl, autotune=False)
    plan.execute()
    plan.status()
    fl.zero_()
    plan.execute()
    with pytest.raises(nat.NativeError):
        plan.status()
