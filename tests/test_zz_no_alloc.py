"""ngc_standard_grid_noalloc, the allocation-free form of ngc_standard_grid,
called through the C-ABI: it gives the existing entry point's results (to
the precision bar of each form) on two calls in a row, the second smaller,
so that a call does not depend on what an earlier call left in the shared
scratch array; it accumulates like the existing entry point; and a call too
large for the scratch is refused with a clear error before its outputs are
touched."""
import numpy as np
import pytest

from conftest import rel_err
from helpers import small_case

pytestmark = pytest.mark.gpu

FORMS = [("fp64", "image", 1e-12), ("fp64", "psf", 1e-12), ("fp64_tiled", "image", 1e-12),
         ("fp32_kahan", "image", 1e-6), ("fp32_kahan", "psf", 1e-6), ("fp32", "image", 1e-6),
         ("fp64", "density", 1e-12)]


def _inputs(c, precision, mode):
    import torch
    single = precision in ("fp32_kahan", "fp32")
    vdt = torch.complex64 if single else torch.complex128
    wdt = torch.float32 if single else torch.float64
    d = dict(uvw=torch.from_numpy(c["uvw"]).cuda(),
             weight=torch.from_numpy(c["weight"]).cuda().to(wdt))
    if mode == "image":
        d["vis"] = torch.from_numpy(c["vis"]).cuda().to(vdt)
    return d


def _grid(nat, c, d, precision, mode, **kw):
    from cngi_prototype_amd import synthetic as syn  # noqa: F401
    import oracle
    n_chan, n_pol = c["weight"].shape[2:]
    if mode == "density":
        cgk, support, os_, n = np.ones(1), 1, 0, c["image_size"]
        m = nat.GRID_DENSITY
    else:
        cgk, support, os_, n = oracle.ps_kernel_1d(100, 7), 7, 100, c["npad"]
        m = nat.GRID_IMAGE if mode == "image" else nat.GRID_PSF
    return nat.standard_grid(d["uvw"], d["weight"], c["freq"], np.arange(n_chan),
                             np.arange(n_pol), cgk, n_chan, [n, n], c["delta"], support, os_, m,
                             precision, vis=d.get("vis"), **kw)


@pytest.mark.parametrize("precision,mode,tol", FORMS)
def test_noalloc_matches_standard_grid_on_two_calls(gpu, precision, mode, tol):
    import torch
    import cngi_prototype_amd._native as nat
    big = small_case(n_time=40, n_chan=4, n_pol=4, image_size=128, nan_frac=0.05)
    small = small_case(n_time=9, n_chan=2, n_pol=4, image_size=64, seed=5, nan_frac=0.05)
    ins = [(c, _inputs(c, precision, mode)) for c in (big, small)]
    refs = [_grid(nat, c, d, precision, mode) for c, d in ins]
    outs = []
    for (c, d), (g_ref, sw_ref) in zip(ins, refs):
        g = torch.full_like(g_ref, 7.0)
        sw = torch.full_like(sw_ref, 7.0)
        print(f"noalloc {precision} {mode} grid {tuple(g.shape)}", flush=True)
        _grid(nat, c, d, precision, mode, grid=g, sum_weight=sw, noalloc=True)
        outs.append((g, sw))
    torch.cuda.synchronize()
    for (g, sw), (g_ref, sw_ref) in zip(outs, refs):
        assert torch.count_nonzero(g_ref) > 0
        assert rel_err(g.cpu().numpy(), g_ref.cpu().numpy()) < tol
        assert rel_err(sw.cpu().numpy(), sw_ref.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32_kahan"])
def test_noalloc_accumulates(gpu, precision):
    """accumulate = 1 adds the call to the grid and sum_weight it is given."""
    import torch
    import cngi_prototype_amd._native as nat
    a = small_case(n_time=20, n_chan=3, n_pol=2, image_size=96)
    b = small_case(n_time=20, n_chan=3, n_pol=2, image_size=96, seed=9)
    da, db = _inputs(a, precision, "image"), _inputs(b, precision, "image")
    ga, swa = _grid(nat, a, da, precision, "image")
    gb, swb = _grid(nat, b, db, precision, "image")
    g, sw = _grid(nat, a, da, precision, "image", noalloc=True)
    _grid(nat, b, db, precision, "image", grid=g, sum_weight=sw, accumulate=True, noalloc=True)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert rel_err(g.cpu().numpy(), (ga + gb).cpu().numpy()) < tol
    assert rel_err(sw.cpu().numpy(), (swa + swb).cpu().numpy()) < 1e-12


def test_noalloc_refuses_a_call_beyond_its_scratch(gpu):
    """About 2.4M bin records of 40 bytes do not fit the 64 MiB scratch: the
    call is refused with NGC_ERR_ARG naming the bin records, the grid and
    sum_weight keep what they held, and the next call that fits works."""
    import torch
    import cngi_prototype_amd._native as nat
    big = small_case(n_time=300, n_chan=16, n_pol=4, image_size=128, n_src=2)
    d = _inputs(big, "fp32_kahan", "image")
    g_ref, sw_ref = _grid(nat, big, d, "fp32_kahan", "image")
    g = torch.full_like(g_ref, 3.0)
    sw = torch.full_like(sw_ref, 3.0)
    with pytest.raises(nat.NativeError) as e:
        _grid(nat, big, d, "fp32_kahan", "image", grid=g, sum_weight=sw, noalloc=True)
    assert e.value.code == -1 and "bin records" in str(e.value), str(e.value)
    torch.cuda.synchronize()
    assert bool(torch.all(g == 3.0)) and bool(torch.all(sw == 3.0))
    small = small_case(n_time=12, n_chan=2, n_pol=4, image_size=64)
    ds = _inputs(small, "fp32_kahan", "image")
    gs_ref, sws_ref = _grid(nat, small, ds, "fp32_kahan", "image")
    gs, sws = _grid(nat, small, ds, "fp32_kahan", "image", noalloc=True)
    assert rel_err(gs.cpu().numpy(), gs_ref.cpu().numpy()) < 1e-6
    assert rel_err(sws.cpu().numpy(), sws_ref.cpu().numpy()) < 1e-12
