"""make_pb: the GPU Airy primary-beam kernel (ngc_make_pb) against the oracle's
_casa_airy_disk / _airy_disk restatement, and the make_pb entry point's
outputs (dims, coordinates, data group)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ARCSEC = np.pi / (180 * 3600)


@pytest.mark.parametrize("casa", [True, False])
@pytest.mark.parametrize("ipower", [1, 2])
def test_make_pb_kernel_matches_oracle(gpu, casa, ipower):
    import cngi_prototype_amd._native as nat
    freq = np.array([86.0e9, 100.0e9, 115.5e9])
    dish = [12.0, 7.0, 25.0]
    blockage = [0.75, 0.0, 2.0]
    size, centre = [96, 81], [41, 40]
    cell = np.array([-0.4 * ARCSEC, 0.35 * ARCSEC])
    got = nat.make_pb(freq, dish, blockage, casa, ipower, size, centre, cell, 2).cpu().numpy()
    ref = oracle.pb_image(freq, 2, dish, blockage, ipower, cell, size, centre, casa=casa)
    assert got.shape == ref.shape == (96, 81, 3, 2, 3)
    # device J1 against scipy's; values are at most 1
    assert np.abs(got - ref).max() < 1e-13
    assert np.all(got[41, 40] == 1.0)


def test_make_pb_refuses_centre_outside_image(gpu):
    import cngi_prototype_amd._native as nat
    with pytest.raises(nat.NativeError):
        nat.make_pb([1e11], [12.0], [0.0], True, 2, [32, 32], [32, 0], [1e-6, 1e-6], 1)


def test_make_pb_entry_point(gpu):
    from cngi_prototype_amd.image import make_empty_sky_image
    from cngi_prototype_amd.imaging import make_pb
    from cngi_prototype_amd.xds import Dataset
    freq = np.array([230.0e9, 230.5e9])
    n = 64
    img = make_empty_sky_image(Dataset(), [0.3, -0.5], [n, n], [0.5 * ARCSEC, 0.5 * ARCSEC],
                               freq, [5e8, 5e8], [9, 12], [0.0])
    pb_parms = {'list_dish_diameters': [12.0, 7.0], 'list_blockage_diameters': [0.75, 0.75]}
    grid_parms = {'image_size': [n, n], 'cell_size': [0.5, 0.5], 'fft_padding': 1.0}
    out = make_pb(img, pb_parms, grid_parms, {})
    pb = np.asarray(out['PB'].values)
    assert out['PB'].dims == ('l', 'm', 'time', 'chan', 'pol', 'dish_type')
    assert pb.shape == (n, n, 1, 2, 2, 2)
    assert list(np.asarray(out.coords['dish_type'].values)) == [0, 1]
    assert 'PB' not in img.data_vars          # the input is not modified
    group = list(out.attrs['data_groups'][0].values())[0]
    assert group['pb'] == 'PB'
    cell = np.array([-0.5 * ARCSEC, 0.5 * ARCSEC])
    ref = oracle.pb_image(freq, 2, [12.0, 7.0], [0.75, 0.75], 2, cell, [n, n], [n // 2, n // 2])
    assert np.abs(pb[:, :, 0] - ref).max() < 1e-13
    with pytest.raises(UnboundLocalError):
        make_pb(img, {**pb_parms, 'function': 'gaussian'}, grid_parms, {})
