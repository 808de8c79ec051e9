"""FITS output (martini/martini.py:939-1152): the file format against the FITS
standard, the header against the reference's card list, and on the GPU the
data unit (mi_cube_to_fits) against numpy's transpose and byte swap.

The reference's own tests open the written files with astropy and run
``hdul.verify("exception")`` (tests/test_fits_compliant.py), check the data
dtype (tests/test_martini.py:246-268) and recover the HI mass from the file's
data and header (tests/test_martini.py:86-130).  astropy is not available
here, so the standard's rules are checked directly and the same mass is
recovered with ``martini_amd.fits.read_fits``.
"""

import numpy as np
import pytest

from martini_amd import DataCube, Martini, SPHSource, fits
from martini_amd.beams import GaussianBeam
from martini_amd.sph_kernels import CubicSplineKernel
from martini_amd.spectral_models import GaussianSpectrum

MANDATORY = ["SIMPLE", "BITPIX", "NAXIS"]


def _check_standard(path):
    """FITS 4.0 rules a reader relies on: whole 2880-byte blocks; 80-column
    ASCII cards; SIMPLE = T, BITPIX, NAXIS, NAXISn first and in order; an END
    card; blank header fill; zero data fill."""
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    cards, data = fits.read_fits(path)
    keys = [k for k, _, _ in cards]
    hdr = {k: v for k, v, _ in cards}
    naxis = hdr["NAXIS"]
    assert keys[:3 + naxis] == MANDATORY + [f"NAXIS{i}" for i in range(1, naxis + 1)]
    assert hdr["SIMPLE"] is True and hdr["BITPIX"] in (-32, -64)
    assert len(set(keys)) == len(keys)
    text = raw[:len(cards) * 80 + 80].decode("ascii")
    assert all(32 <= ord(ch) <= 126 for ch in text)
    end = len(cards) * 80
    assert text[end:end + 80] == "END".ljust(80)
    hlen = -(-(end + 80) // 2880) * 2880
    assert raw[end + 80:hlen] == b" " * (hlen - end - 80)
    nbytes = data.size * abs(hdr["BITPIX"]) // 8
    assert raw[hlen + nbytes:] == b"\0" * (len(raw) - hlen - nbytes)
    return cards, data


def test_card_format():
    assert fits.card("SIMPLE", True) == "SIMPLE  =                    T".ljust(80)
    assert fits.card("NAXIS", 3) == "NAXIS   =                    3".ljust(80)
    assert fits.card("CTYPE1", "RA---TAN") == "CTYPE1  = 'RA---TAN'".ljust(80)
    assert fits.card("BTYPE", "beam    ") == "BTYPE   = 'beam    '".ljust(80)
    assert fits.card("OBJECT", "MOCK") == "OBJECT  = 'MOCK    '".ljust(80)
    assert fits.card("BUNIT", "") == "BUNIT   = '        '".ljust(80)
    assert fits.card("OBSERVER", "K. O'Man") == "OBSERVER= 'K. O''Man'".ljust(80)
    for v in [1.0, -0.0016666666666666668, 1e-5, 6.02e23, 1420405751.0, -2.5e-300, 0.0]:
        c = fits.card("X", v)
        assert len(c) == 80 and c[8:10] == "= " and len(c[10:30].strip()) <= 20
        assert float(c[10:30]) == pytest.approx(v, rel=5e-15, abs=0)  # 15+ digits in 20 columns
        assert "." in c[10:30]  # a FITS real carries a decimal point
    with pytest.raises(ValueError):
        fits.card("TOOLONGKEY", 1)


def test_header_roundtrip(tmp_path):
    cards = [fits.card("SIMPLE", True), fits.card("BITPIX", -32), fits.card("NAXIS", 2),
             fits.card("NAXIS1", 5), fits.card("NAXIS2", 3),
             fits.card("INSTRUME", "MARTINI", "a comment"), fits.card("CDELT1", -1e-3)]
    data = np.arange(15, dtype=">f4").reshape(3, 5)
    path = str(tmp_path / "x.fits")
    fits.write_hdu(path, cards, data.tobytes(), overwrite=True)
    got_cards, got = _check_standard(path)
    assert got_cards[5] == ("INSTRUME", "MARTINI", "a comment")
    assert got_cards[6][1] == -1e-3
    np.testing.assert_array_equal(got, data)
    with pytest.raises(OSError):
        fits.write_hdu(path, cards, data.tobytes(), overwrite=False)


def test_filename_and_dtype_rules():
    assert fits.fits_filename("cube") == "cube.fits"
    assert fits.fits_filename("cube.fits") == "cube.fits"
    assert fits.fits_filename("cube.fits.gz") == "cube.fits.gz"
    assert fits.bitpix_for(None) == -64
    assert fits.bitpix_for(np.float32) == -32
    assert fits.bitpix_for("float64") == -64
    with pytest.raises(NotImplementedError):
        fits.bitpix_for(np.int16)


def _compact_source(n=200, seed=3):
    """A source well inside the cube: particles within 1 kpc of the centre."""
    rng = np.random.default_rng(seed)
    xyz = rng.normal(0.0, 0.3, (n, 3))
    vxyz = rng.normal(0.0, 10.0, (n, 3))
    return SPHSource(distance=3.0, mHI_g=rng.uniform(1e4, 1e5, n), xyz_g=xyz, vxyz_g=vxyz,
                     hsm_g=np.full(n, 0.5), T_g=np.full(n, 1e4))


def _martini(beam=True, stokes=False, freq=False):
    # centred on the source's Hubble-flow velocity, 3 Mpc x 70 km/s/Mpc
    dc = DataCube(n_px_x=40, n_px_y=36, n_channels=24, px_size=15.0,
                  channel_width=(5.0 if not freq else 2.0e4), spectral_centre=210.0,
                  stokes_axis=stokes, freq_channels=freq)
    return Martini(source=_compact_source(), datacube=dc,
                   beam=GaussianBeam(bmaj=45.0, bmin=30.0, bpa=20.0) if beam else None,
                   sph_kernel=CubicSplineKernel(), spectral_model=GaussianSpectrum(sigma="thermal"),
                   quiet=True)


@pytest.mark.parametrize("stokes", [False, True])
@pytest.mark.parametrize("freq", [False, True])
def test_cube_header_cards(stokes, freq):
    m = _martini(stokes=stokes, freq=freq)
    dc = m.datacube
    dc.drop_pad()
    cards = fits.cube_cards(dc, m.beam, "M" * 20, -64, -1.5, 2.5)
    hdr = {}
    for c in cards:
        assert len(c) == 80
        k = c[:8].strip()
        v, _ = fits._parse_value(c[10:])
        hdr[k] = v
    keys = [c[:8].strip() for c in cards]
    # the reference's order (martini.py:984-1043)
    expect = ["SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "NAXIS3"] + (
        ["NAXIS4"] if stokes else []) + ["EXTEND"] + [
        f"{k}{i}" for i in (1, 2, 3) for k in ("CDELT", "CRPIX", "CRVAL", "CTYPE", "CUNIT")] + (
        ["CDELT4", "CRPIX4", "CRVAL4", "CTYPE4", "CUNIT4"] if stokes else []) + [
        "EPOCH", "INSTRUME", "BSCALE", "BZERO", "DATAMAX", "DATAMIN", "ORIGIN", "OBJECT", "BPA",
        "OBSERVER", "BUNIT", "DATE-OBS", "MJD-OBS", "BMAJ", "BMIN", "BTYPE", "SPECSYS", "RESTFRQ"]
    assert keys == expect
    assert hdr["NAXIS"] == (4 if stokes else 3)
    assert (hdr["NAXIS1"], hdr["NAXIS2"], hdr["NAXIS3"]) == (40, 36, 24)
    assert hdr["OBJECT"] == "M" * 16
    assert hdr["CTYPE1"] == "RA---TAN" and hdr["CTYPE2"] == "DEC--TAN"
    assert hdr["CTYPE3"] == ("FREQ" if freq else "VRAD")
    assert hdr["CUNIT3"] == ("Hz" if freq else "m s-1")
    # crpix = n/2 + 0.5 after the pad is dropped; the sky steps in degrees
    assert hdr["CRPIX1"] == 20.5 and hdr["CRPIX2"] == 18.5 and hdr["CRPIX3"] == 12.5
    assert hdr["CDELT1"] == pytest.approx(-15.0 / 3600, rel=1e-15)
    assert hdr["CDELT2"] == pytest.approx(15.0 / 3600, rel=1e-15)
    assert hdr["CDELT3"] == (2.0e4 if freq else -5000.0)
    assert hdr["BMAJ"] == pytest.approx(45.0 / 3600) and hdr["BMIN"] == pytest.approx(30.0 / 3600)
    assert hdr["BPA"] == 20.0
    assert hdr["DATAMIN"] == -1.5 and hdr["DATAMAX"] == 2.5
    assert hdr["BUNIT"] == "pix-2 Jy"  # before insertion the array is in Jy/pix^2
    assert hdr["RESTFRQ"] == 1420405751.0 and hdr["SPECSYS"] == "ICRS"


def test_beam_fits_file(tmp_path):
    """write_beam_fits needs no GPU: the beam image is host data."""
    m = _martini()
    path = str(tmp_path / "beam")
    m.write_beam_fits(path)
    cards, data = _check_standard(path + ".fits")
    hdr = {k: v for k, v, _ in cards}
    k = m.beam.kernel
    assert data.shape == (1, k.shape[1], k.shape[0])
    np.testing.assert_array_equal(data[0].T, k)  # kernel[..., newaxis].T, bit for bit
    assert hdr["NAXIS1"] == k.shape[0] and hdr["CRPIX1"] == k.shape[0] // 2 + 1
    assert hdr["DATAMAX"] == k.max() and hdr["OBJECT"] == "MOCKBEAM"
    with pytest.raises(ValueError):
        _martini(beam=False).write_beam_fits(path)


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("cube_bits,bitpix", [(32, -32), (32, -64), (64, -64), (64, -32)])
@pytest.mark.parametrize("shape,pad", [((70, 45, 67), (3, 5)), ((64, 64, 64), (0, 0)),
                                       ((1, 1, 1), (0, 2))])
def test_cube_to_fits_kernel(cube_bits, bitpix, shape, pad):
    import torch

    from martini_amd import _lib

    nx, ny, nc = shape
    px, py = pad
    rng = np.random.default_rng(nx + 7 * nc)
    full = rng.normal(size=(nx + 2 * px, ny + 2 * py, nc))
    full = full.astype(np.float32 if cube_bits == 32 else np.float64)
    d_cube = torch.from_numpy(full).cuda()
    d_out = torch.empty(nx * ny * nc * (-bitpix // 8), dtype=torch.uint8, device="cuda")
    d_range = torch.empty(2, dtype=torch.float64, device="cuda")
    _lib.cube_to_fits(d_cube, nx, ny, nc, px, py, bitpix, d_out, d_range)
    got = np.frombuffer(d_out.cpu().numpy().tobytes(), ">f4" if bitpix == -32 else ">f8")
    inner = full[px:px + nx, py:py + ny, :]
    ref = np.ascontiguousarray(inner.T).astype(np.float32 if bitpix == -32 else np.float64)
    np.testing.assert_array_equal(got.reshape(nc, ny, nx), ref)  # exact: a pure reshuffle
    lo, hi = d_range.cpu().tolist()
    assert lo == float(inner.min()) and hi == float(inner.max())


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [None, np.float32, np.float64])
def test_write_fits_data_and_dtype(tmp_path, dtype):
    """The written data unit is the unpadded cube transposed (the
    reference's ``_array.T``), in the requested dtype (float64 by default)."""
    m = _martini()
    m.insert_source_in_cube()
    m.convolve_beam()
    path = str(tmp_path / "cube")
    m.write_fits(path, dtype=dtype)
    cards, data = _check_standard(path + ".fits")
    hdr = {k: v for k, v, _ in cards}
    want = np.float64 if dtype is None else dtype
    assert data.dtype == np.dtype(want).newbyteorder(">")
    arr = m.datacube._array
    np.testing.assert_array_equal(data, arr.T.astype(want))
    # the reference takes DATAMAX / DATAMIN from the FP64 array, before the cast
    assert hdr["DATAMAX"] == float(arr.max()) and hdr["DATAMIN"] == float(arr.min())
    assert hdr["BUNIT"] == "beam-1 Jy"


@pytest.mark.gpu
def test_write_fits_recovers_hi_mass(tmp_path):
    """The HI mass from the file's data and header alone (as
    tests/test_martini.py:86-130 does with astropy) matches the source's."""
    m = _martini()
    m.insert_source_in_cube()
    m.convolve_beam()
    path = str(tmp_path / "cube.fits")
    m.write_fits(path)
    cards, data = fits.read_fits(path)
    hdr = {k: v for k, v, _ in cards}
    assert hdr["BUNIT"] == "beam-1 Jy" and hdr["CUNIT1"] == "deg" and hdr["CUNIT3"] == "m s-1"
    irad = data.T.sum((0, 1)).squeeze()  # Jy/beam per channel
    beam_area = np.pi * (hdr["BMAJ"] * 3600) * (hdr["BMIN"] * 3600) / 4 / np.log(2)  # arcsec^2
    px_area = abs(hdr["CDELT1"] * 3600) * abs(hdr["CDELT2"] * 3600)  # arcsec^2
    flux = irad / beam_area * px_area  # Jy per channel
    dv = abs(hdr["CDELT3"]) / 1e3  # km/s (VRAD channels are evenly spaced)
    mhi = np.sum(2.36e5 * m.source.distance**2 * flux * dv)
    assert mhi == pytest.approx(m.source.mHI_g.sum(), rel=1e-2)
