"""Martini.write_hdf5 (martini/martini.py:1154-1366): the reference's datasets,
attributes and grids, read back through the HDF5 C library and, where the
image holds one, through h5py itself (the library the reference writes with),
in a separate interpreter."""

import json
import os
import subprocess

import numpy as np
import pytest

from martini_amd import DataCube, Martini, hdf5, synthetic
from martini_amd.beams import GaussianBeam
from oracle import cube_c
from oracle import martini_oracle as O

H5PY_PYTHON = "/opt/conda/bin/python3.9"

CUBE_ATTRS = ["AxisOrder", "FluxCubeUnit", "deltaRA_in_RAUnit", "RA0_in_px", "RA0_in_RAUnit",
              "RAUnit", "RAProjType", "deltaDec_in_DecUnit", "Dec0_in_px", "Dec0_in_DecUnit",
              "DecUnit", "DecProjType", "deltaV_in_VUnit", "V0_in_px", "V0_in_VUnit", "VUnit",
              "VProjType", "SpecSys", "BeamPA", "BeamMajor_in_deg", "BeamMinor_in_deg",
              "DateCreated", "MartiniVersion", "AstropyVersion"]
BEAM_ATTRS = ["BeamUnit", "deltaRA_in_RAUnit", "RA0_in_px", "RA0_in_RAUnit", "RAUnit",
              "RAProjType", "deltaDec_in_DecUnit", "Dec0_in_px", "Dec0_in_DecUnit", "DecUnit",
              "DecProjType", "deltaV_in_VUnit", "V0_in_px", "V0_in_VUnit", "VUnit", "VProjType",
              "BeamPA", "BeamMajor_in_deg", "BeamMinor_in_deg", "DateCreated", "MartiniVersion",
              "AstropyVersion"]
GRIDS = ["RA", "Dec", "channel_mids", "RA_vertices", "Dec_vertices", "channel_vertices"]
CHANNELS = ["velocity_channel_mids", "velocity_channel_edges", "frequency_channel_mids",
            "frequency_channel_edges"]


def _martini(n=2000, beam=True, freq=False, stokes=False):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, s = synthetic.build(w, n)
    if freq or stokes:
        dc = DataCube(n_px_x=w.n_px, n_px_y=w.n_px, n_channels=w.n_channels, px_size=w.px_size,
                      channel_width=(w.channel_width * 1e3 / 299792458.0 * 1.420405751e9
                                     if freq else w.channel_width),
                      spectral_centre=w.vsys, freq_channels=freq, stokes_axis=stokes)
    b = GaussianBeam(bmaj=30.0, bmin=20.0, bpa=10.0) if beam else None
    return w, Martini(source=src, datacube=dc, beam=b, sph_kernel=k, spectral_model=s, quiet=True)


def _read(path, names, attrs):
    with hdf5.open_file(path) as f:
        return {n: f.read(n, attrs.get(n, ())) for n in names}


@pytest.mark.parametrize("freq", [False, True])
def test_hdf5_grids(tmp_path, freq):
    """The reference's test_hdf5_grids (tests/test_martini.py:742-800): the
    written voxel-centre and vertex coordinates map back to the pixel grid
    (0-based) and to it minus one half, through the cube's WCS."""
    _, m = _martini(freq=freq)
    path = str(tmp_path / "cube")
    m.write_hdf5(path)
    assert os.path.exists(path + ".hdf5")
    d = _read(path + ".hdf5", GRIDS, {g: ["Unit"] for g in GRIDS})
    dc = m.datacube
    assert d["RA"].attrs["Unit"] == "deg" and d["Dec"].attrs["Unit"] == "deg"
    assert d["channel_mids"].attrs["Unit"] == ("Hz" if freq else "m s-1")
    for suffix, off, shape in (("", 0.0, (dc.n_px_x, dc.n_px_y, dc.n_channels)),
                               ("_vertices", -0.5, (dc.n_px_x + 1, dc.n_px_y + 1,
                                                    dc.n_channels + 1))):
        ra = d["RA" + suffix].data
        dec = d["Dec" + suffix].data
        spec = d[("channel_mids" if not suffix else "channel_vertices")].data
        assert ra.shape == dec.shape == spec.shape == shape
        # world -> pixel through world_to_pixel, which takes km/s (radio velocity)
        v_kms = (299792458.0 * (1.0 - spec / 1.420405751e9) if freq else spec) / 1e3
        px = dc.world_to_pixel(ra.ravel(), dec.ravel(), v_kms.ravel())
        expect = np.meshgrid(*(np.arange(n) + off for n in shape), indexing="ij")
        for axis in range(3):
            np.testing.assert_allclose(px[axis].reshape(shape), expect[axis], rtol=0, atol=1e-9)


def test_hdf5_layout_and_attributes(tmp_path):
    """Dataset names, shapes and attributes (names, values, types) as
    martini.py:1209-1360 writes them."""
    _, m = _martini()
    m.datacube._array[...] = np.random.default_rng(1).random(m.datacube._array.shape)
    dc = m.datacube
    pad = (dc.padx, dc.pady)
    path = str(tmp_path / "cube.hdf5")
    m.write_hdf5(path)
    assert (dc.padx, dc.pady) == (0, 0) and pad != (0, 0)  # written without the pad
    d = _read(path, ["FluxCube", "Beam"] + CHANNELS,
              {"FluxCube": CUBE_ATTRS, "Beam": BEAM_ATTRS, **{c: ["Unit"] for c in CHANNELS}})
    assert np.array_equal(d["FluxCube"].data, dc._array)
    a = d["FluxCube"].attrs
    assert a["AxisOrder"] == "(RA,Dec,Channels)" and a["FluxCubeUnit"] == "Jy / pix2"
    assert a["RA0_in_px"] == dc.crpix[0] - 1 and a["Dec0_in_px"] == dc.crpix[1] - 1
    assert a["V0_in_px"] == dc.crpix[2] - 1
    assert (a["deltaRA_in_RAUnit"], a["deltaDec_in_DecUnit"], a["deltaV_in_VUnit"]) == tuple(dc.cdelt)
    assert (a["RA0_in_RAUnit"], a["Dec0_in_DecUnit"], a["V0_in_VUnit"]) == tuple(dc.crval)
    assert (a["RAUnit"], a["DecUnit"], a["VUnit"]) == ("deg", "deg", "m s-1")
    assert (a["RAProjType"], a["DecProjType"], a["VProjType"]) == ("RA---TAN", "DEC--TAN", "VRAD")
    assert a["SpecSys"] == "ICRS"
    assert a["BeamPA"] == 10.0 and a["BeamMajor_in_deg"] == 30.0 / 3600.0
    assert a["BeamMinor_in_deg"] == 20.0 / 3600.0
    for k in ("RA0_in_px", "deltaRA_in_RAUnit", "BeamPA"):
        assert isinstance(a[k], np.floating)
    b = d["Beam"].attrs
    k = m.beam.kernel
    assert np.array_equal(d["Beam"].data, np.asarray(k)[..., None])
    assert b["BeamUnit"] == ""
    assert (b["RA0_in_px"], b["Dec0_in_px"], b["V0_in_px"]) == (k.shape[0] // 2, k.shape[1] // 2, 0)
    for key in ("RA0_in_px", "Dec0_in_px", "V0_in_px"):
        assert isinstance(b[key], np.integer)  # Python ints, as h5py stores them
    assert b["DateCreated"] == a["DateCreated"]
    for name, unit in zip(CHANNELS, ("m / s", "m / s", "Hz", "Hz")):
        assert d[name].attrs["Unit"] == unit
        np.testing.assert_array_equal(d[name].data, getattr(dc, name))


def test_hdf5_options(tmp_path):
    _, m = _martini(beam=False)
    path = str(tmp_path / "c.hdf5")
    m.write_hdf5(path, compact=True)
    with hdf5.open_file(path) as f:
        a = f.read("FluxCube", ["AxisOrder"]).attrs
        assert a["AxisOrder"] == "(RA,Dec,Channels)"
        with pytest.raises(OSError):
            f.read("RA")  # compact: no grids
        with pytest.raises(OSError):
            f.read("Beam")  # no beam
    with pytest.raises(FileExistsError):
        m.write_hdf5(path, overwrite=False)
    m.write_hdf5(path, overwrite=True)
    mem = m.write_hdf5(str(tmp_path / "mem"), memmap=True)
    try:
        assert not os.path.exists(str(tmp_path / "mem.hdf5"))
        assert mem.read("FluxCube").data.shape == (64, 64, 32)
    finally:
        mem.close()
    _, m = _martini()
    m.beam.kernel = None
    with pytest.raises(ValueError, match="beam kernel uninitialized"):
        m.write_hdf5(str(tmp_path / "nokernel"))


def test_hdf5_stokes_axis_squeezed(tmp_path):
    _, m = _martini(beam=False, stokes=True)
    assert m.datacube._array.ndim == 4
    path = str(tmp_path / "s.hdf5")
    m.write_hdf5(path, compact=True)
    with hdf5.open_file(path) as f:
        assert f.read("FluxCube").data.shape == (64, 64, 32)


def test_hdf5_mass_from_file(tmp_path):
    """The reference's check_mass_accuracy on the HDF5 file
    (tests/test_martini.py:146-200), without a beam: a single particle's FP64
    oracle cube (Jy/arcsec^2, the reference's dc_zeros-like 16^3 cube) written
    out, and its 1e4 Msun recovered to 1% from the file's FluxCube, pixel
    steps and velocity channel edges alone."""
    _, m = _martini(beam=False)
    spec = O.CubeSpec(16, 16, 16, 15.0, 4.0, spectral_centre=210.0)
    P = O.prepare(np.full((1, 3), 1e-3), np.zeros((1, 3)), [1e4], [1.0], [1e4], distance_mpc=3.0,
                  cube=spec, kernel=O.adaptive(O.KernelSpec("cubic")))
    dc = DataCube(n_px_x=16, n_px_y=16, n_channels=16, px_size=15.0, channel_width=4.0,
                  spectral_centre=210.0)
    dc._array = O.to_jy_per_arcsec2(cube_c.cube(P), spec)
    dc._array_unit = "Jy/arcsec2"
    m._datacube = dc
    path = str(tmp_path / "mass.hdf5")
    m.write_hdf5(path)
    d = _read(path, ["FluxCube", "velocity_channel_edges"],
              {"FluxCube": CUBE_ATTRS[:18], "velocity_channel_edges": ["Unit"]})
    at = d["FluxCube"].attrs
    assert at["FluxCubeUnit"] == "Jy / arcsec2" and at["RAUnit"] == "deg"
    px_area_arcsec2 = abs(at["deltaRA_in_RAUnit"] * at["deltaDec_in_DecUnit"]) * 3600.0**2
    F = d["FluxCube"].data.sum((0, 1)) * px_area_arcsec2  # Jy per channel
    assert d["velocity_channel_edges"].attrs["Unit"] == "m / s"
    dv = np.abs(np.diff(d["velocity_channel_edges"].data)) / 1e3  # km/s
    MHI = np.sum(2.36e5 * 3.0**2 * F * dv)
    assert MHI == pytest.approx(O.inserted_mass(dc._array, spec, 3.0), rel=1e-12)
    assert MHI == pytest.approx(1e4, rel=1e-2)


_DUMP = r"""
import json, sys, h5py, numpy as np
out = {}
with h5py.File(sys.argv[1], "r") as f:
    for k in f.keys():
        v = f[k][()]
        out[k] = {"shape": list(v.shape), "dtype": str(v.dtype), "sum": float(np.sum(v)),
                  "sq": float(np.sum(np.asarray(v, float) ** 2)),
                  "attrs": {a: (x.decode() if isinstance(x, bytes) else
                                x.item() if hasattr(x, "item") else x)
                            for a, x in f[k].attrs.items()},
                  "types": {a: type(x).__name__ for a, x in f[k].attrs.items()}}
print(json.dumps(out))
"""


@pytest.mark.skipif(not os.path.exists(H5PY_PYTHON), reason="no interpreter with h5py in the image")
def test_h5py_reads_the_same_file(tmp_path):
    """h5py (in the image's separate Python, run as a child process) sees
    every dataset and attribute with the values and types written."""
    _, m = _martini()
    m.datacube._array[...] = np.random.default_rng(2).random(m.datacube._array.shape)
    path = str(tmp_path / "h.hdf5")
    m.write_hdf5(path)
    r = subprocess.run([H5PY_PYTHON, "-s", "-c", _DUMP, path], capture_output=True, text=True,
                       timeout=120, env={"PATH": "/usr/bin:/bin"})
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("h5py is not importable in that interpreter")
    assert r.returncode == 0, r.stderr[-2000:]
    seen = json.loads(r.stdout)
    names = ["FluxCube", "Beam"] + GRIDS + CHANNELS
    assert sorted(seen) == sorted(names)
    attrs = {"FluxCube": CUBE_ATTRS, "Beam": BEAM_ATTRS, **{n: ["Unit"] for n in GRIDS + CHANNELS}}
    mine = _read(path, names, attrs)
    for n in names:
        v = mine[n].data
        assert seen[n]["shape"] == list(v.shape) and seen[n]["dtype"] == "float64"
        assert seen[n]["sum"] == pytest.approx(float(np.sum(v)), rel=1e-15)
        assert seen[n]["sq"] == pytest.approx(float(np.sum(v ** 2)), rel=1e-15)
        assert sorted(seen[n]["attrs"]) == sorted(attrs[n])
        for a in attrs[n]:
            x = mine[n].attrs[a]
            assert seen[n]["attrs"][a] == (x.item() if hasattr(x, "item") else x), (n, a)
            kind = seen[n]["types"][a]
            assert kind == ("str" if isinstance(x, str) else
                            "int64" if isinstance(x, np.integer) else "float64"), (n, a, kind)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_hdf5_mass_after_beam_on_gpu(tmp_path, method):
    """The reference's whole check_mass_accuracy flow in HDF5 mode
    (tests/test_martini.py:20-56, :146-200): one particle inserted and
    convolved with a beam on the GPU, the file written, and the mass recovered
    to 1% from the file's Jy/beam cube, beam axes, pixel steps and channel edges."""
    from martini_amd import SPHSource
    from martini_amd import sph_kernels as K
    from martini_amd.spectral_models import GaussianSpectrum

    src = SPHSource(distance=3.0, T_g=np.array([1e4]), mHI_g=np.array([1e4]),
                    xyz_g=np.full((1, 3), 1e-3), vxyz_g=np.zeros((1, 3)), hsm_g=np.array([1.0]))
    dc = DataCube(n_px_x=16, n_px_y=16, n_channels=16, px_size=15.0, channel_width=4.0,
                  spectral_centre=210.0)
    m = Martini(source=src, datacube=dc, beam=GaussianBeam(bmaj=30.0, bmin=30.0),
                sph_kernel=K.CubicSplineKernel(), spectral_model=GaussianSpectrum(sigma="thermal"),
                quiet=True)
    m.insert_source_in_cube(method=method)
    m.convolve_beam()
    path = str(tmp_path / "beam.hdf5")
    m.write_hdf5(path)
    d = _read(path, ["FluxCube", "velocity_channel_edges"],
              {"FluxCube": CUBE_ATTRS, "velocity_channel_edges": ["Unit"]})
    at = d["FluxCube"].attrs
    assert at["FluxCubeUnit"] == "Jy / beam"
    Irad = d["FluxCube"].data.sum((0, 1))  # Jy/beam
    A = np.pi * at["BeamMajor_in_deg"] * at["BeamMinor_in_deg"] * 3600.0**2 / 4 / np.log(2)
    px_area = abs(at["deltaRA_in_RAUnit"] * at["deltaDec_in_DecUnit"]) * 3600.0**2
    F = Irad / A * px_area  # Jy
    dv = np.abs(np.diff(d["velocity_channel_edges"].data)) / 1e3
    MHI = np.sum(2.36e5 * 3.0**2 * F * dv)
    print(f"{method}: HI mass from the HDF5 file {MHI:.2f} of 1e4")
    assert MHI == pytest.approx(1e4, rel=1e-2)
