"""Host FP64 setup -> ``mi_particle`` records (the host half of the boundary).

Everything that needs FP64 happens here, once per particle:

* pixel position split into an integer part and an FP32 fraction, so that
  offsets x - i are exact to 2^-24 px on the device;
* the mask half-width r = sm_ranges (martini/martini.py:271-274) and the
  concrete kernel of each particle (adaptive choice, sph_kernels.py:1241-1274),
  with h = h_px * rescale (sph_kernels.py:85-119);
* Dirac-kernel particles resolved to their single pixel with the reference's
  strict |dx| < 0.5 test (sph_kernels.py:1136-1165) in FP64;
* the line-of-sight velocity in channel-edge units u = (v - e0)/de split as
  k0 + fu, because channel widths sit near FP32 epsilon of the velocities
  themselves;
* the channel window outside which the Gaussian line carries less than
  erfc(4)/2 (7.7e-9) of its flux on either side, or the Dirac line's
  channel(s) with the reference's inclusive edge test
  (spectral_models.py:510-570);
* amp = mHI / d^2 / |dv| / 2.36e5 (spectral_models.py:92-145) times the
  output scale (1/px^2 converts Jy/pix^2 to Jy/arcsec^2, martini.py:364-366).
"""

from __future__ import annotations

import numpy as np

from . import _lib

WINDOW_ERF_ARG = 4.0
MHI_TO_JY = 2.36e5
_MAX_INDEX = 1 << 21  # |pixel coordinate| bound
_MAX_RANGE = 1 << 22  # r is stored in 24 bits of r_kind; larger r already covers the cube


def channel_grid(datacube):
    """Velocity edges (m/s), first edge e0 and uniform step de (signed)."""
    e = np.asarray(datacube.velocity_channel_edges, float)
    nc = datacube.n_channels
    return e, e[0], (e[-1] - e[0]) / nc


def spectral_fields(vlos_kms, sigma_kms, mHI, dist_mpc, datacube, spectrum_kind,
                    out_scale=1.0):
    """k0, fu, s, amp, c_lo, c_hi per particle."""
    e, e0, de = channel_grid(datacube)
    nc = datacube.n_channels
    v = np.asarray(vlos_kms, float) * 1e3
    n = v.size
    with np.errstate(invalid="ignore"):
        u = (v - e0) / de
    finite = np.isfinite(u)
    uu = np.where(finite, u, 0.0)
    k0 = np.floor(uu)
    fu = uu - k0
    amp = (np.asarray(mHI, float) * np.asarray(dist_mpc, float) ** -2.0
           / (abs(de) / 1e3) / MHI_TO_JY * out_scale)
    if spectrum_kind == _lib.SPECTRUM_GAUSS:
        sig = np.broadcast_to(np.asarray(sigma_kms, float), (n,)) * 1e3
        with np.errstate(divide="ignore"):
            s = abs(de) / (np.sqrt(2.0) * sig)
            half = WINDOW_ERF_ARG / s
        c_lo = np.floor(uu - 1.0 - half) + 1.0
        c_hi = np.ceil(uu + half) - 1.0
    else:
        s = np.zeros(n)
        lo_e = np.minimum(e[:-1], e[1:])
        hi_e = np.maximum(e[:-1], e[1:])
        c_lo = np.full(n, np.inf)
        c_hi = np.full(n, -np.inf)
        for d in (-1.0, 0.0, 1.0):
            c = np.clip(k0 + d, 0, nc - 1).astype(np.int64)
            inside = (v >= lo_e[c]) & (v <= hi_e[c]) & finite & (k0 + d >= 0) & (k0 + d <= nc - 1)
            c_lo = np.where(inside, np.minimum(c_lo, c), c_lo)
            c_hi = np.where(inside, np.maximum(c_hi, c), c_hi)
    c_lo = np.clip(np.where(finite, c_lo, nc), 0, nc)
    c_hi = np.clip(np.where(finite, c_hi, -1), -1, nc - 1)
    empty = c_lo > c_hi
    c_lo = np.where(empty, 1, c_lo)
    c_hi = np.where(empty, 0, c_hi)
    return (k0.astype(np.int64), fu, s, amp, c_lo.astype(np.int64), c_hi.astype(np.int64))


# Pairs of (particle, pixel) per call of a host-evaluated kernel integral.
TABLE_CHUNK_PAIRS = 1 << 22


def device_kind(K) -> int:
    """The device form of a concrete kernel: its own kind when it is one of the
    built-in kernels and keeps their integral, else MI_KERNEL_TABLE (a user
    subclass of _BaseSPHKernel, or a subclass that overrides _kernel_integral:
    its integral is evaluated on the host, as the reference evaluates it)."""
    from .sph_kernels import _BaseSPHKernel

    if K._mi_kind >= 0 and type(K)._kernel_integral is _BaseSPHKernel._kernel_integral:
        return int(K._mi_kind)
    return _lib.KERNEL_TABLE


def table_weights(sph_kernel, kernels, which, kinds, x, y, ix, iy, fx, fy, rr, h_eff, nx, ny):
    """Pixel weights of the particles whose kernel has no device form
    (MI_KERNEL_TABLE), from the kernel's own ``_kernel_integral``.

    Each such particle's weights cover its footprint clipped to the cube,
    rows i_lo..i_hi by columns j_lo..j_hi -- exactly the pixels the
    reference's square mask admits (martini/martini.py:271-274) -- row-major;
    offsets[p] is where they start.  The integral is called as the reference
    calls it from the pixel loop (martini.py:275-277): offsets dij = particle -
    pixel centre in pixels, h = sm_lengths * rescale, and for a kernel that is
    not adaptive the particle mask (here the particle index of each pair); an
    adaptive kernel's tiers get no mask (sph_kernels.py:1358-1362).  Returns
    (offsets, FP32 weights)."""
    n = x.size
    offsets = np.zeros(n, np.int64)
    sel = np.flatnonzero((kinds == _lib.KERNEL_TABLE) & (rr >= 0))
    if sel.size == 0:
        return offsets, np.zeros(0, np.float32)
    r = rr[sel].astype(np.int64)
    fx32 = fx[sel].astype(np.float32)
    fy32 = fy[sel].astype(np.float32)
    # the device's footprint (mi_device.h footprint): ceil(x - r) from the FP32 fraction
    i_lo = np.maximum(ix[sel] - r + (fx32 > 0), 0)
    i_hi = np.minimum(ix[sel] + r, nx - 1)
    j_lo = np.maximum(iy[sel] - r + (fy32 > 0), 0)
    j_hi = np.minimum(iy[sel] + r, ny - 1)
    ni = np.maximum(i_hi - i_lo + 1, 0)
    nj = np.maximum(j_hi - j_lo + 1, 0)
    cnt = ni * nj
    starts = np.concatenate(([0], np.cumsum(cnt)))
    total = int(starts[-1])
    if total >= 1 << 32:
        raise ValueError("host-evaluated kernel weights beyond 2^32 entries: insert the source "
                         "in parts")
    offsets[sel] = starts[:-1]
    out = np.zeros(total, np.float32)
    adaptive = len(kernels) > 1 or hasattr(sph_kernel, "kernel_indices")
    a = 0
    while a < sel.size:
        # a run of particles with at most TABLE_CHUNK_PAIRS pairs (at least one particle)
        b = int(np.searchsorted(starts, starts[a] + TABLE_CHUNK_PAIRS, side="right")) - 1
        b = max(b, a + 1)
        b = min(b, sel.size)
        c = cnt[a:b]
        m = int(c.sum())
        if m:
            loc = np.repeat(np.arange(a, b), c)
            l = np.arange(m) - np.repeat(starts[a:b] - starts[a], c)
            row = l // nj[loc]
            i = i_lo[loc] + row
            j = j_lo[loc] + (l - row * nj[loc])
            part = sel[loc]
            dij = np.vstack((x[part] - i, y[part] - j))
            h = h_eff[part]
            w = np.empty(m)
            for ik in np.unique(which[part]):
                km = which[part] == ik
                K = kernels[ik]
                if adaptive:
                    w[km] = np.asarray(K._kernel_integral(dij[:, km], h[km]), float)
                else:
                    w[km] = np.asarray(K._kernel_integral(dij[:, km], h[km], mask=part[km]),
                                       float)
            out[starts[a]:starts[a] + m] = w
        a = b
    return offsets, out


def kernel_fields(sph_kernel, pixcoords, nx, ny):
    """ix, iy, fx, fy, h, r_kind per particle, the Gaussian (t, norm) table and
    the host-evaluated weights of kernels without a device form (None when
    there are none); for those particles h holds their table offset's bits."""
    kernels, which = sph_kernel._device_kernels()
    kinds = np.array([device_kind(K) for K in kernels], np.int64)[which]
    gauss = []
    gidx_of_kernel = np.zeros(len(kernels), np.int64)
    for ik, K in enumerate(kernels):
        if K._mi_kind == _lib.KERNEL_GAUSS:
            key = (float(K.truncate), float(K.norm))
            if key not in gauss:
                gauss.append(key)
            gidx_of_kernel[ik] = gauss.index(key)
    if len(gauss) > 4:
        raise ValueError("at most 4 distinct Gaussian truncations per insertion")
    gidx = gidx_of_kernel[which]
    h_px = np.asarray(sph_kernel.sm_lengths, float)
    h_eff = h_px * np.asarray(sph_kernel._rescale, float)
    r = np.asarray(sph_kernel.sm_ranges, float)
    x = np.asarray(pixcoords[0], float)
    y = np.asarray(pixcoords[1], float)
    with np.errstate(invalid="ignore"):
        alive = (np.isfinite(x) & np.isfinite(y) & (r >= 0) & (x + r >= 0) & (y + r >= 0)
                 & (x - r <= nx - 1) & (y - r <= ny - 1))
    xs = np.where(alive, x, 0.0)
    ys = np.where(alive, y, 0.0)
    if np.any(np.abs(xs) >= _MAX_INDEX) or np.any(np.abs(ys) >= _MAX_INDEX):
        raise ValueError("pixel coordinates beyond 2^21 cannot be represented")
    ix = np.floor(xs)
    iy = np.floor(ys)
    fx = xs - ix
    fy = ys - iy
    rr = np.where(alive, np.minimum(r, float(_MAX_RANGE)), -1.0)
    dirac = kinds == _lib.KERNEL_DIRAC
    if dirac.any():
        # the one pixel with |x - i| < 0.5 (strict) that the mask also admits
        px_i = np.floor(xs + 0.5)
        px_j = np.floor(ys + 0.5)
        ok = ((np.abs(xs - px_i) < 0.5) & (np.abs(ys - px_j) < 0.5)
              & (np.abs(xs - px_i) <= rr) & (np.abs(ys - px_j) <= rr)
              & (px_i >= 0) & (px_i <= nx - 1) & (px_j >= 0) & (px_j <= ny - 1) & alive)
        ix = np.where(dirac, px_i, ix)
        iy = np.where(dirac, px_j, iy)
        fx = np.where(dirac, 0.0, fx)
        fy = np.where(dirac, 0.0, fy)
        rr = np.where(dirac, np.where(ok, 0.0, -1.0), rr)
    r_kind = (rr.astype(np.int64) << 8) | (gidx << 4) | kinds
    ix = ix.astype(np.int64)
    iy = iy.astype(np.int64)
    table = None
    if np.any(kinds == _lib.KERNEL_TABLE):
        offsets, weights = table_weights(sph_kernel, kernels, which, kinds, xs, ys, ix, iy, fx,
                                         fy, rr, h_eff, nx, ny)
        table = (kinds == _lib.KERNEL_TABLE, offsets, weights)
    return ix, iy, fx, fy, h_eff, r_kind, gauss, table


def set_table_offsets(rec: np.ndarray, table) -> None:
    """Write the weight-table offsets of MI_KERNEL_TABLE records into their h
    field, as uint32 bits (mi_particle.h)."""
    use, offsets, _ = table
    words = rec.view(np.uint32).reshape(rec.size, _lib.PARTICLE_DTYPE.itemsize // 4)
    words[use, _lib.PARTICLE_DTYPE.fields["h"][1] // 4] = offsets[use].astype(np.uint32)


def make_records(n, ix, iy, fx, fy, h, r_kind, k0, fu, s, amp, c_lo, c_hi) -> np.ndarray:
    rec = np.empty(n, dtype=_lib.PARTICLE_DTYPE)
    rec["ix"], rec["iy"] = ix, iy
    rec["fx"], rec["fy"] = fx, fy
    rec["h"] = h
    rec["r_kind"] = r_kind
    rec["c_lo"], rec["c_hi"] = c_lo, c_hi
    rec["k0"] = np.clip(k0, -(1 << 30), 1 << 30)
    rec["fu"] = fu
    rec["s"] = s
    rec["amp"] = amp
    return rec


def table_fields(spectral_model, source, datacube, out_scale=1.0):
    """A spectral model without a device form (MI_SPECTRUM_TABLE): its spectra
    evaluated on the host in FP64 (``_BaseSpectrum._host_spectra``) times the
    output scale, as an FP32 N x Nc table; per particle k0 = its table row,
    amp = 1, and [c_lo, c_hi] the first and last channel with a nonzero entry."""
    S = spectral_model._host_spectra(source, datacube) * out_scale
    n, nc = S.shape
    nz = np.isfinite(S) & (S != 0.0)
    S = np.where(nz, S, 0.0)
    has = nz.any(axis=1)
    c_lo = np.where(has, np.argmax(nz, axis=1), 1)
    c_hi = np.where(has, nc - 1 - np.argmax(nz[:, ::-1], axis=1), 0)
    z = np.zeros(n)
    return (np.arange(n, dtype=np.int64), z, z, np.ones(n), c_lo.astype(np.int64),
            c_hi.astype(np.int64), np.ascontiguousarray(S, np.float32))


def pack_particles(source, sph_kernel, spectral_model, datacube, out_scale=1.0):
    """Records and cube descriptor for an initialised (and pruned) source.
    For a table spectrum the descriptor carries the host table as
    ``host_spectra``, for kernels without a device form the host-evaluated
    pixel weights as ``host_weights``; the inserter uploads them and points
    the descriptor at them."""
    nx, ny, nc = datacube._array.shape[:3]
    ix, iy, fx, fy, h, r_kind, gauss, ktable = kernel_fields(sph_kernel, source.pixcoords, nx, ny)
    sc = source.skycoords
    table = None
    if spectral_model._mi_kind == _lib.SPECTRUM_TABLE:
        k0, fu, s, amp, c_lo, c_hi, table = table_fields(spectral_model, source, datacube,
                                                         out_scale)
    else:
        k0, fu, s, amp, c_lo, c_hi = spectral_fields(
            sc.radial_velocity, spectral_model._sigma_kms(source), source.mHI_g, sc.distance,
            datacube, spectral_model._mi_kind, out_scale)
    rec = make_records(source.npart, ix, iy, fx, fy, h, r_kind, k0, fu, s, amp, c_lo, c_hi)
    desc = _lib.make_desc(nx, ny, nc, spectral_model._mi_kind, gauss,
                          spectra_ld=nc if table is not None else 0)
    if table is not None:
        desc.host_spectra = table
    if ktable is not None:
        set_table_offsets(rec, ktable)
        desc.host_weights = ktable[2]
    return rec, desc


def pack_spectral(source, spectral_model, datacube):
    """Records carrying only spectral fields (for init_spectra)."""
    nc = datacube.n_channels
    sc = source.skycoords
    k0, fu, s, amp, c_lo, c_hi = spectral_fields(
        sc.radial_velocity, spectral_model._sigma_kms(source), source.mHI_g, sc.distance,
        datacube, spectral_model._mi_kind, 1.0)
    n = source.npart
    z = np.zeros(n)
    rec = make_records(n, z, z, z, z, z, np.full(n, -256, np.int64), k0, fu, s, amp, c_lo, c_hi)
    return rec, _lib.make_desc(1, 1, nc, spectral_model._mi_kind)
