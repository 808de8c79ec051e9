"""FITS output of the cube and the beam (martini/martini.py:939-1152).

astropy is not required: the primary HDU is written here to the FITS
standard (version 4.0): 80-character header cards in 2880-byte blocks ending
with an END card, then the big-endian data unit padded with zeros to a whole
block.  The header carries the same cards, in the same order, as the
reference's ``write_fits`` / ``write_beam_fits``; BITPIX and the NAXISn follow
the data, as astropy's PrimaryHDU sets them.

The cube's data unit is made on the GPU (``mi_cube_to_fits``): the unpadded
cube transposed to FITS axis order (x fastest), cast to float32 or float64
and byte-swapped, with DATAMIN and DATAMAX from the same pass.  ``read_fits``
reads such a file back (header cards and data), for tests and for users
without astropy.
"""

from __future__ import annotations

import datetime
import os

import numpy as np

from . import __version__
from .datacube import HIfreq

BLOCK = 2880
CARD = 80

# astropy's FITS strings for the cube's array units
_BUNIT = {"Jy/beam": "beam-1 Jy", "Jy/arcsec2": "arcsec-2 Jy", "Jy/pix2": "pix-2 Jy"}


def _format_value(value) -> str:
    """The value field of a fixed-format card (columns 11-30)."""
    if isinstance(value, (bool, np.bool_)):
        return f"{'T' if value else 'F':>20}"
    if isinstance(value, (int, np.integer)):
        return f"{int(value):>20d}"
    if isinstance(value, (float, np.floating)):
        # the shortest repr that round-trips, else the most digits that fit
        for s in [repr(float(value))] + [f"{float(value):.{p}G}" for p in range(16, 0, -1)]:
            s = s.upper()
            mant, _, exp = s.partition("E")
            if "." not in mant and "N" not in mant:
                mant += ".0"
            s = f"{mant}E{exp}" if exp else mant
            if len(s) <= 20:
                break
        return f"{s:>20}"
    text = str(value).replace("'", "''")
    return f"'{text:<8}'".ljust(20)


def card(key: str, value, comment: str | None = None) -> str:
    """One 80-character header card."""
    if len(key) > 8:
        raise ValueError(f"FITS keyword longer than 8 characters: {key}")
    c = f"{key.upper():<8}= {_format_value(value)}"
    if comment:
        c += f" / {comment}"
    if len(c) > CARD:
        c = c[:CARD]
    return c.ljust(CARD)


def header_bytes(cards: list[str]) -> bytes:
    """Cards, END, and blank padding to a whole number of blocks."""
    text = "".join(cards) + "END".ljust(CARD)
    text += " " * (-len(text) % BLOCK)
    return text.encode("ascii")


def write_hdu(filename: str, cards: list[str], data: bytes, overwrite: bool) -> None:
    """Writes a primary HDU: the header, then the data padded to a block."""
    if os.path.exists(filename) and not overwrite:
        raise OSError(f"File {filename} already exists.")
    with open(filename, "wb") as f:
        f.write(header_bytes(cards))
        f.write(data)
        f.write(b"\0" * (-len(data) % BLOCK))


def _now_fits() -> tuple[str, float]:
    now = datetime.datetime.now(datetime.timezone.utc)
    mjd = (now - datetime.datetime(1858, 11, 17, tzinfo=datetime.timezone.utc)).total_seconds()
    return now.strftime("%Y-%m-%dT%H:%M:%S.") + f"{now.microsecond // 1000:03d}", mjd / 86400.0


def _spectral_axis(dc) -> tuple[str, str]:
    return ("FREQ", "Hz") if dc._freq_channel_mode else ("VRAD", "m s-1")


def cube_cards(dc, beam, obj_name: str, bitpix: int, datamin: float, datamax: float) -> list[str]:
    """The reference's write_fits header (martini.py:984-1043), in its order."""
    ctype3, cunit3 = _spectral_axis(dc)
    crpix, cdelt, crval = dc.crpix, dc.cdelt, dc.crval
    cards = [card("SIMPLE", True), card("BITPIX", bitpix),
             card("NAXIS", 4 if dc.stokes_axis else 3),
             card("NAXIS1", dc.n_px_x), card("NAXIS2", dc.n_px_y), card("NAXIS3", dc.n_channels)]
    if dc.stokes_axis:
        cards.append(card("NAXIS4", 1))
    cards += [card("EXTEND", True),
              card("CDELT1", float(cdelt[0])), card("CRPIX1", float(crpix[0])),
              card("CRVAL1", float(crval[0])), card("CTYPE1", "RA---TAN"), card("CUNIT1", "deg"),
              card("CDELT2", float(cdelt[1])), card("CRPIX2", float(crpix[1])),
              card("CRVAL2", float(crval[1])), card("CTYPE2", "DEC--TAN"), card("CUNIT2", "deg"),
              card("CDELT3", float(cdelt[2])), card("CRPIX3", float(crpix[2])),
              card("CRVAL3", float(crval[2])), card("CTYPE3", ctype3), card("CUNIT3", cunit3)]
    if dc.stokes_axis:
        cards += [card("CDELT4", 1.0), card("CRPIX4", 1.0), card("CRVAL4", 1.0),
                  card("CTYPE4", "STOKES"), card("CUNIT4", "PAR")]
    if len(obj_name) > 16:
        import warnings

        warnings.warn("obj_name longer than 16 characters, truncating", RuntimeWarning)
        obj_name = obj_name[:16]
    date_obs, mjd_obs = _now_fits()
    cards += [card("EPOCH", 2000), card("INSTRUME", "MARTINI", f"martini_amd {__version__}"),
              card("BSCALE", 1.0), card("BZERO", 0.0),
              card("DATAMAX", float(datamax)), card("DATAMIN", float(datamin)),
              card("ORIGIN", f"martini_amd v{__version__}"), card("OBJECT", obj_name)]
    if beam is not None:
        cards.append(card("BPA", float(beam.bpa)))
    cards += [card("OBSERVER", "K. Oman"), card("BUNIT", _BUNIT[dc._array_unit]),
              card("DATE-OBS", date_obs), card("MJD-OBS", mjd_obs)]
    if beam is not None:
        cards += [card("BMAJ", float(beam.bmaj) / 3600.0), card("BMIN", float(beam.bmin) / 3600.0)]
    cards += [card("BTYPE", "Intensity"), card("SPECSYS", dc.specsys.upper()),
              card("RESTFRQ", float(HIfreq))]
    return cards


def beam_cards(dc, beam, bitpix: int) -> list[str]:
    """The reference's write_beam_fits header (martini.py:1103-1142), in its order."""
    ctype3, cunit3 = _spectral_axis(dc)
    cdelt, crval = dc.cdelt, dc.crval
    k = beam.kernel
    return [card("SIMPLE", True), card("BITPIX", bitpix), card("NAXIS", 3),
            card("NAXIS1", k.shape[0]), card("NAXIS2", k.shape[1]), card("NAXIS3", 1),
            card("EXTEND", True), card("BSCALE", 1.0), card("BZERO", 0.0),
            card("BUNIT", ""),  # the beam image is dimensionless (beams.py:150-157)
            card("CRPIX1", k.shape[0] // 2 + 1), card("CDELT1", float(cdelt[0])),
            card("CRVAL1", float(crval[0])), card("CTYPE1", "RA---TAN"), card("CUNIT1", "deg"),
            card("CRPIX2", k.shape[1] // 2 + 1), card("CDELT2", float(cdelt[1])),
            card("CRVAL2", float(crval[1])), card("CTYPE2", "DEC--TAN"), card("CUNIT2", "deg"),
            card("CRPIX3", 1), card("CDELT3", float(cdelt[2])), card("CRVAL3", float(crval[2])),
            card("CTYPE3", ctype3), card("CUNIT3", cunit3), card("SPECSYS", dc.specsys.upper()),
            card("BMAJ", float(beam.bmaj) / 3600.0), card("BMIN", float(beam.bmin) / 3600.0),
            card("BPA", float(beam.bpa)), card("BTYPE", "beam    "), card("EPOCH", 2000),
            card("OBSERVER", "K. Oman"), card("OBJECT", "MOCKBEAM"),
            card("INSTRUME", "MARTINI", f"martini_amd {__version__}"),
            card("DATAMAX", float(np.max(k))), card("DATAMIN", float(np.min(k))),
            card("ORIGIN", f"martini_amd v{__version__}")]


def fits_filename(filename: str) -> str:
    """write_fits's rule (martini.py:980): append '.fits' unless a 'fits' suffix is there."""
    filename = str(filename)
    return filename + ".fits" if "fits" not in filename.split(".") else filename


def bitpix_for(dtype) -> int:
    dt = np.dtype(np.float64 if dtype is None else dtype)
    if dt == np.float64:
        return -64
    if dt == np.float32:
        return -32
    raise NotImplementedError(f"FITS output as {dt} is not supported (float32 or float64)")


# ----------------------------------------------------------------- reading
def _parse_value(field: str):
    field = field.strip()
    if field.startswith("'"):
        end = 1
        out = []
        while end < len(field):
            ch = field[end]
            if ch == "'":
                if end + 1 < len(field) and field[end + 1] == "'":
                    out.append("'")
                    end += 2
                    continue
                break
            out.append(ch)
            end += 1
        return "".join(out).rstrip(), field[end + 1:]
    value, _, rest = field.partition("/")
    v = value.strip()
    if v == "T":
        return True, "/" + rest if rest else ""
    if v == "F":
        return False, "/" + rest if rest else ""
    try:
        return int(v), "/" + rest if rest else ""
    except ValueError:
        return float(v.replace("D", "E")), "/" + rest if rest else ""


def read_fits(filename: str) -> tuple[list[tuple[str, object, str]], np.ndarray]:
    """Primary HDU of a FITS file: its cards (key, value, comment), END
    excluded, and its data with numpy axes in reverse FITS order."""
    with open(filename, "rb") as f:
        raw = f.read()
    if len(raw) % BLOCK:
        raise ValueError("file length is not a whole number of 2880-byte blocks")
    cards, pos = [], 0
    while True:
        if pos + CARD > len(raw):
            raise ValueError("no END card")
        text = raw[pos:pos + CARD].decode("ascii")
        pos += CARD
        key = text[:8].strip()
        if key == "END":
            break
        if text[8:10] == "= ":
            value, rest = _parse_value(text[10:])
            comment = rest.partition("/")[2].strip() if "/" in rest else ""
        else:
            value, comment = None, text[8:].strip()
        cards.append((key, value, comment))
    pos += -pos % BLOCK
    hdr = {k: v for k, v, _ in cards}
    naxis = hdr["NAXIS"]
    shape = tuple(hdr[f"NAXIS{i}"] for i in range(naxis, 0, -1))
    dtype = {-64: ">f8", -32: ">f4", 8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8"}[hdr["BITPIX"]]
    count = int(np.prod(shape)) if shape else 0
    data = np.frombuffer(raw, dtype=dtype, count=count, offset=pos).reshape(shape)
    return cards, data
