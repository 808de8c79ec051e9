"""Records' HBM layout: the fixed pseudo-random order applied at upload
(martini_amd.insertion.shuffle_records), checked on CPU tensors."""
import numpy as np
import pytest
import torch

from martini_amd import _lib
from martini_amd.insertion import layout_records, record_permutation, shuffle_records


def _fmix32(x):
    """murmur3's 32-bit finaliser in numpy uint64 arithmetic."""
    m = np.uint64(0xFFFFFFFF)
    x = x.astype(np.uint64) & m
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x85EBCA6B)) & m
    x ^= x >> np.uint64(13)
    x = (x * np.uint64(0xC2B2AE35)) & m
    return x ^ (x >> np.uint64(16))


@pytest.mark.parametrize("n", [2, 1000, 65537])
def test_record_permutation_is_the_hash_order(n):
    """The permutation is the index order of a bijective 32-bit hash of the
    index, restated here in numpy: the same order on any device."""
    k = np.arange(n, dtype=np.uint64)
    keys = _fmix32(_fmix32(k) ^ np.uint64(0x6D617274))
    assert np.unique(keys).size == n
    want = np.argsort(keys, kind="stable")
    assert np.array_equal(record_permutation(n, "cpu").numpy(), want)


def _morton(x, y):
    """Z order of a pixel (x, y) in a tile of up to 16 x 16, bit by bit: x on even bits."""
    z = 0
    for b in range(4):
        z |= ((x >> b) & 1) << (2 * b) | ((y >> b) & 1) << (2 * b + 1)
    return z


@pytest.mark.parametrize("nc,tx", [(64, 16), (128, 8), (None, 16)])
def test_mfma_layout_is_tile_then_z_order(nc, tx):
    """The MFMA form's records are sorted by the binning's tile (x-major; 16 x 16
    pixels below 128 channels, 8 x 16 from 128 on) and by the Z order of the
    pixel inside it, ties in the given order: checked against a plain Python
    sort of the same keys, negative pixels included."""
    rng = np.random.default_rng(9)
    n = 5000
    rec = np.zeros(n, dtype=_lib.PARTICLE_DTYPE)
    rec["ix"] = rng.integers(-40, 300, n)
    rec["iy"] = rng.integers(-40, 300, n)
    rec["k0"] = np.arange(n)  # the given position
    out = layout_records(torch.from_numpy(rec.view(np.uint8).copy()), "mfma", nc).numpy()
    got = out.view(_lib.PARTICLE_DTYPE)
    want = sorted(range(n), key=lambda i: (int(rec["ix"][i]) // tx, int(rec["iy"][i]) // 16,
                                           _morton(int(rec["ix"][i]) % tx, int(rec["iy"][i]) % 16),
                                           i))
    assert np.array_equal(got["k0"], np.array(want))
    assert np.array_equal(got, rec[want])
    assert not torch.equal(layout_records(torch.from_numpy(rec.view(np.uint8).copy()), "plain"),
                           torch.from_numpy(rec.view(np.uint8).copy()))


def _shuffled_order(n):
    """The order shuffle_records gives n records (record j holds given record
    order[j]), read back from records that carry their own index."""
    rec = np.zeros(n, dtype=_lib.PARTICLE_DTYPE)
    rec["ix"] = np.arange(n)
    out = shuffle_records(torch.from_numpy(rec.view(np.uint8).copy())).numpy()
    return out.view(_lib.PARTICLE_DTYPE)["ix"].astype(np.int64)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 100, 4096, 12345])
def test_shuffle_is_a_fixed_permutation_of_whole_records(n):
    rng = np.random.default_rng(n)
    rec = np.zeros(n, dtype=_lib.PARTICLE_DTYPE)
    rec["ix"] = np.arange(n)
    rec["amp"] = rng.random(n).astype(np.float32)
    rec["k0"] = rng.integers(-1000, 1000, n)
    rec["r_kind"] = rng.integers(0, 1 << 20, n)
    out = shuffle_records(torch.from_numpy(rec.view(np.uint8).copy())).numpy()
    got = out.view(_lib.PARTICLE_DTYPE)
    assert got.size == n
    assert np.array_equal(np.sort(got["ix"]), np.arange(n))  # a permutation
    assert np.array_equal(got, rec[got["ix"]])  # of whole records
    again = shuffle_records(torch.from_numpy(rec.view(np.uint8).copy())).numpy()
    assert np.array_equal(out, again)  # the same order every time
    if n > 100:
        assert np.count_nonzero(got["ix"] == np.arange(n)) < n // 10  # not the identity


def chunk_band_imbalance(ix, iy, order):
    """Mean over the FP32 gather's 32-particle chunks (and over x and y) of the
    largest share of a chunk in one 4-pixel band of its 16-pixel tile, relative
    to an even share (1 = even, 4 = all in one band).  Records are laid out by
    `order`; the binning's coarse sort is emulated for a 256 x 256 px grid of
    16 x 16 px cells: key (cell << 10) | (Fibonacci hash of the record index) >> 24,
    8 cell bits, sorted stably on its top 16 of 18 bits."""
    x, y = ix[order], iy[order]
    n = x.size
    j = np.arange(n, dtype=np.uint64)
    h = ((j * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) >> np.uint64(24)
    cell = (x // 16) * 16 + y // 16
    key = (cell.astype(np.uint64) << np.uint64(10)) | h
    o = np.argsort(key >> np.uint64(2), kind="stable")
    cell, x, y = cell[o], x[o], y[o]
    start = np.flatnonzero(np.r_[True, cell[1:] != cell[:-1]])
    first = np.repeat(start, np.diff(np.r_[start, n]))
    pos = np.arange(n) - first
    full = pos < (np.repeat(np.diff(np.r_[start, n]), np.diff(np.r_[start, n])) // 32) * 32
    chunk = (first + pos // 32 * 32)[full]
    _, chunk = np.unique(chunk, return_inverse=True)
    shares = []
    for coord in (x[full], y[full]):
        counts = np.bincount(chunk * 4 + (coord % 16) // 4, minlength=4 * (chunk.max() + 1))
        shares.append(counts.reshape(-1, 4).max(axis=1) / 8.0)
    return float(np.mean(shares))


def test_shuffle_spreads_spatially_ordered_chunks():
    """Spatially ordered input (sorted by pixel column, then row) gives chunks
    that crowd one band of the tile; the shuffled layout restores the balance
    of random order.  A golden-ratio sweep of the records (structured like the
    index hash) restores only part of it, and is here as the case this test
    tells apart."""
    rng = np.random.default_rng(5)
    n = 2_000_000  # about 120 particles per cell and hash bucket, as on a 1e6-particle disc
    ix, iy = rng.integers(0, 256, n), rng.integers(0, 256, n)
    spatial = np.lexsort((iy, ix))
    random = chunk_band_imbalance(ix, iy, rng.permutation(n))
    crowded = chunk_band_imbalance(ix, iy, spatial)
    shuffled = chunk_band_imbalance(ix, iy, spatial[_shuffled_order(n)])
    t = int(round(n * 0.6180339887498949)) + 1  # coprime to 2e6
    golden = chunk_band_imbalance(ix, iy, spatial[(np.arange(n) * t) % n])
    print(f"random {random:.3f} spatial {crowded:.3f} shuffled {shuffled:.3f} "
          f"golden sweep {golden:.3f}")
    assert crowded > 1.3 * random
    assert shuffled <= 1.02 * random
    assert golden > 1.2 * random
