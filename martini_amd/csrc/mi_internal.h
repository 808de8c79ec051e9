// Internal interfaces between the library's translation units: workspace
// layout and the host-side launchers of each stage.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/martini_mi355x.h"
#include "mi_device.h"

namespace mi {

struct Grid {
    int nx, ny, nc;      // padded cube extent
    bool wide;           // cell shape: GeoWide, else GeoSquare (wide_cells)
    int tx_shift, ty_shift, cb_shift;  // log2 of the tile sides and the block
    int ntx, nty, ncb;   // tiles along x, y; channel blocks
    int64_t ncells;
    int cell_bits;       // bits needed for a cell id
};

Grid make_grid(const mi_cube_desc& d);

// Number of cells particle p lands in (0 if its footprint or window is empty).
__host__ __device__ inline uint32_t cells_of(const mi_particle& p, const Grid& g, int& tx0, int& tx1,
                                    int& ty0, int& ty1, int& cb0, int& cb1) {
    int i_lo, i_hi, j_lo, j_hi;
    if (!footprint(p, g.nx, g.ny, i_lo, i_hi, j_lo, j_hi)) return 0;
    int c_lo = p.c_lo < 0 ? 0 : p.c_lo;
    int c_hi = p.c_hi > g.nc - 1 ? g.nc - 1 : p.c_hi;
    if (c_lo > c_hi) return 0;
    tx0 = i_lo >> g.tx_shift;  // footprints and windows are clipped: non-negative
    tx1 = i_hi >> g.tx_shift;
    ty0 = j_lo >> g.ty_shift;
    ty1 = j_hi >> g.ty_shift;
    cb0 = c_lo >> g.cb_shift;
    cb1 = c_hi >> g.cb_shift;
    return (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1) * (cb1 - cb0 + 1));
}


// Byte offsets of the workspace regions (each 256-byte aligned).
struct Layout {
    size_t counts, offsets;            // uint32[emit blocks + 2] each: block totals, bases
    size_t keys, vals, keys2, vals2;   // uint32[n_pairs] each; vals2: particle index in cell order
    size_t cell_start, cell_end;       // uint32[ncells] each
    size_t item_off, multi_off;        // uint32[ncells + 1] each: exclusive scans of the
                                       // items and the partial-tile slots per cell
    size_t items;                      // int4[items_max], in cell order
    size_t item_keys, item_keys2;      // uint32[items_max]: sort keys of the launch order
    size_t item_idx, item_order;       // uint32[items_max]: item index; launch order
    size_t partials;                   // float[multi_max * CELL_VOXELS]
    size_t part_masks;                 // uint64[multi_max]: the sub-tiles each partial holds
    size_t status;                     // uint32[8]
    size_t temp;                       // rocPRIM temporary storage
    size_t temp_bytes;
    size_t total;
    int64_t items_max, multi_max;
    int item_len;                      // particles per work item
};

// status[0]: number of pairs the device produced; status[1]: overflow flag.
// status[4]: a record with a footprint has an unknown kernel kind or table
// weights beyond the descriptor's table (no work items are then made).
enum { ST_PAIRS = 0, ST_OVERFLOW = 1, ST_ITEMS = 2, ST_MULTI = 3, ST_BADREC = 4 };

hipError_t layout_for(const Grid& g, int64_t n, int64_t n_pairs, Layout* L);

// Binning: count, scan, emit, sort, cell bounds, items.  Asynchronous.
// record_order: a cell's pairs keep the record order (the MFMA form: records
// laid out in tile and pixel Z order make its chunks compact) and are sorted on
// the cell and kernel group only; otherwise the keys end in a hash of the
// particle index, which spreads each chunk over its tile, and are sorted on
// the top 16 bits (two radix passes) when the key is longer.
hipError_t bin_particles(const mi_particle* d_p, int64_t n, int64_t n_pairs, const Grid& g,
                         const mi_cube_desc& d, char* ws, const Layout& L, bool record_order,
                         hipStream_t s);

// Gather over work items (plain FP32 or MFMA), then the partial-tile reduction.
hipError_t gather_plain(const mi_particle* d_p, const Grid& g, const mi_cube_desc& d, char* ws,
                        const Layout& L, float* d_cube, hipStream_t s);
hipError_t gather_mfma(const mi_particle* d_p, const Grid& g, const mi_cube_desc& d, char* ws,
                       const Layout& L, float* d_cube, hipStream_t s);
hipError_t reduce_partials(const Grid& g, char* ws, const Layout& L, float* d_cube,
                           hipStream_t s);

// Devices the library keeps per-device state for (tables, scratch pointers).
constexpr int MAX_DEVICES = 64;

// The per-device weight tables (WTAB_KERNELS x WTAB_N), filled by fill_wtab on
// the device's first use: launched on stream s, which is then synchronised.
hipError_t weight_table(hipStream_t s, const float4** out);
hipError_t launch_weights(int kind, float gt, float ginv, const float* dx, const float* dy,
                          const float* h, float* out, int64_t n, hipStream_t s);
hipError_t launch_spectra(const mi_particle* d_p, int64_t n, const mi_cube_desc& d, float* out,
                          hipStream_t s);

// Device FP64 setup (mi_prepare.hip): workspace layout and launcher.
struct PrepLayout {
    size_t keep, off, temp, temp_bytes, total;
};
hipError_t prep_layout(int64_t n, PrepLayout* L);
hipError_t prepare_particles(const double* xyz, const double* vxyz, const double* mHI,
                             const double* hsm, const double* T, const double* edges, int64_t n,
                             const mi_observation& o, const Grid& g, mi_particle* out,
                             unsigned long long* counts, char* ws, const PrepLayout& L,
                             hipStream_t s);

// Beam convolution of a cube (mi_beam.hip).
size_t beam_lds_bytes(int kx, int ky);
hipError_t cube_to_fits(const void* cube, bool cube_f64, int nx, int ny, int nc, int padx,
                        int pady, bool f64, void* out, double* range, hipStream_t s);
hipError_t convolve_beam(const float* in, float* out, int nx, int ny, int nc, const float* K,
                         int kx, int ky, hipStream_t s);

// One work item: particles [begin, end) of the sorted pair list belonging to
// one cell; slot >= 0 names its partial tile, -1 means it writes the cube.
struct Item {
    int32_t cell, begin, end, slot;
};

}  // namespace mi
