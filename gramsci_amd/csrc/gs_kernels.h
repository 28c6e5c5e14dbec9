// gs_kernels.h -- launch interfaces of the HIP kernels (host side).
#pragma once

#include "gs_common.h"

namespace gs {

// grid.hip
hipError_t launch_bbox(const double* xyz, int64_t n, double* part, int nblk, hipStream_t st);
// weights: per-block maxima of the data and of the random weights (part_max[2 * block + class]) and
// the first invalid (negative or non-finite) index (atomic min; *bad starts at ~0), then per-block
// 128-bit fixed-point sums over data and randoms ([nblk][4]: data lo, hi, randoms lo, hi)
hipError_t launch_wcheck(const double* w, int64_t n, int64_t n_data, double* part_max, unsigned long long* bad,
                         int nblk, hipStream_t st);
hipError_t launch_wsum(const double* w, int64_t n, int64_t n_data, int k, unsigned long long* part, int nblk,
                       hipStream_t st);
hipError_t launch_cell_count(const double* xyz, int64_t n, const GridDesc& g, int32_t* cell_of, int32_t* count,
                             hipStream_t st);
hipError_t exclusive_scan(const int32_t* in, int32_t* out, int64_t n, int32_t* scratch, hipStream_t st);
int64_t scan_scratch_size(int64_t n);
hipError_t launch_scatter(const double* xyz, const double* w, const int32_t* buf, int64_t n, int64_t n_data,
                          const GridDesc& g, const int32_t* cell_of, const int32_t* start, int32_t* fill,
                          double4* pw, double* ws, float4* pl, uint32_t* flags, int32_t* orig, hipStream_t st);
hipError_t launch_row_chunks(const int32_t* start, int nx, int nrow, int32_t* nch, hipStream_t st);

// batched regions (grid.hip)
hipError_t launch_fill_u32(void* p, int64_t n, uint32_t v, hipStream_t st);
hipError_t launch_cell_count_b(const double* xyz, int64_t n, const int32_t* off, int nreg, const RegionDesc* reg,
                               double h, double hx, int32_t* cell_of, int32_t* count, hipStream_t st);
hipError_t launch_scatter_b(const double* xyz, const double* w, const int32_t* buf, int64_t n, const int32_t* off,
                            int nreg, const RegionDesc* reg, double h, double hx, const int32_t* cell_of,
                            const int32_t* start,
                            int32_t* fill, double4* pw, double* ws, float4* pl, uint32_t* flags, int32_t* orig,
                            hipStream_t st);
hipError_t launch_row_chunks_b(const int32_t* start, const RegionDesc* reg, const int32_t* row_base, int nreg,
                               int nrow_tot, int32_t* nch, hipStream_t st);

// domain decomposition (decompose.hip)
hipError_t launch_region_box(const float* xyz, const int32_t* reg, int64_t n, int nreg, int32_t* keys, hipStream_t st);
hipError_t launch_region_flag(const float* xyz, int64_t n, const float* ext, int32_t* flag, hipStream_t st);
hipError_t launch_region_scatter(const float* xyz, int64_t n, const float* box, const int32_t* flag, const int32_t* pos,
                                 int32_t* idx, int32_t* buf, hipStream_t st);

// Arguments of the persistent 2PCF pair kernel (pair2.hip).
struct PairArgs {
    const double4* pw;          // sorted points: x, y, z, raw weight
    const double* ws;           // sorted raw weights alone (the weighted register tile reads 8 B per entry)
    const float4* pl;           // sorted: cell-local FP32 x, y, z and the cell x index
    const uint32_t* flags;      // sorted: bit0 random, bit1 never-apex
    const int32_t* cell_start;  // ncell+1
    const int32_t* row_chunk;   // nrow+1: first work chunk of each x-row of cells
    const Row* rows;            // half stencil
    const BinDesc* bins;        // device copy
    int32_t* queue;             // 8 work counters (one per XCD segment), zeroed before launch
    unsigned long long* out_cnt;    // [3][nb][nm]
    unsigned long long* out_w;      // [3][nb][nm][2] 128-bit fixed point (lo, hi)
    unsigned long long* out_stats;  // [0] candidates, [1] rechecks, [2] overflow pairs
    GridDesc g;
    int32_t nrows;
    int32_t n_chunks;
    int32_t reps;               // LDS histogram replicas (power of two <= 32)
    int32_t mirror;             // (r, mu) with no buffer points: one update per unordered pair
    // batched regions (nreg > 0): per-region grids; each wave keeps a private
    // histogram tagged with its current region; n_chunks is read from n_chunks_dev
    const RegionDesc* regions;
    const int32_t* row_base;    // nreg+1 first global x-row per region
    const int32_t* n_chunks_dev;
    int32_t nreg;
    int32_t pad2;
};

// Arguments of the persistent 3PCF triplet kernel (triplet3.hip).
// Histogram cell h = (cfg * nm + mu) * 4 + class.
struct TripArgs {
    const double4* pw;
    const uint32_t* flags;
    const int32_t* orig;        // original input index of each sorted point (RSD: find_normal argument order)
    const int32_t* cell_start;
    const Row* rows;            // full stencil
    const BinDesc* bins;
    const uint16_t* cfg_tab;    // [nb][nb][nb] -> 0-based configuration (unused when equi)
    const uint8_t* ntab;        // RSD: [nm][nm] 0-based find_normal bin of (mu(k1), mu(k2))
    int32_t* queue;             // 8 work counters
    int32_t* err;               // set to 1 if an apex has more than kmax neighbours
    unsigned long long* out_cnt;    // [ncfg][nm][4]
    unsigned long long* out_w;      // [ncfg][nm][4][2]
    unsigned long long* out_stats;  // candidates, rechecks, overflow, units, binned
    GridDesc g;
    int32_t nrows;
    int32_t n_items;            // apexes (all points; buffer points are skipped)
    int32_t ncfg;               // configurations (nb for equi: the configuration is the common bin)
    int32_t nm;                 // mu bins (1: isotropic)
    int32_t equi;               // only triangles whose three sides share one bin
    int32_t reps;
    int32_t kmax;
    int32_t lds_hist;           // histogram in LDS (else global atomics)
    // batched regions (nreg > 0): each apex uses its region's grid (regions[r].lo/dim/cell_base,
    // g.h shared) and adds to its region's histogram, out_cnt/out_w + r * [ncfg][nm][4]
    int32_t nreg;
    const RegionDesc* regions;
    // two-pass runs: with ovf set, an apex with more than kmax neighbours is appended to
    // ovf (count in *ovf_n) and skipped instead of raising err; with items set, work item
    // t is apex items[t] (the second pass over those apexes, with a larger kmax)
    int32_t* ovf;
    int32_t* ovf_n;
    const int32_t* items;
    const int32_t* n_items_dev;     // if set, the number of work items is read here (captured second pass)
    int32_t log_bins;               // selects the kernel variant (log or linear bin estimate)
    int32_t wg;                     // threads per work-group (one apex each): 64, 128 or 256 (0: 256)
    int32_t batch;                  // apexes taken per queue fetch (0: 1)
    // neighbour counts beyond every LDS list: with ovf_kmax set, the largest neighbour count of a
    // queued apex is kept there (atomic max); with glist set, each work-group keeps its apex's
    // neighbour list in device memory at glist + blockIdx.x * glist_stride bytes (kmax entries)
    // instead of LDS, and the histogram goes to global memory
    int32_t* ovf_kmax;
    void* glist;
    int64_t glist_stride;
};
// bytes per neighbour of a device-memory neighbour list (glist) of the triplet kernel
size_t triplet3_glist_entry_bytes(const TripArgs& a, bool fp64, bool weighted);
size_t triplet3_lds_bytes(const TripArgs& a, int nb, bool fp64, bool weighted);
hipError_t launch_triplet3(const TripArgs& a, int nb, bool fp64, bool weighted, int nblocks, hipStream_t st);
int triplet3_occupancy(const TripArgs& a, int nb, bool fp64, bool weighted);
// neighbour counts of the apexes t * stride (t < nsample) of a single catalogue; -1 for buffer points
hipError_t launch_nbr_count(const TripArgs& a, int stride, int nsample, int after_only, int32_t* out,
                            hipStream_t st);

// Arguments of the persistent 4PCF quad kernel (quad4.hip).
struct QuadArgs {
    const double4* pw;
    const uint32_t* flags;
    const int32_t* orig;        // original input index of each sorted point
    const int32_t* cell_start;
    const Row* rows;            // full stencil
    const BinDesc* bins;
    const uint16_t* cfg_tab;    // [nb]^4 -> 1-based perimeter configuration, 0 = not sorted
    int32_t* queue;
    int32_t* err;
    unsigned long long* out_cnt;    // [ncell][16]
    double* out_w;              // weighted: [ncell][16] quad products, [6][ncell][4] disc pair (tetra: triple) products
    unsigned long long* out_stats;
    GridDesc g;
    int32_t nrows;
    int32_t n_items;
    int32_t ncell;
    int32_t kmax;
    int32_t lds_hist;           // counts in LDS (else global atomics)
    int32_t tetra;              // regular tetrahedra only (all six sides in one bin; cell = that bin)
    // two-pass runs, as for the triplet kernel: with ovf set, an apex with more than kmax
    // neighbours is queued (count in *ovf_n) and skipped; with items set, work item t is apex items[t]
    int32_t* ovf;
    int32_t* ovf_n;
    const int32_t* items;
    int32_t batch;              // apexes taken per queue fetch (0: 1)
    // as for the triplet kernel: with ovf_kmax set, the largest neighbour count of a queued apex is
    // kept there; with glist set, each work-group keeps its apex's neighbour list (kmax entries, a
    // power of two) at glist + blockIdx.x * glist_stride bytes of device memory, with no link matrix
    // (the neighbour-neighbour bins are evaluated where they are needed)
    int32_t* ovf_kmax;
    void* glist;
    int64_t glist_stride;
};
// bytes per neighbour of a device-memory neighbour list (glist) of the quad kernel: x, y, z, w,
// key, permutation, packed bin, and the sort's staging copy of x, y, z, w and the packed bin
constexpr size_t kQuadGlistEntryBytes = 4 * 8 + 3 * 4 + 4 * 8 + 4;
size_t quad4_lds_bytes(int nb, int ncell, int kmax, bool lds_hist, bool tetra, int nrows, bool glist = false);
hipError_t launch_quad4(const QuadArgs& a, int nb, bool weighted, int nblocks, hipStream_t st);
int quad4_occupancy(const QuadArgs& a, int nb, bool weighted);

// Arguments of the persistent 2PCF adjoint kernel (adjoint.hip).
struct AdjArgs {
    const double4* pw;
    const double* ws;           // sorted raw weights alone
    const float4* pl;
    const uint32_t* flags;
    const int32_t* orig;
    const int32_t* cell_start;
    const int32_t* row_chunk;
    const Row* rows;            // full stencil
    const BinDesc* bins;
    const double* g2;           // dL/dN2 [nb][nm]
    const double* g3;           // dL/dN3 [nb][nm]
    int32_t* queue;
    double* G;                  // dL/dwn per point, original order
    unsigned long long* out_stats;
    GridDesc g;
    int32_t nrows;
    int32_t n_chunks;
    double inv_wd, inv_wr;
    // half-stencil FP32 kernel: per sorted point a 128-bit fixed-point sum [n][2] (lo, hi; zeroed before
    // launch) of G * 2^F, with the dL/dN tables scaled by g_scale and the normalised weights by w_scale
    // (g_scale * w_scale = 2^F)
    unsigned long long* Gq;
    double g_scale, w_scale;
};
size_t adjoint2_lds_bytes(int nb, int nm, bool fp64);
hipError_t launch_adjoint2(const AdjArgs& a, int nb, int nm, bool fp64, bool rsd, int nblocks, hipStream_t st);
int adjoint2_occupancy(int nb, int nm, bool fp64, bool rsd);
// half-stencil FP32 adjoint (rows: half stencil): every unordered pair once, its term added to both ends
size_t adjoint2h_lds_bytes(int nb, int nm);
hipError_t launch_adjoint2h(const AdjArgs& a, int nb, int nm, bool rsd, int nblocks, hipStream_t st);
int adjoint2h_occupancy(int nb, int nm, bool rsd);
// G[orig[k]] = Gq[k] * 2^-F (the 128-bit sums as doubles, original order)
hipError_t launch_adj_fx(const unsigned long long* Gq, const int32_t* orig, int64_t n, int F, double* G, hipStream_t st);
hipError_t launch_adj_partials(const double* G, const double* w, int64_t n, int64_t n_data, double inv_wd,
                               double inv_wr, double* part, int nblk, hipStream_t st);
hipError_t launch_adj_finish(const double* G, int64_t n, int64_t n_data, double s_d, double s_r, double inv_wd,
                             double inv_wr, double* out, hipStream_t st);

// Dynamic LDS bytes of the pair kernel for these parameters and work-group size wg (a multiple of 64, <= 1024).
size_t pair2_lds_bytes(int nb, int nm, int reps, bool fp64, bool weighted, bool batched = false, int wg = kWG);
hipError_t launch_pair2(const PairArgs& a, int nb, int nm, bool fp64, bool rsd, bool weighted, bool logb,
                        int nblocks, int wg, hipStream_t st);
// Work-groups of wg threads per CU the pair kernel can keep resident (0 if it cannot run).
int pair2_occupancy(int nb, int nm, int reps, bool fp64, bool rsd, bool weighted, bool logb, bool batched = false,
                    int wg = kWG);

}  // namespace gs
