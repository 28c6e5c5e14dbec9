// gs_kernels.h -- launch interfaces of the HIP kernels (host side).
#pragma once

#include "gs_common.h"

namespace gs {

// grid.hip
hipError_t launch_bbox(const double* xyz, int64_t n, double* part, int nblk, hipStream_t st);
hipError_t launch_cell_count(const double* xyz, int64_t n, const GridDesc& g, int32_t* cell_of, int32_t* count,
                             hipStream_t st);
hipError_t exclusive_scan(const int32_t* in, int32_t* out, int64_t n, int32_t* scratch, hipStream_t st);
int64_t scan_scratch_size(int64_t n);
hipError_t launch_scatter(const double* xyz, const double* w, const int32_t* buf, int64_t n, int64_t n_data,
                          const int32_t* cell_of, const int32_t* start, int32_t* fill, double4* pw,
                          uint32_t* flags, int32_t* orig, hipStream_t st);
hipError_t launch_chunk_counts(const int32_t* start, int ncell, int32_t* nch, hipStream_t st);

// Arguments of the persistent 2PCF pair kernel (pair2.hip).
struct PairArgs {
    const double4* pw;          // sorted points: x, y, z, raw weight
    const uint32_t* flags;      // sorted: bit0 random, bit1 never-apex
    const int32_t* cell_start;  // ncell+1
    const int32_t* chunk_off;   // ncell+1: first work chunk of each cell (chunks of kWG points)
    const Row* rows;            // half stencil
    const BinDesc* bins;        // device copy
    int32_t* queue;             // 8 work counters (one per XCD segment), zeroed before launch
    unsigned long long* out_cnt;    // [3][nb][nm]
    unsigned long long* out_w;      // [3][nb][nm][2] 128-bit fixed point (lo, hi)
    unsigned long long* out_stats;  // [0] candidates, [1] rechecks, [2] overflow pairs
    GridDesc g;
    int32_t nrows;
    int32_t n_chunks;
    int32_t reps;               // LDS histogram replicas (power of two <= 64)
    int32_t pad;
    float oshift;               // unused placeholder (keeps the struct 8-byte aligned)
    float pad2;
};

// Dynamic LDS bytes of the pair kernel for these parameters.
size_t pair2_lds_bytes(int nb, int nm, int reps, bool fp64, bool weighted);
hipError_t launch_pair2(const PairArgs& a, int nb, int nm, bool fp64, bool rsd, bool weighted, int nblocks,
                        hipStream_t st);
// Work-groups per CU the pair kernel can keep resident for this LDS size.
int pair2_occupancy(int nb, int nm, int reps, bool fp64, bool rsd, bool weighted);

}  // namespace gs
