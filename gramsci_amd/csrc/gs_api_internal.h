// gs_api_internal.h -- host-side helpers shared by the translation units that
// implement the C ABI (api.cpp, pcf3.cpp, pcf4.cpp, adjoint.cpp, batched.cpp).
#pragma once

#include <cstdarg>
#include <string>
#include <vector>

#include "../../include/gramsci_hip.h"
#include "gs_common.h"
#include "gs_kernels.h"

namespace gs {

extern thread_local std::string g_err;
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t count);   // grows only; contents are not preserved
    void release();
};

// Buffers owned by a context for the 3PCF/4PCF/adjoint/batched paths.
struct GsExtra {
    DevBuf<unsigned char> params;
    DevBuf<unsigned long long> out;
    DevBuf<int32_t> queue;
    DevBuf<double> dbl;
    DevBuf<int32_t> i32;
    DevBuf<unsigned char> big;      // device-memory neighbour lists of the last 3PCF/4PCF pass
};
void extra_release(GsExtra& x);

// Batched independent regions (batched.cpp): host description of the regions
// and the device buffers / captured hipGraphs of the batched pipeline.
// Launch parameters of a batched plan that the autotuner times (0: the automatic choice): cells per
// rmax in y/z (m, required), x cells per y/z cell (pair kernel), threads per work-group and first-pass
// neighbour-list length (triplet kernel).
struct BatchTune {
    int m = 0, sx = 0, wg = 0, kmax1 = 0;
    bool operator==(const BatchTune& o) const { return m == o.m && sx == o.sx && wg == o.wg && kmax1 == o.kmax1; }
};

struct GsBatch {
    int nreg = 0;
    bool valid = false;
    std::vector<int64_t> off, ndata;
    std::vector<double> bbox;            // [nreg][6]
    std::vector<long double> wd, wr;     // per-region weight sums
    std::vector<double> wmr;             // per-region largest weight (1 for unit weights)
    std::vector<double> wmr_d, wmr_r;    // per-region largest data / random weight (1 for none)
    double wmax = 1.0, wmax_d = 1.0, wmax_r = 1.0, xmax = 0.0;
    DevBuf<int32_t> d_off, d_row_base, cell_of, count, cell_start, nch, row_chunk, scan_tmp, queue, orig;
    DevBuf<int32_t> ovf;                 // 3PCF: queued apexes of the second and the device-memory pass
    DevBuf<unsigned char> big;           // 3PCF: device-memory neighbour lists
    DevBuf<unsigned char> d_regions, params;
    DevBuf<double4> pw;
    DevBuf<double> ws;                   // sorted raw weights
    DevBuf<float4> pl;
    DevBuf<uint32_t> flags;
    DevBuf<unsigned long long> out;
    DevBuf<double> xyz, w;               // the batch's catalogue, input order
    DevBuf<int32_t> buf;
    bool has_w = false, has_buf = false;
    struct Plan {
        gs_params p{};
        BatchTune tune{};                // as requested (0: automatic)
        int m = 0;
        int sx = 1;                      // x cells per y/z cell
        int wg1 = 0, kmax1 = 0;          // triplet first pass: threads per work-group, list length
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int nbt = 0;
        int fx_shift = 0;
        std::vector<int> fx_shift_r;     // per-region fixed-point scale (weighted runs; triples)
        std::vector<int> fx_sd_r, fx_sr_r;   // per-region pair scales of data and random weights
        int mst = 0;
        int kind = 2;                    // 2: pair kernel, 3: triplet kernel
        int nh = 0;                      // triplet histogram entries per region
        int kmax = 0;                    // triplet neighbour capacity per apex
        DevBuf<unsigned char> blob;      // regions, bins and stencil rows read by the captured kernels
        // 3PCF: arguments of the device-memory pass run after each replay (queued apexes, count)
        TripArgs t3{};
        const int32_t* t3_list = nullptr;
        const int32_t* t3_cnt = nullptr;
        int nb = 0;
        bool fp64 = false, weighted = false;
    };
    std::vector<Plan> plans;
    std::vector<std::pair<std::string, BatchTune>> tuned;   // parameter key -> tuned launch parameters
};
void batch_release(GsBatch& b);

struct CtxView {
    hipStream_t stream;
    hipEvent_t* ev;
    int n_cu;
    int64_t n_data, n_rand;
    bool has_w, has_buf;
    long double wsum_d, wsum_r;
    double wmax, wmax_d, wmax_r;    // largest weight overall, of the data, of the randoms (1 if none)
    double bbox[6];
    GridDesc g;
    const double4* pw;
    const double* ws;           // sorted raw weights
    const uint32_t* flags;
    const int32_t* orig;
    const int32_t* cell_start;
    const float4* pl;
    const int32_t* row_chunk;
    const double* w;            // raw weights, original order (NULL for unit weights)
    int32_t n_chunks;
    GsExtra* extra;
    GsBatch* batch;
    const double* xyz;           // catalogue, original order
    const int32_t* buf;
};

CtxView view(gs_ctx* c);
int ctx_set_device(gs_ctx* c);
int ctx_upload(gs_ctx* c, const double* xyz, const double* w, const int32_t* buf, int64_t n);
double round11(double h);
int ctx_build_grid(gs_ctx* c, double rmax, int m, int sx, double* h_out);
// device buffers of a grid build (sizes: cell_of, pw, pl, flags, orig n; cell_start, fill, chunk_nc,
// chunk_off ncell + 1; scan_tmp scan_scratch_size(ncell + 1))
struct GridBufs {
    int32_t *cell_of, *cell_start, *fill, *chunk_nc, *chunk_off, *scan_tmp;
    double4* pw;
    double* ws;
    float4* pl;
    uint32_t* flags;
    int32_t* orig;
};
int grid_build_raw(const double* xyz, const double* w, const int32_t* buf, int64_t n, int64_t n_data,
                   const double* bbox, double h, int sx, const GridBufs& b, int64_t max_cells, GridDesc* g_out,
                   int32_t* nch, hipStream_t st);
int pair2_shape(int nb, int nm, bool fp64, bool weighted, bool logb, int* reps, int* wg, int* waves);
double choose_h_bbox(const double* bbox, double rmax, int m, int64_t max_cells, int sx);
int pair_cell_div(int64_t n, const double* bbox, double rmax);
int pair_cell_sx();
int apex_cell_sx();
void combine2(const unsigned long long* h_out, int nb, int nm, const BinDesc& B, bool weighted, long double Wd,
              long double Wr, uint64_t* cnt, double* wsum, double* n2, double* n3, gs_stats* st);
int check_params(const gs_params* p, int mode);
int fill_bins(const gs_params& p, double rrel, double xmax, double wmax_d, double wmax_r, BinDesc& B);
// fixed-point scales (exponents F of 2^F): of one point's weight in a pair product, for the largest
// weight wmax of its class (wmax * 2^F < 2^26); of a triple product for the largest weight wmax
int fx_shift_point(double wmax);
int fx_shift_triples(double wmax);
// exponent of a pair class sum: class 0 DD, 1 DR, 2 RR
inline int fx_pair_class_shift(const BinDesc& B, int cls) { return cls == 0 ? 2 * B.fx_sd : cls == 1 ? B.fx_sd + B.fx_sr : 2 * B.fx_sr; }
// sum of w[0..n) as the exact sum of the fixed-point integers wsum_fixed(w[i], k), as a long double
long double fixed_sum(const double* w, int64_t n, int k);
std::vector<Row> half_stencil(double h, double hx, double rmax, int m);
std::vector<Row> full_stencil(double h, double hx, double rmax, int m);
double ctx_xmax(gs_ctx* c);
long double fx_value(unsigned long long lo, unsigned long long hi, int shift);
int default_cell_div(const gs_ctx* c, double rmax, int mode);
// 3PCF helpers shared by the single and batched paths (pcf34.cpp)
std::vector<uint16_t> table3(int nb);
std::vector<uint16_t> table3_hist(int nb);   // 4 * configuration, for the triplet kernel
std::vector<uint8_t> normal_table(int nm);
int plan_triplet_lds(TripArgs& a, int nb, bool fp64, bool weighted);
// threads per triplet work-group for a mean neighbour count kmean (GS_T3_WG overrides)
int triplet_wg(double kmean);
// apexes per queue fetch for n work items over nblocks work-groups: about 128 fetches per work-group
// (GS_APEX_BATCH overrides)
int triplet_batch(int64_t n, int nblocks);
void combine3(const long double* S, int nq, long double Wd, long double Wr, bool with_n3, double* n2, double* n3);
// last 3PCF pass over the apexes queued in d_list (count d_cnt[0], largest neighbour count d_cnt[1]),
// with neighbour lists in device memory (pcf34.cpp); *n_apex (may be NULL) receives the count
int triplet_glist_pass(const TripArgs& base, int nb, bool fp64, bool weighted, const int32_t* d_list,
                       const int32_t* d_cnt, int32_t* d_queue, DevBuf<unsigned char>& scratch, int n_cu,
                       hipStream_t st, int* n_apex);

}  // namespace gs
