/*
 * gramsci_hip.h -- C ABI of the MI355X N-point correlation engine
 * (libgramsci_hip.so, built from gramsci_amd/csrc/).
 *
 * The reference program (src/gramsci.F90) has no plugin or FFI layer: its
 * counting path is the pair of contained subroutines
 *     create_graph(istart,iend)                      gramsci.F90:288-348
 *     query_graph_2pcf / query_graph_3pcf_all /
 *     query_4pcf_all2 (istart,iend)                  gramsci.F90:352-399, 463-545, 662-861
 * which read program globals (points(3,N), weights(N), buffer(N), num_data,
 * num_rand, rmin, rmax, nbins, nmu, logbins, RSD, bintable) and write N2, N3,
 * discN, discR.  Each entry point below replaces one create_graph+query pair;
 * the globals become explicit arguments:
 *     points(3,N) real(8)  <->  const double* xyz   (xyz[3*i+k], identical memory layout)
 *     weights(N)           <->  const double* w     (RAW weights; the engine normalises them
 *                                                     as read_files_2 does, gramsci.F90:1390-1391)
 *     buffer(N)            <->  const int32_t* buf  (1 = never an apex, gramsci.F90:363)
 *     num_data, num_rand   <->  n_data, n_rand      (a point is random iff index >= n_data)
 *     rmin,rmax,nbins,nmu,logbins <-> gs_params     (RSD is nmu >= 2, gramsci.F90:1219-1222)
 *
 * Conventions
 *   - Every function returns GS_OK (0) or a negative gs_err; gs_last_error()
 *     gives a message (thread-local).  Nothing aborts or exits the process.
 *   - Host pointers are caller-owned and only read during the call.  Output
 *     arrays are caller-allocated and fully overwritten.
 *   - Counts are exact integers.  Weighted sums are accumulated in 128-bit
 *     fixed point on the GPU and are therefore exact sums of the rounded FP64
 *     weight products (independent of thread order); n2/n3 are formed from
 *     them in extended precision on the host.
 *   - A context owns device buffers and a HIP stream; one context per host
 *     thread.  Calls on one context are serialised by the caller.
 */
#ifndef GRAMSCI_HIP_H
#define GRAMSCI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gs_err {
    GS_OK = 0,
    GS_EINVAL = -1,     /* bad argument (sizes, bins, rmin/rmax, NULL pointer) */
    GS_ENOMEM = -2,     /* device or host allocation failed */
    GS_EHIP = -3,       /* a HIP runtime call failed */
    GS_ENODEV = -4,     /* no usable gfx950 device */
    GS_ELIMIT = -5,     /* a documented capacity limit was exceeded */
};

enum gs_precision {
    GS_FP64 = 0,        /* plain FP64 geometry for every pair */
    GS_FP32_GUARD = 1,  /* FP32 geometry, pairs within a guard band of an edge re-evaluated in FP64 */
};

typedef struct gs_params {
    double  rmin, rmax;    /* -rmin / -rmax (gramsci.F90:1203-1210) */
    int32_t nbins;         /* -nbins, 1..127 (int8 bin storage, node_module.F90:7) */
    int32_t nmu;           /* -nmu, 1..127; >= 2 selects (r, mu) binning */
    int32_t log_bins;      /* -log */
    int32_t precision;     /* enum gs_precision */
    int32_t cell_div;      /* grid cells per rmax; 0 = the engine's default/tuned value */
    int32_t reserved;      /* must be 0 */
} gs_params;

typedef struct gs_stats {
    double   ms_grid;        /* device time of the grid build (HIP events) */
    double   ms_count;       /* device time of the counting kernel(s) */
    double   ms_kernel_avg;  /* average duration of the dominant kernel launch */
    uint64_t n_units;        /* work units: binned ordered pairs (2PCF), apex-pair candidates (3PCF),
                                evaluated quads (4PCF) */
    uint64_t n_binned;       /* binned ordered pairs / apex triplets / quads */
    uint64_t n_candidates;   /* geometric tests performed (stencil candidates) */
    uint64_t n_recheck;      /* FP64 re-evaluations triggered by the guard band */
    uint64_t n_overflow;     /* pairs whose reference bin is nbins+1 (not counted; the reference
                                writes them out of bounds) */
    uint64_t n_evaluated;    /* evaluations the kernel performed: candidate pairs tested (2PCF),
                                neighbour pairs whose third side was tested (3PCF; a quarter of
                                n_units or less where each triangle is visited once), quads (4PCF) */
    int32_t  cell_div;       /* cells per rmax actually used */
    int32_t  reserved;
} gs_stats;

typedef struct gs_ctx gs_ctx;

const char* gs_last_error(void);
const char* gs_version(void);

/* Creates a context on HIP device `device` (must be gfx950). */
int gs_create(gs_ctx** ctx, int device);
int gs_destroy(gs_ctx* ctx);

/* Copies a catalogue (host memory) to the device: xyz[n][3], w[n] or NULL for
 * unit weights, buf[n] or NULL for "every point is an apex".  Replaces
 * read_files_2 + the points/weights/buffer globals (gramsci.F90:1364-1396). */
int gs_set_catalogue(gs_ctx* ctx, const double* xyz, const double* w, const int32_t* buf,
                     int64_t n_data, int64_t n_rand);
/* Same, from device pointers already resident in HBM (copied device-to-device). */
int gs_set_catalogue_device(gs_ctx* ctx, const double* d_xyz, const double* d_w,
                            const int32_t* d_buf, int64_t n_data, int64_t n_rand);

/* 2PCF: create_graph + query_graph_2pcf (gramsci.F90:288-399).
 *   cnt [3][nbins][nmu]  ordered-pair counts; class 0 = DD, 1 = DR+RD, 2 = RR
 *   wsum[3][nbins][nmu]  sum of raw-weight products w_i*w_j per class (may be NULL)
 *   n2, n3 [nbins][nmu]  the reference's N2(:,:,3), N3(:,:,3) (may be NULL)
 *   st                   may be NULL */
int gs_count_2pcf(gs_ctx* ctx, const gs_params* p, uint64_t* cnt, double* wsum,
                  double* n2, double* n3, gs_stats* st);

/* gs_count_2pcf for a catalogue in device memory (d_xyz [n][3], d_w [n] or
 * NULL, d_buf [n] or NULL; data first, then randoms), allocating nothing: the
 * sorted catalogue, grid, parameters, accumulators and counters live in one
 * file-scope device array of the library (about 54 MB per GPU, held until the
 * process ends).  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Limits: at most 524,288 points
 * (GS_ELIMIT beyond) and 1,048,576 grid cells (the cell edge is widened to
 * fit).  Outputs as gs_count_2pcf, in host memory.  The catalogue uploaded to
 * ctx (if any) is neither used nor changed.  Calls that overlap on one GPU
 * race for its array. */
int gs_count_2pcf_noalloc(gs_ctx* ctx, const gs_params* p, const double* d_xyz, const double* d_w,
                          const int32_t* d_buf, int64_t n_data, int64_t n_rand, uint64_t* cnt, double* wsum,
                          double* n2, double* n3, gs_stats* st);

/* 3PCF, all configurations: create_graph + query_graph_3pcf_all (gramsci.F90:463-545).
 * ncfg = C(nbins+2,3), in the reference's i<=j<=k order.  nmu >= 2 is the
 * reference's RSD 3PCF: each triplet also goes to the mu bin
 * find_normal(mu(k1), mu(k2)) (:495-500, :1611-1625).
 *   cnt [ncfg][nmu][4]   apex-ordered triplet counts by number of randoms among (apex, k1, k2)
 *   wsum[ncfg][nmu][4]   sum of raw-weight triple products per class (may be NULL)
 *   n2, n3 [ncfg][nmu]   N2(:,:,3), N3(:,:,3) */
int gs_count_3pcf(gs_ctx* ctx, const gs_params* p, uint64_t* cnt, double* wsum,
                  double* n2, double* n3, gs_stats* st);

/* Equilateral 3PCF: create_graph + query_graph_equilateral_triangle
 * (gramsci.F90:402-461, option -equi): triangles whose three sides share one
 * radial bin b.  Same outputs as gs_count_3pcf with the configuration replaced
 * by b: cnt/wsum [nbins][nmu][4], n2/n3 [nbins][nmu].  With nmu >= 2 the
 * reference accumulates no N3 (:429-431): n3 is 0. */
int gs_count_3pcf_equi(gs_ctx* ctx, const gs_params* p, uint64_t* cnt, double* wsum,
                       double* n2, double* n3, gs_stats* st);

/* 4PCF: create_graph + query_4pcf_all2 (gramsci.F90:662-861).  nmu must be 1.
 * Cells are (cfg, ind2, ind6) with ncfg = C(nbins+3,4): ncell = ncfg*nbins*nbins.
 *   cnt [ncell][16]          quad counts by random pattern (bit0 apex, bit1 k1, bit2 k2, bit3 k3)
 *   n2, n3 [ncell]           N2(bin,ind2,ind6), N3(bin,ind2,ind6)
 *   disc_n, disc_r [6][ncell] discN(bin,ind2,ind6,1..6), discR(...) (may be NULL) */
int gs_count_4pcf(gs_ctx* ctx, const gs_params* p, uint64_t* cnt, double* n2, double* n3,
                  double* disc_n, double* disc_r, gs_stats* st);

/* Regular-tetrahedron 4PCF: create_graph + query_graph_tetrahedron
 * (gramsci.F90:548-660, option -tetra): apex and neighbours k1 < k2 < k3 (by
 * original index) whose six sides share one radial bin b.  nmu must be 1.
 *   cnt [nbins][16]           counts by random pattern (bit0 apex, bit1 k1, bit2 k2, bit3 k3)
 *   n2, n3 [nbins]            N2(b,1,3), N3(b,1,3)
 *   disc_n, disc_r [6][nbins] discN(b,1,1,1..6), discR(b,1,1,1..6) (may be NULL) */
int gs_count_4pcf_tetra(gs_ctx* ctx, const gs_params* p, uint64_t* cnt, double* n2, double* n3,
                        double* disc_n, double* disc_r, gs_stats* st);

/* Batched independent regions (domain-decomposed sub-volumes with their
 * buffer shells, jackknife regions): region r holds points [off[r], off[r+1])
 * of xyz/w/buf, its first ndata[r] points are data and the rest randoms; buf
 * flags the shell points that are neighbours but never apexes.  This is the
 * reference's split-run-resum workflow (domain_decomposition.f90:49-74, -cut
 * mode gramsci.F90:1224-1231, buffer skip :363) with every region in ONE
 * persistent launch, captured in a hipGraph and replayed.  Each region is
 * normalised on its own, exactly as a separate -gal/-ran run would be.
 * Host pointers; off[nreg+1], ndata[nreg]. */
int gs_set_catalogue_batched(gs_ctx* ctx, const double* xyz, const double* w, const int32_t* buf,
                             const int64_t* off, const int64_t* ndata, int nreg);
/* Per-region 2PCF of the batched catalogue: cnt/wsum [nreg][3][nbins][nmu],
 * n2/n3 [nreg][nbins][nmu].  p->cell_div = 0 selects the cell size by timing
 * the captured pipeline for 2..6 cells per rmax (cached per parameter set). */
int gs_count_2pcf_batched(gs_ctx* ctx, const gs_params* p, uint64_t* cnt, double* wsum,
                          double* n2, double* n3, gs_stats* st);
/* Per-region 3PCF of the batched catalogue (query_graph_3pcf_all, gramsci.F90:463-545,
 * on every region, in one captured launch of the triplet kernel):
 * cnt/wsum [nreg][ncfg][nmu][4], n2/n3 [nreg][ncfg][nmu], as gs_count_3pcf per region.
 * p->cell_div = 0 times 1..3 cells per rmax on the batch (cached per parameter set). */
int gs_count_3pcf_batched(gs_ctx* ctx, const gs_params* p, uint64_t* cnt, double* wsum,
                          double* n2, double* n3, gs_stats* st);

/* Domain decomposition on the GPU, replacing the reference's offline tool
 * (domain_decomposition.f90:35-74).  region[i] in 1..nreg labels point i (other
 * labels belong to no region).  For each label r: box = bounding box of the
 * points labelled r; members = every point strictly inside the box grown by
 * rmax on every side; buffer flag 0 for a member inside the closed box, 1
 * otherwise.  As in the tool, coordinates, rmax and the boxes are single
 * precision.  A label with no point has box (+FLT_MAX, -FLT_MAX) and no member.
 *   off[nreg+1]  member offsets (always written)
 *   idx, buf     [off[nreg]] member indices (input order within a region) and
 *                buffer flags, region-major; both NULL to query the sizes only,
 *                else cap >= off[nreg] (GS_ELIMIT otherwise, off still written)
 *   box[nreg][6] xmin, ymin, zmin, xmax, ymax, zmax (may be NULL)
 * The members of region r with their flags are the catalogue the tool writes to
 * "r.loadnodes" and, fed to gs_set_catalogue_batched, one region of a batch. */
int gs_decompose(gs_ctx* ctx, const float* xyz, const int32_t* region, int64_t n, float rmax, int nreg,
                 int64_t* off, int32_t* idx, int32_t* buf, int64_t cap, float* box);

/* Adjoint of the weighted 2PCF: given dl_dn2 = dL/dN2 and dl_dn3 = dL/dN3
 * [nbins][nmu] (host memory) for the catalogue of the context, writes
 * dL/dw_k for the RAW weights w_k (the normalisation of gramsci.F90:1390-1391
 * is differentiated through).  The pairs, bins and buffer rule are those of
 * gs_count_2pcf (query_graph_2pcf, gramsci.F90:352-380); the reference has no
 * derivative.  dl_dw[n] in host memory, or d_dl_dw[n] in device memory.
 * GS_FP32_GUARD: per-point sums in exact fixed point (bit-identical across
 * calls); a non-finite dL/dN entry switches to FP64 sums, so NaN propagates. */
int gs_adjoint_2pcf(gs_ctx* ctx, const gs_params* p, const double* dl_dn2, const double* dl_dn3,
                    double* dl_dw, gs_stats* st);
int gs_adjoint_2pcf_device(gs_ctx* ctx, const gs_params* p, const double* dl_dn2, const double* dl_dn3,
                           double* d_dl_dw, gs_stats* st);

/* One-shot forms used by the Fortran host: create a context on device 0 (kept
 * for the process), upload, count.  Same outputs as above. */
int gs_pcf2(const double* xyz, const double* w, const int32_t* buf, int64_t n_data, int64_t n_rand,
            const gs_params* p, uint64_t* cnt, double* n2, double* n3, gs_stats* st);
int gs_pcf3(const double* xyz, const double* w, const int32_t* buf, int64_t n_data, int64_t n_rand,
            const gs_params* p, uint64_t* cnt, double* n2, double* n3, gs_stats* st);
int gs_pcf4(const double* xyz, const double* w, const int32_t* buf, int64_t n_data, int64_t n_rand,
            const gs_params* p, uint64_t* cnt, double* n2, double* n3, double* disc_n, double* disc_r,
            gs_stats* st);

/* Exact bin thresholds used by every kernel: t[0] = largest s with sqrt(s) <= rmin,
 * t[k] (k=1..nbins) = smallest s whose reference bin is >= k+1 (t[nbins] is the
 * first s whose bin overflows), t[nbins+1] = smallest s not kept by
 * s <= rmax*rmax and sqrt(s) < rmax.  s is the squared distance in FP64.
 * t has nbins+2 entries.  Host-only; no device needed. */
int gs_bin_thresholds(const gs_params* p, double* t);

#ifdef __cplusplus
}
#endif
#endif
