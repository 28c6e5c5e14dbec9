/*
 * gs_oracle.h -- CPU restatement of the Gramsci counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gramsci_amd/, host/)
 * links, loads or calls this library.  It is the checker that the HIP path
 * is compared against in tests/, in __graft_entry__.smoke() and in the
 * cpu_baseline leg of bench.py.
 *
 * Every routine follows the reference Fortran (src/gramsci.F90) line by line
 * in plain FP64 with no contraction (-ffp-contract=off), so that the set of
 * binned pairs / triplets / quads and their bins are exactly the reference's.
 * Weighted sums are accumulated in long double (64-bit mantissa), so the
 * oracle's N2/N3 are the reference's intended sums to ~1e-19 relative.
 *
 * Indexing conventions (0-based here, 1-based in the reference):
 *   points are xyz[n][3] (== Fortran points(3,n)), data first, then randoms;
 *   a point is "random" iff its index >= n_data (gramsci.F90:373);
 *   buf[i] != 0 means "never an apex" (gramsci.F90:363), buf may be NULL.
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double  rmin, rmax;
    int32_t nbins, nmu, log_bins;
} or_params;

/* Bin edges as printed by the reference (gramsci.F90:88-96). edges[nbins+1]. */
void or_radial_bins(const or_params* p, double* edges);

/* Normalised weights (gramsci.F90:1390-1391): data w/sum(w_d), randoms -w/sum(w_r).
 * w may be NULL (unit weights). wn[n]. */
void or_normalise(const double* w, int64_t n_data, int64_t n_rand, double* wn);

/* Radial bin of one pair, 1-based, or 0 if the pair is not kept.  Follows the
 * r-ball inclusion s <= rmax^2 (kdtree2.F90:1689), the strict filter
 * rmin < d < rmax (gramsci.F90:316-317) and the bin formula (:334-338).
 * xi is the apex, xj the neighbour. */
int or_pair_rbin(const or_params* p, const double* xi, const double* xj);

/* mu bin of the ordered pair apex i -> neighbour j, 1-based (gramsci.F90:320-328). */
int or_pair_mubin(const or_params* p, const double* xi, const double* xj);

/* 2PCF (gramsci.F90:288-348 + :352-399).  Outputs (any may be NULL):
 *   cnt [3][nbins][nmu]  ordered-pair counts, class 0 = DD, 1 = DR or RD, 2 = RR
 *   wsum[3][nbins][nmu]  sum of raw-weight products w_i*w_j per class
 *   n2, n3 [nbins][nmu]  the reference's N2(:,:,3), N3(:,:,3)
 *   n_overflow           pairs whose bin came out as nbins+1 (not counted)
 * Returns 0, or a negative error code. */
int or_2pcf(const double* xyz, const double* w, const int32_t* buf,
            int64_t n_data, int64_t n_rand, const or_params* p,
            uint64_t* cnt, double* wsum, double* n2, double* n3,
            uint64_t* n_overflow);

/* Number of 3PCF configurations C(nbins+2,3) and 4PCF perimeter configurations C(nbins+3,4). */
int or_n3cfg(int nbins);
int or_n4cfg(int nbins);
/* bintable lookups (gramsci.F90:1398-1477), 1-based bins, 1-based result, 0 = not filled. */
int or_bintable3(int nbins, int a, int b, int c);
int or_bintable4(int nbins, int a, int b, int c, int d);

/* find_normal (gramsci.F90:1611-1625): mu bin of a triplet from the mu bins
 * mu1, mu2 (1-based) of its two apex sides; 1-based. */
int or_normal_bin(int nmu, int mu1, int mu2);

/* 3PCF, all configurations (gramsci.F90:463-545 with find_dist :1579-1609);
 * nmu >= 2 is the reference's RSD 3PCF (find_normal, :495-500).
 *   cnt[ncfg][nmu][4]  apex-ordered triplet counts by number of randoms among (i,k1,k2)
 *   n2, n3 [ncfg][nmu] the reference's N2(:,:,3), N3(:,:,3)
 *   n_cand             apex-pair candidates, sum_i C(k_i,2) over non-buffer apexes */
int or_3pcf(const double* xyz, const double* w, const int32_t* buf,
            int64_t n_data, int64_t n_rand, const or_params* p,
            uint64_t* cnt, double* n2, double* n3, uint64_t* n_cand);

/* Equilateral 3PCF (gramsci.F90:402-461).  cnt[nbins][nmu][4], n2, n3 [nbins][nmu]
 * (N3 is only accumulated when nmu == 1, as in the reference). */
int or_3pcf_equi(const double* xyz, const double* w, const int32_t* buf,
                 int64_t n_data, int64_t n_rand, const or_params* p,
                 uint64_t* cnt, double* n2, double* n3);

/* 4PCF (gramsci.F90:662-861).  cells are (cfg, ind2, ind6): ncfg4*nbins*nbins.
 *   cnt[cell][16]     quad counts by random pattern (bit0 apex, bit1 k1, bit2 k2, bit3 k3)
 *   n2, n3 [cell]     N2(bin,ind2,ind6), N3(bin,ind2,ind6)
 *   disc_n, disc_r [6][cell]  discN(bin,ind2,ind6,1..6), discR(...)
 *   n_quads           evaluated quads: sum over linked (k1,k2) of candidates k3 */
int or_4pcf(const double* xyz, const double* w, const int32_t* buf,
            int64_t n_data, int64_t n_rand, const or_params* p,
            uint64_t* cnt, double* n2, double* n3, double* disc_n, double* disc_r,
            uint64_t* n_quads);

/* Regular-tetrahedron 4PCF (gramsci.F90:548-660).  cnt[nbins][16],
 * n2, n3 [nbins], disc_n, disc_r [6][nbins] (discN(b,1,1,t), discR(b,1,1,t)). */
int or_4pcf_tetra(const double* xyz, const double* w, const int32_t* buf,
                  int64_t n_data, int64_t n_rand, const or_params* p,
                  uint64_t* cnt, double* n2, double* n3, double* disc_n, double* disc_r);

/* Gradient of a loss L(N2, N3) of the 2PCF with respect to the RAW weights
 * (before the normalisation of gramsci.F90:1390-1391), given gN2 = dL/dN2 and
 * gN3 = dL/dN3 [nbins][nmu].  Brute force over the ordered pairs, long double
 * accumulation.  dldw[n]. */
int or_adjoint_2pcf(const double* xyz, const double* w, const int32_t* buf,
                    int64_t n_data, int64_t n_rand, const or_params* p,
                    const double* gN2, const double* gN3, double* dldw);

#ifdef __cplusplus
}
#endif
#endif
