/*
 * pcs_oracle.h -- CPU oracle for the ABC-PMC simulation hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP library, the C++
 * host, the Python package) may include, link or call this code. It is used
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * always as the checker, never as the thing measured or shipped.
 *
 * It restates, in plain C, the reference's simulate-and-bin routine and its
 * cumulative-histogram distance:
 *   - walk / detect / track / bin: src/cpu/simulate.jl:56-124 and
 *     src/gpu/cusim/kernel-simulate-unroll.inl:99-295 (centred box, single
 *     periodic wrap, inclusive detection, k >= kmin recording),
 *   - histogram record: src/gpu/cusim/histogram2d.inl:98-106 (bin = trunc),
 *   - distance: src/cpu/distance.jl:1-8 and
 *     src/gpu/cusim/kernel-distance.inl:81-241 (2D cumulative sums, sum of
 *     squared differences, log10),
 *   - reference histogram: src/gpu/detail/simulation_impl.inl:605-629.
 * The random stream is not the reference's (LCG48 / Julia MersenneTwister):
 * it is Philox4x32-10 keyed by (seed, simulation id) with the counter layout
 * described in pcs_oracle.c, identical to what the GPU path consumes.
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py):
 *   orc_philox4x32_10     Random123 known-answer vectors + rocRAND header
 *   orc_neg2log_u24       libm FP64 log, ulp bound, all 2^24 inputs
 *   orc_sincos16          libm FP64 sin/cos, error bound, all 2^16 codes; the
 *                         table against 50-digit values (fit_math_coeffs.py)
 *   orc_exp2_01           numpy exp2, ulp bound on a 2^22-point grid; 2^0 = 1
 *   orc_gauss_detect      the exact decision u <= exp(-z^2 / (2 s^2)) in FP64
 *                         away from the boundary; detected fraction of a slab
 *   orc_chunk_normals_n   moments, KS against N(0,1), zero correlation between
 *                         the 12 normals of a chunk; orc_bm_angle16 byte layout
 *   orc_simulate          closed forms (sigma=0 tracks, chi-square DE law,
 *                         detected fraction, Poisson mean), brute-force
 *                         hand-checked tiny cases, FP64 twin agreement
 *   orc_distance          numpy cumsum(cumsum()) on random histograms
 *   orc_reference_hist    test_files fixture totals (36,701 in range)
 *   orc_exp10             mpmath 10^x, ulp bound on a grid over [-30, 30]
 *   orc_poisson           pmf against scipy.stats.poisson (chi-square) for
 *                         small lambda, mean and variance for large lambda
 *   orc_propose           parent frequencies against the weights (chi-square),
 *                         displacements N(0, tau^2) (KS) away from the bounds,
 *                         reflection into the bounds, tau = 0 leaves a member
 *                         unchanged, sorted components, cum_prob / sigma /
 *                         half_az from their closed forms
 */
#ifndef PCS_ORACLE_H
#define PCS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SPECIES 8

/* Philox4x32-10 block function (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* y = -2 ln(u), u = (float)(w | 255) * 2^-32 (the top 24 bits of w), in the
 * project's FP32 op order. */
float orc_neg2log_u24(uint32_t w);
/* Box-Muller radius sqrt(-2 ln u) in FP32 (correctly rounded sqrt of the above). */
float orc_bm_radius(uint32_t w);
/* sin and cos of the Box-Muller angle 2 pi (code + 1/2) 2^-16 for a 16-bit
 * code, in the project's FP32 op order (table + rotation, see pcs_oracle.c). */
void orc_sincos16(uint32_t code, float* s, float* c);

/* 16-bit angle code of Box-Muller pair p of a chunk: the low byte of the
 * pair's radius word above byte p % 4 of the chunk's angle word. */
uint32_t orc_bm_angle16(uint32_t w_radius, uint32_t w_angle, uint32_t p);
/* Unit displacements (sigma = 1) of chunks 0..n_chunks-1 of particles
 * 0..n_particles-1 of one video, through the stream layout of pcs_oracle.c:
 * out[(part * n_chunks + chunk) * 12 + i], i = 3 * step + axis. */
void orc_chunk_normals_n(uint64_t seed, uint32_t sim_id, uint32_t video, uint32_t n_particles, uint32_t n_chunks,
                         int gauss, float* out);

/* 2^f for f in [0, 1) and the Gaussian detection draw of the walk: c =
 * log2(e) / (2 s^2) for s^2 = sig2; detected iff (float)(w | 255) <= 2^(32 - c z^2)
 * (the exponent clamped below at -128). */
float orc_exp2_01(float f);
float orc_gauss_c(float sig2);
int orc_gauss_detect(float z, float c, uint32_t w);

/* Batched versions of the math helpers (for exhaustive tests). */
void orc_bm_radius_n(const uint32_t* w, float* out, uint64_t n);
void orc_sincos_angle_n(const uint32_t* w, float* s, float* c, uint64_t n);  /* code = w >> 16 */

/* Model description shared by all simulations of one ABC run. */
typedef struct orc_system {
    uint32_t n_species;     /* C, 1..ORC_MAX_SPECIES */
    uint32_t kmin;          /* minimum recorded track length (>= 2) */
    uint32_t n_k;           /* histogram rows; row r holds tracks of length r+1 */
    uint32_t n_de;          /* histogram columns */
    float de_bin_width;     /* ub_de / n_de */
    float four_dt;          /* 4 * deltat */
    float half_l[3];        /* half box extents */
    float half_a[2];        /* half detection extents in x, y */
    uint32_t z_boundary;    /* 0 periodic (reference), 1 reflecting */
    uint32_t detect_model;  /* 0 hard slab (reference), 1 gaussian in z */
    float detect_sigma;     /* sigma of the gaussian z detection */
} orc_system;

/* One simulation (ABC proposal). */
typedef struct orc_proposal {
    uint32_t sim_id;
    uint32_t n_videos;
    const uint32_t* frames;     /* [n_videos] */
    const uint32_t* particles;  /* [n_videos] */
    float half_az[ORC_MAX_SPECIES];
    float cum_prob[ORC_MAX_SPECIES];
    float sigma[ORC_MAX_SPECIES];
} orc_proposal;

/* FP32 simulate-and-bin. Adds track counts into H[n_k * n_de] (row major).
 * Returns the number of recorded tracks (including ones dropped because
 * their bin or length is out of range, which are counted in *dropped). */
uint64_t orc_simulate_f32(const orc_system* sys, const orc_proposal* p, uint64_t seed,
                          uint32_t* H, uint64_t* dropped);

/* FP64 twin: the same Philox words, FP64 libm math and FP64 positions.
 * sigma/half_az are taken from the *_d arrays (FP64 proposal). */
typedef struct orc_proposal_f64 {
    double half_az[ORC_MAX_SPECIES];
    double cum_prob[ORC_MAX_SPECIES];
    double sigma[ORC_MAX_SPECIES];
} orc_proposal_f64;
uint64_t orc_simulate_f64(const orc_system* sys, const orc_proposal* p,
                          const orc_proposal_f64* pd, double ub_de, double deltat,
                          uint64_t seed, uint32_t* H, uint64_t* dropped);

/* Track list form (generate_experiment.jl): writes up to cap (K, DE) pairs. */
uint64_t orc_tracks_f32(const orc_system* sys, const orc_proposal* p, uint64_t seed,
                        uint32_t* K, float* DE, uint64_t cap);

/* Trajectory of one particle (FP32 and FP64 twins) for drift reports:
 * writes frames*3 positions. */
void orc_trace_particle(const orc_system* sys, const orc_proposal* p, const orc_proposal_f64* pd,
                        uint64_t seed, uint32_t video, uint32_t particle,
                        float* xyz32, double* xyz64);

/* Dense textbook distance: S = cumsum(cumsum(H, rows), cols), same for the
 * observed histogram Hobs; kind 0: sum (S-R)^2, kind 1: sum |S-R|, summed in
 * 128 bits. Returns the exact integer sum, saturated at 2^64 - 1; *log10_out
 * gets log10 of the 128-bit sum, (2^64 hi + lo) rounded once to FP64. */
uint64_t orc_distance(const uint32_t* H, const uint32_t* Hobs, uint32_t n_k, uint32_t n_de,
                      int kind, double* log10_out);

/* The reference's float distance form (src/gpu/cusim/kernel-distance.inl):
 * FP32 accumulation of (float(S) - R)^2 in row sweep order. */
float orc_distance_ref_f32(const uint32_t* H, const uint32_t* Hobs, uint32_t n_k, uint32_t n_de);

/* Observed histogram from track lists (src/gpu/detail/simulation_impl.inl:605-616):
 * row K-1, column trunc(DE / dDE), skipping DE >= ub_de. Returns the number
 * of tracks counted; tracks outside [1, n_k] rows are counted in *dropped. */
uint64_t orc_reference_hist(const uint64_t* K, const double* DE, uint64_t n, double ub_de,
                            uint32_t n_k, uint32_t n_de, uint32_t* H, uint64_t* dropped);

/* 2D inclusive prefix sum (summed-area table) in uint32 (wrapping like the
 * reference's uint32 SAT). */
void orc_sat_u32(const uint32_t* H, uint32_t n_k, uint32_t n_de, uint32_t* S);

/* 10^x in FP64 by the device's operation sequence (Cody-Waite reduction by
 * log10 2, Taylor series of e^y to degree 16, fma Horner, ldexp). */
double orc_exp10(double x);

/* Prior and model of the ABC-PMC proposal step. */
typedef struct orc_prior {
    uint32_t n_components;  /* C */
    uint32_t n_depths;      /* Z: 1 (fixed depth) or C */
    double m_lo, m_hi, log10c_lo, log10c_hi, az_lo, az_hi;
    double deltat;          /* s */
    double volume;          /* Lx Ly Lz / 1e12 */
} orc_prior;

/* Proposal step of one ABC-PMC round (restating the reference's host loop,
 * src/gpu/detail/simulation_impl.inl:256-341 and src/cpu/estimate.jl:126-150,
 * with the device's Philox stream): for trial t, the parent is the first
 * member whose cumulative weight reaches a 53-bit uniform (wrand_idx_); each
 * of the P = 2C + Z parameters is displaced by tau_p times an FP32 Box-Muller
 * normal and reflected into its closed prior interval (displace_); the
 * components are sorted by m (stable); c = 10^log10c; cum_prob, sigma and
 * half_az as the per-trial precompute (with sigma = sqrtf of the FP32-rounded
 * 2 m dt); each of the J video counts is Poisson(sum c * volume), drawn as
 * the sum over 64 lanes of the arrivals of a unit-rate Poisson process on
 * [0, lambda / 64]. theta is P x n_pop and theta_out P x n_trials, row-major.
 * half_az, cum_prob and sigma get 8 floats per trial, particles J per trial,
 * parent_out one index per trial (each nullable). */
void orc_propose(const orc_prior* pr, const double* theta, const double* cum_w, const double* tau, uint32_t n_pop,
                 uint32_t n_trials, uint32_t J, uint64_t seed, double* theta_out, float* half_az, float* cum_prob,
                 float* sigma, uint32_t* particles, uint32_t* parent_out);

/* Poisson(lambda) count of the proposal step's stream for trial t, video v. */
uint32_t orc_poisson(double lambda, uint64_t seed, uint32_t trial, uint32_t video);

#ifdef __cplusplus
}
#endif

#endif
