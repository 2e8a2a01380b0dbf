/*
 * clv_oracle.h -- FP64 CPU oracle of the complex-Langevin update.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load liboracle.so, and only as the checker
 * or the timed CPU baseline -- never as the product path.
 *
 * A plain C++17 restatement of the reference's per-step update
 * (Propagator_without_cutoff.cu:1-265, 351-477; Lattice.cu:2-11, 135-214;
 * Helpers.cu:1-15; Observables_without_cutoff.cu:35-101, 456-488,
 * 1155-1158, 1274-1334, 1527-1547, 1614-1634), with the reference's cuRAND
 * XORWOW stream replaced by Philox4x32-10 + Box-Muller (the engine's noise
 * stream) and cuFFT replaced by a project Stockham FFT of the same rank and
 * sign conventions.  Fields are complex128 in the reference layout
 * ((((c*Nt+tau)*X+x)*Y+y)*Z+z).  Parameters use the engine's clv_params.
 *
 * Pinning (tests/test_oracle.py): Philox against the published
 * known-answer vectors; the FFT against numpy.fft and an O(n^2)
 * DFT; the kinetic term by plane-wave eigenvalues; the drift against an
 * independent numpy action (-dS/dpsibar, -dS/dpsi) and closed-form special
 * cases; the dt controller by a closed-form table; max/mean by brute force;
 * observables by Parseval, the numpy energy density and Pauli-matrix spin
 * densities on random fields.  The reference ships no golden vectors; no
 * function here is left unpinned.
 */
#ifndef CLV_ORACLE_H_
#define CLV_ORACLE_H_

#include <stdint.h>
#include "../include/clv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): out = bijection(ctr; key). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* The engine's complex normal pair for (key, stream, step, site, pair):
 * out = {re_{2pair}, im_{2pair}, re_{2pair+1}, im_{2pair+1}}, each N(0,1). */
void orc_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t site,
                 uint32_t pair, double out[4]);

/* In-place unnormalised 1-D DFT, X[q] = sum_r x[r] exp(sign*2*pi*i*r*q/n),
 * data = n interleaved complex doubles.  sign = -1 is CUFFT_FORWARD, +1
 * CUFFT_INVERSE (Lattice.cu:158-180). */
void orc_fft1d(double *data, int32_t n, int32_t sign);
/* Rank-d transform over (X,Y,Z) of one (X*Y*Z) plane, z fastest -- the
 * cufftPlanMany plan of Lattice.cu:135-156 applied to one batch member. */
void orc_fft3d(double *data, int32_t nx, int32_t ny, int32_t nz, int32_t sign);

/* compute_new_dt (Propagator_without_cutoff.cu:382-404); clamp = -1/0/+1. */
double orc_new_dt(double mean_drift, double max_drift, const clv_params *p, int32_t *clamp);

orc_ctx *orc_create(const clv_shape *shape, const clv_params *p, int32_t nthreads);
void orc_destroy(orc_ctx *o);
void orc_set_params(orc_ctx *o, const clv_params *p);
void orc_set_fields(orc_ctx *o, const double *psi, const double *psibar);
void orc_get_fields(const orc_ctx *o, double *psi, double *psibar);
void orc_init_fields(orc_ctx *o, double n0, double var);

/* K and Kbar of the current fields (compute_kinetic_part + fft/fft_inv +
 * normalize, Propagator_without_cutoff.cu:410-424). */
void orc_kinetic(orc_ctx *o, double *kin, double *kinbar);
/* D and Dbar of the current fields (calculate_drift_without_noise, :1-129). */
void orc_drift(orc_ctx *o, double *drift, double *driftbar);
/* get_max_and_mean_drift (:351-380) on the current fields. */
void orc_max_mean(orc_ctx *o, double *max_drift, double *mean_drift);
/* One Propagator::propagate (or propagate_without_noise when noise_on == 0),
 * nsteps times; eval_every > 0 evaluates observables after step i when the
 * chain's step index i % eval_every == 0 (main.cu:152-156). */
void orc_propagate(orc_ctx *o, int64_t nsteps, int32_t eval_every);
/* Drift of the current fields, then the update with a caller-given dt
 * (:427, 432-441). */
void orc_apply_update(orc_ctx *o, double dt);
void orc_get_step_stats(const orc_ctx *o, clv_step_stats *out);

void orc_evaluate(orc_ctx *o);
void orc_read_observables(const orc_ctx *o, clv_observables *out, double *Fx,
                          double *Fy, double *Fz);
void orc_reset_observables(orc_ctx *o);
/* The observable sums of one evaluation of the current fields, unaccumulated:
 * out = {n_0..n_3 (k-space, as evaluate_n_*), n_0..n_3 (real space), energy}. */
void orc_observable_sums(orc_ctx *o, double out[9]);

#ifdef __cplusplus
}
#endif

#endif /* CLV_ORACLE_H_ */
