/*
 * mcrt_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference's per-ray transport
 * (src/cpu.jl, src/utils.jl, src/geometries/{plane,sphere}.jl, src/scatterers/{index,
 * rayleigh,mie,spline}.jl, src/config.jl, src/MonteCarloRadiativeTransfer.jl:26-85),
 * with Julia's rand() replaced by a counter-based Philox4x32-10 stream.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker; the product (libmcrt_hip.so) never does.
 *
 * It shares the plain-data structs of include/mcrt.h so that one description
 * of a solve feeds both the checker and the product.
 */
#ifndef MCRT_ORACLE_H
#define MCRT_ORACLE_H

#include <stdint.h>
#include "../include/mcrt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } or_cplx;

/* Philox4x32-10 block function and the uniform conversions. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_uniform(uint64_t seed, uint32_t instance, uint64_t ray, uint32_t draw);
float or_uniform_f32(uint64_t seed, uint32_t instance, uint64_t ray, uint32_t draw);

/* Frames and Stokes algebra (src/utils.jl). out4 = (cosθ, sinθ, cosϕ, sinϕ). */
void or_spherical_from_cartesian(const double x[3], double out4[4]);
void or_ray_to_norm(const double x[3], const double f[4], double out[3]);
void or_norm_to_ray(const double x[3], const double f[4], double out[3]);
void or_get_basis(const double k[3], double Eh[3], double Ev[3]);
void or_rotate(const double I[4], double c2, double s2, double out[4]);
void or_update_I(const mcrt_scatterer* s, const double I[4], const double k[3],
                 const double f[4], double out[4], double Eh[3], double Ev[3]);
void or_E_from_I(const double I[4], or_cplx E[2]);
void or_I_from_E(const or_cplx E[2], double out[4]);
void or_update_E(const mcrt_scatterer* s, const or_cplx E[2], const double k[3],
                 const double f[4], or_cplx out[2], double Eh[3], double Ev[3]);
double or_forward_E(const mcrt_scatterer* s, const or_cplx E[2], const double* kpath,
                    int64_t n, const double f0[4], or_cplx out[2]);
double or_backward_E(const mcrt_scatterer* s, const or_cplx E[2], const double* kpath,
                     int64_t n, const double f0[4], or_cplx out[2]);
double or_kepler_solver(double M, double e);

/* Scatterers.  phase_matrix evaluation at a given cosθ:
 * P = (P11, P12, P22, P33, P34, P44). */
void or_phase_eval(const mcrt_scatterer* s, double mu, double P[6]);
double or_phase_sample(const mcrt_scatterer* s, double u);
void or_amplitude(const mcrt_scatterer* s, double mu, or_cplx S[2]);

/* Table builders. */
int or_bhmie(double x, double m_re, double m_im, int n_theta, int nextra, int custom_nstop,
             double q[5], double* S1, double* S2, double* F11, double* F12, double* F33,
             double* F34);
int or_calculate_csrn(const double* F11, int n_theta, int random_points, double* csrn);
int or_gausslegendre(int n, double* x, double* w);
int or_initial_weights(int kind, int n_theta, int n_phi, double* ct, double* st,
                       double* inorm);

/* Geometry queries (plane.jl / sphere.jl).  idx is 0-based. */
double or_distance_to_boundary(const mcrt_geometry* g, const double x[3],
                               const double k[3], int idx);

/* One ray (cpu_kernel, src/cpu.jl:162-273).  Accumulates raw (un-normalised)
 * sums into A[5], Irt, Icb; counters[6] = events, rt_evals, cb_evals,
 * rejections, nonfinite, single-scattering CB evals. */
int or_trace_ray(const mcrt_config* cfg, const mcrt_geometry* g, const double I0[4],
                 uint64_t seed, uint32_t instance, uint64_t ray, double A[5], double* Irt,
                 double* Icb, uint64_t counters[6]);

/* collect_cb_cpu! (src/cpu.jl:38-160) for one event, for the checks of its
 * n = 1 and n > 1 bodies: I is the event's Stokes vector after absorption, k
 * the direction arriving at x, n the event number; kpath/xpath/idxpath hold
 * the path's first n directions, positions and 0-based layers (kpath[0] must
 * be the incident direction, as in cpu_kernel).  Adds into Icb[4*Nθb*Nϕb]. */
int or_collect_cb_event(const mcrt_config* cfg, const mcrt_geometry* g, const double I0[4],
                        const double I[4], const double k[3], const double x[3], int idx,
                        int64_t n, const double* kpath, const double* xpath, const int* idxpath,
                        double* Icb);

/* simulate's CPU branch (MonteCarloRadiativeTransfer.jl:57-76) over rays
 * [ray_begin, ray_end), split into nbatch contiguous equal batches.  Each
 * batch's raw sums go to batches[b*(5+nrt+ncb) ...] (A, Irt, Icb), so the
 * result is bitwise independent of nthreads.  Totals are the in-order sum of
 * the batches. */
int or_simulate(const mcrt_config* cfg, const mcrt_geometry* g, const double I0[4],
                uint64_t seed, uint32_t instance, int64_t ray_begin, int64_t ray_end,
                int nthreads, int nbatch, double* batches, uint64_t counters[6]);

#ifdef __cplusplus
}
#endif

#endif
