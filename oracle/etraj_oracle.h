/*
 * etraj_oracle.h -- CPU FP64 restatement of eTraj.jl's trajectory path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline of bench.py; never by the product library.
 * Descriptor structs are the product's C-ABI structs (include/etraj.h), so
 * one set of inputs drives both.
 */
#ifndef ETRAJ_ORACLE_H
#define ETRAJ_ORACLE_H

#include "../include/etraj.h"

#ifdef __cplusplus
extern "C" {
#endif

void oracle_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
double oracle_grid_point(double lo, double hi, int n, int i);
int64_t oracle_bin_index(double p, double pmax, int32_t num);
void oracle_tsit5_tableau(double* c, double* a, double* b, double* btilde);

int oracle_laser_field(const etraj_laser_desc* d, const double* t, int64_t n,
                       double* Fx, double* Fy, double* Ax, double* Ay);
double oracle_unit_envelope(const etraj_laser_desc* d, double t);
int oracle_target_eval(const etraj_target_desc* d, const double* xyz, int64_t n, double* V, double* F);
int oracle_eom_rhs(const etraj_laser_desc* ld, const etraj_target_desc* td, int32_t dim, int32_t phase_method,
                   int32_t quirks, double t, const double* u, double* du);

void oracle_ff_mf_euler(const double mol_rot[3], double Fx, double Fy, double out[3]);
void oracle_wfat_G(const etraj_sampler_desc* sd, double kappa, double beta, double gamma, int nxi, int m,
                   double out[2]);
double oracle_wigner_d(int l, int m, int mp, double beta);
void oracle_sph_harm_c(int l, int m, double kappa, const double xyz[6], double out[2]);
int oracle_laser_complex(const etraj_laser_desc* d, double t_re, double t_im, double out[8]);
int oracle_laser_field_deriv(const etraj_laser_desc* d, double t, double out[4]);
int oracle_spa_point(const etraj_laser_desc* ld, const etraj_target_desc* td, double tr, double kd, double kz,
                     double out[12]);
int oracle_batch_times(const etraj_sampler_desc* sd, double* t);
int64_t oracle_batch_max_size(const etraj_sampler_desc* sd);
int oracle_gen_electron_batch(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                              int32_t batch_id, double* out, int64_t capacity, int64_t* count);

int oracle_integrate_state(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_run_desc* rd,
                           double t0, double* u, int64_t* nsteps);
int oracle_post_process(const etraj_target_desc* td, int32_t dim, int32_t phase_method, int32_t quirks,
                        const double* u, double t0, double* rec, int32_t* status);
int oracle_integrate_records(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_run_desc* rd,
                             const double* init, int64_t n, int32_t rows, double* records, int32_t* status);
int oracle_integrate_and_bin(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_run_desc* rd,
                             const double* init, int64_t n, int32_t rows, double* pmd, double* ion_prob,
                             double* ion_prob_uncollected, int64_t* n_failed);
int oracle_run_job(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                   const etraj_run_desc* rd, int nthreads, int32_t batch_begin, int32_t batch_end,
                   int32_t batch_stride, double* pmd_out, etraj_stats* st);
/* oracle_run_job with the batches scheduled dynamically over the threads (timing runs) */
int oracle_run_job_dynamic(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                   const etraj_run_desc* rd, int nthreads, int32_t batch_begin, int32_t batch_end,
                   int32_t batch_stride, double* pmd_out, etraj_stats* st);
int oracle_run_job_amplitude(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                             const etraj_run_desc* rd, int nthreads, double* amp_out, etraj_stats* st);

#ifdef __cplusplus
}
#endif

#endif
