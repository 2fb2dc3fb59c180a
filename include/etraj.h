/*
 * etraj.h -- C ABI of libetraj_hip.so, the gfx950 trajectory engine behind
 * eTraj.jl's `perform_traj_simulation` (sample -> integrate -> phase -> bin).
 *
 * Every entry point takes plain POD descriptors, plain pointers and sizes.
 * Units are atomic units except where a field says otherwise; the caller does
 * the Unitful conversions first, exactly as the reference constructors do
 * (src/Lasers/Cos2Laser.jl:49-55, src/Targets/HydrogenLikeAtom.jl:50-52).
 *
 * Conventions shared by all entry points
 *  - Return value: ETRAJ_OK (0) or a negative ETRAJ_E* code; the message is
 *    available from etraj_last_error(ctx) until the next call on that ctx.
 *    No C++ exception, abort or longjmp crosses this boundary.
 *  - A ctx owns one HIP stream and its device buffers. Calls are blocking
 *    (the stream is synchronised before return) and a ctx must be used by
 *    one host thread at a time; many ctxs may coexist.
 *  - Host buffers are owned by the caller (Julia: keep them alive with
 *    GC.@preserve). Matrices use Julia's column-major layout.
 *  - Per-trajectory failures never fail a call: they are counted, as the
 *    reference skips and warns (src/TrajectorySimulation.jl:347-355).
 */
#ifndef ETRAJ_H
#define ETRAJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ETRAJ_ABI_VERSION 4

/* status codes */
#define ETRAJ_OK             0
#define ETRAJ_EINVAL        (-1)  /* invalid argument (reference: AssertionError / error()) */
#define ETRAJ_EHIP          (-2)  /* HIP runtime failure */
#define ETRAJ_ENOMEM        (-3)  /* device allocation failed */
#define ETRAJ_EUNSUPPORTED  (-4)  /* option the engine does not implement */
#define ETRAJ_ECAPACITY     (-5)  /* caller buffer too small */
#define ETRAJ_ERANGE        (-6)  /* a fixed-point PMD sum left its range (ETRAJ_PMD_FIXED) */

/* ---- lasers: src/Lasers/{Cos2,Cos4,Gaussian,Bichromatic}Laser.jl ---------- */
enum {
    ETRAJ_LASER_COS2     = 1,   /* Cos2Laser      (Cos2Laser.jl:7-15)      */
    ETRAJ_LASER_COS4     = 2,   /* Cos4Laser      (Cos4Laser.jl)           */
    ETRAJ_LASER_GAUSSIAN = 3    /* GaussianLaser  (GaussianLaser.jl)       */
};

/* One MonochromaticLaser, the fields of its Serialize dict
 * (Cos2Laser.jl:192-204). `cyc_num` is cyc_num for Cos2/Cos4 and
 * spread_cyc_num for Gaussian. peak_int in W/cm^2, wave_len in nm, angles in
 * rad, t_shift in a.u. */
typedef struct etraj_mono_laser {
    int32_t kind;
    int32_t reserved;
    double  peak_int;
    double  wave_len;
    double  cyc_num;
    double  ellip;
    double  azi;
    double  cep;
    double  t_shift;
} etraj_mono_laser;

/* n_colors = 1: a MonochromaticLaser (color[0]).
 * n_colors = 2: BichromaticLaser(l1=color[0], l2=color[1], delay=delay21),
 *               F(t) = F1(t) + F2(t - delay21)  (BichromaticLaser.jl:73-85). */
typedef struct etraj_laser_desc {
    int32_t n_colors;
    int32_t reserved;
    double  delay21;
    etraj_mono_laser color[2];
} etraj_laser_desc;

/* ---- targets: src/Targets/{HydrogenLikeAtom,SAEAtom}.jl -------------------- */
enum {
    ETRAJ_TARGET_HYDROGEN_LIKE = 1,  /* HydrogenLikeAtom.jl:7-17 */
    ETRAJ_TARGET_SAE           = 2   /* SAEAtom.jl:7-23 (Tong-Lin potential) */
};

/* asymp_coeff is the resolved C_kl (Hartree formula already applied by the
 * caller when `:hartree` was passed, SAEAtomBase_shared.jl:3-10). */
typedef struct etraj_target_desc {
    int32_t kind;
    int32_t nucl_charge;
    int32_t l;
    int32_t m;
    double  Ip;
    double  asymp_coeff;
    double  quan_ax_theta;
    double  quan_ax_phi;
    double  soft_core;
    double  a1, b1, a2, b2, a3, b3;   /* SAE only; zero for HydrogenLike */
} etraj_target_desc;

/* ---- sampler: src/ElectronSamplers ------------------------------------------ */
enum {
    ETRAJ_INIT_ADK  = 1,   /* ADKSampler.jl                                                  */
    ETRAJ_INIT_SPA  = 2,   /* SPASampler.jl (SFA-SPA): complex saddle point per sample       */
    ETRAJ_INIT_SPANE = 3,  /* SPANESampler.jl (SFA-SPA with non-adiabatic expansion)         */
    ETRAJ_INIT_WFAT = 4    /* WFATSampler.jl, channel (nxi, m) = (0, 0); the structure factor
                              G00(beta, gamma) comes from `wfat_coeff` in one of the forms
                              ETRAJ_WFAT_*                                                   */
};
/* WFAT structure-factor forms (etraj_sampler_desc.wfat_form) */
enum {
    ETRAJ_WFAT_TABLE = 0,    /* channel (0, 0) only, G00 = sum_{a<na, b<nb} c_ab P_a(cos beta)
                                e^{-i b gamma}; c_ab at index a * nb + b                      */
    ETRAJ_WFAT_INTDATA = 1   /* every channel (nxi, m) of the WFAT integrals I (the molecule's
                                wfat_intdata[ridx], GenericMolecule.jl:236-260, WFATSampler.jl:
                                151-169): G_{nxi m} = sum_{l>=|m|} sum_{|m'|<=l}
                                I[nxi, m, l, m'] d^l_{m m'}(beta) e^{-i m' gamma}
                                exp(-kappa mu_z(beta)), mu_z = -sin(beta) mu_x + cos(beta) mu_z;
                                wfat_coeff is Julia's Array{ComplexF64,4} memory as is
                                (column-major, (re, im)), wfat_dims its size
                                (nxi_max+1, 2 m_max+1, l_max+1, 2 l_max+1), each <= 31       */
};
enum { ETRAJ_PHASE_CTMC = 0, ETRAJ_PHASE_QTMC = 1, ETRAJ_PHASE_SCTS = 2 };
/* rate_prefix bit set (ADKSampler.jl:49-77): :Exp = 0, :Full = PRECC|JAC */
enum { ETRAJ_PREFIX_PRE = 1, ETRAJ_PREFIX_PRECC = 2, ETRAJ_PREFIX_JAC = 4 };
enum {
    ETRAJ_SAMPLING_STEP       = 0,  /* equidistant t0 x kd [x kz] grid (ADKSampler.jl:133-139) */
    ETRAJ_SAMPLING_MONTECARLO = 1,  /* sorted random t0, uniform kd [kz] (ADKSampler.jl:140-147),
                                       with Philox4x64-10 keyed by `seed` instead of rand() */
    ETRAJ_SAMPLING_STRATIFIED = 2   /* engine extension: t0 in sample_t_num equal strata with
                                       one seeded jitter per stratum, mc_kt_num Gaussian kd [kz]
                                       (sigma kd_sigma [kz_sigma]) per stratum, weight / pdf */
};

typedef struct etraj_sampler_desc {
    int32_t  init_method;      /* ETRAJ_INIT_*                 */
    int32_t  dimension;        /* 2 or 3                       */
    int32_t  phase_method;     /* ETRAJ_PHASE_*                */
    int32_t  rate_prefix;      /* ETRAJ_PREFIX_* bits          */
    int32_t  sampling;         /* ETRAJ_SAMPLING_*             */
    int32_t  sample_t_num;     /* number of batches (t0 values) */
    int32_t  ss_kd_num;
    int32_t  ss_kz_num;
    int64_t  mc_kt_num;        /* samples per batch for MONTECARLO / STRATIFIED */
    double   sample_t_lo, sample_t_hi;
    double   ss_kd_max, ss_kz_max;
    double   mc_kd_max, mc_kz_max;
    double   kd_sigma, kz_sigma;
    double   cutoff_limit;
    uint64_t seed;
    /* molecules: orientation of the molecular frame (ZYZ Euler angles, as
     * MolRotation, GenericMolecule.jl:315); WFAT: the structure factor */
    double   mol_rot[3];
    int32_t  wfat_na, wfat_nb;
    const double* wfat_coeff;  /* (re, im) pairs, layout per wfat_form                  */
    int32_t  wfat_form;        /* ETRAJ_WFAT_*                                         */
    int32_t  wfat_dims[4];     /* ETRAJ_WFAT_INTDATA: size(wfat_intdata[ridx])         */
    int32_t  wfat_reserved;
    double   wfat_mu[3];       /* ETRAJ_WFAT_INTDATA: wfat_mu[ridx], the orbital dipole */
    double   orbit_ip;         /* molecules: Ip of the ionizing orbital, IonPotential(mol,
                                  ridx) (ADKSampler.jl:177-182, WFATSampler.jl:139); <= 0:
                                  the target's Ip. The EOM keeps the target's (HOMO) Ip,
                                  GenericMolecule.jl:457                                 */
    /* ADK / SPA / SPANE prefactor of a molecule (ADKSampler.jl:188-265, SPASampler.jl:
     * 204-278, SPANESampler.jl:208-266): the orbital's asymptotic coefficients C_lm,
     * MolAsympCoeff(mol, ridx), Julia's Matrix{ComplexF64} memory as is ((re, im),
     * column-major (asymp_lmax+1) x (2 asymp_lmax+1), C_lm at [l, m+l]) in the frame of
     * mol_rot. NULL: an atom's (l, m, asymp_coeff) and quantization axis from the
     * target descriptor (l = 0: the scalar C Y_00).                                    */
    int32_t  asymp_lmax;       /* 0..15 */
    int32_t  asymp_reserved;
    const double* asymp_coeff;
} etraj_sampler_desc;

/* ---- trajectory run ---------------------------------------------------------- */
enum { ETRAJ_INTEG_RK4 = 0, ETRAJ_INTEG_TSIT5 = 1 };
enum {
    ETRAJ_PREC_FP64  = 0,   /* every operation in FP64                                   */
    ETRAJ_PREC_MIXED = 1    /* RK4: FP32 EOM with Kahan-compensated state, Kahan FP64
                               phase; electrons that start any step within restep_radius of
                               the core are integrated again from t0 in FP64; FP64 Kepler
                               asymptote and FP64 PMD sums.
                               Tsit5: FP32 stage derivatives, FP64 state, stage sums and
                               error control; an electron with a stage within
                               restep_radius is integrated again from t0 in FP64 */
};
/* quirk bits: set = reproduce the reference's behaviour (default) */
enum {
    ETRAJ_QUIRK_SCTS3D_XPY = 1,  /* 3D SCTS asinh uses x*py (TrajectorySimulation.jl:463)   */
    ETRAJ_QUIRK_SAE2D_SCTS = 2   /* SAE 2D SCTS divides by (r^2+a) (SAEAtom.jl:136)         */
};
#define ETRAJ_QUIRKS_REFERENCE (ETRAJ_QUIRK_SCTS3D_XPY | ETRAJ_QUIRK_SAE2D_SCTS)
/* PMD accumulator of the device binning (etraj_run_job / etraj_run_scan) */
enum {
    ETRAJ_PMD_FLOAT = 0,    /* FP64 atomic additions: the reference's Float64 sums
                               (TrajectorySimulation.jl:385-390), rounding in arrival order */
    ETRAJ_PMD_FIXED = 1     /* every contribution (CTMC weight, amplitude re/im) and the
                               ion_prob sums rounded to a 128-bit two's-complement fixed
                               point with 100 fraction bits (resolution 2^-100 ~ 7.9e-31,
                               range |sum| < 2^26) and added with exact integer atomics:
                               order-independent, bit-identical from run to run and for any
                               slab split; ETRAJ_ERANGE if a sum leaves the range           */
};
/* How the device binning of etraj_run_job / etraj_run_scan adds the electrons to the PMD */
enum {
    ETRAJ_BIN_AUTO = 0,     /* the context's choice (etraj_tuning.bin_tiles): by default the
                               tiles for Tsit5 and the atomics for RK4                      */
    ETRAJ_BIN_ATOMIC = 1,   /* one global atomic addition per electron and value, from the
                               integrator kernel                                            */
    ETRAJ_BIN_TILES = 2     /* the integrator writes (cell, contribution) records; they are
                               sorted by PMD tile and each tile is accumulated in LDS
                               (privatised per workgroup) and added to the PMD in HBM once
                               per non-zero cell. Same sums (ETRAJ_PMD_FIXED: bit for bit);
                               the FP64 re-integration pass of MIXED runs adds its electrons
                               with atomics. ETRAJ_EUNSUPPORTED if the PMD needs more than
                               8192 tiles or 2^32 - 1 cells                                 */
};

typedef struct etraj_run_desc {
    int32_t dimension;
    int32_t phase_method;
    int32_t integrator;        /* ETRAJ_INTEG_*                                            */
    int32_t precision;         /* ETRAJ_PREC_*                                             */
    double  traj_t_final;
    double  dt;                /* RK4: nominal step; the step is (tf-t0)/ceil((tf-t0)/dt)  */
    double  rtol;              /* Tsit5: traj_rtol                                         */
    double  atol;              /* Tsit5: absolute tolerance (OrdinaryDiffEq default 1e-6)  */
    double  restep_radius;     /* MIXED: FP64 re-integration radius (a.u.)                 */
    double  final_p_max[3];
    int32_t final_p_num[3];
    int32_t quirks;            /* ETRAJ_QUIRK_* bits                                       */
    int64_t maxiters;          /* Tsit5 step-attempt limit (<= 0: 1000000)                  */
    int32_t pmd_mode;          /* ETRAJ_PMD_* (the device binning of run_job / run_scan)    */
    int32_t pmd_binning;       /* ETRAJ_BIN_* (the device binning of run_job / run_scan)   */
    uint64_t* hit_counts;      /* NULL, or host counters, one per PMD cell (px fastest; for
                                  run_scan n_jobs * prod(final_p_num), job-major): the number
                                  of electrons binned into each cell is ADDED to them (the
                                  integer-count PMD of a CTMC run)                          */
} etraj_run_desc;

typedef struct etraj_stats {
    double  ion_prob;               /* sum of W over E >= 0 trajectories  (:371)   */
    double  ion_prob_uncollected;   /* ... whose bin is out of range      (:392)   */
    int64_t num_effective_traj;     /* launched trajectories, failed included (:399) */
    int64_t n_failed;               /* integrator failures (skipped)        */
    int64_t n_bound;                /* E < 0 at the end (dropped)           */
    int64_t n_restep;               /* MIXED: electrons re-integrated in FP64 */
    int64_t n_steps;                /* integrator steps, all passes           */
    int64_t n_steps_fp64;           /* of which steps in FP64 (MIXED: the re-integration pass) */
    double  ms_sample;              /* device time of the sampler stage      */
    double  ms_integrate;           /* device time of integrate+phase+bin    */
    double  ms_integrate_main;      /* of which the main integrator kernel (MIXED: the FP32 pass) */
    double  ms_total;               /* wall time of the call                 */
    int64_t n_slabs;                /* run_job: slabs = launches of the main integrator kernel */
    int64_t n_attempts;             /* integrator step attempts, all passes (Tsit5: accepted +
                                       rejected; RK4: n_steps)                                */
} etraj_stats;

typedef struct etraj_ctx etraj_ctx;

/* ---- entry points ------------------------------------------------------------ */

int etraj_abi_version(void);

/* Opens a context on HIP device `device` (one stream, lazily grown buffers). */
int etraj_ctx_create(int device, etraj_ctx** out);
void etraj_ctx_destroy(etraj_ctx* ctx);
const char* etraj_last_error(const etraj_ctx* ctx);

/* LaserFx/LaserFy/LaserAx/LaserAy evaluated on the device at n times
 * (Lasers/exports.jl:6; e.g. Cos2Laser.jl:97-152). Any output may be NULL. */
int etraj_laser_field(etraj_ctx* ctx, const etraj_laser_desc* laser,
                      const double* t, int64_t n,
                      double* Fx, double* Fy, double* Ax, double* Ay);

/* TargetPotential / TargetForce on the device at n points xyz (3 x n,
 * column-major) (HydrogenLikeAtom.jl:75-76, SAEAtom.jl:84-96).
 * V: n values; F: 3 x n. Either output may be NULL. */
int etraj_target_eval(etraj_ctx* ctx, const etraj_target_desc* target,
                      const double* xyz, int64_t n, double* V, double* F);

/* The t0 of every batch: sampler.t_samples (ADKSampler.jl:133-147).
 * t0 must hold sampler->sample_t_num values. */
int etraj_batch_times(etraj_ctx* ctx, const etraj_sampler_desc* sampler, double* t0);

/* Upper bound of the number of columns gen_electron_batch can return
 * (ADKSampler.jl:158-164). */
int64_t etraj_batch_max_size(const etraj_sampler_desc* sampler);

/* Replaces gen_electron_batch(sp, batch_id) (ADKSampler.jl:167-396): writes the
 * surviving samples of batch `batch_id` (1-based, as in Julia) into `out`
 * (rows x capacity, column-major; rows = 6/7 in 2D, 8/9 in 3D, as the
 * reference), in the reference's sample order, and their number into *count
 * (0 = the reference's `nothing`). */
int etraj_gen_electron_batch(etraj_ctx* ctx, const etraj_laser_desc* laser,
                             const etraj_target_desc* target,
                             const etraj_sampler_desc* sampler, int32_t batch_id,
                             double* out, int64_t capacity, int64_t* count);

/* Replaces the body of launch_and_collect_2D!/3D! after the sampler
 * (TrajectorySimulation.jl:318-400 / 413-499): integrates the n electrons of
 * `init` (Julia rows x n matrix from gen_electron_batch) to traj_t_final,
 * applies the Kepler asymptote and the SCTS post-pulse phase, and ADDS their
 * contributions into `pmd` (prod(final_p_num) values, px fastest; double for
 * CTMC, interleaved complex double for QTMC/SCTS -- the amplitude sum, before
 * abs2). ion_prob/ion_prob_uncollected are added to, n_failed is set. */
int etraj_integrate_and_bin(etraj_ctx* ctx, const etraj_laser_desc* laser,
                            const etraj_target_desc* target, const etraj_run_desc* run,
                            const double* init, int64_t n, int32_t rows,
                            void* pmd, double* ion_prob, double* ion_prob_uncollected,
                            int64_t* n_failed);

/* Per-electron results of the same integration, without binning, for
 * diagnostics and parity checks: for each electron 12 doubles
 * [x y z px py pz phase_ode E_inf pinf_x pinf_y pinf_z phase_final]
 * (state at traj_t_final, reference :345-374) and status[i]
 * (0 collected-able E>=0, 1 integrator failure, 2 bound E<0). */
int etraj_integrate_records(etraj_ctx* ctx, const etraj_laser_desc* laser,
                            const etraj_target_desc* target, const etraj_run_desc* run,
                            const double* init, int64_t n, int32_t rows,
                            double* records, int32_t* status);

/* Whole perform_traj_simulation loop (TrajectorySimulation.jl:150-159 and the
 * sum in write_output! :521): device sampler -> integrator -> binning for all
 * batches. pmd_out (host) receives the final momentum_spec (abs2 already taken
 * for QTMC/SCTS), prod(final_p_num) doubles, px fastest. */
int etraj_run_job(etraj_ctx* ctx, const etraj_laser_desc* laser,
                  const etraj_target_desc* target, const etraj_sampler_desc* sampler,
                  const etraj_run_desc* run, double* pmd_out, etraj_stats* stats);

/* Same as etraj_run_job, with the output left in device memory owned by the
 * ctx: *pmd_dev receives a device pointer to the momentum_spec (valid until
 * the next call on ctx). Used where the result stays in HBM. */
int etraj_run_job_device(etraj_ctx* ctx, const etraj_laser_desc* laser,
                         const etraj_target_desc* target, const etraj_sampler_desc* sampler,
                         const etraj_run_desc* run, const double** pmd_dev, etraj_stats* stats);

/* A parameter scan: n_jobs independent jobs, job j with lasers[j] and
 * samplers[j] (intensity, polarization, orientation mol_rot, seeds, t0 grid
 * or sample counts may differ), sharing the target and the run descriptor.
 * Replaces the reference's sequential perform_traj_simulation calls over a
 * scanned parameter (examples/test_Bichromatic_CCP.jl:4-27 over intensities,
 * examples/test_Molecules.jl:20-36 over targets / orbitals): the batches of
 * all jobs share the sampler and integrator launches, each job keeping its
 * own PMD. Every job's result equals etraj_run_job of that job alone.
 * pmd_out: n_jobs * prod(final_p_num) doubles, job-major. stats: NULL or
 * n_jobs entries (the ms_* fields and n_slabs are those of the whole call). */
int etraj_run_scan(etraj_ctx* ctx, int32_t n_jobs, const etraj_laser_desc* lasers,
                   const etraj_target_desc* target, const etraj_sampler_desc* samplers,
                   const etraj_run_desc* run, double* pmd_out, etraj_stats* stats);

/* Same as etraj_run_scan with the n_jobs PMDs left in device memory owned by
 * the ctx (*pmd_dev, valid until the next call on ctx). */
int etraj_run_scan_device(etraj_ctx* ctx, int32_t n_jobs, const etraj_laser_desc* lasers,
                          const etraj_target_desc* target, const etraj_sampler_desc* samplers,
                          const etraj_run_desc* run, const double** pmd_dev, etraj_stats* stats);

/* ---- launch tuning --------------------------------------------------------------
 * No reference counterpart (the reference runs on CPU threads): the launch
 * choices of a ctx. Each choice changes the speed only; the results are the same
 * up to the order of FP64 sums (bit for bit with ETRAJ_PMD_FIXED). */
typedef struct etraj_tuning {
    int32_t bin_tiles;       /* what ETRAJ_BIN_AUTO means: 0 atomics, 1 LDS tiles, -1 (the
                                default) tiles for Tsit5 and atomics for RK4                  */
    int32_t bin_tile_shift;  /* tiles of 2^shift PMD cells, 6..16; 0: the largest that fits the
                                64 KiB LDS accumulator                                        */
    int32_t bin_chunk;       /* records per accumulating workgroup, 256..1048576               */
    int32_t overlap_fp64;    /* MIXED RK4: 1 runs each slab's FP64 re-integration on a second
                                stream beside the next slab's FP32 pass, 0 right after it      */
    int32_t rk4_block;       /* RK4 threads per workgroup: 64, 128 or 256                      */
    int32_t tsit5_refill;    /* Tsit5: idle lanes of a wavefront before it refills them, 1..64 */
    int32_t tsit5_claim;     /* Tsit5: electrons a wavefront claims from the queue, 64..65536  */
    int32_t graph;           /* 1: a job whose batches fit one slab (at most 2^26 init columns for
                                RK4, 2^29 for Tsit5) runs as a HIP graph, captured at its first
                                call and replayed while the call's launch arguments repeat (its stats then give the graph's
                                device time as ms_integrate, with ms_sample and ms_integrate_main
                                0); 0 (the default) launches every kernel on each call          */
} etraj_tuning;

/* The ctx's current choices. */
int etraj_get_tuning(const etraj_ctx* ctx, etraj_tuning* out);
/* Sets the ctx's choices (ETRAJ_EINVAL if a field is out of its range). */
int etraj_set_tuning(etraj_ctx* ctx, const etraj_tuning* tuning);
/* The autotuner: runs the job (as etraj_run_job_device) with the ctx's choices
 * and with candidates that change one choice at a time (RK4: workgroup size,
 * binning, tile size, FP64-pass overlap; Tsit5: refill threshold, queue claim,
 * binning, tile size; both: graph replay), each after a warm-up run, keeps the
 * fastest set in the ctx and returns it with its time (wall time of one job, ms)
 * and the time of the choices it started from. About 12-16 job runs. */
int etraj_autotune(etraj_ctx* ctx, const etraj_laser_desc* laser, const etraj_target_desc* target,
                   const etraj_sampler_desc* sampler, const etraj_run_desc* run, etraj_tuning* best,
                   double* best_ms, double* start_ms);

/* No reference counterpart: for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * results and errors as etraj_integrate_and_bin, with every device buffer a
 * slice of a static 32 MiB device array that each GPU the library runs on
 * holds until the process ends, and the host staging the ctx set up at
 * creation: the call allocates no device or host memory. Limits: n <= 65536
 * columns (ETRAJ_ECAPACITY beyond), and for RK4 one segment's field table
 * (ceil((traj_t_final - t0)/dt) steps, 48 B each) within about 22 MB. Calls
 * on one device share its array: calls that overlap on one device (two host
 * threads, two ctxs) race; consecutive calls need no clearing. */
int etraj_integrate_and_bin_noalloc(etraj_ctx* ctx, const etraj_laser_desc* laser,
                                    const etraj_target_desc* target, const etraj_run_desc* run,
                                    const double* init, int64_t n, int32_t rows,
                                    void* pmd, double* ion_prob, double* ion_prob_uncollected,
                                    int64_t* n_failed);

#ifdef __cplusplus
}
#endif

#endif /* ETRAJ_H */
