// etraj_engine.hpp -- kernel-side parameter blocks and launcher declarations
// shared by the kernels (etraj_kernels.hip) and the C ABI (etraj_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "etraj_physics.hpp"

namespace etraj {

// Resolved ADK sampler parameters for one job (ADKSampler.jl:19-150 + the
// per-job constants of gen_electron_batch).
// WFAT integral tables (ETRAJ_WFAT_INTDATA): each dimension <= kWfatMaxDim,
// at most kWfatMaxChannels channels (nxi, m), |m| <= kWfatMaxM
constexpr int kWfatMaxDim = 31, kWfatMaxChannels = 512, kWfatMaxM = 15;
// Prefactor beyond l = 0 (ADK / SPA / SPANE): l_max <= kPreMaxL; the per-batch
// table has one complex coefficient per monomial u^a v^b z^c, a+b+c <= l_max
// SFA-SPA tunnelling integrals: Gauss-Legendre points on [0, ti]. The
// integrands are sums of cosh/sinh(omega tau) products (derivatives grow like
// omega^n), and omega ti ~ asinh(omega sqrt(k^2 + 2 Ip) / F) stays below about
// 7 for the batches the sampler keeps (F(tr) > 1e-3, wavelengths down to
// 400 nm); there the 16-point rule's relative error bound
// (omega ti)^33 (16!)^4 / (33 (32!)^3) is below 1e-25.
constexpr int kSpaGlPoints = 16;
constexpr int kPreMaxL = 15, kPreMaxMono = (kPreMaxL + 1) * (kPreMaxL + 2) * (kPreMaxL + 3) / 6;

struct SamplerP {
    int dim, phase, sampling, prefix;
    int t_num, kd_num, kz_num, rows;
    int64_t nk;        // samples per batch
    int64_t max_bs;    // column stride of one batch in the init buffer
    double t_lo, t_hi;
    double kd_max, kz_max, mc_kd_max, mc_kz_max, kd_sigma, kz_sigma;
    double cutoff;
    double dkdt_base;  // dk dt of one sample before any 1/pdf factor (:267-276)
    double c_pre;      // c = 2^(n/2+1) kappa^(2n+1/2) Gamma(n/2+1) (:233)
    double numer;      // C * D^0_00 * Y_00 (:255)
    double nstar;      // n* = Z/kappa (:232)
    uint64_t seed;
    int method;        // ETRAJ_INIT_*
    double omega;      // SPA: initial-guess frequency (SPASampler.jl:160-166)
    const double* glq; // SPA: kSpaGlPoints Gauss-Legendre (node, weight) pairs on [-1, 1]
    int wfat_na, wfat_nb;
    double mol_rot[3];
    const double* wfat_coeff;  // device copy of the structure-factor data (re, im), layout per wfat_form
    int wfat_form;             // ETRAJ_WFAT_*
    int wfat_dims[4];          // ETRAJ_WFAT_INTDATA: (nxi_max+1, 2 m_max+1, l_max+1, 2 l_max+1)
    double wfat_mu[3];
    double Ip;                 // the sampler's Ip: the ionizing orbital's for a molecule, else the target's
    int pre_L;                 // -1: the scalar prefactor numer (atom with l = 0); else l_max of pre_C
    const double* pre_C;       // device: (re, im) of C_lm at [l + (pre_L+1)(m+l)] (column-major)
    double pre_rot[3];         // frame of C_lm: MolRotation, or (phi, theta, 0) of an atom's axis
};

struct IntegP {
    LaserP L;
    TargetP T;
    int dim, phase, mixed, quirks, rows;
    int pnum[3];
    double pmax[3];
    double tf, dt, restep_radius;
    double rtol, atol;
    int64_t maxiters;
    int integrator;
    // launch choices (etraj_tuning): RK4 threads per workgroup; Tsit5 idle lanes
    // before a refill and electrons per queue claim
    int rk4_block, ts5_refill, ts5_claim;
};

// per-electron binning record for host-side accumulation in column order
struct Contrib {
    int64_t lin;   // 0-based linear PMD index, -1 if none
    double re, im; // contribution (CTMC: re = W)
    double prob;
    int32_t status;  // 0 collected, 1 failed, 2 bound, 3 ionized but outside the grid
    int32_t pad;
};

// OUT_STATE (Tsit5 in run_job): the kernel leaves each electron's final state
// (r, p, phase; x = NaN if its integration failed) in its init column for
// k_ts5_finish, which bins it; the start values of the adaptive step come from
// k_ts5_prep.
enum { OUT_PMD = 0, OUT_RECORDS = 1, OUT_CONTRIB = 2, OUT_STATE = 3 };
// CNT_RESTEP counts the electrons of the mixed pass re-integrated in FP64
// CNT_LAUNCHED counts the sampled electrons (the sampler kernel adds each batch's survivors)
// CNT_FIXRANGE counts values too large for the fixed-point accumulator
// CNT_ATTEMPTS counts integrator step attempts (Tsit5: accepted + rejected; RK4: = CNT_STEPS)
enum {
    CNT_FAILED = 0, CNT_BOUND = 1, CNT_RESTEP = 2, CNT_STEPS = 3, CNT_STEPS64 = 4, CNT_LAUNCHED = 5,
    CNT_FIXRANGE = 6, CNT_ATTEMPTS = 7, CNT_N = 8
};

// A job's scalar sums (scal, scalq) and counters (cnt) are spread over
// kAccSlots copies, one per workgroup index mod kAccSlots: a launch's
// per-wavefront atomic additions then go to kAccSlots addresses instead of
// one (same-address atomics serialise at the memory side). The host adds the
// copies. Layout per job: scal [slot][2], scalq [slot][4], cnt [slot][CNT_N].
constexpr int kAccSlots = 32;

// Order-independent sums (ETRAJ_PMD_FIXED): a value v is held as the 128-bit
// two's-complement integer round(v * 2^kFixFrac) in two 64-bit words (lo,
// hi). Integer addition is associative, so the sum of any set of values is
// the same bits in any order. |v| < 2^kFixRange for each value and for every
// partial sum keeps the integer below 2^126.
constexpr int kFixFrac = 100, kFixRange = 26;
struct Fix128 {
    uint64_t lo, hi;
};

// round(v 2^kFixFrac), ties to even; ok = false if |v| >= 2^kFixRange or v is not finite
__host__ __device__ inline Fix128 fix_from_double(double v, bool& ok) {
    Fix128 r{0, 0};
    ok = true;
    if (v == 0.0) return r;
    if (!(fabs(v) < 0x1.0p26)) {
        ok = false;
        return r;
    }
    int e;
    const double f = frexp(fabs(v), &e);                 // |v| = f 2^e, f in [1/2, 1)
    const uint64_t m = (uint64_t)ldexp(f, 53);            // the 53-bit significand, exact
    const int sh = e - 53 + kFixFrac;                     // |v| 2^kFixFrac = m 2^sh, sh <= 73
    if (sh >= 64) {
        r.hi = m << (sh - 64);
    } else if (sh > 0) {
        r.lo = m << sh;
        r.hi = m >> (64 - sh);
    } else if (sh == 0) {
        r.lo = m;
    } else if (sh > -54) {                                // m 2^sh >= 1/2 is possible
        const int rs = -sh;
        uint64_t q = m >> rs;
        const uint64_t rem = m & ((1ull << rs) - 1), half = 1ull << (rs - 1);
        if (rem > half || (rem == half && (q & 1))) ++q;
        r.lo = q;
    }
    if (v < 0) {
        r.lo = ~r.lo + 1;
        r.hi = ~r.hi + (r.lo == 0 ? 1 : 0);
    }
    return r;
}

// the double nearest to the fixed-point value (round to nearest even)
__host__ __device__ inline double fix_to_double(uint64_t lo, uint64_t hi) {
    const bool neg = (int64_t)hi < 0;
    if (neg) {
        lo = ~lo + 1;
        hi = ~hi + (lo == 0 ? 1 : 0);
    }
    double d;
    if (hi == 0) {
        d = (double)lo;
    } else {
        const int n = 64 - __builtin_clzll(hi);           // bits to drop so that the value fits 64
        const uint64_t low_mask = (n == 64) ? ~0ull : ((1ull << n) - 1);
        const uint64_t mant = (n == 64 ? hi : ((hi << (64 - n)) | (lo >> n))) | ((lo & low_mask) ? 1 : 0);
        d = ldexp((double)mant, n);
    }
    d = ldexp(d, -kFixFrac);
    return neg ? -d : d;
}

// Per-segment laser field table (k_field_table): entry s of a segment holds
// the field at half-steps 2s+1 and 2s+2 of that segment's step grid.
struct FieldTab {
    const float4* f32;   // stride entries per segment: (Fx, Fy) at 2s+1, (Fx, Fy) at 2s+2
    const double2* f64;  // 2*stride entries per segment, same values in FP64
    const double2* f0;   // F(t0) per segment
    int64_t stride;      // >= max steps of any segment
};

// Outputs of the integrator kernels. For a parameter scan (several jobs in
// one launch) seg_job maps a segment to its job, whose PMD starts at
// pmd + job * job_stride and whose scalars and counters are scal + 2 * job and
// cnt + job * CNT_N; seg_job == nullptr means one job.
// With pmdq != nullptr (ETRAJ_PMD_FIXED) the PMD and the two scalars go to
// pmdq (two words per value, job stride 2 * job_stride) and scalq (4 words per
// job) instead of pmd and scal. hits (may be null): one counter per cell, job
// stride npix.
// bin_cell (may be null): the PMD contributions of an ETRAJ_PMD_* launch are
// written as records for the tile binning (etraj_bin.hip) instead of being
// added to pmd / pmdq / hits: bin_cell[col] = cell0 + cell (kNoCell if the
// electron is not binned) and bin_amp[col] = (re, im) of its contribution;
// cell0 = job * npix.
struct OutP {
    double* pmd;       // CTMC: npix; coherent: 2*npix (re, im) per job
    double* scal;      // [ion_prob, ion_prob_uncollected] per job and slot (kAccSlots)
    int64_t* cnt;      // CNT_* per job and slot (kAccSlots)
    uint64_t* pmdq;    // fixed-point PMD, 2 words per value
    uint64_t* scalq;   // fixed-point [ion_prob, ion_prob_uncollected], 4 words per job and slot
    uint64_t* hits;    // electrons binned per cell
    int64_t npix;
    double* records;   // 12 per electron
    int32_t* status;
    Contrib* contrib;
    const int32_t* seg_job;
    int64_t job_stride;
    uint32_t* bin_cell;
    double2* bin_amp;
    int64_t cell0;
};

// ---- tile binning (etraj_bin.hip) ----------------------------------------
constexpr uint32_t kNoCell = 0xffffffffu;
constexpr int kBinMaxTiles = 8192;            // tiles of one binning (all jobs of a slab)
constexpr int kBinTileBytes = 64 * 1024;      // LDS of one tile's accumulator
// Tile 2^shift cells; chunk: records per accumulating workgroup
struct BinCfg {
    int shift, chunk;
};
struct BinBufs {
    uint32_t* cell;        // records: n (written by the integrator)
    double2* amp;
    uint32_t* tile_count;  // kBinMaxTiles, zero between binnings
    uint32_t* tile_off;    // kBinMaxTiles + 1
    uint32_t* chunk_off;   // kBinMaxTiles + 1
    uint32_t* cursor;      // kBinMaxTiles
    uint32_t* scell;       // records sorted by tile: n
    double2* samp;
};
struct BinOut {
    double* pmd;       // ETRAJ_PMD_FLOAT: words per cell, all jobs (cell g at words * g)
    uint64_t* pmdq;    // ETRAJ_PMD_FIXED: 2 words per value
    uint64_t* hits;    // may be null
    int64_t* cnt;      // per-job counters (CNT_FIXRANGE)
    int64_t npix;
};
// The largest tile (2^shift cells, shift <= max_shift) whose accumulator fits kBinTileBytes.
int bin_tile_shift(int words, bool fixed, bool hits, int max_shift);
// cell[0, n) = kNoCell, tile_count[0, ntile) = 0 (before the integrator writes the records)
hipError_t launch_bin_reset(hipStream_t st, uint32_t* cell, int64_t n, uint32_t* tile_count, int ntile);
// Adds the records cell/amp[0, n) of Bf into out; total_cells = jobs * npix < 2^32 - 1,
// at most kBinMaxTiles tiles. tile_count must be zero on entry and is zero on return.
hipError_t launch_bin_tiles(hipStream_t st, const BinCfg& cfg, const BinBufs& Bf, int64_t n, int64_t total_cells,
                            int words, bool fixed, const BinOut& out);

// Per-segment (one t0 batch) description for the sampler and the field table:
// the batch's t0, its job, its index within the job (the Philox counter) and
// its first column in the init buffer.
struct SegTab {
    const double* t0;
    const int32_t* job;
    const int32_t* batch;
    const int64_t* start;
};

hipError_t launch_laser_field(hipStream_t st, const LaserP& L, const double* t, int64_t n, double* Fx, double* Fy,
                              double* Ax, double* Ay);
hipError_t launch_target_eval(hipStream_t st, const TargetP& T, const double* xyz, int64_t n, double* V, double* F);
// sps / lps: device arrays of per-job sampler and laser parameters
// cnt: per-job counters (CNT_N per job) receiving CNT_LAUNCHED, or NULL
// pre: the samplers carry a prefactor table (SamplerP::pre_L >= 0), the same for every job
hipError_t launch_adk_sample(hipStream_t st, int method, bool pre, const SamplerP* sps, const LaserP* lps,
                             const TargetP& T, const SegTab& G, int nb, double* out, int32_t* counts, int64_t* cnt);
// Doubles of the prefactor table C_lm of a sampler (0 for the scalar form)
inline size_t pre_doubles(const SamplerP& S) {
    return S.pre_L < 0 ? 0 : (size_t)2 * (S.pre_L + 1) * (2 * S.pre_L + 1);
}
// Doubles of the WFAT structure-factor data of a sampler (0 for other samplers)
inline size_t wfat_doubles(const SamplerP& S) {
    if (S.method != ETRAJ_INIT_WFAT) return 0;
    if (S.wfat_form == ETRAJ_WFAT_INTDATA)
        return (size_t)2 * S.wfat_dims[0] * S.wfat_dims[1] * S.wfat_dims[2] * S.wfat_dims[3];
    return (size_t)2 * S.wfat_na * S.wfat_nb;
}
hipError_t launch_rk4(hipStream_t st, const IntegP& P, const OutP& O, int out, const double* init,
                      const int64_t* seg_start, const int32_t* seg_count, int64_t n_seg, int chunks_per_seg,
                      const FieldTab& FT, uint8_t* flags, const int64_t* out_index);
hipError_t launch_compact_flagged(hipStream_t st, double* init, int rows, const int64_t* seg_start,
                                  const int32_t* seg_count, int64_t n_seg, const uint8_t* flags,
                                  const int64_t* out_index, int32_t* count_out, int64_t* orig);
// lps: device array of per-job lasers; seg_job may be NULL (job 0)
hipError_t launch_field_table(hipStream_t st, const LaserP* lps, const int32_t* seg_job, const double* t0s,
                              int64_t nseg, double tf, double dt, const FieldTab& FT);
// P.mixed: FP32 stage derivatives; flags (required) receives 1 for the columns
// to integrate again in FP64 and 0 for the others. orig (may be null) maps a
// column to its output index. queue: one device counter of the launch's work
// queue (set to zero by the launch itself; launches that may run at the same
// time need counters of their own).
hipError_t launch_tsit5(hipStream_t st, const IntegP& P, const OutP& O, int out, const double* init,
                        const int64_t* seg_start, const int64_t* pre, int64_t n_seg, int64_t max_items,
                        unsigned long long* queue, uint8_t* flags, const int64_t* orig);
// Doubles per column of the Tsit5 start values (ts5_prep_doubles): the first
// step size, the first stage derivative (p part, phase part) and, for a simple
// laser (simple_laser), the carrier's and the envelope's sin and cos at t0.
constexpr int ts5_prep_doubles(int dim) { return dim + 6; }
// The Tsit5 run of run_job in three launches over the columns
// [seg_start[s], seg_start[s] + seg_count[s]) of the segments (pre: the
// exclusive prefix sums of seg_count, pre[n_seg] <= max_items; cps >=
// ceil(max seg_count / 64)): k_ts5_prep (the start values of every electron
// into prep), k_tsit5 with OUT_STATE (final states into the init columns,
// which it writes; P.mixed: flags as launch_tsit5), k_ts5_finish (Kepler
// asymptote and binning of every final state into O, skipping flagged columns
// when flags is not null). main_start / main_done (may be null) are recorded
// right before and after k_tsit5.
hipError_t launch_tsit5_staged(hipStream_t st, const IntegP& P, const OutP& O, double* init, const int64_t* seg_start,
                               const int32_t* seg_count, const int64_t* pre, int64_t n_seg, int cps,
                               int64_t max_items, unsigned long long* queue, double* prep, uint8_t* flags,
                               hipEvent_t main_start, hipEvent_t main_done);
// k_ts5_finish alone (etraj_kernels.hip): the binning of the final states
hipError_t launch_ts5_finish(hipStream_t st, const IntegP& P, const OutP& O, const double* init,
                             const int64_t* seg_start, const int32_t* seg_count, int64_t n_seg, int cps,
                             const uint8_t* flags);
hipError_t launch_scan_counts(hipStream_t st, const int32_t* count, int64_t n, int64_t* pre);
hipError_t launch_finalize_abs2(hipStream_t st, const double* amp, int64_t n, double* out);
// zeroes `bytes` (a multiple of 8) at p with a kernel (capturable, unlike hipMemsetAsync)
hipError_t launch_zero(hipStream_t st, void* p, int64_t bytes);
// out[i] = fix_to_double(q[2i], q[2i+1]), i < n
hipError_t launch_fix_to_f64(hipStream_t st, const uint64_t* q, int64_t n, double* out);

// The no-allocation entry point (etraj_integrate_and_bin_noalloc): its device
// buffers are slices of a static device array of kNoAllocScratchBytes per
// device. Limits: kNoAllocMaxCols columns per call, and the RK4 field table of
// a group of segments within what the array leaves after the column buffers.
constexpr int64_t kNoAllocMaxCols = 65536;
constexpr int64_t kNoAllocScratchBytes = (int64_t)32 << 20;
constexpr int kNoAllocMaxDevices = 64;
// The device address of `device`'s copy of the scratch array (that device must be current).
hipError_t noalloc_scratch(int device, void** out);

}  // namespace etraj
