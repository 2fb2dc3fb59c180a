// etraj_capi.cpp -- the extern "C" boundary of libetraj_hip.so (include/etraj.h).
//
// Validates the descriptors the way the reference's constructors assert
// (TrajectorySimulation.jl:178-186, ADKSampler.jl:43-105), resolves them into
// kernel parameter blocks, owns the device buffers of a context, and drives
// the kernels on the context's stream. No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "etraj_engine.hpp"

using namespace etraj;

namespace {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            n = 0;
            if (e != hipSuccess) return e;
        }
        size_t bytes = std::max<size_t>(want, 1) * sizeof(T);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        n = std::max<size_t>(want, 1);
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

// The buffers of one slab in flight. run_job cycles through kSlabSlots slots
// so that a slab's FP64 re-integration (second stream) overlaps the next
// slab's sampling and FP32 pass (first stream). (Three slots, and slabs
// ordered by expected work, measured no faster on the 1e8-electron job.)
constexpr int kSlabSlots = 2;
struct SlabSlot {
    DevBuf<double> init, seg_t0;
    DevBuf<int32_t> counts, counts2, seg_job, seg_b;
    DevBuf<int64_t> seg_start, pre;
    DevBuf<float4> ft32;
    DevBuf<double2> ft64, ft0;
    DevBuf<uint8_t> flags;
    DevBuf<unsigned long long> queue;   // work-queue counter of the slot's Tsit5 launches
    DevBuf<double> prep;                // Tsit5 start values (k_ts5_prep), ts5_prep_doubles per column
    std::vector<double> h_t0;
    std::vector<int32_t> h_job, h_b;
    std::vector<int64_t> h_start;
    hipEvent_t done = nullptr;   // every kernel of the slot's last slab finished
    bool pending = false;
    void release() {
        init.release(); seg_t0.release(); counts.release(); counts2.release(); seg_job.release(); seg_b.release();
        seg_start.release(); pre.release(); ft32.release(); ft64.release(); ft0.release(); flags.release();
        queue.release(); prep.release();
        if (done) (void)hipEventDestroy(done);
        done = nullptr;
    }
};

struct etraj_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;   // FP64 re-integration passes of run_job (default priority)
    bool overlap_fp64 = true;        // run_job: FP64 passes on stream2, beside the next slab's FP32 pass
    // tile binning (ETRAJ_BIN_TILES): records, their tile-sorted copy and the tile tables
    DevBuf<uint32_t> bin_cell, bin_scell, tile_count, tile_off, chunk_off, cursor;
    DevBuf<double2> bin_amp, bin_samp;
    // what ETRAJ_BIN_AUTO means: 1 tiles, 0 atomics, -1 by integrator (tiles for Tsit5, whose binning
    // runs in its own short kernel, k_ts5_finish; atomics for RK4, whose binning spreads over the
    // integrator kernel and costs nothing there)
    int bin_auto_tiles = -1;
    int bin_shift = 0;               // tile of 2^bin_shift cells (0: the largest that fits the LDS)
    int bin_chunk = 16384;           // records per accumulating workgroup
    int rk4_block = 256;             // RK4 threads per workgroup
    int ts5_refill = 4;              // Tsit5: idle lanes before a wavefront refills
    int ts5_claim = 256;             // Tsit5: electrons per queue claim
    // run_job of a job whose batches fit one slab: its launches captured once into a HIP graph
    // (gexec), replayed while the launch arguments (gkey) repeat
    bool graph = false;
    hipGraphExec_t gexec = nullptr;
    std::string gkey;
    SlabSlot slot[kSlabSlots];
    std::string err;
    DevBuf<double> init, pmd, pmd_out, io1, io2, io3, io4, scal, t0s, seg_t0, wfat, prec;
    DevBuf<float4> ft32;
    DevBuf<double2> ft64, ft0;
    DevBuf<int64_t> seg_start, cnt, pre;
    DevBuf<int32_t> counts, counts2, status, seg_job, seg_b;
    DevBuf<SamplerP> sps;
    DevBuf<double> glq;   // kSpaGlPoints Gauss-Legendre (node, weight) pairs for the SPA sampler
    DevBuf<LaserP> lps;
    DevBuf<uint8_t> flags;
    DevBuf<int64_t> orig;
    DevBuf<unsigned long long> queue;   // work-queue counter of the Tsit5 launches on `stream`
    DevBuf<Contrib> contrib;
    DevBuf<uint64_t> pmdq, scalq, hits;   // ETRAJ_PMD_FIXED sums, hit counts
    // host staging of integrate_columns: kept in the ctx so that every asynchronous
    // host-to-device copy reads live memory until the call synchronises the stream
    std::vector<int64_t> h_ss, h_pre;
    std::vector<int32_t> h_sc;
    std::vector<double> h_t0;
    std::vector<Contrib> h_contrib;   // per-electron records copied back for host binning
    void* scratch = nullptr;          // this device's no-allocation scratch array (noalloc_scratch)
};

namespace {

int fail(etraj_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(etraj_ctx* c, hipError_t e, const char* what) {
    return fail(c, e == hipErrorOutOfMemory ? ETRAJ_ENOMEM : ETRAJ_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                  \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
    } while (0)

etraj_tuning tuning_of(const etraj_ctx* c) {
    etraj_tuning t{};
    t.bin_tiles = c->bin_auto_tiles;
    t.bin_tile_shift = c->bin_shift;
    t.bin_chunk = c->bin_chunk;
    t.overlap_fp64 = c->overlap_fp64 ? 1 : 0;
    t.rk4_block = c->rk4_block;
    t.tsit5_refill = c->ts5_refill;
    t.tsit5_claim = c->ts5_claim;
    t.graph = c->graph ? 1 : 0;
    return t;
}

const char* tuning_invalid(const etraj_tuning& t) {
    if (t.bin_tiles < -1 || t.bin_tiles > 1) return "tuning.bin_tiles must be -1, 0 or 1";
    if (t.bin_tile_shift != 0 && (t.bin_tile_shift < 6 || t.bin_tile_shift > 16))
        return "tuning.bin_tile_shift must be 0 or 6..16";
    if (t.bin_chunk < 256 || t.bin_chunk > (1 << 20)) return "tuning.bin_chunk must be 256..1048576";
    if (t.overlap_fp64 != 0 && t.overlap_fp64 != 1) return "tuning.overlap_fp64 must be 0 or 1";
    if (t.rk4_block != 64 && t.rk4_block != 128 && t.rk4_block != 256) return "tuning.rk4_block must be 64, 128 or 256";
    if (t.tsit5_refill < 1 || t.tsit5_refill > 64) return "tuning.tsit5_refill must be 1..64";
    if (t.tsit5_claim < 64 || t.tsit5_claim > 65536) return "tuning.tsit5_claim must be 64..65536";
    if (t.graph != 0 && t.graph != 1) return "tuning.graph must be 0 or 1";
    return nullptr;
}

void apply_tuning(etraj_ctx* c, const etraj_tuning& t) {
    c->bin_auto_tiles = t.bin_tiles;
    c->bin_shift = t.bin_tile_shift;
    c->bin_chunk = t.bin_chunk;
    c->overlap_fp64 = t.overlap_fp64 != 0;
    c->rk4_block = t.rk4_block;
    c->ts5_refill = t.tsit5_refill;
    c->ts5_claim = t.tsit5_claim;
    c->graph = t.graph != 0;
}

int rows_for(int dim, int phase) {
    return dim == 3 ? (phase == ETRAJ_PHASE_CTMC ? 8 : 9) : (phase == ETRAJ_PHASE_CTMC ? 6 : 7);
}

int check_laser(etraj_ctx* c, const etraj_laser_desc* d, LaserP& L) {
    if (!d) return fail(c, ETRAJ_EINVAL, "laser descriptor is NULL");
    if (make_laser(*d, L)) return fail(c, ETRAJ_EINVAL, "invalid laser descriptor (kind, wave_len>0, cyc_num>0)");
    return ETRAJ_OK;
}

int check_target(etraj_ctx* c, const etraj_target_desc* d, TargetP& T) {
    if (!d) return fail(c, ETRAJ_EINVAL, "target descriptor is NULL");
    if (make_target(*d, T)) return fail(c, ETRAJ_EINVAL, "invalid target descriptor (kind, Ip>0, soft_core>=0)");
    return ETRAJ_OK;
}

int64_t batch_max_size_of(const etraj_sampler_desc& s) {
    if (s.sampling == ETRAJ_SAMPLING_STEP)
        return (int64_t)s.ss_kd_num * (s.dimension == 3 ? (int64_t)s.ss_kz_num : 1);
    return s.mc_kt_num;
}

// sp.t_samples (ADKSampler.jl:133-147): range, sorted seeded uniforms, or
// one seeded jitter per stratum.
int batch_times_of(etraj_ctx* c, const etraj_sampler_desc& s, std::vector<double>& t) {
    const int n = s.sample_t_num;
    if (n <= 0) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid time sample number");
    t.resize(n);
    const double lo = s.sample_t_lo, hi = s.sample_t_hi;
    switch (s.sampling) {
        case ETRAJ_SAMPLING_STEP:
            for (int i = 0; i < n; ++i) t[i] = grid_point(lo, hi, n, i);
            break;
        case ETRAJ_SAMPLING_MONTECARLO:
            if (!(lo < hi)) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid sampling time interval");
            for (int i = 0; i < n; ++i) {
                U64x4 o = philox((uint64_t)i, 0, 0, kTagT0, s.seed, 0);
                t[i] = unit_co(o.v[0]) * (hi - lo) + lo;
            }
            std::sort(t.begin(), t.end());
            break;
        case ETRAJ_SAMPLING_STRATIFIED: {
            if (!(lo < hi)) return fail(c, ETRAJ_EINVAL, "invalid sampling time interval");
            const double w = (hi - lo) / (double)n;
            for (int i = 0; i < n; ++i) {
                U64x4 o = philox((uint64_t)i, 0, 0, kTagT0, s.seed, 0);
                t[i] = lo + ((double)i + unit_co(o.v[0])) * w;
            }
            break;
        }
        default:
            return fail(c, ETRAJ_EINVAL, "unknown sampling mode");
    }
    return ETRAJ_OK;
}

// ADKSampler constructor checks (ADKSampler.jl:43-105) and the per-job
// constants of gen_electron_batch.
// n-point Gauss-Legendre rule on [-1, 1]: Newton iteration on P_n from
// the Chebyshev-like initial guesses; (node, weight) pairs.
static void gauss_legendre(int n, double* q) {
    for (int i = 0; i < n / 2; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        const double w = 2 / ((1 - z * z) * dp * dp);
        q[2 * i] = -z;
        q[2 * i + 1] = w;
        q[2 * (n - 1 - i)] = z;
        q[2 * (n - 1 - i) + 1] = w;
    }
}

// The ADK / SPA / SPANE prefactor beyond l = 0 (ADKSampler.jl:188-265): a
// molecule's C_lm table (asymp_coeff, in the frame of mol_rot), or an atom with
// l > 0 as the table with C at (l, m) in the frame (phi, theta, 0) of its
// quantization axis; an atom with l = 0 keeps the scalar C Y_00 (pre_L = -1).
// C receives the host copy of the table ((re, im), [l + (L+1)(m+l)]).
static int make_prefactor(etraj_ctx* c, const etraj_sampler_desc& s, const etraj_target_desc& td, SamplerP& S,
                          std::vector<double>& C) {
    S.pre_L = -1;
    S.pre_C = nullptr;
    C.clear();
    if (s.init_method == ETRAJ_INIT_WFAT) return ETRAJ_OK;
    if (s.asymp_coeff) {
        if (s.asymp_lmax < 0 || s.asymp_lmax > kPreMaxL)
            return fail(c, ETRAJ_EINVAL, "asymp_lmax must be in 0..15");
        S.pre_L = s.asymp_lmax;
        C.assign(s.asymp_coeff, s.asymp_coeff + pre_doubles(S));
        for (double v : C)
            if (!std::isfinite(v)) return fail(c, ETRAJ_EINVAL, "asymp_coeff is not finite");
        for (int k = 0; k < 3; ++k) S.pre_rot[k] = s.mol_rot[k];
    } else if (td.l != 0) {
        if (td.l < 0 || td.l > kPreMaxL || td.m < -td.l || td.m > td.l)
            return fail(c, ETRAJ_EINVAL, "the target's (l, m) must satisfy 0 <= l <= 15, |m| <= l");
        S.pre_L = td.l;
        C.assign(pre_doubles(S), 0.0);
        C[2 * (td.l + (size_t)(td.l + 1) * (td.m + td.l))] = td.asymp_coeff;
        S.pre_rot[0] = td.quan_ax_phi;                                 // ADKSampler.jl:201-203
        S.pre_rot[1] = td.quan_ax_theta;
        S.pre_rot[2] = 0.0;
    }
    return ETRAJ_OK;
}

int make_sampler(etraj_ctx* c, const etraj_sampler_desc* sd, const LaserP& L, const TargetP& T,
                 const std::vector<double>& t, SamplerP& S) {
    if (!sd) return fail(c, ETRAJ_EINVAL, "sampler descriptor is NULL");
    const etraj_sampler_desc& s = *sd;
    if (s.init_method < ETRAJ_INIT_ADK || s.init_method > ETRAJ_INIT_WFAT)
        return fail(c, ETRAJ_EUNSUPPORTED, "init_cond_method: the engine implements :ADK, :SPA, :SPANE and :WFAT");
    if (s.init_method == ETRAJ_INIT_WFAT) {
        if (s.wfat_na < 1 || s.wfat_nb < 1 || !s.wfat_coeff)
            return fail(c, ETRAJ_EINVAL, "[WFATSampler] the structure-factor table (wfat_na, wfat_nb, wfat_coeff) is missing");
        if (s.wfat_form != ETRAJ_WFAT_TABLE && s.wfat_form != ETRAJ_WFAT_INTDATA)
            return fail(c, ETRAJ_EINVAL, "[WFATSampler] undefined structure-factor form (wfat_form)");
        if (s.wfat_form == ETRAJ_WFAT_INTDATA) {
            const int32_t* d = s.wfat_dims;
            if (d[0] < 1 || d[1] < 1 || (d[1] & 1) == 0 || d[2] < 1 || d[3] != 2 * d[2] - 1 || d[0] > kWfatMaxDim ||
                d[1] > kWfatMaxDim || d[2] > kWfatMaxDim || d[0] * d[1] > kWfatMaxChannels)
                return fail(c, ETRAJ_EINVAL, "[WFATSampler] wfat_dims must be (nxi_max+1, 2 m_max+1, l_max+1, 2 l_max+1), "
                                             "each <= 31, with at most 512 channels");
            if (!(std::isfinite(s.wfat_mu[0]) && std::isfinite(s.wfat_mu[1]) && std::isfinite(s.wfat_mu[2])))
                return fail(c, ETRAJ_EINVAL, "[WFATSampler] the orbital dipole wfat_mu is not finite");
        }
    }
    if (s.phase_method < 0 || s.phase_method > 2) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Undefined phase method");
    if (s.dimension != 2 && s.dimension != 3) return fail(c, ETRAJ_EINVAL, "`dimension` must be either 2 or 3");
    if (s.rate_prefix & ~(ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC | ETRAJ_PREFIX_JAC))
        return fail(c, ETRAJ_EINVAL, "[ADKSampler] Undefined tunneling rate prefix");
    if ((s.rate_prefix & ETRAJ_PREFIX_PRE) && (s.rate_prefix & ETRAJ_PREFIX_PRECC))
        return fail(c, ETRAJ_EINVAL, "[ADKSampler] Rate prefixes [Pre] & [PreCC] conflict.");
    if (!(s.cutoff_limit >= 0)) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid cut-off limit");
    if (s.sampling == ETRAJ_SAMPLING_STEP) {
        if (s.dimension == 3) {
            if (!(s.ss_kd_num > 0 && s.ss_kz_num > 0)) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid kd,kz sample number");
            if (!(s.ss_kd_max > 0 || s.ss_kz_max > 0)) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid kd,kz sample boundaries");
        } else {
            if (!(s.ss_kd_num > 0)) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid kd sample number");
            if (!(s.ss_kd_max > 0)) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid kd sample boundaries");
        }
    } else if (s.sampling == ETRAJ_SAMPLING_MONTECARLO) {
        if (!(s.mc_kt_num > 0)) return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid sampling kt_num");
        if (s.dimension == 3 ? !(s.mc_kd_max > 0 || s.mc_kz_max > 0) : !(s.mc_kd_max > 0))
            return fail(c, ETRAJ_EINVAL, "[ADKSampler] Invalid kd,kz sample boundaries");
    } else if (s.sampling == ETRAJ_SAMPLING_STRATIFIED) {
        if (!(s.mc_kt_num > 0)) return fail(c, ETRAJ_EINVAL, "invalid samples per stratum");
        if (!(s.kd_sigma > 0) || (s.dimension == 3 && !(s.kz_sigma > 0)))
            return fail(c, ETRAJ_EINVAL, "stratified sampling needs kd_sigma > 0 (and kz_sigma > 0 in 3D)");
    } else {
        return fail(c, ETRAJ_EINVAL, "unknown sampling mode");
    }
    if (s.sampling == ETRAJ_SAMPLING_STEP && batch_max_size_of(s) > (int64_t)1 << 31)
        return fail(c, ETRAJ_EINVAL, "batch too large");
    S = SamplerP{};
    S.dim = s.dimension;
    S.phase = s.phase_method;
    S.sampling = s.sampling;
    int pre = s.rate_prefix;
    if ((pre & ETRAJ_PREFIX_PRECC) && L.n != 1) {  // ADKSampler.jl:72-77
        pre &= ~ETRAJ_PREFIX_PRECC;
        pre |= ETRAJ_PREFIX_PRE;
    }
    S.prefix = pre;
    S.t_num = s.sample_t_num;
    S.kd_num = s.sampling == ETRAJ_SAMPLING_STEP ? s.ss_kd_num : 0;
    S.kz_num = (s.sampling == ETRAJ_SAMPLING_STEP && s.dimension == 3) ? s.ss_kz_num : 1;
    S.rows = rows_for(s.dimension, s.phase_method);
    S.nk = batch_max_size_of(s);
    S.max_bs = S.nk;
    S.t_lo = s.sample_t_lo;
    S.t_hi = s.sample_t_hi;
    S.kd_max = std::fabs(s.ss_kd_max);
    S.kz_max = std::fabs(s.ss_kz_max);
    S.mc_kd_max = s.mc_kd_max;
    S.mc_kz_max = s.mc_kz_max;
    S.kd_sigma = s.kd_sigma;
    S.kz_sigma = s.kz_sigma;
    S.cutoff = s.cutoff_limit;
    S.seed = s.seed;
    const double tmax = *std::max_element(t.begin(), t.end()), tmin = *std::min_element(t.begin(), t.end());
    const double tstep = (tmax - tmin) / (double)t.size();      // step(range) of ADKSampler.jl:267
    if (s.sampling == ETRAJ_SAMPLING_STEP) {
        double kd_step = s.ss_kd_num > 1 ? (2 * std::fabs(s.ss_kd_max)) / s.ss_kd_num : 1.0;
        double kz_step = (s.dimension == 3 && s.ss_kz_num > 1) ? (2 * std::fabs(s.ss_kz_max)) / s.ss_kz_num : 1.0;
        S.dkdt_base = tstep * kd_step * kz_step;
    } else if (s.sampling == ETRAJ_SAMPLING_MONTECARLO) {
        double kd_span = s.mc_kd_max > 0 ? 2 * s.mc_kd_max : 1.0;
        double kz_span = s.mc_kz_max > 0 ? 2 * s.mc_kz_max : 1.0;  // no dimension check (:274)
        S.dkdt_base = tstep * kd_span * kz_span / s.mc_kt_num;
    } else {
        S.dkdt_base = (s.sample_t_hi - s.sample_t_lo) / s.sample_t_num / s.mc_kt_num;
    }
    // the sampler's Ip: IonPotential(mol, ridx) of a molecule (ADKSampler.jl:177-182, WFATSampler.jl:139)
    S.Ip = s.orbit_ip > 0 ? s.orbit_ip : T.Ip;
    const double kappa = std::sqrt(2 * S.Ip), nn = T.Zd / kappa;
    S.nstar = nn;
    S.c_pre = std::pow(2.0, nn / 2 + 1) * std::pow(kappa, 2 * nn + 0.5) * std::tgamma(nn / 2 + 1);
    S.numer = T.C * 1.0 * std::sqrt(1.0 / (4 * kPi));
    S.method = s.init_method;
    S.pre_L = -1;
    S.pre_C = nullptr;
    S.omega = L.n == 1 ? L.c[0].w : (L.c[0].w + L.c[1].w) / 2;   // SPASampler.jl:160-166
    S.glq = c->glq.p;
    if (s.init_method == ETRAJ_INIT_WFAT) {   // the table is uploaded by the caller (S.wfat_coeff)
        S.wfat_na = s.wfat_na;
        S.wfat_nb = s.wfat_nb;
        S.wfat_form = s.wfat_form;
        for (int k = 0; k < 4; ++k) S.wfat_dims[k] = s.wfat_dims[k];
        for (int k = 0; k < 3; ++k) {
            S.mol_rot[k] = s.mol_rot[k];
            S.wfat_mu[k] = s.wfat_mu[k];
        }
    }
    return ETRAJ_OK;
}

int make_integ(etraj_ctx* c, const etraj_run_desc* rd, const LaserP& L, const TargetP& T, IntegP& P) {
    if (!rd) return fail(c, ETRAJ_EINVAL, "run descriptor is NULL");
    const etraj_run_desc& r = *rd;
    if (r.dimension != 2 && r.dimension != 3)
        return fail(c, ETRAJ_EINVAL, "[TrajectorySimulationJob] `dimension` must be either 2 or 3.");
    if (r.phase_method < 0 || r.phase_method > 2) return fail(c, ETRAJ_EINVAL, "undefined phase method");
    if (r.integrator != ETRAJ_INTEG_RK4 && r.integrator != ETRAJ_INTEG_TSIT5)
        return fail(c, ETRAJ_EINVAL, "unknown integrator");
    if (r.integrator == ETRAJ_INTEG_RK4 && !(r.dt > 0)) return fail(c, ETRAJ_EINVAL, "RK4 needs dt > 0");
    if (r.integrator == ETRAJ_INTEG_TSIT5) {
        if (!(r.rtol > 0)) return fail(c, ETRAJ_EINVAL, "Tsit5 needs traj_rtol > 0");
    }
    if (r.precision != ETRAJ_PREC_FP64 && r.precision != ETRAJ_PREC_MIXED) return fail(c, ETRAJ_EINVAL, "unknown precision");
    if (r.precision == ETRAJ_PREC_MIXED && !(r.restep_radius >= 0)) return fail(c, ETRAJ_EINVAL, "restep_radius must be >= 0");
    if (r.pmd_mode != ETRAJ_PMD_FLOAT && r.pmd_mode != ETRAJ_PMD_FIXED) return fail(c, ETRAJ_EINVAL, "unknown pmd_mode");
    for (int d = 0; d < r.dimension; ++d)
        if (r.final_p_num[d] < 1 || !(r.final_p_max[d] > 0))
            return fail(c, ETRAJ_EINVAL, "final_p_num must be >= 1 and final_p_max > 0 in every dimension");
    int64_t npix = 1;
    for (int d = 0; d < r.dimension; ++d) npix *= r.final_p_num[d];
    if (npix > ((int64_t)1 << 34)) return fail(c, ETRAJ_EINVAL, "momentum grid too large");
    P = IntegP{};
    P.L = L;
    P.T = T;
    P.dim = r.dimension;
    P.phase = r.phase_method;
    P.mixed = r.precision == ETRAJ_PREC_MIXED;
    P.quirks = r.quirks;
    P.rows = rows_for(r.dimension, r.phase_method);
    for (int d = 0; d < 3; ++d) {
        P.pnum[d] = d < r.dimension ? r.final_p_num[d] : 1;
        P.pmax[d] = d < r.dimension ? r.final_p_max[d] : 1.0;
    }
    P.tf = r.traj_t_final;
    P.dt = r.dt;
    P.restep_radius = r.restep_radius;
    P.integrator = r.integrator;
    P.rtol = r.rtol;
    P.atol = r.atol > 0 ? r.atol : 1e-6;          // OrdinaryDiffEq's default abstol
    P.maxiters = r.maxiters > 0 ? r.maxiters : 1000000;
    P.rk4_block = c->rk4_block;
    P.ts5_refill = c->ts5_refill;
    P.ts5_claim = c->ts5_claim;
    return ETRAJ_OK;
}

// RK4 steps of a segment starting at t0 (identical expression on the device)
int64_t steps_of(double tf, double t0, double dt) {
    int64_t n = (int64_t)std::ceil((tf - t0) / dt);
    return n < 1 ? 1 : n;
}

constexpr int64_t kFieldTabBytesPerStep = sizeof(float4) + 2 * sizeof(double2);
// Bytes of field table per launch and init columns per slab. Slabs this large
// (per buffer slot: 4.8 GB of init columns at 9 rows, 8 GiB of field table)
// keep the FP64 re-integration pass of one slab wide enough to fill the
// device even when only a few per cent of the electrons are flagged.
constexpr int64_t kFieldTabBudget = (int64_t)8 << 30;
constexpr int64_t kMaxSlabCols = (int64_t)1 << 26;
// Tsit5 slabs (one job each) take up to 2^28 columns, split evenly, or the
// whole job up to 2^29 (two slots of half its size would hold as much): the
// work queue of each integrator launch ends in a tail of about 1.5 ms in which
// the slowest electrons run on a few lanes, so fewer, larger slabs save their
// tails (1e8-electron job: 6 slabs 159.3 ms, 2 slabs 152.6 ms). 2^28 columns
// of a 3D job are 39 GB of columns and start values per slot.
constexpr int64_t kMaxTs5SlabCols = (int64_t)1 << 28;

// Device buffers of integrate_columns: the ctx's own (grown on demand) or
// fixed slices of the no-allocation scratch array (`fixed`, with the field
// table's capacity in RK4 steps).
struct ColBufs {
    bool fixed = false;
    double* init = nullptr;
    int64_t* seg_start = nullptr;
    int32_t* counts = nullptr;
    int32_t* counts2 = nullptr;
    double* seg_t0 = nullptr;
    int64_t* pre = nullptr;
    uint8_t* flags = nullptr;
    int64_t* orig = nullptr;
    unsigned long long* queue = nullptr;
    LaserP* lps = nullptr;
    float4* ft32 = nullptr;
    double2* ft64 = nullptr;
    double2* ft0 = nullptr;
    int64_t ft_steps = 0;   // fixed: field-table capacity (segments x stride), in steps
    int64_t ft_segs = 0;    // fixed: capacity of ft0, in segments
};

// Grows the ctx's buffers of a call over n columns and nseg segments (a no-op
// for fixed buffers, whose sizes the caller has checked).
int col_bufs_ensure(etraj_ctx* c, ColBufs& B, int64_t n, int64_t nseg, int rows) {
    if (B.fixed) return ETRAJ_OK;
    HIP_TRY(c, c->init.ensure((size_t)n * rows));
    HIP_TRY(c, c->seg_start.ensure(nseg));
    HIP_TRY(c, c->counts.ensure(nseg));
    HIP_TRY(c, c->counts2.ensure(nseg));
    HIP_TRY(c, c->seg_t0.ensure(nseg));
    HIP_TRY(c, c->pre.ensure(nseg + 1));
    HIP_TRY(c, c->flags.ensure((size_t)n));
    HIP_TRY(c, c->orig.ensure((size_t)n));
    HIP_TRY(c, c->lps.ensure(1));
    HIP_TRY(c, c->queue.ensure(1));
    B.queue = c->queue.p;
    B.init = c->init.p; B.seg_start = c->seg_start.p; B.counts = c->counts.p; B.counts2 = c->counts2.p;
    B.seg_t0 = c->seg_t0.p; B.pre = c->pre.p; B.flags = c->flags.p; B.orig = c->orig.p; B.lps = c->lps.p;
    return ETRAJ_OK;
}

// Fills the field table of nseg segments whose t0 are in t0s_dev and whose
// lasers are lps_dev[seg_job_dev[s]] (seg_job_dev NULL: lps_dev[0]); the ctx's
// table buffers are grown to fit unless B is fixed.
int fill_field_table(etraj_ctx* c, ColBufs& B, const IntegP& P, const LaserP* lps_dev, const int32_t* seg_job_dev,
                     const double* t0s_dev, int64_t nseg, int64_t stride, FieldTab& FT) {
    if (!B.fixed) {
        HIP_TRY(c, c->ft32.ensure((size_t)nseg * stride));
        HIP_TRY(c, c->ft64.ensure((size_t)2 * nseg * stride));
        HIP_TRY(c, c->ft0.ensure((size_t)nseg));
        B.ft32 = c->ft32.p;
        B.ft64 = c->ft64.p;
        B.ft0 = c->ft0.p;
    } else if (nseg * stride > B.ft_steps || nseg > B.ft_segs) {
        return fail(c, ETRAJ_ECAPACITY, "field table of the segment group exceeds the no-allocation scratch");
    }
    FT.f32 = B.ft32;
    FT.f64 = B.ft64;
    FT.f0 = B.ft0;
    FT.stride = stride;
    HIP_TRY(c, launch_field_table(c->stream, lps_dev, seg_job_dev, t0s_dev, nseg, P.tf, P.dt, FT));
    return ETRAJ_OK;
}

// Fixed-step RK4 over nseg segments. FP64: one pass. MIXED: the FP32 pass
// flags the electrons that come within restep_radius of the core, which are
// compacted in place per segment and integrated again in FP64 (their outputs
// at their own indices through B.orig unless out is OUT_PMD).
int rk4_passes(etraj_ctx* c, const ColBufs& B, const IntegP& P, const OutP& O, int out, const int64_t* seg_start,
               int32_t* counts, int32_t* counts2, int64_t nseg, int cps, const FieldTab& FT) {
    if (!P.mixed) {
        HIP_TRY(c, launch_rk4(c->stream, P, O, out, B.init, seg_start, counts, nseg, cps, FT, nullptr, nullptr));
        return ETRAJ_OK;
    }
    const bool need_orig = out != OUT_PMD;
    HIP_TRY(c, launch_rk4(c->stream, P, O, out, B.init, seg_start, counts, nseg, cps, FT, B.flags, nullptr));
    HIP_TRY(c, launch_compact_flagged(c->stream, B.init, P.rows, seg_start, counts, nseg, B.flags, nullptr, counts2,
                                      need_orig ? B.orig : nullptr));
    IntegP P64 = P;
    P64.mixed = 0;
    HIP_TRY(c, launch_rk4(c->stream, P64, O, out, B.init, seg_start, counts2, nseg, cps, FT, nullptr,
                          need_orig ? B.orig : nullptr));
    return ETRAJ_OK;
}

int64_t npix_of(const IntegP& P) { return (int64_t)P.pnum[0] * P.pnum[1] * (P.dim == 3 ? P.pnum[2] : 1); }

// Segments of equal t0 in a Julia init matrix (columns of one
// gen_electron_batch share tr), so that each wavefront shares its field.
// start / count receive the segments (capacity >= n); returns their number.
int64_t build_segments(const double* init, int64_t n, int rows, int dim, int64_t* start, int32_t* count,
                       int64_t& maxlen) {
    int64_t nseg = 0;
    maxlen = 0;
    int64_t i = 0;
    while (i < n) {
        int64_t j = i + 1;
        const double t0 = init[i * rows + 2 * dim];
        while (j < n && init[j * rows + 2 * dim] == t0 && j - i < ((int64_t)1 << 30)) ++j;
        start[nseg] = i;
        count[nseg] = (int32_t)(j - i);
        ++nseg;
        maxlen = std::max(maxlen, j - i);
        i = j;
    }
    return nseg;
}

int check_init(etraj_ctx* c, const IntegP& P, const double* init, int64_t n, int32_t rows) {
    if (n < 0) return fail(c, ETRAJ_EINVAL, "n must be >= 0");
    if (n > 0 && !init) return fail(c, ETRAJ_EINVAL, "init is NULL");
    if (rows != P.rows)
        return fail(c, ETRAJ_EINVAL, "init has " + std::to_string(rows) + " rows, the dimension/phase method needs " +
                                         std::to_string(P.rows));
    for (int64_t i = 0; i < n; ++i)
        if (!(init[i * rows + 2 * P.dim] < P.tf))
            return fail(c, ETRAJ_EINVAL, "[TrajectorySimulationJob] `traj_t_final` must be greater than every t0");
    return ETRAJ_OK;
}

// Integrates n columns; out = OUT_RECORDS or OUT_CONTRIB. The host staging
// arrays are the ctx's (h_ss, h_sc, h_pre, h_t0: capacity >= n + 1 on entry).
// RK4 segments are processed in groups whose field tables fit the budget
// (kFieldTabBudget, or the fixed buffers' capacity).
int integrate_columns(etraj_ctx* c, ColBufs& B, const IntegP& P, const double* init, int64_t n, int out, OutP& O) {
    int64_t* ss = c->h_ss.data();
    int32_t* sc = c->h_sc.data();
    int64_t maxlen = 0;
    const int64_t nseg = build_segments(init, n, P.rows, P.dim, ss, sc, maxlen);
    if (int rc = col_bufs_ensure(c, B, n, nseg, P.rows)) return rc;
    HIP_TRY(c, hipMemcpyAsync(B.init, init, sizeof(double) * n * P.rows, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.seg_start, ss, sizeof(int64_t) * nseg, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.counts, sc, sizeof(int32_t) * nseg, hipMemcpyHostToDevice, c->stream));
    if (P.integrator == ETRAJ_INTEG_TSIT5) {
        int64_t* pre = c->h_pre.data();
        pre[0] = 0;
        for (int64_t k = 0; k < nseg; ++k) pre[k + 1] = pre[k] + sc[k];
        HIP_TRY(c, hipMemcpyAsync(B.pre, pre, sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice, c->stream));
        if (!P.mixed) {
            HIP_TRY(c, launch_tsit5(c->stream, P, O, out, B.init, B.seg_start, B.pre, nseg, n, B.queue, nullptr, nullptr));
            return ETRAJ_OK;
        }
        // mixed: FP32 stage derivatives; the flagged electrons are compacted per
        // segment and integrated again in FP64, their outputs at their own columns
        HIP_TRY(c, launch_tsit5(c->stream, P, O, out, B.init, B.seg_start, B.pre, nseg, n, B.queue, B.flags, nullptr));
        HIP_TRY(c, launch_compact_flagged(c->stream, B.init, P.rows, B.seg_start, B.counts, nseg, B.flags, nullptr,
                                          B.counts2, B.orig));
        HIP_TRY(c, launch_scan_counts(c->stream, B.counts2, nseg, B.pre));
        IntegP P64 = P;
        P64.mixed = 0;
        HIP_TRY(c, launch_tsit5(c->stream, P64, O, out, B.init, B.seg_start, B.pre, nseg, n, B.queue, nullptr, B.orig));
        return ETRAJ_OK;
    }
    double* t0 = c->h_t0.data();
    int64_t stride = 1;
    for (int64_t k = 0; k < nseg; ++k) {
        t0[k] = init[ss[k] * P.rows + 2 * P.dim];
        stride = std::max(stride, steps_of(P.tf, t0[k], P.dt));
    }
    const int64_t budget = B.fixed ? B.ft_steps : kFieldTabBudget / kFieldTabBytesPerStep;
    int64_t group = std::min<int64_t>(nseg, budget / stride);
    if (B.fixed) group = std::min(group, B.ft_segs);
    if (group < 1) {
        if (B.fixed)
            return fail(c, ETRAJ_ECAPACITY, "one segment's field table (" + std::to_string(stride) +
                                                 " steps) exceeds the no-allocation scratch");
        group = 1;
    }
    HIP_TRY(c, hipMemcpyAsync(B.seg_t0, t0, sizeof(double) * nseg, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(B.lps, &P.L, sizeof(LaserP), hipMemcpyHostToDevice, c->stream));
    for (int64_t g0 = 0; g0 < nseg; g0 += group) {
        const int64_t ng = std::min(group, nseg - g0);
        int32_t maxc = 0;
        for (int64_t k = g0; k < g0 + ng; ++k) maxc = std::max(maxc, sc[k]);
        FieldTab FT{};
        if (int rc = fill_field_table(c, B, P, B.lps, nullptr, B.seg_t0 + g0, ng, stride, FT)) return rc;
        if (int rc = rk4_passes(c, B, P, O, out, B.seg_start + g0, B.counts + g0, B.counts2 + g0, ng,
                                (int)((maxc + 63) / 64), FT))
            return rc;
    }
    return ETRAJ_OK;
}

// Grows the ctx's host staging of integrate_columns to n columns (no-op within capacity).
void host_staging_reserve(etraj_ctx* c, int64_t n) {
    if ((int64_t)c->h_ss.size() < n + 1) {
        c->h_ss.resize(n + 1);
        c->h_sc.resize(n + 1);
        c->h_pre.resize(n + 1);
        c->h_t0.resize(n + 1);
    }
}

}  // namespace

bool create_slot_events(etraj_ctx* c) {
    for (SlabSlot& Q : c->slot)
        if (hipEventCreateWithFlags(&Q.done, hipEventDisableTiming) != hipSuccess) return false;
    return true;
}

extern "C" {

int etraj_abi_version(void) { return ETRAJ_ABI_VERSION; }

int etraj_ctx_create(int device, etraj_ctx** out) {
    if (!out) return ETRAJ_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ETRAJ_EHIP;
    if (device < 0 || device >= ndev) return ETRAJ_EINVAL;
    etraj_ctx* c = new (std::nothrow) etraj_ctx();
    if (!c) return ETRAJ_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ETRAJ_EHIP;
    }
    double glq[2 * kSpaGlPoints];
    gauss_legendre(kSpaGlPoints, glq);
    if (c->glq.ensure(2 * kSpaGlPoints) != hipSuccess || hipMemcpy(c->glq.p, glq, sizeof glq, hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        !create_slot_events(c)) {
        (void)hipStreamDestroy(c->stream);
        if (c->stream2) (void)hipStreamDestroy(c->stream2);
        c->glq.release();
        for (SlabSlot& Q : c->slot) Q.release();
        delete c;
        return ETRAJ_EHIP;
    }
    // the no-allocation entry point's scratch (this device's copy of the static
    // array) and host staging, set up once here so that its calls allocate nothing
    if (noalloc_scratch(device, &c->scratch) != hipSuccess) {
        etraj_ctx_destroy(c);
        return ETRAJ_EHIP;
    }
    host_staging_reserve(c, kNoAllocMaxCols);
    c->h_contrib.resize(kNoAllocMaxCols);
    *out = c;
    return ETRAJ_OK;
}

void etraj_ctx_destroy(etraj_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    c->init.release(); c->pmd.release(); c->pmd_out.release(); c->io1.release(); c->io2.release();
    c->io3.release(); c->io4.release(); c->scal.release(); c->t0s.release(); c->seg_start.release();
    c->cnt.release(); c->counts.release(); c->status.release(); c->contrib.release();
    c->pmdq.release(); c->scalq.release(); c->hits.release();
    c->seg_t0.release(); c->ft32.release(); c->ft64.release(); c->ft0.release(); c->pre.release();
    c->counts2.release(); c->flags.release(); c->orig.release(); c->wfat.release(); c->prec.release();
    c->queue.release();
    c->seg_job.release(); c->seg_b.release(); c->sps.release(); c->lps.release(); c->glq.release();
    c->bin_cell.release(); c->bin_scell.release(); c->tile_count.release(); c->tile_off.release();
    c->chunk_off.release(); c->cursor.release(); c->bin_amp.release(); c->bin_samp.release();
    for (SlabSlot& Q : c->slot) Q.release();
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* etraj_last_error(const etraj_ctx* c) { return c ? c->err.c_str() : "null context"; }

int etraj_laser_field(etraj_ctx* c, const etraj_laser_desc* laser, const double* t, int64_t n, double* Fx,
                      double* Fy, double* Ax, double* Ay) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    LaserP L;
    if (int rc = check_laser(c, laser, L)) return rc;
    if (n < 0 || (n > 0 && !t)) return fail(c, ETRAJ_EINVAL, "bad time array");
    if (n == 0) return ETRAJ_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->io1.ensure(n)); HIP_TRY(c, c->io2.ensure(n)); HIP_TRY(c, c->io3.ensure(n));
    HIP_TRY(c, c->io4.ensure(n)); HIP_TRY(c, c->t0s.ensure(n));
    HIP_TRY(c, hipMemcpyAsync(c->t0s.p, t, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_laser_field(c->stream, L, c->t0s.p, n, c->io1.p, c->io2.p, c->io3.p, c->io4.p));
    double* outs[4] = {Fx, Fy, Ax, Ay};
    double* devs[4] = {c->io1.p, c->io2.p, c->io3.p, c->io4.p};
    for (int k = 0; k < 4; ++k)
        if (outs[k]) HIP_TRY(c, hipMemcpyAsync(outs[k], devs[k], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ETRAJ_OK;
}

int etraj_target_eval(etraj_ctx* c, const etraj_target_desc* target, const double* xyz, int64_t n, double* V,
                      double* F) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    TargetP T;
    if (int rc = check_target(c, target, T)) return rc;
    if (n < 0 || (n > 0 && !xyz)) return fail(c, ETRAJ_EINVAL, "bad point array");
    if (n == 0) return ETRAJ_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->io1.ensure(3 * n)); HIP_TRY(c, c->io2.ensure(n)); HIP_TRY(c, c->io3.ensure(3 * n));
    HIP_TRY(c, hipMemcpyAsync(c->io1.p, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_target_eval(c->stream, T, c->io1.p, n, c->io2.p, c->io3.p));
    if (V) HIP_TRY(c, hipMemcpyAsync(V, c->io2.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (F) HIP_TRY(c, hipMemcpyAsync(F, c->io3.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ETRAJ_OK;
}

int etraj_batch_times(etraj_ctx* c, const etraj_sampler_desc* s, double* t0) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    if (!s || !t0) return fail(c, ETRAJ_EINVAL, "NULL argument");
    std::vector<double> t;
    if (int rc = batch_times_of(c, *s, t)) return rc;
    std::memcpy(t0, t.data(), sizeof(double) * t.size());
    return ETRAJ_OK;
}

int64_t etraj_batch_max_size(const etraj_sampler_desc* s) { return s ? batch_max_size_of(*s) : -1; }

int etraj_gen_electron_batch(etraj_ctx* c, const etraj_laser_desc* laser, const etraj_target_desc* target,
                             const etraj_sampler_desc* sampler, int32_t batch_id, double* out, int64_t capacity,
                             int64_t* count) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    LaserP L;
    TargetP T;
    SamplerP S;
    if (int rc = check_laser(c, laser, L)) return rc;
    if (int rc = check_target(c, target, T)) return rc;
    if (!sampler) return fail(c, ETRAJ_EINVAL, "sampler descriptor is NULL");
    std::vector<double> t, preC;
    if (int rc = batch_times_of(c, *sampler, t)) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    if (int rc = make_sampler(c, sampler, L, T, t, S)) return rc;
    if (int rc = make_prefactor(c, *sampler, *target, S, preC)) return rc;
    if (batch_id < 1 || batch_id > S.t_num) return fail(c, ETRAJ_EINVAL, "batch_id out of range");
    if (!count || (capacity > 0 && !out)) return fail(c, ETRAJ_EINVAL, "NULL output");
    HIP_TRY(c, hipSetDevice(c->device));
    if (S.method == ETRAJ_INIT_WFAT) {
        const size_t nc = wfat_doubles(S);
        HIP_TRY(c, c->wfat.ensure(nc));
        HIP_TRY(c, hipMemcpyAsync(c->wfat.p, sampler->wfat_coeff, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream));
        S.wfat_coeff = c->wfat.p;
    }
    if (!preC.empty()) {
        HIP_TRY(c, c->prec.ensure(preC.size()));
        HIP_TRY(c, hipMemcpyAsync(c->prec.p, preC.data(), sizeof(double) * preC.size(), hipMemcpyHostToDevice,
                                  c->stream));
        S.pre_C = c->prec.p;
    }
    HIP_TRY(c, c->init.ensure((size_t)S.max_bs * S.rows));
    HIP_TRY(c, c->counts.ensure(1));
    HIP_TRY(c, c->sps.ensure(1));
    HIP_TRY(c, c->lps.ensure(1));
    HIP_TRY(c, c->seg_t0.ensure(1));
    HIP_TRY(c, c->seg_job.ensure(1));
    HIP_TRY(c, c->seg_b.ensure(1));
    HIP_TRY(c, c->seg_start.ensure(1));
    const double t0 = t[batch_id - 1];
    const int32_t job0 = 0, b = batch_id - 1;
    const int64_t start0 = 0;
    HIP_TRY(c, hipMemcpyAsync(c->sps.p, &S, sizeof(SamplerP), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->lps.p, &L, sizeof(LaserP), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->seg_t0.p, &t0, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->seg_job.p, &job0, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->seg_b.p, &b, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->seg_start.p, &start0, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    const SegTab G{c->seg_t0.p, c->seg_job.p, c->seg_b.p, c->seg_start.p};
    HIP_TRY(c, launch_adk_sample(c->stream, S.method, S.pre_L >= 0, c->sps.p, c->lps.p, T, G, 1, c->init.p, c->counts.p, nullptr));
    int32_t cnt = 0;
    HIP_TRY(c, hipMemcpyAsync(&cnt, c->counts.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *count = cnt;
    if (cnt > capacity) return fail(c, ETRAJ_ECAPACITY, "output capacity too small for " + std::to_string(cnt) + " columns");
    if (cnt > 0) {
        HIP_TRY(c, hipMemcpyAsync(out, c->init.p, sizeof(double) * cnt * S.rows, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return ETRAJ_OK;
}

int etraj_integrate_records(etraj_ctx* c, const etraj_laser_desc* laser, const etraj_target_desc* target,
                            const etraj_run_desc* run, const double* init, int64_t n, int32_t rows, double* records,
                            int32_t* status) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    LaserP L;
    TargetP T;
    IntegP P;
    if (int rc = check_laser(c, laser, L)) return rc;
    if (int rc = check_target(c, target, T)) return rc;
    if (int rc = make_integ(c, run, L, T, P)) return rc;
    if (int rc = check_init(c, P, init, n, rows)) return rc;
    if (n == 0) return ETRAJ_OK;
    if (!records || !status) return fail(c, ETRAJ_EINVAL, "NULL output");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->io1.ensure((size_t)12 * n));
    HIP_TRY(c, c->status.ensure(n));
    OutP O{};
    O.records = c->io1.p;
    O.status = c->status.p;
    ColBufs B;
    host_staging_reserve(c, n);
    if (int rc = integrate_columns(c, B, P, init, n, OUT_RECORDS, O)) return rc;
    HIP_TRY(c, hipMemcpyAsync(records, c->io1.p, sizeof(double) * 12 * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(status, c->status.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ETRAJ_OK;
}

// Adds the per-electron records h[0..n) into the caller's PMD and counters in
// column order, as the reference's per-electron loop (TrajectorySimulation.jl:344-397).
static void bin_contribs_host(const IntegP& P, const etraj_run_desc* run, const Contrib* h, int64_t n, void* pmd,
                              double* ion_prob, double* ion_prob_uncollected, int64_t* n_failed) {
    double* out = static_cast<double*>(pmd);
    int64_t nf = 0;
    for (int64_t i = 0; i < n; ++i) {
        const Contrib& k = h[i];
        if (k.status == 1) { ++nf; continue; }
        if (k.status == 2) continue;
        *ion_prob += k.prob;
        if (k.status == 3) { *ion_prob_uncollected += k.prob; continue; }
        if (run->hit_counts) run->hit_counts[k.lin] += 1;
        if (P.phase == ETRAJ_PHASE_CTMC) {
            out[k.lin] += k.re;
        } else {
            out[2 * k.lin] += k.re;
            out[2 * k.lin + 1] += k.im;
        }
    }
    *n_failed = nf;
}

int etraj_integrate_and_bin(etraj_ctx* c, const etraj_laser_desc* laser, const etraj_target_desc* target,
                            const etraj_run_desc* run, const double* init, int64_t n, int32_t rows, void* pmd,
                            double* ion_prob, double* ion_prob_uncollected, int64_t* n_failed) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    LaserP L;
    TargetP T;
    IntegP P;
    if (int rc = check_laser(c, laser, L)) return rc;
    if (int rc = check_target(c, target, T)) return rc;
    if (int rc = make_integ(c, run, L, T, P)) return rc;
    if (int rc = check_init(c, P, init, n, rows)) return rc;
    if (!pmd || !ion_prob || !ion_prob_uncollected || !n_failed) return fail(c, ETRAJ_EINVAL, "NULL output");
    *n_failed = 0;
    if (n == 0) return ETRAJ_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->contrib.ensure(n));
    OutP O{};
    O.contrib = c->contrib.p;
    ColBufs B;
    host_staging_reserve(c, n);
    if (int rc = integrate_columns(c, B, P, init, n, OUT_CONTRIB, O)) return rc;
    if ((int64_t)c->h_contrib.size() < n) c->h_contrib.resize(n);
    HIP_TRY(c, hipMemcpyAsync(c->h_contrib.data(), c->contrib.p, sizeof(Contrib) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    bin_contribs_host(P, run, c->h_contrib.data(), n, pmd, ion_prob, ion_prob_uncollected, n_failed);
    return ETRAJ_OK;
}

// etraj_integrate_and_bin with every device buffer a slice of the library's
// static scratch array (etraj_noalloc.hip) and the host staging the ctx's
// (sized at ctx creation): no device or host allocation on the way. Calls on
// one device share that device's array, so calls that overlap on one device
// (two host threads with two ctxs) race; consecutive calls need no clearing,
// since every slice a call reads is first written by that call.
int etraj_integrate_and_bin_noalloc(etraj_ctx* c, const etraj_laser_desc* laser, const etraj_target_desc* target,
                                    const etraj_run_desc* run, const double* init, int64_t n, int32_t rows,
                                    void* pmd, double* ion_prob, double* ion_prob_uncollected, int64_t* n_failed) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    LaserP L;
    TargetP T;
    IntegP P;
    if (int rc = check_laser(c, laser, L)) return rc;
    if (int rc = check_target(c, target, T)) return rc;
    if (int rc = make_integ(c, run, L, T, P)) return rc;
    if (n > kNoAllocMaxCols)
        return fail(c, ETRAJ_ECAPACITY, "etraj_integrate_and_bin_noalloc takes at most 65536 columns per call");
    if (int rc = check_init(c, P, init, n, rows)) return rc;
    if (!pmd || !ion_prob || !ion_prob_uncollected || !n_failed) return fail(c, ETRAJ_EINVAL, "NULL output");
    *n_failed = 0;
    if (n == 0) return ETRAJ_OK;
    if (!c->scratch) return fail(c, ETRAJ_EHIP, "the no-allocation scratch of this device is not available");
    HIP_TRY(c, hipSetDevice(c->device));
    // scratch layout: columns, records, flags, original indices, segment tables, laser, field table
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const int64_t M = kNoAllocMaxCols;
    ColBufs B;
    B.fixed = true;
    B.init = reinterpret_cast<double*>(take(sizeof(double) * 9 * M));
    Contrib* contrib = reinterpret_cast<Contrib*>(take(sizeof(Contrib) * M));
    B.flags = take(M);
    B.orig = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * M));
    B.seg_start = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * M));
    B.counts = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * M));
    B.counts2 = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * M));
    B.seg_t0 = reinterpret_cast<double*>(take(sizeof(double) * M));
    B.pre = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * (M + 1)));
    B.lps = reinterpret_cast<LaserP*>(take(sizeof(LaserP)));
    B.queue = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long)));
    // field table of a segment group: F(t0) per segment, then the rest of the array as
    // (float4 + 2 double2) per step; each slice above is rounded up to 256 bytes
    B.ft_segs = M;
    B.ft0 = reinterpret_cast<double2*>(take(sizeof(double2) * M));
    B.ft_steps = (int64_t)(((size_t)kNoAllocScratchBytes - off - 512) / kFieldTabBytesPerStep);
    B.ft32 = reinterpret_cast<float4*>(take(sizeof(float4) * B.ft_steps));
    B.ft64 = reinterpret_cast<double2*>(take(sizeof(double2) * 2 * B.ft_steps));
    if (off > (size_t)kNoAllocScratchBytes) return fail(c, ETRAJ_EHIP, "no-allocation scratch layout overflow");
    OutP O{};
    O.contrib = contrib;
    if (int rc = integrate_columns(c, B, P, init, n, OUT_CONTRIB, O)) {
        (void)hipStreamSynchronize(c->stream);
        return rc;
    }
    HIP_TRY(c, hipMemcpyAsync(c->h_contrib.data(), contrib, sizeof(Contrib) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    bin_contribs_host(P, run, c->h_contrib.data(), n, pmd, ion_prob, ion_prob_uncollected, n_failed);
    return ETRAJ_OK;
}

// One job of a run or a scan: its laser, its sampler and its batch times.
struct JobPlan {
    LaserP L;
    SamplerP S;
    std::vector<double> t;
    std::vector<double> preC;   // host copy of the prefactor table (make_prefactor)
};

// Validates n_jobs (laser, sampler) pairs that share the target and the run
// descriptor, as TrajectorySimulationJob does for one job
// (TrajectorySimulation.jl:178-186), and uploads the WFAT tables of all jobs
// into one device buffer.
static int plan_jobs(etraj_ctx* c, int32_t n_jobs, const etraj_laser_desc* lasers, const etraj_target_desc* target,
                     const TargetP& T, const etraj_sampler_desc* samplers, const etraj_run_desc* run,
                     std::vector<JobPlan>& jobs) {
    if (n_jobs < 1) return fail(c, ETRAJ_EINVAL, "n_jobs must be positive");
    if (!lasers || !samplers) return fail(c, ETRAJ_EINVAL, "laser / sampler descriptor array is NULL");
    if (!run) return fail(c, ETRAJ_EINVAL, "run descriptor is NULL");
    jobs.resize(n_jobs);
    size_t ncoef = 0, npre = 0;
    for (int32_t j = 0; j < n_jobs; ++j) {
        JobPlan& J = jobs[j];
        const etraj_sampler_desc& sd = samplers[j];
        int rc = check_laser(c, &lasers[j], J.L);
        if (!rc && sd.init_method != samplers[0].init_method)
            rc = fail(c, ETRAJ_EINVAL, "the jobs of a scan must share init_cond_method");
        if (!rc) rc = batch_times_of(c, sd, J.t);
        if (!rc) rc = make_sampler(c, &sd, J.L, T, J.t, J.S);
        if (!rc) rc = make_prefactor(c, sd, *target, J.S, J.preC);
        if (!rc && j > 0 && (J.S.pre_L >= 0) != (jobs[0].S.pre_L >= 0))
            rc = fail(c, ETRAJ_EINVAL, "the jobs of a scan must all carry, or all omit, asymp_coeff");
        if (!rc && (sd.dimension != run->dimension || sd.phase_method != run->phase_method))
            rc = fail(c, ETRAJ_EINVAL, "sampler and run disagree on dimension / phase method");
        if (!rc && !(run->traj_t_final > sd.sample_t_hi))
            rc = fail(c, ETRAJ_EINVAL, "[TrajectorySimulationJob] `traj_t_final` must be greater than the last time point of `sample_t_intv`.");
        if (!rc && !(sd.sample_t_lo < sd.sample_t_hi) && sd.sample_t_num > 1)
            rc = fail(c, ETRAJ_EINVAL, "[TrajectorySimulationJob] `sample_t_intv` must be a valid interval.");
        if (rc) {
            if (n_jobs > 1) c->err = "job " + std::to_string(j) + ": " + c->err;
            return rc;
        }
        ncoef += wfat_doubles(J.S);
        npre += J.preC.size();
    }
    if (npre) {
        HIP_TRY(c, c->prec.ensure(npre));
        size_t off = 0;
        for (int32_t j = 0; j < n_jobs; ++j) {
            JobPlan& J = jobs[j];
            if (J.preC.empty()) continue;
            HIP_TRY(c, hipMemcpyAsync(c->prec.p + off, J.preC.data(), sizeof(double) * J.preC.size(),
                                      hipMemcpyHostToDevice, c->stream));
            J.S.pre_C = c->prec.p + off;
            off += J.preC.size();
        }
    }
    if (ncoef) {
        HIP_TRY(c, c->wfat.ensure(ncoef));
        size_t off = 0;
        for (int32_t j = 0; j < n_jobs; ++j) {
            SamplerP& S = jobs[j].S;
            if (S.method != ETRAJ_INIT_WFAT) continue;
            const size_t nc = wfat_doubles(S);
            HIP_TRY(c, hipMemcpyAsync(c->wfat.p + off, samplers[j].wfat_coeff, sizeof(double) * nc,
                                      hipMemcpyHostToDevice, c->stream));
            S.wfat_coeff = c->wfat.p + off;
            off += nc;
        }
    }
    return ETRAJ_OK;
}

// The batch loop of n_jobs independent jobs (TrajectorySimulation.jl:150-159
// per job). The batches of all jobs, job-major, are cut into slabs of at most
// kMaxSlabCols init columns and kFieldTabBudget bytes of field table; each slab is
// one field-table launch, one sampler launch and one integrator launch (per
// pass) over segments of any job. Tsit5 slabs hold one job each (its kernel
// evaluates the job's laser directly). PMDs, scalars and counters are per job.
static int run_scan_impl(etraj_ctx* c, int32_t n_jobs, const etraj_laser_desc* lasers,
                         const etraj_target_desc* target, const etraj_sampler_desc* samplers,
                         const etraj_run_desc* run, double* host_out, const double** dev_out, etraj_stats* stats) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    auto wall0 = std::chrono::steady_clock::now();
    TargetP T{};   // zero-filled, padding included: their bytes key the graph replay
    IntegP P{};
    if (int rc = check_target(c, target, T)) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<JobPlan> jobs;
    if (int rc = plan_jobs(c, n_jobs, lasers, target, T, samplers, run, jobs)) return rc;
    if (int rc = make_integ(c, run, jobs[0].L, T, P)) return rc;
    const bool rk4 = P.integrator == ETRAJ_INTEG_RK4;
    const bool coh = P.phase != ETRAJ_PHASE_CTMC;
    const int64_t npix = npix_of(P);
    const int64_t plen = coh ? 2 * npix : npix;
    const int rows = jobs[0].S.rows;

    // batches of all jobs, job-major
    struct Item { int32_t job, b; double t0; int64_t cap; };
    std::vector<Item> items;
    int64_t stride = 1;
    for (int32_t j = 0; j < n_jobs; ++j) {
        const int64_t cap = std::max<int64_t>(jobs[j].S.max_bs, 1);
        for (int32_t b = 0; b < (int32_t)jobs[j].t.size(); ++b) {
            items.push_back({j, b, jobs[j].t[b], cap});
            stride = std::max(stride, steps_of(P.tf, jobs[j].t[b], P.dt));
        }
    }
    const int64_t seg_max = std::max<int64_t>(1, kFieldTabBudget / (stride * kFieldTabBytesPerStep));
    // Tsit5: a job's batches in one slab up to 2 kMaxTs5SlabCols columns, else in the fewest
    // slabs of at most kMaxTs5SlabCols, each filled to the job's columns / number of slabs
    // (the last takes the rest)
    std::vector<int64_t> ts5_fill(n_jobs, 0);
    if (!rk4) {
        for (const Item& it : items) ts5_fill[it.job] += it.cap;
        for (int64_t& v : ts5_fill) {
            const int64_t n = v <= 2 * kMaxTs5SlabCols ? 1 : (v + kMaxTs5SlabCols - 1) / kMaxTs5SlabCols;
            v = std::max<int64_t>(1, (v + n - 1) / n);
        }
    }
    struct Slab { size_t i0, i1; int64_t cols; int64_t maxcap; };
    std::vector<Slab> slabs;
    for (size_t i = 0; i < items.size();) {
        Slab sl{i, i, 0, 0};
        const int64_t max_cols_slab = rk4 ? kMaxSlabCols : 2 * kMaxTs5SlabCols;
        while (sl.i1 < items.size()) {
            const Item& it = items[sl.i1];
            const bool empty = sl.i1 == sl.i0;
            if (!empty && (sl.cols + it.cap > max_cols_slab || (int64_t)(sl.i1 - sl.i0) >= seg_max)) break;
            if (!empty && !rk4 && (it.job != items[sl.i0].job || sl.cols >= ts5_fill[it.job])) break;
            sl.cols += it.cap;
            sl.maxcap = std::max(sl.maxcap, it.cap);
            ++sl.i1;
        }
        slabs.push_back(sl);
        i = sl.i1;
    }
    int64_t max_cols = 0, max_nb = 0;
    for (const Slab& sl : slabs) {
        max_cols = std::max(max_cols, sl.cols);
        max_nb = std::max<int64_t>(max_nb, (int64_t)(sl.i1 - sl.i0));
    }

    // PMD binning: records + LDS tiles, or atomics from the integrator
    const bool fixed_sums = run->pmd_mode == ETRAJ_PMD_FIXED;
    const int64_t total_cells = (int64_t)n_jobs * npix;
    const int words = coh ? 2 : 1;
    if (run->pmd_binning < ETRAJ_BIN_AUTO || run->pmd_binning > ETRAJ_BIN_TILES)
        return fail(c, ETRAJ_EINVAL, "pmd_binning must be ETRAJ_BIN_AUTO, ETRAJ_BIN_ATOMIC or ETRAJ_BIN_TILES");
    BinCfg bcfg{0, c->bin_chunk};
    int ntile = 0;
    const bool auto_tiles = c->bin_auto_tiles < 0 ? run->integrator == ETRAJ_INTEG_TSIT5 : c->bin_auto_tiles == 1;
    bool tiles = run->pmd_binning == ETRAJ_BIN_TILES || (run->pmd_binning == ETRAJ_BIN_AUTO && auto_tiles);
    if (tiles) {
        const int lds_shift = bin_tile_shift(words, fixed_sums, run->hit_counts != nullptr, 30);
        bcfg.shift = c->bin_shift > 0 ? std::min(c->bin_shift, lds_shift) : lds_shift;
        while (bcfg.shift < lds_shift && ((total_cells + (1ll << bcfg.shift) - 1) >> bcfg.shift) > kBinMaxTiles)
            ++bcfg.shift;   // few enough tiles for the tile tables
        ntile = (int)std::min<int64_t>((total_cells + (1ll << bcfg.shift) - 1) >> bcfg.shift, INT32_MAX);
        if (ntile > kBinMaxTiles || total_cells >= (int64_t)kNoCell) {
            if (run->pmd_binning == ETRAJ_BIN_TILES)
                return fail(c, ETRAJ_EUNSUPPORTED, "tile binning: the PMD needs more than 8192 tiles of LDS size "
                                                   "(or 2^32 - 1 cells); use ETRAJ_BIN_ATOMIC");
            tiles = false;
        }
    }
    if (tiles) {
        HIP_TRY(c, c->bin_cell.ensure(max_cols));
        HIP_TRY(c, c->bin_scell.ensure(max_cols));
        HIP_TRY(c, c->bin_amp.ensure(max_cols));
        HIP_TRY(c, c->bin_samp.ensure(max_cols));
        HIP_TRY(c, c->tile_count.ensure(kBinMaxTiles));
        HIP_TRY(c, c->tile_off.ensure(kBinMaxTiles + 1));
        HIP_TRY(c, c->chunk_off.ensure(kBinMaxTiles + 1));
        HIP_TRY(c, c->cursor.ensure(kBinMaxTiles));
    }
    const BinBufs bbuf{c->bin_cell.p, c->bin_amp.p, c->tile_count.p, c->tile_off.p, c->chunk_off.p, c->cursor.p,
                       c->bin_scell.p, c->bin_samp.p};

    std::vector<SamplerP> hsp(n_jobs);
    std::vector<LaserP> hlp(n_jobs);
    for (int32_t j = 0; j < n_jobs; ++j) {
        hsp[j] = jobs[j].S;
        hlp[j] = jobs[j].L;
    }
    const bool fixed = run->pmd_mode == ETRAJ_PMD_FIXED;
    HIP_TRY(c, c->pmd.ensure((size_t)n_jobs * plen));
    HIP_TRY(c, c->scal.ensure((size_t)2 * kAccSlots * n_jobs));
    HIP_TRY(c, c->cnt.ensure((size_t)CNT_N * kAccSlots * n_jobs));
    if (fixed) {
        HIP_TRY(c, c->pmdq.ensure((size_t)2 * n_jobs * plen));
        HIP_TRY(c, c->scalq.ensure((size_t)4 * kAccSlots * n_jobs));
    }
    if (run->hit_counts) HIP_TRY(c, c->hits.ensure((size_t)n_jobs * npix));
    HIP_TRY(c, c->sps.ensure(n_jobs));
    HIP_TRY(c, c->lps.ensure(n_jobs));
    // restep_radius <= 0: no electron is re-integrated, so no flags, compaction or FP64 pass
    const bool two_pass = rk4 && P.mixed && P.restep_radius > 0;
    const int nslab = (int)slabs.size();
    for (int q = 0; q < std::min(kSlabSlots, nslab); ++q) {
        SlabSlot& Q = c->slot[q];
        HIP_TRY(c, Q.init.ensure((size_t)max_cols * rows));
        HIP_TRY(c, Q.counts.ensure(max_nb));
        HIP_TRY(c, Q.counts2.ensure(max_nb));
        HIP_TRY(c, Q.seg_start.ensure(max_nb));
        HIP_TRY(c, Q.seg_t0.ensure(max_nb));
        HIP_TRY(c, Q.seg_job.ensure(max_nb));
        HIP_TRY(c, Q.seg_b.ensure(max_nb));
        HIP_TRY(c, Q.pre.ensure(max_nb + 1));
        HIP_TRY(c, Q.queue.ensure(1));
        if (rk4) {
            HIP_TRY(c, Q.ft32.ensure((size_t)max_nb * stride));
            HIP_TRY(c, Q.ft64.ensure((size_t)2 * max_nb * stride));
            HIP_TRY(c, Q.ft0.ensure(max_nb));
        }
        if (two_pass || (!rk4 && P.mixed)) HIP_TRY(c, Q.flags.ensure(max_cols));
        if (!rk4) HIP_TRY(c, Q.prep.ensure((size_t)max_cols * ts5_prep_doubles(P.dim)));
        if (!rk4 && P.mixed) HIP_TRY(c, Q.counts2.ensure(max_nb));
    }
    HIP_TRY(c, hipMemcpyAsync(c->sps.p, hsp.data(), sizeof(SamplerP) * n_jobs, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->lps.p, hlp.data(), sizeof(LaserP) * n_jobs, hipMemcpyHostToDevice, c->stream));
    if (coh) HIP_TRY(c, c->pmd_out.ensure((size_t)n_jobs * npix));
    const double* result = coh ? c->pmd_out.p : c->pmd.p;

    // A job of one slab with ctx->graph set runs as a HIP graph: its launches are captured once and
    // replayed while their arguments repeat (the key below: the device buffers, the parameter
    // blocks and the segment table). The host-to-device copies (parameter blocks, segment table)
    // stay outside the graph, made before each launch; the graph holds kernels only (the clearing
    // included) and runs on the first stream alone (no FP64-pass overlap); the per-stage event
    // timings give way to the graph's total.
    const bool use_graph = c->graph && nslab == 1;
    // per slab, first stream: start, sampled, main integrator kernel done, first-stream work done;
    // RK4 two-pass: FP64 pass start and end (second stream); Tsit5: [4] main integrator kernel start
    std::vector<hipEvent_t> ev(6 * nslab, nullptr);
    for (auto& e : ev) HIP_TRY(c, hipEventCreate(&e));
    auto rec = [&](int i, hipStream_t s) -> hipError_t { return use_graph ? hipSuccess : hipEventRecord(ev[i], s); };
    // the segment table of slab k (host staging in its slot, then copied to the slot's device buffers)
    auto stage_slab = [&](int k, SlabSlot& Q) -> hipError_t {
        const Slab& sl = slabs[k];
        const int nb = (int)(sl.i1 - sl.i0);
        Q.h_t0.resize(nb); Q.h_job.resize(nb); Q.h_b.resize(nb); Q.h_start.resize(nb);
        int64_t col = 0;
        for (int i = 0; i < nb; ++i) {
            const Item& it = items[sl.i0 + i];
            Q.h_t0[i] = it.t0;
            Q.h_job[i] = it.job;
            Q.h_b[i] = it.b;
            Q.h_start[i] = col;
            col += it.cap;
        }
        hipStream_t A = c->stream;
        hipError_t e = hipMemcpyAsync(Q.seg_t0.p, Q.h_t0.data(), sizeof(double) * nb, hipMemcpyHostToDevice, A);
        if (e == hipSuccess) e = hipMemcpyAsync(Q.seg_job.p, Q.h_job.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, A);
        if (e == hipSuccess) e = hipMemcpyAsync(Q.seg_b.p, Q.h_b.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, A);
        if (e == hipSuccess)
            e = hipMemcpyAsync(Q.seg_start.p, Q.h_start.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice, A);
        return e;
    };

    const BinOut bout{c->pmd.p, fixed ? c->pmdq.p : nullptr, run->hit_counts ? c->hits.p : nullptr, c->cnt.p, npix};
    // every launch of the job, on the first stream (and the second for overlapped FP64 passes)
    auto enqueue = [&]() -> int {
        hipStream_t A = c->stream;
        HIP_TRY(c, launch_zero(A, c->pmd.p, sizeof(double) * n_jobs * plen));
        HIP_TRY(c, launch_zero(A, c->scal.p, sizeof(double) * 2 * kAccSlots * n_jobs));
        HIP_TRY(c, launch_zero(A, c->cnt.p, sizeof(int64_t) * CNT_N * kAccSlots * n_jobs));
        if (fixed) {
            HIP_TRY(c, launch_zero(A, c->pmdq.p, sizeof(uint64_t) * 2 * n_jobs * plen));
            HIP_TRY(c, launch_zero(A, c->scalq.p, sizeof(uint64_t) * 4 * kAccSlots * n_jobs));
        }
        if (run->hit_counts) HIP_TRY(c, launch_zero(A, c->hits.p, sizeof(uint64_t) * n_jobs * npix));
        OutP O{};
        O.pmd = c->pmd.p;
        O.scal = c->scal.p;
        O.cnt = c->cnt.p;
        O.job_stride = plen;
        O.npix = npix;
        O.pmdq = fixed ? c->pmdq.p : nullptr;
        O.scalq = fixed ? c->scalq.p : nullptr;
        O.hits = run->hit_counts ? c->hits.p : nullptr;
        // The FP64 pass of a slab runs on the second stream (c->overlap_fp64, the
        // default), put there once the next slab has been sampled so that it
        // overlaps the next slab's FP32 pass rather than that slab's short
        // sampling kernels, which it would starve of dispatch slots; or on the
        // first stream right after the slab's FP32 pass and compaction. On the
        // 1e8-electron job the overlap saves about 2 % (A/B on one box).
        struct Fp64Pass { int k = -1, nb = 0, cps = 0; SlabSlot* Q = nullptr; FieldTab FT{}; };
        Fp64Pass prev;
        const bool overlap = c->overlap_fp64 && !use_graph;
        auto launch_fp64 = [&](const Fp64Pass& f, hipEvent_t after) -> hipError_t {
            hipStream_t B = overlap ? c->stream2 : A;
            hipError_t e = overlap ? hipStreamWaitEvent(B, after, 0) : hipSuccess;
            if (e == hipSuccess) e = rec(6 * f.k + 4, B);
            OutP Os = O;
            Os.seg_job = f.Q->seg_job.p;
            IntegP P64 = P;
            P64.mixed = 0;
            if (e == hipSuccess)
                e = launch_rk4(B, P64, Os, OUT_PMD, f.Q->init.p, f.Q->seg_start.p, f.Q->counts2.p, f.nb, f.cps, f.FT,
                               nullptr, nullptr);
            if (e == hipSuccess) e = rec(6 * f.k + 5, B);
            if (e == hipSuccess && !use_graph) e = hipEventRecord(f.Q->done, B);
            return e;
        };
        int rc = ETRAJ_OK;
        for (int k = 0; k < nslab && rc == ETRAJ_OK; ++k) {
            const Slab& sl = slabs[k];
            const int nb = (int)(sl.i1 - sl.i0);
            SlabSlot& Q = c->slot[k % kSlabSlots];
            hipError_t e = hipSuccess;
            if (Q.pending) {   // the slot's previous slab (device buffers and host staging) must be finished
                e = hipEventSynchronize(Q.done);
                Q.pending = false;
            }
            if (e == hipSuccess && !use_graph) e = stage_slab(k, Q);   // graph: staged before the launch
            if (e == hipSuccess) e = rec(6 * k, A);
            FieldTab FT{Q.ft32.p, Q.ft64.p, Q.ft0.p, stride};
            if (e == hipSuccess && rk4) e = launch_field_table(A, c->lps.p, Q.seg_job.p, Q.seg_t0.p, nb, P.tf, P.dt, FT);
            const SegTab G{Q.seg_t0.p, Q.seg_job.p, Q.seg_b.p, Q.seg_start.p};
            if (e == hipSuccess)
                e = launch_adk_sample(A, jobs[0].S.method, jobs[0].S.pre_L >= 0, c->sps.p, c->lps.p, T, G, nb, Q.init.p,
                                      Q.counts.p, c->cnt.p);
            if (e == hipSuccess) e = rec(6 * k + 1, A);
            if (e == hipSuccess && prev.k >= 0) {
                e = launch_fp64(prev, ev[6 * k + 1]);
                prev.k = -1;
            }
            if (tiles && e == hipSuccess) e = launch_bin_reset(A, bbuf.cell, sl.cols, bbuf.tile_count, ntile);
            if (rk4) {
                OutP Os = O;
                Os.seg_job = Q.seg_job.p;
                if (tiles) {
                    Os.bin_cell = bbuf.cell;
                    Os.bin_amp = bbuf.amp;
                }
                const int cps = (int)((sl.maxcap + 63) / 64);
                if (e == hipSuccess)
                    e = launch_rk4(A, P, Os, OUT_PMD, Q.init.p, Q.seg_start.p, Q.counts.p, nb, cps, FT,
                                   two_pass ? Q.flags.p : nullptr, nullptr);
                if (e == hipSuccess) e = rec(6 * k + 2, A);
                if (tiles && e == hipSuccess) e = launch_bin_tiles(A, bcfg, bbuf, sl.cols, total_cells, words, fixed, bout);
                if (two_pass) {
                    // the electrons flagged by the FP32 pass, integrated again in FP64: now, or (overlap)
                    // on the second stream after the next slab's sampling or after the loop
                    if (e == hipSuccess)
                        e = launch_compact_flagged(A, Q.init.p, P.rows, Q.seg_start.p, Q.counts.p, nb, Q.flags.p,
                                                   nullptr, Q.counts2.p, nullptr);
                    if (e == hipSuccess) e = rec(6 * k + 3, A);
                    const Fp64Pass cur{k, nb, cps, &Q, FT};
                    if (overlap) prev = cur;
                    else if (e == hipSuccess) e = launch_fp64(cur, ev[6 * k + 3]);
                }
            } else {
                const int32_t j = items[sl.i0].job;   // a Tsit5 slab holds one job
                IntegP Pj = P;
                Pj.L = jobs[j].L;
                OutP Oj = O;
                Oj.pmd = O.pmd + j * plen;
                Oj.scal = O.scal + 2 * kAccSlots * j;
                Oj.cnt = O.cnt + (int64_t)j * CNT_N * kAccSlots;
                if (fixed) {
                    Oj.pmdq = O.pmdq + 2 * j * plen;
                    Oj.scalq = O.scalq + 4 * kAccSlots * j;
                }
                if (O.hits) Oj.hits = O.hits + j * npix;
                OutP Om = Oj;   // the main launch: records for the tiles, or atomics
                if (tiles) {
                    Om.bin_cell = bbuf.cell;
                    Om.bin_amp = bbuf.amp;
                    Om.cell0 = j * npix;
                }
                // start values of every electron, the integration, then the Kepler asymptote and binning of the
                // final states (launch_tsit5_staged); ev[6k+2] marks the end of the integrator kernel
                const int cps = (int)((sl.maxcap + 63) / 64);
                if (e == hipSuccess) e = launch_scan_counts(A, Q.counts.p, nb, Q.pre.p);
                if (e == hipSuccess)
                    e = launch_tsit5_staged(A, Pj, Om, Q.init.p, Q.seg_start.p, Q.counts.p, Q.pre.p, nb, cps, sl.cols,
                                            Q.queue.p, Q.prep.p, P.mixed ? Q.flags.p : nullptr,
                                            use_graph ? nullptr : ev[6 * k + 4], use_graph ? nullptr : ev[6 * k + 2]);
                if (tiles && e == hipSuccess) e = launch_bin_tiles(A, bcfg, bbuf, sl.cols, total_cells, words, fixed, bout);
                if (P.mixed) {   // the flagged electrons again in FP64, on the same stream (binned with atomics)
                    IntegP P64 = Pj;
                    P64.mixed = 0;
                    if (e == hipSuccess)
                        e = launch_compact_flagged(A, Q.init.p, P.rows, Q.seg_start.p, Q.counts.p, nb, Q.flags.p,
                                                   nullptr, Q.counts2.p, nullptr);
                    if (e == hipSuccess) e = launch_scan_counts(A, Q.counts2.p, nb, Q.pre.p);
                    if (e == hipSuccess)
                        e = launch_tsit5_staged(A, P64, Oj, Q.init.p, Q.seg_start.p, Q.counts2.p, Q.pre.p, nb, cps,
                                                sl.cols, Q.queue.p, Q.prep.p, nullptr, nullptr, nullptr);
                }
            }
            if (!two_pass) {
                if (e == hipSuccess) e = rec(6 * k + 3, A);
                if (e == hipSuccess && !use_graph) e = hipEventRecord(Q.done, A);
            }
            Q.pending = !use_graph;
            if (e != hipSuccess) rc = hip_fail(c, e, "run_job slab");
        }
        if (rc == ETRAJ_OK && prev.k >= 0) {
            hipError_t e = launch_fp64(prev, ev[6 * prev.k + 3]);
            if (e != hipSuccess) rc = hip_fail(c, e, "run_job FP64 pass");
        }
        // the first stream finishes the job after every FP64 pass
        for (int q = 0; q < kSlabSlots; ++q)
            if (c->slot[q].pending && rc == ETRAJ_OK) {
                hipError_t e = hipStreamWaitEvent(A, c->slot[q].done, 0);
                if (e != hipSuccess) rc = hip_fail(c, e, "run_job join");
            }
        if (rc) return rc;
        if (fixed) HIP_TRY(c, launch_fix_to_f64(A, c->pmdq.p, (int64_t)n_jobs * plen, c->pmd.p));
        if (coh) HIP_TRY(c, launch_finalize_abs2(A, c->pmd.p, (int64_t)n_jobs * npix, c->pmd_out.p));
        return ETRAJ_OK;
    };

    int rc = ETRAJ_OK;
    if (!use_graph) {
        rc = enqueue();
    } else {
        hipError_t e = stage_slab(0, c->slot[0]);
        // the graph's launches are fixed by these: a replay needs them all unchanged
        std::string key;
        auto put = [&key](const void* p, size_t n) { key.append(static_cast<const char*>(p), n); };
        auto put_v = [&put](auto v) { put(&v, sizeof v); };
        put(&P, sizeof P);
        put(&T, sizeof T);
        put(hsp.data(), sizeof(SamplerP) * hsp.size());
        put(hlp.data(), sizeof(LaserP) * hlp.size());
        for (const Item& it : items) {
            put_v(it.job); put_v(it.b); put_v(it.t0); put_v(it.cap);
        }
        put_v(n_jobs); put_v(npix); put_v(plen); put_v(stride); put_v(total_cells); put_v(ntile); put_v(bcfg.shift);
        put_v(bcfg.chunk); put_v(words); put_v(fixed); put_v(coh); put_v(tiles); put_v(two_pass); put_v(rk4);
        put_v(run->hit_counts != nullptr);
        const SlabSlot& Q0 = c->slot[0];
        for (const void* ptr : {(const void*)c->pmd.p, (const void*)c->pmd_out.p, (const void*)c->scal.p,
                                (const void*)c->cnt.p, (const void*)c->pmdq.p, (const void*)c->scalq.p,
                                (const void*)c->hits.p, (const void*)c->sps.p, (const void*)c->lps.p,
                                (const void*)Q0.init.p, (const void*)Q0.counts.p, (const void*)Q0.counts2.p,
                                (const void*)Q0.seg_start.p, (const void*)Q0.seg_t0.p, (const void*)Q0.seg_job.p,
                                (const void*)Q0.seg_b.p, (const void*)Q0.pre.p, (const void*)Q0.queue.p,
                                (const void*)Q0.ft32.p, (const void*)Q0.ft64.p, (const void*)Q0.ft0.p,
                                (const void*)Q0.flags.p, (const void*)Q0.prep.p, (const void*)bbuf.cell,
                                (const void*)bbuf.amp, (const void*)bbuf.tile_count, (const void*)bbuf.tile_off,
                                (const void*)bbuf.chunk_off, (const void*)bbuf.cursor, (const void*)bbuf.scell,
                                (const void*)bbuf.samp})
            put_v(ptr);
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "run_job segment table");
        } else if (!c->gexec || key != c->gkey) {
            if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
            c->gexec = nullptr;
            c->gkey.clear();
            e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
            if (e != hipSuccess) {
                rc = hip_fail(c, e, "run_job graph capture");
            } else {
                rc = enqueue();
                hipGraph_t g = nullptr;
                e = hipStreamEndCapture(c->stream, &g);
                if (!rc && e != hipSuccess) rc = hip_fail(c, e, "run_job graph capture");
                if (!rc) {
                    e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
                    if (e != hipSuccess) {
                        c->gexec = nullptr;
                        rc = hip_fail(c, e, "run_job graph instantiate");
                    } else {
                        c->gkey = std::move(key);
                    }
                }
                if (g) (void)hipGraphDestroy(g);
            }
        }
        if (!rc) {
            e = hipEventRecord(ev[0], c->stream);
            if (e == hipSuccess) e = hipGraphLaunch(c->gexec, c->stream);
            if (e == hipSuccess) e = hipEventRecord(ev[3], c->stream);
            if (e != hipSuccess) rc = hip_fail(c, e, "run_job graph launch");
        }
    }
    if (rc) {
        (void)hipStreamSynchronize(c->stream2);
        (void)hipStreamSynchronize(c->stream);
        for (SlabSlot& Q : c->slot) Q.pending = false;
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        return rc;
    }
    // per-slot copies (kAccSlots per job), added on the host below
    std::vector<double> scal_s((size_t)2 * kAccSlots * n_jobs), scal(2 * n_jobs, 0.0);
    std::vector<int64_t> cnt_s((size_t)CNT_N * kAccSlots * n_jobs), cnt((size_t)CNT_N * n_jobs, 0);
    HIP_TRY(c, hipMemcpyAsync(scal_s.data(), c->scal.p, sizeof(double) * scal_s.size(), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipMemcpyAsync(cnt_s.data(), c->cnt.p, sizeof(int64_t) * cnt_s.size(), hipMemcpyDeviceToHost, c->stream));
    std::vector<uint64_t> scalq_s(fixed ? (size_t)4 * kAccSlots * n_jobs : 0), hits(run->hit_counts ? n_jobs * npix : 0);
    if (fixed)
        HIP_TRY(c, hipMemcpyAsync(scalq_s.data(), c->scalq.p, sizeof(uint64_t) * scalq_s.size(), hipMemcpyDeviceToHost,
                                  c->stream));
    if (run->hit_counts)
        HIP_TRY(c, hipMemcpyAsync(hits.data(), c->hits.p, sizeof(uint64_t) * n_jobs * npix, hipMemcpyDeviceToHost,
                                  c->stream));
    if (host_out)
        HIP_TRY(c, hipMemcpyAsync(host_out, result, sizeof(double) * npix * n_jobs, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int32_t j = 0; j < n_jobs; ++j)
        for (int q = 0; q < kAccSlots; ++q) {
            const size_t s0 = (size_t)j * kAccSlots + q;
            scal[2 * j] += scal_s[2 * s0];
            scal[2 * j + 1] += scal_s[2 * s0 + 1];
            for (int k = 0; k < CNT_N; ++k) cnt[(size_t)CNT_N * j + k] += cnt_s[CNT_N * s0 + k];
        }
    if (fixed) {
        // every partial sum of a cell is bounded by the job's sum of |values|: the
        // weights' sum (CTMC) or, for amplitudes sqrt(W), sqrt(n * sum W)
        for (int32_t j = 0; j < n_jobs; ++j) {
            uint64_t q4[4] = {0, 0, 0, 0};   // the 128-bit sums of the slots (exact)
            for (int q = 0; q < kAccSlots; ++q) {
                const uint64_t* w = scalq_s.data() + 4 * ((size_t)j * kAccSlots + q);
                for (int h = 0; h < 4; h += 2) {
                    const uint64_t lo = q4[h] + w[h];
                    q4[h + 1] += w[h + 1] + (lo < q4[h] ? 1 : 0);
                    q4[h] = lo;
                }
            }
            scal[2 * j] = fix_to_double(q4[0], q4[1]);
            scal[2 * j + 1] = fix_to_double(q4[2], q4[3]);
            const int64_t* cj = cnt.data() + (size_t)CNT_N * j;
            const double bound = coh ? std::sqrt((double)cj[CNT_LAUNCHED] * scal[2 * j]) : scal[2 * j];
            if (cj[CNT_FIXRANGE] != 0 || !(bound < 0x1.0p26)) {
                for (SlabSlot& Q : c->slot) Q.pending = false;
                for (auto& e : ev)
                    if (e) (void)hipEventDestroy(e);
                return fail(c, ETRAJ_ERANGE, "fixed-point PMD: a sum reaches 2^26 (use pmd_mode = ETRAJ_PMD_FLOAT)");
            }
        }
    }
    if (run->hit_counts)
        for (int64_t i = 0; i < n_jobs * npix; ++i) run->hit_counts[i] += hits[i];
    for (SlabSlot& Q : c->slot) Q.pending = false;
    double ms_s = 0, ms_i = 0, ms_m = 0;
    if (use_graph) {   // the graph's device time, all of it as integration
        float g = 0;
        (void)hipEventElapsedTime(&g, ev[0], ev[3]);
        ms_i = g;
    }
    for (int k = 0; k < nslab && !use_graph; ++k) {
        float a = 0, b = 0, m = 0, f = 0;
        (void)hipEventElapsedTime(&a, ev[6 * k], ev[6 * k + 1]);
        (void)hipEventElapsedTime(&b, ev[6 * k + 1], ev[6 * k + 3]);
        // the main integrator kernel: RK4 right after sampling; Tsit5 after its start values
        (void)hipEventElapsedTime(&m, ev[6 * k + (rk4 ? 1 : 4)], ev[6 * k + 2]);
        if (two_pass) (void)hipEventElapsedTime(&f, ev[6 * k + 4], ev[6 * k + 5]);
        ms_s += a;
        ms_i += b + f;
        ms_m += m;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (dev_out) *dev_out = result;
    if (stats) {
        const double ms_total =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
        for (int32_t j = 0; j < n_jobs; ++j) {
            etraj_stats& s = stats[j];
            const int64_t* cj = cnt.data() + (size_t)CNT_N * j;
            s.ion_prob = scal[2 * j];
            s.ion_prob_uncollected = scal[2 * j + 1];
            s.num_effective_traj = cj[CNT_LAUNCHED];
            s.n_failed = cj[CNT_FAILED];
            s.n_bound = cj[CNT_BOUND];
            s.n_restep = cj[CNT_RESTEP];
            s.n_steps = cj[CNT_STEPS];
            s.n_steps_fp64 = cj[CNT_STEPS64];
            s.ms_sample = ms_s;
            s.ms_integrate = ms_i;
            s.ms_integrate_main = ms_m;
            s.ms_total = ms_total;
            s.n_slabs = nslab;
            s.n_attempts = cj[CNT_ATTEMPTS];
        }
    }
    return ETRAJ_OK;
}

int etraj_run_job(etraj_ctx* c, const etraj_laser_desc* laser, const etraj_target_desc* target,
                  const etraj_sampler_desc* sampler, const etraj_run_desc* run, double* pmd_out, etraj_stats* stats) {
    if (c && !pmd_out) return fail(c, ETRAJ_EINVAL, "pmd_out is NULL");
    if (c && (!laser || !sampler)) return fail(c, ETRAJ_EINVAL, "laser / sampler descriptor is NULL");
    return run_scan_impl(c, 1, laser, target, sampler, run, pmd_out, nullptr, stats);
}

int etraj_run_job_device(etraj_ctx* c, const etraj_laser_desc* laser, const etraj_target_desc* target,
                         const etraj_sampler_desc* sampler, const etraj_run_desc* run, const double** pmd_dev,
                         etraj_stats* stats) {
    if (c && !pmd_dev) return fail(c, ETRAJ_EINVAL, "pmd_dev is NULL");
    if (c && (!laser || !sampler)) return fail(c, ETRAJ_EINVAL, "laser / sampler descriptor is NULL");
    return run_scan_impl(c, 1, laser, target, sampler, run, nullptr, pmd_dev, stats);
}

int etraj_run_scan(etraj_ctx* c, int32_t n_jobs, const etraj_laser_desc* lasers, const etraj_target_desc* target,
                   const etraj_sampler_desc* samplers, const etraj_run_desc* run, double* pmd_out,
                   etraj_stats* stats) {
    if (c && !pmd_out) return fail(c, ETRAJ_EINVAL, "pmd_out is NULL");
    return run_scan_impl(c, n_jobs, lasers, target, samplers, run, pmd_out, nullptr, stats);
}

int etraj_run_scan_device(etraj_ctx* c, int32_t n_jobs, const etraj_laser_desc* lasers,
                          const etraj_target_desc* target, const etraj_sampler_desc* samplers,
                          const etraj_run_desc* run, const double** pmd_dev, etraj_stats* stats) {
    if (c && !pmd_dev) return fail(c, ETRAJ_EINVAL, "pmd_dev is NULL");
    return run_scan_impl(c, n_jobs, lasers, target, samplers, run, nullptr, pmd_dev, stats);
}

int etraj_get_tuning(const etraj_ctx* c, etraj_tuning* out) {
    if (!c || !out) return ETRAJ_EINVAL;
    *out = tuning_of(c);
    return ETRAJ_OK;
}

int etraj_set_tuning(etraj_ctx* c, const etraj_tuning* t) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    if (!t) return fail(c, ETRAJ_EINVAL, "tuning is NULL");
    if (const char* why = tuning_invalid(*t)) return fail(c, ETRAJ_EINVAL, why);
    apply_tuning(c, *t);
    return ETRAJ_OK;
}

int etraj_autotune(etraj_ctx* c, const etraj_laser_desc* laser, const etraj_target_desc* target,
                   const etraj_sampler_desc* sampler, const etraj_run_desc* run, etraj_tuning* best,
                   double* best_ms, double* start_ms) {
    if (!c) return ETRAJ_EINVAL;
    c->err.clear();
    if (!laser || !target || !sampler || !run) return fail(c, ETRAJ_EINVAL, "descriptor is NULL");
    const etraj_tuning start = tuning_of(c);
    // one candidate: two runs of the job (the first also allocates what the choice needs), the faster counts
    auto time_of = [&](const etraj_tuning& t, double& ms) -> int {
        apply_tuning(c, t);
        ms = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            const double* dev = nullptr;
            etraj_stats st{};
            if (int rc = run_scan_impl(c, 1, laser, target, sampler, run, nullptr, &dev, &st)) return rc;
            ms = std::min(ms, st.ms_total);
        }
        return ETRAJ_OK;
    };
    etraj_tuning top = start;
    double top_ms = 0;
    int rc = time_of(start, top_ms);
    const double first_ms = top_ms;
    // coordinate descent over the choices that matter for the job's integrator: a candidate
    // replaces the best set only if it is faster by more than 1 %
    auto try_field = [&](int32_t etraj_tuning::*field, std::initializer_list<int32_t> values) {
        for (int32_t v : values) {
            if (rc || top.*field == v) continue;
            etraj_tuning cand = top;
            cand.*field = v;
            double ms = 0;
            rc = time_of(cand, ms);
            if (!rc && ms < 0.99 * top_ms) {
                top = cand;
                top_ms = ms;
            }
        }
    };
    const bool rk4 = run->integrator == ETRAJ_INTEG_RK4;
    if (rk4) {
        try_field(&etraj_tuning::rk4_block, {64, 128, 256});
        if (run->precision == ETRAJ_PREC_MIXED && run->restep_radius > 0) try_field(&etraj_tuning::overlap_fp64, {0, 1});
    } else {
        try_field(&etraj_tuning::tsit5_refill, {4, 8, 16, 32, 64});
        try_field(&etraj_tuning::tsit5_claim, {64, 256, 1024, 4096});
    }
    if (run->pmd_binning == ETRAJ_BIN_AUTO) {
        try_field(&etraj_tuning::bin_tiles, {0, 1});
        if (top.bin_tiles == 1 || (top.bin_tiles < 0 && !rk4)) {
            try_field(&etraj_tuning::bin_tile_shift, {0, 10});
            try_field(&etraj_tuning::bin_chunk, {4096, 16384, 65536});
        }
    }
    try_field(&etraj_tuning::graph, {0, 1});   // no effect on a job of several slabs: never kept there
    if (rc) {
        apply_tuning(c, start);
        return rc;
    }
    apply_tuning(c, top);
    if (best) *best = top;
    if (best_ms) *best_ms = top_ms;
    if (start_ms) *start_ms = first_ms;
    return ETRAJ_OK;
}

}  // extern "C"
