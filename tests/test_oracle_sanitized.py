"""The CPU oracle (oracle/etraj_oracle.c) built with AddressSanitizer and
UndefinedBehaviorSanitizer (host code only) and run through its main paths on
small cases: ADK / SFA-SPA / SFA-SPANE / WFAT-table samplers in 2D and 3D,
RK4 and Tsit5 integration with the three phase methods, the threaded job with
thread-private PMDs (static and dynamic schedules). Any out-of-bounds access,
use after free, leak or undefined operation fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CC = shutil.which("gcc")

DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "etraj_oracle.h"

static int check(int rc, const char* what) {
    if (rc) { fprintf(stderr, "%s failed: %d\n", what, rc); exit(1); }
    return rc;
}

int main(void) {
    etraj_laser_desc L; memset(&L, 0, sizeof L);
    L.n_colors = 1;
    L.color[0].kind = ETRAJ_LASER_COS2; L.color[0].peak_int = 4e14; L.color[0].wave_len = 800;
    L.color[0].cyc_num = 2; L.color[0].ellip = 0.5;
    etraj_target_desc T; memset(&T, 0, sizeof T);
    T.kind = ETRAJ_TARGET_HYDROGEN_LIKE; T.nucl_charge = 1; T.Ip = 0.5; T.asymp_coeff = 1; T.soft_core = 1e-10;
    etraj_target_desc A = T;   /* an SAE target (Ar-like Tong-Lin parameters) */
    A.kind = ETRAJ_TARGET_SAE; A.Ip = 0.579; A.a1 = 16.039; A.b1 = 2.007; A.a2 = -25.543; A.b2 = 4.525;
    A.a3 = 0.961; A.b3 = 0.443;
    double coeff[2 * 2 * 3] = {1.0, 0.2, 0.1, -0.3, 0.05, 0.0, 0.4, 0.1, -0.2, 0.0, 0.1, 0.1};
    long total = 0;
    for (int method = ETRAJ_INIT_ADK; method <= ETRAJ_INIT_WFAT; ++method)
    for (int dim = 2; dim <= 3; ++dim)
    for (int phase = 0; phase <= 2; ++phase) {
        if (method == ETRAJ_INIT_SPA && dim == 3 && phase != 1) continue;   /* keep the run short */
        etraj_sampler_desc S; memset(&S, 0, sizeof S);
        S.init_method = method; S.dimension = dim; S.phase_method = phase;
        S.rate_prefix = ETRAJ_PREFIX_PRECC | ETRAJ_PREFIX_JAC;
        S.sampling = (phase == 2) ? ETRAJ_SAMPLING_STRATIFIED : ETRAJ_SAMPLING_STEP;
        S.sample_t_num = 12; S.sample_t_lo = -60; S.sample_t_hi = 60;
        S.ss_kd_num = 10; S.ss_kz_num = dim == 3 ? 4 : 0; S.ss_kd_max = 1.0; S.ss_kz_max = 0.5;
        S.mc_kt_num = 40; S.kd_sigma = 0.4; S.kz_sigma = 0.3;
        S.cutoff_limit = 1e-16; S.seed = 42;
        S.wfat_na = 2; S.wfat_nb = 3; S.wfat_coeff = coeff; S.wfat_form = ETRAJ_WFAT_TABLE;
        if (method == ETRAJ_INIT_WFAT) S.rate_prefix = 0;
        etraj_run_desc R; memset(&R, 0, sizeof R);
        R.dimension = dim; R.phase_method = phase; R.traj_t_final = 130;
        R.integrator = (phase == 1) ? ETRAJ_INTEG_TSIT5 : ETRAJ_INTEG_RK4;
        R.dt = 0.2; R.rtol = 1e-6; R.atol = 1e-6; R.maxiters = 100000;
        R.quirks = ETRAJ_QUIRKS_REFERENCE;
        for (int d = 0; d < 3; ++d) { R.final_p_max[d] = 1.5; R.final_p_num[d] = d < dim ? 31 : 1; }
        const etraj_target_desc* tgt = (method == ETRAJ_INIT_SPANE) ? &A : &T;
        double* pmd = malloc(sizeof(double) * 31 * 31 * 31);
        etraj_stats st;
        check(oracle_run_job(&L, tgt, &S, &R, 3, 0, 0, 1, pmd, &st), "run_job");
        check(oracle_run_job_dynamic(&L, tgt, &S, &R, 2, 1, 0, 2, pmd, &st), "run_job_dynamic");
        total += st.num_effective_traj;
        int64_t cap = oracle_batch_max_size(&S), n = 0;
        int rows = dim == 3 ? (phase ? 9 : 8) : (phase ? 7 : 6);
        double* buf = malloc(sizeof(double) * rows * (cap > 0 ? cap : 1));
        check(oracle_gen_electron_batch(&L, tgt, &S, 6, buf, cap, &n), "gen_electron_batch");
        if (n > 0) {
            double* rec = malloc(sizeof(double) * 12 * n);
            int32_t* sts = malloc(sizeof(int32_t) * n);
            check(oracle_integrate_records(&L, tgt, &R, buf, n, rows, rec, sts), "integrate_records");
            free(rec); free(sts);
        }
        free(buf); free(pmd);
    }
    printf("ok %ld\n", total);
    return 0;
}
"""


def test_oracle_under_asan_and_ubsan(tmp_path):
    if not CC:
        pytest.skip("gcc not available")
    exe = tmp_path / "oracle_san"
    (tmp_path / "driver.c").write_text(DRIVER)
    r = subprocess.run([CC, "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=undefined", "-std=gnu11", "-ffp-contract=off",
                        "-I", os.path.join(ROOT, "oracle"), "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "oracle", "etraj_oracle.c"), str(tmp_path / "driver.c"),
                        "-o", str(exe), "-lm", "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.startswith("ok") and int(r.stdout.split()[1]) > 0
