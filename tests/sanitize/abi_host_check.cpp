// Host-side check of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer.  The library is
// built host-only (no device code) with both sanitizers; this driver makes every entry point reject
// arguments that it must refuse before any device work, exercises the pure-host calls (layouts,
// autotuner text, instrumentation), and, when no GPU is visible, makes the entry points that need a
// device fail with an error code instead of crashing.  Exit status 0 when every expectation holds.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/pylians_hip.h"

static int g_fail = 0, g_checks = 0;

static void expect(const char* what, int rc, int want) {
    ++g_checks;
    const char* msg = pyl_last_error();
    const bool ok = want < 0 ? rc != PYL_OK : rc == want;
    if (!ok) {
        ++g_fail;
        printf("FAIL %s: rc %d, wanted %s%d\n", what, rc, want < 0 ? "nonzero, not " : "", want < 0 ? 0 : want);
    } else if (rc != PYL_OK && (!msg || !msg[0])) {
        ++g_fail;
        printf("FAIL %s: rc %d without a message\n", what, rc);
    }
}
#define REJECT(call, code) expect(#call, (call), (code))
#define FAILS(call) expect(#call, (call), -1)
#define OKAY(call) expect(#call, (call), PYL_OK)

int main() {
    const int E = PYL_E_ARG;
    // fake, never dereferenced: every call below returns before touching its buffers
    std::vector<float> hb(64, 0.5f);
    std::vector<double> hd(64, 0.0);
    float* p = hb.data();
    double* o = hd.data();
    void* v = hb.data();
    if (pyl_abi_version() != PYL_ABI_VERSION) {
        printf("FAIL abi version\n");
        ++g_fail;
    }
    int count = -1;
    OKAY(pyl_device_count(&count));
    REJECT(pyl_device_count(NULL), E);

    // ---- layouts (host only): the documented lengths
    for (int dims = 2; dims <= 1024; ++dims) {
        int kpar = 0, kper = 0, kmax = 0;
        int64_t len = 0;
        const int rc = pyl_pk_layout(dims, &kpar, &kper, &kmax, &len);
        const int64_t n2 = (int64_t)(kpar + 1) * (kper + 1);
        if (rc != PYL_OK || len != 6LL * kmax + 3LL * kpar + 4 * n2 || kmax <= 0 || kpar <= 0 || kper <= 0) {
            printf("FAIL pyl_pk_layout(%d): rc %d len %lld kmax %d kpar %d kper %d\n", dims, rc, (long long)len, kmax,
                   kpar, kper);
            ++g_fail;
            break;
        }
        int xk = 0;
        int64_t xl = 0;
        if (pyl_xi_layout(dims, &xk, &xl) != PYL_OK || xl != 5LL * xk) {
            printf("FAIL pyl_xi_layout(%d)\n", dims);
            ++g_fail;
            break;
        }
        for (int nf = 1; nf <= 8; nf += 7) {
            int64_t ol = 0;
            if (pyl_xpk_layout(dims, nf, &ol) != PYL_OK || ol <= 0) {
                printf("FAIL pyl_xpk_layout(%d, %d)\n", dims, nf);
                ++g_fail;
                break;
            }
        }
        g_checks += 3;
    }
    int64_t len = 0;
    int a = 0, b = 0, c = 0;
    REJECT(pyl_pk_layout(1, &a, &b, &c, &len), E);
    REJECT(pyl_xpk_layout(64, 0, &len), E);
    REJECT(pyl_xpk_layout(64, 9, &len), E);

    // ---- arguments every entry point must refuse
    REJECT(pyl_mas_deposit(NULL, NULL, 10, 3, 16, 1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_deposit(p, NULL, -1, 3, 16, 1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_deposit(p, NULL, 10, 4, 16, 1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_deposit(p, NULL, 10, 3, 16, 1.0, 5, 0, v, NULL), E);
    REJECT(pyl_mas_deposit(p, NULL, 10, 3, 0, 1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_deposit(p, NULL, 10, 3, 16, -1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_deposit(p, NULL, 10, 3, 16, 1.0, 2, 0, NULL, NULL), E);
    REJECT(pyl_mas_overdensity(p, NULL, 0, 16, 1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_overdensity(NULL, NULL, 10, 16, 1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_overdensity(p, NULL, 10, 16, 1.0, 0, 0, v, NULL), E);
    REJECT(pyl_mas_overdensity(p, NULL, 10, 0, 1.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_overdensity(p, NULL, 10, 16, 0.0, 2, 0, v, NULL), E);
    REJECT(pyl_mas_overdensity(p, NULL, 10, 16, 1.0, 2, 0, NULL, NULL), E);
    REJECT(pyl_mas_CIC(p, p, NULL, 10, 16, 4, 1.0f, 1), E);
    REJECT(pyl_mas_NGP(NULL, p, NULL, 10, 16, 3, 1.0f, 1), E);
    REJECT(pyl_mas_TSC(p, NULL, NULL, 10, 16, 3, 1.0f, 1), E);
    REJECT(pyl_mas_PCS(p, p, NULL, -5, 16, 3, 1.0f, 1), E);
    REJECT(pyl_fft_r2c(v, v, 16, 3, NULL), E);
    REJECT(pyl_fft_c2r(v, v, 16, 5, 1, NULL), E);
    REJECT(pyl_fft_r2c(v, v, 17, PYL_F32, NULL), PYL_E_UNSUPPORTED);
    REJECT(pyl_pk(v, 16, 1.0, 3, 2, PYL_F32, 0, o, NULL, NULL), E);
    REJECT(pyl_pk(v, 16, 1.0, 0, 5, PYL_F32, 0, o, NULL, NULL), E);
    REJECT(pyl_pk(v, 16, -1.0, 0, 2, PYL_F32, 0, o, NULL, NULL), E);
    REJECT(pyl_pk(v, 16, 1.0, 0, 2, 7, 0, o, NULL, NULL), E);
    REJECT(pyl_pk(NULL, 16, 1.0, 0, 2, PYL_F32, 0, o, NULL, NULL), E);
    REJECT(pyl_pk(v, 16, 1.0, 0, 2, PYL_F32, 0, NULL, NULL, NULL), E);
    REJECT(pyl_pk_noalloc(v, 256, 1.0, 0, 2, PYL_F32, 0, o, NULL, NULL), PYL_E_UNSUPPORTED);
    REJECT(pyl_pk_particles(p, NULL, 0, 16, 1.0, 2, 0, PYL_F32, 0, o, NULL), E);
    REJECT(pyl_pk_particles(p, NULL, 10, 16, 1.0, 9, 0, PYL_F32, 0, o, NULL), E);
    REJECT(pyl_pk_particles(p, NULL, 10, 16, 1.0, 2, 3, PYL_F32, 0, o, NULL), E);
    REJECT(pyl_pk_particles(p, NULL, 10, 16, 1.0, 2, 0, PYL_F32, 0, NULL, NULL), E);
    REJECT(pyl_pk_particles_rsd(p, NULL, NULL, 10, 16, 1.0, 100.f, 0.f, 0, 2, 0, PYL_F32, 0, o, NULL), E);
    REJECT(pyl_pk_particles_rsd(p, p, NULL, 10, 16, 1.0, 100.f, 0.f, 3, 2, 0, PYL_F32, 0, o, NULL), E);
    REJECT(pyl_pk_particles_rsd(p, p, NULL, 10, 16, 1.0, 0.f, 0.f, 0, 2, 0, PYL_F32, 0, o, NULL), E);
    REJECT(pyl_pos_redshift_space(p, p, 10, 3, 1.f, 100.f, 0.f, 3, NULL), E);
    REJECT(pyl_pos_redshift_space(p, p, 10, 4, 1.f, 100.f, 0.f, 0, NULL), E);
    REJECT(pyl_pos_redshift_space(p, NULL, 10, 3, 1.f, 100.f, 0.f, 0, NULL), E);
    REJECT(pyl_pos_redshift_space(p, p, -1, 3, 1.f, 100.f, 0.f, 0, NULL), E);
    REJECT(pyl_pos_redshift_space(p, p, 10, 3, 1.f, 0.f, 0.f, 0, NULL), E);
    REJECT(pyl_correct_mas(v, v, 16, 2, 3, NULL), E);
    REJECT(pyl_xi(v, 16, 1.0, 2, 0, 3, o, NULL), E);
    REJECT(pyl_field_div_sub(v, 10, 1.0, 1.0, 3, NULL), E);
    REJECT(pyl_field_div_sub(v, -1, 1.0, 1.0, PYL_F32, NULL), E);
    REJECT(pyl_field_axpy(v, v, 10, 1.0, 3, NULL), E);
    const void* fields[2] = {v, v};
    const int mas[2] = {2, 2};
    REJECT(pyl_xpk(fields, 0, 16, 1.0, 0, mas, PYL_F32, o, NULL), E);
    REJECT(pyl_xpk(fields, 9, 16, 1.0, 0, mas, PYL_F32, o, NULL), E);
    REJECT(pyl_xpk(NULL, 2, 16, 1.0, 0, mas, PYL_F32, o, NULL), E);
    REJECT(pyl_xpk(fields, 2, 16, 1.0, 3, mas, PYL_F32, o, NULL), E);
    const double th[3] = {0.0, 1.0, 2.0};
    REJECT(pyl_bk(v, 16, 1.0, 0.1, 0.2, th, -1, 2, PYL_F32, o, o, o, o, NULL), E);
    REJECT(pyl_bk(v, 16, 0.0, 0.1, 0.2, th, 3, 2, PYL_F32, o, o, o, o, NULL), E);
    REJECT(pyl_bk(v, 16, 1.0, 0.1, 0.2, th, 3, 5, PYL_F32, o, o, o, o, NULL), E);
    REJECT(pyl_bk(v, 17, 1.0, 0.1, 0.2, th, 3, 2, PYL_F32, o, o, o, o, NULL), PYL_E_UNSUPPORTED);
    REJECT(pyl_bk(v, 16, 1.0, 0.1, 0.2, NULL, 3, 2, PYL_F32, o, o, o, o, NULL), E);
    REJECT(pyl_bk(v, 16, 1.0, 0.1, 0.2, th, 3, 2, 3, o, o, o, o, NULL), E);
    pyl_pk_job job = {p, NULL, 10, 16, 1.0, 2, 0, PYL_F32, 0, o};
    REJECT(pyl_pk_batch(NULL, 2, NULL), E);
    pyl_pk_job bad = job;
    bad.scheme = 7;
    REJECT(pyl_pk_batch(&bad, 1, NULL), E);
    bad = job;
    bad.n = 0;
    REJECT(pyl_pk_batch(&bad, 1, NULL), E);
    void* graph = NULL;
    REJECT(pyl_graph_create(&job, 1, NULL, NULL), E);
    REJECT(pyl_graph_create(&bad, 1, NULL, &graph), E);
    REJECT(pyl_graph_launch(NULL, NULL), E);
    REJECT(pyl_pk_adjoint(p, 10, 16, 1.0, 0, o, 4, p, o, NULL), E);
    REJECT(pyl_pk_adjoint(p, 0, 16, 1.0, 2, o, 4, p, o, NULL), E);
    REJECT(pyl_pk_adjoint(p, 10, 16, 1.0, 2, o, 100000, p, o, NULL), E);
    REJECT(pyl_interp(p, 16, 1.0, p, 10, 0, p, NULL), E);
    REJECT(pyl_interp(NULL, 16, 1.0, p, 10, 2, p, NULL), E);
    REJECT(pyl_ft_filter(0, 1.f, 1.f, 0, v, NULL), PYL_E_UNSUPPORTED);
    REJECT(pyl_tune_import(NULL), E);
    REJECT(pyl_dev_alloc(NULL, 16), E);
    void* dp = v;
    REJECT(pyl_dev_alloc(&dp, -1), E);
    OKAY(pyl_dev_alloc(&dp, 0));
    if (dp) {
        printf("FAIL zero-byte allocation returned a pointer\n");
        ++g_fail;
    }
    OKAY(pyl_dev_free(NULL));
    REJECT(pyl_memcpy_h2d(NULL, v, 16, NULL), E);
    REJECT(pyl_memcpy_d2h(v, NULL, 16, NULL), E);
    REJECT(pyl_memcpy_d2d(v, v, -1, NULL), E);
    REJECT(pyl_dev_zero(NULL, 16, NULL), E);
    OKAY(pyl_memcpy_h2d(NULL, NULL, 0, NULL));
    OKAY(pyl_dev_zero(NULL, 0, NULL));

    // ---- host-only state: autotuner text and instrumentation, truncated buffers included
    OKAY(pyl_tune(0));
    OKAY(pyl_tune(1));
    char small[2] = {'x', 'x'};
    if (pyl_tune_export(small, 1) < 0 || small[0] != 0) {
        printf("FAIL pyl_tune_export into a 1-byte buffer\n");
        ++g_fail;
    }
    if (pyl_tune_export(NULL, 0) < 0) {
        printf("FAIL pyl_tune_export length query\n");
        ++g_fail;
    }
    OKAY(pyl_profile(0));
    char pb[4] = {1, 1, 1, 1};
    if (pyl_profile_read(pb, 1) < 0 || pb[0] != 0) {
        printf("FAIL pyl_profile_read into a 1-byte buffer\n");
        ++g_fail;
    }

    // ---- without a GPU, entry points that need one report an error instead of crashing
    if (count == 0) {
        FAILS(pyl_dev_alloc(&dp, 1024));
        FAILS(pyl_mas_overdensity(p, NULL, 10, 16, 1.0, 2, 0, v, NULL));
        FAILS(pyl_pk(v, 16, 1.0, 0, 2, PYL_F32, 0, o, NULL, NULL));
        FAILS(pyl_pk_noalloc(v, 16, 1.0, 0, 2, PYL_F32, 0, o, NULL, NULL));
        FAILS(pyl_fft_r2c(v, v, 16, PYL_F32, NULL));
        FAILS(pyl_mas_CIC(p, p, NULL, 10, 4, 3, 1.0f, 1));
        FAILS(pyl_tune_import("16 2 4 0 3 0\n"));
        FAILS(pyl_stream_sync(NULL));
        FAILS(pyl_memcpy_h2d(v, p, 16, NULL));
        FAILS(pyl_pk_batch(&job, 1, NULL));
    } else {
        printf("a GPU is visible: device-path checks skipped\n");
    }
    printf("%d checks, %d failures\n", g_checks, g_fail);
    return g_fail ? 1 : 0;
}
