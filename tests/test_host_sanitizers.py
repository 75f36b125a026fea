"""Host code of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer.

libggbtd.so is rebuilt with ASan + UBSan on its host code (the device code
objects are built as usual, at -O1) and
driven by a small C program that calls every entry point with rejected
arguments, plus the calls that reach the HIP runtime: without a GPU those
fail, so the failure and clean-up paths run.  Any sanitizer report or an
unexpected status fails the test.  Needs no GPU; skipped where hipcc or
ROCm's clang is missing.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CLANG = os.path.join(ROCM, "llvm", "bin", "clang")
SRC = os.path.join(ROOT, "ggbtddft_amd", "csrc", "ggbtd.hip")

DRIVER = r"""
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include "ggbtd.h"

static int fails = 0;
#define EXPECT(cond) do { if (!(cond)) { fprintf(stderr, "line %d: %s\n", __LINE__, #cond); ++fails; } } while (0)

int main(void) {
    ggb_problem good = {GGB_ABI_VERSION, 64, 9, 16, 17, 2, 3, 2, 64, 1, 0, 0, 0.05};
    uint64_t bytes = 0;
    ggb_ctx *ctx = (ggb_ctx *)0x1;
    EXPECT(ggb_abi_version() == GGB_ABI_VERSION);
    EXPECT(ggb_footprint(NULL, &bytes) == GGB_E_ARG);
    EXPECT(ggb_footprint(&good, NULL) == GGB_E_ARG);
    EXPECT(ggb_footprint(&good, &bytes) == GGB_OK && bytes > 0);

    /* every size check of ggb_create, one at a time */
    EXPECT(ggb_create(NULL, &ctx) == GGB_E_ARG);
    EXPECT(ggb_create(&good, NULL) == GGB_E_ARG);
    for (int k = 0; k < 16; ++k) {
        ggb_problem p = good;
        switch (k) {
            case 0: p.abi_version += 1; break;
            case 1: p.n_r = 3; break;
            case 2: p.n_r = 5000; break;
            case 3: p.n_l = 1; break;
            case 4: p.n_l = 129; break;
            case 5: p.n_theta = p.n_l - 1; break;
            case 6: p.n_theta = 257; break;
            case 7: p.n_har = 0; break;
            case 8: p.n_orb = 0; break;
            case 9: p.n_inst = 0; break;
            case 10: p.n_m = p.n_l + 1; break;
            case 11: p.nk_max = 40; break;
            case 12: p.nk_max = 128; break;
            case 13: p.dt = 0.0; break;
            case 14: p.precision = 3; break;
            case 15: p.n_orb = 300; p.n_inst = 300; break;
        }
        ctx = (ggb_ctx *)0x1;
        EXPECT(ggb_create(&p, &ctx) == GGB_E_ARG);
        EXPECT(ctx == NULL);
        EXPECT(strlen(ggb_last_error(NULL)) > 0);
    }

    /* a valid problem reaches the HIP runtime: with no device it fails and
       frees what it had set up (or succeeds on a machine with a GPU) */
    ctx = NULL;
    int rc = ggb_create(&good, &ctx);
    EXPECT(rc == GGB_OK || rc == GGB_E_HIP || rc == GGB_E_NOMEM);
    if (rc == GGB_OK) {
        EXPECT(ctx != NULL);
        EXPECT(ggb_upload_static(ctx, NULL) == GGB_E_ARG);
        EXPECT(ggb_upload_orbitals(ctx, NULL) == GGB_E_ARG);
        EXPECT(ggb_set_field(ctx, NULL, 4) == GGB_E_ARG);
        EXPECT(ggb_run(ctx, 0, 4, NULL, NULL, NULL) != GGB_OK);
        ggb_destroy(ctx);
    } else {
        EXPECT(ctx == NULL);
        EXPECT(strlen(ggb_last_error(NULL)) > 0);
    }

    /* null contexts and arguments */
    EXPECT(ggb_upload_static(NULL, NULL) == GGB_E_ARG);
    EXPECT(ggb_upload_orbitals(NULL, NULL) == GGB_E_ARG);
    EXPECT(ggb_set_field(NULL, NULL, 1) == GGB_E_ARG);
    EXPECT(ggb_field_ring(NULL, 16) == GGB_E_ARG);
    EXPECT(ggb_set_field_slice(NULL, 0, 0, 1, NULL) == GGB_E_ARG);
    EXPECT(ggb_upload_columns(NULL, 0, 1, NULL) == GGB_E_ARG);
    EXPECT(ggb_set_active(NULL, NULL) == GGB_E_ARG);
    EXPECT(ggb_rebuild_potential(NULL) == GGB_E_ARG);
    EXPECT(ggb_run(NULL, 0, 1, NULL, NULL, NULL) == GGB_E_ARG);
    EXPECT(ggb_set_channels(NULL, 1) == GGB_E_ARG);
    EXPECT(ggb_get_channels(NULL, 1, NULL) == GGB_E_ARG);
    EXPECT(ggb_download_orbitals(NULL, NULL) == GGB_E_ARG);
    EXPECT(ggb_download_potential(NULL, NULL) == GGB_E_ARG);
    EXPECT(ggb_debug_substep(NULL, 0, 0) == GGB_E_ARG);
    EXPECT(ggb_set_graph(NULL, 8) == GGB_E_ARG);
    EXPECT(ggb_set_kinetic_variant(NULL, GGB_KIN_NARROW) == GGB_E_ARG);
    EXPECT(ggb_autotune(NULL, 1, NULL, NULL) == GGB_E_ARG);
    EXPECT(ggb_set_timing(NULL, 1) == GGB_E_ARG);
    EXPECT(ggb_get_timing(NULL, NULL, NULL) == GGB_E_ARG);
    EXPECT(strlen(ggb_last_error(NULL)) > 0);
    ggb_destroy(NULL);

    /* groups */
    ggb_group *g = (ggb_group *)0x1;
    EXPECT(ggb_group_create(NULL, 1, 8, &g) == GGB_E_ARG);
    EXPECT(ggb_group_create(NULL, 0, 8, NULL) == GGB_E_ARG);
    EXPECT(strlen(ggb_group_last_error(NULL)) > 0);
    EXPECT(ggb_group_run(NULL, NULL, 1, NULL, NULL, NULL) == GGB_E_ARG);
    ggb_group_destroy(NULL);

    /* spectrum: rejected arguments, then a valid call (no device: GGB_E_HIP) */
    double sig[64], orders[2] = {1.0, 3.0}, inten[2], power[64 * 3];
    for (int i = 0; i < 64; ++i) sig[i] = (double)(i % 7) - 3.0;
    EXPECT(ggb_spectrum(0, NULL, 64, 0.05, 0, 8.0, 1, 0.057, orders, 2, NULL, inten) == GGB_E_ARG);
    EXPECT(ggb_spectrum(0, sig, 1, 0.05, 0, 8.0, 1, 0.057, orders, 2, NULL, inten) == GGB_E_ARG);
    EXPECT(ggb_spectrum(0, sig, 64, -1.0, 0, 8.0, 1, 0.057, orders, 2, NULL, inten) == GGB_E_ARG);
    EXPECT(ggb_spectrum(0, sig, 64, 0.05, 0, 8.0, -1, 0.057, orders, 2, NULL, inten) == GGB_E_ARG);
    EXPECT(ggb_spectrum_noalloc(0, sig, 64, 0.05, 0, 8.0, 1 << 14, 0.057, orders, 2, NULL, inten) == GGB_E_ARG);
    EXPECT(ggb_spectrum_noalloc(0, sig, 64, 0.05, 0, 8.0, 1, 0.057, orders, 300, NULL, inten) == GGB_E_ARG);
    rc = ggb_spectrum(0, sig, 64, 0.05, 1, 8.0, 1, 0.057, orders, 2, power, inten);
    EXPECT(rc == GGB_OK || rc == GGB_E_HIP);
    rc = ggb_spectrum_noalloc(0, sig, 64, 0.05, 1, 8.0, 1, 0.057, orders, 2, power, inten);
    EXPECT(rc == GGB_OK || rc == GGB_E_HIP);

    /* ground state: rejected sizes */
    ggb_gs_problem gp = {64, 2, 8, 10, 1, 0, 2.0, 0.3, 2.0, 1e-11};
    int32_t it = 0, conv = 0;
    EXPECT(ggb_ground_state(NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                            &it, &conv) == GGB_E_ARG);
    gp.n_r = 2;
    EXPECT(ggb_ground_state(&gp, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                            &it, &conv) == GGB_E_ARG);

    /* propagator tables: rejected sizes, then a valid call (no device: GGB_E_HIP) */
    double rr[8], TT[64], vv[8], QQ[64], ll[64];
    int32_t kk[8], need = 0;
    for (int i = 0; i < 8; ++i) { rr[i] = 1.0 + i; vv[i] = -1.0 / rr[i]; }
    for (int i = 0; i < 64; ++i) TT[i] = (i % 9 == 0) ? -2.0 : 0.0;
    EXPECT(ggb_propagator_tables(0, 8, 1, NULL, TT, vv, 10.0, 8, QQ, ll, kk, &need) == GGB_E_ARG);
    EXPECT(ggb_propagator_tables(0, 2, 1, rr, TT, vv, 10.0, 8, QQ, ll, kk, &need) == GGB_E_ARG);
    EXPECT(ggb_propagator_tables(0, 8, 1, rr, TT, vv, 10.0, 0, QQ, ll, kk, &need) == GGB_E_ARG);
    rc = ggb_propagator_tables(0, 8, 1, rr, TT, vv, 10.0, 8, QQ, ll, kk, &need);
    EXPECT(rc == GGB_OK || rc == GGB_E_HIP);

    /* self tests reach the device */
    double out[512];
    rc = ggb_selftest_mfma(0, out);
    EXPECT(rc == GGB_OK || rc == GGB_E_HIP);
    double ms = 0.0;
    EXPECT(ggb_selftest_copy(0, 15, &ms) == GGB_E_ARG);

    printf("checks done, %d failed\n", fails);
    return fails ? 1 : 0;
}
"""


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                    reason="hipcc / ROCm clang not installed")
def test_c_abi_host_code_under_asan_ubsan(tmp_path):
    lib = tmp_path / "libggbtd_san.so"
    san = []
    for flag in ("-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined",
                 "-fno-omit-frame-pointer"):
        san += ["-Xarch_host", flag]  # host code only
    cmd = [HIPCC, "--offload-arch=gfx950", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", *san,
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), SRC, "-L", os.path.join(ROCM, "lib"), "-lrocsolver",
           "-lrocblas", "-o", str(lib)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    drv = tmp_path / "drive.c"
    drv.write_text(DRIVER)
    exe = tmp_path / "drive"
    cmd = [CLANG, "-std=c11", "-g", "-O1", "-Wall", "-Werror", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include"), str(drv),
           "-L", str(tmp_path), "-lggbtd_san", "-Wl,-rpath," + str(tmp_path), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ)
    # the HIP runtime's own start-up allocations are not this library's: no leak check
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["HIP_VISIBLE_DEVICES"] = env.get("HIP_VISIBLE_DEVICES", "")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    report = r.stdout + r.stderr
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert r.returncode == 0, report[-4000:]
    assert "checks done, 0 failed" in r.stdout
