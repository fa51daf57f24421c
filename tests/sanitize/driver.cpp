// Host-side sanitizer driver (AddressSanitizer + UndefinedBehaviorSanitizer): the
// product's host code that needs no GPU -- the seeded fBm generator, the volume
// builder (bricks, apron pools, mips, conservative grids), the .vbx writer and
// readers and the file-set loader with its missing-level derivation -- and the CPU
// oracle (every stage of two frames at path length 3, and every tracking mode).
// Built by tests/sanitize/Makefile and run by tests/test_sanitizers.py; any
// sanitizer report aborts with a non-zero status.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vrestir.h"
#include "../../oracle/vrestir_oracle.h"
#include "../../volumetricrestirrelease_amd/csrc/vr_volume.h"

static int check(bool ok, const char* what) {
    if (!ok) std::fprintf(stderr, "sanitize driver: %s failed\n", what);
    return ok ? 0 : 1;
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    const int n = 32;
    std::vector<float> dens((size_t)n * n * n);
    int bad = check(vr::synthFbm(3, n, n, n, 4, 3.f, 2.f, 0.5f, 1, -0.1f, 1.4f, dens.data()) == 0, "synthFbm");
    float M[16] = {2.f / n, 0, 0, 0, 0, 2.f / n, 0, 0, 0, 0, 2.f / n, 0, -1, -1, -1, 1};
    vrestir_volume_desc d;
    std::memset(&d, 0, sizeof(d));
    d.nx = d.ny = d.nz = n;
    d.density = dens.data();
    d.numMips = 3;
    d.brickFormat = VRESTIR_BRICK_FP16;
    std::memcpy(d.indexToWorld, M, sizeof(M));
    for (int i = 0; i < 3; i++) { d.sigmaA[i] = 2.f; d.sigmaS[i] = 18.f; }
    d.g = 0.3f;
    d.densityScale = 1.f;
    vr::HostVolume hv;
    std::string err;
    bad += check(vr::buildVolume(d, hv, err) == 0, "buildVolume");
    // a .vbx file set with mip 0 and mip 1 only: the loader derives the rest
    const std::string prefix = dir + "/sanitize_cloud";
    bad += check(vrestir_write_vbx((prefix + "_mip0.vbx").c_str(), dens.data(), n, n, n, M) == 0, "write mip0");
    std::vector<float> half((size_t)(n / 2) * (n / 2) * (n / 2), 0.25f);
    float M1[16];
    std::memcpy(M1, M, sizeof(M));
    M1[0] *= 2; M1[5] *= 2; M1[10] *= 2;
    bad += check(vrestir_write_vbx((prefix + "_mip1.vbx").c_str(), half.data(), n / 2, n / 2, n / 2, M1) == 0,
                 "write mip1");
    vr::HostVolume fv;
    float i2w[16];
    bad += check(vr::buildVolumeFromFiles(prefix, 3, true, fv, i2w, err) == 0, "buildVolumeFromFiles");
    bad += check(fv.numMips == 3 && fv.slots[2].valid && fv.slots[10].valid, "derived levels");
    int nx = 0, ny = 0, nz = 0;
    std::vector<float> back((size_t)n * n * n);
    bad += check(vrestir_read_vbx((prefix + "_mip0.vbx").c_str(), back.data(), &nx, &ny, &nz, nullptr) == 0 &&
                     nx == n && back == dens,
                 "read_vbx round trip");
    vr::VbxBricks vb;
    bad += check(vr::readVbxBricks((prefix + "_mip1.vbx").c_str(), vb) == 0 && vb.nx == n / 2, "readVbxBricks");
    std::remove((prefix + "_mip0.vbx").c_str());
    std::remove((prefix + "_mip1.vbx").c_str());

    // the oracle: two frames of every stage, then every tracking mode
    void* o = orc_create();
    vrestir_params p;
    vrestir_default_params_oracle(&p);
    p.maxBounces = 3;
    p.initialM = 5;
    p.spatialReuseRounds = 2;
    d.brickFormat = VRESTIR_BRICK_FP32;
    bad += check(orc_set_volume(o, &d) == 0, "orc_set_volume");
    vrestir_point_light L[2] = {{{1.6f, 2.2f, 1.4f}, {9.f, 8.5f, 8.f}}, {{-2.f, 1.f, 0.5f}, {3.f, 3.f, 3.f}}};
    const float env[3] = {0.35f, 0.42f, 0.55f};
    bad += check(orc_set_lights(o, L, 2, env) == 0, "orc_set_lights");
    bad += check(orc_set_params(o, &p) == 0, "orc_set_params");
    const int W = 20, H = 14;
    std::vector<float> img((size_t)W * H * 4);
    for (int f = 0; f < 2; f++) {
        vrestir_camera c = {{0.35f + 0.1f * f, 0.45f, 3.1f}, {0, 0, 0}, {0, 1, 0}, 0.73f, 0.1f, 1000.f};
        bad += check(orc_set_camera(o, &c, W, H) == 0, "orc_set_camera");
        bad += check(orc_execute(o, W, H, img.data(), 0, 0, 2) == 0, "orc_execute");
    }
    for (float v : img) bad += check(std::isfinite(v), "finite frame");
    const int nr = 64;
    std::vector<float> rays((size_t)nr * 8), out((size_t)nr * 8);
    std::vector<uint32_t> seeds((size_t)nr * 3);
    for (int i = 0; i < nr; i++) {
        const float a = 0.1f * i, b = 0.07f * i;
        float* r = &rays[(size_t)i * 8];
        r[0] = 0.5f * std::sin(a); r[1] = 0.4f * std::cos(b); r[2] = 0.3f * std::sin(a + b);
        r[3] = std::cos(a) * std::cos(b); r[4] = std::sin(b); r[5] = std::sin(a) * std::cos(b);
        if (i % 8 == 0) { r[3] = 1.f; r[4] = r[5] = 0.f; }  // axis-parallel
        r[6] = i % 2 ? 3.4e38f : 2.f;
        seeds[3 * i] = i; seeds[3 * i + 1] = 7; seeds[3 * i + 2] = 3;
    }
    for (int mode = 0; mode <= VRESTIR_TRACK_VISITED_BRICKS; mode++) {
        if (mode == VRESTIR_TRACK_ENV_SAMPLE || mode == VRESTIR_TRACK_EMISSION || mode == VRESTIR_TRACK_VELOCITY)
            continue;  // need an environment map, an emission or a velocity grid
        const int slot = (mode == VRESTIR_TRACK_SAMPLE_ANALYTIC || mode == VRESTIR_TRACK_SAMPLE_CELL_BY_DENSITY) ? 9 : 0;
        bad += check(orc_track_batch(o, mode, slot, slot == 0, 1.f, 4, nr, rays.data(), seeds.data(), out.data()) == 0,
                     "orc_track_batch");
    }
    orc_destroy(o);
    if (bad) return 1;
    std::printf("sanitize driver: ok\n");
    return 0;
}
