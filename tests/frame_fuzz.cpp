// Host fuzzing of the frame decoder (squishy_volumes_amd/csrc/frame_file.hip),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_frame_file.py: 20000 mutated copies (byte flips, truncations,
// extensions) of a valid frame, each in a heap buffer of exactly its size,
// are decoded; the sanitizers report any read outside the buffer.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "mpm_engine.h"
int main() {
    const uint64_t n = 50;
    std::vector<uint32_t> flags(n), bits(n);
    std::vector<mpm_particle_parameters> prm(n);
    std::vector<float> e(n), x(3 * n), F(9 * n), v(3 * n), C(9 * n), x0(3 * n);
    std::mt19937 rng(1);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t choices[4] = {MPM_IS_SOLID, MPM_IS_FLUID, MPM_IS_SOLID | MPM_USE_VISCOSITY, MPM_IS_SOLID | MPM_USE_SAND_ALPHA};
        flags[i] = choices[rng() % 4];
        prm[i].mass = 1.f;
        prm[i].exponent = 7;
    }
    mpm_particles_view p{n, 1.0, flags.data(), prm.data(), e.data(), bits.data(), x.data(), F.data(), v.data(), C.data(), x0.data(), nullptr};
    std::vector<int32_t> ids(3 * 7); std::vector<uint32_t> gb(7); std::vector<float> gm(7), gv(21);
    mpm_grid_view g{7, 7, ids.data(), gb.data(), gm.data(), gv.data()};
    uint64_t size = mpm_frame_file_size(&p, &g), written = 0;
    std::vector<uint8_t> good(size);
    if (mpm_frame_encode(&p, &g, good.data(), size, &written) != 0) return 1;
    int refused = 0, decoded = 0;
    for (int trial = 0; trial < 20000; ++trial) {
        std::vector<uint8_t> bad = good;
        int kind = trial % 3;
        if (kind == 0) for (int k = 0; k < 1 + (int)(rng() % 4); ++k) bad[96 + rng() % (bad.size() - 96)] = rng() & 255;
        else if (kind == 1) bad.resize(96 + rng() % (bad.size() - 96));
        else for (int k = 0; k < 1 + (int)(rng() % 64); ++k) bad.push_back(rng() & 255);
        // exact-size heap copy so that ASan sees any read past it
        uint8_t *buf = (uint8_t *)malloc(bad.size());
        memcpy(buf, bad.data(), bad.size());
        uint64_t pn = 0, gn = 0;
        int rc = mpm_frame_decode_counts(buf, bad.size(), &pn, &gn);
        if (rc == 0 && pn == n) {
            std::vector<uint32_t> f2(n), b2(n); std::vector<mpm_particle_parameters> q2(n);
            std::vector<float> e2(n), x2(3 * n), F2(9 * n), v2(3 * n), C2(9 * n), y2(3 * n);
            mpm_particles_view p2{n, 0.0, f2.data(), q2.data(), e2.data(), b2.data(), x2.data(), F2.data(), v2.data(), C2.data(), y2.data(), nullptr};
            std::vector<int32_t> i2(3 * 64); std::vector<uint32_t> c2(64); std::vector<float> m2(64), w2(3 * 64);
            mpm_grid_view g2{64, 0, i2.data(), c2.data(), m2.data(), w2.data()};
            rc = mpm_frame_decode(buf, bad.size(), &p2, &g2);
        }
        (rc == 0 ? decoded : refused)++;
        free(buf);
    }
    printf("refused %d decoded %d\n", refused, decoded);
    return 0;
}
