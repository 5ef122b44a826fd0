// Drives every entry point of the CPU oracle (oracle/lvr2_oracle.h) on small seeded inputs, for a
// build with AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitize_cpu.py).  Exits
// non-zero if a call fails; the sanitizers abort on the first finding.
#include "lvr2_oracle.h"

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

int main()
{
    std::mt19937 rng(7);
    std::normal_distribution<float> g(0.0f, 1.0f);
    const int64_t n = 3000, m = 200;
    std::vector<float> p(3 * n), q(3 * m);
    for (int64_t i = 0; i < n; i++) {  // a noisy plane with a bump
        const float x = 10.0f * (float)(i % 60) / 60.0f, y = 10.0f * (float)(i / 60) / 50.0f;
        p[3 * i] = x;
        p[3 * i + 1] = y;
        p[3 * i + 2] = 0.02f * g(rng) + (x > 4 && x < 6 ? 0.5f : 0.0f);
    }
    for (int64_t i = 0; i < m; i++)
        for (int j = 0; j < 3; j++) q[3 * i + j] = p[3 * (i * 13 % n) + j] + 0.1f * g(rng);
    int bad = 0;
    for (int metric = 0; metric < 2; metric++) {
        orc_surface* s = orc_surface_create(p.data(), n, metric);
        float box[6];
        orc_surface_bbox(s, box);
        const int k = 12;
        std::vector<int32_t> idx(m * k), cnt(m);
        std::vector<float> d2(m * k);
        bad |= orc_knn(s, q.data(), m, k, idx.data(), d2.data());
        std::vector<int32_t> big(m * 4000);  // k above n: padded rows
        std::vector<float> bigd(m * 4000);
        bad |= orc_knn(s, q.data(), m, 4000, big.data(), bigd.data());
        bad |= orc_radius(s, q.data(), m, k, 0.4f, idx.data(), d2.data(), cnt.data());
        bad |= orc_knn_brute(p.data(), n, q.data(), m, k, idx.data(), d2.data());
        std::vector<float> n32(3 * n), nrm(3 * n), iq(3 * m), proj(m), eucl(m);
        std::vector<double> n64(3 * n), gap(n), i64(3 * n), dp(m), de(m);
        std::vector<int32_t> kused(n);
        const float flip[3] = {0.0f, 0.0f, 10.0f};
        for (int method : {0, 2, 3})
            bad |= orc_normals(s, 10, method, flip, nullptr, 0, n32.data(), n64.data(), kused.data(),
                               method == 3 ? gap.data() : nullptr);
        const int64_t sub[3] = {0, 5, n - 1};
        bad |= orc_normals(s, 10, 3, flip, sub, 3, n32.data(), nullptr, nullptr, nullptr);
        bad |= orc_normals(s, 10, 3, flip, nullptr, 0, n32.data(), n64.data(), kused.data(), gap.data());
        bad |= orc_interpolate(s, n32.data(), 10, nullptr, 0, nrm.data());
        bad |= orc_interpolate_f64(s, n64.data(), 10, nullptr, 0, i64.data());
        bad |= orc_interpolated_normal(s, nrm.data(), 10, q.data(), m, iq.data());
        bad |= orc_distance(s, nrm.data(), 10, q.data(), m, proj.data(), eucl.data());
        bad |= orc_distance_f64(s, i64.data(), 10, q.data(), m, dp.data(), de.data());
        // the lattice, the invalid rule, the meshes and optimizePlanarFaces
        int64_t nv = 0, nc = 0;
        bad |= orc_grid(p.data(), n, box, 0.5f, 1, &nv, &nc, nullptr, nullptr, nullptr);
        std::vector<float> vtx(3 * nv);
        std::vector<int32_t> cells(8 * nc), cidx(3 * nc);
        bad |= orc_grid(p.data(), n, box, 0.5f, 1, &nv, &nc, vtx.data(), cells.data(), cidx.data());
        std::vector<float> vp(nv), ve(nv);
        bad |= orc_distance(s, nrm.data(), 10, vtx.data(), nv, vp.data(), ve.data());
        std::vector<uint8_t> inv(nv), keep(nc);
        bad |= orc_calc_distance_values(ve.data(), nv, 0.5f, cells.data(), nc, inv.data(), keep.data());
        for (int pmc = 0; pmc < 2; pmc++) {
            int64_t mv = 0, mf = 0;
            std::vector<float> mvtx(3 * 3 * nv + 3);
            std::vector<int32_t> mfac(3 * 4 * nc + 3), fcell(4 * nc + 1);
            bad |= orc_mesh(vtx.data(), nv, cells.data(), nc, vp.data(), inv.data(), pmc, 3 * nv + 1, 4 * nc + 1, &mv,
                            &mf, mvtx.data(), mfac.data(), fcell.data());
            // too small output buffers: counts only, return 1
            int64_t mv2 = 0, mf2 = 0;
            if (orc_mesh(vtx.data(), nv, cells.data(), nc, vp.data(), nullptr, pmc, 1, 1, &mv2, &mf2, mvtx.data(),
                         mfac.data(), fcell.data()) != 1 && mf2 > 1)
                bad |= 1;
            int64_t moves = 0, nonman = 0;
            if (pmc) bad |= orc_pmc_optimize(s, 5, mvtx.data(), mv, mfac.data(), fcell.data(), mf, &moves, &nonman);
            std::printf("metric %d pmc %d: %lld vertices %lld faces\n", metric, pmc, (long long)mv, (long long)mf);
        }
        orc_surface_destroy(s);
    }
    const float cov[9] = {2, 0.5f, 0, 0.5f, 1, 0.1f, 0, 0.1f, 1e-6f};
    float ev[3], evec[9];
    orc_eig3_f32(cov, ev, evec);
    const double covd[9] = {2, 0.5, 0, 0.5, 1, 0.1, 0, 0.1, 1e-12};
    double evd[3], evecd[9];
    orc_eig3_f64(covd, evd, evecd);
    std::printf("oracle sanitize run: %s (%d threads)\n", bad ? "FAILED" : "ok", orc_num_threads());
    return bad ? 1 : 0;
}
