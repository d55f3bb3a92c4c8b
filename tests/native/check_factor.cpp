// Host-only check of the supernodal factor (test program): builds an N x N grid
// cloth, factors K, applies the forward/backward sweeps exactly as the device
// kernels do (dense F_s products, extend-add of children's update vectors) and
// prints the relative residual ||K x - b|| / ||b|| and factor statistics.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../climatenet_amd/csrc/host.hpp"

using namespace pdfc;

int main(int argc, char** argv) {
    int N = argc > 1 ? atoi(argv[1]) : 20;
    int leaf = argc > 2 ? atoi(argv[2]) : 64;
    double cross_mb = argc > 3 ? atof(argv[3]) : 0.;  // > 0: crosses below a top block of this budget
    double side = 1.0, dt = 0.005;
    int nv = (N + 1) * (N + 1);
    std::vector<double> pos(3 * nv);
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i) {
            int v = j * (N + 1) + i;
            pos[3 * v] = (float)(-side / 2 + i * side / N);
            pos[3 * v + 1] = (float)(-side / 2 + j * side / N);
            pos[3 * v + 2] = 0.3f;
        }
    std::vector<int32_t> tris;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            int a = j * (N + 1) + i, b = a + 1, c = a + N + 1, d = c + 1;
            int t[6] = {a, b, d, a, d, c};
            tris.insert(tris.end(), t, t + 6);
        }
    pdfc_scene sc{};
    sc.n_vertices = nv; sc.rest_positions = pos.data(); sc.n_triangles = tris.size() / 3; sc.triangles = tris.data();
    sc.area_density = 0.1; sc.bending = 5e-4; sc.stretching = 2e2;
    Model M;
    build_model(sc, M);
    Factor F;
    factorize(M, dt * dt, F, leaf, cross_mb * 1048576.);
    const int n = nv, S = (int)F.ns.size();
    std::mt19937_64 rng(1);
    std::uniform_real_distribution<double> U(-1, 1);
    std::vector<double> b(3 * n), x(3 * n), y(3 * n), u(3 * F.uoff[S]);
    for (auto& t : b) t = U(rng);
    // permuted planar copies
    std::vector<double> bp(3 * n), yp(3 * n, 0.), xp(3 * n, 0.);
    for (int k = 0; k < n; ++k) for (int c = 0; c < 3; ++c) bp[c * n + k] = b[3 * F.perm[k] + c];
    std::vector<std::vector<int>> by_level(F.n_levels);
    for (int s = 0; s < S; ++s) by_level[F.level[s]].push_back(s);
    for (int l = 0; l < F.n_levels; ++l)
        for (int s : by_level[l]) {
            int ns = F.ns[s], nb = F.nb[s], f = ns + nb, c0 = F.c0[s];
            std::vector<double> bt(3 * ns), ua(3 * nb, 0.);
            for (int t = 0; t < ns; ++t) for (int c = 0; c < 3; ++c) bt[3 * t + c] = bp[c * n + c0 + t];
            for (int64_t q = F.choff[s]; q < F.choff[s + 1]; ++q) {
                int ch = F.children[q];
                for (int t = 0; t < F.nb[ch]; ++t) {
                    int idx = F.relmap[F.mapoff[ch] + t];
                    for (int c = 0; c < 3; ++c) {
                        double val = u[3 * (F.uoff[ch] + t) + c];
                        if (idx < ns) bt[3 * idx + c] -= val; else ua[3 * (idx - ns) + c] += val;
                    }
                }
            }
            for (int r = 0; r < f; ++r) {
                double acc[3] = {0, 0, 0};
                for (int j = 0; j < ns; ++j) for (int c = 0; c < 3; ++c) acc[c] += F.F[F.foff[s] + (size_t)j * f + r] * bt[3 * j + c];
                for (int c = 0; c < 3; ++c) {
                    if (r < ns) yp[c * n + c0 + r] = acc[c];
                    else u[3 * (F.uoff[s] + r - ns) + c] = acc[c] + ua[3 * (r - ns) + c];
                }
            }
        }
    for (int l = F.n_levels - 1; l >= 0; --l)
        for (int s : by_level[l]) {
            int ns = F.ns[s], nb = F.nb[s], f = ns + nb, c0 = F.c0[s];
            std::vector<double> g(3 * f);
            for (int t = 0; t < ns; ++t) for (int c = 0; c < 3; ++c) g[3 * t + c] = yp[c * n + c0 + t];
            for (int t = 0; t < nb; ++t) for (int c = 0; c < 3; ++c) g[3 * (ns + t) + c] = -xp[c * n + F.rows[F.roff[s] + t]];
            for (int j = 0; j < ns; ++j) {
                double acc[3] = {0, 0, 0};
                for (int i = 0; i < f; ++i) for (int c = 0; c < 3; ++c) acc[c] += F.FT[F.foff[s] + (size_t)i * ns + j] * g[3 * i + c];
                for (int c = 0; c < 3; ++c) xp[c * n + c0 + j] = acc[c];
            }
        }
    for (int k = 0; k < n; ++k) for (int c = 0; c < 3; ++c) x[3 * F.perm[k] + c] = xp[c * n + k];
    // residual with K = M + dt^2 A^T A
    double rn = 0, bn = 0;
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            double s = M.mass[i] * x[3 * i + c];
            for (int64_t p = M.ata_ptr[i]; p < M.ata_ptr[i + 1]; ++p) s += dt * dt * M.ata_val[p] * x[3 * M.ata_idx[p] + c];
            rn += (s - b[3 * i + c]) * (s - b[3 * i + c]);
            bn += b[3 * i + c] * b[3 * i + c];
        }
    int maxlvl_tasks = 0;
    for (double mb : {16., 64., 256.}) {
        TopPlan top;
        plan_top(F, mb * 1048576., 8, top);
        printf("top budget %.0f MB: level %d of %d, m=%d, contributions=%zu\n", mb, top.level, F.n_levels, top.m,
               top.cidx.size());
    }
    if (cross_mb > 0.) {
        TopPlan top;
        plan_top(F, F.top_bytes, 8, top);
        printf("chosen: top level %d of %d, m=%d\n", top.level, F.n_levels, top.m);
    }
    printf("n=%d supernodes=%d levels=%d nnzL=%lld nnzF=%lld max_ns=%d max_front=%d factor_s=%.3f resid=%.3e\n",
           n, S, F.n_levels, (long long)F.nnz_L, (long long)F.foff[S], F.max_ns, F.max_front, F.seconds, std::sqrt(rn / bn));
    (void)maxlvl_tasks;
    return std::sqrt(rn / bn) < 1e-10 ? 0 : 1;
}
