// Host-only check of the split backward plans (test program). For grid cloths (one sheet, and
// two disjoint sheets), two top-block heights and split configurations of the backward levels
// (parts 1, 2 or 4 per level, the last level whole), it replays the backward sweep launch by
// launch as the device runs it: each task reduces its range of the front's rows; a split task
// writes a partial, a whole task the final value; a below row whose owner was split in the
// previous launch is read as the ordered sum of its partials, any other row from X; the extra
// workgroups of a launch write the previous level's split rows as the sum of their partials.
// It checks that every read sees a value completed by an earlier launch (a workgroup never sees
// another workgroup's writes of the same launch), that every index is in range, that every
// swept row ends final, and that the result equals the unsplit sweep to rounding.
// Prints one line per case; exit 0 when all pass.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../climatenet_amd/csrc/host.hpp"

using namespace pdfc;

static void grid(int N, double x0, std::vector<double>& pos, std::vector<int32_t>& tris) {
    const int base = (int)pos.size() / 3;
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i) {
            pos.push_back(x0 + (double)i / N);
            pos.push_back((double)j / N);
            pos.push_back(0.3);
        }
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            int a = base + j * (N + 1) + i, b = a + 1, c = a + N + 1, d = c + 1;
            int t[6] = {a, b, d, a, d, c};
            tris.insert(tris.end(), t, t + 6);
        }
}

static int failures = 0;
#define CHECK(cond, ...)                      \
    do {                                      \
        if (!(cond)) {                        \
            if (failures < 20) {              \
                printf("FAIL: " __VA_ARGS__); \
                printf("\n");                 \
            }                                 \
            ++failures;                       \
        }                                     \
    } while (0)

// x_s = F_s^T [y_s ; -x_R] over rows [i0, i1) of the front (one component).
static double task_sum(const Factor& F, int s, int j, int i0, int i1, const std::vector<double>& y,
                       const std::vector<double>& g) {
    const int ns = F.ns[s];
    double acc = 0.;
    for (int i = i0; i < i1; ++i) acc += F.FT[F.foff[s] + (size_t)i * ns + j] * (i < ns ? y[F.c0[s] + i] : g[i - ns]);
    return acc;
}

static double max_rel(const std::vector<double>& a, const std::vector<double>& b) {
    double m = 0., s = 0.;
    for (size_t i = 0; i < a.size(); ++i) {
        m = std::max(m, std::fabs(a[i] - b[i]));
        s = std::max(s, std::fabs(b[i]));
    }
    return m / std::max(s, 1e-300);
}

int main() {
    struct Case {
        const char* name;
        std::vector<int> sheets;
        int leaf;
    };
    const Case cases[] = {{"60x60 sheet", {60}, 32}, {"40x40 + 25x25 sheets", {40, 25}, 32}};
    std::mt19937_64 rng(5);
    std::uniform_real_distribution<double> U(-1., 1.);
    for (const Case& cs : cases) {
        std::vector<double> pos;
        std::vector<int32_t> tris;
        double x0 = 0.;
        for (int N : cs.sheets) {
            grid(N, x0, pos, tris);
            x0 += 2.;
        }
        pdfc_scene sc{};
        sc.n_vertices = (int64_t)pos.size() / 3;
        sc.rest_positions = pos.data();
        sc.n_triangles = (int64_t)tris.size() / 3;
        sc.triangles = tris.data();
        sc.area_density = 0.1;
        sc.bending = 5e-4;
        sc.stretching = 2e2;
        Model M;
        build_model(sc, M);
        for (int cross : {0, 1}) {
        // cross: separator levels below a 1 MiB top block merged in pairs, as the library orders them
        Factor F;
        factorize(M, 0.005 * 0.005, F, cs.leaf, cross ? 1048576. : 0.);
        const int n = F.n;
        for (int drop : {0, 3}) {
            const int top_level = F.n_levels - drop;
            SweepPlan fwd, bwd;
            plan_sweeps(F, 64, top_level, fwd, bwd);
            const int L = (int)bwd.level_ptr.size() - 1;
            std::vector<double> y(n), xtop(n, 0.);
            for (auto& v : y) v = U(rng);
            std::vector<char> swept(n, 0);
            for (size_t s = 0; s < F.ns.size(); ++s)
                if (F.level[s] < L)
                    for (int k = 0; k < F.ns[s]; ++k) swept[F.c0[s] + k] = 1;
            for (int r = 0; r < n; ++r)
                if (!swept[r]) xtop[r] = U(rng);
            // unsplit reference
            std::vector<double> xref = xtop;
            for (int l = 0; l < L; ++l)
                for (int t = bwd.level_ptr[l]; t < bwd.level_ptr[l + 1]; ++t) {
                    const int s = bwd.task_snode[t], j0 = bwd.task_row0[t];
                    std::vector<double> g(F.nb[s]);
                    for (int i = 0; i < F.nb[s]; ++i) g[i] = -xref[F.rows[F.roff[s] + i]];
                    for (int j = j0; j < std::min(F.ns[s], j0 + 64); ++j)
                        xref[F.c0[s] + j] = task_sum(F, s, j, 0, F.ns[s] + F.nb[s], y, g);
                }
            // split configurations: every level 2, every level 4, alternating, only the first
            std::vector<std::vector<int>> configs;
            if (L >= 2) {
                std::vector<int> a(L, 2), b(L, 4), c(L, 1), d(L, 1);
                for (int l = 0; l + 1 < L; l += 2) c[l] = 4;
                d[0] = 2;
                a[L - 1] = b[L - 1] = c[L - 1] = 1;
                configs = {a, b, c, d};
            }
            int checked = 0;
            for (const auto& split : configs) {
                SplitPlan sp;
                plan_backward_split(F, bwd, split, sp);
                CHECK((int)sp.tasks.level_ptr.size() == L + 1, "level count");
                std::vector<double> X = xtop;
                std::vector<std::vector<double>> XP(4, std::vector<double>(n, NAN));
                std::vector<int> final_at(n, INT_MAX), xp_at(4 * (size_t)n, -1);
                for (int r = 0; r < n; ++r)
                    if (!swept[r]) final_at[r] = -1;
                for (int l = 0; l < L; ++l) {
                    std::vector<std::pair<int, double>> direct;   // this launch's writes, applied at its end
                    std::vector<std::pair<int, double>> partial;  // (p * n + r, value)
                    for (int t = sp.tasks.level_ptr[l]; t < sp.tasks.level_ptr[l + 1]; ++t) {
                        const int s = sp.tasks.task_snode[t], j0 = sp.tasks.task_row0[t];
                        const auto& pt = sp.part[t];
                        const int p = pt[0], P = pt[1], i0 = pt[2], i1 = pt[3];
                        const int f = F.ns[s] + F.nb[s];
                        CHECK(0 <= i0 && i0 <= i1 && i1 <= f && 0 <= p && p < P && P <= kMaxSplit, "task range");
                        CHECK(P == split[l], "parts of level %d", l);
                        std::vector<double> g(F.nb[s], 0.);
                        for (int i = std::max(i0, F.ns[s]); i < i1; ++i) {
                            const int k = i - F.ns[s];
                            const int r = F.rows[F.roff[s] + k];
                            const int np = sp.rparts[F.roff[s] + k];
                            CHECK(r >= 0 && r < n && np >= 0 && np <= kMaxSplit, "row index");
                            if (np > 1) {
                                double v = 0.;
                                for (int q = 0; q < np; ++q) {
                                    CHECK(xp_at[(size_t)q * n + r] == l - 1, "partial %d of row %d not from launch %d", q, r, l - 1);
                                    v = q == 0 ? XP[q][r] : v + XP[q][r];
                                }
                                g[k] = -v;
                            } else {
                                CHECK(final_at[r] < l, "row %d read at launch %d before it is final", r, l);
                                g[k] = -X[r];
                            }
                        }
                        for (int j = j0; j < std::min(F.ns[s], j0 + 64); ++j) {
                            const double v = task_sum(F, s, j, i0, i1, y, g);
                            if (P > 1)
                                partial.push_back({p * n + F.c0[s] + j, v});
                            else
                                direct.push_back({F.c0[s] + j, v});
                        }
                    }
                    for (int r : sp.fin_rows[l]) {
                        CHECK(r >= 0 && r < n, "finish row");
                        const int np = sp.fin_parts[l];
                        double v = 0.;
                        for (int q = 0; q < np; ++q) {
                            CHECK(xp_at[(size_t)q * n + r] == l - 1, "finish: partial %d of row %d", q, r);
                            v = q == 0 ? XP[q][r] : v + XP[q][r];
                        }
                        direct.push_back({r, v});
                    }
                    for (auto& w : direct) {
                        X[w.first] = w.second;
                        final_at[w.first] = l;
                    }
                    for (auto& w : partial) {
                        XP[w.first / n][w.first % n] = w.second;
                        xp_at[w.first] = l;
                    }
                }
                for (int r = 0; r < n; ++r) CHECK(final_at[r] < L, "row %d never final", r);
                const double err = max_rel(X, xref);
                CHECK(err < 1e-12, "split result differs from the unsplit sweep by %.3e", err);
                ++checked;
            }
            printf("%s%s: top at %d of %d levels, %d swept, %d split configurations replayed\n", cs.name,
                   cross ? " (crosses)" : "", top_level, F.n_levels, L, checked);
        }
        }
    }
    printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
