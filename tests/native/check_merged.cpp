// Host-only check of the merged-launch plans (test program). For grid cloths (one sheet, and
// two disjoint sheets, whose factor is a forest), several top-block heights and every
// partition of each sweep into at most three ranges of consecutive levels, it checks that
// pdfc::plan_merged lists every task of a range exactly once, in well-formed per-group,
// per-level lists, and that every task's inputs are complete when the merged launch runs it:
//   forward:  each child supernode lies below the range (an earlier launch) or in the task's
//             group at an earlier level of the range (the same workgroup, before a barrier);
//   backward: each below row's owner lies above the range (the top block or an earlier
//             launch) or in the task's group at an earlier level of the range.
// A workgroup sees only its own writes between the barriers of a launch, so these two
// conditions are what the merged kernels need. Prints one line per case; exit 0 when all pass.
#include <cstdio>
#include <cstdlib>
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
#define CHECK(cond, ...)                  \
    do {                                  \
        if (!(cond)) {                    \
            if (failures < 20) {          \
                printf("FAIL: " __VA_ARGS__); \
                printf("\n");             \
            }                             \
            ++failures;                   \
        }                                 \
    } while (0)

// One sweep direction, one range [first, first + nl) of sweep levels.
static void check_range(const Factor& F, const SweepPlan& P, bool forward, int first, int nl,
                        const std::vector<int>& owner) {
    const int L = (int)P.level_ptr.size() - 1;
    MergedGroups mg;
    plan_merged(F, P, forward, first, nl, mg);
    const int G = mg.groups;
    CHECK((int)mg.ptr.size() == G * (nl + 1), "ptr size");
    CHECK((int)mg.tmax.size() == nl, "tmax size");
    const int t0 = P.level_ptr[first], t1 = P.level_ptr[first + nl];
    CHECK((int)mg.task.size() == t1 - t0, "task count %zu vs %d", mg.task.size(), t1 - t0);
    std::vector<int> seen(P.task_snode.size(), 0);
    // group and range level of every supernode that has a task in the range
    std::vector<int> sgroup(F.ns.size(), -1), sk(F.ns.size(), -1);
    for (int g = 0; g < G; ++g) {
        const int* gp = mg.ptr.data() + (size_t)g * (nl + 1);
        CHECK(g == 0 ? gp[0] == 0 : gp[0] == mg.ptr[(size_t)g * (nl + 1) - 1], "groups not contiguous");
        for (int k = 0; k < nl; ++k) {
            CHECK(gp[k] <= gp[k + 1], "ptr not monotone");
            CHECK(gp[k + 1] - gp[k] <= mg.tmax[k], "tmax too small");
            for (int q = gp[k]; q < gp[k + 1]; ++q) {
                const int t = mg.task[q];
                CHECK(t >= P.level_ptr[first + k] && t < P.level_ptr[first + k + 1], "task %d outside its level", t);
                if (t < 0 || t >= (int)seen.size()) continue;
                ++seen[t];
                const int s = P.task_snode[t];
                CHECK(sgroup[s] < 0 || (sgroup[s] == g && sk[s] == k), "supernode %d in two groups", s);
                sgroup[s] = g;
                sk[s] = k;
            }
        }
    }
    CHECK(mg.ptr.empty() || mg.ptr.back() == (int)mg.task.size(), "last offset");
    for (int t = t0; t < t1; ++t) CHECK(seen[t] == 1, "task %d listed %d times", t, seen[t]);
    // inputs complete before each task
    auto level_of = [&](int l) { return forward ? l : L - 1 - l; };
    const int lo = std::min(level_of(first), level_of(first + nl - 1));
    const int hi = std::max(level_of(first), level_of(first + nl - 1));
    for (int t = t0; t < t1; ++t) {
        const int s = P.task_snode[t];
        const int g = sgroup[s], k = sk[s];
        if (forward) {
            for (int64_t q = F.choff[s]; q < F.choff[s + 1]; ++q) {
                const int c = F.children[q];
                if (F.level[c] < lo) continue;
                CHECK(sgroup[c] == g && sk[c] >= 0 && sk[c] < k, "forward: child %d of %d not done before it", c, s);
            }
        } else {
            for (int i = 0; i < F.nb[s]; ++i) {
                const int a = owner[F.rows[F.roff[s] + i]];
                if (F.level[a] > hi) continue;
                CHECK(sgroup[a] == g && sk[a] >= 0 && sk[a] < k, "backward: row owner %d of %d not done before it", a, s);
            }
        }
    }
}

int main() {
    struct Case {
        const char* name;
        std::vector<int> sheets;
        int leaf;
    };
    const Case cases[] = {{"60x60 sheet", {60}, 32}, {"120x120 sheet", {120}, 64}, {"40x40 + 25x25 sheets", {40, 25}, 32}};
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
        std::vector<int> owner(F.n);
        for (size_t s = 0; s < F.ns.size(); ++s)
            for (int j = 0; j < F.ns[s]; ++j) owner[F.c0[s] + j] = (int)s;
        for (int drop : {0, 2, 4}) {
            const int top_level = F.n_levels - drop;
            SweepPlan fwd, bwd;
            plan_sweeps(F, 64, top_level, fwd, bwd);  // 64-row tasks, as the library plans them
            const int L = (int)fwd.level_ptr.size() - 1;
            int ranges = 0;
            for (int c1 = 1; c1 <= L; ++c1)
                for (int c2 = c1; c2 <= L; ++c2) {
                    std::vector<int> parts{c1};
                    if (c2 > c1) parts.push_back(c2 - c1);
                    if (L > c2) parts.push_back(L - c2);
                    for (int dir = 0; dir < 2; ++dir) {
                        int first = 0;
                        for (int nl : parts) {
                            check_range(F, dir == 0 ? fwd : bwd, dir == 0, first, nl, owner);
                            first += nl;
                            ++ranges;
                        }
                    }
                }
            printf("%s%s: %zu supernodes, %d levels, top at %d: %d swept levels, %d ranges checked\n", cs.name,
                   cross ? " (crosses)" : "", F.ns.size(), F.n_levels, top_level, L, ranges);
        }
        }
    }
    printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
