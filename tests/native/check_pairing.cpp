// Host-only check of the product's self-contact pairing and level ordering (test program).
// Reads a mesh, a candidate list and obstacle-contact seed vertices from a text file:
//   n T / n lines "x y z" / T lines "a b c" / nc / nc lines "v t b0 b1 b2" / ns / seeds
// runs the flat-workspace pairing, the general pairing (must agree) and the level ordering,
// and prints "pair v partner level" per self contact, then "levels L" and "fallback f".
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <set>
#include <vector>

#include "../../climatenet_amd/csrc/host.hpp"

using namespace pdfc;

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::ifstream in(argv[1]);
    int n = 0, T = 0;
    in >> n >> T;
    std::vector<double> pos(3 * (size_t)n);
    for (auto& x : pos) in >> x;
    std::vector<int32_t> tris(3 * (size_t)T);
    for (auto& t : tris) in >> t;
    size_t nc = 0;
    in >> nc;
    std::vector<SelfCandidate> cand(nc);
    for (auto& c : cand) {
        in >> c.v >> c.t >> c.bary[0] >> c.bary[1] >> c.bary[2];
        c.normal[0] = c.normal[1] = 0.;
        c.normal[2] = 1.;
    }
    size_t ns = 0;
    in >> ns;
    std::set<int> seeds;
    for (size_t k = 0; k < ns; ++k) {
        int s;
        in >> s;
        seeds.insert(s);
    }
    pdfc_scene sc{};
    sc.n_vertices = (size_t)n;
    sc.rest_positions = pos.data();
    sc.n_triangles = (size_t)T;
    sc.triangles = tris.data();
    sc.area_density = 0.1;
    sc.bending = 5e-4;
    sc.stretching = 200.;
    Model M;
    build_model(sc, M);
    SelfWorkspace W;
    std::vector<SelfPair> fast, general;
    // twice on the same workspace: the second call starts from the first call's leftovers
    pair_self_contacts(W, M, cand.data(), cand.size(), fast);
    pair_self_contacts(W, M, cand.data(), cand.size(), fast);
    pair_self_contacts(M, cand.data(), cand.size(), general);
    if (fast.size() != general.size()) {
        printf("mismatch sizes %zu %zu\n", fast.size(), general.size());
        return 1;
    }
    for (size_t k = 0; k < fast.size(); ++k)
        if (fast[k].v != general[k].v || fast[k].partner != general[k].partner || fast[k].best != general[k].best) {
            printf("mismatch at %zu\n", k);
            return 1;
        }
    // most registered partners of one vertex (above SelfWorkspace::kRegCap: the general form ran)
    std::vector<int> reg(n, 0);
    for (const auto& p : fast) {
        ++reg[p.v];
        ++reg[p.partner];
    }
    int overflow = 0;
    for (int v = 0; v < n; ++v) overflow = std::max(overflow, reg[v]);
    std::vector<int> order, level_ptr;
    int L = order_self_contacts(W, [&](int v) { return seeds.count(v) > 0; }, fast, order, level_ptr);
    // and again, on the workspace as the first ordering left it
    L = order_self_contacts(W, [&](int v) { return seeds.count(v) > 0; }, fast, order, level_ptr);
    for (const auto& p : fast) printf("pair %d %d %d\n", p.v, p.partner, p.level);
    for (int l = 0; l < L; ++l)
        for (int j = level_ptr[l]; j < level_ptr[l + 1]; ++j)
            if (fast[order[j]].level != l) {
                printf("order/level mismatch\n");
                return 1;
            }
    printf("levels %d\nmaxreg %d\n", L, overflow);
    return 0;
}
