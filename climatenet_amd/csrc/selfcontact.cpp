// Host half of self-collision handling, run once per step on the device's
// filtered candidate list:
//   greedy node-node pairing   src/geometry/geometry_util.cpp:94-206
//   computation levels (BFS)   src/physics/SimulationFrictionEstimation.cpp:161-293
// Both are order-dependent and sequential in the reference.
#include <algorithm>
#include <cmath>

#include "host.hpp"

namespace pdfc {
namespace {

// Position in the face of the first face vertex that appears in `list`, 3 if none
// (isTriangleVertexIndexIn, geometry_util.cpp:213-227).
int first_face_vertex_in(const std::array<int, 3>& face, const std::vector<int>& list) {
    for (int k = 0; k < 3; ++k)
        if (std::find(list.begin(), list.end(), face[k]) != list.end()) return k;
    return 3;
}

}  // namespace

void pair_self_contacts(const Model& M, const SelfCandidate* cand, size_t ncand, std::vector<SelfPair>& out) {
    out.clear();
    // registered partners of each vertex touched in this step; the lists persist across
    // calls and only the touched ones are cleared
    static thread_local std::vector<std::vector<int>> registered;
    static thread_local std::vector<int> touched;
    if ((int)registered.size() < M.n) registered.resize(M.n);
    for (int v : touched) registered[v].clear();
    touched.clear();
    for (size_t k = 0; k < ncand; ++k) {
        const SelfCandidate& c = cand[k];
        const int v = c.v;
        const std::array<int, 3>& face = M.tris[c.t];
        // the device list holds only candidates with v outside the face and no zero coordinate
        if (first_face_vertex_in(face, registered[v]) < 3) continue;
        bool ok[3] = {true, true, true};
        for (int u : M.neighbors[v]) {
            int j = first_face_vertex_in(face, registered[u]);
            if (j < 3) ok[j] = false;
        }
        int best = 3;
        double bv = 0.;
        for (int j = 0; j < 3; ++j)
            if (ok[j] && c.bary[j] > bv) {
                best = j;
                bv = c.bary[j];
            }
        if (best >= 3) continue;
        const int partner = face[best];
        if (registered[partner].empty()) touched.push_back(partner);
        if (registered[v].empty()) touched.push_back(v);
        registered[partner].push_back(v);
        registered[v].push_back(partner);
        SelfPair p;
        p.v = v;
        p.partner = partner;
        p.face = face;
        p.best = best;
        p.normal = {c.normal[0], c.normal[1], c.normal[2]};
        p.level = -1;
        out.push_back(p);
    }
}

int order_self_contacts(const std::function<bool(int)>& is_seed, std::vector<SelfPair>& pairs,
                        std::vector<std::vector<int>>& levels) {
    levels.clear();
    const int ns = (int)pairs.size();
    if (ns == 0) return 0;
    // The self-contact graph over the vertices that appear in it, numbered in ascending
    // vertex order (the reference's scans over all vertices visit them in that order and
    // skip every other vertex: no adjacency). Per-step work is O(pairs log pairs).
    std::vector<int> verts;
    verts.reserve(2 * (size_t)ns);
    for (const auto& p : pairs) {
        verts.push_back(p.v);
        verts.push_back(p.partner);
    }
    std::sort(verts.begin(), verts.end());
    verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
    const int nv = (int)verts.size();
    auto id = [&](int v) { return (int)(std::lower_bound(verts.begin(), verts.end(), v) - verts.begin()); };
    std::vector<int> pv(ns), pp(ns);
    std::vector<int> ptr(nv + 1, 0);
    for (int s = 0; s < ns; ++s) {
        pv[s] = id(pairs[s].v);
        pp[s] = id(pairs[s].partner);
        ptr[pv[s] + 1]++;
        ptr[pp[s] + 1]++;
    }
    for (int i = 0; i < nv; ++i) ptr[i + 1] += ptr[i];
    std::vector<std::pair<int, int>> adj(ptr[nv]);
    {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int s = 0; s < ns; ++s) {
            adj[fill[pv[s]]++] = {pp[s], s};
            adj[fill[pp[s]]++] = {pv[s], s};
        }
    }
    std::vector<char> visited(ns, 0);
    std::vector<int> last, next;
    auto bfs = [&](std::vector<int> start) {
        size_t lvl = 0;
        last.swap(start);
        while (!last.empty()) {
            if (levels.size() <= lvl) levels.emplace_back();
            next.clear();
            for (int u : last)
                for (int q = ptr[u]; q < ptr[u + 1]; ++q) {
                    int s = adj[q].second;
                    if (visited[s]) continue;
                    visited[s] = 1;
                    next.push_back(adj[q].first);
                    levels[lvl].push_back(s);
                }
            last.swap(next);
            ++lvl;
        }
    };
    auto all_visited = [&](int u) {
        for (int q = ptr[u]; q < ptr[u + 1]; ++q)
            if (!visited[adj[q].second]) return false;
        return true;
    };
    // the first search starts from every obstacle-contact vertex at once; those outside the
    // graph add nothing, and the order of a frontier changes only the order inside a level
    std::vector<int> start;
    for (int u = 0; u < nv; ++u)
        if (is_seed(verts[u])) start.push_back(u);
    bfs(start);
    for (int u = 0; u < nv; ++u)
        if (ptr[u + 1] - ptr[u] == 1 && !all_visited(u)) bfs({u});
    for (int u = 0; u < nv; ++u)
        if (ptr[u + 1] - ptr[u] > 1 && !all_visited(u)) bfs({u});
    // drop empty trailing levels produced by the last BFS frontier
    while (!levels.empty() && levels.back().empty()) levels.pop_back();
    for (size_t l = 0; l < levels.size(); ++l)
        for (int s : levels[l]) pairs[s].level = (int)l;
    return (int)levels.size();
}

}  // namespace pdfc

namespace pdfc {

void host_contact_basis(const double nr[3], double B[9]) {
    auto cross = [](V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
    auto nrm = [](V3 a) {
        double z = (a.x * a.x + a.y * a.y) + a.z * a.z;
        if (z > 0.) {
            double q = std::sqrt(z);
            return V3{a.x / q, a.y / q, a.z / q};
        }
        return a;
    };
    const V3 n{nr[0], nr[1], nr[2]};
    V3 t = cross(n, V3{1., 0., 0.});
    if (std::sqrt((t.x * t.x + t.y * t.y) + t.z * t.z) < 1.e-15) t = cross(n, V3{0., 1., 0.});
    t = nrm(t);
    V3 b = nrm(cross(n, t));
    B[0] = n.x; B[1] = n.y; B[2] = n.z;
    B[3] = t.x; B[4] = t.y; B[5] = t.z;
    B[6] = b.x; B[7] = b.y; B[8] = b.z;
}

}  // namespace pdfc
