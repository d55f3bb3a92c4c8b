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
        // a vertex of the face, or a point on an edge: no collision (geometry_util.cpp:123-132)
        if (face[0] == v || face[1] == v || face[2] == v || c.bary[0] == 0. || c.bary[1] == 0. || c.bary[2] == 0.)
            continue;
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

void SelfWorkspace::init(const Model& M) {
    n = M.n;
    nbr_ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) nbr_ptr[i + 1] = nbr_ptr[i] + (int)M.neighbors[i].size();
    nbr_idx.resize(nbr_ptr[n]);
    for (int i = 0; i < n; ++i) std::copy(M.neighbors[i].begin(), M.neighbors[i].end(), nbr_idx.begin() + nbr_ptr[i]);
    reg_cnt.assign(n, 0);
    reg.assign((size_t)n * kRegCap, 0);
    touched.clear();
    vid.assign(n, -1);
    vbits.assign((size_t)n / 64 + 1, 0ull);
}

namespace {
// Position in the face of the first face vertex registered with u (isTriangleVertexIndexIn,
// find_first_of over the face), 3 if none.
inline int first_face_slot(const SelfWorkspace& W, const std::array<int, 3>& face, int u) {
    const int c = W.reg_cnt[u];
    const int* r = &W.reg[(size_t)u * SelfWorkspace::kRegCap];
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < c; ++j)
            if (r[j] == face[k]) return k;
    return 3;
}
}  // namespace

void pair_self_contacts(SelfWorkspace& W, const Model& M, const SelfCandidate* cand, size_t ncand,
                        std::vector<SelfPair>& out) {
    if (W.n != M.n) W.init(M);
    for (int v : W.touched) W.reg_cnt[v] = 0;
    W.touched.clear();
    out.clear();
    for (size_t k = 0; k < ncand; ++k) {
        const SelfCandidate& c = cand[k];
        const int v = c.v;
        const std::array<int, 3>& face = M.tris[c.t];
        // a vertex of the face, or a point on an edge: no collision (geometry_util.cpp:123-132)
        if (face[0] == v || face[1] == v || face[2] == v || c.bary[0] == 0. || c.bary[1] == 0. || c.bary[2] == 0.)
            continue;
        if (first_face_slot(W, face, v) < 3) continue;
        bool ok[3] = {true, true, true};
        for (int q = W.nbr_ptr[v]; q < W.nbr_ptr[v + 1]; ++q) {
            const int j = first_face_slot(W, face, W.nbr_idx[q]);
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
        if (W.reg_cnt[partner] == SelfWorkspace::kRegCap || W.reg_cnt[v] == SelfWorkspace::kRegCap) {
            // a vertex with more partners than the slots hold: the general form, from the start
            for (int x : W.touched) W.reg_cnt[x] = 0;
            W.touched.clear();
            pair_self_contacts(M, cand, ncand, out);
            return;
        }
        for (int x : {partner, v}) {
            if (W.reg_cnt[x] == 0) W.touched.push_back(x);
        }
        W.reg[(size_t)partner * SelfWorkspace::kRegCap + W.reg_cnt[partner]++] = v;
        W.reg[(size_t)v * SelfWorkspace::kRegCap + W.reg_cnt[v]++] = partner;
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

int order_self_contacts(SelfWorkspace& W, const std::function<bool(int)>& is_seed, std::vector<SelfPair>& pairs,
                        std::vector<int>& order, std::vector<int>& level_ptr) {
    order.clear();
    level_ptr.assign(1, 0);
    const int ns = (int)pairs.size();
    if (ns == 0) return 0;
    // The self-contact graph over the vertices that appear in it, numbered in ascending vertex
    // order (the reference's scans over all vertices visit them in that order and skip every
    // other vertex: no adjacency). Ascending order comes from a bitmap of the vertices.
    size_t wlo = W.vbits.size(), whi = 0;
    for (const auto& p : pairs)
        for (int x : {p.v, p.partner}) {
            const size_t w = (size_t)x >> 6;
            W.vbits[w] |= 1ull << (x & 63);
            wlo = std::min(wlo, w);
            whi = std::max(whi, w);
        }
    W.verts.clear();
    for (size_t w = wlo; w <= whi; ++w) {  // ascending vertex order; the bitmap is left cleared
        for (unsigned long long b = W.vbits[w]; b; b &= b - 1) W.verts.push_back((int)(64 * w) + __builtin_ctzll(b));
        W.vbits[w] = 0ull;
    }
    const int nv = (int)W.verts.size();
    for (int u = 0; u < nv; ++u) W.vid[W.verts[u]] = u;
    W.adj_ptr.assign(nv + 1, 0);
    for (const auto& p : pairs) {
        W.adj_ptr[W.vid[p.v] + 1]++;
        W.adj_ptr[W.vid[p.partner] + 1]++;
    }
    for (int u = 0; u < nv; ++u) W.adj_ptr[u + 1] += W.adj_ptr[u];
    W.adj_fill.assign(W.adj_ptr.begin(), W.adj_ptr.end() - 1);
    W.adj_v.resize(W.adj_ptr[nv]);
    W.adj_s.resize(W.adj_ptr[nv]);
    for (int s = 0; s < ns; ++s) {  // neighbours in contact order, as the reference's graph lists
        const int a = W.vid[pairs[s].v], b = W.vid[pairs[s].partner];
        W.adj_v[W.adj_fill[a]] = b;
        W.adj_s[W.adj_fill[a]++] = s;
        W.adj_v[W.adj_fill[b]] = a;
        W.adj_s[W.adj_fill[b]++] = s;
    }
    W.visited.assign(ns, 0);
    std::vector<int>& lev_of = W.adj_fill;  // reused: level of each contact
    lev_of.assign(ns, -1);
    int nlev = 0;
    // One queue for every search: the frontier of a level is the queue's range between two level
    // ends, and its vertices are expanded in the order they were discovered (as a frontier list
    // rebuilt per level would hold them).
    W.frontier.resize((size_t)nv + ns);
    int* const queue = W.frontier.data();
    auto bfs = [&](int nstart) {  // queue[0, nstart) holds the starting vertices
        int head = 0, tail = nstart, level_end = nstart, lvl = 0;
        if (nstart > 0 && nlev < 1) nlev = 1;
        while (head < tail) {
            const int u = queue[head++];
            for (int q = W.adj_ptr[u]; q < W.adj_ptr[u + 1]; ++q) {
                const int s = W.adj_s[q];
                if (W.visited[s]) continue;
                W.visited[s] = 1;
                queue[tail++] = W.adj_v[q];
                lev_of[s] = lvl;
            }
            if (head == level_end && head < tail) {
                ++lvl;
                level_end = tail;
                if (nlev <= lvl) nlev = lvl + 1;
            }
        }
    };
    auto all_visited = [&](int u) {
        for (int q = W.adj_ptr[u]; q < W.adj_ptr[u + 1]; ++q)
            if (!W.visited[W.adj_s[q]]) return false;
        return true;
    };
    // the first search starts from every obstacle-contact vertex at once; those outside the
    // graph add nothing, and the order of a frontier changes only the order inside a level
    int nseed = 0;
    for (int u = 0; u < nv; ++u)
        if (is_seed(W.verts[u])) queue[nseed++] = u;
    bfs(nseed);
    for (int pass = 0; pass < 2; ++pass)
        for (int u = 0; u < nv; ++u) {
            const int deg = W.adj_ptr[u + 1] - W.adj_ptr[u];
            if ((pass == 0 ? deg == 1 : deg > 1) && !all_visited(u)) {
                queue[0] = u;
                bfs(1);
            }
        }
    for (int u = 0; u < nv; ++u) W.vid[W.verts[u]] = -1;
    // contacts level by level; inside a level in BFS discovery order is not needed for the
    // arithmetic (a level's contacts are computed from the state at its start), contact order is
    // used instead, which is deterministic. Empty trailing levels are dropped.
    level_ptr.assign(nlev + 1, 0);
    for (int s = 0; s < ns; ++s) level_ptr[lev_of[s] + 1]++;
    while (nlev > 0 && level_ptr[nlev] == 0) --nlev;
    level_ptr.resize(nlev + 1);
    for (int l = 0; l < nlev; ++l) level_ptr[l + 1] += level_ptr[l];
    order.resize(ns);
    W.next.assign(level_ptr.begin(), level_ptr.end() - 1);
    for (int s = 0; s < ns; ++s) {
        order[W.next[lev_of[s]]++] = s;
        pairs[s].level = lev_of[s];
    }
    return nlev;
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
