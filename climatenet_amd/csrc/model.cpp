// Scene model: topology, masses, constraint rows and A^T A.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <unordered_map>

#include "host.hpp"

namespace pdfc {
namespace {

inline V3 vat(const std::vector<double>& p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline double vnorm(V3 a) { return std::sqrt(vdot(a, a)); }
inline V3 vcross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// tan(theta/2) = (1 - cos)/sin of the angle at v2 between v1 and v3, 0 when cos == 1
// (geometry_util.cpp:766-791).
double half_angle_tangent(V3 v1, V3 v2, V3 v3) {
    V3 e1 = vsub(v1, v2), e2 = vsub(v3, v2);
    double l = vnorm(e1) * vnorm(e2);
    double c = vdot(e1, e2) / l;
    double num = 1. - c;
    if (num == 0.) return 0.;
    return num / (vnorm(vcross(e1, e2)) / l);
}

}  // namespace

void build_model(const pdfc_scene& sc, Model& M) {
    if (sc.n_vertices == 0 || sc.n_triangles == 0) throw std::invalid_argument("empty mesh");
    if (sc.n_vertices > (size_t)INT32_MAX / 4) throw std::invalid_argument("mesh too large");
    const int n = (int)sc.n_vertices, T = (int)sc.n_triangles;
    M.n = n;
    M.T = T;
    M.rest.assign(sc.rest_positions, sc.rest_positions + 3 * (size_t)n);
    for (size_t k = 0; k < M.rest.size(); ++k)
        if (!std::isfinite(M.rest[k]) || (sc.positions && !std::isfinite(sc.positions[k])))
            throw std::invalid_argument("non-finite vertex position");
    M.tris.resize(T);
    for (int t = 0; t < T; ++t) {
        for (int k = 0; k < 3; ++k) {
            int x = sc.triangles[3 * t + k];
            if (x < 0 || x >= n) throw std::invalid_argument("triangle index out of range");
            M.tris[t][k] = x;
        }
        if (M.tris[t][0] == M.tris[t][1] || M.tris[t][1] == M.tris[t][2] || M.tris[t][0] == M.tris[t][2])
            throw std::invalid_argument("degenerate triangle");
    }
    // Edges in the order CGAL's add_face creates them (Mesh.cpp:31-56): face by face,
    // corner by corner, first time an undirected edge is seen.
    std::unordered_map<uint64_t, int> edge_id;
    edge_id.reserve(3 * (size_t)T);
    std::vector<int> edge_faces;
    M.edges.clear();
    for (const auto& f : M.tris)
        for (int k = 0; k < 3; ++k) {
            int a = f[k], b = f[(k + 1) % 3];
            uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            auto it = edge_id.find(key);
            if (it == edge_id.end()) {
                edge_id.emplace(key, (int)M.edges.size());
                M.edges.push_back({a, b});
                edge_faces.push_back(1);
            } else {
                if (++edge_faces[it->second] > 2) throw std::invalid_argument("non-manifold edge");
            }
        }
    M.border.assign(n, 0);
    M.neighbors.assign(n, {});
    for (size_t e = 0; e < M.edges.size(); ++e) {
        int a = M.edges[e][0], b = M.edges[e][1];
        M.neighbors[a].push_back(b);
        M.neighbors[b].push_back(a);
        if (edge_faces[e] == 1) M.border[a] = M.border[b] = 1;
    }
    M.max_degree = 0;
    for (int i = 0; i < n; ++i) {
        std::sort(M.neighbors[i].begin(), M.neighbors[i].end());
        if (M.neighbors[i].empty()) M.border[i] = 1;
        M.max_degree = std::max(M.max_degree, (int)M.neighbors[i].size());
    }
    // vertex -> incident faces
    std::vector<int> vf_ptr(n + 1, 0), vf;
    for (const auto& f : M.tris) for (int k = 0; k < 3; ++k) vf_ptr[f[k] + 1]++;
    for (int i = 0; i < n; ++i) vf_ptr[i + 1] += vf_ptr[i];
    vf.resize(vf_ptr[n]);
    {
        std::vector<int> fill(vf_ptr.begin(), vf_ptr.end() - 1);
        for (int t = 0; t < T; ++t) for (int k = 0; k < 3; ++k) vf[fill[M.tris[t][k]]++] = t;
    }
    // Lumped masses: rho * area / 3 summed over incident faces (PhysicMesh.cpp:115-151).
    M.mass.assign(n, 0.);
    for (int i = 0; i < n; ++i) {
        double acc = 0.;
        for (int p = vf_ptr[i]; p < vf_ptr[i + 1]; ++p) {
            const auto& f = M.tris[vf[p]];
            V3 a = vat(M.rest, f[0]);
            V3 cr = vcross(vsub(vat(M.rest, f[1]), a), vsub(vat(M.rest, f[2]), a));
            acc += sc.area_density * std::sqrt(vdot(cr, cr) / 4.);
        }
        M.mass[i] = acc / 3.;
        if (!(M.mass[i] > 0.)) throw std::invalid_argument("vertex without mass (isolated or zero-area faces)");
    }
    // Springs
    const double ws = std::sqrt(sc.stretching);
    M.spring_w.assign(M.edges.size(), ws);
    M.spring_L0.resize(M.edges.size());
    for (size_t e = 0; e < M.edges.size(); ++e)
        M.spring_L0[e] = vnorm(vsub(vat(M.rest, M.edges[e][0]), vat(M.rest, M.edges[e][1])));
    // Bending rows, interior vertices ascending. One-ring: after u comes w for every face
    // (.., u, c, w, ..); the ring starts at the smallest neighbour index.
    const double wb = std::sqrt(sc.bending);
    M.bend_ptr.assign(1, 0);
    M.bend_idx.clear();
    M.bend_coef.clear();
    std::vector<int> ring;
    std::vector<std::pair<int, int>> succ;
    for (int c = 0; c < n; ++c) {
        if (M.border[c]) continue;
        succ.clear();
        for (int p = vf_ptr[c]; p < vf_ptr[c + 1]; ++p) {
            const auto& f = M.tris[vf[p]];
            int k = f[0] == c ? 0 : (f[1] == c ? 1 : 2);
            succ.push_back({f[(k + 2) % 3], f[(k + 1) % 3]});
        }
        std::sort(succ.begin(), succ.end());
        auto next_of = [&](int u) {
            auto it = std::lower_bound(succ.begin(), succ.end(), std::make_pair(u, INT32_MIN));
            if (it == succ.end() || it->first != u) throw std::invalid_argument("non-manifold one-ring");
            return it->second;
        };
        ring.clear();
        int start = M.neighbors[c].front(), cur = start;
        do {
            ring.push_back(cur);
            cur = next_of(cur);
            if (ring.size() > M.neighbors[c].size()) throw std::invalid_argument("non-manifold one-ring");
        } while (cur != start);
        if (ring.size() != M.neighbors[c].size() || ring.size() < 3) throw std::invalid_argument("non-manifold one-ring");
        const size_t N = ring.size();
        const V3 cc = vat(M.rest, c);
        double sum = 0.;
        size_t base = M.bend_coef.size();
        M.bend_idx.push_back(c);
        M.bend_coef.push_back(0.);
        for (size_t i = 0; i < N; ++i) {
            V3 vi = vat(M.rest, ring[i]);
            V3 vn = vat(M.rest, ring[(i + 1) % N]);
            V3 vp = vat(M.rest, ring[(i + N - 1) % N]);
            double coef = (half_angle_tangent(vi, cc, vn) + half_angle_tangent(vi, vp, cc)) / vnorm(vsub(vi, cc));
            sum += coef;
            M.bend_idx.push_back(ring[i]);
            M.bend_coef.push_back(wb * -coef);
        }
        M.bend_coef[base] = wb * sum;
        M.bend_ptr.push_back((int)M.bend_idx.size());
    }
    M.att_vertex.clear();
    M.att_w.clear();
    M.att_pos.clear();
    if (sc.n_attachments > 0 && (!sc.attachment_vertices || !sc.attachment_positions || !sc.attachment_weights))
        throw std::invalid_argument("attachment arrays missing");
    for (size_t a = 0; a < sc.n_attachments; ++a) {
        const int v = sc.attachment_vertices[a];
        if (v < 0 || v >= n) throw std::invalid_argument("attachment vertex out of range");
        if (!(sc.attachment_weights[a] > 0.) || !std::isfinite(sc.attachment_weights[a]))
            throw std::invalid_argument("attachment weight must be positive and finite");
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(sc.attachment_positions[3 * a + c]))
                throw std::invalid_argument("non-finite attachment position");
        M.att_vertex.push_back(v);
        M.att_w.push_back(sc.attachment_weights[a]);
        M.att_pos.insert(M.att_pos.end(), sc.attachment_positions + 3 * a, sc.attachment_positions + 3 * a + 3);
    }
    // A^T A: accumulate a_r^T a_r row by row into per-vertex sorted rows.
    std::vector<std::vector<std::pair<int, double>>> rowacc(n);
    auto add_outer = [&](const int* idx, const double* val, int len) {
        for (int a = 0; a < len; ++a)
            for (int b = 0; b < len; ++b) rowacc[idx[a]].push_back({idx[b], val[a] * val[b]});
    };
    for (size_t e = 0; e < M.edges.size(); ++e) {
        int idx[2] = {M.edges[e][0], M.edges[e][1]};
        double val[2] = {M.spring_w[e], -M.spring_w[e]};
        add_outer(idx, val, 2);
    }
    for (size_t b = 0; b + 1 < M.bend_ptr.size(); ++b)
        add_outer(&M.bend_idx[M.bend_ptr[b]], &M.bend_coef[M.bend_ptr[b]], M.bend_ptr[b + 1] - M.bend_ptr[b]);
    for (size_t a = 0; a < M.att_vertex.size(); ++a) add_outer(&M.att_vertex[a], &M.att_w[a], 1);
    M.ata_ptr.assign(n + 1, 0);
    M.ata_idx.clear();
    M.ata_val.clear();
    M.max_ata_row = 0;
    for (int i = 0; i < n; ++i) {
        auto& r = rowacc[i];
        std::stable_sort(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        int start = (int)M.ata_idx.size();
        for (size_t k = 0; k < r.size();) {
            int j = r[k].first;
            double s = 0.;
            for (; k < r.size() && r[k].first == j; ++k) s += r[k].second;
            M.ata_idx.push_back(j);
            M.ata_val.push_back(s);
        }
        if (std::find(M.ata_idx.begin() + start, M.ata_idx.end(), i) == M.ata_idx.end())
            throw std::invalid_argument("vertex without constraints");
        M.ata_ptr[i + 1] = (int64_t)M.ata_idx.size();
        M.max_ata_row = std::max(M.max_ata_row, (int)(M.ata_ptr[i + 1] - M.ata_ptr[i]));
        r.clear();
        r.shrink_to_fit();
    }
}

}  // namespace pdfc

namespace pdfc {

// Disjoint union of independent scenes: vertices, triangles, constraint rows and
// A^T A of part k follow those of parts 0..k-1 (block-diagonal K).
void merge_models(const std::vector<Model>& parts, Model& M) {
    M = Model();
    for (const Model& P : parts) {
        const int v0 = M.n;
        for (const auto& t : P.tris) M.tris.push_back({t[0] + v0, t[1] + v0, t[2] + v0});
        M.rest.insert(M.rest.end(), P.rest.begin(), P.rest.end());
        M.mass.insert(M.mass.end(), P.mass.begin(), P.mass.end());
        for (const auto& e : P.edges) M.edges.push_back({e[0] + v0, e[1] + v0});
        M.spring_w.insert(M.spring_w.end(), P.spring_w.begin(), P.spring_w.end());
        M.spring_L0.insert(M.spring_L0.end(), P.spring_L0.begin(), P.spring_L0.end());
        if (M.bend_ptr.empty()) M.bend_ptr.push_back(0);
        const int b0 = (int)M.bend_idx.size();
        for (size_t b = 1; b < P.bend_ptr.size(); ++b) M.bend_ptr.push_back(P.bend_ptr[b] + b0);
        for (int x : P.bend_idx) M.bend_idx.push_back(x + v0);
        M.bend_coef.insert(M.bend_coef.end(), P.bend_coef.begin(), P.bend_coef.end());
        for (int x : P.att_vertex) M.att_vertex.push_back(x + v0);
        M.att_w.insert(M.att_w.end(), P.att_w.begin(), P.att_w.end());
        M.att_pos.insert(M.att_pos.end(), P.att_pos.begin(), P.att_pos.end());
        M.border.insert(M.border.end(), P.border.begin(), P.border.end());
        for (const auto& nb : P.neighbors) {
            std::vector<int> r(nb);
            for (int& x : r) x += v0;
            M.neighbors.push_back(std::move(r));
        }
        if (M.ata_ptr.empty()) M.ata_ptr.push_back(0);
        const int64_t a0 = (int64_t)M.ata_idx.size();
        for (size_t i = 1; i < P.ata_ptr.size(); ++i) M.ata_ptr.push_back(P.ata_ptr[i] + a0);
        for (int x : P.ata_idx) M.ata_idx.push_back(x + v0);
        M.ata_val.insert(M.ata_val.end(), P.ata_val.begin(), P.ata_val.end());
        M.max_degree = std::max(M.max_degree, P.max_degree);
        M.max_ata_row = std::max(M.max_ata_row, P.max_ata_row);
        M.n += P.n;
        M.T += P.T;
    }
}

}  // namespace pdfc
