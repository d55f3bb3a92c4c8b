/*
 * pd_oracle.cpp — CPU ORACLE (test infrastructure only).
 *
 * A plain C++/OpenMP restatement of the reference's projective-dynamics step
 * with dry frictional contact (guibquatre/ClimateNet = ProjectiveFriction).
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it; the product library never links or calls it.
 *
 * It follows the reference's algorithm and parallel structure:
 *   - OpenMP over constraints (local projections) and over contacts,
 *   - serial base-RHS sparse products, recomputed every iteration,
 *   - three single-RHS sparse Cholesky solves (one per coordinate) in parallel
 *     against one FP64 factor computed once.
 * Function-level citations give the reference file:line each part restates.
 *
 * Deterministic policies where the reference is racy or unspecified:
 *   - candidate triangles of a vertex are visited in ascending triangle index
 *     (the reference visits them in CGAL box_intersection_d callback order);
 *   - a contact that takes off stores r = 0 for the friction residual (the
 *     reference keeps a stale value);
 *   - the fill-reducing ordering is nested dissection by coordinate bisection
 *     (the reference uses Eigen's AMD); only rounding depends on it.
 *
 * Parity pins (tests/test_oracle.py): closed-form bending coefficients of a
 * regular grid, mass conservation, rest lengths, scipy spsolve on the
 * oracle's own K, free fall, the inclined-plane closed form, resting contact,
 * cone/Alart-Curnier invariants, brute-force candidate lists, brute-force
 * closest points. Mesh obstacles: tests/test_oracle_mesh.py. Boxes, rotating
 * spheres and attachments: tests/test_oracle_obstacles.py (nearest-face brute
 * force, omega x r, sticking on a turning surface, the implicit-Euler closed form
 * of an attachment). Self contacts: tests/test_oracle_self.py (greedy pairing on a
 * hand-built candidate list with hand-derived pairs, BFS levels written by hand from
 * the three seeding passes, the two-body closed form of the mass-weighted level
 * update, hand-computed contact bases).
 */
#include "pdfc.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using V3 = std::array<double, 3>;

inline V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 add(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 scale(double s, const V3& a) { return {s * a[0], s * a[1], s * a[2]}; }
inline double dot(const V3& a, const V3& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
inline double sqn(const V3& a) { return dot(a, a); }
inline double norm(const V3& a) { return std::sqrt(sqn(a)); }
inline V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
// Eigen normalized(): divide by sqrt(squaredNorm) when positive.
inline V3 normalized(const V3& a) {
    double z = sqn(a);
    if (z > 0.) {
        double s = std::sqrt(z);
        return {a[0] / s, a[1] / s, a[2] / s};
    }
    return a;
}

thread_local std::string g_err;

struct Basis { V3 n, t, b; };

// computeBasis: src/physics/SimulationFrictionEstimation.cpp:89-105
Basis contact_basis(const V3& normal) {
    Basis B;
    B.n = normal;
    V3 t = cross(normal, V3{1., 0., 0.});
    if (norm(t) < 1.e-15) t = cross(normal, V3{0., 1., 0.});
    t = normalized(t);
    B.t = t;
    B.b = normalized(cross(normal, t));
    return B;
}
inline V3 to_local(const Basis& B, const V3& x) { return {dot(B.n, x), dot(B.t, x), dot(B.b, x)}; }
inline V3 to_global(const Basis& B, const V3& r) {
    V3 o;
    for (int i = 0; i < 3; ++i) o[i] = (B.n[i] * r[0] + B.t[i] * r[1]) + B.b[i] * r[2];
    return o;
}

// Cone projection of the estimated force: SimulationFrictionEstimation.cpp:370-390.
// Returns false on take-off (f_N > 0).
bool cone_project(const V3& f, double mu, V3& r) {
    if (f[0] > 0.) { r = {0., 0., 0.}; return false; }
    r = {-f[0], -f[1], -f[2]};
    double rt = std::sqrt(r[1] * r[1] + r[2] * r[2]);
    if (rt > mu * r[0]) {
        r[1] *= mu * r[0] / rt;
        r[2] *= mu * r[0] / rt;
    }
    return true;
}

// Alart-Curnier function: src/physics/coulomb_util.cpp:101-152 (rho = 1).
V3 alart_curnier(const V3& u, const V3& r, double mu) {
    V3 ac;
    ac[0] = std::max(0., r[0] - u[0]) - r[0];
    double d1 = r[1] - u[1], d2 = r[2] - u[2];
    double dn = std::sqrt(d1 * d1 + d2 * d2);
    double radius = std::max(0., mu * r[0] / dn);
    double s = std::min(1., radius);
    ac[1] = d1 * s - r[1];
    ac[2] = d2 * s - r[2];
    return ac;
}

// getMeanValue/getSine/getCosine: src/geometry/geometry_util.cpp:766-791
double mean_value(const V3& v1, const V3& v2, const V3& v3) {
    V3 e1 = sub(v1, v2), e2 = sub(v3, v2);
    double n1 = norm(e1), n2 = norm(e2);
    double cosv = dot(e1, e2) / (n1 * n2);
    double num = 1. - cosv;
    if (num == 0.) return 0.;
    double sinv = norm(cross(e1, e2)) / (n1 * n2);
    return num / sinv;
}

// getLaplaceBeltramiDiscretisationMeanValueCoefficients: geometry_util.cpp:673-710
std::vector<double> mean_value_coefficients(const V3& c, const std::vector<V3>& ring) {
    size_t N = ring.size();
    std::vector<double> out(N);
    for (size_t i = 0; i < N; ++i) {
        const V3& v = ring[i];
        const V3& cw = ring[(i + 1) % N];
        const V3& ccw = ring[(i + N - 1) % N];
        double dist = norm(sub(v, c));
        double a = mean_value(v, c, cw);
        double b = mean_value(v, ccw, c);
        out[i] = (a + b) / dist;
    }
    return out;
}

// Ericson closest point on triangle with the reference's early exit:
// src/geometry/geometry_util.cpp:506-616 (tol2 receives the plain tolerance).
void closest_point(const V3& p, const V3& a, const V3& b, const V3& c, double tol2, V3& cp, V3& bary) {
    V3 ap = sub(p, a), ab = sub(b, a), ac = sub(c, a);
    if ((sqn(ap) - sqn(ab)) - sqn(ac) > tol2) { cp = a; bary = {1, 0, 0}; return; }
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { cp = a; bary = {1, 0, 0}; return; }
    V3 bp = sub(p, b);
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { cp = b; bary = {0, 1, 0}; return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        cp = add(a, scale(v, ab)); bary = {1 - v, v, 0}; return;
    }
    V3 cpv = sub(p, c);
    double d5 = dot(ab, cpv), d6 = dot(ac, cpv);
    if (d6 >= 0.0 && d5 <= d6) { cp = c; bary = {0, 0, 1}; return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        cp = add(a, scale(w, ac)); bary = {1 - w, 0, w}; return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        cp = add(b, scale(w, sub(c, b))); bary = {0, 1 - w, w}; return;
    }
    double denom = 1.0 / ((va + vb) + vc);
    double v = vb * denom, w = vc * denom;
    cp = add(add(a, scale(v, ab)), scale(w, ac));
    bary = {(1. - v) - w, v, w};
}

// Matrix3d (columns a,b,c) times barycentric vector, Eigen column order.
inline V3 tri_point(const V3& a, const V3& b, const V3& c, const V3& w) {
    V3 o;
    for (int i = 0; i < 3; ++i) o[i] = (a[i] * w[0] + b[i] * w[1]) + c[i] * w[2];
    return o;
}

struct CSR {
    int n = 0;
    std::vector<int64_t> ptr;
    std::vector<int> idx;
    std::vector<double> val;
};

struct ObsContact {
    int v; int obstacle; V3 point; V3 normal; V3 speed; double mu;
    Basis B; V3 r; int active;
};
struct SelfContact {
    int v; int partner; std::array<int, 3> face; V3 bary; V3 normal; V3 point; double mu;
    Basis B; V3 r; int active; int level;
};
struct Candidate { int v; int t; V3 bary; };

// Triangle-mesh obstacle: StaticMesh (one frame) or MovingMesh (key frames,
// MovingMesh.cpp:28-253) with its clock and current (interpolated) positions.
struct MeshObstacle {
    std::vector<std::array<int, 3>> tris;
    std::vector<std::vector<V3>> frames;
    std::vector<double> times;
    double tol = 0.;          // sqrt(tol * tol), as StaticMesh/MovingMesh pass it on
    double time = 0.;         // m_current_time
    int next = 1;             // m_next_main_frame_filename_index
    std::vector<V3> cur;      // m_current_mesh
    bool moving() const { return frames.size() >= 2; }
    // MovingMesh::isInMovement (:249-253)
    bool in_movement() const { return moving() && time >= times[next - 1] && time <= times[next]; }
};

}  // namespace

struct orc_sim {
    // ---- scene ----
    int n = 0, T = 0;
    std::vector<std::array<int, 3>> tris;
    std::vector<V3> rest, q, v;
    std::vector<double> mass, invmass;
    std::vector<pdfc_obstacle> obstacles;
    std::vector<MeshObstacle> mesh_obs;      // per obstacle (empty unless a triangle mesh)
    // rotating spheres: m_rotation_axe (radial speed times the unit axis) and
    // m_time_beforme_rotation_start, decremented by dt at the end of every step
    std::vector<V3> rot_omega;
    std::vector<double> rot_clock;
    // attachment rows (after the mesh's springs and bending rows): vertex, weight, position
    std::vector<int> att_v;
    std::vector<double> att_w;
    std::vector<V3> att_pos;
    int n_mat = 0; std::vector<double> friction; int mesh_material = 0;
    double mesh_area_density = 0., mesh_bend = 0., mesh_stretch = 0.;
    pdfc_params prm{};
    double tol_eff = 0.;  // sqrt(tol*tol): PhysicScene.cpp:183-202 passes sqrt(tol2)
    // ---- warm start of the next step (no reference counterpart): obstacle reactions in
    // contact order, used instead of the cone projection for iterations < warm_hold
    std::vector<V3> warm_r;
    int warm_hold = 0;
    // ---- topology ----
    std::vector<std::array<int, 2>> edges;
    std::vector<char> border;
    std::vector<std::vector<int>> neighbors;  // one ring (any order)
    // ---- constraints: rows of A ----
    // springs: rows [0, E); bending rows [E, E+B)
    std::vector<double> spring_w, spring_L0;
    std::vector<int> bend_center;
    std::vector<std::vector<int>> bend_idx;     // [center, ring...]
    std::vector<std::vector<double>> bend_coef;  // m_beltrami_coefficients (unweighted)
    std::vector<double> bend_w, bend_rest_norm;
    int m_rows = 0;
    CSR A_rows;   // row -> (vertex, value)
    CSR A_cols;   // vertex -> (row, value)  (A^T rows)
    CSR ATA;      // symmetric full pattern
    CSR K;        // M + dt^2 ATA
    // ---- factor ----
    std::vector<int> perm, iperm;
    std::vector<int64_t> Lp; std::vector<int> Li; std::vector<double> Lx;
    double factor_seconds = 0.;
    // ---- per-step state ----
    std::vector<ObsContact> contacts;
    std::vector<SelfContact> selfc;
    std::vector<Candidate> candidates;
    std::vector<std::vector<int>> levels;
    std::vector<double> damping;
    std::vector<double> ac_mean, ac_max;
    double time = 0.;
};

namespace {

// ------------------------------------------------------------------ topology
// Mesh ctor / CGAL Surface_mesh edge creation order: src/geometry/Mesh.cpp:31-56.
// One-ring order (vertices_around_target): for each face (.., u, c, w, ..) the
// ring element after u is w (src/geometry/Mesh.cpp:253-257).
void build_topology(orc_sim& S) {
    const int n = S.n;
    std::map<std::pair<int, int>, int> edge_faces;
    S.edges.clear();
    for (const auto& f : S.tris) {
        for (int i = 0; i < 3; ++i) {
            int a = f[i], b = f[(i + 1) % 3];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = edge_faces.find(key);
            if (it == edge_faces.end()) {
                edge_faces[key] = 1;
                S.edges.push_back({a, b});
            } else {
                it->second++;
            }
        }
    }
    S.border.assign(n, 0);
    S.neighbors.assign(n, {});
    for (auto& kv : edge_faces) {
        int a = kv.first.first, b = kv.first.second;
        S.neighbors[a].push_back(b);
        S.neighbors[b].push_back(a);
        if (kv.second == 1) { S.border[a] = 1; S.border[b] = 1; }
    }
    for (int i = 0; i < n; ++i)
        if (S.neighbors[i].empty()) S.border[i] = 1;  // isolated vertex: no bending row
}

std::vector<int> ring_of(const orc_sim& S, int c, const std::vector<std::vector<int>>& vf) {
    std::unordered_map<int, int> next;
    for (int fi : vf[c]) {
        const auto& f = S.tris[fi];
        int k = (f[0] == c) ? 0 : (f[1] == c ? 1 : 2);
        int u = f[(k + 2) % 3], w = f[(k + 1) % 3];
        next[u] = w;
    }
    int start = *std::min_element(S.neighbors[c].begin(), S.neighbors[c].end());
    std::vector<int> ring;
    int cur = start;
    for (size_t guard = 0; guard <= next.size(); ++guard) {
        ring.push_back(cur);
        auto it = next.find(cur);
        if (it == next.end()) throw std::invalid_argument("non-manifold one-ring");
        cur = it->second;
        if (cur == start) break;
    }
    if (ring.size() != S.neighbors[c].size()) throw std::invalid_argument("non-manifold one-ring");
    return ring;
}

// Masses (PhysicMesh::initializeMasses, src/physics/PhysicMesh.cpp:115-151), springs
// (SpringConstraint, PhysicMesh.cpp:100-106), bending (BendingConstraint ctor,
// src/physics/BendingConstraint.cpp:47-84).
void build_constraints(orc_sim& S) {
    const int n = S.n;
    std::vector<std::vector<int>> vf(n);
    for (int fi = 0; fi < S.T; ++fi)
        for (int k = 0; k < 3; ++k) vf[S.tris[fi][k]].push_back(fi);
    S.mass.assign(n, 0.);
    S.invmass.assign(n, 0.);
    for (int i = 0; i < n; ++i) {
        double acc = 0.;
        for (int fi : vf[i]) {
            const auto& f = S.tris[fi];
            V3 cr = cross(sub(S.rest[f[1]], S.rest[f[0]]), sub(S.rest[f[2]], S.rest[f[0]]));
            double area = std::sqrt(sqn(cr) / 4.);
            acc += S.mesh_area_density * area;
        }
        S.mass[i] = acc / 3.;
        if (!(S.mass[i] > 0.)) throw std::invalid_argument("vertex with zero mass");
        S.invmass[i] = 1. / S.mass[i];
    }
    const double ws = std::sqrt(S.mesh_stretch);
    S.spring_w.clear(); S.spring_L0.clear();
    for (const auto& e : S.edges) {
        S.spring_w.push_back(ws);
        S.spring_L0.push_back(norm(sub(S.rest[e[0]], S.rest[e[1]])));
    }
    const double wb = std::sqrt(S.mesh_bend);
    S.bend_center.clear(); S.bend_idx.clear(); S.bend_coef.clear(); S.bend_w.clear(); S.bend_rest_norm.clear();
    for (int c = 0; c < n; ++c) {
        if (S.border[c]) continue;
        std::vector<int> ring = ring_of(S, c, vf);
        std::vector<V3> rp;
        for (int r : ring) rp.push_back(S.rest[r]);
        std::vector<double> cf = mean_value_coefficients(S.rest[c], rp);
        std::vector<int> idx{c};
        idx.insert(idx.end(), ring.begin(), ring.end());
        std::vector<double> coef{0.};
        for (double x : cf) { coef[0] += x; coef.push_back(-x); }
        S.bend_center.push_back(c);
        S.bend_idx.push_back(idx);
        S.bend_coef.push_back(coef);
        S.bend_w.push_back(wb);
        // BendingConstraint.cpp:71 evaluates the rest Laplacian before the
        // coefficient list is filled (:73-81), so the stored rest norm is 0.
        S.bend_rest_norm.push_back(0.);
    }
}

// A (SimulationBase ctor, SimulationBase.cpp:71-80), A^T A and K
// (SimulationSpeed::computeLhs, SimulationSpeed.cpp:230-236).
void assemble(orc_sim& S) {
    const int n = S.n;
    const int E = (int)S.edges.size(), B = (int)S.bend_center.size(), NA = (int)S.att_v.size();
    S.m_rows = E + B + NA;
    CSR& R = S.A_rows;
    R.n = S.m_rows;
    R.ptr.assign(S.m_rows + 1, 0);
    R.idx.clear(); R.val.clear();
    for (int e = 0; e < E; ++e) {
        R.idx.push_back(S.edges[e][0]); R.val.push_back(S.spring_w[e]);
        R.idx.push_back(S.edges[e][1]); R.val.push_back(-S.spring_w[e]);
        R.ptr[e + 1] = (int64_t)R.idx.size();
    }
    for (int b = 0; b < B; ++b) {
        for (size_t k = 0; k < S.bend_idx[b].size(); ++k) {
            R.idx.push_back(S.bend_idx[b][k]);
            R.val.push_back(S.bend_w[b] * S.bend_coef[b][k]);
        }
        R.ptr[E + b + 1] = (int64_t)R.idx.size();
    }
    // AttachmentConstraint::addConstraintMatrix (AttachmentConstraint.cpp:31-43): one entry w
    for (int a = 0; a < NA; ++a) {
        R.idx.push_back(S.att_v[a]);
        R.val.push_back(S.att_w[a]);
        R.ptr[E + B + a + 1] = (int64_t)R.idx.size();
    }
    // transpose: vertex -> rows (rows ascending)
    CSR& C = S.A_cols;
    C.n = n;
    C.ptr.assign(n + 1, 0);
    for (int x : R.idx) C.ptr[x + 1]++;
    for (int i = 0; i < n; ++i) C.ptr[i + 1] += C.ptr[i];
    C.idx.assign(R.idx.size(), 0); C.val.assign(R.idx.size(), 0.);
    {
        std::vector<int64_t> fill(C.ptr.begin(), C.ptr.end() - 1);
        for (int r = 0; r < S.m_rows; ++r)
            for (int64_t p = R.ptr[r]; p < R.ptr[r + 1]; ++p) {
                int i = R.idx[p];
                C.idx[fill[i]] = r; C.val[fill[i]] = R.val[p]; fill[i]++;
            }
    }
    // Gustavson A^T A: row i = sum over rows r containing i of A_ri * A_r.
    CSR& M = S.ATA;
    M.n = n; M.ptr.assign(n + 1, 0); M.idx.clear(); M.val.clear();
    std::vector<double> acc(n, 0.);
    std::vector<int> mark(n, -1), cols;
    for (int i = 0; i < n; ++i) {
        cols.clear();
        for (int64_t p = C.ptr[i]; p < C.ptr[i + 1]; ++p) {
            int r = C.idx[p]; double ari = C.val[p];
            for (int64_t q = R.ptr[r]; q < R.ptr[r + 1]; ++q) {
                int j = R.idx[q];
                if (mark[j] != i) { mark[j] = i; acc[j] = 0.; cols.push_back(j); }
                acc[j] += ari * R.val[q];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) { M.idx.push_back(j); M.val.push_back(acc[j]); }
        M.ptr[i + 1] = (int64_t)M.idx.size();
    }
    const double dt2 = S.prm.time_step * S.prm.time_step;
    S.K = M;
    for (int i = 0; i < n; ++i)
        for (int64_t p = S.K.ptr[i]; p < S.K.ptr[i + 1]; ++p) {
            double v = dt2 * M.val[p];
            if (S.K.idx[p] == i) v = S.mass[i] + v;
            S.K.val[p] = v;
        }
}

// ---------------------------------------------------------------- ordering
// Nested dissection by recursive coordinate bisection on the graph of K.
void nd_order(const orc_sim& S, std::vector<int>& set, std::vector<int>& out, std::vector<char>& side) {
    if (set.size() <= 64) {
        std::sort(set.begin(), set.end());
        out.insert(out.end(), set.begin(), set.end());
        return;
    }
    V3 lo = S.rest[set[0]], hi = lo;
    for (int v : set)
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], S.rest[v][k]); hi[k] = std::max(hi[k], S.rest[v][k]); }
    int ax = 0;
    for (int k = 1; k < 3; ++k) if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
    std::sort(set.begin(), set.end(), [&](int a, int b) {
        if (S.rest[a][ax] != S.rest[b][ax]) return S.rest[a][ax] < S.rest[b][ax];
        return a < b;
    });
    size_t half = set.size() / 2;
    std::vector<int> left(set.begin(), set.begin() + half), right(set.begin() + half, set.end());
    for (int v : right) side[v] = 2;
    std::vector<int> sep, left2;
    for (int v : left) {
        bool touches = false;
        for (int64_t p = S.K.ptr[v]; p < S.K.ptr[v + 1] && !touches; ++p) touches = (side[S.K.idx[p]] == 2);
        (touches ? sep : left2).push_back(v);
    }
    for (int v : right) side[v] = 0;
    set.clear(); set.shrink_to_fit();
    nd_order(S, left2, out, side);
    nd_order(S, right, out, side);
    std::sort(sep.begin(), sep.end());
    out.insert(out.end(), sep.begin(), sep.end());
}

// ----------------------------------------------------------- factorisation
// Up-looking sparse Cholesky of C = P K P^T (elimination tree, row subtrees).
void factorize(orc_sim& S) {
    const int n = S.n;
    std::vector<int> all(n);
    std::iota(all.begin(), all.end(), 0);
    std::vector<char> side(n, 0);
    S.perm.clear();
    nd_order(S, all, S.perm, side);
    S.iperm.assign(n, 0);
    for (int k = 0; k < n; ++k) S.iperm[S.perm[k]] = k;
    // C in CSC (symmetric, so CSR of rows == CSC), columns in new numbering
    std::vector<int64_t> Cp(n + 1, 0);
    std::vector<int> Ci; std::vector<double> Cx;
    for (int k = 0; k < n; ++k) {
        int old = S.perm[k];
        std::vector<std::pair<int, double>> col;
        for (int64_t p = S.K.ptr[old]; p < S.K.ptr[old + 1]; ++p) col.push_back({S.iperm[S.K.idx[p]], S.K.val[p]});
        std::sort(col.begin(), col.end());
        for (auto& e : col) { Ci.push_back(e.first); Cx.push_back(e.second); }
        Cp[k + 1] = (int64_t)Ci.size();
    }
    std::vector<int> parent(n, -1), anc(n, -1);
    for (int k = 0; k < n; ++k)
        for (int64_t p = Cp[k]; p < Cp[k + 1]; ++p) {
            int i = Ci[p];
            while (i != -1 && i < k) {
                int nx = anc[i];
                anc[i] = k;
                if (nx == -1) parent[i] = k;
                i = nx;
            }
        }
    std::vector<int> w(n, -1), stack(n), path(n), cnt(n, 1);
    // Pattern of row k of L (the row subtree), in topological order in stack[top..n).
    auto ereach = [&](int k) {
        int top = n;
        w[k] = k;
        for (int64_t p = Cp[k]; p < Cp[k + 1]; ++p) {
            int i = Ci[p];
            if (i > k) continue;
            int len = 0;
            for (; w[i] != k; i = parent[i]) { path[len++] = i; w[i] = k; }
            while (len > 0) stack[--top] = path[--len];
        }
        return top;
    };
    for (int k = 0; k < n; ++k) {
        int top = ereach(k);
        for (int t = top; t < n; ++t) cnt[stack[t]]++;
    }
    S.Lp.assign(n + 1, 0);
    for (int j = 0; j < n; ++j) S.Lp[j + 1] = S.Lp[j] + cnt[j];
    S.Li.assign(S.Lp[n], 0); S.Lx.assign(S.Lp[n], 0.);
    std::vector<int64_t> c(S.Lp.begin(), S.Lp.end() - 1);
    std::vector<double> x(n, 0.);
    std::fill(w.begin(), w.end(), -1);
    for (int k = 0; k < n; ++k) {
        int top = ereach(k);
        for (int64_t p = Cp[k]; p < Cp[k + 1]; ++p) if (Ci[p] <= k) x[Ci[p]] = Cx[p];
        double d = x[k];
        x[k] = 0.;
        for (; top < n; ++top) {
            int i = stack[top];
            double lki = x[i] / S.Lx[S.Lp[i]];
            x[i] = 0.;
            for (int64_t p = S.Lp[i] + 1; p < c[i]; ++p) x[S.Li[p]] -= S.Lx[p] * lki;
            d -= lki * lki;
            int64_t p = c[i]++;
            S.Li[p] = k; S.Lx[p] = lki;
        }
        if (!(d > 0.)) throw std::runtime_error("K is not positive definite");
        int64_t p = c[k]++;
        S.Li[p] = k; S.Lx[p] = std::sqrt(d);
    }
}

// One single-RHS solve (SimplicialLLT::solve: permute, L, L^T, permute back).
void solve1(const orc_sim& S, const double* b, double* out, std::vector<double>& y) {
    const int n = S.n;
    y.resize(n);
    for (int k = 0; k < n; ++k) y[k] = b[S.perm[k]];
    for (int j = 0; j < n; ++j) {
        double yj = y[j] / S.Lx[S.Lp[j]];
        y[j] = yj;
        for (int64_t p = S.Lp[j] + 1; p < S.Lp[j + 1]; ++p) y[S.Li[p]] -= S.Lx[p] * yj;
    }
    for (int j = n - 1; j >= 0; --j) {
        double s = y[j];
        for (int64_t p = S.Lp[j] + 1; p < S.Lp[j + 1]; ++p) s -= S.Lx[p] * y[S.Li[p]];
        y[j] = s / S.Lx[S.Lp[j]];
    }
    for (int k = 0; k < n; ++k) out[S.perm[k]] = y[k];
}

// --------------------------------------------------------------- detection
void detect_mesh_obstacle(orc_sim& S, int o, double mu);

// StaticAxisAlignedBox::checkCollision (StaticAxisAlignedBox.cpp:41-60): inside unless some
// |relative_k| > size_k / 2, where relative = x - position - size / 2; the nearest side is
// the first minimum of |relative_axis - sign size_axis / 2| over TOP (+y), BOTTOM (-y),
// LEFT (-x), RIGHT (+x), FRONT (+z), BACK (-z) (Side enum order, std::min_element with <);
// the contact is relative with that axis set to the side, back in world coordinates.
bool box_contact(const pdfc_obstacle& ob, const V3& x, V3& point, V3& normal) {
    V3 rel;
    for (int k = 0; k < 3; ++k) rel[k] = (x[k] - ob.point[k]) - 0.5 * ob.size[k];
    for (int k = 0; k < 3; ++k)
        if (std::abs(rel[k]) > ob.size[k] / 2) return false;
    static const int axis[6] = {1, 1, 0, 0, 2, 2};
    static const double sign[6] = {1., -1., -1., 1., 1., -1.};
    int best = 0;
    double bd = 0.;
    for (int sd = 0; sd < 6; ++sd) {
        double d = std::abs(rel[axis[sd]] - sign[sd] * 0.5 * ob.size[axis[sd]]);
        if (sd == 0 || d < bd) { bd = d; best = sd; }
    }
    V3 res = rel;
    res[axis[best]] = sign[best] * 0.5 * ob.size[axis[best]];
    for (int k = 0; k < 3; ++k) point[k] = (res[k] + ob.point[k]) + 0.5 * ob.size[k];
    normal = {0., 0., 0.};
    normal[axis[best]] = sign[best];
    return true;
}

// PhysicScene::getCollisionsInfoExceptFor (PhysicScene.cpp:115-181): obstacles in
// order, vertices ascending. Sphere: Sphere.cpp:93-113. Plane: StaticPlane.cpp:57-73.
void detect_obstacles(orc_sim& S) {
    S.contacts.clear();
    for (int o = 0; o < (int)S.obstacles.size(); ++o) {
        const pdfc_obstacle& ob = S.obstacles[o];
        double mu = S.friction[(size_t)S.mesh_material * S.n_mat + ob.material];
        if (ob.type == PDFC_OBSTACLE_MESH) {
            detect_mesh_obstacle(S, o, mu);
            continue;
        }
        for (int i = 0; i < S.n; ++i) {
            const V3& x = S.q[i];
            ObsContact c{};
            c.speed = {0, 0, 0};
            if (ob.type == PDFC_OBSTACLE_SPHERE || ob.type == PDFC_OBSTACLE_ROTATING_SPHERE) {
                V3 ctr{ob.point[0], ob.point[1], ob.point[2]};
                V3 r = sub(x, ctr);
                if (!(norm(r) < ob.radius + 1e-2)) continue;
                r = normalized(r);
                c.point = add(ctr, scale(ob.radius, r));
                c.normal = r;
                // RotatingSphere::checkCollision (RotatingSphere.cpp:71-93)
                if (ob.type == PDFC_OBSTACLE_ROTATING_SPHERE && S.rot_clock[o] < 0)
                    c.speed = cross(S.rot_omega[o], sub(c.point, ctr));
            } else if (ob.type == PDFC_OBSTACLE_BOX) {
                if (!box_contact(ob, x, c.point, c.normal)) continue;
            } else {
                V3 p{ob.point[0], ob.point[1], ob.point[2]}, nn{ob.normal[0], ob.normal[1], ob.normal[2]};
                double d = dot(sub(x, p), nn);
                if (!(d < 1.5e-2)) continue;
                c.point = sub(x, scale(d, nn));
                c.normal = nn;
            }
            c.v = i; c.obstacle = o; c.mu = mu;
            c.B = contact_basis(c.normal);
            c.r = {0, 0, 0}; c.active = 0;
            S.contacts.push_back(c);
        }
    }
}

// Broad phase of getMeshPotentialTriangleCollision (geometry_util.cpp:264-327):
// closed boxes, vertex box x +- 1.1 tol, triangle bbox. A uniform grid finds the
// exact same set as CGAL box_intersection_d; the list is sorted by (v, t).
void broad_phase(const orc_sim& S, std::vector<std::pair<int, int>>& pairs, bool brute) {
    const double t11 = 1.1 * S.tol_eff;
    std::vector<V3> tlo(S.T), thi(S.T);
    double ext = 0.;
    for (int t = 0; t < S.T; ++t) {
        const auto& f = S.tris[t];
        for (int k = 0; k < 3; ++k) {
            tlo[t][k] = std::min(std::min(S.q[f[0]][k], S.q[f[1]][k]), S.q[f[2]][k]);
            thi[t][k] = std::max(std::max(S.q[f[0]][k], S.q[f[1]][k]), S.q[f[2]][k]);
            ext = std::max(ext, thi[t][k] - tlo[t][k]);
        }
    }
    auto overlap = [&](int v, int t) {
        for (int k = 0; k < 3; ++k) {
            double lo = S.q[v][k] - t11, hi = S.q[v][k] + t11;
            if (!(tlo[t][k] <= hi && lo <= thi[t][k])) return false;
        }
        return true;
    };
    pairs.clear();
    if (brute) {
        for (int v = 0; v < S.n; ++v)
            for (int t = 0; t < S.T; ++t)
                if (overlap(v, t)) pairs.push_back({v, t});
        return;
    }
    double h = std::max(ext, 2. * t11);
    if (!(h > 0.)) h = 1.;
    V3 org = S.q.empty() ? V3{0, 0, 0} : S.q[0];
    for (const auto& x : S.q) for (int k = 0; k < 3; ++k) org[k] = std::min(org[k], x[k] - t11);
    auto cell = [&](double x, int k) { return (int64_t)std::floor((x - org[k]) / h); };
    auto key = [](int64_t a, int64_t b, int64_t c) { return (a * 73856093LL) ^ (b * 19349663LL) ^ (c * 83492791LL); };
    std::unordered_map<int64_t, std::vector<int>> grid;
    for (int t = 0; t < S.T; ++t) {
        int64_t a0 = cell(tlo[t][0], 0), a1 = cell(thi[t][0], 0);
        int64_t b0 = cell(tlo[t][1], 1), b1 = cell(thi[t][1], 1);
        int64_t c0 = cell(tlo[t][2], 2), c1 = cell(thi[t][2], 2);
        for (int64_t a = a0; a <= a1; ++a)
            for (int64_t b = b0; b <= b1; ++b)
                for (int64_t c = c0; c <= c1; ++c) grid[key(a, b, c)].push_back(t);
    }
    std::vector<int> cand;
    for (int v = 0; v < S.n; ++v) {
        cand.clear();
        int64_t a0 = cell(S.q[v][0] - t11, 0), a1 = cell(S.q[v][0] + t11, 0);
        int64_t b0 = cell(S.q[v][1] - t11, 1), b1 = cell(S.q[v][1] + t11, 1);
        int64_t c0 = cell(S.q[v][2] - t11, 2), c1 = cell(S.q[v][2] + t11, 2);
        for (int64_t a = a0; a <= a1; ++a)
            for (int64_t b = b0; b <= b1; ++b)
                for (int64_t c = c0; c <= c1; ++c) {
                    auto it = grid.find(key(a, b, c));
                    if (it == grid.end()) continue;
                    for (int t : it->second) if (overlap(v, t)) cand.push_back(t);
                }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (int t : cand) pairs.push_back({v, t});
    }
}

// checkMeshCollision (geometry_util.cpp:32-92) on a mesh obstacle: the candidates of
// checkMeshAllCollision (vertex box x +- 1.1 tol against the triangle boxes, exact test
// with the early exit against tol, kept if |cp - x|^2 < tol^2), then per vertex the
// candidate whose point getTriangle(t) * bary is closest (first minimum in ascending
// triangle order), a contact if that distance^2 < tol^2. Normal: Mesh::getTriangleNormal
// (Mesh.cpp:121-126). Speed: MovingMesh::getPointSpeed (MovingMesh.cpp:225-243).
void detect_mesh_obstacle(orc_sim& S, int o, double mu) {
    const MeshObstacle& M = S.mesh_obs[o];
    const double tol = M.tol, t11 = 1.1 * tol, tol_sq = tol * tol;
    const int T = (int)M.tris.size();
    std::vector<V3> tlo(T), thi(T);
    for (int t = 0; t < T; ++t) {
        const auto& f = M.tris[t];
        for (int k = 0; k < 3; ++k) {
            tlo[t][k] = std::min(std::min(M.cur[f[0]][k], M.cur[f[1]][k]), M.cur[f[2]][k]);
            thi[t][k] = std::max(std::max(M.cur[f[0]][k], M.cur[f[1]][k]), M.cur[f[2]][k]);
        }
    }
    for (int i = 0; i < S.n; ++i) {
        const V3& x = S.q[i];
        double best = std::numeric_limits<double>::infinity();
        int bt = -1;
        V3 bbary{0, 0, 0}, bpt{0, 0, 0};
        for (int t = 0; t < T; ++t) {
            bool ov = true;
            for (int k = 0; k < 3 && ov; ++k) ov = tlo[t][k] <= x[k] + t11 && x[k] - t11 <= thi[t][k];
            if (!ov) continue;
            const auto& f = M.tris[t];
            V3 cp, bary;
            closest_point(x, M.cur[f[0]], M.cur[f[1]], M.cur[f[2]], tol, cp, bary);
            if (!(sqn(sub(cp, x)) < tol_sq)) continue;
            V3 pt = tri_point(M.cur[f[0]], M.cur[f[1]], M.cur[f[2]], bary);
            double d = sqn(sub(pt, x));
            if (d < best) { best = d; bt = t; bbary = bary; bpt = pt; }
        }
        if (!(bt >= 0 && best < tol_sq)) continue;
        const auto& f = M.tris[bt];
        ObsContact c{};
        c.v = i; c.obstacle = o; c.mu = mu;
        c.point = bpt;
        c.normal = normalized(cross(sub(M.cur[f[1]], M.cur[f[0]]), sub(M.cur[f[2]], M.cur[f[0]])));
        c.speed = {0, 0, 0};
        if (M.in_movement()) {
            const auto& P = M.frames[M.next - 1];
            const auto& N = M.frames[M.next];
            V3 pp = tri_point(P[f[0]], P[f[1]], P[f[2]], bbary);
            V3 np = tri_point(N[f[0]], N[f[1]], N[f[2]], bbary);
            c.speed = scale(1. / (M.times[M.next] - M.times[M.next - 1]), sub(np, pp));
        }
        c.B = contact_basis(c.normal);
        c.r = {0, 0, 0}; c.active = 0;
        S.contacts.push_back(c);
    }
}

// MovingMesh::update (MovingMesh.cpp:43-90): advance the clock, move to the next pair of
// key frames (one per update), interpolate the current mesh when in movement.
void update_mesh_obstacle(MeshObstacle& M, double dt) {
    if (!M.moving()) return;
    M.time += dt;
    if (M.times[M.next] < M.time && M.next + 1 < (int)M.frames.size()) ++M.next;
    if (!M.in_movement()) return;
    const double t0 = M.times[M.next - 1], t1 = M.times[M.next];
    const double a = (M.time - t0) / (t1 - t0);
    const auto& P = M.frames[M.next - 1];
    const auto& N = M.frames[M.next];
    for (size_t v = 0; v < M.cur.size(); ++v)
        for (int k = 0; k < 3; ++k) M.cur[v][k] = P[v][k] * (1. - a) + N[v][k] * a;
}

// checkMeshAllCollision (geometry_util.cpp:229-262): vertex boxes against triangle boxes,
// exact closest-point test, kept if |cp - x|^2 < tol^2; sorted by (vertex, triangle).
void self_candidates(orc_sim& S, bool brute) {
    std::vector<std::pair<int, int>> pairs;
    broad_phase(S, pairs, brute);
    const double tol = S.tol_eff, tol_sq = tol * tol;
    S.candidates.clear();
    for (auto& pr : pairs) {
        const auto& f = S.tris[pr.second];
        V3 cp, bary;
        closest_point(S.q[pr.first], S.q[f[0]], S.q[f[1]], S.q[f[2]], tol, cp, bary);
        if (sqn(sub(cp, S.q[pr.first])) < tol_sq) S.candidates.push_back({pr.first, pr.second, bary});
    }
}

// checkMeshSelfCollisions greedy node-node pairing (geometry_util.cpp:94-206) over
// S.candidates in list order (vertex ascending, then triangle): skip a candidate whose
// triangle holds the vertex or has a zero barycentric coordinate, or one of whose vertices
// the vertex is already paired with; a triangle vertex already paired with a one-ring
// neighbour of the vertex is excluded; the partner is the remaining triangle vertex of
// largest barycentric coordinate (none: skip). Normal from the unsnapped closest point.
void pair_candidates(orc_sim& S) {
    S.selfc.clear();
    const double mu = S.friction[(size_t)S.mesh_material * S.n_mat + S.mesh_material];
    std::vector<std::vector<int>> registered(S.n);
    // isTriangleVertexIndexIn (geometry_util.cpp:211-227): first triangle slot whose vertex is listed
    auto index_in = [](const std::array<int, 3>& tri, const std::vector<int>& list) {
        for (int k = 0; k < 3; ++k)
            if (std::find(list.begin(), list.end(), tri[k]) != list.end()) return k;
        return 3;
    };
    for (const auto& cd : S.candidates) {
        const int v = cd.v;
        const auto& tri = S.tris[cd.t];
        V3 bary = cd.bary;
        if (tri[0] == v || tri[1] == v || tri[2] == v || bary[0] == 0 || bary[1] == 0 || bary[2] == 0) continue;
        if (index_in(tri, registered[v]) < 3) continue;
        bool ok[3] = {true, true, true};
        for (int u : S.neighbors[v]) {
            int k = index_in(tri, registered[u]);
            if (k < 3) ok[k] = false;
        }
        int best = 3; double bv = 0.;
        for (int k = 0; k < 3; ++k) if (ok[k] && bary[k] > bv) { best = k; bv = bary[k]; }
        if (best >= 3) continue;
        V3 cp = tri_point(S.q[tri[0]], S.q[tri[1]], S.q[tri[2]], bary);
        SelfContact sc{};
        sc.normal = normalized(sub(S.q[v], cp));
        int partner = tri[best];
        bary = {0, 0, 0}; bary[best] = 1.;
        registered[partner].push_back(v);
        registered[v].push_back(partner);
        sc.point = tri_point(S.q[tri[0]], S.q[tri[1]], S.q[tri[2]], bary);
        sc.v = v; sc.partner = partner; sc.face = tri; sc.bary = bary; sc.mu = mu;
        sc.B = contact_basis(sc.normal);
        sc.r = {0, 0, 0}; sc.active = 0; sc.level = -1;
        S.selfc.push_back(sc);
    }
}

void detect_self(orc_sim& S, bool brute) {
    self_candidates(S, brute);
    pair_candidates(S);
}

// computeCollisionComputationOrder and helpers (SimulationFrictionEstimation.cpp:161-293):
// the graph joins each pair's vertex and partner (both directions, contact order); BFS
// fills levels from index 0 for every seed set: first all obstacle-contact vertices
// together, then each vertex of degree 1 not yet fully visited, then each remaining
// vertex of degree > 1, in vertex order. A contact is visited once; every BFS round
// opens its level even when it adds nothing.
std::vector<std::vector<int>> contact_levels(int n, const std::vector<std::pair<int, int>>& pairs,
                                             const std::vector<int>& seeds) {
    std::vector<std::vector<int>> levels;
    const int ns = (int)pairs.size();
    if (ns == 0) return levels;
    std::vector<std::vector<std::pair<int, int>>> graph(n);
    for (int s = 0; s < ns; ++s) {
        graph[pairs[s].first].push_back({pairs[s].second, s});
        graph[pairs[s].second].push_back({pairs[s].first, s});
    }
    std::vector<char> visited(ns, 0);
    auto fill = [&](std::vector<int> last, size_t idx) {
        std::vector<int> visiting;
        while (!last.empty()) {
            if (levels.size() <= idx) levels.emplace_back();
            for (int u : last)
                for (auto& nb : graph[u]) {
                    if (visited[nb.second]) continue;
                    visiting.push_back(nb.first);
                    visited[nb.second] = 1;
                    levels[idx].push_back(nb.second);
                }
            last.swap(visiting);
            visiting.clear();
            ++idx;
        }
    };
    auto vvisited = [&](int v) {
        for (auto& nb : graph[v]) if (!visited[nb.second]) return false;
        return true;
    };
    fill(seeds, 0);
    for (int v = 0; v < n; ++v) {
        if (vvisited(v)) continue;
        if (graph[v].size() == 1) fill({v}, 0);
    }
    for (int v = 0; v < n; ++v) {
        if (vvisited(v)) continue;
        if (graph[v].size() > 1) fill({v}, 0);
    }
    return levels;
}

void self_levels(orc_sim& S) {
    std::vector<std::pair<int, int>> pairs;
    for (const auto& s : S.selfc) pairs.push_back({s.v, s.partner});
    std::vector<int> seeds;
    for (auto& c : S.contacts) seeds.push_back(c.v);
    S.levels = contact_levels(S.n, pairs, seeds);
    for (int l = 0; l < (int)S.levels.size(); ++l)
        for (int s : S.levels[l]) S.selfc[s].level = l;
}

void detect(orc_sim& S, bool brute = false) {
    detect_obstacles(S);
    S.candidates.clear();
    S.selfc.clear();
    S.levels.clear();
    if (S.prm.self_collision) {
        detect_self(S, brute);
        self_levels(S);
    }
    // damping mask: SimulationBase.cpp:347-396
    S.damping.assign(S.n, std::max(S.prm.air_damping, 0.));
    if (S.prm.air_damping > 0.) {
        for (auto& c : S.contacts) S.damping[c.v] = 0.;
        for (auto& s : S.selfc) { S.damping[s.v] = 0.; S.damping[s.partner] = 0.; }
    }
}

// -------------------------------------------------------------------- step
// SimulationBase::step (SimulationBase.cpp:103-173) specialised by
// SimulationSpeed (SimulationSpeed.cpp:62-76,171-187,212-278) and
// SimulationFrictionEstimation (SimulationFrictionEstimation.cpp:64-87,307-588,590-753).
void step(orc_sim& S) {
    const int n = S.n;
    const double dt = S.prm.time_step, dt2 = dt * dt;
    const int K_it = S.prm.local_global_iterations;
    const V3 g{S.prm.gravity[0], S.prm.gravity[1], S.prm.gravity[2]};
    // initializeStep
    std::vector<V3> t(n), vk(n), qk(n);
    for (int i = 0; i < n; ++i) {
        V3 f = scale(S.mass[i], g);
        t[i] = add(S.v[i], scale(dt, scale(S.invmass[i], f)));
        vk[i] = t[i];
        qk[i] = add(S.q[i], scale(dt, vk[i]));
    }
    detect(S);
    const bool self_on = S.prm.self_collision != 0;
    const int nc = (int)S.contacts.size(), nsc = (int)S.selfc.size();
    // double buffers of SimulationFrictionEstimation.hpp:186-212
    std::vector<V3> CF[2], SCF[2];
    int cur = 0, nxt = 1;
    if (self_on) for (int k = 0; k < 2; ++k) { CF[k].assign(n, {0, 0, 0}); SCF[k].assign(n, {0, 0, 0}); }
    std::vector<V3> remember(nsc, {0, 0, 0});
    const int m = S.m_rows, E = (int)S.edges.size();
    std::vector<double> proj(3 * (size_t)m);
    std::vector<double> rhs(3 * (size_t)n), F(3 * (size_t)n), sol(3 * (size_t)n);
    std::vector<double> tmp1(n), tmp2(n);
    S.ac_mean.assign(K_it, 0.);
    S.ac_max.assign(K_it, 0.);
    for (int it = 0; it < K_it; ++it) {
        // computeProjections (SimulationBase.cpp:314-323): omp over constraints
#pragma omp parallel for schedule(static)
        for (int r = 0; r < m; ++r) {
            if (r >= E + (int)S.bend_center.size()) {
                // AttachmentConstraint::addConstraintProjection (:46-52): w * position
                const int a = r - E - (int)S.bend_center.size();
                for (int k = 0; k < 3; ++k) proj[3 * (size_t)r + k] = S.att_w[a] * S.att_pos[a][k];
            } else if (r < E) {
                V3 e = normalized(sub(qk[S.edges[r][0]], qk[S.edges[r][1]]));
                double s = S.spring_w[r] * S.spring_L0[r];
                for (int k = 0; k < 3; ++k) proj[3 * (size_t)r + k] = s * e[k];
            } else {
                int b = r - E;
                V3 lb{0, 0, 0};
                const auto& idx = S.bend_idx[b];
                const auto& cf = S.bend_coef[b];
                for (size_t j = 0; j < idx.size(); ++j) lb = add(lb, scale(cf[j], qk[idx[j]]));
                double ln = norm(lb);
                if (ln > 1e-25) {
                    lb = scale(S.bend_rest_norm[b] / ln, lb);
                    for (int k = 0; k < 3; ++k) proj[3 * (size_t)r + k] = S.bend_w[b] * lb[k];
                } else {
                    for (int k = 0; k < 3; ++k) proj[3 * (size_t)r + k] = 0.;
                }
            }
        }
        // SimulationSpeed::computeRhs: serial products per component
        for (int c = 0; c < 3; ++c) {
            for (int i = 0; i < n; ++i) {  // A^T p
                double s = 0.;
                for (int64_t p = S.A_cols.ptr[i]; p < S.A_cols.ptr[i + 1]; ++p) s += S.A_cols.val[p] * proj[3 * (size_t)S.A_cols.idx[p] + c];
                tmp1[i] = s;
            }
            for (int i = 0; i < n; ++i) {  // A^T A q_n
                double s = 0.;
                for (int64_t p = S.ATA.ptr[i]; p < S.ATA.ptr[i + 1]; ++p) s += S.ATA.val[p] * S.q[S.ATA.idx[p]][c];
                tmp2[i] = s;
            }
            for (int i = 0; i < n; ++i)
                rhs[(size_t)c * n + i] = S.mass[i] * t[i][c] + dt * ((tmp1[i] - S.damping[i] * vk[i][c]) - tmp2[i]);
        }
        const bool any = (nc > 0) || (self_on && nsc > 0);
        if (any) {
            // forces = rhs - dt^2 v A^T A (SimulationFrictionEstimation.cpp:337-347)
#pragma omp parallel for
            for (int c = 0; c < 3; ++c)
                for (int i = 0; i < n; ++i) {
                    double s = 0.;
                    for (int64_t p = S.ATA.ptr[i]; p < S.ATA.ptr[i + 1]; ++p) s += vk[S.ATA.idx[p]][c] * S.ATA.val[p];
                    F[(size_t)c * n + i] = rhs[(size_t)c * n + i] - dt2 * s;
                }
            auto Fv = [&](int i) { return V3{F[i], F[(size_t)n + i], F[2 * (size_t)n + i]}; };
            // obstacle contacts (:350-423). Serial adds keep the reference's per-vertex sum order.
            std::vector<V3> rg(nc);
#pragma omp parallel for
            for (int ci = 0; ci < nc; ++ci) {
                ObsContact& c = S.contacts[ci];
                V3 fv = Fv(c.v);
                if (self_on) fv = add(fv, SCF[cur][c.v]);
                fv = sub(fv, scale(S.mass[c.v], c.speed));
                V3 fl = to_local(c.B, fv);
                V3 r;
                const bool warm = it < S.warm_hold && ci < (int)S.warm_r.size() &&
                                  (S.warm_r[ci][0] != 0. || S.warm_r[ci][1] != 0. || S.warm_r[ci][2] != 0.);
                if (warm) {
                    // warm start: the given reaction, projected on the cone (take-off if r_N <= 0);
                    // a zero reaction (none given) starts plain
                    r = S.warm_r[ci];
                    if (!(r[0] > 0.)) {
                        r = {0, 0, 0};
                        c.active = 0;
                    } else {
                        double rt = std::sqrt(r[1] * r[1] + r[2] * r[2]);
                        if (rt > c.mu * r[0]) {
                            r[1] *= c.mu * r[0] / rt;
                            r[2] *= c.mu * r[0] / rt;
                        }
                        c.active = 1;
                    }
                } else {
                    c.active = cone_project(fl, c.mu, r) ? 1 : 0;
                }
                c.r = r;  // take-off: r = 0 (documented policy)
                rg[ci] = c.active ? to_global(c.B, r) : V3{0, 0, 0};
            }
            for (int ci = 0; ci < nc; ++ci) {
                const ObsContact& c = S.contacts[ci];
                if (!c.active) continue;
                if (!self_on || nsc == 0)
                    for (int k = 0; k < 3; ++k) rhs[(size_t)k * n + c.v] += rg[ci][k];
                if (self_on) CF[nxt][c.v] = rg[ci];
            }
            if (self_on && nsc > 0) {
                for (const auto& lvl : S.levels) {
                    std::vector<V3> add_f(lvl.size());
#pragma omp parallel for
                    for (int li = 0; li < (int)lvl.size(); ++li) {
                        SelfContact& sc = S.selfc[lvl[li]];
                        int a = sc.v, b = sc.partner;
                        double mA = S.mass[a], mB = S.mass[b];
                        V3 fA = sub(add(add(add(Fv(a), CF[nxt][a]), SCF[cur][a]), SCF[nxt][a]), remember[lvl[li]]);
                        V3 fB = add(add(add(add(Fv(b), CF[nxt][b]), SCF[cur][b]), SCF[nxt][b]), remember[lvl[li]]);
                        V3 fl = to_local(sc.B, sub(scale(1. / mA, fA), scale(1. / mB, fB)));
                        V3 r;
                        if (!cone_project(fl, sc.mu, r)) {
                            sc.active = 0; sc.r = {0, 0, 0}; add_f[li] = {0, 0, 0};
                            continue;
                        }
                        r = scale((mA * mB) / (mA + mB), r);
                        sc.active = 1; sc.r = r;
                        add_f[li] = to_global(sc.B, r);
                    }
                    for (int li = 0; li < (int)lvl.size(); ++li) {
                        int s = lvl[li];
                        int a = S.selfc[s].v, b = S.selfc[s].partner;
                        const V3 prev = remember[s];
                        SCF[cur][a] = sub(SCF[cur][a], prev);
                        SCF[cur][b] = add(SCF[cur][b], prev);
                        SCF[nxt][a] = add(SCF[nxt][a], add_f[li]);
                        SCF[nxt][b] = sub(SCF[nxt][b], add_f[li]);
                        remember[s] = add_f[li];
                    }
                }
                for (int i = 0; i < n; ++i)
                    for (int k = 0; k < 3; ++k) rhs[(size_t)k * n + i] += CF[nxt][i][k] + SCF[nxt][i][k];
            }
            if (self_on) {
                cur = nxt;
                nxt = nxt ? 0 : 1;
                std::fill(CF[nxt].begin(), CF[nxt].end(), V3{0, 0, 0});
                std::fill(SCF[nxt].begin(), SCF[nxt].end(), V3{0, 0, 0});
            }
        }
        // solve: three single-RHS solves in parallel (SimulationSpeed.cpp:171-187)
#pragma omp parallel for
        for (int c = 0; c < 3; ++c) {
            std::vector<double> y;
            solve1(S, rhs.data() + (size_t)c * n, sol.data() + (size_t)c * n, y);
        }
        for (int i = 0; i < n; ++i) {
            vk[i] = {sol[i], sol[(size_t)n + i], sol[2 * (size_t)n + i]};
            qk[i] = add(S.q[i], scale(dt, vk[i]));
        }
        // updateIterationData: Alart-Curnier norms (:590-652)
        double sum = 0., mx = 0.;
        for (const auto& c : S.contacts) {
            V3 u = to_local(c.B, sub(vk[c.v], c.speed));
            double a = norm(alart_curnier(u, c.r, c.mu));
            sum += a; mx = std::max(mx, a);
        }
        for (const auto& s : S.selfc) {
            V3 u = to_local(s.B, sub(vk[s.v], vk[s.partner]));
            double a = norm(alart_curnier(u, s.r, s.mu));
            sum += a; mx = std::max(mx, a);
        }
        int tot = nc + nsc;
        S.ac_mean[it] = tot ? sum / tot : 0.;
        S.ac_max[it] = mx;
    }
    // updateScene (:64-87)
    S.q = qk;
    S.v = vk;
    S.time += dt;
    for (auto& M : S.mesh_obs) update_mesh_obstacle(M, dt);
    for (size_t o = 0; o < S.obstacles.size(); ++o)  // RotatingSphere::update (RotatingSphere.cpp:39-43)
        if (S.obstacles[o].type == PDFC_OBSTACLE_ROTATING_SPHERE) S.rot_clock[o] -= dt;
    S.warm_r.clear();
    S.warm_hold = 0;
}

void fill_contact(const orc_sim& S, const ObsContact& c, pdfc_contact& o) {
    o.vertex = c.v; o.obstacle = c.obstacle; o.active = c.active;
    for (int k = 0; k < 3; ++k) {
        o.normal[k] = c.normal[k]; o.point[k] = c.point[k]; o.r_local[k] = c.r[k]; o.speed[k] = c.speed[k];
    }
    o.mu = c.mu;
    (void)S;
}

}  // namespace

// --------------------------------------------------------------------- C ABI
extern "C" {

struct orc_sim;

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_create(const pdfc_scene* sc, const pdfc_params* prm, orc_sim** out) {
    try {
        if (!sc || !prm || !out) throw std::invalid_argument("null argument");
        auto* S = new orc_sim();
        S->n = (int)sc->n_vertices;
        S->T = (int)sc->n_triangles;
        S->prm = *prm;
        S->rest.resize(S->n);
        S->q.resize(S->n);
        S->v.assign(S->n, {0, 0, 0});
        for (int i = 0; i < S->n; ++i)
            for (int k = 0; k < 3; ++k) {
                S->rest[i][k] = sc->rest_positions[3 * i + k];
                S->q[i][k] = sc->positions ? sc->positions[3 * i + k] : sc->rest_positions[3 * i + k];
            }
        S->tris.resize(S->T);
        for (int t = 0; t < S->T; ++t)
            for (int k = 0; k < 3; ++k) {
                int x = sc->triangles[3 * t + k];
                if (x < 0 || x >= S->n) throw std::invalid_argument("triangle index out of range");
                S->tris[t][k] = x;
            }
        S->n_mat = (int)sc->n_materials;
        S->friction.assign(sc->friction, sc->friction + (size_t)S->n_mat * S->n_mat);
        S->mesh_material = sc->material;
        S->mesh_area_density = sc->area_density;
        S->mesh_bend = sc->bending;
        S->mesh_stretch = sc->stretching;
        S->mesh_obs.resize(sc->n_obstacles);
        for (size_t o = 0; o < sc->n_obstacles; ++o) {
            pdfc_obstacle ob = sc->obstacles[o];
            if (ob.type == PDFC_OBSTACLE_PLANE) {
                V3 nn = normalized(V3{ob.normal[0], ob.normal[1], ob.normal[2]});
                for (int k = 0; k < 3; ++k) ob.normal[k] = nn[k];
            }
            if (ob.type == PDFC_OBSTACLE_MESH) {
                MeshObstacle& M = S->mesh_obs[o];
                if (ob.n_frames < 1 || !ob.mesh_frames || !ob.mesh_triangles || ob.n_mesh_triangles == 0)
                    throw std::invalid_argument("mesh obstacle without data");
                const size_t nv = ob.n_mesh_vertices;
                M.tris.resize(ob.n_mesh_triangles);
                for (size_t t = 0; t < ob.n_mesh_triangles; ++t)
                    for (int k = 0; k < 3; ++k) M.tris[t][k] = ob.mesh_triangles[3 * t + k];
                M.frames.assign(ob.n_frames, std::vector<V3>(nv));
                for (size_t f = 0; f < ob.n_frames; ++f)
                    for (size_t v = 0; v < nv; ++v)
                        for (int k = 0; k < 3; ++k) M.frames[f][v][k] = ob.mesh_frames[(f * nv + v) * 3 + k];
                if (ob.n_frames >= 2) M.times.assign(ob.frame_times, ob.frame_times + ob.n_frames);
                M.tol = std::sqrt(ob.tolerance * ob.tolerance);
                M.cur = M.frames[0];
                ob.mesh_triangles = nullptr;  // the copies above are the oracle's
                ob.mesh_frames = nullptr;
                ob.frame_times = nullptr;
            }
            S->obstacles.push_back(ob);
        }
        S->rot_omega.assign(S->obstacles.size(), {0, 0, 0});
        S->rot_clock.assign(S->obstacles.size(), 0.);
        for (size_t o = 0; o < S->obstacles.size(); ++o) {
            const pdfc_obstacle& ob = S->obstacles[o];
            if (ob.type != PDFC_OBSTACLE_ROTATING_SPHERE) continue;
            V3 ax{ob.rotation_axis[0], ob.rotation_axis[1], ob.rotation_axis[2]};
            if (!(norm(ax) > 0.)) throw std::invalid_argument("rotation axis is zero");
            S->rot_omega[o] = scale(ob.angular_speed, normalized(ax));
            S->rot_clock[o] = ob.rotation_start;
        }
        for (size_t a = 0; a < sc->n_attachments; ++a) {
            int v = sc->attachment_vertices[a];
            if (v < 0 || v >= S->n) throw std::invalid_argument("attachment vertex out of range");
            if (!(sc->attachment_weights[a] > 0.)) throw std::invalid_argument("attachment weight must be positive");
            S->att_v.push_back(v);
            S->att_w.push_back(sc->attachment_weights[a]);
            S->att_pos.push_back({sc->attachment_positions[3 * a], sc->attachment_positions[3 * a + 1],
                                  sc->attachment_positions[3 * a + 2]});
        }
        S->tol_eff = std::sqrt(prm->self_collision_tolerance * prm->self_collision_tolerance);
        build_topology(*S);
        build_constraints(*S);
        assemble(*S);
        auto t0 = std::chrono::steady_clock::now();
        factorize(*S);
        S->factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = S;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_destroy(orc_sim* S) { delete S; }

/* Warm start of the next step only: reactions (3 per obstacle contact, contact basis, in the
 * order of that step's contact list) used instead of the cone projection in iterations < hold;
 * a zero reaction leaves its contact plain. */
void orc_set_warm(orc_sim* S, const double* r_local, size_t n, int hold) {
    S->warm_r.resize(n);
    for (size_t i = 0; i < n; ++i) S->warm_r[i] = {r_local[3 * i], r_local[3 * i + 1], r_local[3 * i + 2]};
    S->warm_hold = hold;
}

/* Local-global iterations of the following steps (the parameter "LocalGlobal"). */
void orc_set_iterations(orc_sim* S, int K) { S->prm.local_global_iterations = K; }

int orc_step(orc_sim* S, int nsteps) {
    try {
        for (int s = 0; s < nsteps; ++s) step(*S);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int orc_detect(orc_sim* S, int brute) {
    try { detect(*S, brute != 0); return 0; } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int orc_get_state(orc_sim* S, double* q, double* v) {
    for (int i = 0; i < S->n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (q) q[3 * i + k] = S->q[i][k];
            if (v) v[3 * i + k] = S->v[i][k];
        }
    return 0;
}

int orc_set_state(orc_sim* S, const double* q, const double* v) {
    for (int i = 0; i < S->n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (q) S->q[i][k] = q[3 * i + k];
            if (v) S->v[i][k] = v[3 * i + k];
        }
    return 0;
}

int orc_get_contacts(orc_sim* S, pdfc_contact* out, size_t cap, size_t* n) {
    *n = S->contacts.size();
    if (cap < *n) return -4;
    for (size_t i = 0; i < *n; ++i) fill_contact(*S, S->contacts[i], out[i]);
    return 0;
}

int orc_get_self_contacts(orc_sim* S, pdfc_self_contact* out, size_t cap, size_t* n) {
    *n = S->selfc.size();
    if (cap < *n) return -4;
    for (size_t i = 0; i < *n; ++i) {
        const auto& s = S->selfc[i];
        pdfc_self_contact& o = out[i];
        o.vertex = s.v; o.partner = s.partner;
        for (int k = 0; k < 3; ++k) { o.face[k] = s.face[k]; o.normal[k] = s.normal[k]; o.r_local[k] = s.r[k]; }
        o.level = s.level; o.active = s.active; o.mu = s.mu;
    }
    return 0;
}

int orc_get_self_candidates(orc_sim* S, pdfc_candidate* out, size_t cap, size_t* n) {
    *n = S->candidates.size();
    if (cap < *n) return -4;
    for (size_t i = 0; i < *n; ++i) {
        out[i].vertex = S->candidates[i].v;
        out[i].triangle = S->candidates[i].t;
        for (int k = 0; k < 3; ++k) out[i].barycentric[k] = S->candidates[i].bary[k];
    }
    return 0;
}

int orc_get_iteration_stats(orc_sim* S, double* mean, double* mx, size_t cap, size_t* n) {
    *n = S->ac_mean.size();
    if (cap < *n) return -4;
    for (size_t i = 0; i < *n; ++i) { if (mean) mean[i] = S->ac_mean[i]; if (mx) mx[i] = S->ac_max[i]; }
    return 0;
}

/* K x = b for three components, generalized layout (3n). */
int orc_solve(orc_sim* S, const double* b, double* x) {
    const int n = S->n;
    std::vector<double> bc(3 * (size_t)n), xc(3 * (size_t)n);
    for (int i = 0; i < n; ++i) for (int k = 0; k < 3; ++k) bc[(size_t)k * n + i] = b[3 * i + k];
    for (int k = 0; k < 3; ++k) { std::vector<double> y; solve1(*S, bc.data() + (size_t)k * n, xc.data() + (size_t)k * n, y); }
    for (int i = 0; i < n; ++i) for (int k = 0; k < 3; ++k) x[3 * i + k] = xc[(size_t)k * n + i];
    return 0;
}

/* Sizes: n, T, E, B, nnz(K), nnz(L). */
int orc_get_sizes(orc_sim* S, int64_t* out6) {
    out6[0] = S->n; out6[1] = S->T; out6[2] = (int64_t)S->edges.size(); out6[3] = (int64_t)S->bend_center.size();
    out6[4] = (int64_t)S->K.idx.size(); out6[5] = (int64_t)S->Li.size();
    return 0;
}

/* CSR copy of K (original numbering); call with NULL arrays to size. */
int orc_get_K(orc_sim* S, int64_t* indptr, int32_t* indices, double* data) {
    for (int i = 0; i <= S->n; ++i) if (indptr) indptr[i] = S->K.ptr[i];
    for (size_t p = 0; p < S->K.idx.size(); ++p) {
        if (indices) indices[p] = S->K.idx[p];
        if (data) data[p] = S->K.val[p];
    }
    return 0;
}

int orc_get_masses(orc_sim* S, double* m) { for (int i = 0; i < S->n; ++i) m[i] = S->mass[i]; return 0; }

/* Springs: (i, j, weight, rest length) per row in A order. */
int orc_get_springs(orc_sim* S, int32_t* ij, double* w, double* L0) {
    for (size_t e = 0; e < S->edges.size(); ++e) {
        ij[2 * e] = S->edges[e][0]; ij[2 * e + 1] = S->edges[e][1];
        w[e] = S->spring_w[e]; L0[e] = S->spring_L0[e];
    }
    return 0;
}

/* Bending row b: fills up to cap entries of (index, unweighted coefficient); returns count. */
int orc_get_bending_row(orc_sim* S, int b, int32_t* idx, double* coef, int cap) {
    const auto& I = S->bend_idx[b];
    int c = (int)I.size();
    for (int k = 0; k < c && k < cap; ++k) { idx[k] = I[k]; coef[k] = S->bend_coef[b][k]; }
    return c;
}

double orc_factor_seconds(orc_sim* S) { return S->factor_seconds; }

/* ---- pieces exposed for unit tests ---- */
/* Greedy node-node pairing and level ordering of a given candidate list: the candidates
 * (vertex, triangle, barycentric) replace the detection's, are paired at the current
 * positions and ordered by level (seeded by the current obstacle contacts); read the
 * result with orc_get_self_contacts. */
int orc_pair_candidates(orc_sim* S, const pdfc_candidate* c, size_t nc) {
    S->candidates.clear();
    for (size_t i = 0; i < nc; ++i) {
        if (c[i].vertex < 0 || c[i].vertex >= S->n || c[i].triangle < 0 || c[i].triangle >= S->T) {
            g_err = "candidate index out of range";
            return -1;
        }
        S->candidates.push_back({(int)c[i].vertex, (int)c[i].triangle,
                                 V3{c[i].barycentric[0], c[i].barycentric[1], c[i].barycentric[2]}});
    }
    pair_candidates(*S);
    self_levels(*S);
    return 0;
}

/* Level ordering alone: pairs (v[s], partner[s]) over n vertices, BFS seeds = obstacle-contact
 * vertices. level[s] gets each pair's level, order[] the pair indices level by level (ns
 * entries). Returns the number of levels (empty trailing levels included), or -1. */
int orc_contact_levels(int n, const int64_t* v, const int64_t* partner, size_t ns, const int64_t* seeds,
                       size_t nseeds, int32_t* level, int32_t* order) {
    std::vector<std::pair<int, int>> pairs;
    for (size_t s = 0; s < ns; ++s) {
        if (v[s] < 0 || v[s] >= n || partner[s] < 0 || partner[s] >= n) { g_err = "pair vertex out of range"; return -1; }
        pairs.push_back({(int)v[s], (int)partner[s]});
    }
    std::vector<int> sd;
    for (size_t k = 0; k < nseeds; ++k) {
        if (seeds[k] < 0 || seeds[k] >= n) { g_err = "seed vertex out of range"; return -1; }
        sd.push_back((int)seeds[k]);
    }
    auto L = contact_levels(n, pairs, sd);
    size_t o = 0;
    for (int l = 0; l < (int)L.size(); ++l)
        for (int s : L[l]) { level[s] = l; order[o++] = s; }
    return (int)L.size();
}

void orc_mean_value_coefficients(const double* c, const double* ring, int N, double* out) {
    std::vector<V3> r(N);
    for (int i = 0; i < N; ++i) r[i] = {ring[3 * i], ring[3 * i + 1], ring[3 * i + 2]};
    auto v = mean_value_coefficients(V3{c[0], c[1], c[2]}, r);
    for (int i = 0; i < N; ++i) out[i] = v[i];
}

void orc_closest_point(const double* p, const double* a, const double* b, const double* c, double tol2, double* cp, double* bary) {
    V3 o, w;
    closest_point(V3{p[0], p[1], p[2]}, V3{a[0], a[1], a[2]}, V3{b[0], b[1], b[2]}, V3{c[0], c[1], c[2]}, tol2, o, w);
    for (int k = 0; k < 3; ++k) { cp[k] = o[k]; bary[k] = w[k]; }
}

int orc_cone(const double* f, double mu, double* r) {
    V3 o;
    int active = cone_project(V3{f[0], f[1], f[2]}, mu, o) ? 1 : 0;
    for (int k = 0; k < 3; ++k) r[k] = o[k];
    return active;
}

double orc_alart_curnier(const double* u, const double* r, double mu, double* ac) {
    V3 a = alart_curnier(V3{u[0], u[1], u[2]}, V3{r[0], r[1], r[2]}, mu);
    if (ac) for (int k = 0; k < 3; ++k) ac[k] = a[k];
    return norm(a);
}

void orc_contact_basis(const double* n, double* B9) {
    Basis B = contact_basis(V3{n[0], n[1], n[2]});
    for (int k = 0; k < 3; ++k) { B9[k] = B.n[k]; B9[3 + k] = B.t[k]; B9[6 + k] = B.b[k]; }
}

/* ---- contact MLP of the learned warm start (no reference counterpart):
 * 12 -> 64 -> 64 -> 64 -> 3, SiLU, in FP64. With bf16 != 0 the values the device
 * rounds to BF16 before its MFMA products are rounded the same way here (normalised
 * inputs, weights, hidden activations, backward operands), so the remaining
 * difference to the device is its FP32 accumulation and exponential; with bf16 == 0
 * it is the exact FP64 network, whose gradient tests/test_oracle.py pins by central
 * finite differences. Parameter layout as climatenet_amd/csrc/mlp.hpp. */
static float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) {
        u = (u | 0x400000u) & 0xffff0000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
static const int kF = 12, kO = 3;
static int mk(int l) { return l == 0 ? 16 : 64; }
static int mn(int l) { return l == 3 ? 16 : 64; }
static int offw(int l) { return l == 0 ? 0 : (l == 1 ? 1088 : (l == 2 ? 5248 : 9408)); }
static int offb(int l) { return offw(l) + mk(l) * mn(l); }

struct MlpFwd {
    std::vector<double> x;           // 16 normalised inputs
    std::vector<double> z[3], h[3];  // 64 each: pre-activations and activations
    std::vector<double> y;           // 16
};

static double rnd(bool bf, double v) { return bf ? (double)bf16_round((float)v) : v; }

static void mlp_fwd(const float* P, const float* norm, const float* xr, bool bf, MlpFwd& F) {
    F.x.assign(16, 0.);
    for (int c = 0; c < kF; ++c) F.x[c] = rnd(bf, ((double)xr[c] - norm[c]) / norm[kF + c]);
    const std::vector<double>* in = &F.x;
    for (int l = 0; l < 3; ++l) {
        F.z[l].assign(64, 0.);
        F.h[l].assign(64, 0.);
        for (int n = 0; n < 64; ++n) {
            double s = 0.;
            for (int k = 0; k < mk(l); ++k) s += (*in)[k] * rnd(bf, P[offw(l) + k * mn(l) + n]);
            s += (double)P[offb(l) + n];
            F.z[l][n] = s;
            F.h[l][n] = rnd(bf, s / (1. + std::exp(-s)));
        }
        in = &F.h[l];
    }
    F.y.assign(16, 0.);
    for (int n = 0; n < 16; ++n) {
        double s = 0.;
        for (int k = 0; k < 64; ++k) s += F.h[2][k] * rnd(bf, P[offw(3) + k * 16 + n]);
        F.y[n] = s + (double)P[offb(3) + n];
    }
}

void orc_mlp_forward(const float* P, const float* norm, const float* x, size_t rows, int bf16, double* y) {
    MlpFwd F;
    for (size_t r = 0; r < rows; ++r) {
        mlp_fwd(P, norm, x + r * kF, bf16 != 0, F);
        for (int c = 0; c < kO; ++c) y[r * kO + c] = F.y[c] * norm[2 * kF + c];
    }
}

/* L = mean over rows and the 3 outputs of (y - t/scale)^2, y the unscaled output;
 * grad in the flat parameter layout (padding entries stay 0). Returns L. */
double orc_mlp_gradient(const float* P, const float* norm, const float* x, const float* t, size_t rows, int bf16,
                        double* grad) {
    const bool bf = bf16 != 0;
    const int NP = 10448;
    for (int p = 0; p < NP; ++p) grad[p] = 0.;
    const double inv = 1. / (double)(rows * kO);
    double loss = 0.;
    MlpFwd F;
    std::vector<double> G(16), dz(64);
    for (size_t r = 0; r < rows; ++r) {
        mlp_fwd(P, norm, x + r * kF, bf, F);
        G.assign(16, 0.);
        for (int c = 0; c < kO; ++c) {
            double diff = F.y[c] - (double)t[r * kO + c] / norm[2 * kF + c];
            loss += diff * diff;
            double d = 2. * diff * inv;
            grad[offb(3) + c] += d;
            G[c] = rnd(bf, d);
        }
        for (int l = 3; l >= 0; --l) {
            const std::vector<double>& hp = l == 0 ? F.x : F.h[l - 1];
            const int K = mk(l), N = mn(l);
            for (int i = 0; i < K; ++i) {
                if (l == 0 && i >= kF) continue;
                for (int n = 0; n < N; ++n) {
                    if (l == 3 && n >= kO) continue;
                    grad[offw(l) + i * N + n] += hp[i] * G[n];
                }
            }
            if (l == 0) break;
            for (int i = 0; i < 64; ++i) {
                double s = 0.;
                for (int n = 0; n < N; ++n) s += G[n] * rnd(bf, P[offw(l) + i * N + n]);
                double zz = F.z[l - 1][i];
                double sg = 1. / (1. + std::exp(-zz));
                dz[i] = s * (sg * (1. + zz * (1. - sg)));
            }
            G.assign(64, 0.);
            for (int i = 0; i < 64; ++i) {
                grad[offb(l - 1) + i] += dz[i];
                G[i] = rnd(bf, dz[i]);
            }
        }
    }
    return loss * inv;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int k) {
#ifdef _OPENMP
    omp_set_num_threads(k > 0 ? k : 1);
#else
    (void)k;
#endif
}

}  // extern "C"
