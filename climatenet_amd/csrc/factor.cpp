// Fill-reducing ordering and supernodal multifrontal factorisation of
// K = M + dt^2 A^T A (SimulationSpeed::computeLhs + SimplicialLLT::compute,
// src/physics/SimulationSpeed.cpp:44-53,230-236), done once on the host in FP64.
//
// Output per supernode s (columns [c0, c0+ns), below-rows R_s, |R_s| = nb):
//   F_s = [ L_ss^-1 ; L_bs L_ss^-1 ]        (ns+nb) x ns, column-major
// so one forward-substitution step is y_s = F_s[0:ns] b~_s, u_s = F_s[ns:] b~_s
// and one backward step is x_s = F_s^T [ y_s ; -x_R ]. Inverting the dense
// diagonal blocks on the host turns both substitutions into dense
// matrix-vector products with no sequential dependence inside a supernode.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <stdexcept>

#include "host.hpp"

namespace pdfc {
namespace {

struct NDState {
    const Model* M;
    std::vector<char> side;
    std::vector<int> order;
    std::vector<int> bounds;  // supernode column boundaries
    int leaf;
};

double coord(const Model& M, int v, int ax) { return M.rest[3 * v + ax]; }

// One coordinate bisection of `set` on the graph of K: the smaller of the two one-sided
// boundaries becomes the separator `sep`, the rest of each side L and R. `set` is consumed.
void bisect(NDState& st, std::vector<int>& set, std::vector<int>& L, std::vector<int>& R, std::vector<int>& sep) {
    const Model& M = *st.M;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = coord(M, set[0], k);
    for (int v : set)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], coord(M, v, k));
            hi[k] = std::max(hi[k], coord(M, v, k));
        }
    int ax = 0;
    for (int k = 1; k < 3; ++k) if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
    auto less = [&](int a, int b) {
        double ca = coord(M, a, ax), cb = coord(M, b, ax);
        return ca < cb || (ca == cb && a < b);
    };
    size_t half = set.size() / 2;
    std::nth_element(set.begin(), set.begin() + half, set.end(), less);
    L.assign(set.begin(), set.begin() + half);
    R.assign(set.begin() + half, set.end());
    set.clear();
    set.shrink_to_fit();
    auto boundary = [&](const std::vector<int>& A, char other, std::vector<int>& sp, std::vector<int>& rest) {
        for (int v : A) {
            bool touch = false;
            for (int64_t p = M.ata_ptr[v]; p < M.ata_ptr[v + 1] && !touch; ++p) touch = st.side[M.ata_idx[p]] == other;
            (touch ? sp : rest).push_back(v);
        }
    };
    for (int v : L) st.side[v] = 1;
    for (int v : R) st.side[v] = 2;
    std::vector<int> sepL, restL, sepR, restR;
    boundary(L, 2, sepL, restL);
    boundary(R, 1, sepR, restR);
    for (int v : L) st.side[v] = 0;
    for (int v : R) st.side[v] = 0;
    sep.clear();
    if (sepL.size() <= sepR.size()) {
        sep.swap(sepL);
        L.swap(restL);
    } else {
        sep.swap(sepR);
        R.swap(restR);
    }
}

// Node of the dissection tree: a leaf set, or a separator with the two halves it splits.
struct NDNode {
    std::vector<int> cols;  // ascending
    int kid[2] = {-1, -1};
    bool leaf = false;
    int height = 0;         // leaves 0, a separator one above its higher half
    bool absorb = false;    // its supernode also takes its halves' separators (a cross)
    int snode = -1;         // its supernode in the ordering without crosses (chooser's scratch)
};

// Recursive coordinate bisection into st-owned nodes; returns the node of `set` (-1 if empty).
int build_nd(NDState& st, std::vector<NDNode>& T, std::vector<int>& set) {
    if (set.empty()) return -1;
    if ((int)set.size() <= st.leaf) {
        NDNode nd;
        nd.leaf = true;
        nd.cols.swap(set);
        std::sort(nd.cols.begin(), nd.cols.end());
        T.push_back(std::move(nd));
        return (int)T.size() - 1;
    }
    std::vector<int> L, R, sep;
    bisect(st, set, L, R, sep);
    const int a = build_nd(st, T, L);
    const int b = build_nd(st, T, R);
    NDNode nd;
    nd.cols.swap(sep);
    std::sort(nd.cols.begin(), nd.cols.end());
    nd.kid[0] = a;
    nd.kid[1] = b;
    for (int k : {a, b})
        if (k >= 0) nd.height = std::max(nd.height, T[k].height + 1);
    T.push_back(std::move(nd));
    return (int)T.size() - 1;
}

// Post-order emission of node i's subtree: leaves and separators become supernodes; a node
// that absorbs gives one supernode to its halves' separators (their units) and its own. With
// `out`, node i's unit is appended there instead (it joins its parent's supernode).
void emit_nd(NDState& st, const std::vector<NDNode>& T, int i, std::vector<int>* out) {
    const NDNode& nd = T[i];
    if (nd.leaf) {
        st.order.insert(st.order.end(), nd.cols.begin(), nd.cols.end());
        st.bounds.push_back((int)st.order.size());
        return;
    }
    std::vector<int> unit;
    for (int k : nd.kid)
        if (k >= 0) emit_nd(st, T, k, nd.absorb && !T[k].leaf ? &unit : nullptr);
    unit.insert(unit.end(), nd.cols.begin(), nd.cols.end());
    if (out) {
        out->insert(out->end(), unit.begin(), unit.end());
    } else if (!unit.empty()) {
        st.order.insert(st.order.end(), unit.begin(), unit.end());
        st.bounds.push_back((int)st.order.size());
    }
}

// Below-row sets of the supernodes of an ordering (perm: new -> old, bounds: supernode column
// boundaries), in the new numbering; the symbolic step of the factorisation.
void below_rows(const Model& M, const std::vector<int>& perm, const std::vector<int>& bounds,
                std::vector<std::vector<int>>& R) {
    const int n = (int)perm.size(), S = (int)bounds.size() - 1;
    std::vector<int> iperm(n), snode_of(n), mark(n, -1);
    for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    for (int s = 0; s < S; ++s)
        for (int j = bounds[s]; j < bounds[s + 1]; ++j) snode_of[j] = s;
    std::vector<std::vector<int>> kids(S);
    R.assign(S, {});
    for (int s = 0; s < S; ++s) {
        const int cend = bounds[s + 1];
        std::vector<int>& buf = R[s];
        for (int j = bounds[s]; j < cend; ++j)
            for (int64_t p = M.ata_ptr[perm[j]]; p < M.ata_ptr[perm[j] + 1]; ++p) {
                int i = iperm[M.ata_idx[p]];
                if (i >= cend && mark[i] != s) { mark[i] = s; buf.push_back(i); }
            }
        for (int c : kids[s])
            for (int i : R[c])
                if (i >= cend && mark[i] != s) { mark[i] = s; buf.push_back(i); }
        if (!buf.empty()) kids[snode_of[*std::min_element(buf.begin(), buf.end())]].push_back(s);
    }
}

// Cost model of one substitution (forward + backward sweep and the dense top block) on MI355X,
// in microseconds, fitted to measurements of the 40k-vertex drape's orderings (3 to 5 swept
// levels, top blocks of 1,594 to 3,946 rows) and of the 64-cloth batch (6 and 9 levels): a swept
// level costs a launch-and-dependent-load latency, plus its operator bytes (F forward, F^T
// backward) at a rate that depends on where they stream from. When the sweeps' working set
// (F, F^T and the top block's tiles) fits the 256 MB last-level cache it is replayed from there
// every iteration (the drape: 0.33 us per MB of F); far beyond it, from HBM (the batch: 0.56).
// The top block costs a fixed part (gather, tiles and finish launches) plus its tiles' bytes.
constexpr double kLevelUs = 12.0;           // forward + backward launch of one level
constexpr double kLevelUsPerMBCached = 0.33;
constexpr double kLevelUsPerMBHbm = 0.56;
constexpr double kCachedMB = 192.;          // working set served from the last-level cache
constexpr double kTopUs = 9.5;              // gather, tiles and finish launches
constexpr double kTopUsPerMB = 0.25;

// Chooses the crosses: the top-block height T (its columns m fit m^2 * e <= max_top_bytes) and
// the grouping of the separator heights T - 1 .. 1 into levels (groups of at most 3 consecutive
// heights, from the top; `forced` fixes the group sizes instead), minimising the cost model.
// Marks the absorbing nodes and returns the top block's budget for that choice.
double choose_crosses(const Model& M, int leaf, std::vector<NDNode>& T, const std::vector<int>& roots,
                      double max_top_bytes, int e, const std::vector<int>& forced) {
    int H = 0;
    for (const NDNode& nd : T) H = std::max(H, nd.height);
    std::vector<double> cols_at(H + 1, 0.);
    for (const NDNode& nd : T) cols_at[nd.height] += (double)nd.cols.size();
    // the ordering without crosses and its below-row sets
    NDState st0;
    st0.M = &M;
    st0.leaf = leaf;
    st0.bounds.push_back(0);
    for (NDNode& nd : T) nd.absorb = false;
    for (int r : roots)
        if (r >= 0) emit_nd(st0, T, r, nullptr);
    {
        // node -> supernode of that ordering (emission order: post-order, non-empty nodes)
        int next = 0;
        std::function<void(int)> walk = [&](int i) {
            for (int k : T[i].kid)
                if (k >= 0) walk(k);
            T[i].snode = T[i].cols.empty() ? -1 : next++;
        };
        for (int r : roots)
            if (r >= 0) walk(r);
    }
    std::vector<std::vector<int>> R;
    below_rows(M, st0.order, st0.bounds, R);
    // operator entries (ns + nb) ns of node i's supernode when it absorbs g - 1 levels below it
    const int n = M.n;
    std::vector<int> stamp(n, -1);
    int tick = 0;
    auto group_entries = [&](int i, int g) -> double {
        std::vector<int> members;
        std::function<void(int, int)> collect = [&](int j, int left) {
            if (T[j].leaf) return;
            if (T[j].snode >= 0) members.push_back(T[j].snode);
            if (left > 1)
                for (int k : T[j].kid)
                    if (k >= 0) collect(k, left - 1);
        };
        collect(i, g);
        ++tick;
        double ns = 0.;
        for (int s : members) {
            ns += st0.bounds[s + 1] - st0.bounds[s];
            for (int c = st0.bounds[s]; c < st0.bounds[s + 1]; ++c) stamp[c] = tick;
        }
        double nb = 0.;
        for (int s : members)
            for (int r : R[s])
                if (stamp[r] != tick) {
                    stamp[r] = tick;
                    nb += 1.;
                }
        return (ns + nb) * ns;
    };
    std::vector<std::array<double, 3>> ent(T.size(), {0., 0., 0.});
    double leaf_entries = 0.;
    for (size_t i = 0; i < T.size(); ++i) {
        if (T[i].leaf) {
            const int s = T[i].snode;
            const double ns = st0.bounds[s + 1] - st0.bounds[s];
            leaf_entries += (ns + (double)R[s].size()) * ns;
            continue;
        }
        if (T[i].cols.empty()) continue;
        for (int g = 1; g <= 3; ++g) ent[i][g - 1] = group_entries((int)i, g);
    }
    const double MB = e / 1e6;
    {
        // Merging levels trades operator bytes for launches, which pays while the levels are
        // latency-bound: when the sweeps' working set is replayed from the last-level cache. A
        // scene whose unmerged sweeps stream from HBM keeps one level per bisection (the 64-cloth
        // batch, 1.1 GB per iteration, measured 2% slower with merged levels).
        double mt = 0., bytes0 = leaf_entries;
        int top0 = H + 1;
        for (int h = H; h >= 0; --h) {
            if ((mt + cols_at[h]) * (mt + cols_at[h]) * e > max_top_bytes) break;
            mt += cols_at[h];
            top0 = h;
        }
        for (size_t i = 0; i < T.size(); ++i)
            if (!T[i].leaf && T[i].height < top0) bytes0 += ent[i][0];
        const double nb0 = std::ceil(mt / 64.);
        const double ws0 = 2. * bytes0 * MB + (mt > 0. ? 4096. * nb0 * (nb0 + 1.) / 2. * MB : 0.);
        if (top0 == 0 || ws0 > kCachedMB) {
            for (NDNode& nd : T) nd.absorb = false;
            return max_top_bytes;
        }
    }
    double best_cost = 1e300, best_bytes = 0.;
    std::vector<int> best_tops;   // group-top heights
    std::vector<int> best_sizes;  // their group sizes
    // A scene whose top level fits the budget keeps a top block: on the 64-cloth batch, sweeping
    // the tallest trees' roots instead (merged with the level below) measured 11% slower although
    // the model saw one launch less.
    const bool need_top = cols_at[H] * cols_at[H] * e <= max_top_bytes;
    double m = 0.;
    for (int top = H + 1; top >= 0; --top) {
        if (top <= H) {
            m += cols_at[top];
            if (m * m * e > max_top_bytes) break;
        } else if (need_top) {
            continue;
        }
        const double nbk = std::ceil(m / 64.);
        const double top_cost = m > 0. ? kTopUs + kTopUsPerMB * (4096. * nbk * (nbk + 1.) / 2.) * MB : 0.;
        // compositions of the heights top - 1 .. 1 into groups of 1 to 3
        std::vector<int> sizes;
        std::function<void(int)> rec = [&](int h) {
            if (h < 1) {
                // top 0: every level in the top block, nothing swept
                double bytes = top > 0 ? leaf_entries : 0., levels = top > 0 ? 1. : 0.;
                std::vector<int> tops;
                int hh = top - 1;
                for (int g : sizes) {
                    tops.push_back(hh);
                    for (size_t i = 0; i < T.size(); ++i)
                        if (!T[i].leaf && T[i].height == hh) bytes += ent[i][g - 1];
                    levels += 1.;
                    hh -= g;
                }
                const double tiles_mb = m > 0. ? 4096. * nbk * (nbk + 1.) / 2. * MB : 0.;
                const double ws = 2. * bytes * MB + tiles_mb;
                const double hit = ws > kCachedMB ? kCachedMB / ws : 1.;
                const double per_mb = hit * kLevelUsPerMBCached + (1. - hit) * kLevelUsPerMBHbm;
                const double cost = levels * kLevelUs + per_mb * bytes * MB + top_cost;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_bytes = (m + 0.5) * (m + 0.5) * e;
                    best_tops = tops;
                    best_sizes = sizes;
                }
                return;
            }
            if (!forced.empty()) {
                const int g = std::max(1, std::min(3, forced[std::min(sizes.size(), forced.size() - 1)]));
                sizes.push_back(std::min(g, h));
                rec(h - sizes.back());
                sizes.pop_back();
                return;
            }
            for (int g = 1; g <= std::min(3, h); ++g) {
                sizes.push_back(g);
                rec(h - g);
                sizes.pop_back();
            }
        };
        rec(top - 1);
    }
    // mark the absorbing nodes top-down: a node at a group-top height absorbs size - 1 levels of
    // separators below it (each absorbed separator passes one level less on)
    std::vector<int> size_at(H + 2, 0);
    for (size_t k = 0; k < best_tops.size(); ++k) size_at[best_tops[k]] = best_sizes[k];
    std::function<void(int, int)> mark = [&](int i, int inherited) {
        NDNode& nd = T[i];
        if (nd.leaf) return;
        const int c = inherited >= 0 ? inherited : std::max(0, size_at[nd.height] - 1);
        nd.absorb = c >= 1 && !nd.cols.empty();
        for (int k : nd.kid)
            if (k >= 0) mark(k, nd.absorb ? c - 1 : -1);
    };
    for (int r : roots)
        if (r >= 0) mark(r, -1);
    return best_bytes;
}

// Dense lower-triangular inverse, column-major ld x n block starting at a.
void lower_inverse(const double* Lm, int ld, int n, std::vector<double>& inv) {
    inv.assign((size_t)n * n, 0.);
    for (int j = 0; j < n; ++j) {
        // solve L x = e_j, x stored in column j of inv (rows >= j)
        double* x = &inv[(size_t)j * n];
        x[j] = 1. / Lm[(size_t)j * ld + j];
        for (int i = j + 1; i < n; ++i) {
            double s = 0.;
            for (int k = j; k < i; ++k) s += Lm[(size_t)k * ld + i] * x[k];
            x[i] = -s / Lm[(size_t)i * ld + i];
        }
    }
}

}  // namespace

void factorize(const Model& M, double dt2, Factor& F, int leaf_size, double cross_top_bytes, int entry_bytes,
               const std::vector<int>& groups) {
    auto t0 = std::chrono::steady_clock::now();
    const int n = M.n;
    F.n = n;
    NDState st;
    st.M = &M;
    st.side.assign(n, 0);
    st.leaf = std::max(8, leaf_size);
    st.bounds.push_back(0);
    std::vector<NDNode> T;
    std::vector<int> roots;
    {
        // Connected components of the graph of K (independent cloth pieces, batched
        // instances) are dissected one after another, so each component gets the
        // ordering and supernodes it would get alone and the elimination forest has
        // one tree per component. Components are numbered by their smallest vertex.
        std::vector<int> comp(n, -1), set, queue;
        for (int s0 = 0; s0 < n; ++s0) {
            if (comp[s0] >= 0) continue;
            set.clear();
            queue.assign(1, s0);
            comp[s0] = s0;
            for (size_t h = 0; h < queue.size(); ++h) {
                int v = queue[h];
                set.push_back(v);
                for (int64_t p = M.ata_ptr[v]; p < M.ata_ptr[v + 1]; ++p) {
                    int u = M.ata_idx[p];
                    if (comp[u] < 0) {
                        comp[u] = s0;
                        queue.push_back(u);
                    }
                }
            }
            std::sort(set.begin(), set.end());
            roots.push_back(build_nd(st, T, set));
        }
    }
    F.top_bytes = 0.;
    if (cross_top_bytes > 0.) {
        // Crosses below the dense top block: the separator levels under it are merged in groups of
        // consecutive levels (a separator absorbs its halves' separators, and theirs), each group
        // one level of the substitution. Fewer levels mean fewer latency-bound launches per
        // substitution for denser operators; the top block's size and the grouping minimise the
        // cost model above.
        F.top_bytes = choose_crosses(M, st.leaf, T, roots, cross_top_bytes, entry_bytes, groups);
    }
    for (int r : roots)
        if (r >= 0) emit_nd(st, T, r, nullptr);
    if ((int)st.order.size() != n) throw std::logic_error("ordering lost vertices");
    F.perm = st.order;
    F.iperm.assign(n, 0);
    for (int k = 0; k < n; ++k) F.iperm[F.perm[k]] = k;
    const int S = (int)st.bounds.size() - 1;
    F.c0.resize(S);
    F.ns.resize(S);
    std::vector<int> snode_of(n);
    for (int s = 0; s < S; ++s) {
        F.c0[s] = st.bounds[s];
        F.ns[s] = st.bounds[s + 1] - st.bounds[s];
        for (int j = F.c0[s]; j < F.c0[s] + F.ns[s]; ++j) snode_of[j] = s;
    }
    // ---- symbolic: below-row structure and supernodal elimination tree
    F.nb.assign(S, 0);
    F.parent.assign(S, -1);
    F.roff.assign(S + 1, 0);
    F.rows.clear();
    std::vector<std::vector<int>> kids(S);
    std::vector<int> mark(n, -1), buf;
    for (int s = 0; s < S; ++s) {
        const int cend = F.c0[s] + F.ns[s];
        buf.clear();
        for (int j = F.c0[s]; j < cend; ++j) {
            int old = F.perm[j];
            for (int64_t p = M.ata_ptr[old]; p < M.ata_ptr[old + 1]; ++p) {
                int i = F.iperm[M.ata_idx[p]];
                if (i >= cend && mark[i] != s) { mark[i] = s; buf.push_back(i); }
            }
        }
        for (int c : kids[s])
            for (int64_t p = F.roff[c]; p < F.roff[c + 1]; ++p) {
                int i = F.rows[p];
                if (i >= cend && mark[i] != s) { mark[i] = s; buf.push_back(i); }
            }
        std::sort(buf.begin(), buf.end());
        F.rows.insert(F.rows.end(), buf.begin(), buf.end());
        F.roff[s + 1] = (int64_t)F.rows.size();
        F.nb[s] = (int)buf.size();
        if (!buf.empty()) {
            int par = snode_of[buf.front()];
            F.parent[s] = par;
            kids[par].push_back(s);
        }
    }
    F.choff.assign(S + 1, 0);
    F.children.clear();
    for (int s = 0; s < S; ++s) {
        F.children.insert(F.children.end(), kids[s].begin(), kids[s].end());
        F.choff[s + 1] = (int64_t)F.children.size();
    }
    F.level.assign(S, 0);
    F.n_levels = 0;
    for (int s = 0; s < S; ++s) {
        for (int c : kids[s]) F.level[s] = std::max(F.level[s], F.level[c] + 1);
        F.n_levels = std::max(F.n_levels, F.level[s] + 1);
    }
    F.uoff.assign(S + 1, 0);
    F.foff.assign(S + 1, 0);
    F.max_ns = F.max_front = 0;
    F.nnz_L = 0;
    for (int s = 0; s < S; ++s) {
        F.uoff[s + 1] = F.uoff[s] + F.nb[s];
        F.foff[s + 1] = F.foff[s] + (int64_t)(F.ns[s] + F.nb[s]) * F.ns[s];
        F.max_ns = std::max(F.max_ns, F.ns[s]);
        F.max_front = std::max(F.max_front, F.ns[s] + F.nb[s]);
        F.nnz_L += (int64_t)F.ns[s] * (F.ns[s] + 1) / 2 + (int64_t)F.nb[s] * F.ns[s];
    }
    F.mapoff = F.uoff;  // one relative index per below row
    F.relmap.assign(F.uoff[S], 0);
    F.F.assign(F.foff[S], 0.);
    F.FT.assign(F.foff[S], 0.);
    // ---- numeric multifrontal factorisation
    std::vector<std::vector<double>> update(S);  // nb x nb lower, column-major
    std::vector<int> pos(n, -1);
    std::vector<double> front, inv;
    for (int s = 0; s < S; ++s) {
        const int ns = F.ns[s], nb = F.nb[s], f = ns + nb, c0 = F.c0[s];
        for (int j = 0; j < ns; ++j) pos[c0 + j] = j;
        for (int k = 0; k < nb; ++k) pos[F.rows[F.roff[s] + k]] = ns + k;
        front.assign((size_t)f * f, 0.);
        auto at = [&](int i, int j) -> double& { return front[(size_t)j * f + i]; };
        for (int j = 0; j < ns; ++j) {
            int col = c0 + j, old = F.perm[col];
            for (int64_t p = M.ata_ptr[old]; p < M.ata_ptr[old + 1]; ++p) {
                int i = F.iperm[M.ata_idx[p]];
                if (i < col) continue;
                double v = dt2 * M.ata_val[p];
                if (i == col) v = M.mass[old] + v;
                at(pos[i], j) += v;
            }
        }
        for (int64_t q = F.choff[s]; q < F.choff[s + 1]; ++q) {
            int c = F.children[q];
            const int nbc = F.nb[c];
            const int* Rc = &F.rows[F.roff[c]];
            int* rel = &F.relmap[F.mapoff[c]];
            for (int k = 0; k < nbc; ++k) rel[k] = pos[Rc[k]];
            const std::vector<double>& U = update[c];
            for (int b = 0; b < nbc; ++b)
                for (int a = b; a < nbc; ++a) at(rel[a], rel[b]) += U[(size_t)b * nbc + a];
            std::vector<double>().swap(update[c]);
        }
        // partial dense Cholesky of the first ns columns (right-looking)
        for (int k = 0; k < ns; ++k) {
            double d = at(k, k);
            if (!(d > 0.)) throw std::runtime_error("K is not positive definite");
            d = std::sqrt(d);
            at(k, k) = d;
            for (int i = k + 1; i < f; ++i) at(i, k) /= d;
            for (int j = k + 1; j < f; ++j) {
                double ljk = at(j, k);
                if (ljk == 0.) continue;
                double* cj = &front[(size_t)j * f];
                const double* ck = &front[(size_t)k * f];
                for (int i = j; i < f; ++i) cj[i] -= ck[i] * ljk;
            }
        }
        if (nb > 0) {
            std::vector<double>& U = update[s];
            U.assign((size_t)nb * nb, 0.);
            for (int b = 0; b < nb; ++b)
                for (int a = b; a < nb; ++a) U[(size_t)b * nb + a] = at(ns + a, ns + b);
        }
        // F_s = [Linv ; L_bs Linv], F_s^T
        lower_inverse(front.data(), f, ns, inv);
        double* Fs = &F.F[F.foff[s]];
        double* FTs = &F.FT[F.foff[s]];
        for (int j = 0; j < ns; ++j) {
            for (int i = 0; i < ns; ++i) Fs[(size_t)j * f + i] = inv[(size_t)j * ns + i];
            for (int r = 0; r < nb; ++r) {
                double acc = 0.;
                for (int k = j; k < ns; ++k) acc += at(ns + r, k) * inv[(size_t)j * ns + k];
                Fs[(size_t)j * f + ns + r] = acc;
            }
        }
        for (int j = 0; j < ns; ++j)
            for (int i = 0; i < f; ++i) FTs[(size_t)i * ns + j] = Fs[(size_t)j * f + i];
        for (int j = 0; j < ns; ++j) pos[c0 + j] = -1;
        for (int k = 0; k < nb; ++k) pos[F.rows[F.roff[s] + k]] = -1;
    }
    F.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void plan_top(const Factor& F, double max_bytes, int entry_bytes, TopPlan& top) {
    const int S = (int)F.ns.size();
    std::vector<int64_t> cols_at(F.n_levels + 1, 0);
    for (int s = 0; s < S; ++s) cols_at[F.level[s]] += F.ns[s];
    int64_t m = 0;
    int best = F.n_levels;
    for (int L = F.n_levels - 1; L >= 0; --L) {
        m += cols_at[L];
        if ((double)m * (double)m * entry_bytes <= max_bytes) best = L;
        else break;
    }
    top = TopPlan();
    top.level = best;
    std::vector<int> tpos(F.n, -1);
    for (int s = 0; s < S; ++s)
        if (F.level[s] >= best)
            for (int j = 0; j < F.ns[s]; ++j) top.rows.push_back(F.c0[s] + j);
    std::sort(top.rows.begin(), top.rows.end());
    top.m = (int)top.rows.size();
    for (int t = 0; t < top.m; ++t) tpos[top.rows[t]] = t;
    std::vector<std::vector<int>> contrib(top.m);
    for (int c = 0; c < S; ++c) {
        if (F.level[c] >= best || F.parent[c] < 0 || F.level[F.parent[c]] < best) continue;
        for (int k = 0; k < F.nb[c]; ++k) {
            int t = tpos[F.rows[F.roff[c] + k]];
            if (t < 0) throw std::logic_error("update row outside the top block");
            contrib[t].push_back((int)(F.uoff[c] + k));
        }
    }
    top.cptr.assign(top.m + 1, 0);
    for (int t = 0; t < top.m; ++t) {
        top.cptr[t + 1] = top.cptr[t] + (int)contrib[t].size();
        top.cidx.insert(top.cidx.end(), contrib[t].begin(), contrib[t].end());
    }
}

void plan_sweeps(const Factor& F, int chunk, int top_level, SweepPlan& fwd, SweepPlan& bwd) {
    const int S = (int)F.ns.size();
    const int L = std::min(top_level, F.n_levels);
    std::vector<std::vector<int>> by_level(F.n_levels);
    for (int s = 0; s < S; ++s) by_level[F.level[s]].push_back(s);
    fwd = SweepPlan();
    bwd = SweepPlan();
    fwd.level_ptr.push_back(0);
    for (int l = 0; l < L; ++l) {
        int kmax = 0;
        for (int s : by_level[l]) {
            kmax = std::max(kmax, F.ns[s]);
            for (int r0 = 0; r0 < F.ns[s] + F.nb[s]; r0 += chunk) {
                fwd.task_snode.push_back(s);
                fwd.task_row0.push_back(r0);
            }
        }
        fwd.level_ptr.push_back((int)fwd.task_snode.size());
        fwd.level_kmax.push_back(kmax);
    }
    bwd.level_ptr.push_back(0);
    for (int l = L - 1; l >= 0; --l) {
        int kmax = 0;
        for (int s : by_level[l]) {
            kmax = std::max(kmax, F.ns[s] + F.nb[s]);
            for (int j0 = 0; j0 < F.ns[s]; j0 += chunk) {
                bwd.task_snode.push_back(s);
                bwd.task_row0.push_back(j0);
            }
        }
        bwd.level_ptr.push_back((int)bwd.task_snode.size());
        bwd.level_kmax.push_back(kmax);
    }
}

void plan_merged(const Factor& F, const SweepPlan& P, bool forward, int first, int nl, MergedGroups& out) {
    const int L = (int)P.level_ptr.size() - 1;
    if (first < 0 || nl < 1 || first + nl > L) throw std::invalid_argument("merged range outside the sweep");
    auto level_of = [&](int l) { return forward ? l : L - 1 - l; };
    const int hi = std::max(level_of(first), level_of(first + nl - 1));
    std::vector<int> gid(F.ns.size(), -1);  // root supernode -> group
    std::vector<std::vector<std::vector<int>>> gt;  // group -> level in range -> tasks
    for (int k = 0; k < nl; ++k)
        for (int t = P.level_ptr[first + k]; t < P.level_ptr[first + k + 1]; ++t) {
            int r = P.task_snode[t];
            while (F.parent[r] >= 0 && F.level[F.parent[r]] <= hi) r = F.parent[r];
            if (gid[r] < 0) {
                gid[r] = (int)gt.size();
                gt.emplace_back(nl);
            }
            gt[gid[r]][k].push_back(t);
        }
    out = MergedGroups();
    out.groups = (int)gt.size();
    out.tmax.assign(nl, 0);
    int acc = 0;
    for (const auto& g : gt)
        for (int k = 0; k <= nl; ++k) {
            out.ptr.push_back(acc);
            if (k < nl) {
                acc += (int)g[k].size();
                out.tmax[k] = std::max(out.tmax[k], (int)g[k].size());
            }
        }
    for (const auto& g : gt)
        for (int k = 0; k < nl; ++k) out.task.insert(out.task.end(), g[k].begin(), g[k].end());
}

void plan_backward_split(const Factor& F, const SweepPlan& bwd, const std::vector<int>& split, SplitPlan& out) {
    const int L = (int)bwd.level_ptr.size() - 1;
    if ((int)split.size() != L) throw std::invalid_argument("one split count per backward level");
    for (int l = 0; l < L; ++l)
        if (split[l] < 1 || split[l] > kMaxSplit || (l == L - 1 && split[l] != 1))
            throw std::invalid_argument("split parts must be 1..kMaxSplit, and 1 for the last level");
    out = SplitPlan();
    out.tasks.level_ptr.push_back(0);
    out.tasks.level_kmax = bwd.level_kmax;
    for (int l = 0; l < L; ++l) {
        const int P = split[l];
        for (int t = bwd.level_ptr[l]; t < bwd.level_ptr[l + 1]; ++t) {
            const int s = bwd.task_snode[t];
            const int f = F.ns[s] + F.nb[s];
            const int chunk = (f + P - 1) / P;
            for (int p = 0; p < P; ++p) {
                out.tasks.task_snode.push_back(s);
                out.tasks.task_row0.push_back(bwd.task_row0[t]);
                out.part.push_back({p, P, std::min(f, p * chunk), std::min(f, (p + 1) * chunk)});
            }
        }
        out.tasks.level_ptr.push_back((int)out.tasks.task_snode.size());
    }
    out.rparts.assign(F.rows.size(), 0);
    out.fin_rows.assign(L, {});
    out.fin_parts.assign(L, 0);
    // backward sweep index l holds level L - 1 - l
    std::vector<int> owner(F.n, -1);
    for (size_t s = 0; s < F.ns.size(); ++s)
        for (int k = 0; k < F.ns[s]; ++k) owner[F.c0[s] + k] = (int)s;
    for (size_t s = 0; s < F.ns.size(); ++s) {
        const int lv = F.level[s];
        if (lv + 1 >= L) continue;  // not swept, or its parent level is not
        for (int t = 0; t < F.nb[s]; ++t) {
            const int a = owner[F.rows[F.roff[s] + t]];
            if (F.level[a] == lv + 1 && split[L - 1 - (lv + 1)] > 1) out.rparts[F.roff[s] + t] = split[L - 1 - (lv + 1)];
        }
    }
    for (int l = 1; l < L; ++l) {
        if (split[l - 1] < 2) continue;
        out.fin_parts[l] = split[l - 1];
        for (size_t s = 0; s < F.ns.size(); ++s)
            if (F.level[s] == L - l)
                for (int k = 0; k < F.ns[s]; ++k) out.fin_rows[l].push_back(F.c0[s] + k);
    }
}

}  // namespace pdfc
