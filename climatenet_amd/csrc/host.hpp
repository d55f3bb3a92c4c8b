// Host-side model of one cloth scene: topology, constraint stencils, the
// constant system matrix K = M + dt^2 A^T A, its fill-reducing ordering and
// the supernodal FP64 factorisation that the device substitution kernels use.
//
// Everything here runs once per pdfc_create (except the self-contact pairing,
// which runs once per step when self-collision is on).
#pragma once

#include <array>
#include <functional>
#include <cstdint>
#include <string>
#include <vector>

#include "pdfc.h"

namespace pdfc {

struct V3 {
    double x, y, z;
};

// Scene-level data in the scene's own (OBJ) vertex numbering.
struct Model {
    int n = 0, T = 0;
    std::vector<std::array<int, 3>> tris;
    std::vector<double> rest;  // 3n
    std::vector<double> mass;  // n
    // springs (rows 0..E-1 of A): SpringConstraint, src/physics/SpringConstraint.cpp:42-69
    std::vector<std::array<int, 2>> edges;
    std::vector<double> spring_w;   // sqrt(k_s)
    std::vector<double> spring_L0;  // rest length
    // bending rows (E..E+B-1): BendingConstraint, src/physics/BendingConstraint.cpp:47-104
    std::vector<int> bend_ptr;      // B+1, into bend_idx/bend_coef
    std::vector<int> bend_idx;      // [center, ring...]
    std::vector<double> bend_coef;  // weighted row entries sqrt(k_b) * beltrami coefficient
    // attachment rows (after the bending rows): AttachmentConstraint,
    // src/physics/AttachmentConstraint.cpp:23-55 (row weight w, projection w * position)
    std::vector<int> att_vertex;
    std::vector<double> att_w;
    std::vector<double> att_pos;    // 3 per attachment
    std::vector<char> border;
    std::vector<std::vector<int>> neighbors;  // one ring, ascending
    // A^T A, symmetric, CSR with ascending columns
    std::vector<int64_t> ata_ptr;
    std::vector<int> ata_idx;
    std::vector<double> ata_val;
    int max_degree = 0;    // springs per vertex
    int max_ata_row = 0;   // non-zeros per row of A^T A
};

// Builds topology, masses, springs, bending rows and A^T A. Throws std::invalid_argument.
void build_model(const pdfc_scene& sc, Model& M);
// Disjoint union of independently built scenes (batched instances).
void merge_models(const std::vector<Model>& parts, Model& M);

// Supernodal factor of P K P^T.
struct Factor {
    int n = 0;
    std::vector<int> perm;   // new -> old
    std::vector<int> iperm;  // old -> new
    // supernode s owns columns [c0[s], c0[s]+ns[s]) and below-rows R (ascending, new numbering)
    std::vector<int> c0, ns, nb, parent, level;
    std::vector<int64_t> roff;       // into rows
    std::vector<int> rows;
    std::vector<int64_t> uoff;       // update-vector offset (nb entries)
    std::vector<int64_t> mapoff;     // relative map of s's below rows into parent front (nb entries)
    std::vector<int> relmap;
    std::vector<int64_t> choff;      // children list offsets (n_snodes+1)
    std::vector<int> children;
    // F_s = [L_ss^-1 ; L_bs L_ss^-1], column-major (ns+nb) x ns, and its transpose
    // F_s^T column-major ns x (ns+nb).
    std::vector<int64_t> foff;       // same offsets for F and F^T
    std::vector<double> F, FT;
    int64_t nnz_L = 0;
    int n_levels = 0;
    int max_ns = 0, max_front = 0;
    double top_bytes = 0.;           // with crosses: the top-block budget the ordering was chosen for
    double seconds = 0.;
};

// Nested dissection (coordinate bisection on the graph of K), symbolic and
// numeric multifrontal factorisation. K = diag(mass) + dt2 * A^T A.
// With cross_top_bytes > 0, the separator levels below a dense top block of at most that budget
// (entries of entry_bytes) are merged in groups of consecutive levels, each group one level of
// cross-shaped supernodes; the top block's size and the groups are chosen by a cost model of the
// substitution (F.top_bytes: the budget to plan the top block with). `groups` fixes the group
// sizes from the top instead (the last repeating).
void factorize(const Model& M, double dt2, Factor& F, int leaf_size = 64, double cross_top_bytes = 0.,
               int entry_bytes = 8, const std::vector<int>& groups = {});

// Work lists of one substitution sweep: (supernode, first row) tasks of at most
// `chunk` rows, grouped by level (forward: leaves first; backward: root first).
struct SweepPlan {
    std::vector<int> task_snode, task_row0;
    std::vector<int> level_ptr;  // tasks of level l are [level_ptr[l], level_ptr[l+1])
    std::vector<int> level_kmax; // longest reduction of the level's tasks
};
// Levels below `top_level` only; supernodes at levels >= top_level form the dense top block.
void plan_sweeps(const Factor& F, int chunk, int top_level, SweepPlan& fwd, SweepPlan& bwd);

// Groups of one merged launch over the sweep levels [first, first + nl) of P (forward: leaves
// first; backward: root side first). A group is the supernodes of the range under one root,
// their highest ancestor whose level is inside the range; groups are numbered in order of
// first appearance in P's task order and keep P's task order at each level. Every child
// (forward) or ancestor (backward) a task depends on is then in the task's group at an
// earlier level of the range, or outside the range.
struct MergedGroups {
    int groups = 0;
    std::vector<int> ptr;    // groups x (nl + 1): group g's tasks at level k are task[ptr[g (nl+1) + k] ..
    std::vector<int> task;   //   ptr[g (nl+1) + k + 1]) (indices into P's task list)
    std::vector<int> tmax;   // nl: most tasks of one group at each level
};
void plan_merged(const Factor& F, const SweepPlan& P, bool forward, int first, int nl, MergedGroups& out);

// Backward sweep with split reductions: level l's tasks (supernode, 64 columns) each become
// split[l] tasks ("parts") over contiguous ranges of the front's rows, whose partial results
// the next level's tasks sum where they gather those rows and whose final values extra
// workgroups of the next level's launch write. The last level (the leaves) is never split.
constexpr int kMaxSplit = 4;  // parts of a split backward reduction
struct SplitPlan {
    SweepPlan tasks;                          // the tasks as launched, grouped by level as bwd
    std::vector<std::array<int, 4>> part;     // per task: part index, parts, first row, end row
    std::vector<int> rparts;                  // parallel to Factor::rows: partials to sum (0: final)
    std::vector<std::vector<int>> fin_rows;   // per level: rows the launch finishes (previous level's)
    std::vector<int> fin_parts;               // per level: their parts
};
void plan_backward_split(const Factor& F, const SweepPlan& bwd, const std::vector<int>& split, SplitPlan& out);

// Dense top block: the columns T of all supernodes at levels >= level. Both
// substitutions over T collapse into x_T = S^-1 b~_T with S the Schur complement
// of K on T (S^-1 = (K^-1)_TT), where b~_T subtracts the update vectors of the
// lower supernodes whose parent is in T.
struct TopPlan {
    int level = 0;             // first collapsed level; == n_levels when there is no top block
    int m = 0;
    std::vector<int> rows;     // supernodal indices of T, ascending
    std::vector<int> cptr;     // m + 1
    std::vector<int> cidx;     // update-vector entries (index into U) to subtract from each top row
};
// Largest top block with m^2 * entry_bytes <= max_bytes (at least one level collapsed
// when possible; the whole matrix for small meshes).
void plan_top(const Factor& F, double max_bytes, int entry_bytes, TopPlan& top);

// One self-collision candidate that survived the exact distance test and the
// pairing pre-filter (vertex outside the face, no zero barycentric coordinate).
// Written by the device, read by the host pairing: layout shared with the kernels.
struct SelfCandidate {
    int v, t;          // scene vertex and triangle index
    double bary[3];    // barycentric coordinates of the closest point
    double normal[3];  // (x_v - closest point) / |.|
};

struct SelfPair {
    int v, partner;
    std::array<int, 3> face;
    int best;       // which face corner is the partner
    V3 normal;
    int level;
};
// Contact basis rows [n, t, b] (computeBasis, SimulationFrictionEstimation.cpp:89-105).
void host_contact_basis(const double n[3], double B[9]);

// Host workspace of the per-step self-contact pairing and level ordering: flat arrays sized
// once per model and reused, so a step allocates nothing once they have grown.
struct SelfWorkspace {
    static constexpr int kRegCap = 16;  // registered partners per vertex held in place
    int n = 0;
    std::vector<int> nbr_ptr, nbr_idx;  // one ring, CSR (ascending)
    std::vector<int> reg_cnt, reg;      // registered partners: count and kRegCap slots per vertex
    std::vector<int> touched;           // vertices with a non-zero count
    std::vector<int> vid;               // level ordering: vertex -> graph id (-1 outside the graph)
    std::vector<unsigned long long> vbits;  // graph vertices, for ascending order
    std::vector<int> verts, adj_ptr, adj_fill, adj_v, adj_s, frontier, next;
    std::vector<char> visited;
    void init(const Model& M);
};

// Greedy node-node pairing over candidates sorted by (vertex, triangle), general form.
void pair_self_contacts(const Model& M, const SelfCandidate* cand, size_t ncand, std::vector<SelfPair>& out);
// The same pairing on the workspace's flat arrays; falls back to the general form when a vertex
// collects more than kRegCap registered partners (identical results either way).
void pair_self_contacts(SelfWorkspace& W, const Model& M, const SelfCandidate* cand, size_t ncand,
                        std::vector<SelfPair>& out);
// Computation levels: BFS from the obstacle-contact vertices (is_seed), then from degree-1
// and remaining vertices in ascending order. Fills pair.level, the contact indices level by
// level (order) and level_ptr (levels + 1 entries); returns the number of levels. Work is
// O(pairs + n / 64): is_seed is asked only for vertices of the pairs.
int order_self_contacts(SelfWorkspace& W, const std::function<bool(int)>& is_seed, std::vector<SelfPair>& pairs,
                        std::vector<int>& order, std::vector<int>& level_ptr);

}  // namespace pdfc
