// Device-side data layout and kernel launchers (gfx950).
//
// All per-vertex vectors are planar SoA in the supernodal (nested-dissection)
// numbering: component c of vertex i lives at vec[c*n + i]. Stencils are ELL,
// slot-major (slot*n + i), so a wave reading 64 consecutive vertices' slot k
// issues one coalesced 256/512-byte load.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <vector>

#include "host.hpp"

namespace pdfc {

constexpr int kMaxObstacles = 16;  // per scene (one contact bit each)
constexpr int kSolveChunk = 64;  // rows per substitution task (one wave)

// Triangle-mesh obstacle (StaticMesh / MovingMesh): current positions, the key frames
// around the current time (point speeds), the triangle grid and the per-vertex result
// of the closest-triangle query of the step.
constexpr int kMaxMeshObstacles = 4;
struct DevMeshObs {
    int obstacle;           // index among the obstacles
    int nv, T;
    int moving;             // 1: the clock is between two key frames (point speeds)
    const int* tris;        // 3T
    const double* cur;      // 3 nv, xyz interleaved
    const double* prev;     // key frames around the current time (3 nv each)
    const double* next;
    double inv_dt;          // 1 / (t_next - t_prev)
    double tol, tol_sq, t11;
    const unsigned long long* ext_bits;  // largest grown triangle box extent (bit pattern of a
                                         // non-negative double); cell size max(ext (1 + 1e-9), 1e-12)
    int P;                  // hash buckets (power of two)
    const int* bucket_off;  // P + 1
    const int* bucket_tris;
    const double* bucket_pos;  // 9 per bucket slot: the triangle's vertex positions
    int* hit_tri;           // per cloth vertex: closest triangle within the tolerance, or -1
    double* hit_bary;       // 3 per cloth vertex
};

// Obstacles in device memory. With batched scenes a vertex meets only the
// obstacles of its own scene.
struct DevObstacles {
    int count;
    const int* type;       // enum pdfc_obstacle_type
    const double* p;       // 3 per obstacle: sphere centre, plane point, box corner (smallest coordinates)
    const double* nrm;     // 3 per obstacle: plane normal; box edge lengths; rotating sphere's
                           // angular velocity (speed times the unit axis)
    double* clock;         // per obstacle: a rotating sphere's time before it turns (the step's
                           // last kernel subtracts dt)
    const double* radius;
    const double* mu;      // friction coefficient against the scene's cloth
    const int* vinst;      // per vertex: its scene; null when there is one scene
    const int* obs_ptr;    // per scene k: obstacles [obs_ptr[k], obs_ptr[k+1])
    int nmesh;
    DevMeshObs mesh[kMaxMeshObstacles];
};
// Closest triangle of mesh obstacle m for every cloth vertex (positions qn, planar
// supernodal layout): hit_tri / hit_bary.
void launch_mesh_query(const DevObstacles& ob, int m, const double* qn, int n, hipStream_t st);

// Device build of a mesh obstacle's triangle grid from its current positions: every
// triangle is registered in each cell its box, grown by grow, overlaps (at most 2 x 2 x 2
// cells, since the cell is wider than any grown box).
struct DevMeshBuild {
    int T, P;
    const int* tris;                // 3T
    const double* cur;              // 3 nv, xyz interleaved
    double grow;                    // 1.1 tol (1 + 1e-9)
    double* tbox;                   // 6T: grown lo xyz, hi xyz
    unsigned long long* ext_bits;   // largest grown extent
    int* off;                       // P + 1: counts, then their exclusive scan
    int* fill;                      // P
    int* tris_sorted;               // 8T
    double* pos_sorted;             // 9 per slot of tris_sorted
    int* scan_sums;                 // one per 256 entries of off, plus one
    int* scan_total;
};
void launch_mesh_build(const DevMeshBuild& B, hipStream_t st);
// cur = prev (1 - a) + next a, elementwise over count doubles (MovingMesh::update's
// interpolation between two key frames, the host's operation order).
void launch_mesh_lerp(double* cur, const double* prev, const double* next, double a, int count, hipStream_t st);

// Obstacle contact record (structure of arrays, capacity n * n_obstacles).
struct DevContacts {
    int* vertex;     // supernodal vertex index
    int* obstacle;
    int* active;
    double* normal;  // 3 * cap, xyz interleaved
    double* point;   // 3 * cap
    double* mu;
    double* basis;   // 9 * cap: n, t, b
    double* r;       // 3 * cap, local frame
    double* speed;   // 3 * cap: obstacle velocity at the contact point
};

struct DevModel {
    int n;
    int sp_slots, ata_slots;
    const double* mass;
    const double* invmass;
    const int* sp_cnt;
    const int* sp_idx;       // sp_slots * n
    const double* sp_kappa;  // w^2 L0 per incident spring
    const int* ata_cnt;
    const int* ata_idx;      // ata_slots * n
    const double* ata_val;
    // attached vertices: (A^T p) of their attachment rows, sum of w^2 * position (constant)
    int n_att;
    const int* att_vertex;   // n_att distinct vertices
    const double* att_term;  // 3 * n_att, xyz interleaved
};

// One substitution task (supernode s, rows or columns r0 .. r0+63), everything the
// kernel needs in one 64-byte load. A backward task whose reduction is split over several
// workgroups ("parts") covers rows [i0, i1) of the front and writes a partial result:
// split = part index | parts << 16 (0: the whole reduction), rng = i0 | i1 << 32.
struct TaskDesc {
    int s, r0, ns, nb;
    int c0, f, pull_base, split;
    int64_t foff;
    int64_t uoff;
    int64_t roff;
    int64_t rng;
};

// Logical block of physical workgroup b of nb. The dispatcher hands workgroups to the 8 XCDs in
// turn (workgroup b to XCD b % 8), and each XCD has its own L2; this gives each XCD a contiguous
// range of logical blocks (b / 8-th of those of XCD b % 8), so that blocks whose gathers touch the
// same lines (neighbouring vertices, sibling supernodes) share one L2. A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b & 7, k = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + k;
}

// Kernel-duration timing. While a launch timer is current on the thread and has a family set,
// the launches made through launch_k record their own start and end (the events that
// hipExtLaunchKernelGGL attaches to the dispatch itself, so no separate markers are timed) under
// that family; the owner reads them after synchronising.
struct LaunchTimer {
    std::vector<hipEvent_t> ev;  // pairs
    std::vector<int> fam;        // family of each recorded pair
    size_t used = 0;             // events in use
    int family = -1;             // < 0: launches are not timed
};
extern thread_local LaunchTimer* g_launch_timer;
inline void launch_timer_events(hipEvent_t& a, hipEvent_t& b) {
    a = b = nullptr;
    LaunchTimer* t = g_launch_timer;
    if (!t || t->family < 0) return;
    if (t->used + 2 > t->ev.size())
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            t->ev.push_back(e);
        }
    a = t->ev[t->used];
    b = t->ev[t->used + 1];
    t->used += 2;
    t->fam.push_back(t->family);
}
template <typename... KArgs, typename... Args>
inline void launch_k(void (*kernel)(KArgs...), dim3 g, dim3 b, size_t lds, hipStream_t st, Args... args) {
    hipEvent_t e0, e1;
    launch_timer_events(e0, e1);
    if (e0)
        hipExtLaunchKernelGGL(kernel, g, b, (std::uint32_t)lds, st, e0, e1, 0u, static_cast<KArgs>(args)...);
    else
        hipLaunchKernelGGL(kernel, g, b, lds, st, static_cast<KArgs>(args)...);
}

// Device layout of the substitution operators: per supernode its F_s block, column-major with
// leading dimension op_ld(ns + nb), then its F_s^T block, column-major with leading dimension
// op_ld(ns), each block starting on a 128-byte boundary (pad entries zero). A dimension above 64
// is split over several 64-row tasks, so it is rounded up to 128 bytes: every task's 64 entries
// of a column (or of a row) then start on a cache-line boundary and no line is fetched by two
// tasks; a dimension of at most 64 is read whole by one task and stays unpadded.
__host__ __device__ inline int op_round(int64_t x, int entry_bytes) {
    const int a = 128 / entry_bytes;
    return (int)((x + a - 1) / a * a);
}
__host__ __device__ inline int op_ld(int x, int entry_bytes) { return x > 64 ? op_round(x, entry_bytes) : x; }
// offset of the F_s^T block inside a supernode's blocks
__host__ __device__ inline int64_t op_t_offset(int ns, int f, int entry_bytes) {
    const int a = 128 / entry_bytes;
    const int64_t b = (int64_t)op_ld(f, entry_bytes) * ns;
    return (b + a - 1) / a * a;
}

// Supernodal substitution operators (see factor.cpp).
struct DevFactor {
    int n;
    const int* c0;
    const int* ns;
    const int* nb;
    const int64_t* foff;
    const int64_t* roff;
    const int64_t* uoff;
    const int64_t* choff;
    const int* children;
    const int* relmap;   // indexed by uoff of the child
    const int* rows;
    const void* F;       // double or float: the padded F_s / F_s^T blocks (TaskDesc::foff)
    const void* FT;      //   the same array
    double* U;           // 3 * total below rows, interleaved
    const int* task_s;
    const int* task_r0;
    const TaskDesc* desc;   // per task of this sweep
    const int* pull_ptr;    // per supernode, f + 1 entries from pull_base: children's update
    const int* pull_idx;    //   entries (index into U) contributing to each front row, child order
    const int4* pull_rec;   // per front row (same index as pull_ptr): entries, the first two entries'
                            //   U indices, its list start (one load instead of a pointer-list chain)
    // backward sweep with split reductions: partial results XP[p][c][row] (parts x 3n), and per
    // below row (parallel to rows) the number of partials to sum for it (0: its value is in X)
    double* XP;
    const int* rparts;
};

struct StepArgs {
    double dt;
    double g[3];
    double damping;
    int record;        // accumulate Alart-Curnier norms
    double* ac_sum;    // per iteration
    unsigned long long* ac_max;  // per iteration, bit pattern of a non-negative double
    int ac_cap;        // entries of ac_sum / ac_max (cleared by k_step_begin when record is set)
    // warm start of obstacle contacts in the iterations launched with warm_mode != 0:
    // r = m_i * warm_y[cid * warm_stride + 0..2] (contact basis, per unit mass), projected on the cone
    int warm_mode;     // 0 plain, 1 previous step (plain where it is zero), 2 learned
    int warm_stride;
    const float* warm_y;
    float* prev_g;     // end of step: world reaction per (scene obstacle slot, vertex), 3 * n each; null: not kept
    int* nonfinite;    // set to 1 by the step's last kernel when a new position or velocity is not finite
};

// Self-collision detection buffers (scene numbering for vertices and triangles).
// A bucket slot's prefilter data: the cell of the triangle box's lower corner (21 bits per
// axis) and the box rounded outwards to float, so a test that the double box passes also passes.
struct alignas(16) BoxKey {
    long long key;
    float lo[3], hi[3];
};

struct DevSelfDetect {
    int n, T;
    const double* q;        // positions, supernodal planar layout
    const int* iperm;       // scene vertex -> supernodal index
    const int* tris;        // 3T scene vertex indices
    double tol, tol_sq, t11;  // tol = sqrt(tol*tol) of the parameters, t11 = 1.1 tol
    double* tbox;           // 6T: lo xyz, hi xyz
    unsigned long long* ext_bits;  // largest box extent (bit pattern)
    int P;                  // hash buckets (power of two)
    int* bucket_cnt;        // P + 1; zero between detections
    int* bucket_off;        // P + 1
    int* bucket_tris;       // T
    int* cnt_full;
    int* cnt_filt;
    int* off_full;
    int* off_filt;
    int* totals;            // [0] candidates, [1] filtered candidates, [2] triangles hashed
    int* scan_sums;
    int* scan_sums2;
    int* full_v;
    int* full_t;
    double* full_bary;
    SelfCandidate* filt;
    int cap_full, cap_filt;
    int write_full;         // also write the complete candidate list
    // single-pass path (filtered list only): triangle data sorted by bucket and a fixed number
    // of candidate slots per vertex
    struct BoxKey* bk_key;  // per bucket slot: the lower corner's cell and the box rounded outwards to float
    double* bk_pos;         // 9 per bucket slot: the three vertex positions
    int* bk_vid;            // 3 per bucket slot: scene vertex indices
    int* bk_tri;            // triangle index per bucket slot
    SelfCandidate* vslots;  // n * kSlotsPerVertex
    int* overflow;          // set when a vertex has more candidates than slots
};
constexpr int kSlotsPerVertex = 8;

// Self contacts of the current step, sorted by computation level.
struct DevSelf {
    int nsc, nslots;
    int ngroups;            // vertex groups over all levels (grp_ptr[levels])
    const int* a;           // contact vertex (supernodal index)
    const int* b;           // paired face vertex
    const int* sa;          // slot of a in the involved-vertex buffers
    const int* sb;
    const double* normal;   // 3 per contact, uploaded
    double* basis;          // 9 per contact: n, t, b, built from the normal by k_self_prepare
    const double* mu;
    double* r;              // local reaction (mass-weighted)
    int* active;
    double* rem;            // force applied in the previous iteration (3 per contact)
    double* addv;           // force of this iteration
    double* Fc;             // force estimate F at the involved vertices (3 per slot)
    double* CFc;            // obstacle contact forces, current / next buffer
    double* CFn;
    double* SCFc;           // self contact forces, current / next buffer
    double* SCFn;
    const int* slot_vertex; // slot -> supernodal vertex
    const double* mass2;    // per contact: the masses of its two vertices
    double* x2;             // per contact: F + CF at its two vertices (3 each), written by the local kernel
    double* b0;             // per slot: the local kernel's right-hand side (3)
    const int* slot_ent_ptr;  // per slot: its (contact, side) entries slot_ent[ptr[slot] .. ptr[slot + 1])
    const int* slot_ent;      //   as 2 contact + side
    const int* lvl_ptr;     // contacts of level l: [lvl_ptr[l], lvl_ptr[l+1])
    const int* grp_ptr;     // vertex groups of level l
    const int* grp_slot;
    const int* ent_ptr;     // entries of group g
    const int* ent_contact;
    const int* ent_sign;    // +1: contact vertex, -1: paired vertex
};

void launch_fill(double* p, size_t count, double value, hipStream_t st);
void launch_fill_int(int* p, int count, int value, hipStream_t st);
void launch_scan_int(int* dst, const int* src, int count, int* sums, int* total, int* tail, hipStream_t st);
void launch_self_detect(const DevSelfDetect& D, const int* ccnt, unsigned long long* seed_bits, hipStream_t st);
// Single pass: filtered candidates only (the hot path); falls back to launch_self_detect
// when *overflow is set.
void launch_self_detect_filtered(const DevSelfDetect& D, const int* ccnt, unsigned long long* seed_bits,
                                 hipStream_t st);
void launch_self_prepare(const DevSelf& S, int* vslot, hipStream_t st);
void launch_self_level(const DevSelf& S, int l, int count, const double* v, int n, const double* mass,
                       const StepArgs& a, int iteration, hipStream_t st);
void launch_self_apply(const DevSelf& S, int l, int groups, hipStream_t st);
// All self-contact levels of one iteration in a single workgroup (for few self contacts: one
// launch instead of two per level and one to finish).
constexpr int kSelfFusedThreads = 1024;
constexpr int kSelfFusedMaxContacts = 4096;
// The same with the level-to-level state in LDS: one contact per thread, 6 doubles per
// involved vertex and per contact (144 KiB at the limits).
constexpr int kSelfLdsMaxContacts = kSelfFusedThreads;
constexpr int kSelfLdsMaxSlots = 2 * kSelfLdsMaxContacts;
constexpr int kSelfLdsMaxLevels = 255;  // the levels' contact and group offsets are held in LDS too
void launch_self_fused_lds(const DevSelf& S, int nlev, const double* v, int n, const double* mass, const StepArgs& a,
                           int iteration, double* rhs, hipStream_t st);
// Levels l0 .. nlev - 1 in one workgroup through global memory, then (finish) the right-hand side.
void launch_self_fused(const DevSelf& S, int l0, int nlev, const double* v, int n, const double* mass,
                       const StepArgs& a, int iteration, double* rhs, bool finish, hipStream_t st);
constexpr int kSelfTailContacts = kSelfFusedThreads;  // a level of at most this many contacts is "small"
void launch_self_finalize(const DevSelf& S, double* rhs, int n, hipStream_t st);
// End of step: the last iteration's self-contact friction residual, and the slot map cleared.
void launch_self_end(const DevSelf& S, int* vslot, const double* v, int n, const StepArgs& a, int iteration,
                     hipStream_t st);
void launch_fill_u64(unsigned long long* p, size_t count, hipStream_t st);
void launch_step_begin(const DevModel& M, const double* qn, const double* vstate, double* v0, double* cvec,
                       const StepArgs& a, hipStream_t st);
void launch_detect(const DevModel& M, const double* qn, const DevObstacles& ob, int* cflags, int* ccnt,
                   int* cstart, int* blocksum, int* n_contacts, DevContacts C, hipStream_t st);
// Self-mode arguments of the local kernel (selfmode = 0: obstacle reactions go
// straight into b; 1: they go to the involved-vertex buffers, see DevSelf).
struct LocalSelf {
    int selfmode;
    const int* vslot;       // n: slot or -1
    double* Fc;
    double* CFn;
    const double* SCFc;
    double* x2;             // DevSelf::x2, b0, slot_ent_ptr, slot_ent
    double* b0;
    const int* slot_ent_ptr;
    const int* slot_ent;
};
void launch_local(const DevModel& M, const double* qn, const double* vcur, const double* cvec, double* rhs,
                  const int* ccnt, const int* cstart, DevContacts C, const StepArgs& a, int iteration,
                  const LocalSelf& ls, hipStream_t st);
void launch_finish(const DevModel& M, double* qn, const double* vlast, double* vstate, const int* ccnt,
                   const int* cstart, DevContacts C, const DevObstacles& ob, const StepArgs& a, int iteration,
                   hipStream_t st);
// Per obstacle contact: the 12 warm-start features (layout of pdfc_get_warm_data) at v^0.
constexpr int kWarmFeatures = 12;
void launch_warm_features(const DevModel& M, const double* qn, const double* v0, const double* cvec,
                          const int* ncontacts, int cap, DevContacts C, const DevObstacles& ob, const StepArgs& a,
                          float* feat, hipStream_t st);
// The top block's ordered finish folded into the first backward level (the one right below the
// top block, whose supernodes have only top-block ancestors): its tasks read the top rows they
// gather as the chunk-ordered sum of the matvec's partial products, and `extra` workgroups of
// the same launch write every top row's final value for the later levels and the result.
struct TopFold {
    const double* part;   // partial products (layout of DevTop::part); null: no fold
    const int* rows;      // m supernodal indices of the top rows
    const int* top_of;    // n: top index of a supernodal row, -1 elsewhere
    int m, kchunks, ntasks, extra;
    int nb;               // DevTop::nb
};
// Rows of the previous backward level's split supernodes, whose final values (the ordered sum
// of their partials) extra workgroups of the next backward launch write into X.
struct SplitFin {
    const int* rows;
    int count, parts;
    int first;            // first extra workgroup (set by the launcher)
};
// One level of the forward (B -> Y, U) or backward (Y -> X) substitution.
// waves: 1, 2, 4, 8 or 16 waves per task splitting the reduction (ns forward, ns + nb backward).
void launch_forward_level(const DevFactor& F, bool fp32, int t0, int ntasks, const double* B, double* Y,
                          int max_ns, int waves, hipStream_t st);
void launch_backward_level(const DevFactor& F, bool fp32, int t0, int ntasks, const double* Y, double* X,
                           int max_front, int waves, hipStream_t st, const TopFold* fold = nullptr,
                           const SplitFin* fin = nullptr);
// Heuristic wave count: chains of about 8 loads, at most 16 waves.
int default_waves(int kmax);

// Consecutive levels of one sweep in a single launch, one 16-wave workgroup per group (the
// supernodes of the range under one root inside the range).
constexpr int kMergedLdsMax = 40960;  // dynamic LDS bytes of a merged launch (24 KiB static besides)
struct MergedRange {
    int nl;            // levels in the range, in sweep order
    int groups;        // workgroups
    const int* ptr;    // groups x (nl + 1): group g's tasks at its k-th level are
    const int* task;   //   task[ptr[g (nl + 1) + k] .. ptr[g (nl + 1) + k + 1]) (indices into DevFactor::desc)
    const int* info;   // 2 nl: waves per task (1, 2, 4, 8, 16) and LDS doubles per slot, per level
    size_t lds_bytes;  // (16 / waves) x slot doubles x 8 at the widest level, <= kMergedLdsMax
};
void launch_sweep_merged(const DevFactor& F, bool fp32, bool forward, const MergedRange& R, const double* In,
                         double* Out, hipStream_t st);
// Dense top block of the substitution (see TopPlan).
struct DevTop {
    int m;
    int kchunks;            // reduction split of the matvec across workgroups
    const int* rows;        // m supernodal indices
    const int* cptr;        // m + 1
    const int* cidx;        // update-vector entries to subtract
    const void* Sinv;       // m x m, double or float
    double* bt;             // 3 m: right-hand side of the top block (planar)
    // Partial products. Split matvec: kchunks x 3 x m. Symmetric tiles (nb > 0): for row
    // block R and source block C, 3 x 64 values at ((R nb + C) 3 + component) 64.
    double* part;
    const int* top_of;      // n: top index of a supernodal row, -1 elsewhere
    int nb;                 // > 0: symmetric tile products over nb blocks of 64 rows; 0: split matvec
    int ntiles;             // nb (nb + 1) / 2 block pairs I <= J
    const int* tiles;       // 2 per tile: I, J
    const void* Stile;      // the tiles' blocks of S^-1, 4096 entries each (double or float)
    int pregather;          // tiles: b_T gathered into bt by its own launch first (else by every tile)
    const int4* trec;       // per top row: supernodal row, update entries, the first two entries' indices
};
// x_T = S^-1 (b_T - lower updates). Split matvec: gather, partial products over kchunks
// column ranges, ordered reduction + scatter. Symmetric tiles: one workgroup per block pair
// I <= J reads the 64 x 64 block S_IJ once, gathers its own b entries and writes both
// S_IJ b_J (rows I) and S_IJ^T b_I (rows J; S^-1 is symmetric), so only the upper half of
// S^-1 is read. The reduction + scatter is left out with finish = false: the first backward
// level does it, see TopFold.
void launch_top(const DevTop& T, bool fp32, const double* U, const double* B, double* X, int n, hipStream_t st,
                bool finish = true);
// Unit right-hand sides e_{rows[j+c]} in plane c (c = 0..2) and extraction of the
// solution's top rows into columns j..j+2 of S^-1 (build of the top block).
void launch_unit3(double* B, int n, const int* rows, int m, int j, hipStream_t st);
void launch_take_cols(const double* X, int n, const int* rows, int m, int j, double* S, hipStream_t st);
void launch_to_float(const double* src, float* dst, size_t count, hipStream_t st);
// The blocks of S^-1 (m x m, column-major) at the tile list's block pairs, tile by tile.
void launch_pack_tiles(const double* S, void* P, bool fp32, int m, const int* tiles, int ntiles, hipStream_t st);

// r = b - K x with K = diag(mass) + dt2 * A^T A
void launch_residual(const DevModel& M, double dt2, const double* b, const double* x, double* r, hipStream_t st);
void launch_axpy(double* x, const double* dx, size_t count, hipStream_t st);
void launch_copy(double* dst, const double* src, size_t count, hipStream_t st);

}  // namespace pdfc
