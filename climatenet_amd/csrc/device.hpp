// Device-side data layout and kernel launchers (gfx950).
//
// All per-vertex vectors are planar SoA in the supernodal (nested-dissection)
// numbering: component c of vertex i lives at vec[c*n + i]. Stencils are ELL,
// slot-major (slot*n + i), so a wave reading 64 consecutive vertices' slot k
// issues one coalesced 256/512-byte load.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pdfc {

constexpr int kMaxObstacles = 16;
constexpr int kSolveChunk = 64;  // rows per substitution task (one wave)

struct DevObstacles {
    int count;
    int type[kMaxObstacles];
    double p[kMaxObstacles][3];
    double nrm[kMaxObstacles][3];
    double radius[kMaxObstacles];
    double mu[kMaxObstacles];
};

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
};

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
    const void* F;       // double or float
    const void* FT;
    double* U;           // 3 * total below rows, interleaved
    const int* task_s;
    const int* task_r0;
};

struct StepArgs {
    double dt;
    double g[3];
    double damping;
    int record;        // accumulate Alart-Curnier norms
    double* ac_sum;    // per iteration
    unsigned long long* ac_max;  // per iteration, bit pattern of a non-negative double
};

void launch_fill(double* p, size_t count, double value, hipStream_t st);
void launch_fill_u64(unsigned long long* p, size_t count, hipStream_t st);
void launch_step_begin(const DevModel& M, const double* qn, const double* vstate, double* v0, double* cvec,
                       const StepArgs& a, hipStream_t st);
void launch_detect(const DevModel& M, const double* qn, const DevObstacles& ob, int* cflags, int* ccnt,
                   int* cstart, int* blocksum, int* n_contacts, DevContacts C, hipStream_t st);
void launch_local(const DevModel& M, const double* qn, const double* vcur, const double* cvec, double* rhs,
                  const int* ccnt, const int* cstart, DevContacts C, const StepArgs& a, int iteration,
                  hipStream_t st);
void launch_finish(const DevModel& M, double* qn, const double* vlast, double* vstate, const int* ccnt,
                   const int* cstart, DevContacts C, const StepArgs& a, int iteration, hipStream_t st);
// One level of the forward (B -> Y, U) or backward (Y -> X) substitution.
void launch_forward_level(const DevFactor& F, bool fp32, int t0, int ntasks, const double* B, double* Y,
                          int max_ns, hipStream_t st);
void launch_backward_level(const DevFactor& F, bool fp32, int t0, int ntasks, const double* Y, double* X,
                           int max_front, hipStream_t st);
// r = b - K x with K = diag(mass) + dt2 * A^T A
void launch_residual(const DevModel& M, double dt2, const double* b, const double* x, double* r, hipStream_t st);
void launch_axpy(double* x, const double* dx, size_t count, hipStream_t st);
void launch_copy(double* dst, const double* src, size_t count, hipStream_t st);

}  // namespace pdfc
