// gfx950 kernels of the projective-dynamics step with dry frictional contact.
//
// Per time step:   step_begin -> detect (count, scan, write) -> K x [local ->
//                  forward levels -> backward levels] -> finish.
// Predicates (contact tests) are evaluated without fused multiply-adds
// (-ffp-contract=off) and with correctly rounded sqrt/div so that contact sets
// match the CPU oracle bit for bit; the dense substitution products use
// explicit fma.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <stdexcept>

#include "device.hpp"

namespace pdfc {
thread_local LaunchTimer* g_launch_timer = nullptr;
namespace {

constexpr int kBlock = 256;
constexpr int kAtaChunk = 8;  // A^T A gathers in flight at a time

__device__ __forceinline__ double wave_sum(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ double wave_max(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 ld3(const double* v, int n, int i) { return {v[i], v[n + i], v[2 * n + i]}; }
__device__ __forceinline__ void st3(double* v, int n, int i, D3 a) {
    v[i] = a.x;
    v[n + i] = a.y;
    v[2 * n + i] = a.z;
}
__device__ __forceinline__ double dot3(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ D3 cross3(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ D3 normalized3(D3 a) {
    double z = dot3(a, a);
    if (z > 0.) {
        double s = sqrt(z);
        return {a.x / s, a.y / s, a.z / s};
    }
    return a;
}

// computeBasis, src/physics/SimulationFrictionEstimation.cpp:89-105
__device__ void contact_basis(D3 nrm, double* B) {
    D3 t = cross3(nrm, D3{1., 0., 0.});
    if (sqrt(dot3(t, t)) < 1.e-15) t = cross3(nrm, D3{0., 1., 0.});
    t = normalized3(t);
    D3 b = normalized3(cross3(nrm, t));
    B[0] = nrm.x; B[1] = nrm.y; B[2] = nrm.z;
    B[3] = t.x;   B[4] = t.y;   B[5] = t.z;
    B[6] = b.x;   B[7] = b.y;   B[8] = b.z;
}

// Block-wide accumulation of (sum, max) into per-iteration slots.
// Per wave (no block barrier: a wave that is done need not wait for the block's others).
__device__ void block_ac_accumulate(double s, double m, double* sum_slot, unsigned long long* max_slot) {
    s = wave_sum(s);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) {
        if (s != 0.) atomicAdd(sum_slot, s);
        if (m > 0.) atomicMax(max_slot, (unsigned long long)__double_as_longlong(m));
    }
}

// Alart-Curnier norm, src/physics/coulomb_util.cpp:101-152.
__device__ double ac_norm(D3 u, const double* r, double mu) {
    double an = fmax(0., r[0] - u.x) - r[0];
    double d1 = r[1] - u.y, d2 = r[2] - u.z;
    double dn = sqrt(d1 * d1 + d2 * d2);
    double s = fmin(1., fmax(0., mu * r[0] / dn));
    double a1 = d1 * s - r[1], a2 = d2 * s - r[2];
    return sqrt((an * an + a1 * a1) + a2 * a2);
}

__global__ void k_fill(double* p, size_t count, double value) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}
__global__ void k_fill_u64(unsigned long long* p, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0ull;
}
__global__ void k_copy(double* dst, const double* src, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i];
}
__global__ void k_axpy(double* x, const double* dx, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) x[i] += dx[i];
}

// (A^T A x)_i with kAtaChunk index and gather loads in flight at a time (clamped indices:
// every load in bounds and unconditional), summed in entry order.
__device__ __forceinline__ D3 ata_product(const DevModel& M, const double* __restrict__ x, int i) {
    const int n = M.n;
    D3 s = {0., 0., 0.};
    const int cnt = M.ata_cnt[i];
    for (int k0 = 0; k0 < cnt; k0 += kAtaChunk) {
        int jj[kAtaChunk];
        double w[kAtaChunk], g[kAtaChunk][3];
#pragma unroll
        for (int k = 0; k < kAtaChunk; ++k) {
            const int e = min(k0 + k, cnt - 1);
            jj[k] = M.ata_idx[(size_t)e * n + i];
            w[k] = M.ata_val[(size_t)e * n + i];
        }
#pragma unroll
        for (int k = 0; k < kAtaChunk; ++k) {
            g[k][0] = x[jj[k]];
            g[k][1] = x[n + jj[k]];
            g[k][2] = x[2 * n + jj[k]];
        }
#pragma unroll
        for (int k = 0; k < kAtaChunk; ++k)
            if (k0 + k < cnt) {
                s.x += w[k] * g[k][0];
                s.y += w[k] * g[k][1];
                s.z += w[k] * g[k][2];
            }
    }
    return s;
}

// t_n = v_n + dt M^-1 (M g); v^0 = t_n; c = M t_n - dt A^T A q_n
// (SimulationSpeed::initializeStep, SimulationSpeed.cpp:62-76; constant part of
// SimulationSpeed::computeRhs, :263-277, hoisted out of the iteration loop).
__global__ void k_step_begin(DevModel M, const double* __restrict__ qn, const double* __restrict__ vstate,
                             double* __restrict__ v0, double* __restrict__ cvec, StepArgs a) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;  // the A^T A gathers share an L2
    const int n = M.n;
    if (a.record)  // the step's residual records start from zero
        for (int k = i; k < a.ac_cap; k += gridDim.x * blockDim.x) {
            a.ac_sum[k] = 0.;
            a.ac_max[k] = 0ull;
        }
    if (i >= n) return;
    double m = M.mass[i], im = M.invmass[i];
    D3 v = ld3(vstate, n, i);
    D3 t = {v.x + a.dt * (im * (m * a.g[0])), v.y + a.dt * (im * (m * a.g[1])), v.z + a.dt * (im * (m * a.g[2]))};
    st3(v0, n, i, t);
    const D3 s = ata_product(M, qn, i);  // (A^T A q_n)_i
    st3(cvec, n, i, D3{m * t.x - a.dt * s.x, m * t.y - a.dt * s.y, m * t.z - a.dt * s.z});
}

// Attachment rows: their A^T p is constant (w^2 * position), so it joins the step's
// constant part c += dt (A^T p)_attachments (AttachmentConstraint.cpp:46-52).
__global__ void k_attach(DevModel M, double* __restrict__ cvec, double dt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M.n_att) return;
    const int n = M.n, i = M.att_vertex[k];
    cvec[i] += dt * M.att_term[3 * k];
    cvec[n + i] += dt * M.att_term[3 * k + 1];
    cvec[2 * n + i] += dt * M.att_term[3 * k + 2];
}

// Obstacles vertex i meets: all of them, or those of its scene when scenes are batched.
__device__ __forceinline__ void obstacle_range(const DevObstacles& ob, int i, int& o0, int& o1) {
    o0 = 0;
    o1 = ob.count;
    if (ob.vinst) {
        const int k = ob.vinst[i];
        o0 = ob.obs_ptr[k];
        o1 = ob.obs_ptr[k + 1];
    }
}

__device__ __forceinline__ int mesh_slot(const DevObstacles& ob, int o) {
    int m = 0;
    while (m + 1 < ob.nmesh && ob.mesh[m].obstacle != o) ++m;
    return m;
}
__device__ __forceinline__ D3 ld_aos(const double* p, int v) { return {p[3 * v], p[3 * v + 1], p[3 * v + 2]}; }
__device__ __forceinline__ D3 tri_point3(D3 a, D3 b, D3 c, const double* w) {
    return {(a.x * w[0] + b.x * w[1]) + c.x * w[2], (a.y * w[0] + b.y * w[1]) + c.y * w[2],
            (a.z * w[0] + b.z * w[1]) + c.z * w[2]};
}

// StaticAxisAlignedBox::checkCollision (StaticAxisAlignedBox.cpp:41-60): inside unless
// some |relative_k| > size_k / 2 (relative = x - corner - size / 2).
__device__ __forceinline__ bool box_inside(const double* corner, const double* size, D3 x, D3& rel) {
    rel = {(x.x - corner[0]) - 0.5 * size[0], (x.y - corner[1]) - 0.5 * size[1], (x.z - corner[2]) - 0.5 * size[2]};
    return !(fabs(rel.x) > size[0] / 2 || fabs(rel.y) > size[1] / 2 || fabs(rel.z) > size[2] / 2);
}
// Nearest side: first minimum of |relative_axis - sign size_axis / 2| in the Side enum
// order TOP (+y), BOTTOM (-y), LEFT (-x), RIGHT (+x), FRONT (+z), BACK (-z); the contact
// is relative with that axis moved onto the side, back in world coordinates.
__device__ __forceinline__ void box_contact(const double* corner, const double* size, D3 rel, D3& pt, D3& nrm) {
    const int axis[6] = {1, 1, 0, 0, 2, 2};
    const double sign[6] = {1., -1., -1., 1., 1., -1.};
    double r[3] = {rel.x, rel.y, rel.z};
    int best = 0;
    double bd = 0.;
    for (int sd = 0; sd < 6; ++sd) {
        double d = fabs(r[axis[sd]] - sign[sd] * 0.5 * size[axis[sd]]);
        if (sd == 0 || d < bd) {
            bd = d;
            best = sd;
        }
    }
    r[axis[best]] = sign[best] * 0.5 * size[axis[best]];
    pt = {(r[0] + corner[0]) + 0.5 * size[0], (r[1] + corner[1]) + 0.5 * size[1], (r[2] + corner[2]) + 0.5 * size[2]};
    double nn[3] = {0., 0., 0.};
    nn[axis[best]] = sign[best];
    nrm = {nn[0], nn[1], nn[2]};
}

// Obstacle detection at q_n: StaticSphere/RotatingSphere via Sphere::isInside
// (Sphere.cpp:93-113), StaticPlane (StaticPlane.cpp:57-73), boxes, mesh hits from
// k_mesh_query. One bit per obstacle.
__global__ void k_detect_count(DevModel M, const double* __restrict__ qn, DevObstacles ob, int* cflags, int* ccnt,
                               int* blocksum) {
    __shared__ int red[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = M.n;
    int cnt = 0;
    if (i < n) {
        D3 x = ld3(qn, n, i);
        int flags = 0;
        int o0, o1;
        obstacle_range(ob, i, o0, o1);
        for (int o = o0; o < o1; ++o) {
            bool hit;
            const int ty = ob.type[o];
            D3 r = {x.x - ob.p[3 * o], x.y - ob.p[3 * o + 1], x.z - ob.p[3 * o + 2]};
            if (ty == PDFC_OBSTACLE_SPHERE || ty == PDFC_OBSTACLE_ROTATING_SPHERE)
                hit = sqrt(dot3(r, r)) < ob.radius[o] + 1e-2;
            else if (ty == PDFC_OBSTACLE_PLANE)
                hit = dot3(r, D3{ob.nrm[3 * o], ob.nrm[3 * o + 1], ob.nrm[3 * o + 2]}) < 1.5e-2;
            else if (ty == PDFC_OBSTACLE_BOX)
                hit = box_inside(&ob.p[3 * o], &ob.nrm[3 * o], x, r);
            else
                hit = ob.mesh[mesh_slot(ob, o)].hit_tri[i] >= 0;
            if (hit) flags |= 1 << (o - o0);
        }
        cnt = __popc(flags);
        cflags[i] = flags;
        ccnt[i] = cnt;
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
        blocksum[blockIdx.x] = t;
    }
}

// Exclusive scan of the per-block counts (one block); total -> *total.
__global__ void k_scan_blocks(int* blocksum, int nblocks, int* total) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblocks; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < nblocks ? blocksum[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks) blocksum[i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_detect_write(DevModel M, const double* __restrict__ qn, DevObstacles ob, const int* cflags,
                               const int* ccnt, const int* blockoff, int* cstart, DevContacts C) {
    __shared__ int buf[kBlock];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = M.n;
    int v = i < n ? ccnt[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    if (i >= n) return;
    int start = blockoff[blockIdx.x] + buf[threadIdx.x] - v;
    cstart[i] = start;
    int flags = cflags[i];
    if (!flags) return;
    D3 x = ld3(qn, n, i);
    int slot = start;
    int o0, o1;
    obstacle_range(ob, i, o0, o1);
    for (int o = o0; o < o1; ++o) {
        if (!(flags & (1 << (o - o0)))) continue;
        D3 nrm, pt, spd = {0., 0., 0.};
        const D3 c = {ob.p[3 * o], ob.p[3 * o + 1], ob.p[3 * o + 2]};
        const int ty = ob.type[o];
        if (ty == PDFC_OBSTACLE_SPHERE || ty == PDFC_OBSTACLE_ROTATING_SPHERE) {
            D3 r = normalized3(D3{x.x - c.x, x.y - c.y, x.z - c.z});
            double R = ob.radius[o];
            pt = {c.x + R * r.x, c.y + R * r.y, c.z + R * r.z};
            nrm = r;
            // RotatingSphere::checkCollision (RotatingSphere.cpp:71-93): omega x (contact - centre)
            // once the time before the rotation starts is negative
            if (ty == PDFC_OBSTACLE_ROTATING_SPHERE && ob.clock[o] < 0.) {
                const D3 w = {ob.nrm[3 * o], ob.nrm[3 * o + 1], ob.nrm[3 * o + 2]};
                spd = cross3(w, D3{pt.x - c.x, pt.y - c.y, pt.z - c.z});
            }
        } else if (ty == PDFC_OBSTACLE_BOX) {
            D3 rel;
            box_inside(&ob.p[3 * o], &ob.nrm[3 * o], x, rel);
            box_contact(&ob.p[3 * o], &ob.nrm[3 * o], rel, pt, nrm);
        } else if (ty == PDFC_OBSTACLE_PLANE) {
            nrm = {ob.nrm[3 * o], ob.nrm[3 * o + 1], ob.nrm[3 * o + 2]};
            D3 r = {x.x - c.x, x.y - c.y, x.z - c.z};
            double d = dot3(r, nrm);
            pt = {x.x - nrm.x * d, x.y - nrm.y * d, x.z - nrm.z * d};
        } else {
            // checkMeshCollision (geometry_util.cpp:32-92): the closest triangle's point,
            // its normal (Mesh::getTriangleNormal) and MovingMesh::getPointSpeed
            const DevMeshObs& Mo = ob.mesh[mesh_slot(ob, o)];
            const int t = Mo.hit_tri[i];
            const double* w = &Mo.hit_bary[3 * i];
            const int* f = &Mo.tris[3 * t];
            D3 a = ld_aos(Mo.cur, f[0]), b = ld_aos(Mo.cur, f[1]), cc = ld_aos(Mo.cur, f[2]);
            pt = tri_point3(a, b, cc, w);
            nrm = normalized3(cross3(D3{b.x - a.x, b.y - a.y, b.z - a.z}, D3{cc.x - a.x, cc.y - a.y, cc.z - a.z}));
            if (Mo.moving) {
                D3 pp = tri_point3(ld_aos(Mo.prev, f[0]), ld_aos(Mo.prev, f[1]), ld_aos(Mo.prev, f[2]), w);
                D3 np = tri_point3(ld_aos(Mo.next, f[0]), ld_aos(Mo.next, f[1]), ld_aos(Mo.next, f[2]), w);
                spd = {Mo.inv_dt * (np.x - pp.x), Mo.inv_dt * (np.y - pp.y), Mo.inv_dt * (np.z - pp.z)};
            }
        }
        C.vertex[slot] = i;
        C.obstacle[slot] = o;
        C.active[slot] = 0;
        C.normal[3 * slot] = nrm.x; C.normal[3 * slot + 1] = nrm.y; C.normal[3 * slot + 2] = nrm.z;
        C.point[3 * slot] = pt.x;   C.point[3 * slot + 1] = pt.y;   C.point[3 * slot + 2] = pt.z;
        C.mu[slot] = ob.mu[o];
        contact_basis(nrm, &C.basis[9 * slot]);
        C.r[3 * slot] = 0.; C.r[3 * slot + 1] = 0.; C.r[3 * slot + 2] = 0.;
        C.speed[3 * slot] = spd.x; C.speed[3 * slot + 1] = spd.y; C.speed[3 * slot + 2] = spd.z;
        ++slot;
    }
}

// Stencil widths handled with all of a vertex's loads in flight at once (indices first,
// then the gathers); wider stencils take the loop below.
constexpr int kSpringCap = 8;
constexpr int kAtaCap = 24;

// (A^T p)_i at q = q_n + dt v: sum over the incident springs of w^2 L0 (q_i - q_j)/|q_i - q_j|.
__device__ __forceinline__ D3 spring_sum(const DevModel& M, const double* __restrict__ qn,
                                         const double* __restrict__ vcur, double dt, int i, D3 q) {
    const int n = M.n;
    D3 s = {0., 0., 0.};
    const int deg = M.sp_cnt[i];
    if (M.sp_slots <= kSpringCap) {
        int jj[kSpringCap];
        double kap[kSpringCap];
#pragma unroll
        for (int k = 0; k < kSpringCap; ++k)
            if (k < deg) {
                jj[k] = M.sp_idx[(size_t)k * n + i];
                kap[k] = M.sp_kappa[(size_t)k * n + i];
            }
        D3 qj[kSpringCap];
#pragma unroll
        for (int k = 0; k < kSpringCap; ++k)
            if (k < deg) {
                const int j = jj[k];
                qj[k] = {qn[j] + dt * vcur[j], qn[n + j] + dt * vcur[n + j], qn[2 * n + j] + dt * vcur[2 * n + j]};
            }
#pragma unroll
        for (int k = 0; k < kSpringCap; ++k)
            if (k < deg) {
                D3 e = normalized3(D3{q.x - qj[k].x, q.y - qj[k].y, q.z - qj[k].z});
                s.x += kap[k] * e.x;
                s.y += kap[k] * e.y;
                s.z += kap[k] * e.z;
            }
        return s;
    }
    for (int k = 0; k < deg; ++k) {
        int j = M.sp_idx[(size_t)k * n + i];
        double kap = M.sp_kappa[(size_t)k * n + i];
        D3 qj = {qn[j] + dt * vcur[j], qn[n + j] + dt * vcur[n + j], qn[2 * n + j] + dt * vcur[2 * n + j]};
        D3 e = normalized3(D3{q.x - qj.x, q.y - qj.y, q.z - qj.z});
        s.x += kap * e.x;
        s.y += kap * e.y;
        s.z += kap * e.z;
    }
    return s;
}

// (A^T A v)_i
__device__ __forceinline__ D3 ata_row(const DevModel& M, const double* __restrict__ vcur, int i) {
    const int n = M.n;
    D3 av = {0., 0., 0.};
    const int cnt = M.ata_cnt[i];
    if (M.ata_slots <= kAtaCap) {
        int jj[kAtaCap];
        double w[kAtaCap];
#pragma unroll
        for (int k = 0; k < kAtaCap; ++k)
            if (k < cnt) {
                jj[k] = M.ata_idx[(size_t)k * n + i];
                w[k] = M.ata_val[(size_t)k * n + i];
            }
#pragma unroll
        for (int k = 0; k < kAtaCap; ++k)
            if (k < cnt) {
                const int j = jj[k];
                av.x += w[k] * vcur[j];
                av.y += w[k] * vcur[n + j];
                av.z += w[k] * vcur[2 * n + j];
            }
        return av;
    }
    for (int k = 0; k < cnt; ++k) {
        int j = M.ata_idx[(size_t)k * n + i];
        double w = M.ata_val[(size_t)k * n + i];
        av.x += w * vcur[j];
        av.y += w * vcur[n + j];
        av.z += w * vcur[2 * n + j];
    }
    return av;
}

// A vertex's stencils held in registers between the loads that fetch them and the sums
// that use them (spring and A^T A rows up to kSpringCap / kAtaCap entries). Slots beyond
// a row's length point at the vertex itself, so every gather is issued without a branch.
struct SpringRow {
    int deg;
    int j[kSpringCap];
    double kap[kSpringCap];
};
struct AtaRowRegs {
    int cnt;
    int j[kAtaCap];
    double w[kAtaCap];
};

__device__ __forceinline__ void spring_row_load(const DevModel& M, int i, SpringRow& R) {
    const int n = M.n;
    R.deg = M.sp_cnt[i];
#pragma unroll
    for (int k = 0; k < kSpringCap; ++k) {
        R.j[k] = i;
        R.kap[k] = 0.;
        if (k < M.sp_slots) {  // uniform: the loads do not wait for deg
            R.j[k] = M.sp_idx[(size_t)k * n + i];
            R.kap[k] = M.sp_kappa[(size_t)k * n + i];
        }
    }
}

// Same sum, same order as spring_sum.
__device__ __forceinline__ D3 spring_row_sum(const SpringRow& R, const double* __restrict__ qn,
                                             const double* __restrict__ vcur, int n, int i, double dt, D3 q) {
    double g[kSpringCap][6];
#pragma unroll
    for (int k = 0; k < kSpringCap; ++k) {
        const int j = k < R.deg ? R.j[k] : i;
        g[k][0] = qn[j];
        g[k][1] = qn[n + j];
        g[k][2] = qn[2 * n + j];
        g[k][3] = vcur[j];
        g[k][4] = vcur[n + j];
        g[k][5] = vcur[2 * n + j];
    }
    D3 s = {0., 0., 0.};
#pragma unroll
    for (int k = 0; k < kSpringCap; ++k)
        if (k < R.deg) {
            const D3 qj = {g[k][0] + dt * g[k][3], g[k][1] + dt * g[k][4], g[k][2] + dt * g[k][5]};
            const D3 e = normalized3(D3{q.x - qj.x, q.y - qj.y, q.z - qj.z});
            s.x += R.kap[k] * e.x;
            s.y += R.kap[k] * e.y;
            s.z += R.kap[k] * e.z;
        }
    return s;
}

__device__ __forceinline__ void ata_row_load(const DevModel& M, int i, AtaRowRegs& R) {
    const int n = M.n;
    R.cnt = M.ata_cnt[i];
#pragma unroll
    for (int k = 0; k < kAtaCap; ++k) {
        R.j[k] = i;
        R.w[k] = 0.;
        if (k < M.ata_slots) {
            R.j[k] = M.ata_idx[(size_t)k * n + i];
            R.w[k] = M.ata_val[(size_t)k * n + i];
        }
    }
}

// Same sum, same order as ata_row; the gathers go out kAtaChunk entries at a time (fewer
// registers held than all kAtaCap at once).
__device__ __forceinline__ D3 ata_row_sum(const AtaRowRegs& R, const double* __restrict__ vcur, int n, int i) {
    D3 av = {0., 0., 0.};
#pragma unroll
    for (int k0 = 0; k0 < kAtaCap; k0 += kAtaChunk) {
        if (k0 >= R.cnt) break;
        double g[kAtaChunk][3];
#pragma unroll
        for (int k = 0; k < kAtaChunk; ++k) {
            const int j = k0 + k < R.cnt ? R.j[k0 + k] : i;
            g[k][0] = vcur[j];
            g[k][1] = vcur[n + j];
            g[k][2] = vcur[2 * n + j];
        }
#pragma unroll
        for (int k = 0; k < kAtaChunk; ++k)
            if (k0 + k < R.cnt) {
                av.x += R.w[k0 + k] * g[k][0];
                av.y += R.w[k0 + k] * g[k][1];
                av.z += R.w[k0 + k] * g[k][2];
            }
    }
    return av;
}

// One obstacle contact's inputs to the cone projection.
struct ContactIn {
    double B[9];     // n, t, b
    double sp[3];    // obstacle speed at the contact point
    double r[3];     // previous iteration's reaction (residual bookkeeping)
    double mu;
};
__device__ __forceinline__ void contact_load(const DevContacts& C, int cid, bool with_r, ContactIn& c) {
#pragma unroll
    for (int k = 0; k < 9; ++k) c.B[k] = C.basis[9 * (size_t)cid + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) c.sp[k] = C.speed[3 * (size_t)cid + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) c.r[k] = with_r ? C.r[3 * (size_t)cid + k] : 0.;
    c.mu = C.mu[cid];
}

// Local step + right-hand side + obstacle friction, one thread per vertex:
//   q = q_n + dt v;  s_i = sum_springs w^2 L0 (q_i - q_j)/|q_i - q_j|  (A^T p; bending
//   projections vanish because the reference's rest Laplacian norm is 0);
//   b = c + dt (s - D v)                        (SimulationSpeed::computeRhs)
//   F = b - dt^2 (A^T A v)_i at contact vertices, r = cone(B^T F), b += B r
//                                               (SimulationFrictionEstimation.cpp:337-423)
// The friction residual of the previous iteration (r^{k-1}, v^k) is folded in.
// kLowLatency: the loads are issued in three dependent rounds (everything indexed by the
// vertex with the stencil indices; the spring gathers, the A^T A indices and the first
// contact; the A^T A gathers), at the price of registers (one wave per SIMD). For grids of
// many blocks the other instance (spring_sum / ata_row: a gather per stencil entry, 111
// VGPRs) keeps four waves per SIMD resident instead.
template <bool kLowLatency>
__global__ void __launch_bounds__(kBlock) k_local(DevModel M, const double* __restrict__ qn,
                                                  const double* __restrict__ vcur, const double* __restrict__ cvec,
                                                  double* __restrict__ rhs, const int* __restrict__ ccnt,
                                                  const int* __restrict__ cstart, DevContacts C, StepArgs a,
                                                  int iteration, LocalSelf ls) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int n = M.n;
    double ac_s = 0., ac_m = 0.;
    if (i < n) {
        const double dt = a.dt;
        const bool regs = kLowLatency && M.sp_slots <= kSpringCap && M.ata_slots <= kAtaCap;
        const bool rec = a.record && iteration > 0;
        const int nc = ccnt[i];
        const int slot = ls.selfmode ? ls.vslot[i] : -1;
        const int c0 = cstart[i];
        const D3 v = ld3(vcur, n, i);
        const D3 qi = ld3(qn, n, i);
        const D3 c = ld3(cvec, n, i);
        const double mi = M.mass[i];
        SpringRow SR;
        if (regs) spring_row_load(M, i, SR);
        const bool contact = nc > 0 || slot >= 0;
        AtaRowRegs AR;
        AR.cnt = 0;
        ContactIn first;
        if (contact && regs) ata_row_load(M, i, AR);
        if (kLowLatency && nc > 0) contact_load(C, c0, rec, first);
        // a slot vertex's (contact, side) entries: their range and first two now, so that their
        // stores after the contacts wait for no load
        int se0 = 0, se1 = 0, ent0 = 0, ent1 = 0;
        if (slot >= 0) {
            se0 = ls.slot_ent_ptr[slot];
            se1 = ls.slot_ent_ptr[slot + 1];
            ent0 = ls.slot_ent[se0];  // every slot has an entry
            ent1 = ls.slot_ent[min(se0 + 1, se1 - 1)];
        }
        const D3 q = {qi.x + dt * v.x, qi.y + dt * v.y, qi.z + dt * v.z};
        const D3 s = regs ? spring_row_sum(SR, qn, vcur, n, i, dt, q) : spring_sum(M, qn, vcur, dt, i, q);
        double damp = contact ? 0. : a.damping;
        D3 b = {c.x + dt * (s.x - damp * v.x), c.y + dt * (s.y - damp * v.y), c.z + dt * (s.z - damp * v.z)};
        if (contact) {
            // self-collision mode, vertex without a self contact: only its last active obstacle
            // reaction reaches the right-hand side (m_contact_forces is assigned, not summed,
            // SimulationFrictionEstimation.cpp:410-418; rhs += CF + SCF with SCF = 0 there)
            D3 glast = {0., 0., 0.};
            D3 cfl = {0., 0., 0.};  // slot vertex: its last active obstacle reaction
            const D3 av = regs ? ata_row_sum(AR, vcur, n, i) : ata_row(M, vcur, i);
            const double dt2 = dt * dt;
            D3 F = {b.x - dt2 * av.x, b.y - dt2 * av.y, b.z - dt2 * av.z};
            D3 Fs = F;  // what the obstacle contacts see
            if (slot >= 0) {
                ls.Fc[3 * slot] = F.x;
                ls.Fc[3 * slot + 1] = F.y;
                ls.Fc[3 * slot + 2] = F.z;
                Fs = {F.x + ls.SCFc[3 * slot], F.y + ls.SCFc[3 * slot + 1], F.z + ls.SCFc[3 * slot + 2]};
            }
            for (int k = 0; k < nc; ++k) {
                const int cid = c0 + k;
                ContactIn ci;
                if (kLowLatency && k == 0)
                    ci = first;
                else
                    contact_load(C, cid, rec, ci);
                double* r = &C.r[3 * cid];
                const D3 bn = {ci.B[0], ci.B[1], ci.B[2]}, bt = {ci.B[3], ci.B[4], ci.B[5]},
                         bb = {ci.B[6], ci.B[7], ci.B[8]};
                const double mu = ci.mu;
                const D3 sp = {ci.sp[0], ci.sp[1], ci.sp[2]};
                if (rec) {
                    const D3 u = {v.x - sp.x, v.y - sp.y, v.z - sp.z};
                    double an = ac_norm(D3{dot3(bn, u), dot3(bt, u), dot3(bb, u)}, ci.r, mu);
                    ac_s += an;
                    ac_m = fmax(ac_m, an);
                }
                // f = B^T (F - m u_obstacle)  (SimulationFrictionEstimation.cpp:360-372)
                const D3 Fo = {Fs.x - mi * sp.x, Fs.y - mi * sp.y, Fs.z - mi * sp.z};
                D3 fl = {dot3(bn, Fo), dot3(bt, Fo), dot3(bb, Fo)};
                double r0, r1, r2;
                bool warm = false;
                if (a.warm_mode != 0) {  // the host clears warm_mode once the hold ends
                    const float* y = a.warm_y + (size_t)cid * a.warm_stride;
                    r0 = mi * (double)y[0];
                    r1 = mi * (double)y[1];
                    r2 = mi * (double)y[2];
                    warm = r0 != 0. || r1 != 0. || r2 != 0.;  // no warm value: the contact starts plain
                }
                if (warm ? !(r0 > 0.) : fl.x > 0.) {
                    r[0] = 0.; r[1] = 0.; r[2] = 0.;
                    C.active[cid] = 0;
                    continue;
                }
                if (!warm) {
                    r0 = -fl.x;
                    r1 = -fl.y;
                    r2 = -fl.z;
                }
                double rt = sqrt(r1 * r1 + r2 * r2);
                if (rt > mu * r0) {
                    r1 *= mu * r0 / rt;
                    r2 *= mu * r0 / rt;
                }
                r[0] = r0; r[1] = r1; r[2] = r2;
                C.active[cid] = 1;
                D3 g = {(bn.x * r0 + bt.x * r1) + bb.x * r2, (bn.y * r0 + bt.y * r1) + bb.y * r2,
                        (bn.z * r0 + bt.z * r1) + bb.z * r2};
                if (slot >= 0) {
                    // self-collision mode: the last active contact of the vertex is kept and
                    // added after the self-contact levels
                    ls.CFn[3 * slot] = g.x;
                    ls.CFn[3 * slot + 1] = g.y;
                    ls.CFn[3 * slot + 2] = g.z;
                    cfl = g;
                } else if (ls.selfmode) {
                    glast = g;
                } else {
                    b.x += g.x;
                    b.y += g.y;
                    b.z += g.z;
                }
            }
            if (ls.selfmode && slot < 0) {
                b.x += glast.x;
                b.y += glast.y;
                b.z += glast.z;
            }
            if (slot >= 0) {
                // F + CF for each of the slot's (contact, side) entries and the right-hand side,
                // so that the self-contact levels read their inputs by contact and by slot
                // without a dependent load
                const D3 fcf = {F.x + cfl.x, F.y + cfl.y, F.z + cfl.z};
                for (int e = se0; e < se1; ++e) {
                    const int code = e == se0 ? ent0 : (e == se0 + 1 ? ent1 : ls.slot_ent[e]);
                    double* x = ls.x2 + 3 * (size_t)code;
                    x[0] = fcf.x;
                    x[1] = fcf.y;
                    x[2] = fcf.z;
                }
                ls.b0[3 * slot] = b.x;
                ls.b0[3 * slot + 1] = b.y;
                ls.b0[3 * slot + 2] = b.z;
            }
        }
        st3(rhs, n, i, b);
    }
    if (a.record && iteration > 0) block_ac_accumulate(ac_s, ac_m, &a.ac_sum[iteration - 1], &a.ac_max[iteration - 1]);
}

// q_{n+1} = q_n + dt v^K, v_{n+1} = v^K (SimulationFrictionEstimation::updateScene,
// :64-87) and the friction residual of the last iteration.
__global__ void k_finish(DevModel M, double* __restrict__ qn, const double* __restrict__ vlast,
                         double* __restrict__ vstate, const int* __restrict__ ccnt, const int* __restrict__ cstart,
                         DevContacts C, DevObstacles ob, StepArgs a, int iteration) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = M.n;
    double ac_s = 0., ac_m = 0.;
    // RotatingSphere::update (RotatingSphere.cpp:39-43), once per step; detection of this
    // step has read the clock already
    if (i < ob.count && ob.type[i] == PDFC_OBSTACLE_ROTATING_SPHERE) ob.clock[i] -= a.dt;
    if (i < n) {
        D3 v = ld3(vlast, n, i);
        if (a.record) {
            int nc = ccnt[i], c0 = cstart[i];
            for (int k = 0; k < nc; ++k) {
                const int cid = c0 + k;
                const double* B = &C.basis[9 * cid];
                D3 bn = {B[0], B[1], B[2]}, bt = {B[3], B[4], B[5]}, bb = {B[6], B[7], B[8]};
                const D3 u = {v.x - C.speed[3 * cid], v.y - C.speed[3 * cid + 1], v.z - C.speed[3 * cid + 2]};
                double an = ac_norm(D3{dot3(bn, u), dot3(bt, u), dot3(bb, u)}, &C.r[3 * cid], C.mu[cid]);
                ac_s += an;
                ac_m = fmax(ac_m, an);
            }
        }
        const double qx = qn[i] + a.dt * v.x, qy = qn[n + i] + a.dt * v.y, qz = qn[2 * n + i] + a.dt * v.z;
        qn[i] = qx;
        qn[n + i] = qy;
        qn[2 * n + i] = qz;
        st3(vstate, n, i, v);
        // failure detection (no counterpart in the reference, which steps on with NaNs)
        if (a.nonfinite && !(isfinite(qx) && isfinite(qy) && isfinite(qz) && isfinite(v.x) && isfinite(v.y) &&
                             isfinite(v.z)))
            *a.nonfinite = 1;
        if (a.prev_g) {
            int o0, o1;
            obstacle_range(ob, i, o0, o1);
            for (int o = 0; o < o1 - o0; ++o) {
                float* g = a.prev_g + 3 * ((size_t)o * n + i);
                g[0] = 0.f; g[1] = 0.f; g[2] = 0.f;
            }
            int nc = ccnt[i], c0 = cstart[i];
            for (int k = 0; k < nc; ++k) {
                const int cid = c0 + k;
                if (!C.active[cid]) continue;
                const double* B = &C.basis[9 * cid];
                const double* r = &C.r[3 * cid];
                float* g = a.prev_g + 3 * ((size_t)(C.obstacle[cid] - o0) * n + i);
                g[0] = (float)((B[0] * r[0] + B[3] * r[1]) + B[6] * r[2]);
                g[1] = (float)((B[1] * r[0] + B[4] * r[1]) + B[7] * r[2]);
                g[2] = (float)((B[2] * r[0] + B[5] * r[1]) + B[8] * r[2]);
            }
        }
    }
    if (a.record) block_ac_accumulate(ac_s, ac_m, &a.ac_sum[iteration], &a.ac_max[iteration]);
}

// Warm-start features of each obstacle contact at the start of the step (v = v^0):
// relative velocity and previous reaction in the contact basis, normal, mu, gap at q_n,
// and the normal contact force of the plain warm start, the last two per unit mass.
__global__ void k_warm_features(DevModel M, const double* __restrict__ qn, const double* __restrict__ v0,
                                const double* __restrict__ cvec, const int* __restrict__ ncontacts, DevContacts C,
                                DevObstacles ob, StepArgs a, float* __restrict__ feat) {
    const int cid = blockIdx.x * blockDim.x + threadIdx.x;
    if (cid >= *ncontacts) return;
    const int n = M.n;
    const int i = C.vertex[cid], o = C.obstacle[cid];
    const double dt = a.dt;
    const double m = M.mass[i], im = 1. / m;
    const double* B = &C.basis[9 * cid];
    D3 bn = {B[0], B[1], B[2]}, bt = {B[3], B[4], B[5]}, bb = {B[6], B[7], B[8]};
    D3 v = ld3(v0, n, i);
    D3 x = ld3(qn, n, i);
    D3 q = {x.x + dt * v.x, x.y + dt * v.y, x.z + dt * v.z};
    D3 s = spring_sum(M, qn, v0, dt, i, q);
    D3 c = ld3(cvec, n, i);
    D3 av = ata_row(M, v0, i);
    const double dt2 = dt * dt;
    D3 F = {c.x + dt * s.x - dt2 * av.x, c.y + dt * s.y - dt2 * av.y, c.z + dt * s.z - dt2 * av.z};
    double gap;
    {
        D3 r = {x.x - ob.p[3 * o], x.y - ob.p[3 * o + 1], x.z - ob.p[3 * o + 2]};
        if (ob.type[o] == PDFC_OBSTACLE_SPHERE || ob.type[o] == PDFC_OBSTACLE_ROTATING_SPHERE) {
            gap = sqrt(dot3(r, r)) - ob.radius[o];
        } else if (ob.type[o] == PDFC_OBSTACLE_PLANE) {
            gap = dot3(r, D3{ob.nrm[3 * o], ob.nrm[3 * o + 1], ob.nrm[3 * o + 2]});
        } else {  // signed distance to the contact point along the normal
            const D3 cp = {C.point[3 * cid], C.point[3 * cid + 1], C.point[3 * cid + 2]};
            gap = dot3(D3{x.x - cp.x, x.y - cp.y, x.z - cp.z}, bn);
        }
    }
    D3 g = {0., 0., 0.};
    if (a.prev_g) {
        int o0, o1;
        obstacle_range(ob, i, o0, o1);
        const float* pg = a.prev_g + 3 * ((size_t)(o - o0) * n + i);
        g = {(double)pg[0], (double)pg[1], (double)pg[2]};
    }
    const D3 sp = {C.speed[3 * cid], C.speed[3 * cid + 1], C.speed[3 * cid + 2]};
    const D3 u = {v.x - sp.x, v.y - sp.y, v.z - sp.z};
    F = {F.x - m * sp.x, F.y - m * sp.y, F.z - m * sp.z};
    float* f = feat + (size_t)cid * kWarmFeatures;
    f[0] = (float)dot3(bn, u);
    f[1] = (float)dot3(bt, u);
    f[2] = (float)dot3(bb, u);
    f[3] = (float)bn.x;
    f[4] = (float)bn.y;
    f[5] = (float)bn.z;
    f[6] = (float)(im * dot3(bn, g));
    f[7] = (float)(im * dot3(bt, g));
    f[8] = (float)(im * dot3(bb, g));
    f[9] = (float)C.mu[cid];
    f[10] = (float)gap;
    f[11] = (float)(im * dot3(bn, F));
}

template <typename T>
__device__ __forceinline__ double as_d(T x) { return (double)x; }

// Operator entries per thread loaded before a task's gather barrier.
constexpr int kPrefetch = 4;

// x op= U[idx[q]] for q = q0 .. q1-1 in order (op: -= when kSub, else +=), with the index and
// entry loads of kPullBatch entries in flight at a time (clamped indices keep every load
// in bounds and unconditional; entries past q1 are loaded but not used). Forward tasks pull
// 2 at a time (most rows have one or two entries; more loads cost issue slots on large
// levels), the top block's gathers 4 (their chains sit in every tile's critical path).
template <bool kSub, int kPullBatch>
__device__ __forceinline__ void pull_entries(const double* __restrict__ U, const int* __restrict__ idx, int q0, int q1,
                                             double& x, double& y, double& z) {
    for (int q = q0; q < q1; q += kPullBatch) {
        int j[kPullBatch];
#pragma unroll
        for (int k = 0; k < kPullBatch; ++k) j[k] = idx[min(q + k, q1 - 1)];
        double u[kPullBatch][3];
#pragma unroll
        for (int k = 0; k < kPullBatch; ++k) {
            const double* p = U + 3 * (size_t)j[k];
            u[k][0] = p[0];
            u[k][1] = p[1];
            u[k][2] = p[2];
        }
#pragma unroll
        for (int k = 0; k < kPullBatch; ++k)
            if (q + k < q1) {
                if (kSub) {
                    x -= u[k][0];
                    y -= u[k][1];
                    z -= u[k][2];
                } else {
                    x += u[k][0];
                    y += u[k][1];
                    z += u[k][2];
                }
            }
    }
}

// x op= the first two of a row's update entries (op: -= when kSub, else +=), named by its record
// (cnt entries, U indices i0 and i1), in list order; both loads are issued before either add.
template <bool kSub>
__device__ __forceinline__ void pull_two(const double* __restrict__ U, int cnt, int i0, int i1, double& x, double& y,
                                         double& z) {
    double a0 = 0., a1 = 0., a2 = 0., b0 = 0., b1 = 0., b2 = 0.;
    if (cnt > 0) {
        const double* p = U + 3 * (size_t)i0;
        a0 = p[0];
        a1 = p[1];
        a2 = p[2];
    }
    if (cnt > 1) {
        const double* p = U + 3 * (size_t)i1;
        b0 = p[0];
        b1 = p[1];
        b2 = p[2];
    }
    if (cnt > 0) {
        if (kSub) {
            x -= a0;
            y -= a1;
            z -= a2;
        } else {
            x += a0;
            y += a1;
            z += a2;
        }
    }
    if (cnt > 1) {
        if (kSub) {
            x -= b0;
            y -= b1;
            z -= b2;
        } else {
            x += b0;
            y += b1;
            z += b2;
        }
    }
}

// All of a front row's update entries in list order: the first two from its record, any
// further ones from the pull list.
template <bool kSub>
__device__ __forceinline__ void pull_row(const DevFactor& Fa, int4 rc, double& x, double& y, double& z) {
    pull_two<kSub>(Fa.U, rc.x, rc.y, rc.z, x, y, z);
    if (rc.x > 2) pull_entries<kSub, 2>(Fa.U, Fa.pull_idx, rc.w + 2, rc.w + rc.x, x, y, z);
}

// Forward substitution, one task = (supernode s, 64 rows of F_s), W waves:
//   b~_s = b_s - sum_children u_c|own ;  [y_s ; u_s] = F_s b~_s (+ children's u_c|below)
// The reduction over the ns columns is split across the W waves (wave w takes
// columns w, w+W, ...) so each wave has a short chain of independent loads; the
// partial sums are combined in LDS in wave order (deterministic).
template <typename T, int W>
__device__ __forceinline__ void forward_task(const DevFactor& Fa, const TaskDesc& d, const double* __restrict__ Bv,
                                             double* __restrict__ Y, double* lds, double (*part)[3][64]) {
    const int ns = d.ns, f = d.f, c0 = d.c0, r0 = d.r0;
    const int n = Fa.n;
    double* bx = lds;
    double* by = lds + ns;
    double* bz = lds + 2 * ns;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int4* rec = Fa.pull_rec + d.pull_base;
    // The first kPrefetch operator entries of this thread's column slice do not depend on
    // the gather: issue them now so their latency overlaps the gather's.
    const int my_row = r0 + lane;
    const int ldf = op_ld(f, sizeof(T));
    const T* Fs = static_cast<const T*>(Fa.F) + d.foff + my_row;
    double pre[kPrefetch];
#pragma unroll
    for (int q = 0; q < kPrefetch; ++q) {
        const int j = wave + q * W;
        pre[q] = (my_row < f && j < ns) ? as_d(Fs[(size_t)j * ldf]) : 0.;
    }
    // own rows: b~ = b - sum of the children's update entries (pull, child order)
    for (int t = tid; t < ns; t += 64 * W) {
        double x = Bv[c0 + t], y = Bv[n + c0 + t], z = Bv[2 * n + c0 + t];
        pull_row<true>(Fa, rec[t], x, y, z);
        bx[t] = x;
        by[t] = y;
        bz[t] = z;
    }
    // this task's below rows: children's entries to pass on
    double ux = 0., uy = 0., uz = 0.;
    if (wave == 0 && my_row >= ns && my_row < f) pull_row<false>(Fa, rec[my_row], ux, uy, uz);
    __syncthreads();
    double ax = 0., ay = 0., az = 0.;
    if (my_row < f) {
#pragma unroll
        for (int q = 0; q < kPrefetch; ++q) {
            const int j = wave + q * W;
            if (j < ns) {
                ax = fma(pre[q], bx[j], ax);
                ay = fma(pre[q], by[j], ay);
                az = fma(pre[q], bz[j], az);
            }
        }
#pragma unroll 8
        for (int j = wave + kPrefetch * W; j < ns; j += W) {
            double fv = as_d(Fs[(size_t)j * ldf]);
            ax = fma(fv, bx[j], ax);
            ay = fma(fv, by[j], ay);
            az = fma(fv, bz[j], az);
        }
    }
    if (W > 1) {
        part[wave][0][lane] = ax;
        part[wave][1][lane] = ay;
        part[wave][2][lane] = az;
        __syncthreads();
        if (wave == 0) {
            ax = part[0][0][lane];
            ay = part[0][1][lane];
            az = part[0][2][lane];
            for (int w = 1; w < W; ++w) {
                ax += part[w][0][lane];
                ay += part[w][1][lane];
                az += part[w][2][lane];
            }
        }
    }
    if (wave == 0 && my_row < f) {
        if (my_row < ns) {
            Y[c0 + my_row] = ax;
            Y[n + c0 + my_row] = ay;
            Y[2 * n + c0 + my_row] = az;
        } else {
            double* u = Fa.U + 3 * (d.uoff + my_row - ns);
            u[0] = ax + ux;
            u[1] = ay + uy;
            u[2] = az + uz;
        }
    }
}

template <typename T, int W>
__global__ void __launch_bounds__(64 * W) k_forward(DevFactor Fa, int t0, const double* __restrict__ Bv,
                                                    double* __restrict__ Y) {
    extern __shared__ double lds[];
    __shared__ double part[W > 1 ? W : 1][3][64];
    forward_task<T, W>(Fa, Fa.desc[t0 + xcd_block(blockIdx.x, gridDim.x)], Bv, Y, lds, part);
}

// Backward substitution, one task = (supernode s, 64 columns), W waves:
//   x_s = F_s^T [ y_s ; -x_R ]      (reduction over the f = ns + nb rows split across waves)
// Final value of top row t: the matvec's partial products summed in chunk order, exactly as
// k_top_finish sums them.
// All the loads are issued before the first add (at most kTopChunksMax chunks).
constexpr int kTopChunksMax = 16;
// Symmetric tiles: the nb source blocks of row t in order, 16 loads in flight at a time.
// Group g's sum (source blocks 16 g .. 16 g + 15 of row t, in order, starting from 0).
__device__ __forceinline__ void top_group_sum(const double* part, int nb, int t, int g, double& x, double& y,
                                              double& z) {
    const double* P = part + (size_t)(t >> 6) * nb * 192 + (t & 63);
    const int c0 = g * kTopChunksMax;
    double px[kTopChunksMax], py[kTopChunksMax], pz[kTopChunksMax];
#pragma unroll
    for (int c = 0; c < kTopChunksMax; ++c)
        if (c0 + c < nb) {
            const double* Q = P + (size_t)(c0 + c) * 192;
            px[c] = Q[0];
            py[c] = Q[64];
            pz[c] = Q[128];
        }
    x = 0.;
    y = 0.;
    z = 0.;
#pragma unroll
    for (int c = 0; c < kTopChunksMax; ++c)
        if (c0 + c < nb) {
            x += px[c];
            y += py[c];
            z += pz[c];
        }
}

// Symmetric tiles: row t's nb source blocks in groups of 16, each group summed in block
// order, then the group sums added in group order (k_top_finish computes the groups on
// separate lanes and combines them in the same order).
__device__ __forceinline__ void top_sum_blocks(const double* part, int nb, int t, double& x, double& y, double& z) {
    top_group_sum(part, nb, t, 0, x, y, z);
    for (int g = 1; g * kTopChunksMax < nb; ++g) {
        double gx, gy, gz;
        top_group_sum(part, nb, t, g, gx, gy, gz);
        x += gx;
        y += gy;
        z += gz;
    }
}

__device__ __forceinline__ void top_sum(const TopFold& fo, int t, double& x, double& y, double& z) {
    if (fo.nb > 0) {
        top_sum_blocks(fo.part, fo.nb, t, x, y, z);
        return;
    }
    double px[kTopChunksMax], py[kTopChunksMax], pz[kTopChunksMax];
#pragma unroll
    for (int c = 0; c < kTopChunksMax; ++c)
        if (c < fo.kchunks) {
            const double* P = fo.part + (size_t)c * 3 * fo.m;
            px[c] = P[t];
            py[c] = P[fo.m + t];
            pz[c] = P[2 * fo.m + t];
        }
    x = 0.;
    y = 0.;
    z = 0.;
#pragma unroll
    for (int c = 0; c < kTopChunksMax; ++c)
        if (c < fo.kchunks) {
            x += px[c];
            y += py[c];
            z += pz[c];
        }
}

// Ordered sum of the partials of row r (parts >= 2): XP[0] + XP[1] + ... per component.
__device__ __forceinline__ void split_sum(const double* __restrict__ XP, int n, int parts, int r, double& x, double& y,
                                          double& z) {
    const size_t n3 = 3 * (size_t)n;
    double px[kMaxSplit], py[kMaxSplit], pz[kMaxSplit];
#pragma unroll
    for (int p = 0; p < kMaxSplit; ++p)
        if (p < parts) {
            px[p] = XP[p * n3 + r];
            py[p] = XP[p * n3 + n + r];
            pz[p] = XP[p * n3 + 2 * n + r];
        }
    x = px[0];
    y = py[0];
    z = pz[0];
#pragma unroll
    for (int p = 1; p < kMaxSplit; ++p)
        if (p < parts) {
            x += px[p];
            y += py[p];
            z += pz[p];
        }
}

template <typename T, int W>
__device__ __forceinline__ void backward_task(const DevFactor& Fa, const TaskDesc& d, const double* __restrict__ Y,
                                              double* __restrict__ X, double* lds, double (*part)[3][64],
                                              const TopFold& fo) {
    const int ns = d.ns, f = d.f, c0 = d.c0, j0 = d.r0;
    const int n = Fa.n;
    // rows [i0, i1) of the front: all of them, or this part's share of a split reduction
    const int parts = d.split >> 16, mypart = d.split & 0xffff;
    const int i0 = parts ? (int)(d.rng & 0xffffffff) : 0;
    const int i1 = parts ? (int)(d.rng >> 32) : f;
    const int len = i1 - i0;
    double* gx = lds;
    double* gy = lds + len;
    double* gz = lds + 2 * len;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = j0 + lane;
    const int ldt = op_ld(ns, sizeof(T));
    const T* Ft = static_cast<const T*>(Fa.FT) + d.foff + op_t_offset(ns, f, sizeof(T)) + j;
    double pre[kPrefetch];  // operator entries independent of the gather, issued first
#pragma unroll
    for (int q = 0; q < kPrefetch; ++q) {
        const int i = i0 + wave + q * W;
        pre[q] = (j < ns && i < i1) ? as_d(Ft[(size_t)i * ldt]) : 0.;
    }
    const int a1 = min(i1, ns);
    for (int i = i0 + tid; i < a1; i += 64 * W) {
        gx[i - i0] = Y[c0 + i];
        gy[i - i0] = Y[n + c0 + i];
        gz[i - i0] = Y[2 * n + c0 + i];
    }
    const int* R = Fa.rows + d.roff;
    const int* RP = Fa.rparts ? Fa.rparts + d.roff : nullptr;
    for (int i = max(i0, ns) + tid; i < i1; i += 64 * W) {
        const int t = i - ns;
        const int r = R[t];
        const int np = RP ? RP[t] : 0;
        const int tr = fo.part ? fo.top_of[r] : -1;
        double x, y, z;
        if (tr >= 0) {
            top_sum(fo, tr, x, y, z);
        } else if (np > 1) {
            split_sum(Fa.XP, n, np, r, x, y, z);
        } else {
            x = X[r];
            y = X[n + r];
            z = X[2 * n + r];
        }
        gx[i - i0] = -x;
        gy[i - i0] = -y;
        gz[i - i0] = -z;
    }
    __syncthreads();
    double ax = 0., ay = 0., az = 0.;
    if (j < ns) {
#pragma unroll
        for (int q = 0; q < kPrefetch; ++q) {
            const int i = i0 + wave + q * W;
            if (i < i1) {
                ax = fma(pre[q], gx[i - i0], ax);
                ay = fma(pre[q], gy[i - i0], ay);
                az = fma(pre[q], gz[i - i0], az);
            }
        }
#pragma unroll 8
        for (int i = i0 + wave + kPrefetch * W; i < i1; i += W) {
            double fv = as_d(Ft[(size_t)i * ldt]);
            ax = fma(fv, gx[i - i0], ax);
            ay = fma(fv, gy[i - i0], ay);
            az = fma(fv, gz[i - i0], az);
        }
    }
    if (W > 1) {
        part[wave][0][lane] = ax;
        part[wave][1][lane] = ay;
        part[wave][2][lane] = az;
        __syncthreads();
        if (wave == 0) {
            ax = part[0][0][lane];
            ay = part[0][1][lane];
            az = part[0][2][lane];
            for (int w = 1; w < W; ++w) {
                ax += part[w][0][lane];
                ay += part[w][1][lane];
                az += part[w][2][lane];
            }
        }
    }
    if (wave == 0 && j < ns) {
        double* out = parts ? Fa.XP + (size_t)mypart * 3 * n : X;
        out[c0 + j] = ax;
        out[n + c0 + j] = ay;
        out[2 * n + c0 + j] = az;
    }
}

template <typename T, int W>
__global__ void __launch_bounds__(64 * W) k_backward(DevFactor Fa, int t0, const double* __restrict__ Y,
                                                     double* __restrict__ X, TopFold fo, SplitFin fin) {
    extern __shared__ double lds[];
    __shared__ double part[W > 1 ? W : 1][3][64];
    if (fo.part && (int)blockIdx.x >= fo.ntasks) {  // folded top-block finish: write the top rows
        const int t = ((int)blockIdx.x - fo.ntasks) * (int)blockDim.x + (int)threadIdx.x;
        if (t < fo.m) {
            double x, y, z;
            top_sum(fo, t, x, y, z);
            const int r = fo.rows[t];
            X[r] = x;
            X[Fa.n + r] = y;
            X[2 * Fa.n + r] = z;
        }
        return;
    }
    if (fin.count > 0 && (int)blockIdx.x >= fin.first) {  // final values of the previous level's split rows
        const int t = ((int)blockIdx.x - fin.first) * (int)blockDim.x + (int)threadIdx.x;
        if (t < fin.count) {
            const int r = fin.rows[t];
            double x, y, z;
            split_sum(Fa.XP, Fa.n, fin.parts, r, x, y, z);
            X[r] = x;
            X[Fa.n + r] = y;
            X[2 * Fa.n + r] = z;
        }
        return;
    }
    // the task workgroups come first; the extra ones above (fold, split finish) keep their order
    const int ntask = fo.part ? fo.ntasks : (fin.count > 0 ? fin.first : (int)gridDim.x);
    backward_task<T, W>(Fa, Fa.desc[t0 + xcd_block(blockIdx.x, ntask)], Y, X, lds, part, fo);
}

// ---------------------------------------------------------------- merged levels
// Consecutive levels in one launch, one 16-wave workgroup per group: the supernodes of the
// range under one root (their highest ancestor inside the range). A group's supernode reads
// its children (forward) or ancestors (backward) only inside the group or outside the range,
// where an earlier launch wrote them, so the levels are separated by the workgroup barrier
// alone and every group's data stays in its CU's cache hierarchy. At each level the group's
// tasks run W waves each, 16 / W at a time ("slots"); a slot computes exactly what the level
// kernel's task with W waves computes, in the same operation order, so results are
// bit-identical to the level-by-level sweep with the same waves per task. Idle slots (past
// the level's last task) take part in every barrier.

// forward_task for one slot: W waves (runtime), wv = wave within the slot, task < 0 idle.
template <typename T>
__device__ __forceinline__ void forward_slot(const DevFactor& Fa, int task, int W, int wv, const double* __restrict__ Bv,
                                             double* __restrict__ Y, double* lds, double (*part)[3][64]) {
    TaskDesc d{};
    if (task >= 0) d = Fa.desc[task];
    const int ns = d.ns, f = d.f, c0 = d.c0, r0 = d.r0;
    const int n = Fa.n;
    double* bx = lds;
    double* by = lds + ns;
    double* bz = lds + 2 * ns;
    const int lane = threadIdx.x & 63, tid = wv * 64 + lane, nthr = 64 * W;
    const int4* rec = Fa.pull_rec + d.pull_base;
    const int my_row = r0 + lane;
    const int ldf = op_ld(f, sizeof(T));
    const T* Fs = static_cast<const T*>(Fa.F) + d.foff + my_row;
    double pre[kPrefetch];
#pragma unroll
    for (int q = 0; q < kPrefetch; ++q) {
        const int j = wv + q * W;
        pre[q] = (my_row < f && j < ns) ? as_d(Fs[(size_t)j * ldf]) : 0.;
    }
    for (int t = tid; t < ns; t += nthr) {
        double x = Bv[c0 + t], y = Bv[n + c0 + t], z = Bv[2 * n + c0 + t];
        pull_row<true>(Fa, rec[t], x, y, z);
        bx[t] = x;
        by[t] = y;
        bz[t] = z;
    }
    double ux = 0., uy = 0., uz = 0.;
    if (wv == 0 && my_row >= ns && my_row < f) pull_row<false>(Fa, rec[my_row], ux, uy, uz);
    __syncthreads();
    double ax = 0., ay = 0., az = 0.;
    if (my_row < f) {
#pragma unroll
        for (int q = 0; q < kPrefetch; ++q) {
            const int j = wv + q * W;
            if (j < ns) {
                ax = fma(pre[q], bx[j], ax);
                ay = fma(pre[q], by[j], ay);
                az = fma(pre[q], bz[j], az);
            }
        }
#pragma unroll 8
        for (int j = wv + kPrefetch * W; j < ns; j += W) {
            double fv = as_d(Fs[(size_t)j * ldf]);
            ax = fma(fv, bx[j], ax);
            ay = fma(fv, by[j], ay);
            az = fma(fv, bz[j], az);
        }
    }
    if (W > 1) {
        part[wv][0][lane] = ax;
        part[wv][1][lane] = ay;
        part[wv][2][lane] = az;
        __syncthreads();
        if (wv == 0) {
            ax = part[0][0][lane];
            ay = part[0][1][lane];
            az = part[0][2][lane];
            for (int w = 1; w < W; ++w) {
                ax += part[w][0][lane];
                ay += part[w][1][lane];
                az += part[w][2][lane];
            }
        }
    }
    if (wv == 0 && my_row < f) {
        if (my_row < ns) {
            Y[c0 + my_row] = ax;
            Y[n + c0 + my_row] = ay;
            Y[2 * n + c0 + my_row] = az;
        } else {
            double* u = Fa.U + 3 * (d.uoff + my_row - ns);
            u[0] = ax + ux;
            u[1] = ay + uy;
            u[2] = az + uz;
        }
    }
}

// backward_task for one slot (no folded top-block finish: the top rows are in X).
template <typename T>
__device__ __forceinline__ void backward_slot(const DevFactor& Fa, int task, int W, int wv, const double* __restrict__ Y,
                                              double* __restrict__ X, double* lds, double (*part)[3][64]) {
    TaskDesc d{};
    if (task >= 0) d = Fa.desc[task];
    const int ns = d.ns, nb = d.nb, f = d.f, c0 = d.c0, j0 = d.r0;
    const int n = Fa.n;
    double* gx = lds;
    double* gy = lds + f;
    double* gz = lds + 2 * f;
    const int lane = threadIdx.x & 63, tid = wv * 64 + lane, nthr = 64 * W;
    const int j = j0 + lane;
    const int ldt = op_ld(ns, sizeof(T));
    const T* Ft = static_cast<const T*>(Fa.FT) + d.foff + op_t_offset(ns, f, sizeof(T)) + j;
    double pre[kPrefetch];
#pragma unroll
    for (int q = 0; q < kPrefetch; ++q) {
        const int i = wv + q * W;
        pre[q] = (j < ns && i < f) ? as_d(Ft[(size_t)i * ldt]) : 0.;
    }
    for (int t = tid; t < ns; t += nthr) {
        gx[t] = Y[c0 + t];
        gy[t] = Y[n + c0 + t];
        gz[t] = Y[2 * n + c0 + t];
    }
    const int* R = Fa.rows + d.roff;
    for (int t = tid; t < nb; t += nthr) {
        const int r = R[t];
        gx[ns + t] = -X[r];
        gy[ns + t] = -X[n + r];
        gz[ns + t] = -X[2 * n + r];
    }
    __syncthreads();
    double ax = 0., ay = 0., az = 0.;
    if (j < ns) {
#pragma unroll
        for (int q = 0; q < kPrefetch; ++q) {
            const int i = wv + q * W;
            if (i < f) {
                ax = fma(pre[q], gx[i], ax);
                ay = fma(pre[q], gy[i], ay);
                az = fma(pre[q], gz[i], az);
            }
        }
#pragma unroll 8
        for (int i = wv + kPrefetch * W; i < f; i += W) {
            double fv = as_d(Ft[(size_t)i * ldt]);
            ax = fma(fv, gx[i], ax);
            ay = fma(fv, gy[i], ay);
            az = fma(fv, gz[i], az);
        }
    }
    if (W > 1) {
        part[wv][0][lane] = ax;
        part[wv][1][lane] = ay;
        part[wv][2][lane] = az;
        __syncthreads();
        if (wv == 0) {
            ax = part[0][0][lane];
            ay = part[0][1][lane];
            az = part[0][2][lane];
            for (int w = 1; w < W; ++w) {
                ax += part[w][0][lane];
                ay += part[w][1][lane];
                az += part[w][2][lane];
            }
        }
    }
    if (wv == 0 && j < ns) {
        X[c0 + j] = ax;
        X[n + c0 + j] = ay;
        X[2 * n + c0 + j] = az;
    }
}

// The control flow below is uniform across the workgroup (group offsets and level info are
// the same for every wave), so every wave meets the same barriers.
template <typename T, bool kForward>
__global__ void __launch_bounds__(1024) k_sweep_merged(DevFactor Fa, MergedRange R, const double* __restrict__ In,
                                                       double* __restrict__ Out) {
    extern __shared__ double lds[];
    __shared__ double part[16][3][64];
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int* gp = R.ptr + (size_t)blockIdx.x * (R.nl + 1);
    for (int k = 0; k < R.nl; ++k) {
        const int W = R.info[2 * k], stride = R.info[2 * k + 1];
        const int per = 16 / W;
        const int slot = wave / W, wv = wave - slot * W;
        const int a = gp[k], b = gp[k + 1];
        for (int base = a; base < b; base += per) {
            const int t = base + slot;
            const int task = t < b ? R.task[t] : -1;
            if (kForward)
                forward_slot<T>(Fa, task, W, wv, In, Out, lds + (size_t)slot * stride, part + slot * W);
            else
                backward_slot<T>(Fa, task, W, wv, In, Out, lds + (size_t)slot * stride, part + slot * W);
            // tasks of one level are independent; a slot's LDS and partial sums are reused only
            // after its own waves passed the slot's last barrier (W = 1: one wave, in order)
        }
        __syncthreads();  // this level's results are read by the next level's tasks
    }
}

// ---------------------------------------------------------------- top block
__global__ void k_top_gather(DevTop T, const double* __restrict__ U, const double* __restrict__ B, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.m) return;
    const int4 tr = T.trec[t];
    double bx = B[tr.x], by = B[n + tr.x], bz = B[2 * n + tr.x];
    pull_two<true>(U, tr.y, tr.z, tr.w, bx, by, bz);
    if (tr.y > 2) pull_entries<true, 4>(U, T.cidx, T.cptr[t] + 2, T.cptr[t + 1], bx, by, bz);
    T.bt[t] = bx;
    T.bt[T.m + t] = by;
    T.bt[2 * T.m + t] = bz;
}

// Partial products of S^-1 b~: workgroup (row block of 64, reduction chunk), 16 waves
// splitting the chunk; b~ entries are wave-uniform loads.
template <typename T>
__global__ void __launch_bounds__(1024) k_top_matvec(DevTop Tp) {
    __shared__ double part[16][3][64];
    const int m = Tp.m;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = blockIdx.x * 64 + lane;
    const int klen = (m + Tp.kchunks - 1) / Tp.kchunks;
    const int k0 = blockIdx.y * klen, k1 = min(m, k0 + klen);
    double ax = 0., ay = 0., az = 0.;
    if (row < m) {
        const T* S = static_cast<const T*>(Tp.Sinv) + row;
#pragma unroll 8
        for (int k = k0 + wave; k < k1; k += 16) {
            double f = (double)S[(size_t)k * m];
            ax = fma(f, Tp.bt[k], ax);
            ay = fma(f, Tp.bt[m + k], ay);
            az = fma(f, Tp.bt[2 * m + k], az);
        }
    }
    part[wave][0][lane] = ax;
    part[wave][1][lane] = ay;
    part[wave][2][lane] = az;
    __syncthreads();
    if (wave != 0 || row >= m) return;
    for (int w = 1; w < 16; ++w) {
        ax += part[w][0][lane];
        ay += part[w][1][lane];
        az += part[w][2][lane];
    }
    double* P = Tp.part + (size_t)blockIdx.y * 3 * m;
    P[row] = ax;
    P[m + row] = ay;
    P[2 * m + row] = az;
}

// Symmetric tile products, one workgroup (4 waves) per block pair I <= J of S^-1 (A below;
// A(r, k) = Sinv[k m + r]). Wave w holds columns w, w + 4, ... of the 64 x 64 block in
// registers (lane = row), so the block is read from memory once:
//   rows I:  y_I(r) = sum_c A(r0 + r, c0 + c) b_J(c)      (waves summed 0..3, then the block's
//   rows J:  y_J(c) = sum_r A(r0 + r, c0 + c) b_I(r)       partial goes to (I, J) / (J, I))
// b entries are gathered here (b_T - lower updates, as k_top_gather).
template <typename T>
__global__ void __launch_bounds__(256) k_top_tile(DevTop Tp, const double* __restrict__ U,
                                                  const double* __restrict__ B, int n) {
    __shared__ double blk[64][65];
    __shared__ double xb[2][3][64];
    __shared__ double red[4][3][64];
    const int I = Tp.tiles[2 * blockIdx.x], J = Tp.tiles[2 * blockIdx.x + 1];
    const int m = Tp.m, nb = Tp.nb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = 64 * I, c0 = 64 * J;
    // this tile's 64 x 64 block, contiguous in the packed copy (zero outside the matrix)
    const T* S = static_cast<const T*>(Tp.Stile) + (size_t)blockIdx.x * 4096 + lane;
    double a[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) a[s] = (double)S[64 * (wave + 4 * s)];
    if (wave < 2 && (wave == 0 || I != J)) {
        const int t = (wave == 0 ? r0 : c0) + lane;
        double bx = 0., by = 0., bz = 0.;
        if (t < m && Tp.pregather) {
            bx = Tp.bt[t];
            by = Tp.bt[m + t];
            bz = Tp.bt[2 * m + t];
        } else if (t < m) {
            const int4 tr = Tp.trec[t];
            bx = B[tr.x];
            by = B[n + tr.x];
            bz = B[2 * n + tr.x];
            pull_two<true>(U, tr.y, tr.z, tr.w, bx, by, bz);
            if (tr.y > 2) pull_entries<true, 4>(U, Tp.cidx, Tp.cptr[t] + 2, Tp.cptr[t + 1], bx, by, bz);
        }
        xb[wave][0][lane] = bx;
        xb[wave][1][lane] = by;
        xb[wave][2][lane] = bz;
    }
    __syncthreads();
    const int jb = I != J ? 1 : 0;  // b_J
    double ax = 0., ay = 0., az = 0.;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int c = wave + 4 * s;
        ax = fma(a[s], xb[jb][0][c], ax);
        ay = fma(a[s], xb[jb][1][c], ay);
        az = fma(a[s], xb[jb][2][c], az);
        if (I != J) blk[lane][c] = a[s];
    }
    red[wave][0][lane] = ax;
    red[wave][1][lane] = ay;
    red[wave][2][lane] = az;
    __syncthreads();
    if (wave == 0) {
        double* P = Tp.part + ((size_t)I * nb + J) * 192 + lane;
        for (int k = 0; k < 3; ++k) P[64 * k] = ((red[0][k][lane] + red[1][k][lane]) + red[2][k][lane]) + red[3][k][lane];
    }
    if (I == J) return;
    __syncthreads();  // red is reused
    // rows J: lane = column c, wave w sums rows 16 w .. 16 w + 15
    ax = 0.;
    ay = 0.;
    az = 0.;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int r = 16 * wave + s;
        const double f = blk[r][lane];
        ax = fma(f, xb[0][0][r], ax);
        ay = fma(f, xb[0][1][r], ay);
        az = fma(f, xb[0][2][r], az);
    }
    red[wave][0][lane] = ax;
    red[wave][1][lane] = ay;
    red[wave][2][lane] = az;
    __syncthreads();
    if (wave == 0) {
        double* P = Tp.part + ((size_t)J * nb + I) * 192 + lane;
        for (int k = 0; k < 3; ++k) P[64 * k] = ((red[0][k][lane] + red[1][k][lane]) + red[2][k][lane]) + red[3][k][lane];
    }
}

// Symmetric tiles: one workgroup per 64-row block, wave g sums group g of the block's rows
// (as top_sum_blocks; lane = row, so every load instruction reads 64 consecutive partials),
// wave 0 adds the group sums in group order and scatters.
constexpr int kFinishLanes = 8;  // groups (waves) per row block
__global__ void __launch_bounds__(64 * kFinishLanes) k_top_finish_blocks(DevTop T, double* __restrict__ X, int n) {
    __shared__ double gsum[kFinishLanes][3][64];
    const int g = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = 64 * blockIdx.x + lane;
    const int groups = (T.nb + kTopChunksMax - 1) / kTopChunksMax;
    double x = 0., y = 0., z = 0.;
    if (t < T.m && g < groups) top_group_sum(T.part, T.nb, t, g, x, y, z);
    gsum[g][0][lane] = x;
    gsum[g][1][lane] = y;
    gsum[g][2][lane] = z;
    __syncthreads();
    if (g != 0 || t >= T.m) return;
    double sx = gsum[0][0][lane], sy = gsum[0][1][lane], sz = gsum[0][2][lane];
    for (int k = 1; k < groups; ++k) {
        sx += gsum[k][0][lane];
        sy += gsum[k][1][lane];
        sz += gsum[k][2][lane];
    }
    const int row = T.rows[t];
    X[row] = sx;
    X[n + row] = sy;
    X[2 * n + row] = sz;
}

__global__ void k_top_finish(DevTop T, double* __restrict__ X, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.m) return;
    double ax = 0., ay = 0., az = 0.;
    for (int c = 0; c < T.kchunks; ++c) {
        const double* P = T.part + (size_t)c * 3 * T.m;
        ax += P[t];
        ay += P[T.m + t];
        az += P[2 * T.m + t];
    }
    int row = T.rows[t];
    X[row] = ax;
    X[n + row] = ay;
    X[2 * n + row] = az;
}

__global__ void k_unit3(double* B, int n, const int* rows, int m, int j) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * n) B[i] = 0.;
}
__global__ void k_unit3_set(double* B, int n, const int* rows, int m, int j) {
    int c = threadIdx.x;
    if (c < 3 && j + c < m) B[(size_t)c * n + rows[j + c]] = 1.;
}
__global__ void k_take_cols(const double* X, int n, const int* rows, int m, int j, double* S) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    for (int c = 0; c < 3; ++c)
        if (j + c < m) S[(size_t)(j + c) * m + t] = X[(size_t)c * n + rows[t]];
}
// Upper block triangle of S^-1 in tile order: tile p = (I, J) holds A(64 I + r, 64 J + c) at
// 4096 p + 64 c + r, zero outside the m x m matrix.
template <typename T>
__global__ void k_pack_tiles(const double* __restrict__ S, T* __restrict__ P, int m, const int* __restrict__ tiles,
                             int ntiles) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)ntiles * 4096) return;
    const int p = (int)(e >> 12), c = (int)((e >> 6) & 63), r = (int)(e & 63);
    const int row = 64 * tiles[2 * p] + r, col = 64 * tiles[2 * p + 1] + c;
    P[e] = (row < m && col < m) ? (T)S[(size_t)col * m + row] : (T)0;
}

__global__ void k_to_float(const double* src, float* dst, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = (float)src[i];
}

// r = b - (M x + dt2 A^T A x)
__global__ void k_residual(DevModel M, double dt2, const double* __restrict__ b, const double* __restrict__ x,
                           double* __restrict__ r) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = M.n;
    if (i >= n) return;
    const D3 s = ata_product(M, x, i);
    double m = M.mass[i];
    r[i] = b[i] - (m * x[i] + dt2 * s.x);
    r[n + i] = b[n + i] - (m * x[n + i] + dt2 * s.y);
    r[2 * n + i] = b[2 * n + i] - (m * x[2 * n + i] + dt2 * s.z);
}

inline unsigned blocks_for(size_t count, int bs = kBlock) { return (unsigned)((count + bs - 1) / bs); }
// Up to two blocks per CU (256 CUs) the local kernel's latency matters more than its occupancy.
constexpr unsigned kLocalLowLatencyBlocks = 512;

}  // namespace

void launch_fill(double* p, size_t count, double value, hipStream_t st) {
    if (count) hipLaunchKernelGGL(k_fill, dim3(blocks_for(count)), dim3(kBlock), 0, st, p, count, value);
}
void launch_fill_u64(unsigned long long* p, size_t count, hipStream_t st) {
    if (count) hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(count)), dim3(kBlock), 0, st, p, count);
}
void launch_copy(double* dst, const double* src, size_t count, hipStream_t st) {
    if (count) hipLaunchKernelGGL(k_copy, dim3(blocks_for(count)), dim3(kBlock), 0, st, dst, src, count);
}
void launch_axpy(double* x, const double* dx, size_t count, hipStream_t st) {
    if (count) hipLaunchKernelGGL(k_axpy, dim3(blocks_for(count)), dim3(kBlock), 0, st, x, dx, count);
}
void launch_step_begin(const DevModel& M, const double* qn, const double* vstate, double* v0, double* cvec,
                       const StepArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_step_begin, dim3(blocks_for(M.n)), dim3(kBlock), 0, st, M, qn, vstate, v0, cvec, a);
    if (M.n_att > 0)
        hipLaunchKernelGGL(k_attach, dim3(blocks_for(M.n_att)), dim3(kBlock), 0, st, M, cvec, a.dt);
}
void launch_detect(const DevModel& M, const double* qn, const DevObstacles& ob, int* cflags, int* ccnt, int* cstart,
                   int* blocksum, int* n_contacts, DevContacts C, hipStream_t st) {
    unsigned nb = blocks_for(M.n);
    hipLaunchKernelGGL(k_detect_count, dim3(nb), dim3(kBlock), 0, st, M, qn, ob, cflags, ccnt, blocksum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, blocksum, (int)nb, n_contacts);
    hipLaunchKernelGGL(k_detect_write, dim3(nb), dim3(kBlock), 0, st, M, qn, ob, cflags, ccnt, blocksum, cstart, C);
}
void launch_local(const DevModel& M, const double* qn, const double* vcur, const double* cvec, double* rhs,
                  const int* ccnt, const int* cstart, DevContacts C, const StepArgs& a, int iteration,
                  const LocalSelf& ls, hipStream_t st) {
    const unsigned blocks = blocks_for(M.n);
    if (blocks <= kLocalLowLatencyBlocks) {
        static const int lb = [] {
            const char* e = std::getenv("PDFC_LOCAL_BLOCK");
            const int v = e ? std::atoi(e) : 64;
            return (v == 64 || v == 128 || v == 256) ? v : 64;
        }();
        launch_k(k_local<true>, dim3((M.n + lb - 1) / lb), dim3(lb), 0, st, M, qn, vcur, cvec, rhs, ccnt, cstart, C, a,
                 iteration, ls);
    } else
        launch_k(k_local<false>, dim3(blocks), dim3(kBlock), 0, st, M, qn, vcur, cvec, rhs, ccnt, cstart, C, a,
                 iteration, ls);
}
void launch_finish(const DevModel& M, double* qn, const double* vlast, double* vstate, const int* ccnt,
                   const int* cstart, DevContacts C, const DevObstacles& ob, const StepArgs& a, int iteration,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_finish, dim3(blocks_for(std::max(M.n, ob.count))), dim3(kBlock), 0, st, M, qn, vlast, vstate,
                       ccnt, cstart, C, ob, a, iteration);
}
void launch_warm_features(const DevModel& M, const double* qn, const double* v0, const double* cvec,
                          const int* ncontacts, int cap, DevContacts C, const DevObstacles& ob, const StepArgs& a,
                          float* feat, hipStream_t st) {
    if (cap > 0)
        hipLaunchKernelGGL(k_warm_features, dim3(blocks_for(cap)), dim3(kBlock), 0, st, M, qn, v0, cvec, ncontacts, C,
                           ob, a, feat);
}
// Waves per task: the reduction length (ns forward, ns+nb backward) divided into
// chains of about 8 loads, at most 16 waves (1024 threads).
int default_waves(int kmax) {
    int w = 1;
    while (w < 16 && w * 8 < kmax) w <<= 1;
    return w;
}

template <typename T>
static void forward_dispatch(int W, dim3 g, size_t lds, hipStream_t st, const DevFactor& F, int t0, const double* B,
                             double* Y) {
    switch (W) {
        case 1: launch_k(k_forward<T, 1>, g, dim3(64), lds, st, F, t0, B, Y); break;
        case 2: launch_k(k_forward<T, 2>, g, dim3(128), lds, st, F, t0, B, Y); break;
        case 4: launch_k(k_forward<T, 4>, g, dim3(256), lds, st, F, t0, B, Y); break;
        case 8: launch_k(k_forward<T, 8>, g, dim3(512), lds, st, F, t0, B, Y); break;
        default: launch_k(k_forward<T, 16>, g, dim3(1024), lds, st, F, t0, B, Y); break;
    }
}
template <typename T>
static void backward_dispatch(int W, dim3 g, size_t lds, hipStream_t st, const DevFactor& F, int t0, const double* Y,
                              double* X, const TopFold& fo, const SplitFin& fin) {
    switch (W) {
        case 1: launch_k(k_backward<T, 1>, g, dim3(64), lds, st, F, t0, Y, X, fo, fin); break;
        case 2: launch_k(k_backward<T, 2>, g, dim3(128), lds, st, F, t0, Y, X, fo, fin); break;
        case 4: launch_k(k_backward<T, 4>, g, dim3(256), lds, st, F, t0, Y, X, fo, fin); break;
        case 8: launch_k(k_backward<T, 8>, g, dim3(512), lds, st, F, t0, Y, X, fo, fin); break;
        default: launch_k(k_backward<T, 16>, g, dim3(1024), lds, st, F, t0, Y, X, fo, fin); break;
    }
}

void launch_forward_level(const DevFactor& F, bool fp32, int t0, int ntasks, const double* B, double* Y, int max_ns,
                          int W, hipStream_t st) {
    if (ntasks <= 0) return;
    size_t lds = (size_t)3 * max_ns * sizeof(double);
    if (fp32)
        forward_dispatch<float>(W, dim3(ntasks), lds, st, F, t0, B, Y);
    else
        forward_dispatch<double>(W, dim3(ntasks), lds, st, F, t0, B, Y);
}
void launch_backward_level(const DevFactor& F, bool fp32, int t0, int ntasks, const double* Y, double* X,
                           int max_front, int W, hipStream_t st, const TopFold* fold, const SplitFin* fin) {
    TopFold fo{};
    if (fold && fold->part && fold->m > 0) {
        fo = *fold;
        fo.ntasks = ntasks;
        fo.extra = (fo.m + 64 * W - 1) / (64 * W);
    }
    SplitFin sf{};
    int fin_extra = 0;
    if (fin && fin->count > 0) {
        if (fo.extra > 0) throw std::logic_error("a backward launch has either the top fold or split rows to finish");
        if (fin->parts < 2 || fin->parts > kMaxSplit) throw std::logic_error("split parts out of range");
        sf = *fin;
        sf.first = ntasks;
        fin_extra = (sf.count + 64 * W - 1) / (64 * W);
    }
    const int blocks = ntasks + fo.extra + fin_extra;
    if (blocks <= 0) return;
    size_t lds = (size_t)3 * max_front * sizeof(double);
    if (fp32)
        backward_dispatch<float>(W, dim3(blocks), lds, st, F, t0, Y, X, fo, sf);
    else
        backward_dispatch<double>(W, dim3(blocks), lds, st, F, t0, Y, X, fo, sf);
}
void launch_sweep_merged(const DevFactor& F, bool fp32, bool forward, const MergedRange& R, const double* In,
                         double* Out, hipStream_t st) {
    if (R.groups <= 0) return;
    const dim3 g(R.groups), b(1024);
    if (forward) {
        if (fp32)
            launch_k(k_sweep_merged<float, true>, g, b, R.lds_bytes, st, F, R, In, Out);
        else
            launch_k(k_sweep_merged<double, true>, g, b, R.lds_bytes, st, F, R, In, Out);
    } else {
        if (fp32)
            launch_k(k_sweep_merged<float, false>, g, b, R.lds_bytes, st, F, R, In, Out);
        else
            launch_k(k_sweep_merged<double, false>, g, b, R.lds_bytes, st, F, R, In, Out);
    }
}
void launch_top(const DevTop& T, bool fp32, const double* U, const double* B, double* X, int n, hipStream_t st,
                bool finish) {
    if (T.m <= 0) return;
    if (T.nb > 0) {
        if (T.pregather) launch_k(k_top_gather, dim3(blocks_for(T.m)), dim3(kBlock), 0, st, T, U, B, n);
        if (fp32)
            launch_k(k_top_tile<float>, dim3(T.ntiles), dim3(256), 0, st, T, U, B, n);
        else
            launch_k(k_top_tile<double>, dim3(T.ntiles), dim3(256), 0, st, T, U, B, n);
        // nb <= kFinishLanes * kTopChunksMax (the host offers the tiles for m <= 8192 only)
        static_assert(64 % kFinishLanes == 0 && kFinishLanes * kTopChunksMax == 128, "finish lanes");
        if (finish)
            launch_k(k_top_finish_blocks, dim3(T.nb), dim3(64 * kFinishLanes), 0, st, T, X, n);
        return;
    }
    launch_k(k_top_gather, dim3(blocks_for(T.m)), dim3(kBlock), 0, st, T, U, B, n);
    dim3 g((T.m + 63) / 64, T.kchunks);
    if (fp32)
        launch_k(k_top_matvec<float>, g, dim3(1024), 0, st, T);
    else
        launch_k(k_top_matvec<double>, g, dim3(1024), 0, st, T);
    if (finish) launch_k(k_top_finish, dim3(blocks_for(T.m)), dim3(kBlock), 0, st, T, X, n);
}
void launch_unit3(double* B, int n, const int* rows, int m, int j, hipStream_t st) {
    hipLaunchKernelGGL(k_unit3, dim3(blocks_for(3 * (size_t)n)), dim3(kBlock), 0, st, B, n, rows, m, j);
    hipLaunchKernelGGL(k_unit3_set, dim3(1), dim3(64), 0, st, B, n, rows, m, j);
}
void launch_take_cols(const double* X, int n, const int* rows, int m, int j, double* S, hipStream_t st) {
    hipLaunchKernelGGL(k_take_cols, dim3(blocks_for(m)), dim3(kBlock), 0, st, X, n, rows, m, j, S);
}
void launch_pack_tiles(const double* S, void* P, bool fp32, int m, const int* tiles, int ntiles, hipStream_t st) {
    const size_t count = (size_t)ntiles * 4096;
    if (!count) return;
    if (fp32)
        hipLaunchKernelGGL(k_pack_tiles<float>, dim3(blocks_for(count)), dim3(kBlock), 0, st, S,
                           static_cast<float*>(P), m, tiles, ntiles);
    else
        hipLaunchKernelGGL(k_pack_tiles<double>, dim3(blocks_for(count)), dim3(kBlock), 0, st, S,
                           static_cast<double*>(P), m, tiles, ntiles);
}
void launch_to_float(const double* src, float* dst, size_t count, hipStream_t st) {
    if (count) hipLaunchKernelGGL(k_to_float, dim3(blocks_for(count)), dim3(kBlock), 0, st, src, dst, count);
}
void launch_residual(const DevModel& M, double dt2, const double* b, const double* x, double* r, hipStream_t st) {
    hipLaunchKernelGGL(k_residual, dim3(blocks_for(M.n)), dim3(kBlock), 0, st, M, dt2, b, x, r);
}

}  // namespace pdfc
