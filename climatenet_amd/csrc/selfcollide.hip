// Self-collision on gfx950: detection (hash grid + exact closest-point test) and
// the level-scheduled self-contact friction of the local-global iteration.
//
// Detection restates getMeshPotentialTriangleCollision / checkMeshAllCollision
// (src/geometry/geometry_util.cpp:229-327): closed boxes, vertex box x +- 1.1 tol,
// triangle bounding box, exact test |cp - x|^2 < tol^2 with Ericson's closest
// point (geometry_util.cpp:506-616, early exit against the plain tolerance).
// Candidate lists come out sorted by (vertex, triangle) in scene numbering and
// match the CPU oracle bit for bit (-ffp-contract=off, correctly rounded div/sqrt).
//
// Friction restates SimulationFrictionEstimation::computeRhs self-contact levels
// (src/physics/SimulationFrictionEstimation.cpp:431-587) with deterministic
// per-vertex accumulation in the reference's serial order.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace pdfc {
namespace {

constexpr int kB = 256;

struct P3 {
    double x, y, z;
};
__device__ __forceinline__ P3 at_perm(const double* q, int n, int p) { return {q[p], q[n + p], q[2 * n + p]}; }
__device__ __forceinline__ P3 psub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ P3 padd(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ P3 pscale(double s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double pdot(P3 a, P3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

__device__ __forceinline__ double cell_size(const DevSelfDetect& D) {
    double E = __longlong_as_double((long long)*D.ext_bits);
    double c = E * (1. + 1e-9) + 2. * D.t11 * (1. + 1e-9);
    return c > 1e-12 ? c : 1e-12;
}
__device__ __forceinline__ long long cell_of(double x, double c) { return (long long)floor(x / c); }
__device__ __forceinline__ unsigned cell_hash(long long ix, long long iy, long long iz, int P) {
    unsigned h = ((unsigned)ix * 73856093u) ^ ((unsigned)iy * 19349663u) ^ ((unsigned)iz * 83492791u);
    return h & (unsigned)(P - 1);
}

// Ericson closest point, barycentric output (same operation order as the CPU oracle).
__device__ void closest_point(P3 p, P3 a, P3 b, P3 c, double tol2, double* bary, P3& cp) {
    P3 ap = psub(p, a), ab = psub(b, a), ac = psub(c, a);
    if ((pdot(ap, ap) - pdot(ab, ab)) - pdot(ac, ac) > tol2) { cp = a; bary[0] = 1; bary[1] = 0; bary[2] = 0; return; }
    double d1 = pdot(ab, ap), d2 = pdot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { cp = a; bary[0] = 1; bary[1] = 0; bary[2] = 0; return; }
    P3 bp = psub(p, b);
    double d3 = pdot(ab, bp), d4 = pdot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { cp = b; bary[0] = 0; bary[1] = 1; bary[2] = 0; return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        cp = padd(a, pscale(v, ab)); bary[0] = 1 - v; bary[1] = v; bary[2] = 0; return;
    }
    P3 cpv = psub(p, c);
    double d5 = pdot(ab, cpv), d6 = pdot(ac, cpv);
    if (d6 >= 0.0 && d5 <= d6) { cp = c; bary[0] = 0; bary[1] = 0; bary[2] = 1; return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        cp = padd(a, pscale(w, ac)); bary[0] = 1 - w; bary[1] = 0; bary[2] = w; return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        cp = padd(b, pscale(w, psub(c, b))); bary[0] = 0; bary[1] = 1 - w; bary[2] = w; return;
    }
    double denom = 1.0 / ((va + vb) + vc);
    double v = vb * denom, w = vc * denom;
    cp = padd(padd(a, pscale(v, ab)), pscale(w, ac));
    bary[0] = (1. - v) - w; bary[1] = v; bary[2] = w;
}

__global__ void k_fill_int(int* p, int count, int value) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

// Triangle bounding boxes and the largest box extent (for the cell size). Also clears the
// slot-overflow flag of the query that follows (the largest extent starts from zero: the
// previous detection's k_seed_bits, or the allocation, cleared it).
__global__ void k_tri_box(DevSelfDetect D) {
    __shared__ double red[kB / 64];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *D.overflow = 0;
    double ext = 0.;
    if (t < D.T) {
        P3 a = at_perm(D.q, D.n, D.iperm[D.tris[3 * t]]);
        P3 b = at_perm(D.q, D.n, D.iperm[D.tris[3 * t + 1]]);
        P3 c = at_perm(D.q, D.n, D.iperm[D.tris[3 * t + 2]]);
        double lo[3] = {fmin(fmin(a.x, b.x), c.x), fmin(fmin(a.y, b.y), c.y), fmin(fmin(a.z, b.z), c.z)};
        double hi[3] = {fmax(fmax(a.x, b.x), c.x), fmax(fmax(a.y, b.y), c.y), fmax(fmax(a.z, b.z), c.z)};
        for (int k = 0; k < 3; ++k) {
            D.tbox[6 * t + k] = lo[k];
            D.tbox[6 * t + 3 + k] = hi[k];
            ext = fmax(ext, hi[k] - lo[k]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) ext = fmax(ext, __shfl_xor(ext, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ext;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, red[k]);
        atomicMax(D.ext_bits, (unsigned long long)__double_as_longlong(m));
    }
}

__global__ void k_tri_count(DevSelfDetect D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.T) return;
    double c = cell_size(D);
    unsigned h = cell_hash(cell_of(D.tbox[6 * t], c), cell_of(D.tbox[6 * t + 1], c), cell_of(D.tbox[6 * t + 2], c), D.P);
    atomicAdd(&D.bucket_cnt[h], 1);
}

__global__ void k_tri_scatter(DevSelfDetect D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.T) return;
    double c = cell_size(D);
    unsigned h = cell_hash(cell_of(D.tbox[6 * t], c), cell_of(D.tbox[6 * t + 1], c), cell_of(D.tbox[6 * t + 2], c), D.P);
    // positions from the bucket's end down; the count returns to zero for the next detection
    int slot = atomicSub(&D.bucket_cnt[h], 1) - 1;
    D.bucket_tris[D.bucket_off[h] + slot] = t;
}

// pass 0: count survivors per vertex; pass 1: write them (sorted by triangle).
__global__ void k_vertex_query(DevSelfDetect D, int pass) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= D.n) return;
    const double c = cell_size(D);
    const double E = __longlong_as_double((long long)*D.ext_bits) * (1. + 1e-9);
    P3 x = at_perm(D.q, D.n, D.iperm[v]);
    const double lo[3] = {x.x - D.t11, x.y - D.t11, x.z - D.t11};
    const double hi[3] = {x.x + D.t11, x.y + D.t11, x.z + D.t11};
    long long c0[3], c1[3];
    bool sane = isfinite(x.x) && isfinite(x.y) && isfinite(x.z);
    for (int k = 0; k < 3; ++k) {
        c0[k] = sane ? cell_of(lo[k] - E, c) : 0;
        c1[k] = sane ? cell_of(hi[k], c) : -1;
        // the query range is never longer than one cell, so it spans at most 3 cells
        if (c1[k] - c0[k] > 2) c1[k] = c0[k] - 1;
    }
    int nfull = 0, nfilt = 0;
    int wf = 0, wp = 0;
    const int base_f = pass ? D.off_full[v] : 0;
    const int base_p = pass ? D.off_filt[v] : 0;
    for (long long ix = c0[0]; ix <= c1[0]; ++ix)
        for (long long iy = c0[1]; iy <= c1[1]; ++iy)
            for (long long iz = c0[2]; iz <= c1[2]; ++iz) {
                unsigned h = cell_hash(ix, iy, iz, D.P);
                int b0 = D.bucket_off[h], b1 = D.bucket_off[h + 1];
                for (int q = b0; q < b1; ++q) {
                    int t = D.bucket_tris[q];
                    const double* bx = &D.tbox[6 * t];
                    if (cell_of(bx[0], c) != ix || cell_of(bx[1], c) != iy || cell_of(bx[2], c) != iz) continue;
                    if (!(bx[0] <= hi[0] && lo[0] <= bx[3] && bx[1] <= hi[1] && lo[1] <= bx[4] && bx[2] <= hi[2] &&
                          lo[2] <= bx[5]))
                        continue;
                    int ia = D.tris[3 * t], ib = D.tris[3 * t + 1], ic = D.tris[3 * t + 2];
                    P3 a = at_perm(D.q, D.n, D.iperm[ia]);
                    P3 b = at_perm(D.q, D.n, D.iperm[ib]);
                    P3 cc = at_perm(D.q, D.n, D.iperm[ic]);
                    double bary[3];
                    P3 cp;
                    closest_point(x, a, b, cc, D.tol, bary, cp);
                    P3 d = psub(cp, x);
                    if (!(pdot(d, d) < D.tol_sq)) continue;
                    bool keep = !(ia == v || ib == v || ic == v || bary[0] == 0 || bary[1] == 0 || bary[2] == 0);
                    if (pass == 0) {
                        ++nfull;
                        nfilt += keep;
                        continue;
                    }
                    if (D.write_full) {
                        int k = base_f + wf++;
                        if (k < D.cap_full) {
                            D.full_v[k] = v;
                            D.full_t[k] = t;
                            D.full_bary[3 * k] = bary[0];
                            D.full_bary[3 * k + 1] = bary[1];
                            D.full_bary[3 * k + 2] = bary[2];
                        }
                    }
                    if (keep) {
                        int k = base_p + wp++;
                        if (k < D.cap_filt) {
                            SelfCandidate& o = D.filt[k];
                            o.v = v;
                            o.t = t;
                            o.bary[0] = bary[0];
                            o.bary[1] = bary[1];
                            o.bary[2] = bary[2];
                            // normal = (x - cp) / |x - cp| with cp = a w0 + b w1 + c w2
                            P3 cpb = {(a.x * bary[0] + b.x * bary[1]) + cc.x * bary[2],
                                      (a.y * bary[0] + b.y * bary[1]) + cc.y * bary[2],
                                      (a.z * bary[0] + b.z * bary[1]) + cc.z * bary[2]};
                            P3 nv = psub(x, cpb);
                            double z = pdot(nv, nv);
                            if (z > 0.) {
                                double s = sqrt(z);
                                nv = {nv.x / s, nv.y / s, nv.z / s};
                            }
                            o.normal[0] = nv.x;
                            o.normal[1] = nv.y;
                            o.normal[2] = nv.z;
                        }
                    }
                }
            }
    if (pass == 0) {
        D.cnt_full[v] = nfull;
        D.cnt_filt[v] = nfilt;
        return;
    }
    // insertion sort of this vertex's segments by triangle index
    if (D.write_full && base_f + wf <= D.cap_full) {
        for (int i = base_f + 1; i < base_f + wf; ++i) {
            int t = D.full_t[i];
            double b0 = D.full_bary[3 * i], b1 = D.full_bary[3 * i + 1], b2 = D.full_bary[3 * i + 2];
            int j = i - 1;
            while (j >= base_f && D.full_t[j] > t) {
                D.full_t[j + 1] = D.full_t[j];
                D.full_bary[3 * (j + 1)] = D.full_bary[3 * j];
                D.full_bary[3 * (j + 1) + 1] = D.full_bary[3 * j + 1];
                D.full_bary[3 * (j + 1) + 2] = D.full_bary[3 * j + 2];
                --j;
            }
            D.full_t[j + 1] = t;
            D.full_bary[3 * (j + 1)] = b0;
            D.full_bary[3 * (j + 1) + 1] = b1;
            D.full_bary[3 * (j + 1) + 2] = b2;
        }
    }
    if (base_p + wp <= D.cap_filt) {
        for (int i = base_p + 1; i < base_p + wp; ++i) {
            SelfCandidate key = D.filt[i];
            int j = i - 1;
            while (j >= base_p && D.filt[j].t > key.t) {
                D.filt[j + 1] = D.filt[j];
                --j;
            }
            D.filt[j + 1] = key;
        }
    }
}

// Bucket-sorted copy of the triangle data (box, vertex positions, vertex ids).
// Single-pass path: triangle index, vertices and positions into the bucket of the cell of the
// triangle box's lower corner (bucket slots sorted by bucket).
__device__ __forceinline__ long long cell_key(long long ix, long long iy, long long iz) {
    const long long m = (1ll << 21) - 1;
    return ((ix & m) << 42) | ((iy & m) << 21) | (iz & m);
}

__global__ void k_tri_scatter_packed(DevSelfDetect D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.T) return;
    double c = cell_size(D);
    const long long ix = cell_of(D.tbox[6 * t], c), iy = cell_of(D.tbox[6 * t + 1], c), iz = cell_of(D.tbox[6 * t + 2], c);
    unsigned h = cell_hash(ix, iy, iz, D.P);
    // positions from the bucket's end down; the count returns to zero for the next detection
    int q = D.bucket_off[h] + atomicSub(&D.bucket_cnt[h], 1) - 1;
    BoxKey e;
    e.key = cell_key(ix, iy, iz);
    for (int k = 0; k < 3; ++k) {
        e.lo[k] = __double2float_rd(D.tbox[6 * t + k]);
        e.hi[k] = __double2float_ru(D.tbox[6 * t + 3 + k]);
    }
    D.bk_key[q] = e;
    D.bk_tri[q] = t;
    for (int k = 0; k < 3; ++k) {
        int v = D.tris[3 * t + k];
        D.bk_vid[3 * q + k] = v;
        P3 p = at_perm(D.q, D.n, D.iperm[v]);
        D.bk_pos[9 * q + 3 * k] = p.x;
        D.bk_pos[9 * q + 3 * k + 1] = p.y;
        D.bk_pos[9 * q + 3 * k + 2] = p.z;
    }
}

// One pass over the grid: filtered candidates into the vertex's slots. The vertex's query box
// x - 1.1 tol - E .. x + 1.1 tol (E the largest triangle extent) is shorter than a cell, so it
// spans at most 2 x 2 x 2 cells; their buckets are taken as one list that kFlatLanes lanes split
// (lane j takes entries j, j + kFlatLanes, ...), so each lane's chain of dependent loads is a few
// entries long whatever the buckets' sizes. A bucket entry counts for a cell only if its triangle's
// lower corner lies in that cell (buckets shared by several cells through the hash), and only if
// its float box, rounded outwards, meets the query box; then its positions are read and the
// exact tests run on the double box (the min / max of the positions, the values k_tri_box
// computes). Slot order is arbitrary, the compaction sorts by triangle.
constexpr int kFlatLanes = 16;
constexpr int kQueryCells = 8;
__global__ void k_vertex_query_flat(DevSelfDetect D) {
    __shared__ int slot_count[kB / kFlatLanes];
    const int local_v = threadIdx.x / kFlatLanes, sub = threadIdx.x % kFlatLanes;
    const int v = xcd_block(blockIdx.x, gridDim.x) * (kB / kFlatLanes) + local_v;
    if (sub == 0) slot_count[local_v] = 0;
    __syncthreads();
    if (v < D.n) {
        const double c = cell_size(D);
        const double E = __longlong_as_double((long long)*D.ext_bits) * (1. + 1e-9);
        const P3 x = at_perm(D.q, D.n, D.iperm[v]);
        const double lo[3] = {x.x - D.t11, x.y - D.t11, x.z - D.t11};
        const double hi[3] = {x.x + D.t11, x.y + D.t11, x.z + D.t11};
        long long c0[3];
        int span[3];
        const bool sane = isfinite(x.x) && isfinite(x.y) && isfinite(x.z);
        for (int k = 0; k < 3; ++k) {
            c0[k] = sane ? cell_of(lo[k] - E, c) : 0;
            const long long sp = sane ? cell_of(hi[k], c) - c0[k] + 1 : 0;
            span[k] = (sp < 0 || sp > 3) ? 0 : (int)sp;
        }
        int ncell = span[0] * span[1] * span[2];
        if (ncell > kQueryCells) {  // not for a finite vertex (the box is shorter than a cell)
            atomicOr(D.overflow, 1);
            ncell = 0;
        }
        int bstart[kQueryCells], blen[kQueryCells];
        int total = 0;
#pragma unroll
        for (int ci = 0; ci < kQueryCells; ++ci) {
            bstart[ci] = 0;
            blen[ci] = 0;
            if (ci < ncell) {
                const long long ix = c0[0] + ci % span[0], iy = c0[1] + (ci / span[0]) % span[1],
                                iz = c0[2] + ci / (span[0] * span[1]);
                const unsigned h = cell_hash(ix, iy, iz, D.P);
                bstart[ci] = D.bucket_off[h];
                blen[ci] = D.bucket_off[h + 1] - bstart[ci];
            }
            total += blen[ci];
        }
        SelfCandidate* out = D.vslots + (size_t)v * kSlotsPerVertex;
        for (int k = sub; k < total; k += kFlatLanes) {
            int q = -1, cell = 0, base = 0;
#pragma unroll
            for (int ci = 0; ci < kQueryCells; ++ci) {
                if (q < 0 && k < base + blen[ci]) {
                    q = bstart[ci] + (k - base);
                    cell = ci;
                }
                base += blen[ci];
            }
            const long long ix = c0[0] + cell % span[0], iy = c0[1] + (cell / span[0]) % span[1],
                            iz = c0[2] + cell / (span[0] * span[1]);
            // prefilter on 32 bytes: the entry's cell, then its box rounded outwards
            const BoxKey e = D.bk_key[q];
            if (e.key != cell_key(ix, iy, iz)) continue;
            if (!((double)e.lo[0] <= hi[0] && lo[0] <= (double)e.hi[0] && (double)e.lo[1] <= hi[1] &&
                  lo[1] <= (double)e.hi[1] && (double)e.lo[2] <= hi[2] && lo[2] <= (double)e.hi[2]))
                continue;
            const double* pp = &D.bk_pos[9 * (size_t)q];
            const P3 a = {pp[0], pp[1], pp[2]}, b = {pp[3], pp[4], pp[5]}, cc = {pp[6], pp[7], pp[8]};
            const double bx[6] = {fmin(fmin(a.x, b.x), cc.x), fmin(fmin(a.y, b.y), cc.y), fmin(fmin(a.z, b.z), cc.z),
                                  fmax(fmax(a.x, b.x), cc.x), fmax(fmax(a.y, b.y), cc.y), fmax(fmax(a.z, b.z), cc.z)};
            if (!(bx[0] <= hi[0] && lo[0] <= bx[3] && bx[1] <= hi[1] && lo[1] <= bx[4] && bx[2] <= hi[2] &&
                  lo[2] <= bx[5]))
                continue;
            const int* vid = &D.bk_vid[3 * (size_t)q];
            if (vid[0] == v || vid[1] == v || vid[2] == v) continue;  // never kept by the pairing
            double bary[3];
            P3 cp;
            closest_point(x, a, b, cc, D.tol, bary, cp);
            P3 d = psub(cp, x);
            if (!(pdot(d, d) < D.tol_sq)) continue;
            if (bary[0] == 0 || bary[1] == 0 || bary[2] == 0) continue;
            int slot = atomicAdd(&slot_count[local_v], 1);
            if (slot >= kSlotsPerVertex) {
                atomicOr(D.overflow, 1);
                continue;
            }
            SelfCandidate& o = out[slot];
            o.v = v;
            o.t = D.bk_tri[q];
            o.bary[0] = bary[0];
            o.bary[1] = bary[1];
            o.bary[2] = bary[2];
            P3 cpb = {(a.x * bary[0] + b.x * bary[1]) + cc.x * bary[2],
                      (a.y * bary[0] + b.y * bary[1]) + cc.y * bary[2],
                      (a.z * bary[0] + b.z * bary[1]) + cc.z * bary[2]};
            P3 nv = psub(x, cpb);
            double z = pdot(nv, nv);
            if (z > 0.) {
                double sq = sqrt(z);
                nv = {nv.x / sq, nv.y / sq, nv.z / sq};
            }
            o.normal[0] = nv.x;
            o.normal[1] = nv.y;
            o.normal[2] = nv.z;
        }
    }
    __syncthreads();
    if (v < D.n && sub == 0) {
        int cnt = slot_count[local_v];
        D.cnt_filt[v] = cnt < kSlotsPerVertex ? cnt : kSlotsPerVertex;
    }
}

// Compaction of the slots, sorted by triangle index.
__global__ void k_compact_slots(DevSelfDetect D) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= D.n) return;
    const int cnt = D.cnt_filt[v];
    const int base = D.off_filt[v];
    if (cnt == 0 || base + cnt > D.cap_filt) return;
    const SelfCandidate* in = D.vslots + (size_t)v * kSlotsPerVertex;
    int order[kSlotsPerVertex];
    for (int i = 0; i < cnt; ++i) order[i] = i;
    for (int i = 1; i < cnt; ++i) {
        int k = order[i], j = i - 1;
        while (j >= 0 && in[order[j]].t > in[k].t) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = k;
    }
    for (int i = 0; i < cnt; ++i) D.filt[base + i] = in[order[i]];
}

// Bitset (permuted numbering) of vertices with at least one obstacle contact. The last kernel
// of a detection: it also clears the largest box extent for the next one.
__global__ void k_seed_bits(const int* ccnt, int n, unsigned long long* bits, unsigned long long* ext_bits) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *ext_bits = 0ull;
    bool on = i < n && ccnt[i] > 0;
    unsigned long long m = __ballot(on);
    if ((threadIdx.x & 63) == 0 && i < n) bits[i >> 6] = m;
}

// Exclusive scan of int counts, dst = scan(src) (dst may be src); *total (and *tail when given)
// = the sum. Block-wise scans with per-block totals, then each block adds the sum of the totals
// of the blocks before it, which it reduces itself (the last block writes the sum).
__global__ void k_scan_local(int* dst, const int* src, int count, int* sums) {
    __shared__ int buf[kB];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int v = i < count ? src[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kB; off <<= 1) {
        int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < count) dst[i] = buf[threadIdx.x] - v;
    if (threadIdx.x == kB - 1) sums[blockIdx.x] = buf[kB - 1];
}
__global__ void k_scan_add(int* a, int count, const int* sums, int* total, int* tail) {
    __shared__ int red[kB / 64];
    const int b = blockIdx.x;
    int s = 0;
    for (int k = threadIdx.x; k < b; k += kB) s += sums[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    s = 0;
#pragma unroll
    for (int k = 0; k < kB / 64; ++k) s += red[k];
    const int i = b * kB + threadIdx.x;
    if (i < count) a[i] += s;
    if (b == (int)gridDim.x - 1 && threadIdx.x == 0) {
        const int sum = s + sums[b];
        *total = sum;
        if (tail) *tail = sum;
    }
}

// ---------------------------------------------------------------- friction
__device__ __forceinline__ P3 ld_slot(const double* A, int s) { return {A[3 * s], A[3 * s + 1], A[3 * s + 2]}; }
__device__ __forceinline__ void st_slot(double* A, int s, P3 v) {
    A[3 * s] = v.x;
    A[3 * s + 1] = v.y;
    A[3 * s + 2] = v.z;
}

__device__ double ac_norm3(P3 u, const double* r, double mu) {
    double an = fmax(0., r[0] - u.x) - r[0];
    double d1 = r[1] - u.y, d2 = r[2] - u.z;
    double dn = sqrt(d1 * d1 + d2 * d2);
    double s = fmin(1., fmax(0., mu * r[0] / dn));
    double a1 = d1 * s - r[1], a2 = d2 * s - r[2];
    return sqrt((an * an + a1 * a1) + a2 * a2);
}

// Per wave (no block barrier), as the obstacle contacts' accumulation.
__device__ void ac_block(double s, double m, double* sum_slot, unsigned long long* max_slot) {
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        m = fmax(m, __shfl_xor(m, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (s != 0.) atomicAdd(sum_slot, s);
        if (m > 0.) atomicMax(max_slot, (unsigned long long)__double_as_longlong(m));
    }
}

// One self contact's new reaction from the state at its level's start (the reference's first
// parallel loop, SimulationFrictionEstimation.cpp:431-520); accumulates its friction residual.
__device__ __forceinline__ void self_contact_update(const DevSelf& S, int s, const double* __restrict__ v, int n,
                                                    const double* __restrict__ mass, const StepArgs& a, int iteration,
                                                    double& acs, double& acm) {
    const double* B = &S.basis[9 * s];
    P3 bn = {B[0], B[1], B[2]}, bt = {B[3], B[4], B[5]}, bb = {B[6], B[7], B[8]};
    double* r = &S.r[3 * s];
    const int pa = S.a[s], pb = S.b[s];
    if (a.record && iteration > 0) {
        P3 u = psub(at_perm(v, n, pa), at_perm(v, n, pb));
        const double an = ac_norm3(P3{pdot(bn, u), pdot(bt, u), pdot(bb, u)}, r, S.mu[s]);
        acs += an;
        acm = fmax(acm, an);
    }
    const int sa = S.sa[s], sb = S.sb[s];
    P3 rem = ld_slot(S.rem, s);
    P3 fA = psub(padd(padd(padd(ld_slot(S.Fc, sa), ld_slot(S.CFn, sa)), ld_slot(S.SCFc, sa)), ld_slot(S.SCFn, sa)), rem);
    P3 fB = padd(padd(padd(padd(ld_slot(S.Fc, sb), ld_slot(S.CFn, sb)), ld_slot(S.SCFc, sb)), ld_slot(S.SCFn, sb)), rem);
    const double mA = mass[pa], mB = mass[pb];
    P3 f = psub(pscale(1. / mA, fA), pscale(1. / mB, fB));
    P3 fl = {pdot(bn, f), pdot(bt, f), pdot(bb, f)};
    P3 add = {0., 0., 0.};
    if (fl.x > 0.) {
        r[0] = 0.; r[1] = 0.; r[2] = 0.;
        S.active[s] = 0;
    } else {
        double mu = S.mu[s];
        double r0 = -fl.x, r1 = -fl.y, r2 = -fl.z;
        double rt = sqrt(r1 * r1 + r2 * r2);
        if (rt > mu * r0) {
            r1 *= mu * r0 / rt;
            r2 *= mu * r0 / rt;
        }
        double w = (mA * mB) / (mA + mB);
        r0 *= w; r1 *= w; r2 *= w;
        r[0] = r0; r[1] = r1; r[2] = r2;
        S.active[s] = 1;
        add = {(bn.x * r0 + bt.x * r1) + bb.x * r2, (bn.y * r0 + bt.y * r1) + bb.y * r2,
               (bn.z * r0 + bt.z * r1) + bb.z * r2};
    }
    st_slot(S.addv, s, add);
}

// Apply group g of a level (the second parallel loop, :522-560): per touched vertex, in the
// level's contact order, SCF_cur -= sign * previous force, SCF_next += sign * new force.
__device__ __forceinline__ void self_group_apply(const DevSelf& S, int g) {
    const int slot = S.grp_slot[g];
    P3 cur = ld_slot(S.SCFc, slot), nxt = ld_slot(S.SCFn, slot);
    for (int e = S.ent_ptr[g]; e < S.ent_ptr[g + 1]; ++e) {
        int s = S.ent_contact[e];
        P3 prev = ld_slot(S.rem, s), add = ld_slot(S.addv, s);
        if (S.ent_sign[e] > 0) {
            cur = psub(cur, prev);
            nxt = padd(nxt, add);
        } else {
            cur = padd(cur, prev);
            nxt = psub(nxt, add);
        }
    }
    st_slot(S.SCFc, slot, cur);
    st_slot(S.SCFn, slot, nxt);
}

// b += CF_next + SCF_next at an involved vertex (:566-568); clear the buffers that become
// "next" after the swap.
__device__ __forceinline__ void self_slot_finalize(const DevSelf& S, int s, double* __restrict__ rhs, int n) {
    int v = S.slot_vertex[s];
    P3 c = ld_slot(S.CFn, s), sc = ld_slot(S.SCFn, s);
    rhs[v] += c.x + sc.x;
    rhs[n + v] += c.y + sc.y;
    rhs[2 * n + v] += c.z + sc.z;
    st_slot(S.CFc, s, P3{0., 0., 0.});
    st_slot(S.SCFc, s, P3{0., 0., 0.});
}

// One level: new reaction of every self contact of the level from the state at the
// level's start (computed-then-applied, as in the reference's two parallel loops).
__global__ void k_self_level(DevSelf S, int l, const double* __restrict__ v, int n, const double* __restrict__ mass,
                             StepArgs a, int iteration) {
    int s = S.lvl_ptr[l] + blockIdx.x * blockDim.x + threadIdx.x;
    double acs = 0., acm = 0.;
    if (s < S.lvl_ptr[l + 1]) self_contact_update(S, s, v, n, mass, a, iteration, acs, acm);
    if (a.record && iteration > 0) ac_block(acs, acm, &a.ac_sum[iteration - 1], &a.ac_max[iteration - 1]);
}

__global__ void k_self_apply(DevSelf S, int l) {
    int g = S.grp_ptr[l] + blockIdx.x * blockDim.x + threadIdx.x;
    if (g < S.grp_ptr[l + 1]) self_group_apply(S, g);
}

__global__ void k_self_finalize(DevSelf S, double* __restrict__ rhs, int n) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S.nslots) self_slot_finalize(S, s, rhs, n);
}

// The new reaction of a self contact from its inputs (the arithmetic of self_contact_update):
// f = B^T (fA / mA - fB / mB), take-off if f_N > 0, else the cone projection of -f scaled by
// mA mB / (mA + mB). Writes r and active; returns the global force to apply.
__device__ __forceinline__ P3 self_reaction(const DevSelf& S, int s, P3 bn, P3 bt, P3 bb, P3 fA, P3 fB, double mA,
                                            double mB, double mu) {
    double* r = &S.r[3 * s];
    P3 f = psub(pscale(1. / mA, fA), pscale(1. / mB, fB));
    P3 fl = {pdot(bn, f), pdot(bt, f), pdot(bb, f)};
    if (fl.x > 0.) {
        r[0] = 0.; r[1] = 0.; r[2] = 0.;
        S.active[s] = 0;
        return P3{0., 0., 0.};
    }
    double r0 = -fl.x, r1 = -fl.y, r2 = -fl.z;
    double rt = sqrt(r1 * r1 + r2 * r2);
    if (rt > mu * r0) {
        r1 *= mu * r0 / rt;
        r2 *= mu * r0 / rt;
    }
    double w = (mA * mB) / (mA + mB);
    r0 *= w; r1 *= w; r2 *= w;
    r[0] = r0; r[1] = r1; r[2] = r2;
    S.active[s] = 1;
    return P3{(bn.x * r0 + bt.x * r1) + bb.x * r2, (bn.y * r0 + bt.y * r1) + bb.y * r2,
              (bn.z * r0 + bt.z * r1) + bb.z * r2};
}

// All levels of one iteration in one workgroup with the level-to-level state in LDS (at most
// kSelfLdsMaxContacts self contacts, one per thread, and kSelfLdsMaxSlots involved vertices):
// the self-contact forces of the involved vertices (current and next buffers) and each
// contact's previous and new force live in LDS; each thread loads its contact's read-only data
// (basis, vertices, masses, F and the obstacle reaction at both vertices, residual inputs) once.
// The per-vertex groups of every level (their slots and entries, 16-bit) are copied to LDS
// with those loads, and each thread's involved vertices for the right-hand side are read there
// too, so no global load waits inside the level loop or before the final update.
// A level is then its updates from LDS, a barrier, the per-vertex applies in LDS, a barrier.
// Same operations in the same order as self_contact_update / self_group_apply /
// self_slot_finalize, so the results are those of the other paths bit for bit.
__global__ void __launch_bounds__(kSelfFusedThreads) k_self_fused_lds(DevSelf S, int nlev,
                                                                     const double* __restrict__ v, int n,
                                                                     const double* __restrict__ mass, StepArgs a,
                                                                     int iteration, double* __restrict__ rhs) {
    __shared__ double sh[6 * (kSelfLdsMaxSlots + kSelfLdsMaxContacts)];
    // groups of all levels: at most one per (contact, side), so 2 kSelfLdsMaxContacts
    __shared__ unsigned short g_slot[2 * kSelfLdsMaxContacts];
    __shared__ unsigned short g_ent[2 * kSelfLdsMaxContacts + 1];
    __shared__ unsigned short e_code[2 * kSelfLdsMaxContacts];  // contact * 2 + (sign > 0)
    // each level's first contact and first group: read from LDS in the level loop (a global load
    // there puts a memory round trip into every level)
    __shared__ unsigned short l_ptr[kSelfLdsMaxLevels + 1];
    __shared__ unsigned short l_grp[kSelfLdsMaxLevels + 1];
    double* scf_c = sh;                                  // 3 per slot
    double* scf_n = sh + 3 * kSelfLdsMaxSlots;           // 3 per slot
    double* c_rem = sh + 6 * kSelfLdsMaxSlots;           // 3 per contact
    double* c_add = c_rem + 3 * kSelfLdsMaxContacts;     // 3 per contact
    const int t = threadIdx.x;
    // Every load of the LDS-resident data is issued before the first LDS store (a loop that
    // stores each value as it arrives waits a memory round trip per element it handles):
    // at most 6 force components per buffer, 2 groups and 2 entries per thread.
    static_assert(3 * kSelfLdsMaxSlots <= 6 * kSelfFusedThreads && 2 * kSelfLdsMaxContacts <= 2 * kSelfFusedThreads,
                  "LDS-resident data per thread");
    // The "next" self-contact forces start at zero: k_self_prepare clears both buffers, and each
    // launch of this kernel clears the one (SCFc) that the swap after it makes the next one. The
    // group count comes with the arguments, so no load waits for it.
    const int n3s = 3 * S.nslots, G = S.ngroups, E = 2 * S.nsc;
    double fc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int k = t + q * kSelfFusedThreads;
        fc[q] = k < n3s ? S.SCFc[k] : 0.;
    }
    int gs[2], ge[2], ec[2], es[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int k = t + q * kSelfFusedThreads;
        gs[q] = k < G ? S.grp_slot[k] : 0;
        ge[q] = k <= G ? S.ent_ptr[k] : 0;
        ec[q] = k < E ? S.ent_contact[k] : 0;
        es[q] = k < E ? S.ent_sign[k] : 0;
    }
    static_assert(kSelfLdsMaxSlots <= 2 * kSelfFusedThreads, "two involved vertices per thread");
    int vv[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int sl = t + q * kSelfFusedThreads;
        vv[q] = sl < S.nslots ? S.slot_vertex[sl] : 0;
    }
    static_assert(kSelfLdsMaxLevels < kSelfFusedThreads, "one level offset per thread");
    int lp = 0, gp = 0;
    if (t <= nlev) {
        lp = S.lvl_ptr[t];
        gp = S.grp_ptr[t];
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int k = t + q * kSelfFusedThreads;
        if (k < n3s) {
            scf_c[k] = fc[q];
            scf_n[k] = 0.;
        }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int k = t + q * kSelfFusedThreads;
        if (k < G) g_slot[k] = (unsigned short)gs[q];
        if (k <= G) g_ent[k] = (unsigned short)ge[q];
        if (k < E) e_code[k] = (unsigned short)(2 * ec[q] + (es[q] > 0 ? 1 : 0));
    }
    if (t <= nlev) {
        l_ptr[t] = (unsigned short)lp;
        l_grp[t] = (unsigned short)gp;
    }
    // Every per-contact and per-slot input is indexed by the thread (one round of loads): the
    // masses of the contact's vertices, F + CF at both (stored by contact by the local kernel,
    // the value fA / fB start from), and the first slot's right-hand side and obstacle reaction
    // for the final update (which then waits for no load; they do not change during the levels).
    const int s = t;
    const bool mine = s < S.nsc;
    int sa = 0, sb = 0, pa = 0, pb = 0;
    double mA = 1., mB = 1., mu = 0.;
    P3 bn{}, bt{}, bb{}, fa{}, fb{}, rem{}, rprev{};
    if (mine) {
        const double* B = &S.basis[9 * s];
        bn = P3{B[0], B[1], B[2]};
        bt = P3{B[3], B[4], B[5]};
        bb = P3{B[6], B[7], B[8]};
        pa = S.a[s];
        pb = S.b[s];
        sa = S.sa[s];
        sb = S.sb[s];
        mu = S.mu[s];
        mA = S.mass2[2 * s];
        mB = S.mass2[2 * s + 1];
        fa = ld_slot(S.x2, 2 * s);
        fb = ld_slot(S.x2, 2 * s + 1);
        rem = ld_slot(S.rem, s);
        rprev = ld_slot(S.r, s);
        c_rem[3 * s] = rem.x;
        c_rem[3 * s + 1] = rem.y;
        c_rem[3 * s + 2] = rem.z;
    }
    P3 rhs0{}, cf0{};
    if (t < S.nslots) {
        rhs0 = ld_slot(S.b0, t);
        cf0 = ld_slot(S.CFn, t);
    }
    // the residual's relative velocity (a dependent load) is used after the levels
    P3 u{};
    const bool res = mine && a.record && iteration > 0;
    if (res) u = psub(at_perm(v, n, pa), at_perm(v, n, pb));
    __syncthreads();
    for (int l = 0; l < nlev; ++l) {
        if (mine && s >= l_ptr[l] && s < l_ptr[l + 1]) {
            const P3 ca = {scf_c[3 * sa], scf_c[3 * sa + 1], scf_c[3 * sa + 2]};
            const P3 na = {scf_n[3 * sa], scf_n[3 * sa + 1], scf_n[3 * sa + 2]};
            const P3 cb = {scf_c[3 * sb], scf_c[3 * sb + 1], scf_c[3 * sb + 2]};
            const P3 nb = {scf_n[3 * sb], scf_n[3 * sb + 1], scf_n[3 * sb + 2]};
            const P3 fA = psub(padd(padd(fa, ca), na), rem);
            const P3 fB = padd(padd(padd(fb, cb), nb), rem);
            const P3 add = self_reaction(S, s, bn, bt, bb, fA, fB, mA, mB, mu);
            c_add[3 * s] = add.x;
            c_add[3 * s + 1] = add.y;
            c_add[3 * s + 2] = add.z;
        }
        __syncthreads();
        const int g1 = l_grp[l + 1];
        for (int g = l_grp[l] + t; g < g1; g += kSelfFusedThreads) {
            const int slot = g_slot[g];
            P3 cur = {scf_c[3 * slot], scf_c[3 * slot + 1], scf_c[3 * slot + 2]};
            P3 nxt = {scf_n[3 * slot], scf_n[3 * slot + 1], scf_n[3 * slot + 2]};
            for (int e = g_ent[g]; e < g_ent[g + 1]; ++e) {
                const int code = e_code[e];
                const int c = code >> 1;
                const P3 prev = {c_rem[3 * c], c_rem[3 * c + 1], c_rem[3 * c + 2]};
                const P3 add = {c_add[3 * c], c_add[3 * c + 1], c_add[3 * c + 2]};
                if (code & 1) {
                    cur = psub(cur, prev);
                    nxt = padd(nxt, add);
                } else {
                    cur = padd(cur, prev);
                    nxt = psub(nxt, add);
                }
            }
            scf_c[3 * slot] = cur.x; scf_c[3 * slot + 1] = cur.y; scf_c[3 * slot + 2] = cur.z;
            scf_n[3 * slot] = nxt.x; scf_n[3 * slot + 1] = nxt.y; scf_n[3 * slot + 2] = nxt.z;
        }
        __syncthreads();
    }
    if (mine) st_slot(S.addv, s, P3{c_add[3 * s], c_add[3 * s + 1], c_add[3 * s + 2]});
    if (a.record && iteration > 0) {
        // the friction residual of the previous iteration (r^{k-1}, v^k), as self_contact_update
        double acs = 0.;
        if (res) {
            const double rp[3] = {rprev.x, rprev.y, rprev.z};
            acs = ac_norm3(P3{pdot(bn, u), pdot(bt, u), pdot(bb, u)}, rp, mu);
        }
        ac_block(acs, acs, &a.ac_sum[iteration - 1], &a.ac_max[iteration - 1]);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int sl = t + q * kSelfFusedThreads;
        if (sl >= S.nslots) break;
        const int w = vv[q];
        const P3 c = q == 0 ? cf0 : ld_slot(S.CFn, sl);
        const P3 b0 = q == 0 ? rhs0 : ld_slot(S.b0, sl);
        const P3 sc = {scf_n[3 * sl], scf_n[3 * sl + 1], scf_n[3 * sl + 2]};
        rhs[w] = b0.x + (c.x + sc.x);
        rhs[n + w] = b0.y + (c.y + sc.y);
        rhs[2 * n + w] = b0.z + (c.z + sc.z);
        st_slot(S.SCFn, sl, sc);
        st_slot(S.CFc, sl, P3{0., 0., 0.});
        st_slot(S.SCFc, sl, P3{0., 0., 0.});
    }
}

// Levels l0 .. nlev - 1 of one iteration in one workgroup (few self contacts, or the tail of
// small levels after per-level launches of the large ones): each level's updates, a barrier,
// its applies, a barrier; then, with finish, the right-hand side (without, k_self_finalize
// follows over all CUs: past a few thousand involved vertices one workgroup's rounds cost more).
// The barriers order the workgroup's global-memory writes and reads as the kernel boundaries of
// the per-level launches do; the arithmetic per contact, group and vertex is theirs.
__global__ void __launch_bounds__(kSelfFusedThreads) k_self_fused(DevSelf S, int l0, int nlev,
                                                                 const double* __restrict__ v, int n,
                                                                 const double* __restrict__ mass, StepArgs a,
                                                                 int iteration, double* __restrict__ rhs,
                                                                 int finish) {
    double acs = 0., acm = 0.;
    for (int l = l0; l < nlev; ++l) {
        const int s1 = S.lvl_ptr[l + 1];
        for (int s = S.lvl_ptr[l] + (int)threadIdx.x; s < s1; s += kSelfFusedThreads)
            self_contact_update(S, s, v, n, mass, a, iteration, acs, acm);
        __syncthreads();
        const int g1 = S.grp_ptr[l + 1];
        for (int g = S.grp_ptr[l] + (int)threadIdx.x; g < g1; g += kSelfFusedThreads) self_group_apply(S, g);
        __syncthreads();
    }
    if (finish)
        for (int s = threadIdx.x; s < S.nslots; s += kSelfFusedThreads) self_slot_finalize(S, s, rhs, n);
    if (a.record && iteration > 0) ac_block(acs, acm, &a.ac_sum[iteration - 1], &a.ac_max[iteration - 1]);
}

// End of step: the friction residual of the last iteration's self contacts, and the slot map off
// (one launch; neither part reads what the other writes).
__global__ void k_self_end(DevSelf S, int* vslot, const double* __restrict__ v, int n, StepArgs a, int iteration) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    double acs = 0.;
    if (s < S.nsc) {
        const double* B = &S.basis[9 * s];
        P3 bn = {B[0], B[1], B[2]}, bt = {B[3], B[4], B[5]}, bb = {B[6], B[7], B[8]};
        P3 u = psub(at_perm(v, n, S.a[s]), at_perm(v, n, S.b[s]));
        acs = ac_norm3(P3{pdot(bn, u), pdot(bt, u), pdot(bb, u)}, &S.r[3 * s], S.mu[s]);
    }
    ac_block(acs, acs, &a.ac_sum[iteration], &a.ac_max[iteration]);
    if (s < S.nslots) vslot[S.slot_vertex[s]] = -1;
}

// Start of step: slot map on, buffers zeroed (the end of step is k_self_end).
// computeBasis (SimulationFrictionEstimation.cpp:89-105): n, t = n x e_x (n x e_y when that is
// shorter than 1e-15), b = n x t, t and b normalised; the operations of the host's and the
// oracle's, in their order (IEEE square root and division).
__device__ void self_contact_basis(P3 n, double* B) {
    auto cross = [](P3 a, P3 b) { return P3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
    auto nrm = [](P3 a) {
        const double z = (a.x * a.x + a.y * a.y) + a.z * a.z;
        if (z > 0.) {
            const double q = sqrt(z);
            return P3{a.x / q, a.y / q, a.z / q};
        }
        return a;
    };
    P3 t = cross(n, P3{1., 0., 0.});
    if (sqrt((t.x * t.x + t.y * t.y) + t.z * t.z) < 1.e-15) t = cross(n, P3{0., 1., 0.});
    t = nrm(t);
    const P3 b = nrm(cross(n, t));
    B[0] = n.x; B[1] = n.y; B[2] = n.z;
    B[3] = t.x; B[4] = t.y; B[5] = t.z;
    B[6] = b.x; B[7] = b.y; B[8] = b.z;
}

__global__ void k_self_prepare(DevSelf S, int* vslot) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S.nslots) {
        vslot[S.slot_vertex[s]] = s;
        st_slot(S.CFc, s, P3{0., 0., 0.});
        st_slot(S.CFn, s, P3{0., 0., 0.});
        st_slot(S.SCFc, s, P3{0., 0., 0.});
        st_slot(S.SCFn, s, P3{0., 0., 0.});
        st_slot(S.Fc, s, P3{0., 0., 0.});
    }
    if (s < S.nsc) {
        self_contact_basis(ld_slot(S.normal, s), &S.basis[9 * s]);
        st_slot(S.rem, s, P3{0., 0., 0.});
        st_slot(S.addv, s, P3{0., 0., 0.});
        S.r[3 * s] = 0.; S.r[3 * s + 1] = 0.; S.r[3 * s + 2] = 0.;
        S.active[s] = 0;
    }
}

inline unsigned nblk(long long c) { return (unsigned)((c + kB - 1) / kB); }

// Closest triangle of a mesh obstacle for every cloth vertex (checkMeshAllCollision +
// checkMeshCollision, geometry_util.cpp:32-92,229-262): the triangles registered in the
// vertex's grid cell pass the closed box test against x +- 1.1 tol and the exact test
// |cp - x|^2 < tol^2 (early exit against tol); the closest of their points
// getTriangle(t) * bary wins, the lower triangle index on ties.
__device__ __forceinline__ double mesh_cell(const unsigned long long* ext_bits) {
    const double h = __longlong_as_double((long long)*ext_bits) * (1. + 1e-9);
    return h > 1e-12 ? h : 1e-12;
}

__global__ void k_mesh_lerp(double* __restrict__ cur, const double* __restrict__ prev,
                            const double* __restrict__ next, double a, int count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) cur[k] = prev[k] * (1. - a) + next[k] * a;
}

__global__ void k_mesh_clear(DevMeshBuild B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= B.P) B.off[i] = 0;
    if (i < B.P) B.fill[i] = 0;
    if (i == 0) *B.ext_bits = 0ull;
}

// Grown triangle boxes and their largest extent.
__global__ void k_mesh_box(DevMeshBuild B) {
    __shared__ double red[kB / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double ext = 0.;
    if (t < B.T) {
        const int* f = &B.tris[3 * t];
        for (int k = 0; k < 3; ++k) {
            const double a = B.cur[3 * f[0] + k], b = B.cur[3 * f[1] + k], c = B.cur[3 * f[2] + k];
            const double lo = fmin(fmin(a, b), c) - B.grow, hi = fmax(fmax(a, b), c) + B.grow;
            B.tbox[6 * t + k] = lo;
            B.tbox[6 * t + 3 + k] = hi;
            ext = fmax(ext, hi - lo);
        }
    }
    for (int o = 32; o > 0; o >>= 1) ext = fmax(ext, __shfl_xor(ext, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ext;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, red[k]);
        atomicMax(B.ext_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// The distinct buckets of the cells a grown box overlaps (at most 8).
__device__ __forceinline__ int mesh_buckets(const DevMeshBuild& B, int t, double h, unsigned* bk) {
    long long c0[3], c1[3];
    for (int k = 0; k < 3; ++k) {
        c0[k] = cell_of(B.tbox[6 * t + k], h);
        c1[k] = cell_of(B.tbox[6 * t + 3 + k], h);
        if (c1[k] > c0[k] + 1) c1[k] = c0[k] + 1;  // h exceeds every grown extent
    }
    int nb = 0;
    for (long long a = c0[0]; a <= c1[0]; ++a)
        for (long long b = c0[1]; b <= c1[1]; ++b)
            for (long long c = c0[2]; c <= c1[2]; ++c) {
                const unsigned x = cell_hash(a, b, c, B.P);
                bool seen = false;
                for (int j = 0; j < nb; ++j) seen |= bk[j] == x;
                if (!seen) bk[nb++] = x;
            }
    return nb;
}

__global__ void k_mesh_count(DevMeshBuild B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B.T) return;
    unsigned bk[8];
    const int nb = mesh_buckets(B, t, mesh_cell(B.ext_bits), bk);
    for (int j = 0; j < nb; ++j) atomicAdd(&B.off[bk[j]], 1);
}

// Triangle index and vertex positions into every bucket slot of the triangle. Bucket
// order is arbitrary: the query keeps the nearest triangle, ties to the lowest index.
__global__ void k_mesh_scatter(DevMeshBuild B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B.T) return;
    unsigned bk[8];
    const int nb = mesh_buckets(B, t, mesh_cell(B.ext_bits), bk);
    const int* f = &B.tris[3 * t];
    double x[9];
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) x[3 * v + k] = B.cur[3 * f[v] + k];
    for (int j = 0; j < nb; ++j) {
        const int s = B.off[bk[j]] + atomicAdd(&B.fill[bk[j]], 1);
        B.tris_sorted[s] = t;
        for (int k = 0; k < 9; ++k) B.pos_sorted[9 * (size_t)s + k] = x[k];
    }
}

// kMeshLanes lanes per cloth vertex share its bucket (lane j takes slots j, j + kMeshLanes,
// ...); the nearest hit, ties to the lowest triangle index, is reduced across the lanes.
constexpr int kMeshLanes = 16;
__global__ void k_mesh_query(DevMeshObs Mo, const double* __restrict__ qn, int n) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = g / kMeshLanes, lane = g % kMeshLanes;
    const bool live = i < n;  // every lane of the wave takes part in the reduction
    double best = 1.0 / 0.0;
    int bt = -1;
    double bw[3] = {0., 0., 0.};
    if (live) {
        const P3 x = at_perm(qn, n, i);
        const double h = mesh_cell(Mo.ext_bits);
        const long long ix = cell_of(x.x, h), iy = cell_of(x.y, h), iz = cell_of(x.z, h);
        const unsigned bkt = cell_hash(ix, iy, iz, Mo.P);
        const int k1 = Mo.bucket_off[bkt + 1];
        for (int k = Mo.bucket_off[bkt] + lane; k < k1; k += kMeshLanes) {
            const double* r = &Mo.bucket_pos[9 * (size_t)k];
            const P3 a = {r[0], r[1], r[2]}, b = {r[3], r[4], r[5]}, c = {r[6], r[7], r[8]};
            const int t = Mo.bucket_tris[k];
            const P3 lo = {fmin(fmin(a.x, b.x), c.x), fmin(fmin(a.y, b.y), c.y), fmin(fmin(a.z, b.z), c.z)};
            const P3 hi = {fmax(fmax(a.x, b.x), c.x), fmax(fmax(a.y, b.y), c.y), fmax(fmax(a.z, b.z), c.z)};
            if (!(lo.x <= x.x + Mo.t11 && x.x - Mo.t11 <= hi.x && lo.y <= x.y + Mo.t11 && x.y - Mo.t11 <= hi.y &&
                  lo.z <= x.z + Mo.t11 && x.z - Mo.t11 <= hi.z))
                continue;
            double w[3];
            P3 cp;
            closest_point(x, a, b, c, Mo.tol, w, cp);
            const P3 dcp = psub(cp, x);
            if (!(pdot(dcp, dcp) < Mo.tol_sq)) continue;
            const P3 pt = {(a.x * w[0] + b.x * w[1]) + c.x * w[2], (a.y * w[0] + b.y * w[1]) + c.y * w[2],
                           (a.z * w[0] + b.z * w[1]) + c.z * w[2]};
            const P3 dp = psub(pt, x);
            const double d = pdot(dp, dp);
            if (d < best || (d == best && t < bt)) {
                best = d;
                bt = t;
                bw[0] = w[0];
                bw[1] = w[1];
                bw[2] = w[2];
            }
        }
    }
    for (int o = 1; o < kMeshLanes; o <<= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int ot = __shfl_xor(bt, o, 64);
        const double o0 = __shfl_xor(bw[0], o, 64), o1 = __shfl_xor(bw[1], o, 64), o2 = __shfl_xor(bw[2], o, 64);
        // a lane without a hit holds (inf, -1) and never wins against a hit
        if (ot >= 0 && (bt < 0 || ob < best || (ob == best && ot < bt))) {
            best = ob;
            bt = ot;
            bw[0] = o0;
            bw[1] = o1;
            bw[2] = o2;
        }
    }
    if (!live || lane != 0) return;
    const bool hit = bt >= 0 && best < Mo.tol_sq;
    Mo.hit_tri[i] = hit ? bt : -1;
    Mo.hit_bary[3 * i] = bw[0];
    Mo.hit_bary[3 * i + 1] = bw[1];
    Mo.hit_bary[3 * i + 2] = bw[2];
}

}  // namespace

void launch_mesh_query(const DevObstacles& ob, int m, const double* qn, int n, hipStream_t st) {
    static_assert(kB % kMeshLanes == 0 && 64 % kMeshLanes == 0, "a vertex's lanes share one wave");
    if (n > 0) hipLaunchKernelGGL(k_mesh_query, dim3(nblk((long long)n * kMeshLanes)), dim3(kB), 0, st, ob.mesh[m], qn, n);
}

void launch_mesh_lerp(double* cur, const double* prev, const double* next, double a, int count, hipStream_t st) {
    if (count > 0) hipLaunchKernelGGL(k_mesh_lerp, dim3(nblk(count)), dim3(kB), 0, st, cur, prev, next, a, count);
}

void launch_mesh_build(const DevMeshBuild& B, hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_clear, dim3(nblk(B.P + 1)), dim3(kB), 0, st, B);
    if (B.T > 0) {
        hipLaunchKernelGGL(k_mesh_box, dim3(nblk(B.T)), dim3(kB), 0, st, B);
        hipLaunchKernelGGL(k_mesh_count, dim3(nblk(B.T)), dim3(kB), 0, st, B);
    }
    launch_scan_int(B.off, B.off, B.P + 1, B.scan_sums, B.scan_total, nullptr, st);
    if (B.T > 0) hipLaunchKernelGGL(k_mesh_scatter, dim3(nblk(B.T)), dim3(kB), 0, st, B);
}

void launch_fill_int(int* p, int count, int value, hipStream_t st) {
    if (count > 0) hipLaunchKernelGGL(k_fill_int, dim3(nblk(count)), dim3(kB), 0, st, p, count, value);
}

void launch_scan_int(int* dst, const int* src, int count, int* sums, int* total, int* tail, hipStream_t st) {
    const unsigned nb = nblk(count);
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kB), 0, st, dst, src, count, sums);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kB), 0, st, dst, count, sums, total, tail);
}

// The bucket counts are zero on entry (the previous detection's scatter took them back to zero,
// or the allocation cleared them), and so is the largest box extent (k_seed_bits).
void launch_self_detect(const DevSelfDetect& D, const int* ccnt, unsigned long long* seed_bits, hipStream_t st) {
    hipLaunchKernelGGL(k_tri_box, dim3(nblk(D.T)), dim3(kB), 0, st, D);
    hipLaunchKernelGGL(k_tri_count, dim3(nblk(D.T)), dim3(kB), 0, st, D);
    // bucket_off = exclusive scan of bucket_cnt (P+1 entries, the last one becomes T)
    launch_scan_int(D.bucket_off, D.bucket_cnt, D.P + 1, D.scan_sums, D.totals + 2, nullptr, st);
    hipLaunchKernelGGL(k_tri_scatter, dim3(nblk(D.T)), dim3(kB), 0, st, D);
    hipLaunchKernelGGL(k_vertex_query, dim3(nblk(D.n)), dim3(kB), 0, st, D, 0);
    launch_scan_int(D.off_full, D.cnt_full, D.n, D.scan_sums, D.totals + 0, nullptr, st);
    launch_scan_int(D.off_filt, D.cnt_filt, D.n, D.scan_sums2, D.totals + 1, nullptr, st);
    hipLaunchKernelGGL(k_vertex_query, dim3(nblk(D.n)), dim3(kB), 0, st, D, 1);
    hipLaunchKernelGGL(k_seed_bits, dim3(nblk(D.n)), dim3(kB), 0, st, ccnt, D.n, seed_bits, D.ext_bits);
}

void launch_self_detect_filtered(const DevSelfDetect& D, const int* ccnt, unsigned long long* seed_bits,
                                 hipStream_t st) {
    hipLaunchKernelGGL(k_tri_box, dim3(nblk(D.T)), dim3(kB), 0, st, D);
    hipLaunchKernelGGL(k_tri_count, dim3(nblk(D.T)), dim3(kB), 0, st, D);
    launch_scan_int(D.bucket_off, D.bucket_cnt, D.P + 1, D.scan_sums, D.totals + 2, nullptr, st);
    hipLaunchKernelGGL(k_tri_scatter_packed, dim3(nblk(D.T)), dim3(kB), 0, st, D);
    hipLaunchKernelGGL(k_vertex_query_flat, dim3((unsigned)((D.n + kB / kFlatLanes - 1) / (kB / kFlatLanes))),
                       dim3(kB), 0, st, D);
    launch_scan_int(D.off_filt, D.cnt_filt, D.n, D.scan_sums2, D.totals + 1, nullptr, st);
    hipLaunchKernelGGL(k_compact_slots, dim3(nblk(D.n)), dim3(kB), 0, st, D);
    hipLaunchKernelGGL(k_seed_bits, dim3(nblk(D.n)), dim3(kB), 0, st, ccnt, D.n, seed_bits, D.ext_bits);
}

void launch_self_prepare(const DevSelf& S, int* vslot, hipStream_t st) {
    int c = S.nslots > S.nsc ? S.nslots : S.nsc;
    if (c > 0) hipLaunchKernelGGL(k_self_prepare, dim3(nblk(c)), dim3(kB), 0, st, S, vslot);
}
void launch_self_level(const DevSelf& S, int l, int count, const double* v, int n, const double* mass,
                       const StepArgs& a, int iteration, hipStream_t st) {
    if (count > 0) launch_k(k_self_level, dim3(nblk(count)), dim3(kB), 0, st, S, l, v, n, mass, a, iteration);
}
void launch_self_apply(const DevSelf& S, int l, int groups, hipStream_t st) {
    if (groups > 0) launch_k(k_self_apply, dim3(nblk(groups)), dim3(kB), 0, st, S, l);
}
void launch_self_fused(const DevSelf& S, int l0, int nlev, const double* v, int n, const double* mass,
                       const StepArgs& a, int iteration, double* rhs, bool finish, hipStream_t st) {
    launch_k(k_self_fused, dim3(1), dim3(kSelfFusedThreads), 0, st, S, l0, nlev, v, n, mass, a, iteration, rhs,
             finish ? 1 : 0);
}
void launch_self_fused_lds(const DevSelf& S, int nlev, const double* v, int n, const double* mass, const StepArgs& a,
                           int iteration, double* rhs, hipStream_t st) {
    launch_k(k_self_fused_lds, dim3(1), dim3(kSelfFusedThreads), 0, st, S, nlev, v, n, mass, a, iteration, rhs);
}
void launch_self_finalize(const DevSelf& S, double* rhs, int n, hipStream_t st) {
    if (S.nslots > 0) launch_k(k_self_finalize, dim3(nblk(S.nslots)), dim3(kB), 0, st, S, rhs, n);
}
void launch_self_end(const DevSelf& S, int* vslot, const double* v, int n, const StepArgs& a, int iteration,
                     hipStream_t st) {
    const int c = S.nslots > S.nsc ? S.nslots : S.nsc;
    if (c > 0) hipLaunchKernelGGL(k_self_end, dim3(nblk(c)), dim3(kB), 0, st, S, vslot, v, n, a, iteration);
}

}  // namespace pdfc
