// GPU build of the reference binary BVH and its traversal layout.
//
// The tree is the one source/bvh.cpp:102-210 builds, node for node:
//   * a node of n >= 3 primitives takes axis and interval from the single
//     primitive box of largest extent (bvh.cpp:119-139, first maximum wins),
//   * evaluates 10 planes min + (max-min)/10*i with the SAH cost
//     1 + (nL*SA_L + nR*SA_R)/SA (bvh.cpp:46-98, first strict minimum),
//   * partitions stably by centroid < plane, or by even/odd position when every
//     plane leaves a side empty (bvh.cpp:156-180);
//   * n == 2 gives a parent of both, n == 1 the primitive leaf itself.
// It is built level-synchronously: every pending node of a level is one
// workgroup; the preorder index of each node follows from the leaf counts
// (left child p+1, right child p+2*nL), so the serialised layout of
// bvh.cpp:213-260 (3 x vec4 per node, hit/miss links) is written in place.
//
// Candidate statistics are gathered in 11 centroid bins (the planes are
// nondecreasing, so "left of plane i" == "bin <= i"); only their surface areas
// are used, which do not depend on the order of the min/max folds.  Node boxes
// are written from an order-preserving fold (glm::min/max keep the first of
// equal values), so the serialised bits equal a sequential build's.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>
#include <cstring>

#include "krt_device.h"
#include "krt_internal.h"

namespace krt {

namespace {

constexpr int NBIN = 11;

struct Seg {
    int start, count, pre, miss, rank;
};

__global__ void tri_boxes_kernel(const float *pos, const uint32_t *idx, uint32_t nt, float4 *bmin, float4 *bmax) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const float *p1 = pos + 3 * idx[3 * t], *p2 = pos + 3 * idx[3 * t + 1], *p3 = pos + 3 * idx[3 * t + 2];
    // include/mesh.hpp:66-71
    bmin[t] = make_float4(gmin(p1[0], gmin(p2[0], p3[0])), gmin(p1[1], gmin(p2[1], p3[1])),
                          gmin(p1[2], gmin(p2[2], p3[2])), 0.0f);
    bmax[t] = make_float4(gmax(p1[0], gmax(p2[0], p3[0])), gmax(p1[1], gmax(p2[1], p3[1])),
                          gmax(p1[2], gmax(p2[2], p3[2])), 0.0f);
}

__global__ void iota_kernel(int *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

__device__ __forceinline__ float comp(const float4 &v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

__device__ __forceinline__ float surface_area(float mnx, float mny, float mnz, float mxx, float mxy, float mxz) {
    float dx = mxx - mnx, dy = mxy - mny, dz = mxz - mnz;  // include/bbox.hpp:16-26
    float xy = dx * dy, yz = dy * dz, xz = dx * dz;
    return 2.0f * (xy + yz + xz);
}

// Serialised node writers (bvh.cpp:241-253)
__device__ void write_node(float4 *ser, int pre, bool leaf, int object, int hit, int miss, float4 mn, float4 mx) {
    ser[3 * pre] = make_float4(leaf ? 1.0f : 0.0f, __int_as_float(object), __int_as_float(hit), __int_as_float(miss));
    ser[3 * pre + 1] = make_float4(mn.x, mn.y, mn.z, 0.0f);
    ser[3 * pre + 2] = make_float4(mx.x, mx.y, mx.z, 0.0f);
}

struct BuildOut {
    float4 *ser;
    int *leaf_order;
    int *leafcount;  // per preorder node
    int *rank;       // first leaf rank per preorder node
};

__device__ void emit_leaf(const BuildOut &o, int pre, int miss, int rank, int prim, const float4 *bmin,
                          const float4 *bmax) {
    write_node(o.ser, pre, true, prim, miss, miss, bmin[prim], bmax[prim]);
    o.leaf_order[rank] = prim;
    o.leafcount[pre] = 1;
    o.rank[pre] = rank;
}

__device__ void emit_pair(const BuildOut &o, int pre, int miss, int rank, int a, int b, const float4 *bmin,
                          const float4 *bmax) {
    float4 mn = bmin[a], mx = bmax[a];
    float4 bn = bmin[b], bx = bmax[b];
    mn = make_float4(gmin(mn.x, bn.x), gmin(mn.y, bn.y), gmin(mn.z, bn.z), 0.0f);
    mx = make_float4(gmax(mx.x, bx.x), gmax(mx.y, bx.y), gmax(mx.z, bx.z), 0.0f);
    write_node(o.ser, pre, false, -1, 3 * (pre + 1), miss, mn, mx);
    o.leafcount[pre] = 2;
    o.rank[pre] = rank;
    emit_leaf(o, pre + 1, 3 * (pre + 2), rank, a, bmin, bmax);
    emit_leaf(o, pre + 2, miss, rank + 1, b, bmin, bmax);
}

__global__ void root_small_kernel(BuildOut o, int n, const float4 *bmin, const float4 *bmax) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (n == 1) emit_leaf(o, 0, -1, 0, 0, bmin, bmax);
    else emit_pair(o, 0, -1, 0, 0, 1, bmin, bmax);
}

// Primitive boxes in the order of a permutation array (perm[i]'s box at i):
// each level reads its segments' boxes as contiguous runs instead of
// gathering them through the permutation, and writes them in the order of
// its output permutation.
struct PermBoxes {
    float4 *mn, *mx;
};

// The bvh.cpp tree's first levels: pass 1 (largest single extent, union
// box) and pass 2 (the bins) by kPreWide workgroups per node.  Every ordered
// choice becomes a 64-bit key that one atomic min / max settles in any order:
// {value, index} with ties to the lower index, +0 and -0 compared equal (the
// winning element's own bits are read back), so the level kernel that takes
// the results over makes the sequential scan's choices.
constexpr int kPreWideSegs = 16;
constexpr uint32_t kPreMinPrims = 1u << 16;  // both trees' pre-passes: builds of at least this many
constexpr int kPreWide = 64;
struct PreExt {
    unsigned long long ext;    // {extent bits, ~index}: max
    unsigned long long mn[3];  // {order key of the coordinate, index}: min
    unsigned long long mx[3];  // {order key, ~index}: max
    int cnt[NBIN];
    uint32_t bmn[NBIN][3], bmx[NBIN][3];  // order keys
};

__device__ __forceinline__ uint32_t okey(float v) {  // order-preserving, +0 == -0
    const uint32_t u = __float_as_uint(v == 0.0f ? 0.0f : v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ounkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void pre_ext_clear_kernel(PreExt *pe) {
    PreExt &e = pe[blockIdx.x];
    const int t = threadIdx.x;
    if (t == 0) e.ext = 0ull;
    if (t < 3) {
        e.mn[t] = ~0ull;
        e.mx[t] = 0ull;
    }
    if (t < NBIN) {
        e.cnt[t] = 0;
        for (int a = 0; a < 3; a++) {
            e.bmn[t][a] = 0xffffffffu;
            e.bmx[t][a] = 0u;
        }
    }
}

__global__ __launch_bounds__(256) void pre_ext_kernel(const Seg *segs, PermBoxes bin_, PreExt *pe) {
    const int seg = blockIdx.x / kPreWide, part = blockIdx.x % kPreWide;
    const Seg sg = segs[seg];
    __shared__ unsigned long long s_ext, s_mn[3], s_mx[3];
    if (threadIdx.x == 0) s_ext = 0ull;
    if (threadIdx.x < 3) {
        s_mn[threadIdx.x] = ~0ull;
        s_mx[threadIdx.x] = 0ull;
    }
    __syncthreads();
    unsigned long long ext = 0ull, mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0ull, 0ull, 0ull};
    for (int j = part * 256 + threadIdx.x; j < sg.count; j += kPreWide * 256) {
        const float4 a = bin_.mn[sg.start + j], b = bin_.mx[sg.start + j];
        const float pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z};
        float e = 0.0f;
        for (int k = 0; k < 3; k++) e = fmaxf(e, fabsf(pb[k] - pa[k]));
        const unsigned long long ke = (unsigned long long)__float_as_uint(e) << 32 | (0xffffffffu - (uint32_t)j);
        ext = ke > ext ? ke : ext;
        for (int k = 0; k < 3; k++) {
            const unsigned long long kn = (unsigned long long)okey(pa[k]) << 32 | (uint32_t)j;
            const unsigned long long kx = (unsigned long long)okey(pb[k]) << 32 | (0xffffffffu - (uint32_t)j);
            mn[k] = kn < mn[k] ? kn : mn[k];
            mx[k] = kx > mx[k] ? kx : mx[k];
        }
    }
    atomicMax(&s_ext, ext);
    for (int k = 0; k < 3; k++) {
        atomicMin(&s_mn[k], mn[k]);
        atomicMax(&s_mx[k], mx[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&pe[seg].ext, s_ext);
    if (threadIdx.x < 3) {
        atomicMin(&pe[seg].mn[threadIdx.x], s_mn[threadIdx.x]);
        atomicMax(&pe[seg].mx[threadIdx.x], s_mx[threadIdx.x]);
    }
}

// The winner of pass 1 from the keys: the element of largest extent (lowest
// index on ties) and its first axis of that extent; extent 0 everywhere
// keeps the scan's initial values.
__device__ void pre_ext_result(const PreExt &e, const Seg &sg, PermBoxes bin_, int &axis, float &lo, float &hi) {
    axis = 0;
    lo = FLT_MAX;
    hi = -FLT_MAX;
    if ((e.ext >> 32) == 0ull) return;
    const int j = (int)(0xffffffffu - (uint32_t)(e.ext & 0xffffffffull));
    const float4 a = bin_.mn[sg.start + j], b = bin_.mx[sg.start + j];
    float best = 0.0f;
    for (int k = 0; k < 3; k++) {
        const float ev = fabsf(comp(b, k) - comp(a, k));
        if (ev > best) {
            best = ev;
            lo = comp(a, k);
            hi = comp(b, k);
            axis = k;
        }
    }
}

__global__ __launch_bounds__(256) void pre_split_bins_kernel(const Seg *segs, PermBoxes bin_, PreExt *pe) {
    const int seg = blockIdx.x / kPreWide, part = blockIdx.x % kPreWide;
    const Seg sg = segs[seg];
    __shared__ int s_cnt[NBIN];
    __shared__ uint32_t s_mn[NBIN][3], s_mx[NBIN][3];
    if (threadIdx.x < NBIN) {
        s_cnt[threadIdx.x] = 0;
        for (int a = 0; a < 3; a++) {
            s_mn[threadIdx.x][a] = 0xffffffffu;
            s_mx[threadIdx.x][a] = 0u;
        }
    }
    __syncthreads();
    int axis;
    float lo, hi;
    pre_ext_result(pe[seg], sg, bin_, axis, lo, hi);
    float split[10];
    for (int i = 0; i < 10; i++) split[i] = (hi - lo) / 10.0f * (float)i + lo;
    for (int j = part * 256 + threadIdx.x; j < sg.count; j += kPreWide * 256) {
        const float4 mn = bin_.mn[sg.start + j], mx = bin_.mx[sg.start + j];
        const float c = (comp(mn, axis) + comp(mx, axis)) / 2.0f;
        int bin = 0;
        for (int i = 0; i < 10; i++) bin += (c < split[i]) ? 0 : 1;
        atomicAdd(&s_cnt[bin], 1);
        const float pa[3] = {mn.x, mn.y, mn.z}, pb[3] = {mx.x, mx.y, mx.z};
        for (int a = 0; a < 3; a++) {
            atomicMin(&s_mn[bin][a], okey(pa[a]));
            atomicMax(&s_mx[bin][a], okey(pb[a]));
        }
    }
    __syncthreads();
    if (threadIdx.x < NBIN && s_cnt[threadIdx.x]) {
        const int b = threadIdx.x;
        atomicAdd(&pe[seg].cnt[b], s_cnt[b]);
        for (int a = 0; a < 3; a++) {
            atomicMin(&pe[seg].bmn[b][a], s_mn[b][a]);
            atomicMax(&pe[seg].bmx[b][a], s_mx[b][a]);
        }
    }
}

// One workgroup of NT lanes per pending node (count >= 3).  The result does
// not depend on NT or on the order of the reductions: candidates carry their
// index (ties to the lower), the bins are order-free and the partition is
// stable.
template <int NT>
__global__ __launch_bounds__(NT) void level_kernel(const Seg *segs, const int *perm_in, int *perm_out,
                                                     const float4 *bmin, const float4 *bmax, BuildOut o, Seg *next,
                                                     int *next_count, PermBoxes bin_, PermBoxes bout,
                                                     const PreExt *pre) {
    const Seg sg = segs[blockIdx.x];
    const int tid = threadIdx.x;
    const int n = sg.count;

    __shared__ float s_ext[NT], s_lo[NT], s_hi[NT];
    __shared__ int s_axis[NT], s_jb[NT], s_has[NT];
    __shared__ float s_umn[3][NT], s_umx[3][NT];
    __shared__ int s_jmn[3][NT], s_jmx[3][NT];

    // ---- pass 1: largest single extent and union box.  The lanes stride the
    // segment (coalesced reads); every candidate carries its index and ties
    // go to the lower one, which is what the sequential scan of bvh.cpp
    // gives (a later element replaces the current one only if strictly
    // larger, or smaller for the box minimum) -- also between +0 and -0.
    float best = 0.0f, lo = FLT_MAX, hi = -FLT_MAX;
    int axis = 0, jb = INT_MAX;
    if (pre) {  // pass 1 was done by the pre-pass: take its winners
        if (tid == 0) {
            const PreExt &e = pre[blockIdx.x];
            pre_ext_result(e, sg, bin_, s_axis[0], s_lo[0], s_hi[0]);
            for (int a = 0; a < 3; a++) {
                const int jn = (int)(uint32_t)(e.mn[a] & 0xffffffffull);
                const int jx = (int)(0xffffffffu - (uint32_t)(e.mx[a] & 0xffffffffull));
                s_umn[a][0] = comp(bin_.mn[sg.start + jn], a);  // the winner's own bits (+0 or -0)
                s_umx[a][0] = comp(bin_.mx[sg.start + jx], a);
            }
        }
        __syncthreads();
    } else {
    float umn[3] = {0, 0, 0}, umx[3] = {0, 0, 0};
    int jmn[3] = {0, 0, 0}, jmx[3] = {0, 0, 0};
    for (int j = tid; j < n; j += NT) {
        float4 mn = bin_.mn[sg.start + j], mx = bin_.mx[sg.start + j];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float e = fabsf(comp(mx, a) - comp(mn, a));
            if (e > best) {
                best = e;
                lo = comp(mn, a);
                hi = comp(mx, a);
                axis = a;
                jb = j;
            }
        }
        const float pm[3] = {mn.x, mn.y, mn.z}, px[3] = {mx.x, mx.y, mx.z};
        if (j == tid) {
            for (int a = 0; a < 3; a++) {
                umn[a] = pm[a];
                umx[a] = px[a];
                jmn[a] = jmx[a] = j;
            }
        } else {
            for (int a = 0; a < 3; a++) {
                if (pm[a] < umn[a]) {
                    umn[a] = pm[a];
                    jmn[a] = j;
                }
                if (px[a] > umx[a]) {
                    umx[a] = px[a];
                    jmx[a] = j;
                }
            }
        }
    }
    s_ext[tid] = best;
    s_lo[tid] = lo;
    s_hi[tid] = hi;
    s_axis[tid] = axis;
    s_jb[tid] = jb;
    s_has[tid] = tid < n;
    for (int a = 0; a < 3; a++) {
        s_umn[a][tid] = umn[a];
        s_umx[a][tid] = umx[a];
        s_jmn[a][tid] = jmn[a];
        s_jmx[a][tid] = jmx[a];
    }
    __syncthreads();
    for (int s = 1; s < NT; s <<= 1) {
        if ((tid % (2 * s)) == 0 && tid + s < NT) {
            int r = tid + s;
            if (s_ext[r] > s_ext[tid] || (s_ext[r] == s_ext[tid] && s_jb[r] < s_jb[tid])) {
                s_ext[tid] = s_ext[r];
                s_lo[tid] = s_lo[r];
                s_hi[tid] = s_hi[r];
                s_axis[tid] = s_axis[r];
                s_jb[tid] = s_jb[r];
            }
            if (s_has[r]) {
                if (s_has[tid]) {
                    for (int a = 0; a < 3; a++) {
                        const float vn = s_umn[a][r], vx = s_umx[a][r];
                        if (vn < s_umn[a][tid] || (vn == s_umn[a][tid] && s_jmn[a][r] < s_jmn[a][tid])) {
                            s_umn[a][tid] = vn;
                            s_jmn[a][tid] = s_jmn[a][r];
                        }
                        if (vx > s_umx[a][tid] || (vx == s_umx[a][tid] && s_jmx[a][r] < s_jmx[a][tid])) {
                            s_umx[a][tid] = vx;
                            s_jmx[a][tid] = s_jmx[a][r];
                        }
                    }
                } else {
                    for (int a = 0; a < 3; a++) {
                        s_umn[a][tid] = s_umn[a][r];
                        s_umx[a][tid] = s_umx[a][r];
                        s_jmn[a][tid] = s_jmn[a][r];
                        s_jmx[a][tid] = s_jmx[a][r];
                    }
                    s_has[tid] = 1;
                }
            }
        }
        __syncthreads();
    }
    }  // pass 1
    axis = s_axis[0];
    lo = s_lo[0];
    hi = s_hi[0];
    const float4 nmn = make_float4(s_umn[0][0], s_umn[1][0], s_umn[2][0], 0.0f);
    const float4 nmx = make_float4(s_umx[0][0], s_umx[1][0], s_umx[2][0], 0.0f);

    // ---- pass 2: centroid bins over the 10 planes
    float split[10];
#pragma unroll
    for (int i = 0; i < 10; i++) split[i] = (hi - lo) / 10.0f * (float)i + lo;
    // wave reduction, then across the waves through LDS
    __shared__ int w_c[NT / 64][NBIN];
    __shared__ float w_mn[NT / 64][NBIN][3], w_mx[NT / 64][NBIN][3];
    const int lane = tid & 63, wv = tid >> 6;
    if constexpr (NT == 64) {
        // one wavefront per node (the deep levels: few elements each): LDS
        // atomics on order keys keep the bins out of registers
        __shared__ int l_c[NBIN];
        __shared__ uint32_t l_mn[NBIN][3], l_mx[NBIN][3];
        if (tid < NBIN) {
            l_c[tid] = 0;
            for (int a = 0; a < 3; a++) {
                l_mn[tid][a] = 0xffffffffu;
                l_mx[tid][a] = 0u;
            }
        }
        __syncthreads();
        for (int j = pre ? n : tid; j < n; j += NT) {
            float4 mn = bin_.mn[sg.start + j], mx = bin_.mx[sg.start + j];
            float c = (comp(mn, axis) + comp(mx, axis)) / 2.0f;
            int bin = 0;
#pragma unroll
            for (int i = 0; i < 10; i++) bin += (c < split[i]) ? 0 : 1;
            atomicAdd(&l_c[bin], 1);
            const float pa[3] = {mn.x, mn.y, mn.z}, pb[3] = {mx.x, mx.y, mx.z};
            for (int a = 0; a < 3; a++) {
                atomicMin(&l_mn[bin][a], okey(pa[a]));
                atomicMax(&l_mx[bin][a], okey(pb[a]));
            }
        }
        __syncthreads();
        if (tid < NBIN) {  // an empty bin keeps +-FLT_MAX
            w_c[0][tid] = l_c[tid];
            for (int a = 0; a < 3; a++) {
                w_mn[0][tid][a] = l_c[tid] ? ounkey(l_mn[tid][a]) : FLT_MAX;
                w_mx[0][tid][a] = l_c[tid] ? ounkey(l_mx[tid][a]) : -FLT_MAX;
            }
        }
    } else {
    int bc[NBIN];
    float bmn[NBIN][3], bmx[NBIN][3];
#pragma unroll
    for (int b = 0; b < NBIN; b++) {
        bc[b] = 0;
        for (int a = 0; a < 3; a++) {
            bmn[b][a] = FLT_MAX;
            bmx[b][a] = -FLT_MAX;
        }
    }
    for (int j = pre ? n : tid; j < n; j += NT) {  // order-free: counts and box unions per bin
        float4 mn = bin_.mn[sg.start + j], mx = bin_.mx[sg.start + j];
        float c = (comp(mn, axis) + comp(mx, axis)) / 2.0f;
        int bin = 0;
#pragma unroll
        for (int i = 0; i < 10; i++) bin += (c < split[i]) ? 0 : 1;
#pragma unroll
        for (int b = 0; b < NBIN; b++) {
            if (bin == b) {
                bc[b]++;
                bmn[b][0] = fminf(bmn[b][0], mn.x); bmn[b][1] = fminf(bmn[b][1], mn.y); bmn[b][2] = fminf(bmn[b][2], mn.z);
                bmx[b][0] = fmaxf(bmx[b][0], mx.x); bmx[b][1] = fmaxf(bmx[b][1], mx.y); bmx[b][2] = fmaxf(bmx[b][2], mx.z);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < NBIN; b++) {
        for (int off = 32; off > 0; off >>= 1) {
            bc[b] += __shfl_xor(bc[b], off, 64);
            for (int a = 0; a < 3; a++) {
                bmn[b][a] = fminf(bmn[b][a], __shfl_xor(bmn[b][a], off, 64));
                bmx[b][a] = fmaxf(bmx[b][a], __shfl_xor(bmx[b][a], off, 64));
            }
        }
        if (lane == 0) {
            w_c[wv][b] = bc[b];
            for (int a = 0; a < 3; a++) {
                w_mn[wv][b][a] = bmn[b][a];
                w_mx[wv][b][a] = bmx[b][a];
            }
        }
    }
    }  // pass 2
    __shared__ int s_nl, s_fallback;
    __shared__ float s_split;
    __syncthreads();
    if (tid < 64) {
        // wave 0: lane b holds bin b (merged over the waves), a scan over the
        // lanes gives every plane its two sides, lane i prices plane i.  The
        // counts are exact; min / max in the scan's order can differ from
        // the sequential one only in the sign of a zero, which no surface
        // area (differences of the bounds) and no cost (1 + ...) keeps.
        int c = 0;
        float bm[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        if (lane < NBIN) {
            for (int w = 0; w < NT / 64; w++) {
                c += w_c[w][lane];
                for (int a = 0; a < 3; a++) {
                    bm[a] = fminf(bm[a], w_mn[w][lane][a]);
                    bx[a] = fmaxf(bx[a], w_mx[w][lane][a]);
                }
            }
            if (pre) {  // the bins of the pre-pass (an empty bin keeps +-FLT_MAX)
                c = pre[blockIdx.x].cnt[lane];
                if (c)
                    for (int a = 0; a < 3; a++) {
                        bm[a] = ounkey(pre[blockIdx.x].bmn[lane][a]);
                        bx[a] = ounkey(pre[blockIdx.x].bmx[lane][a]);
                    }
            }
        }
        int lc = c, rc = c;  // bins 0..lane, bins lane..NBIN-1
        float lm[3], lx[3], rm[3], rx[3];
        for (int a = 0; a < 3; a++) {
            lm[a] = rm[a] = bm[a];
            lx[a] = rx[a] = bx[a];
        }
        for (int d = 1; d < 16; d <<= 1) {
            const int tl = __shfl_up(lc, d, 64), tr = __shfl_down(rc, d, 64);
            if (lane >= d) lc += tl;
            if (lane + d < 64) rc += tr;
            for (int a = 0; a < 3; a++) {
                const float ulm = __shfl_up(lm[a], d, 64), ulx = __shfl_up(lx[a], d, 64);
                const float drm = __shfl_down(rm[a], d, 64), drx = __shfl_down(rx[a], d, 64);
                if (lane >= d) {
                    lm[a] = fminf(lm[a], ulm);
                    lx[a] = fmaxf(lx[a], ulx);
                }
                if (lane + d < 64) {
                    rm[a] = fminf(rm[a], drm);
                    rx[a] = fmaxf(rx[a], drx);
                }
            }
        }
        // the right side of plane i: bins i+1..NBIN-1
        const int nl = lc, nr = __shfl_down(rc, 1, 64);
        float q[6];
        for (int a = 0; a < 3; a++) {
            q[a] = __shfl_down(rm[a], 1, 64);
            q[3 + a] = __shfl_down(rx[a], 1, 64);
        }
        const float sa_total = surface_area(nmn.x, nmn.y, nmn.z, nmx.x, nmx.y, nmx.z);
        float cost = FLT_MAX;
        if (lane < 10 && nl != 0 && nr != 0)
            cost = 1.0f + ((float)nl * surface_area(lm[0], lm[1], lm[2], lx[0], lx[1], lx[2]) +
                           (float)nr * surface_area(q[0], q[1], q[2], q[3], q[4], q[5])) /
                              sa_total;
        // the planes in order, a later one only if strictly cheaper
        float min_cost = FLT_MAX, min_split = 0.0f;
        int best_nl = 0;
#pragma unroll
        for (int k = 0; k < 10; k++) {
            const float ck = __shfl(cost, k, 64);
            const int nk = __shfl(nl, k, 64);
            if (ck < min_cost) {
                min_cost = ck;
                min_split = split[k];
                best_nl = nk;
            }
        }
        if (tid == 0) {
            s_fallback = (min_cost == FLT_MAX);
            s_split = min_split;
            s_nl = s_fallback ? (n + 1) / 2 : best_nl;
        }
    }
    __syncthreads();
    const bool fallback = s_fallback;
    const float msplit = s_split;
    const int nl = s_nl;

    // ---- pass 3: stable partition into perm_out
    __shared__ int s_wl[NT / 64];
    int lbase = 0, rbase = 0;
    for (int t0 = 0; t0 < n; t0 += NT) {
        int j = t0 + tid;
        bool valid = j < n;
        int prim = valid ? perm_in[sg.start + j] : 0;
        float4 pmn = make_float4(0, 0, 0, 0), pmx = pmn;
        bool left = false;
        if (valid) {
            pmn = bin_.mn[sg.start + j];
            pmx = bin_.mx[sg.start + j];
            if (fallback) left = (j % 2) == 0;
            else {
                float c = (comp(pmn, axis) + comp(pmx, axis)) / 2.0f;
                left = c < msplit;
            }
        }
        unsigned long long ball = __ballot(left);
        int lpre = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) s_wl[wv] = __popcll(ball);
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < NT / 64; w++) {
            if (w < wv) woff += s_wl[w];
            tot += s_wl[w];
        }
        int tile_valid = min(NT, n - t0);
        if (valid) {
            int lrank = woff + lpre;  // left items before me in the tile
            const int at = sg.start + (left ? lbase + lrank : nl + rbase + (tid - lrank));
            perm_out[at] = prim;
            bout.mn[at] = pmn;
            bout.mx[at] = pmx;
        }
        lbase += tot;
        rbase += tile_valid - tot;
        __syncthreads();
    }
    __syncthreads();

    // ---- emit this node and the children that finish here
    if (tid == 0) {
        const int p = sg.pre;
        write_node(o.ser, p, false, -1, 3 * (p + 1), sg.miss, nmn, nmx);
        o.leafcount[p] = n;
        o.rank[p] = sg.rank;
        const int nr = n - nl;
        Seg L{sg.start, nl, p + 1, 3 * (p + 2 * nl), sg.rank};
        Seg R{sg.start + nl, nr, p + 2 * nl, sg.miss, sg.rank + nl};
        const Seg ch[2] = {L, R};
        // one reservation for both pending children (the deep levels have
        // ~10^5 nodes on this counter)
        int slot = (L.count >= 3) + (R.count >= 3);
        if (slot) slot = atomicAdd(next_count, slot);
        for (int k = 0; k < 2; k++) {
            const Seg &c = ch[k];
            if (c.count >= 3) {
                next[slot++] = c;
            } else if (c.count == 2) {
                emit_pair(o, c.pre, c.miss, c.rank, perm_out[c.start], perm_out[c.start + 1], bmin, bmax);
            } else {
                emit_leaf(o, c.pre, c.miss, c.rank, perm_out[c.start], bmin, bmax);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Traversal layout: binary nodes storing both child boxes.  Subtrees of at most
// leaf_max primitives become leaf ranges of consecutive ranks (a subtree's
// leaves are consecutive in preorder).  Links: internal node index (>= 0) or
// 0x80000000 | (count-1) << 26 | first_rank.
// ---------------------------------------------------------------------------
__global__ void mark_kernel(const float4 *ser, const int *leafcount, uint32_t nser, int leaf_max, int *flag) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nser) return;
    bool internal = ser[3 * p].x == 0.0f;
    flag[p] = (p == 0) || (internal && leafcount[p] > leaf_max);
}

__device__ __forceinline__ int leaf_link(int first, int count) {
    return (int)(0x80000000u | ((uint32_t)(count - 1) << 26) | (uint32_t)first);
}

__global__ void fill_nodes_kernel(const float4 *ser, const int *leafcount, const int *rank, const int *flag,
                                  const int *node_id, uint32_t nser, int leaf_max, float4 *nodes) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nser || !flag[p]) return;
    int id = node_id[p];
    float4 mn[2], mx[2];
    int link[2];
    bool internal = ser[3 * p].x == 0.0f;
    if (!internal || leafcount[p] <= leaf_max) {
        // whole scene in one leaf range (tiny scenes): both slots carry it
        for (int k = 0; k < 2; k++) {
            mn[k] = ser[3 * p + 1];
            mx[k] = ser[3 * p + 2];
            link[k] = leaf_link(rank[p], leafcount[p]);
        }
    } else {
        int c0 = (int)p + 1;
        int c1 = __float_as_int(ser[3 * c0].w) / 3;  // left child's miss link = right child
        int cs[2] = {c0, c1};
        for (int k = 0; k < 2; k++) {
            int c = cs[k];
            mn[k] = ser[3 * c + 1];
            mx[k] = ser[3 * c + 2];
            link[k] = flag[c] ? node_id[c] : leaf_link(rank[c], leafcount[c]);
        }
    }
    nodes[4 * id + 0] = make_float4(mn[0].x, mx[0].x, mn[1].x, mx[1].x);
    nodes[4 * id + 1] = make_float4(mn[0].y, mx[0].y, mn[1].y, mx[1].y);
    nodes[4 * id + 2] = make_float4(mn[0].z, mx[0].z, mn[1].z, mx[1].z);
    nodes[4 * id + 3] = make_float4(__int_as_float(link[0]), __int_as_float(link[1]), 0.0f, 0.0f);
}

// Single-workgroup exclusive scan of the node flags: each thread owns one
// contiguous chunk (sum, block scan of the 1024 partials, then write-out).
__global__ __launch_bounds__(1024) void scan_kernel(const int *flag, int *out, uint32_t n, int *total) {
    __shared__ int s[1024];
    const int tid = threadIdx.x;
    const uint32_t chunk = (n + 1023) / 1024;
    const uint32_t b0 = min((uint32_t)tid * chunk, n), b1 = min(b0 + chunk, n);
    int sum = 0;
    for (uint32_t i = b0; i < b1; i++) sum += flag[i];
    s[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int x = tid >= off ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += x;
        __syncthreads();
    }
    int run = s[tid] - sum;
    for (uint32_t i = b0; i < b1; i++) {
        out[i] = run;
        run += flag[i];
    }
    if (tid == 1023) *total = s[1023];
}

// Triangles in traversal order: order[p] = primitive at position p; the
// bvh.cpp rank of each primitive travels with it for the tie-break.
__global__ void tri_pack_kernel(const float *pos, const uint32_t *idx, const int *order, const int *rank_of,
                                uint32_t nt, float4 *tri) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    int prim = order[r];
    f3 v1 = ld3(pos + 3 * idx[3 * prim]), v2 = ld3(pos + 3 * idx[3 * prim + 1]), v3 = ld3(pos + 3 * idx[3 * prim + 2]);
    f3 e1 = v2 - v1, e2 = v3 - v1;  // as intersect.glsl:16-17 computes them
    tri[3 * r] = make_float4(v1.x, v1.y, v1.z, __int_as_float(prim));
    tri[3 * r + 1] = make_float4(e1.x, e1.y, e1.z, __int_as_float(rank_of[prim]));
    tri[3 * r + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
}

__global__ void invert_kernel(const int *order, uint32_t n, int *rank_of) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) rank_of[order[r]] = (int)r;
}

// ---------------------------------------------------------------------------
// Traversal tree: binned SAH (16 centroid bins on the axis of largest centroid
// extent), level-synchronous, one workgroup per pending node.  Subtrees of at
// most leaf_max primitives become leaf ranges.  Child boxes are the exact
// unions of their primitives' boxes.
// ---------------------------------------------------------------------------
constexpr int SB = 16;

struct SahSeg {
    int start, count, node;
};

__device__ __forceinline__ uint32_t f2u(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float u2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

struct LdsBox {
    uint32_t mn[3], mx[3];
    __device__ void clear() {
        for (int a = 0; a < 3; a++) {
            mn[a] = 0xffffffffu;
            mx[a] = 0u;
        }
    }
    __device__ void add(const float *lo, const float *hi) {
        for (int a = 0; a < 3; a++) {
            atomicMin(&mn[a], f2u(lo[a]));
            atomicMax(&mx[a], f2u(hi[a]));
        }
    }
};

__device__ __forceinline__ float sa_of(const float *mn, const float *mx) {
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    return 2.0f * (dx * dy + dy * dz + dx * dz);
}

// The top levels' centroid bounds and bins, computed by kPreBlocks
// workgroups per node (per-workgroup LDS results merged with global
// atomics): min / max / counts do not depend on the order, so the level
// kernel that takes them over gives the same split.
constexpr int kPreSegs = 16;    // levels with at most this many nodes
constexpr int kPreBlocks = 64;  // workgroups per node
struct PreBins {
    LdsBox cb[kPreSegs];
    LdsBox bin[kPreSegs][SB];
    int cnt[kPreSegs][SB];
};

__global__ void pre_clear_kernel(PreBins *pb) {
    const int s = blockIdx.x, t = threadIdx.x;
    if (t == 0) pb->cb[s].clear();
    if (t < SB) {
        pb->bin[s][t].clear();
        pb->cnt[s][t] = 0;
    }
}

__device__ void pre_axis(const LdsBox &cb, int &axis, float &clo, float &scale) {
    float ext[3];
    for (int k = 0; k < 3; k++) ext[k] = u2f(cb.mx[k]) - u2f(cb.mn[k]);
    axis = (ext[0] >= ext[1] && ext[0] >= ext[2]) ? 0 : (ext[1] >= ext[2] ? 1 : 2);
    clo = u2f(cb.mn[axis]);
    scale = ext[axis] > 0.0f ? (float)SB / ext[axis] : 0.0f;
}

__global__ __launch_bounds__(256) void pre_bounds_kernel(const SahSeg *segs, PermBoxes bin_, PreBins *pb) {
    const int seg = blockIdx.x / kPreBlocks, part = blockIdx.x % kPreBlocks;
    const SahSeg sg = segs[seg];
    __shared__ LdsBox s_cb;
    if (threadIdx.x == 0) s_cb.clear();
    __syncthreads();
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    bool any = false;
    for (int j = part * 256 + threadIdx.x; j < sg.count; j += kPreBlocks * 256) {
        float4 a = bin_.mn[sg.start + j], b = bin_.mx[sg.start + j];
        float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], c[k]);
            hi[k] = fmaxf(hi[k], c[k]);
        }
        any = true;
    }
    if (any) s_cb.add(lo, hi);
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&pb->cb[seg].mn[threadIdx.x], s_cb.mn[threadIdx.x]);
        atomicMax(&pb->cb[seg].mx[threadIdx.x], s_cb.mx[threadIdx.x]);
    }
}

__global__ __launch_bounds__(256) void pre_bins_kernel(const SahSeg *segs, PermBoxes bin_, PreBins *pb) {
    const int seg = blockIdx.x / kPreBlocks, part = blockIdx.x % kPreBlocks;
    const SahSeg sg = segs[seg];
    __shared__ LdsBox s_bin[SB];
    __shared__ int s_cnt[SB];
    if (threadIdx.x < SB) {
        s_bin[threadIdx.x].clear();
        s_cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    int axis;
    float clo, scale;
    pre_axis(pb->cb[seg], axis, clo, scale);
    for (int j = part * 256 + threadIdx.x; j < sg.count; j += kPreBlocks * 256) {
        float4 a = bin_.mn[sg.start + j], b = bin_.mx[sg.start + j];
        float c = 0.5f * (comp(a, axis) + comp(b, axis));
        int bin = min(SB - 1, max(0, (int)((c - clo) * scale)));
        atomicAdd(&s_cnt[bin], 1);
        float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
        s_bin[bin].add(lo, hi);
    }
    __syncthreads();
    if (threadIdx.x < SB && s_cnt[threadIdx.x]) {
        const int b = threadIdx.x;
        atomicAdd(&pb->cnt[seg][b], s_cnt[b]);
        for (int k = 0; k < 3; k++) {
            atomicMin(&pb->bin[seg][b].mn[k], s_bin[b].mn[k]);
            atomicMax(&pb->bin[seg][b].mx[k], s_bin[b].mx[k]);
        }
    }
}

// pb != nullptr: the level's centroid bounds and bins come from the pre-pass.
template <int NT>
__global__ __launch_bounds__(NT) void sah_level_kernel(const SahSeg *segs, const int *perm_in, int *perm_out,
                                                         int *perm_final, PermBoxes bin_, PermBoxes bout,
                                                         float4 *nodes, int *node_ctr, SahSeg *next,
                                                         int *next_count, int leaf_max, const PreBins *pb) {
    const SahSeg sg = segs[blockIdx.x];
    const int tid = threadIdx.x;
    const int n = sg.count;
    __shared__ LdsBox s_cb;  // centroid bounds
    __shared__ LdsBox s_bin[SB];
    __shared__ int s_cnt[SB];
    __shared__ LdsBox s_child[2];
    __shared__ int s_nl, s_axis_split, s_median;
    __shared__ float s_lo, s_scale;
    __shared__ int s_axis;
    __shared__ int s_wl[NT / 64];
    if (tid == 0) s_cb.clear();
    if (tid < SB) {
        s_bin[tid].clear();
        s_cnt[tid] = 0;
    }
    if (tid < 2) s_child[tid].clear();
    __syncthreads();
    // centroid bounds
    if (pb) {
        if (tid == 0) s_cb = pb->cb[blockIdx.x];
    } else {
        float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        for (int j = tid; j < n; j += NT) {
            float4 a = bin_.mn[sg.start + j], b = bin_.mx[sg.start + j];
            float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(lo[k], c[k]);
                hi[k] = fmaxf(hi[k], c[k]);
            }
        }
        if (n > 0) s_cb.add(lo, hi);
    }
    __syncthreads();
    if (tid == 0) {
        float ext[3];
        for (int k = 0; k < 3; k++) ext[k] = u2f(s_cb.mx[k]) - u2f(s_cb.mn[k]);
        int ax = (ext[0] >= ext[1] && ext[0] >= ext[2]) ? 0 : (ext[1] >= ext[2] ? 1 : 2);
        s_axis = ax;
        s_lo = u2f(s_cb.mn[ax]);
        s_scale = ext[ax] > 0.0f ? (float)SB / ext[ax] : 0.0f;
    }
    __syncthreads();
    const int axis = s_axis;
    const float clo = s_lo, scale = s_scale;
    // bins
    if (pb) {
        if (tid < SB) {
            s_bin[tid] = pb->bin[blockIdx.x][tid];
            s_cnt[tid] = pb->cnt[blockIdx.x][tid];
        }
    } else {
        for (int j = tid; j < n; j += NT) {
            float4 a = bin_.mn[sg.start + j], b = bin_.mx[sg.start + j];
            float c = 0.5f * (comp(a, axis) + comp(b, axis));
            int bin = min(SB - 1, max(0, (int)((c - clo) * scale)));
            atomicAdd(&s_cnt[bin], 1);
            float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
            s_bin[bin].add(lo, hi);
        }
    }
    __syncthreads();
    if (tid == 0) {
        float pmn[SB][3], pmx[SB][3];
        int pc[SB];
        // prefix (left) sweep
        float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        int c = 0;
        for (int b = 0; b < SB; b++) {
            c += s_cnt[b];
            if (s_cnt[b])
                for (int k = 0; k < 3; k++) {
                    mn[k] = fminf(mn[k], u2f(s_bin[b].mn[k]));
                    mx[k] = fmaxf(mx[k], u2f(s_bin[b].mx[k]));
                }
            for (int k = 0; k < 3; k++) {
                pmn[b][k] = mn[k];
                pmx[b][k] = mx[k];
            }
            pc[b] = c;
        }
        float best = 3.4e38f;
        int bsplit = -1;
        float rmn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, rmx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        int rc = 0;
        for (int b = SB - 1; b >= 1; b--) {
            rc += s_cnt[b];
            if (s_cnt[b])
                for (int k = 0; k < 3; k++) {
                    rmn[k] = fminf(rmn[k], u2f(s_bin[b].mn[k]));
                    rmx[k] = fmaxf(rmx[k], u2f(s_bin[b].mx[k]));
                }
            int lc = pc[b - 1];
            if (lc == 0 || rc == 0) continue;
            float cost = (float)lc * sa_of(pmn[b - 1], pmx[b - 1]) + (float)rc * sa_of(rmn, rmx);
            if (cost < best) {
                best = cost;
                bsplit = b - 1;
            }
        }
        s_median = bsplit < 0;
        s_axis_split = bsplit;
        s_nl = bsplit < 0 ? n / 2 : pc[bsplit];
    }
    __syncthreads();
    const bool median = s_median;
    const int bsplit = s_axis_split;
    const int nl = s_nl;
    const int lane = tid & 63, wv = tid >> 6;
    // stable partition (median split: by position)
    int lbase = 0, rbase = 0;
    for (int t0 = 0; t0 < n; t0 += NT) {
        int j = t0 + tid;
        bool valid = j < n;
        int prim = valid ? perm_in[sg.start + j] : 0;
        float4 pmn = make_float4(0, 0, 0, 0), pmx = pmn;
        bool left = false;
        if (valid) {
            pmn = bin_.mn[sg.start + j];
            pmx = bin_.mx[sg.start + j];
            if (median) left = j < nl;
            else {
                float c = 0.5f * (comp(pmn, axis) + comp(pmx, axis));
                int bin = min(SB - 1, max(0, (int)((c - clo) * scale)));
                left = bin <= bsplit;
            }
        }
        unsigned long long ball = __ballot(left);
        int lpre = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) s_wl[wv] = __popcll(ball);
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < NT / 64; w++) {
            if (w < wv) woff += s_wl[w];
            tot += s_wl[w];
        }
        if (valid) {
            int lrank = woff + lpre;
            const int at = sg.start + (left ? lbase + lrank : nl + rbase + (tid - lrank));
            perm_out[at] = prim;
            bout.mn[at] = pmn;
            bout.mx[at] = pmx;
        }
        lbase += tot;
        rbase += min(NT, n - t0) - tot;
        __syncthreads();
    }
    // child boxes
    {
        float lo[2][3], hi[2][3];
        for (int c = 0; c < 2; c++)
            for (int k = 0; k < 3; k++) {
                lo[c][k] = 3.4e38f;
                hi[c][k] = -3.4e38f;
            }
        bool any[2] = {false, false};
        __syncthreads();  // the partition's box copies are read back below
        for (int j = tid; j < n; j += NT) {
            int c = j < nl ? 0 : 1;
            float4 a = bout.mn[sg.start + j], b = bout.mx[sg.start + j];
            lo[c][0] = fminf(lo[c][0], a.x); lo[c][1] = fminf(lo[c][1], a.y); lo[c][2] = fminf(lo[c][2], a.z);
            hi[c][0] = fmaxf(hi[c][0], b.x); hi[c][1] = fmaxf(hi[c][1], b.y); hi[c][2] = fmaxf(hi[c][2], b.z);
            any[c] = true;
        }
        for (int c = 0; c < 2; c++)
            if (any[c]) s_child[c].add(lo[c], hi[c]);
    }
    __syncthreads();
    // leaves: copy their primitives to the final order
    const int cnt[2] = {nl, n - nl};
    const int st[2] = {sg.start, sg.start + nl};
    for (int c = 0; c < 2; c++)
        if (cnt[c] <= leaf_max)
            for (int j = tid; j < cnt[c]; j += NT) perm_final[st[c] + j] = perm_out[st[c] + j];
    if (tid == 0) {
        int link[2];
        const int inner = (cnt[0] > leaf_max) + (cnt[1] > leaf_max);  // one reservation for both
        int id = inner ? atomicAdd(node_ctr, inner) : 0, slot = inner ? atomicAdd(next_count, inner) : 0;
        for (int c = 0; c < 2; c++) {
            if (cnt[c] <= leaf_max) {
                link[c] = leaf_link(st[c], cnt[c]);
            } else {
                link[c] = id;
                next[slot++] = SahSeg{st[c], cnt[c], id++};
            }
        }
        float mn[2][3], mx[2][3];
        for (int c = 0; c < 2; c++)
            for (int k = 0; k < 3; k++) {
                mn[c][k] = u2f(s_child[c].mn[k]);
                mx[c][k] = u2f(s_child[c].mx[k]);
            }
        float4 *nd = nodes + 4 * sg.node;
        nd[0] = make_float4(mn[0][0], mx[0][0], mn[1][0], mx[1][0]);
        nd[1] = make_float4(mn[0][1], mx[0][1], mn[1][1], mx[1][1]);
        nd[2] = make_float4(mn[0][2], mx[0][2], mn[1][2], mx[1][2]);
        nd[3] = make_float4(__int_as_float(link[0]), __int_as_float(link[1]), 0.0f, 0.0f);
    }
}

// Root of a scene with at most leaf_max primitives: one node, both slots the
// same leaf range.
__global__ void sah_root_leaf_kernel(const float4 *bmin, const float4 *bmax, int n, int *perm_final,
                                     float4 *nodes) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int i = 0; i < n; i++) {
        perm_final[i] = i;
        float4 a = bmin[i], b = bmax[i];
        mn[0] = fminf(mn[0], a.x); mn[1] = fminf(mn[1], a.y); mn[2] = fminf(mn[2], a.z);
        mx[0] = fmaxf(mx[0], b.x); mx[1] = fmaxf(mx[1], b.y); mx[2] = fmaxf(mx[2], b.z);
    }
    int l = leaf_link(0, n);
    nodes[0] = make_float4(mn[0], mx[0], mn[0], mx[0]);
    nodes[1] = make_float4(mn[1], mx[1], mn[1], mx[1]);
    nodes[2] = make_float4(mn[2], mx[2], mn[2], mx[2]);
    nodes[3] = make_float4(__int_as_float(l), __int_as_float(l), 0.0f, 0.0f);
}

inline unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// 4-wide collapse of a binary traversal tree (either layout).  A wide node
// starts from a binary node's two children and twice replaces the internal
// child of largest surface area by its two children; every internal child
// that remains becomes a wide node of the next level.  Node: 8 float4 (one
// 128 B cache line) = {min.x}, {max.x}, {min.y}, {max.y}, {min.z}, {max.z} of
// the four children, their links (wide node index, leaf range, or 0 = empty:
// the root is nobody's child), pad.
// ---------------------------------------------------------------------------
struct WideItem {
    int bin;   // binary node
    int wide;  // its wide node
};

KRT_DEV float half_area(float4 mn, float4 mx) {
    float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
    return dx * dy + dy * dz + dz * dx;
}

__global__ void collapse_kernel(const float4 *__restrict__ bin, float4 *__restrict__ wide,
                                const WideItem *__restrict__ front_in, const int *n_in, WideItem *front_out,
                                int *n_out, int *wcount) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_in) return;
    const WideItem it = front_in[i];
    float4 mn[4], mx[4];
    int link[4];
    auto child = [&](int b, int k, float4 &a, float4 &c, int &l) {
        const float4 X = bin[4 * b], Y = bin[4 * b + 1], Z = bin[4 * b + 2], L = bin[4 * b + 3];
        a = k ? make_float4(X.z, Y.z, Z.z, 0.0f) : make_float4(X.x, Y.x, Z.x, 0.0f);
        c = k ? make_float4(X.w, Y.w, Z.w, 0.0f) : make_float4(X.y, Y.y, Z.y, 0.0f);
        l = __float_as_int(k ? L.y : L.x);
    };
    child(it.bin, 0, mn[0], mx[0], link[0]);
    child(it.bin, 1, mn[1], mx[1], link[1]);
    int nc = 2;
    for (int round = 0; round < 2; round++) {
        int pick = -1;
        float best = -1.0f;
        for (int j = 0; j < nc; j++) {
            if (link[j] <= 0) continue;  // leaf range
            const float a = half_area(mn[j], mx[j]);
            if (a > best) {
                best = a;
                pick = j;
            }
        }
        if (pick < 0) break;
        const int b = link[pick];
        child(b, 0, mn[pick], mx[pick], link[pick]);
        child(b, 1, mn[nc], mx[nc], link[nc]);
        nc++;
    }
    int ninternal = 0;
    for (int j = 0; j < nc; j++) ninternal += link[j] > 0 ? 1 : 0;
    int base = ninternal ? atomicAdd(wcount, ninternal) : 0;
    int slot = ninternal ? atomicAdd(n_out, ninternal) : 0;
    int out[4] = {0, 0, 0, 0};
    for (int j = 0; j < nc; j++) {
        if (link[j] > 0) {
            out[j] = base;
            front_out[slot++] = WideItem{link[j], base};
            base++;
        } else {
            out[j] = link[j];
        }
    }
    for (int j = nc; j < 4; j++) {
        mn[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        mx[j] = mn[j];
    }
    float4 *w = wide + 8 * (size_t)it.wide;
    w[0] = make_float4(mn[0].x, mn[1].x, mn[2].x, mn[3].x);
    w[1] = make_float4(mx[0].x, mx[1].x, mx[2].x, mx[3].x);
    w[2] = make_float4(mn[0].y, mn[1].y, mn[2].y, mn[3].y);
    w[3] = make_float4(mx[0].y, mx[1].y, mx[2].y, mx[3].y);
    w[4] = make_float4(mn[0].z, mn[1].z, mn[2].z, mn[3].z);
    w[5] = make_float4(mx[0].z, mx[1].z, mx[2].z, mx[3].z);
    w[6] = make_float4(__int_as_float(out[0]), __int_as_float(out[1]), __int_as_float(out[2]),
                       __int_as_float(out[3]));
    w[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

// Returns the number of wide levels; A.wide / A.nwide filled.
static uint32_t collapse_wide(Context &ctx, Accel &A, int nbin) {
    hipStream_t st = ctx.stream;
    A.wide.alloc(sizeof(float4) * 8 * (size_t)nbin);
    DevBuf front[2] = {DevBuf(sizeof(WideItem) * (size_t)nbin), DevBuf(sizeof(WideItem) * (size_t)nbin)};
    DevBuf counts(sizeof(int) * 3);  // n[0], n[1], wide count
    const int init[3] = {1, 0, 1};
    const WideItem root{0, 0};
    KRT_HIP(hipMemcpyAsync(front[0].p, &root, sizeof(root), hipMemcpyHostToDevice, st));
    KRT_HIP(hipMemcpyAsync(counts.p, init, sizeof(init), hipMemcpyHostToDevice, st));
    int *cnt = counts.as<int>();
    int n = 1;
    uint32_t lv = 0;
    while (n > 0) {
        KRT_HIP(hipMemsetAsync(cnt + ((lv + 1) & 1), 0, sizeof(int), st));
        collapse_kernel<<<blocks((size_t)n, 256), 256, 0, st>>>(A.nodes.as<float4>(), A.wide.as<float4>(),
                                                              front[lv & 1].as<WideItem>(), cnt + (lv & 1),
                                                              front[(lv + 1) & 1].as<WideItem>(), cnt + ((lv + 1) & 1),
                                                              cnt + 2);
        KRT_HIP(hipGetLastError());
        KRT_HIP(hipMemcpyAsync(&n, cnt + ((lv + 1) & 1), sizeof(int), hipMemcpyDeviceToHost, st));
        KRT_HIP(hipStreamSynchronize(st));
        lv++;
        KRT_REQUIRE(lv < 100000, KRT_ERR_HIP, "wide collapse did not converge");
    }
    int nw = 0;
    KRT_HIP(hipMemcpyAsync(&nw, cnt + 2, sizeof(int), hipMemcpyDeviceToHost, st));
    KRT_HIP(hipStreamSynchronize(st));
    A.nwide = (uint32_t)nw;
    return lv;
}

// The bvh.cpp tree of n primitive boxes (A.ser, A.leaf_order; rank_of = the
// preorder rank of every primitive), the traversal tree of `layout` (A.nodes)
// and its 4-wide collapse (A.wide).  order receives the primitive at every
// position of the traversal tree's leaf ranges.
// Workgroup size of a level: the top levels have few nodes with many
// primitives each, where more lanes per node hide more of the memory latency;
// the deep levels have tens of thousands of nodes of a few dozen primitives,
// where one wavefront per node keeps the per-node reductions short.
constexpr int kBuildWide = 1024;
constexpr int kBuildNarrow = 256;
constexpr int kBuildWave = 64;

static void build_trees(Context &ctx, const float4 *bmin, const float4 *bmax, uint32_t n, int leaf_max, int layout,
                        Accel &A, DevBuf &order_out, DevBuf &rank_of) {
    hipStream_t st = ctx.stream;
    // the binned-SAH tree builds on the second stream, level by level beside
    // the bvh.cpp tree: both trees' top levels occupy few CUs
    hipStream_t st2 = ctx.side ? ctx.side : ctx.stream;
    const uint32_t nser = 2 * n - 1;
    DevBuf perm[2] = {DevBuf(sizeof(int) * n), DevBuf(sizeof(int) * n)};
    DevBuf segs[2] = {DevBuf(sizeof(Seg) * (n / 3 + 2)), DevBuf(sizeof(Seg) * (n / 3 + 2))};
    DevBuf cnt(sizeof(int)), cnt2(sizeof(int));
    // boxes in permutation order, double-buffered per tree (level 0 reads
    // bmin / bmax: its permutation is the identity)
    DevBuf pbox[2][2][2];  // [tree][buffer][min / max]
    DevBuf pre_buf;        // PreBins of the binned-SAH tree's first levels
    DevBuf pre_ext_buf;    // PreExt of the bvh.cpp tree's first levels
    auto boxes = [&](int tree, uint32_t level, bool out) {
        const uint32_t b = (level + (out ? 1u : 0u)) & 1u;
        if (!out && level == 0) return PermBoxes{const_cast<float4 *>(bmin), const_cast<float4 *>(bmax)};
        DevBuf *d = pbox[tree][b];
        return PermBoxes{d[0].as<float4>(), d[1].as<float4>()};
    };
    DevBuf leafcount(sizeof(int) * nser), rank(sizeof(int) * nser);
    A.ser.alloc(sizeof(float4) * 3 * nser);
    A.leaf_order.alloc(sizeof(int) * n);
    order_out.alloc(sizeof(int) * n);
    BuildOut o{A.ser.as<float4>(), A.leaf_order.as<int>(), leafcount.as<int>(), rank.as<int>()};
    iota_kernel<<<blocks(n, 256), 256, 0, st>>>(perm[0].as<int>(), n);
    KRT_HIP(hipGetLastError());
    hipEvent_t boxes_ready, sah_done;
    KRT_HIP(hipEventCreate(&boxes_ready));
    KRT_HIP(hipEventCreate(&sah_done));
    KRT_HIP(hipEventRecord(boxes_ready, st));
    KRT_HIP(hipStreamWaitEvent(st2, boxes_ready, 0));

    const bool sah = layout != KRT_LAYOUT_REFERENCE;
    DevBuf sperm[2], ssegs[2], nctr;
    int nseg = 0, nseg2 = 0, nnodes = 0;
    uint32_t levels = 0, lv = 0, depth = 0;
    if (n <= 2) {
        root_small_kernel<<<1, 64, 0, st>>>(o, (int)n, bmin, bmax);
        KRT_HIP(hipGetLastError());
        levels = 1;
    } else {
        Seg root{0, (int)n, 0, -1, 0};
        KRT_HIP(hipMemcpyAsync(segs[0].p, &root, sizeof(Seg), hipMemcpyHostToDevice, st));
        nseg = 1;
    }
    if (sah) {
        A.nodes.alloc(sizeof(float4) * 4 * (size_t)(n + 1));
        if ((int)n <= leaf_max) {
            sah_root_leaf_kernel<<<1, 64, 0, st2>>>(bmin, bmax, (int)n, order_out.as<int>(), A.nodes.as<float4>());
            KRT_HIP(hipGetLastError());
            nnodes = 1;
            depth = 1;
        } else {
            for (int k = 0; k < 2; k++) {
                sperm[k].alloc(sizeof(int) * n);
                ssegs[k].alloc(sizeof(SahSeg) * (n / 2 + 2));
            }
            nctr.alloc(sizeof(int));
            iota_kernel<<<blocks(n, 256), 256, 0, st2>>>(sperm[0].as<int>(), n);
            KRT_HIP(hipGetLastError());
            SahSeg root{0, (int)n, 0};
            int one = 1;
            KRT_HIP(hipMemcpyAsync(ssegs[0].p, &root, sizeof(SahSeg), hipMemcpyHostToDevice, st2));
            KRT_HIP(hipMemcpyAsync(nctr.p, &one, sizeof(int), hipMemcpyHostToDevice, st2));
            nseg2 = 1;
        }
    }
    // the next level's segment counts land in pinned memory, so the copies
    // stay asynchronous and both streams' levels are in flight together
    struct Pinned {
        int *p = nullptr;
        ~Pinned() {
            if (p) (void)hipHostFree(p);
        }
    } pinned;
    KRT_HIP(hipHostMalloc((void **)&pinned.p, 2 * sizeof(int), hipHostMallocDefault));
    int &next = pinned.p[0], &next2 = pinned.p[1];
    next = next2 = 0;
    for (int tree = 0; tree < 2; tree++)
        if (tree == 0 ? nseg > 0 : nseg2 > 0)
            for (auto &b : pbox[tree])
                for (auto &d : b) d.alloc(sizeof(float4) * n);
    while (nseg > 0 || nseg2 > 0) {
        if (nseg > 0) {
            KRT_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), st));
            const Seg *sin = segs[levels & 1].as<Seg>();
            Seg *sout = segs[(levels + 1) & 1].as<Seg>();
            const int *pin = perm[levels & 1].as<int>();
            int *pout = perm[(levels + 1) & 1].as<int>();
            const PermBoxes bi = boxes(0, levels, false), bo = boxes(0, levels, true);
            const PreExt *pre = nullptr;
            if (nseg <= kPreWideSegs && n >= kPreMinPrims) {
                if (!pre_ext_buf.p) pre_ext_buf.alloc(sizeof(PreExt) * kPreWideSegs);
                pre = pre_ext_buf.as<PreExt>();
                pre_ext_clear_kernel<<<nseg, 64, 0, st>>>(pre_ext_buf.as<PreExt>());
                pre_ext_kernel<<<nseg * kPreWide, 256, 0, st>>>(sin, bi, pre_ext_buf.as<PreExt>());
                pre_split_bins_kernel<<<nseg * kPreWide, 256, 0, st>>>(sin, bi, pre_ext_buf.as<PreExt>());
            }
            if (nseg < 1024)
                level_kernel<kBuildWide><<<nseg, kBuildWide, 0, st>>>(sin, pin, pout, bmin, bmax, o, sout, cnt.as<int>(),
                                                                      bi, bo, pre);
            else if (nseg < 8192)
                level_kernel<kBuildNarrow><<<nseg, kBuildNarrow, 0, st>>>(sin, pin, pout, bmin, bmax, o, sout,
                                                                          cnt.as<int>(), bi, bo, pre);
            else
                level_kernel<kBuildWave><<<nseg, kBuildWave, 0, st>>>(sin, pin, pout, bmin, bmax, o, sout,
                                                                      cnt.as<int>(), bi, bo, pre);
            KRT_HIP(hipGetLastError());
            KRT_HIP(hipMemcpyAsync(&next, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
        }
        if (nseg2 > 0) {
            KRT_HIP(hipMemsetAsync(cnt2.p, 0, sizeof(int), st2));
            const SahSeg *sin = ssegs[lv & 1].as<SahSeg>();
            SahSeg *sout = ssegs[(lv + 1) & 1].as<SahSeg>();
            const int *pin = sperm[lv & 1].as<int>();
            int *pout = sperm[(lv + 1) & 1].as<int>();
            const PermBoxes bi = boxes(1, lv, false), bo = boxes(1, lv, true);
            // the first levels of a large tree: bounds and bins by many workgroups per node
            const PreBins *pre = nullptr;
            if (nseg2 <= kPreSegs && n >= kPreMinPrims) {
                if (!pre_buf.p) pre_buf.alloc(sizeof(PreBins));
                pre = pre_buf.as<PreBins>();
                pre_clear_kernel<<<nseg2, 64, 0, st2>>>(pre_buf.as<PreBins>());
                pre_bounds_kernel<<<nseg2 * kPreBlocks, 256, 0, st2>>>(sin, bi, pre_buf.as<PreBins>());
                pre_bins_kernel<<<nseg2 * kPreBlocks, 256, 0, st2>>>(sin, bi, pre_buf.as<PreBins>());
                KRT_HIP(hipGetLastError());
            }
            if (nseg2 < 1024)
                sah_level_kernel<kBuildWide><<<nseg2, kBuildWide, 0, st2>>>(sin, pin, pout, order_out.as<int>(), bi, bo,
                                                                           A.nodes.as<float4>(), nctr.as<int>(),
                                                                           sout, cnt2.as<int>(), leaf_max, pre);
            else if (nseg2 < 8192)
                sah_level_kernel<kBuildNarrow><<<nseg2, kBuildNarrow, 0, st2>>>(
                    sin, pin, pout, order_out.as<int>(), bi, bo, A.nodes.as<float4>(), nctr.as<int>(), sout,
                    cnt2.as<int>(), leaf_max, nullptr);
            else
                sah_level_kernel<kBuildWave><<<nseg2, kBuildWave, 0, st2>>>(
                    sin, pin, pout, order_out.as<int>(), bi, bo, A.nodes.as<float4>(), nctr.as<int>(), sout,
                    cnt2.as<int>(), leaf_max, nullptr);
            KRT_HIP(hipGetLastError());
            KRT_HIP(hipMemcpyAsync(&next2, cnt2.p, sizeof(int), hipMemcpyDeviceToHost, st2));
        }
        if (nseg > 0) {
            KRT_HIP(hipStreamSynchronize(st));
            nseg = next;
            levels++;
            KRT_REQUIRE(levels < 1000000, KRT_ERR_HIP, "BVH build did not converge");
        }
        if (nseg2 > 0) {
            KRT_HIP(hipStreamSynchronize(st2));
            nseg2 = next2;
            lv++;
            KRT_REQUIRE(lv < 100000, KRT_ERR_HIP, "SAH build did not converge");
        }
    }
    if (sah && (int)n > leaf_max) {
        KRT_HIP(hipMemcpyAsync(&nnodes, nctr.p, sizeof(int), hipMemcpyDeviceToHost, st2));
        KRT_HIP(hipStreamSynchronize(st2));
        depth = lv + 1;
    }
    KRT_HIP(hipEventRecord(sah_done, st2));
    KRT_HIP(hipStreamWaitEvent(st, sah_done, 0));
    (void)hipEventDestroy(boxes_ready);
    (void)hipEventDestroy(sah_done);

    // rank of each primitive in the bvh.cpp preorder (closest-hit tie-break)
    rank_of.alloc(sizeof(int) * n);
    invert_kernel<<<blocks(n, 256), 256, 0, st>>>(A.leaf_order.as<int>(), n, rank_of.as<int>());
    if (!sah) {
        depth = levels + 2;
        // collapse of the bvh.cpp tree itself
        DevBuf flag(sizeof(int) * nser), node_id(sizeof(int) * nser), total(sizeof(int));
        mark_kernel<<<blocks(nser, 256), 256, 0, st>>>(A.ser.as<float4>(), leafcount.as<int>(), nser, leaf_max,
                                                       flag.as<int>());
        scan_kernel<<<1, 1024, 0, st>>>(flag.as<int>(), node_id.as<int>(), nser, total.as<int>());
        KRT_HIP(hipMemcpyAsync(&nnodes, total.p, sizeof(int), hipMemcpyDeviceToHost, st));
        KRT_HIP(hipStreamSynchronize(st));
        A.nodes.alloc(sizeof(float4) * 4 * (size_t)nnodes);
        fill_nodes_kernel<<<blocks(nser, 256), 256, 0, st>>>(A.ser.as<float4>(), leafcount.as<int>(),
                                                             rank.as<int>(), flag.as<int>(), node_id.as<int>(), nser,
                                                             leaf_max, A.nodes.as<float4>());
        KRT_HIP(hipMemcpyAsync(order_out.p, A.leaf_order.p, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    }
    // each wide level pushes at most three siblings
    const uint32_t wide_levels = collapse_wide(ctx, A, nnodes);
    A.nt = n;
    A.ser_depth_off.clear();  // a new tree: the refit schedule is made again
    A.ser_by_depth.release();
    A.nser = nser;
    A.nnodes = (uint32_t)nnodes;
    A.levels = levels;
    A.depth_bound = std::max(depth, 3 * wide_levels + 1);
    A.bin_depth = depth;
    A.layout = layout;
}

constexpr bool kTwoLevelEnabled = false;

// Host side of the two-level build.  Row-major 3x4 affine transform of a
// point with glm's pairing of mat4 x vec4, (c0 x + c1 y) + (c2 z + c3 w).
static void xform_point_host(const float *M, const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++)
        out[r] = (M[4 * r] * p[0] + M[4 * r + 1] * p[1]) + (M[4 * r + 2] * p[2] + M[4 * r + 3] * 1.0f);
}

// Inverse of a row-major 3x4 affine transform: cofactors of the 3x3 part in
// FP64, translation -R^-1 t, rounded to float once.
void affine_inverse(const float *M, float *Minv) {
    const double a = M[0], b = M[1], c = M[2], d = M[4], e = M[5], f = M[6], g = M[8], h = M[9], k = M[10];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    KRT_REQUIRE(std::isfinite(det) && det != 0.0, KRT_ERR_INVALID, "instance transform is singular");
    const double r[9] = {A / det, -(b * k - c * h) / det, (b * f - c * e) / det,
                         B / det, (a * k - c * g) / det, -(a * f - c * d) / det,
                         C / det, -(a * h - b * g) / det, (a * e - b * d) / det};
    const double t[3] = {M[3], M[7], M[11]};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) Minv[4 * i + j] = (float)r[3 * i + j];
        Minv[4 * i + 3] = (float)(-(r[3 * i] * t[0] + r[3 * i + 1] * t[1] + r[3 * i + 2] * t[2]));
    }
}

// World box of an instance: the 8 corners of its mesh's root box through M,
// widened by 1e-4 of the largest extent + 1e-5 of the largest coordinate so
// that rounding of the ray transforms can never cull a hit.
static void instance_box(const float *root_min, const float *root_max, const float *M, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; a++) {
        lo[a] = FLT_MAX;
        hi[a] = -FLT_MAX;
    }
    for (int k = 0; k < 8; k++) {
        const float p[3] = {(k & 1) ? root_max[0] : root_min[0], (k & 2) ? root_max[1] : root_min[1],
                            (k & 4) ? root_max[2] : root_min[2]};
        float w[3];
        xform_point_host(M, p, w);
        for (int a = 0; a < 3; a++) {
            lo[a] = (w[a] < lo[a]) ? w[a] : lo[a];
            hi[a] = (hi[a] < w[a]) ? w[a] : hi[a];
        }
    }
    float ext = 0.0f, mag = 0.0f;
    for (int a = 0; a < 3; a++) {
        const float e = hi[a] - lo[a];
        ext = (ext < e) ? e : ext;
        const float m = (std::fabs(lo[a]) < std::fabs(hi[a])) ? std::fabs(hi[a]) : std::fabs(lo[a]);
        mag = (mag < m) ? m : mag;
    }
    const float pad = 1e-4f * ext + 1e-5f * mag;
    for (int a = 0; a < 3; a++) {
        lo[a] = lo[a] - pad;
        hi[a] = hi[a] + pad;
    }
}

// Two-level build (Accelerator::update + build_tlas,
// include/amadeus/accelerator.cuh:145-277): one bvh.cpp tree, traversal tree
// and 4-wide collapse per base mesh, back to back in s.accel.wide / tri; then
// the same build over the instances' world boxes with one instance per leaf.
static void accel_build_instanced(Context &ctx, Scene &s, int leaf_max, int layout) {
    hipStream_t st = ctx.stream;
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    KRT_HIP(hipEventRecord(e0, st));
    std::vector<Accel> blas(s.nmesh);
    std::vector<uint32_t> wide_base(s.nmesh), ser_base(s.nmesh);
    uint32_t nwide = 0, nser = 0, depth = 0, levels = 0, nbin = 0;
    for (uint32_t m = 0; m < s.nmesh; m++) {
        const uint32_t n = s.mesh_nt[m];
        const float *pos = s.pos.as<float>() + 3 * (size_t)s.mesh_vbase[m];
        const uint32_t *idx = s.idx.as<uint32_t>() + 3 * (size_t)s.mesh_tbase[m];
        DevBuf bmin(sizeof(float4) * n), bmax(sizeof(float4) * n);
        tri_boxes_kernel<<<blocks(n, 256), 256, 0, st>>>(pos, idx, n, bmin.as<float4>(), bmax.as<float4>());
        KRT_HIP(hipGetLastError());
        DevBuf order, rank_of;
        build_trees(ctx, bmin.as<float4>(), bmax.as<float4>(), n, leaf_max, layout, blas[m], order, rank_of);
        blas[m].tri.alloc(sizeof(float4) * 3 * (size_t)n);
        tri_pack_kernel<<<blocks(n, 256), 256, 0, st>>>(pos, idx, order.as<int>(), rank_of.as<int>(), n,
                                                        blas[m].tri.as<float4>());
        KRT_HIP(hipGetLastError());
        wide_base[m] = nwide;
        ser_base[m] = nser;
        nwide += blas[m].nwide;
        nser += blas[m].nser;
        depth = std::max(depth, blas[m].depth_bound);
        levels = std::max(levels, blas[m].levels);
        nbin += blas[m].nnodes;
    }
    Accel &A = s.accel;
    A.wide.alloc(sizeof(float4) * 8 * (size_t)nwide);
    A.tri.alloc(sizeof(float4) * 3 * (size_t)s.mesh_tbase.back() + sizeof(float4) * 3 * (size_t)s.mesh_nt.back());
    s.mesh_ser.alloc(sizeof(float4) * 3 * (size_t)nser);
    std::vector<float> roots(12 * (size_t)s.nmesh);
    for (uint32_t m = 0; m < s.nmesh; m++) {
        KRT_HIP(hipMemcpyAsync(A.wide.as<float4>() + 8 * (size_t)wide_base[m], blas[m].wide.p,
                               sizeof(float4) * 8 * (size_t)blas[m].nwide, hipMemcpyDeviceToDevice, st));
        KRT_HIP(hipMemcpyAsync(A.tri.as<float4>() + 3 * (size_t)s.mesh_tbase[m], blas[m].tri.p,
                               sizeof(float4) * 3 * (size_t)s.mesh_nt[m], hipMemcpyDeviceToDevice, st));
        KRT_HIP(hipMemcpyAsync(s.mesh_ser.as<float4>() + 3 * (size_t)ser_base[m], blas[m].ser.p,
                               sizeof(float4) * 3 * (size_t)blas[m].nser, hipMemcpyDeviceToDevice, st));
        KRT_HIP(hipMemcpyAsync(&roots[12 * m], blas[m].ser.p, sizeof(float) * 12, hipMemcpyDeviceToHost, st));
    }
    KRT_HIP(hipStreamSynchronize(st));
    s.mesh_ser_base = ser_base;
    // instance tree over the padded world boxes, one instance per leaf
    const uint32_t ni = s.ninst;
    std::vector<float4> ilo(ni), ihi(ni);
    std::vector<uint32_t> prim_base(ni);
    uint32_t base = 0;
    for (uint32_t i = 0; i < ni; i++) {
        const uint32_t m = (uint32_t)s.inst_mesh[i];
        float lo[3], hi[3];
        instance_box(&roots[12 * m + 4], &roots[12 * m + 8], &s.xform[12 * (size_t)i], lo, hi);
        ilo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        ihi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        prim_base[i] = base;
        base += s.mesh_nt[m];
    }
    DevBuf dlo(sizeof(float4) * ni), dhi(sizeof(float4) * ni);
    KRT_HIP(hipMemcpyAsync(dlo.p, ilo.data(), sizeof(float4) * ni, hipMemcpyHostToDevice, st));
    KRT_HIP(hipMemcpyAsync(dhi.p, ihi.data(), sizeof(float4) * ni, hipMemcpyHostToDevice, st));
    DevBuf iorder, irank;
    build_trees(ctx, dlo.as<float4>(), dhi.as<float4>(), ni, 1, layout, s.tlas, iorder, irank);
    std::vector<int> order(ni), rank(ni);
    KRT_HIP(hipMemcpyAsync(order.data(), iorder.p, sizeof(int) * ni, hipMemcpyDeviceToHost, st));
    KRT_HIP(hipMemcpyAsync(rank.data(), irank.p, sizeof(int) * ni, hipMemcpyDeviceToHost, st));
    KRT_HIP(hipStreamSynchronize(st));
    // instance records in the instance tree's traversal order
    std::vector<float> rec(32 * (size_t)ni, 0.0f);
    for (uint32_t p = 0; p < ni; p++) {
        const int i = order[p];
        const uint32_t m = (uint32_t)s.inst_mesh[i];
        float *r = &rec[32 * (size_t)p];
        affine_inverse(&s.xform[12 * (size_t)i], r);
        std::memcpy(r + 12, &s.xform[12 * (size_t)i], sizeof(float) * 12);
        const int32_t meta[8] = {(int32_t)wide_base[m], (int32_t)s.mesh_tbase[m], (int32_t)prim_base[i], rank[i],
                                 (int32_t)s.mesh_vbase[m], (int32_t)s.mesh_tbase[m], (int32_t)m, 0};
        std::memcpy(r + 24, meta, sizeof(meta));
    }
    s.inst_rec.alloc(sizeof(float) * rec.size());
    KRT_HIP(hipMemcpyAsync(s.inst_rec.p, rec.data(), sizeof(float) * rec.size(), hipMemcpyHostToDevice, st));
    KRT_HIP(hipEventRecord(e1, st));
    KRT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    KRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    A.built = true;
    A.nt = s.nt;
    A.nser = nser;
    A.nnodes = nbin;
    A.nwide = nwide;
    A.levels = std::max(levels, s.tlas.levels);
    A.depth_bound = depth + s.tlas.depth_bound;
    A.layout = layout;
    A.build_ms = ms;
    s.tlas.built = true;
    s.version++;
}

// ---------------------------------------------------------------------------
// Refit (moved geometry, same topology): Accelerator::update rebuilds the
// acceleration structure when an entity moves (accelerator.cuh:145-196,
// armada.cu:575-602, 716-727); here the trees keep their shape and their
// boxes are recomputed bottom-up from the new vertex positions: the bvh.cpp
// tree one depth per launch (its million-node chains would otherwise pay a
// device-scope fence per step), the traversal tree by arrival counters -- a
// node is computed by the last of its children to finish: leaves start,
// each finished node counts itself at its parent, and the thread whose
// arrival completes a parent goes on with it.  Boxes are
// unions (min / max) of triangle boxes, so a refit node's box is exactly the
// one a build of the same shape would give.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 vload(const float4 *p) {  // bypasses the (non-coherent) vector L1
    const volatile float *v = reinterpret_cast<const volatile float *>(p);
    return make_float4(v[0], v[1], v[2], v[3]);
}

__global__ void ser_parent_kernel(const float4 *ser, uint32_t nser, int *parent) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nser) return;
    if (p == 0) parent[0] = -1;
    if (ser[3 * p].x != 0.0f) return;  // leaf
    const int c0 = (int)p + 1, c1 = __float_as_int(ser[3 * c0].w) / 3;  // left child's miss link = right child
    parent[c0] = (int)p;
    parent[c1] = (int)p;
}

// bvh.cpp layout: leaf box = its primitive's box, internal = union of both children

// Level-synchronous refit of the bvh.cpp tree: the depth of every node from
// the parent links (-1 for leaves), once per built tree; then per refit the
// leaves, and the internal nodes one depth per launch from the deepest up,
// each the union of its two children (finished by the launch before).  No
// device-scope fences between the steps: the launches order them.
__global__ void ser_depth_kernel(const float4 *ser, const int *parent, uint32_t nser, int *depth) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nser) return;
    int d = 0;
    for (int p = parent[i]; p >= 0; p = parent[p]) d++;
    depth[i] = ser[3 * i].x != 0.0f ? -1 : d;
}

// Counting sort of the internal nodes by depth on the device: a histogram
// (per-workgroup in LDS, then added), the offsets, and the scatter (any
// order within a depth: its nodes are independent).
constexpr int kMaxSerDepth = 1024;
__global__ void depth_hist_kernel(const int *depth, uint32_t nser, uint32_t *hist) {
    __shared__ uint32_t h[kMaxSerDepth];
    for (int k = threadIdx.x; k < kMaxSerDepth; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nser && depth[i] >= 0) atomicAdd(&h[depth[i]], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxSerDepth; k += blockDim.x)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}

__global__ void depth_offsets_kernel(const uint32_t *hist, uint32_t *off, uint32_t *fill) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t run = 0;
    for (int d = 0; d < kMaxSerDepth; d++) {
        off[d] = fill[d] = run;
        run += hist[d];
    }
    off[kMaxSerDepth] = run;
}

__global__ void depth_scatter_kernel(const int *depth, uint32_t nser, uint32_t *fill, int *list) {
    // a workgroup reserves one range per depth it holds, then fills it
    __shared__ uint32_t cnt[kMaxSerDepth], base[kMaxSerDepth];
    for (int k = threadIdx.x; k < kMaxSerDepth; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int d = i < nser ? depth[i] : -1;
    const uint32_t local = d >= 0 ? atomicAdd(&cnt[d], 1u) : 0u;
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxSerDepth; k += blockDim.x)
        if (cnt[k]) base[k] = atomicAdd(&fill[k], cnt[k]);
    __syncthreads();
    if (d >= 0) list[base[d] + local] = (int)i;
}

__global__ void ser_leaf_kernel(float4 *ser, uint32_t nser, const float4 *bmin, const float4 *bmax) {
    const uint32_t leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= nser || ser[3 * leaf].x == 0.0f) return;
    const int prim = __float_as_int(ser[3 * leaf].y);
    ser[3 * leaf + 1] = make_float4(bmin[prim].x, bmin[prim].y, bmin[prim].z, 0.0f);
    ser[3 * leaf + 2] = make_float4(bmax[prim].x, bmax[prim].y, bmax[prim].z, 0.0f);
}

__global__ void ser_level_refit_kernel(float4 *ser, const int *nodes, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int p = nodes[k];
    const int c0 = p + 1, c1 = __float_as_int(ser[3 * c0].w) / 3;
    const float4 a0 = ser[3 * c0 + 1], a1 = ser[3 * c0 + 2];
    const float4 b0 = ser[3 * c1 + 1], b1 = ser[3 * c1 + 2];
    ser[3 * p + 1] = make_float4(gmin(a0.x, b0.x), gmin(a0.y, b0.y), gmin(a0.z, b0.z), 0.0f);
    ser[3 * p + 2] = make_float4(gmax(a1.x, b1.x), gmax(a1.y, b1.y), gmax(a1.z, b1.z), 0.0f);
}

// Binary traversal tree (either layout): links are node ids (> 0) or leaf
// ranges of traversal positions; a node's arrival target is its number of
// internal children.
__global__ void bin_parent_kernel(const float4 *nodes, uint32_t nn, int *parent, int *need) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    if (i == 0) parent[0] = -1;
    const float4 L = nodes[4 * i + 3];
    const int l0 = __float_as_int(L.x), l1 = __float_as_int(L.y);
    need[i] = (l0 > 0 ? 1 : 0) + (l1 > 0 ? 1 : 0);
    if (l0 > 0) parent[l0] = (int)i;
    if (l1 > 0) parent[l1] = (int)i;
}

__device__ void range_box(const float4 *tri, const float4 *bmin, const float4 *bmax, int link, float4 &mn,
                          float4 &mx) {
    const int first = link & 0x03FFFFFF, cnt = ((link >> 26) & 31) + 1;
    mn = make_float4(3.4e38f, 3.4e38f, 3.4e38f, 0.0f);
    mx = make_float4(-3.4e38f, -3.4e38f, -3.4e38f, 0.0f);
    for (int r = first; r < first + cnt; r++) {
        const int prim = __float_as_int(tri[3 * r].w);
        const float4 a = bmin[prim], b = bmax[prim];
        mn = make_float4(fminf(mn.x, a.x), fminf(mn.y, a.y), fminf(mn.z, a.z), 0.0f);
        mx = make_float4(fmaxf(mx.x, b.x), fmaxf(mx.y, b.y), fmaxf(mx.z, b.z), 0.0f);
    }
}

__device__ void bin_node_refit(float4 *nodes, int i, const float4 *tri, const float4 *bmin, const float4 *bmax) {
    const float4 L = nodes[4 * i + 3];
    const int link[2] = {__float_as_int(L.x), __float_as_int(L.y)};
    float4 mn[2], mx[2];
    for (int k = 0; k < 2; k++) {
        if (link[k] > 0) {  // union of the child node's two slots
            const int c = link[k];
            const float4 X = vload(nodes + 4 * c), Y = vload(nodes + 4 * c + 1), Z = vload(nodes + 4 * c + 2);
            mn[k] = make_float4(fminf(X.x, X.z), fminf(Y.x, Y.z), fminf(Z.x, Z.z), 0.0f);
            mx[k] = make_float4(fmaxf(X.y, X.w), fmaxf(Y.y, Y.w), fmaxf(Z.y, Z.w), 0.0f);
        } else {
            range_box(tri, bmin, bmax, link[k], mn[k], mx[k]);
        }
    }
    nodes[4 * i + 0] = make_float4(mn[0].x, mx[0].x, mn[1].x, mx[1].x);
    nodes[4 * i + 1] = make_float4(mn[0].y, mx[0].y, mn[1].y, mx[1].y);
    nodes[4 * i + 2] = make_float4(mn[0].z, mx[0].z, mn[1].z, mx[1].z);
}

__global__ void bin_refit_kernel(float4 *nodes, uint32_t nn, const int *parent, const int *need, int *arrivals,
                                 const float4 *tri, const float4 *bmin, const float4 *bmax) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 >= nn || need[i0] != 0) return;  // start at nodes whose children are all leaf ranges
    int i = (int)i0;
    for (;;) {
        bin_node_refit(nodes, i, tri, bmin, bmax);
        const int p = parent[i];
        if (p < 0) return;
        __threadfence();
        if (atomicAdd(&arrivals[p], 1) + 1 < need[p]) return;  // another child finishes the parent
        i = p;
    }
}

// The triangle records in traversal order keep their primitive and rank
__global__ void tri_repack_kernel(const float *pos, const uint32_t *idx, uint32_t nt, float4 *tri) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    const int prim = __float_as_int(tri[3 * r].w);
    const f3 v1 = ld3(pos + 3 * idx[3 * prim]), v2 = ld3(pos + 3 * idx[3 * prim + 1]),
             v3 = ld3(pos + 3 * idx[3 * prim + 2]);
    const f3 e1 = v2 - v1, e2 = v3 - v1;
    tri[3 * r] = make_float4(v1.x, v1.y, v1.z, __int_as_float(prim));
    tri[3 * r + 1] = make_float4(e1.x, e1.y, e1.z, tri[3 * r + 1].w);
    tri[3 * r + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
}

// Triangle light rows of the moved geometry (the expressions of the upload's
// light_row: armada.cu:117-140 with lighting.cuh:102-110)
__global__ void light_rows_kernel(const float *pos, const uint32_t *idx, const uint32_t *tri_mat, const float4 *mats,
                                  const int *light_tri, uint32_t nl, float4 *lights) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl) return;
    const int t = light_tri[l];
    const float *a = pos + 3 * idx[3 * t], *b = pos + 3 * idx[3 * t + 1], *c = pos + 3 * idx[3 * t + 2];
    const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
    const float dd = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
    const float area = sqrtf(dd) * 0.5f;
    const float inv = 1.0f / sqrtf(dd);
    const float4 e = mats[4 * tri_mat[t] + 2];
    lights[4 * l + 0] = make_float4(a[0], a[1], a[2], area);
    lights[4 * l + 1] = make_float4(ab[0], ab[1], ab[2], cr[0] * inv);
    lights[4 * l + 2] = make_float4(ac[0], ac[1], ac[2], cr[1] * inv);
    lights[4 * l + 3] = make_float4(e.x, e.y, e.z, cr[2] * inv);
}

void accel_refit(Context &ctx, Scene &s) {
    KRT_REQUIRE(!s.instanced, KRT_ERR_STATE, "refit takes flat scenes (krt_scene_upload)");
    KRT_REQUIRE(s.accel.built && s.accel.nser > 0, KRT_ERR_STATE, "scene has no acceleration structure to refit");
    hipStream_t st = ctx.stream;
    Accel &A = s.accel;
    const uint32_t n = s.nt;
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    KRT_HIP(hipEventRecord(e0, st));
    DevBuf bmin(sizeof(float4) * n), bmax(sizeof(float4) * n);
    tri_boxes_kernel<<<blocks(n, 256), 256, 0, st>>>(s.pos.as<float>(), s.idx.as<uint32_t>(), n, bmin.as<float4>(),
                                                     bmax.as<float4>());
    KRT_HIP(hipGetLastError());
    // bvh.cpp tree: its schedule (internal nodes by depth) once per build
    if (A.ser_depth_off.empty()) {
        DevBuf parent(sizeof(int) * A.nser), depth(sizeof(int) * A.nser);
        ser_parent_kernel<<<blocks(A.nser, 256), 256, 0, st>>>(A.ser.as<float4>(), A.nser, parent.as<int>());
        KRT_HIP(hipGetLastError());
        ser_depth_kernel<<<blocks(A.nser, 256), 256, 0, st>>>(A.ser.as<float4>(), parent.as<int>(), A.nser,
                                                              depth.as<int>());
        KRT_HIP(hipGetLastError());
        KRT_REQUIRE(A.levels + 2 < (uint32_t)kMaxSerDepth, KRT_ERR_STATE,
                    "bvh.cpp tree deeper than 1024 levels: rebuild (krt_accel_build) instead of refit");
        DevBuf hist(sizeof(uint32_t) * kMaxSerDepth), off(sizeof(uint32_t) * (kMaxSerDepth + 1)),
            fill(sizeof(uint32_t) * kMaxSerDepth);
        KRT_HIP(hipMemsetAsync(hist.p, 0, hist.bytes, st));
        depth_hist_kernel<<<blocks(A.nser, 256), 256, 0, st>>>(depth.as<int>(), A.nser, hist.as<uint32_t>());
        depth_offsets_kernel<<<1, 1, 0, st>>>(hist.as<uint32_t>(), off.as<uint32_t>(), fill.as<uint32_t>());
        A.ser_by_depth.alloc(sizeof(int) * A.nser);
        depth_scatter_kernel<<<blocks(A.nser, 256), 256, 0, st>>>(depth.as<int>(), A.nser, fill.as<uint32_t>(),
                                                                  A.ser_by_depth.as<int>());
        KRT_HIP(hipGetLastError());
        std::vector<uint32_t> oh(kMaxSerDepth + 1);
        KRT_HIP(hipMemcpyAsync(oh.data(), off.p, sizeof(uint32_t) * oh.size(), hipMemcpyDeviceToHost, st));
        KRT_HIP(hipStreamSynchronize(st));
        size_t used = 1;
        while (used < (size_t)kMaxSerDepth && oh[used] < oh[kMaxSerDepth]) used++;
        A.ser_depth_off.assign(oh.begin(), oh.begin() + (long)used + 1);  // depths 0 .. used-1
    }
    ser_leaf_kernel<<<blocks(A.nser, 256), 256, 0, st>>>(A.ser.as<float4>(), A.nser, bmin.as<float4>(),
                                                         bmax.as<float4>());
    KRT_HIP(hipGetLastError());
    for (size_t d = A.ser_depth_off.size() - 1; d-- > 0;) {
        const uint32_t first = A.ser_depth_off[d], count = A.ser_depth_off[d + 1] - first;
        if (!count) continue;
        ser_level_refit_kernel<<<blocks(count, 256), 256, 0, st>>>(A.ser.as<float4>(),
                                                                   A.ser_by_depth.as<int>() + first, count);
        KRT_HIP(hipGetLastError());
    }
    // triangle records, then the binary traversal tree over them, then its 4-wide collapse
    tri_repack_kernel<<<blocks(n, 256), 256, 0, st>>>(s.pos.as<float>(), s.idx.as<uint32_t>(), n, A.tri.as<float4>());
    KRT_HIP(hipGetLastError());
    DevBuf bparent(sizeof(int) * A.nnodes), need(sizeof(int) * A.nnodes), barr(sizeof(int) * A.nnodes);
    KRT_HIP(hipMemsetAsync(barr.p, 0, barr.bytes, st));
    bin_parent_kernel<<<blocks(A.nnodes, 256), 256, 0, st>>>(A.nodes.as<float4>(), A.nnodes, bparent.as<int>(),
                                                            need.as<int>());
    KRT_HIP(hipGetLastError());
    bin_refit_kernel<<<blocks(A.nnodes, 256), 256, 0, st>>>(A.nodes.as<float4>(), A.nnodes, bparent.as<int>(),
                                                           need.as<int>(), barr.as<int>(), A.tri.as<float4>(),
                                                           bmin.as<float4>(), bmax.as<float4>());
    KRT_HIP(hipGetLastError());
    const uint32_t wide_levels = collapse_wide(ctx, A, (int)A.nnodes);
    A.depth_bound = std::max(A.bin_depth, 3 * wide_levels + 1);
    if (s.nl > 0) {
        light_rows_kernel<<<blocks(s.nl, 256), 256, 0, st>>>(s.pos.as<float>(), s.idx.as<uint32_t>(),
                                                            s.tri_mat.as<uint32_t>(), s.mats.as<float4>(),
                                                            s.light_tri.as<int>(), s.nl, s.lights.as<float4>());
        KRT_HIP(hipGetLastError());
    }
    KRT_HIP(hipEventRecord(e1, st));
    KRT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    KRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    A.refit_ms = ms;
    s.stale = false;
    s.version++;
}

void accel_build(Context &ctx, Scene &s, int leaf_max, int layout) {
    KRT_REQUIRE(leaf_max >= 1 && leaf_max <= 32, KRT_ERR_INVALID, "leaf_max must be in [1, 32]");
    s.stale = false;
    KRT_REQUIRE(layout == KRT_LAYOUT_SAH || layout == KRT_LAYOUT_REFERENCE, KRT_ERR_INVALID, "unknown BVH layout");
    KRT_REQUIRE(s.nt >= 1, KRT_ERR_STATE, "scene has no triangles");
    KRT_REQUIRE(s.nt < (1u << 26), KRT_ERR_INVALID, "too many triangles for the leaf link encoding");
    if (s.instanced) {
        // The two-level path is switched off: its first GPU test run left the
        // device with a memory fault whose cause has not been found (its
        // traversal and frame results matched the oracle bit for bit up to
        // that point).  Scenes can still be uploaded and flattened by the
        // caller; the oracle keeps the two-level restatement.
        KRT_REQUIRE(kTwoLevelEnabled, KRT_ERR_STATE,
                    "two-level acceleration structures are disabled in this build; upload the flattened scene");
        accel_build_instanced(ctx, s, leaf_max, layout);
        return;
    }
    hipStream_t st = ctx.stream;
    Accel &A = s.accel;
    const uint32_t n = s.nt;
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    KRT_HIP(hipEventRecord(e0, st));
    DevBuf bmin(sizeof(float4) * n), bmax(sizeof(float4) * n);
    tri_boxes_kernel<<<blocks(n, 256), 256, 0, st>>>(s.pos.as<float>(), s.idx.as<uint32_t>(), n, bmin.as<float4>(),
                                                     bmax.as<float4>());
    KRT_HIP(hipGetLastError());
    DevBuf order, rank_of;
    build_trees(ctx, bmin.as<float4>(), bmax.as<float4>(), n, leaf_max, layout, A, order, rank_of);
    A.tri.alloc(sizeof(float4) * 3 * (size_t)n);
    tri_pack_kernel<<<blocks(n, 256), 256, 0, st>>>(s.pos.as<float>(), s.idx.as<uint32_t>(), order.as<int>(),
                                                    rank_of.as<int>(), n, A.tri.as<float4>());
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(e1, st));
    KRT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    KRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    A.built = true;
    A.build_ms = ms;
    s.version++;
}

}  // namespace krt
