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

#include <cfloat>
#include <cstring>

#include "krt_device.h"
#include "krt_internal.h"

namespace krt {

namespace {

constexpr int BB = 256;  // build workgroup
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

// One workgroup per pending node (count >= 3).
__global__ __launch_bounds__(BB) void level_kernel(const Seg *segs, const int *perm_in, int *perm_out,
                                                     const float4 *bmin, const float4 *bmax, BuildOut o, Seg *next,
                                                     int *next_count) {
    const Seg sg = segs[blockIdx.x];
    const int tid = threadIdx.x;
    const int n = sg.count;
    const int chunk = (n + BB - 1) / BB;
    const int j0 = min(tid * chunk, n), j1 = min(j0 + chunk, n);

    __shared__ float s_ext[BB], s_lo[BB], s_hi[BB];
    __shared__ int s_axis[BB], s_has[BB];
    __shared__ float s_umn[3][BB], s_umx[3][BB];

    // ---- pass 1: largest single extent (ordered) and ordered union box
    float best = 0.0f, lo = FLT_MAX, hi = -FLT_MAX;
    int axis = 0;
    float umn[3] = {0, 0, 0}, umx[3] = {0, 0, 0};
    for (int j = j0; j < j1; j++) {
        int prim = perm_in[sg.start + j];
        float4 mn = bmin[prim], mx = bmax[prim];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float e = fabsf(comp(mx, a) - comp(mn, a));
            if (e > best) {
                best = e;
                lo = comp(mn, a);
                hi = comp(mx, a);
                axis = a;
            }
        }
        if (j == j0) {
            umn[0] = mn.x; umn[1] = mn.y; umn[2] = mn.z;
            umx[0] = mx.x; umx[1] = mx.y; umx[2] = mx.z;
        } else {
            umn[0] = gmin(umn[0], mn.x); umn[1] = gmin(umn[1], mn.y); umn[2] = gmin(umn[2], mn.z);
            umx[0] = gmax(umx[0], mx.x); umx[1] = gmax(umx[1], mx.y); umx[2] = gmax(umx[2], mx.z);
        }
    }
    s_ext[tid] = best;
    s_lo[tid] = lo;
    s_hi[tid] = hi;
    s_axis[tid] = axis;
    s_has[tid] = j1 > j0;
    for (int a = 0; a < 3; a++) {
        s_umn[a][tid] = umn[a];
        s_umx[a][tid] = umx[a];
    }
    __syncthreads();
    for (int s = 1; s < BB; s <<= 1) {
        if ((tid % (2 * s)) == 0 && tid + s < BB) {
            int r = tid + s;
            if (s_ext[r] > s_ext[tid]) {  // later wins only if strictly larger
                s_ext[tid] = s_ext[r];
                s_lo[tid] = s_lo[r];
                s_hi[tid] = s_hi[r];
                s_axis[tid] = s_axis[r];
            }
            if (s_has[r]) {
                if (s_has[tid]) {
                    for (int a = 0; a < 3; a++) {
                        s_umn[a][tid] = gmin(s_umn[a][tid], s_umn[a][r]);
                        s_umx[a][tid] = gmax(s_umx[a][tid], s_umx[a][r]);
                    }
                } else {
                    for (int a = 0; a < 3; a++) {
                        s_umn[a][tid] = s_umn[a][r];
                        s_umx[a][tid] = s_umx[a][r];
                    }
                    s_has[tid] = 1;
                }
            }
        }
        __syncthreads();
    }
    axis = s_axis[0];
    lo = s_lo[0];
    hi = s_hi[0];
    const float4 nmn = make_float4(s_umn[0][0], s_umn[1][0], s_umn[2][0], 0.0f);
    const float4 nmx = make_float4(s_umx[0][0], s_umx[1][0], s_umx[2][0], 0.0f);

    // ---- pass 2: centroid bins over the 10 planes
    float split[10];
#pragma unroll
    for (int i = 0; i < 10; i++) split[i] = (hi - lo) / 10.0f * (float)i + lo;
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
    for (int j = j0; j < j1; j++) {
        int prim = perm_in[sg.start + j];
        float4 mn = bmin[prim], mx = bmax[prim];
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
    // wave reduction, then across the 4 waves through LDS
    __shared__ int w_c[BB / 64][NBIN];
    __shared__ float w_mn[BB / 64][NBIN][3], w_mx[BB / 64][NBIN][3];
    const int lane = tid & 63, wv = tid >> 6;
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
    __shared__ int s_nl, s_fallback;
    __shared__ float s_split;
    __syncthreads();
    if (tid == 0) {
        int cnt[NBIN];
        float mn[NBIN][3], mx[NBIN][3];
        for (int b = 0; b < NBIN; b++) {
            cnt[b] = 0;
            for (int a = 0; a < 3; a++) {
                mn[b][a] = FLT_MAX;
                mx[b][a] = -FLT_MAX;
            }
            for (int w = 0; w < BB / 64; w++) {
                cnt[b] += w_c[w][b];
                for (int a = 0; a < 3; a++) {
                    mn[b][a] = fminf(mn[b][a], w_mn[w][b][a]);
                    mx[b][a] = fmaxf(mx[b][a], w_mx[w][b][a]);
                }
            }
        }
        float sa_total = surface_area(nmn.x, nmn.y, nmn.z, nmx.x, nmx.y, nmx.z);
        float min_cost = FLT_MAX, min_split = 0.0f;
        int best_nl = 0;
        for (int i = 0; i < 10; i++) {
            int nl = 0, nr = 0;
            float lmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            float rmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (int b = 0; b < NBIN; b++) {
                if (b <= i) {
                    nl += cnt[b];
                    for (int a = 0; a < 3; a++) {
                        lmn[a] = fminf(lmn[a], mn[b][a]);
                        lmx[a] = fmaxf(lmx[a], mx[b][a]);
                    }
                } else {
                    nr += cnt[b];
                    for (int a = 0; a < 3; a++) {
                        rmn[a] = fminf(rmn[a], mn[b][a]);
                        rmx[a] = fmaxf(rmx[a], mx[b][a]);
                    }
                }
            }
            float cost;
            if (nl == 0 || nr == 0) cost = FLT_MAX;
            else
                cost = 1.0f + ((float)nl * surface_area(lmn[0], lmn[1], lmn[2], lmx[0], lmx[1], lmx[2]) +
                               (float)nr * surface_area(rmn[0], rmn[1], rmn[2], rmx[0], rmx[1], rmx[2])) /
                                  sa_total;
            if (cost < min_cost) {
                min_cost = cost;
                min_split = split[i];
                best_nl = nl;
            }
        }
        s_fallback = (min_cost == FLT_MAX);
        s_split = min_split;
        s_nl = s_fallback ? (n + 1) / 2 : best_nl;
    }
    __syncthreads();
    const bool fallback = s_fallback;
    const float msplit = s_split;
    const int nl = s_nl;

    // ---- pass 3: stable partition into perm_out
    __shared__ int s_wl[BB / 64];
    int lbase = 0, rbase = 0;
    for (int t0 = 0; t0 < n; t0 += BB) {
        int j = t0 + tid;
        bool valid = j < n;
        int prim = valid ? perm_in[sg.start + j] : 0;
        bool left = false;
        if (valid) {
            if (fallback) left = (j % 2) == 0;
            else {
                float c = (comp(bmin[prim], axis) + comp(bmax[prim], axis)) / 2.0f;
                left = c < msplit;
            }
        }
        unsigned long long ball = __ballot(left);
        int lpre = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) s_wl[wv] = __popcll(ball);
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < BB / 64; w++) {
            if (w < wv) woff += s_wl[w];
            tot += s_wl[w];
        }
        int tile_valid = min(BB, n - t0);
        if (valid) {
            int lrank = woff + lpre;  // left items before me in the tile
            if (left) perm_out[sg.start + lbase + lrank] = prim;
            else perm_out[sg.start + nl + rbase + (tid - lrank)] = prim;
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
        for (int k = 0; k < 2; k++) {
            const Seg &c = ch[k];
            if (c.count >= 3) {
                int slot = atomicAdd(next_count, 1);
                next[slot] = c;
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

__global__ void tri_pack_kernel(const float *pos, const uint32_t *idx, const int *leaf_order, uint32_t nt,
                                float4 *tri) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    int prim = leaf_order[r];
    f3 v1 = ld3(pos + 3 * idx[3 * prim]), v2 = ld3(pos + 3 * idx[3 * prim + 1]), v3 = ld3(pos + 3 * idx[3 * prim + 2]);
    f3 e1 = v2 - v1, e2 = v3 - v1;  // as intersect.glsl:16-17 computes them
    tri[3 * r] = make_float4(v1.x, v1.y, v1.z, __int_as_float(prim));
    tri[3 * r + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
    tri[3 * r + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
}

inline unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void accel_build(Context &ctx, Scene &s, int leaf_max) {
    KRT_REQUIRE(leaf_max >= 1 && leaf_max <= 32, KRT_ERR_INVALID, "leaf_max must be in [1, 32]");
    KRT_REQUIRE(s.nt >= 1, KRT_ERR_STATE, "scene has no triangles");
    KRT_REQUIRE(s.nt < (1u << 26), KRT_ERR_INVALID, "too many triangles for the leaf link encoding");
    hipStream_t st = ctx.stream;
    Accel &A = s.accel;
    const uint32_t n = s.nt;
    const uint32_t nser = 2 * n - 1;
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    KRT_HIP(hipEventRecord(e0, st));

    DevBuf bmin(sizeof(float4) * n), bmax(sizeof(float4) * n);
    DevBuf perm[2] = {DevBuf(sizeof(int) * n), DevBuf(sizeof(int) * n)};
    DevBuf segs[2] = {DevBuf(sizeof(Seg) * (n / 3 + 2)), DevBuf(sizeof(Seg) * (n / 3 + 2))};
    DevBuf cnt(sizeof(int));
    DevBuf leafcount(sizeof(int) * nser), rank(sizeof(int) * nser);
    A.ser.alloc(sizeof(float4) * 3 * nser);
    A.leaf_order.alloc(sizeof(int) * n);
    BuildOut o{A.ser.as<float4>(), A.leaf_order.as<int>(), leafcount.as<int>(), rank.as<int>()};

    tri_boxes_kernel<<<blocks(n, 256), 256, 0, st>>>(s.pos.as<float>(), s.idx.as<uint32_t>(), n, bmin.as<float4>(),
                                                     bmax.as<float4>());
    iota_kernel<<<blocks(n, 256), 256, 0, st>>>(perm[0].as<int>(), n);
    KRT_HIP(hipGetLastError());

    uint32_t levels = 0;
    if (n <= 2) {
        root_small_kernel<<<1, 64, 0, st>>>(o, (int)n, bmin.as<float4>(), bmax.as<float4>());
        KRT_HIP(hipGetLastError());
        levels = 1;
    } else {
        Seg root{0, (int)n, 0, -1, 0};
        KRT_HIP(hipMemcpyAsync(segs[0].p, &root, sizeof(Seg), hipMemcpyHostToDevice, st));
        int nseg = 1;
        while (nseg > 0) {
            KRT_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), st));
            level_kernel<<<nseg, BB, 0, st>>>(segs[levels & 1].as<Seg>(), perm[levels & 1].as<int>(),
                                              perm[(levels + 1) & 1].as<int>(), bmin.as<float4>(), bmax.as<float4>(),
                                              o, segs[(levels + 1) & 1].as<Seg>(), cnt.as<int>());
            KRT_HIP(hipGetLastError());
            KRT_HIP(hipMemcpyAsync(&nseg, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
            KRT_HIP(hipStreamSynchronize(st));
            levels++;
            KRT_REQUIRE(levels < 1000000, KRT_ERR_HIP, "BVH build did not converge");
        }
    }

    // traversal layout
    DevBuf flag(sizeof(int) * nser), node_id(sizeof(int) * nser), total(sizeof(int));
    mark_kernel<<<blocks(nser, 256), 256, 0, st>>>(A.ser.as<float4>(), leafcount.as<int>(), nser, leaf_max,
                                                   flag.as<int>());
    scan_kernel<<<1, 1024, 0, st>>>(flag.as<int>(), node_id.as<int>(), nser, total.as<int>());
    int nnodes = 0;
    KRT_HIP(hipMemcpyAsync(&nnodes, total.p, sizeof(int), hipMemcpyDeviceToHost, st));
    KRT_HIP(hipStreamSynchronize(st));
    A.nodes.alloc(sizeof(float4) * 4 * (size_t)nnodes);
    fill_nodes_kernel<<<blocks(nser, 256), 256, 0, st>>>(A.ser.as<float4>(), leafcount.as<int>(), rank.as<int>(),
                                                         flag.as<int>(), node_id.as<int>(), nser, leaf_max,
                                                         A.nodes.as<float4>());
    A.tri.alloc(sizeof(float4) * 3 * (size_t)n);
    tri_pack_kernel<<<blocks(n, 256), 256, 0, st>>>(s.pos.as<float>(), s.idx.as<uint32_t>(), A.leaf_order.as<int>(),
                                                    n, A.tri.as<float4>());
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(e1, st));
    KRT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    KRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    A.built = true;
    A.nt = n;
    A.nser = nser;
    A.nnodes = (uint32_t)nnodes;
    A.levels = levels;
    A.depth_bound = levels + 2;
    A.build_ms = ms;
    s.version++;
}

}  // namespace krt
