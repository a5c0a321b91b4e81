// Closest-hit and any-hit traversal of the krt BVH layout on gfx950.
//
// Node (4-wide, one 128 B line): 8 x float4 = the four children's min.x,
// max.x, min.y, max.y, min.z, max.z as SoA float4s, their links (wide node
// index, a leaf range 0x80000000 | (count-1) << 26 | first_rank, or 0 for an
// empty slot), pad.  Triangles sit in leaf-preorder rank order as
// {v1, prim}, {e1}, {e2}.
//
// Ordered descent (nearest child next, the other hit children on a per-lane
// stack kept in LDS, far first; entries past the LDS depth spill to a
// per-thread global slab).
//
// Result contract (the reference GLSL walk, shaders/rt/modules/bvh.glsl:52-127,
// keeps the first leaf in preorder among equal t): the closest hit is the
// lexicographic minimum of (t, rank) over triangles with 0 < t < tmax, with t,
// b1, b2 from Möller–Trumbore exactly as intersect.glsl:14-33, and rank the
// triangle's leaf position in the preorder of the bvh.cpp tree (stored with
// the triangle, so any traversal tree gives the same answer).  The box test is
// conservative (relative slack; a zero direction component admits every t
// while the origin lies inside that slab), so traversal order never changes
// that minimum.
#pragma once

#include "krt_device.h"

namespace krt {

struct Hit {
    float t, b1, b2;
    int prim;
    int rank;
    int inst;  // instance record of a two-level scene's hit (-1 otherwise)
};

// LDS-typed pointer: keeps stack traffic on ds_read/ds_write instead of flat
// loads through the generic aperture.
typedef __attribute__((address_space(3))) int lds_int;

struct TraceStack {
    lds_int *lds;  // this lane's column: lds[k * stride]
    int stride;
    int lds_depth;
    int *gmem;     // this lane's spill column: gmem[k * gstride]
    int gstride;
    __device__ __forceinline__ TraceStack(int *lds_col, int stride_, int depth, int *gcol, int gstride_)
        : lds((lds_int *)lds_col), stride(stride_), lds_depth(depth), gmem(gcol), gstride(gstride_) {}
    __device__ __forceinline__ void put(int k, int v) const {
        if (k < lds_depth) lds[k * stride] = v;
        else gmem[(k - lds_depth) * gstride] = v;
    }
    __device__ __forceinline__ int get(int k) const {
        return k < lds_depth ? lds[k * stride] : gmem[(k - lds_depth) * gstride];
    }
};

struct TravStats {
    unsigned nodes = 0, tris = 0;
};

KRT_DEV f3 safe_inv(f3 d) {
    float x = d.x != 0.0f ? d.x : copysignf(1e-30f, d.x);
    float y = d.y != 0.0f ? d.y : copysignf(1e-30f, d.y);
    float z = d.z != 0.0f ? d.z : copysignf(1e-30f, d.z);
    return mk(1.0f / x, 1.0f / y, 1.0f / z);
}

// Per-ray mask of zero direction components (bit a set when d[a] == 0).
KRT_DEV int zero_mask(f3 d) { return (d.x == 0.0f ? 1 : 0) | (d.y == 0.0f ? 2 : 0) | (d.z == 0.0f ? 4 : 0); }

// A ray with d[a] == 0 stays in the plane o[a]: the slab of that axis admits
// every t when lo <= o[a] <= hi and none otherwise.
KRT_DEV void zero_axis(float lo, float hi, float o, float &tn_a, float &tf_a) {
    bool in = o >= lo && o <= hi;
    tn_a = in ? -3.4e38f : 3.4e38f;
    tf_a = in ? 3.4e38f : -3.4e38f;
}

// Conservative slab test of one child box (lo/hi per axis).
KRT_DEV bool slab(float lox, float hix, float loy, float hiy, float loz, float hiz, f3 o, f3 inv, int zmask,
                  float tlim, float &tnear) {
    float tx0 = (lox - o.x) * inv.x, tx1 = (hix - o.x) * inv.x;
    float ty0 = (loy - o.y) * inv.y, ty1 = (hiy - o.y) * inv.y;
    float tz0 = (loz - o.z) * inv.z, tz1 = (hiz - o.z) * inv.z;
    float nx = fminf(tx0, tx1), fx = fmaxf(tx0, tx1);
    float ny = fminf(ty0, ty1), fy = fmaxf(ty0, ty1);
    float nz = fminf(tz0, tz1), fz = fmaxf(tz0, tz1);
    if (zmask) {  // rare: axis-parallel rays
        if (zmask & 1) zero_axis(lox, hix, o.x, nx, fx);
        if (zmask & 2) zero_axis(loy, hiy, o.y, ny, fy);
        if (zmask & 4) zero_axis(loz, hiz, o.z, nz, fz);
    }
    float tn = fmaxf(fmaxf(nx, ny), nz);
    float tf = fminf(fminf(fx, fy), fz);
    float tn_s = tn - (fabsf(tn) * 1e-5f + 1e-6f);
    float tf_s = tf + (fabsf(tf) * 1e-5f + 1e-6f);
    tnear = tn;
    return tn_s <= tf_s && tf_s >= 0.0f && tn_s <= tlim;
}

// Möller–Trumbore on the packed triangle (intersect.glsl:14-33, e1/e2 precomputed
// with the same subtraction).  Returns t or -1.
KRT_DEV float mt_test(const float4 *tri, int r, f3 o, f3 d, float &b1, float &b2, int &prim, int &rank) {
    float4 A = tri[3 * r], B = tri[3 * r + 1], C = tri[3 * r + 2];
    prim = __float_as_int(A.w);
    rank = __float_as_int(B.w);
    f3 v1 = mk(A.x, A.y, A.z), e1 = mk(B.x, B.y, B.z), e2 = mk(C.x, C.y, C.z);
    f3 s1 = cross(d, e2);
    float divisor = dot(s1, e1);
    if (divisor == 0.0f) return -1.0f;
    f3 s = o - v1;
    float inv = 1.0f / divisor;
    b1 = dot(s, s1) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return -1.0f;
    f3 s2 = cross(s, e1);
    b2 = dot(d, s2) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return -1.0f;
    return dot(e2, s2) * inv;
}

KRT_DEV bool is_leaf_link(int l) { return l < 0; }

// Ascending compare-exchange of (key, link) pairs.
KRT_DEV void cswap(float &ka, int &la, float &kb, int &lb) {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    const int l = sw ? lb : la;
    kb = sw ? ka : kb;
    lb = sw ? la : lb;
    ka = k;
    la = l;
}

template <bool ANY>
KRT_DEV bool traverse4(const float4 *__restrict__ nodes, const float4 *__restrict__ tri, f3 o, f3 d, float tmax,
                      const TraceStack &stk, Hit &hit, TravStats &ts) {
    const f3 inv = safe_inv(d);
    const int zmask = zero_mask(d);
    // Only (t, rank, position) of the best candidate live through the walk;
    // the hit record is recomputed from the winner at the end (the same
    // Möller–Trumbore expressions give the same bits).
    float best = tmax;
    int best_rank = 0x7fffffff;
    int best_pos = -1;
    int sp = 0;
    int cur = 0;  // the root; no child link is 0, so 0 also marks an empty slot
    for (;;) {
        if (!is_leaf_link(cur)) {
            const float4 *nd = nodes + 8 * (size_t)cur;
            const float4 MNX = nd[0], MXX = nd[1], MNY = nd[2], MXY = nd[3], MNZ = nd[4], MXZ = nd[5], L = nd[6];
            ts.nodes++;
            float k0, k1, k2, k3;
            int l0 = __float_as_int(L.x), l1 = __float_as_int(L.y), l2 = __float_as_int(L.z),
                l3 = __float_as_int(L.w);
            const float inf = 3.4e38f;
            bool h;
            h = l0 != 0 && slab(MNX.x, MXX.x, MNY.x, MXY.x, MNZ.x, MXZ.x, o, inv, zmask, best, k0);
            k0 = h ? k0 : inf;
            h = l1 != 0 && slab(MNX.y, MXX.y, MNY.y, MXY.y, MNZ.y, MXZ.y, o, inv, zmask, best, k1);
            k1 = h ? k1 : inf;
            h = l2 != 0 && slab(MNX.z, MXX.z, MNY.z, MXY.z, MNZ.z, MXZ.z, o, inv, zmask, best, k2);
            k2 = h ? k2 : inf;
            h = l3 != 0 && slab(MNX.w, MXX.w, MNY.w, MXY.w, MNZ.w, MXZ.w, o, inv, zmask, best, k3);
            k3 = h ? k3 : inf;
            // near first: sort the four (entry, link) pairs, missed boxes last
            cswap(k0, l0, k1, l1);
            cswap(k2, l2, k3, l3);
            cswap(k0, l0, k2, l2);
            cswap(k1, l1, k3, l3);
            cswap(k1, l1, k2, l2);
            if (k0 == inf) {
                if (sp == 0) break;
                cur = stk.get(--sp);
                continue;
            }
            if (k3 != inf) stk.put(sp++, l3);
            if (k2 != inf) stk.put(sp++, l2);
            if (k1 != inf) stk.put(sp++, l1);
            cur = l0;
        } else {
            const int first = cur & 0x03FFFFFF;
            const int cnt = ((cur >> 26) & 31) + 1;
            for (int k = 0; k < cnt; k++) {
                float b1, b2;
                int prim, r;
                float t = mt_test(tri, first + k, o, d, b1, b2, prim, r);
                ts.tris++;
                if (ANY) {
                    if (t > 0.0f && t < tmax) return true;
                } else {
                    bool take = t > 0.0f && (t < best || (best_pos >= 0 && t == best && r < best_rank));
                    best = take ? t : best;
                    best_rank = take ? r : best_rank;
                    best_pos = take ? first + k : best_pos;
                }
            }
            if (sp == 0) break;
            cur = stk.get(--sp);
        }
    }
    hit.prim = -1;
    hit.rank = 0x7fffffff;
    hit.inst = -1;
    hit.t = 0.0f;
    hit.b1 = hit.b2 = 0.0f;
    if (ANY || best_pos < 0) return false;
    float b1 = 0.0f, b2 = 0.0f;
    int prim, r;
    hit.t = mt_test(tri, best_pos, o, d, b1, b2, prim, r);
    hit.b1 = b1;
    hit.b2 = b2;
    hit.prim = prim;
    hit.rank = r;
    return true;
}

// Binary walk: node = 4 x float4 {c0.min, c0.max, c1.min, c1.max} per axis
// + two links; near child next, far child on the stack.
template <bool ANY>
KRT_DEV bool traverse2(const float4 *__restrict__ nodes, const float4 *__restrict__ tri, f3 o, f3 d, float tmax,
                      const TraceStack &stk, Hit &hit, TravStats &ts) {
    const f3 inv = safe_inv(d);
    const int zmask = zero_mask(d);
    // Only (t, rank, position) of the best candidate live through the walk;
    // the hit record is recomputed from the winner at the end (the same
    // Möller–Trumbore expressions give the same bits).
    float best = tmax;
    int best_rank = 0x7fffffff;
    int best_pos = -1;
    int sp = 0;
    int cur = 0;
    for (;;) {
        if (!is_leaf_link(cur)) {
            const float4 *nd = nodes + 4 * cur;
            float4 X = nd[0], Y = nd[1], Z = nd[2], L = nd[3];
            ts.nodes++;
            float t0, t1;
            bool h0 = slab(X.x, X.y, Y.x, Y.y, Z.x, Z.y, o, inv, zmask, best, t0);
            bool h1 = slab(X.z, X.w, Y.z, Y.w, Z.z, Z.w, o, inv, zmask, best, t1);
            int l0 = __float_as_int(L.x), l1 = __float_as_int(L.y);
            if (h0 && h1) {
                bool first0 = t0 <= t1;
                stk.put(sp++, first0 ? l1 : l0);
                cur = first0 ? l0 : l1;
            } else if (h0) {
                cur = l0;
            } else if (h1) {
                cur = l1;
            } else {
                if (sp == 0) break;
                cur = stk.get(--sp);
            }
        } else {
            const int first = cur & 0x03FFFFFF;
            const int cnt = ((cur >> 26) & 31) + 1;
            for (int k = 0; k < cnt; k++) {
                float b1, b2;
                int prim, r;
                float t = mt_test(tri, first + k, o, d, b1, b2, prim, r);
                ts.tris++;
                if (ANY) {
                    if (t > 0.0f && t < tmax) return true;
                } else {
                    bool take = t > 0.0f && (t < best || (best_pos >= 0 && t == best && r < best_rank));
                    best = take ? t : best;
                    best_rank = take ? r : best_rank;
                    best_pos = take ? first + k : best_pos;
                }
            }
            if (sp == 0) break;
            cur = stk.get(--sp);
        }
    }
    hit.prim = -1;
    hit.rank = 0x7fffffff;
    hit.inst = -1;
    hit.t = 0.0f;
    hit.b1 = hit.b2 = 0.0f;
    if (ANY || best_pos < 0) return false;
    float b1 = 0.0f, b2 = 0.0f;
    int prim, r;
    hit.t = mt_test(tri, best_pos, o, d, b1, b2, prim, r);
    hit.b1 = b1;
    hit.b2 = b2;
    hit.prim = prim;
    hit.rank = r;
    return true;
}

// Row-major 3x4 affine transform of a point (w = 1) or direction (w = 0)
// with glm's pairing of mat4 x vec4: (c0 x + c1 y) + (c2 z + c3 w).
KRT_DEV f3 xform(const float4 &r0, const float4 &r1, const float4 &r2, f3 p, float w) {
    return mk((r0.x * p.x + r0.y * p.y) + (r0.z * p.z + r0.w * w), (r1.x * p.x + r1.y * p.y) + (r1.z * p.z + r1.w * w),
              (r2.x * p.x + r2.y * p.y) + (r2.z * p.z + r2.w * w));
}

// Instance record (8 float4, in instance-tree traversal order): rows of the
// inverse transform, rows of the transform, {4-wide node base, packed
// triangle base, global id of the first triangle, bvh.cpp rank of the
// instance}, {vertex base, triangle base of the mesh, mesh, 0}.
constexpr int kInstRec = 8;

// Two-level walk (Accelerator::build_tlas + optixTrace semantics,
// include/amadeus/accelerator.cuh:197-277): the world ray walks the 4-wide
// instance tree (leaves hold one instance); at an instance the ray is taken
// into the mesh's space by the inverse transform (the direction is not
// renormalised, so t is shared) and walks that mesh's 4-wide tree on the same
// stack.  The closest hit is the lexicographic minimum of (t, instance rank,
// triangle rank): the first instance in instance-tree preorder, then the first
// triangle in mesh preorder, among equal t.
template <bool ANY>
KRT_DEV bool traverse_inst(const float4 *__restrict__ tlas, const float4 *__restrict__ inst,
                           const float4 *__restrict__ wide, const float4 *__restrict__ tri, const f3 wo, const f3 wd,
                           float tmax, const TraceStack &stk, Hit &hit, TravStats &ts) {
    const f3 winv = safe_inv(wd);
    const int wz = zero_mask(wd);
    f3 o = wo, d = wd, inv = winv;
    int zmask = wz;
    const float4 *nodes = tlas;
    const float4 *tris = tri;
    float best = tmax;
    int best_rank = 0x7fffffff, best_irank = 0x7fffffff;
    int best_pos = -1, best_inst = -1;
    int cur_inst = -1, cur_irank = 0;
    int base_sp = -1;  // stack height when the current instance was entered (-1: in the instance tree)
    int sp = 0;
    int cur = 0;
    for (;;) {
        if (!is_leaf_link(cur)) {
            const float4 *nd = nodes + 8 * (size_t)cur;
            const float4 MNX = nd[0], MXX = nd[1], MNY = nd[2], MXY = nd[3], MNZ = nd[4], MXZ = nd[5], L = nd[6];
            ts.nodes++;
            float k0, k1, k2, k3;
            int l0 = __float_as_int(L.x), l1 = __float_as_int(L.y), l2 = __float_as_int(L.z),
                l3 = __float_as_int(L.w);
            const float inf = 3.4e38f;
            bool h;
            h = l0 != 0 && slab(MNX.x, MXX.x, MNY.x, MXY.x, MNZ.x, MXZ.x, o, inv, zmask, best, k0);
            k0 = h ? k0 : inf;
            h = l1 != 0 && slab(MNX.y, MXX.y, MNY.y, MXY.y, MNZ.y, MXZ.y, o, inv, zmask, best, k1);
            k1 = h ? k1 : inf;
            h = l2 != 0 && slab(MNX.z, MXX.z, MNY.z, MXY.z, MNZ.z, MXZ.z, o, inv, zmask, best, k2);
            k2 = h ? k2 : inf;
            h = l3 != 0 && slab(MNX.w, MXX.w, MNY.w, MXY.w, MNZ.w, MXZ.w, o, inv, zmask, best, k3);
            k3 = h ? k3 : inf;
            cswap(k0, l0, k1, l1);
            cswap(k2, l2, k3, l3);
            cswap(k0, l0, k2, l2);
            cswap(k1, l1, k3, l3);
            cswap(k1, l1, k2, l2);
            if (k0 != inf) {
                if (k3 != inf) stk.put(sp++, l3);
                if (k2 != inf) stk.put(sp++, l2);
                if (k1 != inf) stk.put(sp++, l1);
                cur = l0;
                continue;
            }
        } else if (base_sp < 0) {
            // instance leaf: enter the mesh of instance record `first`
            const int r = cur & 0x03FFFFFF;
            const float4 *rec = inst + kInstRec * (size_t)r;
            const float4 i0 = rec[0], i1 = rec[1], i2 = rec[2], meta = rec[6];
            o = xform(i0, i1, i2, wo, 1.0f);
            d = xform(i0, i1, i2, wd, 0.0f);
            inv = safe_inv(d);
            zmask = zero_mask(d);
            nodes = wide + 8 * (size_t)__float_as_int(meta.x);
            tris = tri + 3 * (size_t)__float_as_int(meta.y);
            cur_inst = r;
            cur_irank = __float_as_int(meta.w);
            base_sp = sp;
            cur = 0;
            continue;
        } else {
            const int first = cur & 0x03FFFFFF;
            const int cnt = ((cur >> 26) & 31) + 1;
            for (int k = 0; k < cnt; k++) {
                float b1, b2;
                int prim, rk;
                float t = mt_test(tris, first + k, o, d, b1, b2, prim, rk);
                ts.tris++;
                if (ANY) {
                    if (t > 0.0f && t < tmax) return true;
                } else {
                    bool take = t > 0.0f &&
                                (t < best || (best_pos >= 0 && t == best &&
                                              (cur_irank < best_irank || (cur_irank == best_irank && rk < best_rank))));
                    best = take ? t : best;
                    best_rank = take ? rk : best_rank;
                    best_irank = take ? cur_irank : best_irank;
                    best_pos = take ? first + k : best_pos;
                    best_inst = take ? cur_inst : best_inst;
                }
            }
        }
        // pop; leaving an instance when its part of the stack is exhausted
        if (base_sp >= 0 && sp == base_sp) {
            o = wo;
            d = wd;
            inv = winv;
            zmask = wz;
            nodes = tlas;
            base_sp = -1;
        }
        if (sp == 0) break;
        cur = stk.get(--sp);
    }
    hit.prim = -1;
    hit.rank = 0x7fffffff;
    hit.inst = -1;
    hit.t = 0.0f;
    hit.b1 = hit.b2 = 0.0f;
    if (ANY || best_pos < 0) return false;
    const float4 *rec = inst + kInstRec * (size_t)best_inst;
    const float4 i0 = rec[0], i1 = rec[1], i2 = rec[2], meta = rec[6];
    const f3 bo = xform(i0, i1, i2, wo, 1.0f), bd = xform(i0, i1, i2, wd, 0.0f);
    float b1 = 0.0f, b2 = 0.0f;
    int prim, rk;
    hit.t = mt_test(tri + 3 * (size_t)__float_as_int(meta.y), best_pos, bo, bd, b1, b2, prim, rk);
    hit.b1 = b1;
    hit.b2 = b2;
    hit.prim = __float_as_int(meta.z) + prim;
    hit.rank = rk;
    hit.inst = best_inst;
    return true;
}

// Which tree a query walks.  Closest-hit rays keep the binary tree.  Shadow
// rays take the 4-wide tree (four boxes per 128 B fetch, fewer dependent
// loads) unless the kernel opts out: the wide walk needs more registers, and a
// kernel already near the occupancy cliff (ReSTIR's initial pass, the cached
// path kernel) runs faster on the binary tree at higher occupancy.
template <bool ANY, bool WIDE = ANY>
KRT_DEV bool traverse(const float4 *__restrict__ bin, const float4 *__restrict__ wide, const float4 *__restrict__ tri,
                      f3 o, f3 d, float tmax, const TraceStack &stk, Hit &hit, TravStats &ts) {
    if (WIDE) return traverse4<ANY>(wide, tri, o, d, tmax, stk, hit, ts);
    return traverse2<ANY>(bin, tri, o, d, tmax, stk, hit, ts);
}

}  // namespace krt
