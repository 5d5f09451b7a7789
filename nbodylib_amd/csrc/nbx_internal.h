// nbx_internal.h -- device-side layout of the KDTree hot path on gfx950.
//
// Particles live in HBM as SoA arrays in kd split order ("slots").  Leaves
// are fixed buckets of NBX_LEAF consecutive slots; internal nodes form an
// implicit complete binary tree in heap order over the leaf count padded to a
// power of two (nleafp): the root is node 1, node n has children 2n and 2n+1,
// and leaf L is node nleafp + L.  Padding leaves carry an empty box
// (lo = +inf, hi = -inf), whose lower bound is +inf, so they are never opened.
// Every node stores its tight FP64 bounding box, so all pruning bounds are
// computed with the same roundings as the particle distances (box_lb/box_ub).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NBX_LEAF 32
#define NBX_WAVE 64
#define NBX_MAXLEV 40
#define NBX_MAXVALUE 1e32  // MAXVALUE, Precision.h:48
// Coordinate arrays are padded to whole leaves with this far sentinel: any
// squared distance to it overflows to +inf, so padding never passes a test.
#define NBX_PAD_COORD 1e300
#define NBX_PAD_COORD_F 1e30f

typedef double dx8 __attribute__((ext_vector_type(8)));
typedef const __attribute__((address_space(4))) dx8 *cd8ptr;
typedef double dx4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) dx4 *cd4ptr;

typedef const __attribute__((address_space(4))) double *cdptr;  // scalar-cache (read-only) view
typedef const __attribute__((address_space(4))) float *cfptr;
typedef const __attribute__((address_space(4))) int32_t *ciptr;

// Kernel-side view of a built tree (passed by value).
struct DevTree {
    int32_t n;      // particles
    int32_t nleaf;  // leaves = ceil(n / NBX_LEAF)
    int32_t nd;     // 3 or 6
    int32_t pad;
    const double *x[6];    // sorted coordinates (x y z [vx vy vz])
    const double *mass;    // sorted masses
    const int32_t *perm;   // slot -> ID
    const int32_t *inv;    // ID -> slot
    const double *box;     // node boxes, [node][lo[nd], hi[nd]]
    const float *boxf;     // precision form: the node boxes rounded outward to FP32 (same layout)
    const float *plane;    // internal nodes: child-order plane [node][n[nd], off] (4 or 8 floats)
    const double *org;     // mixed form: per-leaf FP64 origin [leaf][nd]
    const float *xf[6];    // mixed form, 6-D: FP32 offsets from the leaf origin
    const float4 *xq;      // mixed form, 3-D: (x, y, z, |.|^2) FP32 offsets per slot (padding: 1e30, +inf)
    const double *kern;    // kernel table
    const float *kernf;    // kernel table in FP32 (mixed form)
    int32_t kernres;
    int32_t nleafp;        // leaves padded to a power of two
    int32_t depth;         // log2(nleafp): depth of the leaf level (root depth 0)
    int32_t pad2;
};

// box of heap node n (1 <= n < 2 nleafp)
__device__ __forceinline__ const double *node_box(const DevTree &T, int n, int nd) {
    return T.box + (size_t)n * 2 * nd;
}
__device__ __forceinline__ int node_depth(int n) { return 31 - __clz(n); }
// first slot of heap node n
__device__ __forceinline__ int32_t node_first(const DevTree &T, int n) {
    return ((n << (T.depth - node_depth(n))) - T.nleafp) * NBX_LEAF;
}
// one past the last slot of heap node n (clamped to the particle count)
__device__ __forceinline__ int32_t node_end(const DevTree &T, int n) {
    int32_t e = (((n + 1) << (T.depth - node_depth(n))) - T.nleafp) * NBX_LEAF;
    return e < T.n ? e : T.n;
}

// Lower bound on the squared distance from q to any point of box b, summed
// in coordinate order from 0.  Because IEEE subtraction, squaring and
// addition are monotone, the computed bound never exceeds the computed
// distance to a particle inside the box (reference: DistFunc.h:14-19 sums
// (p1-p2)^2 in the same order), so a search pruned by it is exact.
template <int ND>
__device__ __forceinline__ double box_lb(const double *q, cdptr b) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        // at most one of (lo - q), (q - hi) is positive; max() is exact
        double t = fmax(fmax(b[d] - q[d], q[d] - b[ND + d]), 0.0);
        s += t * t;
    }
    return s;
}
// A lower bound on the squared distance from q to any point of the FP32
// box b (rounded outward from the FP64 box), never above the FP64 box_lb of
// the same node, so a walk that opens a node when this bound reaches its
// threshold opens every node the FP64 test would.  q32 = fl32(q), with
// |q| <= 1e30 (the tree's coordinates are checked when the FP32 boxes are
// made).  Per dimension the FP32 gap t over-states the exact gap g by at most
// (g + u|q|)(1 + u), u = 2^-24, because the outward box only shrinks the gap
// and q32 and the subtraction add one rounding each; t (1 - 2^-20) -
// 2^-20 |q32|, rounded once by the FMA, is then at most g (1 - 14u), and the
// squares and their sum (at most 2 ND roundings) stay below the exact bound
// times (1 - 2^-20), itself below the FP64 value, which is within 2^-50
// relative of exact.  An empty box (+inf, -inf) gives +inf.
template <int ND>
__device__ __forceinline__ float box_lb32(const float *q32, cfptr b) {
    float s = 0.0f;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const float t = fmaxf(b[d] - q32[d], q32[d] - b[ND + d]);
        const float tl = fmaxf(__builtin_fmaf(t, 1.0f - 0x1p-20f, -0x1p-20f * fabsf(q32[d])), 0.0f);
        s = __builtin_fmaf(tl, tl, s);
    }
    return s;
}
// Upper bound on the squared distance from q to any point of box b (farthest
// corner, KDNode.h:279-291 computes the same for whole-node links).
template <int ND>
__device__ __forceinline__ double box_ub(const double *q, cdptr b) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double a = q[d] - b[d], c = b[ND + d] - q[d];
        double a2 = a * a, c2 = c * c;
        s += (a2 > c2) ? a2 : c2;
    }
    return s;
}
// Upper bound on the squared distance between any point of box a and any
// point of box b (monotone for the same reason).
template <int ND>
__device__ __forceinline__ double boxbox_ub(cdptr a, cdptr b) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double u = b[ND + d] - a[d], v = a[ND + d] - b[d];
        double m = u > v ? u : v;
        s += m * m;
    }
    return s;
}
template <int ND>
__device__ __forceinline__ double boxbox_lb(cdptr a, cdptr b) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double t = 0.0;
        if (a[ND + d] < b[d]) t = b[d] - a[ND + d];
        else if (b[ND + d] < a[d]) t = a[d] - b[ND + d];
        s += t * t;
    }
    return s;
}

// Which child of internal node p to descend first, seen from the uniform
// point qc: the one whose box centre is nearer (a heuristic for the visiting
// order only; pruning never depends on it).
template <int ND>
__device__ __forceinline__ bool first_is_right(const float *plane, int p, const float *qc) {
    const cfptr pl = (cfptr)(plane + (size_t)p * (ND == 6 ? 8 : 4));
    float s = 0.0f;
#pragma unroll
    for (int d = 0; d < ND; ++d) s = __builtin_fmaf(pl[d], qc[d], s);
    return s > pl[ND];
}
template <int ND>
__device__ __forceinline__ bool first_is_right(const DevTree &T, int p, const float *qc) {
    return first_is_right<ND>(T.plane, p, qc);
}

// Workgroups of a launch are dealt round-robin to the 8 XCDs (workgroup b
// runs on XCD b % 8), each with its own L2.  Renumber them so that XCD x
// walks one contiguous run of tree-ordered waves: concurrently resident
// waves on an XCD then share tree nodes and leaves in that XCD's L2 instead
// of every L2 fetching every neighbourhood.  A bijection for any grid size.
__device__ __forceinline__ int xcd_wave(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p ? 1 : 0); }

__device__ __forceinline__ double readfirst_d(double v) {
    union { double d; int32_t i[2]; } u;
    u.d = v;
    u.i[0] = __builtin_amdgcn_readfirstlane(u.i[0]);
    u.i[1] = __builtin_amdgcn_readfirstlane(u.i[1]);
    return u.d;
}

// Wsm, KDCalcSmoothQuantities.cxx:12-15: linear interpolation in the table,
// indexed as (int)(r/h*0.5*(kernres-1)).  Past the table's end (r/h >= 2,
// which CalcVelDensityWithPhysDensity's h without a square root reaches) the
// last entry is taken; that is decided on the double before any conversion,
// so a huge r/h never meets an out-of-range float-to-int conversion.
__device__ __forceinline__ double wsm_at(double rij, double hi, const double *kern, int kernres) {
    const double r = rij / hi;
    const double fi = rij / hi * 0.5 * (kernres - 1);
    if (!(fi < (double)(kernres - 1))) return kern[kernres - 1];
    int i = (int)fi;
    if (i < 0) i = 0;
    const double delta = 2.0 / (double)(kernres - 1);
    return kern[i] + (kern[i + 1] - kern[i]) * (r - delta * i) / delta;
}

// Lock-free union-find over slots (hook the larger root under the smaller).
//
// A parent pointer only ever points to a smaller slot of the same set, and
// only ever changes to another such slot (a root is hooked under a smaller
// slot; path splitting skips to a grandparent).  So a parent value read
// late, from a cache that another XCD's writes have not reached, still
// names a slot of the same set, and a chain of such values still descends
// and ends.  The links made while walking (uf_union_r) therefore take
// ordinary cached loads first (the wavefront-scope load hits the CU's and
// the XCD's caches) and leave device-scope loads, each a round trip past
// the XCD's L2, for the cases that need the present state:
//  * two chains that meet prove the sets joined, since sets only merge;
//  * a hook is a device-scope compare-and-swap that succeeds only on a
//    true root, and hangs it under a smaller slot, so no cycle can form;
//    when it fails the union runs again on device-scope finds.
// uf_union, which joins whole nodes slot by slot, and the finds after the
// linking, which need the final root, use device-scope loads throughout
// (measured the better form for both).
__device__ __forceinline__ int32_t uf_load(int32_t *P, int32_t x) {
    return __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t uf_load_cached(int32_t *P, int32_t x) {
    return __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
template <bool CACHED>
__device__ __forceinline__ int32_t uf_find_t(int32_t *P, int32_t x) {
    int32_t p = CACHED ? uf_load_cached(P, x) : uf_load(P, x);
    while (p != x) {
        int32_t gp = CACHED ? uf_load_cached(P, p) : uf_load(P, p);
        if (gp != p) __hip_atomic_store(P + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = gp;
    }
    return x;
}
__device__ __forceinline__ int32_t uf_find(int32_t *P, int32_t x) { return uf_find_t<false>(P, x); }

// One attempt to hook the larger of two roots under the smaller.
__device__ __forceinline__ bool uf_hook(int32_t *P, int32_t &a, int32_t &b) {
    if (a < b) { int32_t t = a; a = b; b = t; }
    int32_t expected = a;
    return __hip_atomic_compare_exchange_strong(P + a, &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}

// Union of the sets of a and b; returns a root the joined set had when the
// link was made or found (a valid starting point for later finds from a).
__device__ __forceinline__ int32_t uf_union_r(int32_t *P, int32_t a, int32_t b) {
    a = uf_find_t<true>(P, a);
    b = uf_find_t<true>(P, b);
    if (a == b) return a;
    if (uf_hook(P, a, b)) return b;
    for (;;) {
        a = uf_find(P, a);
        b = uf_find(P, b);
        if (a == b) return a;
        if (uf_hook(P, a, b)) return b;
    }
}
__device__ __forceinline__ void uf_union(int32_t *P, int32_t a, int32_t b) {
    for (;;) {
        a = uf_find(P, a);
        b = uf_find(P, b);
        if (a == b || uf_hook(P, a, b)) return;
    }
}
