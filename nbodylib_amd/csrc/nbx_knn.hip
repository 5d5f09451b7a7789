// nbx_knn.hip -- exact FP64 k-nearest-neighbour search and SPH densities.
//
// Replaces KDTree::FindNearest* (KDFindNearest.cxx:247-361, :444-564),
// SplitNode/LeafNode::FindNearestPos/Phase (KDSplitNode.cxx:15-95,
// KDLeafNode.cxx:15-57, :119-154), the periodic image passes
// (KDSplitNode.cxx:1119-1179), CalcDensity / CalcDensityParticle /
// CalcVelDensity (KDCalcSmoothQuantities.cxx:203-389, :768-844).
//
// Mapping to CDNA4: one wavefront owns 64 slot-consecutive queries (two
// sibling leaves).  The wave walks the implicit tree once for all 64 lanes
// (a node is opened when any lane's bound reaches it), so node boxes and
// leaf particles arrive through the scalar cache as 64-byte s_load bursts
// and are shared by the whole wave.  Each lane keeps its own 4-ary max-heap
// of (d2, particle ID) in LDS -- the reference's PriorityQueue
// (PriorityQueue.h:47-84), with three dependent levels instead of five for
// k = 32 -- laid out [position][lane] so same-position accesses are
// conflict-free.  A leaf is consumed in two phases: (1) all lanes test every
// candidate against their current k-th distance, building a 32-bit accept
// mask; (2) lanes insert their accepted candidates one at a time from an LDS
// copy of the leaf with a branch-free root replacement, so the wave pays for
// as many heap insertions as its busiest lane, not the sum over lanes.
//
// Exactness: distances are summed in the reference's order with no FMA
// (this file is built with -ffp-contract=off); pruning uses box_lb, which is
// never larger than a computed particle distance.  The neighbour set is the k
// smallest (d2, ID) pairs: ties at the k-th distance go to the smaller ID
// (the reference's choice there depends on its traversal order).
#include "nbx_kernels.h"

// Heap entries are (d2, particle ID) pairs ordered by d2, then ID; the ID is
// carried in the heap so ties never need a lookup.  Empty entries are
// (NBX_MAXVALUE, NBX_NOID); filler entries (-1, -1) pad the heap to a
// capacity of 1 (mod 4) and, being smaller than every key, are never evicted.
#define NBX_NOID 0x7fffffff

// (da, ia) > (db, ib)
__device__ __forceinline__ bool kgt(double da, int32_t ia, double db, int32_t ib) {
    return da > db || (da == db && ia > ib);
}

#define HD(p) hd[(p) * NBX_WAVE + lane]
#define HI(p) hid[(p) * NBX_WAVE + lane]

// Heap capacity for kh kept entries: kh rounded up to 1 (mod 4), so every
// parent of a full 4-ary heap has four children.
__host__ __device__ __forceinline__ int heap_cap4(int kh) { return ((kh + 2) & ~3) + 1; }

// Largest of four children (and its index 0..3).
__device__ __forceinline__ void max4(const double *d, const int32_t *id, double &bd, int32_t &bi, int &bx) {
    const bool s01 = kgt(d[1], id[1], d[0], id[0]), s23 = kgt(d[3], id[3], d[2], id[2]);
    const double m01 = s01 ? d[1] : d[0], m23 = s23 ? d[3] : d[2];
    const int32_t j01 = s01 ? id[1] : id[0], j23 = s23 ? id[3] : id[2];
    const bool s = kgt(m23, j23, m01, j01);
    bd = s ? m23 : m01;
    bi = s ? j23 : j01;
    bx = s ? (s23 ? 3 : 2) : (s01 ? 1 : 0);
}

// 4-ary max-heap, 0-based, children of p at 4p+1 .. 4p+4 (the reference's
// PriorityQueue, PriorityQueue.h:47-84, with three dependent levels instead
// of five for k = 32).  Replace the root of the full heap of hc entries by
// (dv, iv) for the lanes with ins set and sift it down; the other lanes
// rewrite the root with itself.  LV unrolled levels cover hc <= 85 (LV = 3)
// or hc <= 341 (LV = 4) and nothing branches.  (kth, kid) is the root afterwards.
template <int LV>
__device__ __forceinline__ void heap_replace_root(double *hd, int32_t *hid, int lane, int hc, bool ins, double dn,
                                                  int32_t in, double &kth, int32_t &kid) {
    const double dv = ins ? dn : kth;
    const int32_t iv = ins ? in : kid;
    double td = dv;
    int32_t ti = iv;
    int p = 0;
    bool go = ins;
#pragma unroll
    for (int lv = 0; lv < LV; ++lv) {
        const int c = 4 * p + 1;
        go = go && c < hc;
        const int cc = go ? c : 1;
        double d[4];
        int32_t id[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { d[u] = HD(cc + u); id[u] = HI(cc + u); }
        double bd;
        int32_t bi;
        int bx;
        max4(d, id, bd, bi, bx);
        go = go && kgt(bd, bi, dv, iv);
        HD(p) = go ? bd : dv;
        HI(p) = go ? bi : iv;
        if (lv == 0) { td = go ? bd : dv; ti = go ? bi : iv; }
        p = go ? cc + bx : p;
    }
    HD(p) = dv;
    HI(p) = iv;
    kth = td;
    kid = ti;
}

// Sift (dv, iv) down from the root of a heap of m <= hc entries; children
// at positions >= m read as (-inf, min) and never outrank an entry.
template <int LV>
__device__ __forceinline__ void heap_sift_bounded(double *hd, int32_t *hid, int lane, int m, double dv, int32_t iv) {
    int p = 0;
    bool go = true;
#pragma unroll
    for (int lv = 0; lv < LV; ++lv) {
        const int c = 4 * p + 1;
        go = go && c < m;
        const int cc = go ? c : 1;
        double d[4];
        int32_t id[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double t = HD(cc + u);
            const int32_t ti = HI(cc + u);
            const bool in = cc + u < m;
            d[u] = in ? t : -INFINITY;
            id[u] = in ? ti : (int32_t)0x80000000;
        }
        double bd;
        int32_t bi;
        int bx;
        max4(d, id, bd, bi, bx);
        go = go && kgt(bd, bi, dv, iv);
        HD(p) = go ? bd : dv;
        HI(p) = go ? bi : iv;
        p = go ? cc + bx : p;
    }
    HD(p) = dv;
    HI(p) = iv;
}

// General sift from position p in a heap of size m (short selection of the
// velocity density only).
__device__ __forceinline__ void heap_sift(double *hd, int32_t *hid, int lane, int m, int p, double dv, int32_t iv) {
    for (;;) {
        int c = 4 * p + 1;
        if (c >= m) break;
        double dc = HD(c);
        int32_t ic = HI(c);
        int best = c;
        for (int u = 1; u < 4; ++u) {
            if (c + u < m && kgt(HD(c + u), HI(c + u), dc, ic)) { best = c + u; dc = HD(c + u); ic = HI(c + u); }
        }
        if (!kgt(dc, ic, dv, iv)) break;
        HD(p) = dc;
        HI(p) = ic;
        p = best;
    }
    HD(p) = dv;
    HI(p) = iv;
}

// Heap sort in place: positions 0..hc-1 ascending afterwards (fillers first).
template <int LV>
__device__ __forceinline__ void heap_sort(double *hd, int32_t *hid, int lane, int hc) {
    for (int m = hc - 1; m > 0; --m) {
        const double dt = HD(0), dm = HD(m);
        const int32_t it = HI(0), im = HI(m);
        HD(m) = dt;
        HI(m) = it;
        heap_sift_bounded<LV>(hd, hid, lane, m, dm, im);
    }
}

// heap levels below the root that a capacity of KMAX entries needs
template <int KMAX>
struct HeapLv {
    static constexpr int LV = KMAX <= 85 ? 3 : 4;
    static_assert(KMAX <= 341, "four levels hold at most 341 entries");
};

template <int ND>
struct LeafCtx {
    double q[ND];
    bool act;
    bool exz;
    double kth;   // root of the heap: the kh-th smallest (d2, ID) so far
    int32_t kid;
};

// Consume one leaf for every active lane (two-phase insertion, see header).
template <int ND, int LV>
__device__ __forceinline__ void knn_leaf(const DevTree &T, int leaf, LeafCtx<ND> &c, double *hd, int32_t *hid,
                                         double *cx, int32_t *cid, int lane, int hc) {
    const int base = leaf * NBX_LEAF;
    if (lane < NBX_LEAF) {
#pragma unroll
        for (int d = 0; d < ND; ++d) cx[d * NBX_LEAF + lane] = T.x[d][base + lane];
        cid[lane] = T.perm[base + lane];
    }
    uint32_t mask = 0;
    // candidates arrive in chunks through the scalar cache: 8 per dimension
    // (24 SGPR pairs) in 3-D, 4 in 6-D
    constexpr int CH = ND == 6 ? 4 : 8;
    typedef double dxc __attribute__((ext_vector_type(CH)));
    typedef const __attribute__((address_space(4))) dxc *cdcptr;
#pragma unroll
    for (int jj = 0; jj < NBX_LEAF; jj += CH) {
        dxc cv[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) cv[d] = *(cdcptr)(T.x[d] + base + jj);
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                double t = c.q[d] - cv[d][u];
                d2 += t * t;
            }
            // the query itself has d2 == 0: excluded by the d2 > 0 rule of
            // the target search and kept by the position search, as in the
            // reference (KDLeafNode.cxx:19-21, :123-124)
            bool ok = (d2 <= c.kth) && (!c.exz || d2 > 0.0);
            mask |= ok ? (1u << (jj + u)) : 0u;
        }
    }
    if (!c.act) mask = 0;
    __syncthreads();
    while (ballot(mask != 0)) {
        const bool has = mask != 0;
        const int j = __builtin_ctz(mask | 0x80000000u);
        mask &= mask - 1;
        double d2 = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double t = c.q[d] - cx[d * NBX_LEAF + j];
            d2 += t * t;
        }
        const int32_t id = cid[j];
        heap_replace_root<LV>(hd, hid, lane, hc, has && kgt(c.kth, c.kid, d2, id), d2, id, c.kth, c.kid);
    }
    __syncthreads();
}

// Uniform stackless depth-first walk of the heap-ordered tree.  The first
// descent follows the ancestors of the wave's own leaf node (anc; -1 when
// the queries are not tree particles or are image-shifted), so every heap
// is filled from the query's own neighbourhood before anything else is
// opened; below that children are taken nearest first from qc.  A node is
// opened when any active lane has box_lb <= its current k-th distance.
// State is uniform (SGPRs): node, depth and one "right child first" bit per
// depth.
template <int ND, int LV>
__device__ __forceinline__ void knn_walk(const DevTree &T, LeafCtx<ND> &c, const float *qc, int anc, double *hd,
                                         int32_t *hid, double *cx, int32_t *cid, int lane, int hc) {
    const int D = T.depth;
    int node = 1, dep = 0;
    uint64_t firstright = 0;
    bool down = true;
    for (;;) {
        if (down) {
            cdptr bx = (cdptr)node_box(T, node, ND);
            bool want = c.act && box_lb<ND>(c.q, bx) <= c.kth;
            if (ballot(want)) {
                if (dep == D) {
                    knn_leaf<ND, LV>(T, node - T.nleafp, c, hd, hid, cx, cid, lane, hc);
                    down = false;
                } else {
                    int c0 = 2 * node;
                    int first = c0;
                    if (anc >= 0 && (anc >> (D - dep)) == node) first = anc >> (D - dep - 1);
                    else if (first_is_right<ND>(T, node, qc)) first = c0 + 1;
                    dep += 1;
                    if (first & 1) firstright |= (1ULL << dep);
                    else firstright &= ~(1ULL << dep);
                    node = first;
                    continue;
                }
            } else {
                down = false;
            }
        }
        if (node == 1) break;
        if ((uint64_t)(node & 1) == ((firstright >> dep) & 1ULL)) {
            node ^= 1;  // the sibling always exists in the padded tree
            down = true;
        } else {
            node >>= 1;
            dep -= 1;
        }
    }
}

// Image point and test for periodic pass p (1..7), DistFunc.h:326-355 and
// KDSplitNode.cxx:1119-1148: 1-D images need sqrt(Top) > s, the 2-D and 3-D
// ones compare Top (a squared distance) with s (a distance), as the
// reference does.  Phase-space searches take every image (:1153-1179).
__device__ __forceinline__ bool image_pass(int pass, const double *x0, const double *per, double kth, bool uncond,
                                           double *xp) {
    const bool u0 = pass == 1 || pass == 4 || pass == 5 || pass == 7;
    const bool u1 = pass == 2 || pass == 4 || pass == 6 || pass == 7;
    const bool u2 = pass == 3 || pass == 5 || pass == 6 || pass == 7;
    const bool use[3] = {u0, u1, u2};
    double sval = 0.0, s1 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xp[k] = x0[k];
        if (use[k]) {
            if (x0[k] < per[k] / 2.0) {
                xp[k] = x0[k] + per[k];
                sval += x0[k] * x0[k];
                s1 = x0[k];
            } else {
                xp[k] = x0[k] - per[k];
                sval += xp[k] * xp[k];
                s1 = -xp[k];
            }
        }
    }
    if (pass <= 3) return uncond || sqrt(kth) > s1;
    return uncond || kth > sqrt(sval);
}

template <int ND, int KMAX, int QSRC>
__global__ __launch_bounds__(64) void k_knn(DevTree T, KnnArgs A) {
    __shared__ double hd[KMAX * NBX_WAVE];
    __shared__ int32_t hid[KMAX * NBX_WAVE];
    __shared__ double cx[ND * NBX_LEAF];
    __shared__ int32_t cid[NBX_LEAF];
    __shared__ double racc[NBX_WAVE];  // symmetric scatter into the wave's own queries
    const int lane = threadIdx.x;
    const int g = xcd_wave(blockIdx.x, gridDim.x) + A.wbeg;
    const int kh = A.kh;
    const int hc = heap_cap4(kh);
    const int nfill = hc - kh;  // filler entries, first after sorting
    const bool listed = (QSRC != NBX_Q_POINTS) && A.qlist != nullptr;
    const int wend = A.wbeg + A.wcnt;
    const int nq = (QSRC == NBX_Q_POINTS || listed) ? A.m
                   : (A.wcnt > 0 && wend < (T.n + NBX_WAVE - 1) / NBX_WAVE ? wend * NBX_WAVE : T.n);
    const int gi = g * NBX_WAVE + lane;
    const bool valid = gi < nq;
    const int qi = listed ? (valid ? A.qlist[gi] : 0) : gi;
    double q0[ND];
    if (QSRC == NBX_Q_POINTS) {
        // phase points (ND = 6) bring their velocities in qv
#pragma unroll
        for (int d = 0; d < ND; ++d) q0[d] = valid ? (d < 3 ? A.q[d][qi] : A.qv[d - 3][qi]) : 0.0;
    } else {
#pragma unroll
        for (int d = 0; d < ND; ++d) q0[d] = valid ? T.x[d][qi] : 0.0;
    }
    // the fillers sit at the bottom of the heap array, where a 4-ary heap of
    // empty entries stays valid with them as leaves
    for (int p = 0; p < hc; ++p) {
        const bool fill = p >= kh;
        HD(p) = fill ? -1.0 : NBX_MAXVALUE;
        HI(p) = fill ? -1 : NBX_NOID;
    }
    LeafCtx<ND> c;
    c.exz = (QSRC == NBX_Q_SLOTS) && A.excl_zero;
    c.kth = NBX_MAXVALUE;
    c.kid = NBX_NOID;
    const int npass = (QSRC != NBX_Q_SLOTS && A.periodic) ? 8 : 1;
    for (int pass = 0; pass < npass; ++pass) {
#pragma unroll
        for (int d = 0; d < ND; ++d) c.q[d] = q0[d];
        c.act = valid;
        if (pass > 0) {
            double xp[3];
            c.act = valid && image_pass(pass, q0, A.period, c.kth, ND == 6, xp);
            c.q[0] = xp[0]; c.q[1] = xp[1]; c.q[2] = xp[2];
        }
        if (!ballot(c.act)) continue;
        float qc[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) qc[d] = (float)readfirst_d(c.q[d]);
        const int anc = (pass == 0 && QSRC != NBX_Q_POINTS && !listed) ? T.nleafp + 2 * g : -1;
        knn_walk<ND, HeapLv<KMAX>::LV>(T, c, qc, anc, hd, hid, cx, cid, lane, hc);
    }
    const double kth = c.kth;

    // ------------------------------------------------------------ densities
    // Symmetric scatter: contributions to the wave's own 64 queries (slot
    // runs only) are summed in LDS and flushed with one atomic per query.
    const bool sym = QSRC == NBX_Q_SLOTS && A.dens == NBX_DO_SYM;
    const bool own = sym && !listed;
    double racc_own = 0.0;
    if (own) racc[lane] = 0.0;
    __syncthreads();
    // position queries take the gather forms only (mass and slot of the
    // query are undefined there)
    if (A.dens != NBX_DO_NONE && valid &&
        (QSRC == NBX_Q_SLOTS || A.dens == NBX_DO_GATHER || (QSRC == NBX_Q_POINTS && A.dens == NBX_DO_VEL))) {
        bool bad = false;
        if (A.dens == NBX_DO_VEL || A.dens == NBX_DO_VELPHYS) {
            // CalcVelDensity: |dv| of the Nsearch physical neighbours, the
            // Nsmooth smallest of them, h_v = max/2, sum W(|dv|/h_v) h_v^-3.
            // The kh entries move to positions 0..kh-1 first (sorted order).
            // (CalcVelDensityPosition: the query's velocity comes with it)
            const int ns = A.nsmooth;
            const bool pts = QSRC == NBX_Q_POINTS;
            const double qvx = pts ? A.qv[0][qi] : T.x[3][qi];
            const double qvy = pts ? A.qv[1][qi] : T.x[4][qi];
            const double qvz = pts ? A.qv[2][qi] : T.x[5][qi];
            heap_sort<HeapLv<KMAX>::LV>(hd, hid, lane, hc);
            for (int p = 0; p < kh; ++p) {
                const int32_t id = HI(p + nfill);
                if (id == NBX_NOID) { bad = true; break; }
                const int32_t s = T.inv[id];
                double vx = qvx - T.x[3][s], vy = qvy - T.x[4][s], vz = qvz - T.x[5][s];
                HD(p) = sqrt(vx * vx + vy * vy + vz * vz);
                HI(p) = id;
            }
            if (!bad) {
                for (int p = (ns - 2) / 4; p >= 0; --p) heap_sift(hd, hid, lane, ns, p, HD(p), HI(p));
                for (int p = ns; p < kh; ++p) {
                    const double dv = HD(p);
                    const int32_t iv = HI(p);
                    if (kgt(HD(0), HI(0), dv, iv)) heap_sift(hd, hid, lane, ns, 0, dv, iv);
                }
                double hv = 0.5 * HD(0);
                double norm = 1.0 / pow(hv, 3.0);
                double den = 0.0;
                if (A.dens == NBX_DO_VEL) {
                    for (int p = 0; p < ns; ++p) den = den + wsm_at(HD(p), hv, T.kern, T.kernres) * norm;
                } else {
                    // CalcVelDensityWithPhysDensity (KDCalcSmoothQuantities.cxx:428-466):
                    // each term times W(r_ij / h) h^-3 / rho_i * m_i, with
                    // h = d2 of the Nsearch-th neighbour / 2 (no square root)
                    const double hp = 0.5 * kth;
                    const double pnorm = 1.0 / pow(hp, 3.0);
                    const double rhoi = A.rho_in[qi], mi = T.mass[qi];
                    for (int p = 0; p < ns; ++p) {
                        const int32_t s = T.inv[HI(p)];
                        double dx = T.x[0][qi] - T.x[0][s], dy = T.x[1][qi] - T.x[1][s], dz = T.x[2][qi] - T.x[2][s];
                        double r = sqrt(dx * dx + dy * dy + dz * dz);
                        double vW = wsm_at(HD(p), hv, T.kern, T.kernres) * norm;
                        double W = wsm_at(r, hp, T.kern, T.kernres) * pnorm;
                        den = den + vW * (W / rhoi * mi);
                    }
                }
                A.rho[qi] = den;
                if (A.hout) A.hout[A.hout_by_id ? T.perm[qi] : qi] = hv;
            }
        } else {
            double hi = 0.5 * sqrt(kth);
            double norm = 1.0 / pow(hi, (double)ND);
            const double mi = QSRC == NBX_Q_SLOTS ? T.mass[qi] : 0.0;
            double acc = 0.0;
            // entries in groups of 8 so their slot/mass gathers are in flight
            // together; summed in heap-position order
            for (int p0 = 0; p0 < hc; p0 += 8) {
                int32_t idv[8], sv[8];
                double dv[8], mv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool in = p0 + u < hc;
                    idv[u] = in ? HI(p0 + u) : -1;
                    dv[u] = in ? HD(p0 + u) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) sv[u] = (idv[u] >= 0 && idv[u] != NBX_NOID) ? T.inv[idv[u]] : 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) mv[u] = T.mass[sv[u]];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int32_t id = idv[u];
                    if (id < 0) continue;  // filler or past the end
                    if (id == NBX_NOID) { bad = true; continue; }
                    const int32_t s = sv[u];
                    double W = wsm_at(sqrt(dv[u]), hi, T.kern, T.kernres);
                    if (A.dens == NBX_DO_SYM) {
                        double Wij = 0.5 * W * norm;
                        acc = acc + Wij * mv[u];
                        const uint32_t r = (uint32_t)(s - g * NBX_WAVE);
                        if (own && r < NBX_WAVE) atomicAdd(&racc[r], Wij * mi);
                        else unsafeAtomicAdd(A.rho + s, Wij * mi);
                    } else {
                        double Wij = W * norm;
                        acc += Wij * mv[u];
                    }
                }
            }
            if (A.dens == NBX_DO_SYM) racc_own = acc;
            else A.rho[qi] = acc;
            if (A.hout) A.hout[A.hout_by_id ? T.perm[qi] : qi] = hi;
        }
        if (bad) atomicOr(A.err, 1);
    }
    __syncthreads();
    if (sym && valid) unsafeAtomicAdd(A.rho + qi, racc_own + (own ? racc[lane] : 0.0));

    // ------------------------------------------------------------ outputs
    if (A.nn == nullptr && A.d2 == nullptr) return;
    // (the velocity density sorted already and reused the entries)
    if (valid && A.dens != NBX_DO_VEL && A.dens != NBX_DO_VELPHYS) heap_sort<HeapLv<KMAX>::LV>(hd, hid, lane, hc);
    __syncthreads();
    const int kout = A.kout;
    const int off = ((A.drop == NBX_DROP_SMALLEST) ? 1 : 0) + ((A.dens == NBX_DO_VEL || A.dens == NBX_DO_VELPHYS) ? 0 : nfill);
    const int total = NBX_WAVE * kout;
    for (int e = lane; e < total; e += NBX_WAVE) {
        int r = e / kout, t = e - r * kout;
        int qr = g * NBX_WAVE + r;
        if (qr >= nq) break;
        const int lr = qr;
        if (listed) qr = A.qlist[qr];
        int64_t row;
        if (QSRC == NBX_Q_POINTS) row = qr;
        else if (listed && A.list_rows) row = lr;
        else if (A.layout == 1) row = qr - A.row0;
        else row = T.perm[qr];
        const int32_t id = hid[(t + off) * NBX_WAVE + r];
        if (A.nn) {
            int32_t v = -1;
            if (id != NBX_NOID) v = (QSRC == NBX_Q_POINTS || A.layout == 0) ? id : T.inv[id];
            A.nn[row * kout + t] = v;
        }
        if (A.d2) A.d2[row * kout + t] = hd[(t + off) * NBX_WAVE + r];
    }
}

template <int ND, int KMAX>
static void launch_q(const DevTree &T, const KnnArgs &A, int blocks, hipStream_t st) {
    if (A.qsrc == NBX_Q_SLOTS)
        hipLaunchKernelGGL((k_knn<ND, KMAX, NBX_Q_SLOTS>), dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
    else if (A.qsrc == NBX_Q_PERIODIC)
        hipLaunchKernelGGL((k_knn<ND, KMAX, NBX_Q_PERIODIC>), dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
    else
        hipLaunchKernelGGL((k_knn<ND, KMAX, NBX_Q_POINTS>), dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
}

hipError_t nbx_launch_knn(const DevTree &T, const KnnArgs &A, int nqueries, hipStream_t st) {
    int blocks = (nqueries + NBX_WAVE - 1) / NBX_WAVE;
    if (blocks == 0) return hipSuccess;
    ProfScope ps(A.qlist ? "k_knn_fallback" : "k_knn_fp64", st);
    if (T.nd == 6) {
        if (A.kh <= 65) launch_q<6, 65>(T, A, blocks, st);
        else if (A.kh <= 129) launch_q<6, 129>(T, A, blocks, st);
        else return hipErrorInvalidValue;
    } else {
        if (A.kh <= 33) launch_q<3, 33>(T, A, blocks, st);
        else if (A.kh <= 65) launch_q<3, 65>(T, A, blocks, st);
        else if (A.kh <= 129) launch_q<3, 129>(T, A, blocks, st);
        else return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
