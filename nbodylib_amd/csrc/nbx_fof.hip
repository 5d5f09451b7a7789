// nbx_fof.hip -- friends-of-friends on gfx950.
//
// Replaces KDTree::FOF (KDFOF.cxx:29-153) with its node routines
// SplitNode/LeafNode::FOFSearchBall (KDSplitNode.cxx:921-988,
// KDLeafNode.cxx:527-590) and the periodic image search
// (KDSplitNode.cxx:1602-1646).  The reference grows one group at a time by a
// serial breadth-first search; here every particle links concurrently into a
// lock-free union-find over slots, which yields the same partition (the
// transitive closure of d2 < b^2), then groups are ranked canonically.
//
// Exact link shortcuts (the reference's enclosure test, KDNode.h:279-303):
//  * a leaf whose box diagonal is < b is internally connected (pre-pass);
//  * a query whose farthest-corner distance to such a leaf is < b links to
//    the whole leaf with one union;
//  * when the farthest distance between the wave's 64 particles and a node
//    is < b, all cross pairs link, so one union joins the node to the wave.
// All bounds are monotone in the roundings (nbx_internal.h), so a shortcut
// is taken only when every computed particle distance is below b^2.
//
// On a 6-D (TPHS) tree the link distance is the reference's
// DistanceSqd(positions) + DistanceSqd(velocities) (KDLeafNode.cxx:570-571):
// two sums of three squares, added.  Every particle distance and every box
// bound here is summed with that same association, so the bounds stay
// monotone.  Periodic images shift the positions only (KDSplitNode.cxx:
// 1609-1641).
//
// The same kernels run KDTree::FOFCriterion (KDFOF.cxx:157-265) for the
// built-in comparators FOF3d and FOF6d (FOFFunc.h:30-55): the pair value is
// then the comparator's scaled sum, computed in its order, and links when
// below 1 (fof_val).  Its periodic search (KDSplitNode.cxx:1649-1680) tries
// all seven images; here they are tried when the root's bound allows a link,
// which changes no link decision.
#include "nbx_kernels.h"

// Sum of ND per-dimension squares in the reference's association: one
// running sum for 3-D, position and velocity sums added for 6-D.
template <int ND>
__device__ __forceinline__ double fof_sum(const double (&t2)[ND]) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) s += t2[d];
    if constexpr (ND == 6) {
        double v = 0.0;
#pragma unroll
        for (int d = 3; d < 6; ++d) v += t2[d];
        s += v;
    }
    return s;
}

// Link value of one pair from its per-dimension squares; a pair links when
// the value is below L.lim.  MT 0: the plain search, d2 against b^2.  MT 1
// (FOF3d) and 2 (FOF6d): each square divided by its comparator scale
// (params[6] for positions, params[7] for velocities) and summed in the
// comparator's order (x0 x1 x2, or x0 v0 x1 v1 x2 v2), against 1.  Every
// step is monotone, so a box bound run through it stays a bound.
template <int ND, int MT>
__device__ __forceinline__ double fof_val(const double (&t2)[ND], const FofLink &L) {
    if constexpr (MT == 0) {
        return fof_sum<ND>(t2);
    } else if constexpr (MT == 1) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) s += t2[d] / L.s6;
        return s;
    } else {
        static_assert(ND == 6, "FOF6d links in phase space");
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s += t2[d] / L.s6;
            s += t2[3 + d] / L.s7;
        }
        return s;
    }
}

// The comparator sum with each division replaced by a multiplication by the
// correctly rounded reciprocal.  Term by term that is within 3.01u of the
// quotient (u = 2^-53), and the two sums, of at most six nonnegative terms
// added in the same order, are within 13.2u of each other.  So a value below
// 1 - 2^-46 proves the comparator's value below 1, and one at or above
// 1 + 2^-46 proves it not; only values in between need the divisions.
template <int ND, int MT>
__device__ __forceinline__ double fof_val_fast(const double (&t2)[ND], const FofLink &L) {
    double s = 0.0;
    if constexpr (MT == 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) s += t2[d] * L.r6;
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s += t2[d] * L.r6;
            s += t2[3 + d] * L.r7;
        }
    }
    return s;
}
#define NBX_FOF_BAND_LO (1.0 - 0x1p-46)
#define NBX_FOF_BAND_HI (1.0 + 0x1p-46)
// Box bounds for the comparators use the reciprocal sum too, read through
// the band: a node is skipped only when its lower bound is surely >= 1
// (fof_maybe fails), and a wholesale link is taken only when the upper bound
// is surely < 1 (fof_surely).  Pair decisions stay exact either way.
template <int ND, int MT>
__device__ __forceinline__ double fof_bval(const double (&t2)[ND], const FofLink &L) {
    if constexpr (MT == 0) return fof_sum<ND>(t2);
    else return fof_val_fast<ND, MT>(t2, L);
}
template <int MT>
__device__ __forceinline__ bool fof_maybe(double v, const FofLink &L) {
    return MT == 0 ? v < L.lim : v < NBX_FOF_BAND_HI;
}
template <int MT>
__device__ __forceinline__ bool fof_surely(double v, const FofLink &L) {
    return MT == 0 ? v < L.lim : v < NBX_FOF_BAND_LO;
}

template <int ND, int MT>
__device__ __forceinline__ double fof_lb(const double *q, cdptr b, const FofLink &L) {
    double t2[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const double t = fmax(fmax(b[d] - q[d], q[d] - b[ND + d]), 0.0);
        t2[d] = t * t;
    }
    return fof_bval<ND, MT>(t2, L);
}
template <int ND, int MT>
__device__ __forceinline__ double fof_ub(const double *q, cdptr b, const FofLink &L) {
    double t2[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const double a = q[d] - b[d], c = b[ND + d] - q[d];
        const double a2 = a * a, c2 = c * c;
        t2[d] = (a2 > c2) ? a2 : c2;
    }
    return fof_bval<ND, MT>(t2, L);
}
template <int ND, int MT>
__device__ __forceinline__ double fof_bbub(cdptr a, cdptr b, const FofLink &L) {
    double t2[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const double u = b[ND + d] - a[d], v = a[ND + d] - b[d];
        const double m = u > v ? u : v;
        t2[d] = m * m;
    }
    return fof_bval<ND, MT>(t2, L);
}

template <int ND, int MT>
__global__ __launch_bounds__(256) void k_fof_leafinit(DevTree T, FofLink L, int32_t *P, uint8_t *conn) {
    const int32_t leaf = blockIdx.x * 256 + threadIdx.x;
    if (leaf >= T.nleaf) return;
    const int node = T.nleafp + leaf;
    cdptr bx = (cdptr)node_box(T, node, ND);
    int32_t s0 = leaf * NBX_LEAF;
    int32_t s1 = s0 + NBX_LEAF < T.n ? s0 + NBX_LEAF : T.n;
    bool c = fof_surely<MT>(fof_bbub<ND, MT>(bx, bx, L), L);
    conn[node] = c ? 1 : 0;
    for (int32_t s = s0; s < s1; ++s) P[s] = c ? s0 : s;
}

// Internal node whose diagonal is < b: both children are connected (their
// diagonals are smaller) and every cross pair links, so one union connects
// it.  Nodes [first, first+count).  An empty box (padding) has
// an infinite diagonal and is never connected.
template <int ND, int MT>
__global__ __launch_bounds__(256) void k_fof_nodeinit(DevTree T, FofLink L, int32_t *P, uint8_t *conn, int32_t first,
                                                       int32_t count) {
    int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const int node = first + i;
    cdptr bx = (cdptr)node_box(T, node, ND);
    bool c = fof_surely<MT>(fof_bbub<ND, MT>(bx, bx, L), L);
    conn[node] = c ? 1 : 0;
    if (!c) return;
    const int right = 2 * node + 1;
    cdptr br = (cdptr)node_box(T, right, ND);
    if (br[0] <= br[ND]) uf_union(P, node_first(T, node), node_first(T, right));
}

// Leaf scan for a wave in which few lanes want the leaf: their positions are
// compacted into LDS and each step tests two of them against the leaf's 32
// candidates, one candidate per lane (lanes 0-31 the first position, 32-63
// the second), so the wave takes ceil(few / 2) steps instead of 32.  Each
// pair is computed as the scan by lanes computes it (q - x, squared, summed
// in the same order; band values decided by the divisions), so the link
// decisions are the same.  Returns the lane's mask, candidate j at bit j.
template <int ND, int MT>
__device__ __forceinline__ uint32_t fof_scan_few(const DevTree &T, int base, const double *q, bool pair,
                                                 uint64_t pm, const FofLink &L) {
    __shared__ double qs[ND][NBX_WAVE];
    const int lane = threadIdx.x;  // one wave per workgroup
    const int npair = __builtin_popcountll(pm);
    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
    if (pair) {
#pragma unroll
        for (int d = 0; d < ND; ++d) qs[d][rank] = q[d];
    }
    __syncthreads();
    const int c = lane & (NBX_LEAF - 1), h = lane >> 5;
    double cx[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) cx[d] = T.x[d][base + c];
    uint32_t own = 0;
    for (int s = 0; s < npair; s += 2) {
        const int r = s + h;
        const int rr = r < npair ? r : npair - 1;  // the second half idles on an odd tail
        double t2[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const double t = qs[d][rr] - cx[d];
            t2[d] = t * t;
        }
        bool near;
        if constexpr (MT == 0) {
            near = fof_val<ND, MT>(t2, L) < L.lim;
        } else {
            const double v = fof_val_fast<ND, MT>(t2, L);
            near = v < NBX_FOF_BAND_LO;
            const bool uns = v >= NBX_FOF_BAND_LO && v < NBX_FOF_BAND_HI;
            if (ballot(uns) && uns) near = fof_val<ND, MT>(t2, L) < 1.0;
        }
        const uint64_t b = ballot(near && r < npair);
        if (rank == s) own = (uint32_t)b;
        if (rank == s + 1) own = (uint32_t)(b >> 32);
    }
    __syncthreads();  // qs is written again at the next leaf
    return pair ? own : 0u;
}
// Largest count of wanting lanes for which the leaf is scanned by
// fof_scan_few: the measured crossover with the scan by lanes on 3-D trees;
// on 6-D trees, where a leaf is wanted by 6 lanes on average, every count
// (measured faster throughout).
template <int ND>
__device__ __forceinline__ constexpr int fof_few_max() { return ND == 6 ? NBX_WAVE : 24; }

// Links from the (possibly image-shifted) position q to the particles of one
// leaf.  act is the lane's own bound test for the leaf: a lane whose bound
// rules out every link takes no part.  A leaf that is internally connected
// needs one union with any particle it links to; otherwise every link is a
// union (in the first pass only toward larger slots: the pair is met from
// both sides).  rq is a hint for the root of qi (any former root of its
// set), kept across unions so each find starts near the top.
template <int PASS0, int ND, int MT>
__device__ __forceinline__ void fof_leaf(const DevTree &T, int leaf, const double *q, bool act, int32_t qi,
                                         const FofLink &L, int32_t *P, const uint8_t *conn, int32_t &rq) {
    const int base = leaf * NBX_LEAF;
    const bool lconn = conn[T.nleafp + leaf];
    cdptr bx = (cdptr)node_box(T, T.nleafp + leaf, ND);
    const bool encl = act && lconn && fof_surely<MT>(fof_ub<ND, MT>(q, bx, L), L);
    // The members of a connected leaf are one set (from the pre-pass, or
    // joined by the shortcut that flagged it before the pass ends), so once
    // its first slot's parent is rq, or it is rq, the lane is joined to all
    // of it and neither scans the leaf nor makes a union; in dense regions
    // most visits to connected leaves are such.  A stale parent only misses
    // the skip, since sets only merge.
    const int32_t pb = lconn ? P[base] : -1;
    const bool joined = lconn && (pb == rq || base == rq);
    if (encl && !joined) rq = uf_union_r(P, rq, base);
    const bool pair = act && !encl && !joined;
    const uint64_t pm = ballot(pair);
    if (!pm) return;
    uint32_t mask = 0;
    if (__builtin_popcountll(pm) <= fof_few_max<ND>()) {
        mask = fof_scan_few<ND, MT>(T, base, q, pair, pm, L);
    } else {
        uint32_t unsure = 0;
        // the scan by lanes: candidates arrive as 64-byte (3-D) or 32-byte
        // (6-D) scalar-load chunks per dimension (padding slots are 1e300
        // away and never link)
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
                double t2[ND];
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double t = q[d] - cv[d][u];
                    t2[d] = t * t;
                }
                // decisions are shifted in as they are made (a bit per candidate
                // held in scalar masks until the end would spill them all); the
                // comparators' values inside the reciprocal band are decided
                // with the divisions after the scan, or at once in the 6-D
                // persistent kernel (each the measured better form)
                bool near;
                if constexpr (MT == 0) {
                    near = fof_val<ND, MT>(t2, L) < L.lim;
                } else if constexpr (ND == 3) {
                    const double v = fof_val_fast<ND, MT>(t2, L);
                    near = v < NBX_FOF_BAND_LO;
                    unsure = unsure + unsure + ((v >= NBX_FOF_BAND_LO && v < NBX_FOF_BAND_HI) ? 1u : 0u);
                } else {
                    const double v = fof_val_fast<ND, MT>(t2, L);
                    near = v < NBX_FOF_BAND_LO;
                    const bool uns = v >= NBX_FOF_BAND_LO && v < NBX_FOF_BAND_HI;
                    if (ballot(uns) && uns) near = fof_val<ND, MT>(t2, L) < 1.0;
                }
                mask = mask + mask + (near ? 1u : 0u);
            }
        }
        if constexpr (MT != 0 && ND == 3) {
            // unsure holds candidate j at bit 31 - j, as mask does here
            if (ballot(unsure != 0)) {
                uint32_t um = unsure;
                while (um) {
                    const int b = 31 - __builtin_clz(um);
                    um &= ~(1u << b);
                    const int32_t cs = base + (31 - b);
                    double t2[ND];
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        const double t = q[d] - T.x[d][cs];
                        t2[d] = t * t;
                    }
                    if (fof_val<ND, MT>(t2, L) < 1.0) mask |= 1u << b;
                }
            }
        }
        // candidate j was shifted in at bit 31 - j
        mask = __builtin_bitreverse32(mask);
    }
    // Never the query itself; in the first pass, in a leaf not internally
    // connected, only candidates at larger slots (the pair is met from both
    // sides).
    const int qrel = qi - base;
    const uint32_t self = (qrel >= 0 && qrel < NBX_LEAF) ? (1u << qrel) : 0u;
    const uint32_t above = qrel < 0 ? ~0u : (qrel >= NBX_LEAF - 1 ? 0u : (~0u << (qrel + 1)));
    mask &= (PASS0 && !lconn) ? above : ~self;
    if (!pair) mask = 0;
    if (!ballot(mask != 0)) return;
    // The leaf's parents, read once with ordinary loads.  A candidate whose
    // parent is the lane's root hint rq (a former root of its set) is
    // already in that set, because sets only ever merge, so its link needs
    // no union.  A stale parent costs only the skip.
    __shared__ int32_t par[NBX_LEAF];
    const int ln = threadIdx.x;  // one wave per workgroup
    if (ln < NBX_LEAF) par[ln] = base + ln < T.n ? P[base + ln] : -1;
    __syncthreads();
    if (lconn) {
        // one link joins the whole leaf; the union goes through the first
        // slot, whose chain to the root is the shortest
        if (mask) {
            const int j = __builtin_ctz(mask);
            if (par[j] != rq && par[0] != rq && base != rq) rq = uf_union_r(P, rq, base);
        }
    } else {
        while (mask) {
            const int j = __builtin_ctz(mask);
            mask &= mask - 1;
            if (par[j] == rq) continue;
            rq = uf_union_r(P, rq, base + j);
        }
    }
    __syncthreads();
}

template <int PASS0, int ND, int MT>
__device__ __forceinline__ void fof_walk(const DevTree &T, const double *q, bool act, int32_t qi, int g,
                                         const FofLink &L, int32_t *P, uint8_t *conn, int32_t &rq) {
    const int D = T.depth;
    // the wave's own 64 particles form the parent of leaves 2g, 2g+1 (or the
    // single leaf when the tree has one)
    const int gnode = D >= 1 ? (T.nleafp + 2 * g) >> 1 : 1;
    cdptr gb = (cdptr)node_box(T, gnode, ND);
    // In the first pass a node whose slots all lie below the wave's first
    // slot need not be entered.  Each of its links with a wave particle is
    // also met from the lower particle's wave, whose walk reaches the wave's
    // leaves (a node holding a slot above that wave's own is never skipped
    // there): the pair's computed distance is the same from both sides,
    // fl(a - b) = -fl(b - a), and each route that side can take (scan,
    // enclosure of a connected leaf, wave-node shortcut) puts the pair in
    // one set.  On a 6-D tree every such node is skipped; on a 3-D tree
    // only such leaves, because there the wave-node shortcuts met in the
    // lower subtrees join dense regions in fewer unions than their pairs'
    // links do (the measured better form for each).  sh is the node's
    // height, so its leaves are [node << sh, (node + 1) << sh) in heap
    // order.
    const int lfirst = T.nleafp + 2 * g;
    int node = 1, sh = D;
    bool down = true;
    for (;;) {
        if (down && PASS0 && (ND == 6 || sh == 0) && ((node + 1) << sh) <= lfirst) down = false;
        if (down) {
            cdptr bx = (cdptr)node_box(T, node, ND);
            bool want = act && fof_maybe<MT>(fof_lb<ND, MT>(q, bx, L), L);
            if (ballot(want)) {
                if (PASS0 && fof_surely<MT>(fof_bbub<ND, MT>(gb, bx, L), L)) {
                    // every wave particle links to every particle of the node:
                    // wave and node form one component
                    const int32_t nfirst = node_first(T, node);
                    if (act) rq = uf_union_r(P, rq, nfirst);
                    if (!conn[node]) {
                        const int32_t nlast = node_end(T, node);
                        for (int32_t s = nfirst + 1 + threadIdx.x; s < nlast; s += NBX_WAVE) uf_union(P, s, nfirst);
                        conn[node] = 1;
                    }
                    down = false;
                } else if (node >= T.nleafp) {
                    fof_leaf<PASS0, ND, MT>(T, node - T.nleafp, q, want, qi, L, P, conn, rq);
                    down = false;
                } else {
                    node = 2 * node;
                    --sh;
                    continue;
                }
            } else {
                down = false;
            }
        }
        if (node == 1) break;
        if (!(node & 1)) {
            node ^= 1;
            down = true;
        } else {
            node >>= 1;
            ++sh;
        }
    }
}

// Links of one wave of 64 slot-consecutive particles (wave index g).
template <int ND, int MT>
__device__ __forceinline__ void fof_task(const DevTree &T, const FofArgs &A, uint8_t *conn, int g) {
    const int lane = threadIdx.x;
    const int32_t qi = g * NBX_WAVE + lane;
    const bool valid = qi < T.n;
    double q0[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) q0[d] = valid ? T.x[d][qi] : 0.0;
    int32_t rq = valid ? qi : 0;
    const FofLink L = A.link;
    fof_walk<1, ND, MT>(T, q0, valid, qi, g, L, A.parent, conn, rq);
    if (!A.periodic) return;
    // FOFSearchBallPeriodic: image toward the far wall when b^2 > s^2;
    // FOFSearchCriterionPeriodic: every image, left to the walk's bounds
    const double *per = A.period;
    for (int pass = 1; pass < 8; ++pass) {
        double xp[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) xp[d] = q0[d];
        double sval = 0.0;
        if (pass <= 3) {
            int k = pass - 1;
            if (q0[k] < per[k] / 2.0) { xp[k] = q0[k] + per[k]; sval = q0[k]; }
            else { xp[k] = q0[k] - per[k]; sval = -xp[k]; }
        } else {
            int k1 = (pass == 6) ? 1 : 0, k2 = (pass == 4) ? 1 : 2;
            for (int k = 0; k < 3; ++k) {
                bool use = (pass == 7) || k == k1 || k == k2;
                if (!use) continue;
                if (q0[k] < per[k] / 2.0) { xp[k] = q0[k] + per[k]; sval += q0[k] * q0[k]; }
                else { xp[k] = q0[k] - per[k]; sval += xp[k] * xp[k]; }
            }
            sval = sqrt(sval);
        }
        bool act = valid && (MT != 0 || L.lim > sval * sval);
        if (!ballot(act)) continue;
        fof_walk<0, ND, MT>(T, xp, act, qi, g, L, A.parent, conn, rq);
    }
}

// Linking kernel for 3-D trees: one wave per workgroup in launch order, so
// the hardware deals consecutive waves round robin to the XCDs.
template <int ND, int MT>
__global__ __launch_bounds__(64) void k_fof_waves(DevTree T, FofArgs A, uint8_t *conn) {
    fof_task<ND, MT>(T, A, conn, blockIdx.x);
}

// Persistent linking kernel for 6-D trees, one wave per workgroup.  Workgroup b stays on
// XCD (b mod 8) and takes waves from that XCD's contiguous share of the tree
// order through the XCD's counter, so the nodes and leaves a neighbourhood's
// waves share are fetched into one XCD's L2; once its share is drained it
// takes waves from the other XCDs' counters in turn, which evens out the
// shares' unequal cost (dense halos).  Every ticket names one wave, so every
// wave runs once; a workgroup exits when all eight counters are past their
// shares' ends.
#define FOF_CTR_STRIDE 16  // XCD counters 64 bytes apart
template <int ND, int MT>
__global__ __launch_bounds__(64) void k_fof(DevTree T, FofArgs A, uint8_t *conn, int32_t *ctr) {
    const int nw = (T.n + NBX_WAVE - 1) / NBX_WAVE;
    const int x = blockIdx.x & 7, q8 = nw >> 3, r8 = nw & 7;
    for (int s = 0; s < 8; ++s) {
        const int y = (x + s) & 7;
        const int ybeg = y * q8 + (y < r8 ? y : r8), ycnt = q8 + (y < r8 ? 1 : 0);
        for (;;) {
            int t = 0;
            if (threadIdx.x == 0) t = atomicAdd(ctr + y * FOF_CTR_STRIDE, 1);
            t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
            if (t >= ycnt) break;
            fof_task<ND, MT>(T, A, conn, ybeg + t);
        }
    }
}

// Resident workgroups of one k_fof instance, a multiple of 8 (every XCD
// gets workgroups) and at least 8.
template <int ND, int MT>
static int fof_grid() {
    static int grid = 0;
    if (grid == 0) {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_fof<ND, MT>, NBX_WAVE, 0) != hipSuccess || per <= 0)
            per = 8;
        const int gsz = (per * ncu) & ~7;
        grid = gsz < 8 ? 8 : gsz;
    }
    return grid;
}

// Group sizes and smallest member IDs.  Neighbouring slots mostly share a
// root, so the lanes whose root equals the first active lane's are summed
// in the wave and counted with one atomic pair; the rest count alone.
__global__ __launch_bounds__(256) void k_fof_compress(int32_t *P, int64_t n, const int32_t *perm, int32_t *cnt,
                                                        int32_t *minid) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool in = i < n;
    const int32_t r = in ? uf_find(P, (int32_t)i) : -1;
    const int32_t id = in ? perm[i] : 0x7fffffff;
    const uint64_t act = ballot(in);
    if (!act) return;
    const int lead = __builtin_ctzll(act);
    const int32_t r0 = __shfl(r, lead);
    const bool same = in && r == r0;
    const uint64_t m = ballot(same);
    int32_t mn = same ? id : 0x7fffffff;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mn = min(mn, __shfl_xor(mn, off));
    if ((int)(threadIdx.x & 63) == lead) {
        atomicAdd(cnt + r0, (int32_t)__popcll(m));
        atomicMin(minid + r0, mn);
    }
    if (in && !same) {
        atomicAdd(cnt + r, 1);
        atomicMin(minid + r, id);
    }
}
__global__ __launch_bounds__(256) void k_fof_store_root(int32_t *P, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    P[i] = uf_find(P, (int32_t)i);
}
__global__ __launch_bounds__(256) void k_fof_collect(const int32_t *P, const int32_t *cnt, const int32_t *minid,
                                                       int64_t n, int32_t minnum, uint64_t *keys, int32_t *vals,
                                                       int32_t *ng) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (P[i] != i || cnt[i] < minnum) return;
    int32_t at = atomicAdd(ng, 1);
    keys[at] = ((uint64_t)(0xffffffffu - (uint32_t)cnt[i]) << 32) | (uint32_t)minid[i];
    vals[at] = (int32_t)i;
}
__global__ __launch_bounds__(256) void k_fof_rank(const int32_t *vals, int64_t ng, int32_t *lab) {
    int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= ng) return;
    lab[vals[r]] = (int32_t)(r + 1);
}
__global__ __launch_bounds__(256) void k_fof_out(const int32_t *P, const int32_t *lab, const int32_t *perm, int64_t n,
                                                   int32_t *out) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[perm[i]] = lab[P[i]];
}

static inline int nblk(int64_t n) { return (int)((n + 255) / 256); }

// ---------------------------------------------------------------- ball search, many positions
// SearchBallPosTagged(x, r2) (KDFindNearest.cxx:618-688) for m positions at
// once: one wave per 64 positions walks the union of the nodes any lane's
// ball reaches (per-lane FP64 box bounds, the reference's strict d2 < r2 with
// x -> z sums), then its periodic images (SearchBallPosPeriodicTagged,
// KDSplitNode.cxx:1388-1414: image when r2 > s^2, the position shifted
// first).  FILL 0 counts hits per position; FILL 1 writes each position's IDs
// (and d2) from its offset.  Positions are taken in the given walk order
// (Morton order of the positions, so a wave's balls are neighbours).
template <int FILL>
__device__ __forceinline__ void ball_leaf(const DevTree &T, int leaf, const double *q, bool act, double r2,
                                          const BallManyArgs &A, int64_t &at) {
    const int base = leaf * NBX_LEAF;
    typedef double dxc __attribute__((ext_vector_type(8)));
    typedef const __attribute__((address_space(4))) dxc *cdcptr;
#pragma unroll
    for (int jj = 0; jj < NBX_LEAF; jj += 8) {
        dxc cv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) cv[d] = *(cdcptr)(T.x[d] + base + jj);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            double s = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double t = q[d] - cv[d][u];
                s += t * t;
            }
            if (act && s < r2) {  // padding slots are 1e300 away
                if (FILL) {
                    A.ids[at] = T.perm[base + jj + u];
                    if (A.d2) A.d2[at] = s;
                }
                ++at;
            }
        }
    }
}

template <int FILL>
__device__ __forceinline__ void ball_walk(const DevTree &T, const double *q, bool act, double r2,
                                          const BallManyArgs &A, int64_t &at) {
    int node = 1;
    bool down = true;
    for (;;) {
        if (down) {
            cdptr bx = (cdptr)node_box(T, node, 3);
            const bool want = act && box_lb<3>(q, bx) < r2;
            if (ballot(want)) {
                if (node >= T.nleafp) {
                    ball_leaf<FILL>(T, node - T.nleafp, q, want, r2, A, at);
                    down = false;
                } else {
                    node = 2 * node;
                    continue;
                }
            } else {
                down = false;
            }
        }
        if (node == 1) break;
        if (!(node & 1)) {
            node ^= 1;
            down = true;
        } else {
            node >>= 1;
        }
    }
}

__device__ __forceinline__ uint64_t morton_spread(uint64_t v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}
__global__ __launch_bounds__(256) void k_morton_keys(const double *qx, const double *qy, const double *qz, int64_t m,
                                                     double lo0, double lo1, double lo2, double s0, double s1,
                                                     double s2, uint64_t *keys, int32_t *idx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const double q[3] = {qx[i], qy[i], qz[i]}, lo[3] = {lo0, lo1, lo2}, sc[3] = {s0, s1, s2};
    uint64_t key = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double t = fmin(fmax((q[d] - lo[d]) * sc[d], 0.0), 2097151.0);
        key |= morton_spread((uint64_t)t) << (2 - d);
    }
    keys[i] = key;
    idx[i] = (int32_t)i;
}

template <int FILL>
__global__ __launch_bounds__(64) void k_ball_many(DevTree T, BallManyArgs A) {
    const int64_t r = (int64_t)blockIdx.x * NBX_WAVE + threadIdx.x;
    const bool valid = r < A.m;
    const int64_t qi = valid ? (A.order ? (int64_t)A.order[r] : r) : 0;
    double q0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) q0[d] = valid ? A.q[d][qi] : 0.0;
    int64_t at = (FILL && valid) ? A.off[qi] : 0;
    ball_walk<FILL>(T, q0, valid, A.r2, A, at);
    if (A.periodic) {
        const double *per = A.period;
        for (int pass = 1; pass < 8; ++pass) {
            double xp[3] = {q0[0], q0[1], q0[2]};
            double sval = 0.0;
            if (pass <= 3) {
                const int k = pass - 1;
                if (q0[k] < per[k] / 2.0) { xp[k] = q0[k] + per[k]; sval = q0[k]; }
                else { xp[k] = q0[k] - per[k]; sval = -xp[k]; }
            } else {
                const int k1 = (pass == 6) ? 1 : 0, k2 = (pass == 4) ? 1 : 2;
                for (int k = 0; k < 3; ++k) {
                    if (!(pass == 7 || k == k1 || k == k2)) continue;
                    if (q0[k] < per[k] / 2.0) { xp[k] = q0[k] + per[k]; sval += q0[k] * q0[k]; }
                    else { xp[k] = q0[k] - per[k]; sval += xp[k] * xp[k]; }
                }
                sval = sqrt(sval);
            }
            const bool act = valid && A.r2 > sval * sval;
            if (!ballot(act)) continue;
            ball_walk<FILL>(T, xp, act, A.r2, A, at);
        }
    }
    if (!FILL && valid) A.count[qi] = at;
}

hipError_t nbx_launch_ball_many(const DevTree &T, const BallManyArgs &A, bool fill, hipStream_t st) {
    const int blocks = (int)((A.m + NBX_WAVE - 1) / NBX_WAVE);
    if (blocks == 0) return hipSuccess;
    ProfScope ps("k_ball_many", st);
    if (fill) hipLaunchKernelGGL(k_ball_many<1>, dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
    else hipLaunchKernelGGL(k_ball_many<0>, dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
    return hipGetLastError();
}

template <int ND, int MT>
static void launch_fof_t(const DevTree &T, const FofArgs &A, uint8_t *conn, hipStream_t st) {
    {
        ProfScope ps("k_fof_init", st);
        hipLaunchKernelGGL((k_fof_leafinit<ND, MT>), dim3(nblk(T.nleaf)), dim3(256), 0, st, T, A.link, A.parent, conn);
        // every internal node at once: a node's union depends only on its own
        // box, and unions commute
        if (T.nleafp > 1)
            hipLaunchKernelGGL((k_fof_nodeinit<ND, MT>), dim3(nblk((int64_t)T.nleafp - 1)), dim3(256), 0, st, T, A.link,
                               A.parent, conn, 1, T.nleafp - 1);
    }
    ProfScope ps("k_fof_link", st);
    if (ND == 3) {
        // measured on 16.8M clustered particles: 10.2 ms per call in this
        // order, 11.1 ms with the XCD shares below (before the skip of
        // joined leaves); on the 6-D set the other way round (FOF6d 15.3 ms
        // with the shares, 17.0 ms in this order)
        hipLaunchKernelGGL((k_fof_waves<ND, MT>), dim3((T.n + NBX_WAVE - 1) / NBX_WAVE), dim3(NBX_WAVE), 0, st, T, A,
                           conn);
        return;
    }
    // the XCD queue counters sit after the node flags, 64-byte aligned
    int32_t *ctr = (int32_t *)(((uintptr_t)(conn + 2 * (int64_t)T.nleafp) + 63) & ~(uintptr_t)63);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, ctr, 0, (int64_t)(8 * FOF_CTR_STRIDE));
    hipLaunchKernelGGL((k_fof<ND, MT>), dim3(fof_grid<ND, MT>()), dim3(NBX_WAVE), 0, st, T, A, conn, ctr);
}

// Bytes of the parent array, node flags and queue counters that
// nbx_launch_fof expects at A.parent.
size_t nbx_fof_work_bytes(const DevTree &T) {
    return sizeof(int32_t) * (size_t)T.n + 2 * (size_t)T.nleafp + 64 + 8 * FOF_CTR_STRIDE * sizeof(int32_t);
}

hipError_t nbx_launch_fof(const DevTree &T, const FofArgs &A, hipStream_t st) {
    // per-node connectivity flags live right after the parent array (allocated by the caller)
    uint8_t *conn = (uint8_t *)(A.parent + T.n);
    if (A.metric == NBX_FOFM_PLAIN && T.nd == 3) launch_fof_t<3, NBX_FOFM_PLAIN>(T, A, conn, st);
    else if (A.metric == NBX_FOFM_PLAIN && T.nd == 6) launch_fof_t<6, NBX_FOFM_PLAIN>(T, A, conn, st);
    else if (A.metric == NBX_FOFM_3D && T.nd == 3) launch_fof_t<3, NBX_FOFM_3D>(T, A, conn, st);
    else if (A.metric == NBX_FOFM_6D && T.nd == 6) launch_fof_t<6, NBX_FOFM_6D>(T, A, conn, st);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// The canonical labelling's work arrays, carved in order from one block:
// per-particle counts, minimum IDs and labels, the group counter, and the
// (key, group) pairs of at most n / minnum + 1 groups with their sort's
// temporary storage.
struct FofCanonWork {
    int64_t cap;
    size_t off[9], sort_bytes, total;
    FofCanonWork(int64_t n, int32_t minnum) {
        cap = n / (minnum > 1 ? minnum : 1) + 1;
        sort_bytes = 0;
        nbx_sort_pairs_u64(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, cap, nullptr);
        const size_t sz[9] = {sizeof(int32_t) * (size_t)n, sizeof(int32_t) * (size_t)n,
                              sizeof(int32_t) * (size_t)n, sizeof(int32_t),
                              sizeof(uint64_t) * (size_t)cap, sizeof(uint64_t) * (size_t)cap,
                              sizeof(int32_t) * (size_t)cap, sizeof(int32_t) * (size_t)cap, sort_bytes};
        size_t at = 0;
        for (int i = 0; i < 9; ++i) {
            off[i] = at;
            at = (at + sz[i] + 255) & ~(size_t)255;
        }
        total = at;
    }
};

size_t nbx_fof_canonical_bytes(int64_t n, int32_t minnum) { return FofCanonWork(n, minnum).total; }

// Canonical labels: groups of >= minnum members numbered 1..G by size
// descending, ties by smaller minimum member ID; everything else 0.  work:
// device memory of nbx_fof_canonical_bytes(n, minnum), 256-byte aligned.
hipError_t nbx_fof_canonical(const DevTree &T, int32_t *P, int32_t minnum, int32_t *labels_by_id, int64_t *ngroups,
                             void *work, hipStream_t st) {
    int64_t n = T.n;
    const FofCanonWork W(n, minnum);
    char *wb = (char *)work;
    int32_t *cnt = (int32_t *)(wb + W.off[0]), *minid = (int32_t *)(wb + W.off[1]), *lab = (int32_t *)(wb + W.off[2]);
    int32_t *ng = (int32_t *)(wb + W.off[3]);
    uint64_t *keys = (uint64_t *)(wb + W.off[4]), *keys2 = (uint64_t *)(wb + W.off[5]);
    int32_t *vals = (int32_t *)(wb + W.off[6]), *vals2 = (int32_t *)(wb + W.off[7]);
    void *tmp = wb + W.off[8];
    hipError_t e = hipSuccess;
#define CK(x) do { e = (x); if (e != hipSuccess) return e; } while (0)
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(256), 0, st, cnt, 0, n);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(256), 0, st, minid, 0x7fffffff, n);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(256), 0, st, lab, 0, n);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, ng, 0, (int64_t)1);
    {
        ProfScope ps("k_fof_finish", st);
        hipLaunchKernelGGL(k_fof_compress, dim3(nblk(n)), dim3(256), 0, st, P, n, T.perm, cnt, minid);
        hipLaunchKernelGGL(k_fof_store_root, dim3(nblk(n)), dim3(256), 0, st, P, n);
    }
    {
        hipLaunchKernelGGL(k_fof_collect, dim3(nblk(n)), dim3(256), 0, st, P, cnt, minid, n, minnum > 1 ? minnum : 1,
                           keys, vals, ng);
        int32_t hng = 0;
        CK(hipMemcpyAsync(&hng, ng, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        *ngroups = hng;
        if (hng > 0) {
            // the sort's storage was sized for cap pairs; a sort of fewer
            // that asks for more gets its own block
            size_t tb = 0;
            CK(nbx_sort_pairs_u64(nullptr, tb, keys, keys2, vals, vals2, hng, st));
            void *own = nullptr;
            if (tb > W.sort_bytes) CK(hipMallocAsync(&own, tb, st));
            e = nbx_sort_pairs_u64(own ? own : tmp, tb, keys, keys2, vals, vals2, hng, st);
            if (own) hipFreeAsync(own, st);
            CK(e);
            hipLaunchKernelGGL(k_fof_rank, dim3(nblk(hng)), dim3(256), 0, st, vals2, (int64_t)hng, lab);
        }
        hipLaunchKernelGGL(k_fof_out, dim3(nblk(n)), dim3(256), 0, st, P, lab, T.perm, n, labels_by_id);
        CK(hipGetLastError());
    }
#undef CK
    return e;
}

// ------------------------------------------------------------------ small kernels
__global__ __launch_bounds__(256) void k_scatter_by_id(const double *src, const int32_t *perm, double *dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}
__global__ __launch_bounds__(256) void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ __launch_bounds__(256) void k_fill_f64(double *p, double v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ __launch_bounds__(256) void k_iota(int32_t *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// tree slots of the particles with the given IDs (-1 for an ID out of range)
__global__ __launch_bounds__(256) void k_ids_to_slots(const int32_t *ids, const int32_t *inv, int64_t m,
                                                      int32_t *slots) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) slots[i] = inv[ids[i]];
}

// ------------------------------------------------------------------ ball search
// SearchBallPosTagged(x) (KDFindNearest.cxx:618-688, KDLeafNode.cxx:336-356)
// and SearchBallPos(x) (KDFindNearest.cxx:575-582, KDLeafNode.cxx:279-300):
// every particle with d2 < r2 around one position, its ID and (dout, when
// given) its d2 from that position; one wave walks the tree and each lane
// tests one candidate of a leaf.
__global__ __launch_bounds__(64) void k_ball(DevTree T, double qx, double qy, double qz, double r2, int32_t *out,
                                             double *dout, int64_t cap, unsigned long long *count) {
    const double q[3] = {qx, qy, qz};
    int node = 1;
    bool down = true;
    for (;;) {
        if (down) {
            cdptr bx = (cdptr)node_box(T, node, 3);
            if (box_lb<3>(q, bx) < r2) {
                if (node >= T.nleafp) {
                    int32_t cs = (node - T.nleafp) * NBX_LEAF + (int)threadIdx.x;
                    if (threadIdx.x < NBX_LEAF && cs < T.n) {
                        double d2 = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            double t = q[d] - T.x[d][cs];
                            d2 += t * t;
                        }
                        if (d2 < r2) {
                            unsigned long long at = atomicAdd(count, 1ULL);
                            if ((int64_t)at < cap) {
                                out[at] = T.perm[cs];
                                if (dout) dout[at] = d2;
                            }
                        }
                    }
                    down = false;
                } else {
                    node = 2 * node;
                    continue;
                }
            } else {
                down = false;
            }
        }
        if (node == 1) break;
        if (!(node & 1)) {
            node ^= 1;
            down = true;
        } else {
            node >>= 1;
        }
    }
}
