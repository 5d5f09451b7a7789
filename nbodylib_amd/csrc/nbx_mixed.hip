// nbx_mixed.hip -- precision form of kNN + SPH density (FP32 search, FP64
// re-check band, exact neighbour sets).
//
// Same call as CalcDensity (KDCalcSmoothQuantities.cxx:203-305) and the
// target-based FindNearest / FindNearestPhase (KDFindNearest.cxx:247-318,
// :347-361) on a non-periodic search, in 3-D or on the 6-D phase-space
// tree, with the same wave/leaf structure as the FP64 kernel (nbx_knn.hip).
//
// Search: particle coordinates are stored as FP32 offsets from their leaf's
// FP64 centre; a query is re-expressed per leaf as fl32(q - centre) so every
// difference is formed between two small FP32 numbers.  Each lane keeps the
// k+1 smallest FP32 d2: their bits in LDS (4 bytes per entry, so LDS does
// not limit how many waves a CU holds) and their particle slots in a
// per-workgroup scratch array in device memory.  The kernel is persistent:
// each workgroup takes waves of queries from its XCD's share of the tree
// order through an atomic counter, so the scratch is sized by the resident
// workgroups, not by n.
//
// Error bound: with u = 2^-24 and E the half-diagonal of the leaf a
// candidate came from, |dist32 - dist64| <= 12 u dist32 + 5 u E.  Per
// component the two offset roundings and the difference add at most
// u(|q - o| + |x - o| + |t|) <= u(dist + 2E + dist); the FMA chain over ND
// <= 6 squares adds ND u relative to the sum, sqrt halves it and adds u:
// at most (2 + ND/2 + 1) u dist + 2 u E = 6 u dist + 2 u E in 6-D, so the
// constants keep a factor 2 of margin.  A kept key adds its tag error (2^TB
// u dist, see the kept keys below) to the 12 u dist term.  Each lane tracks E_max over the leaves it evaluated
// as reachable or accepted a candidate from (candidates of other leaves are
// outside by the FP64 node bound).
//
// Re-check band: after the walk the k+1 keys are sorted; the candidates
// whose distance interval overlaps the k-th one's (plus any that could be a
// d2 == 0 duplicate) are recomputed in FP64 and the k-th neighbour is chosen
// among them by (d2, ID) as in the FP64 form.  Certificate: the largest
// selected FP64 d2 must be strictly below the squared lower bound of every
// particle outside the heap (bounded through the (k+1)-th key).  Lanes that
// fail it (the bound is loose when a lane has evaluated a very large leaf)
// are verified exactly: the k smallest (d2, ID) of the k+1 candidates are
// taken in FP64 and an FP64 walk of that ball counts the particles below the
// k-th; exactly k-1 proves the set.  Only lanes that fail this too are
// appended to a fallback list and recomputed by the FP64 kernel, so
// neighbour sets are always exact.
// h = |x - x_k| / 2 comes from the FP64 k-th distance, and the kernel
// weights of the k selected neighbours use FP64 distances (an FP32 offset
// carries a storage error proportional to its leaf's size, which in
// clustered data can reach 1e-5 of the neighbour distance); the search --
// every candidate test and heap operation -- stays in FP32.
#include "nbx_kernels.h"

typedef float fx8 __attribute__((ext_vector_type(8)));
typedef const __attribute__((address_space(4))) fx8 *cf8ptr;

#define MX_U 5.9604644775390625e-8  // 2^-24
#define MX_DB 4               // neighbours gathered together in the density pass
#define MX_EMPTY 0x7F7FFFFFu  // empty kept entry: the bits of FLT_MAX

#define HK(p) hk[(p) * NBX_WAVE + lane]   // kept keys (LDS)
typedef __attribute__((address_space(1))) uint16_t gu16;
typedef __attribute__((address_space(1))) uint32_t gu32;
#define SLH(p) ((gu16 *)sl + lane)[(p) * NBX_WAVE]  // their handles (scratch, by position)
#define MX_V 2048  // leaves a wave can record: handle = leaf's visit number << 5 | index in leaf

// Kept keys.  The kh = k+1 kept candidates of a lane sit at positions
// p < kh of KMAX = G * S (G groups of S positions; positions p >= kh hold
// 0, which never outranks a key).  A kept key is the FP32 d2's bits with the
// low TB bits replaced by a tag: its position p when KMAX <= 33 (keys are
// then unique), else its position within its group (the group holding the
// largest key is found among the group maxima).  Keys compare as unsigned
// integers in d2 order up to the TB replaced bits, and the largest key names
// its own position.  A key overstates or understates d2 by less than 2^TB
// ulps -- a relative 2^TB * 2^-23 in d2, 2^TB * u in the distance -- which
// the error bound below carries.  Each group's largest key stays in a
// register; replacing the largest kept key is one LDS store, one scratch
// store and one S-wide LDS reload of its group.
template <int KMAX>
struct GShape {
    static constexpr int G = KMAX <= 33 ? 3 : 5;
    static constexpr int S = KMAX / G;
    static constexpr bool GT = KMAX <= 33;    // tag = position p (else position in group)
    static constexpr int TB = GT ? 6 : 4;      // tag bits
    static constexpr uint32_t TM = (1u << TB) - 1u;
    static_assert(G * S == KMAX && (GT ? KMAX : S) <= (1 << TB), "group shape must tile KMAX");
};

__device__ __forceinline__ uint32_t umax3(uint32_t a, uint32_t b, uint32_t c) { return max(max(a, b), c); }

// Batcher's odd-even merge sorting network over 64 inputs, keeping only the
// comparators whose inputs are both below lim (the inputs from lim on are
// +inf, which those comparators would leave in place).  Built at compile
// time; a fully unrolled pass over it sorts keys held in registers.
struct SortNet {
    int n;
    uint8_t a[256], b[256];
};
constexpr void sn_add(SortNet &s, int i, int j, int lim) {
    if (j < lim) {
        s.a[s.n] = (uint8_t)i;
        s.b[s.n] = (uint8_t)j;
        s.n += 1;
    }
}
constexpr void sn_merge(SortNet &s, int lo, int hi, int r, int lim) {
    const int step = 2 * r;
    if (step < hi - lo) {
        sn_merge(s, lo, hi, step, lim);
        sn_merge(s, lo + r, hi, step, lim);
        for (int i = lo + r; i < hi - r; i += step) sn_add(s, i, i + r, lim);
    } else {
        sn_add(s, lo, lo + r, lim);
    }
}
constexpr void sn_sort(SortNet &s, int lo, int hi, int lim) {
    if (hi - lo >= 1) {
        const int mid = lo + (hi - lo) / 2;
        sn_sort(s, lo, mid, lim);
        sn_sort(s, mid + 1, hi, lim);
        sn_merge(s, lo, hi, 1, lim);
    }
}
constexpr SortNet sort_net(int lim) {
    SortNet s{};
    sn_sort(s, 0, 63, lim);
    return s;
}
constexpr SortNet MX_NET33 = sort_net(33);  // 246 comparators

// Largest of S keys (a max tree).
template <int S>
__device__ __forceinline__ uint32_t gmaxu(const uint32_t (&x)[S]) {
    uint32_t v[S];
#pragma unroll
    for (int u = 0; u < S; ++u) v[u] = x[u];
#pragma unroll
    for (int w = 1; w < S; w *= 3)
#pragma unroll
        for (int u = 0; u + w < S; u += 3 * w) v[u] = u + 2 * w < S ? umax3(v[u], v[u + w], v[u + 2 * w]) : max(v[u], v[u + w]);
    return v[0];
}

template <int KMAX>
__device__ __forceinline__ uint32_t gtop_init(uint32_t *hk, int lane, int kh, uint32_t (&gm)[GShape<KMAX>::G]) {
    constexpr int G = GShape<KMAX>::G, S = GShape<KMAX>::S;
    constexpr uint32_t TM = GShape<KMAX>::TM;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        uint32_t m = 0u;
        for (int u = 0; u < S; ++u) {
            const int p = g * S + u;
            const uint32_t v = p < kh ? ((MX_EMPTY & ~TM) | (uint32_t)(GShape<KMAX>::GT ? p : u)) : 0u;
            HK(p) = v;
            m = max(m, v);
        }
        gm[g] = m;
    }
    return gmaxu<G>(gm);
}

// Replace the largest kept key (root) by key in the lanes with ins set (key
// carries the root's position); the others store the root back, so nothing
// but the handle store branches.  slot is the candidate's 16-bit handle.
// Returns the new largest key.
template <int KMAX>
__device__ __forceinline__ uint32_t gtop_replace(uint32_t *hk, uint32_t *sl, int lane,
                                                 uint32_t (&gm)[GShape<KMAX>::G], bool ins, uint32_t key,
                                                 int32_t slot, uint32_t root) {
    constexpr int G = GShape<KMAX>::G, S = GShape<KMAX>::S;
    constexpr uint32_t TM = GShape<KMAX>::TM;
    int g0 = 0, p;
    if constexpr (GShape<KMAX>::GT) {
        p = (int)(root & TM);
#pragma unroll
        for (int g = 1; g < G; ++g) g0 += p >= g * S ? 1 : 0;
    } else {
#pragma unroll
        for (int g = G - 1; g > 0; --g) g0 = gm[g] == root ? g : g0;
        p = g0 * S + (int)(root & TM);
    }
    const int b = g0 * S;
    HK(p) = ins ? key : root;
    if (ins) SLH(p) = (uint16_t)slot;
    uint32_t x[S];
#pragma unroll
    for (int u = 0; u < S; ++u) x[u] = HK(b + u);
    const uint32_t bc = gmaxu<S>(x);
#pragma unroll
    for (int g = 0; g < G; ++g) gm[g] = g0 == g ? bc : gm[g];
    return gmaxu<G>(gm);
}

template <int ND>
struct MixCtx {
    double q[ND];
    bool act;
    int32_t self;
    uint32_t root;  // largest kept key
    double emax;    // largest leaf half-diagonal that matters for this lane
    int nvis;       // leaves recorded (wave-uniform), the next handle's visit number
    int fillkh;     // kh while the heap is still empty and kh >= 32 (wave-uniform), else 0
};

// The leaf data a visit reads, each base a value of its own (see mix_walk):
// origins and half-diagonals (scalar loads), and the FP32 offsets, 3-D
// interleaved (scalar and vector loads) or 6-D per dimension (vector loads
// only, so those bases are held in vector registers).
template <int ND>
struct MixSrc {
    const double *org, *ext;
    const float4 *xq;
    const float *xf[ND == 6 ? 6 : 1];
};

template <int ND, int KMAX>
__device__ __forceinline__ void mix_leaf(const MixSrc<ND> &S, int leaf, MixCtx<ND> &c, bool want,
                                         uint32_t *hk, uint32_t *sl, float *cx, int lane,
                                         uint32_t (&gm)[GShape<KMAX>::G]) {
    const int base = leaf * NBX_LEAF;
    // A kept candidate is stored as a 16-bit handle, its leaf's visit number
    // << 5 | its index in the leaf; the wave's visit table (vt, right after
    // the handles in scratch) maps visit numbers to leaves.  A 3-D wave
    // numbers every leaf it visits (about 36); a 6-D wave visits hundreds,
    // so it numbers only the leaves where some lane accepts a candidate.
    uint32_t *vt = sl + KMAX * (NBX_WAVE / 2);
    int rbase = c.nvis << 5;
    if constexpr (ND == 3) {
        if (lane == 0) vt[c.nvis] = (uint32_t)leaf;
        c.nvis += 1;
    }
    // candidate (lane mod 32) of this leaf: the tested point in phase 1 and
    // this lane's part of the LDS copy used in phase 2
    float st[ND];
    if constexpr (ND == 3) {
        const float4 v = S.xq[base + (lane & (NBX_LEAF - 1))];
        st[0] = v.x;
        st[1] = v.y;
        st[2] = v.z;
    } else {
#pragma unroll
        for (int d = 0; d < ND; ++d) st[d] = S.xf[d][base + (lane & (NBX_LEAF - 1))];
    }
    // query relative to this leaf's centre, rounded once to FP32
    const double *o = S.org + (size_t)leaf * ND;
    float qv[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) qv[d] = (float)(c.q[d] - ((cdptr)o)[d]);
    const float th = __uint_as_float(c.root);
    uint32_t mask = 0u;
    if constexpr (ND == 3) {
        // Phase 1: every lane tests the 32 candidates, read in 64-byte scalar
        // loads, in the expanded form
        //     w + 2 qv.p >= |p|^2,  w = th + m - |qv|^2
        // (|p|^2 stored per slot): three FMAs and a compare per candidate,
        // the accept bit shifted in with a carry.  In exact arithmetic this
        // is |qv - p|^2 <= th + m.  Phase 2 accepts only candidates whose
        // difference-form d2 (3 roundings, relative error < 6u, u = 2^-24)
        // is <= th, i.e. with |qv - p|^2 <= th (1 + 6u); the rounding of the
        // expanded form is below 11u th + 5u m + 7u (|qv| + |p|)^2, so the
        // margin m = 2^-20 (th + (|qv| + e)^2) (e >= |p|: the leaf's
        // half-diagonal) lets every such candidate through.  The few extra
        // ones fail phase 2's test, so the kept keys are those of the
        // difference form alone.
        typedef float fxc __attribute__((ext_vector_type(16)));
        typedef const __attribute__((address_space(4))) fxc *cfcptr;
        const float q2 = __builtin_fmaf(qv[2], qv[2], __builtin_fmaf(qv[1], qv[1], qv[0] * qv[0]));
        const float r = sqrtf(q2) + (float)(((cdptr)S.ext)[leaf] * (1.0 + 0x1p-20));
        const float m = 0x1p-20f * __builtin_fmaf(r, r, th);
        // FLT_MAX (no bound yet): every real candidate passes
        const float w = fminf((th + m) - q2, 3.402823466e38f);
        const float a0 = 2.0f * qv[0], a1 = 2.0f * qv[1], a2 = 2.0f * qv[2];
        // two halves of 16 candidates (x, y, z, |p|^2), 64 scalar registers
        // each (the whole leaf at once would not fit the scalar register file)
        const cfcptr lp = (cfcptr)(S.xq + base);
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            fxc c4[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) c4[i] = lp[4 * h + i];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int i = u >> 2, o = (u & 3) * 4;
                float v = __builtin_fmaf(a0, c4[i][o], w);
                v = __builtin_fmaf(a1, c4[i][o + 1], v);
                v = __builtin_fmaf(a2, c4[i][o + 2], v);
                mask = mask + mask + (v >= c4[i][o + 3] ? 1u : 0u);
            }
        }
        // candidate j was shifted in at bit 31 - j
        mask = __builtin_bitreverse32(mask);
    } else {
        // Phase 1, transposed: the lanes that want this leaf put their query
        // offsets and thresholds in LDS (qb, tb, which alias cx), and each
        // step tests two of them against the 32 candidates, one candidate per
        // lane; the ballot's halves are the two queries' accept masks,
        // selected into the owning lanes' mask register.  Lanes that do not
        // want the leaf have no candidate in it: their FP64 bound already
        // exceeds their largest kept key.
        const uint64_t wm = ballot(want);
        const int nwq = __popcll(wm);
        float *qb = cx;
        float *tb = cx + ND * NBX_WAVE;
        if (want) {
            const int wr = __builtin_amdgcn_mbcnt_hi((uint32_t)(wm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)wm, 0u));
#pragma unroll
            for (int d = 0; d < ND; ++d) qb[d * NBX_WAVE + wr] = qv[d];
            tb[wr] = th;
        }
        __syncthreads();
        uint64_t wrest = wm;  // wanting lanes not yet given a mask, in rank order
        for (int it = 0; it < nwq; it += 2) {
            const int qs = it + (lane >> 5);
            const int qc = qs < nwq ? qs : it;
            float t = qb[qc] - st[0];
            float dd = t * t;
#pragma unroll
            for (int d = 1; d < ND; ++d) {
                t = qb[d * NBX_WAVE + qc] - st[d];
                dd = __builtin_fmaf(t, t, dd);
            }
            const uint64_t pm = ballot(qs < nwq && dd <= tb[qc]);
            const int l0 = (int)__builtin_ctzll(wrest);
            wrest &= wrest - 1;
            mask = lane == l0 ? (uint32_t)pm : mask;
            if (it + 1 < nwq) {
                const int l1 = (int)__builtin_ctzll(wrest);
                wrest &= wrest - 1;
                mask = lane == l1 ? (uint32_t)(pm >> 32) : mask;
            }
        }
        __syncthreads();
    }
    if (c.self >= base && c.self < base + NBX_LEAF) mask &= ~(1u << (c.self - base));
    if (!c.act) mask = 0;
    if (c.act && (want || mask)) c.emax = fmax(c.emax, ((cdptr)S.ext)[leaf]);
    if constexpr (ND == 6) {
        // no lane accepts anything: nothing to insert, no visit number
        if (!ballot(mask != 0)) return;
        if (lane == 0) vt[c.nvis] = (uint32_t)leaf;
        c.nvis += 1;
    }
    if (lane < NBX_LEAF) {
#pragma unroll
        for (int d = 0; d < ND; ++d) cx[d * NBX_LEAF + lane] = st[d];
    }
    __syncthreads();
    if constexpr (GShape<KMAX>::GT) {
        if (c.fillkh) {
            // First leaf of the walk, every heap still empty and at least 32
            // positions: the replacement would put the m-th accepted
            // candidate at position kh - 1 - m (the largest empty key's), so
            // the keys go straight there in one pass over the leaf, and the
            // group maxima are taken once at the end.
            const int kh = c.fillkh;
            c.fillkh = 0;
            constexpr uint32_t TM = GShape<KMAX>::TM;
            constexpr int G = GShape<KMAX>::G, S = GShape<KMAX>::S;
            int cnt = 0;
            for (int j = 0; j < NBX_LEAF; ++j) {
                float t = qv[0] - cx[j];
                float dd = t * t;
#pragma unroll
                for (int d = 1; d < ND; ++d) {
                    t = qv[d] - cx[d * NBX_LEAF + j];
                    dd = __builtin_fmaf(t, t, dd);
                }
                const int p = kh - 1 - cnt;
                const uint32_t key = (__float_as_uint(dd) & ~TM) | (uint32_t)p;
                const bool ins = ((mask >> j) & 1u) && key < ((MX_EMPTY & ~TM) | (uint32_t)p);
                if (ins) {
                    HK(p) = key;
                    SLH(p) = (uint16_t)(rbase + j);
                }
                cnt += ins ? 1 : 0;
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                uint32_t x[S];
#pragma unroll
                for (int u = 0; u < S; ++u) x[u] = HK(g * S + u);
                gm[g] = gmaxu<S>(x);
            }
            c.root = gmaxu<G>(gm);
            __syncthreads();
            return;
        }
    }
    // as many steps as the lane with the most candidates needs (a counted
    // loop keeps the kept-key state in place across steps)
    int nit = __popc(mask);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nit = max(nit, __shfl_xor(nit, off));
    nit = __builtin_amdgcn_readfirstlane(nit);
    for (int it = 0; it < nit; ++it) {
        const bool has = mask != 0;
        const int j = __builtin_ctz(mask | 0x80000000u);
        mask &= mask - 1;
        float t = qv[0] - cx[j];
        float dd = t * t;
#pragma unroll
        for (int d = 1; d < ND; ++d) {
            t = qv[d] - cx[d * NBX_LEAF + j];
            dd = __builtin_fmaf(t, t, dd);
        }
        constexpr uint32_t TM = GShape<KMAX>::TM;
        const uint32_t key = (__float_as_uint(dd) & ~TM) | (c.root & TM);
        c.root = gtop_replace<KMAX>(hk, sl, lane, gm, has && key < c.root, key, rbase + j, c.root);
    }
    __syncthreads();
}

// Stackless depth-first walk: the ancestor of the wave's own leaves first,
// else the child nearer to the wave's first query; firstright[dep] records
// which child was entered first at each depth.  Two loop forms visit the
// nodes in the same order; the compiler schedules them differently, and the
// measured faster one is used for each dimensionality (3-D: climb to the
// next unvisited sibling in an inner loop; 6-D: one node step per trip).
template <int ND, int KMAX>
__device__ __forceinline__ void mix_walk(const DevTree &T, const double *ext, MixCtx<ND> &c, const float *qc, int anc,
                                         uint32_t *hk, uint32_t *sl, float *cx, int lane,
                                         uint32_t (&gm)[GShape<KMAX>::G]) {
    // The depth and the box and plane bases as scalar values of their own.
    // The tree view arrives as wide blocks of scalar registers, and the
    // register allocator spills and restores such a block whole: a field
    // read from it in the loop brought back 16 registers (16 v_readlane)
    // per node in 6-D.
    int D = T.depth, nleafp = T.nleafp;
    const float *boxp = T.boxf;
    const float *planep = T.plane;
    asm volatile("" : "+s"(D), "+s"(nleafp), "+s"(boxp), "+s"(planep));
    // node tests in FP32 on the outward-rounded boxes: never above the FP64
    // bound, so every node the FP64 test opens is opened (a lane past the end
    // sits at 1e30, beyond every box)
    float q32[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) q32[d] = c.act ? (float)c.q[d] : NBX_PAD_COORD_F;
    MixSrc<ND> S;
    S.org = T.org;
    S.ext = ext;
    asm volatile("" : "+s"(S.org), "+s"(S.ext));
    if constexpr (ND == 3) {
        S.xq = T.xq;
        asm volatile("" : "+s"(S.xq));
    } else {
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            S.xf[d] = T.xf[d];
            asm volatile("" : "+v"(S.xf[d]));
        }
    }
    int node = 1, dep = 0;
    uint64_t firstright = 0;
    if constexpr (ND == 3) {
        for (;;) {
            cfptr bx = (cfptr)(boxp + (size_t)node * 2 * ND);
            const bool want = box_lb32<ND>(q32, bx) <= __uint_as_float(c.root);
            if (ballot(want)) {
                if (dep == D) {
                    if (c.nvis >= MX_V) break;  // visit table full: lanes fall back
                    mix_leaf<ND, KMAX>(S, node - nleafp, c, want, hk, sl, cx, lane, gm);
                } else {
                    int first;
                    if ((anc >> (D - dep)) == node) first = anc >> (D - dep - 1);
                    else first = 2 * node + (first_is_right<ND>(planep, node, qc) ? 1 : 0);
                    dep += 1;
                    firstright = (firstright & ~(1ULL << dep)) | ((uint64_t)(first & 1) << dep);
                    node = first;
                    continue;
                }
            }
            while (node != 1 && (uint64_t)(node & 1) != ((firstright >> dep) & 1ULL)) {
                node >>= 1;
                dep -= 1;
            }
            if (node == 1) break;
            node ^= 1;
        }
    } else {
        bool down = true;
        for (;;) {
            if (down) {
                cfptr bx = (cfptr)(boxp + (size_t)node * 2 * ND);
                const bool want = box_lb32<ND>(q32, bx) <= __uint_as_float(c.root);
                if (ballot(want)) {
                    if (dep == D) {
                        if (c.nvis >= MX_V) break;  // visit table full: lanes fall back
                        mix_leaf<ND, KMAX>(S, node - nleafp, c, want, hk, sl, cx, lane, gm);
                        down = false;
                    } else {
                        const int c0 = 2 * node;
                        int first = c0;
                        if (anc >= 0 && (anc >> (D - dep)) == node) {
                            first = anc >> (D - dep - 1);
                        } else if (first_is_right<ND>(planep, node, qc)) {
                            first = c0 + 1;
                        }
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
                node ^= 1;
                down = true;
            } else {
                node >>= 1;
                dep -= 1;
            }
        }
    }
}

template <int ND>
__device__ __forceinline__ double d64(const DevTree &T, const double *q, int32_t s) {
    double d2 = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double t = q[d] - T.x[d][s];
        d2 += t * t;
    }
    return d2;
}

// Exact FP64 check of a candidate set: counts the particles whose key
// (d2, ID) is below (Dk, IDk) -- self and d2 == 0 excluded, as in the target
// search -- within the ball of radius^2 Dk, for the lanes with act set.
// Wave-uniform left-first walk; leaves through 64-byte scalar loads.
template <int ND>
__device__ __forceinline__ int verify_walk(const DevTree &T, const double *q, bool act, double Dk, int32_t IDk,
                                           int32_t self) {
    constexpr int CH = ND == 6 ? 4 : 8;
    typedef double dxc __attribute__((ext_vector_type(CH)));
    typedef const __attribute__((address_space(4))) dxc *cdcptr;
    int cnt = 0;
    int node = 1;
    bool down = true;
    for (;;) {
        if (down) {
            cdptr bx = (cdptr)node_box(T, node, ND);
            bool want = act && box_lb<ND>(q, bx) <= Dk;
            if (ballot(want)) {
                if (node >= T.nleafp) {
                    const int base = (node - T.nleafp) * NBX_LEAF;
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
                                const double t = q[d] - cv[d][u];
                                d2 += t * t;
                            }
                            const int32_t cs = base + jj + u;
                            bool below = d2 < Dk;
                            if (d2 == Dk && act && cs != self) below = T.perm[cs] < IDk;
                            cnt += (act && want && d2 > 0.0 && cs != self && below) ? 1 : 0;
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
    return cnt;
}

// After the sort: kept keys in ascending order (SKC) and their slots (SLOT).
// When KMAX <= 33 the keys are sorted in place in LDS and a key's tag is
// its slot's scratch position; else keys and slots are sorted into two more
// scratch regions.
#define SKC(p) kb[(p) * NBX_WAVE + lane]
#define SLB(p) slb[(p) * NBX_WAVE + lane]

// One wave of 64 consecutive queries (tree order), wave index g.
// k = kh - 1 neighbours; A.dens: NBX_DO_NONE / NBX_DO_SYM / NBX_DO_GATHER.
template <int ND, int KMAX>
__device__ __forceinline__ void mix_task(const DevTree &T, const KnnArgs &A, const double *ext, int32_t *fb_list,
                                         int32_t *fb_count, uint32_t *hk, float *cx, double *racc, uint32_t *sl,
                                         uint32_t *slb, uint32_t *skc, int g, int lane) {
    constexpr bool GT = GShape<KMAX>::GT;
    constexpr uint32_t TM = GShape<KMAX>::TM;
    uint32_t *const kb = GT ? hk : skc;
    // slot of sorted entry p; move sorted entry src to dst
    uint32_t *const vt = sl + KMAX * (NBX_WAVE / 2);  // visit table
    auto slot_at = [&](int p) -> int32_t {
        if constexpr (GT) {
            const uint32_t h = SLH(SKC(p) & TM);
            return (int32_t)(vt[h >> 5] * NBX_LEAF + (h & 31u));
        } else {
            return (int32_t)SLB(p);
        }
    };
    auto move_entry = [&](int dst, int src) {
        if constexpr (GT) SKC(dst) = SKC(src);
        else SLB(dst) = SLB(src);
    };
    // k + 1, opaque to the compiler so nothing derived from it is hoisted
    // out of the workgroup's loop into registers
    int kh = A.kh;
    asm volatile("" : "+s"(kh));
    const int k = kh - 1;
    const int qi = g * NBX_WAVE + lane;
    const bool valid = qi < T.n;
    const int kres = T.kernres;
    MixCtx<ND> c;
    // A lane past the end queries from the padding sentinel: node boxes
    // bound real particles only, so every box bound from there is +inf and
    // the walk needs no activity test (the leaf code still masks the lane).
#pragma unroll
    for (int d = 0; d < ND; ++d) c.q[d] = valid ? T.x[d][qi] : NBX_PAD_COORD;
    c.act = valid;
    c.self = qi;
    c.root = 0u;
    c.emax = 0.0;
    uint32_t gm[GShape<KMAX>::G];
    c.root = gtop_init<KMAX>(hk, lane, kh, gm);
    // every slot read later is a valid index: handle 0 names the leaf in
    // visit table entry 0, first the wave's own leaf, then the first leaf
    // with an accepted candidate
    for (int p = 0; p < kh; ++p) SLH(p) = 0;
    if (lane == 0) vt[0] = (uint32_t)(2 * g);
    c.nvis = 0;
    c.fillkh = (GT && kh >= NBX_LEAF) ? kh : 0;
    float qc[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) qc[d] = (float)readfirst_d(c.q[d]);
    mix_walk<ND, KMAX>(T, ext, c, qc, T.nleafp + 2 * g, hk, sl, cx, lane, gm);

    // ---------------------------------------------------------------- band
    const bool ovf = c.nvis >= MX_V;  // walk stopped early: no certificate
    bool ok = valid && !ovf;
    if (valid) {
        // rank sort of the kh kept keys: keys to SKC, slots to SLB
        if constexpr (GT) {
            // unique keys (unused positions +inf) sorted in registers by the
            // network, then stored back in order
            static_assert(KMAX == 33, "the sorting network is built for 33 keys");
            uint32_t kq[KMAX];
#pragma unroll
            for (int p = 0; p < KMAX; ++p) kq[p] = p < kh ? HK(p) : 0xffffffffu;
#pragma unroll
            for (int e = 0; e < MX_NET33.n; ++e) {
                const uint32_t x = kq[MX_NET33.a[e]], y = kq[MX_NET33.b[e]];
                kq[MX_NET33.a[e]] = min(x, y);
                kq[MX_NET33.b[e]] = max(x, y);
            }
#pragma unroll
            for (int p = 0; p < KMAX; ++p)
                if (p < kh) HK(p) = kq[p];
        } else {
            // ties (possible only between groups) by position
            uint32_t kq[KMAX];
#pragma unroll
            for (int p = 0; p < KMAX; ++p) kq[p] = p < kh ? HK(p) : 0xffffffffu;
            for (int i = 0; i < kh; ++i) {
                const uint32_t ki = HK(i);
                int r = 0;
#pragma unroll
                for (int j = 0; j < KMAX; ++j) r += (kq[j] < ki || (kq[j] == ki && j < i)) ? 1 : 0;
                const uint32_t h = SLH(i);
                SKC(r) = ki;
                SLB(r) = vt[h >> 5] * NBX_LEAF + (h & 31u);
            }
        }
        if ((SKC(kh - 1) & ~GShape<KMAX>::TM) == (MX_EMPTY & ~GShape<KMAX>::TM)) ok = false;  // < k+1 candidates
    }
    // FP32 search error plus the key's position bits, or the caller's
    // re-check band where that is wider (a wider band only sends more
    // entries to the FP64 re-check and the exact verification)
    const double eu = fmax((12.0 + (double)(1 << GShape<KMAX>::TB)) * MX_U, A.band_rel), ea = 5.0 * MX_U * c.emax;
    int i0 = k - 1;
    int32_t bs[4];
    uint32_t bk[4];
    double bd[4];
    double Dk = 0.0;
    if (ok) {
        // band: entries whose interval overlaps the k-th entry's
        double dk = sqrt((double)__uint_as_float(SKC(k - 1)));
        double lo_k = dk - (eu * dk + ea);
        while (i0 > 0) {
            double di = sqrt((double)__uint_as_float(SKC(i0 - 1)));
            if (di + (eu * di + ea) < lo_k) break;
            --i0;
        }
        // a certain entry that could be an exact duplicate (d2 == 0) must be checked too
        double d0 = sqrt((double)__uint_as_float(SKC(0)));
        if (d0 - (eu * d0 + ea) <= 0.0 && i0 > 0) ok = false;
        if (kh - i0 > 4) ok = false;
    }
    if (ok) {
        // recompute the band in FP64; drop one entry: a d2 == 0 duplicate if
        // present, else the largest (d2, ID)
        const int nb = kh - i0;
        int drop = -1, nzero = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            bs[b] = -1;
            bd[b] = 0.0;
            if (b < nb) {
                bk[b] = SKC(i0 + b);
                bs[b] = slot_at(i0 + b);
                bd[b] = d64<ND>(T, c.q, bs[b]);
                if (bd[b] == 0.0) { nzero++; drop = b; }
            }
        }
        if (nzero > 1) ok = false;
        if (drop < 0) {
            // the largest (d2, ID) of the band (static indices only)
            drop = 0;
            double dd = bd[0];
            int32_t pd = T.perm[bs[0]];
#pragma unroll
            for (int b = 1; b < 4; ++b) {
                const int32_t pb = b < nb ? T.perm[bs[b]] : 0;
                if (b < nb && (bd[b] > dd || (bd[b] == dd && pb > pd))) {
                    drop = b;
                    dd = bd[b];
                    pd = pb;
                }
            }
        }
        Dk = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (b < nb && b != drop && bd[b] > Dk) Dk = bd[b];
        // certificate against everything outside the heap
        double dl = sqrt((double)__uint_as_float(SKC(kh - 1)));
        double lout = dl - (eu * dl + ea);
        if (!(lout > 0.0 && Dk < lout * lout)) ok = false;
        if (ok) {
            // the kept band entries fill positions i0 .. k-1 (FP32 order)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (b < nb && b != drop) {
                    const int dst = i0 + b - (b > drop ? 1 : 0);
                    if constexpr (GT) SKC(dst) = bk[b];
                    else SLB(dst) = (uint32_t)bs[b];
                }
        }
    }
    // exact verification for lanes the error bound could not decide: keep
    // the k smallest (d2, ID) of the k+1 candidates in FP64, then count in
    // FP64 the particles below the k-th of them
    bool verify = valid && !ok && !ovf && (SKC(kh - 1) & ~GShape<KMAX>::TM) != (MX_EMPTY & ~GShape<KMAX>::TM);
    int32_t IDk = 0;
    if (verify) {
        int zero_at = -1, nzero = 0, maxp = -1;
        double dmax = -1.0;
        int32_t imax = -1;
        for (int p = 0; p < kh; ++p) {
            int32_t s = slot_at(p);
            double d = d64<ND>(T, c.q, s);
            if (d == 0.0) {
                nzero++;
                zero_at = p;
                continue;
            }
            int32_t id = T.perm[s];
            if (maxp < 0 || d > dmax || (d == dmax && id > imax)) { maxp = p; dmax = d; imax = id; }
        }
        const int drop = nzero == 1 ? zero_at : maxp;
        if (nzero > 1) verify = false;
        if (verify) {
            for (int p = drop; p < kh - 1; ++p) move_entry(p, p + 1);
            Dk = 0.0;
            IDk = -1;
            for (int p = 0; p < k; ++p) {
                int32_t s = slot_at(p);
                double d = d64<ND>(T, c.q, s);
                int32_t id = T.perm[s];
                if (d > Dk || (d == Dk && id > IDk)) { Dk = d; IDk = id; }
            }
        }
    }
    if (ballot(verify)) {
        int cnt = verify_walk<ND>(T, c.q, verify, Dk, IDk, qi);
        if (verify && cnt == k - 1) ok = true;
    }
    // fallback list for the FP64 kernel
    uint64_t fbm = ballot(valid && !ok);
    if (fbm) {
        int base = 0;
        if (lane == 0) base = atomicAdd(fb_count, __popcll(fbm));
        base = __shfl(base, 0);
        if (valid && !ok) fb_list[base + __popcll(fbm & ((1ULL << lane) - 1))] = qi;
    }

    // k <= 32: the slots of the k kept entries, decoded in one batch (every
    // handle load, then every visit-table load, instead of a dependent pair
    // per entry) and stored over their keys, which nothing reads after this;
    // the density and the output rows take them from there
    if constexpr (GT) {
        if (ok) {
            const gu32 *vtg = (const gu32 *)vt;
            uint32_t hv[KMAX];
#pragma unroll
            for (int p = 0; p < KMAX; ++p) hv[p] = p < k ? (uint32_t)SLH(SKC(p) & TM) : 0u;
#pragma unroll
            for (int p = 0; p < KMAX; ++p)
                if (p < k) SKC(p) = vtg[hv[p] >> 5] * NBX_LEAF + (hv[p] & 31u);
        }
    }
    auto slot_final = [&](int p) -> int32_t {
        if constexpr (GT) return (int32_t)SKC(p);
        else return (int32_t)SLB(p);
    };

    // ---------------------------------------------------------------- density
    // Symmetric scatter: contributions to the wave's own 64 queries (about
    // half of all neighbour entries) are summed in LDS and flushed with one
    // atomic per query; the rest go straight to HBM.
    const bool sym = A.dens == NBX_DO_SYM;
    double racc_own = 0.0;
    const bool dens_here = ok && A.dens != NBX_DO_NONE;
    if (sym) racc[lane] = 0.0;
    __syncthreads();
    if (dens_here) {
        // weights from FP64 distances of the k selected neighbours (the FP32
        // offsets carry a storage error proportional to the leaf size)
        const double hi = 0.5 * sqrt(Dk);
        const double norm = 1.0 / pow(hi, (double)ND);
        const double mi = T.mass[qi];
        // the table lookup of Wsm (KDCalcSmoothQuantities.cxx:12-15) with
        // its division by the table step turned into a product: index
        // floor(r/h * c), c = (kernres - 1) / 2 = 1 / step, and the linear
        // weight of the next entry the index's fraction.  r/h stays a
        // division, so the k-th neighbour's is exactly 2 and it takes the
        // table's last entry, as in the reference.
        const double ck = 0.5 * (kres - 1);
        double acc = 0.0;
        // neighbours in groups of MX_DB so their coordinate/mass gathers are in
        // flight together
        for (int p0 = 0; p0 < k; p0 += MX_DB) {
            int32_t sv[MX_DB];
            double d2v[MX_DB], mv[MX_DB];
#pragma unroll
            for (int u = 0; u < MX_DB; ++u) {
                sv[u] = p0 + u < k ? slot_final(p0 + u) : qi;
                d2v[u] = d64<ND>(T, c.q, sv[u]);
                mv[u] = T.mass[sv[u]];
            }
#pragma unroll
            for (int u = 0; u < MX_DB; ++u) {
                if (p0 + u >= k) break;
                const int32_t s = sv[u];
                const double fi = sqrt(d2v[u]) / hi * ck;
                double W;
                if (!(fi < (double)(kres - 1))) {
                    W = T.kern[kres - 1];
                } else {
                    const int i = (int)fi;
                    W = T.kern[i] + (T.kern[i + 1] - T.kern[i]) * (fi - (double)i);
                }
                if (sym) {
                    double Wij = 0.5 * W * norm;
                    acc = acc + Wij * mv[u];
                    const uint32_t r = (uint32_t)(s - g * NBX_WAVE);
                    if (r < NBX_WAVE) atomicAdd(&racc[r], Wij * mi);
                    else unsafeAtomicAdd(A.rho + s, Wij * mi);
                } else {
                    acc += W * norm * mv[u];
                }
            }
        }
        if (sym) racc_own = acc;
        else A.rho[qi] = acc;
        if (A.hout) A.hout[A.hout_by_id ? T.perm[qi] : qi] = hi;
    }
    __syncthreads();
    if (sym && valid) unsafeAtomicAdd(A.rho + qi, racc_own + racc[lane]);

    // ---------------------------------------------------------------- outputs
    if (A.nn == nullptr) return;
    const uint64_t okm = ballot(ok);
    __syncthreads();
    // entry e = lane + 64 i of the wave's rows is row r = e / k, column
    // t = e % k: stepped without a division in 3-D, divided in 6-D (each the
    // measured faster form of its instance)
    const int q64 = NBX_WAVE / k, m64 = NBX_WAVE - q64 * k;
    int r = lane / k, t = lane - (lane / k) * k;
    for (int i = 0; i < k; ++i, t += m64, r += q64) {
        if constexpr (ND == 3) {
            if (t >= k) {
                t -= k;
                r += 1;
            }
        } else {
            const int e = lane + NBX_WAVE * i;
            r = e / k;
            t = e - r * k;
        }
        int qr = g * NBX_WAVE + r;
        if (qr >= T.n) break;
        if (!((okm >> r) & 1ULL)) continue;
        int32_t s;
        if constexpr (GT) s = (int32_t)hk[t * NBX_WAVE + r];  // decoded above
        else s = (int32_t)slb[t * NBX_WAVE + r];
        int64_t row = A.layout == 1 ? qr - A.row0 : T.perm[qr];
        A.nn[row * k + t] = A.layout == 1 ? s : T.perm[s];
    }
}

// Persistent precision-form kernel, one wave per workgroup.  Workgroup b
// stays on XCD (b mod 8) (workgroups are dealt to XCDs round robin), takes
// waves of queries from that XCD's contiguous share of the tree order (the
// xcd_wave split) through the XCD's counter, and owns one scratch block of
// MX_BLKW(KMAX) words (16-bit handles by position and the visit table; for
// KMAX > 33 also slots and keys in ascending order) for its whole life, so
// the scratch is sized by the resident workgroups and only ever touched from
// one XCD.  Once its own
// share is drained a workgroup takes waves from the other XCDs' counters in
// turn (each ticket names one wave, so every wave runs once), which evens out
// shares of unequal cost; it exits when all eight counters have passed their
// shares' ends.
#define MX_CTR_STRIDE 16  // XCD counters 64 bytes apart
// scratch words per workgroup: 16-bit handles and the visit table, and for
// KMAX > 33 the slots and keys in ascending order
#define MX_BLKW(KM) ((KM) * (NBX_WAVE / 2) + MX_V + ((KM) <= 33 ? 0 : 2 * (KM) * NBX_WAVE))
template <int ND, int KMAX>
__global__ __launch_bounds__(64) void k_knn_mixed(DevTree T, KnnArgs A, const double *ext, int32_t *fb_list,
                                                   int32_t *fb_count, uint32_t *scratch, int32_t *ctr) {
    __shared__ uint32_t hk[KMAX * NBX_WAVE];
    // one buffer for three uses at different times: the leaf copy (phase 2),
    // the compacted queries and masks (phase 1) and the symmetric scatter into
    // the wave's own queries (density)
    constexpr int UB1 = ND == 3 ? 0 : ND * NBX_WAVE + NBX_WAVE;  // transposed phase 1 (6-D only)
    constexpr int UBN = UB1 > ND * NBX_LEAF + 2 * NBX_WAVE ? UB1 : ND * NBX_LEAF + 2 * NBX_WAVE;
    __shared__ __attribute__((aligned(8))) float ub[UBN];
    float *cx = ub;
    double *racc = (double *)(ub + ND * NBX_LEAF);
    const int lane = threadIdx.x;
    uint32_t *sl = scratch + (size_t)blockIdx.x * MX_BLKW(KMAX);
    // kept in a VGPR pair: the scalar registers are oversubscribed, and a
    // scalar base would be reloaded from its spill slot at every handle store
    asm volatile("" : "+v"(sl));
    uint32_t *slb = sl + KMAX * (NBX_WAVE / 2) + MX_V, *skc = slb + KMAX * NBX_WAVE;
    const int nw = A.wcnt > 0 ? A.wcnt : (T.n + NBX_WAVE - 1) / NBX_WAVE;
    const int x = blockIdx.x & 7, q8 = nw >> 3, r8 = nw & 7;
    for (int s = 0; s < 8; ++s) {
        const int y = (x + s) & 7;
        const int ybeg = y * q8 + (y < r8 ? y : r8), ycnt = q8 + (y < r8 ? 1 : 0);
        for (;;) {
            int t = 0;
            if (lane == 0) t = atomicAdd(ctr + y * MX_CTR_STRIDE, 1);
            t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
            if (t >= ycnt) break;
            mix_task<ND, KMAX>(T, A, ext, fb_list, fb_count, hk, cx, racc, sl, slb, skc, A.wbeg + ybeg + t, lane);
        }
    }
}

__global__ void k_mx_zero(int32_t *p, int cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = 0;
}

// Resident workgroups of one instance over the device, a multiple of 8
// (every XCD gets workgroups) and at least 8.
template <int ND, int KMAX>
static int mx_grid() {
    static int grid = 0;
    if (grid == 0) {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_knn_mixed<ND, KMAX>, NBX_WAVE, 0) != hipSuccess ||
            per <= 0)
            per = 8;
        const int gsz = (per * ncu) & ~7;
        grid = gsz < 8 ? 8 : gsz;
    }
    return grid;
}

static int mx_launch_grid(int nd, int km) {
    return nd == 3 ? (km == 33 ? mx_grid<3, 33>() : mx_grid<3, 65>())
                   : (km == 33 ? mx_grid<6, 33>() : mx_grid<6, 65>());
}

size_t nbx_mixed_scratch_bytes(int nd, int kh) {
    const int km = kh <= 33 ? 33 : 65;
    return (size_t)mx_launch_grid(nd, km) * MX_BLKW(km) * sizeof(uint32_t) + 8 * MX_CTR_STRIDE * sizeof(int32_t);
}

hipError_t nbx_launch_knn_mixed(const DevTree &T, const KnnArgs &A, const double *ext, int32_t *fallback_list,
                                int32_t *fallback_count, hipStream_t st, void *work, size_t work_bytes) {
    const int nw = A.wcnt > 0 ? A.wcnt : (T.n + NBX_WAVE - 1) / NBX_WAVE;
    if (nw == 0) return hipSuccess;
    if ((T.nd != 3 && T.nd != 6) || A.kh > 65 || A.kh < 2 || !T.kernf || !T.org || !T.boxf || !ext)
        return hipErrorInvalidValue;
    const int km = A.kh <= 33 ? 33 : 65;
    const int grid = mx_launch_grid(T.nd, km);
    const size_t nscr = (size_t)grid * MX_BLKW(km);
    const size_t need = nscr * sizeof(uint32_t) + 8 * MX_CTR_STRIDE * sizeof(int32_t);
    void *mem = work;
    hipError_t e = hipSuccess;
    if (mem) {
        if (work_bytes < need) return hipErrorInvalidValue;
    } else {
        e = hipMallocAsync(&mem, need, st);
        if (e != hipSuccess) return e;
    }
    uint32_t *scr = (uint32_t *)mem;
    int32_t *ctr = (int32_t *)(scr + nscr);
    hipLaunchKernelGGL(k_mx_zero, dim3(1), dim3(8 * MX_CTR_STRIDE), 0, st, ctr, 8 * MX_CTR_STRIDE);
    {
        ProfScope ps("k_knn_mixed", st);
#define MIXL(ND, KM)                                                                                            \
    hipLaunchKernelGGL((k_knn_mixed<ND, KM>), dim3(grid), dim3(NBX_WAVE), 0, st, T, A, ext, fallback_list, \
                       fallback_count, scr, ctr)
        if (T.nd == 3) {
            if (km == 33) MIXL(3, 33);
            else MIXL(3, 65);
        } else {
            if (km == 33) MIXL(6, 33);
            else MIXL(6, 65);
        }
#undef MIXL
    }
    e = hipGetLastError();
    if (!work) hipFreeAsync(mem, st);
    return e;
}

// FP64 distances for rows written by the precision-form kernel, whose rows
// are in FP32 key order: each row's entries get their FP64 d2 (summed over
// the dimensions in order, no FMA, as every FP64 distance here) and are
// sorted ascending by (d2, ID) -- the LoadNN order of the FP64 form
// (KDFindNearest.cxx:11-24) with ties to the smaller ID.  Rows and entries
// are by ID (layout 0) or by slot (layout 1).  One thread per row.
template <int ND, int KM>
__global__ __launch_bounds__(256) void k_rows_fp64(DevTree T, int32_t *nn, double *d2, int k, int layout) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= T.n) return;
    const int32_t qs = layout == 0 ? T.inv[r] : (int32_t)r;
    double q[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) q[d] = T.x[d][qs];
    double dv[KM];
    int32_t iv[KM], ev[KM];
    int32_t *row = nn + r * k;
    for (int j = 0; j < k; ++j) {
        const int32_t e = row[j];
        // a missing neighbour (-1) keeps d2 = MAXVALUE and sorts last
        const bool miss = e < 0;
        const int32_t s = miss ? 0 : (layout == 0 ? T.inv[e] : e);
        const double v = miss ? NBX_MAXVALUE : d64<ND>(T, q, s);
        const int32_t id = miss ? 0x7fffffff : (layout == 0 ? e : T.perm[e]);
        // insertion into the sorted prefix (the rows come nearly sorted)
        int p = j;
        while (p > 0 && (dv[p - 1] > v || (dv[p - 1] == v && iv[p - 1] > id))) {
            dv[p] = dv[p - 1];
            iv[p] = iv[p - 1];
            ev[p] = ev[p - 1];
            --p;
        }
        dv[p] = v;
        iv[p] = id;
        ev[p] = e;
    }
    for (int j = 0; j < k; ++j) {
        row[j] = ev[j];
        d2[r * k + j] = dv[j];
    }
}

hipError_t nbx_launch_rows_fp64(const DevTree &T, int32_t *nn, double *d2, int k, int layout, hipStream_t st) {
    if (T.n == 0) return hipSuccess;
    if (k < 1 || k > 64) return hipErrorInvalidValue;
    const int blocks = (int)((T.n + 255) / 256);
    ProfScope ps("k_rows_fp64", st);
    if (T.nd == 3) {
        if (k <= 32) hipLaunchKernelGGL((k_rows_fp64<3, 32>), dim3(blocks), dim3(256), 0, st, T, nn, d2, k, layout);
        else hipLaunchKernelGGL((k_rows_fp64<3, 64>), dim3(blocks), dim3(256), 0, st, T, nn, d2, k, layout);
    } else {
        if (k <= 32) hipLaunchKernelGGL((k_rows_fp64<6, 32>), dim3(blocks), dim3(256), 0, st, T, nn, d2, k, layout);
        else hipLaunchKernelGGL((k_rows_fp64<6, 64>), dim3(blocks), dim3(256), 0, st, T, nn, d2, k, layout);
    }
    return hipGetLastError();
}
