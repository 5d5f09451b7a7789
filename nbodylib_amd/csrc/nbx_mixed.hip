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
// k+1 smallest packed keys (bits(d32) << 32 | slot) in a 4-ary LDS heap --
// 8-byte entries compared with one 64-bit integer compare.
//
// Error bound: with u = 2^-24 and E the half-diagonal of the leaf a
// candidate came from, |dist32 - dist64| <= 12 u dist32 + 5 u E.  Per
// component the two offset roundings and the difference add at most
// u(|q - o| + |x - o| + |t|) <= u(dist + 2E + dist); the FMA chain over ND
// <= 6 squares adds ND u relative to the sum, sqrt halves it and adds u:
// at most (2 + ND/2 + 1) u dist + 2 u E = 6 u dist + 2 u E in 6-D, so the
// constants keep a factor 2 of margin.  Each lane tracks E_max over the leaves it evaluated
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
// Heap keys: (bits(d32) << 32 | slot) + 2^52.  The offset adds one to the
// exponent field of the key read as an FP64 number, so every key is a
// normal, finite, positive double whose order is the integer order: the
// largest of several keys is one v_max_f64 per pair.
#define MX_KOFF (1ULL << 52)
#define MX_EMPTY (0x7F7FFFFFFFFFFFFFULL + MX_KOFF)  // empty heap entry (FLT_MAX, slot ~0)

__device__ __forceinline__ uint64_t mkkey(float d, int32_t s) {
    return (((uint64_t)__float_as_uint(d) << 32) | (uint32_t)s) + MX_KOFF;
}
__device__ __forceinline__ float keyd(uint64_t k) { return __uint_as_float((uint32_t)((k - MX_KOFF) >> 32)); }
__device__ __forceinline__ int32_t keys(uint64_t k) { return (int32_t)(uint32_t)k; }
__device__ __forceinline__ double kasd(uint64_t k) { return __longlong_as_double((long long)k); }
__device__ __forceinline__ uint64_t dask(double v) { return (uint64_t)__double_as_longlong(v); }

// Largest of four keys and its index 0..3.
__device__ __forceinline__ void kmax4(uint64_t x0, uint64_t x1, uint64_t x2, uint64_t x3, uint64_t &bc, int &bi) {
    bc = dask(fmax(fmax(kasd(x0), kasd(x1)), fmax(kasd(x2), kasd(x3))));
    bi = bc == x1 ? 1 : (bc == x2 ? 2 : (bc == x3 ? 3 : 0));
}

#define HK(p) hk[(p) * NBX_WAVE + lane]

// The kh kept keys live in G groups of S LDS slots (slot p = g * S + s,
// G * S = KMAX >= kh; slots p >= kh hold 0, which never outranks a key),
// with every group's largest key and its slot in registers.  Replacing the
// largest kept key is one LDS store and one S-wide reload of its group --
// one LDS round trip, where a 4-ary heap of 33 needs three dependent ones.
template <int KMAX>
struct GShape {
    static constexpr int G = KMAX <= 33 ? 3 : 5;
    static constexpr int S = KMAX / G;
    static_assert(G * S == KMAX, "group shape must tile KMAX");
};

template <int G>
struct GTop {
    uint64_t gm[G];  // largest key of each group
    int gs[G];       // its slot within the group
    int rg;          // group holding the largest kept key
};

template <int G, int S>
__device__ __forceinline__ void gtop_init(uint64_t *hk, int lane, int kh, GTop<G> &H) {
    for (int p = 0; p < G * S; ++p) HK(p) = p < kh ? MX_EMPTY : 0ULL;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        H.gm[g] = g * S < kh ? MX_EMPTY : 0ULL;
        H.gs[g] = 0;
    }
    H.rg = 0;
}

// Largest of S keys (a max tree) and the first slot holding it.
template <int S>
__device__ __forceinline__ void gmax(const uint64_t (&x)[S], uint64_t &bc, int &bs) {
    double v[S];
#pragma unroll
    for (int u = 0; u < S; ++u) v[u] = kasd(x[u]);
#pragma unroll
    for (int w = 1; w < S; w <<= 1)
#pragma unroll
        for (int u = 0; u + w < S; u += 2 * w) v[u] = fmax(v[u], v[u + w]);
    bc = dask(v[0]);
    bs = S - 1;
#pragma unroll
    for (int u = S - 2; u >= 0; --u) bs = x[u] == bc ? u : bs;
}

// Replace the largest kept key (root) by `key` in the lanes with ins set;
// the others store the root back into its own slot, so nothing branches.
// Returns the new largest.
template <int G, int S>
__device__ __forceinline__ uint64_t gtop_replace(uint64_t *hk, int lane, GTop<G> &H, bool ins, uint64_t key,
                                                 uint64_t root) {
    const int g0 = H.rg;
    int s0 = H.gs[0];
#pragma unroll
    for (int g = 1; g < G; ++g) s0 = g0 == g ? H.gs[g] : s0;
    const int b = g0 * S;
    HK(b + s0) = ins ? key : root;
    uint64_t x[S];
#pragma unroll
    for (int u = 0; u < S; ++u) x[u] = HK(b + u);
    uint64_t bc;
    int bs;
    gmax<S>(x, bc, bs);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        H.gm[g] = g0 == g ? bc : H.gm[g];
        H.gs[g] = g0 == g ? bs : H.gs[g];
    }
    uint64_t top;
    gmax<G>(H.gm, top, H.rg);
    return top;
}

// Sift v down from slot p of a 4-ary max-heap of m <= KMAX entries; KMAX is
// 1 mod 4, so the four children of any parent below m lie inside the KMAX
// slots; children at positions >= m read as 0, which never outranks a key.
__device__ __forceinline__ void kheap_sift_at(uint64_t *hk, int lane, int m, int p, uint64_t v) {
    bool go = true;
#pragma unroll
    for (int lv = 0; lv < 3; ++lv) {
        const int c = 4 * p + 1;
        go = go && c < m;
        const int cc = go ? c : 1;
        uint64_t x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t t = HK(cc + u);
            x[u] = cc + u < m ? t : 0ULL;
        }
        uint64_t bc;
        int bi;
        kmax4(x[0], x[1], x[2], x[3], bc, bi);
        go = go && bc > v;
        HK(p) = go ? bc : v;
        p = go ? cc + bi : p;
    }
    HK(p) = v;
}


template <int ND>
struct MixCtx {
    double q[ND];
    bool act;
    int32_t self;
    uint64_t root;  // current largest key kept
    double emax;    // largest leaf half-diagonal that matters for this lane
};

template <int ND, int G, int S>
__device__ __forceinline__ void mix_leaf(const DevTree &T, const double *ext, int leaf, MixCtx<ND> &c, bool want,
                                         uint64_t *hk, float *cx, int lane, GTop<G> &H) {
    const int base = leaf * NBX_LEAF;
    // the LDS copy of the leaf for phase 2: loaded now, stored after phase 1
    // so the vector-load latency overlaps the candidate tests
    float st[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) st[d] = T.xf[d][base + (lane & (NBX_LEAF - 1))];
    // query relative to this leaf's centre, rounded once to FP32
    const double *o = T.org + (size_t)leaf * ND;
    float qv[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) qv[d] = (float)(c.q[d] - ((cdptr)o)[d]);
    const float th = keyd(c.root);
    uint32_t mask = 0;
    // candidates in 32- or 16-byte scalar-load chunks per dimension
    constexpr int CH = ND == 6 ? 4 : 8;
    typedef float fxc __attribute__((ext_vector_type(CH)));
    typedef const __attribute__((address_space(4))) fxc *cfcptr;
#pragma unroll
    for (int jj = 0; jj < NBX_LEAF; jj += CH) {
        fxc pv[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) pv[d] = *(cfcptr)(T.xf[d] + base + jj);
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            float t = qv[0] - pv[0][u];
            float dd = t * t;
#pragma unroll
            for (int d = 1; d < ND; ++d) {
                t = qv[d] - pv[d][u];
                dd = __builtin_fmaf(t, t, dd);
            }
            mask |= (dd <= th) ? (1u << (jj + u)) : 0u;
        }
    }
    if (c.self >= base && c.self < base + NBX_LEAF) mask &= ~(1u << (c.self - base));
    if (!c.act) mask = 0;
    if (c.act && (want || mask)) c.emax = fmax(c.emax, ((cdptr)ext)[leaf]);
    if (lane < NBX_LEAF) {
#pragma unroll
        for (int d = 0; d < ND; ++d) cx[d * NBX_LEAF + lane] = st[d];
    }
    __syncthreads();
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
        const uint64_t key = mkkey(dd, base + j);
        c.root = gtop_replace<G, S>(hk, lane, H, has && key < c.root, key, c.root);
    }
    __syncthreads();
}

// Stackless depth-first walk: the ancestor of the wave's own leaves first,
// else the child nearer to the wave's first query; firstright[dep] records
// which child was entered first at each depth.  Two loop forms visit the
// nodes in the same order; the compiler schedules them differently, and the
// measured faster one is used for each dimensionality (3-D: climb to the
// next unvisited sibling in an inner loop; 6-D: one node step per trip).
template <int ND, int G, int S>
__device__ __forceinline__ void mix_walk(const DevTree &T, const double *ext, MixCtx<ND> &c, const float *qc, int anc,
                                         uint64_t *hk, float *cx, int lane, GTop<G> &H) {
    const int D = T.depth;
    int node = 1, dep = 0;
    uint64_t firstright = 0;
    if constexpr (ND == 3) {
        for (;;) {
            cdptr bx = (cdptr)node_box(T, node, ND);
            const bool want = c.act && box_lb<ND>(c.q, bx) <= (double)keyd(c.root);
            if (ballot(want)) {
                if (dep == D) {
                    mix_leaf<ND, G, S>(T, ext, node - T.nleafp, c, want, hk, cx, lane, H);
                } else {
                    int first;
                    if ((anc >> (D - dep)) == node) first = anc >> (D - dep - 1);
                    else first = 2 * node + (first_is_right<ND>(T, node, qc) ? 1 : 0);
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
                cdptr bx = (cdptr)node_box(T, node, ND);
                const bool want = c.act && box_lb<ND>(c.q, bx) <= (double)keyd(c.root);
                if (ballot(want)) {
                    if (dep == D) {
                        mix_leaf<ND, G, S>(T, ext, node - T.nleafp, c, want, hk, cx, lane, H);
                        down = false;
                    } else {
                        const int c0 = 2 * node;
                        int first = c0;
                        if (anc >= 0 && (anc >> (D - dep)) == node) {
                            first = anc >> (D - dep - 1);
                        } else if (first_is_right<ND>(T, node, qc)) {
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

// k = kh - 1 neighbours; A.dens: NBX_DO_NONE / NBX_DO_SYM / NBX_DO_GATHER.
template <int ND, int KMAX>
__global__ __launch_bounds__(64) void k_knn_mixed(DevTree T, KnnArgs A, const double *ext, int32_t *fb_list,
                                                   int32_t *fb_count) {
    __shared__ uint64_t hk[KMAX * NBX_WAVE];
    __shared__ float cx[ND * NBX_LEAF];
    __shared__ double racc[NBX_WAVE];  // symmetric scatter into the wave's own queries
    const int lane = threadIdx.x;
    const int g = xcd_wave(blockIdx.x, gridDim.x);
    const int kh = A.kh;  // k + 1
    const int k = kh - 1;
    const int qi = g * NBX_WAVE + lane;
    const bool valid = qi < T.n;
    const int kres = T.kernres;
    MixCtx<ND> c;
#pragma unroll
    for (int d = 0; d < ND; ++d) c.q[d] = valid ? T.x[d][qi] : 0.0;
    c.act = valid;
    c.self = qi;
    c.root = MX_EMPTY;
    c.emax = 0.0;
    constexpr int G = GShape<KMAX>::G, S = GShape<KMAX>::S;
    GTop<G> H;
    gtop_init<G, S>(hk, lane, kh, H);
    float qc[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) qc[d] = (float)readfirst_d(c.q[d]);
    mix_walk<ND, G, S>(T, ext, c, qc, T.nleafp + 2 * g, hk, cx, lane, H);

    // ---------------------------------------------------------------- band
    bool ok = valid;
    if (valid) {
        // heap sort of the kh kept keys: ascending keys in slots 0 .. kh-1
        for (int p = (kh - 2) >> 2; p >= 0; --p) kheap_sift_at(hk, lane, kh, p, HK(p));
        for (int m = kh - 1; m > 0; --m) {
            uint64_t top = HK(0), last = HK(m);
            HK(m) = top;
            kheap_sift_at(hk, lane, m, 0, last);
        }
        if (HK(kh - 1) == MX_EMPTY) ok = false;  // fewer than k+1 candidates
    }
    const double eu = 12.0 * MX_U, ea = 5.0 * MX_U * c.emax;
    int i0 = k - 1;
    int32_t bs[4];
    double bd[4];
    double Dk = 0.0;
    if (ok) {
        // band: entries whose interval overlaps the k-th entry's
        double dk = sqrt((double)keyd(HK(k - 1)));
        double lo_k = dk - (eu * dk + ea);
        while (i0 > 0) {
            double di = sqrt((double)keyd(HK(i0 - 1)));
            if (di + (eu * di + ea) < lo_k) break;
            --i0;
        }
        // a certain entry that could be an exact duplicate (d2 == 0) must be checked too
        double d0 = sqrt((double)keyd(HK(0)));
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
                bs[b] = keys(HK(i0 + b));
                bd[b] = d64<ND>(T, c.q, bs[b]);
                if (bd[b] == 0.0) { nzero++; drop = b; }
            }
        }
        if (nzero > 1) ok = false;
        if (drop < 0) {
            drop = 0;
#pragma unroll
            for (int b = 1; b < 4; ++b)
                if (b < nb && (bd[b] > bd[drop] || (bd[b] == bd[drop] && T.perm[bs[b]] > T.perm[bs[drop]]))) drop = b;
        }
        Dk = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (b < nb && b != drop && bd[b] > Dk) Dk = bd[b];
        // certificate against everything outside the heap
        double dl = sqrt((double)keyd(HK(kh - 1)));
        double lout = dl - (eu * dl + ea);
        if (!(lout > 0.0 && Dk < lout * lout)) ok = false;
        if (ok) {
            // the kept band entries fill positions i0 .. k-1 (FP32 order)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (b < nb && b != drop) HK(i0 + b - (b > drop ? 1 : 0)) = (uint64_t)(uint32_t)bs[b];
        }
    }
    // exact verification for lanes the error bound could not decide: keep
    // the k smallest (d2, ID) of the k+1 candidates in FP64, then count in
    // FP64 the particles below the k-th of them
    bool verify = valid && !ok && HK(kh - 1) != MX_EMPTY;
    int32_t IDk = 0;
    if (verify) {
        int zero_at = -1, nzero = 0, maxp = -1;
        double dmax = -1.0;
        int32_t imax = -1;
        for (int p = 0; p < kh; ++p) {
            int32_t s = keys(HK(p));
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
            for (int p = drop; p < kh - 1; ++p) HK(p) = HK(p + 1);
            Dk = 0.0;
            IDk = -1;
            for (int p = 0; p < k; ++p) {
                int32_t s = keys(HK(p));
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
        double acc = 0.0;
        // neighbours in groups of 8 so their coordinate/mass gathers are in
        // flight together
        for (int p0 = 0; p0 < k; p0 += 8) {
            int32_t sv[8];
            double d2v[8], mv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                sv[u] = p0 + u < k ? keys(HK(p0 + u)) : qi;
                d2v[u] = d64<ND>(T, c.q, sv[u]);
                mv[u] = T.mass[sv[u]];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (p0 + u >= k) break;
                const int32_t s = sv[u];
                double W = wsm_at(sqrt(d2v[u]), hi, T.kern, kres);
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
        if (A.hout) A.hout[qi] = hi;
    }
    __syncthreads();
    if (sym && valid) unsafeAtomicAdd(A.rho + qi, racc_own + racc[lane]);

    // ---------------------------------------------------------------- outputs
    if (A.nn == nullptr) return;
    const uint64_t okm = ballot(ok);
    __syncthreads();
    const int total = NBX_WAVE * k;
    for (int e = lane; e < total; e += NBX_WAVE) {
        int r = e / k, t = e - r * k;
        int qr = g * NBX_WAVE + r;
        if (qr >= T.n) break;
        if (!((okm >> r) & 1ULL)) continue;
        int32_t s = keys(hk[t * NBX_WAVE + r]);
        int64_t row = A.layout == 1 ? qr : T.perm[qr];
        A.nn[row * k + t] = A.layout == 1 ? s : T.perm[s];
    }
}

hipError_t nbx_launch_knn_mixed(const DevTree &T, const KnnArgs &A, const double *ext, int32_t *fallback_list,
                                int32_t *fallback_count, hipStream_t st) {
    int blocks = (T.n + NBX_WAVE - 1) / NBX_WAVE;
    if (blocks == 0) return hipSuccess;
    if ((T.nd != 3 && T.nd != 6) || A.kh > 65 || A.kh < 2 || !T.kernf || !T.org || !ext) return hipErrorInvalidValue;
    ProfScope ps("k_knn_mixed", st);
#define MIXL(ND, KM)                                                                                                  \
    hipLaunchKernelGGL((k_knn_mixed<ND, KM>), dim3(blocks), dim3(NBX_WAVE), 0, st, T, A, ext, fallback_list, \
                       fallback_count)
    if (T.nd == 3) {
        if (A.kh <= 33) MIXL(3, 33);
        else MIXL(3, 65);
    } else {
        if (A.kh <= 33) MIXL(6, 33);
        else MIXL(6, 65);
    }
#undef MIXL
    return hipGetLastError();
}
