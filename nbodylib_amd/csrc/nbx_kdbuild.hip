// nbx_kdbuild.hip -- particle order of the implicit kd tree on gfx950.
//
// Replaces the recursive median split of KDTree::BuildNodes
// (KDTree.cxx:970-1055) with DetermineSplitDim/SpreadestPos (:459-504,
// :27-48) and MedianPos (:328-370): every node of the implicit heap-ordered
// tree over NBX_LEAF-slot leaves is split at its slot midpoint along the
// dimension of largest extent, with the lower half (by that coordinate) going
// left -- the reference's "spread" criterion and median split, applied to a
// balanced tree whose node ranges are fixed powers of two of leaves.
//
// Top levels (node ranges longer than KD_LOCAL slots) run over the whole set
// with one sorted index list per dimension: the split dimension of a node is
// read from the ends of its lists, the lower half of the chosen list flags
// the particles that go left (one bit each), and every list is stably
// partitioned by that flag (a global exclusive scan gives each element its
// rank on its side).
// Lower levels run inside one workgroup per KD_LOCAL-slot range with the
// same scheme on FP32 copies of the coordinates and 16-bit index lists in LDS.
//
// Split coordinates are compared as FP32 (order-preserving 32-bit keys): a
// split only decides which particles share a node, and every box is later
// computed from the FP64 coordinates of its particles, so the tree is exact
// whatever the rounding of the split.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "nbx_internal.h"
#include "nbx_kernels.h"

#define KD_LOCAL 2048  // slots per workgroup in the local stage (64 leaves)

__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_kd_key(const double *__restrict__ c, int64_t n, uint32_t *__restrict__ key,
                                                int32_t *__restrict__ idx) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    key[i] = fkey((float)c[i]);
    idx[i] = (int32_t)i;
}

struct KdLists {
    const double *c[6];
    int32_t *S[6];   // per-dimension index lists (current)
    int32_t *S2[6];  // next
};

// Split dimension of each node at this level: largest extent, read from the
// ends of the node's sorted lists.
__global__ __launch_bounds__(256) void k_kd_dims(KdLists L, int nd, int64_t n, int lseg, int64_t nseg,
                                                 uint8_t *__restrict__ dsel) {
    int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= nseg) return;
    const int64_t a = s << lseg;
    if (a >= n) return;
    const int64_t b = min(a + ((int64_t)1 << lseg), n) - 1;
    int best = 0;
    double bext = -1.0;
    for (int d = 0; d < nd; ++d) {
        const double e = L.c[d][L.S[d][b]] - L.c[d][L.S[d][a]];
        if (e > bext) { bext = e; best = d; }
    }
    dsel[s] = (uint8_t)best;
}

// Bit id of fw set for the particles of the lower half of each node along
// its split dimension (fw cleared first).  One bit per particle keeps the
// flag array at N/8 bytes, so the gathers below mostly hit in L2.
__global__ __launch_bounds__(256) void k_kd_clear(uint32_t *__restrict__ fw, int64_t nw) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nw) fw[i] = 0u;
}
__global__ __launch_bounds__(256) void k_kd_flag(KdLists L, int64_t n, int lseg, const uint8_t *__restrict__ dsel,
                                                 uint32_t *__restrict__ fw) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t s = i >> lseg;
    const int64_t a = s << lseg;
    const int64_t m = min(a + ((int64_t)1 << (lseg - 1)), n);
    if (i < m) {
        const int32_t id = L.S[dsel[s]][i];
        atomicOr(fw + (id >> 5), 1u << (id & 31));
    }
}

__device__ __forceinline__ int32_t kd_bit(const uint32_t *fw, int32_t id) { return (int32_t)((fw[id >> 5] >> (id & 31)) & 1u); }

// flag of the i-th element of a list, as the scan input
struct KdFlagOf {
    const int32_t *S;
    const uint32_t *fw;
    __device__ __forceinline__ int32_t operator()(int64_t i) const { return kd_bit(fw, S[i]); }
};

// Stable partition of one list inside every node: left elements keep their
// order in [a, m), right ones in [m, b); pos is the exclusive scan of the
// flags in list order.
__global__ __launch_bounds__(256) void k_kd_scatter(const int32_t *__restrict__ S, int32_t *__restrict__ S2, int64_t n,
                                                    int lseg, const uint32_t *__restrict__ fw,
                                                    const int32_t *__restrict__ pos) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int32_t id = S[i];
    const int64_t a = (i >> lseg) << lseg;
    const int64_t m = min(a + ((int64_t)1 << (lseg - 1)), n);
    const int64_t nl = (int64_t)pos[i] - pos[a];
    const int64_t dst = kd_bit(fw, id) ? a + nl : m + (i - a - nl);
    S2[dst] = id;
}

// loc[id] = position of particle id in its block of the first list, the
// local element index the local stage works with.
__global__ __launch_bounds__(256) void k_kd_loc(const int32_t *__restrict__ S0, int64_t n, int lblk,
                                                int32_t *__restrict__ loc) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    loc[S0[i]] = (int32_t)(i & (((int64_t)1 << lblk) - 1));
}

// Local stage: the node ranges of one KD_LOCAL-slot block (or the whole set
// when it is smaller), split level by level down to single leaves with the
// same per-dimension sorted lists as the global levels, here as 16-bit local
// element indices in LDS; the stable partitions use block-wide scans.
template <int ND>
__global__ __launch_bounds__(256) void k_kd_local(KdLists L, const int32_t *__restrict__ loc, int64_t n, int lblk,
                                                  int32_t *__restrict__ order) {
    __shared__ float cf[ND][KD_LOCAL];
    __shared__ uint16_t lst[2][ND][KD_LOCAL];
    __shared__ uint16_t pre[KD_LOCAL];
    __shared__ uint8_t fl[KD_LOCAL];
    __shared__ uint8_t dsel[KD_LOCAL / (2 * NBX_LEAF)];
    __shared__ int32_t wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int ss = 1 << lblk;  // block span in slots (power of two, <= KD_LOCAL)
    const int64_t a = (int64_t)blockIdx.x << lblk;
    if (a >= n) return;
    const int cnt = (int)min((int64_t)ss, n - a);
    for (int j = tid; j < cnt; j += 256) {
        const int32_t id = L.S[0][a + j];
#pragma unroll
        for (int d = 0; d < ND; ++d) cf[d][j] = (float)L.c[d][id];
        lst[0][0][j] = (uint16_t)j;
#pragma unroll
        for (int d = 1; d < ND; ++d) lst[0][d][j] = (uint16_t)loc[L.S[d][a + j]];
    }
    __syncthreads();
    const int per = ss >= 256 ? ss / 256 : 1;  // <= KD_LOCAL / 256 = 8
    const int base = tid * per;
    int cur = 0;
    for (int s = ss; s >= 2 * NBX_LEAF; s >>= 1) {
        const int nsub = ss / s;
        // split dimension of each node range: largest extent (list ends)
        for (int t = tid; t < nsub; t += 256) {
            const int u = t * s;
            int best = 0;
            if (u < cnt) {
                const int e = min(u + s, cnt) - 1;
                float bext = -1.0f;
                for (int d = 0; d < ND; ++d) {
                    const float x = cf[d][lst[cur][d][e]] - cf[d][lst[cur][d][u]];
                    if (x > bext) { bext = x; best = d; }
                }
            }
            dsel[t] = (uint8_t)best;
        }
        __syncthreads();
        for (int j = tid; j < cnt; j += 256) {
            const int sub = j / s;
            const int d = dsel[sub];
            fl[lst[cur][d][j]] = j < sub * s + s / 2 ? 1 : 0;
        }
        __syncthreads();
        for (int d = 0; d < ND; ++d) {
            int fv[KD_LOCAL / 256];
            int tsum = 0;
#pragma unroll
            for (int q = 0; q < KD_LOCAL / 256; ++q) {
                const int j = base + q;
                fv[q] = (q < per && j < cnt) ? fl[lst[cur][d][j]] : 0;
                tsum += fv[q];
            }
            int incl = tsum;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(incl, off);
                if (lane >= off) incl += y;
            }
            if (lane == 63) wsum[w] = incl;
            __syncthreads();
            int run = incl - tsum;
            for (int v = 0; v < w; ++v) run += wsum[v];
#pragma unroll
            for (int q = 0; q < KD_LOCAL / 256; ++q) {
                const int j = base + q;
                if (q < per && j < ss) pre[j] = (uint16_t)run;
                run += fv[q];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < KD_LOCAL / 256; ++q) {
                const int j = base + q;
                if (q < per && j < cnt) {
                    const int u = (j / s) * s;
                    const int m = min(u + s / 2, cnt);
                    const int nl = (int)pre[j] - (int)pre[u];
                    const int dst = fv[q] ? u + nl : m + (j - u - nl);
                    lst[cur ^ 1][d][dst] = lst[cur][d][j];
                }
            }
            __syncthreads();
        }
        cur ^= 1;
    }
    for (int j = tid; j < cnt; j += 256) order[a + j] = L.S[0][a + lst[cur][0][j]];
}

static inline int nblk256(int64_t n) { return (int)((n + 255) / 256); }

hipError_t nbx_kd_order(const double *const *c, int nd, int64_t n, int32_t nleafp, int32_t *order, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (nd != 3 && nd != 6) return hipErrorInvalidValue;
    int ltot = 0;  // log2 of the padded slot span
    while (((int64_t)1 << ltot) < (int64_t)nleafp * NBX_LEAF) ++ltot;
    int lloc = 0;
    while ((1 << lloc) < KD_LOCAL) ++lloc;
    const int lblk = ltot < lloc ? ltot : lloc;

    KdLists Lh;
    memset(&Lh, 0, sizeof(Lh));
    hipError_t e = hipSuccess;
    uint32_t *k0 = nullptr, *k1 = nullptr;
    int32_t *iota = nullptr, *pos = nullptr, *loc = nullptr;
    uint32_t *fw = nullptr;
    uint8_t *dsel = nullptr;
    const int64_t nw = (n + 31) / 32;
    typedef rocprim::transform_iterator<rocprim::counting_iterator<int64_t>, KdFlagOf, int32_t> FlagIt;
    void *tmp = nullptr;
    size_t tb = 0, tb2 = 0;
#define KDCK(x)                          \
    do {                                 \
        e = (x);                         \
        if (e != hipSuccess) goto done;  \
    } while (0)
    for (int d = 0; d < nd; ++d) {
        Lh.c[d] = c[d];
        KDCK(hipMallocAsync((void **)&Lh.S[d], sizeof(int32_t) * n, st));
        if (ltot > lblk) KDCK(hipMallocAsync((void **)&Lh.S2[d], sizeof(int32_t) * n, st));
    }
    KDCK(hipMallocAsync((void **)&k0, sizeof(uint32_t) * n, st));
    KDCK(hipMallocAsync((void **)&k1, sizeof(uint32_t) * n, st));
    KDCK(hipMallocAsync((void **)&iota, sizeof(int32_t) * n, st));
    KDCK(rocprim::radix_sort_pairs((void *)nullptr, tb, k0, k1, iota, Lh.S[0], (unsigned int)n, 0, 32, st, false));
    KDCK(rocprim::exclusive_scan((void *)nullptr, tb2, FlagIt(rocprim::counting_iterator<int64_t>(0), KdFlagOf{}),
                                 (int32_t *)nullptr, 0, (size_t)n, rocprim::plus<int32_t>(), st, false));
    KDCK(hipMallocAsync(&tmp, tb > tb2 ? tb : tb2, st));
    // one index list per dimension, sorted by that coordinate (FP32 keys)
    for (int d = 0; d < nd; ++d) {
        hipLaunchKernelGGL(k_kd_key, dim3(nblk256(n)), dim3(256), 0, st, c[d], n, k0, iota);
        KDCK(rocprim::radix_sort_pairs(tmp, tb, k0, k1, iota, Lh.S[d], (unsigned int)n, 0, 32, st, false));
    }
    if (ltot > lblk) {
        KDCK(hipMallocAsync((void **)&fw, sizeof(uint32_t) * nw, st));
        KDCK(hipMallocAsync((void **)&pos, sizeof(int32_t) * n, st));
        KDCK(hipMallocAsync((void **)&dsel, (size_t)1 << (ltot - lblk), st));
        for (int lseg = ltot; lseg > lblk; --lseg) {
            const int64_t nseg = (int64_t)1 << (ltot - lseg);
            hipLaunchKernelGGL(k_kd_dims, dim3(nblk256(nseg)), dim3(256), 0, st, Lh, nd, n, lseg, nseg, dsel);
            hipLaunchKernelGGL(k_kd_clear, dim3(nblk256(nw)), dim3(256), 0, st, fw, nw);
            hipLaunchKernelGGL(k_kd_flag, dim3(nblk256(n)), dim3(256), 0, st, Lh, n, lseg, dsel, fw);
            for (int d = 0; d < nd; ++d) {
                FlagIt fi(rocprim::counting_iterator<int64_t>(0), KdFlagOf{Lh.S[d], fw});
                KDCK(rocprim::exclusive_scan(tmp, tb2, fi, pos, 0, (size_t)n, rocprim::plus<int32_t>(), st, false));
                hipLaunchKernelGGL(k_kd_scatter, dim3(nblk256(n)), dim3(256), 0, st, Lh.S[d], Lh.S2[d], n, lseg, fw,
                                   pos);
            }
            for (int d = 0; d < nd; ++d) std::swap(Lh.S[d], Lh.S2[d]);
        }
    }
    {
        KDCK(hipMallocAsync((void **)&loc, sizeof(int32_t) * n, st));
        hipLaunchKernelGGL(k_kd_loc, dim3(nblk256(n)), dim3(256), 0, st, Lh.S[0], n, lblk, loc);
        const int64_t nb = (n + ((int64_t)1 << lblk) - 1) >> lblk;
        if (nd == 3)
            hipLaunchKernelGGL(k_kd_local<3>, dim3((unsigned)nb), dim3(256), 0, st, Lh, loc, n, lblk, order);
        else
            hipLaunchKernelGGL(k_kd_local<6>, dim3((unsigned)nb), dim3(256), 0, st, Lh, loc, n, lblk, order);
    }
    e = hipGetLastError();
done:
    for (int d = 0; d < 6; ++d) {
        if (Lh.S[d]) hipFreeAsync(Lh.S[d], st);
        if (Lh.S2[d]) hipFreeAsync(Lh.S2[d], st);
    }
    if (k0) hipFreeAsync(k0, st);
    if (k1) hipFreeAsync(k1, st);
    if (iota) hipFreeAsync(iota, st);
    if (pos) hipFreeAsync(pos, st);
    if (fw) hipFreeAsync(fw, st);
    if (dsel) hipFreeAsync(dsel, st);
    if (loc) hipFreeAsync(loc, st);
    if (tmp) hipFreeAsync(tmp, st);
#undef KDCK
    return e;
}
