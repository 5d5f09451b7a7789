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
// Particles (ID + one order-preserving 32-bit FP32 key per dimension) are
// moved physically, level by level, so every pass streams through HBM:
//  * top levels (node ranges longer than KD_LOCAL slots): per-node extents
//    by block reductions and atomics, the exact median of the split
//    coordinate by a 4-pass 8-bit radix select (block histograms in LDS,
//    per-node histograms in HBM), then a stable partition: one exclusive
//    scan of packed (below, equal) counts gives every particle its rank on
//    its side, equal keys filling the left side in order;
//  * lower levels run inside one workgroup per KD_LOCAL-slot range on FP32
//    coordinates and per-dimension lists of 16-bit local indices in LDS
//    (sorted once by bitonic sorts, then stably partitioned by block scans).
//
// Split coordinates are compared as FP32: a split only decides which
// particles share a node, and every box is later computed from the FP64
// coordinates of its particles, so the tree is exact whatever the rounding.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "nbx_internal.h"
#include "nbx_kernels.h"

#define KD_LOCAL 2048  // slots per workgroup in the local stage (64 leaves)
#define KD_T 256       // threads per block
#define KD_PER 8       // elements per thread in the global passes (KD_T * KD_PER = KD_LOCAL)
#define KD_LT 1024     // threads per block in the local stage
#define KD_LPER 2      // elements per thread there (KD_LT * KD_LPER = KD_LOCAL)

__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdekey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

struct KdState {
    int32_t *id;       // particle ID at each position (current order)
    uint32_t *key[6];  // split keys per dimension (current order)
    int32_t *id2;      // next order
    uint32_t *key2[6];
};

struct KdCoords {
    const double *c[6];
};

__global__ __launch_bounds__(KD_T) void k_kd_init(KdCoords C, int nd, int64_t n, KdState S) {
    int64_t i = (int64_t)blockIdx.x * KD_T + threadIdx.x;
    if (i >= n) return;
    S.id[i] = (int32_t)i;
    for (int d = 0; d < nd; ++d) S.key[d][i] = fkey((float)C.c[d][i]);
}

// ext holds (min, max) key pairs: min slots start at ~0, max slots at 0.
__global__ __launch_bounds__(KD_T) void k_kd_ext_init(uint32_t *__restrict__ ext, int64_t npairs) {
    int64_t i = (int64_t)blockIdx.x * KD_T + threadIdx.x;
    if (i < npairs) {
        ext[2 * i] = 0xffffffffu;
        ext[2 * i + 1] = 0u;
    }
}

__global__ __launch_bounds__(KD_T) void k_kd_zero(uint32_t *__restrict__ p, int64_t cnt) {
    int64_t i = (int64_t)blockIdx.x * KD_T + threadIdx.x;
    if (i < cnt) p[i] = 0u;
}

// Per-node key extents.  A block covers KD_LOCAL consecutive positions, all
// in one node (node ranges at these levels are longer).
template <int ND>
__global__ __launch_bounds__(KD_T) void k_kd_ext(KdState S, int64_t n, int lseg, uint32_t *__restrict__ ext) {
    __shared__ uint32_t smn[KD_T / 64][ND], smx[KD_T / 64][ND];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * KD_LOCAL;
    uint32_t mn[ND], mx[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        mn[d] = 0xffffffffu;
        mx[d] = 0u;
    }
#pragma unroll
    for (int q = 0; q < KD_PER; ++q) {
        const int64_t i = i0 + q * KD_T + t;
        if (i < n) {
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const uint32_t k = S.key[d][i];
                mn[d] = min(mn[d], k);
                mx[d] = max(mx[d], k);
            }
        }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn[d] = min(mn[d], (uint32_t)__shfl_xor((int)mn[d], off));
            mx[d] = max(mx[d], (uint32_t)__shfl_xor((int)mx[d], off));
        }
        if (lane == 0) {
            smn[w][d] = mn[d];
            smx[w][d] = mx[d];
        }
    }
    __syncthreads();
    if (t < ND && i0 < n) {
        uint32_t a = smn[0][t], b = smx[0][t];
        for (int v = 1; v < KD_T / 64; ++v) {
            a = min(a, smn[v][t]);
            b = max(b, smx[v][t]);
        }
        const int64_t s = i0 >> lseg;
        atomicMin(ext + (s * ND + t) * 2, a);
        atomicMax(ext + (s * ND + t) * 2 + 1, b);
    }
}

// Split dimension (largest extent) and the left count of every node; the
// radix select state starts with prefix 0 and rank = left count.
__global__ __launch_bounds__(KD_T) void k_kd_dims(const uint32_t *__restrict__ ext, int nd, int64_t n, int lseg,
                                                  int64_t nseg, uint8_t *__restrict__ dsel, uint32_t *__restrict__ pre,
                                                  int32_t *__restrict__ need) {
    int64_t s = (int64_t)blockIdx.x * KD_T + threadIdx.x;
    if (s >= nseg) return;
    const int64_t a = s << lseg;
    if (a >= n) return;
    int best = 0;
    float bext = -1.0f;
    for (int d = 0; d < nd; ++d) {
        const float e = fdekey(ext[(s * nd + d) * 2 + 1]) - fdekey(ext[(s * nd + d) * 2]);
        if (e > bext) {
            bext = e;
            best = d;
        }
    }
    dsel[s] = (uint8_t)best;
    pre[s] = 0u;
    const int64_t m = min(a + ((int64_t)1 << (lseg - 1)), n);
    need[s] = (int32_t)(m - a);
}

// Radix select pass: histogram of the 8-bit digit at `sh` of the split keys
// whose higher digits equal the node's prefix.
__global__ __launch_bounds__(KD_T) void k_kd_hist(KdState S, int64_t n, int lseg, int sh,
                                                  const uint8_t *__restrict__ dsel, const uint32_t *__restrict__ pre,
                                                  uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0u;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * KD_LOCAL;
    const int64_t s = i0 >> lseg;
    const uint32_t *K = S.key[dsel[s]];
    const uint32_t p = pre[s];
    const uint32_t hm = sh >= 24 ? 0u : (0xffffffffu << (sh + 8));  // higher digits already fixed
    const int lane = t & 63;
#pragma unroll
    for (int q = 0; q < KD_PER; ++q) {
        const int64_t i = i0 + q * KD_T + t;
        uint32_t k = 0;
        if (i < n) k = K[i];
        const bool in = i < n && (k & hm) == (p & hm);
        const uint32_t bin = (k >> sh) & 255u;
        // the high digits of a node's keys are mostly one bin: a wave whose
        // counted keys all share it adds them with one LDS atomic
        const uint64_t m = ballot(in);
        if (!m) continue;
        const uint32_t b0 = (uint32_t)__shfl((int)bin, __builtin_ctzll(m));
        if (!ballot(in && bin != b0)) {
            if (lane == 0) atomicAdd(&h[b0], (uint32_t)__popcll(m));
        } else if (in) {
            atomicAdd(&h[bin], 1u);
        }
    }
    __syncthreads();
    if (h[t]) atomicAdd(hist + s * 256 + t, h[t]);
}

// Radix select step: the digit bin holding the node's remaining rank.  One
// wave per node, four bins per lane and a wave prefix sum; the node's bins
// are cleared for the next pass.
__global__ __launch_bounds__(KD_T) void k_kd_select(int64_t n, int lseg, int64_t nseg, int sh,
                                                    uint32_t *__restrict__ hist, uint32_t *__restrict__ pre,
                                                    int32_t *__restrict__ need) {
    const int lane = threadIdx.x & 63;
    const int64_t s = (int64_t)blockIdx.x * (KD_T / 64) + (threadIdx.x >> 6);
    if (s >= nseg || (s << lseg) >= n) return;
    uint32_t *h = hist + s * 256;
    const uint4 v = *(const uint4 *)(h + 4 * lane);
    const uint32_t c4 = v.x + v.y + v.z + v.w;
    uint32_t inc = c4;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, off);
        if (lane >= off) inc += y;
    }
    const uint32_t exc = inc - c4;
    const uint32_t r = (uint32_t)need[s];
    const uint64_t bm = ballot(exc < r && r <= inc);
    const int L = bm ? __builtin_ctzll(bm) : 63;
    if (lane == L) {
        uint32_t rr = r - exc;
        int b = 4 * lane;
        if (rr > v.x) {
            rr -= v.x;
            ++b;
            if (rr > v.y) {
                rr -= v.y;
                ++b;
                if (rr > v.z) {
                    rr -= v.z;
                    ++b;
                }
            }
        }
        pre[s] |= (uint32_t)b << sh;
        need[s] = (int32_t)rr;
    }
    *(uint4 *)(h + 4 * lane) = make_uint4(0u, 0u, 0u, 0u);
}

// Per-block (below, equal) counts of the split keys against the node's
// pivot, packed as below << 32 | equal.  A block covers KD_LOCAL positions
// of one node.
__global__ __launch_bounds__(KD_T) void k_kd_count(KdState S, int64_t n, int lseg, const uint8_t *__restrict__ dsel,
                                                   const uint32_t *__restrict__ pre, uint64_t *__restrict__ blk) {
    __shared__ uint32_t wb[KD_T / 64], we[KD_T / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * KD_LOCAL;
    const int64_t s = i0 >> lseg;
    const uint32_t *K = S.key[dsel[s]];
    const uint32_t p = pre[s];
    uint32_t nb = 0, ne = 0;
#pragma unroll
    for (int q = 0; q < KD_PER; ++q) {
        const int64_t i = i0 + q * KD_T + t;
        if (i < n) {
            const uint32_t k = K[i];
            nb += k < p;
            ne += k == p;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        nb += (uint32_t)__shfl_xor((int)nb, off);
        ne += (uint32_t)__shfl_xor((int)ne, off);
    }
    if (lane == 0) {
        wb[w] = nb;
        we[w] = ne;
    }
    __syncthreads();
    if (t == 0) {
        uint64_t b = 0, e = 0;
        for (int v = 0; v < KD_T / 64; ++v) {
            b += wb[v];
            e += we[v];
        }
        blk[blockIdx.x] = (b << 32) | e;
    }
}

// Exclusive scan of the packed block counts (one workgroup; both halves are
// counts below 2^32, so the packed sums never carry between them).
#define KD_ST 1024
__global__ __launch_bounds__(KD_ST) void k_kd_bscan(const uint64_t *__restrict__ blk, int64_t nb,
                                                    uint64_t *__restrict__ out) {
    __shared__ uint64_t ws[KD_ST / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t per = (nb + KD_ST - 1) / KD_ST;
    const int64_t b0 = t * per, b1 = min(b0 + per, nb);
    uint64_t sum = 0;
    for (int64_t b = b0; b < b1; ++b) sum += blk[b];
    uint64_t incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint64_t run = incl - sum;
    for (int v = 0; v < w; ++v) run += ws[v];
    for (int64_t b = b0; b < b1; ++b) {
        out[b] = run;
        run += blk[b];
    }
}

template <int ND>
__device__ __forceinline__ void kd_load8(const uint32_t *__restrict__ p, int64_t i, int64_t n, uint32_t (&v)[KD_PER]) {
    if (i + KD_PER <= n) {
        const uint4 a = *(const uint4 *)(p + i), b = *(const uint4 *)(p + i + 4);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else {
#pragma unroll
        for (int q = 0; q < KD_PER; ++q) v[q] = i + q < n ? p[i + q] : 0u;
    }
}

// Stable partition of one KD_LOCAL block: below-pivot keys and the first
// `need` equal ones (in order) fill the node's left half [a, m), the rest
// fill [m, b) in order.  Each thread moves KD_PER consecutive positions; its
// rank on each side is the block's packed prefix plus a block scan.  When
// the next level also runs here (ext != null) the per-dimension key extents
// of the two child nodes are reduced on the way.
template <int ND>
__global__ __launch_bounds__(KD_T) void k_kd_move(KdState S, int64_t n, int lseg, const uint8_t *__restrict__ dsel,
                                                  const uint32_t *__restrict__ pre, const int32_t *__restrict__ need,
                                                  const uint64_t *__restrict__ bpre, uint32_t *__restrict__ ext) {
    __shared__ uint32_t ws[KD_T / 64];
    __shared__ uint32_t red[KD_T / 64][4 * ND];
    __shared__ uint32_t stg[(1 + ND) * KD_LOCAL];  // the block's elements in destination order
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * KD_LOCAL;
    const int64_t s = i0 >> lseg;
    const int64_t a = s << lseg;
    const int64_t m = min(a + ((int64_t)1 << (lseg - 1)), n);
    const int ds = dsel[s];
    const uint32_t p = pre[s];
    const int64_t ne = need[s];
    const uint64_t base = bpre[blockIdx.x] - bpre[a / KD_LOCAL];  // packed counts of the node before this block
    const int64_t i = i0 + (int64_t)t * KD_PER;
    uint32_t kk[ND][KD_PER], id[KD_PER];
    kd_load8<ND>((const uint32_t *)S.id, i, n, id);
#pragma unroll
    for (int d = 0; d < ND; ++d) kd_load8<ND>(S.key[d], i, n, kk[d]);
    // packed (below << 16 | equal) flags, thread-local exclusive prefix
    uint32_t fl[KD_PER], tsum = 0;
#pragma unroll
    for (int q = 0; q < KD_PER; ++q) {
        uint32_t k = kk[0][q];
#pragma unroll
        for (int d = 1; d < ND; ++d)
            if (d == ds) k = kk[d][q];
        const uint32_t f = (i + q < n) ? (k < p ? 0x10000u : (k == p ? 1u : 0u)) : 0u;
        fl[q] = tsum;
        tsum += f;
    }
    uint32_t incl = tsum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint32_t run = incl - tsum, tot = 0;
    for (int v = 0; v < KD_T / 64; ++v) {
        run += v < w ? ws[v] : 0u;
        tot += ws[v];
    }
    const int64_t lb0 = (int64_t)(base >> 32) + (run >> 16), eb0 = (int64_t)(base & 0xffffffffu) + (run & 0xffffu);
    // left elements of the node before this block and through its end: the
    // block's left elements go to [a + lrb, a + lre), its right ones follow
    // the node's earlier right elements from m
    const int64_t bb = (int64_t)(base >> 32), be = (int64_t)(base & 0xffffffffu);
    const int64_t lrb = bb + (be < ne ? be : ne);
    const int64_t lre = bb + (tot >> 16) + (be + (tot & 0xffffu) < ne ? be + (tot & 0xffffu) : ne);
    const int nl = (int)(lre - lrb);
    uint32_t mn[2][ND], mx[2][ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) mn[0][d] = mn[1][d] = 0xffffffffu, mx[0][d] = mx[1][d] = 0u;
#pragma unroll
    for (int q = 0; q < KD_PER; ++q) {
        if (i + q < n) {
            uint32_t k = kk[0][q];
#pragma unroll
            for (int d = 1; d < ND; ++d)
                if (d == ds) k = kk[d][q];
            const int64_t lb = lb0 + (fl[q] >> 16), eb = eb0 + (fl[q] & 0xffffu);
            const bool left = k < p || (k == p && eb < ne);
            const int64_t lr = lb + (eb < ne ? eb : ne);
            const int j = (int)(i + q - i0), jl = (int)(lr - lrb);
            const int sp = left ? jl : nl + (j - jl);
            stg[sp] = id[q];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                stg[(1 + d) * KD_LOCAL + sp] = kk[d][q];
                const uint32_t lo = min(mn[0][d], kk[d][q]), hi = max(mx[0][d], kk[d][q]);
                const uint32_t lo1 = min(mn[1][d], kk[d][q]), hi1 = max(mx[1][d], kk[d][q]);
                mn[0][d] = left ? lo : mn[0][d];
                mx[0][d] = left ? hi : mx[0][d];
                mn[1][d] = left ? mn[1][d] : lo1;
                mx[1][d] = left ? mx[1][d] : hi1;
            }
        }
    }
    __syncthreads();
    // coalesced write-out of the two runs
    const int cnt = (int)min((int64_t)KD_LOCAL, n - i0);
    const int64_t rb = m + (i0 - a - lrb);
    for (int j = t; j < cnt; j += KD_T) {
        const int64_t dst = j < nl ? a + lrb + j : rb + (j - nl);
        S.id2[dst] = (int32_t)stg[j];
#pragma unroll
        for (int d = 0; d < ND; ++d) S.key2[d][dst] = stg[(1 + d) * KD_LOCAL + j];
    }
    if (ext == nullptr) return;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                mn[c][d] = min(mn[c][d], (uint32_t)__shfl_xor((int)mn[c][d], off));
                mx[c][d] = max(mx[c][d], (uint32_t)__shfl_xor((int)mx[c][d], off));
            }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                red[w][(c * ND + d) * 2] = mn[c][d];
                red[w][(c * ND + d) * 2 + 1] = mx[c][d];
            }
    }
    __syncthreads();
    if (t < 4 * ND) {
        const int c = t / (2 * ND), d = (t >> 1) % ND, hi = t & 1;
        uint32_t v = red[0][t];
        for (int u = 1; u < KD_T / 64; ++u) v = hi ? max(v, red[u][t]) : min(v, red[u][t]);
        const int64_t cs = 2 * s + c;  // child node index on the next level
        const bool empty = hi ? v == 0u : v == 0xffffffffu;
        if (!empty && (cs << (lseg - 1)) < n) {
            uint32_t *e = ext + (cs * ND + d) * 2 + hi;
            if (hi)
                atomicMax(e, v);
            else
                atomicMin(e, v);
        }
    }
}

// Local stage: the node ranges of one KD_LOCAL-slot block (or the whole set
// when it is smaller), split level by level down to single leaves with
// per-dimension sorted lists of 16-bit local indices in LDS; the stable
// partitions use block-wide scans.
template <int ND>
__global__ __launch_bounds__(KD_LT) void k_kd_local(KdState S, int64_t n, int lblk, int32_t *__restrict__ order) {
    __shared__ float cf[ND][KD_LOCAL];
    __shared__ uint16_t lst[2][ND][KD_LOCAL];
    __shared__ uint64_t srt[KD_LOCAL];
    __shared__ uint16_t pref[KD_LOCAL];
    __shared__ uint8_t fl[KD_LOCAL];
    __shared__ uint8_t dsel[KD_LOCAL / (2 * NBX_LEAF)];
    __shared__ int32_t wsum[KD_LT / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int ss = 1 << lblk;  // block span in slots (power of two, <= KD_LOCAL)
    const int64_t a = (int64_t)blockIdx.x << lblk;
    if (a >= n) return;
    const int cnt = (int)min((int64_t)ss, n - a);
    for (int j = tid; j < cnt; j += KD_LT) {
#pragma unroll
        for (int d = 0; d < ND; ++d) cf[d][j] = fdekey(S.key[d][a + j]);
    }
    // one sorted list per dimension: bitonic sort of (key, element) over the
    // block (padding keys sort last)
    for (int d = 0; d < ND; ++d) {
        for (int j = tid; j < ss; j += KD_LT)
            srt[j] = j < cnt ? (((uint64_t)S.key[d][a + j] << 16) | (uint64_t)j) : ~0ULL;
        __syncthreads();
        for (int k = 2; k <= ss; k <<= 1) {
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                // one compare-exchange per pair (i, i | jj), i with bit jj clear
                for (int pi = tid; pi < ss / 2; pi += KD_LT) {
                    const int i = ((pi & ~(jj - 1)) << 1) | (pi & (jj - 1)), ixj = i | jj;
                    const bool up = (i & k) == 0;
                    const uint64_t x = srt[i], y = srt[ixj];
                    if ((x > y) == up) {
                        srt[i] = y;
                        srt[ixj] = x;
                    }
                }
                __syncthreads();
            }
        }
        for (int j = tid; j < cnt; j += KD_LT) lst[0][d][j] = (uint16_t)(srt[j] & 0xffffu);
        __syncthreads();
    }
    const int per = ss >= KD_LT ? ss / KD_LT : 1;  // <= KD_LPER
    const int base = tid * per;
    int cur = 0;
    for (int s = ss; s >= 2 * NBX_LEAF; s >>= 1) {
        const int nsub = ss / s;
        // split dimension of each node range: largest extent (list ends)
        for (int t = tid; t < nsub; t += KD_LT) {
            const int u = t * s;
            int best = 0;
            if (u < cnt) {
                const int e = min(u + s, cnt) - 1;
                float bext = -1.0f;
                for (int d = 0; d < ND; ++d) {
                    const float x = cf[d][lst[cur][d][e]] - cf[d][lst[cur][d][u]];
                    if (x > bext) {
                        bext = x;
                        best = d;
                    }
                }
            }
            dsel[t] = (uint8_t)best;
        }
        __syncthreads();
        for (int j = tid; j < cnt; j += KD_LT) {
            const int sub = j / s;
            const int d = dsel[sub];
            fl[lst[cur][d][j]] = j < sub * s + s / 2 ? 1 : 0;
        }
        __syncthreads();
        for (int d = 0; d < ND; ++d) {
            int fv[KD_LPER];
            int tsum = 0;
#pragma unroll
            for (int q = 0; q < KD_LPER; ++q) {
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
            for (int q = 0; q < KD_LPER; ++q) {
                const int j = base + q;
                if (q < per && j < ss) pref[j] = (uint16_t)run;
                run += fv[q];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < KD_LPER; ++q) {
                const int j = base + q;
                if (q < per && j < cnt) {
                    const int u = (j / s) * s;
                    const int m = min(u + s / 2, cnt);
                    const int nl = (int)pref[j] - (int)pref[u];
                    const int dst = fv[q] ? u + nl : m + (j - u - nl);
                    lst[cur ^ 1][d][dst] = lst[cur][d][j];
                }
            }
            __syncthreads();
        }
        cur ^= 1;
    }
    for (int j = tid; j < cnt; j += KD_LT) order[a + j] = S.id[a + lst[cur][0][j]];
}

static inline int nblk(int64_t n, int b = KD_T) { return (int)((n + b - 1) / b); }

hipError_t nbx_kd_order(const double *const *c, int nd, int64_t n, int32_t nleafp, int32_t *order, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (nd != 3 && nd != 6) return hipErrorInvalidValue;
    int ltot = 0;  // log2 of the padded slot span
    while (((int64_t)1 << ltot) < (int64_t)nleafp * NBX_LEAF) ++ltot;
    int lloc = 0;
    while ((1 << lloc) < KD_LOCAL) ++lloc;
    const int lblk = ltot < lloc ? ltot : lloc;
    const bool global = ltot > lblk;
    // the most nodes any global level has (its node ranges are 2 * KD_LOCAL)
    const int64_t nsegmax = global ? ((int64_t)1 << (ltot - lblk - 1)) : 1;

    KdState Sh;
    memset(&Sh, 0, sizeof(Sh));
    KdCoords C;
    memset(&C, 0, sizeof(C));
    for (int d = 0; d < nd; ++d) C.c[d] = c[d];
    hipError_t e = hipSuccess;
    uint32_t *ext = nullptr, *pre = nullptr, *hist = nullptr;
    int32_t *need = nullptr;
    uint8_t *dsel = nullptr;
    uint64_t *blk = nullptr, *bpre = nullptr;
#define KDCK(x)                          \
    do {                                 \
        e = (x);                         \
        if (e != hipSuccess) goto done;  \
    } while (0)
    KDCK(hipMallocAsync((void **)&Sh.id, sizeof(int32_t) * n, st));
    for (int d = 0; d < nd; ++d) KDCK(hipMallocAsync((void **)&Sh.key[d], sizeof(uint32_t) * n, st));
    hipLaunchKernelGGL(k_kd_init, dim3(nblk(n)), dim3(KD_T), 0, st, C, nd, n, Sh);
    if (global) {
        const int nb = nblk(n, KD_LOCAL);
        KDCK(hipMallocAsync((void **)&Sh.id2, sizeof(int32_t) * n, st));
        for (int d = 0; d < nd; ++d) KDCK(hipMallocAsync((void **)&Sh.key2[d], sizeof(uint32_t) * n, st));
        KDCK(hipMallocAsync((void **)&ext, sizeof(uint32_t) * 2 * nd * nsegmax, st));
        KDCK(hipMallocAsync((void **)&pre, sizeof(uint32_t) * nsegmax, st));
        KDCK(hipMallocAsync((void **)&need, sizeof(int32_t) * nsegmax, st));
        KDCK(hipMallocAsync((void **)&dsel, (size_t)nsegmax, st));
        KDCK(hipMallocAsync((void **)&hist, sizeof(uint32_t) * 256 * nsegmax, st));
        hipLaunchKernelGGL(k_kd_zero, dim3(nblk(256 * nsegmax)), dim3(KD_T), 0, st, hist, (int64_t)256 * nsegmax);
        KDCK(hipMallocAsync((void **)&blk, sizeof(uint64_t) * nb, st));
        KDCK(hipMallocAsync((void **)&bpre, sizeof(uint64_t) * nb, st));
        // root extents; deeper levels get theirs from the previous move
        hipLaunchKernelGGL(k_kd_ext_init, dim3(nblk(nd)), dim3(KD_T), 0, st, ext, (int64_t)nd);
        if (nd == 3)
            hipLaunchKernelGGL(k_kd_ext<3>, dim3(nb), dim3(KD_T), 0, st, Sh, n, ltot, ext);
        else
            hipLaunchKernelGGL(k_kd_ext<6>, dim3(nb), dim3(KD_T), 0, st, Sh, n, ltot, ext);
        for (int lseg = ltot; lseg > lblk; --lseg) {
            const int64_t nseg = (int64_t)1 << (ltot - lseg);
            hipLaunchKernelGGL(k_kd_dims, dim3(nblk(nseg)), dim3(KD_T), 0, st, ext, nd, n, lseg, nseg, dsel, pre,
                               need);
            const bool next = lseg - 1 > lblk;
            if (next)
                hipLaunchKernelGGL(k_kd_ext_init, dim3(nblk(2 * nd * nseg)), dim3(KD_T), 0, st, ext,
                                   (int64_t)2 * nd * nseg);
            for (int sh = 24; sh >= 0; sh -= 8) {
                // the select step leaves every bin it read at zero
                hipLaunchKernelGGL(k_kd_hist, dim3(nb), dim3(KD_T), 0, st, Sh, n, lseg, sh, dsel, pre, hist);
                hipLaunchKernelGGL(k_kd_select, dim3(nblk(nseg, KD_T / 64)), dim3(KD_T), 0, st, n, lseg, nseg, sh,
                                   hist, pre, need);
            }
            hipLaunchKernelGGL(k_kd_count, dim3(nb), dim3(KD_T), 0, st, Sh, n, lseg, dsel, pre, blk);
            hipLaunchKernelGGL(k_kd_bscan, dim3(1), dim3(KD_ST), 0, st, blk, (int64_t)nb, bpre);
            uint32_t *enext = next ? ext : nullptr;
            if (nd == 3)
                hipLaunchKernelGGL(k_kd_move<3>, dim3(nb), dim3(KD_T), 0, st, Sh, n, lseg, dsel, pre, need, bpre,
                                   enext);
            else
                hipLaunchKernelGGL(k_kd_move<6>, dim3(nb), dim3(KD_T), 0, st, Sh, n, lseg, dsel, pre, need, bpre,
                                   enext);
            std::swap(Sh.id, Sh.id2);
            for (int d = 0; d < nd; ++d) std::swap(Sh.key[d], Sh.key2[d]);
        }
    }
    {
        const int64_t nb = (n + ((int64_t)1 << lblk) - 1) >> lblk;
        if (nd == 3)
            hipLaunchKernelGGL(k_kd_local<3>, dim3((unsigned)nb), dim3(KD_LT), 0, st, Sh, n, lblk, order);
        else
            hipLaunchKernelGGL(k_kd_local<6>, dim3((unsigned)nb), dim3(KD_LT), 0, st, Sh, n, lblk, order);
    }
    e = hipGetLastError();
done:
    if (Sh.id) hipFreeAsync(Sh.id, st);
    if (Sh.id2) hipFreeAsync(Sh.id2, st);
    for (int d = 0; d < 6; ++d) {
        if (Sh.key[d]) hipFreeAsync(Sh.key[d], st);
        if (Sh.key2[d]) hipFreeAsync(Sh.key2[d], st);
    }
    if (ext) hipFreeAsync(ext, st);
    if (pre) hipFreeAsync(pre, st);
    if (need) hipFreeAsync(need, st);
    if (dsel) hipFreeAsync(dsel, st);
    if (hist) hipFreeAsync(hist, st);
    if (blk) hipFreeAsync(blk, st);
    if (bpre) hipFreeAsync(bpre, st);
#undef KDCK
    return e;
}
