// nbx_build.hip -- tree construction on gfx950.
//
// Replaces KDTree::BuildNodes (KDTree.cxx:970-1055): instead of a recursive
// median split that swaps 88-byte particles in place, the split order is
// computed on index lists (nbx_kdbuild.hip), particles are gathered into SoA
// arrays in that order, cut into fixed buckets of NBX_LEAF slots and boxed
// bottom-up.  All passes here are streaming and coalesced (HBM-bound).
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "nbx_internal.h"
#include "nbx_kernels.h"

// ---------------------------------------------------------------- bounding box
__global__ __launch_bounds__(256) void k_bbox(const double *__restrict__ c, int64_t n, double *__restrict__ part) {
    __shared__ double smn[256], smx[256];
    double mn = 1.0 / 0.0, mx = -1.0 / 0.0;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double v = c[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + s]);
            smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smn[0];
        part[2 * blockIdx.x + 1] = smx[0];
    }
}

// ---------------------------------------------------------------- gather
__global__ __launch_bounds__(256) void k_gather(GatherArgs a) {
    int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= a.npad) return;
    if (s >= a.n) {  // padding up to a whole leaf
        for (int d = 0; d < a.nd; ++d) a.out[d][s] = NBX_PAD_COORD;
        a.mass_out[s] = 0.0;
        a.perm[s] = -1;
        return;
    }
    int32_t id = a.order[s];
    for (int d = 0; d < a.nd; ++d) a.out[d][s] = a.in[d][id];
    a.mass_out[s] = a.mass_in ? a.mass_in[id] : a.mass_default;
    a.perm[s] = id;
    if (a.inv) a.inv[id] = (int32_t)s;
}

// ---------------------------------------------------------------- boxes
// One thread per (padded) leaf: tight min/max over its NBX_LEAF slots;
// padding leaves get the empty box (+inf, -inf).
__global__ __launch_bounds__(256) void k_leafbox(LeafBoxArgs a) {
    int32_t L = blockIdx.x * 256 + threadIdx.x;
    if (L >= a.nleafp) return;
    double *b = a.box + (size_t)(a.nleafp + L) * 2 * a.nd;
    if (L >= a.nleaf) {
        for (int d = 0; d < a.nd; ++d) { b[d] = 1.0 / 0.0; b[a.nd + d] = -1.0 / 0.0; }
        return;
    }
    int64_t s0 = (int64_t)L * NBX_LEAF, s1 = s0 + NBX_LEAF;
    if (s1 > a.n) s1 = a.n;
    for (int d = 0; d < a.nd; ++d) {
        const double *c = a.c[d];
        double mn = c[s0], mx = mn;
        for (int64_t s = s0 + 1; s < s1; ++s) {
            double v = c[s];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
        b[d] = mn;
        b[a.nd + d] = mx;
        if (a.org) a.org[(size_t)L * a.nd + d] = 0.5 * (mn + mx);
    }
    if (a.ext) {
        // half-diagonal from the centre, rounded up
        double e2 = 0.0;
        for (int d = 0; d < a.nd; ++d) {
            double o = 0.5 * (b[d] + b[a.nd + d]);
            double w = fmax(o - b[d], b[a.nd + d] - o);
            e2 += w * w;
        }
        a.ext[L] = sqrt(e2) * (1.0 + 1e-12);
    }
}

// Internal nodes [first, first+count) of one depth: union of the two children.
__global__ __launch_bounds__(256) void k_levelbox(double *__restrict__ box, float *__restrict__ plane, int32_t first,
                                                  int32_t count, int nd) {
    int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    int32_t nn = first + i;
    const double *a = box + (size_t)(2 * nn) * 2 * nd;
    const double *b = a + 2 * nd;
    double *o = box + (size_t)nn * 2 * nd;
    for (int d = 0; d < nd; ++d) {
        o[d] = fmin(a[d], b[d]);
        o[nd + d] = fmax(a[nd + d], b[nd + d]);
    }
    // Child order plane: the right child is nearer to a point q than the left
    // when n.q > off, n = c_right - c_left (box centres), off = n.(c_l + c_r)/2.
    // An empty right child (padding) gives n = 0, off = 1: left first.
    const int plw = nd == 6 ? 8 : 4;
    float *pl = plane + (size_t)nn * plw;
    const bool empty = !(a[0] <= a[nd]) || !(b[0] <= b[nd]);
    double off = 0.0;
    for (int d = 0; d < nd; ++d) {
        const double ca = 0.5 * (a[d] + a[nd + d]), cb = 0.5 * (b[d] + b[nd + d]);
        const double nv = empty ? 0.0 : cb - ca;
        pl[d] = (float)nv;
        off += empty ? 0.0 : nv * 0.5 * (ca + cb);
    }
    pl[nd] = empty ? 1.0f : (float)off;
    for (int d = nd + 1; d < plw; ++d) pl[d] = 0.0f;
}

// Mixed form: FP32 offsets of each particle from its leaf's FP64 origin.
__global__ __launch_bounds__(256) void k_mixed_offsets(MixedArgs a) {
    int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= a.npad) return;
    int64_t L = s / NBX_LEAF;
    for (int d = 0; d < a.nd; ++d)
        a.xf[d][s] = s < a.n ? (float)(a.c[d][s] - a.org[L * a.nd + d]) : NBX_PAD_COORD_F;
}

// ---------------------------------------------------------------- sort
// sort of 64-bit keys with 32-bit values (FOF group ranking)
hipError_t nbx_sort_pairs_u64(void *temp, size_t &temp_bytes, uint64_t *kin, uint64_t *kout, int32_t *vin,
                              int32_t *vout, int64_t n, hipStream_t st) {
    return rocprim::radix_sort_pairs(temp, temp_bytes, kin, kout, vin, vout, (unsigned int)n, 0, 64, st, false);
}
