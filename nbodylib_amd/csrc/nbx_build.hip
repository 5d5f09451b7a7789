// nbx_build.hip -- tree construction on gfx950.
//
// Replaces KDTree::BuildNodes (KDTree.cxx:970-1055): instead of a recursive
// median split that swaps 88-byte particles in place, the split order is
// computed by physical partitions (nbx_kdbuild.hip), particles are gathered
// into SoA arrays in that order, cut into fixed buckets of NBX_LEAF slots and
// boxed bottom-up.  All passes here are streaming and coalesced (HBM-bound)
// apart from the gather's read of the input by ID.
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
// One thread per (padded) slot, gathering its particle into tree order.  A
// half-wave is exactly one leaf of NBX_LEAF slots, so the leaf's tight
// FP64 box is a 32-lane min/max reduction over the values just gathered
// (padding slots count as +inf / -inf), and for the precision form the
// leaf origin (box centre), half-diagonal and FP32 offsets follow from it
// without reading the coordinates again.
__global__ __launch_bounds__(256) void k_gather_leaf(GatherArgs a) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool in = s < a.npad;
    const bool real = s < a.n;
    const int lane = threadIdx.x & 63;
    double v[6];
    if (real) {
        const int32_t id = a.order[s];
        for (int d = 0; d < a.ncoord; ++d) v[d] = a.in[d][id];
        a.mass_out[s] = a.mass_in ? a.mass_in[id] : a.mass_default;
        a.perm[s] = id;
        if (a.inv) a.inv[id] = (int32_t)s;
    } else if (in) {  // padding up to a whole leaf
        for (int d = 0; d < a.ncoord; ++d) v[d] = NBX_PAD_COORD;
        a.mass_out[s] = 0.0;
        a.perm[s] = -1;
    }
    if (in)
        for (int d = 0; d < a.ncoord; ++d) a.out[d][s] = v[d];
    const int64_t L = s / NBX_LEAF;
    double mn[6], mx[6];
    for (int d = 0; d < a.nd; ++d) {
        mn[d] = real ? v[d] : 1.0 / 0.0;
        mx[d] = real ? v[d] : -1.0 / 0.0;
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) {
            mn[d] = fmin(mn[d], __shfl_xor(mn[d], off));
            mx[d] = fmax(mx[d], __shfl_xor(mx[d], off));
        }
    }
    if (!in || L >= a.nleaf) return;
    if ((lane & 31) == 0) {
        double *b = a.box + (size_t)(a.nleafp + L) * 2 * a.nd;
        double e2 = 0.0;
        for (int d = 0; d < a.nd; ++d) {
            b[d] = mn[d];
            b[a.nd + d] = mx[d];
            if (a.org) a.org[(size_t)L * a.nd + d] = 0.5 * (mn[d] + mx[d]);
            // half-diagonal from the centre, rounded up
            const double o = 0.5 * (mn[d] + mx[d]);
            const double w = fmax(o - mn[d], mx[d] - o);
            e2 += w * w;
        }
        if (a.ext) a.ext[L] = sqrt(e2) * (1.0 + 1e-12);
    }
    if (a.org) {
        float f[6];
        for (int d = 0; d < a.nd; ++d) {
            const double o = 0.5 * (mn[d] + mx[d]);  // the same in every lane of the leaf
            f[d] = real ? (float)(v[d] - o) : NBX_PAD_COORD_F;
        }
        if (a.nd == 3) {
            const float f2 = __builtin_fmaf(f[2], f[2], __builtin_fmaf(f[1], f[1], f[0] * f[0]));
            a.xq[s] = make_float4(f[0], f[1], f[2], real ? f2 : 1.0f / 0.0f);
        } else {
            for (int d = 0; d < a.nd; ++d) a.xf[d][s] = f[d];
        }
    }
}

// Padding leaves [nleaf, nleafp) hold the empty box (+inf, -inf).
__global__ __launch_bounds__(256) void k_leafbox_pad(double *box, int32_t nleaf, int32_t nleafp, int nd) {
    const int32_t L = nleaf + (int32_t)(blockIdx.x * 256 + threadIdx.x);
    if (L >= nleafp) return;
    double *b = box + (size_t)(nleafp + L) * 2 * nd;
    for (int d = 0; d < nd; ++d) {
        b[d] = 1.0 / 0.0;
        b[nd + d] = -1.0 / 0.0;
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

// the next FP32 value above f (f finite or -inf; +inf and NaN unchanged)
__device__ __forceinline__ float f32_next_up(float f) {
    if (!(f < 1.0f / 0.0f)) return f;
    if (f == 0.0f) return __uint_as_float(1u);
    const uint32_t b = __float_as_uint(f);
    return __uint_as_float(f > 0.0f ? b + 1u : b - 1u);
}

// FP32 copy of the node boxes rounded outward (lower corners down, upper
// corners up), for the precision form's node tests; count = 2 nd per node.
__global__ __launch_bounds__(256) void k_boxf(const double *__restrict__ box, float *__restrict__ boxf, int64_t count,
                                              int nd) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const double v = box[i];
    const bool upper = (int)(i % (2 * nd)) >= nd;
    float f = (float)v;
    if (upper && (double)f < v) f = f32_next_up(f);
    if (!upper && (double)f > v) f = -f32_next_up(-f);
    boxf[i] = f;
}

// ---------------------------------------------------------------- sort
// sort of 64-bit keys with 32-bit values (FOF group ranking)
hipError_t nbx_sort_pairs_u64(void *temp, size_t &temp_bytes, uint64_t *kin, uint64_t *kout, int32_t *vin,
                              int32_t *vout, int64_t n, hipStream_t st) {
    return rocprim::radix_sort_pairs(temp, temp_bytes, kin, kout, vin, vout, (unsigned int)n, 0, 64, st, false);
}
