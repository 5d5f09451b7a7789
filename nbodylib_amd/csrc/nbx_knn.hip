// nbx_knn.hip -- exact FP64 k-nearest-neighbour search and SPH densities.
//
// Replaces KDTree::FindNearest* (KDFindNearest.cxx:247-361, :444-564),
// SplitNode/LeafNode::FindNearestPos/Phase (KDSplitNode.cxx:15-95,
// KDLeafNode.cxx:15-57, :119-154), the periodic image passes
// (KDSplitNode.cxx:1119-1179), CalcDensity / CalcDensityParticle /
// CalcVelDensity (KDCalcSmoothQuantities.cxx:203-389, :768-844).
//
// Mapping to CDNA4: one wavefront owns 64 Morton-consecutive queries (two
// sibling leaves).  The wave walks the implicit tree once for all 64 lanes
// (a node is opened when any lane's bound reaches it), so node boxes and
// leaf particles arrive through the scalar cache and are shared by the
// whole wave.  Each lane keeps its own max-heap of (d2, slot) in LDS -- the
// reference's PriorityQueue (PriorityQueue.h:47-84) -- laid out
// [position][lane] so same-position accesses are conflict-free.  A leaf is
// consumed in two phases: (1) all lanes test every candidate against their
// current k-th distance, building a 32-bit accept mask; (2) lanes insert
// their accepted candidates one at a time, so the wave pays for as many heap
// insertions as its busiest lane, not as the sum over lanes.
//
// Exactness: distances are summed in the reference's order with no FMA
// (this file is built with -ffp-contract=off); pruning uses box_lb, which is
// never larger than a computed particle distance.  The neighbour set is the k
// smallest (d2, ID) pairs: ties at the k-th distance go to the smaller ID
// (the reference's choice there depends on its traversal order).
#include "nbx_kernels.h"

__device__ __forceinline__ int32_t sid(const int32_t *perm, int32_t s) { return s < 0 ? 0x7fffffff : perm[s]; }

// (da, ID(sa)) > (db, ID(sb))
__device__ __forceinline__ bool key_gt(double da, int32_t sa, double db, int32_t sb, const int32_t *perm) {
    if (da > db) return true;
    if (da < db) return false;
    if (sa == sb) return false;
    return sid(perm, sa) > sid(perm, sb);
}

// Place (dv, sv) at heap position p of a max-heap of size kh and sift down.
__device__ __forceinline__ void heap_sift(double *hd, int32_t *hs, int lane, int kh, int p, double dv, int32_t sv,
                                          const int32_t *perm) {
    for (;;) {
        int l = 2 * p;
        if (l > kh) break;
        double dc = hd[(l - 1) * NBX_WAVE + lane];
        int32_t sc = hs[(l - 1) * NBX_WAVE + lane];
        int c = l;
        if (l + 1 <= kh) {
            double dr = hd[l * NBX_WAVE + lane];
            int32_t sr = hs[l * NBX_WAVE + lane];
            if (key_gt(dr, sr, dc, sc, perm)) { c = l + 1; dc = dr; sc = sr; }
        }
        if (!key_gt(dc, sc, dv, sv, perm)) break;
        hd[(p - 1) * NBX_WAVE + lane] = dc;
        hs[(p - 1) * NBX_WAVE + lane] = sc;
        p = c;
    }
    hd[(p - 1) * NBX_WAVE + lane] = dv;
    hs[(p - 1) * NBX_WAVE + lane] = sv;
}

// Heap sort in place: positions 1..kh ascending afterwards.
__device__ __forceinline__ void heap_sort(double *hd, int32_t *hs, int lane, int kh, const int32_t *perm) {
    for (int m = kh; m > 1; --m) {
        double dt = hd[lane];
        int32_t st = hs[lane];
        double dm = hd[(m - 1) * NBX_WAVE + lane];
        int32_t sm = hs[(m - 1) * NBX_WAVE + lane];
        hd[(m - 1) * NBX_WAVE + lane] = dt;
        hs[(m - 1) * NBX_WAVE + lane] = st;
        heap_sift(hd, hs, lane, m - 1, 1, dm, sm, perm);
    }
}

// Consume one leaf for every active lane (two-phase insertion, see header).
template <int ND>
__device__ __forceinline__ void knn_leaf(const DevTree &T, int leaf, const double *q, bool act, int32_t self,
                                         bool exz, double &kth, int32_t &kslot, double *hd, int32_t *hs, int lane,
                                         int kh) {
    const int base = leaf * NBX_LEAF;
    int cnt = T.n - base;
    cnt = cnt < NBX_LEAF ? cnt : NBX_LEAF;
    uint32_t mask = 0;
    for (int j = 0; j < cnt; ++j) {
        double d2 = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double c = ((cdptr)T.x[d])[base + j];
            double t = q[d] - c;
            d2 += t * t;
        }
        bool ok = act && (d2 <= kth) && (!exz || d2 > 0.0) && (base + j != self);
        mask |= ok ? (1u << j) : 0u;
    }
    while (ballot(mask != 0)) {
        if (mask) {
            int j = __builtin_ctz(mask);
            mask &= mask - 1;
            int32_t cs = base + j;
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                double t = q[d] - T.x[d][cs];
                d2 += t * t;
            }
            if (key_gt(kth, kslot, d2, cs, T.perm)) {
                heap_sift(hd, hs, lane, kh, 1, d2, cs, T.perm);
                kth = hd[lane];
                kslot = hs[lane];
            }
        }
    }
}

// Uniform stackless depth-first walk of the implicit tree, near child first
// (ordered by the uniform point qc).  Nodes are opened when any active lane
// has box_lb <= its current k-th distance.
template <int ND>
__device__ __forceinline__ void knn_walk(const DevTree &T, const double *q, const double *qc, bool act, int32_t self,
                                         bool exz, double &kth, int32_t &kslot, double *hd, int32_t *hs, int lane,
                                         int kh) {
    const int top = T.lv.nlev - 1;
    int level = top, idx = 0;
    uint64_t firstright = 0;
    bool down = true;
    for (;;) {
        if (down) {
            cdptr bx = (cdptr)node_box(T, level, idx, ND);
            bool want = act && box_lb<ND>(q, bx) <= kth;
            if (ballot(want)) {
                if (level == 0) {
                    knn_leaf<ND>(T, idx, q, act, self, exz, kth, kslot, hd, hs, lane, kh);
                    down = false;
                } else {
                    int c0 = 2 * idx, c1 = c0 + 1;
                    int first = c0;
                    if (c1 < T.lv.cnt[level - 1]) {
                        double d0 = box_lb<ND>(qc, (cdptr)node_box(T, level - 1, c0, ND));
                        double d1 = box_lb<ND>(qc, (cdptr)node_box(T, level - 1, c1, ND));
                        if (d1 < d0) first = c1;
                    }
                    if (first & 1) firstright |= (1ULL << (level - 1));
                    else firstright &= ~(1ULL << (level - 1));
                    level -= 1;
                    idx = first;
                    continue;
                }
            } else {
                down = false;
            }
        }
        if (level == top) break;
        bool wasfirst = (uint64_t)(idx & 1) == ((firstright >> level) & 1ULL);
        int sib = idx ^ 1;
        if (wasfirst && sib < T.lv.cnt[level]) {
            idx = sib;
            down = true;
        } else {
            idx >>= 1;
            level += 1;
        }
    }
}

// Image point and test for periodic pass p (1..7), DistFunc.h:326-355 and
// KDSplitNode.cxx:1119-1148: 1-D images need sqrt(Top) > s, the 2-D and 3-D
// ones compare Top (a squared distance) with s (a distance), as the reference does.
__device__ __forceinline__ bool image_pass(int pass, const double *x0, const double *per, double kth, bool uncond,
                                           double *xp) {
    xp[0] = x0[0]; xp[1] = x0[1]; xp[2] = x0[2];
    if (pass <= 3) {
        int k = pass - 1;
        double sval;
        if (x0[k] < per[k] / 2.0) { xp[k] = x0[k] + per[k]; sval = x0[k]; }
        else { xp[k] = x0[k] - per[k]; sval = -xp[k]; }
        return uncond || sqrt(kth) > sval;
    }
    int k1, k2;
    double sval = 0.0;
    if (pass < 7) {
        k1 = (pass == 6) ? 1 : 0;
        k2 = (pass == 4) ? 1 : 2;
        if (x0[k1] < per[k1] / 2.0) { xp[k1] = x0[k1] + per[k1]; sval += x0[k1] * x0[k1]; }
        else { xp[k1] = x0[k1] - per[k1]; sval += xp[k1] * xp[k1]; }
        if (x0[k2] < per[k2] / 2.0) { xp[k2] = x0[k2] + per[k2]; sval += x0[k2] * x0[k2]; }
        else { xp[k2] = x0[k2] - per[k2]; sval += xp[k2] * xp[k2]; }
    } else {
        for (int k = 0; k < 3; ++k) {
            if (x0[k] < per[k] / 2.0) { xp[k] = x0[k] + per[k]; sval += x0[k] * x0[k]; }
            else { xp[k] = x0[k] - per[k]; sval += xp[k] * xp[k]; }
        }
    }
    sval = sqrt(sval);
    return uncond || kth > sval;
}

template <int ND, int KMAX>
__global__ __launch_bounds__(64) void k_knn(DevTree T, KnnArgs A) {
    __shared__ double hd[KMAX * NBX_WAVE];
    __shared__ int32_t hs[KMAX * NBX_WAVE];
    const int lane = threadIdx.x;
    const int g = blockIdx.x;
    const int kh = A.kh;
    const int nq = (A.qsrc == NBX_Q_POINTS) ? A.m : T.n;
    const int qi = g * NBX_WAVE + lane;
    const bool valid = qi < nq;
    double q0[ND];
    int32_t self = -1;
    if (A.qsrc == NBX_Q_POINTS) {
#pragma unroll
        for (int d = 0; d < ND; ++d) q0[d] = valid ? A.q[d][qi] : 0.0;
    } else {
#pragma unroll
        for (int d = 0; d < ND; ++d) q0[d] = valid ? T.x[d][qi] : 0.0;
        if (A.qsrc == NBX_Q_SLOTS) self = qi;
    }
    const bool exz = (A.qsrc == NBX_Q_SLOTS) && A.excl_zero;
    for (int p = 0; p < kh; ++p) {
        hd[p * NBX_WAVE + lane] = NBX_MAXVALUE;
        hs[p * NBX_WAVE + lane] = -1;
    }
    double kth = NBX_MAXVALUE;
    int32_t kslot = -1;
    const int npass = (A.periodic && A.qsrc != NBX_Q_SLOTS) ? 8 : 1;
    for (int pass = 0; pass < npass; ++pass) {
        double q[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) q[d] = q0[d];
        bool act = valid;
        if (pass > 0) {
            double xp[3];
            act = valid && image_pass(pass, q0, A.period, kth, ND == 6, xp);
            q[0] = xp[0]; q[1] = xp[1]; q[2] = xp[2];
        }
        if (!ballot(act)) continue;
        double qc[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) qc[d] = readfirst_d(q[d]);
        knn_walk<ND>(T, q, qc, act, self, exz, kth, kslot, hd, hs, lane, kh);
    }

    // ------------------------------------------------------------ densities
    if (A.dens != NBX_DO_NONE && valid) {
        bool bad = false;
        if (A.dens == NBX_DO_VEL) {
            // CalcVelDensity: |dv| of the Nsearch physical neighbours, the
            // Nsmooth smallest of them, h_v = max/2, sum W(|dv|/h_v) h_v^-3.
            const int ns = A.nsmooth;
            for (int p = 1; p <= kh; ++p) {
                int32_t s = hs[(p - 1) * NBX_WAVE + lane];
                if (s < 0) { bad = true; break; }
                double vx = T.x[3][qi] - T.x[3][s], vy = T.x[4][qi] - T.x[4][s], vz = T.x[5][qi] - T.x[5][s];
                hd[(p - 1) * NBX_WAVE + lane] = sqrt(vx * vx + vy * vy + vz * vz);
            }
            if (!bad) {
                for (int p = ns / 2; p >= 1; --p)
                    heap_sift(hd, hs, lane, ns, p, hd[(p - 1) * NBX_WAVE + lane], hs[(p - 1) * NBX_WAVE + lane], T.perm);
                for (int p = ns + 1; p <= kh; ++p) {
                    double dv = hd[(p - 1) * NBX_WAVE + lane];
                    int32_t sv = hs[(p - 1) * NBX_WAVE + lane];
                    if (key_gt(hd[lane], hs[lane], dv, sv, T.perm)) heap_sift(hd, hs, lane, ns, 1, dv, sv, T.perm);
                }
                double hv = 0.5 * hd[lane];
                double norm = 1.0 / pow(hv, 3.0);
                double den = 0.0;
                for (int p = 1; p <= ns; ++p) den = den + wsm_at(hd[(p - 1) * NBX_WAVE + lane], hv, T.kern, T.kernres) * norm;
                A.rho[qi] = den;
                if (A.hout) A.hout[qi] = hv;
            }
        } else {
            double hi = 0.5 * sqrt(kth);
            double norm = 1.0 / pow(hi, (double)ND);
            double mi = T.mass[qi];
            double acc = 0.0;
            for (int p = 1; p <= kh; ++p) {
                int32_t s = hs[(p - 1) * NBX_WAVE + lane];
                if (s < 0) { bad = true; continue; }
                double rij = sqrt(hd[(p - 1) * NBX_WAVE + lane]);
                double W = wsm_at(rij, hi, T.kern, T.kernres);
                if (A.dens == NBX_DO_SYM) {
                    double Wij = 0.5 * W * norm;
                    acc = acc + Wij * T.mass[s];
                    unsafeAtomicAdd(A.rho + s, Wij * mi);
                } else {
                    double Wij = W * norm;
                    acc += Wij * T.mass[s];
                }
            }
            if (A.dens == NBX_DO_SYM) unsafeAtomicAdd(A.rho + qi, acc);
            else A.rho[qi] = acc;
            if (A.hout) A.hout[qi] = hi;
        }
        if (bad) atomicOr(A.err, 1);
    }

    // ------------------------------------------------------------ outputs
    if (A.nn == nullptr && A.d2 == nullptr) return;
    if (valid) heap_sort(hd, hs, lane, kh, T.perm);
    __syncthreads();
    const int kout = A.kout;
    const int off = (A.drop == NBX_DROP_SMALLEST) ? 1 : 0;
    const int total = NBX_WAVE * kout;
    for (int e = lane; e < total; e += NBX_WAVE) {
        int r = e / kout, t = e - r * kout;
        int qr = g * NBX_WAVE + r;
        if (qr >= nq) break;
        int64_t row;
        if (A.qsrc == NBX_Q_POINTS || A.layout == 1) row = qr;
        else row = T.perm[qr];
        int32_t s = hs[(t + off) * NBX_WAVE + r];
        if (A.nn) {
            int32_t v = s;
            if (s >= 0 && (A.qsrc == NBX_Q_POINTS || A.layout == 0)) v = T.perm[s];
            A.nn[row * kout + t] = v;
        }
        if (A.d2) A.d2[row * kout + t] = hd[(t + off) * NBX_WAVE + r];
    }
}

hipError_t nbx_launch_knn(const DevTree &T, const KnnArgs &A, int nqueries, hipStream_t st) {
    int blocks = (nqueries + NBX_WAVE - 1) / NBX_WAVE;
    if (blocks == 0) return hipSuccess;
    ProfScope ps("k_knn", st);
    if (T.nd == 6) {
        if (A.kh <= 65) hipLaunchKernelGGL((k_knn<6, 65>), dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
        else return hipErrorInvalidValue;
    } else {
        if (A.kh <= 33) hipLaunchKernelGGL((k_knn<3, 33>), dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
        else if (A.kh <= 65) hipLaunchKernelGGL((k_knn<3, 65>), dim3(blocks), dim3(NBX_WAVE), 0, st, T, A);
        else return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
