// nbx_adjoint.hip -- differentiation form: SPH density over a frozen
// neighbour list and its hand-written adjoint.
//
// No reference counterpart: NBodylib allocates derKernel but never fills it
// (KDTree.cxx:1150, :1345).  The forward is CalcDensity's symmetric formula
// (KDCalcSmoothQuantities.cxx:260-286) with the table replaced by the
// analytic cubic spline WSPH (SmoothingKernels.h:32-37), 3-D norm 1/pi
// (KDTree.cxx:1158-1164), so the derivative is continuous:
//   w_ij = (sigma/2) W(r_ij/h_i) h_i^-3,  rho_i += w_ij m_j,  rho_j += w_ij m_i,
//   h_i  = |x_i - x_K(i)| / 2  (K(i) = last column of the list).
// Adjoint, with a_ij = g_i m_j + g_j m_i and q = r/h:
//   dL/dm_j += w_ij g_i,  dL/dm_i += w_ij g_j,
//   dL/dx_i += a_ij (sigma/2) W'(q) h^-4 (x_i-x_j)/r   (and minus on x_j),
//   dL/dh_i  = sum_j a_ij (-(sigma/2) h^-4 (q W'(q) + 3 W(q))),
//   dL/dx_i += dL/dh_i (x_i-x_K)/(2 r_iK)               (and minus on x_K).
// One thread per particle i; contributions to other particles use FP64
// atomics (LDS ones for particles of the same block).
#include "nbx_kernels.h"

// Cubic spline without branches: with t = max(2 - q, 0), s = max(1 - q, 0),
// W = t^3/4 - s^3 (= 1 - 1.5q^2 + 0.75q^3 on [0,1], (2-q)^3/4 on [1,2], 0
// beyond) and W' = -0.75 t^2 + 3 s^2.
__device__ __forceinline__ double w_cubic(double q) {
    const double t = fmax(2.0 - q, 0.0), s = fmax(1.0 - q, 0.0);
    return 0.25 * t * t * t - s * s * s;
}
__device__ __forceinline__ double dw_cubic(double q) {
    const double t = fmax(2.0 - q, 0.0), s = fmax(1.0 - q, 0.0);
    return -0.75 * t * t + 3.0 * s * s;
}

#define SIGMA3 0.31830988618379067154  // 1/pi

// Contributions to particles of the same 1024-particle block (most of them
// when the arrays are in tree order) are summed in LDS and flushed with one
// atomic per particle and component; the rest go straight to HBM.
#define ADJ_B 1024
#define ADJ_G 8   // neighbours whose gathers are issued together (forward)
#define ADJ_GB 4  // (adjoint: more live values per neighbour)

// Rows i in [r0, r0 + nr) of the n particles: row i of the list is
// nbr[(i - r0) k ...]; positions X, Y, Z and the other arrays have all n.
__global__ __launch_bounds__(ADJ_B) void k_sph_fwd(const double *__restrict__ X, const double *__restrict__ Y,
                                                   const double *__restrict__ Z, const double *__restrict__ mass,
                                                   const int32_t *__restrict__ nbr, int64_t n, int64_t r0, int64_t nr,
                                                   int k, double *rho, double *hout) {
    __shared__ double lrho[ADJ_B];
    const int t = threadIdx.x;
    const int64_t i0 = r0 + (int64_t)blockIdx.x * ADJ_B;
    const int64_t i = i0 + t;
    lrho[t] = 0.0;
    __syncthreads();
    double acc = 0.0;
    if (i < r0 + nr) {
        const double xi = X[i], yi = Y[i], zi = Z[i], mi = mass[i];
        const int32_t *row = nbr + (i - r0) * k;
        const int64_t K = row[k - 1];
        const double dx = xi - X[K], dy = yi - Y[K], dz = zi - Z[K];
        const double h = 0.5 * sqrt(dx * dx + dy * dy + dz * dz);
        if (hout) hout[i] = h;
        const double ih = 1.0 / h;
        const double norm = 0.5 * SIGMA3 * ih * ih * ih;
        for (int j0 = 0; j0 < k; j0 += ADJ_G) {
            int64_t qv[ADJ_G];
            double rv[ADJ_G], mv[ADJ_G];
#pragma unroll
            for (int u = 0; u < ADJ_G; ++u) {
                qv[u] = j0 + u < k ? row[j0 + u] : i;
                const double ex = xi - X[qv[u]], ey = yi - Y[qv[u]], ez = zi - Z[qv[u]];
                rv[u] = sqrt(ex * ex + ey * ey + ez * ez);
                mv[u] = mass[qv[u]];
            }
#pragma unroll
            for (int u = 0; u < ADJ_G; ++u) {
                if (j0 + u >= k) break;
                const double w = norm * w_cubic(rv[u] * ih);
                acc += w * mv[u];
                const uint64_t r = (uint64_t)(qv[u] - i0);
                if (r < ADJ_B) atomicAdd(&lrho[r], w * mi);
                else unsafeAtomicAdd(rho + qv[u], w * mi);
            }
        }
    }
    __syncthreads();
    if (i < n) unsafeAtomicAdd(rho + i, acc + lrho[t]);
}

__global__ __launch_bounds__(ADJ_B) void k_sph_adj(const double *__restrict__ X, const double *__restrict__ Y,
                                                   const double *__restrict__ Z, const double *__restrict__ mass,
                                                   const int32_t *__restrict__ nbr, int64_t n, int64_t r0, int64_t nr,
                                                   int k, const double *__restrict__ g, double *DX, double *DY,
                                                   double *DZ, double *dmass) {
    __shared__ double lx[ADJ_B], ly[ADJ_B], lz[ADJ_B], lm[ADJ_B];
    const int t = threadIdx.x;
    const int64_t i0 = r0 + (int64_t)blockIdx.x * ADJ_B;
    const int64_t i = i0 + t;
    lx[t] = 0.0;
    ly[t] = 0.0;
    lz[t] = 0.0;
    lm[t] = 0.0;
    __syncthreads();
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
    if (i < r0 + nr) {
        const double xi = X[i], yi = Y[i], zi = Z[i], mi = mass[i], gi = g[i];
        const int32_t *row = nbr + (i - r0) * k;
        const int64_t K = row[k - 1];
        const double kx = xi - X[K], ky = yi - Y[K], kz = zi - Z[K];
        const double rk = sqrt(kx * kx + ky * ky + kz * kz);
        const double h = 0.5 * rk;
        const double ih = 1.0 / h;
        const double ih3 = ih * ih * ih, ih4 = ih3 * ih;
        double Hs = 0.0;
        for (int j0 = 0; j0 < k; j0 += ADJ_GB) {
            int64_t qv[ADJ_GB];
            double exv[ADJ_GB], eyv[ADJ_GB], ezv[ADJ_GB], mv[ADJ_GB], gv[ADJ_GB];
#pragma unroll
            for (int u = 0; u < ADJ_GB; ++u) {
                qv[u] = j0 + u < k ? row[j0 + u] : i;
                exv[u] = xi - X[qv[u]];
                eyv[u] = yi - Y[qv[u]];
                ezv[u] = zi - Z[qv[u]];
                mv[u] = mass[qv[u]];
                gv[u] = g[qv[u]];
            }
#pragma unroll
            for (int u = 0; u < ADJ_GB; ++u) {
                if (j0 + u >= k) break;
                const int64_t q = qv[u];
                const double ex = exv[u], ey = eyv[u], ez = ezv[u];
                const double r = sqrt(ex * ex + ey * ey + ez * ez);
                const double qq = r * ih;
                const double W = w_cubic(qq), dW = dw_cubic(qq);
                const double w = 0.5 * SIGMA3 * W * ih3;
                const double a = gi * mv[u] + gv[u] * mi;
                const uint64_t lr = (uint64_t)(q - i0);
                const bool loc = lr < ADJ_B;
                if (loc) atomicAdd(&lm[lr], w * gi);
                else unsafeAtomicAdd(dmass + q, w * gi);
                gm += w * gv[u];
                Hs += a * (-0.5 * SIGMA3 * (qq * dW + 3.0 * W) * ih4);
                if (r > 0.0) {
                    const double f = a * 0.5 * SIGMA3 * dW * ih4 / r;
                    gx += f * ex;
                    gy += f * ey;
                    gz += f * ez;
                    if (loc) {
                        atomicAdd(&lx[lr], -f * ex);
                        atomicAdd(&ly[lr], -f * ey);
                        atomicAdd(&lz[lr], -f * ez);
                    } else {
                        unsafeAtomicAdd(DX + q, -f * ex);
                        unsafeAtomicAdd(DY + q, -f * ey);
                        unsafeAtomicAdd(DZ + q, -f * ez);
                    }
                }
            }
        }
        if (rk > 0.0) {
            const double f = Hs * 0.5 / rk;
            gx += f * kx;
            gy += f * ky;
            gz += f * kz;
            const uint64_t lr = (uint64_t)(K - i0);
            if (lr < ADJ_B) {
                atomicAdd(&lx[lr], -f * kx);
                atomicAdd(&ly[lr], -f * ky);
                atomicAdd(&lz[lr], -f * kz);
            } else {
                unsafeAtomicAdd(DX + K, -f * kx);
                unsafeAtomicAdd(DY + K, -f * ky);
                unsafeAtomicAdd(DZ + K, -f * kz);
            }
        }
    }
    __syncthreads();
    if (i < n) {
        unsafeAtomicAdd(DX + i, gx + lx[t]);
        unsafeAtomicAdd(DY + i, gy + ly[t]);
        unsafeAtomicAdd(DZ + i, gz + lz[t]);
        unsafeAtomicAdd(dmass + i, gm + lm[t]);
    }
}

hipError_t nbx_launch_sph_fwd_rows(const double *const *X, const double *mass, const int32_t *nbr, int64_t n,
                                   int64_t r0, int64_t nr, int k, double *rho, double *h, hipStream_t st) {
    int blocks = (int)((nr + ADJ_B - 1) / ADJ_B);
    if (!blocks) return hipSuccess;
    ProfScope ps("k_sph_fwd", st);
    hipLaunchKernelGGL(k_sph_fwd, dim3(blocks), dim3(ADJ_B), 0, st, X[0], X[1], X[2], mass, nbr, n, r0, nr, k, rho, h);
    return hipGetLastError();
}

hipError_t nbx_launch_sph_adj_rows(const double *const *X, const double *mass, const int32_t *nbr, int64_t n,
                                   int64_t r0, int64_t nr, int k, const double *g, double *const *D, double *dmass,
                                   hipStream_t st) {
    int blocks = (int)((nr + ADJ_B - 1) / ADJ_B);
    if (!blocks) return hipSuccess;
    ProfScope ps("k_sph_adj", st);
    hipLaunchKernelGGL(k_sph_adj, dim3(blocks), dim3(ADJ_B), 0, st, X[0], X[1], X[2], mass, nbr, n, r0, nr, k, g, D[0],
                       D[1], D[2], dmass);
    return hipGetLastError();
}

hipError_t nbx_launch_sph_fwd(const double *pos, const double *mass, const int32_t *nbr, int64_t n, int k, double *rho,
                              double *h, hipStream_t st) {
    const double *X[3] = {pos, pos + n, pos + 2 * n};
    return nbx_launch_sph_fwd_rows(X, mass, nbr, n, 0, n, k, rho, h, st);
}

hipError_t nbx_launch_sph_adj(const double *pos, const double *mass, const int32_t *nbr, int64_t n, int k,
                              const double *g, double *dpos, double *dmass, hipStream_t st) {
    const double *X[3] = {pos, pos + n, pos + 2 * n};
    double *D[3] = {dpos, dpos + n, dpos + 2 * n};
    return nbx_launch_sph_adj_rows(X, mass, nbr, n, 0, n, k, g, D, dmass, st);
}

// Move each row's k-th neighbour -- its largest (d2, ID), the one that sets
// h -- into the last column (rows of slots from a search whose rows are not
// in exact FP64 order).  d2 summed x -> z without FMA, as the search's.
// A row with fewer than k neighbours (entries -1, possible when exact
// duplicates are excluded) raises *err and its missing entries become the
// row's own slot, so no later pass reads outside the arrays.
__global__ void k_kth_last(const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
                           const int32_t *__restrict__ perm, int32_t *nbr, int64_t r0, int64_t nr, int k,
                           int32_t *err) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const int64_t i = r0 + r;
    int32_t *row = nbr + r * k;
    const double xi = X[i], yi = Y[i], zi = Z[i];
    int best = 0;
    double bd = -1.0;
    int32_t bid = -1;
    for (int j = 0; j < k; ++j) {
        if (row[j] < 0) {
            row[j] = (int32_t)i;
            atomicOr(err, 1);
        }
        const int32_t s = row[j];
        const double ex = xi - X[s], ey = yi - Y[s], ez = zi - Z[s];
        const double d = ex * ex + ey * ey + ez * ez;
        const int32_t id = perm[s];
        if (d > bd || (d == bd && id > bid)) { best = j; bd = d; bid = id; }
    }
    const int32_t tmp = row[k - 1];
    row[k - 1] = row[best];
    row[best] = tmp;
}

hipError_t nbx_launch_kth_last(const double *const *X, const int32_t *perm, int32_t *nbr, int64_t r0, int64_t nr, int k,
                               int32_t *err, hipStream_t st) {
    const int blocks = (int)((nr + 255) / 256);
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_kth_last, dim3(blocks), dim3(256), 0, st, X[0], X[1], X[2], perm, nbr, r0, nr, k, err);
    return hipGetLastError();
}
