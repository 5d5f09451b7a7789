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
// One thread per particle i; contributions to other particles use hardware
// FP64 atomics.
#include "nbx_kernels.h"

__device__ __forceinline__ double w_cubic(double q) {
    if (q <= 1.0) return 1.0 - 1.5 * q * q + 0.75 * q * q * q;
    if (q <= 2.0) { double t = 2.0 - q; return 0.25 * t * t * t; }
    return 0.0;
}
__device__ __forceinline__ double dw_cubic(double q) {
    if (q <= 1.0) return -3.0 * q + 2.25 * q * q;
    if (q <= 2.0) { double t = 2.0 - q; return -0.75 * t * t; }
    return 0.0;
}

#define SIGMA3 0.31830988618379067154  // 1/pi

__global__ __launch_bounds__(256) void k_sph_fwd(const double *__restrict__ pos, const double *__restrict__ mass,
                                                  const int32_t *__restrict__ nbr, int64_t n, int k, double *rho,
                                                  double *hout) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double *X = pos, *Y = pos + n, *Z = pos + 2 * n;
    double xi = X[i], yi = Y[i], zi = Z[i], mi = mass[i];
    const int32_t *row = nbr + i * k;
    int64_t K = row[k - 1];
    double dx = xi - X[K], dy = yi - Y[K], dz = zi - Z[K];
    double h = 0.5 * sqrt(dx * dx + dy * dy + dz * dz);
    if (hout) hout[i] = h;
    double ih = 1.0 / h;
    double norm = 0.5 * SIGMA3 * ih * ih * ih;
    double acc = 0.0;
    for (int j = 0; j < k; ++j) {
        int64_t q = row[j];
        double ex = xi - X[q], ey = yi - Y[q], ez = zi - Z[q];
        double r = sqrt(ex * ex + ey * ey + ez * ez);
        double w = norm * w_cubic(r * ih);
        acc += w * mass[q];
        unsafeAtomicAdd(rho + q, w * mi);
    }
    unsafeAtomicAdd(rho + i, acc);
}

__global__ __launch_bounds__(256) void k_sph_adj(const double *__restrict__ pos, const double *__restrict__ mass,
                                                  const int32_t *__restrict__ nbr, int64_t n, int k,
                                                  const double *__restrict__ g, double *dpos, double *dmass) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double *X = pos, *Y = pos + n, *Z = pos + 2 * n;
    double *DX = dpos, *DY = dpos + n, *DZ = dpos + 2 * n;
    double xi = X[i], yi = Y[i], zi = Z[i], mi = mass[i], gi = g[i];
    const int32_t *row = nbr + i * k;
    int64_t K = row[k - 1];
    double kx = xi - X[K], ky = yi - Y[K], kz = zi - Z[K];
    double rk = sqrt(kx * kx + ky * ky + kz * kz);
    double h = 0.5 * rk;
    double ih = 1.0 / h;
    double ih3 = ih * ih * ih, ih4 = ih3 * ih;
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0, Hs = 0.0;
    for (int j = 0; j < k; ++j) {
        int64_t q = row[j];
        double ex = xi - X[q], ey = yi - Y[q], ez = zi - Z[q];
        double r = sqrt(ex * ex + ey * ey + ez * ez);
        double qq = r * ih;
        double W = w_cubic(qq), dW = dw_cubic(qq);
        double w = 0.5 * SIGMA3 * W * ih3;
        double gq = g[q];
        double a = gi * mass[q] + gq * mi;
        unsafeAtomicAdd(dmass + q, w * gi);
        gm += w * gq;
        Hs += a * (-0.5 * SIGMA3 * (qq * dW + 3.0 * W) * ih4);
        if (r > 0.0) {
            double f = a * 0.5 * SIGMA3 * dW * ih4 / r;
            gx += f * ex; gy += f * ey; gz += f * ez;
            unsafeAtomicAdd(DX + q, -f * ex);
            unsafeAtomicAdd(DY + q, -f * ey);
            unsafeAtomicAdd(DZ + q, -f * ez);
        }
    }
    if (rk > 0.0) {
        double f = Hs * 0.5 / rk;
        gx += f * kx; gy += f * ky; gz += f * kz;
        unsafeAtomicAdd(DX + K, -f * kx);
        unsafeAtomicAdd(DY + K, -f * ky);
        unsafeAtomicAdd(DZ + K, -f * kz);
    }
    unsafeAtomicAdd(DX + i, gx);
    unsafeAtomicAdd(DY + i, gy);
    unsafeAtomicAdd(DZ + i, gz);
    unsafeAtomicAdd(dmass + i, gm);
}

hipError_t nbx_launch_sph_fwd(const double *pos, const double *mass, const int32_t *nbr, int64_t n, int k, double *rho,
                              double *h, hipStream_t st) {
    int blocks = (int)((n + 255) / 256);
    if (!blocks) return hipSuccess;
    ProfScope ps("k_sph_fwd", st);
    hipLaunchKernelGGL(k_sph_fwd, dim3(blocks), dim3(256), 0, st, pos, mass, nbr, n, k, rho, h);
    return hipGetLastError();
}

hipError_t nbx_launch_sph_adj(const double *pos, const double *mass, const int32_t *nbr, int64_t n, int k,
                              const double *g, double *dpos, double *dmass, hipStream_t st) {
    int blocks = (int)((n + 255) / 256);
    if (!blocks) return hipSuccess;
    ProfScope ps("k_sph_adj", st);
    hipLaunchKernelGGL(k_sph_adj, dim3(blocks), dim3(256), 0, st, pos, mass, nbr, n, k, g, dpos, dmass);
    return hipGetLastError();
}
