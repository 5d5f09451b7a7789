// nbx_smooth.hip -- smoothed velocity mean and dispersion on gfx950.
//
// Replaces KDTree::CalcSmoothVel (KDCalcSmoothQuantities.cxx:480-540),
// CalcSmoothVelDisp (:542-614), CalcSmoothVelSkew (:617-690) and
// CalcSmoothVelKurtosis (:692-767) on TPHYS trees, and the local mean and
// dispersion of a per-particle scalar (k_smooth_local).  For every particle i the
// reference takes the non-periodic target kNN of CalcDensity, h_i = r_k / 2,
// and for each neighbour j adds half-weight shares to both ends:
//   mean_i += W_ij / rho_j * m_j * v_j,   mean_j += W_ij / rho_i * m_i * v_i,
// with W_ij = W(r_ij / h_i) h_i^-3 / 2 from the kernel table (the weights of
// CalcDensity's symmetric sum).  The dispersion adds the same weights times
// (v - mean)(v - mean)^T about the receiving particle's finished mean.
//
// The neighbour lists come from the fused kNN + density pass (rows by slot,
// entries slots).  One thread per particle walks its row: its own share sums
// in registers, the other end's share goes to LDS when that particle is in
// the same block and to an FP64 atomic otherwise.  Sums are
// order-independent up to rounding.
#include "nbx_kernels.h"

__device__ __forceinline__ double smooth_d2(const DevTree &T, const double *xi, int32_t j) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double t = xi[d] - T.x[d][j];
        s += t * t;
    }
    return s;
}

// h of row i: half the largest neighbour distance (the k-th)
__device__ __forceinline__ double smooth_h(const DevTree &T, const double *xi, const int32_t *row, int k) {
    double m = 0.0;
    for (int t = 0; t < k; ++t) m = fmax(m, smooth_d2(T, xi, row[t]));
    return 0.5 * sqrt(m);
}

// Blocks of SM_B consecutive slots: a share for a particle of the same block
// (most neighbours are) is summed with LDS atomics, and each particle's total
// goes out with one global atomic per component; shares for particles of
// other blocks go out as global atomics at once.
#define SM_B 1024

__global__ __launch_bounds__(SM_B) void k_smooth_vel(DevTree T, SmoothArgs A) {
    __shared__ double lv[3][SM_B];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * SM_B, i = i0 + tid;
    const int64_t n = T.n;
    const bool valid = i < n;
#pragma unroll
    for (int d = 0; d < 3; ++d) lv[d][tid] = 0.0;
    __syncthreads();
    double acc[3] = {0.0, 0.0, 0.0};
    if (valid) {
        const int k = A.k;
        const int32_t *row = A.nn + i * k;
        double xi[3], vi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xi[d] = T.x[d][i];
            vi[d] = T.x[3 + d][i];
        }
        const double hi = smooth_h(T, xi, row, k);
        const double norm = 1.0 / pow(hi, 3.0);
        const double rhoi = A.rho[i], mi = T.mass[i];
        for (int t = 0; t < k; ++t) {
            const int32_t j = row[t];
            const double rij = sqrt(smooth_d2(T, xi, j));
            const double w = 0.5 * wsm_at(rij, hi, T.kern, T.kernres) * norm;
            const double tj = w / A.rho[j] * T.mass[j];
            const double ti = w / rhoi * mi;
            const uint32_t r = (uint32_t)(j - i0);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                acc[d] += tj * T.x[3 + d][j];
                if (r < SM_B) atomicAdd(&lv[d][r], ti * vi[d]);
                else unsafeAtomicAdd(A.mean + d * n + j, ti * vi[d]);
            }
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int d = 0; d < 3; ++d) unsafeAtomicAdd(A.mean + d * n + i, acc[d] + lv[d][tid]);
    }
}

// Second pass, over the finished means.  The 3x3 products are symmetric
// (dv_a * dv_b == dv_b * dv_a exactly), so six are summed and mirrored.
__global__ __launch_bounds__(SM_B) void k_smooth_disp(DevTree T, SmoothArgs A) {
    __shared__ double ld[6][SM_B];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * SM_B, i = i0 + tid;
    const int64_t n = T.n;
    const bool valid = i < n;
#pragma unroll
    for (int c = 0; c < 6; ++c) ld[c][tid] = 0.0;
    __syncthreads();
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (valid) {
        const int k = A.k;
        const int32_t *row = A.nn + i * k;
        double xi[3], vi[3], ui[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xi[d] = T.x[d][i];
            vi[d] = T.x[3 + d][i];
            ui[d] = A.mean[d * n + i];
        }
        const double hi = smooth_h(T, xi, row, k);
        const double norm = 1.0 / pow(hi, 3.0);
        const double rhoi = A.rho[i], mi = T.mass[i];
        for (int t = 0; t < k; ++t) {
            const int32_t j = row[t];
            const double rij = sqrt(smooth_d2(T, xi, j));
            const double w = 0.5 * wsm_at(rij, hi, T.kern, T.kernres) * norm;
            const double tj = w / A.rho[j] * T.mass[j];
            const double ti = w / rhoi * mi;
            const uint32_t r = (uint32_t)(j - i0);
            double a[3], b[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                a[d] = T.x[3 + d][j] - ui[d];       // neighbour about my mean
                b[d] = vi[d] - A.mean[d * n + j];  // me about the neighbour's mean
            }
            int c = 0;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = p; q < 3; ++q, ++c) {
                    acc[c] += tj * a[p] * a[q];
                    const double v = ti * b[p] * b[q];
                    if (r < SM_B) atomicAdd(&ld[c][r], v);
                    else unsafeAtomicAdd(A.disp + (3 * p + q) * n + j, v);
                }
        }
    }
    __syncthreads();
    if (valid) {
        int c = 0;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = p; q < 3; ++q, ++c) unsafeAtomicAdd(A.disp + (3 * p + q) * n + i, acc[c] + ld[c][tid]);
    }
}

// Third pass: per-component skewness and kurtosis (CalcSmoothVelSkew,
// KDCalcSmoothQuantities.cxx:617-690; CalcSmoothVelKurtosis, :692-767) about
// the receiving particle's mean, scaled by its dispersion diagonal s:
// share * dv^3 / s^1.5 and share * dv^4 / (s * s) - 3, the 3 taken off every
// share as the reference does.
__global__ __launch_bounds__(SM_B) void k_smooth_moments(DevTree T, SmoothArgs A) {
    __shared__ double lm[6][SM_B];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * SM_B, i = i0 + tid;
    const int64_t n = T.n;
    const bool valid = i < n;
#pragma unroll
    for (int c = 0; c < 6; ++c) lm[c][tid] = 0.0;
    __syncthreads();
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (valid) {
        const int k = A.k;
        const int32_t *row = A.nn + i * k;
        double xi[3], vi[3], ui[3], si[3], pi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xi[d] = T.x[d][i];
            vi[d] = T.x[3 + d][i];
            ui[d] = A.mean[d * n + i];
            si[d] = A.disp[4 * d * n + i];
            pi[d] = pow(si[d], 1.5);
        }
        const double hi = smooth_h(T, xi, row, k);
        const double norm = 1.0 / pow(hi, 3.0);
        const double rhoi = A.rho[i], mi = T.mass[i];
        for (int t = 0; t < k; ++t) {
            const int32_t j = row[t];
            const double rij = sqrt(smooth_d2(T, xi, j));
            const double w = 0.5 * wsm_at(rij, hi, T.kern, T.kernres) * norm;
            const double tj = w / A.rho[j] * T.mass[j];
            const double ti = w / rhoi * mi;
            const uint32_t r = (uint32_t)(j - i0);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double a = T.x[3 + d][j] - ui[d];
                acc[d] += tj * a * a * a / pi[d];
                acc[3 + d] += tj * a * a * a * a / (si[d] * si[d]) - 3.;
                const double b = vi[d] - A.mean[d * n + j], sj = A.disp[4 * d * n + j];
                const double vs = ti * b * b * b / pow(sj, 1.5);
                const double vk = ti * b * b * b * b / (sj * sj) - 3.;
                if (r < SM_B) {
                    atomicAdd(&lm[d][r], vs);
                    atomicAdd(&lm[3 + d][r], vk);
                } else {
                    unsafeAtomicAdd(A.skew + d * n + j, vs);
                    unsafeAtomicAdd(A.kurt + d * n + j, vk);
                }
            }
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            unsafeAtomicAdd(A.skew + d * n + i, acc[d] + lm[d][tid]);
            unsafeAtomicAdd(A.kurt + d * n + i, acc[3 + d] + lm[3 + d][tid]);
        }
    }
}

// CalcSmoothLocalMeanParticle / CalcSmoothLocalDispParticle (:1539-1617) and
// the Position forms (:1619-1693): the gather sum over a row with the full
// weight W(r/h) h^-3 (no half, no scatter), each neighbour counting
// W / rho_j * m_j * w_j, and for the dispersion W / rho_j * m_j * (w_j - mean)^2.
__global__ __launch_bounds__(256) void k_smooth_local(DevTree T, LocalArgs A) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= A.m) return;
    const int k = A.k;
    const int32_t *row = A.nn + r * k;
    double xi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xi[d] = A.q[0] ? A.q[d][r] : T.x[d][r];
    double m2 = 0.0;
    for (int t = 0; t < k; ++t) {
        const int32_t j = A.nn_ids ? T.inv[row[t]] : row[t];
        m2 = fmax(m2, smooth_d2(T, xi, j));
    }
    const double hi = 0.5 * sqrt(m2);
    const double norm = 1.0 / pow(hi, 3.0);
    const double mu = A.mean_in ? A.mean_in[r] : 0.0;
    double mean = 0.0, disp = 0.0;
    for (int t = 0; t < k; ++t) {
        const int32_t j = A.nn_ids ? T.inv[row[t]] : row[t];
        const double rij = sqrt(smooth_d2(T, xi, j));
        const double w = wsm_at(rij, hi, T.kern, T.kernres) * norm;
        const double temp = w / A.rho[j] * T.mass[j];
        mean += temp * A.w[j];
        const double delta = A.w[j] - mu;
        disp += temp * delta * delta;
    }
    if (A.mean) A.mean[r] = mean;
    if (A.disp && A.mean_in) A.disp[r] = disp;
}

hipError_t nbx_launch_smooth_local(const DevTree &T, const LocalArgs &A, hipStream_t st) {
    if (A.m == 0) return hipSuccess;
    ProfScope ps("k_smooth_local", st);
    hipLaunchKernelGGL(k_smooth_local, dim3((unsigned)((A.m + 255) / 256)), dim3(256), 0, st, T, A);
    return hipGetLastError();
}

// Slot-major SoA to ID-major rows: out[id * ncomp + c] = in[c * n + slot];
// the lower triangle of a 3x3 dispersion (ncomp 9) mirrors the upper.
__global__ __launch_bounds__(256) void k_smooth_out(const double *in, const int32_t *perm, int64_t n, int ncomp,
                                                    double *out) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int64_t id = perm[s];
    for (int c = 0; c < ncomp; ++c) {
        int src = c;
        if (ncomp == 9 && (c / 3) > (c % 3)) src = 3 * (c % 3) + c / 3;
        out[id * ncomp + c] = in[(int64_t)src * n + s];
    }
}

__global__ __launch_bounds__(256) void k_gather_by_id(const double *by_id, const int32_t *perm, int64_t n,
                                                      double *by_slot) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s < n) by_slot[s] = by_id[perm[s]];
}

static inline int nblk(int64_t n) { return (int)((n + 255) / 256); }

hipError_t nbx_launch_smooth(const DevTree &T, const SmoothArgs &A, bool with_mean, bool with_disp, hipStream_t st) {
    const int64_t n = T.n;
    if (n == 0) return hipSuccess;
    if (T.nd != 3 || !T.x[3] || A.k < 1) return hipErrorInvalidValue;
    if ((with_disp || A.skew) && !A.disp) return hipErrorInvalidValue;
    if (A.skew && !A.kurt) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((n + SM_B - 1) / SM_B));
    if (with_mean) {
        ProfScope ps("k_smooth_vel", st);
        hipLaunchKernelGGL(k_smooth_vel, grid, dim3(SM_B), 0, st, T, A);
    }
    if (with_disp) {
        ProfScope ps("k_smooth_disp", st);
        hipLaunchKernelGGL(k_smooth_disp, grid, dim3(SM_B), 0, st, T, A);
    }
    if (A.skew) {
        ProfScope ps("k_smooth_moments", st);
        hipLaunchKernelGGL(k_smooth_moments, grid, dim3(SM_B), 0, st, T, A);
    }
    return hipGetLastError();
}
