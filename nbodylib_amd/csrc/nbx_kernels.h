// nbx_kernels.h -- argument blocks and host-side launchers shared between the
// HIP translation units and the C ABI (nbx_api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nbx_internal.h"

// Gather into tree order fused with the leaf boxes and, for the precision
// form, the leaf origins, half-diagonals and FP32 offsets.
struct GatherArgs {
    int64_t n, npad;
    int ncoord;  // coordinates gathered (3, or 6 with velocities)
    int nd;      // tree dimensions boxed
    int32_t nleaf, nleafp;
    const int32_t *order;
    const double *in[6];
    double *out[6];
    const double *mass_in;
    double mass_default;
    double *mass_out;
    int32_t *perm, *inv;
    double *box;
    double *org;  // nullable: leaf centre (precision-form origin)
    double *ext;  // nullable: leaf half-diagonal (precision-form error bound)
    float *xf[6];  // precision form, 6-D: FP32 offsets from the leaf origin
    float4 *xq;    // precision form, 3-D: the offsets and fl(fl(fl(x^2) + y^2) + z^2) interleaved
};

__global__ void k_bbox(const double *c, int64_t n, double *part);
__global__ void k_gather_leaf(GatherArgs a);
__global__ void k_leafbox_pad(double *box, int32_t nleaf, int32_t nleafp, int nd);
__global__ void k_levelbox(double *box, float *plane, int32_t first, int32_t count, int nd);
__global__ void k_boxf(const double *box, float *boxf, int64_t count, int nd);

// order (slot -> ID) of the implicit kd tree: midpoint splits along the
// widest dimension of every node (nbx_kdbuild.hip)
hipError_t nbx_kd_order(const double *const *c, int nd, int64_t n, int32_t nleafp, int32_t *order, hipStream_t st);
hipError_t nbx_sort_pairs_u64(void *temp, size_t &temp_bytes, uint64_t *kin, uint64_t *kout, int32_t *vin,
                              int32_t *vout, int64_t n, hipStream_t st);

// ------------------------------------------------------------------ kNN
// Query source
#define NBX_Q_SLOTS 0   // every particle, target-based (self / d2==0 excluded)
#define NBX_Q_PERIODIC 1  // every particle, position-based + sequential periodic images
#define NBX_Q_POINTS 2  // external points, position-based (+ images if periodic)
// Density fused into the search
#define NBX_DO_NONE 0
#define NBX_DO_SYM 1     // CalcDensity: 1/2 W scattered to both ends
#define NBX_DO_GATHER 2  // CalcDensityParticle: W m_j gathered
#define NBX_DO_VEL 3     // CalcVelDensity
#define NBX_DO_VELPHYS 4  // CalcVelDensityWithPhysDensity: CalcVelDensity weighted by W(r) m_i / rho_i
// Output drop rule for the k+1 periodic search
#define NBX_DROP_NONE 0
#define NBX_DROP_SMALLEST 1
#define NBX_DROP_LARGEST 2

struct KnnArgs {
    int32_t qsrc;       // NBX_Q_*
    int32_t kh;         // heap size (k, or k+1 for periodic)
    int32_t kout;       // entries written per row
    int32_t drop;       // NBX_DROP_*
    int32_t excl_zero;  // target rule: skip d2 == 0
    int32_t layout;     // NBX_LAYOUT_BY_ID / BY_SLOT
    int32_t dens;       // NBX_DO_*
    int32_t nsmooth;    // velocity density
    int32_t m;          // external points
    int32_t periodic;
    double period[3];
    const double *q[3];  // external points (SoA)
    const double *qv[3];  // external points: their velocities (NBX_DO_VEL)
    const double *rho_in;  // NBX_DO_VELPHYS: densities by slot
    int32_t *nn;        // nullable
    double *d2;         // nullable
    double *rho;        // by slot (density); by query row for position queries
    double *hout;       // as rho, or by ID with hout_by_id (nullable)
    int32_t hout_by_id;  // particle queries: hout[ID] rather than hout[slot]
    int32_t *err;       // set to 1 if a row has fewer than kh neighbours
    const int32_t *qlist;  // nullable: query slots (FP64 re-run of mixed-form fallbacks)
    // particle queries of a run of whole waves only: slots [64 wbeg, 64 (wbeg + wcnt)) (wcnt 0: every slot)
    int32_t wbeg, wcnt;
    int32_t row0;  // subtracted from the slot of a row written in the by-slot layout
    int32_t list_rows;  // with qlist: row r of the output is the query qlist[r]
    double band_rel;    // precision form: least relative half-width of the FP64 re-check band
};

hipError_t nbx_launch_knn(const DevTree &T, const KnnArgs &A, int nqueries, hipStream_t st);
// work (nullable): device memory of at least nbx_mixed_scratch_bytes(nd, kh)
// bytes for the per-workgroup scratch and the queue counters; without it the
// launcher allocates them on the stream.
hipError_t nbx_launch_knn_mixed(const DevTree &T, const KnnArgs &A, const double *ext, int32_t *fallback_list,
                                int32_t *fallback_count, hipStream_t st, void *work = nullptr, size_t work_bytes = 0);
size_t nbx_mixed_scratch_bytes(int nd, int kh);
// FP64 d2 of every entry of the n rows nn (by ID, layout 0, or by slot,
// layout 1) written by the precision-form kernel, rows re-sorted ascending
// by (d2, ID); k <= 64.
hipError_t nbx_launch_rows_fp64(const DevTree &T, int32_t *nn, double *d2, int k, int layout, hipStream_t st);

// ------------------------------------------------------------------ FOF
// Link rule of a FOF run: pairs link when their value is below lim.  Plain
// FOF: d2 against b^2.  FOFCriterion with FOF3d / FOF6d: the comparator's sum
// of squares divided by s6 (positions) and s7 (velocities), against 1.
#define NBX_FOFM_PLAIN 0
#define NBX_FOFM_3D 1
#define NBX_FOFM_6D 2
struct FofLink {
    double lim;
    double s6, s7;
    double r6, r7;  // 1 / s6, 1 / s7 (correctly rounded)
};
struct FofArgs {
    FofLink link;
    int32_t metric;  // NBX_FOFM_*
    int32_t periodic;
    double period[3];
    int32_t *parent;
};
// A.parent: nbx_fof_work_bytes(T) bytes (parents, node flags, queue counters)
hipError_t nbx_launch_fof(const DevTree &T, const FofArgs &A, hipStream_t st);
size_t nbx_fof_work_bytes(const DevTree &T);
hipError_t nbx_launch_fof_finish(const DevTree &T, int32_t *parent, int32_t *count, int32_t *minid, int64_t n,
                                 hipStream_t st);

// ------------------------------------------------------------------ ball search, many positions
struct BallManyArgs {
    const double *q[3];  // positions, SoA
    int64_t m;
    double r2;
    int32_t periodic;
    double period[3];
    int64_t *count;      // counting pass: hits per position
    const int64_t *off;  // filling pass: first output index per position
    int32_t *ids;        // filling pass
    double *d2;          // filling pass, nullable
    const int32_t *order;  // positions in walk order (nullable: as given)
};
hipError_t nbx_launch_ball_many(const DevTree &T, const BallManyArgs &A, bool fill, hipStream_t st);
// Morton keys (21 bits a dimension over the box lo + [0, 2^21 / s)) and
// their indices, for ordering query positions along the space-filling curve
__global__ void k_morton_keys(const double *qx, const double *qy, const double *qz, int64_t m, double lo0, double lo1,
                              double lo2, double s0, double s1, double s2, uint64_t *keys, int32_t *idx);

// ------------------------------------------------------------------ smoothed velocities
struct SmoothArgs {
    const int32_t *nn;  // n x k neighbour slots, rows by slot
    int32_t k;
    const double *rho;  // densities by slot
    double *mean;       // smoothed velocity by slot, [3][n]; zeroed by the caller
    double *disp;       // dispersion by slot, [9][n] (upper triangle used) or null; zeroed by the caller
    double *skew;       // skewness by slot, [3][n], or null; zeroed by the caller
    double *kurt;       // kurtosis by slot, [3][n] (with skew); zeroed by the caller
};
// with_mean / with_disp: compute those first (else A.mean / A.disp hold the
// caller's); the moments run when A.skew is set.
hipError_t nbx_launch_smooth(const DevTree &T, const SmoothArgs &A, bool with_mean, bool with_disp, hipStream_t st);
// Local mean / dispersion of a per-particle scalar (gather form): row r of
// nn holds the k neighbours of query r, whose coordinates are q (position
// queries) or the particle at slot r (q[0] null).
struct LocalArgs {
    const double *q[3];     // query coordinates, or null for particle rows
    int64_t m;              // rows
    const int32_t *nn;      // m x k neighbours
    int32_t k;
    int32_t nn_ids;         // entries are IDs (else slots)
    const double *rho;      // by slot
    const double *w;        // weights by slot
    const double *mean_in;  // by row, for the dispersion (null: no dispersion)
    double *mean;           // by row, nullable
    double *disp;           // by row, nullable
};
hipError_t nbx_launch_smooth_local(const DevTree &T, const LocalArgs &A, hipStream_t st);
__global__ void k_smooth_out(const double *in, const int32_t *perm, int64_t n, int ncomp, double *out);
__global__ void k_gather_by_id(const double *by_id, const int32_t *perm, int64_t n, double *by_slot);

// ------------------------------------------------------------------ small kernels
__global__ void k_scatter_by_id(const double *src, const int32_t *perm, double *dst, int64_t n);
__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n);
__global__ void k_fill_f64(double *p, double v, int64_t n);
__global__ void k_iota(int32_t *p, int64_t n);
__global__ void k_ids_to_slots(const int32_t *ids, const int32_t *inv, int64_t m, int32_t *slots);

// ------------------------------------------------------------------ adjoint
hipError_t nbx_launch_sph_fwd(const double *pos, const double *mass, const int32_t *nbr, int64_t n, int k, double *rho,
                              double *h, hipStream_t st);
hipError_t nbx_launch_sph_adj(const double *pos, const double *mass, const int32_t *nbr, int64_t n, int k,
                              const double *g, double *dpos, double *dmass, hipStream_t st);
// The same over rows [r0, r0 + nr) of n particles, row i of the list at
// nbr[(i - r0) k]; X[3] / D[3] are the coordinate / gradient arrays of all n.
hipError_t nbx_launch_sph_fwd_rows(const double *const *X, const double *mass, const int32_t *nbr, int64_t n,
                                   int64_t r0, int64_t nr, int k, double *rho, double *h, hipStream_t st);
hipError_t nbx_launch_sph_adj_rows(const double *const *X, const double *mass, const int32_t *nbr, int64_t n,
                                   int64_t r0, int64_t nr, int k, const double *g, double *const *D, double *dmass,
                                   hipStream_t st);
// Each row's largest (d2, ID) neighbour moved into its last column; rows
// with a missing entry (-1) set *err and have it replaced by their own slot.
hipError_t nbx_launch_kth_last(const double *const *X, const int32_t *perm, int32_t *nbr, int64_t r0, int64_t nr, int k,
                               int32_t *err, hipStream_t st);

// ------------------------------------------------------------------ profiling
struct ProfScope {
    hipStream_t st;
    hipEvent_t e0, e1;
    const char *name;
    bool on;
    ProfScope(const char *name, hipStream_t st);
    ~ProfScope();
};
