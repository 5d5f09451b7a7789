// nbx_host_stub.cpp -- TEST CODE ONLY: a host implementation of the nbx_*
// entry points the C++ drop-in (kdtree_shim.cpp) calls, answered by the CPU
// oracle (oracle/nbx_oracle.c).  It lets the drop-in's host logic -- the ID
// overwrite, the in-place reorder and its undoing, the per-particle caches
// called from OpenMP threads, FOF with a check function -- run under the
// address / undefined-behaviour sanitizers on a machine without a GPU.  It is
// never linked into libnbx.so.  Its tree order is the oracle's median-split
// order, so the drop-in's reordering is exercised with a real permutation.
// Entry points the driver does not use return NBX_ERR_UNSUPPORTED.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "nbx.h"

extern "C" {
void *orc_tree_build(const double *phase, const double *mass, int64_t n, int b, int treetype, int kerntype,
                     int kernres, const double *period);
void orc_tree_free(void *h);
void orc_tree_order(void *h, int32_t *slot_to_id);
int orc_kernel(int nd, int kerntype, int kernres, double *kernnorm_out, double *table);
void orc_knn_tree(void *h, int k, int variant, int32_t *ids, double *d2, int nthreads);
void orc_knn_points(void *h, const double *q, int64_t m, int k, int32_t *ids, double *d2, int nthreads);
int orc_calc_density(void *h, int k, double *rho_by_id);
int orc_calc_density_gather(void *h, int k, double *rho_by_id, double *h_by_id, int nthreads);
int orc_calc_density_position(void *h, const double *q, int64_t m, int k, double *rho, double *hout, int nthreads);
int orc_calc_veldensity(void *h, int nsmooth, int nsearch, double *rho_by_id, int nthreads);
int64_t orc_fof_tree(void *h, double fdist, int minnum, int order, int32_t *labels_by_id);
int64_t orc_fof_canonical(int64_t n, const int32_t *labels, int minnum, int32_t *out);
void orc_ball_mark(const double *phase, int64_t n, const double *x, int64_t exclude, double r2, const double *period,
                   int32_t imark, int32_t *group, double *d2out);
}

static thread_local std::string g_err;
static int fail(int code, const char *msg) {
    g_err = msg;
    return code;
}
extern "C" const char *nbx_last_error(void) { return g_err.c_str(); }
extern "C" int nbx_host_fail(int code, const char *msg) { return fail(code, msg); }

struct nbx_tree {
    nbx_params p;
    int64_t n;
    void *h;
    std::vector<double> phase;  // n x 6 by ID
    std::vector<int32_t> order, inv;
    const double *period() const { return p.periodic ? p.period : nullptr; }
};

extern "C" void nbx_params_default(nbx_params *p) {
    memset(p, 0, sizeof(*p));
    p->bucket_size = 16;
    p->treetype = NBX_TPHYS;
    p->kerntype = NBX_KEPAN;
    p->kernres = 1000;
    p->form = NBX_FORM_PLAIN;
    p->recheck_rel = 1e-5;
}

extern "C" int nbx_tree_create(const nbx_params *p, const double *pos, const double *vel, const double *mass,
                               int64_t n, nbx_tree **out) {
    if (!p || !pos || !out || n < 1) return fail(NBX_ERR_ARG, "bad argument");
    if (p->treetype != NBX_TPHYS) return fail(NBX_ERR_UNSUPPORTED, "the host stub builds TPHYS trees");
    nbx_tree *t = new nbx_tree();
    t->p = *p;
    t->n = n;
    t->phase.assign(6 * (size_t)n, 0.0);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            t->phase[6 * i + d] = pos[d * n + i];
            t->phase[6 * i + 3 + d] = vel ? vel[d * n + i] : 0.0;
        }
    t->h = orc_tree_build(t->phase.data(), mass, n, p->bucket_size, p->treetype, p->kerntype, p->kernres,
                          t->period());
    t->order.resize(n);
    t->inv.resize(n);
    orc_tree_order(t->h, t->order.data());
    for (int64_t s = 0; s < n; ++s) t->inv[t->order[s]] = (int32_t)s;
    *out = t;
    return NBX_OK;
}
extern "C" int nbx_tree_destroy(nbx_tree *t) {
    if (t) orc_tree_free(t->h);
    delete t;
    return NBX_OK;
}
extern "C" int nbx_tree_order(nbx_tree *t, int32_t *slot_to_id) {
    memcpy(slot_to_id, t->order.data(), sizeof(int32_t) * t->n);
    return NBX_OK;
}
extern "C" int nbx_tree_kernel(const nbx_tree *t, double *kernnorm, double *, int32_t) {
    return orc_kernel(3, t->p.kerntype, t->p.kernres, kernnorm, nullptr);
}

extern "C" int nbx_knn(nbx_tree *t, int32_t k, int32_t variant, int32_t layout, int32_t *nn, double *d2) {
    if (layout != NBX_LAYOUT_BY_SLOT || variant > NBX_KNN_NONPERIODIC) return fail(NBX_ERR_UNSUPPORTED, "stub");
    const int64_t n = t->n;
    std::vector<int32_t> ids((size_t)n * k);
    std::vector<double> dd((size_t)n * k);
    orc_knn_tree(t->h, k, variant, ids.data(), dd.data(), 0);
    for (int64_t s = 0; s < n; ++s) {
        const int32_t id = t->order[s];
        for (int j = 0; j < k; ++j) {
            const int32_t e = ids[(size_t)id * k + j];
            nn[s * k + j] = e >= 0 ? t->inv[e] : -1;
            if (d2) d2[s * k + j] = dd[(size_t)id * k + j];
        }
    }
    return NBX_OK;
}
extern "C" int nbx_knn_points(nbx_tree *t, const double *q, int64_t m, int32_t k, int32_t *nn, double *d2) {
    std::vector<double> rows(3 * (size_t)m);
    for (int64_t i = 0; i < m; ++i)
        for (int d = 0; d < 3; ++d) rows[3 * i + d] = q[d * m + i];
    std::vector<double> dd((size_t)m * k);
    orc_knn_points(t->h, rows.data(), m, k, nn, dd.data(), 0);
    if (d2) memcpy(d2, dd.data(), sizeof(double) * dd.size());
    return NBX_OK;
}
extern "C" int nbx_knn_phase_points(nbx_tree *, const double *, const double *, int64_t, int32_t, int32_t *,
                                    double *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}

extern "C" int nbx_density(nbx_tree *t, int32_t k, int32_t variant, int32_t nsearch, double *rho_by_id,
                           double *h_by_id) {
    int rc = -1;
    if (variant == NBX_DENS_SYMMETRIC && !h_by_id) rc = orc_calc_density(t->h, k, rho_by_id);
    else if (variant == NBX_DENS_GATHER) rc = orc_calc_density_gather(t->h, k, rho_by_id, h_by_id, 0);
    else if (variant == NBX_DENS_VEL && !h_by_id) rc = orc_calc_veldensity(t->h, k, nsearch, rho_by_id, 0);
    else return fail(NBX_ERR_UNSUPPORTED, "stub");
    return rc == 0 ? NBX_OK : fail(NBX_ERR_SEARCH, "CalcDensity failed");
}
extern "C" int nbx_density_points(nbx_tree *t, const double *q, int64_t m, int32_t k, double *rho, double *h) {
    std::vector<double> rows(3 * (size_t)m);
    for (int64_t i = 0; i < m; ++i)
        for (int d = 0; d < 3; ++d) rows[3 * i + d] = q[d * m + i];
    return orc_calc_density_position(t->h, rows.data(), m, k, rho, h, 0) == 0 ? NBX_OK
                                                                              : fail(NBX_ERR_SEARCH, "failed");
}

// ball searches by brute force over the particles (IDs), with the image rule
static int ball_ids(nbx_tree *t, const double *x, double r2, std::vector<int32_t> &ids, std::vector<double> &d2) {
    const int64_t n = t->n;
    std::vector<int32_t> g(n, -1);
    std::vector<double> dd(n, 0.0);
    orc_ball_mark(t->phase.data(), n, x, -1, r2, t->period(), 1, g.data(), dd.data());
    ids.clear();
    d2.clear();
    for (int64_t i = 0; i < n; ++i)
        if (g[i] == 1) {
            ids.push_back((int32_t)i);
            d2.push_back(dd[i]);
        }
    return NBX_OK;
}
extern "C" int nbx_ball_points(nbx_tree *t, const double *x, double r2, int32_t *out, int64_t cap, int64_t *nout) {
    std::vector<int32_t> ids;
    std::vector<double> d2;
    ball_ids(t, x, r2, ids, d2);
    *nout = (int64_t)ids.size();
    for (int64_t i = 0; i < cap && i < *nout; ++i) out[i] = ids[i];
    return NBX_OK;
}
extern "C" int nbx_ball_points_many(nbx_tree *t, const double *q, int64_t m, double r2, int64_t *offsets,
                                    int32_t *ids, double *d2, int64_t cap) {
    std::vector<std::vector<int32_t>> rows(m);
    std::vector<std::vector<double>> rd(m);
    offsets[0] = 0;
    for (int64_t i = 0; i < m; ++i) {
        const double x[3] = {q[i], q[m + i], q[2 * m + i]};
        ball_ids(t, x, r2, rows[i], rd[i]);
        offsets[i + 1] = offsets[i] + (int64_t)rows[i].size();
    }
    if (offsets[m] > cap || !ids) return NBX_OK;
    for (int64_t i = 0; i < m; ++i)
        for (size_t e = 0; e < rows[i].size(); ++e) {
            ids[offsets[i] + e] = rows[i][e];
            if (d2) d2[offsets[i] + e] = rd[i][e];
        }
    return NBX_OK;
}
extern "C" int nbx_ball_mark(nbx_tree *t, const double *x, int64_t exclude_id, double r2, int32_t imark,
                             int32_t *group_by_id, double *d2_by_id) {
    orc_ball_mark(t->phase.data(), t->n, x, exclude_id, r2, t->period(), imark, group_by_id, d2_by_id);
    return NBX_OK;
}

extern "C" int nbx_fof(nbx_tree *t, double fdist, int32_t minnum, int32_t *labels_by_id, int64_t *ngroups) {
    std::vector<int32_t> raw(t->n);
    orc_fof_tree(t->h, fdist, minnum, 0, raw.data());
    *ngroups = orc_fof_canonical(t->n, raw.data(), 1, labels_by_id);
    return NBX_OK;
}

// not used by the driver
extern "C" int nbx_fof_criterion(nbx_tree *, int32_t, const double *, int32_t, int32_t *, int64_t *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}
extern "C" int nbx_smooth_vel(nbx_tree *, int32_t, const double *, const double *, double *, double *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}
extern "C" int nbx_smooth_vel_moments(nbx_tree *, int32_t, const double *, const double *, const double *, double *,
                                      double *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}
extern "C" int nbx_smooth_local(nbx_tree *, int32_t, const double *, const double *, const double *, double *,
                                double *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}
extern "C" int nbx_smooth_local_points(nbx_tree *, const double *, int64_t, int32_t, const double *, const double *,
                                       const double *, double *, double *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}
extern "C" int nbx_veldensity_phys(nbx_tree *, int32_t, int32_t, const double *, double *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}
extern "C" int nbx_veldensity_points(nbx_tree *, const double *, const double *, int64_t, int32_t, int32_t, double *,
                                     double *) {
    return fail(NBX_ERR_UNSUPPORTED, "stub");
}
