// nbx_api.hip -- host side of the C ABI declared in include/nbx.h.
//
// Owns the device-resident tree (sorted SoA particles, node boxes, kernel
// table) and turns each NBodylib call into kernel launches on the caller's
// stream.  No CPU fallback exists: every result comes from the HIP kernels.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nbx.h"
#include "nbx_kernels.h"

size_t nbx_fof_canonical_bytes(int64_t n, int32_t minnum);
hipError_t nbx_fof_canonical(const DevTree &T, int32_t *P, int32_t minnum, int32_t *labels_by_id, int64_t *ngroups,
                             void *work, hipStream_t st);
__global__ void k_ball(DevTree T, double qx, double qy, double qz, double r2, int32_t *out, double *dout,
                       int64_t cap, unsigned long long *count);

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
static int hipfail(hipError_t e, const char *where) {
    return fail(e == hipErrorOutOfMemory ? NBX_ERR_NOMEM : NBX_ERR_HIP,
                std::string(where) + ": " + hipGetErrorString(e));
}
#define HIPCK(x)                                  \
    do {                                          \
        hipError_t _e = (x);                      \
        if (_e != hipSuccess) return hipfail(_e, #x); \
    } while (0)

extern "C" const char *nbx_last_error(void) { return g_err.c_str(); }
extern "C" int nbx_host_fail(int code, const char *msg) { return fail(code, msg); }

// Temporary device buffers of one host call, freed when the call returns on
// any path (the HIPCK early returns included).  Constructed with a stream, it
// allocates and frees in that stream's order (hipMallocAsync / hipFreeAsync).
struct DevBufs {
    std::vector<void *> held;
    bool ordered = false;  // stream-ordered allocations on st
    hipStream_t st = nullptr;
    DevBufs() = default;
    explicit DevBufs(hipStream_t s) : ordered(true), st(s) {}
    DevBufs(const DevBufs &) = delete;
    DevBufs &operator=(const DevBufs &) = delete;
    template <class T>
    hipError_t alloc(T **ptr, size_t bytes) {
        *ptr = nullptr;
        const size_t b = bytes > 0 ? bytes : 1;
        hipError_t e = ordered ? hipMallocAsync((void **)ptr, b, st) : hipMalloc((void **)ptr, b);
        if (e == hipSuccess) held.push_back((void *)*ptr);
        else *ptr = nullptr;
        return e;
    }
    ~DevBufs() {
        for (void *x : held) {
            if (ordered) hipFreeAsync(x, st);
            else hipFree(x);
        }
    }
};

// Device memory handed out from one caller-owned block (the no-alloc entry
// point's scratch): every carve is 256-byte aligned; carve() returns null
// once the block is exhausted.
struct Workspace {
    char *base;
    size_t cap, used;
    template <class T>
    T *carve(size_t bytes) {
        const size_t a = (used + 255) & ~(size_t)255;
        if (a + bytes > cap) return nullptr;
        used = a + bytes;
        return (T *)(base + a);
    }
};

// ------------------------------------------------------------------ profiling
struct ProfRec {
    std::string name;
    hipEvent_t e0, e1;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof_pending;
struct ProfAcc {
    std::string name;
    double ms;
    int64_t n;
};
static std::vector<ProfAcc> g_prof_acc;

ProfScope::ProfScope(const char *nm, hipStream_t s) : st(s), e0(nullptr), e1(nullptr), name(nm), on(g_prof_on) {
    if (!on) return;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, st);
}
ProfScope::~ProfScope() {
    if (!on) return;
    hipEventRecord(e1, st);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_pending.push_back({name, e0, e1});
}
static void prof_drain() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &r : g_prof_pending) {
        hipEventSynchronize(r.e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.e0, r.e1);
        bool found = false;
        for (auto &a : g_prof_acc)
            if (a.name == r.name) { a.ms += ms; a.n += 1; found = true; break; }
        if (!found) g_prof_acc.push_back({r.name, (double)ms, 1});
        hipEventDestroy(r.e0);
        hipEventDestroy(r.e1);
    }
    g_prof_pending.clear();
}
extern "C" void nbx_prof_enable(int on) { g_prof_on = on != 0; }
extern "C" void nbx_prof_reset(void) {
    prof_drain();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_acc.clear();
}
extern "C" int nbx_prof_get(const char *name, double *ms, int64_t *launches) {
    prof_drain();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    double t = 0;
    int64_t c = 0;
    for (auto &a : g_prof_acc)
        if (strstr(a.name.c_str(), name)) { t += a.ms; c += a.n; }
    if (ms) *ms = t;
    if (launches) *launches = c;
    return NBX_OK;
}

// ------------------------------------------------------------------ tree
struct nbx_tree {
    nbx_params p;
    int64_t n;
    int nd;        // tree dimensionality (3 or 6)
    bool has_vel;  // TPHYS tree that also keeps sorted velocities (x[3..5])
    bool tvel;     // TVEL: the tree's coordinates are the velocities
    double *xs[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double *mass = nullptr;
    int32_t *perm = nullptr;
    int32_t *inv = nullptr;
    double *box = nullptr;
    float *boxf = nullptr;  // precision form: outward FP32 node boxes (null when coordinates exceed 1e30)
    float *plane = nullptr;
    double *org = nullptr;
    float *xf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float4 *xq = nullptr;  // precision form, 3-D: FP32 offsets and their squared norms, interleaved
    double *kern = nullptr;
    float *kernf = nullptr;
    double *ext = nullptr;
    std::atomic<int64_t> last_fallbacks{0};  // set by every precision-form search, from any thread
    int kernres;
    double kernnorm;
    std::vector<double> hkern;
    int32_t nleaf, nleafp, depth;
    int64_t nnodes;
    // Device scratch kept from one call to the next (densities by slot, the
    // fallback list, the search scratch, FOF's work arrays), so a repeated
    // call allocates nothing: grown when a call needs more, freed with the
    // tree.  One call holds it at a time (scr_mu); scr_done is recorded on
    // the holder's stream after its last use, and the next holder's stream
    // waits for it.
    std::mutex scr_mu;
    char *scr = nullptr;
    size_t scr_cap = 0;
    hipEvent_t scr_done = nullptr;
    bool scr_recorded = false;
    DevTree view() const {
        DevTree T;
        memset(&T, 0, sizeof(T));
        T.n = (int32_t)n;
        T.nleaf = nleaf;
        T.nd = nd;
        for (int d = 0; d < 6; ++d) { T.x[d] = xs[d]; T.xf[d] = xf[d]; }
        T.xq = xq;
        T.mass = mass;
        T.perm = perm;
        T.inv = inv;
        T.box = box;
        T.boxf = boxf;
        T.plane = plane;
        T.org = org;
        T.kern = kern;
        T.kernf = kernf;
        T.kernres = kernres;
        T.nleafp = nleafp;
        T.depth = depth;
        return T;
    }
};

extern "C" void nbx_params_default(nbx_params *p) {
    memset(p, 0, sizeof(*p));
    p->bucket_size = 16;
    p->treetype = NBX_TPHYS;
    p->kerntype = NBX_KEPAN;
    p->kernres = 1000;
    p->splitting = 0;
    p->aniso = 0;
    p->scalespace = 0;
    p->periodic = 0;
    p->form = NBX_FORM_PLAIN;
    p->recheck_rel = 1e-5;
}

// SmoothingKernels.h:32-56 and KernelConstruction (KDTree.cxx:1144-1183)
static double kw_sph(double r, double h) {
    double rh = r / h;
    if (rh <= 1.0) return (1.0 - 0.75 * (2.0 - rh) * (rh * rh));
    else if (rh > 1.0 && rh <= 2.0) return 0.25 * (2.0 - rh) * (2.0 - rh) * (2.0 - rh);
    return 0;
}
static double kw_gauss(double r, double h) { double rh = r / h * 2.0; return exp(-0.5 * pow(rh, 2.0)); }
static double kw_epan(double r, double h) { double rh = r / h * 0.5; return rh <= 1.0 ? (1.0 - rh * rh) : 0.; }
static double kw_th(double r, double h) { double rh = r / h * 0.5; return (rh <= 1.0); }
static void make_kernel(int nd, int type, int &res, double &norm, std::vector<double> &tab) {
    if (res < 100) res = 100;
    norm = 1.0 / nd * pow(M_1_PI, nd / 2.) * tgamma(nd / 2. + 1.);
    double (*f)(double, double) = kw_sph;
    if (type == NBX_KGAUSS) { f = kw_gauss; norm *= pow(0.5 * M_1_PI, nd / 2.); }
    else if (type == NBX_KEPAN) { f = kw_epan; norm *= nd * (nd + 2.) * pow(0.5, nd + 1.); }
    else if (type == NBX_KTH) { f = kw_th; }
    else { norm *= nd * (nd + 1.) * (nd + 2.) * (nd + 3.) / (6 * (pow(2., nd + 1) - 1.)); }
    double delta = 2.0 / (double)(res - 1);
    tab.resize(res);
    for (int i = 0; i < res; i++) tab[i] = norm * f(i * delta, 1.0);
}

static inline int nblk(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

static void free_tree(nbx_tree *t) {
    if (!t) return;
    for (int d = 0; d < 6; ++d) { if (t->xs[d]) hipFree(t->xs[d]); if (t->xf[d]) hipFree(t->xf[d]); }
    if (t->mass) hipFree(t->mass);
    if (t->perm) hipFree(t->perm);
    if (t->inv) hipFree(t->inv);
    if (t->box) hipFree(t->box);
    if (t->boxf) hipFree(t->boxf);
    if (t->plane) hipFree(t->plane);
    if (t->org) hipFree(t->org);
    if (t->kern) hipFree(t->kern);
    if (t->kernf) hipFree(t->kernf);
    if (t->ext) hipFree(t->ext);
    if (t->xq) hipFree(t->xq);
    if (t->scr) hipFree(t->scr);
    if (t->scr_done) hipEventDestroy(t->scr_done);
    delete t;
}

static int build_tree(nbx_tree *t, const double *pos, const double *vel, const double *mass, hipStream_t st) {
    const int64_t n = t->n;
    const int nd = t->nd;
    const int ncoord = (nd == 6 || t->has_vel) ? 6 : 3;
    const double *c[6];
    for (int d = 0; d < 3; ++d) c[d] = pos + d * n;
    for (int d = 0; d < 3; ++d) c[3 + d] = vel ? vel + d * n : nullptr;

    // coordinates must be finite (a NaN or inf has no place in a split)
    {
        const int nb = 1024;
        double *part = nullptr;
        DevBufs pb(st);
        HIPCK(pb.alloc(&part, sizeof(double) * 2 * nb * nd));
        for (int d = 0; d < nd; ++d)
            hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, st, c[d], n, part + 2 * nb * d);
        std::vector<double> hp(2 * nb * nd);
        HIPCK(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        for (int d = 0; d < nd; ++d) {
            double mn = INFINITY, mx = -INFINITY;
            for (int b = 0; b < nb; ++b) { mn = std::min(mn, hp[2 * nb * d + 2 * b]); mx = std::max(mx, hp[2 * nb * d + 2 * b + 1]); }
            if (!std::isfinite(mn) || !std::isfinite(mx)) return fail(NBX_ERR_ARG, "non-finite coordinates");
        }
    }
    // implicit heap-ordered tree over the leaves padded to a power of two
    t->nleaf = (int32_t)((n + NBX_LEAF - 1) / NBX_LEAF);
    t->nleafp = 1;
    t->depth = 0;
    while (t->nleafp < t->nleaf) { t->nleafp <<= 1; t->depth++; }
    t->nnodes = 2 * (int64_t)t->nleafp;
    int32_t *order = nullptr;
    DevBufs ob(st);  // order is freed (stream-ordered) when build_tree returns
    HIPCK(ob.alloc(&order, sizeof(int32_t) * n));
    {
        ProfScope ps("k_build_kdorder", st);
        HIPCK(nbx_kd_order(c, nd, n, t->nleafp, order, st));
    }
    const int64_t npad = (n + NBX_LEAF - 1) / NBX_LEAF * NBX_LEAF;
    for (int d = 0; d < ncoord; ++d) HIPCK(hipMalloc((void **)&t->xs[d], sizeof(double) * npad));
    HIPCK(hipMalloc((void **)&t->mass, sizeof(double) * npad));
    HIPCK(hipMalloc((void **)&t->perm, sizeof(int32_t) * npad));
    HIPCK(hipMalloc((void **)&t->inv, sizeof(int32_t) * n));
    HIPCK(hipMalloc((void **)&t->box, sizeof(double) * 2 * nd * (size_t)t->nnodes));
    HIPCK(hipMalloc((void **)&t->plane, sizeof(float) * (nd == 6 ? 8 : 4) * (size_t)t->nleafp));
    if (t->p.form == NBX_FORM_MIXED) {
        HIPCK(hipMalloc((void **)&t->org, sizeof(double) * nd * (size_t)t->nleaf));
        HIPCK(hipMalloc((void **)&t->ext, sizeof(double) * (size_t)t->nleaf));
        if (nd == 3) HIPCK(hipMalloc((void **)&t->xq, sizeof(float4) * npad));
        else for (int d = 0; d < nd; ++d) HIPCK(hipMalloc((void **)&t->xf[d], sizeof(float) * npad));
    }
    {
        ProfScope ps("k_build_gather", st);
        GatherArgs g;
        memset(&g, 0, sizeof(g));
        g.n = n;
        g.npad = npad;
        g.ncoord = ncoord;
        g.nd = nd;
        g.nleaf = t->nleaf;
        g.nleafp = t->nleafp;
        g.order = order;
        for (int d = 0; d < ncoord; ++d) { g.in[d] = c[d]; g.out[d] = t->xs[d]; }
        g.mass_in = mass;
        g.mass_default = 1.0 / (double)n;
        g.mass_out = t->mass;
        g.perm = t->perm;
        g.inv = t->inv;
        g.box = t->box;
        g.org = t->org;
        g.ext = t->ext;
        for (int d = 0; d < nd; ++d) g.xf[d] = t->xf[d];
        g.xq = t->xq;
        hipLaunchKernelGGL(k_gather_leaf, dim3(nblk(npad)), dim3(256), 0, st, g);
    }
    {
        ProfScope ps("k_build_boxes", st);
        if (t->nleafp > t->nleaf)
            hipLaunchKernelGGL(k_leafbox_pad, dim3(nblk(t->nleafp - t->nleaf)), dim3(256), 0, st, t->box, t->nleaf,
                               t->nleafp, nd);
        for (int dep = t->depth - 1; dep >= 0; --dep)
            hipLaunchKernelGGL(k_levelbox, dim3(nblk((int64_t)1 << dep)), dim3(256), 0, st, t->box, t->plane,
                               1 << dep, 1 << dep, nd);
    }
    HIPCK(hipGetLastError());
    if (t->p.form == NBX_FORM_MIXED) {
        // FP32 node boxes for the precision form's walk, when every
        // coordinate lies within 1e30 (the bound its FP32 node test assumes;
        // otherwise the precision-form calls run the FP64 kernel)
        double rb[12];
        HIPCK(hipMemcpyAsync(rb, t->box + 2 * nd, sizeof(double) * 2 * nd, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        bool fits = true;
        for (int d = 0; d < 2 * nd; ++d) fits = fits && fabs(rb[d]) <= 1e30;
        if (fits) {
            const int64_t cnt = 2 * nd * (int64_t)t->nnodes;
            HIPCK(hipMalloc((void **)&t->boxf, sizeof(float) * cnt));
            hipLaunchKernelGGL(k_boxf, dim3(nblk(cnt)), dim3(256), 0, st, t->box, t->boxf, cnt, nd);
            HIPCK(hipGetLastError());
        }
    }
    // kernel table
    make_kernel(nd, t->p.kerntype, t->kernres, t->kernnorm, t->hkern);
    HIPCK(hipMalloc((void **)&t->kern, sizeof(double) * t->kernres));
    HIPCK(hipMemcpyAsync(t->kern, t->hkern.data(), sizeof(double) * t->kernres, hipMemcpyHostToDevice, st));
    {
        std::vector<float> kf(t->hkern.begin(), t->hkern.end());
        HIPCK(hipMalloc((void **)&t->kernf, sizeof(float) * t->kernres));
        HIPCK(hipMemcpyAsync(t->kernf, kf.data(), sizeof(float) * t->kernres, hipMemcpyHostToDevice, st));
        HIPCK(hipStreamSynchronize(st));
    }
    HIPCK(hipStreamSynchronize(st));
    return NBX_OK;
}

static int check_params(const nbx_params *p, int64_t n, bool has_vel) {
    if (!p) return fail(NBX_ERR_ARG, "params is NULL");
    if (n < 1 || n > 0x7fffffffLL - 4096) return fail(NBX_ERR_ARG, "number of particles out of range");
    if (p->treetype != NBX_TPHYS && p->treetype != NBX_TVEL && p->treetype != NBX_TPHS)
        return fail(NBX_ERR_UNSUPPORTED, "Error in type of tree specified (TPHYS, TVEL or TPHS supported)");
    if (p->splitting != 0) return fail(NBX_ERR_UNSUPPORTED, "only the spread splitting criterion is supported");
    if (p->scalespace != 0) return fail(NBX_ERR_UNSUPPORTED, "ScaleSpace is not supported; prescale coordinates");
    if (p->treetype == NBX_TPHS && p->aniso != -1)
        return fail(NBX_ERR_UNSUPPORTED, "TPHS trees need aniso = -1 (metric factor 1 on prescaled coordinates)");
    if (p->treetype == NBX_TPHS && !has_vel) return fail(NBX_ERR_ARG, "TPHS tree needs velocities");
    if (p->treetype == NBX_TVEL && !has_vel) return fail(NBX_ERR_ARG, "TVEL tree needs velocities");
    if (p->treetype == NBX_TVEL && p->periodic)
        return fail(NBX_ERR_UNSUPPORTED, "periodic TVEL trees are not supported (the period applies to positions)");
    if (p->periodic)
        for (int k = 0; k < 3; ++k)
            if (!(p->period[k] > 0)) return fail(NBX_ERR_ARG, "period must be positive");
    if (p->form != NBX_FORM_PLAIN && p->form != NBX_FORM_MIXED) return fail(NBX_ERR_ARG, "unknown form");
    return NBX_OK;
}

extern "C" int nbx_tree_create_dev(const nbx_params *p, const double *pos, const double *vel, const double *mass,
                                   int64_t n, void *stream, nbx_tree **out) {
    if (!out || !p) return fail(NBX_ERR_ARG, "NULL argument");
    const bool tvel = p->treetype == NBX_TVEL;
    if (!(tvel ? vel : pos)) return fail(NBX_ERR_ARG, "NULL argument");
    int rc = check_params(p, n, vel != nullptr);
    if (rc) return rc;
    nbx_tree *t = new nbx_tree();
    t->p = *p;
    if (t->p.recheck_rel <= 0) t->p.recheck_rel = 1e-5;
    t->n = n;
    t->nd = p->treetype == NBX_TPHS ? 6 : 3;
    t->tvel = tvel;
    t->has_vel = (t->nd == 3 && !tvel && vel != nullptr);
    t->kernres = p->kernres;
    // a TVEL tree is the 3-D tree over the velocities
    rc = tvel ? build_tree(t, vel, nullptr, mass, (hipStream_t)stream)
              : build_tree(t, pos, vel, mass, (hipStream_t)stream);
    if (rc) { free_tree(t); return rc; }
    *out = t;
    return NBX_OK;
}

extern "C" int nbx_tree_create(const nbx_params *p, const double *pos, const double *vel, const double *mass, int64_t n,
                               nbx_tree **out) {
    if (!out || !p) return fail(NBX_ERR_ARG, "NULL argument");
    if (!(p->treetype == NBX_TVEL ? vel : pos)) return fail(NBX_ERR_ARG, "NULL argument");
    int rc0 = check_params(p, n, vel != nullptr);
    if (rc0) return rc0;
    double *dpos = nullptr, *dvel = nullptr, *dm = nullptr;
    DevBufs bufs;
    if (pos && p->treetype != NBX_TVEL) {
        HIPCK(bufs.alloc(&dpos, sizeof(double) * 3 * n));
        HIPCK(hipMemcpy(dpos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    }
    if (vel) {
        HIPCK(bufs.alloc(&dvel, sizeof(double) * 3 * n));
        HIPCK(hipMemcpy(dvel, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    }
    if (mass) {
        HIPCK(bufs.alloc(&dm, sizeof(double) * n));
        HIPCK(hipMemcpy(dm, mass, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    return nbx_tree_create_dev(p, dpos, dvel, dm, n, nullptr, out);
}

extern "C" int nbx_tree_destroy(nbx_tree *t) {
    if (!t) return NBX_OK;
    free_tree(t);
    return NBX_OK;
}
extern "C" int64_t nbx_tree_size(const nbx_tree *t) { return t ? t->n : 0; }
extern "C" int nbx_tree_order(nbx_tree *t, int32_t *slot_to_id) {
    if (!t || !slot_to_id) return fail(NBX_ERR_ARG, "NULL argument");
    HIPCK(hipMemcpy(slot_to_id, t->perm, sizeof(int32_t) * t->n, hipMemcpyDeviceToHost));
    return NBX_OK;
}
extern "C" int nbx_tree_kernel(const nbx_tree *t, double *kernnorm, double *table, int32_t cap) {
    if (!t) return fail(NBX_ERR_ARG, "NULL tree");
    if (kernnorm) *kernnorm = t->kernnorm;
    if (table) memcpy(table, t->hkern.data(), sizeof(double) * std::min<int64_t>(cap, t->kernres));
    return t->kernres;
}

// The tree's kept scratch, held for one call as a Workspace of at least need
// bytes.  A larger need replaces the block (after the work that used it has
// finished).  On destruction the holder's stream records that its work on
// the block is queued.  A call that already holds the lease passes its
// Workspace down instead of taking another (the lock is not recursive).
struct ScratchLease {
    nbx_tree *t;
    hipStream_t st;
    std::unique_lock<std::mutex> lk;
    Workspace ws = {nullptr, 0, 0};
    hipError_t err = hipSuccess;
    ScratchLease(nbx_tree *tr, size_t need, hipStream_t s) : t(tr), st(s), lk(tr->scr_mu) {
        if (!t->scr_done) {
            err = hipEventCreateWithFlags(&t->scr_done, hipEventDisableTiming);
            if (err != hipSuccess) { t->scr_done = nullptr; return; }
        }
        if (need > t->scr_cap) {
            if (t->scr) {
                if (t->scr_recorded) err = hipEventSynchronize(t->scr_done);
                if (err != hipSuccess) return;
                hipFree(t->scr);
                t->scr = nullptr;
                t->scr_cap = 0;
                t->scr_recorded = false;
            }
            err = hipMalloc((void **)&t->scr, need);
            if (err != hipSuccess) { t->scr = nullptr; return; }
            t->scr_cap = need;
        }
        if (t->scr_recorded) err = hipStreamWaitEvent(st, t->scr_done, 0);
        if (err == hipSuccess) ws = {t->scr, t->scr_cap, 0};
    }
    ~ScratchLease() {
        if (err == hipSuccess && hipEventRecord(t->scr_done, st) == hipSuccess) t->scr_recorded = true;
    }
    ScratchLease(const ScratchLease &) = delete;
    ScratchLease &operator=(const ScratchLease &) = delete;
};

// Bytes the fused kNN + density pass carves, in its order: densities by slot,
// the error flag and, for the precision form, the fallback list and the
// search scratch.
static size_t density_ws_bytes(const nbx_tree *t, int32_t k, bool mixed) {
    Workspace dry = {nullptr, ~(size_t)0, 0};
    dry.carve<char>(sizeof(double) * t->n);
    dry.carve<char>(sizeof(int32_t));
    if (mixed) {
        dry.carve<char>(sizeof(int32_t) * (t->n + 1));
        dry.carve<char>(nbx_mixed_scratch_bytes(t->nd, k + 1));
    }
    return dry.used;
}

// ------------------------------------------------------------------ precision form
// FP32 search with the FP64 certificate; queries whose certificate fails are
// re-run by the FP64 kernel from a list, so results equal the plain form's
// neighbour sets.  A.dens / A.rho / A.nn / A.layout as for nbx_launch_knn.
// The fallback list and the search scratch come from ws, or from the tree's
// kept scratch when ws is null.
static hipError_t run_mixed(nbx_tree *t, KnnArgs A, int k, hipStream_t st, Workspace *ws = nullptr) {
    const int64_t n = t->n;
    // the fallback list holds at most the queries of the launch
    const int64_t nfb = A.wcnt > 0 ? std::min<int64_t>(n, (int64_t)A.wcnt * NBX_WAVE) : n;
    const size_t mbytes = nbx_mixed_scratch_bytes(t->nd, k + 1);
    std::unique_ptr<ScratchLease> lease;
    if (!ws) {
        Workspace dry = {nullptr, ~(size_t)0, 0};
        dry.carve<char>(sizeof(int32_t) * (nfb + 1));
        dry.carve<char>(mbytes);
        lease.reset(new ScratchLease(t, dry.used, st));
        if (lease->err != hipSuccess) return lease->err;
        ws = &lease->ws;
    }
    int32_t *fb = ws->carve<int32_t>(sizeof(int32_t) * (nfb + 1));
    void *mwork = ws->carve<void>(mbytes);
    if (!fb || !mwork) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    int32_t *cnt = fb + nfb;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, cnt, 0, (int64_t)1);
    KnnArgs M = A;
    M.kh = k + 1;
    M.kout = k;
    M.band_rel = t->p.recheck_rel;
    DevTree T = t->view();
    e = nbx_launch_knn_mixed(T, M, t->ext, fb, cnt, st, mwork, mbytes);
    int32_t hc = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&hc, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && hc > 0) {
        KnnArgs F = A;
        F.qsrc = NBX_Q_SLOTS;
        F.kh = k;
        F.kout = k;
        F.excl_zero = 1;
        F.qlist = fb;
        F.m = hc;
        F.wbeg = F.wcnt = 0;
        e = nbx_launch_knn(T, F, hc, st);
    }
    t->last_fallbacks.store(hc);
    return e;
}

extern "C" int nbx_tree_last_fallbacks(const nbx_tree *t, int64_t *fallbacks) {
    if (!t || !fallbacks) return fail(NBX_ERR_ARG, "NULL argument");
    *fallbacks = t->last_fallbacks.load();
    return NBX_OK;
}

// ------------------------------------------------------------------ kNN
// Largest heap (kept entries per query) the FP64 kernel holds: k + 1 for
// the periodic searches, so k <= 128 there and k <= 129 otherwise.  The
// precision-form kernel keeps k + 1 entries up to 65 (k <= 64); larger k
// run on the FP64 kernel, whose neighbour sets the precision form equals.
static int heap_cap(const nbx_tree *) { return 129; }
static bool mixed_ok(const nbx_tree *t, int32_t k) {
    return t->p.form == NBX_FORM_MIXED && k + 1 <= 65 && t->boxf != nullptr;
}

static int setup_knn(nbx_tree *t, int32_t k, int32_t variant, KnnArgs &A, int &nq) {
    memset(&A, 0, sizeof(A));
    if (!t) return fail(NBX_ERR_ARG, "NULL tree");
    if (k < 1) return fail(NBX_ERR_ARG, "k must be >= 1");
    const bool per = t->p.periodic != 0;
    const bool rootleaf = t->n <= t->p.bucket_size;  // the reference's root is a leaf
    A.excl_zero = 1;
    A.kout = k;
    nq = (int)t->n;
    switch (variant) {
        case NBX_KNN_FINDNEAREST:
        case NBX_KNN_FINDNEARESTPHASE:
            if (variant == NBX_KNN_FINDNEARESTPHASE && t->nd != 6)
                return fail(NBX_ERR_UNSUPPORTED, "FindNearestPhase needs a TPHS tree");
            if (!per) { A.qsrc = NBX_Q_SLOTS; A.kh = k; A.drop = NBX_DROP_NONE; }
            else if (rootleaf && t->nd == 3) { A.qsrc = NBX_Q_SLOTS; A.kh = k + 1; A.drop = NBX_DROP_SMALLEST; }
            else { A.qsrc = NBX_Q_PERIODIC; A.kh = k + 1; A.drop = NBX_DROP_SMALLEST; A.periodic = !rootleaf; }
            break;
        case NBX_KNN_FINDNEARESTPOS:
            if (t->nd != 3 || t->tvel) return fail(NBX_ERR_UNSUPPORTED, "FindNearestPos needs a TPHYS tree");
            if (!per) { A.qsrc = NBX_Q_SLOTS; A.kh = k; A.drop = NBX_DROP_NONE; }
            else if (rootleaf) { A.qsrc = NBX_Q_SLOTS; A.kh = k + 1; A.drop = NBX_DROP_LARGEST; }
            else { A.qsrc = NBX_Q_PERIODIC; A.kh = k + 1; A.drop = NBX_DROP_LARGEST; A.periodic = 1; }
            break;
        case NBX_KNN_NONPERIODIC:
            A.qsrc = NBX_Q_SLOTS; A.kh = k; A.drop = NBX_DROP_NONE;
            break;
        default:
            return fail(NBX_ERR_ARG, "unknown kNN variant");
    }
    if (A.kh > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "k too large (k <= 128 supported)");
    if (A.periodic) for (int d = 0; d < 3; ++d) A.period[d] = t->p.period[d];
    return NBX_OK;
}

extern "C" int nbx_knn_dev(nbx_tree *t, int32_t k, int32_t variant, int32_t layout, int32_t *nn_dev, double *d2_dev,
                           void *stream) {
    KnnArgs A;
    int nq;
    int rc = setup_knn(t, k, variant, A, nq);
    if (rc) return rc;
    if (layout != NBX_LAYOUT_BY_ID && layout != NBX_LAYOUT_BY_SLOT) return fail(NBX_ERR_ARG, "unknown layout");
    A.layout = layout;
    A.nn = nn_dev;
    A.d2 = d2_dev;
    hipStream_t st = (hipStream_t)stream;
    if (mixed_ok(t, k) && A.qsrc == NBX_Q_SLOTS && A.kh == k) {
        // FP32 search; when distances are asked for, the rows' FP64 d2 and
        // their FP64 order come from a row pass over the selected entries
        DevBufs bufs(st);
        if (d2_dev && !A.nn) HIPCK(bufs.alloc(&A.nn, sizeof(int32_t) * (size_t)t->n * k));
        A.d2 = nullptr;
        HIPCK(run_mixed(t, A, k, st));
        if (d2_dev) HIPCK(nbx_launch_rows_fp64(t->view(), A.nn, d2_dev, k, layout, st));
        return NBX_OK;
    }
    HIPCK(nbx_launch_knn(t->view(), A, nq, st));
    return NBX_OK;
}

extern "C" int nbx_knn(nbx_tree *t, int32_t k, int32_t variant, int32_t layout, int32_t *nn, double *d2) {
    if (!t) return fail(NBX_ERR_ARG, "NULL tree");
    int32_t *dnn = nullptr;
    double *dd2 = nullptr;
    size_t cnt = (size_t)t->n * (size_t)k;
    DevBufs bufs;
    if (nn) HIPCK(bufs.alloc(&dnn, sizeof(int32_t) * cnt));
    if (d2) HIPCK(bufs.alloc(&dd2, sizeof(double) * cnt));
    int rc = nbx_knn_dev(t, k, variant, layout, dnn, dd2, nullptr);
    if (rc == NBX_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess && nn) e = hipMemcpy(nn, dnn, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost);
        if (e == hipSuccess && d2) e = hipMemcpy(d2, dd2, sizeof(double) * cnt, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(e, "nbx_knn");
    }
    return rc;
}

// FindNearest(tt) / FindNearestPos(tt) / FindNearestPhase(tt) for a list of
// particles at once: the batch variant's rules, rows in the list's order.
extern "C" int nbx_knn_targets(nbx_tree *t, const int32_t *ids, int64_t m, int32_t k, int32_t variant, int32_t layout,
                               int32_t *nn, double *d2) {
    KnnArgs A;
    int nq;
    int rc = setup_knn(t, k, variant, A, nq);
    if (rc) return rc;
    if (!nn || (m > 0 && !ids)) return fail(NBX_ERR_ARG, "NULL argument");
    if (layout != NBX_LAYOUT_BY_ID && layout != NBX_LAYOUT_BY_SLOT) return fail(NBX_ERR_ARG, "unknown layout");
    if (m < 0 || m > INT32_MAX) return fail(NBX_ERR_ARG, "number of targets out of range");
    if (m == 0) return NBX_OK;
    for (int64_t i = 0; i < m; ++i)
        if (ids[i] < 0 || ids[i] >= t->n) return fail(NBX_ERR_ARG, "target ID out of range");
    DevBufs bufs;
    int32_t *did = nullptr, *dsl = nullptr, *dnn = nullptr;
    double *dd2 = nullptr;
    HIPCK(bufs.alloc(&did, sizeof(int32_t) * m));
    HIPCK(bufs.alloc(&dsl, sizeof(int32_t) * m));
    HIPCK(bufs.alloc(&dnn, sizeof(int32_t) * m * k));
    if (d2) HIPCK(bufs.alloc(&dd2, sizeof(double) * m * k));
    HIPCK(hipMemcpy(did, ids, sizeof(int32_t) * m, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_ids_to_slots, dim3(nblk(m)), dim3(256), 0, nullptr, did, t->inv, m, dsl);
    A.layout = layout;
    A.nn = dnn;
    A.d2 = dd2;
    A.qlist = dsl;
    A.m = (int32_t)m;
    A.list_rows = 1;
    hipError_t e = nbx_launch_knn(t->view(), A, (int)m, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(nn, dnn, sizeof(int32_t) * m * k, hipMemcpyDeviceToHost);
    if (e == hipSuccess && d2) e = hipMemcpy(d2, dd2, sizeof(double) * m * k, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipfail(e, "nbx_knn_targets");
    return NBX_OK;
}

// Position queries through the plain kernel: kNN (nn/d2, nullable) and/or the
// gather density with h = r_k/2 (rho/h, nullable).  periodic: the image passes
// of the period-aware search.
// qv (velocities of the points) selects the velocity density over the k
// nearest positions, with nsmooth of them in velocity (else the gather SPH
// density when rho is given).
static int points_impl(nbx_tree *t, const double *q, int64_t m, int32_t k, bool periodic, int32_t *nn, double *d2,
                       double *rho, double *h, const char *what, const double *qv = nullptr, int32_t nsmooth = 0) {
    DevBufs bufs;
    if (m <= 0) return NBX_OK;
    if (m > INT32_MAX) return fail(NBX_ERR_ARG, "too many positions (m < 2^31)");
    double *dq = nullptr, *dqv = nullptr, *dd2 = nullptr, *drho = nullptr, *dh = nullptr;
    int32_t *dnn = nullptr, *err = nullptr;
    HIPCK(bufs.alloc(&dq, sizeof(double) * 3 * m));
    HIPCK(hipMemcpy(dq, q, sizeof(double) * 3 * m, hipMemcpyHostToDevice));
    if (qv) {
        HIPCK(bufs.alloc(&dqv, sizeof(double) * 3 * m));
        HIPCK(hipMemcpy(dqv, qv, sizeof(double) * 3 * m, hipMemcpyHostToDevice));
    }
    if (nn) HIPCK(bufs.alloc(&dnn, sizeof(int32_t) * m * k));
    if (d2) HIPCK(bufs.alloc(&dd2, sizeof(double) * m * k));
    if (rho) HIPCK(bufs.alloc(&drho, sizeof(double) * m));
    if (h) HIPCK(bufs.alloc(&dh, sizeof(double) * m));
    HIPCK(bufs.alloc(&err, sizeof(int32_t)));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, nullptr, err, 0, (int64_t)1);
    KnnArgs A;
    memset(&A, 0, sizeof(A));
    A.qsrc = NBX_Q_POINTS;
    A.kh = k;
    A.kout = k;
    A.m = (int32_t)m;
    A.periodic = periodic && t->p.periodic && t->n > t->p.bucket_size;
    for (int d = 0; d < 3; ++d) { A.period[d] = t->p.period[d]; A.q[d] = dq + d * m; }
    A.nn = dnn;
    A.d2 = dd2;
    A.dens = rho ? (qv ? NBX_DO_VEL : NBX_DO_GATHER) : NBX_DO_NONE;
    if (qv) {
        for (int d = 0; d < 3; ++d) A.qv[d] = dqv + d * m;
        A.nsmooth = nsmooth;
    }
    A.rho = drho;
    A.hout = dh;
    A.err = err;
    int32_t herr = 0;
    hipError_t e = nbx_launch_knn(t->view(), A, (int)m, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && nn) e = hipMemcpy(nn, dnn, sizeof(int32_t) * m * k, hipMemcpyDeviceToHost);
    if (e == hipSuccess && d2) e = hipMemcpy(d2, dd2, sizeof(double) * m * k, hipMemcpyDeviceToHost);
    if (e == hipSuccess && rho) e = hipMemcpy(rho, drho, sizeof(double) * m, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h) e = hipMemcpy(h, dh, sizeof(double) * m, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipfail(e, what);
    if (herr) return fail(NBX_ERR_SEARCH, qv ? "CalcVelDensityPosition failed: fewer than Nsearch particles"
                                             : "CalcDensityPosition failed: fewer than Nsmooth particles");
    return NBX_OK;
}

extern "C" int nbx_knn_points(nbx_tree *t, const double *q, int64_t m, int32_t k, int32_t *nn, double *d2) {
    if (!t || !q || !nn) return fail(NBX_ERR_ARG, "NULL argument");
    if (t->nd != 3) return fail(NBX_ERR_UNSUPPORTED, "position queries need a TPHYS tree");
    if (k < 1 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "k out of range (1..129)");
    return points_impl(t, q, m, k, true, nn, d2, nullptr, nullptr, "nbx_knn_points");
}

// FindNearestPhase(x, v) (KDFindNearest.cxx:528-538) and, on a TPHS tree,
// FindNearest(x) (:462-509): the 6-D search from phase points, no exclusion,
// every periodic image taken (KDSplitNode.cxx:1153-1179)
extern "C" int nbx_knn_phase_points(nbx_tree *t, const double *x, const double *v, int64_t m, int32_t k, int32_t *nn,
                                    double *d2) {
    if (!t || !x || !v || !nn) return fail(NBX_ERR_ARG, "NULL argument");
    if (t->nd != 6) return fail(NBX_ERR_UNSUPPORTED, "phase-space point queries need a TPHS tree");
    if (k < 1 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "k out of range");
    return points_impl(t, x, m, k, true, nn, d2, nullptr, nullptr, "nbx_knn_phase_points", v, 0);
}

extern "C" int nbx_density_points(nbx_tree *t, const double *q, int64_t m, int32_t k, double *rho, double *h) {
    if (!t || !q || !rho) return fail(NBX_ERR_ARG, "NULL argument");
    if (t->nd != 3) return fail(NBX_ERR_UNSUPPORTED, "CalcDensityPosition here needs a TPHYS tree");
    if (k < 1 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "Nsmooth out of range (1..129)");
    if (k > t->n) return fail(NBX_ERR_SEARCH, "CalcDensityPosition failed: fewer than Nsmooth particles");
    return points_impl(t, q, m, k, false, nullptr, nullptr, rho, h, "nbx_density_points");
}

extern "C" int nbx_veldensity_points(nbx_tree *t, const double *x, const double *v, int64_t m, int32_t nsmooth,
                                     int32_t nsearch, double *rho, double *hv) {
    if (!t || !x || !v || !rho) return fail(NBX_ERR_ARG, "NULL argument");
    if (t->nd != 3 || !t->has_vel)
        return fail(NBX_ERR_UNSUPPORTED, "CalcVelDensityPosition needs a TPHYS tree built with velocities");
    if (nsmooth < 1 || nsearch < 1) return fail(NBX_ERR_ARG, "Nsmooth and Nsearch must be >= 1");
    if (nsearch > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "Nsearch out of range (1..129)");
    if (nsearch > t->n) return fail(NBX_ERR_SEARCH, "CalcVelDensityPosition failed: fewer than Nsearch particles");
    // Nsmooth > Nsearch is taken as Nsearch (KDCalcSmoothQuantities.cxx:1160-1163)
    return points_impl(t, x, m, nsearch, false, nullptr, nullptr, rho, hv, "nbx_veldensity_points", v,
                       nsmooth > nsearch ? nsearch : nsmooth);
}

// ------------------------------------------------------------------ density
// variant DENS_VELPHYS (internal): CalcVelDensityWithPhysDensity over the
// densities rho_in_id (device, by ID)
static const int32_t DENS_VELPHYS = 100;
static int density_impl(nbx_tree *t, int32_t k, int32_t variant, int32_t nsearch, double *rho_id, double *h_id,
                        int32_t *nn_slot, hipStream_t st, const double *rho_in_id = nullptr, Workspace *ws = nullptr) {
    if (!t) return fail(NBX_ERR_ARG, "NULL tree");
    if (k < 1) return fail(NBX_ERR_ARG, "k must be >= 1");
    KnnArgs A;
    memset(&A, 0, sizeof(A));
    A.qsrc = NBX_Q_SLOTS;
    A.excl_zero = 1;
    A.kh = k;
    A.kout = k;
    A.layout = NBX_LAYOUT_BY_SLOT;
    if (variant == NBX_DENS_SYMMETRIC) A.dens = NBX_DO_SYM;
    else if (variant == NBX_DENS_GATHER) A.dens = NBX_DO_GATHER;
    else if (variant == NBX_DENS_VEL || (variant == DENS_VELPHYS && rho_in_id)) {
        if (t->nd != 3 || !t->has_vel)
            return fail(NBX_ERR_UNSUPPORTED, "CalcVelDensity needs a TPHYS tree built with velocities");
        A.dens = variant == NBX_DENS_VEL ? NBX_DO_VEL : NBX_DO_VELPHYS;
        if (nsearch < 1) return fail(NBX_ERR_ARG, "Nsearch must be >= 1");
        A.kh = nsearch;
        A.nsmooth = k > nsearch ? nsearch : k;
    } else
        return fail(NBX_ERR_ARG, "unknown density variant");
    if (A.kh > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "k too large (k <= 129 supported)");
    const bool mixed = mixed_ok(t, k) && A.dens != NBX_DO_VEL && A.dens != NBX_DO_VELPHYS;
    const int64_t n = t->n;
    DevBufs bufs(st);
    // temporaries from the caller's workspace, else from the tree's kept scratch
    std::unique_ptr<ScratchLease> lease;
    if (ws) {
        if (A.dens == NBX_DO_VELPHYS) return fail(NBX_ERR_UNSUPPORTED, "no workspace form of this density");
    } else {
        lease.reset(new ScratchLease(t, density_ws_bytes(t, k, mixed), st));
        if (lease->err != hipSuccess) return hipfail(lease->err, "density scratch");
        ws = &lease->ws;
    }
    double *rho = ws->carve<double>(sizeof(double) * n);
    int32_t *err = ws->carve<int32_t>(sizeof(int32_t));
    if (!rho || !err) return fail(NBX_ERR_ARG, "scratch too small");
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n)), dim3(256), 0, st, rho, 0.0, n);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, err, 0, (int64_t)1);
    if (A.dens == NBX_DO_VELPHYS) {
        double *rin = nullptr;
        HIPCK(bufs.alloc(&rin, sizeof(double) * n));
        hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, rho_in_id, t->perm, n, rin);
        A.rho_in = rin;
    }
    A.rho = rho;
    A.hout = h_id;  // written by ID in the search
    A.hout_by_id = 1;
    A.err = err;
    A.nn = nn_slot;
    hipError_t e = mixed ? run_mixed(t, A, k, st, ws) : nbx_launch_knn(t->view(), A, (int)n, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_scatter_by_id, dim3(nblk(n)), dim3(256), 0, st, rho, t->perm, rho_id, n);
        e = hipGetLastError();
    }
    int32_t herr = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hipfail(e, "density");
    if (herr) return fail(NBX_ERR_SEARCH, "CalcDensity failed: fewer than k neighbours");
    return NBX_OK;
}

extern "C" int nbx_density_dev(nbx_tree *t, int32_t k, int32_t variant, int32_t nsearch, double *rho_by_id_dev,
                               double *h_by_id_dev, void *stream) {
    if (!rho_by_id_dev) return fail(NBX_ERR_ARG, "NULL output");
    return density_impl(t, k, variant, nsearch, rho_by_id_dev, h_by_id_dev, nullptr, (hipStream_t)stream);
}

extern "C" int nbx_knn_density_dev(nbx_tree *t, int32_t k, int32_t *nn_by_slot_dev, double *rho_by_id_dev,
                                   double *h_by_id_dev, void *stream) {
    if (!rho_by_id_dev) return fail(NBX_ERR_ARG, "NULL output");
    return density_impl(t, k, NBX_DENS_SYMMETRIC, 0, rho_by_id_dev, h_by_id_dev, nn_by_slot_dev, (hipStream_t)stream);
}

// ------------------------------------------------------------------ entry point that allocates nothing
// nbx_knn_density_noalloc: nbx_knn_density_dev with every temporary (rho by
// slot, the error flag, and for the precision form the fallback list,
// the per-workgroup search scratch and the queue counters) carved from one
// file-scope device array instead of allocated.  Each GPU holds its own copy
// of the array from the library's load until the process ends; host code
// reaches it only through the address hipGetSymbolAddress gives on that
// device.  Every call on a GPU shares that GPU's copy: it clears what must
// start at zero itself (fill kernels), and calls that overlap on one GPU race.
#define NBX_NOALLOC_BYTES ((size_t)256 << 20)
static __device__ __align__(256) unsigned char scratch[NBX_NOALLOC_BYTES];
#define NBX_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NBX_MAX_DEVICES];  // per device: the address of `scratch` there

static int noalloc_scratch(void **ptr) {
    int dev = 0;
    HIPCK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NBX_MAX_DEVICES) return fail(NBX_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return NBX_OK;
}

extern "C" int nbx_knn_density_noalloc(nbx_tree *t, int32_t k, int32_t *nn_by_slot_dev, double *rho_by_id_dev,
                                       double *h_by_id_dev, void *stream) {
    if (!t || !rho_by_id_dev) return fail(NBX_ERR_ARG, "NULL argument");
    if (k < 1 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "k out of range");
    // the carves density_impl and run_mixed will make
    const int64_t n = t->n;
    const size_t need = density_ws_bytes(t, k, mixed_ok(t, k));
    if (need > NBX_NOALLOC_BYTES)
        return fail(NBX_ERR_UNSUPPORTED, "nbx_knn_density_noalloc: " + std::to_string(n) + " particles need " +
                                             std::to_string(need >> 20) + " MiB of scratch; the maximum is " +
                                             std::to_string(NBX_NOALLOC_BYTES >> 20) + " MiB");
    void *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    Workspace ws = {(char *)base, NBX_NOALLOC_BYTES, 0};
    return density_impl(t, k, NBX_DENS_SYMMETRIC, 0, rho_by_id_dev, h_by_id_dev, nn_by_slot_dev, (hipStream_t)stream,
                        nullptr, &ws);
}

// Smoothed velocity mean, dispersion and moments: the fused kNN + CalcDensity
// pass gives neighbour rows by slot and rho, then the smoothing kernels run
// over those rows.  A caller's rho (the reference's densityset = 1), means
// (meanvelset = 1) or dispersions (veldispset = 1) replace computed ones.
// Host arrays by ID: rows of ncomp values -> [ncomp][n] by slot on the device.
static hipError_t upload_rows(const nbx_tree *t, const double *rows, int ncomp, double *tmp, double *by_slot) {
    const int64_t n = t->n;
    std::vector<double> soa((size_t)ncomp * n);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < ncomp; ++c) soa[(size_t)c * n + i] = rows[ncomp * i + c];
    hipError_t e = hipMemcpy(tmp, soa.data(), sizeof(double) * ncomp * n, hipMemcpyHostToDevice);
    for (int c = 0; c < ncomp && e == hipSuccess; ++c)
        hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, nullptr, tmp + c * n, t->perm, n,
                           by_slot + c * n);
    return e;
}

static int smooth_impl(nbx_tree *t, int32_t k, const double *rho_by_id, const double *mean_in, const double *disp_in,
                       double *mean_out, double *disp_out, double *skew_out, double *kurt_out) {
    DevBufs bufs;
    if (t->nd != 3 || !t->has_vel)
        return fail(NBX_ERR_UNSUPPORTED, "CalcSmoothVel needs a TPHYS tree built with velocities");
    if (k < 1 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "Nsmooth out of range (1..129)");
    const int64_t n = t->n;
    const bool moments = skew_out || kurt_out;
    const bool need_disp = disp_out || moments;
    hipStream_t st = nullptr;
    int32_t *nn = nullptr;
    double *rid = nullptr, *rslot = nullptr, *mean = nullptr, *disp = nullptr, *mom = nullptr, *buf = nullptr;
    HIPCK(bufs.alloc(&nn, sizeof(int32_t) * (size_t)n * k));
    HIPCK(bufs.alloc(&rid, sizeof(double) * n));
    HIPCK(bufs.alloc(&rslot, sizeof(double) * n));
    HIPCK(bufs.alloc(&mean, sizeof(double) * 3 * n));
    if (need_disp) HIPCK(bufs.alloc(&disp, sizeof(double) * 9 * n));
    if (moments) HIPCK(bufs.alloc(&mom, sizeof(double) * 6 * n));
    HIPCK(bufs.alloc(&buf, sizeof(double) * 9 * n));
    int rc = density_impl(t, k, NBX_DENS_SYMMETRIC, 0, rid, nullptr, nn, st);
    hipError_t e = hipSuccess;
    if (rc == NBX_OK) {
        if (rho_by_id) e = hipMemcpy(rid, rho_by_id, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, rid, t->perm, n, rslot);
        if (e == hipSuccess) {
            if (mean_in) e = upload_rows(t, mean_in, 3, buf, mean);
            else hipLaunchKernelGGL(k_fill_f64, dim3(nblk(3 * n)), dim3(256), 0, st, mean, 0.0, 3 * n);
        }
        if (e == hipSuccess && need_disp) {
            if (disp_in) e = upload_rows(t, disp_in, 9, buf, disp);
            else hipLaunchKernelGGL(k_fill_f64, dim3(nblk(9 * n)), dim3(256), 0, st, disp, 0.0, 9 * n);
        }
        if (e == hipSuccess && moments)
            hipLaunchKernelGGL(k_fill_f64, dim3(nblk(6 * n)), dim3(256), 0, st, mom, 0.0, 6 * n);
        if (e == hipSuccess) {
            SmoothArgs A = {nn, k, rslot, mean, disp, moments ? mom : nullptr, moments ? mom + 3 * n : nullptr};
            e = nbx_launch_smooth(t->view(), A, mean_in == nullptr, need_disp && disp_in == nullptr, st);
        }
        struct Out {
            double *dev;
            int ncomp;
            double *host;
        } outs[4] = {{mean, 3, mean_out}, {disp, 9, disp_out}, {mom, 3, skew_out}, {mom + 3 * n, 3, kurt_out}};
        for (const Out &o : outs) {
            if (e != hipSuccess || !o.host) continue;
            hipLaunchKernelGGL(k_smooth_out, dim3(nblk(n)), dim3(256), 0, st, o.dev, t->perm, n, o.ncomp, buf);
            e = hipMemcpy(o.host, buf, sizeof(double) * o.ncomp * n, hipMemcpyDeviceToHost);
        }
        if (e != hipSuccess) rc = hipfail(e, "smoothed velocities");
    }
    return rc;
}

extern "C" int nbx_smooth_vel(nbx_tree *t, int32_t k, const double *rho_by_id, const double *mean_in_by_id,
                              double *smvel_by_id, double *disp_by_id) {
    if (!t || (!smvel_by_id && !disp_by_id)) return fail(NBX_ERR_ARG, "NULL argument");
    return smooth_impl(t, k, rho_by_id, mean_in_by_id, nullptr, smvel_by_id, disp_by_id, nullptr, nullptr);
}

extern "C" int nbx_smooth_vel_moments(nbx_tree *t, int32_t k, const double *rho_by_id, const double *mean_in_by_id,
                                      const double *disp_in_by_id, double *skew_by_id, double *kurt_by_id) {
    if (!t || (!skew_by_id && !kurt_by_id)) return fail(NBX_ERR_ARG, "NULL argument");
    return smooth_impl(t, k, rho_by_id, mean_in_by_id, disp_in_by_id, nullptr, nullptr, skew_by_id, kurt_by_id);
}

// CalcSmoothLocalMeanParticle / CalcSmoothLocalDispParticle for every
// particle: CalcDensity's neighbour rows (fused pass), then the gather sums.
extern "C" int nbx_smooth_local(nbx_tree *t, int32_t k, const double *weight_by_id, const double *rho_by_id,
                                const double *mean_in_by_id, double *mean_by_id, double *disp_by_id) {
    DevBufs bufs;
    if (!t || !weight_by_id || (!mean_by_id && !disp_by_id)) return fail(NBX_ERR_ARG, "NULL argument");
    if (disp_by_id && !mean_in_by_id) return fail(NBX_ERR_ARG, "the dispersion needs the means (mean_in_by_id)");
    if (t->nd != 3 || t->tvel) return fail(NBX_ERR_UNSUPPORTED, "CalcSmoothLocalMean here needs a TPHYS tree");
    if (k < 1 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "Nsmooth out of range (1..129)");
    const int64_t n = t->n;
    hipStream_t st = nullptr;
    int32_t *nn = nullptr;
    double *b[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    HIPCK(bufs.alloc(&nn, sizeof(int32_t) * (size_t)n * k));
    for (int i = 0; i < 8; ++i) HIPCK(bufs.alloc(&b[i], sizeof(double) * n));
    double *rid = b[0], *rslot = b[1], *wid = b[2], *wslot = b[3], *mid = b[4], *mslot = b[5], *om = b[6], *od = b[7];
    int rc = density_impl(t, k, NBX_DENS_SYMMETRIC, 0, rid, nullptr, nn, st);
    hipError_t e = hipSuccess;
    if (rc == NBX_OK) {
        if (rho_by_id) e = hipMemcpy(rid, rho_by_id, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(wid, weight_by_id, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess && mean_in_by_id) e = hipMemcpy(mid, mean_in_by_id, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, rid, t->perm, n, rslot);
            hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, wid, t->perm, n, wslot);
            if (mean_in_by_id) hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, mid, t->perm, n, mslot);
            LocalArgs A = {{nullptr, nullptr, nullptr}, n, nn, k, 0, rslot, wslot, mean_in_by_id ? mslot : nullptr,
                           om, od};
            e = nbx_launch_smooth_local(t->view(), A, st);
        }
        if (e == hipSuccess && mean_by_id) {
            hipLaunchKernelGGL(k_scatter_by_id, dim3(nblk(n)), dim3(256), 0, st, om, t->perm, rid, n);
            e = hipMemcpy(mean_by_id, rid, sizeof(double) * n, hipMemcpyDeviceToHost);
        }
        if (e == hipSuccess && disp_by_id) {
            hipLaunchKernelGGL(k_scatter_by_id, dim3(nblk(n)), dim3(256), 0, st, od, t->perm, rid, n);
            e = hipMemcpy(disp_by_id, rid, sizeof(double) * n, hipMemcpyDeviceToHost);
        }
        if (e != hipSuccess) rc = hipfail(e, "nbx_smooth_local");
    }
    return rc;
}

// CalcSmoothLocalMeanPosition / CalcSmoothLocalDispPosition for m positions:
// the node search from x (non-periodic, no exclusion) and the gather sums.
extern "C" int nbx_smooth_local_points(nbx_tree *t, const double *q, int64_t m, int32_t k, const double *weight_by_id,
                                       const double *rho_by_id, const double *mean_in, double *mean_out,
                                       double *disp_out) {
    DevBufs bufs;
    if (!t || !q || !weight_by_id || (!mean_out && !disp_out)) return fail(NBX_ERR_ARG, "NULL argument");
    if (disp_out && !mean_in) return fail(NBX_ERR_ARG, "the dispersion needs the means (mean_in)");
    if (t->nd != 3 || t->tvel)
        return fail(NBX_ERR_UNSUPPORTED, "CalcSmoothLocalMeanPosition here needs a TPHYS tree");
    if (k < 1 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "Nsmooth out of range (1..129)");
    if (k > t->n) return fail(NBX_ERR_SEARCH, "CalcSmoothLocalMean failed: fewer than Nsmooth particles");
    if (m <= 0) return NBX_OK;
    if (m > INT32_MAX) return fail(NBX_ERR_ARG, "too many positions (m < 2^31)");
    const int64_t n = t->n;
    hipStream_t st = nullptr;
    double *dq = nullptr, *rid = nullptr, *rslot = nullptr, *wid = nullptr, *wslot = nullptr, *dmi = nullptr,
           *dm = nullptr, *dd = nullptr;
    int32_t *dnn = nullptr;
    HIPCK(bufs.alloc(&dq, sizeof(double) * 3 * m));
    HIPCK(hipMemcpy(dq, q, sizeof(double) * 3 * m, hipMemcpyHostToDevice));
    HIPCK(bufs.alloc(&dnn, sizeof(int32_t) * m * k));
    HIPCK(bufs.alloc(&rid, sizeof(double) * n));
    HIPCK(bufs.alloc(&rslot, sizeof(double) * n));
    HIPCK(bufs.alloc(&wid, sizeof(double) * n));
    HIPCK(bufs.alloc(&wslot, sizeof(double) * n));
    HIPCK(bufs.alloc(&dm, sizeof(double) * m));
    HIPCK(bufs.alloc(&dd, sizeof(double) * m));
    if (mean_in) {
        HIPCK(bufs.alloc(&dmi, sizeof(double) * m));
        HIPCK(hipMemcpy(dmi, mean_in, sizeof(double) * m, hipMemcpyHostToDevice));
    }
    int rc = NBX_OK;
    if (!rho_by_id) rc = density_impl(t, k, NBX_DENS_SYMMETRIC, 0, rid, nullptr, nullptr, st);
    hipError_t e = hipSuccess;
    if (rc == NBX_OK) {
        if (rho_by_id) e = hipMemcpy(rid, rho_by_id, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(wid, weight_by_id, sizeof(double) * n, hipMemcpyHostToDevice);
        KnnArgs K;
        memset(&K, 0, sizeof(K));
        K.qsrc = NBX_Q_POINTS;
        K.kh = k;
        K.kout = k;
        K.m = (int32_t)m;
        for (int d = 0; d < 3; ++d) K.q[d] = dq + d * m;
        K.nn = dnn;
        if (e == hipSuccess) e = nbx_launch_knn(t->view(), K, (int)m, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, rid, t->perm, n, rslot);
            hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, wid, t->perm, n, wslot);
            LocalArgs A = {{dq, dq + m, dq + 2 * m}, m, dnn, k, 1, rslot, wslot, dmi, dm, dd};
            e = nbx_launch_smooth_local(t->view(), A, st);
        }
        if (e == hipSuccess && mean_out) e = hipMemcpy(mean_out, dm, sizeof(double) * m, hipMemcpyDeviceToHost);
        if (e == hipSuccess && disp_out) e = hipMemcpy(disp_out, dd, sizeof(double) * m, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(e, "nbx_smooth_local_points");
    }
    return rc;
}

extern "C" int nbx_density(nbx_tree *t, int32_t k, int32_t variant, int32_t nsearch, double *rho_by_id,
                           double *h_by_id) {
    if (!t || !rho_by_id) return fail(NBX_ERR_ARG, "NULL argument");
    double *dr = nullptr, *dh = nullptr;
    DevBufs bufs;
    HIPCK(bufs.alloc(&dr, sizeof(double) * t->n));
    if (h_by_id) HIPCK(bufs.alloc(&dh, sizeof(double) * t->n));
    int rc = density_impl(t, k, variant, nsearch, dr, dh, nullptr, nullptr);
    if (rc == NBX_OK) {
        hipError_t e = hipMemcpy(rho_by_id, dr, sizeof(double) * t->n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && h_by_id) e = hipMemcpy(h_by_id, dh, sizeof(double) * t->n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(e, "nbx_density");
    }
    return rc;
}

extern "C" int nbx_veldensity_phys(nbx_tree *t, int32_t nsmooth, int32_t nsearch, const double *rho_by_id,
                                   double *vden_by_id) {
    if (!t || !rho_by_id || !vden_by_id) return fail(NBX_ERR_ARG, "NULL argument");
    if (nsmooth < 1) return fail(NBX_ERR_ARG, "Nsmooth must be >= 1");
    DevBufs bufs;
    double *rin = nullptr, *dv = nullptr;
    HIPCK(bufs.alloc(&rin, sizeof(double) * t->n));
    HIPCK(bufs.alloc(&dv, sizeof(double) * t->n));
    HIPCK(hipMemcpy(rin, rho_by_id, sizeof(double) * t->n, hipMemcpyHostToDevice));
    int rc = density_impl(t, nsmooth, DENS_VELPHYS, nsearch, dv, nullptr, nullptr, nullptr, rin);
    if (rc == NBX_OK) {
        hipError_t e = hipMemcpy(vden_by_id, dv, sizeof(double) * t->n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(e, "nbx_veldensity_phys");
    }
    return rc;
}

// ------------------------------------------------------------------ FOF
static int fof_impl(nbx_tree *t, int metric, const FofLink &link, int32_t minnum, int32_t *labels_by_id_dev,
                    int64_t *ngroups, hipStream_t st) {
    const int64_t n = t->n;
    DevTree T = t->view();
    // parents, node flags and queue counters, then the labelling's arrays,
    // from the tree's kept scratch
    const size_t wb = nbx_fof_work_bytes(T), cb = nbx_fof_canonical_bytes(n, minnum);
    Workspace dry = {nullptr, ~(size_t)0, 0};
    dry.carve<char>(wb);
    dry.carve<char>(cb);
    ScratchLease lease(t, dry.used, st);
    if (lease.err != hipSuccess) return hipfail(lease.err, "fof scratch");
    int32_t *P = lease.ws.carve<int32_t>(wb);
    void *cw = lease.ws.carve<void>(cb);
    if (!P || !cw) return fail(NBX_ERR_ARG, "fof scratch too small");
    FofArgs A;
    memset(&A, 0, sizeof(A));
    A.link = link;
    A.metric = metric;
    A.periodic = t->p.periodic && n > t->p.bucket_size;
    for (int d = 0; d < 3; ++d) A.period[d] = t->p.period[d];
    A.parent = P;
    hipError_t e = nbx_launch_fof(T, A, st);
    if (e == hipSuccess) e = nbx_fof_canonical(T, P, minnum, labels_by_id_dev, ngroups, cw, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hipfail(e, "fof");
    return NBX_OK;
}

extern "C" int nbx_fof_dev(nbx_tree *t, double fdist, int32_t minnum, int32_t *labels_by_id_dev, int64_t *ngroups,
                           void *stream) {
    if (!t || !labels_by_id_dev || !ngroups) return fail(NBX_ERR_ARG, "NULL argument");
    if (!(fdist >= 0)) return fail(NBX_ERR_ARG, "linking length must be >= 0");
    if (t->tvel) return fail(NBX_ERR_UNSUPPORTED, "FOF needs a TPHYS or TPHS tree");
    FofLink L = {fdist * fdist, 0.0, 0.0, 0.0, 0.0};
    return fof_impl(t, NBX_FOFM_PLAIN, L, minnum, labels_by_id_dev, ngroups, (hipStream_t)stream);
}

// FOFCriterion's tree search scales offsets by params[1] (positions) and
// params[2] (velocities) and its comparator by params[6] / params[7].  The
// groups are the comparator's transitive closure exactly when the search is
// no narrower than the comparator, which is what the check below requires.
extern "C" int nbx_fof_criterion_dev(nbx_tree *t, int32_t comparator, const double *params, int32_t minnum,
                                     int32_t *labels_by_id_dev, int64_t *ngroups, void *stream) {
    if (!t || !params || !labels_by_id_dev || !ngroups) return fail(NBX_ERR_ARG, "NULL argument");
    FofLink L = {1.0, params[6], comparator == NBX_FOF_CMP_6D ? params[7] : 1.0, 0.0, 0.0};
    L.r6 = 1.0 / L.s6;
    L.r7 = 1.0 / L.s7;
    int metric;
    if (comparator == NBX_FOF_CMP_3D) {
        if (t->nd != 3 || t->tvel || (int)params[0] != NBX_TPHYS)
            return fail(NBX_ERR_UNSUPPORTED, "FOF3d runs on a TPHYS tree with params[0] = TPHYS");
        if (!(params[6] > 0) || !(params[1] >= params[6]))
            return fail(NBX_ERR_UNSUPPORTED, "FOF3d needs params[6] > 0 and a search scale params[1] >= params[6]");
        metric = NBX_FOFM_3D;
    } else if (comparator == NBX_FOF_CMP_6D) {
        if (t->nd != 6 || (int)params[0] != NBX_TPHS)
            return fail(NBX_ERR_UNSUPPORTED, "FOF6d runs on a TPHS tree with params[0] = TPHS");
        if (!(params[6] > 0) || !(params[7] > 0) || !(params[1] >= params[6]) || !(params[2] >= params[7]))
            return fail(NBX_ERR_UNSUPPORTED,
                        "FOF6d needs params[6], params[7] > 0 and search scales params[1] >= params[6], "
                        "params[2] >= params[7]");
        metric = NBX_FOFM_6D;
    } else {
        return fail(NBX_ERR_UNSUPPORTED, "only the built-in comparators FOF3d and FOF6d are provided");
    }
    return fof_impl(t, metric, L, minnum, labels_by_id_dev, ngroups, (hipStream_t)stream);
}

extern "C" int nbx_fof_criterion(nbx_tree *t, int32_t comparator, const double *params, int32_t minnum,
                                 int32_t *labels_by_id, int64_t *ngroups) {
    if (!t || !labels_by_id || !ngroups) return fail(NBX_ERR_ARG, "NULL argument");
    int32_t *d = nullptr;
    DevBufs bufs;
    HIPCK(bufs.alloc(&d, sizeof(int32_t) * t->n));
    int rc = nbx_fof_criterion_dev(t, comparator, params, minnum, d, ngroups, nullptr);
    if (rc == NBX_OK) {
        hipError_t e = hipMemcpy(labels_by_id, d, sizeof(int32_t) * t->n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(e, "nbx_fof_criterion");
    }
    return rc;
}

extern "C" int nbx_fof(nbx_tree *t, double fdist, int32_t minnum, int32_t *labels_by_id, int64_t *ngroups) {
    if (!t || !labels_by_id || !ngroups) return fail(NBX_ERR_ARG, "NULL argument");
    int32_t *d = nullptr;
    DevBufs bufs;
    HIPCK(bufs.alloc(&d, sizeof(int32_t) * t->n));
    int rc = nbx_fof_dev(t, fdist, minnum, d, ngroups, nullptr);
    if (rc == NBX_OK) {
        hipError_t e = hipMemcpy(labels_by_id, d, sizeof(int32_t) * t->n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(e, "nbx_fof");
    }
    return rc;
}

// ------------------------------------------------------------------ ball search, many positions
extern "C" int nbx_ball_points_many(nbx_tree *t, const double *q, int64_t m, double r2, int64_t *offsets,
                                    int32_t *ids, double *d2, int64_t cap) {
    DevBufs bufs;
    if (!t || (m > 0 && !q) || !offsets) return fail(NBX_ERR_ARG, "NULL argument");
    if (t->nd != 3 || t->tvel) return fail(NBX_ERR_UNSUPPORTED, "ball search needs a TPHYS tree");
    if (m < 0 || cap < 0 || !(r2 >= 0)) return fail(NBX_ERR_ARG, "bad sizes or radius");
    offsets[0] = 0;
    if (m == 0) return NBX_OK;
    const bool per = t->p.periodic && t->n > t->p.bucket_size;
    if (per)
        for (int d = 0; d < 3; ++d)
            if (!(4.0 * r2 < t->p.period[d] * t->p.period[d]))
                return fail(NBX_ERR_UNSUPPORTED, "periodic ball search needs a radius below half the period");
    if (m > INT32_MAX) return fail(NBX_ERR_ARG, "too many positions (m < 2^31)");
    double *dq = nullptr;
    int64_t *cnt = nullptr, *doff = nullptr;
    uint64_t *keys = nullptr;
    int32_t *idx = nullptr;
    void *tmp = nullptr;
    HIPCK(bufs.alloc(&dq, sizeof(double) * 3 * m));
    HIPCK(hipMemcpy(dq, q, sizeof(double) * 3 * m, hipMemcpyHostToDevice));
    HIPCK(bufs.alloc(&cnt, sizeof(int64_t) * m));
    // walk order: Morton order of the positions over the root box
    double rb[6];
    HIPCK(bufs.alloc(&keys, sizeof(uint64_t) * 2 * m));
    HIPCK(bufs.alloc(&idx, sizeof(int32_t) * 2 * m));
    HIPCK(hipMemcpy(rb, t->box + 6, sizeof(rb), hipMemcpyDeviceToHost));
    double sc[3];
    for (int d = 0; d < 3; ++d) sc[d] = rb[3 + d] > rb[d] ? 2097151.0 / (rb[3 + d] - rb[d]) : 0.0;
    hipLaunchKernelGGL(k_morton_keys, dim3(nblk(m)), dim3(256), 0, nullptr, dq, dq + m, dq + 2 * m, m, rb[0], rb[1],
                       rb[2], sc[0], sc[1], sc[2], keys, idx);
    size_t tb = 0;
    hipError_t e = nbx_sort_pairs_u64(nullptr, tb, keys, keys + m, idx, idx + m, m, nullptr);
    if (e == hipSuccess) e = bufs.alloc((char **)&tmp, tb > 0 ? tb : 1);
    if (e == hipSuccess) e = nbx_sort_pairs_u64(tmp, tb, keys, keys + m, idx, idx + m, m, nullptr);
    BallManyArgs A;
    memset(&A, 0, sizeof(A));
    for (int d = 0; d < 3; ++d) { A.q[d] = dq + d * m; A.period[d] = t->p.period[d]; }
    A.m = m;
    A.r2 = r2;
    A.periodic = per;
    A.count = cnt;
    A.order = idx + m;
    DevTree T = t->view();
    if (e == hipSuccess) e = nbx_launch_ball_many(T, A, false, nullptr);
    std::vector<int64_t> hc((size_t)m);
    if (e == hipSuccess) e = hipMemcpy(hc.data(), cnt, sizeof(int64_t) * m, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        for (int64_t i = 0; i < m; ++i) offsets[i + 1] = offsets[i] + hc[i];
        const int64_t total = offsets[m];
        if (total > 0 && total <= cap && ids) {
            int32_t *dids = nullptr;
            double *dd2 = nullptr;
            e = bufs.alloc(&doff, sizeof(int64_t) * m);
            if (e == hipSuccess) e = hipMemcpy(doff, offsets, sizeof(int64_t) * m, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = bufs.alloc(&dids, sizeof(int32_t) * total);
            if (e == hipSuccess && d2) e = bufs.alloc(&dd2, sizeof(double) * total);
            if (e == hipSuccess) {
                A.off = doff;
                A.ids = dids;
                A.d2 = dd2;
                e = nbx_launch_ball_many(T, A, true, nullptr);
            }
            if (e == hipSuccess) e = hipMemcpy(ids, dids, sizeof(int32_t) * total, hipMemcpyDeviceToHost);
            if (e == hipSuccess && d2) e = hipMemcpy(d2, dd2, sizeof(double) * total, hipMemcpyDeviceToHost);
        }
    }
    if (e != hipSuccess) return hipfail(e, "nbx_ball_points_many");
    return NBX_OK;
}

// ------------------------------------------------------------------ ball search
// IDs (and d2, when d2out is given) of the particles with d2 < r2 around x,
// through the images of SearchBallPosPeriodicTagged (KDSplitNode.cxx:1388-1414).
static int ball_impl(nbx_tree *t, const double *x, double r2, int32_t *out, double *d2out, int64_t cap,
                     int64_t *nout) {
    if (!t || !x || !nout) return fail(NBX_ERR_ARG, "NULL argument");
    if (t->nd != 3 || t->tvel) return fail(NBX_ERR_UNSUPPORTED, "ball search needs a TPHYS tree");
    int32_t *dout = nullptr;
    double *dd2 = nullptr;
    unsigned long long *dc = nullptr;
    int64_t dcap = cap > 0 ? cap : 1;
    DevBufs bufs;
    HIPCK(bufs.alloc(&dout, sizeof(int32_t) * dcap));
    if (d2out) HIPCK(bufs.alloc(&dd2, sizeof(double) * dcap));
    HIPCK(bufs.alloc(&dc, sizeof(unsigned long long)));
    HIPCK(hipMemset(dc, 0, sizeof(unsigned long long)));
    DevTree T = t->view();
    // primary position and the images of SearchBallPosPeriodicTagged
    // (KDSplitNode.cxx:1388-1414): image when r2 > s^2
    double pts[8][3];
    int np = 0;
    pts[np][0] = x[0]; pts[np][1] = x[1]; pts[np][2] = x[2]; np++;
    if (t->p.periodic && t->n > t->p.bucket_size) {
        const double *p = t->p.period;
        for (int k = 0; k < 3; ++k) {
            double xp[3] = {x[0], x[1], x[2]}, s;
            if (x[k] < p[k] / 2.0) { xp[k] = x[k] + p[k]; s = x[k]; } else { xp[k] = x[k] - p[k]; s = -xp[k]; }
            if (r2 > s * s) { memcpy(pts[np], xp, sizeof(xp)); np++; }
        }
        const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int c = 0; c < 4; ++c) {
            double xp[3] = {x[0], x[1], x[2]}, s = 0;
            for (int k = 0; k < 3; ++k) {
                bool use = c == 3 || k == pr[c][0] || k == pr[c][1];
                if (!use) continue;
                if (x[k] < p[k] / 2.0) { xp[k] = x[k] + p[k]; s += x[k] * x[k]; } else { xp[k] = x[k] - p[k]; s += xp[k] * xp[k]; }
            }
            s = sqrt(s);
            if (r2 > s * s) { memcpy(pts[np], xp, sizeof(xp)); np++; }
        }
    }
    for (int i = 0; i < np; ++i)
        hipLaunchKernelGGL(k_ball, dim3(1), dim3(NBX_WAVE), 0, 0, T, pts[i][0], pts[i][1], pts[i][2], r2, dout, dd2, cap,
                           dc);
    unsigned long long hc = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(&hc, dc, sizeof(hc), hipMemcpyDeviceToHost);
    if (e == hipSuccess && cap > 0) e = hipMemcpy(out, dout, sizeof(int32_t) * std::min<int64_t>(cap, (int64_t)hc), hipMemcpyDeviceToHost);
    if (e == hipSuccess && cap > 0 && d2out)
        e = hipMemcpy(d2out, dd2, sizeof(double) * std::min<int64_t>(cap, (int64_t)hc), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipfail(e, "ball");
    *nout = (int64_t)hc;
    return NBX_OK;
}

extern "C" int nbx_ball_points(nbx_tree *t, const double *x, double r2, int32_t *out, int64_t cap, int64_t *nout) {
    return ball_impl(t, x, r2, out, nullptr, cap, nout);
}

extern "C" int nbx_ball_mark(nbx_tree *t, const double *x, int64_t exclude_id, double r2, int32_t imark,
                             int32_t *group_by_id, double *d2_by_id) {
    if (!t || !x || !group_by_id || !d2_by_id) return fail(NBX_ERR_ARG, "NULL argument");
    if (t->p.periodic) {
        for (int d = 0; d < 3; ++d)
            if (!(4.0 * r2 < t->p.period[d] * t->p.period[d]))
                return fail(NBX_ERR_UNSUPPORTED, "periodic ball search needs a radius below half the period");
    }
    // with r below half the period a particle lies within r of at most one
    // image of x, so its d2 is the one the reference's image pass writes
    int64_t cap = std::min<int64_t>(t->n, 1 << 20), nout = 0;
    std::vector<int32_t> ids;
    std::vector<double> d2;
    for (;;) {
        ids.resize(cap);
        d2.resize(cap);
        int rc = ball_impl(t, x, r2, ids.data(), d2.data(), cap, &nout);
        if (rc != NBX_OK) return rc;
        if (nout <= cap) break;
        cap = nout;
    }
    for (int64_t i = 0; i < nout; ++i) {
        if (ids[i] == exclude_id) continue;
        group_by_id[ids[i]] = imark;
        d2_by_id[ids[i]] = d2[i];
    }
    return NBX_OK;
}

// ------------------------------------------------------------------ differentiation form over the tree
// The frozen list is the tree's own: each particle's k nearest under
// CalcDensity's search (KDCalcSmoothQuantities.cxx:227; self and d2 = 0
// excluded, no period), recomputed for one block of tree slots at a time
// into a buffer of that block and consumed there, so no list of all n is
// kept.  Forward and adjoint are row-separable (every row adds its own
// terms), so the blocks sum to the whole.  Work runs in slot order on the
// tree's sorted coordinates; results are scattered back by ID.
static int sph_tree_impl(nbx_tree *t, int32_t k, int64_t block, const double *g_id, double *rho_id, double *h_id,
                         double *dpos_id, double *dmass_id, hipStream_t st) {
    if (!t) return fail(NBX_ERR_ARG, "NULL tree");
    if (t->nd != 3 || t->tvel) return fail(NBX_ERR_UNSUPPORTED, "the differentiation form needs a TPHYS tree");
    const bool mixed = mixed_ok(t, k);
    if (k < 2 || k > heap_cap(t)) return fail(NBX_ERR_UNSUPPORTED, "k out of range");
    const int64_t n = t->n;
    if (n <= k) return fail(NBX_ERR_SEARCH, "fewer than k + 1 particles");
    if (block <= 0) block = n;
    block = (block + 1023) / 1024 * 1024;  // whole waves and whole 1024-row windows
    if (block > n) block = (n + 1023) / 1024 * 1024;
    DevBufs bufs(st);
    int32_t *nbr = nullptr, *err = nullptr;
    HIPCK(bufs.alloc(&nbr, sizeof(int32_t) * (size_t)block * k));
    HIPCK(bufs.alloc(&err, sizeof(int32_t)));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, err, 0, (int64_t)1);
    const double *X[3] = {t->xs[0], t->xs[1], t->xs[2]};
    double *rs = nullptr, *hs = nullptr, *gs = nullptr, *ds = nullptr, *ms = nullptr;
    const bool adj = g_id != nullptr;
    if (adj) {
        HIPCK(bufs.alloc(&gs, sizeof(double) * n));
        HIPCK(bufs.alloc(&ds, sizeof(double) * 3 * n));
        HIPCK(bufs.alloc(&ms, sizeof(double) * n));
        hipLaunchKernelGGL(k_gather_by_id, dim3(nblk(n)), dim3(256), 0, st, g_id, t->perm, n, gs);
        hipLaunchKernelGGL(k_fill_f64, dim3(nblk(3 * n)), dim3(256), 0, st, ds, 0.0, 3 * n);
        hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n)), dim3(256), 0, st, ms, 0.0, n);
    } else {
        HIPCK(bufs.alloc(&rs, sizeof(double) * n));
        if (h_id) HIPCK(bufs.alloc(&hs, sizeof(double) * n));
        hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n)), dim3(256), 0, st, rs, 0.0, n);
    }
    // the search's temporaries, reused by every block
    Workspace ws = {nullptr, 0, 0};
    if (mixed) {
        ws.cap = 256 + sizeof(int32_t) * (block + 1) + 256 + nbx_mixed_scratch_bytes(t->nd, k + 1);
        HIPCK(bufs.alloc(&ws.base, ws.cap));
    }
    double *D[3] = {ds, ds ? ds + n : nullptr, ds ? ds + 2 * n : nullptr};
    for (int64_t r0 = 0; r0 < n; r0 += block) {
        const int64_t nr = std::min<int64_t>(block, n - r0);
        KnnArgs A;
        memset(&A, 0, sizeof(A));
        A.qsrc = NBX_Q_SLOTS;
        A.excl_zero = 1;
        A.kh = k;
        A.kout = k;
        A.layout = NBX_LAYOUT_BY_SLOT;
        A.dens = NBX_DO_NONE;
        A.nn = nbr;
        A.wbeg = (int32_t)(r0 / NBX_WAVE);
        A.wcnt = (int32_t)((nr + NBX_WAVE - 1) / NBX_WAVE);
        A.row0 = (int32_t)r0;
        ws.used = 0;
        hipError_t e = mixed ? run_mixed(t, A, k, st, &ws) : nbx_launch_knn(t->view(), A, A.wcnt * NBX_WAVE, st);
        if (e == hipSuccess) e = nbx_launch_kth_last(X, t->perm, nbr, r0, nr, k, err, st);
        if (e == hipSuccess)
            e = adj ? nbx_launch_sph_adj_rows(X, t->mass, nbr, n, r0, nr, k, gs, D, ms, st)
                    : nbx_launch_sph_fwd_rows(X, t->mass, nbr, n, r0, nr, k, rs, hs, st);
        if (e != hipSuccess) return hipfail(e, "differentiation form");
    }
    if (adj) {
        for (int c = 0; c < 3; ++c)
            hipLaunchKernelGGL(k_scatter_by_id, dim3(nblk(n)), dim3(256), 0, st, ds + c * n, t->perm, dpos_id + c * n, n);
        hipLaunchKernelGGL(k_scatter_by_id, dim3(nblk(n)), dim3(256), 0, st, ms, t->perm, dmass_id, n);
    } else {
        hipLaunchKernelGGL(k_scatter_by_id, dim3(nblk(n)), dim3(256), 0, st, rs, t->perm, rho_id, n);
        if (h_id) hipLaunchKernelGGL(k_scatter_by_id, dim3(nblk(n)), dim3(256), 0, st, hs, t->perm, h_id, n);
    }
    int32_t herr = 0;
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (herr) return fail(NBX_ERR_SEARCH, "differentiation form: fewer than k neighbours for some particle");
    return NBX_OK;
}

extern "C" int nbx_sph_density_fwd_tree_dev(nbx_tree *t, int32_t k, int64_t block_rows, double *rho_by_id_dev,
                                            double *h_by_id_dev, void *stream) {
    if (!rho_by_id_dev) return fail(NBX_ERR_ARG, "NULL output");
    return sph_tree_impl(t, k, block_rows, nullptr, rho_by_id_dev, h_by_id_dev, nullptr, nullptr, (hipStream_t)stream);
}
extern "C" int nbx_sph_density_adjoint_tree_dev(nbx_tree *t, int32_t k, int64_t block_rows, const double *g_by_id_dev,
                                                double *dpos_by_id_dev, double *dmass_by_id_dev, void *stream) {
    if (!g_by_id_dev || !dpos_by_id_dev || !dmass_by_id_dev) return fail(NBX_ERR_ARG, "NULL argument");
    return sph_tree_impl(t, k, block_rows, g_by_id_dev, nullptr, nullptr, dpos_by_id_dev, dmass_by_id_dev,
                         (hipStream_t)stream);
}

// ------------------------------------------------------------------ differentiation form
extern "C" int nbx_sph_density_fwd_dev(const double *pos, const double *mass, const int32_t *nbr, int64_t n, int32_t k,
                                       double *rho, double *h, void *stream) {
    if (!pos || !mass || !nbr || !rho) return fail(NBX_ERR_ARG, "NULL argument");
    if (n < 1 || k < 1) return fail(NBX_ERR_ARG, "bad sizes");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n)), dim3(256), 0, st, rho, 0.0, n);
    HIPCK(nbx_launch_sph_fwd(pos, mass, nbr, n, k, rho, h, st));
    return NBX_OK;
}
extern "C" int nbx_sph_density_adjoint_dev(const double *pos, const double *mass, const int32_t *nbr, int64_t n,
                                           int32_t k, const double *g, double *dpos, double *dmass, void *stream) {
    if (!pos || !mass || !nbr || !g || !dpos || !dmass) return fail(NBX_ERR_ARG, "NULL argument");
    if (n < 1 || k < 1) return fail(NBX_ERR_ARG, "bad sizes");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(3 * n)), dim3(256), 0, st, dpos, 0.0, 3 * n);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n)), dim3(256), 0, st, dmass, 0.0, n);
    HIPCK(nbx_launch_sph_adj(pos, mass, nbr, n, k, g, dpos, dmass, st));
    return NBX_OK;
}
