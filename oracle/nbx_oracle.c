/*
 * nbx_oracle.c -- CPU oracle for the KDTree hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (nbodylib_amd/,
 * include/) links, loads or calls this file.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it, and there
 * only as the checker (or as the timed CPU baseline), never as a product path.
 *
 * Two independent restatements of the reference's algorithm live here:
 *
 *  (1) A faithful restatement of NBodylib's KDTree (src/KDTree/):
 *      median-split tree built by the same Hoare-style quickselect
 *      (KDTree.cxx:328-370, BuildNodes :970-1055, DetermineSplitDim :459-504),
 *      the same 1-based max-heap PriorityQueue (PriorityQueue.h:47-84),
 *      the same near-first recursion with the incremental offset bound
 *      (KDSplitNode.cxx:15-95), leaf rules (KDLeafNode.cxx:15-57, :119-154),
 *      the sequential periodic image rule (KDSplitNode.cxx:1119-1179,
 *      DistFunc.h:326-355), LoadNN (KDFindNearest.cxx:11-24), CalcDensity
 *      (KDCalcSmoothQuantities.cxx:203-305), CalcDensityParticle (:768-844),
 *      CalcVelDensity (:309-389) and the serial BFS FOF (KDFOF.cxx:18-153,
 *      KDSplitNode.cxx:921-988, :1602-1646, KDLeafNode.cxx:527-590,
 *      KDNode.h:249-332).  Compiled with -O2 -fopenmp -ffp-contract=off it is
 *      also the CPU baseline ("port").
 *
 *  (2) Brute-force restatements that share no code with (1): all-pairs kNN with
 *      the same acceptance rules, the sequential image rule applied to all
 *      points, all-pairs union-find FOF (3-D, and 6-D phase space with the
 *      reference's |dx|^2 + |dv|^2, KDLeafNode.cxx:570-571), long-double density from neighbour
 *      lists, the FOF canonicaliser, the analytic (C1) cubic-spline density used
 *      by the differentiation form and its hand-derived adjoint.
 *
 * Pins (tests/test_oracle_*.py): (1) == (2) on seeded inputs and adversarial
 * cases; kernel normalisation against closed forms (1/pi in 3-D,
 * 3024/(762 pi^3) in 6-D) and against the normalisation integral; mean
 * density of a uniform box against sum(m)/V; FOF on constructed chains;
 * 6-D FOF with zero velocities against 3-D FOF;
 * adjoint against central finite differences.
 *
 * Arithmetic follows the reference's x86-64 build: IEEE double, no FMA
 * contraction (build with -ffp-contract=off), sums in the reference's order.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXVALUE 1e32 /* Precision.h:48 */
#define ORC_TPHYS 0
#define ORC_TPHS 3
#define ORC_KSPH 0
#define ORC_KGAUSS 1
#define ORC_KEPAN 2
#define ORC_KTH 3

/* optional planted mistakes, used by tests to show the pins catch them */
static int g_mutation = 0;
void orc_set_mutation(int m) { g_mutation = m; }

/* ------------------------------------------------------------------ kernels */
/* SmoothingKernels.h:32-56 */
static double w_sph(double r, double h) {
    double rh = r / h;
    if (rh <= 1.0) return (1.0 - 0.75 * (2.0 - rh) * (rh * rh));
    else if (rh > 1.0 && rh <= 2.0) return 0.25 * (2.0 - rh) * (2.0 - rh) * (2.0 - rh);
    else return 0;
}
static double w_th(double r, double h) { double rh = r / h * 0.5; return (rh <= 1.0); }
static double w_gauss(double r, double h) { double rh = r / h * 2.0; return exp(-0.5 * pow(rh, 2.0)); }
static double w_epan(double r, double h) {
    double rh = r / h * 0.5;
    if (rh <= 1.0) return (1.0 - rh * rh);
    return 0.;
}

/* KernelConstruction, KDTree.cxx:1144-1183.  Returns the (possibly clamped)
 * kernres; writes kernnorm and table[kernres] when non-NULL. */
int orc_kernel(int nd, int kerntype, int kernres, double *kernnorm_out, double *table) {
    if (kernres < 100) kernres = 100;
    double kernnorm = 1.0 / nd * pow(M_1_PI, nd / 2.) * tgamma(nd / 2. + 1.);
    double (*f)(double, double) = w_sph;
    if (kerntype == ORC_KSPH) {
        kernnorm *= nd * (nd + 1.) * (nd + 2.) * (nd + 3.) / (6 * (pow(2., nd + 1) - 1.));
    } else if (kerntype == ORC_KGAUSS) {
        f = w_gauss;
        kernnorm *= pow(0.5 * M_1_PI, nd / 2.);
    } else if (kerntype == ORC_KEPAN) {
        f = w_epan;
        kernnorm *= nd * (nd + 2.) * pow(0.5, nd + 1.);
    } else if (kerntype == ORC_KTH) {
        f = w_th;
    } else {
        kernnorm *= nd * (nd + 1.) * (nd + 2.) * (nd + 3.) / (6 * (pow(2., nd + 1) - 1.));
    }
    if (g_mutation == 1) kernnorm *= 0.5; /* planted: dropped factor */
    if (kernnorm_out) *kernnorm_out = kernnorm;
    if (table) {
        double delta = 2.0 / (double)(kernres - 1);
        for (int i = 0; i < kernres; i++) table[i] = kernnorm * f(i * delta, 1.0);
    }
    return kernres;
}

/* Wsm, KDCalcSmoothQuantities.cxx:12-15.  The reference returns x[i] past
 * the interpolation range, which reads beyond the table once r/h > 2; the
 * searches that set h = r_k / 2 never go there (i <= size - 1), but
 * CalcVelDensityWithPhysDensity's h without a square root does.  There the
 * last entry W(2) = 0 is taken. */
static inline double wsm(double r, int i, int size, double delta, const double *x) {
    if (i < size - 1) return (x[i] + (x[i + 1] - x[i]) * (r - delta * i) / delta);
    return x[size - 1];
}
static inline double wsm_at(double rij, double hi, int kernres, const double *kern) {
    const double fi = rij / hi * 0.5 * (kernres - 1);
    /* past the end: decided before the conversion, which a huge r/h would overflow */
    if (!(fi < (double)(kernres - 1))) return kern[kernres - 1];
    return wsm(rij / hi, (int)fi, kernres, 2.0 / (double)(kernres - 1), kern);
}

/* ---------------------------------------------------------- priority queue */
/* PriorityQueue.h:14-85: 1-based binary max-heap of (index, key). */
typedef struct { int32_t q; double p; } pqm;
typedef struct { int32_t n, max; pqm *a; } pq_t;
static void pq_init(pq_t *h, int32_t max) { h->n = 0; h->max = max; h->a = (pqm *)malloc(sizeof(pqm) * (size_t)(max + 1)); }
static void pq_free(pq_t *h) { free(h->a); }
static void pq_push(pq_t *h, int32_t q, double dist) {
    if (++h->n > h->max) { fprintf(stderr, "oracle: priority queue full\n"); abort(); }
    int32_t r = h->n;
    while (r > 1) {
        int32_t p = r >> 1;
        if (h->a[p].p >= dist) break;
        h->a[r] = h->a[p];
        r = p;
    }
    h->a[r].p = dist;
    h->a[r].q = q;
}
static void pq_pop(pq_t *h) {
    double kn = h->a[h->n--].p;
    int32_t p = 1, r = 2;
    while (r <= h->n) {
        if (r < h->n && h->a[r].p < h->a[r + 1].p) r++;
        if (kn >= h->a[r].p) break;
        h->a[p] = h->a[r];
        p = r;
        r = p << 1;
    }
    h->a[p] = h->a[h->n + 1];
}
static inline double pq_top(const pq_t *h) { return h->a[1].p; }
static inline int32_t pq_topq(const pq_t *h) { return h->a[1].q; }

/* -------------------------------------------------------------------- tree */
typedef struct {
    double ph[6]; /* x y z vx vy vz */
    double m;
    int32_t id; /* input index, KDTree.cxx:1291 */
    double rho;
} opart;

typedef struct {
    int64_t start, end;
    int isleaf, cut_dim;
    double cut_val;
    int32_t left, right, nid;
    double bnd[6][2];
    double maxextent2;
} onode;

typedef struct {
    int64_t n;
    int nd, treetype, b;
    int periodic;
    double period[3];
    opart *p;
    onode *nodes;
    int32_t nnodes, cap, nleaf, root;
    int kerntype, kernres;
    double kernnorm;
    double *kern;
} otree;

static inline double coord(const otree *t, int64_t s, int d) { return t->p[s].ph[d]; }

static int32_t new_node(otree *t) {
    if (t->nnodes == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->nodes = (onode *)realloc(t->nodes, sizeof(onode) * (size_t)t->cap);
    }
    memset(&t->nodes[t->nnodes], 0, sizeof(onode));
    return t->nnodes++;
}

/* SpreadestPos/BoundaryandMeanPos (KDTree.cxx:27-48, :96-119): min/max scan */
static double spread(const otree *t, int j, int64_t start, int64_t end, double *bnd) {
    double mn = coord(t, start, j), mx = mn;
    for (int64_t i = start + 1; i < end; i++) {
        double v = coord(t, i, j);
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
    bnd[0] = mn;
    bnd[1] = mx;
    return mx - mn;
}

/* MedianPos, KDTree.cxx:328-370 (balanced branch), swapping whole particles */
static double median(otree *t, int d, int64_t k, int64_t start, int64_t end) {
    int64_t left = start, right = end - 1, i, j;
    opart *P = t->p, w, tmp;
    while (left < right) {
        double x = P[k].ph[d];
        tmp = P[right]; P[right] = P[k]; P[k] = tmp;
        i = left - 1;
        j = right;
        for (;;) {
            while (i < j) if (P[++i].ph[d] >= x) break;
            while (i < j) if (P[--j].ph[d] <= x) break;
            tmp = P[i]; P[i] = P[j]; P[j] = tmp;
            if (j <= i) break;
        }
        w = P[j];
        P[j] = P[i];
        P[i] = P[right];
        P[right] = w;
        if (i >= k) right = i - 1;
        if (i <= k) left = i + 1;
    }
    return P[k].ph[d];
}

static void set_extent(onode *nd, int ndim) {
    double me = 0;
    for (int j = 0; j < ndim; j++) { double e = nd->bnd[j][1] - nd->bnd[j][0]; if (e > me) me = e; }
    nd->maxextent2 = me * me; /* KDNode.h:366-370, :505-509 */
}

/* BuildNodes, KDTree.cxx:970-1055 (non-adaptive, spread criterion) */
static int32_t build(otree *t, int64_t start, int64_t end) {
    int64_t size = end - start;
    int32_t me = new_node(t);
    onode nd;
    memset(&nd, 0, sizeof(nd));
    nd.start = start;
    nd.end = end;
    nd.left = nd.right = -1;
    if (size <= t->b) {
        nd.isleaf = 1;
        for (int j = 0; j < t->nd; j++) spread(t, j, start, end, nd.bnd[j]);
        set_extent(&nd, t->nd);
        t->nodes[me] = nd;
        t->nleaf++;
        return me;
    }
    double sv[6];
    for (int j = 0; j < t->nd; j++) sv[j] = spread(t, j, start, end, nd.bnd[j]);
    int sd = 0;
    double cur = sv[0];
    for (int j = 1; j < t->nd; j++) if (sv[j] > cur) { sd = j; cur = sv[j]; }
    int64_t splitindex = start + (size - 1) / 2;
    nd.cut_dim = sd;
    nd.cut_val = median(t, sd, splitindex, start, end);
    set_extent(&nd, t->nd);
    t->nodes[me] = nd;
    int32_t l = build(t, start, splitindex + 1);
    int32_t r = build(t, splitindex + 1, end);
    t->nodes[me].left = l;
    t->nodes[me].right = r;
    return me;
}

static void number_nodes(otree *t, int32_t node, int32_t *c) { /* BuildNodeIDs, :1185-1204 */
    t->nodes[node].nid = (*c)++;
    if (!t->nodes[node].isleaf) {
        number_nodes(t, t->nodes[node].left, c);
        number_nodes(t, t->nodes[node].right, c);
    }
}

/* phase: n x 6 row-major (x y z vx vy vz); mass may be NULL (then 1/n) */
void *orc_tree_build(const double *phase, const double *mass, int64_t n, int b, int treetype,
                     int kerntype, int kernres, const double *period) {
    if (treetype != ORC_TPHYS && treetype != ORC_TPHS) return NULL;
    otree *t = (otree *)calloc(1, sizeof(otree));
    t->n = n;
    t->b = b;
    t->treetype = treetype;
    t->nd = (treetype == ORC_TPHS) ? 6 : 3;
    t->kerntype = kerntype;
    t->kernres = orc_kernel(t->nd, kerntype, kernres, &t->kernnorm, NULL);
    t->kern = (double *)malloc(sizeof(double) * (size_t)t->kernres);
    orc_kernel(t->nd, kerntype, t->kernres, NULL, t->kern);
    if (period) { t->periodic = 1; for (int k = 0; k < 3; k++) t->period[k] = period[k]; }
    t->p = (opart *)malloc(sizeof(opart) * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        for (int d = 0; d < 6; d++) t->p[i].ph[d] = phase[i * 6 + d];
        t->p[i].m = mass ? mass[i] : 1.0 / (double)n;
        t->p[i].id = (int32_t)i;
        t->p[i].rho = 0;
    }
    t->root = build(t, 0, n);
    int32_t c = 0;
    number_nodes(t, t->root, &c);
    return t;
}
void orc_tree_free(void *h) {
    otree *t = (otree *)h;
    if (!t) return;
    free(t->p); free(t->nodes); free(t->kern); free(t);
}
int64_t orc_tree_numnodes(void *h) { return ((otree *)h)->nnodes; }
int64_t orc_tree_numleaves(void *h) { return ((otree *)h)->nleaf; }
void orc_tree_order(void *h, int32_t *slot_to_id) {
    otree *t = (otree *)h;
    for (int64_t i = 0; i < t->n; i++) slot_to_id[i] = t->p[i].id;
}

/* ---------------------------------------------------------- kNN traversal */
/* DistanceSqd(p1,p2,dim) DistFunc.h:14-19, loop form; PhaseDistSqd :41-49 */
static inline double dist_dim(const double *a, const double *b, int dim) {
    double total = 0;
    for (int i = 0; i < dim; i++) total += (a[i] - b[i]) * (a[i] - b[i]);
    return total;
}
static inline double dist_phase(const double *a, const double *b) {
    return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]) +
           (a[3] - b[3]) * (a[3] - b[3]) + (a[4] - b[4]) * (a[4] - b[4]) + (a[5] - b[5]) * (a[5] - b[5]);
}

/* target-based search: FindNearestPos(tt) / FindNearestPhase(tt)
 * (KDSplitNode.cxx:15-41, :69-95; KDLeafNode.cxx:15-28, :43-57) */
static void nn_target(const otree *t, int32_t node, double rd, pq_t *pq, double *off, int64_t target, int phase, int dim) {
    const onode *nd = &t->nodes[node];
    if (nd->isleaf) {
        const double *xt = t->p[target].ph;
        for (int64_t i = nd->start; i < nd->end; i++) {
            if (i != target) {
                double d2 = phase ? dist_phase(xt, t->p[i].ph) : dist_dim(xt, t->p[i].ph, dim);
                if (d2 < pq_top(pq) && (d2 > 0 || g_mutation == 2)) { pq_pop(pq); pq_push(pq, (int32_t)i, d2); }
            }
        }
        return;
    }
    double old_off = off[nd->cut_dim];
    double new_off = t->p[target].ph[nd->cut_dim] - nd->cut_val;
    if (new_off < 0) {
        nn_target(t, nd->left, rd, pq, off, target, phase, dim);
        rd += -old_off * old_off + new_off * new_off;
        if (rd < pq_top(pq)) {
            off[nd->cut_dim] = new_off;
            nn_target(t, nd->right, rd, pq, off, target, phase, dim);
            off[nd->cut_dim] = old_off;
        }
    } else {
        nn_target(t, nd->right, rd, pq, off, target, phase, dim);
        rd += -old_off * old_off + new_off * new_off;
        if (rd < pq_top(pq)) {
            off[nd->cut_dim] = new_off;
            nn_target(t, nd->left, rd, pq, off, target, phase, dim);
            off[nd->cut_dim] = old_off;
        }
    }
}

/* position-based search: FindNearestPos(x) / FindNearestPhase(x,v)
 * (KDSplitNode.cxx:206-289; KDLeafNode.cxx:119-154): no exclusions */
static void nn_pos(const otree *t, int32_t node, double rd, pq_t *pq, double *off, const double *x, int phase, int dim) {
    const onode *nd = &t->nodes[node];
    if (nd->isleaf) {
        for (int64_t i = nd->start; i < nd->end; i++) {
            double d2 = phase ? dist_phase(x, t->p[i].ph) : dist_dim(x, t->p[i].ph, dim);
            if (d2 < pq_top(pq)) { pq_pop(pq); pq_push(pq, (int32_t)i, d2); }
        }
        return;
    }
    double old_off = off[nd->cut_dim];
    double new_off = x[nd->cut_dim] - nd->cut_val;
    if (new_off < 0) {
        nn_pos(t, nd->left, rd, pq, off, x, phase, dim);
        rd += -old_off * old_off + new_off * new_off;
        if (rd < pq_top(pq)) {
            off[nd->cut_dim] = new_off;
            nn_pos(t, nd->right, rd, pq, off, x, phase, dim);
            off[nd->cut_dim] = old_off;
        }
    } else {
        nn_pos(t, nd->right, rd, pq, off, x, phase, dim);
        rd += -old_off * old_off + new_off * new_off;
        if (rd < pq_top(pq)) {
            off[nd->cut_dim] = new_off;
            nn_pos(t, nd->left, rd, pq, off, x, phase, dim);
            off[nd->cut_dim] = old_off;
        }
    }
}

/* PeriodicReflection1D/2D/ND, DistFunc.h:326-355 (3-vector Coordinate) */
static double refl1(const double *x0, double *xp, const double *p, int k) {
    double sval = 0;
    xp[0] = x0[0]; xp[1] = x0[1]; xp[2] = x0[2];
    if (x0[k] < p[k] / 2.0) { xp[k] = x0[k] + p[k]; sval = x0[k]; }
    else { xp[k] = x0[k] - p[k]; sval = -xp[k]; }
    return sval;
}
static double refl2(const double *x0, double *xp, const double *p, int k1, int k2) {
    double sval = 0;
    xp[0] = x0[0]; xp[1] = x0[1]; xp[2] = x0[2];
    if (x0[k1] < p[k1] / 2.0) { xp[k1] = x0[k1] + p[k1]; sval += x0[k1] * x0[k1]; }
    else { xp[k1] = x0[k1] - p[k1]; sval += xp[k1] * xp[k1]; }
    if (x0[k2] < p[k2] / 2.0) { xp[k2] = x0[k2] + p[k2]; sval += x0[k2] * x0[k2]; }
    else { xp[k2] = x0[k2] - p[k2]; sval += xp[k2] * xp[k2]; }
    return sqrt(sval);
}
static double reflN(const double *x0, double *xp, const double *p, int ndim) {
    double sval = 0;
    xp[0] = x0[0]; xp[1] = x0[1]; xp[2] = x0[2];
    for (int k = 0; k < ndim; k++) {
        if (x0[k] < p[k] / 2.0) { xp[k] = x0[k] + p[k]; sval += x0[k] * x0[k]; }
        else { xp[k] = x0[k] - p[k]; sval += xp[k] * xp[k]; }
    }
    return sqrt(sval);
}

/* SplitNode::FindNearestPosPeriodic(x), KDSplitNode.cxx:1119-1148 */
static void nn_pos_periodic(const otree *t, pq_t *pq, const double *x) {
    double off[6] = {0, 0, 0, 0, 0, 0}, xp[3], sval;
    const double *p = t->period;
    const int dim = 3;
    nn_pos(t, t->root, 0.0, pq, off, x, 0, dim);
    for (int k = 0; k < dim; k++) {
        for (int j = 0; j < dim; j++) off[j] = 0.0;
        sval = refl1(x, xp, p, k);
        if (sqrt(pq_top(pq)) > sval) nn_pos(t, t->root, 0.0, pq, off, xp, 0, dim);
    }
    for (int j = 0; j < dim; j++) off[j] = 0.0;
    sval = refl2(x, xp, p, 0, 1);
    if (pq_top(pq) > sval) nn_pos(t, t->root, 0.0, pq, off, xp, 0, dim);
    for (int j = 0; j < dim; j++) off[j] = 0.0;
    sval = refl2(x, xp, p, 0, 2);
    if (pq_top(pq) > sval) nn_pos(t, t->root, 0.0, pq, off, xp, 0, dim);
    for (int j = 0; j < dim; j++) off[j] = 0.0;
    sval = refl2(x, xp, p, 1, 2);
    if (pq_top(pq) > sval) nn_pos(t, t->root, 0.0, pq, off, xp, 0, dim);
    for (int j = 0; j < dim; j++) off[j] = 0.0;
    sval = reflN(x, xp, p, dim);
    if (pq_top(pq) > sval) nn_pos(t, t->root, 0.0, pq, off, xp, 0, dim);
}

/* SplitNode::FindNearestPhasePeriodic(x,v), KDSplitNode.cxx:1153-1179:
 * all 7 spatial images unconditionally. */
static void nn_phase_periodic(const otree *t, pq_t *pq, const double *ph) {
    double off[6] = {0, 0, 0, 0, 0, 0}, xp[3], q[6];
    const double *p = t->period;
    nn_pos(t, t->root, 0.0, pq, off, ph, 1, 6);
    q[3] = ph[3]; q[4] = ph[4]; q[5] = ph[5];
    for (int k = 0; k < 3; k++) {
        memset(off, 0, sizeof(off));
        refl1(ph, xp, p, k);
        q[0] = xp[0]; q[1] = xp[1]; q[2] = xp[2];
        nn_pos(t, t->root, 0.0, pq, off, q, 1, 6);
    }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int c = 0; c < 3; c++) {
        memset(off, 0, sizeof(off));
        refl2(ph, xp, p, pairs[c][0], pairs[c][1]);
        q[0] = xp[0]; q[1] = xp[1]; q[2] = xp[2];
        nn_pos(t, t->root, 0.0, pq, off, q, 1, 6);
    }
    memset(off, 0, sizeof(off));
    reflN(ph, xp, p, 3);
    q[0] = xp[0]; q[1] = xp[1]; q[2] = xp[2];
    nn_pos(t, t->root, 0.0, pq, off, q, 1, 6);
}

/* LoadNN, KDFindNearest.cxx:11-24 (tree indices, ascending d2) */
static void load_nn(int32_t ns, pq_t *pq, int32_t *nn, double *d) {
    for (int32_t i = ns - 1; i >= 0; i--) {
        nn[i] = pq_topq(pq);
        d[i] = pq_top(pq);
        pq_pop(pq);
    }
}

/* One query for tree slot tt.  variant:
 *   0 FindNearest(tt)        KDFindNearest.cxx:247-318 (TPHYS, or TPHS aniso=-1)
 *   1 FindNearestPos(tt)     :320-334
 *   2 FindNearestPhase(tt)   :347-361
 *   3 non-periodic FindNearestPos(tt) as called by CalcDensity (:227) -- also
 *     FindNearestPhase when the tree is TPHS. */
static void query_one(const otree *t, int64_t tt, int k, int variant, pq_t *pq, int32_t *nn, double *d2) {
    int periodic = t->periodic && variant != 3;
    int ns = k + (periodic ? 1 : 0);
    double off[6] = {0, 0, 0, 0, 0, 0};
    pq->n = 0;
    pq->max = ns;
    for (int i = 0; i < ns; i++) pq_push(pq, -1, ORC_MAXVALUE);
    int phase = (variant == 2) || ((variant == 0 || variant == 3) && t->treetype == ORC_TPHS);
    int dim = phase ? 6 : ((variant == 0 || variant == 3) ? t->nd : 3);
    const onode *root = &t->nodes[t->root];
    if (!periodic) {
        nn_target(t, t->root, 0.0, pq, off, tt, phase, dim);
    } else if (root->isleaf) {
        /* LeafNode::Find*Periodic(target) fall back to the target search
         * (KDLeafNode.cxx:658-670), except the phase form which searches by
         * position without images. */
        if (phase) nn_pos(t, t->root, 0.0, pq, off, t->p[tt].ph, 1, 6);
        else nn_target(t, t->root, 0.0, pq, off, tt, 0, dim);
    } else {
        if (phase) nn_phase_periodic(t, pq, t->p[tt].ph);
        else nn_pos_periodic(t, pq, t->p[tt].ph);
    }
    if (periodic) {
        if (variant == 1) pq_pop(pq); /* FindNearestPos(tt): drop the largest */
        ns -= 1;                      /* the rest: leave the smallest in the heap */
    }
    load_nn(ns, pq, nn, d2);
}

/* All particles (the reference's caller-side omp loop, test_kdtree.cxx:278-301).
 * ids[id*k+j] = ID of j-th neighbour of particle ID id (-1 if not found),
 * d2[id*k+j] ascending. */
void orc_knn_tree(void *h, int k, int variant, int32_t *ids, double *d2, int nthreads) {
    otree *t = (otree *)h;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        pq_t pq;
        pq_init(&pq, k + 1);
        int32_t *nn = (int32_t *)malloc(sizeof(int32_t) * (size_t)(k + 1));
        double *dd = (double *)malloc(sizeof(double) * (size_t)(k + 1));
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < t->n; i++) {
            query_one(t, i, k, variant, &pq, nn, dd);
            int64_t id = t->p[i].id;
            for (int j = 0; j < k; j++) {
                ids[id * k + j] = nn[j] >= 0 ? t->p[nn[j]].id : -1;
                if (d2) d2[id * k + j] = dd[j];
            }
        }
        free(nn); free(dd); pq_free(&pq);
    }
}

/* kNN for arbitrary positions (FindNearestPos(x)/FindNearest(x)): period-aware
 * for a periodic TPHYS tree (KDFindNearest.cxx:462-564). */
void orc_knn_points(void *h, const double *q, int64_t m, int k, int32_t *ids, double *d2, int nthreads) {
    otree *t = (otree *)h;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        pq_t pq;
        pq_init(&pq, k);
        int32_t *nn = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
        double *dd = (double *)malloc(sizeof(double) * (size_t)k);
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < m; i++) {
            double off[6] = {0, 0, 0, 0, 0, 0};
            pq.n = 0;
            for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            if (t->periodic && !t->nodes[t->root].isleaf) nn_pos_periodic(t, &pq, q + 3 * i);
            else nn_pos(t, t->root, 0.0, &pq, off, q + 3 * i, 0, 3);
            load_nn(k, &pq, nn, dd);
            for (int j = 0; j < k; j++) {
                ids[i * k + j] = nn[j] >= 0 ? t->p[nn[j]].id : -1;
                if (d2) d2[i * k + j] = dd[j];
            }
        }
        free(nn); free(dd); pq_free(&pq);
    }
}

/* ----------------------------------------------------------------- density */
/* CalcDensity, KDCalcSmoothQuantities.cxx:203-305, serial as shipped.
 * TPHYS: 3-D.  TPHS: 6-D phase search with the metric factor taken as 1
 * (the reference multiplies by uninitialised m0 for aniso=-1).
 * Returns 0, or -1 when a neighbour is missing (the reference exits). */
int orc_calc_density(void *h, int k, double *rho_by_id) {
    otree *t = (otree *)h;
    pq_t pq;
    pq_init(&pq, k);
    int nd = t->nd, rc = 0;
    for (int64_t i = 0; i < t->n; i++) t->p[i].rho = 0;
    for (int64_t i = 0; i < t->n && rc == 0; i++) {
        double off[6] = {0, 0, 0, 0, 0, 0};
        pq.n = 0;
        for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
        nn_target(t, t->root, 0.0, &pq, off, i, t->treetype == ORC_TPHS, nd);
        double hi = 0.5 * sqrt(pq_top(&pq));
        double norm = 1.0 / pow(hi, (double)(nd * 1.));
        for (int j = 0; j < k; j++) {
            if (pq_topq(&pq) == -1) { rc = -1; break; }
            double rij = sqrt(pq_top(&pq));
            double Wij = 0.5 * wsm_at(rij, hi, t->kernres, t->kern) * norm;
            int32_t id = pq_topq(&pq);
            t->p[i].rho = t->p[i].rho + Wij * t->p[id].m;
            if (g_mutation != 3) t->p[id].rho = t->p[id].rho + Wij * t->p[i].m; /* planted: dropped scatter */
            pq_pop(&pq);
        }
    }
    pq_free(&pq);
    for (int64_t i = 0; i < t->n; i++) rho_by_id[t->p[i].id] = t->p[i].rho;
    return rc;
}

/* CalcSmoothVel (KDCalcSmoothQuantities.cxx:480-540) and, when disp is not
 * NULL, CalcSmoothVelDisp (:542-614) on a TPHYS tree with velocities,
 * serial as shipped.  rho_by_id gives the densities divided by (the
 * reference's densityset = 1); NULL runs CalcDensity(k) first.  mean: n x 3
 * by ID; disp: n x 9 (row-major 3x3) by ID.  Returns 0, or -1 when a
 * neighbour is missing (the reference exits). */
int orc_smooth_vel(void *h, int k, const double *rho_by_id, double *mean, double *disp) {
    otree *t = (otree *)h;
    const int64_t n = t->n;
    if (rho_by_id) {
        for (int64_t i = 0; i < n; i++) t->p[i].rho = rho_by_id[t->p[i].id];
    } else {
        double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
        int rc = orc_calc_density(h, k, tmp);
        free(tmp);
        if (rc) return rc;
    }
    pq_t pq;
    pq_init(&pq, k);
    int rc = 0;
    for (int pass = 0; pass < (disp ? 2 : 1) && rc == 0; pass++) {
        double *out = pass ? disp : mean;
        const int nc = pass ? 9 : 3;
        for (int64_t i = 0; i < n * nc; i++) out[i] = 0;
        for (int64_t i = 0; i < n && rc == 0; i++) {
            double off[6] = {0, 0, 0, 0, 0, 0};
            pq.n = 0;
            for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            nn_target(t, t->root, 0.0, &pq, off, i, 0, 3);
            double hi = 0.5 * sqrt(pq_top(&pq));
            double norm = 1.0 / pow(hi, 3.0);
            for (int j = 0; j < k; j++) {
                if (pq_topq(&pq) == -1) { rc = -1; break; }
                double rij = sqrt(pq_top(&pq));
                double Wij = 0.5 * wsm_at(rij, hi, t->kernres, t->kern) * norm;
                int32_t x = pq_topq(&pq);
                const int64_t id1 = t->p[i].id, id2 = t->p[x].id;
                double temp = Wij / t->p[x].rho * t->p[x].m;
                if (!pass) {
                    for (int d = 0; d < 3; d++) mean[id1 * 3 + d] += temp * t->p[x].ph[3 + d];
                } else {
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++)
                            disp[id1 * 9 + 3 * a + b] += temp * (t->p[x].ph[3 + a] - mean[id1 * 3 + a]) *
                                                         (t->p[x].ph[3 + b] - mean[id1 * 3 + b]);
                }
                temp = Wij / t->p[i].rho * t->p[i].m;
                if (!pass) {
                    for (int d = 0; d < 3; d++) mean[id2 * 3 + d] += temp * t->p[i].ph[3 + d];
                } else {
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++)
                            disp[id2 * 9 + 3 * a + b] += temp * (t->p[i].ph[3 + a] - mean[id2 * 3 + a]) *
                                                         (t->p[i].ph[3 + b] - mean[id2 * 3 + b]);
                }
                pq_pop(&pq);
            }
        }
    }
    pq_free(&pq);
    return rc;
}

/* CalcSmoothVelSkew (KDCalcSmoothQuantities.cxx:617-690) and
 * CalcSmoothVelKurtosis (:692-767) given the means (n x 3) and dispersions
 * (n x 9) by ID, serial as shipped: per component, share * dv^3 /
 * pow(s, 1.5) and share * dv^4 / (s * s) - 3 for every share, with s the
 * receiving particle's dispersion diagonal.  rho_by_id as orc_smooth_vel.
 * skew, kurt: n x 3 by ID. */
int orc_smooth_moments(void *h, int k, const double *rho_by_id, const double *mean, const double *disp, double *skew,
                       double *kurt) {
    otree *t = (otree *)h;
    const int64_t n = t->n;
    if (rho_by_id) {
        for (int64_t i = 0; i < n; i++) t->p[i].rho = rho_by_id[t->p[i].id];
    } else {
        double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
        int rc = orc_calc_density(h, k, tmp);
        free(tmp);
        if (rc) return rc;
    }
    for (int64_t i = 0; i < 3 * n; i++) skew[i] = kurt[i] = 0;
    pq_t pq;
    pq_init(&pq, k);
    int rc = 0;
    for (int64_t i = 0; i < n && rc == 0; i++) {
        double off[6] = {0, 0, 0, 0, 0, 0};
        pq.n = 0;
        for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
        nn_target(t, t->root, 0.0, &pq, off, i, 0, 3);
        double hi = 0.5 * sqrt(pq_top(&pq));
        double norm = 1.0 / pow(hi, 3.0);
        for (int j = 0; j < k; j++) {
            if (pq_topq(&pq) == -1) { rc = -1; break; }
            double rij = sqrt(pq_top(&pq));
            double Wij = 0.5 * wsm_at(rij, hi, t->kernres, t->kern) * norm;
            int32_t x = pq_topq(&pq);
            const int64_t id1 = t->p[i].id, id2 = t->p[x].id;
            double temp = Wij / t->p[x].rho * t->p[x].m;
            for (int d = 0; d < 3; d++) {
                double dv = t->p[x].ph[3 + d] - mean[id1 * 3 + d], s = disp[id1 * 9 + 4 * d];
                skew[id1 * 3 + d] += temp * dv * dv * dv / pow(s, 1.5);
                kurt[id1 * 3 + d] += temp * dv * dv * dv * dv / (s * s) - 3.;
            }
            temp = Wij / t->p[i].rho * t->p[i].m;
            for (int d = 0; d < 3; d++) {
                double dv = t->p[i].ph[3 + d] - mean[id2 * 3 + d], s = disp[id2 * 9 + 4 * d];
                skew[id2 * 3 + d] += temp * dv * dv * dv / pow(s, 1.5);
                kurt[id2 * 3 + d] += temp * dv * dv * dv * dv / (s * s) - 3.;
            }
            pq_pop(&pq);
        }
    }
    pq_free(&pq);
    return rc;
}

/* CalcSmoothLocalMeanParticle / CalcSmoothLocalDispParticle
 * (KDCalcSmoothQuantities.cxx:1539-1617) for every particle (q == NULL), or
 * the Position forms (:1619-1693) for m positions q (3 per row): the k
 * nearest (target search, or the non-periodic search from x), h = sqrt(top)/2,
 * in pop order temp = W(r/h) h^-3 / rho_j * m_j, mean += temp * w_j and
 * disp += temp * (w_j - mean_in)^2.  Weights, rho and (particle form) means by
 * ID; rho NULL runs CalcDensity(k) first; outputs by ID or by row. */
int orc_smooth_local(void *h, const double *q, int64_t m, int k, const double *weight, const double *rho_by_id,
                     const double *mean_in, double *mean_out, double *disp_out) {
    otree *t = (otree *)h;
    const int64_t n = t->n;
    if (rho_by_id) {
        for (int64_t i = 0; i < n; i++) t->p[i].rho = rho_by_id[t->p[i].id];
    } else {
        double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
        int rc = orc_calc_density(h, k, tmp);
        free(tmp);
        if (rc) return rc;
    }
    const int64_t rows = q ? m : n;
    pq_t pq;
    pq_init(&pq, k);
    int rc = 0;
    for (int64_t r = 0; r < rows && rc == 0; r++) {
        double off[6] = {0, 0, 0, 0, 0, 0};
        pq.n = 0;
        for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
        if (q) nn_pos(t, t->root, 0.0, &pq, off, q + 3 * r, 0, 3);
        else nn_target(t, t->root, 0.0, &pq, off, r, 0, 3);
        const int64_t out = q ? r : t->p[r].id;
        const double mu = mean_in ? mean_in[out] : 0.0;
        double hi = 0.5 * sqrt(pq_top(&pq)), norm = 1.0 / pow(hi, 3.0), mean = 0., disp = 0.;
        for (int j = 0; j < k; j++) {
            if (pq_topq(&pq) == -1) { rc = -1; break; }
            double rij = sqrt(pq_top(&pq));
            double Wij = wsm_at(rij, hi, t->kernres, t->kern) * norm;
            int32_t x = pq_topq(&pq);
            double temp = Wij / t->p[x].rho * t->p[x].m, wj = weight[t->p[x].id];
            mean += temp * wj;
            disp += temp * (wj - mu) * (wj - mu);
            pq_pop(&pq);
        }
        if (mean_out) mean_out[out] = mean;
        if (disp_out) disp_out[out] = disp;
    }
    pq_free(&pq);
    return rc;
}

/* CalcDensityParticle for every particle, KDCalcSmoothQuantities.cxx:768-844
 * (gather form, no 1/2, no scatter).  Also the all-cores CPU restatement. */
int orc_calc_density_gather(void *h, int k, double *rho_by_id, double *h_by_id, int nthreads) {
    otree *t = (otree *)h;
    int nd = t->nd, rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(min : rc)
    {
        pq_t pq;
        pq_init(&pq, k);
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < t->n; i++) {
            double off[6] = {0, 0, 0, 0, 0, 0}, den = 0.;
            pq.n = 0;
            for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            nn_target(t, t->root, 0.0, &pq, off, i, t->treetype == ORC_TPHS, nd);
            double hi = 0.5 * sqrt(pq_top(&pq));
            double norm = 1.0 / pow(hi, (double)(nd * 1.));
            for (int j = 0; j < k; j++) {
                if (pq_topq(&pq) == -1) { rc = -1; break; }
                double rij = sqrt(pq_top(&pq));
                double Wij = wsm_at(rij, hi, t->kernres, t->kern) * norm;
                den += Wij * t->p[pq_topq(&pq)].m;
                pq_pop(&pq);
            }
            rho_by_id[t->p[i].id] = den;
            if (h_by_id) h_by_id[t->p[i].id] = hi;
        }
        pq_free(&pq);
    }
    return rc;
}

/* CalcDensityPosition(x, Nsmooth) for m positions (TPHYS tree),
 * KDCalcSmoothQuantities.cxx:1092-1149: the node search root->FindNearestPos
 * (non-periodic, no exclusion: a particle at x is its own neighbour, d2 = 0),
 * h = sqrt(top)/2, then the gather sum of W(r/h) h^-3 m_j in pop order
 * (farthest first).  Returns 0, or -1 when fewer than Nsmooth were found. */
int orc_calc_density_position(void *h, const double *q, int64_t m, int k, double *rho, double *hout, int nthreads) {
    otree *t = (otree *)h;
    int rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(min : rc)
    {
        pq_t pq;
        pq_init(&pq, k);
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < m; i++) {
            double off[6] = {0, 0, 0, 0, 0, 0}, den = 0.;
            pq.n = 0;
            for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            nn_pos(t, t->root, 0.0, &pq, off, q + 3 * i, 0, 3);
            double hi = 0.5 * sqrt(pq_top(&pq));
            double norm = 1.0 / pow(hi, 3.0);
            for (int j = 0; j < k; j++) {
                if (pq_topq(&pq) == -1) { rc = -1; break; }
                double rij = sqrt(pq_top(&pq));
                den += wsm_at(rij, hi, t->kernres, t->kern) * norm * t->p[pq_topq(&pq)].m;
                pq_pop(&pq);
            }
            rho[i] = den;
            if (hout) hout[i] = hi;
        }
        pq_free(&pq);
    }
    return rc;
}

/* CalcVelDensity, KDCalcSmoothQuantities.cxx:309-389 (TPHYS tree):
 * Nsearch physical neighbours, the Nsmooth nearest of them in |dv|,
 * h_v = max|dv|/2, sum of W(|dv|/h_v) h_v^-3 (no mass, no 1/2). */
int orc_calc_veldensity(void *h, int nsmooth, int nsearch, double *rho_by_id, int nthreads) {
    otree *t = (otree *)h;
    if (nsmooth > nsearch) nsmooth = nsearch;
    int nd = t->nd, rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(min : rc)
    {
        pq_t pq, pq2;
        pq_init(&pq, nsearch);
        pq_init(&pq2, nsmooth);
        int32_t *nnid = (int32_t *)malloc(sizeof(int32_t) * (size_t)nsearch);
        double *vdist = (double *)malloc(sizeof(double) * (size_t)nsearch);
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < t->n; i++) {
            double off[6] = {0, 0, 0, 0, 0, 0};
            pq.n = 0;
            for (int j = 0; j < nsearch; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            nn_target(t, t->root, 0.0, &pq, off, i, 0, nd);
            int bad = 0;
            for (int j = 0; j < nsearch; j++) {
                if (pq_topq(&pq) == -1) { bad = 1; break; }
                nnid[j] = pq_topq(&pq);
                vdist[j] = sqrt(dist_dim(t->p[i].ph + 3, t->p[nnid[j]].ph + 3, nd));
                pq_pop(&pq);
            }
            if (bad) { rc = -1; continue; }
            pq2.n = 0;
            for (int j = 0; j < nsmooth; j++) pq_push(&pq2, -1, ORC_MAXVALUE);
            for (int j = 0; j < nsearch; j++)
                if (vdist[j] < pq_top(&pq2)) { pq_pop(&pq2); pq_push(&pq2, nnid[j], vdist[j]); }
            double hi = 0.5 * pq_top(&pq2);
            double norm = 1.0 / pow(hi, (double)(nd * 1.));
            double den = 0;
            for (int j = 0; j < nsmooth; j++) {
                double rij = pq_top(&pq2);
                den = den + wsm_at(rij, hi, t->kernres, t->kern) * norm;
                pq_pop(&pq2);
            }
            rho_by_id[t->p[i].id] = den;
        }
        free(nnid); free(vdist); pq_free(&pq); pq_free(&pq2);
    }
    return rc;
}

/* CalcVelDensityWithPhysDensity(Nsmooth, Nsearch, densityset = 1),
 * KDCalcSmoothQuantities.cxx:393-476: CalcVelDensity's selection, each term
 * times W(r_ij / h) h^-3 / rho_i * m_i with h = 0.5 * (the Nsearch-th d2)
 * (no square root, :428).  rho_by_id: the densities read from Particle::rho. */
int orc_calc_veldensity_phys(void *h, int nsmooth, int nsearch, const double *rho_by_id, double *out_by_id,
                             int nthreads) {
    otree *t = (otree *)h;
    if (nsmooth > nsearch) nsmooth = nsearch;
    int rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(min : rc)
    {
        pq_t pq, pq2;
        pq_init(&pq, nsearch);
        pq_init(&pq2, nsmooth);
        int32_t *nnid = (int32_t *)malloc(sizeof(int32_t) * (size_t)nsearch);
        double *vdist = (double *)malloc(sizeof(double) * (size_t)nsearch);
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < t->n; i++) {
            double off[6] = {0, 0, 0, 0, 0, 0};
            pq.n = 0;
            for (int j = 0; j < nsearch; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            nn_target(t, t->root, 0.0, &pq, off, i, 0, 3);
            double hp = 0.5 * pq_top(&pq);
            double pnorm = 1.0 / pow(hp, 3.0);
            int bad = 0;
            for (int j = 0; j < nsearch; j++) {
                if (pq_topq(&pq) == -1) { bad = 1; break; }
                nnid[j] = pq_topq(&pq);
                vdist[j] = sqrt(dist_dim(t->p[i].ph + 3, t->p[nnid[j]].ph + 3, 3));
                pq_pop(&pq);
            }
            if (bad) { rc = -1; continue; }
            pq2.n = 0;
            for (int j = 0; j < nsmooth; j++) pq_push(&pq2, -1, ORC_MAXVALUE);
            for (int j = 0; j < nsearch; j++)
                if (vdist[j] < pq_top(&pq2)) { pq_pop(&pq2); pq_push(&pq2, nnid[j], vdist[j]); }
            double hv = 0.5 * pq_top(&pq2);
            double vnorm = 1.0 / pow(hv, 3.0);
            double rhoi = rho_by_id[t->p[i].id], den = 0;
            for (int j = 0; j < nsmooth; j++) {
                int32_t jj = pq_topq(&pq2);
                double rij = sqrt(dist_dim(t->p[i].ph, t->p[jj].ph, 3));
                double vW = wsm_at(pq_top(&pq2), hv, t->kernres, t->kern) * vnorm;
                double W = wsm_at(rij, hp, t->kernres, t->kern) * pnorm;
                den += vW * (W / rhoi * t->p[i].m);
                pq_pop(&pq2);
            }
            out_by_id[t->p[i].id] = den;
        }
        free(nnid); free(vdist); pq_free(&pq); pq_free(&pq2);
    }
    return rc;
}

/* CalcVelDensityPosition(x, v, Nsmooth, Nsearch), KDCalcSmoothQuantities.cxx:1150-1208:
 * the Nsearch nearest of position x (position search: nothing excluded, no
 * period), their |v - v_j|, the Nsmooth smallest, h_v = max / 2 and the sum of
 * W(|dv| / h_v) h_v^-3 without masses.  q: m rows of (x, y, z); qv: m rows of
 * velocities.  A TPHYS tree built with velocities. */
int orc_calc_veldensity_position(void *h, const double *q, const double *qv, int64_t m, int nsmooth, int nsearch,
                                 double *rho, double *hout, int nthreads) {
    otree *t = (otree *)h;
    if (nsmooth > nsearch) nsmooth = nsearch;
    int rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(min : rc)
    {
        pq_t pq, pq2;
        pq_init(&pq, nsearch);
        pq_init(&pq2, nsmooth);
        int32_t *nnid = (int32_t *)malloc(sizeof(int32_t) * (size_t)nsearch);
        double *vdist = (double *)malloc(sizeof(double) * (size_t)nsearch);
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < m; i++) {
            double off[6] = {0, 0, 0, 0, 0, 0};
            pq.n = 0;
            for (int j = 0; j < nsearch; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            nn_pos(t, t->root, 0.0, &pq, off, q + 3 * i, 0, 3);
            int bad = 0;
            for (int j = 0; j < nsearch; j++) {
                if (pq_topq(&pq) == -1) { bad = 1; break; }
                nnid[j] = pq_topq(&pq);
                vdist[j] = sqrt(dist_dim(qv + 3 * i, t->p[nnid[j]].ph + 3, 3));
                pq_pop(&pq);
            }
            if (bad) { rc = -1; continue; }
            pq2.n = 0;
            for (int j = 0; j < nsmooth; j++) pq_push(&pq2, -1, ORC_MAXVALUE);
            for (int j = 0; j < nsearch; j++)
                if (vdist[j] < pq_top(&pq2)) { pq_pop(&pq2); pq_push(&pq2, nnid[j], vdist[j]); }
            double hi = 0.5 * pq_top(&pq2);
            double norm = 1.0 / pow(hi, 3.0);
            double den = 0;
            for (int j = 0; j < nsmooth; j++) {
                den += wsm_at(pq_top(&pq2), hi, t->kernres, t->kern) * norm;
                pq_pop(&pq2);
            }
            rho[i] = den;
            if (hout) hout[i] = hi;
        }
        free(nnid); free(vdist); pq_free(&pq); pq_free(&pq2);
    }
    return rc;
}

/* --------------------------------------------------------------------- FOF */
typedef struct {
    const otree *t;
    double fdist2;
    int32_t iGroup, nActive;
    int32_t *Group, *Len, *Head, *Tail, *Next, *Fifo;
    short *BucketFlag;
    int32_t iTail;
} fofctx;

/* __checkboundaries, KDNode.h:279-291 (farthest2 == -1 for these trees) */
static int flag_node(const onode *nd, double fdist2, const double *x, int numdim) {
    if (nd->maxextent2 > fdist2) return 0;
    double maxr2 = 0;
    for (int j = 0; j < numdim; j++) {
        double dista = x[j] - nd->bnd[j][0], distb = nd->bnd[j][1] - x[j];
        double dista2 = dista * dista, distb2 = distb * distb;
        maxr2 += (dista2 > distb2) ? dista2 : distb2;
    }
    return (int)(maxr2 < fdist2);
}

static inline void fof_link(fofctx *c, int64_t i, int64_t target) {
    int32_t id = c->t->p[i].id;
    c->Group[id] = c->iGroup;
    c->Fifo[c->iTail++] = (int32_t)i;
    c->Len[c->iGroup]++;
    c->Next[c->Tail[c->Head[target]]] = c->Head[i];
    c->Tail[c->Head[target]] = c->Tail[c->Head[i]];
    c->Head[i] = c->Head[target];
    if (c->iTail == c->nActive) c->iTail = 0;
}

/* FOFSearchBall: SplitNode KDSplitNode.cxx:921-988, LeafNode KDLeafNode.cxx:527-590.
 * xt: the (possibly image-shifted) target position. */
static void fof_search(fofctx *c, int32_t node, double rd, double *off, int64_t target, const double *xt) {
    const otree *t = c->t;
    const onode *nd = &t->nodes[node];
    int numdim = t->nd;
    if (nd->isleaf) {
        if (c->BucketFlag[nd->nid]) return;
        int flag = 1;
        int fl = flag_node(nd, c->fdist2, xt, numdim);
        if (fl == 1) {
            for (int64_t i = nd->start; i < nd->end; i++) {
                if (c->Group[t->p[i].id]) continue;
                fof_link(c, i, target);
            }
            c->BucketFlag[nd->nid] = 1;
        } else {
            for (int64_t i = nd->start; i < nd->end; i++) {
                if (c->Group[t->p[i].id]) continue;
                const double *xi = t->p[i].ph;
                double d2 = (xt[0] - xi[0]) * (xt[0] - xi[0]) + (xt[1] - xi[1]) * (xt[1] - xi[1]) +
                            (xt[2] - xi[2]) * (xt[2] - xi[2]);
                if (numdim == 6)
                    d2 += (xt[3] - xi[3]) * (xt[3] - xi[3]) + (xt[4] - xi[4]) * (xt[4] - xi[4]) +
                          (xt[5] - xi[5]) * (xt[5] - xi[5]);
                if (d2 < c->fdist2) fof_link(c, i, target);
                else flag = 0;
            }
            if (flag) c->BucketFlag[nd->nid] = 1;
        }
        return;
    }
    if (c->BucketFlag[nd->nid] && c->Head[target] == c->Head[nd->start]) return;
    int fl = flag_node(nd, c->fdist2, xt, numdim);
    if (fl == 1) {
        for (int64_t i = nd->start; i < nd->end; i++) {
            if (c->Group[t->p[i].id]) continue;
            int32_t id = t->p[i].id;
            c->Group[id] = c->iGroup;
            c->Fifo[c->iTail++] = (int32_t)i;
            c->Len[c->iGroup]++;
            c->Next[c->Tail[c->Head[target]]] = c->Head[i];
            c->Tail[c->Head[target]] = c->Tail[c->Head[i]];
            c->Head[i] = c->Head[target];
            c->iTail %= c->nActive;
        }
        c->BucketFlag[nd->nid] = 1;
        return;
    }
    double old_off = off[nd->cut_dim];
    double new_off = xt[nd->cut_dim] - nd->cut_val;
    if (new_off < 0) {
        fof_search(c, nd->left, rd, off, target, xt);
        rd += -old_off * old_off + new_off * new_off;
        if (rd < c->fdist2) {
            off[nd->cut_dim] = new_off;
            fof_search(c, nd->right, rd, off, target, xt);
            off[nd->cut_dim] = old_off;
        }
    } else {
        fof_search(c, nd->right, rd, off, target, xt);
        rd += -old_off * old_off + new_off * new_off;
        if (rd < c->fdist2) {
            off[nd->cut_dim] = new_off;
            fof_search(c, nd->left, rd, off, target, xt);
            off[nd->cut_dim] = old_off;
        }
    }
    if (c->BucketFlag[t->nodes[nd->left].nid] == 1 && c->BucketFlag[t->nodes[nd->right].nid] == 1)
        c->BucketFlag[nd->nid] = 1;
}

/* FOFSearchBallPeriodic, KDSplitNode.cxx:1602-1646 (image iff fdist2 > s^2) */
static void fof_search_periodic(fofctx *c, int64_t target) {
    const otree *t = c->t;
    double off[6] = {0, 0, 0, 0, 0, 0}, x0[6], xp[6];
    memcpy(x0, t->p[target].ph, sizeof(x0));
    fof_search(c, t->root, 0.0, off, target, x0);
    if (t->nodes[t->root].isleaf) return; /* LeafNode periodic == plain search */
    const double *p = t->period;
    double sval;
    memcpy(xp, x0, sizeof(xp));
    for (int k = 0; k < 3; k++) {
        memset(off, 0, sizeof(off));
        sval = refl1(x0, xp, p, k);
        if (c->fdist2 > sval * sval) fof_search(c, t->root, 0.0, off, target, xp);
    }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int q = 0; q < 3; q++) {
        memset(off, 0, sizeof(off));
        sval = refl2(x0, xp, p, pairs[q][0], pairs[q][1]);
        if (c->fdist2 > sval * sval) fof_search(c, t->root, 0.0, off, target, xp);
    }
    memset(off, 0, sizeof(off));
    sval = reflN(x0, xp, p, 3);
    if (c->fdist2 > sval * sval) fof_search(c, t->root, 0.0, off, target, xp);
}

/* splay, KDFOF.cxx:18-26 */
static void splay(int32_t *Fifo, int32_t iTail, int32_t iHead, int64_t n) {
    if (iHead == iTail) return;
    int32_t iTail2 = iTail - 1;
    if (iTail2 == -1) iTail2 = (int32_t)n - 1;
    int32_t d = Fifo[iTail2];
    Fifo[iTail2] = Fifo[iHead];
    Fifo[iHead] = d;
}

/* FOF, KDFOF.cxx:29-153.  labels_by_id: the reference's raw labels (discovery
 * order, or size order through the heap when order=1).  Returns ngroups. */
int64_t orc_fof_tree(void *h, double fdist, int minnum, int order, int32_t *labels_by_id) {
    otree *t = (otree *)h;
    int64_t n = t->n;
    fofctx c;
    memset(&c, 0, sizeof(c));
    c.t = t;
    c.fdist2 = fdist * fdist;
    c.nActive = (int32_t)n;
    c.Group = labels_by_id;
    int32_t *GroupHead = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    c.Fifo = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    c.BucketFlag = (short *)calloc((size_t)t->nnodes, sizeof(short));
    c.Head = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    c.Next = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    c.Len = (int32_t *)calloc((size_t)(n + 1), sizeof(int32_t));
    c.Tail = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t iHead = 0;
    for (int64_t i = 0; i < n; i++) {
        c.Head[i] = c.Tail[i] = (int32_t)i;
        c.Next[i] = -1;
        c.Group[t->p[i].id] = 0;
    }
    for (int64_t i = 0; i < n; i++) {
        int32_t id = t->p[i].id;
        if (c.Group[id] != 0) continue;
        c.Group[id] = ++c.iGroup;
        c.Len[c.iGroup] = 1;
        GroupHead[c.iGroup] = (int32_t)i;
        c.Fifo[c.iTail++] = (int32_t)i;
        if (c.iTail == n) c.iTail = 0;
        while (iHead != c.iTail) {
            int32_t iid = c.Fifo[iHead++];
            if (iHead == n) iHead = 0;
            if (!t->periodic) {
                double off[6] = {0, 0, 0, 0, 0, 0};
                fof_search(&c, t->root, 0.0, off, iid, t->p[iid].ph);
            } else {
                fof_search_periodic(&c, iid);
            }
            splay(c.Fifo, c.iTail, iHead, n);
        }
        if (c.Len[c.iGroup] < minnum) {
            int32_t ii = c.Head[GroupHead[c.iGroup]];
            do {
                c.Group[t->p[ii].id] = -1;
            } while ((ii = c.Next[ii]) != -1);
            c.Len[c.iGroup--] = 0;
        }
    }
    for (int64_t i = 0; i < n; i++) if (c.Group[t->p[i].id] == -1) c.Group[t->p[i].id] = 0;
    int64_t ng = c.iGroup;
    if (ng > 0 && order) {
        /* size ordering through the PriorityQueue, KDFOF.cxx:117-145 */
        int32_t **plist = (int32_t **)malloc(sizeof(int32_t *) * (size_t)(ng + 1));
        int32_t *pcount = (int32_t *)calloc((size_t)(ng + 1), sizeof(int32_t));
        for (int64_t g = 1; g <= ng; g++) plist[g] = (int32_t *)malloc(sizeof(int32_t) * (size_t)c.Len[g]);
        for (int64_t i = 0; i < n; i++) {
            int32_t gid = c.Group[t->p[i].id];
            if (gid > 0) plist[gid][pcount[gid]++] = (int32_t)i;
        }
        pq_t pq;
        pq_init(&pq, (int32_t)ng);
        for (int64_t g = 1; g <= ng; g++) pq_push(&pq, (int32_t)g, (double)c.Len[g]);
        for (int64_t g = 1; g <= ng; g++) {
            int32_t gid = pq_topq(&pq);
            pq_pop(&pq);
            for (int32_t j = 0; j < c.Len[gid]; j++) c.Group[t->p[plist[gid][j]].id] = (int32_t)g;
            free(plist[gid]);
        }
        pq_free(&pq);
        free(plist);
        free(pcount);
    }
    free(GroupHead); free(c.Fifo); free(c.BucketFlag); free(c.Head); free(c.Next); free(c.Len); free(c.Tail);
    return ng;
}

/* ============================================================ brute force */
/* Candidate list kept sorted ascending by (d2, idx) */
typedef struct { double d; int32_t j; } cand;
static inline int cand_less(double d, int32_t j, const cand *c) { return d < c->d || (d == c->d && j < c->j); }
static void cand_insert(cand *list, int K, double d, int32_t j) {
    if (!cand_less(d, j, &list[K - 1])) return;
    int p = K - 1;
    while (p > 0 && cand_less(d, j, &list[p - 1])) { list[p] = list[p - 1]; p--; }
    list[p].d = d;
    list[p].j = j;
}

static inline double bf_d2(const double *a, const double *b, int nd) {
    return nd == 6 ? dist_phase(a, b) : dist_dim(a, b, 3);
}

/* mode 0: FindNearest(tt) non-periodic (also CalcDensity's search);
 * mode 1: periodic FindNearest(tt); mode 2: periodic FindNearestPos(tt).
 * phase: n x 6 row-major.  nd: 3 or 6 (6 = FindNearestPhase metric).
 * Output by input index, ascending (d2, id). */
void orc_knn_brute(const double *phase, int64_t n, int nd, int k, int mode, const double *period,
                   int32_t *ids, double *d2out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        int K = k + (mode ? 1 : 0);
        cand *L = (cand *)malloc(sizeof(cand) * (size_t)K);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; i++) {
            for (int j = 0; j < K; j++) { L[j].d = ORC_MAXVALUE; L[j].j = -1; }
            const double *xi = phase + 6 * i;
            if (mode == 0) {
                for (int64_t j = 0; j < n; j++) {
                    if (j == i) continue;
                    double d = bf_d2(xi, phase + 6 * j, nd);
                    if (d > 0) cand_insert(L, K, d, (int32_t)j);
                }
            } else {
                /* sequential image rule of KDSplitNode.cxx:1119-1148 applied to
                 * every point: the best K over the union is order independent */
                double q[6], xp[3];
                memcpy(q, xi, sizeof(q));
                /* 6-D (FindNearestPhasePeriodic, KDSplitNode.cxx:1153-1179)
                 * searches every image without a test */
                const int uncond = (nd == 6);
                for (int j = 0; j < n; j++) cand_insert(L, K, bf_d2(q, phase + 6 * j, nd), (int32_t)j);
                for (int kk = 0; kk < 3; kk++) {
                    double s = refl1(xi, xp, period, kk);
                    if (uncond || sqrt(L[K - 1].d) > s) {
                        q[0] = xp[0]; q[1] = xp[1]; q[2] = xp[2];
                        for (int64_t j = 0; j < n; j++) cand_insert(L, K, bf_d2(q, phase + 6 * j, nd), (int32_t)j);
                    }
                }
                const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
                for (int c = 0; c < 3; c++) {
                    double s = refl2(xi, xp, period, pairs[c][0], pairs[c][1]);
                    if (uncond || L[K - 1].d > s) {
                        q[0] = xp[0]; q[1] = xp[1]; q[2] = xp[2];
                        for (int64_t j = 0; j < n; j++) cand_insert(L, K, bf_d2(q, phase + 6 * j, nd), (int32_t)j);
                    }
                }
                double s = reflN(xi, xp, period, 3);
                if (uncond || L[K - 1].d > s) {
                    q[0] = xp[0]; q[1] = xp[1]; q[2] = xp[2];
                    for (int64_t j = 0; j < n; j++) cand_insert(L, K, bf_d2(q, phase + 6 * j, nd), (int32_t)j);
                }
            }
            int off = 0;
            if (mode == 1) off = 1; /* drop the smallest */
            for (int j = 0; j < k; j++) {
                ids[i * k + j] = L[j + off].j;
                if (d2out) d2out[i * k + j] = L[j + off].d;
            }
        }
        free(L);
    }
}

/* Symmetric density from neighbour lists by ID (rows ascending, last = k-th),
 * long-double accumulation (KDCalcSmoothQuantities.cxx:260-286 restated). */
void orc_density_from_lists(const double *phase, const double *mass, int64_t n, int nd, int k, const int32_t *ids,
                            const double *kern, int kernres, int symmetric, double *rho, double *hout) {
    long double *acc = (long double *)calloc((size_t)n, sizeof(long double));
    for (int64_t i = 0; i < n; i++) {
        const double *xi = phase + 6 * i;
        double dk = bf_d2(xi, phase + 6 * (int64_t)ids[i * k + k - 1], nd);
        double hi = 0.5 * sqrt(dk);
        double norm = 1.0 / pow(hi, (double)nd);
        if (hout) hout[i] = hi;
        for (int j = 0; j < k; j++) {
            int32_t q = ids[i * k + j];
            double rij = sqrt(bf_d2(xi, phase + 6 * (int64_t)q, nd));
            double W = wsm_at(rij, hi, kernres, kern) * norm;
            double mi = mass ? mass[i] : 1.0 / (double)n, mq = mass ? mass[q] : 1.0 / (double)n;
            if (symmetric) {
                acc[i] += (long double)(0.5 * W) * mq;
                acc[q] += (long double)(0.5 * W) * mi;
            } else {
                acc[i] += (long double)W * mq;
            }
        }
    }
    for (int64_t i = 0; i < n; i++) rho[i] = (double)acc[i];
    free(acc);
}

/* union-find helpers */
static int64_t uf_find(int64_t *par, int64_t x) {
    while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
    return x;
}
static void uf_union(int64_t *par, int64_t a, int64_t b) {
    a = uf_find(par, a); b = uf_find(par, b);
    if (a == b) return;
    if (a < b) par[b] = a; else par[a] = b;
}

/* Canonical FOF labels: 0 = in no group of size >= minnum; groups numbered
 * 1..G by size descending, ties by smaller minimum member ID.  Input labels:
 * any integers, equal label = same group, label <= 0 = ungrouped singleton. */
int64_t orc_fof_canonical(int64_t n, const int32_t *labels, int minnum, int32_t *out) {
    int64_t maxl = 0;
    for (int64_t i = 0; i < n; i++) if (labels[i] > maxl) maxl = labels[i];
    int64_t *size = (int64_t *)calloc((size_t)(maxl + 1), sizeof(int64_t));
    int64_t *minid = (int64_t *)malloc(sizeof(int64_t) * (size_t)(maxl + 1));
    for (int64_t g = 0; g <= maxl; g++) minid[g] = INT64_MAX;
    for (int64_t i = 0; i < n; i++) {
        int32_t g = labels[i];
        if (g <= 0) continue;
        size[g]++;
        if (i < minid[g]) minid[g] = i;
    }
    int64_t G = 0;
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(maxl + 1));
    for (int64_t g = 1; g <= maxl; g++) if (size[g] >= minnum && size[g] > 0) order[G++] = g;
    /* insertion-free: simple sort by (size desc, minid asc) */
    for (int64_t a = 1; a < G; a++) {
        int64_t v = order[a], b = a - 1;
        while (b >= 0 && (size[order[b]] < size[v] || (size[order[b]] == size[v] && minid[order[b]] > minid[v]))) {
            order[b + 1] = order[b];
            b--;
        }
        order[b + 1] = v;
    }
    int32_t *map = (int32_t *)calloc((size_t)(maxl + 1), sizeof(int32_t));
    for (int64_t r = 0; r < G; r++) map[order[r]] = (int32_t)(r + 1);
    for (int64_t i = 0; i < n; i++) out[i] = labels[i] > 0 ? map[labels[i]] : 0;
    free(size); free(minid); free(order); free(map);
    return G;
}

/* link test with the reference arithmetic: primary distance, or an image of
 * the first particle shifted toward the far wall on the axes in `mask`
 * (DistFunc.h:326-355 forms the image point first, then the distance). */
/* d2 between phase points a and b in the reference FOF's association
 * (KDLeafNode.cxx:570-571): DistanceSqd of the positions, plus DistanceSqd of
 * the velocities on a 6-D tree. */
static inline double fof_d2(const double *a, const double *b, int nd) {
    double d2 = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
    if (nd == 6)
        d2 += (a[3] - b[3]) * (a[3] - b[3]) + (a[4] - b[4]) * (a[4] - b[4]) + (a[5] - b[5]) * (a[5] - b[5]);
    return d2;
}

/* Does a link to b, directly or through one of the images the reference
 * searches (KDSplitNode.cxx:1602-1646: image iff b^2 > s^2; only the position
 * is shifted)? */
static int link_dir(const double *a, const double *b, double b2, const double *p, int periodic, int nd) {
    if (fof_d2(a, b, nd) < b2) return 1;
    if (!periodic) return 0;
    double xp[6], s;
    for (int j = 3; j < 6; j++) xp[j] = a[j];
    for (int k = 0; k < 3; k++) {
        s = refl1(a, xp, p, k);
        if (b2 > s * s && fof_d2(xp, b, nd) < b2) return 1;
    }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int c = 0; c < 3; c++) {
        s = refl2(a, xp, p, pairs[c][0], pairs[c][1]);
        if (b2 > s * s && fof_d2(xp, b, nd) < b2) return 1;
    }
    s = reflN(a, xp, p, 3);
    if (b2 > s * s && fof_d2(xp, b, nd) < b2) return 1;
    return 0;
}

/* SearchBallPos(x | tt, fdist2, imark, nn, dist2) restated over all particles
 * (KDFindNearest.cxx:567-583, leaf rules KDLeafNode.cxx:247-300): every
 * particle with d2 < r2 gets group[id] = imark and d2out[id] = d2, where d2 =
 * sum over x, y, z of (x - p)^2.  Periodic (SearchBallPosPeriodic,
 * KDSplitNode.cxx:1341-1459): the images with r2 > s^2 are searched after the
 * primary pass, in the reference's order, and a later pass overwrites.  The
 * particle with ID exclude is skipped (the target of the tt form). */
void orc_ball_mark(const double *phase, int64_t n, const double *x, int64_t exclude, double r2,
                   const double *period, int32_t imark, int32_t *group, double *d2out) {
    double pts[8][3];
    int np = 0;
    pts[np][0] = x[0]; pts[np][1] = x[1]; pts[np][2] = x[2]; np++;
    if (period) {
        double xp[3], s;
        for (int k = 0; k < 3; k++) {
            s = refl1(x, xp, period, k);
            if (r2 > s * s) { memcpy(pts[np], xp, sizeof(xp)); np++; }
        }
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int c = 0; c < 3; c++) {
            s = refl2(x, xp, period, pairs[c][0], pairs[c][1]);
            if (r2 > s * s) { memcpy(pts[np], xp, sizeof(xp)); np++; }
        }
        s = reflN(x, xp, period, 3);
        if (r2 > s * s) { memcpy(pts[np], xp, sizeof(xp)); np++; }
    }
    for (int q = 0; q < np; q++)
        for (int64_t i = 0; i < n; i++) {
            if (i == exclude) continue;
            const double *pi = phase + 6 * i;
            double d2 = 0.0;
            for (int d = 0; d < 3; d++) d2 += (pts[q][d] - pi[d]) * (pts[q][d] - pi[d]);
            if (d2 < r2) { group[i] = imark; d2out[i] = d2; }
        }
}

/* All-pairs FOF on nd = 3 or 6 phase coordinates (n x 6 rows), canonical labels. */
int64_t orc_fof_brute_nd(const double *phase, int64_t n, int nd, double b, int minnum, const double *period,
                         int32_t *labels) {
    int64_t *par = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) par[i] = i;
    double b2 = b * b;
    int periodic = period != NULL;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            const double *a = phase + 6 * i, *c = phase + 6 * j;
            if (link_dir(a, c, b2, period, periodic, nd) || (periodic && link_dir(c, a, b2, period, periodic, nd)))
                uf_union(par, i, j);
        }
    int32_t *raw = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) raw[i] = (int32_t)(uf_find(par, i) + 1);
    int64_t G = orc_fof_canonical(n, raw, minnum, labels);
    free(par); free(raw);
    return G;
}

/* The built-in FOFCriterion comparators (FOFFunc.h:30-55) on phase rows:
 * FOF3d (cmp 1) sums (x_j - y_j)^2 / p[6] over j; FOF6d (cmp 2) adds
 * (v_j - w_j)^2 / p[7] after each position term; both link when the total is
 * below 1. */
static int cmp_link(const double *a, const double *b, int cmp, const double *p) {
    double total = 0;
    for (int j = 0; j < 3; j++) {
        total += (a[j] - b[j]) * (a[j] - b[j]) / p[6];
        if (cmp == 2) total += (a[3 + j] - b[3 + j]) * (a[3 + j] - b[3 + j]) / p[7];
    }
    return total < 1;
}
/* FOFSearchCriterionPeriodic (KDSplitNode.cxx:1649-1680): the target, then its
 * seven images (positions moved by PeriodicReflection*, velocities kept), all
 * tried without a distance test. */
static int cmp_link_dir(const double *a, const double *b, int cmp, const double *p, const double *period) {
    if (cmp_link(a, b, cmp, p)) return 1;
    if (!period) return 0;
    double xp[6];
    for (int j = 3; j < 6; j++) xp[j] = a[j];
    for (int k = 0; k < 3; k++) {
        refl1(a, xp, period, k);
        if (cmp_link(xp, b, cmp, p)) return 1;
    }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int c = 0; c < 3; c++) {
        refl2(a, xp, period, pairs[c][0], pairs[c][1]);
        if (cmp_link(xp, b, cmp, p)) return 1;
    }
    reflN(a, xp, period, 3);
    return cmp_link(xp, b, cmp, p);
}

/* KDTree::FOFCriterion (KDFOF.cxx:157-265) restated as the transitive closure
 * of the comparator over all pairs, canonical labels.  The reference's tree
 * search reaches every such pair when its search scales params[1] / params[2]
 * are no smaller than the comparator's params[6] / params[7]. */
int64_t orc_fof_criterion_brute(const double *phase, int64_t n, int cmp, const double *params, int minnum,
                                const double *period, int32_t *labels) {
    int64_t *par = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) par[i] = i;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            const double *a = phase + 6 * i, *c = phase + 6 * j;
            if (cmp_link_dir(a, c, cmp, params, period) || (period && cmp_link_dir(c, a, cmp, params, period)))
                uf_union(par, i, j);
        }
    int32_t *raw = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) raw[i] = (int32_t)(uf_find(par, i) + 1);
    int64_t G = orc_fof_canonical(n, raw, minnum, labels);
    free(par); free(raw);
    return G;
}

int64_t orc_fof_brute(const double *phase, int64_t n, double b, int minnum, const double *period, int32_t *labels) {
    return orc_fof_brute_nd(phase, n, 3, b, minnum, period, labels);
}

/* ===================================================== differentiation form */
/* Analytic C1 cubic spline in 3-D, kernnorm 1/pi (SmoothingKernels.h:32-37
 * with the KSPH normalisation of KDTree.cxx:1158-1164), and its derivative. */
static inline double wq(double q) {
    if (q <= 1.0) return 1.0 - 1.5 * q * q + 0.75 * q * q * q;
    if (q <= 2.0) { double t = 2.0 - q; return 0.25 * t * t * t; }
    return 0.0;
}
static inline double dwq(double q) {
    if (q <= 1.0) return -3.0 * q + 2.25 * q * q;
    if (q <= 2.0) { double t = 2.0 - q; return -0.75 * t * t; }
    return 0.0;
}

/* Symmetric density over a frozen list (pos: n x 3 row-major; nbr[i*k+j] by
 * index; the last column is the k-th neighbour that sets h_i = |x_i-x_K|/2). */
void orc_density_analytic(const double *pos, const double *mass, int64_t n, const int32_t *nbr, int k, double *rho,
                          double *hout) {
    const double sig = M_1_PI;
    for (int64_t i = 0; i < n; i++) rho[i] = 0;
    for (int64_t i = 0; i < n; i++) {
        const double *xi = pos + 3 * i;
        const double *xk = pos + 3 * (int64_t)nbr[i * k + k - 1];
        double dk2 = (xi[0] - xk[0]) * (xi[0] - xk[0]) + (xi[1] - xk[1]) * (xi[1] - xk[1]) + (xi[2] - xk[2]) * (xi[2] - xk[2]);
        double h = 0.5 * sqrt(dk2);
        if (hout) hout[i] = h;
        double norm = sig / (h * h * h);
        for (int j = 0; j < k; j++) {
            int64_t q = nbr[i * k + j];
            const double *xj = pos + 3 * q;
            double r = sqrt((xi[0] - xj[0]) * (xi[0] - xj[0]) + (xi[1] - xj[1]) * (xi[1] - xj[1]) + (xi[2] - xj[2]) * (xi[2] - xj[2]));
            double w = 0.5 * wq(r / h) * norm;
            rho[i] += w * mass[q];
            rho[q] += w * mass[i];
        }
    }
}

/* Hand-derived adjoint of orc_density_analytic: given g = dL/drho, returns
 * dL/dpos (n x 3) and dL/dmass (n).  w_ij = (sig/2) W(r/h) h^-3;
 * dw/dr = (sig/2) W'(q) h^-4; dw/dh = -(sig/2) h^-4 (q W'(q) + 3 W(q));
 * h = |x_i - x_K|/2. */
void orc_density_adjoint(const double *pos, const double *mass, int64_t n, const int32_t *nbr, int k, const double *g,
                         double *dpos, double *dmass) {
    const double sig = M_1_PI;
    for (int64_t i = 0; i < 3 * n; i++) dpos[i] = 0;
    for (int64_t i = 0; i < n; i++) dmass[i] = 0;
    for (int64_t i = 0; i < n; i++) {
        const double *xi = pos + 3 * i;
        int64_t K = nbr[i * k + k - 1];
        const double *xk = pos + 3 * K;
        double dxk[3] = {xi[0] - xk[0], xi[1] - xk[1], xi[2] - xk[2]};
        double rk = sqrt(dxk[0] * dxk[0] + dxk[1] * dxk[1] + dxk[2] * dxk[2]);
        double h = 0.5 * rk;
        double h4 = h * h * h * h;
        double Hs = 0; /* dL/dh_i */
        for (int j = 0; j < k; j++) {
            int64_t q = nbr[i * k + j];
            const double *xj = pos + 3 * q;
            double dx[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
            double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double qq = r / h;
            double W = wq(qq), dW = dwq(qq);
            double w = 0.5 * sig * W / (h * h * h);
            double a = g[i] * mass[q] + g[q] * mass[i];
            dmass[q] += w * g[i];
            dmass[i] += w * g[q];
            double dwdr = 0.5 * sig * dW / h4;
            double dwdh = -0.5 * sig * (qq * dW + 3.0 * W) / h4;
            Hs += a * dwdh;
            if (r > 0) {
                double f = a * dwdr / r;
                for (int c = 0; c < 3; c++) { dpos[3 * i + c] += f * dx[c]; dpos[3 * q + c] -= f * dx[c]; }
            }
        }
        if (rk > 0) {
            double f = Hs * 0.5 / rk;
            for (int c = 0; c < 3; c++) { dpos[3 * i + c] += f * dxk[c]; dpos[3 * K + c] -= f * dxk[c]; }
        }
    }
}

/* ============================================================ CPU baseline */
/* kNN + gather density (CalcDensityParticle, KDCalcSmoothQuantities.cxx:768-844)
 * for m sampled tree slots on all host cores: the bounded CPU sample bench.py
 * times.  Returns the number of queries done. */
int64_t orc_knn_density_sample(void *h, int k, const int64_t *slots, int64_t m, double *rho, int nthreads) {
    otree *t = (otree *)h;
    int nd = t->nd;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        pq_t pq;
        pq_init(&pq, k);
#pragma omp for schedule(guided)
        for (int64_t q = 0; q < m; q++) {
            int64_t i = slots[q];
            double off[6] = {0, 0, 0, 0, 0, 0}, den = 0.;
            pq.n = 0;
            for (int j = 0; j < k; j++) pq_push(&pq, -1, ORC_MAXVALUE);
            nn_target(t, t->root, 0.0, &pq, off, i, t->treetype == ORC_TPHS, nd);
            double hi = 0.5 * sqrt(pq_top(&pq));
            double norm = 1.0 / pow(hi, (double)(nd * 1.));
            for (int j = 0; j < k; j++) {
                if (pq_topq(&pq) < 0) break;
                double rij = sqrt(pq_top(&pq));
                den += wsm_at(rij, hi, t->kernres, t->kern) * norm * t->p[pq_topq(&pq)].m;
                pq_pop(&pq);
            }
            rho[q] = den;
        }
        pq_free(&pq);
    }
    return m;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
