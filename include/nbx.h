/*
 * nbx.h -- C ABI of the MI355X (gfx950) KDTree hot path.
 *
 * Plain C types only (pointers, sizes, status codes).  Every entry point
 * names the NBodylib interface it replaces (file:line in pelahi/NBodylib).
 * Host-pointer entry points copy in/out; the *_dev variants take device
 * pointers and a hipStream_t passed as void* and leave results in HBM.
 *
 * Index spaces (as in the reference):
 *   ID    = position of a particle in the caller's input array (the reference
 *           overwrites Particle::id with it, KDTree.cxx:1291).
 *   slot  = position in this library's tree order (kd split order).  It plays
 *           the role of the reference's "tree index" (KDTree.h:295).
 *
 * All functions return NBX_OK (0) or a negative status; nbx_last_error()
 * gives the message (thread-local).  The reference prints and exit()s in
 * these cases (KDCalcSmoothQuantities.cxx:205-207, :273-276,
 * PriorityQueue.h:49-53, KDTree.cxx:1096-1105).
 *
 * Scratch kept by a tree: the density passes, the precision-form searches
 * and FOF take their device temporaries from one block that the tree keeps
 * from call to call, so repeated calls allocate nothing after the first.
 * The block grows to the largest need seen (densities: 8 bytes per particle,
 * plus 4 bytes per particle and about 50 MB of search scratch in the
 * precision form; FOF: 16 bytes per particle and 24 per possible group of
 * minnum members) and is freed by
 * nbx_tree_destroy.  Calls on one tree from several threads or streams take
 * turns with the block; each waits on the device for the previous one's use.
 */
#ifndef NBX_H
#define NBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBX_OK 0
#define NBX_ERR_ARG (-1)         /* invalid argument */
#define NBX_ERR_HIP (-2)         /* HIP runtime error */
#define NBX_ERR_UNSUPPORTED (-3) /* option the reference has but this path does not implement */
#define NBX_ERR_NOMEM (-4)       /* device allocation failed */
#define NBX_ERR_SEARCH (-5)      /* fewer than k neighbours ("CalcDensity failed") */

/* KDTree::TPHYS/TVEL/TPHS (KDTree.h:90-91) */
#define NBX_TPHYS 0
#define NBX_TVEL 2 /* 3-D tree on the velocities (KDTree.cxx:1116-1122, ND = 3 at :1154) */
#define NBX_TPHS 3
/* KDTree::KSPH/KGAUSS/KEPAN/KTH (KDTree.h:97) */
#define NBX_KSPH 0
#define NBX_KGAUSS 1
#define NBX_KEPAN 2
#define NBX_KTH 3
/* arithmetic forms */
#define NBX_FORM_PLAIN 0 /* FP64 throughout: the drop-in */
#define NBX_FORM_MIXED 1 /* FP32 node-relative search + FP64 re-check band */

typedef struct nbx_tree nbx_tree;

/* Mirrors the KDTree constructor options (KDTree.h:229-245). */
typedef struct nbx_params {
    int32_t bucket_size; /* bucket_size (default 16); only the N<=b edge case depends on it */
    int32_t treetype;    /* NBX_TPHYS, NBX_TVEL (non-periodic) or NBX_TPHS */
    int32_t kerntype;    /* NBX_KSPH ... (reference default KEPAN) */
    int32_t kernres;     /* kernel table size (default 1000; <100 clamps to 100) */
    int32_t splitting;   /* only KDTREE_SPLIT_SPREAD (0) */
    int32_t aniso;       /* TPHS: -1 only (metric factor 1 on caller-prescaled x/sx, v/sv) */
    int32_t scalespace;  /* must be 0 (the reference's ScaleSpace reads uninitialised xmean) */
    int32_t periodic;    /* 0 or 1 */
    double period[3];    /* Period[3] when periodic */
    int32_t form;        /* NBX_FORM_PLAIN or NBX_FORM_MIXED */
    double recheck_rel;  /* mixed form: least relative half-width of the FP64 re-check band around the
                          * k-th distance (default 1e-5; the proven FP32 error bound is used where wider);
                          * a wider band re-checks more candidates and never changes a result */
} nbx_params;

/* Defaults of KDTree::KDTree (KDTree.h:229-245). */
void nbx_params_default(nbx_params *p);

/* Builds the tree.  Replaces KDTree::KDTree(Particle*,N,...) KDTree.cxx:1238-1306
 * (BuildNodes :970-1055, KernelConstruction :1144-1183).
 * pos: 3*n SoA (x[0..n), y[..], z[..]); vel: 3*n SoA or NULL (needed for TPHS
 * and velocity density); mass: n or NULL (then 1/n each).
 * TVEL: the tree is built on vel (pos is not read and may be NULL); kNN
 * (NBX_KNN_FINDNEAREST = FindNearestVel(tt), KDFindNearest.cxx:335-346,
 * :258; nbx_knn_points = FindNearestVel(v), :530-541), CalcDensity,
 * CalcDensityParticle and CalcDensityPosition then run on velocity
 * coordinates as in the reference (KDCalcSmoothQuantities.cxx:228, :783,
 * :1107).  Calls that need positions (FOF, ball searches, smoothing,
 * CalcVelDensity, FindNearestPos(tt)) return NBX_ERR_UNSUPPORTED on it, and
 * so does a periodic TVEL tree (the reference applies the period to
 * velocity images there). */
int nbx_tree_create(const nbx_params *p, const double *pos, const double *vel, const double *mass, int64_t n,
                    nbx_tree **out);
int nbx_tree_create_dev(const nbx_params *p, const double *pos_dev, const double *vel_dev, const double *mass_dev,
                        int64_t n, void *stream, nbx_tree **out);
/* Replaces KDTree::~KDTree (KDTree.cxx:1340-1356); caller-side reorder is the
 * shim's job.  The tree's arrays go back with hipFree, which waits for work
 * still running on the device, so no launch on any stream may still be using
 * the tree. */
int nbx_tree_destroy(nbx_tree *t);
/* slot -> ID map (host array of n).  The C++ shim uses it to reorder the
 * caller's Particle array as the reference does in BuildNodes. */
int nbx_tree_order(nbx_tree *t, int32_t *slot_to_id);
int64_t nbx_tree_size(const nbx_tree *t);
/* Kernel normalisation and table (KDTree::GetKernNorm, KDTree.h:283). */
int nbx_tree_kernel(const nbx_tree *t, double *kernnorm, double *table, int32_t cap);

/* kNN of every particle.  Replaces the batch loops
 * KDTree::FindNearest(Int_t**,Double_t**,k) KDFindNearest.cxx:444-459 and
 * the per-particle FindNearest(tt) :247-318, FindNearestPos(tt) :320-334,
 * FindNearestPhase(tt) :347-361. */
#define NBX_KNN_FINDNEAREST 0      /* self and d2==0 excluded; periodic: k+1 search, drop smallest */
#define NBX_KNN_FINDNEARESTPOS 1   /* periodic: k+1 search, drop largest (returns self first) */
#define NBX_KNN_FINDNEARESTPHASE 2 /* 6-D phase metric (TPHS trees) */
#define NBX_KNN_NONPERIODIC 3      /* the search CalcDensity makes (KDCalcSmoothQuantities.cxx:227) */
#define NBX_LAYOUT_BY_ID 0   /* row r = particle with ID r; entries are IDs */
#define NBX_LAYOUT_BY_SLOT 1 /* row r = slot r; entries are slots ("tree indices") */
/* nn: n*k int32, d2: n*k double or NULL; rows ascending in d2 (LoadNN,
 * KDFindNearest.cxx:11-24); missing neighbours are -1 with d2 = 1e32. */
int nbx_knn(nbx_tree *t, int32_t k, int32_t variant, int32_t layout, int32_t *nn, double *d2);
int nbx_knn_dev(nbx_tree *t, int32_t k, int32_t variant, int32_t layout, int32_t *nn_dev, double *d2_dev,
                void *stream);
/* kNN of a list of particles (ids: m IDs).  Replaces m per-particle calls
 * FindNearest(tt) / FindNearestPos(tt) / FindNearestPhase(tt)
 * (KDFindNearest.cxx:247-361) with the variant's rules as in nbx_knn; row r
 * belongs to particle ids[r], entries are IDs (NBX_LAYOUT_BY_ID) or slots
 * (NBX_LAYOUT_BY_SLOT).  nn: m*k; d2: m*k or NULL. */
int nbx_knn_targets(nbx_tree *t, const int32_t *ids, int64_t m, int32_t k, int32_t variant, int32_t layout,
                    int32_t *nn, double *d2);
/* kNN of arbitrary positions (q: 3*m SoA).  Replaces FindNearestPos(Double_t*x,...)
 * KDFindNearest.cxx:462-564 (period-aware like the reference).  Entries are IDs. */
int nbx_knn_points(nbx_tree *t, const double *q, int64_t m, int32_t k, int32_t *nn, double *d2);
/* kNN of arbitrary phase-space points (x, v: 3*m SoA each), TPHS trees.
 * Replaces FindNearestPhase(Double_t *x, Double_t *v, ...) KDFindNearest.cxx:528-538
 * and FindNearest(Double_t *x, ...) on a TPHS tree (:462-509, x[3..5] = v):
 * 6-D distances, nothing excluded, and on a periodic tree every one of the 7
 * spatial images searched (KDSplitNode.cxx:1153-1179).  Entries are IDs. */
int nbx_knn_phase_points(nbx_tree *t, const double *x, const double *v, int64_t m, int32_t k, int32_t *nn,
                         double *d2);
/* Gather SPH density at arbitrary positions (q: 3*m SoA), TPHYS trees.
 * Replaces CalcDensityPosition(Double_t *x, Nsmooth, v) KDCalcSmoothQuantities.cxx:1092-1149
 * for m positions: non-periodic search with no exclusion (a particle at x is
 * its own neighbour), h = r_k/2, rho = sum_j W(r_j/h) h^-3 m_j.  rho[m];
 * h[m] nullable.  NBX_ERR_SEARCH when the tree holds fewer than k particles
 * (the reference exits). */
int nbx_density_points(nbx_tree *t, const double *q, int64_t m, int32_t k, double *rho, double *h);

/* Velocity density at arbitrary phase points (x, v: 3*m SoA each), TPHYS
 * trees built with velocities.  Replaces CalcVelDensityPosition(x, v,
 * Nsmooth, Nsearch) KDCalcSmoothQuantities.cxx:1150-1208 for m points: the
 * Nsearch nearest of x (non-periodic, nothing excluded), the Nsmooth of them
 * nearest to v, h_v = max |v - v_j| / 2, rho = sum W(|v - v_j| / h_v) h_v^-3
 * (no masses).  Nsmooth > Nsearch is taken as Nsearch.  rho[m]; hv[m]
 * nullable.  NBX_ERR_SEARCH when the tree holds fewer than Nsearch particles. */
int nbx_veldensity_points(nbx_tree *t, const double *x, const double *v, int64_t m, int32_t nsmooth, int32_t nsearch,
                          double *rho, double *hv);

/* Velocity density weighted by the physical overlap, every particle, keyed
 * by ID.  Replaces CalcVelDensityWithPhysDensity(Nsmooth, Nsearch, densityset)
 * KDCalcSmoothQuantities.cxx:393-476 (and, per particle,
 * CalcVelDensityWithPhysDensityParticle :922-990): CalcVelDensity's selection
 * (Nsearch physical neighbours, self and d2 == 0 excluded, no period; the
 * Nsmooth nearest in velocity), each term W(|dv| / h_v) h_v^-3 times
 * W(r / h) h^-3 / rho_i * m_i, where h = (d2 of the Nsearch-th neighbour) / 2
 * -- the reference takes no square root there.  rho_by_id: the densities the
 * reference reads from Particle::rho (CalcDensity(Nsmooth) when densityset
 * != 1; the caller runs it).  TPHYS trees built with velocities. */
int nbx_veldensity_phys(nbx_tree *t, int32_t nsmooth, int32_t nsearch, const double *rho_by_id, double *vden_by_id);

/* SPH densities, results keyed by ID.
 * SYMMETRIC replaces KDTree::CalcDensity KDCalcSmoothQuantities.cxx:203-305;
 * GATHER replaces CalcDensityParticle :768-844 for every particle;
 * VEL replaces CalcVelDensity :309-389 (k = Nsmooth, nsearch = Nsearch).
 * h_by_id (nullable) receives h_i = r_k/2 (the reference computes but never stores it). */
#define NBX_DENS_SYMMETRIC 0
#define NBX_DENS_GATHER 1
#define NBX_DENS_VEL 2
int nbx_density(nbx_tree *t, int32_t k, int32_t variant, int32_t nsearch, double *rho_by_id, double *h_by_id);
int nbx_density_dev(nbx_tree *t, int32_t k, int32_t variant, int32_t nsearch, double *rho_by_id_dev,
                    double *h_by_id_dev, void *stream);
/* Smoothed velocity mean and dispersion (TPHYS trees built with velocities).
 * Replaces CalcSmoothVel(Nsmooth, densityset) KDCalcSmoothQuantities.cxx:480-540
 * and CalcSmoothVelDisp(smvel, Nsmooth, densityset, meanvelset) :542-614:
 * CalcDensity's kNN (non-periodic, h = r_k/2) and half weights W_ij/2, added
 * to both ends of each pair as W_ij/2 * m/rho * v (mean) and
 * W_ij/2 * m/rho * (v - mean)(v - mean)^T (dispersion, about the receiving
 * particle's mean).  rho_by_id: the densities to divide by (densityset = 1),
 * or NULL to run CalcDensity(k) first.  mean_in_by_id: n x 3 means for the
 * dispersion (CalcSmoothVelDisp's smvel with meanvelset = 1), or NULL to
 * compute them.  smvel_by_id: n x 3 by ID, or NULL; disp_by_id: n x 9
 * (row-major 3x3) by ID, or NULL; at least one of the two. */
int nbx_smooth_vel(nbx_tree *t, int32_t k, const double *rho_by_id, const double *mean_in_by_id, double *smvel_by_id,
                   double *disp_by_id);
/* Per-component velocity skewness and kurtosis over the same pairs.  Replaces
 * CalcSmoothVelSkew(smvel, smveldisp, Nsmooth, densityset, meanvelset,
 * veldispset) :617-690 and CalcSmoothVelKurtosis :692-767: each share adds
 * share * dv^3 / s^1.5 and share * dv^4 / s^2 - 3, with dv about the receiving
 * particle's mean and s its dispersion diagonal (the 3 is taken off every
 * share, as the reference does).  mean_in_by_id (n x 3) and disp_in_by_id
 * (n x 9) replace the computed ones (meanvelset / veldispset = 1), or NULL.
 * skew_by_id, kurt_by_id: n x 3 by ID, either may be NULL (not both). */
int nbx_smooth_vel_moments(nbx_tree *t, int32_t k, const double *rho_by_id, const double *mean_in_by_id,
                           const double *disp_in_by_id, double *skew_by_id, double *kurt_by_id);
/* Local mean and dispersion of a per-particle scalar (TPHYS trees), the
 * gather form with the full weight W(r/h) h^-3: mean = sum_j W/rho_j * m_j *
 * w_j and disp = sum_j W/rho_j * m_j * (w_j - mean_in)^2 over the k nearest.
 * nbx_smooth_local replaces CalcSmoothLocalMeanParticle(target, weight,
 * Nsmooth, densityset) KDCalcSmoothQuantities.cxx:1539-1577 and
 * CalcSmoothLocalDispParticle :1579-1617 for every particle (CalcDensity's
 * neighbours); nbx_smooth_local_points replaces the Position forms
 * :1619-1693 for m positions (q: 3*m SoA; the non-periodic search from x, no
 * exclusion).  weight_by_id: n values; rho_by_id: the densities divided by
 * (densityset = 1), or NULL to run CalcDensity(k) first; mean_in: the mean
 * each dispersion is taken about (by ID, or by position row), needed only
 * for the dispersion.  Outputs by ID (or by row), each nullable. */
int nbx_smooth_local(nbx_tree *t, int32_t k, const double *weight_by_id, const double *rho_by_id,
                     const double *mean_in_by_id, double *mean_by_id, double *disp_by_id);
int nbx_smooth_local_points(nbx_tree *t, const double *q, int64_t m, int32_t k, const double *weight_by_id,
                            const double *rho_by_id, const double *mean_in, double *mean_out, double *disp_out);
/* Fused: kNN (NBX_KNN_NONPERIODIC semantics) + symmetric density in one pass,
 * neighbour IDs by slot.  The batch form of FindNearest + CalcDensity. */
int nbx_knn_density_dev(nbx_tree *t, int32_t k, int32_t *nn_by_slot_dev, double *rho_by_id_dev,
                        double *h_by_id_dev, void *stream);

/* nbx_knn_density_dev for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  No
 * counterpart in the reference.  Same arguments and results; the call makes
 * no hipMalloc / hipFree or other allocation: its temporaries come from a
 * 256 MiB device array that each GPU the library runs on holds until the
 * process ends.  n particles need 8n bytes plus, for the precision form,
 * 4n bytes and the search scratch (about 50 MB); larger trees return
 * NBX_ERR_UNSUPPORTED before anything runs.  Calls that overlap
 * on one GPU share that array and race. */
int nbx_knn_density_noalloc(nbx_tree *t, int32_t k, int32_t *nn_by_slot_dev, double *rho_by_id_dev,
                            double *h_by_id_dev, void *stream);

/* Ball search around a position: IDs of all particles with d2 < r2 (strict).
 * Replaces SearchBallPosTagged(Double_t*x, fdist2, Int_t*tagged)
 * KDFindNearest.cxx:618-688 (period-aware).  out: capacity cap; *nout = total found. */
int nbx_ball_points(nbx_tree *t, const double *x, double r2, int32_t *out, int64_t cap, int64_t *nout);
/* Ball search around m positions (q: 3*m SoA) at once: SearchBallPosTagged
 * (x, fdist2, tagged) KDFindNearest.cxx:618-688 for each row (period-aware
 * when the tree is, with a radius below half the period).  offsets[m+1]: CSR
 * offsets, offsets[m] = total hits; ids (IDs) and d2 (nullable) are written
 * only when total <= cap, else call again with a larger cap.  Order within a
 * row is unspecified (the reference's is its traversal order). */
int nbx_ball_points_many(nbx_tree *t, const double *q, int64_t m, double r2, int64_t *offsets, int32_t *ids,
                         double *d2, int64_t cap);
/* Ball search that marks: group_by_id[id] = imark and d2_by_id[id] = d2 for
 * every particle with d2 < r2 around x, except the particle with ID
 * exclude_id (pass -1 to exclude none); other entries are left as they are.
 * Replaces SearchBallPos(Double_t *x, fdist2, imark, nn, dist2) and, with
 * x = the target's position and exclude_id = its ID, SearchBallPos(Int_t tt,
 * ...) KDFindNearest.cxx:567-583 (leaf rules KDLeafNode.cxx:247-300).  On a
 * periodic tree r must be below half the period (then a particle is within r
 * of at most one image and its d2 is that image's). */
int nbx_ball_mark(nbx_tree *t, const double *x, int64_t exclude_id, double r2, int32_t imark, int32_t *group_by_id,
                  double *d2_by_id);

/* Friends-of-friends.  Replaces KDTree::FOF(fdist, numgroup, minnum, order)
 * KDFOF.cxx:29-153 (periodic when the tree is).  labels_by_id: canonical labels
 * -- 0 = in no group of >= minnum members; groups 1..G by size descending,
 * ties by smaller minimum member ID (the reference's order=1 numbering up to
 * its heap tie order). */
int nbx_fof(nbx_tree *t, double fdist, int32_t minnum, int32_t *labels_by_id, int64_t *ngroups);
int nbx_fof_dev(nbx_tree *t, double fdist, int32_t minnum, int32_t *labels_by_id_dev, int64_t *ngroups, void *stream);
/* FOF with a built-in comparator.  Replaces KDTree::FOFCriterion(cmp, params,
 * numgroups, minnum, order) KDFOF.cxx:157-265 for cmp = FOF3d or FOF6d
 * (FOFFunc.h:30-55), periodic when the tree is (KDSplitNode.cxx:1649-1680).
 * params[8] as the reference's: [0] treetype, [1] / [2] search scales for
 * positions / velocities, [6] / [7] comparator scales.  FOF3d: TPHYS tree,
 * sum_j dx_j^2 / params[6] < 1; FOF6d: TPHS tree, sum_j (dx_j^2 / params[6] +
 * dv_j^2 / params[7]) < 1.  The search scales must be >= the comparator's
 * (else the reference's groups depend on its traversal); NBX_ERR_UNSUPPORTED
 * otherwise and for any other comparator.  Labels as nbx_fof. */
/* FOF's linked-list outputs (KDTree::FOF's pHead, pNext, pTail, pLen,
 * KDFOF.cxx:29-31, :52-55) built from canonical labels on the host: arrays of n
 * by tree index (slot; slot_to_id from nbx_tree_order).  Members of group g
 * form one list in slot order: head[s] = its first slot, next[s] = the next
 * slot or -1, tail[head] = its last slot; a particle in no group is a list
 * of its own (head = tail = itself).  len[g] = the size of group g (1..G),
 * other entries 0.  Any output may be NULL.  The reference's lists run in
 * its search order.  The reference also keeps the members of groups pruned
 * below minnum linked (KDFOF.cxx:98-104): pass the labels of a minnum-1 run
 * for those lists; its groups of >= minnum members are its first labels
 * (canonical order is by size), the rest become 0. */
int nbx_fof_lists(const int32_t *labels_by_id, const int32_t *slot_to_id, int64_t n, int64_t ngroups, int32_t *head,
                  int32_t *next, int32_t *tail, int32_t *len);
#define NBX_FOF_CMP_3D 1
#define NBX_FOF_CMP_6D 2
int nbx_fof_criterion(nbx_tree *t, int32_t comparator, const double *params, int32_t minnum, int32_t *labels_by_id,
                      int64_t *ngroups);
int nbx_fof_criterion_dev(nbx_tree *t, int32_t comparator, const double *params, int32_t minnum,
                          int32_t *labels_by_id_dev, int64_t *ngroups, void *stream);

/* Differentiation form (no reference counterpart: the reference allocates
 * derKernel but never fills it, KDTree.cxx:1150).  Symmetric SPH density with
 * the analytic C1 cubic spline (SmoothingKernels.h:32-37, 3-D norm 1/pi) over a
 * frozen neighbour list nbr (n*k indices, last column = k-th neighbour that
 * sets h_i = |x_i - x_K|/2).  pos: 3*n SoA.  All device pointers. */
int nbx_sph_density_fwd_dev(const double *pos_dev, const double *mass_dev, const int32_t *nbr_dev, int64_t n,
                            int32_t k, double *rho_dev, double *h_dev, void *stream);
/* Adjoint: given g = dL/drho, writes dL/dpos (3*n SoA) and dL/dmass (n);
 * both outputs are cleared by the call first. */
int nbx_sph_density_adjoint_dev(const double *pos_dev, const double *mass_dev, const int32_t *nbr_dev, int64_t n,
                                int32_t k, const double *g_dev, double *dpos_dev, double *dmass_dev, void *stream);

/* The same forward and adjoint over the tree's own neighbour lists, with no
 * list of all n kept: each particle's k nearest under CalcDensity's search
 * (KDCalcSmoothQuantities.cxx:227: self and d2 == 0 excluded, no period) are
 * recomputed for one block of block_rows tree slots at a time (rounded up
 * to a multiple of 1024; <= 0 takes every slot at once) into a buffer of
 * block_rows * k * 4 bytes and consumed there.  The list memory is that
 * buffer instead of n * k * 4 bytes.  TPHYS trees, either form (the
 * precision form's neighbour sets are the plain form's).  All by ID, device
 * pointers: rho_by_id, h_by_id (nullable), g_by_id = dL/drho, dpos_by_id
 * (3*n SoA) and dmass_by_id, the last two written (not accumulated). */
int nbx_sph_density_fwd_tree_dev(nbx_tree *t, int32_t k, int64_t block_rows, double *rho_by_id_dev,
                                 double *h_by_id_dev, void *stream);
int nbx_sph_density_adjoint_tree_dev(nbx_tree *t, int32_t k, int64_t block_rows, const double *g_by_id_dev,
                                     double *dpos_by_id_dev, double *dmass_by_id_dev, void *stream);

/* Precision-form statistics of the last kNN/density call on this tree:
 * queries whose FP32 certificate failed and were recomputed in FP64. */
int nbx_tree_last_fallbacks(const nbx_tree *t, int64_t *fallbacks);

/* Per-kernel timing with HIP events recorded on the launch stream.  When on,
 * every library kernel launch is bracketed; nbx_prof_get returns the summed
 * milliseconds and launch count of kernels whose name contains `name`. */
void nbx_prof_enable(int on);
void nbx_prof_reset(void);
int nbx_prof_get(const char *name, double *ms, int64_t *launches);

const char *nbx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* NBX_H */
