// kdtree_shim.cpp -- NBody::KDTree drop-in over the C ABI (include/nbx_kdtree.h).
//
// Reproduces the reference's host-side contract around the hot path:
//  * KDTree::KDTree (KDTree.cxx:1238-1306): ids overwritten with the input
//    index (:1291) and the caller's array reordered into tree order;
//  * ~KDTree (:1340-1356): array sorted back by id unless SetResetOrder(false);
//  * OverWriteInputOrder (:1358-1361): ids rewritten to the current order;
//  * CalcDensity / CalcVelDensity write Particle::rho (SetDensity);
//  * FOF returns new Int_t[N] keyed by id (KDFOF.cxx:35, :152).
// All computation is in libnbx.so; this file only moves data.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../../include/nbx_kdtree.h"

namespace NBody {

static bool g_throw = false;
static bool g_mixed = false;

void KDTree::SetThrowOnError(bool on) { g_throw = on; }
void KDTree::SetPrecisionForm(bool mixed) { g_mixed = mixed; }

static void die(const std::string &msg) {
    if (g_throw) throw std::runtime_error(msg);
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(1);
}

static void check(int rc, const char *what) {
    if (rc >= 0) return;
    die(std::string(what) + ": " + nbx_last_error());
}

// tagged ball rows are cached for blocks of this many consecutive tree indices
static const int BALL_BLOCK_LOG2 = 16;
static const size_t BALL_BLOCKS_KEPT = 64;

void KDTree::init(Particle *p, Int_t n, Int_t bucket_size, int TreeType, int KernType, int KernRes, int Splitting,
                  int Aniso, int ScaleSpace, Double_t *Period) {
    bucket = p;
    numparts = n;
    b = bucket_size;
    treetype = TreeType;
    kernfunctype = KernType;
    iresetorder = true;
    tree = nullptr;
    cache_variant = cache_k = dcache_k = vcache_k = vcache_s = -1;
    pcache_k = pcache_s = pcache_set = -1;
    ball_r2 = -1.0;
    ball_tick = 0;
    periodic = Period != nullptr;
    for (int k = 0; k < 3; ++k) period[k] = periodic ? Period[k] : 0.0;
    nbx_params_default(&prm);
    prm.bucket_size = bucket_size;
    prm.treetype = TreeType;
    prm.kerntype = KernType;
    prm.kernres = KernRes;
    prm.splitting = Splitting;
    // TPHS trees: Aniso = -1 only (metric factor 1 on coordinates the caller
    // has scaled); the reference's Aniso = 0 and 1 search with a local metric
    // tensor (KDCalcSmoothQuantities.cxx:236-262), which is not provided, so
    // nbx_tree_create reports them as unsupported
    prm.aniso = Aniso;
    prm.scalespace = ScaleSpace;
    prm.periodic = periodic ? 1 : 0;
    for (int k = 0; k < 3; ++k) prm.period[k] = period[k];
    prm.form = g_mixed ? NBX_FORM_MIXED : NBX_FORM_PLAIN;
    // ids := input index (KDTree.cxx:1291); SoA copies for the device
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), mass(n);
    for (Int_t i = 0; i < n; ++i) {
        p[i].SetID(i);
        for (int d = 0; d < 3; ++d) {
            pos[d * (size_t)n + i] = p[i].GetPosition(d);
            vel[d * (size_t)n + i] = p[i].GetVelocity(d);
        }
        mass[i] = p[i].GetMass();
    }
    check(nbx_tree_create(&prm, pos.data(), vel.data(), mass.data(), n, &tree), "KDTree");
    // reorder the caller's array into tree order, as BuildNodes does in place
    std::vector<int32_t> order(n);
    check(nbx_tree_order(tree, order.data()), "KDTree");
    std::vector<Particle> copy(p, p + n);
    slot_of_id.assign(n, 0);
    for (Int_t s = 0; s < n; ++s) {
        p[s] = copy[order[s]];
        slot_of_id[order[s]] = s;
    }
    numleafnodes = (n + 31) / 32;
    int lp = 1;
    while (lp < numleafnodes) lp <<= 1;
    numnodes = 2 * lp - 1;
}

KDTree::KDTree(Particle *p, Int_t n, Int_t bucket_size, int TreeType, int KernType, int KernRes, int Splitting,
               int Aniso, int ScaleSpace, Double_t *Period, Double_t **, bool, bool, Double_t, Double_t, Int_t) {
    init(p, n, bucket_size, TreeType, KernType, KernRes, Splitting, Aniso, ScaleSpace, Period);
}

KDTree::KDTree(std::vector<Particle> &p, Int_t bucket_size, int TreeType, int KernType, int KernRes, int Splitting,
               int Aniso, int ScaleSpace, Double_t *Period, Double_t **, bool, bool, Double_t, Double_t, Int_t) {
    init(p.data(), (Int_t)p.size(), bucket_size, TreeType, KernType, KernRes, Splitting, Aniso, ScaleSpace, Period);
}

KDTree::~KDTree() {
    if (tree) nbx_tree_destroy(tree);
    tree = nullptr;
    if (iresetorder)
        std::sort(bucket, bucket + numparts, [](const Particle &a, const Particle &c) { return a.GetID() < c.GetID(); });
}

void KDTree::OverWriteInputOrder() {
    iresetorder = false;
    for (Int_t i = 0; i < numparts; ++i) bucket[i].SetID(i);
}
void KDTree::SetResetOrder(bool a) { iresetorder = a; }

Double_t KDTree::GetKernNorm() {
    double kn = 0;
    check(nbx_tree_kernel(tree, &kn, nullptr, 0), "GetKernNorm");
    return kn;
}

// ------------------------------------------------------------------ kNN
void KDTree::batch_knn(int variant, Int_t k, Int_t *nn, Double_t *d2) {
    check(nbx_knn(tree, k, variant, NBX_LAYOUT_BY_SLOT, nn, d2), "FindNearest");
}

static void scatter_rows(const std::vector<Int_t> &nn, const std::vector<Double_t> &d2, Int_t n, Int_t k, Int_t **out,
                         Double_t **dout) {
    for (Int_t i = 0; i < n; ++i)
        for (Int_t j = 0; j < k; ++j) {
            out[i][j] = nn[(size_t)i * k + j];
            dout[i][j] = d2[(size_t)i * k + j];
        }
}

void KDTree::FindNearest(Int_t **nn, Double_t **dist2, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    ensure_cache(treetype == TPHS ? NBX_KNN_FINDNEARESTPHASE : NBX_KNN_FINDNEAREST, Nsearch);
    scatter_rows(cache_nn, cache_d2, numparts, Nsearch, nn, dist2);
}
void KDTree::FindNearestPos(Int_t **nn, Double_t **dist2, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    ensure_cache(NBX_KNN_FINDNEARESTPOS, Nsearch);
    scatter_rows(cache_nn, cache_d2, numparts, Nsearch, nn, dist2);
}
void KDTree::FindNearestPhase(Int_t **nn, Double_t **dist2, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    ensure_cache(NBX_KNN_FINDNEARESTPHASE, Nsearch);
    scatter_rows(cache_nn, cache_d2, numparts, Nsearch, nn, dist2);
}

void KDTree::ensure_cache(int variant, Int_t k) {
    if (cache_variant == variant && cache_k == k) return;
    cache_variant = cache_k = -1;
    cache_nn.assign((size_t)numparts * k, -1);
    cache_d2.assign((size_t)numparts * k, 0.0);
    batch_knn(variant, k, cache_nn.data(), cache_d2.data());
    cache_variant = variant;
    cache_k = k;
}

void KDTree::FindNearest(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    ensure_cache(treetype == TPHS ? NBX_KNN_FINDNEARESTPHASE : NBX_KNN_FINDNEAREST, Nsearch);
    std::copy(cache_nn.begin() + (size_t)tt * Nsearch, cache_nn.begin() + (size_t)(tt + 1) * Nsearch, nn);
    std::copy(cache_d2.begin() + (size_t)tt * Nsearch, cache_d2.begin() + (size_t)(tt + 1) * Nsearch, dist2);
}
void KDTree::FindNearestPos(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    ensure_cache(NBX_KNN_FINDNEARESTPOS, Nsearch);
    std::copy(cache_nn.begin() + (size_t)tt * Nsearch, cache_nn.begin() + (size_t)(tt + 1) * Nsearch, nn);
    std::copy(cache_d2.begin() + (size_t)tt * Nsearch, cache_d2.begin() + (size_t)(tt + 1) * Nsearch, dist2);
}
void KDTree::FindNearestPhase(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    ensure_cache(NBX_KNN_FINDNEARESTPHASE, Nsearch);
    std::copy(cache_nn.begin() + (size_t)tt * Nsearch, cache_nn.begin() + (size_t)(tt + 1) * Nsearch, nn);
    std::copy(cache_d2.begin() + (size_t)tt * Nsearch, cache_d2.begin() + (size_t)(tt + 1) * Nsearch, dist2);
}

// one point in the tree's own 3-D space (positions, or velocities on TVEL)
void KDTree::FindNearest(Double_t *x, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    if (treetype == TPHS) {  // x[0..5] is the phase point (KDFindNearest.cxx:475-478, :500-503)
        FindNearestPhase(x, x + 3, nn, dist2, Nsearch);
        return;
    }
    std::lock_guard<std::mutex> lk(mu);
    double q[3] = {x[0], x[1], x[2]};
    std::vector<int32_t> ids(Nsearch);
    check(nbx_knn_points(tree, q, 1, Nsearch, ids.data(), dist2), "FindNearest");
    for (Int_t j = 0; j < Nsearch; ++j) nn[j] = ids[j] >= 0 ? slot_of_id[ids[j]] : -1;
}
void KDTree::FindNearestPos(Double_t *x, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    if (treetype != TPHYS) die("FindNearestPos(x) here needs a TPHYS tree");
    FindNearest(x, nn, dist2, Nsearch);
}
void KDTree::FindNearestPhase(Double_t *x, Double_t *v, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    double qx[3] = {x[0], x[1], x[2]}, qv[3] = {v[0], v[1], v[2]};
    std::vector<int32_t> ids(Nsearch);
    check(nbx_knn_phase_points(tree, qx, qv, 1, Nsearch, ids.data(), dist2), "FindNearestPhase");
    for (Int_t j = 0; j < Nsearch; ++j) nn[j] = ids[j] >= 0 ? slot_of_id[ids[j]] : -1;
}

// velocity space: the TVEL tree's own coordinates (KDFindNearest.cxx:335-346,
// :451-453, :530-541, :559-561)
void KDTree::FindNearestVel(Int_t **nn, Double_t **dist2, Int_t Nsearch) {
    if (treetype != TVEL) die("FindNearestVel needs a TVEL tree");
    FindNearest(nn, dist2, Nsearch);
}
void KDTree::FindNearestVel(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    if (treetype != TVEL) die("FindNearestVel needs a TVEL tree");
    FindNearest(tt, nn, dist2, Nsearch);
}
void KDTree::FindNearestVel(Double_t *v, Int_t *nn, Double_t *dist2, Int_t Nsearch) {
    if (treetype != TVEL) die("FindNearestVel needs a TVEL tree");
    FindNearest(v, nn, dist2, Nsearch);
}

// ------------------------------------------------------------------ ball search
// The rows of the block of tree indices holding tt, for radius^2 fdist2:
// one batched search around every position of the block (the caller holds
// the mutex).  Rows hold every particle with d2 < fdist2, the block's own
// particle included.
const KDTree::BallBlock &KDTree::ball_block(Int_t tt, Double_t fdist2) {
    if (fdist2 != ball_r2) {
        ball_cache.clear();
        ball_r2 = fdist2;
    }
    const Int_t blk = tt >> BALL_BLOCK_LOG2;
    auto it = ball_cache.find(blk);
    if (it != ball_cache.end()) {
        it->second->used = ++ball_tick;
        return *it->second;
    }
    if (ball_cache.size() >= BALL_BLOCKS_KEPT) {
        auto old = ball_cache.begin();
        for (auto j = ball_cache.begin(); j != ball_cache.end(); ++j)
            if (j->second->used < old->second->used) old = j;
        ball_cache.erase(old);
    }
    const Int_t s0 = blk << BALL_BLOCK_LOG2;
    const int64_t m = std::min<int64_t>((int64_t)1 << BALL_BLOCK_LOG2, (int64_t)numparts - s0);
    std::vector<double> q(3 * (size_t)m);
    for (int64_t i = 0; i < m; ++i)
        for (int d = 0; d < 3; ++d) q[(size_t)d * m + i] = bucket[s0 + i].GetPosition(d);
    std::unique_ptr<BallBlock> B(new BallBlock());
    B->off.assign(m + 1, 0);
    check(nbx_ball_points_many(tree, q.data(), m, fdist2, B->off.data(), nullptr, nullptr, 0), "SearchBallPosTagged");
    const int64_t total = B->off[m];
    std::vector<int32_t> ids(total > 0 ? total : 1);
    B->d2.assign(total > 0 ? total : 1, 0.0);
    if (total > 0)
        check(nbx_ball_points_many(tree, q.data(), m, fdist2, B->off.data(), ids.data(), B->d2.data(), total),
              "SearchBallPosTagged");
    B->idx.resize(total);
    for (int64_t e = 0; e < total; ++e) B->idx[e] = slot_of_id[ids[e]];
    B->used = ++ball_tick;
    const BallBlock &ref = *B;
    ball_cache[blk] = std::move(B);
    return ref;
}

void KDTree::SearchBallPos(Double_t *x, Double_t fdist2, Int_t imark, Int_t *nn, Double_t *dist2) {
    std::lock_guard<std::mutex> lk(mu);
    check(nbx_ball_mark(tree, x, -1, fdist2, imark, nn, dist2), "SearchBallPos");
}
void KDTree::SearchBallPos(Int_t tt, Double_t fdist2, Int_t imark, Int_t *nn, Double_t *dist2) {
    if (treetype != TPHYS) die("SearchBallPos needs a TPHYS tree");
    std::lock_guard<std::mutex> lk(mu);
    const BallBlock &B = ball_block(tt, fdist2);
    const Int_t r = tt & ((1 << BALL_BLOCK_LOG2) - 1);
    for (int64_t e = B.off[r]; e < B.off[r + 1]; ++e) {
        const Int_t i = B.idx[e];
        if (!periodic && i == tt) continue;
        const Int_t id = bucket[i].GetID();
        nn[id] = imark;
        dist2[id] = B.d2[e];
    }
}
std::vector<Int_t> KDTree::SearchBallPosTagged(Double_t *x, Double_t fdist2) {
    std::lock_guard<std::mutex> lk(mu);
    int64_t cap = std::min<int64_t>(numparts, 1 << 16), nout = 0;
    std::vector<int32_t> ids;
    for (;;) {
        ids.resize(cap);
        check(nbx_ball_points(tree, x, fdist2, ids.data(), cap, &nout), "SearchBallPosTagged");
        if (nout <= cap) break;
        cap = nout;
    }
    std::vector<Int_t> out(nout);
    for (int64_t i = 0; i < nout; ++i) out[i] = slot_of_id[ids[i]];
    return out;
}
std::vector<Int_t> KDTree::SearchBallPosTagged(Int_t tt, Double_t fdist2) {
    if (treetype != TPHYS) die("SearchBallPosTagged needs a TPHYS tree");
    std::lock_guard<std::mutex> lk(mu);
    const BallBlock &B = ball_block(tt, fdist2);
    const Int_t r = tt & ((1 << BALL_BLOCK_LOG2) - 1);
    std::vector<Int_t> out;
    out.reserve(B.off[r + 1] - B.off[r]);
    for (int64_t e = B.off[r]; e < B.off[r + 1]; ++e)
        if (periodic || B.idx[e] != tt) out.push_back(B.idx[e]);
    return out;
}
Int_t KDTree::SearchBallPosTagged(Double_t *x, Double_t fdist2, Int_t *tagged) {
    std::vector<Int_t> v = SearchBallPosTagged(x, fdist2);
    std::copy(v.begin(), v.end(), tagged);
    return (Int_t)v.size();
}
Int_t KDTree::SearchBallPosTagged(Int_t tt, Double_t fdist2, Int_t *tagged) {
    std::vector<Int_t> v = SearchBallPosTagged(tt, fdist2);
    std::copy(v.begin(), v.end(), tagged);
    return (Int_t)v.size();
}

// ------------------------------------------------------------------ densities
// caches that read Particle::rho (the caller holds the mutex)
void KDTree::drop_density_caches() { pcache_k = pcache_s = pcache_set = -1; }

void KDTree::density_to_bucket(Int_t Nsmooth) {
    std::vector<double> rho(numparts);
    check(nbx_density(tree, Nsmooth, NBX_DENS_SYMMETRIC, 0, rho.data(), nullptr), "CalcDensity failed");
    for (Int_t s = 0; s < numparts; ++s) bucket[s].SetDensity(rho[bucket[s].GetID()]);
    drop_density_caches();
}
std::vector<double> KDTree::rho_by_id() {
    std::vector<double> rho(numparts);
    for (Int_t s = 0; s < numparts; ++s) rho[bucket[s].GetID()] = bucket[s].GetDensity();
    return rho;
}
void KDTree::CalcDensity(Int_t Nsmooth) {
    std::lock_guard<std::mutex> lk(mu);
    density_to_bucket(Nsmooth);
}
Coordinate *KDTree::CalcSmoothVel(Int_t Nsmooth, int densityset) {
    std::lock_guard<std::mutex> lk(mu);
    if (densityset != 1) density_to_bucket(Nsmooth);
    const std::vector<double> rho = rho_by_id();
    std::vector<double> mean((size_t)3 * numparts);
    check(nbx_smooth_vel(tree, Nsmooth, rho.data(), nullptr, mean.data(), nullptr), "CalcSmoothVel failed");
    Coordinate *out = new Coordinate[numparts];
    for (Int_t i = 0; i < numparts; ++i) out[i] = Coordinate(mean[3 * i], mean[3 * i + 1], mean[3 * i + 2]);
    return out;
}
Matrix *KDTree::CalcSmoothVelDisp(Coordinate *smvel, Int_t Nsmooth, int densityset, int meanvelset) {
    const int own = meanvelset != 1;
    if (own) {
        delete[] smvel;
        smvel = CalcSmoothVel(Nsmooth, densityset);
        densityset = 1;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (densityset != 1) density_to_bucket(Nsmooth);
    const std::vector<double> rho = rho_by_id();
    std::vector<double> mean((size_t)3 * numparts), disp((size_t)9 * numparts);
    for (Int_t i = 0; i < numparts; ++i)
        for (int d = 0; d < 3; ++d) mean[3 * i + d] = smvel[i][d];
    check(nbx_smooth_vel(tree, Nsmooth, rho.data(), mean.data(), nullptr, disp.data()), "CalcSmoothVelDisp failed");
    Matrix *out = new Matrix[numparts];
    for (Int_t i = 0; i < numparts; ++i)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) out[i](a, b) = disp[9 * i + 3 * a + b];
    if (own) delete[] smvel;  // the means computed here (the caller's were deleted, as the reference does)
    return out;
}
void KDTree::CalcVelDensity(Int_t Nsmooth, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    std::vector<double> rho(numparts);
    check(nbx_density(tree, Nsmooth, NBX_DENS_VEL, Nsearch, rho.data(), nullptr), "CalcVelDensity failed");
    for (Int_t s = 0; s < numparts; ++s) bucket[s].SetDensity(rho[bucket[s].GetID()]);
    drop_density_caches();
}
Double_t KDTree::CalcDensityParticle(Int_t target, Int_t Nsmooth) {
    std::lock_guard<std::mutex> lk(mu);
    if (dcache_k != Nsmooth) {
        dcache_k = -1;
        dcache.assign(numparts, 0.0);
        check(nbx_density(tree, Nsmooth, NBX_DENS_GATHER, 0, dcache.data(), nullptr), "CalcDensity failed");
        dcache_k = Nsmooth;
    }
    return dcache[bucket[target].GetID()];
}
Double_t KDTree::CalcVelDensityParticle(Int_t target, Int_t Nsmooth, Int_t Nsearch, int, PriorityQueue *,
                                        PriorityQueue *, Int_t *, Double_t *) {
    std::lock_guard<std::mutex> lk(mu);
    if (Nsmooth > Nsearch) Nsmooth = Nsearch;  // as the reference does (:855-858)
    if (vcache_k != Nsmooth || vcache_s != Nsearch) {
        vcache_k = vcache_s = -1;
        vcache.assign(numparts, 0.0);
        check(nbx_density(tree, Nsmooth, NBX_DENS_VEL, Nsearch, vcache.data(), nullptr), "CalcDensity failed");
        vcache_k = Nsmooth;
        vcache_s = Nsearch;
    }
    return vcache[bucket[target].GetID()];
}
Double_t KDTree::CalcDensityPosition(Double_t *x, Int_t Nsmooth, Double_t *) {
    if (treetype == TPHS) die("CalcDensityPosition here needs a TPHYS or TVEL tree");
    std::lock_guard<std::mutex> lk(mu);
    double den = 0.0;
    check(nbx_density_points(tree, x, 1, Nsmooth, &den, nullptr), "CalcDensity failed");
    return den;
}
Double_t *KDTree::CalcVelDensityWithPhysDensity(Int_t Nsmooth, Int_t Nsearch, int densityset) {
    if (Nsmooth > Nsearch) Nsmooth = Nsearch;  // before the densities, as the reference (:405-409)
    std::lock_guard<std::mutex> lk(mu);
    if (densityset != 1) density_to_bucket(Nsmooth);
    const std::vector<double> rho = rho_by_id();
    Double_t *out = new Double_t[numparts];
    const int rc = nbx_veldensity_phys(tree, Nsmooth, Nsearch, rho.data(), out);
    if (rc < 0) delete[] out;
    check(rc, "CalcVelDensitywithPhysDensity failed");
    return out;
}
Double_t KDTree::CalcVelDensityWithPhysDensityParticle(Int_t target, Int_t Nsmooth, Int_t Nsearch, int densityset) {
    if (Nsmooth > Nsearch) Nsmooth = Nsearch;  // :933-936
    const int set = densityset == 1 ? 1 : 0;
    std::lock_guard<std::mutex> lk(mu);
    if (pcache_k != Nsmooth || pcache_s != Nsearch || pcache_set != set) {
        if (!set) density_to_bucket(Nsmooth);  // :937
        pcache.assign(numparts, 0.0);
        const std::vector<double> rho = rho_by_id();
        check(nbx_veldensity_phys(tree, Nsmooth, Nsearch, rho.data(), pcache.data()),
              "CalcVelDensitywithPhysDensity failed");
        pcache_k = Nsmooth;
        pcache_s = Nsearch;
        pcache_set = set;
    }
    return pcache[bucket[target].GetID()];
}
Double_t KDTree::CalcVelDensityPosition(Double_t *x, Double_t *v, Int_t Nsmooth, Int_t Nsearch) {
    std::lock_guard<std::mutex> lk(mu);
    double den = 0.0;
    check(nbx_veldensity_points(tree, x, v, 1, Nsmooth, Nsearch, &den, nullptr), "CalcVelDensityPosition failed");
    return den;
}

// which: 0 skewness, 1 kurtosis.  The reference's order of preparation:
// densities (densityset), then means (meanvelset), then dispersions
// (veldispset), replacing the caller's arrays it recomputes.
Coordinate *KDTree::smooth_moment(int which, Coordinate *smvel, Matrix *smveldisp, Int_t Nsmooth, int densityset,
                                  int meanvelset, int veldispset) {
    if (densityset != 1) CalcDensity(Nsmooth);
    const bool own_mean = meanvelset != 1, own_disp = veldispset != 1;
    if (own_mean) {
        delete[] smvel;
        smvel = CalcSmoothVel(Nsmooth, 1);
    }
    if (own_disp) {
        delete[] smveldisp;
        smveldisp = CalcSmoothVelDisp(smvel, Nsmooth, 1, 1);
    }
    std::lock_guard<std::mutex> lk(mu);
    const std::vector<double> rho = rho_by_id();
    std::vector<double> mean((size_t)3 * numparts), disp((size_t)9 * numparts), out((size_t)3 * numparts);
    for (Int_t i = 0; i < numparts; ++i)
        for (int a = 0; a < 3; ++a) {
            mean[3 * i + a] = smvel[i][a];
            for (int b = 0; b < 3; ++b) disp[9 * i + 3 * a + b] = smveldisp[i](a, b);
        }
    check(nbx_smooth_vel_moments(tree, Nsmooth, rho.data(), mean.data(), disp.data(), which ? nullptr : out.data(),
                                 which ? out.data() : nullptr),
          which ? "CalcSmoothVelKurtosis failed" : "CalcSmoothVelSkew failed");
    if (own_mean) delete[] smvel;
    if (own_disp) delete[] smveldisp;
    Coordinate *res = new Coordinate[numparts];
    for (Int_t i = 0; i < numparts; ++i) res[i] = Coordinate(out[3 * i], out[3 * i + 1], out[3 * i + 2]);
    return res;
}
Coordinate *KDTree::CalcSmoothVelSkew(Coordinate *smvel, Matrix *smveldisp, Int_t Nsmooth, int densityset,
                                      int meanvelset, int veldispset) {
    return smooth_moment(0, smvel, smveldisp, Nsmooth, densityset, meanvelset, veldispset);
}
Coordinate *KDTree::CalcSmoothVelKurtosis(Coordinate *smvel, Matrix *smveldisp, Int_t Nsmooth, int densityset,
                                          int meanvelset, int veldispset) {
    return smooth_moment(1, smvel, smveldisp, Nsmooth, densityset, meanvelset, veldispset);
}

// One local mean or dispersion: at particle `target` (x null) or position x.
// The weights are by ID, as the reference indexes them (weight[GetID()]).
Double_t KDTree::smooth_local_one(Int_t target, const Double_t *x, Double_t *weight, const Double_t *mean,
                                  Int_t Nsmooth, int densityset) {
    if (densityset != 1) CalcDensity(Nsmooth);
    std::lock_guard<std::mutex> lk(mu);
    const std::vector<double> rho = rho_by_id();
    double out = 0.0;
    if (x) {
        const double q[3] = {x[0], x[1], x[2]};
        check(nbx_smooth_local_points(tree, q, 1, Nsmooth, weight, rho.data(), mean, mean ? nullptr : &out,
                                      mean ? &out : nullptr),
              "CalcSmoothLocalMean failed");
        return out;
    }
    std::vector<double> res(numparts), mi;
    const Int_t id = bucket[target].GetID();
    if (mean) mi.assign(numparts, *mean);
    check(nbx_smooth_local(tree, Nsmooth, weight, rho.data(), mean ? mi.data() : nullptr, mean ? nullptr : res.data(),
                           mean ? res.data() : nullptr),
          "CalcSmoothLocalMean failed");
    return res[id];
}
Coordinate KDTree::CalcSmoothVelParticle(Int_t target, Int_t Nsmooth, int densityset) {
    std::vector<Double_t> w(numparts);
    Coordinate out;
    for (int d = 0; d < 3; ++d) {
        for (Int_t s = 0; s < numparts; ++s) w[bucket[s].GetID()] = bucket[s].GetVelocity(d);
        out[d] = smooth_local_one(target, nullptr, w.data(), nullptr, Nsmooth, d == 0 ? densityset : 1);
    }
    return out;
}
Coordinate KDTree::CalcSmoothVelPosition(Double_t *x, Int_t Nsmooth, int densityset) {
    std::vector<Double_t> w(numparts);
    Coordinate out;
    for (int d = 0; d < 3; ++d) {
        for (Int_t s = 0; s < numparts; ++s) w[bucket[s].GetID()] = bucket[s].GetVelocity(d);
        out[d] = smooth_local_one(0, x, w.data(), nullptr, Nsmooth, d == 0 ? densityset : 1);
    }
    return out;
}
Double_t KDTree::CalcSmoothLocalMeanParticle(Int_t target, Double_t *weight, Int_t Nsmooth, int densityset) {
    return smooth_local_one(target, nullptr, weight, nullptr, Nsmooth, densityset);
}
Double_t KDTree::CalcSmoothLocalDispParticle(Int_t target, Double_t *weight, Double_t localmean, Int_t Nsmooth,
                                             int densityset) {
    return smooth_local_one(target, nullptr, weight, &localmean, Nsmooth, densityset);
}
Double_t KDTree::CalcSmoothLocalMeanPosition(Double_t *x, Double_t *weight, Int_t Nsmooth, int densityset) {
    return smooth_local_one(0, x, weight, nullptr, Nsmooth, densityset);
}
Double_t KDTree::CalcSmoothLocalDispPosition(Double_t *x, Double_t *weight, Double_t localmean, Int_t Nsmooth,
                                             int densityset) {
    return smooth_local_one(0, x, weight, &localmean, Nsmooth, densityset);
}

// ------------------------------------------------------------------ FOF
// One FOF run (cmp 0: FOF with linking length fdist; else FOFCriterion with
// that built-in comparator).  With ipcheckflag, the particles whose check is
// nonzero take no part: the links are those of a tree over the others, built
// with this tree's options and the same periodic rule (a reference tree of
// N <= bucket_size particles searches without images, KDLeafNode.cxx:841-844).
Int_t *KDTree::fof_run(int cmp, Double_t fdist, Double_t *params, Int_t &numgroup, Int_t minnum, Int_tree_t *pHead,
                       Int_tree_t *pNext, Int_tree_t *pTail, Int_tree_t *pLen, int ipcheckflag, FOFcheckfunc chk,
                       const char *what) {
    std::lock_guard<std::mutex> lk(mu);
    std::vector<int> val;  // check values by ID (empty: every particle takes part)
    if (ipcheckflag && chk != &Pnocheck) {
        val.assign(numparts, 0);
        bool any = false;
        for (Int_t s = 0; s < numparts; ++s) {
            const int c = chk(bucket[s], params);
            val[bucket[s].GetID()] = c;
            any = any || c != 0;
            if (cmp && c > 0)
                die(std::string(what) + ": check functions returning a positive value are not supported "
                                        "(the reference lets such particles join groups but not start one)");
        }
        if (!any) val.clear();
    }
    // With lists requested every component is found (minnum 1): the
    // reference leaves the members of groups pruned below minnum linked
    // (KDFOF.cxx:98-104).  Canonical labels run by size descending, so the
    // groups of >= minnum members are the first ones and keep their labels.
    const bool lists = pHead || pNext || pTail || pLen;
    const Int_t mrun = lists ? 1 : minnum;
    Int_t *pGroup = new Int_t[numparts];
    int64_t ng = 0;
    int rc = NBX_OK;
    if (val.empty()) {
        rc = cmp ? nbx_fof_criterion(tree, cmp, params, mrun, pGroup, &ng) : nbx_fof(tree, fdist, mrun, pGroup, &ng);
    } else {
        std::vector<Int_t> act;  // IDs taking part, ascending (so canonical ties follow the IDs)
        for (Int_t i = 0; i < numparts; ++i)
            if (val[i] == 0) act.push_back(i);
        for (Int_t i = 0; i < numparts; ++i) pGroup[i] = 0;
        const int64_t m = (int64_t)act.size();
        if (m > 0) {
            std::vector<double> pos(3 * (size_t)m), vel(3 * (size_t)m), mass(m);
            for (int64_t j = 0; j < m; ++j) {
                const Particle &p = bucket[slot_of_id[act[j]]];
                for (int d = 0; d < 3; ++d) {
                    pos[d * (size_t)m + j] = p.GetPosition(d);
                    vel[d * (size_t)m + j] = p.GetVelocity(d);
                }
                mass[j] = p.GetMass();
            }
            nbx_params sp = prm;
            sp.form = NBX_FORM_PLAIN;
            sp.periodic = periodic && numparts > b ? 1 : 0;
            sp.bucket_size = 1;
            nbx_tree *sub = nullptr;
            rc = nbx_tree_create(&sp, pos.data(), vel.data(), mass.data(), m, &sub);
            std::vector<int32_t> lab(m, 0);
            if (rc == NBX_OK)
                rc = cmp ? nbx_fof_criterion(sub, cmp, params, mrun, lab.data(), &ng)
                         : nbx_fof(sub, fdist, mrun, lab.data(), &ng);
            if (sub) nbx_tree_destroy(sub);
            if (rc == NBX_OK)
                for (int64_t j = 0; j < m; ++j) pGroup[act[j]] = lab[j];
        }
    }
    if (rc < 0) delete[] pGroup;
    check(rc, what);
    // pHead/pNext/pTail/pLen (KDFOF.cxx:29-31), by tree index, from every
    // component; then the groups below minnum are dropped from the labels
    // and pLen (the reference zeroes a pruned group's pLen, KDFOF.cxx:103)
    if (lists) {
        std::vector<int32_t> order(numparts);
        check(nbx_tree_order(tree, order.data()), what);
        check(nbx_fof_lists(pGroup, order.data(), numparts, ng, pHead, pNext, pTail, pLen), what);
        std::vector<int64_t> size((size_t)ng + 1, 0);
        for (Int_t i = 0; i < numparts; ++i) size[pGroup[i]]++;
        int64_t kept = 0;
        while (kept < ng && size[kept + 1] >= minnum) ++kept;
        for (Int_t i = 0; i < numparts; ++i)
            if (pGroup[i] > kept) pGroup[i] = 0;
        if (pLen)
            for (int64_t g = kept + 1; g <= ng && g < numparts; ++g) pLen[g] = 0;
        ng = kept;
    }
    numgroup = (Int_t)ng;
    // particles left out keep their check value, -1 becoming 0 (KDFOF.cxx:63-64, :107)
    if (!val.empty())
        for (Int_t i = 0; i < numparts; ++i)
            if (val[i] != 0) pGroup[i] = val[i] == -1 ? 0 : val[i];
    return pGroup;
}

Int_t *KDTree::FOF(Double_t fdist, Int_t &numgroup, Int_t minnum, int, Int_tree_t *pHead, Int_tree_t *pNext,
                   Int_tree_t *pTail, Int_tree_t *pLen, int ipcheckflag, FOFcheckfunc chk, Double_t *params) {
    return fof_run(0, fdist, params, numgroup, minnum, pHead, pNext, pTail, pLen, ipcheckflag, chk, "FOF");
}

Int_t *KDTree::FOFCriterion(FOFcompfunc cmp, Double_t *params, Int_t &numgroups, Int_t minnum, int, int ipcheckflag,
                            FOFcheckfunc chk, Int_tree_t *pHead, Int_tree_t *pNext, Int_tree_t *pTail,
                            Int_tree_t *pLen) {
    const int32_t c = cmp == &FOF3d ? NBX_FOF_CMP_3D : cmp == &FOF6d ? NBX_FOF_CMP_6D : 0;
    if (!c) die("FOFCriterion: only the built-in comparators FOF3d and FOF6d are provided");
    return fof_run(c, 0.0, params, numgroups, minnum, pHead, pNext, pTail, pLen, ipcheckflag, chk, "FOFCriterion");
}

}  // namespace NBody
