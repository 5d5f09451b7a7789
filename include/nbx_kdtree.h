/*
 * nbx_kdtree.h -- drop-in NBody::KDTree / NBody::Particle for the KDTree hot
 * path, backed by the MI355X C ABI (nbx.h, libnbx.so).
 *
 * Mirrors the reference's call surface for the hot path:
 *   Particle      src/NBody/Particle.h:264-293 (88-byte layout, default build)
 *   KDTree ctor   src/KDTree/KDTree.h:229-274
 *   FindNearest*  KDTree.h:300-331      SearchBallPosTagged  KDTree.h:363-365
 *   SearchBallPos KDTree.h:342, :348
 *   CalcDensity*  KDTree.h:411-416      FOF                  KDTree.h:480-482
 *   FOFCriterion  KDTree.h:503 with the comparators of FOFFunc.h:24-55
 *   CalcSmoothVel / CalcSmoothVelDisp  KDTree.h:414-418
 *   OverWriteInputOrder / SetResetOrder KDTree.h:515-518
 *
 * Semantics kept from the reference: the constructor overwrites Particle::id
 * with the input index and reorders the caller's array into tree order (here
 * kd split order); "tree indices" returned by the kNN calls index that array;
 * Group/FOF arrays are keyed by ID; ~KDTree sorts the array back by ID.
 * Errors print a message and exit(1) as the reference does, unless
 * KDTree::SetThrowOnError(true) selects std::runtime_error instead.
 *
 * Single-particle queries (FindNearest(tt,...), CalcDensityParticle) are
 * answered from one batched GPU call per (variant, Nsearch) whose result is
 * cached, so OpenMP loops over particles (the reference's usage,
 * test_kdtree.cxx:278-301) run at batch speed; the cache is dropped when the
 * tree is destroyed.
 */
#ifndef NBX_KDTREE_H
#define NBX_KDTREE_H

#include <cstdio>
#include <mutex>
#include <vector>

#include "nbx.h"

namespace NBody {

typedef double Double_t;  // Precision.h:38-52 default build
typedef float Real_t;
typedef int Int_t;
typedef int Int_tree_t;

class Particle {
  protected:
    Double_t mass;
    Double_t position[3];
    Double_t velocity[3];
    Int_t pid;
    Int_t id;
    int type;
    Double_t rho;
    Double_t phi;

  public:
    Particle(Double_t Mass = 0, Double_t x = 0, Double_t y = 0, Double_t z = 0, Double_t vx = 0, Double_t vy = 0,
             Double_t vz = 0, Int_t ID = 0, int Type = 0, Double_t Rho = 0, Double_t Phi = 0, Int_t PID = 0)
        : mass(Mass), position{x, y, z}, velocity{vx, vy, vz}, pid(PID), id(ID), type(Type), rho(Rho), phi(Phi) {}
    Particle(Double_t Mass, const Double_t *NewPos, const Double_t *NewVel, Int_t ID = 0, int Type = 0,
             Double_t Rho = 0, Double_t Phi = 0, Int_t PID = 0)
        : mass(Mass), position{NewPos[0], NewPos[1], NewPos[2]}, velocity{NewVel[0], NewVel[1], NewVel[2]},
          pid(PID), id(ID), type(Type), rho(Rho), phi(Phi) {}

    Double_t GetMass() const { return mass; }
    void SetMass(const Double_t &m) { mass = m; }
    const Double_t *GetPosition() const { return position; }
    Double_t GetPosition(const int &i) const { return position[i]; }
    void SetPosition(const int &i, const Double_t &x) { position[i] = x; }
    void SetPosition(const Double_t *p) { position[0] = p[0]; position[1] = p[1]; position[2] = p[2]; }
    const Double_t *GetVelocity() const { return velocity; }
    Double_t GetVelocity(const int &i) const { return velocity[i]; }
    void SetVelocity(const int &i, const Double_t &v) { velocity[i] = v; }
    Double_t GetPhase(const int &i) const { return i < 3 ? position[i] : velocity[i - 3]; }
    void SetPhase(const int &i, const Double_t &x) { if (i < 3) position[i] = x; else velocity[i - 3] = x; }
    Int_t GetPID() const { return pid; }
    void SetPID(const Int_t &i) { pid = i; }
    Int_t GetID() const { return id; }
    void SetID(const Int_t &i) { id = i; }
    int GetType() const { return type; }
    void SetType(int t) { type = t; }
    Double_t GetDensity() const { return rho; }
    void SetDensity(const Double_t &r) { rho = r; }
    Double_t GetPotential() const { return phi; }
    void SetPotential(const Double_t &p) { phi = p; }
};

// The reference driver's raw particle file (src/tests/test_kdtree.cxx:60-73,
// :181-208): six doubles per particle, GetPhase(0..5), no header.  Reading
// sets mass 1 and ID / PID = input index, as the driver does; n = 0 reads
// every record.  Return false on a short or unreadable file.
inline bool ReadPartsOut(const char *fname, std::vector<Particle> &parts, size_t n = 0) {
    std::FILE *f = std::fopen(fname, "rb");
    if (!f) return false;
    if (n == 0) {
        std::fseek(f, 0, SEEK_END);
        n = (size_t)std::ftell(f) / (6 * sizeof(Double_t));
        std::fseek(f, 0, SEEK_SET);
    }
    parts.assign(n, Particle());
    bool ok = true;
    for (size_t i = 0; i < n && ok; ++i) {
        Double_t x[6];
        ok = std::fread(x, sizeof(Double_t), 6, f) == 6;
        for (int j = 0; j < 6 && ok; ++j) parts[i].SetPhase(j, x[j]);
        parts[i].SetMass(1.0);
        parts[i].SetID((Int_t)i);
        parts[i].SetPID((Int_t)i);
    }
    std::fclose(f);
    return ok;
}
inline bool WritePartsOut(const char *fname, const std::vector<Particle> &parts) {
    std::FILE *f = std::fopen(fname, "wb");
    if (!f) return false;
    bool ok = true;
    for (const Particle &p : parts) {
        const Double_t x[6] = {p.GetPhase(0), p.GetPhase(1), p.GetPhase(2), p.GetPhase(3), p.GetPhase(4), p.GetPhase(5)};
        ok = ok && std::fwrite(x, sizeof(Double_t), 6, f) == 6;
    }
    return std::fclose(f) == 0 && ok;
}

// The results of CalcSmoothVel and CalcSmoothVelDisp: the forms of
// NBody::Coordinate (src/Math/Coordinate.h:46-78) and the 3x3 NBody::Matrix
// (src/Math/Matrix.h:23-41) that callers index, c[i] and m(i, j).
class Coordinate {
    Double_t c[3];

  public:
    Coordinate(const Double_t x = 0, const Double_t y = 0, const Double_t z = 0) { c[0] = x; c[1] = y; c[2] = z; }
    Double_t &operator[](int i) { return c[i]; }
    const Double_t &operator[](int i) const { return c[i]; }
    Double_t *GetCoord() { return c; }
};
class Matrix {
    Double_t m[9];

  public:
    Matrix(Double_t a = 0) {
        for (int i = 0; i < 9; ++i) m[i] = a;
    }
    Double_t &operator()(int i, int j) { return m[3 * i + j]; }
    const Double_t &operator()(int i, int j) const { return m[3 * i + j]; }
};

class PriorityQueue;  // src/KDTree/PriorityQueue.h; only passed through as workspace here

// FOFFunc.h:24-55: comparator and check-function types and the built-in
// functions.  A comparator links a and b when it returns nonzero; params[6]
// and params[7] are the position and velocity scales (squared linking
// lengths).  KDTree::FOFCriterion runs FOF3d and FOF6d on the GPU; it
// recognises them by address, so pass these functions themselves.
typedef int (*FOFcompfunc)(Particle &, Particle &, Double_t *);
typedef int (*FOFcheckfunc)(Particle &, Double_t *);
inline int FOF3d(Particle &a, Particle &b, Double_t *params) {
    Double_t total = 0;
    for (int j = 0; j < 3; j++) {
        const Double_t t = a.GetPosition(j) - b.GetPosition(j);
        total += t * t / params[6];
    }
    return total < 1;
}
inline int FOFVel(Particle &a, Particle &b, Double_t *params) {
    Double_t total = 0;
    for (int j = 0; j < 3; j++) {
        const Double_t t = a.GetVelocity(j) - b.GetVelocity(j);
        total += t * t / params[6];
    }
    return total < 1;
}
inline int FOF6d(Particle &a, Particle &b, Double_t *params) {
    Double_t total = 0;
    for (int j = 0; j < 3; j++) {
        const Double_t t = a.GetPosition(j) - b.GetPosition(j), u = a.GetVelocity(j) - b.GetVelocity(j);
        total += t * t / params[6];
        total += u * u / params[7];
    }
    return total < 1;
}
inline int Pnocheck(Particle &, Double_t *) { return 0; }

class KDTree {
  public:
    const static int TPHYS = 0, TPROJ = 1, TVEL = 2, TPHS = 3, TMETRIC = 4;
    const static int KSPH = 0, KGAUSS = 1, KEPAN = 2, KTH = 3;
    const static int KDTREE_SPLIT_ENTROPY = 1, KDTREE_SPLIT_DISPERSION = 2, KDTREE_SPLIT_MAXINTERPARTICLESPACING = 3,
                     KDTREE_SPLIT_SPREAD = 0;

    KDTree(Particle *p, Int_t numparts, Int_t bucket_size = 16, int TreeType = TPHYS, int KernType = KEPAN,
           int KernRes = 1000, int SplittingCriterion = KDTREE_SPLIT_SPREAD, int Aniso = 0, int ScaleSpace = 0,
           Double_t *Period = nullptr, Double_t **metric = nullptr, bool iBuildInParallel = true,
           bool iKeepInputOrder = false, Double_t Rdistadapt = -1, Double_t AdaptiveMedianFac = 0.0,
           Int_t min_bucket_size = 16);
    KDTree(std::vector<Particle> &p, Int_t bucket_size = 16, int TreeType = TPHYS, int KernType = KEPAN,
           int KernRes = 1000, int SplittingCriterion = KDTREE_SPLIT_SPREAD, int Aniso = 0, int ScaleSpace = 0,
           Double_t *Period = nullptr, Double_t **metric = nullptr, bool iBuildInParallel = true,
           bool iKeepInputOrder = false, Double_t Rdistadapt = -1, Double_t AdaptiveMedianFac = 0.0,
           Int_t min_bucket_size = 16);
    ~KDTree();
    KDTree(const KDTree &) = delete;
    KDTree &operator=(const KDTree &) = delete;

    Int_t GetNumNodes() { return numnodes; }
    Int_t GetNumLeafNodes() { return numleafnodes; }
    Int_t GetBucketSize() { return b; }
    Int_t GetTreeType() { return treetype; }
    Int_t GetKernType() { return kernfunctype; }
    Double_t GetKernNorm();
    Double_t GetPeriod(int j) { return period[j]; }

    // kNN for every particle: nn[i][*] are tree indices, ascending in dist2
    void FindNearest(Int_t **nn, Double_t **dist2, Int_t Nsearch = 64);
    void FindNearestPos(Int_t **nn, Double_t **dist2, Int_t Nsearch = 64);
    void FindNearestPhase(Int_t **nn, Double_t **dist2, Int_t Nsearch = 64);
    // one particle (tree index tt)
    void FindNearest(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestPos(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestPhase(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    // one position
    void FindNearestPos(Double_t *x, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    // velocity space, TVEL trees only (KDTree.h FindNearestVel; KDFindNearest.cxx:335-346,
    // :451-453, :530-541, :559-561): every particle, one particle, one velocity
    void FindNearestVel(Int_t **nn, Double_t **dist2, Int_t Nsearch = 64);
    void FindNearestVel(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestVel(Double_t *v, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestVel(Coordinate v, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64) {
        FindNearestVel(v.GetCoord(), nn, dist2, Nsearch);
    }

    // tree indices of all particles with d2 < fdist2 around a position / particle
    // nn[id] = imark, dist2[id] = d2 for every particle with d2 < fdist2
    // (KDTree.h:342, :348; KDFindNearest.cxx:567-583; the target itself is not
    // marked)
    void SearchBallPos(Int_t tt, Double_t fdist2, Int_t imark, Int_t *nn, Double_t *dist2);
    void SearchBallPos(Double_t *x, Double_t fdist2, Int_t imark, Int_t *nn, Double_t *dist2);
    Int_t SearchBallPosTagged(Double_t *x, Double_t fdist2, Int_t *tagged);
    Int_t SearchBallPosTagged(Int_t tt, Double_t fdist2, Int_t *tagged);
    std::vector<Int_t> SearchBallPosTagged(Double_t *x, Double_t fdist2);

    void CalcDensity(Int_t Nsmooth = 64);
    // KDTree.h:414-418, KDCalcSmoothQuantities.cxx:480-614 (TPHYS trees with
    // velocities): new[] arrays of numparts by ID.  densityset = 1 divides by
    // Particle::rho, else CalcDensity(Nsmooth) runs first; meanvelset = 1
    // takes smvel as the means, else CalcSmoothVel runs first.
    Coordinate *CalcSmoothVel(Int_t Nsmooth = 64, int densityset = 1);
    Matrix *CalcSmoothVelDisp(Coordinate *smvel, Int_t Nsmooth = 64, int densityset = 1, int meanvelset = 1);
    // KDTree.h:420-422, KDCalcSmoothQuantities.cxx:617-767; veldispset = 1
    // takes smveldisp as the dispersions, else CalcSmoothVelDisp runs first.
    Coordinate *CalcSmoothVelSkew(Coordinate *smvel, Matrix *smveldisp, Int_t Nsmooth = 64, int densityset = 1,
                                  int meanvelset = 1, int veldispset = 1);
    Coordinate *CalcSmoothVelKurtosis(Coordinate *smvel, Matrix *smveldisp, Int_t Nsmooth = 64, int densityset = 1,
                                      int meanvelset = 1, int veldispset = 1);
    // KDTree.h:427, :435, KDCalcSmoothQuantities.cxx:992-1035, :1209-1253: the
    // gather form of the smoothed velocity at one particle or position (full
    // weight, no scatter), i.e. the local mean of each velocity component.
    Coordinate CalcSmoothVelParticle(Int_t target, Int_t Nsmooth = 64, int densityset = 1);
    Coordinate CalcSmoothVelPosition(Double_t *x, Int_t Nsmooth = 64, int densityset = 1);
    // KDTree.h:454-459, KDCalcSmoothQuantities.cxx:1539-1693: weight by ID.
    // Each call runs one GPU pass (every particle, or the one position); for
    // many targets call nbx_smooth_local / nbx_smooth_local_points once.
    Double_t CalcSmoothLocalMeanParticle(Int_t target, Double_t *weight, Int_t Nsmooth = 64, int densityset = 1);
    Double_t CalcSmoothLocalDispParticle(Int_t target, Double_t *weight, Double_t localmean, Int_t Nsmooth = 64,
                                         int densityset = 1);
    Double_t CalcSmoothLocalMeanPosition(Double_t *x, Double_t *weight, Int_t Nsmooth = 64, int densityset = 1);
    Double_t CalcSmoothLocalDispPosition(Double_t *x, Double_t *weight, Double_t localmean, Int_t Nsmooth = 64,
                                         int densityset = 1);
    void CalcVelDensity(Int_t Nsmooth = 64, Int_t Nsearch = 64);
    Double_t CalcDensityParticle(Int_t target, Int_t Nsmooth = 64);
    // KDTree.h:425, KDCalcSmoothQuantities.cxx:845-921 (TPHYS trees with
    // velocities): served like CalcDensityParticle from one batched
    // CalcVelDensity per (Nsmooth, Nsearch); the workspace arguments (iflag,
    // pq, pq2, nnIDs, vdist) are accepted and not needed.
    Double_t CalcVelDensityParticle(Int_t target, Int_t Nsmooth = 64, Int_t Nsearch = 64, int iflag = 0,
                                    PriorityQueue *pq = nullptr, PriorityQueue *pq2 = nullptr, Int_t *nnIDs = nullptr,
                                    Double_t *vdist = nullptr);
    // KDTree.h:432 (TPHYS trees; v is unused there); KDCalcSmoothQuantities.cxx:1092-1149
    Double_t CalcDensityPosition(Double_t *x, Int_t Nsmooth = 64, Double_t *v = NULL);
    // KDCalcSmoothQuantities.cxx:393-476 (new[] array by ID; the caller delete[]s it) and,
    // per particle, :922-990, served from one batched call per (Nsmooth, Nsearch)
    Double_t *CalcVelDensityWithPhysDensity(Int_t Nsmooth = 64, Int_t Nsearch = 64, int densityset = 0);
    Double_t CalcVelDensityWithPhysDensityParticle(Int_t target, Int_t Nsmooth = 64, Int_t Nsearch = 64,
                                                   int densityset = 0);
    // KDTree.h:433, :438 (TPHYS trees with velocities); KDCalcSmoothQuantities.cxx:1150-1208
    Double_t CalcVelDensityPosition(Double_t *x, Double_t *v, Int_t Nsmooth = 64, Int_t Nsearch = 64);
    Double_t CalcVelDensityPosition(Coordinate x, Coordinate v, Int_t Nsmooth = 64, Int_t Nsearch = 64) {
        return CalcVelDensityPosition(x.GetCoord(), v.GetCoord(), Nsmooth, Nsearch);
    }

    Int_t *FOF(Double_t fdist, Int_t &numgroup, Int_t minnum = 8, int order = 0, Int_tree_t *pHead = nullptr,
               Int_tree_t *pNext = nullptr, Int_tree_t *pTail = nullptr, Int_tree_t *pLen = nullptr);
    // KDTree.h:503, KDFOF.cxx:157-265: cmp = FOF3d (TPHYS tree) or FOF6d (TPHS
    // tree) with params as the reference's; any other comparator or a check
    // function other than Pnocheck is reported as unsupported.  Labels are
    // canonical, as FOF's; the linked lists, when given, are nbx_fof_lists'
    // (members in tree-index order).
    Int_t *FOFCriterion(FOFcompfunc cmp, Double_t *params, Int_t &numgroups, Int_t minnum = 8, int order = 0,
                        int ipcheckflag = 0, FOFcheckfunc check = Pnocheck, Int_tree_t *pHead = nullptr,
                        Int_tree_t *pNext = nullptr, Int_tree_t *pTail = nullptr, Int_tree_t *pLen = nullptr);

    void OverWriteInputOrder();
    void SetResetOrder(bool);

    static void SetThrowOnError(bool on);
    // Precision form for the batched calls (NBX_FORM_MIXED): exact neighbour
    // sets, densities within 2e-6 of FP64 (see nbx.h).  Off by default.
    static void SetPrecisionForm(bool mixed);

  private:
    void density_to_bucket(Int_t Nsmooth);
    Double_t smooth_local_one(Int_t target, const Double_t *x, Double_t *weight, const Double_t *mean, Int_t Nsmooth,
                              int densityset);
    Coordinate *smooth_moment(int which, Coordinate *smvel, Matrix *smveldisp, Int_t Nsmooth, int densityset,
                              int meanvelset, int veldispset);
    std::vector<double> rho_by_id();
    Particle *bucket;
    Int_t numparts;
    Int_t b;
    int treetype, kernfunctype;
    Int_t numnodes, numleafnodes;
    Double_t period[3];
    bool periodic;
    bool iresetorder;
    nbx_tree *tree;
    std::vector<Int_t> slot_of_id;  // ID -> tree index
    std::mutex mu;
    // cached batched single-particle answers
    int cache_variant, cache_k;
    std::vector<Int_t> cache_nn;
    std::vector<Double_t> cache_d2;
    int dcache_k;
    std::vector<Double_t> dcache;
    int vcache_k, vcache_s;
    std::vector<Double_t> vcache;

    void init(Particle *p, Int_t n, Int_t bucket_size, int TreeType, int KernType, int KernRes, int Splitting,
              int Aniso, int ScaleSpace, Double_t *Period);
    void batch_knn(int variant, Int_t k, Int_t *nn, Double_t *d2);
    void ensure_cache(int variant, Int_t k);
};

}  // namespace NBody

#endif
