/*
 * nbx_kdtree.h -- drop-in NBody::KDTree / NBody::Particle for the KDTree hot
 * path, backed by the MI355X C ABI (nbx.h, libnbx.so).
 *
 * Mirrors the reference's call surface for the hot path:
 *   Particle      src/NBody/Particle.h:264-293 (88-byte layout, default build)
 *   KDTree ctor   src/KDTree/KDTree.h:229-264 (the KDTree(System&) form,
 *                 KDTree.h:266-274, is not provided: System is outside the
 *                 hot path)
 *   FindNearest*  KDTree.h:300-319      SearchBallPos        KDTree.h:342-358
 *   SearchBallPosTagged  KDTree.h:363-373
 *   CalcDensity*  KDTree.h:407-441      FOF                  KDTree.h:480-482
 *   FOFCriterion  KDTree.h:503 with the comparators of FOFFunc.h:24-55
 *   OverWriteInputOrder / SetResetOrder KDTree.h:522-524
 *
 * Semantics kept from the reference: the constructor overwrites Particle::id
 * with the input index and reorders the caller's array into tree order (here
 * kd split order); "tree indices" returned by the kNN and tagged ball calls
 * index that array; Group/FOF arrays are keyed by ID; ~KDTree sorts the array
 * back by ID.  Errors print a message and exit(1) as the reference does,
 * unless KDTree::SetThrowOnError(true) selects std::runtime_error instead.
 *
 * Single-particle queries (FindNearest(tt,...), CalcDensityParticle,
 * CalcVelDensityParticle, CalcVelDensityWithPhysDensityParticle,
 * SearchBallPosTagged(tt,...), SearchBallPos(tt,...)) are answered from
 * batched GPU calls whose results are cached, so OpenMP loops over particles
 * (the reference's usage, test_kdtree.cxx:278-301, :324-341) run at batch
 * speed.  The kNN and density caches hold every particle; the ball caches
 * hold blocks of 65536 consecutive tree indices per radius.  Every call is
 * thread-safe (one mutex per tree).  The caches are dropped when a call
 * rewrites Particle::rho and when the tree is destroyed; a per-particle call
 * with densityset = 1 reads the densities as they were at the first call of
 * a run with the same arguments.
 */
#ifndef NBX_KDTREE_H
#define NBX_KDTREE_H

#include <cstddef>
#include <cstdio>
#include <map>
#include <memory>
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
    friend struct ParticleLayout;

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

// The reference's default-build layout (Particle.h:264-293: mass, position[3],
// velocity[3], pid, id, type, rho, phi with Double_t = double, Int_t = int):
// arrays of it are passed between the caller and the tree as they are.
struct ParticleLayout {
    static_assert(sizeof(Particle) == 88, "Particle must be the reference's 88-byte record");
    static_assert(offsetof(Particle, mass) == 0 && offsetof(Particle, position) == 8 &&
                      offsetof(Particle, velocity) == 32 && offsetof(Particle, pid) == 56 &&
                      offsetof(Particle, id) == 60 && offsetof(Particle, type) == 64 && offsetof(Particle, rho) == 72 &&
                      offsetof(Particle, phi) == 80,
                  "Particle field offsets must match the reference's");
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

// NBody::Coordinate (src/Math/Coordinate.h:46-103) as the hot path's calls
// take it: three doubles, built from numbers or from an array of at least
// three; and the 3x3 NBody::Matrix (src/Math/Matrix.h:23-41) that
// CalcSmoothVelDisp returns, indexed m(i, j).
class Coordinate {
    Double_t c[3];

  public:
    Coordinate(const Double_t x = 0, const Double_t y = 0, const Double_t z = 0) { c[0] = x; c[1] = y; c[2] = z; }
    Coordinate(const Double_t *a) { c[0] = a[0]; c[1] = a[1]; c[2] = a[2]; }
    Coordinate(const Real_t *a) { c[0] = a[0]; c[1] = a[1]; c[2] = a[2]; }
    Double_t X() const { return c[0]; }
    Double_t Y() const { return c[1]; }
    Double_t Z() const { return c[2]; }
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
// lengths).  A check function returns -1 for a particle the search should
// leave out (FOFFunc.h:25).  KDTree::FOFCriterion runs FOF3d and FOF6d on the
// GPU; it recognises them by address, so pass these functions themselves.
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

    // Rdistadapt and AdaptiveMedianFac shape the reference's tree only; the
    // results they give do not depend on it, so they are accepted and unused,
    // as are metric, iBuildInParallel, iKeepInputOrder and min_bucket_size.
    // A TPHS tree needs Aniso = -1 (metric factor 1 on coordinates the caller
    // has scaled); the reference's Aniso = 0 and 1 use a local metric tensor,
    // which is not provided, and are reported as unsupported.
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
    // one point (KDFindNearest.cxx:462-542): FindNearest(x) searches the
    // tree's own space -- positions (TPHYS), velocities (TVEL), or the phase
    // point x[0..5] (TPHS); FindNearestPos(x) positions and FindNearestPhase(x, v)
    // phase space, each period-aware as the reference's
    void FindNearest(Double_t *x, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestPos(Double_t *x, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestPhase(Double_t *x, Double_t *v, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestPos(Coordinate x, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64) {
        FindNearestPos(x.GetCoord(), nn, dist2, Nsearch);
    }
    void FindNearestPhase(Coordinate x, Coordinate v, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64) {
        FindNearestPhase(x.GetCoord(), v.GetCoord(), nn, dist2, Nsearch);
    }
    // velocity space, TVEL trees only (KDFindNearest.cxx:335-346, :451-453,
    // :530-541, :559-561): every particle, one particle, one velocity
    void FindNearestVel(Int_t **nn, Double_t **dist2, Int_t Nsearch = 64);
    void FindNearestVel(Int_t tt, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestVel(Double_t *v, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64);
    void FindNearestVel(Coordinate v, Int_t *nn, Double_t *dist2, Int_t Nsearch = 64) {
        FindNearestVel(v.GetCoord(), nn, dist2, Nsearch);
    }

    // nn[id] = imark, dist2[id] = d2 for every particle with d2 < fdist2
    // (KDTree.h:342-354; KDFindNearest.cxx:567-590).  The target form leaves
    // the target itself unmarked on a non-periodic tree; on a periodic tree it
    // is the search around the target's position (KDSplitNode.cxx:1341-1346),
    // which marks the target too.
    void SearchBallPos(Int_t tt, Double_t fdist2, Int_t imark, Int_t *nn, Double_t *dist2);
    void SearchBallPos(Double_t *x, Double_t fdist2, Int_t imark, Int_t *nn, Double_t *dist2);
    void SearchBallPos(Coordinate x, Double_t fdist2, Int_t imark, Int_t *nn, Double_t *dist2) {
        SearchBallPos(x.GetCoord(), fdist2, imark, nn, dist2);
    }
    // tree indices of the particles with d2 < fdist2 (KDTree.h:363-373,
    // KDFindNearest.cxx:618-706).  The target form leaves out the target
    // itself on a non-periodic tree (KDLeafNode.cxx:314-333,
    // KDSplitNode.cxx:542-548) and is the search around its position on a
    // periodic one (KDSplitNode.cxx:1380-1385), which includes it.
    Int_t SearchBallPosTagged(Int_t tt, Double_t fdist2, Int_t *tagged);
    Int_t SearchBallPosTagged(Double_t *x, Double_t fdist2, Int_t *tagged);
    Int_t SearchBallPosTagged(Coordinate x, Double_t fdist2, Int_t *tagged) {
        return SearchBallPosTagged(x.GetCoord(), fdist2, tagged);
    }
    std::vector<Int_t> SearchBallPosTagged(Int_t tt, Double_t fdist2);
    std::vector<Int_t> SearchBallPosTagged(Double_t *x, Double_t fdist2);
    std::vector<Int_t> SearchBallPosTagged(Coordinate x, Double_t fdist2) {
        return SearchBallPosTagged(x.GetCoord(), fdist2);
    }

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
    // KDTree.h:427, :435, :439, :441, KDCalcSmoothQuantities.cxx:992-1035,
    // :1209-1253: the gather form of the smoothed velocity at one particle or
    // position (full weight, no scatter), i.e. the local mean of each velocity
    // component.
    Coordinate CalcSmoothVelParticle(Int_t target, Int_t Nsmooth = 64, int densityset = 1);
    Coordinate CalcSmoothVelPosition(Double_t *x, Int_t Nsmooth = 64, int densityset = 1);
    Coordinate CalcSmoothVelPosition(Coordinate x, Int_t Nsmooth = 64, int densityset = 1) {
        return CalcSmoothVelPosition(x.GetCoord(), Nsmooth, densityset);
    }
    Coordinate CalcSmoothVelPosition(Real_t *x, Int_t Nsmooth = 64, int densityset = 1) {
        Double_t d[3] = {x[0], x[1], x[2]};
        return CalcSmoothVelPosition(d, Nsmooth, densityset);
    }
    // KDTree.h:454-461, KDCalcSmoothQuantities.cxx:1539-1701: weight by ID.
    // Each call runs one GPU pass (every particle, or the one position); for
    // many targets call nbx_smooth_local / nbx_smooth_local_points once.
    Double_t CalcSmoothLocalMeanParticle(Int_t target, Double_t *weight, Int_t Nsmooth = 64, int densityset = 1);
    Double_t CalcSmoothLocalDispParticle(Int_t target, Double_t *weight, Double_t localmean, Int_t Nsmooth = 64,
                                         int densityset = 1);
    Double_t CalcSmoothLocalMeanPosition(Double_t *x, Double_t *weight, Int_t Nsmooth = 64, int densityset = 1);
    Double_t CalcSmoothLocalDispPosition(Double_t *x, Double_t *weight, Double_t localmean, Int_t Nsmooth = 64,
                                         int densityset = 1);
    Double_t CalcSmoothLocalMeanPosition(Coordinate x, Double_t *weight, Int_t Nsmooth = 64, int densityset = 1) {
        return CalcSmoothLocalMeanPosition(x.GetCoord(), weight, Nsmooth, densityset);
    }
    Double_t CalcSmoothLocalDispPosition(Coordinate x, Double_t *weight, Double_t localmean, Int_t Nsmooth = 64,
                                         int densityset = 1) {
        return CalcSmoothLocalDispPosition(x.GetCoord(), weight, localmean, Nsmooth, densityset);
    }
    void CalcVelDensity(Int_t Nsmooth = 64, Int_t Nsearch = 64);
    Double_t CalcDensityParticle(Int_t target, Int_t Nsmooth = 64);
    // KDTree.h:425, KDCalcSmoothQuantities.cxx:845-921 (TPHYS trees with
    // velocities): served like CalcDensityParticle from one batched
    // CalcVelDensity per (Nsmooth, Nsearch); the workspace arguments (iflag,
    // pq, pq2, nnIDs, vdist) are accepted and not needed.
    Double_t CalcVelDensityParticle(Int_t target, Int_t Nsmooth = 64, Int_t Nsearch = 64, int iflag = 0,
                                    PriorityQueue *pq = nullptr, PriorityQueue *pq2 = nullptr, Int_t *nnIDs = nullptr,
                                    Double_t *vdist = nullptr);
    // KDTree.h:432, :436, :442 (TPHYS and TVEL trees; v is unused there);
    // KDCalcSmoothQuantities.cxx:1092-1149, :1307-1364, :1521-1524
    Double_t CalcDensityPosition(Double_t *x, Int_t Nsmooth = 64, Double_t *v = NULL);
    Double_t CalcDensityPosition(Coordinate x, Int_t Nsmooth = 64, Coordinate v = Coordinate(0.)) {
        return CalcDensityPosition(x.GetCoord(), Nsmooth, v.GetCoord());
    }
    Double_t CalcDensityPosition(Real_t *x, Int_t Nsmooth = 64, Real_t *v = NULL) {
        Double_t d[3] = {x[0], x[1], x[2]};
        return CalcDensityPosition(d, Nsmooth, nullptr);
    }
    // KDCalcSmoothQuantities.cxx:393-476 (new[] array by ID; the caller
    // delete[]s it) and, per particle, :922-990.  Nsmooth > Nsearch is taken
    // as Nsearch before anything else, as the reference does (:405-409).
    Double_t *CalcVelDensityWithPhysDensity(Int_t Nsmooth = 64, Int_t Nsearch = 64, int densityset = 1);
    Double_t CalcVelDensityWithPhysDensityParticle(Int_t target, Int_t Nsmooth = 64, Int_t Nsearch = 64,
                                                   int densityset = 1);
    // KDTree.h:433, :437, :443 (TPHYS trees with velocities);
    // KDCalcSmoothQuantities.cxx:1150-1208, :1365-1423, :1525-1528
    Double_t CalcVelDensityPosition(Double_t *x, Double_t *v, Int_t Nsmooth = 64, Int_t Nsearch = 64);
    Double_t CalcVelDensityPosition(Coordinate x, Coordinate v, Int_t Nsmooth = 64, Int_t Nsearch = 64) {
        return CalcVelDensityPosition(x.GetCoord(), v.GetCoord(), Nsmooth, Nsearch);
    }
    Double_t CalcVelDensityPosition(Real_t *x, Real_t *v, Int_t Nsmooth = 64, Int_t Nsearch = 64) {
        Double_t dx[3] = {x[0], x[1], x[2]}, dv[3] = {v[0], v[1], v[2]};
        return CalcVelDensityPosition(dx, dv, Nsmooth, Nsearch);
    }

    // KDTree.h:480-482, KDFOF.cxx:29-153.  Labels are canonical (groups by
    // size, descending; ties by smallest member ID).  With ipcheckflag, the
    // particles whose check(p, params) is nonzero take no part in the links
    // and keep that value as their label, -1 becoming 0 (KDFOF.cxx:63-64,
    // :107).  The linked lists, when given, hold every linked component's
    // members in tree-index order, including those of components pruned
    // below minnum, as the reference leaves them (KDFOF.cxx:98-104); a
    // particle linked to none is a list of its own.  pLen[g] is the size of
    // group g for g = 1..numgroup and 0 after it.
    Int_t *FOF(Double_t fdist, Int_t &numgroup, Int_t minnum = 8, int order = 0, Int_tree_t *pHead = nullptr,
               Int_tree_t *pNext = nullptr, Int_tree_t *pTail = nullptr, Int_tree_t *pLen = nullptr,
               int ipcheckflag = 0, FOFcheckfunc check = Pnocheck, Double_t *params = nullptr);
    // KDTree.h:503, KDFOF.cxx:157-265: cmp = FOF3d (TPHYS tree) or FOF6d (TPHS
    // tree) with params as the reference's; any other comparator is reported
    // as unsupported.  With ipcheckflag, particles whose check returns -1 take
    // no part (label 0); a check returning a positive value (which the
    // reference lets be linked but not start a group) is reported as
    // unsupported.  Labels and lists as FOF's.
    Int_t *FOFCriterion(FOFcompfunc cmp, Double_t *params, Int_t &numgroups, Int_t minnum = 8, int order = 0,
                        int ipcheckflag = 0, FOFcheckfunc check = Pnocheck, Int_tree_t *pHead = nullptr,
                        Int_tree_t *pNext = nullptr, Int_tree_t *pTail = nullptr, Int_tree_t *pLen = nullptr);

    void OverWriteInputOrder();
    void SetResetOrder(bool);

    static void SetThrowOnError(bool on);
    // Precision form for the batched calls (NBX_FORM_MIXED): the FP64 form's
    // neighbour sets (certified, else re-run in FP64) and distances, with the
    // densities summed from FP64 distances (tests hold them within 2e-6 of
    // the FP64 form; about 1e-15 is measured).  Off by default.
    static void SetPrecisionForm(bool mixed);

  private:
    struct BallBlock {  // tagged ball rows of one block of tree indices
        std::vector<int64_t> off;
        std::vector<Int_t> idx;  // tree indices
        std::vector<Double_t> d2;
        unsigned long long used;
    };
    void density_to_bucket(Int_t Nsmooth);
    Double_t smooth_local_one(Int_t target, const Double_t *x, Double_t *weight, const Double_t *mean, Int_t Nsmooth,
                              int densityset);
    Coordinate *smooth_moment(int which, Coordinate *smvel, Matrix *smveldisp, Int_t Nsmooth, int densityset,
                              int meanvelset, int veldispset);
    std::vector<double> rho_by_id();
    void drop_density_caches();
    const BallBlock &ball_block(Int_t tt, Double_t fdist2);
    Int_t *fof_run(int cmp, Double_t fdist, Double_t *params, Int_t &numgroup, Int_t minnum, Int_tree_t *pHead,
                   Int_tree_t *pNext, Int_tree_t *pTail, Int_tree_t *pLen, int ipcheckflag, FOFcheckfunc check,
                   const char *what);
    Particle *bucket;
    Int_t numparts;
    Int_t b;
    int treetype, kernfunctype;
    Int_t numnodes, numleafnodes;
    Double_t period[3];
    bool periodic;
    bool iresetorder;
    nbx_params prm;
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
    int pcache_k, pcache_s, pcache_set;
    std::vector<Double_t> pcache;
    Double_t ball_r2;
    unsigned long long ball_tick;
    std::map<Int_t, std::unique_ptr<BallBlock>> ball_cache;

    void init(Particle *p, Int_t n, Int_t bucket_size, int TreeType, int KernType, int KernRes, int Splitting,
              int Aniso, int ScaleSpace, Double_t *Period);
    void batch_knn(int variant, Int_t k, Int_t *nn, Double_t *d2);
    void ensure_cache(int variant, Int_t k);
};

}  // namespace NBody

#endif
