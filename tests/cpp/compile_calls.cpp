// Compiles (never runs) the call forms reference callers write against
// NBody::KDTree, to show they build unchanged against the drop-in header:
// the driver's constructor call (src/tests/test_kdtree.cxx:236-245, with the
// kernel named through a null tree pointer as there), FindNearest(i, ...) from
// an OpenMP loop (:290), SearchBallPosTagged(i, r2) into an auto vector
// (:334), FOF(rdist, numgroup, minnum) (:363), and the other overloads of
// KDTree.h:300-373, :407-443, :480-503.
#include <vector>

#include "nbx_kdtree.h"

using namespace NBody;

static int no_check(Particle &, Double_t *) { return 0; }

double compile_calls(std::vector<Particle> &parts, int treetype, double rdist2fac, double adaptivefactor) {
    NBody::KDTree *tree = nullptr;
    const auto N = parts.size();
    Int_t b = 16;
    Double_t *P = NULL;
    Double_t **m = nullptr;
    int split = 0, aniso = 0, scalespace = 0;
    bool iBuildInParallel = true, iKeepInputOrder = false;
    tree = new NBody::KDTree(parts.data(), N, b, treetype, tree->KEPAN, 100, split, aniso, scalespace, P, m,
                             iBuildInParallel, iKeepInputOrder, rdist2fac * 1.0, adaptivefactor);
    double acc = tree->GetNumNodes() + tree->GetNumLeafNodes() + tree->GetKernNorm();

    const int num_nn = 16;
    std::vector<Int_t> nn(num_nn);
    std::vector<Double_t> nnd2(num_nn);
    for (Int_t i = 0; i < (Int_t)N; i++) tree->FindNearest(i, nn.data(), nnd2.data(), num_nn);
    const double r2 = 0.01;
    for (Int_t i = 0; i < (Int_t)N; i++) {
        auto neighbours = tree->SearchBallPosTagged(i, r2);
        acc += neighbours.size();
    }
    Int_t numgroup = 0;
    auto fofarr = tree->FOF(1.0, numgroup, 20);
    delete[] fofarr;

    // the other forms
    Double_t x[6] = {0.5, 0.5, 0.5, 0, 0, 0}, v[3] = {0, 0, 0};
    Real_t xf[3] = {0.5f, 0.5f, 0.5f}, vf[3] = {0.f, 0.f, 0.f};
    Coordinate cx(x), cv(v);
    std::vector<Int_t *> nnp(N);
    std::vector<Double_t *> d2p(N);
    tree->FindNearest(nnp.data(), d2p.data(), num_nn);
    tree->FindNearestPos(nnp.data(), d2p.data(), num_nn);
    tree->FindNearestPhase(nnp.data(), d2p.data(), num_nn);
    tree->FindNearestPos(0, nn.data(), nnd2.data(), num_nn);
    tree->FindNearestPhase(0, nn.data(), nnd2.data(), num_nn);
    tree->FindNearest(x, nn.data(), nnd2.data(), num_nn);
    tree->FindNearestPos(x, nn.data(), nnd2.data(), num_nn);
    tree->FindNearestPos(cx, nn.data(), nnd2.data(), num_nn);
    tree->FindNearestVel(v, nn.data(), nnd2.data(), num_nn);
    tree->FindNearestVel(cv, nn.data(), nnd2.data(), num_nn);
    tree->FindNearestPhase(x, v, nn.data(), nnd2.data(), num_nn);
    tree->FindNearestPhase(cx, cv, nn.data(), nnd2.data(), num_nn);
    std::vector<Int_t> group(N);
    std::vector<Double_t> gd2(N);
    tree->SearchBallPos(0, r2, 1, group.data(), gd2.data());
    tree->SearchBallPos(x, r2, 1, group.data(), gd2.data());
    tree->SearchBallPos(cx, r2, 1, group.data(), gd2.data());
    std::vector<Int_t> tagged(N);
    acc += tree->SearchBallPosTagged(0, r2, tagged.data());
    acc += tree->SearchBallPosTagged(x, r2, tagged.data());
    acc += tree->SearchBallPosTagged(cx, r2, tagged.data());
    acc += tree->SearchBallPosTagged(x, r2).size() + tree->SearchBallPosTagged(cx, r2).size();

    tree->CalcDensity(32);
    tree->CalcVelDensity(32, 64);
    acc += tree->CalcDensityParticle(0, 32) + tree->CalcVelDensityParticle(0, 32, 64);
    acc += tree->CalcDensityPosition(x, 32) + tree->CalcDensityPosition(cx, 32) + tree->CalcDensityPosition(xf, 32);
    acc += tree->CalcVelDensityPosition(x, v, 32, 64) + tree->CalcVelDensityPosition(cx, cv, 32, 64) +
           tree->CalcVelDensityPosition(xf, vf, 32, 64);
    Double_t *vden = tree->CalcVelDensityWithPhysDensity(32, 64, 0);
    acc += vden[0] + tree->CalcVelDensityWithPhysDensityParticle(0, 32, 64);
    delete[] vden;
    Coordinate *smvel = tree->CalcSmoothVel(32, 1);
    Matrix *disp = tree->CalcSmoothVelDisp(smvel, 32, 1, 1);
    Coordinate *skew = tree->CalcSmoothVelSkew(smvel, disp, 32);
    Coordinate *kurt = tree->CalcSmoothVelKurtosis(smvel, disp, 32);
    acc += smvel[0][0] + disp[0](0, 0) + skew[0][0] + kurt[0][0];
    delete[] smvel;
    delete[] disp;
    delete[] skew;
    delete[] kurt;
    acc += tree->CalcSmoothVelParticle(0, 32)[0] + tree->CalcSmoothVelPosition(x, 32)[0] +
           tree->CalcSmoothVelPosition(cx, 32)[1] + tree->CalcSmoothVelPosition(xf, 32)[2];
    std::vector<Double_t> w(N, 1.0);
    acc += tree->CalcSmoothLocalMeanParticle(0, w.data(), 32) + tree->CalcSmoothLocalDispParticle(0, w.data(), 1.0) +
           tree->CalcSmoothLocalMeanPosition(x, w.data()) + tree->CalcSmoothLocalMeanPosition(cx, w.data()) +
           tree->CalcSmoothLocalDispPosition(x, w.data(), 1.0) + tree->CalcSmoothLocalDispPosition(cx, w.data(), 1.0);

    std::vector<Int_tree_t> head(N), next(N), tail(N), len(N);
    Double_t params[8] = {0, 0.01, 0, 0, 0, 0, 0.01, 0};
    fofarr = tree->FOF(0.1, numgroup, 8, 1, head.data(), next.data(), tail.data(), len.data(), 1, no_check, params);
    delete[] fofarr;
    fofarr = tree->FOFCriterion(FOF3d, params, numgroup, 8, 0, 0, Pnocheck, head.data(), next.data(), tail.data(),
                                len.data());
    delete[] fofarr;
    tree->SetResetOrder(true);
    tree->OverWriteInputOrder();
    delete tree;
    return acc;
}
