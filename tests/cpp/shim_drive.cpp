// Drives NBody::KDTree (include/nbx_kdtree.h) the way reference callers do
// (src/tests/test_kdtree.cxx, src/Analysis/Density.cxx): build from a Particle
// array, batched and per-particle FindNearest inside an OpenMP loop,
// per-particle SearchBallPosTagged inside an OpenMP loop, CalcDensity, FOF
// with its linked lists, FOF with a check function, point queries, and
// destruction restoring the input order.
//
// nbx_shim_drive(argc, argv) takes the arguments
//   <in.bin> <out.bin> <k> <fof_b> <periodic:0|1> <L> <ball_r>
// in.bin : int32 n, then n x (x y z vx vy vz m) doubles
// out.bin: int32 n, k; n*k int32 neighbour IDs (by ID); n doubles rho (input
//          order); n int32 FOF labels; n int32 FOF labels with the check
//          (every 5th PID left out); n int32 ball-row counts (by ID), then the
//          rows as IDs (by ID order); 8 points x k neighbour IDs of
//          FindNearest(x); int32 status flags[8]; n int32 FOF list heads
//          (the ID at the head of each ID's list)
// It returns 0, or a nonzero code after printing what failed; errors of the
// tree are thrown (SetThrowOnError), so a harness in another program survives.
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "nbx_kdtree.h"

using namespace NBody;

// leaves out every fifth particle by PID (the check returns -1 for it)
static int skip_fifth(Particle &p, Double_t *) { return p.GetPID() % 5 == 0 ? -1 : 0; }

static int drive(int argc, char **argv) {
    if (argc < 8) return 2;
    const int k = atoi(argv[3]);
    const double b = atof(argv[4]);
    const int periodic = atoi(argv[5]);
    const double L = atof(argv[6]);
    const double rball = atof(argv[7]);
    FILE *f = fopen(argv[1], "rb");
    int n = 0;
    if (!f || fread(&n, 4, 1, f) != 1) return 3;
    std::vector<double> raw(7 * (size_t)n);
    if (fread(raw.data(), 8, raw.size(), f) != raw.size()) return 3;
    fclose(f);
    std::vector<Particle> parts(n);
    for (int i = 0; i < n; ++i) {
        const double *r = &raw[7 * (size_t)i];
        parts[i] = Particle(r[6], r[0], r[1], r[2], r[3], r[4], r[5], 12345, 1, 0.0, 0.0, i);
    }
    double period[3] = {L, L, L};
    std::vector<int> ids((size_t)n * k), flags(8, 0), lab1(n), lab2(n), ballcnt(n), lhead(n);
    std::vector<std::vector<int>> ballrow(n);
    std::vector<int> ptnn(8 * (size_t)k);
    std::vector<double> rho(n);
    int ngroup = 0, ngroup2 = 0;
    {
        fprintf(stderr, "stage: build n=%d periodic=%d\n", n, periodic);
        KDTree tree(parts.data(), n, 16, KDTree::TPHYS, KDTree::KSPH, 1000, 0, 0, 0, periodic ? period : nullptr);
        fprintf(stderr, "stage: batched FindNearest k=%d\n", k);
        // batched kNN, mapped to IDs through the reordered particle array
        std::vector<int> nnflat((size_t)n * k);
        std::vector<double> d2flat((size_t)n * k);
        std::vector<int *> nn(n);
        std::vector<double *> d2(n);
        for (int i = 0; i < n; ++i) { nn[i] = &nnflat[(size_t)i * k]; d2[i] = &d2flat[(size_t)i * k]; }
        tree.FindNearest(nn.data(), d2.data(), k);
        for (int i = 0; i < n; ++i) {
            const int id = parts[i].GetID();
            for (int j = 0; j < k; ++j) {
                const int t = nnflat[(size_t)i * k + j];
                ids[(size_t)id * k + j] = t >= 0 ? parts[t].GetID() : -1;
            }
        }
        fprintf(stderr, "stage: per-particle FindNearest and SearchBallPosTagged (OpenMP)\n");
        // per-particle calls from OpenMP loops, as test_kdtree.cxx:278-301 and :324-341
        int bad = 0;
        const double r2 = rball * rball;
#pragma omp parallel for schedule(guided) reduction(+ : bad)
        for (int i = 0; i < n; ++i) {
            std::vector<int> one(k);
            std::vector<double> od(k);
            tree.FindNearest(i, one.data(), od.data(), k);
            for (int j = 0; j < k; ++j)
                bad += (one[j] != nnflat[(size_t)i * k + j]) || (od[j] != d2flat[(size_t)i * k + j]);
            auto neighbours = tree.SearchBallPosTagged(i, r2);
            std::vector<int> row(neighbours.size());
            for (size_t e = 0; e < neighbours.size(); ++e) row[e] = parts[neighbours[e]].GetID();
            ballrow[parts[i].GetID()] = row;
        }
        flags[0] = bad;
        // the array form agrees with the vector form
        {
            std::vector<int> tagged(n);
            const int nt = tree.SearchBallPosTagged(n / 2, r2, tagged.data());
            auto v = tree.SearchBallPosTagged(n / 2, r2);
            flags[4] = nt != (int)v.size();
            for (int e = 0; e < nt && !flags[4]; ++e) flags[4] = tagged[e] != v[e];
        }
        fprintf(stderr, "stage: FindNearest(x)\n");
        for (int p = 0; p < 8; ++p) {
            Double_t x[3] = {raw[7 * (size_t)(p * 97 % n)] + 1e-3, raw[7 * (size_t)(p * 97 % n) + 1],
                             raw[7 * (size_t)(p * 97 % n) + 2]};
            std::vector<int> one(k);
            std::vector<double> od(k);
            tree.FindNearest(x, one.data(), od.data(), k);
            for (int j = 0; j < k; ++j) ptnn[(size_t)p * k + j] = one[j] >= 0 ? parts[one[j]].GetID() : -1;
        }
        fprintf(stderr, "stage: CalcDensity\n");
        tree.CalcDensity(k);
        fprintf(stderr, "stage: FOF b=%g\n", b);
        std::vector<Int_tree_t> head(n), next(n), tail(n), len(n);
        int *l1 = tree.FOF(b, ngroup, 8, 0, head.data(), next.data(), tail.data(), len.data());
        for (int i = 0; i < n; ++i) lab1[i] = l1[i];
        delete[] l1;
        // the lists: every linked component once from its head, in tree-index
        // order; a group's list holds its label and pLen members, a pruned
        // component's (label 0) fewer than minnum = 8
        {
            std::vector<int> seen(ngroup + 1, 0);
            int lbad = 0, listed = 0;
            for (int s = 0; s < n; ++s) {
                lhead[parts[s].GetID()] = parts[head[s]].GetID();
                if (head[s] != s) continue;
                const int g = lab1[parts[s].GetID()];
                if (g > 0 && seen[g]++) ++lbad;
                int c = 0, last = -1;
                for (int e = s; e != -1; e = next[e]) {
                    lbad += lab1[parts[e].GetID()] != g || e <= last || head[e] != s;
                    last = e;
                    ++c;
                }
                listed += c;
                lbad += tail[s] != last || (g > 0 ? c != len[g] : c >= 8);
            }
            for (int g = 1; g <= ngroup; ++g) lbad += seen[g] != 1;
            for (int g = ngroup + 1; g < n; ++g) lbad += len[g] != 0;
            flags[5] = lbad + (listed != n);
        }
        fprintf(stderr, "stage: FOF with a check function\n");
        int *l2 = tree.FOF(b, ngroup2, 8, 0, nullptr, nullptr, nullptr, nullptr, 1, skip_fifth, nullptr);
        for (int i = 0; i < n; ++i) lab2[i] = l2[i];
        delete[] l2;
        fprintf(stderr, "stage: destroy\n");
        // pid and type ride along with the reordering
        int moved = 0;
        for (int i = 0; i < n; ++i) moved += (parts[i].GetPID() != parts[i].GetID()) || parts[i].GetType() != 1;
        flags[1] = moved;
    }
    // ~KDTree sorted the array back into input order
    int order = 0;
    for (int i = 0; i < n; ++i) {
        order += parts[i].GetID() != i;
        rho[i] = parts[i].GetDensity();
    }
    flags[2] = order;
    flags[3] = ngroup;
    flags[6] = ngroup2;
    for (int i = 0; i < n; ++i) ballcnt[i] = (int)ballrow[i].size();
    fprintf(stderr, "stage: write\n");
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 4;
    fwrite(&n, 4, 1, o);
    fwrite(&k, 4, 1, o);
    fwrite(ids.data(), 4, ids.size(), o);
    fwrite(rho.data(), 8, n, o);
    fwrite(lab1.data(), 4, n, o);
    fwrite(lab2.data(), 4, n, o);
    fwrite(ballcnt.data(), 4, n, o);
    for (int i = 0; i < n; ++i)
        if (!ballrow[i].empty()) fwrite(ballrow[i].data(), 4, ballrow[i].size(), o);
    fwrite(ptnn.data(), 4, ptnn.size(), o);
    fwrite(flags.data(), 4, flags.size(), o);
    fwrite(lhead.data(), 4, n, o);
    if (fclose(o) != 0) return 4;
    fprintf(stderr, "shim driver ok n=%d k=%d groups=%d flags=%d %d %d %d %d\n", n, k, ngroup, flags[0], flags[1],
            flags[2], flags[4], flags[5]);
    return 0;
}

extern "C" int nbx_shim_drive(int argc, char **argv) {
    KDTree::SetThrowOnError(true);
    try {
        return drive(argc, argv);
    } catch (const std::exception &e) {
        fprintf(stderr, "shim driver: %s\n", e.what());
        return 5;
    }
}
