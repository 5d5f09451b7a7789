// Drives NBody::KDTree (include/nbx_kdtree.h) the way reference callers do
// (src/tests/test_kdtree.cxx, src/Analysis/Density.cxx): build from a Particle
// array, batched and per-particle FindNearest inside an OpenMP loop,
// CalcDensity, FOF, destruction restoring the input order.
// usage: test_kdtree_shim <in.bin> <out.bin> <k> <fof_b> <periodic:0|1> <L>
// in.bin : int32 n, then n x (x y z vx vy vz m) doubles
// out.bin: int32 n, k; n*k int32 neighbour IDs (by ID); n doubles rho (input
//          order); n int32 FOF labels; int32 status flags
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nbx_kdtree.h"

using namespace NBody;

int main(int argc, char **argv) {
    if (argc < 7) return 2;
    int k = atoi(argv[3]);
    double b = atof(argv[4]);
    int periodic = atoi(argv[5]);
    double L = atof(argv[6]);
    FILE *f = fopen(argv[1], "rb");
    int n = 0;
    if (!f || fread(&n, 4, 1, f) != 1) return 3;
    std::vector<double> raw(7 * (size_t)n);
    if (fread(raw.data(), 8, raw.size(), f) != raw.size()) return 3;
    fclose(f);
    std::vector<Particle> parts(n);
    for (int i = 0; i < n; ++i) {
        const double *r = &raw[7 * (size_t)i];
        parts[i] = Particle(r[6], r[0], r[1], r[2], r[3], r[4], r[5], 12345, 1, 0.0, 0.0, 1000 + i);
    }
    double period[3] = {L, L, L};
    std::vector<int> ids((size_t)n * k);
    std::vector<int> flags(4, 0);
    std::vector<double> rho(n);
    int *labels = nullptr;
    int ngroup = 0;
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
            int id = parts[i].GetID();
            for (int j = 0; j < k; ++j) ids[(size_t)id * k + j] = nnflat[(size_t)i * k + j] >= 0 ? parts[nnflat[(size_t)i * k + j]].GetID() : -1;
        }
        fprintf(stderr, "stage: per-particle FindNearest (OpenMP)\n");
        // per-particle calls from an OpenMP loop agree with the batch rows
        int bad = 0;
#pragma omp parallel for schedule(guided) reduction(+ : bad)
        for (int i = 0; i < n; ++i) {
            std::vector<int> one(k);
            std::vector<double> od(k);
            tree.FindNearest(i, one.data(), od.data(), k);
            for (int j = 0; j < k; ++j) bad += (one[j] != nnflat[(size_t)i * k + j]) || (od[j] != d2flat[(size_t)i * k + j]);
        }
        flags[0] = bad;
        fprintf(stderr, "stage: CalcDensity\n");
        tree.CalcDensity(k);
        fprintf(stderr, "stage: FOF b=%g\n", b);
        labels = tree.FOF(b, ngroup, 8);
        fprintf(stderr, "stage: destroy\n");
        // pid and type ride along with the reordering
        int moved = 0;
        for (int i = 0; i < n; ++i) moved += (parts[i].GetPID() != 1000 + parts[i].GetID()) || parts[i].GetType() != 1;
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
    fprintf(stderr, "stage: write\n");
    FILE *o = fopen(argv[2], "wb");
    fwrite(&n, 4, 1, o);
    fwrite(&k, 4, 1, o);
    fwrite(ids.data(), 4, ids.size(), o);
    fwrite(rho.data(), 8, n, o);
    fwrite(labels, 4, n, o);
    fwrite(flags.data(), 4, 4, o);
    fclose(o);
    delete[] labels;
    printf("shim driver ok n=%d k=%d groups=%d flags=%d %d %d\n", n, k, ngroup, flags[0], flags[1], flags[2]);
    return 0;
}
