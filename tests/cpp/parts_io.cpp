// Reads parts.out with NBody::ReadPartsOut, checks masses and IDs, writes it
// back with NBody::WritePartsOut.  Usage: parts_io in.out out.out
#include <cstdio>

#include "nbx_kdtree.h"

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    std::vector<NBody::Particle> parts;
    if (!NBody::ReadPartsOut(argv[1], parts)) return 3;
    for (size_t i = 0; i < parts.size(); ++i)
        if (parts[i].GetMass() != 1.0 || parts[i].GetID() != (int)i || parts[i].GetPID() != (int)i) return 4;
    if (!NBody::WritePartsOut(argv[2], parts)) return 5;
    std::printf("%zu\n", parts.size());
    return 0;
}
