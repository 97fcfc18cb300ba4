// Writes the reference's own Owen-scrambled Sobol points (Sobol2D,
// src/math/Sobol.cpp) for a list of seeds, as raw doubles: for each seed,
// n x-values then n y-values.  Linked against the reference library built by
// oracle/ref.mk in the container that holds the reference sources; the output
// is converted to tests/golden/sobol_points.npz by make_fixtures.py.
//   usage: ref_sobol_dump OUT N SEED...
#include "Sobol.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char **argv)
{
    if (argc < 4)
    {
        std::fprintf(stderr, "usage: %s OUT N SEED...\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "wb");
    if (!f)
        return 3;
    const int n = std::atoi(argv[2]);
    for (int a = 3; a < argc; ++a)
    {
        Sobol2D s((uint32_t)std::strtoul(argv[a], nullptr, 10));
        std::vector<double> x, y;
        s.generate(n, x, y);
        std::fwrite(x.data(), sizeof(double), n, f);
        std::fwrite(y.data(), sizeof(double), n, f);
    }
    std::fclose(f);
    return 0;
}
