/* Driver for the reference's cubic-spline routines spline and splint
 * (src/montecarlo/spline.cpp:180-226), compiled from that file as it lies
 * (myfloat = float, param.h floatpoint 0): natural end conditions, as
 * interp3D calls them.
 * usage: ref_spline in.bin out.bin
 * in.bin (little-endian): int32 n, float x[n], float y[n], int32 m, float xs[m]
 * out.bin: float y2[n], then float ys[m] (splint at each xs) */
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "param.h"
#include "definitions.h"
#include "functions.h"

int main(int argc, char **argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: ref_spline in.bin out.bin\n"); return 2; }
    FILE *fi = std::fopen(argv[1], "rb");
    if (!fi) return 1;
    int n = 0, m = 0;
    if (std::fread(&n, sizeof(int), 1, fi) != 1 || n < 2) return 1;
    std::vector<myfloat> x(n + 1), y(n + 1), y2(n + 1);     /* 1-based, as the routines index */
    if (std::fread(&x[1], sizeof(myfloat), n, fi) != (size_t)n) return 1;
    if (std::fread(&y[1], sizeof(myfloat), n, fi) != (size_t)n) return 1;
    if (std::fread(&m, sizeof(int), 1, fi) != 1 || m < 0) return 1;
    std::vector<myfloat> xs(m), ys(m);
    if (m && std::fread(xs.data(), sizeof(myfloat), m, fi) != (size_t)m) return 1;
    std::fclose(fi);
    spline(x.data(), y.data(), n, 1.e30, 1.0e30, y2.data());
    for (int q = 0; q < m; ++q) ys[q] = splint(x.data(), y.data(), y2.data(), n, xs[q]);
    FILE *fo = std::fopen(argv[2], "wb");
    if (!fo) return 1;
    std::fwrite(&y2[1], sizeof(myfloat), n, fo);
    if (m) std::fwrite(ys.data(), sizeof(myfloat), m, fo);
    std::fclose(fo);
    return 0;
}
