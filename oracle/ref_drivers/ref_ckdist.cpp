/* Driver for the reference generator's correlated-k and table steps,
 * correlatedKT (nbcorrk/correlatedK.cpp:145-301) and calcProbT
 * (nbcorrk/calc_prob.cpp:17-151), linked with the generator's own sources
 * (its main renamed at compile time and never called), at the compile-time
 * sizes of nbcorrk/param.h (nT = 2 temperatures, n_pwrk = 20000, nQuad = 16).
 *
 * usage: ref_ckdist in.bin out.bin   (run where errors/ exists: correlatedKT
 * writes its diagnostics there)
 * in.bin (little-endian): int32 nB; double T[nT]; then per temperature i and
 * band j: int32 n, double wv_left, wv_right, wvn[n], kvn[n].
 * out.bin: per i: double kP; per (i, j): double kq[nQuad], k_avg, cuprob,
 * cuprob_new2[nQuad].  Fields are those the reference fills. */
#include <cstdio>
#include <cstdlib>
#include "param.h"
#include "functions.h"

int main(int argc, char **argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: ref_ckdist in.bin out.bin\n"); return 2; }
    FILE *fi = std::fopen(argv[1], "rb");
    if (!fi) return 1;
    int nB = 0;
    if (std::fread(&nB, sizeof(int), 1, fi) != 1 || nB < 1) return 1;
    TempBand *tb = new TempBand[nT];
    for (int i = 0; i < nT; ++i) {
        if (std::fread(&tb[i].T, sizeof(double), 1, fi) != 1) return 1;
        tb[i].totbands = nB;
        tb[i].allocNarrBands(nB);
    }
    for (int i = 0; i < nT; ++i)
        for (int j = 0; j < nB; ++j) {
            NarrowBand2 &b = tb[i].narrband[j];
            int n = 0;
            if (std::fread(&n, sizeof(int), 1, fi) != 1 || n < 1) return 1;
            b.allocVar(n);
            b.size = n;
            if (std::fread(&b.wv_left, sizeof(double), 1, fi) != 1) return 1;
            if (std::fread(&b.wv_right, sizeof(double), 1, fi) != 1) return 1;
            if (std::fread(b.wvn, sizeof(double), n, fi) != (size_t)n) return 1;
            if (std::fread(b.kvn, sizeof(double), n, fi) != (size_t)n) return 1;
            b.wvc = (b.wv_left + b.wv_right) / 2;   /* as ckdist_hitran_part.cpp:112 */
        }
    std::fclose(fi);
    correlatedKT(tb);
    calcProbT(tb);
    FILE *fo = std::fopen(argv[2], "wb");
    if (!fo) return 1;
    for (int i = 0; i < nT; ++i) {
        std::fwrite(&tb[i].kP, sizeof(double), 1, fo);
        for (int j = 0; j < nB; ++j) {
            const NarrowBand2 &b = tb[i].narrband[j];
            std::fwrite(b.kq, sizeof(double), nQuad, fo);
            std::fwrite(&b.k_avg, sizeof(double), 1, fo);
            std::fwrite(&b.cuprob, sizeof(double), 1, fo);
            std::fwrite(b.cuprob_new2, sizeof(double), nQuad, fo);
        }
    }
    std::fclose(fo);
    return 0;
}
