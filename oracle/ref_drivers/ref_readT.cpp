/* Driver for the reference table reader readT (src/montecarlo/read_tables.cpp
 * :21-95), run in a directory holding properties/ with the reference's
 * compile-time sizes (nB, nT, nQ of src/param/NarrowBand.h:10-12).
 * Writes the parsed tables as raw little-endian float32 to the file named by
 * argv[2], in this order: Tnb[nT], kP[nT], kappamax, then per band
 * wvl, wvr, wvc, kavg, kq[nT][nQ]; then prob_h[nT*nB], probg_h[nT*nB*nQ].
 * usage: ref_readT Tmax out.bin
 * The two arrays below are the storage that mc_gpu_noint.cu:23 defines for
 * read_tables.cpp's extern declarations (read_tables.cpp:8). */
#include <cstdio>
#include <cstdlib>
#include "param.h"
#include "definitions.h"
#include "functions.h"

myfloat prob[nT][nB], probg[nT][nB][nQ];

int main(int argc, char **argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: ref_readT Tmax out.bin\n"); return 2; }
    const myfloat Tmax = (myfloat)std::atof(argv[1]);
    NarrowBand *nb = (NarrowBand *)std::malloc(nB * sizeof(NarrowBand));
    myfloat *Tnb = (myfloat *)std::malloc(nT * sizeof(myfloat));
    myfloat *kP = (myfloat *)std::malloc(nT * sizeof(myfloat));
    myfloat *prob_h = (myfloat *)std::malloc((size_t)nB * nT * sizeof(myfloat));
    myfloat *probg_h = (myfloat *)std::malloc((size_t)nB * nT * nQ * sizeof(myfloat));
    myfloat kappamax = 0;
    readT(nb, Tnb, kP, prob_h, probg_h, Tmax, &kappamax);
    FILE *fp = std::fopen(argv[2], "wb");
    if (!fp) return 1;
    std::fwrite(Tnb, sizeof(myfloat), nT, fp);
    std::fwrite(kP, sizeof(myfloat), nT, fp);
    std::fwrite(&kappamax, sizeof(myfloat), 1, fp);
    for (int b = 0; b < nB; ++b) {
        std::fwrite(&nb[b].wvl, sizeof(myfloat), 1, fp);
        std::fwrite(&nb[b].wvr, sizeof(myfloat), 1, fp);
        std::fwrite(&nb[b].wvc, sizeof(myfloat), 1, fp);
        std::fwrite(&nb[b].kavg, sizeof(myfloat), 1, fp);
        std::fwrite(&nb[b].kq[0][0], sizeof(myfloat), (size_t)nT * nQ, fp);
    }
    std::fwrite(prob_h, sizeof(myfloat), (size_t)nT * nB, fp);
    std::fwrite(probg_h, sizeof(myfloat), (size_t)nT * nB * nQ, fp);
    std::fclose(fp);
    return 0;
}
