/* Driver for the reference generator's front end: absWideBand
 * (nbcorrk/read_files.cpp:22-73) on a reference spectrum file, and allocT
 * + readFillT (alloc_bands.cpp:17-51, read_files.cpp:76-114) filling given
 * narrow bands from the temperature files allocT names, linked with the
 * generator's own sources (its main renamed at compile time and never
 * called), at the compile-time constants of nbcorrk/param.h (toll 1e-6,
 * wvtoll 25000, dwv 0.01, nT = 2 temperatures from Tinit = 2500 K by
 * DT = 10 K).
 *
 * usage: ref_lbl ref_spectrum.txt bands.bin out.bin
 *   run in a directory holding CO2-data/CO2_2500K_HITRAN.txt and
 *   CO2-data/CO2_2510K_HITRAN.txt (the names allocT builds).
 * bands.bin (little-endian): int32 nb; per band: int32 n, double wvn[n].
 * out.bin: int32 nwide; per wide band: int32 size, double wv_first,
 *   wv_last; then per temperature: double T, kPr, kvn of every band. */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include "param.h"
#include "functions.h"

int main(int argc, char **argv) {
    if (argc != 4) { std::fprintf(stderr, "usage: ref_lbl ref.txt bands.bin out.bin\n"); return 2; }
    std::deque<PartSpecBand> wide;
    absWideBand(wide, argv[1], 2500.);
    FILE *fb = std::fopen(argv[2], "rb");
    if (!fb) return 1;
    int nb = 0;
    if (std::fread(&nb, sizeof(int), 1, fb) != 1 || nb < 1) return 1;
    /* the given narrow bands as the narrow bands of one wide band, the form
     * allocT copies them from */
    std::deque<PartSpecBand> bands(1);
    bands[0].numbands = nb;
    bands[0].alloc_narrbands();
    for (int h = 0; h < nb; ++h) {
        int n = 0;
        if (std::fread(&n, sizeof(int), 1, fb) != 1 || n < 1) return 1;
        for (int t = 0; t < n; ++t) {
            double w;
            if (std::fread(&w, sizeof(double), 1, fb) != 1) return 1;
            bands[0].narrband[h].wvn.push_back(w);
        }
        bands[0].narrband[h].wv_left = bands[0].narrband[h].wvn.front();
        bands[0].narrband[h].wv_right = bands[0].narrband[h].wvn.back();
    }
    std::fclose(fb);
    TempBand *tb = new TempBand[nT];
    allocT(tb, bands, nb);
    readFillT(tb, bands);
    FILE *fo = std::fopen(argv[3], "wb");
    if (!fo) return 1;
    const int nw = (int)wide.size();
    std::fwrite(&nw, sizeof(int), 1, fo);
    for (int i = 0; i < nw; ++i) {
        const int sz = (int)wide[i].wv.size();
        const double a = wide[i].wv.front(), b = wide[i].wv.back();
        std::fwrite(&sz, sizeof(int), 1, fo);
        std::fwrite(&a, sizeof(double), 1, fo);
        std::fwrite(&b, sizeof(double), 1, fo);
    }
    for (int i = 0; i < nT; ++i) {
        std::fwrite(&tb[i].T, sizeof(double), 1, fo);
        std::fwrite(&tb[i].kPr, sizeof(double), 1, fo);
        for (int h = 0; h < nb; ++h) std::fwrite(tb[i].narrband[h].kvn, sizeof(double), tb[i].narrband[h].size, fo);
    }
    std::fclose(fo);
    return 0;
}
