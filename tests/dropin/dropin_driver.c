/* A C caller of the drop-in entry points, driven the way the DNS drives
 * them (src/dns/channel.f:237-269): per radiative step MC_GPU(TMC,
 * xstart(3), nrank) launches the solve and GET_RESULTS(qr, vr, nrank)
 * harvests it, every argument passed by reference as gfortran passes the
 * implicit F77 calls.  A harvest before any launch must leave its outputs
 * untouched and report an error.
 *
 * usage: dropin_driver table_dir field.bin out.bin imax jmax kmax p_col
 *                      zstart photmax nVar calcwalls seed steps
 * field.bin: (imax+2)(jmax+2)(kmax+2) floats, i fastest.
 * out.bin:   res then var of the last step, (imax+2)(jmax+2)(kmax/p_col+2)
 *            floats each.  Exit status 0 on success. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "radchan.h"

int main(int argc, char **argv) {
    if (argc != 14) {
        fprintf(stderr, "usage: dropin_driver table_dir field.bin out.bin imax jmax kmax p_col zstart photmax nVar "
                        "calcwalls seed steps\n");
        return 2;
    }
    rc_params p;
    rc_default_params(&p);
    p.imax = atoi(argv[4]); p.jmax = atoi(argv[5]); p.kmax = atoi(argv[6]); p.p_col = atoi(argv[7]);
    int zstart = atoi(argv[8]);
    p.photmax = atoi(argv[9]); p.nVar = atoi(argv[10]); p.calcwalls = atoi(argv[11]);
    p.seed = (uint64_t)strtoull(argv[12], NULL, 10);
    const int steps = atoi(argv[13]);
    /* the channel of the tests: 2 x 2 pi x 2 pi, one grid level, black x
     * walls (adiabatic), periodic y and z, the 360-band tables */
    p.Lx = 2.0f; p.Ly = 6.28318530717958648f; p.Lz = 6.28318530717958648f;
    p.fact_mesh = 1.5; p.chanHeight = 0.5;
    p.grid_num = 1;
    for (int g = 0; g < RC_MAXLEV; ++g) { p.maxi[g] = p.maxj[g] = p.maxk[g] = 1; p.maxs[g] = 1; }
    p.maxi[0] = p.imax; p.maxj[0] = p.jmax; p.maxk[0] = p.kmax; p.maxs[0] = 100000;
    int bd[6] = {2, 2, 1, 1, 1, 1};
    for (int f = 0; f < 6; ++f) { p.bd[f] = bd[f]; p.eps[f] = 1.0f; p.Tw[f] = 0.0f; }
    p.adiabw = p.adiabe = 1;
    p.nB = 360; p.nT = 20; p.nQ = 16;
    p.sampler = 0;
    if (rc_set_params(&p) != 0) { fprintf(stderr, "rc_set_params: %s\n", rc_last_error()); return 1; }
    if (rc_set_table_dir(argv[1]) != 0) { fprintf(stderr, "rc_set_table_dir: %s\n", rc_last_error()); return 1; }
    const size_t nt = (size_t)(p.imax + 2) * (p.jmax + 2) * (p.kmax + 2);
    const size_t no = (size_t)(p.imax + 2) * (p.jmax + 2) * (p.kmax / p.p_col + 2);
    float *T = malloc(nt * sizeof(float)), *res = malloc(no * sizeof(float)), *var = malloc(no * sizeof(float));
    if (!T || !res || !var) return 1;
    FILE *f = fopen(argv[2], "rb");
    if (!f || fread(T, sizeof(float), nt, f) != nt) { fprintf(stderr, "cannot read %s\n", argv[2]); return 1; }
    fclose(f);
    int rank = 0;
    /* a harvest without a launch: outputs untouched, an error reported */
    for (size_t i = 0; i < no; ++i) res[i] = var[i] = -7.0f;
    get_results_(res, var, &rank);
    for (size_t i = 0; i < no; ++i)
        if (res[i] != -7.0f || var[i] != -7.0f) { fprintf(stderr, "outputs touched without a launch\n"); return 3; }
    if (strlen(rc_last_error()) == 0) { fprintf(stderr, "no error reported\n"); return 3; }
    for (int s = 0; s < steps; ++s) {
        mc_gpu_(T, &zstart, &rank);
        get_results_(res, var, &rank);
    }
    f = fopen(argv[3], "wb");
    if (!f) return 1;
    fwrite(res, sizeof(float), no, f);
    fwrite(var, sizeof(float), no, f);
    fclose(f);
    free(T); free(res); free(var);
    return 0;
}
