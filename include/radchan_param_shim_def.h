/*
 * radchan_param_shim_def.h -- included after src/param/param.h and
 * src/param/NarrowBand.h (see radchan_param_shim.h): defines the load-time
 * constructor that hands the compiled-in parameters to the library.
 */
#ifndef RADCHAN_PARAM_SHIM_DEF_H
#define RADCHAN_PARAM_SHIM_DEF_H

#ifndef RADCHAN_PARAM_SHIM_H
#error "include radchan_param_shim.h before param.h"
#endif

/* the grey block of NarrowBand.h:34-43 exists only when grey == 1 */
#if defined(abscoeff)
#define RADCHAN_GREY_VALUES abscoeff, Tmins, wvmax
#else
#define RADCHAN_GREY_VALUES 1.0f, 490.0f, 5000.0f
#endif

__attribute__((constructor)) static void radchan_params_from_macros(void) {
    rc_params prm;
    const float e6[6] = {epsw, epse, epss, epsn, epsb, epst};
    const float t6[6] = {Tww, Twe, Tws, Twn, Twb, Twt};
    const int b6[6] = {bdw, bde, bds, bdn, bdb, bdt};
    if (rc_params_compiled(&prm, imax, jmax, kmax, p_col, Lx, Ly, Lz, fact_mesh, nVar, photmax, chanHeight,
                           grid_num, maxi, maxj, maxk, maxs, srt, grey, calcwalls, e6, t6, adiabw, adiabe, toll,
                           tollS, b6, nT, nQ, nB, RADCHAN_GREY_VALUES) != 0) {
        fprintf(stderr, "radchan: compiled-in parameters rejected: %s\n", rc_last_error());
        return;
    }
#if defined(random)
    prm.sampler = (random == 1) ? 1 : 0;   /* param.h:62-63: 1 Sobol, other pseudo-random */
#endif
#if defined(floatpoint)
    prm.fp64 = (floatpoint == 1) ? 1 : 0;   /* param.h:64-65: 1 double, 0 float */
#endif
    if (rc_set_params(&prm) != 0)
        fprintf(stderr, "radchan: compiled-in parameters rejected: %s\n", rc_last_error());
}

#endif
