/*
 * rc_internal.h -- host-side internals shared by the table/grid code and the
 * HIP solver.
 */
#ifndef RC_INTERNAL_H
#define RC_INTERNAL_H

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/radchan.h"

namespace rc {

constexpr double kStefan = 5.670373e-08;                 /* param.h:85 */
constexpr double kPi = 3.14159265358979323846;           /* param.h:88 */
constexpr double kC1 = 3.741771790075259e-16;            /* memory_copy_noint.cu:292 */
constexpr double kC2 = 0.014387741858429;                /* memory_copy_noint.cu:293 */

enum Face { FW = 0, FE = 1, FS = 2, FN = 3, FB = 4, FT = 5 };

/* error channel behind rc_last_error() */
void set_error(const std::string &msg);
const char *last_error();

/* owned copy of an rc_tables */
struct Tables {
    int nB = 0, nT = 0, nQ = 0;
    std::vector<float> Tnb, kP, wvl, wvr, wvc, kavg, kq, prob, probg, wq;
    void resize(int nb, int nt, int nq);
    rc_tables view();
};

struct Grid {
    int im, jm, km;
    std::vector<float> x, xu, y, yu, z, zu;
};

float planck(float T, float nu);
void build_grid(const rc_params &p, int level, Grid &g);
int read_tables(const std::string &dir, double chanHeight, Tables &t);
int grey_tables(const rc_params &p, float Tmax, Tables &t);
int write_tables(const std::string &dir, const Tables &t, double chanHeight);
/* the generator's front end: wide and narrow bands of a line-by-line
 * spectrum, and the fill of the bands at one temperature */
int lbl_bands(int64_t n, const double *wv, const double *kv, double toll, double wvtoll, double dwv,
              const double *deltawv, std::vector<int64_t> &first, std::vector<int64_t> &count,
              std::vector<double> &wvl, std::vector<double> &wvr, std::vector<double> &wvc, int &nwide);
int64_t lbl_fill(int64_t n, const double *wv, const double *kv, double T, double dwv, int64_t ntot,
                 const double *wvn, double *kvn, double &kPr);
int ck_tables(int nB, int nT, int nQ, const double *T, const double *wvl, const double *wvr, const double *kq,
              const double *kavg, const double *wq, double chanHeight, Tables &t);
void sort_idx(const float *kavg, int nB, int32_t *idx);
int check_params(const rc_params &p);
void default_params(rc_params &p);
/* quadrature weights of NarrowBand.h:21-27 for nQ in {7,16,17} (else empty) */
std::vector<float> reference_wq(int nQ);

}  // namespace rc

#endif
