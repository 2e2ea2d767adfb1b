/*
 * rc_tables.cpp -- host-side set-up of the radiative solver: parameters,
 * grids, the spectral-table reader/writer and grey synthesis, band sort.
 *
 * Reference behaviour followed:
 *   parameters   src/param/param.h:8-100, src/param/NarrowBand.h:10-43
 *   grid         src/montecarlo/definitions.h:68-136 (Gridn::mk_grid)
 *   Tmax         src/montecarlo/mc_gpu_noint.cu:696-710
 *   tables       src/montecarlo/read_tables.cpp:21-95 (narrow band),
 *                :98-175 (grey); writer format nbcorrk/write_tables.cpp:18-109
 *   band sort    src/montecarlo/mc_gpu_noint.cu:1270-1294
 *   Planck       src/montecarlo/memory_copy_noint.cu:288-297
 */
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

#include "rc_internal.h"

namespace rc {

namespace {
std::mutex g_err_mu;
std::string g_err;
}  // namespace

void set_error(const std::string &msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
}
const char *last_error() {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

void Tables::resize(int nb, int nt, int nq) {
    nB = nb; nT = nt; nQ = nq;
    Tnb.assign(nt, 0.f); kP.assign(nt, 0.f);
    wvl.assign(nb, 0.f); wvr.assign(nb, 0.f); wvc.assign(nb, 0.f); kavg.assign(nb, 0.f);
    kq.assign((size_t)nb * nt * nq, 0.f);
    prob.assign((size_t)nt * nb, 0.f);
    probg.assign((size_t)nt * nb * nq, 0.f);
    wq.assign(nq, 0.f);
}

rc_tables Tables::view() {
    rc_tables t;
    t.nB = nB; t.nT = nT; t.nQ = nQ; t.pad = 0;
    t.Tnb = Tnb.data(); t.kP = kP.data();
    t.wvl = wvl.data(); t.wvr = wvr.data(); t.wvc = wvc.data(); t.kavg = kavg.data();
    t.kq = kq.data(); t.prob = prob.data(); t.probg = probg.data(); t.wq = wq.data();
    return t;
}

/* Planck spectral intensity per unit wavenumber nu [cm^-1] at T [K], in
 * W m^-2 sr^-1 per m^-1: (1/pi) C1 (100 nu)^3 / (exp(C2 100 nu / T) - 1).
 * Evaluated in FP64 with expm1 and rounded once to FP32. */
float planck(float T, float nu) {
    const double w = 100.0 * (double)nu;
    return (float)(1.0 / kPi * kC1 * w * w * w / std::expm1(kC2 * w / (double)T));
}

std::vector<float> reference_wq(int nQ) {
    /* NarrowBand.h:21-27 */
    static const float w16[16] = {0.09654009f, 0.09563872f, 0.09384440f, 0.09117387f, 0.08765209f, 0.08331193f,
                                  0.07819389f, 0.07234580f, 0.06582222f, 0.05868409f, 0.05099806f, 0.04283589f,
                                  0.03427387f, 0.02539207f, 0.01627439f, 0.00701862f};
    static const float w17[17] = {0.045f, 0.245f, 0.32f, 0.245f, 0.0495f, 0.0245f, 0.032f, 0.0245f, 0.00495f,
                                  0.00245f, 0.0032f, 0.00245f, 0.0005f, 0.0002722f, 0.00035556f, 0.000272222f, 0.00005f};
    static const float w7[7] = {0.04500f, 0.24500f, 0.32000f, 0.24500f, 0.05611f, 0.05125f, 0.03764f};
    if (nQ == 16) return std::vector<float>(w16, w16 + 16);
    if (nQ == 17) return std::vector<float>(w17, w17 + 17);
    if (nQ == 7) return std::vector<float>(w7, w7 + 7);
    if (nQ == 1) return std::vector<float>(1, 1.0f);
    return {};
}

void default_params(rc_params &p) {
    std::memset(&p, 0, sizeof p);
    p.imax = 406; p.jmax = 576; p.kmax = 2304; p.p_col = 48;
    p.Lx = 2.f; p.Ly = 6.283185307179586f; p.Lz = 37.69911184307751f;
    p.fact_mesh = 1.6; p.nVar = 1; p.photmax = 6000; p.chanHeight = 0.5;
    p.grid_num = 9;
    const int mi[9] = {1, 1, 1, 1, 1, 29, 29, 29, 203};
    const int mj[9] = {1, 3, 6, 12, 24, 48, 96, 96, 192};
    const int mk[9] = {1, 2, 4, 8, 16, 32, 64, 128, 256};
    for (int g = 0; g < 9; ++g) {
        p.maxi[g] = p.imax / mi[g]; p.maxj[g] = p.jmax / mj[g]; p.maxk[g] = p.kmax / mk[g];
        p.maxs[g] = (g < 8) ? 5 : 100000;
    }
    p.srt = 1; p.grey = 0; p.calcwalls = 0; p.grey_true = 0;
    for (int f = 0; f < 6; ++f) { p.eps[f] = 1.f; p.Tw[f] = 0.f; }
    p.adiabw = 1; p.adiabe = 1;
    p.toll = 1.0e-3f; p.tollS = 1.0e-3f;
    const int bd[6] = {2, 2, 1, 1, 4, 4};
    for (int f = 0; f < 6; ++f) p.bd[f] = bd[f];
    p.nT = 135; p.nQ = 16; p.nB = 999;
    p.abscoeff = 1.f; p.Tmins = 490.f; p.wvmax = 5000.f;
    p.seed = 1234;
    p.sampler = 1;      /* param.h:62-63: random 1 (Sobol directions) */
    p.fp64 = 0;   /* param.h:64-65: floatpoint 0 (single precision) */
}

int check_params(const rc_params &p) {
    auto fail = [](const std::string &m) { set_error("invalid parameters: " + m); return -1; };
    if (p.imax < 1 || p.jmax < 1 || p.kmax < 1) return fail("imax, jmax, kmax must be positive");
    if (p.p_col < 1 || p.kmax % p.p_col) return fail("kmax must be a multiple of p_col");
    if (p.grid_num < 1 || p.grid_num > RC_MAXLEV) return fail("grid_num must be in 1..9");
    if (p.maxi[0] != p.imax || p.maxj[0] != p.jmax || p.maxk[0] != p.kmax)
        return fail("level 0 must be the full grid (maxi[0]=imax, ...)");
    for (int g = 0; g < p.grid_num; ++g) {
        if (p.maxi[g] < 1 || p.maxj[g] < 1 || p.maxk[g] < 1) return fail("level sizes must be positive");
        if (p.imax % p.maxi[g] || p.jmax % p.maxj[g] || p.kmax % p.maxk[g])
            return fail("every level must divide the full grid");
        if (g > 0 && (p.maxi[g - 1] % p.maxi[g] || p.maxj[g - 1] % p.maxj[g] || p.maxk[g - 1] % p.maxk[g]))
            return fail("each level must divide the previous one");
        if (p.maxs[g] < 1) return fail("maxs must be positive");
    }
    if (p.photmax < 1 || p.nVar < 1) return fail("photmax and nVar must be positive");
    if ((int64_t)p.jmax * p.kmax >= (1 << 24) || (int64_t)p.imax * p.jmax * p.kmax >= ((int64_t)1 << 31))
        return fail("grid too large: jmax kmax must stay below 2^24 and imax jmax kmax below 2^31 (cell indices)");
    if (p.srt == 1 && p.photmax > 16384)
        return fail("srt: photmax above 16384 (the plain variant sorts a cell's band keys in one workgroup's LDS)");
    if (!(p.Lx > 0) || !(p.Ly > 0) || !(p.Lz > 0)) return fail("domain lengths must be positive");
    for (int f = 0; f < 4; ++f)
        if (p.bd[f] < 1 || p.bd[f] > 3) return fail("x and y faces take bd 1 (periodic), 2 (black) or 3 (specular)");
    for (int f = 4; f < 6; ++f)
        if (p.bd[f] < 1 || p.bd[f] > 5) return fail("z faces take bd 1..5");
    if ((p.bd[FW] == 1) != (p.bd[FE] == 1) || (p.bd[FS] == 1) != (p.bd[FN] == 1) || (p.bd[FB] == 1) != (p.bd[FT] == 1))
        return fail("periodic faces come in pairs");
    if (p.nT < 2 || p.nQ < 1 || p.nB < 1) return fail("table sizes nT>=2, nQ>=1, nB>=1");
    if ((int64_t)p.nB * p.nQ > 65535) return fail("nB*nQ must fit 16 bits");
    if (p.grey && p.nQ != 1) return fail("grey runs use nQ=1");
    if (p.grey && p.grey_true && p.nB != 1) return fail("true grey uses nB=1");
    if (!(p.toll > 0) || !(p.tollS > 0) || !(p.toll < 1) || !(p.tollS < 1))
        return fail("toll and tollS must lie in (0, 1)");
    if (p.calcwalls == 1 && !(p.adiabw == 1 || p.adiabe == 1)) return fail("calcwalls needs an adiabatic x wall");
    if (p.sampler != 0 && p.sampler != 1) return fail("sampler: 0 (Philox) or 1 (Sobol)");
    if (p.fp64 != 0 && p.fp64 != 1) return fail("floatpoint: 0 (single) or 1 (double precision)");
    return 0;
}

/* Gridn::mk_grid (definitions.h:68-136) for level `level`. */
void build_grid(const rc_params &p, int level, Grid &g) {
    const int im = p.maxi[level], jm = p.maxj[level], km = p.maxk[level], im0 = p.maxi[0];
    g.im = im; g.jm = jm; g.km = km;
    g.x.resize(im + 2); g.xu.resize(im + 1);
    g.y.resize(jm + 2); g.yu.resize(jm + 1);
    g.z.resize(km + 2); g.zu.resize(km + 1);
    /* uniform y and z: faces j*dy, centres halfway, ghost centres mirrored */
    const float dy = p.Ly * 1.0f / (float)jm;
    const float dz = p.Lz * 1.0f / (float)km;
    for (int j = 0; j <= jm; ++j) g.yu[j] = (float)j * dy;
    for (int k = 0; k <= km; ++k) g.zu[k] = (float)k * dz;
    g.y[0] = -dy / 2.0f;
    g.z[0] = -dz / 2.0f;
    for (int j = 1; j <= jm; ++j) g.y[j] = (g.yu[j - 1] + g.yu[j]) / 2.0f;
    for (int k = 1; k <= km; ++k) g.z[k] = (g.zu[k - 1] + g.zu[k]) / 2.0f;
    g.y[jm + 1] = g.y[jm] + dy;
    g.z[km + 1] = g.z[km] + dz;
    /* x: cell widths 0.5 - f (r - 1/2)^2 accumulated to the mid-plane,
     * scaled to Lx/2 and mirrored; coarse levels take every (im0/im)-th
     * fine face */
    std::vector<float> f(im0 + 1);
    f[0] = 0.f;
    const int half = im0 / 2;
    for (int i = 1; i <= half; ++i) {
        const double r = (double)((float)i / (float)im0) - 0.5;
        f[i] = f[i - 1] + (float)(0.5 - p.fact_mesh * r * r);
    }
    const double norm = (double)f[half];
    const double scale = (double)p.Lx / 2.;
    for (int i = 1; i <= half; ++i) f[i] = (float)(scale * (double)f[i] / norm);
    for (int i = im0; i > half; --i) f[i] = p.Lx - f[im0 - i];
    g.xu[0] = 0.f;
    for (int i = 1; i <= im; ++i) g.xu[i] = f[(size_t)i * im0 / im];
    for (int i = 1; i <= im; ++i) g.x[i] = (float)(0.5 * (double)(g.xu[i] + g.xu[i - 1]));
    g.x[0] = -g.x[1];
    g.x[im + 1] = g.xu[im] + (g.xu[im] - g.x[im]);
}

namespace {

/* whitespace-separated numeric tokens of a text file */
struct TokenFile {
    std::string buf;
    const char *cur = nullptr;
    bool open(const std::string &path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) return false;
        std::ostringstream ss;
        ss << in.rdbuf();
        buf = ss.str();
        cur = buf.c_str();
        return true;
    }
    bool next_float(float &v) {
        char *end = nullptr;
        errno = 0;
        v = std::strtof(cur, &end);
        if (end == cur) return false;
        cur = end;
        return true;
    }
    bool next_int(long &v) {
        char *end = nullptr;
        v = std::strtol(cur, &end, 10);
        if (end == cur) return false;
        cur = end;
        return true;
    }
};

}  // namespace

/* readT, narrow-band tables (read_tables.cpp:21-95) */
int read_tables(const std::string &dir, double chanHeight, Tables &t) {
    const float scale = (float)chanHeight;
    TokenFile f;
    const std::string pm = dir + "/planck-mean.txt";
    if (!f.open(pm)) { set_error("cannot open " + pm); return -2; }
    for (int i = 0; i < t.nT; ++i) {
        float d1, d2;
        if (!f.next_float(t.Tnb[i]) || !f.next_float(t.kP[i]) || !f.next_float(d1) || !f.next_float(d2)) {
            set_error("malformed " + pm); return -3;
        }
        t.kP[i] *= scale;
    }
    for (int b = 0; b < t.nB; ++b) {
        const std::string nbf = dir + "/NarrBand" + std::to_string(b) + ".txt";
        if (!f.open(nbf)) { set_error("cannot open " + nbf); return -2; }
        if (!f.next_float(t.wvl[b]) || !f.next_float(t.wvr[b]) || !f.next_float(t.wvc[b])) {
            set_error("malformed " + nbf); return -3;
        }
        for (int r = 0; r < t.nT; ++r) {
            float T, cp;
            /* each row: T, cumulative probability, kavg, kq[nQ]; kavg keeps
             * the last row's value (read_tables.cpp:48-52) */
            if (!f.next_float(T) || !f.next_float(cp) || !f.next_float(t.kavg[b])) { set_error("malformed " + nbf); return -3; }
            for (int g = 0; g < t.nQ; ++g) {
                float &k = t.kq[((size_t)b * t.nT + r) * t.nQ + g];
                if (!f.next_float(k)) { set_error("malformed " + nbf); return -3; }
                k *= scale;
            }
        }
    }
    const std::string pf = dir + "/prob_new2.txt";
    if (!f.open(pf)) { set_error("cannot open " + pf); return -2; }
    for (int r = 0; r < t.nT; ++r)
        for (int b = 0; b < t.nB; ++b) {
            float T;
            long idx;
            if (!f.next_float(T) || !f.next_int(idx) || !f.next_float(t.prob[(size_t)r * t.nB + b])) {
                set_error("malformed " + pf); return -3;
            }
            for (int g = 0; g < t.nQ; ++g)
                if (!f.next_float(t.probg[((size_t)r * t.nB + b) * t.nQ + g])) { set_error("malformed " + pf); return -3; }
        }
    for (int i = 1; i < t.nT; ++i)
        if (!(t.Tnb[i] > t.Tnb[i - 1])) { set_error("temperature table is not increasing"); return -4; }
    return 0;
}

/* writer in the nbcorrk/write_tables.cpp:18-109 formats.  kP and kq are
 * divided by chanHeight so that reading them back restores the values. */
int write_tables(const std::string &dir, const Tables &t, double chanHeight) {
    const double s = 1.0 / chanHeight;
    const std::string pm = dir + "/planck-mean.txt";
    FILE *fp = std::fopen(pm.c_str(), "w");
    if (!fp) { set_error("cannot write " + pm); return -2; }
    for (int i = 0; i < t.nT; ++i)
        std::fprintf(fp, "%.9g %.9e %.9e %.9e\n", (double)t.Tnb[i], (double)t.kP[i] * s, 0.0, 0.0);
    std::fclose(fp);
    const std::string pf = dir + "/prob_new2.txt";
    fp = std::fopen(pf.c_str(), "w");
    if (!fp) { set_error("cannot write " + pf); return -2; }
    for (int r = 0; r < t.nT; ++r)
        for (int b = 0; b < t.nB; ++b) {
            std::fprintf(fp, "%.9g \t %-5d \t%30.20e \t", (double)t.Tnb[r], b, (double)t.prob[(size_t)r * t.nB + b]);
            for (int g = 0; g < t.nQ; ++g)
                std::fprintf(fp, "%30.20e\t", (double)t.probg[((size_t)r * t.nB + b) * t.nQ + g]);
            std::fprintf(fp, "\n");
        }
    std::fclose(fp);
    for (int b = 0; b < t.nB; ++b) {
        const std::string nbf = dir + "/NarrBand" + std::to_string(b) + ".txt";
        fp = std::fopen(nbf.c_str(), "w");
        if (!fp) { set_error("cannot write " + nbf); return -2; }
        std::fprintf(fp, "%.9e\t%.9e\t%.9e \n", (double)t.wvl[b], (double)t.wvr[b], (double)t.wvc[b]);
        for (int r = 0; r < t.nT; ++r) {
            /* kavg is written on every row; readers keep the last one */
            std::fprintf(fp, " %.9g\t %.9e\t %.9e\t", (double)t.Tnb[r], (double)t.prob[(size_t)r * t.nB + b],
                         (double)t.kavg[b]);
            for (int g = 0; g < t.nQ; ++g)
                std::fprintf(fp, " %.9e\t", (double)t.kq[((size_t)b * t.nT + r) * t.nQ + g] * s);
            std::fprintf(fp, "\n");
        }
        std::fclose(fp);
    }
    return 0;
}

/* The generator's front end on a line-by-line spectrum read in file order
 * (wavenumber wv in cm^-1, absorption coefficient kv in cm^-1).
 *
 * Wide bands (absWideBand, nbcorrk/read_files.cpp:22-73): a point with
 * kv > toll opens a new wide band when the previous point had kv <= toll
 * and at least wvtoll of "weak" wavenumber (dwv per point with kv <= toll)
 * has passed since the last strong point; otherwise it joins the current
 * band; a weak point joins the current band while that weak wavenumber is
 * below wvtoll.  Every wide band is thus a contiguous run of points.
 * Narrow bands (ckdist_hitran_part.cpp:55-114): wide band i is split into
 * numbands = (int)((wv_last - wv_first) / deltawv[min(i / 3, 2)]) (at least
 * one) bands by point count: with comp = points / numbands, point j goes to
 * band h, and h moves on after every point j > 0 with j % comp == 0 until
 * the last band, which takes the rest (so the first band holds comp + 1
 * points).  A band's edges are its first and last wavenumber, wvc their
 * mean.  Outputs per narrow band: its first point and point count in the
 * input, wv_left, wv_right, wvc; returns the number of narrow bands (or
 * < 0 on error), nwide the number of wide bands. */
int lbl_bands(int64_t n, const double *wv, const double *kv, double toll, double wvtoll, double dwv,
              const double *deltawv, std::vector<int64_t> &first, std::vector<int64_t> &count,
              std::vector<double> &wvl, std::vector<double> &wvr, std::vector<double> &wvc, int &nwide) {
    std::vector<int64_t> ws, we;                 /* wide bands [ws, we) */
    double kvprev = 0.0, wvcu = wvtoll + 1.0;
    for (int64_t i = 0; i < n; ++i) {
        if (kv[i] > toll) {
            if (kvprev <= toll && wvcu >= wvtoll) { ws.push_back(i); we.push_back(i + 1); }
            else if (!ws.empty()) we.back() = i + 1;
            wvcu = 0.0;
        } else if (wvcu < wvtoll) {
            if (!ws.empty()) we.back() = i + 1;
            wvcu = wvcu + dwv;
        }
        kvprev = kv[i];
    }
    nwide = (int)ws.size();
    first.clear(); count.clear(); wvl.clear(); wvr.clear(); wvc.clear();
    for (size_t b = 0; b < ws.size(); ++b) {
        const int64_t a = ws[b], m = we[b] - ws[b];
        const double range = wv[we[b] - 1] - wv[a];
        int numbands = (int)(range / deltawv[std::min<size_t>(b / 3, 2)]);
        if (numbands == 0) numbands = 1;
        const int64_t comp = m / numbands;
        if (comp < 1) {
            set_error("wide band " + std::to_string(b) + " holds fewer points than narrow bands");
            return -3;
        }
        /* band h holds points [lo_h, hi_h): band 0 ends after point comp,
         * band h < numbands - 1 after point (h + 1) comp, the last at m */
        for (int64_t h = 0; h < numbands; ++h) {
            const int64_t lo = (h == 0) ? 0 : h * comp + 1;
            const int64_t hi = (h == numbands - 1) ? m : (h + 1) * comp + 1;
            if (hi <= lo) { set_error("narrow band " + std::to_string(h) + " of wide band " + std::to_string(b) +
                                      " is empty"); return -3; }
            first.push_back(a + lo);
            count.push_back(hi - lo);
            const double l = wv[a + lo], r = wv[a + hi - 1];
            wvl.push_back(l);
            wvr.push_back(r);
            wvc.push_back((l + r) / 2);
        }
    }
    return (int)first.size();
}

/* readFillT (nbcorrk/read_files.cpp:76-114) for one temperature T: the
 * spectrum (wv, kv, n points in file order) fills the narrow bands'
 * absorption coefficients kvn (times 100: m^-1) wherever its wavenumber
 * equals the next wavenumber awaited (wvn, the bands' points in order),
 * and accumulates the line-by-line Planck mean
 * kPr = sum pi kv 100 I_b(T, wv) dwv 100 / (sigma T^4) over the points read
 * until the last band is full (the reading stops there).  Returns the
 * number of band points filled. */
int64_t lbl_fill(int64_t n, const double *wv, const double *kv, double T, double dwv, int64_t ntot,
                 const double *wvn, double *kvn, double &kPr) {
    const double pi = kPi, stefan = kStefan;
    const double C1 = kC1, C2 = kC2;
    auto iblack = [&](double Tr, double nu) {     /* nbcorrk/ckdist_hitran_part.cpp:229-242 */
        return 1.0 / pi * C1 * std::pow((nu * 100), 3) / (std::exp((float)(C2 * nu * 100 / Tr)) - 1);
    };
    kPr = 0.0;
    int64_t q = 0;
    for (int64_t i = 0; i < n && q < ntot; ++i) {
        kPr = kPr + pi * kv[i] * 100 * iblack(T, wv[i]) * dwv * 100 / (stefan * std::pow(T, 4));
        if (wv[i] == wvn[q]) {
            kvn[q] = kv[i] * 100;
            ++q;
        }
    }
    return q;
}

/* calcProbT (nbcorrk/calc_prob.cpp:17-106) from the distributions of every
 * (temperature row i, band j): kq[i][j][nQ] and kavg[i][j] in FP64.  Per
 * row, the correlated-k Planck mean kP and the band CDF come from the band
 * means k_avg (the reference sums kk ff over the bins of each band inside
 * the same products; the sum over bins is k_avg), the g CDF of each band
 * from kq Ib dnu 100 wq cumulated and divided by its last value.  The
 * tables are filled as the reader leaves them: FP32, k values times
 * chanHeight, kavg from the last row.  kP also stands where the solver
 * reads the line-by-line Planck mean (planck-mean.txt column 2), so the
 * bound kappa_max and the band CDF share one normalisation. */
int ck_tables(int nB, int nT, int nQ, const double *T, const double *wvl, const double *wvr, const double *kq,
              const double *kavg, const double *wq, double chanHeight, Tables &t) {
    const double pi = kPi, stefan = kStefan;
    const double C1 = kC1, C2 = kC2;
    auto iblack = [&](double Tr, double nu) {     /* nbcorrk/ckdist_hitran_part.cpp:229-242 */
        return 1.0 / pi * C1 * std::pow((nu * 100), 3) / (std::exp((float)(C2 * nu * 100 / Tr)) - 1);
    };
    t.resize(nB, nT, nQ);
    for (int j = 0; j < nB; ++j) {
        if (!(wvr[j] > wvl[j])) { set_error("band " + std::to_string(j) + " has wv_right <= wv_left"); return -2; }
        t.wvl[j] = (float)wvl[j];
        t.wvr[j] = (float)wvr[j];
        t.wvc[j] = (float)((wvl[j] + wvr[j]) / 2);
        t.kavg[j] = (float)kavg[(size_t)(nT - 1) * nB + j];
    }
    const float sc = (float)chanHeight;
    std::vector<double> ib(nB);
    for (int i = 0; i < nT; ++i) {
        if (!(T[i] > 0.0) || (i > 0 && !(T[i] > T[i - 1]))) {
            set_error("temperatures must be positive and increasing");
            return -3;
        }
        const double T4 = stefan * std::pow(T[i], 4);
        double kP = 0.0;
        for (int j = 0; j < nB; ++j) {
            ib[j] = iblack(T[i], (wvl[j] + wvr[j]) / 2);
            kP += kavg[(size_t)i * nB + j] * ib[j] * (wvr[j] - wvl[j]) * 100 * pi / T4;
        }
        if (!(kP > 0.0)) { set_error("row " + std::to_string(i) + " has no absorption (kP = 0)"); return -4; }
        t.Tnb[i] = (float)T[i];
        t.kP[i] = (float)kP * sc;
        double cu = 0.0;
        for (int j = 0; j < nB; ++j) {
            cu += kavg[(size_t)i * nB + j] * ib[j] * (wvr[j] - wvl[j]) * 100 * pi / (kP * stefan * std::pow(T[i], 4));
            t.prob[(size_t)i * nB + j] = (float)cu;
            const double *k = kq + ((size_t)i * nB + j) * nQ;
            double cg[1024];
            for (int g = 0; g < nQ; ++g) {
                const double p2 = k[g] * ib[j] * (wvr[j] - wvl[j]) * 100 * wq[g];
                cg[g] = g == 0 ? p2 : cg[g - 1] + p2;
            }
            for (int g = 0; g < nQ; ++g) {
                t.probg[((size_t)i * nB + j) * nQ + g] = (float)(cg[g] / cg[nQ - 1]);
                t.kq[((size_t)j * nT + i) * nQ + g] = (float)k[g] * sc;
            }
        }
    }
    t.wq.resize(nQ);
    for (int g = 0; g < nQ; ++g) t.wq[g] = (float)wq[g];
    return 0;
}

/* readT, grey tables (read_tables.cpp:98-175) */
int grey_tables(const rc_params &p, float Tmax, Tables &t) {
    if (t.nQ != 1) { set_error("grey tables have nQ=1"); return -1; }
    const double lo = (double)p.Tmins * 0.9, hi = (double)Tmax * 1.1;
    for (int i = 0; i < t.nT; ++i) {
        t.kP[i] = p.abscoeff;
        t.Tnb[i] = (float)(lo + (hi - lo) * (double)(float)i / (double)(float)t.nT);
    }
    t.wq.assign(1, 1.0f);
    if (p.grey_true) {
        if (t.nB != 1) { set_error("true grey has one band"); return -1; }
        t.wvl[0] = 0.f; t.wvr[0] = p.wvmax; t.wvc[0] = p.wvmax / 2.f; t.kavg[0] = p.abscoeff;
        for (int r = 0; r < t.nT; ++r) { t.kq[r] = p.abscoeff; t.prob[r] = 1.0f; t.probg[r] = 1.0f; }
        return 0;
    }
    const int nB = t.nB;
    const float dw = p.wvmax / (float)nB;
    float l = 0.f, r = dw;
    for (int b = 0; b < nB; ++b) {
        if (b > 0) { l += dw; r += dw; }
        t.wvl[b] = l; t.wvr[b] = r;
        t.wvc[b] = (float)((double)(l + r) / 2.);
        t.kavg[b] = p.abscoeff;
        for (int q = 0; q < t.nT; ++q) t.kq[(size_t)b * t.nT + q] = p.abscoeff;
    }
    /* band CDF of the Planck function on each temperature row */
    for (int q = 0; q < t.nT; ++q) {
        float *row = &t.prob[(size_t)q * nB];
        float acc = 0.f;
        for (int b = 0; b < nB; ++b) {
            const float w = planck(t.Tnb[q], t.wvc[b]) * (t.wvr[b] - t.wvl[b]) * 100.f;
            acc = (b == 0) ? w : acc + w;
            row[b] = acc;
        }
        for (int b = 0; b < nB; ++b) { row[b] = row[b] / acc; t.probg[(size_t)q * nB + b] = 1.0f; }
    }
    return 0;
}

/* idx[rank] = band with the rank-th smallest kavg; equal keys keep band order */
void sort_idx(const float *kavg, int nB, int32_t *idx) {
    std::vector<std::pair<float, int32_t>> v((size_t)nB);
    for (int b = 0; b < nB; ++b) v[b] = {kavg[b], b};
    std::sort(v.begin(), v.end(), [](const std::pair<float, int32_t> &a, const std::pair<float, int32_t> &b) {
        if (a.first < b.first) return true;
        if (b.first < a.first) return false;
        return a.second < b.second;
    });
    for (int r = 0; r < nB; ++r) idx[r] = v[r].second;
}

}  // namespace rc
