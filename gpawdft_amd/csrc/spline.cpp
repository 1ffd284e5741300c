// Radial cubic splines and their tabulation on the grid (host, FP64).
//
// Set-up code, run once per geometry like the reference's: the spline
// construction of c/bmgs/spline.c:11-58 (natural cubic spline with zero
// slope at r = 0, four polynomial coefficients per bin) and the
// spline_to_grid tabulation of c/lfc.c:785-875 with the real solid
// harmonics r^l Y_lm of c/bmgs/spherical_harmonics.h.  Built with
// -ffp-contract=off and no FMA, so values match the reference bit for bit.
#include <cmath>
#include <new>
#include <vector>

#include "../../include/gpawdft_amd.h"

namespace gpd {
void set_last_error(const char* msg);
}

struct gpd_spline {
    int l = 0;
    int nbins = 0;
    double dr = 0.0;
    std::vector<double> data;   // 4 * (nbins + 1)
};

static double spline_eval(const gpd_spline* s, double r)
{
    int b = (int)(r / s->dr);
    if (b >= s->nbins) return 0.0;
    double u = r - b * s->dr;
    const double* c = s->data.data() + 4 * b;
    return c[0] + u * (c[1] + u * (c[2] + u * c[3]));
}

// f * r^l Y_lm for l <= 4, m in the reference's order.  The prefactors are
// the real-harmonic normalisations sqrt((2l+1)/4pi) x const.
static void solid_harmonics(int l, double f, double x, double y, double z, double r2, double* p)
{
    switch (l) {
    case 0:
        p[0] = f * 0.28209479177387814;
        break;
    case 1:
        p[0] = f * 0.48860251190291992 * y;
        p[1] = f * 0.48860251190291992 * z;
        p[2] = f * 0.48860251190291992 * x;
        break;
    case 2:
        p[0] = f * 1.0925484305920792 * x * y;
        p[1] = f * 1.0925484305920792 * y * z;
        p[2] = f * 0.31539156525252005 * (-r2 + 3 * z * z);
        p[3] = f * 1.0925484305920792 * x * z;
        p[4] = f * 0.54627421529603959 * (-y * y + x * x);
        break;
    case 3:
        p[0] = f * 0.59004358992664352 * (-y * y * y + 3 * x * x * y);
        p[1] = f * 2.8906114426405538 * x * y * z;
        p[2] = f * 0.45704579946446577 * (5 * y * z * z - y * r2);
        p[3] = f * 0.3731763325901154 * (-3 * z * r2 + 5 * z * z * z);
        p[4] = f * 0.45704579946446577 * (-x * r2 + 5 * x * z * z);
        p[5] = f * 1.4453057213202769 * (-y * y * z + x * x * z);
        p[6] = f * 0.59004358992664352 * (x * x * x - 3 * x * y * y);
        break;
    default:
        p[0] = f * 2.5033429417967046 * (x * x * x * y - x * y * y * y);
        p[1] = f * 1.7701307697799307 * (3 * x * x * y * z - y * y * y * z);
        p[2] = f * 0.94617469575756008 * (-x * y * r2 + 7 * x * y * z * z);
        p[3] = f * 0.66904654355728921 * (-3 * y * z * r2 + 7 * y * z * z * z);
        p[4] = f * 0.10578554691520431 * (3 * r2 * r2 - 30 * z * z * r2 + 35 * z * z * z * z);
        p[5] = f * 0.66904654355728921 * (7 * x * z * z * z - 3 * x * z * r2);
        p[6] = f * 0.47308734787878004 *
               (y * y * r2 + 7 * x * x * z * z - x * x * r2 - 7 * y * y * z * z);
        p[7] = f * 1.7701307697799307 * (x * x * x * z - 3 * x * y * y * z);
        p[8] = f * 0.62583573544917614 * (-6 * x * x * y * y + x * x * x * x + y * y * y * y);
        break;
    }
}

extern "C" int gpd_spline_create(int l, double rcut, const double* f, int npts, gpd_spline** out)
{
    if (!out || !f || npts < 2 || l < 0 || l > 4 || !(rcut > 0.0)) {
        gpd::set_last_error("bad spline arguments (need npts >= 2, 0 <= l <= 4, rcut > 0)");
        return GPD_ERR_ARG;
    }
    gpd_spline* s = new (std::nothrow) gpd_spline();
    if (!s) return GPD_ERR_NOMEM;
    const int nb = npts - 1;
    const double dr = rcut / nb;
    s->l = l;
    s->nbins = nb;
    s->dr = dr;
    // second derivatives by the tridiagonal recursion; f2[0] = -0.5 and
    // u[0] encode a zero first derivative at r = 0
    const double c = 3.0 / (dr * dr);
    std::vector<double> f2(nb + 1), u(nb);
    f2[0] = -0.5;
    u[0] = (f[1] - f[0]) * c;
    for (int b = 1; b < nb; b++) {
        double p = 0.5 * f2[b - 1] + 2.0;
        f2[b] = -0.5 / p;
        u[b] = ((f[b + 1] - 2.0 * f[b] + f[b - 1]) * c - 0.5 * u[b - 1]) / p;
    }
    f2[nb] = ((f[nb - 1] * c - 0.5 * u[nb - 1]) / (0.5 * f2[nb - 1] + 1.0));
    for (int b = nb - 1; b >= 0; b--) f2[b] = f2[b] * f2[b + 1] + u[b];
    s->data.assign(4 * (nb + 1), 0.0);
    for (int b = 0; b < nb; b++) {
        double* d = s->data.data() + 4 * b;
        d[0] = f[b];
        d[1] = (f[b + 1] - f[b]) / dr - (f2[b] / 3 + f2[b + 1] / 6) * dr;
        d[2] = 0.5 * f2[b];
        d[3] = (f2[b + 1] - f2[b]) / (6 * dr);
    }
    *out = s;
    return GPD_OK;
}

extern "C" double gpd_spline_value(const gpd_spline* s, double r) { return s ? spline_eval(s, r) : 0.0; }

extern "C" double gpd_spline_get_cutoff(const gpd_spline* s) { return s ? s->dr * s->nbins : 0.0; }

extern "C" int gpd_spline_get_angular_momentum_number(const gpd_spline* s) { return s ? s->l : -1; }

extern "C" int gpd_spline_destroy(gpd_spline* s)
{
    delete s;
    return GPD_OK;
}

extern "C" int gpd_spline_to_grid(const gpd_spline* s, const long beg[3], const long end[3],
                                  const double pos_v[3], const double h_cv[9], const long n[3],
                                  const long gdcorner[3], double* A_gm, int* G_B, long* ng,
                                  long* nB_out)
{
    if (!s || !beg || !end || !pos_v || !h_cv || !n || !gdcorner || !A_gm || !G_B || !ng ||
        !nB_out) {
        gpd::set_last_error("null argument");
        return GPD_ERR_ARG;
    }
    const int nm = 2 * s->l + 1;
    const double rcut = s->dr * s->nbins;
    long ngm = 0, nB = 0;
    long G = -gdcorner[2] + n[2] * (beg[1] - gdcorner[1] + n[1] * (beg[0] - gdcorner[0]));
    for (long g0 = beg[0]; g0 < end[0]; g0++) {
        for (long g1 = beg[1]; g1 < end[1]; g1++) {
            long first = -1, last = -1;
            for (long g2 = beg[2]; g2 < end[2]; g2++) {
                double x = h_cv[0] * g0 + h_cv[3] * g1 + h_cv[6] * g2 - pos_v[0];
                double y = h_cv[1] * g0 + h_cv[4] * g1 + h_cv[7] * g2 - pos_v[1];
                double z = h_cv[2] * g0 + h_cv[5] * g1 + h_cv[8] * g2 - pos_v[2];
                double r2 = x * x + y * y + z * z;
                double r = std::sqrt(r2);
                if (r < rcut) {
                    if (first < 0) first = g2;
                    last = g2;
                    solid_harmonics(s->l, spline_eval(s, r), x, y, z, r2, A_gm + ngm);
                    ngm += nm;
                }
            }
            if (last >= 0) {
                G_B[nB++] = (int)(G + first);
                G_B[nB++] = (int)(G + last + 1);
            }
            G += n[2];
        }
        G += n[2] * (n[1] - end[1] + beg[1]);
    }
    *ng = ngm / nm;
    *nB_out = nB;
    return GPD_OK;
}
