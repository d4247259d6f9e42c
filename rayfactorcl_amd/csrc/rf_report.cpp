// rf_report.cpp -- surface areas of the primitives and the closure /
// reciprocity checks of a view-factor matrix (the report beside the
// output.txt writer).  Host code only.

#include <cmath>
#include <vector>

#include "rayfactor.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// Gauss-Legendre nodes and weights on [-1, 1] (Newton iteration on P_n).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const double dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

struct Lin {
    double a[3][3];   // linear part of the object -> world map (rows)
    void apply(const double v[3], double o[3]) const
    {
        for (int r = 0; r < 3; ++r) o[r] = a[r][0] * v[0] + a[r][1] * v[1] + a[r][2] * v[2];
    }
};

double cross_norm(const double u[3], const double v[3])
{
    const double c0 = u[1] * v[2] - u[2] * v[1];
    const double c1 = u[2] * v[0] - u[0] * v[2];
    const double c2 = u[0] * v[1] - u[1] * v[0];
    return std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
}

// Area element |M du x M dv| of the parametrisation at (s, t).
template <class Tangents>
double integrate(const Lin &M, double s0, double s1, bool s_periodic, double t0, double t1, Tangents tang)
{
    constexpr int NS = 256;   // periodic direction: trapezoid rule, spectrally accurate
    constexpr int NT = 24;
    static std::vector<double> gx, gw;
    if (gx.empty()) gauss_legendre(NT, gx, gw);
    double sum = 0.0;
    const int ns = s_periodic ? NS : NT;
    for (int i = 0; i < ns; ++i) {
        double s, ws;
        if (s_periodic) {
            s = s0 + (s1 - s0) * i / NS;
            ws = (s1 - s0) / NS;
        } else {
            s = 0.5 * (s0 + s1) + 0.5 * (s1 - s0) * gx[i];
            ws = 0.5 * (s1 - s0) * gw[i];
        }
        for (int j = 0; j < NT; ++j) {
            const double t = 0.5 * (t0 + t1) + 0.5 * (t1 - t0) * gx[j];
            const double wt = 0.5 * (t1 - t0) * gw[j];
            double du[3], dv[3], Mu[3], Mv[3];
            tang(s, t, du, dv);
            M.apply(du, Mu);
            M.apply(dv, Mv);
            sum += ws * wt * cross_norm(Mu, Mv);
        }
    }
    return sum;
}

}  // namespace

extern "C" int32_t rf_areas(const rf_primitive *prims, int32_t n, double *area)
{
    if (n < 0 || (n > 0 && (!prims || !area))) return RF_ERR_INVALID_ARG;
    for (int32_t i = 0; i < n; ++i) {
        const rf_primitive &p = prims[i];
        Lin M;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) M.a[r][c] = p.m[4 * r + c];
        double A = 0.0;
        switch (p.type) {
        case RF_CYLINDER:   // (cos s, sin s, t), t in [0, 1]
            A = integrate(M, 0.0, 2 * kPi, true, 0.0, 1.0, [](double s, double, double *du, double *dv) {
                du[0] = -std::sin(s); du[1] = std::cos(s); du[2] = 0.0;
                dv[0] = 0.0; dv[1] = 0.0; dv[2] = 1.0;
            });
            break;
        case RF_FRUSTUM: {  // (r cos s, r sin s, t), r = 1 + k t
            const double k = (double)p.util1 - 1.0;
            A = integrate(M, 0.0, 2 * kPi, true, 0.0, 1.0, [k](double s, double t, double *du, double *dv) {
                const double r = 1.0 + k * t;
                du[0] = -r * std::sin(s); du[1] = r * std::cos(s); du[2] = 0.0;
                dv[0] = k * std::cos(s); dv[1] = k * std::sin(s); dv[2] = 1.0;
            });
            break;
        }
        case RF_SPHERE:     // (sin t cos s, sin t sin s, cos t), t in [0, pi]
            A = integrate(M, 0.0, 2 * kPi, true, 0.0, kPi, [](double s, double t, double *du, double *dv) {
                du[0] = -std::sin(t) * std::sin(s); du[1] = std::sin(t) * std::cos(s); du[2] = 0.0;
                dv[0] = std::cos(t) * std::cos(s); dv[1] = std::cos(t) * std::sin(s); dv[2] = -std::sin(t);
            });
            break;
        case RF_DISC:
        case RF_ANNULUS: {  // (r cos s, r sin s, 0), r in [r_in, r_out]
            double r0 = 0.0, r1 = 1.0;
            if (p.type == RF_ANNULUS) {
                r0 = std::sqrt(std::fmax(0.0, (double)p.util1));
                r1 = std::sqrt(std::fmax(0.0, (double)p.util2));
            }
            A = integrate(M, 0.0, 2 * kPi, true, r0, r1, [](double s, double r, double *du, double *dv) {
                du[0] = -r * std::sin(s); du[1] = r * std::cos(s); du[2] = 0.0;
                dv[0] = std::cos(s); dv[1] = std::sin(s); dv[2] = 0.0;
            });
            break;
        }
        case RF_RECTANGLE: { // [-1, 1]^2 in z = 0: the map is linear, the element constant
            const double u[3] = {1.0, 0.0, 0.0}, v[3] = {0.0, 1.0, 0.0};
            double Mu[3], Mv[3];
            M.apply(u, Mu);
            M.apply(v, Mv);
            A = 4.0 * cross_norm(Mu, Mv);
            break;
        }
        default:
            return RF_ERR_INVALID_SCENE;
        }
        area[i] = A;
    }
    return RF_OK;
}

extern "C" int32_t rf_vf_check(int32_t n, const double *vf, const double *area, double min_product, double *closure,
                               rf_vf_report *report)
{
    if (n <= 0 || !vf || !area || !report) return RF_ERR_INVALID_ARG;
    rf_vf_report r{};
    r.min_closure = INFINITY;
    r.max_closure = -INFINITY;
    r.worst_i = r.worst_j = -1;
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t j = 0; j < n; ++j) s += vf[(size_t)i * n + j];
        if (closure) closure[i] = s;
        r.min_closure = std::fmin(r.min_closure, s);
        r.max_closure = std::fmax(r.max_closure, s);
    }
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = i + 1; j < n; ++j) {
            const double a = area[i] * vf[(size_t)i * n + j];
            const double b = area[j] * vf[(size_t)j * n + i];
            if (a < min_product || b < min_product) continue;
            ++r.pairs_compared;
            const double d = std::fabs(a - b);
            const double rel = d / std::fmax(a, b);
            r.max_recip_abs = std::fmax(r.max_recip_abs, d);
            if (rel > r.max_recip_rel || r.worst_i < 0) {
                r.max_recip_rel = rel;
                r.worst_i = i;
                r.worst_j = j;
            }
        }
    *report = r;
    return RF_OK;
}
