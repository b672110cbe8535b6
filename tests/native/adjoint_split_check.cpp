// Host check of the quad adjoint in split form (forward mode in the in-plane
// coordinates and properties, reverse mode for the frame, chain rule through
// the frame's Jacobian: quad_phi, quad_frame_adjoint, quad_frame in
// element_math.hpp) against forward mode through the whole element once per
// coordinate (dual_per_coordinate, built from the same element routines).  Built and run
// on the CPU by tests/test_adjoint_split.py.
#include "element_math.hpp"
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
using namespace jsso;

static double dual_per_coordinate(const double* crds, const int nd[4], const double pr[5], const double* u, const double* lam,
                      double ls, int nrhs, int rhs, int p) {
    Dual c[12];
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) c[3 * a + k] = Dual(crds[3 * nd[a] + k]);
    Dual t(pr[0]), E(pr[1]), nu(pr[2]), kx(pr[3]), ky(pr[4]);
    if (p < 12) c[p].d = 1.0;
    else if (p == 12) t.d = 1.0;
    else if (p == 13) E.d = 1.0;
    else if (p == 14) nu.d = 1.0;
    else if (p == 15) kx.d = 1.0;
    else ky.d = 1.0;
    V3<Dual> R[3];
    Dual xy[8];
    quad_frame<Dual>(c, R, xy);
    Dual cb, cs, Cm[3][3];
    quad_plate_consts<Dual>(t, E, nu, cb, cs);
    quad_cm<Dual>(E, nu, kx, ky, Cm);
    Dual diag[8];
    quad_rot_diag<Dual, Dual>(xy, cb, cs, nu, diag);
    const Dual krz = quad_drill<Dual>(diag);
    Dual ul[24], ll[24];
    for (int a = 0; a < 4; ++a)
        for (int h = 0; h < 2; ++h) {
            double uv[3], lv[3];
            for (int k = 0; k < 3; ++k) {
                const int64_t gi = ((int64_t)6 * nd[a] + 3 * h + k) * nrhs + rhs;
                uv[k] = u[gi];
                lv[k] = lam ? lam[gi] : ls * u[gi];
            }
            for (int i = 0; i < 3; ++i) {
                ul[6 * a + 3 * h + i] = R[i].x * uv[0] + R[i].y * uv[1] + R[i].z * uv[2];
                ll[6 * a + 3 * h + i] = R[i].x * lv[0] + R[i].y * lv[1] + R[i].z * lv[2];
            }
        }
    Dual em(0.0), eb(0.0), es(0.0), ed(0.0);
    for (int g = 0; g < 4; ++g) {
        QuadGP<Dual> G;
        quad_gp<Dual>(xy, g, G);
        Dual mu[3], ml[3], ku[3], kl[3], gu[2], gl[2];
        for (int a = 0; a < 4; ++a) {
            const Dual h0 = G.dH[0][a], h1 = G.dH[1][a];
            mu[0] = mu[0] + h0 * ul[6 * a];
            mu[1] = mu[1] + h1 * ul[6 * a + 1];
            mu[2] = mu[2] + h1 * ul[6 * a] + h0 * ul[6 * a + 1];
            ml[0] = ml[0] + h0 * ll[6 * a];
            ml[1] = ml[1] + h1 * ll[6 * a + 1];
            ml[2] = ml[2] + h1 * ll[6 * a] + h0 * ll[6 * a + 1];
            ku[0] = ku[0] - h0 * ul[6 * a + 4];
            ku[1] = ku[1] + h1 * ul[6 * a + 3];
            ku[2] = ku[2] + h0 * ul[6 * a + 3] - h1 * ul[6 * a + 4];
            kl[0] = kl[0] - h0 * ll[6 * a + 4];
            kl[1] = kl[1] + h1 * ll[6 * a + 3];
            kl[2] = kl[2] + h0 * ll[6 * a + 3] - h1 * ll[6 * a + 4];
            for (int cc = 0; cc < 3; ++cc) {
                Dual b0, b1;
                quad_bgamma_col<Dual>(xy, G, 3 * a + cc, b0, b1);
                gu[0] = gu[0] + b0 * ul[6 * a + 2 + cc];
                gu[1] = gu[1] + b1 * ul[6 * a + 2 + cc];
                gl[0] = gl[0] + b0 * ll[6 * a + 2 + cc];
                gl[1] = gl[1] + b1 * ll[6 * a + 2 + cc];
            }
        }
        Dual mform(0.0);
        for (int r = 0; r < 3; ++r) mform = mform + ml[r] * (Cm[r][0] * mu[0] + Cm[r][1] * mu[1] + Cm[r][2] * mu[2]);
        em = em + mform * G.detJ;
        eb = eb + cb_form(cb, nu, kl[0], kl[1], kl[2], ku[0], ku[1], ku[2]) * G.detJ;
        es = es + cs * (gl[0] * gu[0] + gl[1] * gu[1]) * G.detJ;
    }
    for (int a = 0; a < 4; ++a) ed = ed + ll[6 * a + 5] * ul[6 * a + 5];
    const Dual phi = t * em + eb + es + krz * ed;
    return -phi.d;
}

static const int XYF[6] = {0, 1, 2, 3, 6, 7};

// plant: 0 none, 1 flips the sign of dphi/dR_12, 2 scales dphi/dxy_3 by 1 + 1e-6
// (the check must fail for 1 and 2)
static int plant = 0;

static void split_form(const double* crds, const int nd[4], const double pr[5], const double* u, const double* lam,
                       double ls, int nrhs, int rhs, double out[17]) {
    double c[12];
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) c[3 * a + k] = crds[3 * nd[a] + k];
    V3<double> R[3];
    double xy[8];
    quad_frame<double>(c, R, xy);
    double dxy[6], dR[3][3];
    for (int k = 0; k < 6; ++k) {
        Dual xd[8];
        for (int i = 0; i < 8; ++i) xd[i] = Dual(xy[i], i == XYF[k] ? 1.0 : 0.0);
        dxy[k] = quad_phi<Dual, double>(R, xd, pr[0], pr[1], pr[2], pr[3], pr[4], nd, u, lam, ls, nrhs, rhs).d;
    }
    for (int k = 0; k < 5; ++k) {
        Dual m[5];
        for (int i = 0; i < 5; ++i) m[i] = Dual(pr[i], i == k ? 1.0 : 0.0);
        out[12 + k] = -quad_phi<double, Dual>(R, xy, m[0], m[1], m[2], m[3], m[4], nd, u, lam, ls, nrhs, rhs).d;
    }
    quad_frame_adjoint(R, xy, pr[0], pr[1], pr[2], pr[3], pr[4], nd, u, lam, ls, nrhs, rhs, dR);
    if (plant == 1) dR[1][2] = -dR[1][2];
    if (plant == 2) dxy[3] *= 1.0 + 1e-6;
    for (int p = 0; p < 12; ++p) {
        Dual cd[12];
        for (int i = 0; i < 12; ++i) cd[i] = Dual(c[i], i == p ? 1.0 : 0.0);
        V3<Dual> Rd[3];
        Dual xyd[8];
        quad_frame<Dual>(cd, Rd, xyd);
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += dxy[k] * xyd[XYF[k]].d;
        for (int i = 0; i < 3; ++i) s += dR[i][0] * Rd[i].x.d + dR[i][1] * Rd[i].y.d + dR[i][2] * Rd[i].z.d;
        out[p] = -s;
    }
}

int main(int argc, char** argv) {
    if (argc > 1) plant = atoi(argv[1]);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double worst = 0.0;
    for (int trial = 0; trial < 200; ++trial) {
        const int nrhs = 1 + trial % 3;
        double crds[12];
        const double base[4][2] = {{1, 1}, {0, 1}, {0, 0}, {1, 0}};
        const double sc = 0.2 + 0.8 * (trial % 5) / 4.0;
        for (int a = 0; a < 4; ++a) {
            crds[3 * a] = sc * (base[a][0] + 0.15 * U(rng));
            crds[3 * a + 1] = sc * (base[a][1] + 0.15 * U(rng));
            crds[3 * a + 2] = sc * 0.2 * U(rng) + (trial % 2 ? 0.0 : 0.3 * crds[3 * a]);
        }
        if (trial % 7 == 0)  // a perfect square: many ties in the drilling minimum
            for (int a = 0; a < 4; ++a) { crds[3 * a] = base[a][0]; crds[3 * a + 1] = base[a][1]; crds[3 * a + 2] = 0.0; }
        const int nd[4] = {0, 1, 2, 3};
        const double pr[5] = {0.05 + 0.1 * (U(rng) + 1), 3e10 * (1 + 0.2 * U(rng)), 0.2 + 0.1 * U(rng), 1.0, 1.0 + 0.3 * (trial % 3 == 0)};
        std::vector<double> u(24 * nrhs), lam(24 * nrhs);
        for (auto& v : u) v = 1e-3 * U(rng);
        for (auto& v : lam) v = 1e-3 * U(rng);
        for (int rhs = 0; rhs < nrhs; ++rhs)
            for (int mode = 0; mode < 2; ++mode) {
                const double* lp = mode ? lam.data() : nullptr;
                double o[17], nw[17], scale = 0.0;
                for (int p = 0; p < 17; ++p) o[p] = dual_per_coordinate(crds, nd, pr, u.data(), lp, 0.5, nrhs, rhs, p);
                split_form(crds, nd, pr, u.data(), lp, 0.5, nrhs, rhs, nw);
                for (int p = 0; p < 12; ++p) scale = std::max(scale, std::fabs(o[p]));
                for (int p = 0; p < 17; ++p) {
                    const double ref = p < 12 ? scale : std::fabs(o[p]) + 1e-300;
                    const double e = std::fabs(o[p] - nw[p]) / ref;
                    if (e > worst) worst = e;
                    if (e > 1e-10 && plant == 0)
                        printf("trial %d rhs %d mode %d p %d per-coordinate %.16e split %.16e rel %.2e\n", trial, rhs,
                               mode, p, o[p], nw[p], e);
                }
            }
    }
    printf("worst relative difference %.3e\n", worst);
    return worst < 1e-10 ? 0 : 1;
}
