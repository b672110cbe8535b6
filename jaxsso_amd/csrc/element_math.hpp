// Element routines of the beam-column and the MITC-4 shell, templated on the
// scalar type: `double` for stiffness matrices, `Dual` (value + one tangent)
// for the forward-mode adjoint.  Each function cites the reference routine of
// JaxSSO/element.py it restates.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define JDEV __host__ __device__ __forceinline__

namespace jsso {

// ---------------------------------------------------------------------------
// dual numbers (one tangent direction)
// ---------------------------------------------------------------------------
struct Dual {
    double v, d;
    JDEV Dual() : v(0.0), d(0.0) {}
    JDEV Dual(double a) : v(a), d(0.0) {}
    JDEV Dual(double a, double b) : v(a), d(b) {}
};
JDEV Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
JDEV Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
JDEV Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
JDEV Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
JDEV Dual operator/(Dual a, Dual b) {
    double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
JDEV Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
JDEV Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
JDEV Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
JDEV Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
JDEV Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
JDEV Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
JDEV Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
JDEV Dual operator/(double a, Dual b) {
    double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}
JDEV Dual sqrt(Dual a) {
    double s = ::sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}

JDEV double val(double a) { return a; }
JDEV double val(Dual a) { return a.v; }
JDEV double tan_(double) { return 0.0; }
JDEV double tan_(Dual a) { return a.d; }
JDEV double dsqrt(double a) { return ::sqrt(a); }
JDEV void set_tan(double&, double) {}
JDEV void set_tan(Dual& a, double d) { a.d = d; }
JDEV Dual dsqrt(Dual a) { return sqrt(a); }

template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T>
JDEV V3<T> sub(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
JDEV V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T>
JDEV T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
JDEV T norm(const V3<T>& a) { return dsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
template <typename T>
JDEV V3<T> scale(const V3<T>& a, T s) { return {a.x / s, a.y / s, a.z / s}; }

// Relative tolerance under which drilling-spring candidates count as tied (the
// derivative then averages their tangents, JAX's jnp.min rule on exact values).
constexpr double TIE_RTOL = 1e-10;
constexpr double GP = 0.57735026918962576451;  // 1/sqrt(3)
__host__ __device__ constexpr double gauss_r(int g) { return (g == 0 || g == 3) ? GP : -GP; }
__host__ __device__ constexpr double gauss_s(int g) { return (g < 2) ? GP : -GP; }

// ---------------------------------------------------------------------------
// Beam-column: BeamCol.T / K_local (element.py:66-128)
// ---------------------------------------------------------------------------

// Direction-cosine rows of BeamCol.T (element.py:66-105, sin_alpha = 0,
// cos_alpha = 1).  Returns true for the Cxz == 0 branch (member parallel to
// global Y), whose matrix [[0,Cy,0],[-Cy,0,-1],[-Cy,0,0]] is not orthogonal.
template <typename T>
JDEV bool beam_dircos(const T c[6], T D[3][3], T& L) {
    T dx = c[3] - c[0], dy = c[4] - c[1], dz = c[5] - c[2];
    L = dsqrt((c[0] - c[3]) * (c[0] - c[3]) + (c[1] - c[4]) * (c[1] - c[4]) +
              (c[2] - c[5]) * (c[2] - c[5]));
    T Cx = dx / L, Cy = dy / L, Cz = dz / L;
    T Cxz = dsqrt(Cx * Cx + Cz * Cz);
    if (val(Cxz) == 0.0) {
        D[0][0] = T(0.0); D[0][1] = Cy;        D[0][2] = T(0.0);
        D[1][0] = -Cy;    D[1][1] = T(0.0);    D[1][2] = T(-1.0);
        D[2][0] = -Cy;    D[2][1] = T(0.0);    D[2][2] = T(0.0);
        return true;
    }
    D[0][0] = Cx;              D[0][1] = Cy;  D[0][2] = Cz;
    D[1][0] = Cz / Cxz;        D[1][1] = T(0.0); D[1][2] = -Cx / Cxz;
    D[2][0] = -Cx * Cy / Cxz;  D[2][1] = Cxz; D[2][2] = -Cy * Cz / Cxz;
    return false;
}

// Entry (I,J) of the 12x12 local Euler-Bernoulli stiffness (element.py:107-128).
template <typename T>
JDEV T beam_kloc(int I, int J, T L, T E, T G, T Iy, T Iz, T Jt, T A) {
    if (I > J) { int t = I; I = J; J = t; }
    switch (I * 12 + J) {
        case 0 * 12 + 0: case 6 * 12 + 6: return A * E / L;
        case 0 * 12 + 6: return -(A * E / L);
        case 1 * 12 + 1: case 7 * 12 + 7: return 12.0 * E * Iz / (L * L * L);
        case 1 * 12 + 7: return -(12.0 * E * Iz / (L * L * L));
        case 1 * 12 + 5: case 1 * 12 + 11: return 6.0 * E * Iz / (L * L);
        case 5 * 12 + 7: case 7 * 12 + 11: return -(6.0 * E * Iz / (L * L));
        case 5 * 12 + 5: case 11 * 12 + 11: return 4.0 * E * Iz / L;
        case 5 * 12 + 11: return 2.0 * E * Iz / L;
        case 2 * 12 + 2: case 8 * 12 + 8: return 12.0 * E * Iy / (L * L * L);
        case 2 * 12 + 8: return -(12.0 * E * Iy / (L * L * L));
        case 2 * 12 + 4: case 2 * 12 + 10: return -(6.0 * E * Iy / (L * L));
        case 4 * 12 + 8: case 8 * 12 + 10: return 6.0 * E * Iy / (L * L);
        case 4 * 12 + 4: case 10 * 12 + 10: return 4.0 * E * Iy / L;
        case 4 * 12 + 10: return 2.0 * E * Iy / L;
        case 3 * 12 + 3: case 9 * 12 + 9: return G * Jt / L;
        case 3 * 12 + 9: return -(G * Jt / L);
        default: return T(0.0);
    }
}

JDEV void inv3(const double D[3][3], double Di[3][3]) {
    double c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
    double c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
    double c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
    double det = D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02;
    double id = 1.0 / det;
    Di[0][0] = c00 * id;
    Di[1][0] = c01 * id;
    Di[2][0] = c02 * id;
    Di[0][1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * id;
    Di[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
    Di[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
    Di[0][2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * id;
    Di[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
    Di[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
}

// ---------------------------------------------------------------------------
// MITC-4 quad (element.py:471-1106)
// ---------------------------------------------------------------------------

// Local frame and in-plane coordinates of Quad.loc_crds / Quad.T
// (element.py:488-539, 643-696): x along node3->node1, z = x cross (node4->node2),
// y = z cross x; node 3 is the origin.
template <typename T>
JDEV void quad_frame(const T c[12], V3<T> R[3], T xy[8]) {
    V3<T> X0{c[0], c[1], c[2]}, X1{c[3], c[4], c[5]}, X2{c[6], c[7], c[8]},
        X3{c[9], c[10], c[11]};
    V3<T> v31 = sub(X0, X2), v32 = sub(X1, X2), v34 = sub(X3, X2), v42 = sub(X1, X3);
    V3<T> x = v31;
    V3<T> z = cross(x, v42);
    V3<T> y = cross(z, x);
    R[0] = scale(x, norm(x));
    R[1] = scale(y, norm(y));
    R[2] = scale(z, norm(z));
    xy[0] = dot(v31, R[0]);
    xy[1] = dot(v31, R[1]);
    xy[2] = dot(v32, R[0]);
    xy[3] = dot(v32, R[1]);
    xy[4] = T(0.0);
    xy[5] = T(0.0);
    xy[6] = dot(v34, R[0]);
    xy[7] = dot(v34, R[1]);
}

// Per-Gauss-point quantities: det J, shape-function derivatives dH = J^-1 dN
// (element.py:698-733, 803-818) and the MITC-4 assumed-shear rows
// (Quad.B_gamma_MITC4, element.py:735-801).
template <typename T>
struct QuadGP {
    T detJ;
    T dH[2][4];
    T gr, gs;            // MITC-4 scale factors
    T sa, ca, sb, cb;    // sin/cos alpha, beta
    double r, s;
};

template <typename T>
JDEV void quad_gp(const T xy[8], int g, QuadGP<T>& G) {
    const double r = gauss_r(g), s = gauss_s(g);
    G.r = r;
    G.s = s;
    const T x1 = xy[0], y1 = xy[1], x2 = xy[2], y2 = xy[3], x3 = xy[4], y3 = xy[5],
            x4 = xy[6], y4 = xy[7];
    T J00 = 0.25 * (x1 * (s + 1) - x2 * (s + 1) + x3 * (s - 1) - x4 * (s - 1));
    T J01 = 0.25 * (y1 * (s + 1) - y2 * (s + 1) + y3 * (s - 1) - y4 * (s - 1));
    T J10 = 0.25 * (x1 * (r + 1) - x2 * (r - 1) + x3 * (r - 1) - x4 * (r + 1));
    T J11 = 0.25 * (y1 * (r + 1) - y2 * (r - 1) + y3 * (r - 1) - y4 * (r + 1));
    T det = J00 * J11 - J01 * J10;
    G.detJ = det;
    const double dN0[4] = {0.25 * (1 + s), 0.25 * (-1 - s), 0.25 * (-1 + s), 0.25 * (1 - s)};
    const double dN1[4] = {0.25 * (1 + r), 0.25 * (1 - r), 0.25 * (-1 + r), 0.25 * (-1 - r)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        G.dH[0][k] = (J11 * dN0[k] - J01 * dN1[k]) / det;
        G.dH[1][k] = (J00 * dN1[k] - J10 * dN0[k]) / det;
    }
    T Ax = x1 - x2 - x3 + x4, Bx = x1 - x2 + x3 - x4, Cx = x1 + x2 - x3 - x4;
    T Ay = y1 - y2 - y3 + y4, By = y1 - y2 + y3 - y4, Cy = y1 + y2 - y3 - y4;
    T rx = (x1 + x4) / 2.0 - (x2 + x3) / 2.0, ry = (y1 + y4) / 2.0 - (y2 + y3) / 2.0;
    T sx = (x1 + x2) / 2.0 - (x3 + x4) / 2.0, sy = (y1 + y2) / 2.0 - (y3 + y4) / 2.0;
    T rn = dsqrt(rx * rx + ry * ry), sn = dsqrt(sx * sx + sy * sy);
    rx = rx / rn; ry = ry / rn;
    sx = sx / sn; sy = sy / sn;
    T cr = Cx + r * Bx, cry = Cy + r * By, as_ = Ax + s * Bx, asy = Ay + s * By;
    G.gr = dsqrt(cr * cr + cry * cry) / (8.0 * det);
    G.gs = dsqrt(as_ * as_ + asy * asy) / (8.0 * det);
    G.ca = rx;
    G.cb = sx;
    G.sa = -dsqrt(ry * ry);  // -|r_axis x x_axis|, sign of element.py:795
    G.sb = dsqrt(sy * sy);
}

// k-th entry (k = 3*node + {w, theta_x, theta_y}) of the MITC-4 covariant shear
// rows gamma_rz / gamma_sz (without the gr, gs factors), element.py:774-775.
template <typename T>
JDEV void quad_gamma_coef(const T xy[8], double r, double s, int k, T& crz, T& csz) {
    const T x1 = xy[0], y1 = xy[1], x2 = xy[2], y2 = xy[3], x3 = xy[4], y3 = xy[5],
            x4 = xy[6], y4 = xy[7];
    switch (k) {
        case 0: crz = T((1 + s) / 2); csz = T((1 + r) / 2); break;
        case 1: crz = -(y1 - y2) * (1 + s) / 4; csz = -(y1 - y4) * (1 + r) / 4; break;
        case 2: crz = (x1 - x2) * (1 + s) / 4; csz = (x1 - x4) * (1 + r) / 4; break;
        case 3: crz = T(-(1 + s) / 2); csz = T((1 - r) / 2); break;
        case 4: crz = -(y1 - y2) * (1 + s) / 4; csz = -(y2 - y3) * (1 - r) / 4; break;
        case 5: crz = (x1 - x2) * (1 + s) / 4; csz = (x2 - x3) * (1 - r) / 4; break;
        case 6: crz = T(-(1 - s) / 2); csz = T(-(1 - r) / 2); break;
        case 7: crz = -(y4 - y3) * (1 - s) / 4; csz = -(y2 - y3) * (1 - r) / 4; break;
        case 8: crz = (x4 - x3) * (1 - s) / 4; csz = (x2 - x3) * (1 - r) / 4; break;
        case 9: crz = T((1 - s) / 2); csz = T(-(1 + r) / 2); break;
        case 10: crz = -(y4 - y3) * (1 - s) / 4; csz = -(y1 - y4) * (1 + r) / 4; break;
        default: crz = (x4 - x3) * (1 - s) / 4; csz = (x1 - x4) * (1 + r) / 4; break;
    }
}

// Shear B column k (2 rows) at a Gauss point: B_gamma = [g_rz sin b - g_sz sin a;
// -g_rz cos b + g_sz cos a] (element.py:798).
template <typename T>
JDEV void quad_bgamma_col(const T xy[8], const QuadGP<T>& G, int k, T& b0, T& b1) {
    T crz, csz;
    quad_gamma_coef(xy, G.r, G.s, k, crz, csz);
    T grz = G.gr * crz, gsz = G.gs * csz;
    b0 = grz * G.sb - gsz * G.sa;
    b1 = -(grz * G.cb) + gsz * G.ca;
}

// Bending B column for (node a, comp c in {0:w, 1:theta_x, 2:theta_y}),
// Quad.B_kappa (element.py:728-730).
template <typename T>
JDEV void quad_bkappa_col(const QuadGP<T>& G, int a, int c, T& k0, T& k1, T& k2) {
    if (c == 0) {
        k0 = T(0.0); k1 = T(0.0); k2 = T(0.0);
    } else if (c == 1) {
        k0 = T(0.0); k1 = G.dH[1][a]; k2 = G.dH[0][a];
    } else {
        k0 = -G.dH[0][a]; k1 = T(0.0); k2 = -G.dH[1][a];
    }
}

// Plate constants: Cb = E h^3 / (12 (1 - nu^2)) [[1,nu,0],[nu,1,0],[0,0,(1-nu)/2]],
// Cs = 5/6 E h / (2 (1 + nu)) I (element.py:820-855).
template <typename T>
JDEV void quad_plate_consts(T t, T E, T nu, T& cb, T& cs) {
    cb = E * t * t * t / (12.0 * (1.0 - nu * nu));
    cs = E * t * (5.0 / 6.0) / (2.0 * (1.0 + nu));
}

// x^T Cb y for 3-vectors.
template <typename T>
JDEV T cb_form(T cb, T nu, T a0, T a1, T a2, T b0, T b1, T b2) {
    return cb * (a0 * b0 + nu * a0 * b1 + nu * a1 * b0 + a1 * b1 + (1.0 - nu) / 2.0 * a2 * b2);
}

// Membrane constitutive Cm (element.py:857-875): 1/(1-nu^2) [[Ex, nu Ex, 0],
// [nu Ey, Ey, 0], [0, 0, (1-nu^2) G]], Ex = E kx, Ey = E ky, G = E / (2 (1+nu)).
template <typename T>
JDEV void quad_cm(T E, T nu, T kx, T ky, T Cm[3][3]) {
    T f = 1.0 / (1.0 - nu * nu);
    T Ex = E * kx, Ey = E * ky, G = E / (2.0 * (1.0 + nu));
    Cm[0][0] = Ex * f; Cm[0][1] = nu * Ex * f; Cm[0][2] = T(0.0);
    Cm[1][0] = nu * Ey * f; Cm[1][1] = Ey * f; Cm[1][2] = T(0.0);
    Cm[2][0] = T(0.0); Cm[2][1] = T(0.0); Cm[2][2] = (1.0 - nu * nu) * G * f;
}

// Diagonal entry d (of the 12x12 bending+shear block, d in {1,2,4,5,7,8,10,11})
// summed over the Gauss points: candidates of the drilling spring
// k_rz = min |diag| / 1000 (element.py:953-978).
template <typename T>
JDEV void quad_rot_diag(const T xy[8], T cb, T cs, T nu, T diag[8]) {
#pragma unroll
    for (int m = 0; m < 8; ++m) diag[m] = T(0.0);
    for (int g = 0; g < 4; ++g) {
        QuadGP<T> G;
        quad_gp(xy, g, G);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int d = (m / 2) * 3 + 1 + (m & 1);
            T k0, k1, k2, b0, b1;
            quad_bkappa_col(G, d / 3, d % 3, k0, k1, k2);
            quad_bgamma_col(xy, G, d, b0, b1);
            diag[m] = diag[m] + (cb_form(cb, nu, k0, k1, k2, k0, k1, k2) + cs * (b0 * b0 + b1 * b1)) * G.detJ;
        }
    }
}

// k_rz with the tie-averaged tangent.
template <typename T>
JDEV T quad_drill(const T diag[8]) {
    double m = fabs(val(diag[0]));
#pragma unroll
    for (int i = 1; i < 8; ++i) m = fmin(m, fabs(val(diag[i])));
    double tsum = 0.0;
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double a = fabs(val(diag[i]));
        if (a - m <= TIE_RTOL * m) {
            tsum += (val(diag[i]) >= 0.0 ? 1.0 : -1.0) * tan_(diag[i]);
            cnt++;
        }
    }
    T out = T(m / 1000.0);
    set_tan(out, tsum / cnt / 1000.0);
    return out;
}

}  // namespace jsso
