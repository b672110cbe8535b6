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
JDEV void quad_gp_rs(const T xy[8], double r, double s, QuadGP<T>& G, T Jm[2][2] = nullptr);

template <typename T>
JDEV void quad_gp(const T xy[8], int g, QuadGP<T>& G) {
    quad_gp_rs<T>(xy, gauss_r(g), gauss_s(g), G);
}

// The same at any natural coordinates (r, s); Jm, if given, receives the 2x2
// Jacobian of Quad.J (element.py:698-709).
template <typename T>
JDEV void quad_gp_rs(const T xy[8], double r, double s, QuadGP<T>& G, T Jm[2][2]) {
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
    if (Jm) {
        Jm[0][0] = J00;
        Jm[0][1] = J01;
        Jm[1][0] = J10;
        Jm[1][1] = J11;
    }
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

// x^T Cb y for 3-vectors (constants of type TC, vectors of type T).
template <typename TC, typename T>
JDEV auto cb_form(TC cb, TC nu, T a0, T a1, T a2, T b0, T b1, T b2) -> decltype(TC(0.0) * T(0.0)) {
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
template <typename T, typename TM>
JDEV void quad_rot_diag(const T xy[8], TM cb, TM cs, TM nu, decltype(T(0.0) * TM(0.0)) diag[8]) {
    using P = decltype(T(0.0) * TM(0.0));
#pragma unroll
    for (int m = 0; m < 8; ++m) diag[m] = P(0.0);
#pragma unroll 1
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

// ---------------------------------------------------------------------------
// MITC-4 adjoint: phi = lambda_e^T K_e u_e for one load case (the element term
// of the reference's VJP through K, solver.py:221-250, with K_e of
// element.py:880-1084)
// ---------------------------------------------------------------------------
// phi depends on the 12 nodal coordinates only through the frame R (3 x 3) and
// the in-plane coordinates xy (six free entries: node 3 is the origin), so
// dphi/dc = dphi/dxy . dxy/dc + dphi/dR : dR/dc.

// An empty statement the compiler must assume reads and writes memory: placed
// at the top of a node's loop iteration it keeps that node's loads and products
// inside the iteration (hoisted, the four nodes' values of every pass stay live
// together and the adjoint kernels run out of registers).
JDEV void keep_loads_here() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::: "memory");
#endif
}

// Local values of node `node`: ul = R u, ll = R lambda per translation and
// rotation triple (lambda = lam, or lam_scale u when lam is null).
JDEV void quad_local_vals(const V3<double> R[3], const double* u, const double* lam, double lam_scale, int node,
                          int nrhs, int rhs, double ul[6], double ll[6]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        double uv[3], lv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int64_t gi = ((int64_t)6 * node + 3 * h + k) * nrhs + rhs;
            uv[k] = u[gi];
            lv[k] = lam ? lam[gi] : lam_scale * u[gi];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ul[3 * h + i] = R[i].x * uv[0] + R[i].y * uv[1] + R[i].z * uv[2];
            ll[3 * h + i] = R[i].x * lv[0] + R[i].y * lv[1] + R[i].z * lv[2];
        }
    }
}

// phi with the frame and the nodal values fixed; TX is the type of xy, TM that
// of the properties (Dual for the one direction being differentiated).
template <typename TX, typename TM>
JDEV auto quad_phi(const V3<double> R[3], const TX xy[8], TM t, TM E, TM nu, TM kx, TM ky, const int nd[4],
                   const double* u, const double* lam, double lam_scale, int nrhs, int rhs)
    -> decltype(TX(0.0) * TM(0.0)) {
    using P = decltype(TX(0.0) * TM(0.0));
    TM cb, cs, Cm[3][3];
    quad_plate_consts<TM>(t, E, nu, cb, cs);
    quad_cm<TM>(E, nu, kx, ky, Cm);
    P diag[8];
    quad_rot_diag<TX, TM>(xy, cb, cs, nu, diag);
    const P krz = quad_drill<P>(diag);
    P em(0.0), eb(0.0), es(0.0);
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
        QuadGP<TX> G;
        quad_gp<TX>(xy, g, G);
        TX mu[3] = {TX(0.0), TX(0.0), TX(0.0)}, ml[3] = {TX(0.0), TX(0.0), TX(0.0)};
        TX ku[3] = {TX(0.0), TX(0.0), TX(0.0)}, kl[3] = {TX(0.0), TX(0.0), TX(0.0)};
        TX gu[2] = {TX(0.0), TX(0.0)}, gl[2] = {TX(0.0), TX(0.0)};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            keep_loads_here();
            double ul[6], ll[6];
            quad_local_vals(R, u, lam, lam_scale, nd[a], nrhs, rhs, ul, ll);
            const TX h0 = G.dH[0][a], h1 = G.dH[1][a];
            mu[0] = mu[0] + h0 * ul[0];
            mu[1] = mu[1] + h1 * ul[1];
            mu[2] = mu[2] + h1 * ul[0] + h0 * ul[1];
            ml[0] = ml[0] + h0 * ll[0];
            ml[1] = ml[1] + h1 * ll[1];
            ml[2] = ml[2] + h1 * ll[0] + h0 * ll[1];
            ku[0] = ku[0] - h0 * ul[4];
            ku[1] = ku[1] + h1 * ul[3];
            ku[2] = ku[2] + h0 * ul[3] - h1 * ul[4];
            kl[0] = kl[0] - h0 * ll[4];
            kl[1] = kl[1] + h1 * ll[3];
            kl[2] = kl[2] + h0 * ll[3] - h1 * ll[4];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                TX b0, b1;
                quad_bgamma_col<TX>(xy, G, 3 * a + cc, b0, b1);
                gu[0] = gu[0] + b0 * ul[2 + cc];
                gu[1] = gu[1] + b1 * ul[2 + cc];
                gl[0] = gl[0] + b0 * ll[2 + cc];
                gl[1] = gl[1] + b1 * ll[2 + cc];
            }
        }
        P mform(0.0);
#pragma unroll
        for (int r = 0; r < 3; ++r) mform = mform + ml[r] * (Cm[r][0] * mu[0] + Cm[r][1] * mu[1] + Cm[r][2] * mu[2]);
        em = em + mform * G.detJ;
        eb = eb + cb_form(cb, nu, kl[0], kl[1], kl[2], ku[0], ku[1], ku[2]) * G.detJ;
        es = es + cs * (gl[0] * gu[0] + gl[1] * gu[1]) * G.detJ;
    }
    double ed = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double ul[6], ll[6];
        quad_local_vals(R, u, lam, lam_scale, nd[a], nrhs, rhs, ul, ll);
        ed += ll[5] * ul[5];
    }
    return t * em + eb + es + krz * ed;
}

// dphi/dR (reverse mode through the local values): phi is bilinear in ul and ll,
// dphi/dul = sum over the Gauss points of B^T (dphi/dstrain) plus the drilling
// term krz ll, likewise for ll, and ul = R u gives dphi/dR_ij = sum over the
// nodes and triples of dphi/dul_i u_j + dphi/dll_i lambda_j.
JDEV void quad_frame_adjoint(const V3<double> R[3], const double xy[8], double t, double E, double nu, double kx,
                             double ky, const int nd[4], const double* u, const double* lam, double lam_scale,
                             int nrhs, int rhs, double dR[3][3]) {
    double cb, cs, Cm[3][3], diag[8];
    quad_plate_consts<double>(t, E, nu, cb, cs);
    quad_cm<double>(E, nu, kx, ky, Cm);
    quad_rot_diag<double, double>(xy, cb, cs, nu, diag);
    const double krz = quad_drill<double>(diag);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dR[i][j] = 0.0;
    // adds dphi/d(ul, ll) of node a (gu, gl: the six local components) into dR
    auto scatter = [&](int a, const double gu[6], const double gl[6]) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int64_t gi = ((int64_t)6 * nd[a] + 3 * h + j) * nrhs + rhs;
                const double uj = u[gi], lj = lam ? lam[gi] : lam_scale * u[gi];
#pragma unroll
                for (int i = 0; i < 3; ++i) dR[i][j] += gu[3 * h + i] * uj + gl[3 * h + i] * lj;
            }
    };
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
        QuadGP<double> G;
        quad_gp<double>(xy, g, G);
        double mu[3] = {0.0, 0.0, 0.0}, ml[3] = {0.0, 0.0, 0.0}, ku[3] = {0.0, 0.0, 0.0}, kl[3] = {0.0, 0.0, 0.0};
        double gu[2] = {0.0, 0.0}, gl[2] = {0.0, 0.0};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            keep_loads_here();
            double ul[6], ll[6];
            quad_local_vals(R, u, lam, lam_scale, nd[a], nrhs, rhs, ul, ll);
            const double h0 = G.dH[0][a], h1 = G.dH[1][a];
            mu[0] += h0 * ul[0];
            mu[1] += h1 * ul[1];
            mu[2] += h1 * ul[0] + h0 * ul[1];
            ml[0] += h0 * ll[0];
            ml[1] += h1 * ll[1];
            ml[2] += h1 * ll[0] + h0 * ll[1];
            ku[0] -= h0 * ul[4];
            ku[1] += h1 * ul[3];
            ku[2] += h0 * ul[3] - h1 * ul[4];
            kl[0] -= h0 * ll[4];
            kl[1] += h1 * ll[3];
            kl[2] += h0 * ll[3] - h1 * ll[4];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                double b0, b1;
                quad_bgamma_col<double>(xy, G, 3 * a + cc, b0, b1);
                gu[0] += b0 * ul[2 + cc];
                gu[1] += b1 * ul[2 + cc];
                gl[0] += b0 * ll[2 + cc];
                gl[1] += b1 * ll[2 + cc];
            }
        }
        // dphi/dstrain of u (s*u) and of lambda (s*l) at this Gauss point
        const double tw = t * G.detJ, bw = cb * G.detJ, sw = cs * G.detJ;
        double smu[3], sml[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            smu[c] = tw * (ml[0] * Cm[0][c] + ml[1] * Cm[1][c] + ml[2] * Cm[2][c]);
            sml[c] = tw * (Cm[c][0] * mu[0] + Cm[c][1] * mu[1] + Cm[c][2] * mu[2]);
        }
        const double sku[3] = {bw * (kl[0] + nu * kl[1]), bw * (nu * kl[0] + kl[1]), bw * (1.0 - nu) / 2.0 * kl[2]};
        const double skl[3] = {bw * (ku[0] + nu * ku[1]), bw * (nu * ku[0] + ku[1]), bw * (1.0 - nu) / 2.0 * ku[2]};
        const double sgu[2] = {sw * gl[0], sw * gl[1]}, sgl[2] = {sw * gu[0], sw * gu[1]};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            keep_loads_here();
            const double h0 = G.dH[0][a], h1 = G.dH[1][a];
            double du[6], dl[6];
            du[0] = h0 * smu[0] + h1 * smu[2];
            du[1] = h1 * smu[1] + h0 * smu[2];
            du[3] = h1 * sku[1] + h0 * sku[2];
            du[4] = -h0 * sku[0] - h1 * sku[2];
            dl[0] = h0 * sml[0] + h1 * sml[2];
            dl[1] = h1 * sml[1] + h0 * sml[2];
            dl[3] = h1 * skl[1] + h0 * skl[2];
            dl[4] = -h0 * skl[0] - h1 * skl[2];
            du[2] = dl[2] = du[5] = dl[5] = 0.0;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {   // shear: w, theta_x, theta_y
                double b0, b1;
                quad_bgamma_col<double>(xy, G, 3 * a + cc, b0, b1);
                du[2 + cc] += b0 * sgu[0] + b1 * sgu[1];
                dl[2 + cc] += b0 * sgl[0] + b1 * sgl[1];
            }
            scatter(a, du, dl);
        }
    }
    // drilling: dphi/dul_5 = krz ll_5, dphi/dll_5 = krz ul_5
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double ul[6], ll[6];
        quad_local_vals(R, u, lam, lam_scale, nd[a], nrhs, rhs, ul, ll);
        const double du[6] = {0.0, 0.0, 0.0, 0.0, 0.0, krz * ll[5]}, dl[6] = {0.0, 0.0, 0.0, 0.0, 0.0, krz * ul[5]};
        scatter(a, du, dl);
    }
}

}  // namespace jsso
