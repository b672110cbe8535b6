// Element stiffness kernels (beam-column, MITC-4) and the forward-mode adjoint
// kernels.  Element math lives in element_math.hpp.
#include <mutex>

#include "element_math.hpp"
#include "jsso_internal.hpp"

namespace jsso {

// ---------------------------------------------------------------------------
// Beam-column stiffness: one thread per (element, node pair a, b).
// Ke_ab = blockdiag(D^-1, D^-1) K_loc,ab blockdiag(D, D)  (solve(T,K)@T,
// element.py:130-139); D^-1 = D^T for the orthogonal general branch.
// BLOCKS: output [e][a][b][6x6] (assembly), else row-major 12x12 per element.
// ---------------------------------------------------------------------------
template <bool BLOCKS>
__global__ __launch_bounds__(256) void ke_beam_kernel(int32_t n, const double* __restrict__ crds_e,
                                                      const int32_t* __restrict__ cnct,
                                                      const double* __restrict__ crds_nodes,
                                                      const double* __restrict__ props,
                                                      double* __restrict__ ke, int32_t* flags,
                                                      int32_t* status) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = tid >> 2;
    if (e >= n) return;
    const int a = (tid >> 1) & 1, b = tid & 1;
    double c[6];
    if (cnct) {
        const int n0 = cnct[2 * e], n1 = cnct[2 * e + 1];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c[k] = crds_nodes[3 * n0 + k];
            c[3 + k] = crds_nodes[3 * n1 + k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) c[k] = crds_e[6 * e + k];
    }
    const double E = props[6 * e + 0], G = props[6 * e + 1], Iy = props[6 * e + 2],
                 Iz = props[6 * e + 3], Jt = props[6 * e + 4], A = props[6 * e + 5];
    double D[3][3], Di[3][3], L;
    const bool special = beam_dircos<double>(c, D, L);
    if (special) {
        inv3(D, Di);
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Di[i][j] = D[j][i];
    }
    if (flags && a == 0 && b == 0) flags[e] = special ? 1 : 0;
    if (status && special && a == 0 && b == 0) atomicOr(status, STATUS_NONSYM | STATUS_Y_BEAM);
    double out[6][6];
#pragma unroll
    for (int pq = 0; pq < 4; ++pq) {
        const int p = pq >> 1, q = pq & 1;  // translation / rotation sub-blocks
        double S[3][3], tmp[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                S[i][j] = beam_kloc<double>(6 * a + 3 * p + i, 6 * b + 3 * q + j, L, E, G, Iy, Iz, Jt, A);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                tmp[i][j] = S[i][0] * D[0][j] + S[i][1] * D[1][j] + S[i][2] * D[2][j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                out[3 * p + i][3 * q + j] = Di[i][0] * tmp[0][j] + Di[i][1] * tmp[1][j] + Di[i][2] * tmp[2][j];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (BLOCKS)
                ke[((int64_t)e * 4 + a * 2 + b) * 36 + i * 6 + j] = out[i][j];
            else
                ke[(int64_t)e * 144 + (6 * a + i) * 12 + 6 * b + j] = out[i][j];
        }
}

// ---------------------------------------------------------------------------
// MITC-4 stiffness: 16 lanes per element (one per node pair), 16 elements per
// 256-thread workgroup.  Lanes 0-3 stage the per-Gauss-point tiles (det J, dH,
// shear B) in LDS, lanes 0-7 the drilling candidates; then every lane forms
// its 6x6 block of T^T (k_m + k_b) T (element.py:1073-1084).
// MODE: QK_BLOCKS output [e][a][b][6x6] (assembly), QK_GLOBAL row-major 24x24
// per element, QK_LOCAL row-major 24x24 of k_m + k_b before the rotation
// (Quad.element_K_quad_local, element.py:1086-1095).  EPW elements per
// workgroup (4, 8 or 16: 64, 128 or 256 threads), chosen per plan by timing on
// the live assembly (tune_ke_quad).
// ---------------------------------------------------------------------------
enum { QK_BLOCKS = 0, QK_GLOBAL = 1, QK_LOCAL = 2 };

struct QuadTile {
    double R[9];
    double t, E, nu, kx, ky;
    double detJ[4];
    double dH[4][8];
    double Bg[4][24];
    double diag[8];
};

template <int MODE, int EPW = 16>
__global__ __launch_bounds__(16 * EPW) void ke_quad_kernel(int32_t n, const double* __restrict__ crds_e,
                                                           const int32_t* __restrict__ cnct,
                                                           const double* __restrict__ crds_nodes,
                                                           const double* __restrict__ props,
                                                           double* __restrict__ ke, int32_t* status = nullptr) {
    __shared__ QuadTile sh[EPW];
    const int le = threadIdx.x >> 4, lane = threadIdx.x & 15;
    const int e = blockIdx.x * EPW + le;
    const bool valid = e < n;
    QuadTile& T = sh[le];
    if (valid && lane < 4) {
        double c[12];
        if (cnct) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int nd = cnct[4 * e + a];
#pragma unroll
                for (int k = 0; k < 3; ++k) c[3 * a + k] = crds_nodes[3 * nd + k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) c[k] = crds_e[12 * e + k];
        }
        V3<double> R[3];
        double xy[8];
        quad_frame<double>(c, R, xy);
        QuadGP<double> G;
        quad_gp<double>(xy, lane, G);
        T.detJ[lane] = G.detJ;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            T.dH[lane][k] = G.dH[0][k];
            T.dH[lane][4 + k] = G.dH[1][k];
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) quad_bgamma_col<double>(xy, G, k, T.Bg[lane][k], T.Bg[lane][12 + k]);
        if (lane == 0) {
            T.R[0] = R[0].x; T.R[1] = R[0].y; T.R[2] = R[0].z;
            T.R[3] = R[1].x; T.R[4] = R[1].y; T.R[5] = R[1].z;
            T.R[6] = R[2].x; T.R[7] = R[2].y; T.R[8] = R[2].z;
            T.t = props[5 * e + 0];
            T.E = props[5 * e + 1];
            T.nu = props[5 * e + 2];
            T.kx = props[5 * e + 3];
            T.ky = props[5 * e + 4];
            // Cm of element.py:852-875 is [[Ex, nu Ex, 0], [nu Ey, Ey, 0], ...]: not
            // symmetric when kx_mod != ky_mod (and nu != 0), so neither is this Ke
            if (status && T.nu != 0.0 && T.kx != T.ky) atomicOr(status, STATUS_NONSYM | STATUS_ORTHO_QUAD);
        }
    }
    __syncthreads();
    if (valid && lane < 8) {
        double cb, cs;
        quad_plate_consts<double>(T.t, T.E, T.nu, cb, cs);
        const int d = (lane / 2) * 3 + 1 + (lane & 1);
        const int a = d / 3, comp = d % 3;
        double acc = 0.0;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            double k0 = 0, k1 = 0, k2 = 0;
            if (comp == 1) { k1 = T.dH[g][4 + a]; k2 = T.dH[g][a]; }
            else { k0 = -T.dH[g][a]; k2 = -T.dH[g][4 + a]; }
            const double b0 = T.Bg[g][d], b1 = T.Bg[g][12 + d];
            acc += (cb_form<double>(cb, T.nu, k0, k1, k2, k0, k1, k2) + cs * (b0 * b0 + b1 * b1)) * T.detJ[g];
        }
        T.diag[lane] = acc;
    }
    __syncthreads();
    if (!valid) return;
    const int a = lane >> 2, b = lane & 3;
    double cb, cs, Cm[3][3];
    quad_plate_consts<double>(T.t, T.E, T.nu, cb, cs);
    quad_cm<double>(T.E, T.nu, T.kx, T.ky, Cm);
    double mem[2][2] = {{0, 0}, {0, 0}}, ben[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double dJ = T.detJ[g];
        const double ha0 = T.dH[g][a], ha1 = T.dH[g][4 + a], hb0 = T.dH[g][b], hb1 = T.dH[g][4 + b];
        // membrane columns: u -> (h0, 0, h1), v -> (0, h1, h0)
        const double Ba[2][3] = {{ha0, 0.0, ha1}, {0.0, ha1, ha0}};
        const double Bb[2][3] = {{hb0, 0.0, hb1}, {0.0, hb1, hb0}};
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    s += Ba[i][r] * (Cm[r][0] * Bb[j][0] + Cm[r][1] * Bb[j][1] + Cm[r][2] * Bb[j][2]);
                mem[i][j] += s * dJ;
            }
        // bending columns w, theta_x, theta_y
        const double Ka[3][3] = {{0, 0, 0}, {0, ha1, ha0}, {-ha0, 0, -ha1}};
        const double Kb[3][3] = {{0, 0, 0}, {0, hb1, hb0}, {-hb0, 0, -hb1}};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double kb_ = cb_form<double>(cb, T.nu, Ka[i][0], Ka[i][1], Ka[i][2], Kb[j][0], Kb[j][1], Kb[j][2]);
                const double sh_ = cs * (T.Bg[g][3 * a + i] * T.Bg[g][3 * b + j] +
                                         T.Bg[g][12 + 3 * a + i] * T.Bg[g][12 + 3 * b + j]);
                ben[i][j] += (kb_ + sh_) * dJ;
            }
    }
    // local 6x6 block, DOFs [u, v, w, rx, ry, rz]
    double Lb[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) Lb[i][j] = 0.0;
    Lb[0][0] = T.t * mem[0][0]; Lb[0][1] = T.t * mem[0][1];
    Lb[1][0] = T.t * mem[1][0]; Lb[1][1] = T.t * mem[1][1];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Lb[2 + i][2 + j] = ben[i][j];
    if (a == b) {
        double m = fabs(T.diag[0]);
#pragma unroll
        for (int i = 1; i < 8; ++i) m = fmin(m, fabs(T.diag[i]));
        Lb[5][5] = m / 1000.0;
    }
    if (MODE == QK_LOCAL) {
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) ke[(int64_t)e * 576 + (6 * a + i) * 24 + 6 * b + j] = Lb[i][j];
        return;
    }
    double R[3][3];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i / 3][i % 3] = T.R[i];
    double out[6][6];
#pragma unroll
    for (int pq = 0; pq < 4; ++pq) {
        const int p = pq >> 1, q = pq & 1;
        double tmp[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                tmp[i][j] = Lb[3 * p + i][3 * q + 0] * R[0][j] + Lb[3 * p + i][3 * q + 1] * R[1][j] +
                            Lb[3 * p + i][3 * q + 2] * R[2][j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                out[3 * p + i][3 * q + j] = R[0][i] * tmp[0][j] + R[1][i] * tmp[1][j] + R[2][i] * tmp[2][j];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (MODE == QK_BLOCKS)
                ke[((int64_t)e * 16 + lane) * 36 + i * 6 + j] = out[i][j];
            else
                ke[(int64_t)e * 576 + (6 * a + i) * 24 + 6 * b + j] = out[i][j];
        }
}

int launch_ke_beam(int32_t n, const double* crds_e, const int32_t* cnct, const double* crds_nodes,
                   const double* props, double* ke, int32_t* flags, int32_t* status, hipStream_t s) {
    if (n <= 0) return JSSO_OK;
    const int threads = 256;
    const int64_t total = (int64_t)n * 4;
    const int grid = (int)((total + threads - 1) / threads);
    if (cnct)
        hipLaunchKernelGGL(ke_beam_kernel<true>, dim3(grid), dim3(threads), 0, s, n, crds_e, cnct,
                           crds_nodes, props, ke, flags, status);
    else
        hipLaunchKernelGGL(ke_beam_kernel<false>, dim3(grid), dim3(threads), 0, s, n, crds_e, cnct,
                           crds_nodes, props, ke, flags, status);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

int launch_ke_quad(int32_t n, const double* crds_e, const int32_t* cnct, const double* crds_nodes,
                   const double* props, double* ke, hipStream_t s, int epw, int32_t* status) {
    if (n <= 0) return JSSO_OK;
    const int grid = (n + 15) / 16;
    if (cnct) {
        switch (epw) {
            case 4:
                hipLaunchKernelGGL((ke_quad_kernel<QK_BLOCKS, 4>), dim3((n + 3) / 4), dim3(64), 0, s, n, crds_e, cnct,
                                   crds_nodes, props, ke, status);
                break;
            case 8:
                hipLaunchKernelGGL((ke_quad_kernel<QK_BLOCKS, 8>), dim3((n + 7) / 8), dim3(128), 0, s, n, crds_e,
                                   cnct, crds_nodes, props, ke, status);
                break;
            default:
                hipLaunchKernelGGL((ke_quad_kernel<QK_BLOCKS, 16>), dim3(grid), dim3(256), 0, s, n, crds_e, cnct,
                                   crds_nodes, props, ke, status);
        }
    } else
        hipLaunchKernelGGL((ke_quad_kernel<QK_GLOBAL, 16>), dim3(grid), dim3(256), 0, s, n, crds_e, cnct,
                           crds_nodes, props, ke, status);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

// ---------------------------------------------------------------------------
// Element-level parts of the reference's element API, one thread per element.
// ---------------------------------------------------------------------------

// BeamCol.T (12x12 = blockdiag of the direction cosines, element.py:66-105) and
// BeamCol.K_local (element.py:107-128), row-major; either output may be NULL.
__global__ __launch_bounds__(256) void beam_local_kernel(int32_t n, const double* __restrict__ crds,
                                                         const double* __restrict__ props, double* __restrict__ T,
                                                         double* __restrict__ kl) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double c[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = crds[6 * e + k];
    double D[3][3], L;
    beam_dircos<double>(c, D, L);
    if (T) {
        double* t = T + (int64_t)e * 144;
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) t[i * 12 + j] = (i / 3 == j / 3) ? D[i % 3][j % 3] : 0.0;
    }
    if (kl) {
        const double E = props[6 * e + 0], G = props[6 * e + 1], Iy = props[6 * e + 2], Iz = props[6 * e + 3],
                     Jt = props[6 * e + 4], A = props[6 * e + 5];
        double* k = kl + (int64_t)e * 144;
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) k[i * 12 + j] = beam_kloc<double>(i, j, L, E, G, Iy, Iz, Jt, A);
    }
}

// Quad.loc_crds (in-plane coordinates, the normal of the element's plane and of
// the triangles 1-2-3 and 1-3-4, element.py:488-539) and Quad.T (24x24 =
// blockdiag of the frame, element.py:643-696); each output may be NULL.
__global__ __launch_bounds__(256) void quad_frame_kernel(int32_t n, const double* __restrict__ crds,
                                                         double* __restrict__ xy_out, double* __restrict__ normals,
                                                         double* __restrict__ T) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double c[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) c[k] = crds[12 * e + k];
    V3<double> R[3];
    double xy[8];
    quad_frame<double>(c, R, xy);
    if (xy_out)
#pragma unroll
        for (int k = 0; k < 8; ++k) xy_out[8 * e + k] = xy[k];
    if (normals) {
        const V3<double> X0{c[0], c[1], c[2]}, X1{c[3], c[4], c[5]}, X2{c[6], c[7], c[8]}, X3{c[9], c[10], c[11]};
        const V3<double> v31 = sub(X0, X2), v32 = sub(X1, X2), v34 = sub(X3, X2);
        V3<double> t1 = cross(v31, v32), t2 = cross(v34, v31);
        t1 = scale(t1, norm(t1));
        t2 = scale(t2, norm(t2));
        const double nv[9] = {R[2].x, R[2].y, R[2].z, t1.x, t1.y, t1.z, t2.x, t2.y, t2.z};
#pragma unroll
        for (int k = 0; k < 9; ++k) normals[9 * e + k] = nv[k];
    }
    if (T) {
        const double D[3][3] = {{R[0].x, R[0].y, R[0].z}, {R[1].x, R[1].y, R[1].z}, {R[2].x, R[2].y, R[2].z}};
        double* t = T + (int64_t)e * 576;
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) t[i * 24 + j] = (i / 3 == j / 3) ? D[i % 3][j % 3] : 0.0;
    }
}

// At natural coordinates (r, s): Quad.J (2x2), B_kappa (3x12), B_gamma_MITC4
// (2x12), B_m (3x8) from the in-plane coordinates (element.py:698-818), and the
// constitutive Cb (3x3), Cs (2x2), Cm (3x3) from [t, E, nu, kx_mod, ky_mod]
// (element.py:820-875); row-major, each output may be NULL.
__global__ __launch_bounds__(256) void quad_gauss_kernel(int32_t n, const double* __restrict__ xy_in,
                                                         const double* __restrict__ props, double r, double s,
                                                         double* __restrict__ Jo, double* __restrict__ Bk,
                                                         double* __restrict__ Bg, double* __restrict__ Bm,
                                                         double* __restrict__ Cb, double* __restrict__ Cs,
                                                         double* __restrict__ Cm) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    if (xy_in) {
        double xy[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) xy[k] = xy_in[8 * e + k];
        QuadGP<double> G;
        double Jm[2][2];
        quad_gp_rs<double>(xy, r, s, G, Jm);
        if (Jo)
#pragma unroll
            for (int k = 0; k < 4; ++k) Jo[4 * e + k] = Jm[k / 2][k % 2];
        if (Bk) {
            double* b = Bk + (int64_t)e * 36;
            for (int a = 0; a < 4; ++a)
                for (int c = 0; c < 3; ++c) {
                    double k0, k1, k2;
                    quad_bkappa_col<double>(G, a, c, k0, k1, k2);
                    b[0 * 12 + 3 * a + c] = k0;
                    b[1 * 12 + 3 * a + c] = k1;
                    b[2 * 12 + 3 * a + c] = k2;
                }
        }
        if (Bg) {
            double* b = Bg + (int64_t)e * 24;
            for (int k = 0; k < 12; ++k) quad_bgamma_col<double>(xy, G, k, b[k], b[12 + k]);
        }
        if (Bm) {
            double* b = Bm + (int64_t)e * 24;
            for (int a = 0; a < 4; ++a) {
                b[0 * 8 + 2 * a] = G.dH[0][a];
                b[0 * 8 + 2 * a + 1] = 0.0;
                b[1 * 8 + 2 * a] = 0.0;
                b[1 * 8 + 2 * a + 1] = G.dH[1][a];
                b[2 * 8 + 2 * a] = G.dH[1][a];
                b[2 * 8 + 2 * a + 1] = G.dH[0][a];
            }
        }
    }
    if (props) {
        const double t = props[5 * e + 0], E = props[5 * e + 1], nu = props[5 * e + 2], kx = props[5 * e + 3],
                     ky = props[5 * e + 4];
        double cb, cs, M[3][3];
        quad_plate_consts<double>(t, E, nu, cb, cs);
        if (Cb) {
            const double v[9] = {cb, cb * nu, 0.0, cb * nu, cb, 0.0, 0.0, 0.0, cb * (1.0 - nu) / 2.0};
#pragma unroll
            for (int k = 0; k < 9; ++k) Cb[9 * e + k] = v[k];
        }
        if (Cs) {
            Cs[4 * e + 0] = cs;
            Cs[4 * e + 1] = 0.0;
            Cs[4 * e + 2] = 0.0;
            Cs[4 * e + 3] = cs;
        }
        if (Cm) {
            quad_cm<double>(E, nu, kx, ky, M);
#pragma unroll
            for (int k = 0; k < 9; ++k) Cm[9 * e + k] = M[k / 3][k % 3];
        }
    }
}

// Elements per workgroup of the assembly's MITC-4 kernel, timed on the live
// plan at its first assembly (every variant writes the same element blocks,
// bit for bit): 4, 8 and 16 elements (64, 128, 256 threads) per workgroup.
int tune_ke_quad(Plan* p, const double* crds, const double* props, double* ke, hipStream_t s, int32_t& out) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    JSSO_HIP(hipEventCreate(&e0));
    JSSO_HIP(hipEventCreate(&e1));
    float best = 0.f;
    int pick = 16;
    for (int epw : {16, 8, 4}) {
        int rc = launch_ke_quad(p->n_quad, nullptr, p->d_cnct_quad, crds, props, ke, s, epw, p->d_status);   // warm
        if (rc) return rc;
        JSSO_HIP(hipEventRecord(e0, s));
        for (int rep = 0; rep < 2; ++rep)
            if ((rc = launch_ke_quad(p->n_quad, nullptr, p->d_cnct_quad, crds, props, ke, s, epw, p->d_status)))
                return rc;
        JSSO_HIP(hipEventRecord(e1, s));
        JSSO_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        JSSO_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (epw == 16 || ms < best) {
            best = ms;
            pick = epw;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out = pick;
    return JSSO_OK;
}

int launch_beam_local(int32_t n, const double* crds, const double* props, double* T, double* kl, hipStream_t s) {
    if (n <= 0) return JSSO_OK;
    hipLaunchKernelGGL(beam_local_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, crds, props, T, kl);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

int launch_quad_local(int32_t n, const double* crds, const double* props, double* xy, double* normals, double* T,
                      double* kl, hipStream_t s) {
    if (n <= 0) return JSSO_OK;
    if (xy || normals || T)
        hipLaunchKernelGGL(quad_frame_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, crds, xy, normals, T);
    if (kl)
        hipLaunchKernelGGL((ke_quad_kernel<QK_LOCAL, 16>), dim3((n + 15) / 16), dim3(256), 0, s, n, crds,
                           (const int32_t*)nullptr, (const double*)nullptr, props, kl, (int32_t*)nullptr);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

int launch_quad_gauss(int32_t n, const double* xy, const double* props, double r, double s_, double* J, double* Bk,
                      double* Bg, double* Bm, double* Cb, double* Cs, double* Cm, hipStream_t s) {
    if (n <= 0) return JSSO_OK;
    hipLaunchKernelGGL(quad_gauss_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, xy, props, r, s_, J, Bk, Bg,
                       Bm, Cb, Cs, Cm);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

// ---------------------------------------------------------------------------
// Adjoint: one thread per (element, input direction).  The element energy
// phi = -lam_e^T K_e u_e is evaluated in dual numbers seeded on that input;
// its tangent is dphi/dp.  Beam directions: 6 coordinates, then E,G,Iy,Iz,J,A.
// Quad directions: 12 coordinates, then t,E,nu,kx_mod,ky_mod.
// gbuf[(elem_dir) * nrhs + rhs] receives dphi/dp (coordinate directions);
// property directions go straight to dprop (layout [e*nprop + k][rhs]).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double load_lam(const double* u, const double* lam, double scale, int64_t i) {
    return lam ? lam[i] : scale * u[i];
}

__global__ __launch_bounds__(256) void adj_beam_kernel(int32_t n, const int32_t* __restrict__ cnct,
                                                       const double* __restrict__ crds,
                                                       const double* __restrict__ props,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ lam, double lam_scale,
                                                       int nrhs, double* __restrict__ gbuf,
                                                       double* __restrict__ dprop) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)n * BDIR) return;
    const int e = (int)(tid / BDIR), p = (int)(tid % BDIR);
    const int n0 = cnct[2 * e], n1 = cnct[2 * e + 1];
    Dual c[6], pr[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c[k] = Dual(crds[3 * n0 + k]);
        c[3 + k] = Dual(crds[3 * n1 + k]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) pr[k] = Dual(props[6 * e + k]);
    if (p < 6) c[p].d = 1.0; else pr[p - 6].d = 1.0;
    Dual D[3][3], L;
    const bool special = beam_dircos<Dual>(c, D, L);
    double Dm[3][3], Dti[3][3];
    if (special) {
        double Dv[3][3], Di[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Dv[i][j] = D[i][j].v;
        inv3(Dv, Di);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) { Dm[i][j] = Dv[i][j]; Dti[i][j] = Di[j][i]; }
    }
    for (int rhs = 0; rhs < nrhs; ++rhs) {
        double out;
        if (special && p < 6) {
            out = __builtin_nan("");
        } else {
            Dual ul[12], ll[12];
            for (int a = 0; a < 2; ++a) {
                const int nd = a ? n1 : n0;
                for (int h = 0; h < 2; ++h) {
                    double uv[3], lv[3];
                    for (int k = 0; k < 3; ++k) {
                        const int64_t gi = ((int64_t)6 * nd + 3 * h + k) * nrhs + rhs;
                        uv[k] = u[gi];
                        lv[k] = load_lam(u, lam, lam_scale, gi);
                    }
                    for (int i = 0; i < 3; ++i) {
                        if (special) {
                            ul[6 * a + 3 * h + i] = Dual(Dm[i][0] * uv[0] + Dm[i][1] * uv[1] + Dm[i][2] * uv[2]);
                            ll[6 * a + 3 * h + i] = Dual(Dti[i][0] * lv[0] + Dti[i][1] * lv[1] + Dti[i][2] * lv[2]);
                        } else {
                            ul[6 * a + 3 * h + i] = D[i][0] * uv[0] + D[i][1] * uv[1] + D[i][2] * uv[2];
                            ll[6 * a + 3 * h + i] = D[i][0] * lv[0] + D[i][1] * lv[1] + D[i][2] * lv[2];
                        }
                    }
                }
            }
            Dual phi(0.0);
            for (int I = 0; I < 12; ++I) {
                Dual row(0.0);
                for (int J = 0; J < 12; ++J) {
                    const Dual k = beam_kloc<Dual>(I, J, L, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5]);
                    if (k.v != 0.0 || k.d != 0.0) row = row + k * ul[J];
                }
                phi = phi + ll[I] * row;
            }
            out = -phi.d;
        }
        if (p < 6)
            gbuf[((int64_t)e * BDIR + p) * nrhs + rhs] = out;
        else if (dprop)
            dprop[((int64_t)e * 6 + (p - 6)) * nrhs + rhs] = out;
    }
}

// Quad adjoint for load case `rhs`.  phi = lambda_e^T K_e u_e depends on the
// coordinates only through the frame R and the in-plane coordinates xy
// (element_math.hpp), so it is differentiated in three kernels that each run
// uniform wavefronts:
//   adj_quad_fwd:     forward mode in one free in-plane coordinate (k < 6: dphi/dxy
//                     into the element's gbuf slot k) or one property (k >= 6:
//                     dprop), with R and the local nodal values plain doubles;
//   adj_quad_frame:   dphi/dR by reverse mode (gbuf slots 6..14);
//   adj_quad_combine: thread (element, coordinate p) applies the frame's Jacobian
//                     column dc_p (forward mode through quad_frame) to those 15
//                     partials and overwrites slot p.
// The result equals forward mode through the whole element per coordinate (to
// rounding) at about half the work, and without the dual frame and nodal values
// that made that form spill.
__device__ constexpr int QA_XY_FREE[6] = {0, 1, 2, 3, 6, 7};   // xy[4] = xy[5] = 0: node 3 is the origin
constexpr int QA_FWD = 11;                                      // 6 in-plane coordinates + 5 properties
constexpr int QA_E = 32;                                        // elements per combine workgroup

__device__ __forceinline__ void quad_geometry(int64_t e, const int32_t* __restrict__ cnct,
                                              const double* __restrict__ crds, int nd[4], V3<double> R[3],
                                              double xy[8]) {
    double c[12];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        nd[a] = cnct[4 * e + a];
#pragma unroll
        for (int q = 0; q < 3; ++q) c[3 * a + q] = crds[3 * nd[a] + q];
    }
    quad_frame<double>(c, R, xy);
}

__device__ __forceinline__ int64_t quad_slot(int32_t n_beam, int64_t e, int slot, int nrhs, int rhs) {
    return ((int64_t)n_beam * BDIR + e * QDIR + slot) * nrhs + rhs;
}

__global__ __launch_bounds__(256) void adj_quad_fwd(int32_t n, int32_t n_beam, const int32_t* __restrict__ cnct,
                                                    const double* __restrict__ crds,
                                                    const double* __restrict__ props, const double* __restrict__ u,
                                                    const double* __restrict__ lam, double lam_scale, int nrhs,
                                                    int rhs, double* __restrict__ gbuf,
                                                    double* __restrict__ dprop) {
    // workgroup g: direction g % QA_FWD of the 256 elements of chunk g / QA_FWD, so
    // that the 11 workgroups of a chunk (one XCD, in sequence) share its inputs in L2
    const int64_t g = xcd_wg();
    const int k = (int)(g % QA_FWD);
    const int64_t e = (g / QA_FWD) * 256 + threadIdx.x;
    if (e >= n) return;
    int nd[4];
    V3<double> R[3];
    double xy[8];
    quad_geometry(e, cnct, crds, nd, R, xy);
    const double* pr = props + 5 * e;
    if (k < 6) {
        Dual xd[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) xd[i] = Dual(xy[i], i == QA_XY_FREE[k] ? 1.0 : 0.0);
        gbuf[quad_slot(n_beam, e, k, nrhs, rhs)] =
            quad_phi<Dual, double>(R, xd, pr[0], pr[1], pr[2], pr[3], pr[4], nd, u, lam, lam_scale, nrhs, rhs).d;
    } else {
        Dual m[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) m[i] = Dual(pr[i], i == k - 6 ? 1.0 : 0.0);
        const double d =
            quad_phi<double, Dual>(R, xy, m[0], m[1], m[2], m[3], m[4], nd, u, lam, lam_scale, nrhs, rhs).d;
        if (dprop) dprop[(e * 5 + (k - 6)) * nrhs + rhs] = -d;
    }
}

__global__ __launch_bounds__(256) void adj_quad_frame(int32_t n, int32_t n_beam, const int32_t* __restrict__ cnct,
                                                      const double* __restrict__ crds,
                                                      const double* __restrict__ props,
                                                      const double* __restrict__ u, const double* __restrict__ lam,
                                                      double lam_scale, int nrhs, int rhs,
                                                      double* __restrict__ gbuf) {
    const int64_t e = xcd_wg() * blockDim.x + threadIdx.x;
    if (e >= n) return;
    int nd[4];
    V3<double> R[3];
    double xy[8];
    quad_geometry(e, cnct, crds, nd, R, xy);
    const double* pr = props + 5 * e;
    double dR[3][3];
    quad_frame_adjoint(R, xy, pr[0], pr[1], pr[2], pr[3], pr[4], nd, u, lam, lam_scale, nrhs, rhs, dR);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) gbuf[quad_slot(n_beam, e, 6 + 3 * i + j, nrhs, rhs)] = dR[i][j];
}

__global__ __launch_bounds__(12 * QA_E) void adj_quad_combine(int32_t n, int32_t n_beam,
                                                              const int32_t* __restrict__ cnct,
                                                              const double* __restrict__ crds, int nrhs, int rhs,
                                                              double* __restrict__ gbuf) {
    __shared__ double part[QA_E][15];   // dphi/dxy (6 free entries), dphi/dR (9)
    const int tid = threadIdx.x;
    const int64_t e0 = xcd_wg() * QA_E;
    for (int v = tid; v < QA_E * 15; v += blockDim.x) {
        const int el = v / 15, q = v % 15;
        if (e0 + el < n) part[el][q] = gbuf[quad_slot(n_beam, e0 + el, q, nrhs, rhs)];
    }
    __syncthreads();   // every partial is read before any slot is overwritten
    const int el = tid % QA_E, p = tid / QA_E;   // coordinate p = 3 * node + component
    const int64_t e = e0 + el;
    if (e >= n) return;
    Dual c[12];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int node = cnct[4 * e + a];
#pragma unroll
        for (int q = 0; q < 3; ++q) c[3 * a + q] = Dual(crds[3 * node + q], 3 * a + q == p ? 1.0 : 0.0);
    }
    V3<Dual> Rd[3];
    Dual xyd[8];
    quad_frame<Dual>(c, Rd, xyd);
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) sum += part[el][q] * xyd[QA_XY_FREE[q]].d;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        sum += part[el][6 + 3 * i] * Rd[i].x.d + part[el][7 + 3 * i] * Rd[i].y.d + part[el][8 + 3 * i] * Rd[i].z.d;
    gbuf[quad_slot(n_beam, e, p, nrhs, rhs)] = -sum;
}

// node gradient: sum of the coordinate derivatives of the elements incident to
// each node, in a fixed order (no atomics).
__global__ void gather_grad_kernel(int32_t n_nodes, int nrhs, const int32_t* __restrict__ nptr,
                                   const int32_t* __restrict__ nsrc, const double* __restrict__ gbuf,
                                   double* __restrict__ dcrds) {
    const int64_t tid = xcd_wg() * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)n_nodes * 3 * nrhs) return;
    const int rhs = (int)(tid % nrhs);
    const int64_t nc = tid / nrhs;
    const int node = (int)(nc / 3), comp = (int)(nc % 3);
    double s = 0.0;
    for (int k = nptr[node]; k < nptr[node + 1]; ++k) s += gbuf[((int64_t)nsrc[k] + comp) * nrhs + rhs];
    dcrds[tid] = s;
}

int launch_adjoint(Plan* p, const double* crds, const double* prop_beam, const double* prop_quad,
                        const double* u, const double* lam, double lam_scale, int nrhs, double* gbuf,
                        double* dprop_beam, double* dprop_quad, hipStream_t s) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    if (p->n_beam > 0) {
        const int64_t total = (int64_t)p->n_beam * BDIR;
        hipLaunchKernelGGL(adj_beam_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           p->n_beam, p->d_cnct_beam, crds, prop_beam, u, lam, lam_scale, nrhs, gbuf,
                           dprop_beam);
        JSSO_HIP(hipGetLastError());
    }
    if (p->n_quad > 0) {
        const int nq = p->n_quad;
        for (int rhs = 0; rhs < nrhs; ++rhs) {
            hipLaunchKernelGGL(adj_quad_fwd, xcd_grid((int64_t)QA_FWD * ((nq + 255) / 256)), dim3(256), 0, s,
                               nq, p->n_beam, p->d_cnct_quad, crds, prop_quad, u, lam, lam_scale, nrhs, rhs, gbuf,
                               dprop_quad);
            hipLaunchKernelGGL(adj_quad_frame, xcd_grid((nq + 255) / 256), dim3(256), 0, s, nq, p->n_beam,
                               p->d_cnct_quad, crds, prop_quad, u, lam, lam_scale, nrhs, rhs, gbuf);
            hipLaunchKernelGGL(adj_quad_combine, xcd_grid((nq + QA_E - 1) / QA_E), dim3(12 * QA_E), 0, s,
                               nq, p->n_beam, p->d_cnct_quad, crds, nrhs, rhs, gbuf);
        }
        JSSO_HIP(hipGetLastError());
    }
    return JSSO_OK;
}

// ---------------------------------------------------------------------------
// Fixed gradient scratch for jsso_adjoint_grad_noalloc: a file-scope device
// array (one copy per GPU, held until the process ends), so that callers that
// cannot allocate (graph capture, serving loops) get the adjoint without a
// hipMalloc.  Its device address is looked up once per device with
// hipGetSymbolAddress; host code never passes the array's name itself (that is
// a host-side stand-in, not a device address).  Every call on a GPU shares
// that GPU's array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

int noalloc_scratch(double** ptr, size_t* bytes) {
    static std::mutex mu;
    static void* dev_ptr[64] = {};
    int dev = -1;
    JSSO_HIP(hipGetDevice(&dev));
    JSSO_REQUIRE(dev >= 0 && dev < 64, JSSO_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!dev_ptr[dev]) {
        void* p = nullptr;
        JSSO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        JSSO_REQUIRE(p, JSSO_EHIP, "hipGetSymbolAddress gave no address for the adjoint scratch");
        dev_ptr[dev] = p;
    }
    *ptr = (double*)dev_ptr[dev];
    *bytes = sizeof(double) * (size_t)NOALLOC_SCRATCH_DOUBLES;
    return JSSO_OK;
}

int launch_gather_grad(Plan* p, const double* gbuf, int nrhs, double* dcrds, hipStream_t s) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    if (!dcrds || p->n_nodes == 0) return JSSO_OK;
    const int64_t total = (int64_t)p->n_nodes * 3 * nrhs;
    hipLaunchKernelGGL(gather_grad_kernel, xcd_grid((total + 255) / 256), dim3(256), 0, s,
                       p->n_nodes, nrhs, p->d_nptr, p->d_nsrc, gbuf, dcrds);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

}  // namespace jsso
