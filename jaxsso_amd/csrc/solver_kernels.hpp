// Kernels shared by the batched solvers: initialisation, the FP64 residual
// replacement, the per-system convergence decision, the best-iterate copy and
// the final unscaling u = L^-T x.
#pragma once
#include "solver_common.hpp"

namespace jsso {

// b = L^-1 f (constrained DOFs zeroed), x = 0, r_0 = b (buffer 0); partials <b,b>.
template <typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_init(Slices sl, Partials P, const uint8_t* __restrict__ fixed,
                                                    const double* __restrict__ Linv, const double* __restrict__ f,
                                                    double* __restrict__ b64, double* __restrict__ x64,
                                                    double* __restrict__ xbest, VT* __restrict__ xlo,
                                                    VT* __restrict__ r0, VT* __restrict__ p) {
    __shared__ double lds[NWAVES * NR];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    const bool on = R < 6 * (int64_t)sl.row1[w];
    double vb[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) vb[h] = 0.0;
    if (on) {
        const int64_t i = R / 6;
        const int r = (int)(R % 6);
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const int64_t dof = 6 * i + c;
                const double fv = fixed[dof] ? 0.0 : f[dof * NR + h];
                s += Linv[i * 36 + r * 6 + c] * fv;
            }
            const int64_t o = R * NR + h;
            b64[o] = s;
            x64[o] = 0.0;
            xbest[o] = 0.0;
            xlo[o] = VT(0);
            r0[o] = VT(s);
            p[o] = VT(0);
            vb[h] = s * s;
        }
    }
    write_partials<NR, NWAVES>(vb, lds, P.b, w);
}

// One wavefront per structure: convergence bookkeeping of its load cases from
// the slice partials, and the per-slice activity masks the CG kernels test.
// mode 0 (after k_init): bnorm2, activity.  mode 1 (after k_replace): true
// relative residual, best iterate, stagnation, status, after `steps` more
// iterations (counted per system, so the kernel can sit in a captured graph).
// A system whose true residual has not improved by 10% for max(min_window,
// it/2) iterations has reached its attainable accuracy and stops.
// Attainable-accuracy floor (rho given: an estimate of ||A||_2 per structure):
// the FP64 residual of any iterate x carries rounding errors of about
// eps (||A|| ||x|| + ||b||), so a system whose relative residual is within
// floor_factor of eps (rho ||x|| + ||b||) / ||b|| and has not improved by 10%
// for floor_window iterations stops there as SYS_FLOOR (converged as far as
// FP64 allows) instead of waiting out the long window; so does one whose best
// residual is in that band when the iteration diverges.
// Every lane computes the same (butterfly) totals; lane 0 writes the state.
template <int NR>
__global__ __launch_bounds__(64) void k_decide(Slices sl, Partials P, SysState st, int nsys_struct, int mode,
                                               double tol, int steps, int min_window, const double* __restrict__ rho,
                                               double floor_factor, int floor_window) {
    const int sys = blockIdx.x;
    if (sys >= nsys_struct) return;
    const int lane = threadIdx.x;
    const int s0 = sl.sys_off[sys], s1 = sl.sys_off[sys + 1];
    double ta[NR], tb[NR], tx[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) ta[h] = tb[h] = tx[h] = 0.0;
    for (int w = s0 + lane; w < s1; w += 64) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            if (mode == 0) {
                ta[h] += P.b[(int64_t)w * NR + h];
            } else {
                ta[h] += P.t[(int64_t)w * NR + h];
                tb[h] += P.rtrue[(int64_t)w * NR + h];
                if (rho) tx[h] += P.xx[(int64_t)w * NR + h];
            }
        }
    }
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        ta[h] = wave_sum(ta[h]);
        tb[h] = wave_sum(tb[h]);
        tx[h] = wave_sum(tx[h]);
    }
    int mask = 0;
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        const int k = sys * NR + h;
        if (mode == 0) {
            const double bb = ta[h];
            if (lane == 0) {
                st.bnorm2[k] = bb;
                st.relres[k] = bb > 0.0 ? 1.0 : 0.0;
                st.best[k] = bb > 0.0 ? 1.0 : 0.0;
                st.active[k] = bb > 0.0 ? 1 : 0;
                st.status[k] = bb > 0.0 ? SYS_RUNNING : SYS_CONVERGED;
                st.stall[k] = 0;
                st.save[k] = 0;
                st.iters[k] = 0;
            }
            mask |= bb > 0.0 ? 1 << h : 0;
            continue;
        }
        if (!st.active[k]) {
            if (lane == 0) st.save[k] = 0;
            continue;
        }
        const int it = st.iters[k] + steps;
        const double rel = sqrt(ta[h] / st.bnorm2[k]);
        const double best = st.best[k];
        const bool improved = rel < best;
        const double nbest = improved ? rel : best;
        const int stall = rel <= 0.9 * best ? it : st.stall[k];   // iteration of the last 10% gain
        // block-Jacobi CG on strongly heterogeneous systems (SIMP densities)
        // plateaus for a few thousand iterations before it moves on: only a
        // longer flat stretch is taken as the attainable-accuracy floor there
        const int window = it / 2 > min_window ? it / 2 : min_window;
        const double bn = sqrt(st.bnorm2[k]);
        const double floor_rel = rho ? 2.220446049250313e-16 * (rho[sys] * sqrt(tx[h]) + bn) / bn : 0.0;
        // the best iterate lies within the floor band (x of a diverging iterate is
        // still close to the best one while rel stays small, so ||x|| holds)
        const bool near_floor = rho && isfinite(rel) && rel < 1e-3 && nbest <= floor_factor * floor_rel;
        int active = 1, status = SYS_RUNNING;
        if (rel <= tol || tb[h] == 0.0) {
            active = 0;
            status = SYS_CONVERGED;
        } else if (near_floor && it - stall > floor_window) {
            active = 0;
            status = SYS_FLOOR;
        } else if (!isfinite(rel) || rel > 1e3 * nbest || it - stall > window) {
            // past the floor, the replaced residual (rounding noise) no longer fits
            // the recurrences and the iteration can jump away: the best iterate
            // is then as accurate as FP64 allows
            active = 0;
            status = near_floor ? SYS_FLOOR : SYS_STALLED;
        }
        if (lane == 0) {
            st.relres[k] = rel;
            st.save[k] = improved ? 1 : 0;
            st.best[k] = nbest;
            st.stall[k] = stall;
            st.active[k] = active;
            st.status[k] = status;
            st.iters[k] = it;
        }
        mask |= active << h;
    }
    for (int w = s0 + lane; w < s1; w += 64) st.slice_act[w] = mask;
}

// r64 = b - A64 x64; r_k = r64 (low precision, buffer k & 1); partials
// <r64,r64> (FP64, convergence) and <r_k,r_k>.  Same slice layout, wave split
// and straight-line front end as k_cgf.
template <typename VT, int NR, int WPS>
__global__ __launch_bounds__(64 * WPS) void k_replace(Slices sl, Partials P, SysState st,
                                                      const double* __restrict__ A, const double* __restrict__ b64,
                                                      const double* __restrict__ x64, VT* __restrict__ rcur) {
    __shared__ double red[WPS > 1 ? WPS - 1 : 1][6 * NR][64];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int4 ma = sl.meta[3 * w];
    const int4 mb = sl.meta[3 * w + 1];
    const int amask = st.slice_act[w];
    if ((amask | ((ma.x | ma.y | ma.z | ma.w | mb.x | mb.y | mb.z | mb.w) < 0)) == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)ma.y + lane;
    const bool on = i < ma.z;
    const int64_t ii = on ? i : (int64_t)ma.y;
    const int pup = mb.x, nup = mb.y, plo = mb.z, nlo = mb.w;
    const int nslot = nup + nlo;
    const double* Aw = A + 36 * (int64_t)ma.w;
    double xown[6][NR], bown[6][NR];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int64_t o = (6 * ii + c) * NR + h;
            xown[c][h] = x64[o];
            bown[c][h] = b64[o];
        }
    double acc[6][NR];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[c][h] = 0.0;
    // the next slot's column and value position are requested before the
    // current slot's products (a slot past the end reads the metadata word)
    const int32_t* dummy = (const int32_t*)&sl.meta[3 * w];
    auto cpos = [&](int jj) -> int64_t {
        return jj < nup ? (int64_t)pup + (int64_t)jj * ROWS_PER_WG + lane
                        : (int64_t)plo + (int64_t)(jj - nup) * ROWS_PER_WG + lane;
    };
    int64_t col = *(wave < nslot ? sl.ell_col + cpos(wave) : dummy);
    int vp = *(wave < nslot ? sl.ell_vpos + cpos(wave) : dummy);
    for (int j = wave; j < nslot; j += WPS) {
        const bool upper = j < nup;
        double xj[6][NR];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) xj[c][h] = x64[(6 * col + c) * NR + h];
        double av[36];
        if (upper)
            load_block(Aw + (int64_t)j * 36 * ROWS_PER_WG, lane, av);
        else
            load_block(A + (int64_t)(vp >> 6) * 36 * ROWS_PER_WG, vp & (ROWS_PER_WG - 1), av);
        const int jn = j + WPS;
        const int64_t coln = *(jn < nslot ? sl.ell_col + cpos(jn) : dummy);
        const int vpn = *(jn < nslot ? sl.ell_vpos + cpos(jn) : dummy);
        if (upper) {
#pragma unroll
            for (int e = 0; e < 36; ++e)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[e / 6][h] += av[e] * xj[e % 6][h];
        } else {
#pragma unroll
            for (int e = 0; e < 36; ++e)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[e % 6][h] += av[e] * xj[e / 6][h];
        }
        col = coln;
        vp = vpn;
    }
    if (WPS > 1) {
        if (wave > 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) red[wave - 1][c * NR + h][lane] = acc[c][h];
        }
        __syncthreads();
        if (wave > 0) return;
#pragma unroll
        for (int u = 0; u < WPS - 1; ++u)
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[c][h] += red[u][c * NR + h][lane];
    }
    double vt[NR], vr[NR], vx[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) vt[h] = vr[h] = vx[h] = 0.0;
    if (on) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) {
                if (!((amask >> h) & 1)) continue;
                const int64_t o = (6 * i + c) * NR + h;
                const double rv = bown[c][h] - (xown[c][h] + acc[c][h]);
                const VT rl = VT(rv);
                rcur[o] = rl;
                vt[h] += rv * rv;
                vr[h] += (double)rl * (double)rl;
                vx[h] += xown[c][h] * xown[c][h];
            }
    }
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        vt[h] = wave_sum(vt[h]);
        vr[h] = wave_sum(vr[h]);
        vx[h] = wave_sum(vx[h]);
    }
    if (lane == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            P.t[(int64_t)w * NR + h] = vt[h];
            P.rtrue[(int64_t)w * NR + h] = vr[h];
            P.xx[(int64_t)w * NR + h] = vx[h];
        }
    }
}

// keep the best iterate: xbest = x64 where the last replacement improved it
template <int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_save(Slices sl, const int32_t* __restrict__ save,
                                                    const double* __restrict__ x64, double* __restrict__ xbest) {
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int sys = sl.sys[w];
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    if (R >= 6 * (int64_t)sl.row1[w]) return;
#pragma unroll
    for (int h = 0; h < NR; ++h)
        if (save[sys * NR + h]) xbest[R * NR + h] = x64[R * NR + h];
}

// u = L^-T x
template <int NR>
__global__ __launch_bounds__(256) void k_unscale(int32_t n_nodes, const double* __restrict__ Linv,
                                                 const double* __restrict__ x64, double* __restrict__ u) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= 6 * (int64_t)n_nodes) return;
    const int64_t i = R / 6;
    const int r = (int)(R % 6);
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) s += Linv[i * 36 + c * 6 + r] * x64[(6 * i + c) * NR + h];
        u[R * NR + h] = s;
    }
}

}  // namespace jsso
