// vanka_device.h - the local 5x5 block of a cell and its solve, FP64.
//
// Unknowns of cell (i, j): (u_l, u_r, v_t, v_b, H_c) = (u[i,j], u[i,j+1],
// v[i,j], v[i+1,j], H[i,j]); neighbours are read from the level state and the
// cell viscosities are frozen at the sweep's input state.  The assembly
// follows glide/cuda/vanka.cu:56-411 (forward) and :485-780 (adjoint) term
// by term: the block keeps only the terms the reference keeps (the shear
// cross-couplings between u rows and v unknowns are not part of it).
#pragma once

#include "ssa_device.h"

namespace glide {

// 5x5 Doolittle LU without pivoting, then forward/back substitution
// (common.cu:132-175).  The reference builds this code with --use_fast_math
// (kernels.py:72), i.e. with contracted multiply-adds and reciprocal-based
// division; the FP64 form here fixes those choices explicitly: every
// update c - a*b is one fma, and back substitution multiplies by the pivot
// reciprocals of the elimination.  Each substitution step sums the terms of
// the later (earlier-known) unknowns first, so only one fma and one multiply
// wait on the unknown just found.
__host__ __device__ __forceinline__ void lu5_solve(const double *A, const double *b, double *x)
{
    double LU[5][5], inv[5];
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++) LU[i][j] = A[i * 5 + j];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        inv[k] = 1.0 / LU[k][k];
#pragma unroll
        for (int i = k + 1; i < 5; i++) {
            LU[i][k] *= inv[k];
#pragma unroll
            for (int j = k + 1; j < 5; j++) LU[i][j] = fma(-LU[i][k], LU[k][j], LU[i][j]);
        }
    }
    double y[5];
    y[0] = b[0];
    y[1] = fma(-LU[1][0], y[0], b[1]);
    y[2] = fma(-LU[2][1], y[1], fma(-LU[2][0], y[0], b[2]));
    y[3] = fma(-LU[3][2], y[2], fma(-LU[3][1], y[1], fma(-LU[3][0], y[0], b[3])));
    y[4] = fma(-LU[4][3], y[3], fma(-LU[4][2], y[2], fma(-LU[4][1], y[1], fma(-LU[4][0], y[0], b[4]))));
    x[4] = y[4] * inv[4];
    x[3] = fma(-LU[3][4], x[4], y[3]) * inv[3];
    x[2] = fma(-LU[2][3], x[3], fma(-LU[2][4], x[4], y[2])) * inv[2];
    x[1] = fma(-LU[1][2], x[2], fma(-LU[1][3], x[3], fma(-LU[1][4], x[4], y[1]))) * inv[1];
    x[0] = fma(-LU[0][1], x[1], fma(-LU[0][2], x[2], fma(-LU[0][3], x[3], fma(-LU[0][4], x[4], y[0])))) * inv[0];
}

// Neighbourhood of one cell that stays fixed during its Newton loop.
struct CellNbhd {
    // viscosities: centre, left, right, top, bottom, and the four diagonals
    double e_c, e_l, e_r, e_t, e_b, e_tl, e_tr, e_bl, e_br;
    // thickness of the 8 neighbours
    double H_l, H_r, H_t, H_b, H_tl, H_tr, H_bl, H_br;
    double bed_c, bed_l, bed_r, bed_t, bed_b;
    double beta_c, beta_l, beta_r, beta_t, beta_b;
    // neighbouring velocities
    double u_ll, u_rr, u_tl, u_tr, u_bl, u_br;   // u[i,j-1], u[i,j+2], u[i-1,j], u[i-1,j+1], u[i+1,j], u[i+1,j+1]
    double v_lt, v_lb, v_rt, v_rb, v_tt, v_bb;   // v[i,j-1], v[i+1,j-1], v[i,j+1], v[i+1,j+1], v[i-1,j], v[i+2,j]
    double f0, f1, f2, f3, f4;                   // right-hand side at the 5 dofs
    // grounding sigmoids of the 4 edge neighbours (fixed while the cell iterates)
    Ground g_l, g_r, g_t, g_b;
};

__host__ __device__ __forceinline__ void nbhd_grounding(CellNbhd &N, const Params &p)
{
    N.g_l = ground(N.bed_l, N.H_l, p);
    N.g_r = ground(N.bed_r, N.H_r, p);
    N.g_t = ground(N.bed_t, N.H_t, p);
    N.g_b = ground(N.bed_b, N.H_b, p);
}

// MODE 0: forward Vanka (residual, active-set test, diagonal shift)
// MODE 1: adjoint block (no residual, mass row replaced where adj_masked)
// ROWS: bit k set = assemble row k (J[5k..5k+4], r[k]); each entry receives
// exactly the operations it receives in the full assembly (ROWS = 31).
template <int MODE, int ROWS = 31>
__host__ __device__ __forceinline__ void vanka_assemble(const CellNbhd &N, const Params &p, int i, int j, int ny, int nx,
                                               double dx_inv, double dt, double u_l, double u_r, double v_t,
                                               double v_b, double H_c, double thklim, bool adj_masked,
                                               double J[25], double r[5])
{
#pragma unroll
    for (int k = 0; k < 25; k++) J[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) r[k] = 0.0;
    const Ground g_c = ground(N.bed_c, H_c, p);   // the only sigmoid that depends on the iterate
    if (MODE == 0) {
        if (ROWS & 1) r[0] -= N.f0;
        if (ROWS & 2) r[1] -= N.f1;
        if (ROWS & 4) r[2] -= N.f2;
        if (ROWS & 8) r[3] -= N.f3;
        if (ROWS & 16) r[4] -= N.f4;
    }
    {
        if (ROWS & 16) J[24] = 1.0 / dt;
        if (ROWS & 16) r[4] += H_c / dt;
        Calv jc = calving_g(H_c, g_c, p);
        if (ROWS & 16) J[24] -= jc.d_H;
        if (ROWS & 16) r[4] -= jc.res;
        Flux q = flux_x(u_l, N.H_l, H_c, j, nx);
        if (ROWS & 16) J[20] = fma(-q.d_u, dx_inv, J[20]);
        if (ROWS & 16) J[24] = fma(-q.d_H_r, dx_inv, J[24]);
        if (ROWS & 16) r[4] = fma(-q.res, dx_inv, r[4]);
        q = flux_x(u_r, H_c, N.H_r, j + 1, nx);
        if (ROWS & 16) J[21] = fma(q.d_u, dx_inv, J[21]);
        if (ROWS & 16) J[24] = fma(q.d_H_l, dx_inv, J[24]);
        if (ROWS & 16) r[4] = fma(q.res, dx_inv, r[4]);
        q = flux_y(v_t, N.H_t, H_c, i, ny);
        if (ROWS & 16) J[22] = fma(q.d_u, dx_inv, J[22]);
        if (ROWS & 16) J[24] = fma(q.d_H_r, dx_inv, J[24]);
        if (ROWS & 16) r[4] = fma(q.res, dx_inv, r[4]);
        q = flux_y(v_b, H_c, N.H_b, i + 1, ny);
        if (ROWS & 16) J[23] = fma(-q.d_u, dx_inv, J[23]);
        if (ROWS & 16) J[24] = fma(-q.d_H_l, dx_inv, J[24]);
        if (ROWS & 16) r[4] = fma(-q.res, dx_inv, r[4]);
        bool constrained = (ROWS & 16) && ((MODE == 0) ? ((H_c - dt * r[4]) <= thklim) : adj_masked);
        if ((ROWS & 16) && constrained) {
#pragma unroll
            for (int k = 0; k < 5; k++) J[20 + k] = 0.0;
            if (ROWS & 16) J[24] = 1.0;
            if (MODE == 0) if (ROWS & 16) r[4] = H_c - thklim;
        }
    }
    {
        double eH = H_c * N.e_c;
        SigN s = sigma_xx(u_l, u_r, v_t, v_b, eH, dx_inv, j, nx);
        if (ROWS & 1) r[0] = fma(s.res, dx_inv, r[0]);
        if (ROWS & 1) J[0] = fma(s.d_u_l, dx_inv, J[0]); if (ROWS & 1) J[1] = fma(s.d_u_r, dx_inv, J[1]); if (ROWS & 1) J[2] = fma(s.d_v_t, dx_inv, J[2]);
        if (ROWS & 1) J[3] = fma(s.d_v_b, dx_inv, J[3]); if (ROWS & 1) J[4] = fma(s.d_eta_H * N.e_c, dx_inv, J[4]);
        if (ROWS & 2) r[1] = fma(-s.res, dx_inv, r[1]);
        if (ROWS & 2) J[5] = fma(-s.d_u_l, dx_inv, J[5]); if (ROWS & 2) J[6] = fma(-s.d_u_r, dx_inv, J[6]); if (ROWS & 2) J[7] = fma(-s.d_v_t, dx_inv, J[7]);
        if (ROWS & 2) J[8] = fma(-s.d_v_b, dx_inv, J[8]); if (ROWS & 2) J[9] = fma(-(s.d_eta_H * N.e_c), dx_inv, J[9]);
        SigN t = sigma_yy(u_l, u_r, v_t, v_b, eH, dx_inv, i, ny);
        if (ROWS & 4) r[2] = fma(-t.res, dx_inv, r[2]);
        if (ROWS & 4) J[10] = fma(-t.d_u_l, dx_inv, J[10]); if (ROWS & 4) J[11] = fma(-t.d_u_r, dx_inv, J[11]); if (ROWS & 4) J[12] = fma(-t.d_v_t, dx_inv, J[12]);
        if (ROWS & 4) J[13] = fma(-t.d_v_b, dx_inv, J[13]); if (ROWS & 4) J[14] = fma(-(t.d_eta_H * N.e_c), dx_inv, J[14]);
        if (ROWS & 8) r[3] = fma(t.res, dx_inv, r[3]);
        if (ROWS & 8) J[15] = fma(t.d_u_l, dx_inv, J[15]); if (ROWS & 8) J[16] = fma(t.d_u_r, dx_inv, J[16]); if (ROWS & 8) J[17] = fma(t.d_v_t, dx_inv, J[17]);
        if (ROWS & 8) J[18] = fma(t.d_v_b, dx_inv, J[18]); if (ROWS & 8) J[19] = fma(t.d_eta_H * N.e_c, dx_inv, J[19]);
    }
    {
        SigN s = sigma_xx(N.u_ll, u_l, N.v_lt, N.v_lb, N.e_l * N.H_l, dx_inv, j - 1, nx);
        if (ROWS & 1) r[0] = fma(-s.res, dx_inv, r[0]);
        if (ROWS & 1) J[0] = fma(-s.d_u_r, dx_inv, J[0]);
    }
    {
        SigN s = sigma_xx(u_r, N.u_rr, N.v_rt, N.v_rb, N.e_r * N.H_r, dx_inv, j + 1, nx);
        if (ROWS & 2) r[1] = fma(s.res, dx_inv, r[1]);
        if (ROWS & 2) J[6] = fma(s.d_u_l, dx_inv, J[6]);
    }
    {
        SigN s = sigma_yy(N.u_tl, N.u_tr, N.v_tt, v_t, N.e_t * N.H_t, dx_inv, i - 1, ny);
        if (ROWS & 4) r[2] = fma(s.res, dx_inv, r[2]);
        if (ROWS & 4) J[12] = fma(s.d_v_b, dx_inv, J[12]);
    }
    {
        SigN s = sigma_yy(N.u_bl, N.u_br, v_b, N.v_bb, N.e_b * N.H_b, dx_inv, i + 1, ny);
        if (ROWS & 8) r[3] = fma(-s.res, dx_inv, r[3]);
        if (ROWS & 8) J[18] = fma(-s.d_v_t, dx_inv, J[18]);
    }
    {   // top-left vertex (i, j)
        double eH = etaH_vertex(N.e_tl, N.e_t, N.e_l, N.e_c, N.H_tl, N.H_t, N.H_l, H_c);
        double dH = 0.25 * N.e_c;
        SigS s = sigma_xy(N.u_tl, u_l, N.v_lt, v_t, eH, dx_inv, i, j, ny, nx);
        if (ROWS & 1) r[0] = fma(s.res, dx_inv, r[0]);
        if (ROWS & 1) J[0] = fma(s.d_u_b, dx_inv, J[0]);
        if (ROWS & 1) J[4] = fma(s.d_eta_H * dH, dx_inv, J[4]);
        if (ROWS & 4) r[2] = fma(-s.res, dx_inv, r[2]);
        if (ROWS & 4) J[12] = fma(-s.d_v_r, dx_inv, J[12]);
        if (ROWS & 4) J[14] = fma(-(s.d_eta_H * dH), dx_inv, J[14]);
    }
    {   // top-right vertex (i, j+1)
        double eH = etaH_vertex(N.e_t, N.e_tr, N.e_c, N.e_r, N.H_t, N.H_tr, H_c, N.H_r);
        double dH = 0.25 * N.e_c;
        SigS s = sigma_xy(N.u_tr, u_r, v_t, N.v_rt, eH, dx_inv, i, j + 1, ny, nx);
        if (ROWS & 2) r[1] = fma(s.res, dx_inv, r[1]);
        if (ROWS & 2) J[6] = fma(s.d_u_b, dx_inv, J[6]);
        if (ROWS & 2) J[9] = fma(s.d_eta_H * dH, dx_inv, J[9]);
        if (ROWS & 4) r[2] = fma(s.res, dx_inv, r[2]);
        if (ROWS & 4) J[12] = fma(s.d_v_l, dx_inv, J[12]);
        if (ROWS & 4) J[14] = fma(s.d_eta_H * dH, dx_inv, J[14]);
    }
    {   // bottom-left vertex (i+1, j)
        double eH = etaH_vertex(N.e_l, N.e_c, N.e_bl, N.e_b, N.H_l, H_c, N.H_bl, N.H_b);
        double dH = 0.25 * N.e_c;
        SigS s = sigma_xy(u_l, N.u_bl, N.v_lb, v_b, eH, dx_inv, i + 1, j, ny, nx);
        if (ROWS & 1) r[0] = fma(-s.res, dx_inv, r[0]);
        if (ROWS & 1) J[0] = fma(-s.d_u_t, dx_inv, J[0]);
        if (ROWS & 1) J[4] = fma(-(s.d_eta_H * dH), dx_inv, J[4]);
        if (ROWS & 8) r[3] = fma(-s.res, dx_inv, r[3]);
        if (ROWS & 8) J[18] = fma(-s.d_v_r, dx_inv, J[18]);
        if (ROWS & 8) J[19] = fma(-(s.d_eta_H * dH), dx_inv, J[19]);
    }
    {   // bottom-right vertex (i+1, j+1)
        double eH = etaH_vertex(N.e_c, N.e_r, N.e_b, N.e_br, H_c, N.H_r, N.H_b, N.H_br);
        double dH = 0.25 * N.e_c;
        SigS s = sigma_xy(u_r, N.u_br, v_b, N.v_rb, eH, dx_inv, i + 1, j + 1, ny, nx);
        if (ROWS & 2) r[1] = fma(-s.res, dx_inv, r[1]);
        if (ROWS & 2) J[6] = fma(-s.d_u_t, dx_inv, J[6]);
        if (ROWS & 2) J[9] = fma(-(s.d_eta_H * dH), dx_inv, J[9]);
        if (ROWS & 8) r[3] = fma(s.res, dx_inv, r[3]);
        if (ROWS & 8) J[18] = fma(s.d_v_l, dx_inv, J[18]);
        if (ROWS & 8) J[19] = fma(s.d_eta_H * dH, dx_inv, J[19]);
    }
    {
        TauB t = tau_b_g(u_l, N.g_l.g, g_c.g, N.g_l.gd, g_c.gd, N.beta_l, N.beta_c, p);
        if (ROWS & 1) r[0] += t.res; if (ROWS & 1) J[0] += t.d_u; if (ROWS & 1) J[4] += t.d_H_r;
    }
    {
        TauB t = tau_b_g(u_r, g_c.g, N.g_r.g, g_c.gd, N.g_r.gd, N.beta_c, N.beta_r, p);
        if (ROWS & 2) r[1] += t.res; if (ROWS & 2) J[6] += t.d_u; if (ROWS & 2) J[9] += t.d_H_l;
    }
    {
        TauB t = tau_b_g(v_t, N.g_t.g, g_c.g, N.g_t.gd, g_c.gd, N.beta_t, N.beta_c, p);
        if (ROWS & 4) r[2] += t.res; if (ROWS & 4) J[12] += t.d_u; if (ROWS & 4) J[14] += t.d_H_r;
    }
    {
        TauB t = tau_b_g(v_b, g_c.g, N.g_b.g, g_c.gd, N.g_b.gd, N.beta_c, N.beta_b, p);
        if (ROWS & 8) r[3] += t.res; if (ROWS & 8) J[18] += t.d_u; if (ROWS & 8) J[19] += t.d_H_l;
    }
    {
        TauD t = tau_dx(N.H_l, H_c, N.bed_l, N.bed_c, dx_inv, j, nx);
        if (ROWS & 1) r[0] -= t.res; if (ROWS & 1) J[4] -= t.d_H_r;
    }
    {
        TauD t = tau_dx(H_c, N.H_r, N.bed_c, N.bed_r, dx_inv, j + 1, nx);
        if (ROWS & 2) r[1] -= t.res; if (ROWS & 2) J[9] -= t.d_H_l;
    }
    {
        TauD t = tau_dy(N.H_t, H_c, N.bed_t, N.bed_c, dx_inv, i, ny);
        if (ROWS & 4) r[2] -= t.res; if (ROWS & 4) J[14] -= t.d_H_r;
    }
    {
        TauD t = tau_dy(H_c, N.H_b, N.bed_c, N.bed_b, dx_inv, i + 1, ny);
        if (ROWS & 8) r[3] -= t.res; if (ROWS & 8) J[19] -= t.d_H_l;
    }
    if (MODE == 0) {
        if (ROWS & 1) J[0] -= 1.0; if (ROWS & 2) J[6] -= 1.0; if (ROWS & 4) J[12] -= 1.0; if (ROWS & 8) J[18] -= 1.0; if (ROWS & 16) J[24] += 1.0;
    }
}

}  // namespace glide
