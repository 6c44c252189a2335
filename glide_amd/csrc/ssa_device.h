// ssa_device.h - pointwise SSA + mass-conservation physics on the MAC grid,
// FP64, for gfx950.
//
// Each function evaluates one local term (a stress, a drag, a flux) with its
// Jacobian row, with the reference's algebra and operation order
// (glide/cuda/stress.cu, flux.cu, viscosity.cu, common.cu).  The kernels
// compose them; out-of-domain terms return all-zero rows exactly where the
// reference's do.  Compiled with -ffp-contract=off, so every product and sum
// rounds as written except where an explicit fma() fuses them: at the
// multiply-adds that the reference's --use_fast_math build contracts
// (kernels.py:72), written out so that the host oracle rounds identically
// (the dual-number product's fma is explicit in the reference, common.cu:20).
#pragma once

#include <hip/hip_runtime.h>

#include "detmath.h"
#include "glide_hip.h"

namespace glide {

struct Params {
    double n, eps_reg, water_drag, calving_rate, gl_c;
    int gl_derivatives;
};

__host__ __device__ inline Params to_params(const GlideParams &p)
{
    return Params{p.n, p.eps_reg, p.water_drag, p.calving_rate, p.gl_sigmoid_c, p.gl_derivatives};
}

__host__ __device__ __forceinline__ int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

// Clamped MAC accessors (common.cu:78-125).  u: ny x (nx+1), v: (ny+1) x nx, cells: ny x nx.
__host__ __device__ __forceinline__ double ld_u(const double *__restrict__ u, int i, int j, int ny, int nx)
{
    return u[(size_t)clampi(i, 0, ny - 1) * (nx + 1) + clampi(j, 0, nx)];
}
__host__ __device__ __forceinline__ double ld_v(const double *__restrict__ v, int i, int j, int ny, int nx)
{
    return v[(size_t)clampi(i, 0, ny) * nx + clampi(j, 0, nx - 1)];
}
__host__ __device__ __forceinline__ double ld_c(const double *__restrict__ a, int i, int j, int ny, int nx)
{
    return a[(size_t)clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1)];
}

// Grounding-line sigmoid and its derivative (common.cu:67-76)
__host__ __device__ __forceinline__ double sigm(double z, double c)
{
    double s = fmin(fmax(c * z, -20.0), 20.0);
    return 1.0 / (1.0 + dm_exp(-s));
}
__host__ __device__ __forceinline__ double sigm_d(double z, double c)
{
    double s = sigm(z, c);
    return c * s * (1.0 - s);
}

struct Dual {
    double v, d;
};
__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, fma(a.v, b.d, a.d * b.v)}; }
__host__ __device__ __forceinline__ Dual operator*(Dual a, double s) { return {a.v * s, a.d * s}; }
__host__ __device__ __forceinline__ Dual operator+(Dual a, double s) { return {a.v + s, a.d}; }

// Cell viscosity eta = 1/2 B (eps_II^2 + eps_reg)^((1-n)/2n) with the four
// corner shear rates masked at the domain edge (viscosity.cu:61-112).  U(a, b)
// and V(a, b) return the clamped u[a, b] and v[a, b]; Bc is B at the cell.
template <class UA, class VA>
__host__ __device__ __forceinline__ double eta_cell_acc(int i, int j, const UA &U, const VA &V, double Bc,
                                                        const Params &p, double dx, int ny, int nx)
{
    double dx_inv = 1.0 / dx;
    double glen_exp = (1.0 - p.n) / (2.0 * p.n);
    double u_l = U(i, j), u_r = U(i, j + 1);
    double v_t = V(i, j), v_b = V(i + 1, j);
    double dudx = (u_r - u_l) * dx_inv;
    double dvdy = (v_t - v_b) * dx_inv;
    double tl_mask = (i > 0 && j > 0);
    double e_tl = 0.5 * fma(U(i - 1, j) - u_l, dx_inv, (v_t - V(i, j - 1)) * dx_inv) * tl_mask;
    double tr_mask = (i > 0 && j < (nx - 1));
    double e_tr = 0.5 * fma(U(i - 1, j + 1) - u_r, dx_inv, (V(i, j + 1) - v_t) * dx_inv) * tr_mask;
    double bl_mask = (i < (ny - 1) && j > 0);
    double e_bl = 0.5 * fma(u_l - U(i + 1, j), dx_inv, (v_b - V(i + 1, j - 1)) * dx_inv) * bl_mask;
    double br_mask = (i < (ny - 1) && j < (nx - 1));
    double e_br = 0.5 * fma(u_r - U(i + 1, j + 1), dx_inv, (V(i + 1, j + 1) - v_b) * dx_inv) * br_mask;
    double exy2 = 0.25 * fma(e_br, e_br, fma(e_bl, e_bl, fma(e_tr, e_tr, e_tl * e_tl)));
    double eII = fma(dudx, dvdy, fma(dvdy, dvdy, dudx * dudx)) + exy2 + p.eps_reg;
    return 0.5 * Bc * dm_pow(eII, glen_exp);
}
__host__ __device__ __forceinline__ double eta_cell(int i, int j, const double *__restrict__ u,
                                           const double *__restrict__ v, const double *__restrict__ B,
                                           const Params &p, double dx, int ny, int nx)
{
    auto U = [&](int a, int b) { return ld_u(u, a, b, ny, nx); };
    auto V = [&](int a, int b) { return ld_v(v, a, b, ny, nx); };
    return eta_cell_acc(i, j, U, V, ld_c(B, i, j, ny, nx), p, dx, ny, nx);
}

// Dual-number viscosity: value and directional derivative along (du, dv)
// (viscosity.cu:4-59).  U(a, b) and V(a, b) return the clamped Dual
// (u[a, b], du[a, b]) and (v[a, b], dv[a, b]); Bc is B at the cell.
template <class UA, class VA>
__host__ __device__ __forceinline__ Dual eta_cell_dual_acc(int i, int j, const UA &U, const VA &V, double Bc,
                                                           const Params &p, double dx, int ny, int nx)
{
    double dx_inv = 1.0 / dx;
    double glen_exp = (1.0 - p.n) / (2.0 * p.n);
    Dual u_l = U(i, j), u_r = U(i, j + 1), v_t = V(i, j), v_b = V(i + 1, j);
    Dual dudx = (u_r - u_l) * dx_inv;
    Dual dvdy = (v_t - v_b) * dx_inv;
    double tl_mask = (i > 0 && j > 0);
    Dual e_tl = (((U(i - 1, j) - u_l) * dx_inv + (v_t - V(i, j - 1)) * dx_inv) * 0.5) * tl_mask;
    double tr_mask = (i > 0 && j < (nx - 1));
    Dual e_tr = (((U(i - 1, j + 1) - u_r) * dx_inv + (V(i, j + 1) - v_t) * dx_inv) * 0.5) * tr_mask;
    double bl_mask = (i < (ny - 1) && j > 0);
    Dual e_bl = (((u_l - U(i + 1, j)) * dx_inv + (v_b - V(i + 1, j - 1)) * dx_inv) * 0.5) * bl_mask;
    double br_mask = (i < (ny - 1) && j < (nx - 1));
    Dual e_br = (((u_r - U(i + 1, j + 1)) * dx_inv + (V(i + 1, j + 1) - v_b) * dx_inv) * 0.5) * br_mask;
    Dual exy2 = (e_tl * e_tl + e_tr * e_tr + e_bl * e_bl + e_br * e_br) * 0.25;
    Dual eII = dudx * dudx + dvdy * dvdy + dudx * dvdy + exy2 + p.eps_reg;
    double s = 0.5 * Bc;
    // x^p and x^(p-1) from one logarithm: dm_pow(x, q) = dm_exp(q * dm_log(x)),
    // so the values are those of two dm_pow calls
    Dual pw;
    if (eII.v == 0.0) {   // dm_pow's own value at 0 (only reachable with eps_reg = 0)
        pw.v = dm_pow(eII.v, glen_exp);
        pw.d = glen_exp * dm_pow(eII.v, glen_exp - 1.0) * eII.d;
    } else {
        const double lg = dm_log(eII.v);
        pw.v = dm_exp(glen_exp * lg);
        pw.d = glen_exp * dm_exp((glen_exp - 1.0) * lg) * eII.d;
    }
    return pw * s;
}
__host__ __device__ __forceinline__ Dual eta_cell_dual(int i, int j, const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ du,
                                              const double *__restrict__ dv, const double *__restrict__ B,
                                              const Params &p, double dx, int ny, int nx)
{
    auto U = [&](int a, int b) { return Dual{ld_u(u, a, b, ny, nx), ld_u(du, a, b, ny, nx)}; };
    auto V = [&](int a, int b) { return Dual{ld_v(v, a, b, ny, nx), ld_v(dv, a, b, ny, nx)}; };
    return eta_cell_dual_acc(i, j, U, V, ld_c(B, i, j, ny, nx), p, dx, ny, nx);
}

// eta*H at a vertex from its four cells (viscosity.cu:201-217)
__host__ __device__ __forceinline__ double etaH_vertex(double etl, double etr, double ebl, double ebr, double Htl,
                                              double Htr, double Hbl, double Hbr)
{
    return 0.25 * (etl * Htl + etr * Htr + ebl * Hbl + ebr * Hbr);
}

struct SigN { double res, d_u_l, d_u_r, d_v_t, d_v_b, d_eta_H; };
struct SigS { double res, d_u_t, d_u_b, d_v_l, d_v_r, d_eta_H; };
struct TauB { double res, d_u, d_H_l, d_H_r, d_beta_l, d_beta_r; };
struct TauD { double res, d_H_l, d_H_r; };
struct Flux { double res, d_u, d_H_l, d_H_r; };
struct Calv { double res, d_H; };

// sigma_xx at cell (i, j) (stress.cu:44-66)
__host__ __device__ __forceinline__ SigN sigma_xx(double u_l, double u_r, double v_t, double v_b, double eta_H,
                                         double dx_inv, int j, int nx)
{
    SigN J = {0, 0, 0, 0, 0, 0};
    if (j < 0 || j >= nx) return J;
    double eps_xx = fma(2.0 * (u_r - u_l), dx_inv, (v_t - v_b) * dx_inv);
    double pre = 2.0 * eta_H * dx_inv;
    J.res = 2.0 * eta_H * eps_xx;
    J.d_u_l = -2.0 * pre;
    J.d_u_r = 2.0 * pre;
    J.d_v_t = pre;
    J.d_v_b = -pre;
    J.d_eta_H = 2.0 * eps_xx;
    return J;
}
// sigma_yy at cell (i, j) (stress.cu:79-102)
__host__ __device__ __forceinline__ SigN sigma_yy(double u_l, double u_r, double v_t, double v_b, double eta_H,
                                         double dx_inv, int i, int ny)
{
    SigN J = {0, 0, 0, 0, 0, 0};
    if (i < 0 || i >= ny) return J;
    double eps_yy = fma(u_r - u_l, dx_inv, 2.0 * (v_t - v_b) * dx_inv);
    double pre = 2.0 * eta_H * dx_inv;
    J.res = 2.0 * eta_H * eps_yy;
    J.d_u_l = -pre;
    J.d_u_r = pre;
    J.d_v_t = 2.0 * pre;
    J.d_v_b = -2.0 * pre;
    J.d_eta_H = 2.0 * eps_yy;
    return J;
}
// sigma_xy at vertex (i, j), zero on boundary vertices (stress.cu:160-183)
__host__ __device__ __forceinline__ SigS sigma_xy(double u_t, double u_b, double v_l, double v_r, double eta_H,
                                         double dx_inv, int i, int j, int ny, int nx)
{
    SigS J = {0, 0, 0, 0, 0, 0};
    if (i <= 0 || i >= ny || j <= 0 || j >= nx) return J;
    double eps_xy = 0.5 * fma(u_t - u_b, dx_inv, (v_r - v_l) * dx_inv);
    double pre = eta_H * dx_inv;
    J.res = 2.0 * eta_H * eps_xy;
    J.d_u_t = pre;
    J.d_u_b = -pre;
    J.d_v_l = -pre;
    J.d_v_r = pre;
    J.d_eta_H = 2.0 * eps_xy;
    return J;
}
// Basal drag on a face between cells l and r (or t and b), given the two
// cells' grounding sigmoids g = sigm(bed + 0.917 H) and, with gl_derivatives,
// their derivatives sigm_d (stress.cu:245-361).
__host__ __device__ __forceinline__ TauB tau_b_g(double vel, double g_l, double g_r, double gd_l, double gd_r,
                                                 double beta_l, double beta_r, const Params &p)
{
    TauB J = {0, 0, 0, 0, 0, 0};
    double be_l = fma(g_l, beta_l, (1.0 - g_l) * p.water_drag);
    double be_r = fma(g_r, beta_r, (1.0 - g_r) * p.water_drag);
    double be = 0.5 * (be_l + be_r);
    J.res = -be * vel;
    J.d_u = -be;
    J.d_beta_l = -0.5 * g_l * vel;
    J.d_beta_r = -0.5 * g_r * vel;
    if (p.gl_derivatives) {
        double dg_l = 0.917 * gd_l;
        double dg_r = 0.917 * gd_r;
        double dbe_l = dg_l * (beta_l - p.water_drag);
        double dbe_r = dg_r * (beta_r - p.water_drag);
        J.d_H_l = -0.5 * dbe_l * vel;
        J.d_H_r = -0.5 * dbe_r * vel;
    }
    return J;
}
// grounding sigmoid of one cell and its derivative (0 without gl_derivatives)
struct Ground { double g, gd; };
__host__ __device__ __forceinline__ Ground ground(double bed, double H, const Params &p)
{
    double z = fma(0.917, H, bed);
    Ground o;
    o.g = sigm(z, p.gl_c);
    o.gd = p.gl_derivatives ? sigm_d(z, p.gl_c) : 0.0;
    return o;
}
__host__ __device__ __forceinline__ TauB tau_b(double vel, double H_l, double H_r, double bed_l, double bed_r,
                                               double beta_l, double beta_r, const Params &p)
{
    Ground a = ground(bed_l, H_l, p), b = ground(bed_r, H_r, p);
    return tau_b_g(vel, a.g, b.g, a.gd, b.gd, beta_l, beta_r, p);
}
// Driving stress on u face column j (stress.cu:407-436)
__host__ __device__ __forceinline__ TauD tau_dx(double H_l, double H_r, double bed_l, double bed_r, double dx_inv,
                                       int j, int nx)
{
    TauD J = {0, 0, 0};
    if (j <= 0 || j >= nx) return J;
    double H_avg = 0.5 * (H_l + H_r);
    double base_l = fmax(bed_l, -0.917 * H_l);
    double base_r = fmax(bed_r, -0.917 * H_r);
    double db_l = bed_l > -0.917 * H_l ? 0.0 : -0.917;
    double db_r = bed_r > -0.917 * H_r ? 0.0 : -0.917;
    double S_l = base_l + H_l;
    double S_r = base_r + H_r;
    J.res = H_avg * (S_r - S_l) * dx_inv;
    J.d_H_l = fma(-(H_avg * (1.0 + db_l)), dx_inv, 0.5 * (S_r - S_l) * dx_inv);
    J.d_H_r = fma(H_avg * (1.0 + db_r), dx_inv, 0.5 * (S_r - S_l) * dx_inv);
    return J;
}
// Driving stress on v face row i, cells t (above) and b (below); d_H_l = d/dH_t
// (stress.cu:482-510)
__host__ __device__ __forceinline__ TauD tau_dy(double H_t, double H_b, double bed_t, double bed_b, double dx_inv,
                                       int i, int ny)
{
    TauD J = {0, 0, 0};
    if (i <= 0 || i >= ny) return J;
    double H_avg = 0.5 * (H_t + H_b);
    double base_t = fmax(bed_t, -0.917 * H_t);
    double base_b = fmax(bed_b, -0.917 * H_b);
    double db_t = bed_t > -0.917 * H_t ? 0.0 : -0.917;
    double db_b = bed_b > -0.917 * H_b ? 0.0 : -0.917;
    double S_t = base_t + H_t;
    double S_b = base_b + H_b;
    J.res = H_avg * (S_t - S_b) * dx_inv;
    J.d_H_l = fma(H_avg * (1.0 + db_t), dx_inv, 0.5 * (S_t - S_b) * dx_inv);
    J.d_H_r = fma(-(H_avg * (1.0 + db_b)), dx_inv, 0.5 * (S_t - S_b) * dx_inv);
    return J;
}
// Upwind mass flux through u face column j (flux.cu:40-64)
__host__ __device__ __forceinline__ Flux flux_x(double u, double H_l, double H_r, int j, int nx)
{
    Flux J = {0, 0, 0, 0};
    if (j <= 0 || j >= nx) return J;
    double H_avg = 0.5 * (H_l + H_r);
    double um = fabs(u);
    double us = copysign(1.0, u);
    J.res = fma(H_avg, u, -(0.5 * um * (H_r - H_l)));
    J.d_H_l = 0.5 * (u + um);
    J.d_H_r = 0.5 * (u - um);
    J.d_u = fma(-(0.5 * us), H_r - H_l, H_avg);
    return J;
}
// Upwind mass flux through v face row i; d_H_l = d/dH_t, d_H_r = d/dH_b (flux.cu:111-135)
__host__ __device__ __forceinline__ Flux flux_y(double v, double H_t, double H_b, int i, int ny)
{
    Flux J = {0, 0, 0, 0};
    if (i <= 0 || i >= ny) return J;
    double H_avg = 0.5 * (H_t + H_b);
    double vm = fabs(v);
    double vs = copysign(1.0, v);
    J.res = fma(H_avg, v, -(0.5 * vm * (H_t - H_b)));
    J.d_H_l = 0.5 * (v - vm);
    J.d_H_r = 0.5 * (v + vm);
    J.d_u = fma(-(0.5 * vs), H_t - H_b, H_avg);
    return J;
}
// Calving sink of floating ice (flux.cu:188-211), given the cell's sigmoid
__host__ __device__ __forceinline__ Calv calving_g(double H, const Ground &gr, const Params &p)
{
    Calv J;
    J.res = -p.calving_rate * (1.0 - gr.g) * H;
    J.d_H = -p.calving_rate * (1.0 - gr.g);
    if (p.gl_derivatives) {
        double dg = 0.917 * gr.gd;
        J.d_H = fma(p.calving_rate * H, dg, J.d_H);
    }
    return J;
}
__host__ __device__ __forceinline__ Calv calving(double H, double bed, const Params &p)
{
    return calving_g(H, ground(bed, H, p), p);
}

__host__ __device__ __forceinline__ double sigN_jvp(const SigN &J, double ul, double ur, double vt, double vb, double eH)
{
    return J.d_u_l * ul + J.d_u_r * ur + J.d_v_t * vt + J.d_v_b * vb + J.d_eta_H * eH;
}
__host__ __device__ __forceinline__ double sigS_jvp(const SigS &J, double ut, double ub, double vl, double vr, double eH)
{
    return J.d_u_t * ut + J.d_u_b * ub + J.d_v_l * vl + J.d_v_r * vr + J.d_eta_H * eH;
}

}  // namespace glide
