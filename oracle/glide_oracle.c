/*
 * ORACLE - TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of GLIDE's SSA / FASCD solver.  Only tests/, the
 * smoke() check and bench.py's cpu_baseline leg may load this library, and
 * only as the reference results (or as the timed CPU baseline).  The product path
 * (glide_amd/) never links, imports or calls it.
 *
 * Every function follows the reference's algorithm in FP64 with the
 * reference's operation order (the reference computes in FP32 with
 * --use_fast_math, glide/kernels.py:72).  Built with -ffp-contract=off so no
 * multiply-add is fused behind the source's back.  Where the fast-math build
 * contracts a product and a sum (the LU updates, the Horner steps of exp and
 * log, the Vanka block's accumulations, the stress, drag, flux and viscosity
 * sums), this restatement writes an explicit fma(), at the same places as the
 * device code, so both round identically; the dual-number product's fma() is
 * explicit in the reference too (common.cu:20).
 *
 * Transcendentals: by default the deterministic exp/log/pow of
 * glide_amd/csrc/detmath.h (same bits on host and gfx950); building with
 * -DORACLE_LIBM switches to glibc exp/pow (used by a test that shows the
 * choice of exp/pow moves a forward step by round-off only).
 *
 * Parity pins: see tests/test_oracle_*.py (finite-difference Jacobians,
 * dot-product test of JVP/VJP, closed forms of the transfer operators, LU vs
 * numpy.linalg.solve, FAS fixed-point, flux telescoping, ...).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_LIBM
#define OEXP(x) exp(x)
#define OPOW(x, p) pow(x, p)
#else
#include "../glide_amd/csrc/detmath.h"
#define OEXP(x) dm_exp(x)
#define OPOW(x, p) dm_pow(x, p)
#endif

/* ------------------------------------------------------------------ */
/* Parameters (common.cu:240-247, kernels.py:10-49)                    */
/* ------------------------------------------------------------------ */
typedef struct {
    double n, eps_reg, water_drag, calving_rate, gl_sigmoid_c;
    int gl_derivatives;
} OParams;

typedef struct { int ny, nx; double dx, dt; } OGeom;

/* ------------------------------------------------------------------ */
/* Clamped MAC accessors (common.cu:78-125): these define the ghosts   */
/* ------------------------------------------------------------------ */
static inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }
static inline double VF(const double *u, int i, int j, int ny, int nx)
{ i = clampi(i, 0, ny - 1); j = clampi(j, 0, nx); return u[(size_t)i * (nx + 1) + j]; }
static inline double HF(const double *v, int i, int j, int ny, int nx)
{ i = clampi(i, 0, ny); j = clampi(j, 0, nx - 1); return v[(size_t)i * nx + j]; }
static inline double CE(const double *a, int i, int j, int ny, int nx)
{ i = clampi(i, 0, ny - 1); j = clampi(j, 0, nx - 1); return a[(size_t)i * nx + j]; }

/* Out-of-domain viscosity slots only ever meet zeroed Jacobian rows in the
 * reference (stress.cu:52,88,168 return before using them); 0 keeps them finite. */
static inline double ETA(const double *e, int i, int j, int ny, int nx)
{ return (i < 0 || i >= ny || j < 0 || j >= nx) ? 0.0 : e[(size_t)i * nx + j]; }

/* ------------------------------------------------------------------ */
/* Dual numbers (common.cu:4-65)                                       */
/* ------------------------------------------------------------------ */
typedef struct { double v, d; } Dual;
static inline Dual dadd(Dual a, Dual b) { Dual r = {a.v + b.v, a.d + b.d}; return r; }
static inline Dual dsub(Dual a, Dual b) { Dual r = {a.v - b.v, a.d - b.d}; return r; }
static inline Dual dmul(Dual a, Dual b) { Dual r = {a.v * b.v, fma(a.v, b.d, a.d * b.v)}; return r; }
static inline Dual dscale(Dual a, double s) { Dual r = {a.v * s, a.d * s}; return r; }
static inline Dual daddc(Dual a, double s) { Dual r = {a.v + s, a.d}; return r; }
static inline Dual dpow(Dual u, double p)
{
    Dual r;
    r.v = OPOW(u.v, p);
    r.d = p * OPOW(u.v, p - 1.0) * u.d;
    return r;
}
static inline Dual VFd(const double *u, const double *du, int i, int j, int ny, int nx)
{ i = clampi(i, 0, ny - 1); j = clampi(j, 0, nx); size_t k = (size_t)i * (nx + 1) + j; Dual r = {u[k], du[k]}; return r; }
static inline Dual HFd(const double *v, const double *dv, int i, int j, int ny, int nx)
{ i = clampi(i, 0, ny); j = clampi(j, 0, nx - 1); size_t k = (size_t)i * nx + j; Dual r = {v[k], dv[k]}; return r; }
static inline Dual CEd(const double *a, const double *da, int i, int j, int ny, int nx)
{ i = clampi(i, 0, ny - 1); j = clampi(j, 0, nx - 1); size_t k = (size_t)i * nx + j; Dual r = {a[k], da[k]}; return r; }
static inline Dual ETAd(const Dual *e, int i, int j, int ny, int nx)
{ Dual z = {0.0, 0.0}; return (i < 0 || i >= ny || j < 0 || j >= nx) ? z : e[(size_t)i * nx + j]; }

/* ------------------------------------------------------------------ */
/* Sigmoid (common.cu:67-76)                                           */
/* ------------------------------------------------------------------ */
static inline double sigm(double z, double c)
{
    double s = fmin(fmax(c * z, -20.0), 20.0);
    return 1.0 / (1.0 + OEXP(-s));
}
static inline double sigm_d(double z, double c)
{
    double s = sigm(z, c);
    return c * s * (1.0 - s);
}

/* ------------------------------------------------------------------ */
/* Viscosity (viscosity.cu:61-112 float, :4-59 dual)                   */
/* ------------------------------------------------------------------ */
static double eta_cell(int i, int j, const double *u, const double *v, const double *B,
                       double n, double eps_reg, double dx, int ny, int nx)
{
    double dx_inv = 1.0 / dx;
    double glen_exp = (1.0 - n) / (2.0 * n);
    double u_l = VF(u, i, j, ny, nx), u_r = VF(u, i, j + 1, ny, nx);
    double v_t = HF(v, i, j, ny, nx), v_b = HF(v, i + 1, j, ny, nx);
    double dudx = (u_r - u_l) * dx_inv;
    double dvdy = (v_t - v_b) * dx_inv;
    double tl_mask = (i > 0 && j > 0);
    double u_tl = VF(u, i - 1, j, ny, nx), v_lt = HF(v, i, j - 1, ny, nx);
    double e_tl = 0.5 * fma(u_tl - u_l, dx_inv, (v_t - v_lt) * dx_inv) * tl_mask;
    double tr_mask = (i > 0 && j < (nx - 1));
    double u_tr = VF(u, i - 1, j + 1, ny, nx), v_rt = HF(v, i, j + 1, ny, nx);
    double e_tr = 0.5 * fma(u_tr - u_r, dx_inv, (v_rt - v_t) * dx_inv) * tr_mask;
    double bl_mask = (i < (ny - 1) && j > 0);
    double u_bl = VF(u, i + 1, j, ny, nx), v_lb = HF(v, i + 1, j - 1, ny, nx);
    double e_bl = 0.5 * fma(u_l - u_bl, dx_inv, (v_b - v_lb) * dx_inv) * bl_mask;
    double br_mask = (i < (ny - 1) && j < (nx - 1));
    double u_br = VF(u, i + 1, j + 1, ny, nx), v_rb = HF(v, i + 1, j + 1, ny, nx);
    double e_br = 0.5 * fma(u_r - u_br, dx_inv, (v_rb - v_b) * dx_inv) * br_mask;
    double exy2 = 0.25 * fma(e_br, e_br, fma(e_bl, e_bl, fma(e_tr, e_tr, e_tl * e_tl)));
    double eII = fma(dudx, dvdy, fma(dvdy, dvdy, dudx * dudx)) + exy2 + eps_reg;
    return 0.5 * CE(B, i, j, ny, nx) * OPOW(eII, glen_exp);
}

static Dual eta_cell_dual(int i, int j, const double *u, const double *v,
                          const double *du, const double *dv, const double *B,
                          double n, double eps_reg, double dx, int ny, int nx)
{
    double dx_inv = 1.0 / dx;
    double glen_exp = (1.0 - n) / (2.0 * n);
    Dual u_l = VFd(u, du, i, j, ny, nx), u_r = VFd(u, du, i, j + 1, ny, nx);
    Dual v_t = HFd(v, dv, i, j, ny, nx), v_b = HFd(v, dv, i + 1, j, ny, nx);
    Dual dudx = dscale(dsub(u_r, u_l), dx_inv);
    Dual dvdy = dscale(dsub(v_t, v_b), dx_inv);
    double tl_mask = (i > 0 && j > 0);
    Dual u_tl = VFd(u, du, i - 1, j, ny, nx), v_lt = HFd(v, dv, i, j - 1, ny, nx);
    Dual e_tl = dscale(dscale(dadd(dscale(dsub(u_tl, u_l), dx_inv), dscale(dsub(v_t, v_lt), dx_inv)), 0.5), tl_mask);
    double tr_mask = (i > 0 && j < (nx - 1));
    Dual u_tr = VFd(u, du, i - 1, j + 1, ny, nx), v_rt = HFd(v, dv, i, j + 1, ny, nx);
    Dual e_tr = dscale(dscale(dadd(dscale(dsub(u_tr, u_r), dx_inv), dscale(dsub(v_rt, v_t), dx_inv)), 0.5), tr_mask);
    double bl_mask = (i < (ny - 1) && j > 0);
    Dual u_bl = VFd(u, du, i + 1, j, ny, nx), v_lb = HFd(v, dv, i + 1, j - 1, ny, nx);
    Dual e_bl = dscale(dscale(dadd(dscale(dsub(u_l, u_bl), dx_inv), dscale(dsub(v_b, v_lb), dx_inv)), 0.5), bl_mask);
    double br_mask = (i < (ny - 1) && j < (nx - 1));
    Dual u_br = VFd(u, du, i + 1, j + 1, ny, nx), v_rb = HFd(v, dv, i + 1, j + 1, ny, nx);
    Dual e_br = dscale(dscale(dadd(dscale(dsub(u_r, u_br), dx_inv), dscale(dsub(v_rb, v_b), dx_inv)), 0.5), br_mask);
    Dual s4 = dadd(dadd(dadd(dmul(e_tl, e_tl), dmul(e_tr, e_tr)), dmul(e_bl, e_bl)), dmul(e_br, e_br));
    Dual exy2 = dscale(s4, 0.25);
    Dual eII = daddc(dadd(dadd(dadd(dmul(dudx, dudx), dmul(dvdy, dvdy)), dmul(dudx, dvdy)), exy2), eps_reg);
    double s = 0.5 * CE(B, i, j, ny, nx);
    return dscale(dpow(eII, glen_exp), s);
}

/* ------------------------------------------------------------------ */
/* eta*H products (viscosity.cu:150-217)                               */
/* ------------------------------------------------------------------ */
static inline double etaH_vertex(double etl, double etr, double ebl, double ebr,
                                 double Htl, double Htr, double Hbl, double Hbr)
{
    return 0.25 * (etl * Htl + etr * Htr + ebl * Hbl + ebr * Hbr);
}

/* ------------------------------------------------------------------ */
/* Stresses (stress.cu)                                                */
/* ------------------------------------------------------------------ */
typedef struct { double res, d_u_l, d_u_r, d_v_t, d_v_b, d_eta_H; } SigN;
typedef struct { double res, d_u_t, d_u_b, d_v_l, d_v_r, d_eta_H; } SigS;
typedef struct { double res, d_u, d_H_l, d_H_r, d_beta_l, d_beta_r; } TauB;
typedef struct { double res, d_H_l, d_H_r; } TauD;
typedef struct { double res, d_u, d_H_l, d_H_r; } Flux;
typedef struct { double res, d_H; } Calv;

/* stress.cu:44-66 */
static inline SigN sigma_xx(double u_l, double u_r, double v_t, double v_b, double eta_H,
                            double dx_inv, int i, int j, int ny, int nx)
{
    SigN J = {0, 0, 0, 0, 0, 0};
    (void)i; (void)ny;
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
/* stress.cu:79-102 */
static inline SigN sigma_yy(double u_l, double u_r, double v_t, double v_b, double eta_H,
                            double dx_inv, int i, int j, int ny, int nx)
{
    SigN J = {0, 0, 0, 0, 0, 0};
    (void)j; (void)nx;
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
/* stress.cu:160-183, (i, j) is a vertex */
static inline SigS sigma_xy(double u_t, double u_b, double v_l, double v_r, double eta_H,
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
/* stress.cu:245-276 (x) and :331-361 (y) have identical algebra with
 * (l, r) -> (t, b); vel is u or v. */
static inline TauB tau_b(double vel, double H_l, double H_r, double bed_l, double bed_r,
                         double beta_l, double beta_r, const OParams *p)
{
    TauB J = {0, 0, 0, 0, 0, 0};
    double z_l = fma(0.917, H_l, bed_l);
    double z_r = fma(0.917, H_r, bed_r);
    double g_l = sigm(z_l, p->gl_sigmoid_c);
    double g_r = sigm(z_r, p->gl_sigmoid_c);
    double be_l = fma(g_l, beta_l, (1.0 - g_l) * p->water_drag);
    double be_r = fma(g_r, beta_r, (1.0 - g_r) * p->water_drag);
    double be = 0.5 * (be_l + be_r);
    J.res = -be * vel;
    J.d_u = -be;
    J.d_beta_l = -0.5 * g_l * vel;
    J.d_beta_r = -0.5 * g_r * vel;
    if (p->gl_derivatives) {
        double dg_l = 0.917 * sigm_d(z_l, p->gl_sigmoid_c);
        double dg_r = 0.917 * sigm_d(z_r, p->gl_sigmoid_c);
        double dbe_l = dg_l * (beta_l - p->water_drag);
        double dbe_r = dg_r * (beta_r - p->water_drag);
        J.d_H_l = -0.5 * dbe_l * vel;
        J.d_H_r = -0.5 * dbe_r * vel;
    }
    return J;
}
/* stress.cu:407-436: face (i, j) of the u grid, l/r cells */
static inline TauD tau_dx(double H_l, double H_r, double bed_l, double bed_r,
                          double dx_inv, int i, int j, int ny, int nx)
{
    TauD J = {0, 0, 0};
    (void)i; (void)ny;
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
/* stress.cu:482-510: face (i, j) of the v grid, t/b cells; d_H_l=d_H_t, d_H_r=d_H_b */
static inline TauD tau_dy(double H_t, double H_b, double bed_t, double bed_b,
                          double dx_inv, int i, int j, int ny, int nx)
{
    TauD J = {0, 0, 0};
    (void)j; (void)nx;
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

/* ------------------------------------------------------------------ */
/* Mass fluxes and calving (flux.cu)                                   */
/* ------------------------------------------------------------------ */
/* flux.cu:40-64: u face (i, j), cells l = (i, j-1), r = (i, j) */
static inline Flux flux_x(double u, double H_l, double H_r, int i, int j, int ny, int nx)
{
    Flux J = {0, 0, 0, 0};
    (void)i; (void)ny;
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
/* flux.cu:111-135: v face (i, j), cells t = (i-1, j), b = (i, j); d_H_l=d_H_t, d_H_r=d_H_b */
static inline Flux flux_y(double v, double H_t, double H_b, int i, int j, int ny, int nx)
{
    Flux J = {0, 0, 0, 0};
    (void)j; (void)nx;
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
/* flux.cu:188-211 */
static inline Calv calving(double H, double bed, const OParams *p)
{
    Calv J;
    double z = fma(0.917, H, bed);
    double g = sigm(z, p->gl_sigmoid_c);
    J.res = -p->calving_rate * (1.0 - g) * H;
    J.d_H = -p->calving_rate * (1.0 - g);
    if (p->gl_derivatives) {
        double dg = 0.917 * sigm_d(z, p->gl_sigmoid_c);
        J.d_H = fma(p->calving_rate * H, dg, J.d_H);
    }
    return J;
}

/* ------------------------------------------------------------------ */
/* Field layout helpers                                                */
/* ------------------------------------------------------------------ */
#define NU(g) ((size_t)(g)->ny * ((g)->nx + 1))
#define NV(g) ((size_t)((g)->ny + 1) * (g)->nx)
#define NC(g) ((size_t)(g)->ny * (g)->nx)
#define NT(g) (NU(g) + NV(g) + NC(g))

/* eta for every cell of the level */
void or_eta(const OGeom *g, const OParams *p, double *eta, const double *U, const double *B)
{
    int ny = g->ny, nx = g->nx;
    const double *u = U, *v = U + NU(g);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++)
            eta[(size_t)i * nx + j] = eta_cell(i, j, u, v, B, p->n, p->eps_reg, g->dx, ny, nx);
}

/* ------------------------------------------------------------------ */
/* Residual r = F(U) - f (residuals.cu:6-292)                          */
/* ------------------------------------------------------------------ */
static double res_cell(int i, int j, const OGeom *g, const OParams *p, const double *u,
                       const double *v, const double *H, const double *f_H,
                       const double *bed, const double *mask, const double *gamma)
{
    int ny = g->ny, nx = g->nx;
    double dx_inv = 1.0 / g->dx, dt = g->dt;
    double H_c = CE(H, i, j, ny, nx);
    double rH = H_c / dt - CE(f_H, i, j, ny, nx);
    double bed_c = CE(bed, i, j, ny, nx);
    Calv jc = calving(H_c, bed_c, p);
    rH -= jc.res;
    double H_l = CE(H, i, j - 1, ny, nx), u_l = VF(u, i, j, ny, nx);
    Flux jl = flux_x(u_l, H_l, H_c, i, j, ny, nx);
    rH -= jl.res * dx_inv;
    double H_r = CE(H, i, j + 1, ny, nx), u_r = VF(u, i, j + 1, ny, nx);
    Flux jr = flux_x(u_r, H_c, H_r, i, j + 1, ny, nx);
    rH += jr.res * dx_inv;
    double H_t = CE(H, i - 1, j, ny, nx), v_t = HF(v, i, j, ny, nx);
    Flux jt = flux_y(v_t, H_t, H_c, i, j, ny, nx);
    rH += jt.res * dx_inv;
    double H_b = CE(H, i + 1, j, ny, nx), v_b = HF(v, i + 1, j, ny, nx);
    Flux jb = flux_y(v_b, H_c, H_b, i + 1, j, ny, nx);
    rH -= jb.res * dx_inv;
    double m = CE(mask, i, j, ny, nx);
    double th = CE(gamma, i, j, ny, nx);
    return (1.0 - m) * rH + m * (H_c - th);
}

static double res_u(int i, int j, const OGeom *g, const OParams *p, const double *eta,
                    const double *u, const double *v, const double *H, const double *f_u,
                    const double *bed, const double *beta)
{
    int ny = g->ny, nx = g->nx;
    double dx_inv = 1.0 / g->dx;
    double ru = -VF(f_u, i, j, ny, nx);
    {
        double eH = ETA(eta, i, j, ny, nx) * CE(H, i, j, ny, nx);
        SigN s = sigma_xx(VF(u, i, j, ny, nx), VF(u, i, j + 1, ny, nx), HF(v, i, j, ny, nx),
                          HF(v, i + 1, j, ny, nx), eH, dx_inv, i, j, ny, nx);
        ru += s.res * dx_inv;
    }
    {
        double eH = ETA(eta, i, j - 1, ny, nx) * CE(H, i, j - 1, ny, nx);
        SigN s = sigma_xx(VF(u, i, j - 1, ny, nx), VF(u, i, j, ny, nx), HF(v, i, j - 1, ny, nx),
                          HF(v, i + 1, j - 1, ny, nx), eH, dx_inv, i, j - 1, ny, nx);
        ru -= s.res * dx_inv;
    }
    {
        double eH = etaH_vertex(ETA(eta, i - 1, j - 1, ny, nx), ETA(eta, i - 1, j, ny, nx),
                                ETA(eta, i, j - 1, ny, nx), ETA(eta, i, j, ny, nx),
                                CE(H, i - 1, j - 1, ny, nx), CE(H, i - 1, j, ny, nx),
                                CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx));
        SigS s = sigma_xy(VF(u, i - 1, j, ny, nx), VF(u, i, j, ny, nx), HF(v, i, j - 1, ny, nx),
                          HF(v, i, j, ny, nx), eH, dx_inv, i, j, ny, nx);
        ru += s.res * dx_inv;
    }
    {
        double eH = etaH_vertex(ETA(eta, i, j - 1, ny, nx), ETA(eta, i, j, ny, nx),
                                ETA(eta, i + 1, j - 1, ny, nx), ETA(eta, i + 1, j, ny, nx),
                                CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx),
                                CE(H, i + 1, j - 1, ny, nx), CE(H, i + 1, j, ny, nx));
        SigS s = sigma_xy(VF(u, i, j, ny, nx), VF(u, i + 1, j, ny, nx), HF(v, i + 1, j - 1, ny, nx),
                          HF(v, i + 1, j, ny, nx), eH, dx_inv, i + 1, j, ny, nx);
        ru -= s.res * dx_inv;
    }
    {
        TauB t = tau_b(VF(u, i, j, ny, nx), CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx),
                       CE(bed, i, j - 1, ny, nx), CE(bed, i, j, ny, nx),
                       CE(beta, i, j - 1, ny, nx), CE(beta, i, j, ny, nx), p);
        ru += t.res;
    }
    {
        TauD t = tau_dx(CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx), CE(bed, i, j - 1, ny, nx),
                        CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
        ru -= t.res;
    }
    return ru;
}

static double res_v(int i, int j, const OGeom *g, const OParams *p, const double *eta,
                    const double *u, const double *v, const double *H, const double *f_v,
                    const double *bed, const double *beta)
{
    int ny = g->ny, nx = g->nx;
    double dx_inv = 1.0 / g->dx;
    double rv = -HF(f_v, i, j, ny, nx);
    {
        double eH = ETA(eta, i - 1, j, ny, nx) * CE(H, i - 1, j, ny, nx);
        SigN s = sigma_yy(VF(u, i - 1, j, ny, nx), VF(u, i - 1, j + 1, ny, nx), HF(v, i - 1, j, ny, nx),
                          HF(v, i, j, ny, nx), eH, dx_inv, i - 1, j, ny, nx);
        rv += s.res * dx_inv;
    }
    {
        double eH = ETA(eta, i, j, ny, nx) * CE(H, i, j, ny, nx);
        SigN s = sigma_yy(VF(u, i, j, ny, nx), VF(u, i, j + 1, ny, nx), HF(v, i, j, ny, nx),
                          HF(v, i + 1, j, ny, nx), eH, dx_inv, i, j, ny, nx);
        rv -= s.res * dx_inv;
    }
    {
        double eH = etaH_vertex(ETA(eta, i - 1, j - 1, ny, nx), ETA(eta, i - 1, j, ny, nx),
                                ETA(eta, i, j - 1, ny, nx), ETA(eta, i, j, ny, nx),
                                CE(H, i - 1, j - 1, ny, nx), CE(H, i - 1, j, ny, nx),
                                CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx));
        SigS s = sigma_xy(VF(u, i - 1, j, ny, nx), VF(u, i, j, ny, nx), HF(v, i, j - 1, ny, nx),
                          HF(v, i, j, ny, nx), eH, dx_inv, i, j, ny, nx);
        rv -= s.res * dx_inv;
    }
    {
        double eH = etaH_vertex(ETA(eta, i - 1, j, ny, nx), ETA(eta, i - 1, j + 1, ny, nx),
                                ETA(eta, i, j, ny, nx), ETA(eta, i, j + 1, ny, nx),
                                CE(H, i - 1, j, ny, nx), CE(H, i - 1, j + 1, ny, nx),
                                CE(H, i, j, ny, nx), CE(H, i, j + 1, ny, nx));
        SigS s = sigma_xy(VF(u, i - 1, j + 1, ny, nx), VF(u, i, j + 1, ny, nx), HF(v, i, j, ny, nx),
                          HF(v, i, j + 1, ny, nx), eH, dx_inv, i, j + 1, ny, nx);
        rv += s.res * dx_inv;
    }
    {
        TauB t = tau_b(HF(v, i, j, ny, nx), CE(H, i - 1, j, ny, nx), CE(H, i, j, ny, nx),
                       CE(bed, i - 1, j, ny, nx), CE(bed, i, j, ny, nx),
                       CE(beta, i - 1, j, ny, nx), CE(beta, i, j, ny, nx), p);
        rv += t.res;
    }
    {
        TauD t = tau_dy(CE(H, i - 1, j, ny, nx), CE(H, i, j, ny, nx), CE(bed, i - 1, j, ny, nx),
                        CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
        rv -= t.res;
    }
    return rv;
}

/* r = F(U) - f over the whole level.  eta is workspace of NC doubles. */
void or_residual(const OGeom *g, const OParams *p, double *r, const double *U, const double *f,
                 const double *bed, const double *B, const double *beta, const double *mask,
                 const double *gamma, double *eta)
{
    int ny = g->ny, nx = g->nx;
    const double *u = U, *v = U + NU(g), *H = U + NU(g) + NV(g);
    const double *f_u = f, *f_v = f + NU(g), *f_H = f + NU(g) + NV(g);
    double *r_u = r, *r_v = r + NU(g), *r_H = r + NU(g) + NV(g);
    or_eta(g, p, eta, U, B);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i <= ny; i++) {
        for (int j = 0; j <= nx; j++) {
            if (i < ny && j < nx)
                r_H[(size_t)i * nx + j] = res_cell(i, j, g, p, u, v, H, f_H, bed, mask, gamma);
            if (i < ny)
                r_u[(size_t)i * (nx + 1) + j] = res_u(i, j, g, p, eta, u, v, H, f_u, bed, beta);
            if (j < nx)
                r_v[(size_t)i * nx + j] = res_v(i, j, g, p, eta, u, v, H, f_v, bed, beta);
        }
    }
}

/* ------------------------------------------------------------------ */
/* JVP  j = J(U) dU  (residuals.cu:299-579), dual-number forward mode  */
/* ------------------------------------------------------------------ */
static inline double sigN_jvp(SigN J, double ul, double ur, double vt, double vb, double eH)
{ return J.d_u_l * ul + J.d_u_r * ur + J.d_v_t * vt + J.d_v_b * vb + J.d_eta_H * eH; }
static inline double sigS_jvp(SigS J, double ut, double ub, double vl, double vr, double eH)
{ return J.d_u_t * ut + J.d_u_b * ub + J.d_v_l * vl + J.d_v_r * vr + J.d_eta_H * eH; }
/* get_eta_H_cell_dual: res = H*eta, d = H*eta' + eta*H' (viscosity.cu:144-163) */
static inline Dual etaH_cell_dual(Dual eta, Dual H)
{ Dual r; r.v = H.v * eta.v; r.d = H.v * eta.d + eta.v * H.d; return r; }
/* get_eta_H_vertex_dual (viscosity.cu:191-222) */
static inline Dual etaH_vertex_dual(Dual etl, Dual etr, Dual ebl, Dual ebr,
                                    Dual Htl, Dual Htr, Dual Hbl, Dual Hbr)
{
    Dual r;
    r.v = etaH_vertex(etl.v, etr.v, ebl.v, ebr.v, Htl.v, Htr.v, Hbl.v, Hbr.v);
    r.d = 0.25 * Htl.v * etl.d + 0.25 * etl.v * Htl.d +
          0.25 * Htr.v * etr.d + 0.25 * etr.v * Htr.d +
          0.25 * Hbl.v * ebl.d + 0.25 * ebl.v * Hbl.d +
          0.25 * Hbr.v * ebr.d + 0.25 * ebr.v * Hbr.d;
    return r;
}

void or_jvp(const OGeom *g, const OParams *p, double *jv, const double *U, const double *dU,
            const double *bed, const double *B, const double *beta, const double *mask,
            const double *gamma, Dual *eta)
{
    int ny = g->ny, nx = g->nx;
    double dx_inv = 1.0 / g->dx, dt = g->dt;
    const double *u = U, *v = U + NU(g), *H = U + NU(g) + NV(g);
    const double *du = dU, *dv = dU + NU(g), *dH = dU + NU(g) + NV(g);
    double *j_u = jv, *j_v = jv + NU(g), *j_H = jv + NU(g) + NV(g);
    (void)gamma;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++)
            eta[(size_t)i * nx + j] = eta_cell_dual(i, j, u, v, du, dv, B, p->n, p->eps_reg, g->dx, ny, nx);

    #pragma omp parallel for schedule(static)
    for (int i = 0; i <= ny; i++) {
        for (int j = 0; j <= nx; j++) {
            if (i < ny && j < nx) {
                Dual H_c = CEd(H, dH, i, j, ny, nx);
                double d = H_c.d / dt;
                Calv jc = calving(H_c.v, CE(bed, i, j, ny, nx), p);
                d -= jc.d_H * H_c.d;
                Dual H_l = CEd(H, dH, i, j - 1, ny, nx), u_l = VFd(u, du, i, j, ny, nx);
                Flux q = flux_x(u_l.v, H_l.v, H_c.v, i, j, ny, nx);
                d -= (q.d_u * u_l.d + q.d_H_l * H_l.d + q.d_H_r * H_c.d) * dx_inv;
                Dual H_r = CEd(H, dH, i, j + 1, ny, nx), u_r = VFd(u, du, i, j + 1, ny, nx);
                q = flux_x(u_r.v, H_c.v, H_r.v, i, j + 1, ny, nx);
                d += (q.d_u * u_r.d + q.d_H_l * H_c.d + q.d_H_r * H_r.d) * dx_inv;
                Dual H_t = CEd(H, dH, i - 1, j, ny, nx), v_t = HFd(v, dv, i, j, ny, nx);
                q = flux_y(v_t.v, H_t.v, H_c.v, i, j, ny, nx);
                d += (q.d_u * v_t.d + q.d_H_l * H_t.d + q.d_H_r * H_c.d) * dx_inv;
                Dual H_b = CEd(H, dH, i + 1, j, ny, nx), v_b = HFd(v, dv, i + 1, j, ny, nx);
                q = flux_y(v_b.v, H_c.v, H_b.v, i + 1, j, ny, nx);
                d -= (q.d_u * v_b.d + q.d_H_l * H_c.d + q.d_H_r * H_b.d) * dx_inv;
                double m = CE(mask, i, j, ny, nx);
                j_H[(size_t)i * nx + j] = (1.0 - m) * d + m * H_c.d;
            }
            if (i < ny) {
                double d = 0.0;
                {
                    Dual eH = etaH_cell_dual(ETAd(eta, i, j, ny, nx), CEd(H, dH, i, j, ny, nx));
                    Dual ul = VFd(u, du, i, j, ny, nx), ur = VFd(u, du, i, j + 1, ny, nx);
                    Dual vt = HFd(v, dv, i, j, ny, nx), vb = HFd(v, dv, i + 1, j, ny, nx);
                    SigN s = sigma_xx(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, i, j, ny, nx);
                    d += sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d) * dx_inv;
                }
                {
                    Dual eH = etaH_cell_dual(ETAd(eta, i, j - 1, ny, nx), CEd(H, dH, i, j - 1, ny, nx));
                    Dual ul = VFd(u, du, i, j, ny, nx), ull = VFd(u, du, i, j - 1, ny, nx);
                    Dual vt = HFd(v, dv, i, j - 1, ny, nx), vb = HFd(v, dv, i + 1, j - 1, ny, nx);
                    SigN s = sigma_xx(ull.v, ul.v, vt.v, vb.v, eH.v, dx_inv, i, j - 1, ny, nx);
                    d -= sigN_jvp(s, ull.d, ul.d, vt.d, vb.d, eH.d) * dx_inv;
                }
                {
                    Dual eH = etaH_vertex_dual(ETAd(eta, i - 1, j - 1, ny, nx), ETAd(eta, i - 1, j, ny, nx),
                                               ETAd(eta, i, j - 1, ny, nx), ETAd(eta, i, j, ny, nx),
                                               CEd(H, dH, i - 1, j - 1, ny, nx), CEd(H, dH, i - 1, j, ny, nx),
                                               CEd(H, dH, i, j - 1, ny, nx), CEd(H, dH, i, j, ny, nx));
                    Dual ut = VFd(u, du, i - 1, j, ny, nx), ub = VFd(u, du, i, j, ny, nx);
                    Dual vl = HFd(v, dv, i, j - 1, ny, nx), vr = HFd(v, dv, i, j, ny, nx);
                    SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i, j, ny, nx);
                    d += sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
                }
                {
                    Dual eH = etaH_vertex_dual(ETAd(eta, i, j - 1, ny, nx), ETAd(eta, i, j, ny, nx),
                                               ETAd(eta, i + 1, j - 1, ny, nx), ETAd(eta, i + 1, j, ny, nx),
                                               CEd(H, dH, i, j - 1, ny, nx), CEd(H, dH, i, j, ny, nx),
                                               CEd(H, dH, i + 1, j - 1, ny, nx), CEd(H, dH, i + 1, j, ny, nx));
                    Dual ut = VFd(u, du, i, j, ny, nx), ub = VFd(u, du, i + 1, j, ny, nx);
                    Dual vl = HFd(v, dv, i + 1, j - 1, ny, nx), vr = HFd(v, dv, i + 1, j, ny, nx);
                    SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i + 1, j, ny, nx);
                    d -= sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
                }
                {
                    Dual ul = VFd(u, du, i, j, ny, nx);
                    Dual Hl = CEd(H, dH, i, j - 1, ny, nx), Hc = CEd(H, dH, i, j, ny, nx);
                    TauB t = tau_b(ul.v, Hl.v, Hc.v, CE(bed, i, j - 1, ny, nx), CE(bed, i, j, ny, nx),
                                   CE(beta, i, j - 1, ny, nx), CE(beta, i, j, ny, nx), p);
                    d += t.d_u * ul.d + t.d_H_l * Hl.d + t.d_H_r * Hc.d;
                }
                {
                    Dual Hl = CEd(H, dH, i, j - 1, ny, nx), Hc = CEd(H, dH, i, j, ny, nx);
                    TauD t = tau_dx(Hl.v, Hc.v, CE(bed, i, j - 1, ny, nx), CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
                    d -= t.d_H_l * Hl.d + t.d_H_r * Hc.d;
                }
                j_u[(size_t)i * (nx + 1) + j] = d;
            }
            if (j < nx) {
                double d = 0.0;
                {
                    Dual eH = etaH_cell_dual(ETAd(eta, i - 1, j, ny, nx), CEd(H, dH, i - 1, j, ny, nx));
                    Dual ul = VFd(u, du, i - 1, j, ny, nx), ur = VFd(u, du, i - 1, j + 1, ny, nx);
                    Dual vt = HFd(v, dv, i - 1, j, ny, nx), vb = HFd(v, dv, i, j, ny, nx);
                    SigN s = sigma_yy(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, i - 1, j, ny, nx);
                    d += sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d) * dx_inv;
                }
                {
                    Dual eH = etaH_cell_dual(ETAd(eta, i, j, ny, nx), CEd(H, dH, i, j, ny, nx));
                    Dual ul = VFd(u, du, i, j, ny, nx), ur = VFd(u, du, i, j + 1, ny, nx);
                    Dual vt = HFd(v, dv, i, j, ny, nx), vb = HFd(v, dv, i + 1, j, ny, nx);
                    SigN s = sigma_yy(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, i, j, ny, nx);
                    d -= sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d) * dx_inv;
                }
                {
                    Dual eH = etaH_vertex_dual(ETAd(eta, i - 1, j - 1, ny, nx), ETAd(eta, i - 1, j, ny, nx),
                                               ETAd(eta, i, j - 1, ny, nx), ETAd(eta, i, j, ny, nx),
                                               CEd(H, dH, i - 1, j - 1, ny, nx), CEd(H, dH, i - 1, j, ny, nx),
                                               CEd(H, dH, i, j - 1, ny, nx), CEd(H, dH, i, j, ny, nx));
                    Dual ut = VFd(u, du, i - 1, j, ny, nx), ub = VFd(u, du, i, j, ny, nx);
                    Dual vl = HFd(v, dv, i, j - 1, ny, nx), vr = HFd(v, dv, i, j, ny, nx);
                    SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i, j, ny, nx);
                    d -= sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
                }
                {
                    Dual eH = etaH_vertex_dual(ETAd(eta, i - 1, j, ny, nx), ETAd(eta, i - 1, j + 1, ny, nx),
                                               ETAd(eta, i, j, ny, nx), ETAd(eta, i, j + 1, ny, nx),
                                               CEd(H, dH, i - 1, j, ny, nx), CEd(H, dH, i - 1, j + 1, ny, nx),
                                               CEd(H, dH, i, j, ny, nx), CEd(H, dH, i, j + 1, ny, nx));
                    Dual ut = VFd(u, du, i - 1, j + 1, ny, nx), ub = VFd(u, du, i, j + 1, ny, nx);
                    Dual vl = HFd(v, dv, i, j, ny, nx), vr = HFd(v, dv, i, j + 1, ny, nx);
                    SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i, j + 1, ny, nx);
                    d += sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
                }
                {
                    Dual vt = HFd(v, dv, i, j, ny, nx);
                    Dual Ht = CEd(H, dH, i - 1, j, ny, nx), Hc = CEd(H, dH, i, j, ny, nx);
                    TauB t = tau_b(vt.v, Ht.v, Hc.v, CE(bed, i - 1, j, ny, nx), CE(bed, i, j, ny, nx),
                                   CE(beta, i - 1, j, ny, nx), CE(beta, i, j, ny, nx), p);
                    d += t.d_u * vt.d + t.d_H_l * Ht.d + t.d_H_r * Hc.d;
                }
                {
                    Dual Ht = CEd(H, dH, i - 1, j, ny, nx), Hc = CEd(H, dH, i, j, ny, nx);
                    TauD t = tau_dy(Ht.v, Hc.v, CE(bed, i - 1, j, ny, nx), CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
                    d -= t.d_H_l * Ht.d + t.d_H_r * Hc.d;
                }
                j_v[(size_t)i * nx + j] = d;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* VJP  l = J(U)^T lambda  (residuals.cu:586-1023), scatter form in    */
/* the reference's per-thread term order; the (u,v)-(u,v) block is a   */
/* JVP along lambda (symmetric SSA Hessian), H and mass rows are true  */
/* transposes; masked mass rows are identity rows.                     */
/* ------------------------------------------------------------------ */
void or_vjp(const OGeom *g, const OParams *p, double *l, const double *U, const double *lam,
            const double *bed, const double *B, const double *beta, const double *mask,
            const double *gamma, Dual *eta)
{
    int ny = g->ny, nx = g->nx;
    double dx_inv = 1.0 / g->dx, dt = g->dt;
    const double *u = U, *v = U + NU(g), *H = U + NU(g) + NV(g);
    const double *lu = lam, *lv = lam + NU(g), *lH = lam + NU(g) + NV(g);
    double *a_u = l, *a_v = l + NU(g), *a_H = l + NU(g) + NV(g);
    (void)gamma;
    memset(l, 0, NT(g) * sizeof(double));
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++)
            eta[(size_t)i * nx + j] = eta_cell_dual(i, j, u, v, lu, lv, B, p->n, p->eps_reg, g->dx, ny, nx);

#define AU(ii, jj, val) do { int _i = (ii), _j = (jj); if (_i >= 0 && _i < ny && _j >= 0 && _j <= nx) a_u[(size_t)_i * (nx + 1) + _j] += (val); } while (0)
#define AV(ii, jj, val) do { int _i = (ii), _j = (jj); if (_i >= 0 && _i <= ny && _j >= 0 && _j < nx) a_v[(size_t)_i * nx + _j] += (val); } while (0)
#define AH(ii, jj, val) do { int _i = (ii), _j = (jj); if (_i >= 0 && _i < ny && _j >= 0 && _j < nx) a_H[(size_t)_i * nx + _j] += (val); } while (0)

    for (int i = 0; i <= ny; i++) {
        for (int j = 0; j <= nx; j++) {
            if (i < ny && j < nx) {
                double H_c = CE(H, i, j, ny, nx);
                double lc = CE(lH, i, j, ny, nx) * (1.0 - CE(mask, i, j, ny, nx));
                AH(i, j, lc / dt);
                Calv jc = calving(H_c, CE(bed, i, j, ny, nx), p);
                AH(i, j, -lc * jc.d_H);
                Flux q = flux_x(VF(u, i, j, ny, nx), CE(H, i, j - 1, ny, nx), H_c, i, j, ny, nx);
                AH(i, j, -lc * q.d_H_r * dx_inv);
                AH(i, j - 1, -lc * q.d_H_l * dx_inv);
                AU(i, j, -lc * q.d_u * dx_inv);
                q = flux_x(VF(u, i, j + 1, ny, nx), H_c, CE(H, i, j + 1, ny, nx), i, j + 1, ny, nx);
                AH(i, j, lc * q.d_H_l * dx_inv);
                AH(i, j + 1, lc * q.d_H_r * dx_inv);
                AU(i, j + 1, lc * q.d_u * dx_inv);
                q = flux_y(HF(v, i, j, ny, nx), CE(H, i - 1, j, ny, nx), H_c, i, j, ny, nx);
                AH(i, j, lc * q.d_H_r * dx_inv);
                AH(i - 1, j, lc * q.d_H_l * dx_inv);
                AV(i, j, lc * q.d_u * dx_inv);
                q = flux_y(HF(v, i + 1, j, ny, nx), H_c, CE(H, i + 1, j, ny, nx), i + 1, j, ny, nx);
                AH(i, j, -lc * q.d_H_l * dx_inv);
                AH(i + 1, j, -lc * q.d_H_r * dx_inv);
                AV(i + 1, j, -lc * q.d_u * dx_inv);
                double m = CE(mask, i, j, ny, nx);
                AH(i, j, CE(lH, i, j, ny, nx) * m);
            }
            if (i < ny) {
                double lul = VF(lu, i, j, ny, nx);
                {
                    Dual ec = ETAd(eta, i, j, ny, nx);
                    double Hc = CE(H, i, j, ny, nx);
                    SigN s = sigma_xx(VF(u, i, j, ny, nx), VF(u, i, j + 1, ny, nx), HF(v, i, j, ny, nx),
                                      HF(v, i + 1, j, ny, nx), ec.v * Hc, dx_inv, i, j, ny, nx);
                    double x = sigN_jvp(s, lul, VF(lu, i, j + 1, ny, nx), HF(lv, i, j, ny, nx),
                                        HF(lv, i + 1, j, ny, nx), Hc * ec.d + ec.v * 0.0);
                    AU(i, j, x * dx_inv);
                    AH(i, j, lul * s.d_eta_H * ec.v * dx_inv);
                }
                {
                    Dual el = ETAd(eta, i, j - 1, ny, nx);
                    double Hl = CE(H, i, j - 1, ny, nx);
                    SigN s = sigma_xx(VF(u, i, j - 1, ny, nx), VF(u, i, j, ny, nx), HF(v, i, j - 1, ny, nx),
                                      HF(v, i + 1, j - 1, ny, nx), el.v * Hl, dx_inv, i, j - 1, ny, nx);
                    double x = sigN_jvp(s, VF(lu, i, j - 1, ny, nx), lul, HF(lv, i, j - 1, ny, nx),
                                        HF(lv, i + 1, j - 1, ny, nx), Hl * el.d + el.v * 0.0);
                    AU(i, j, -x * dx_inv);
                    AH(i, j - 1, -lul * s.d_eta_H * el.v * dx_inv);
                }
                {
                    Dual e0 = ETAd(eta, i - 1, j - 1, ny, nx), e1 = ETAd(eta, i - 1, j, ny, nx);
                    Dual e2 = ETAd(eta, i, j - 1, ny, nx), e3 = ETAd(eta, i, j, ny, nx);
                    double H0 = CE(H, i - 1, j - 1, ny, nx), H1 = CE(H, i - 1, j, ny, nx);
                    double H2 = CE(H, i, j - 1, ny, nx), H3 = CE(H, i, j, ny, nx);
                    double eH = etaH_vertex(e0.v, e1.v, e2.v, e3.v, H0, H1, H2, H3);
                    double deH = 0.25 * H0 * e0.d + 0.25 * e0.v * 0.0 + 0.25 * H1 * e1.d + 0.25 * e1.v * 0.0 +
                                 0.25 * H2 * e2.d + 0.25 * e2.v * 0.0 + 0.25 * H3 * e3.d + 0.25 * e3.v * 0.0;
                    SigS s = sigma_xy(VF(u, i - 1, j, ny, nx), VF(u, i, j, ny, nx), HF(v, i, j - 1, ny, nx),
                                      HF(v, i, j, ny, nx), eH, dx_inv, i, j, ny, nx);
                    double x = sigS_jvp(s, VF(lu, i - 1, j, ny, nx), lul, HF(lv, i, j - 1, ny, nx),
                                        HF(lv, i, j, ny, nx), deH);
                    AU(i, j, x * dx_inv);
                    AH(i - 1, j - 1, lul * s.d_eta_H * (0.25 * e0.v) * dx_inv);
                    AH(i - 1, j, lul * s.d_eta_H * (0.25 * e1.v) * dx_inv);
                    AH(i, j - 1, lul * s.d_eta_H * (0.25 * e2.v) * dx_inv);
                    AH(i, j, lul * s.d_eta_H * (0.25 * e3.v) * dx_inv);
                }
                {
                    Dual e0 = ETAd(eta, i, j - 1, ny, nx), e1 = ETAd(eta, i, j, ny, nx);
                    Dual e2 = ETAd(eta, i + 1, j - 1, ny, nx), e3 = ETAd(eta, i + 1, j, ny, nx);
                    double H0 = CE(H, i, j - 1, ny, nx), H1 = CE(H, i, j, ny, nx);
                    double H2 = CE(H, i + 1, j - 1, ny, nx), H3 = CE(H, i + 1, j, ny, nx);
                    double eH = etaH_vertex(e0.v, e1.v, e2.v, e3.v, H0, H1, H2, H3);
                    double deH = 0.25 * H0 * e0.d + 0.25 * e0.v * 0.0 + 0.25 * H1 * e1.d + 0.25 * e1.v * 0.0 +
                                 0.25 * H2 * e2.d + 0.25 * e2.v * 0.0 + 0.25 * H3 * e3.d + 0.25 * e3.v * 0.0;
                    SigS s = sigma_xy(VF(u, i, j, ny, nx), VF(u, i + 1, j, ny, nx), HF(v, i + 1, j - 1, ny, nx),
                                      HF(v, i + 1, j, ny, nx), eH, dx_inv, i + 1, j, ny, nx);
                    double x = sigS_jvp(s, lul, VF(lu, i + 1, j, ny, nx), HF(lv, i + 1, j - 1, ny, nx),
                                        HF(lv, i + 1, j, ny, nx), deH);
                    AU(i, j, -x * dx_inv);
                    AH(i, j - 1, -lul * s.d_eta_H * (0.25 * e0.v) * dx_inv);
                    AH(i, j, -lul * s.d_eta_H * (0.25 * e1.v) * dx_inv);
                    AH(i + 1, j - 1, -lul * s.d_eta_H * (0.25 * e2.v) * dx_inv);
                    AH(i + 1, j, -lul * s.d_eta_H * (0.25 * e3.v) * dx_inv);
                }
                {
                    TauB t = tau_b(VF(u, i, j, ny, nx), CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx),
                                   CE(bed, i, j - 1, ny, nx), CE(bed, i, j, ny, nx),
                                   CE(beta, i, j - 1, ny, nx), CE(beta, i, j, ny, nx), p);
                    AU(i, j, lul * t.d_u);
                    AH(i, j - 1, lul * t.d_H_l);
                    AH(i, j, lul * t.d_H_r);
                }
                {
                    TauD t = tau_dx(CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx), CE(bed, i, j - 1, ny, nx),
                                    CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
                    AH(i, j - 1, -lul * t.d_H_l);
                    AH(i, j, -lul * t.d_H_r);
                }
            }
            if (j < nx) {
                double lvt = HF(lv, i, j, ny, nx);
                {
                    Dual et = ETAd(eta, i - 1, j, ny, nx);
                    double Ht = CE(H, i - 1, j, ny, nx);
                    SigN s = sigma_yy(VF(u, i - 1, j, ny, nx), VF(u, i - 1, j + 1, ny, nx), HF(v, i - 1, j, ny, nx),
                                      HF(v, i, j, ny, nx), et.v * Ht, dx_inv, i - 1, j, ny, nx);
                    double x = sigN_jvp(s, VF(lu, i - 1, j, ny, nx), VF(lu, i - 1, j + 1, ny, nx),
                                        HF(lv, i - 1, j, ny, nx), lvt, Ht * et.d + et.v * 0.0);
                    AV(i, j, x * dx_inv);
                    AH(i - 1, j, lvt * s.d_eta_H * et.v * dx_inv);
                }
                {
                    Dual ec = ETAd(eta, i, j, ny, nx);
                    double Hc = CE(H, i, j, ny, nx);
                    SigN s = sigma_yy(VF(u, i, j, ny, nx), VF(u, i, j + 1, ny, nx), HF(v, i, j, ny, nx),
                                      HF(v, i + 1, j, ny, nx), ec.v * Hc, dx_inv, i, j, ny, nx);
                    double x = sigN_jvp(s, VF(lu, i, j, ny, nx), VF(lu, i, j + 1, ny, nx), lvt,
                                        HF(lv, i + 1, j, ny, nx), Hc * ec.d + ec.v * 0.0);
                    AV(i, j, -x * dx_inv);
                    AH(i, j, -lvt * s.d_eta_H * ec.v * dx_inv);
                }
                {
                    Dual e0 = ETAd(eta, i - 1, j - 1, ny, nx), e1 = ETAd(eta, i - 1, j, ny, nx);
                    Dual e2 = ETAd(eta, i, j - 1, ny, nx), e3 = ETAd(eta, i, j, ny, nx);
                    double H0 = CE(H, i - 1, j - 1, ny, nx), H1 = CE(H, i - 1, j, ny, nx);
                    double H2 = CE(H, i, j - 1, ny, nx), H3 = CE(H, i, j, ny, nx);
                    double eH = etaH_vertex(e0.v, e1.v, e2.v, e3.v, H0, H1, H2, H3);
                    double deH = 0.25 * H0 * e0.d + 0.25 * e0.v * 0.0 + 0.25 * H1 * e1.d + 0.25 * e1.v * 0.0 +
                                 0.25 * H2 * e2.d + 0.25 * e2.v * 0.0 + 0.25 * H3 * e3.d + 0.25 * e3.v * 0.0;
                    SigS s = sigma_xy(VF(u, i - 1, j, ny, nx), VF(u, i, j, ny, nx), HF(v, i, j - 1, ny, nx),
                                      HF(v, i, j, ny, nx), eH, dx_inv, i, j, ny, nx);
                    double x = sigS_jvp(s, VF(lu, i - 1, j, ny, nx), VF(lu, i, j, ny, nx),
                                        HF(lv, i, j - 1, ny, nx), lvt, deH);
                    AV(i, j, -x * dx_inv);
                    AH(i - 1, j - 1, -lvt * s.d_eta_H * (0.25 * e0.v) * dx_inv);
                    AH(i - 1, j, -lvt * s.d_eta_H * (0.25 * e1.v) * dx_inv);
                    AH(i, j - 1, -lvt * s.d_eta_H * (0.25 * e2.v) * dx_inv);
                    AH(i, j, -lvt * s.d_eta_H * (0.25 * e3.v) * dx_inv);
                }
                {
                    Dual e0 = ETAd(eta, i - 1, j, ny, nx), e1 = ETAd(eta, i - 1, j + 1, ny, nx);
                    Dual e2 = ETAd(eta, i, j, ny, nx), e3 = ETAd(eta, i, j + 1, ny, nx);
                    double H0 = CE(H, i - 1, j, ny, nx), H1 = CE(H, i - 1, j + 1, ny, nx);
                    double H2 = CE(H, i, j, ny, nx), H3 = CE(H, i, j + 1, ny, nx);
                    double eH = etaH_vertex(e0.v, e1.v, e2.v, e3.v, H0, H1, H2, H3);
                    double deH = 0.25 * H0 * e0.d + 0.25 * e0.v * 0.0 + 0.25 * H1 * e1.d + 0.25 * e1.v * 0.0 +
                                 0.25 * H2 * e2.d + 0.25 * e2.v * 0.0 + 0.25 * H3 * e3.d + 0.25 * e3.v * 0.0;
                    SigS s = sigma_xy(VF(u, i - 1, j + 1, ny, nx), VF(u, i, j + 1, ny, nx), HF(v, i, j, ny, nx),
                                      HF(v, i, j + 1, ny, nx), eH, dx_inv, i, j + 1, ny, nx);
                    double x = sigS_jvp(s, VF(lu, i - 1, j + 1, ny, nx), VF(lu, i, j + 1, ny, nx), lvt,
                                        HF(lv, i, j + 1, ny, nx), deH);
                    AV(i, j, x * dx_inv);
                    AH(i - 1, j, lvt * s.d_eta_H * (0.25 * e0.v) * dx_inv);
                    AH(i - 1, j + 1, lvt * s.d_eta_H * (0.25 * e1.v) * dx_inv);
                    AH(i, j, lvt * s.d_eta_H * (0.25 * e2.v) * dx_inv);
                    AH(i, j + 1, lvt * s.d_eta_H * (0.25 * e3.v) * dx_inv);
                }
                {
                    TauB t = tau_b(HF(v, i, j, ny, nx), CE(H, i - 1, j, ny, nx), CE(H, i, j, ny, nx),
                                   CE(bed, i - 1, j, ny, nx), CE(bed, i, j, ny, nx),
                                   CE(beta, i - 1, j, ny, nx), CE(beta, i, j, ny, nx), p);
                    AV(i, j, lvt * t.d_u);
                    AH(i - 1, j, lvt * t.d_H_l);
                    AH(i, j, lvt * t.d_H_r);
                }
                {
                    TauD t = tau_dy(CE(H, i - 1, j, ny, nx), CE(H, i, j, ny, nx), CE(bed, i - 1, j, ny, nx),
                                    CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
                    AH(i - 1, j, -lvt * t.d_H_l);
                    AH(i, j, -lvt * t.d_H_r);
                }
            }
        }
    }
#undef AU
#undef AV
#undef AH
}

/* ------------------------------------------------------------------ */
/* 5x5 Doolittle LU without pivoting (common.cu:132-175).  The          */
/* reference's module is built with --use_fast_math (kernels.py:72):    */
/* contracted multiply-adds, reciprocal-based division.  Here every     */
/* update c - a*b is one fma and back substitution multiplies by the    */
/* elimination's pivot reciprocals; substitution sums the terms of the  */
/* later unknowns first.                                                */
/* ------------------------------------------------------------------ */
void or_lu5(const double *A, const double *b, double *x)
{
    double LU[5][5], inv[5];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) LU[i][j] = A[i * 5 + j];
    for (int k = 0; k < 5; k++) {
        inv[k] = 1.0 / LU[k][k];
        for (int i = k + 1; i < 5; i++) {
            LU[i][k] *= inv[k];
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

/* ------------------------------------------------------------------ */
/* Local 5x5 block assembly shared by the forward Vanka Newton loop     */
/* (vanka.cu:56-411) and the adjoint smoother (vanka.cu:485-780).      */
/* Unknowns (u_l, u_r, v_t, v_b, H_c) of cell (i, j); neighbours come   */
/* from the global state; eta is frozen.  mode 0: forward (residual,    */
/* active-set row test, diagonal shift), mode 1: adjoint (no residual,  */
/* mass row replaced where adj_masked).                                */
/* ------------------------------------------------------------------ */
typedef struct {
    const OGeom *g; const OParams *p;
    const double *u, *v, *H, *f_u, *f_v, *f_H, *bed, *beta, *gamma, *eta;
} VCtx;

/* Experimental variant (not the reference algorithm): add the shear
 * cross-couplings that the reference's block leaves out. */
static int g_full_block = 0;
void or_set_full_block(int on) { g_full_block = on; }

static void vanka_assemble(const VCtx *c, int i, int j, double u_l, double u_r, double v_t,
                           double v_b, double H_c, double thklim, int mode, int adj_masked,
                           double J[25], double r[5])
{
    const OGeom *g = c->g; const OParams *p = c->p;
    int ny = g->ny, nx = g->nx;
    double dx_inv = 1.0 / g->dx, dt = g->dt;
    const double *u = c->u, *v = c->v, *H = c->H, *bed = c->bed, *beta = c->beta, *eta = c->eta;
    for (int k = 0; k < 25; k++) J[k] = 0.0;
    for (int k = 0; k < 5; k++) r[k] = 0.0;
    if (mode == 0) {
        r[0] -= VF(c->f_u, i, j, ny, nx);
        r[1] -= VF(c->f_u, i, j + 1, ny, nx);
        r[2] -= HF(c->f_v, i, j, ny, nx);
        r[3] -= HF(c->f_v, i + 1, j, ny, nx);
        r[4] -= HF(c->f_H, i, j, ny, nx);
    }
    /* mass conservation row */
    {
        J[24] = 1.0 / dt;
        r[4] += H_c / dt;
        double bed_c = CE(bed, i, j, ny, nx);
        Calv jc = calving(H_c, bed_c, p);
        J[24] -= jc.d_H;
        r[4] -= jc.res;
        double H_l = CE(H, i, j - 1, ny, nx);
        Flux q = flux_x(u_l, H_l, H_c, i, j, ny, nx);
        J[20] = fma(-q.d_u, dx_inv, J[20]);
        J[24] = fma(-q.d_H_r, dx_inv, J[24]);
        r[4] = fma(-q.res, dx_inv, r[4]);
        double H_r = CE(H, i, j + 1, ny, nx);
        q = flux_x(u_r, H_c, H_r, i, j + 1, ny, nx);
        J[21] = fma(q.d_u, dx_inv, J[21]);
        J[24] = fma(q.d_H_l, dx_inv, J[24]);
        r[4] = fma(q.res, dx_inv, r[4]);
        double H_t = CE(H, i - 1, j, ny, nx);
        q = flux_y(v_t, H_t, H_c, i, j, ny, nx);
        J[22] = fma(q.d_u, dx_inv, J[22]);
        J[24] = fma(q.d_H_r, dx_inv, J[24]);
        r[4] = fma(q.res, dx_inv, r[4]);
        double H_b = CE(H, i + 1, j, ny, nx);
        q = flux_y(v_b, H_c, H_b, i + 1, j, ny, nx);
        J[23] = fma(-q.d_u, dx_inv, J[23]);
        J[24] = fma(-q.d_H_l, dx_inv, J[24]);
        r[4] = fma(-q.res, dx_inv, r[4]);
        int constrained = (mode == 0) ? ((H_c - dt * r[4]) <= thklim) : (adj_masked != 0);
        if (constrained) {
            for (int k = 0; k < 5; k++) J[20 + k] = 0.0;
            J[24] = 1.0;
            if (mode == 0) r[4] = H_c - thklim;
        }
    }
    /* centre-cell normal stresses into all four face rows */
    {
        double ec = ETA(eta, i, j, ny, nx);
        double eH = H_c * ec;   /* get_eta_H_cell_jac: res = H*eta, d_H = eta */
        SigN s = sigma_xx(u_l, u_r, v_t, v_b, eH, dx_inv, i, j, ny, nx);
        r[0] = fma(s.res, dx_inv, r[0]);
        J[0] = fma(s.d_u_l, dx_inv, J[0]); J[1] = fma(s.d_u_r, dx_inv, J[1]); J[2] = fma(s.d_v_t, dx_inv, J[2]);
        J[3] = fma(s.d_v_b, dx_inv, J[3]); J[4] = fma(s.d_eta_H * ec, dx_inv, J[4]);
        r[1] = fma(-s.res, dx_inv, r[1]);
        J[5] = fma(-s.d_u_l, dx_inv, J[5]); J[6] = fma(-s.d_u_r, dx_inv, J[6]); J[7] = fma(-s.d_v_t, dx_inv, J[7]);
        J[8] = fma(-s.d_v_b, dx_inv, J[8]); J[9] = fma(-(s.d_eta_H * ec), dx_inv, J[9]);
        SigN t = sigma_yy(u_l, u_r, v_t, v_b, eH, dx_inv, i, j, ny, nx);
        r[2] = fma(-t.res, dx_inv, r[2]);
        J[10] = fma(-t.d_u_l, dx_inv, J[10]); J[11] = fma(-t.d_u_r, dx_inv, J[11]); J[12] = fma(-t.d_v_t, dx_inv, J[12]);
        J[13] = fma(-t.d_v_b, dx_inv, J[13]); J[14] = fma(-(t.d_eta_H * ec), dx_inv, J[14]);
        r[3] = fma(t.res, dx_inv, r[3]);
        J[15] = fma(t.d_u_l, dx_inv, J[15]); J[16] = fma(t.d_u_r, dx_inv, J[16]); J[17] = fma(t.d_v_t, dx_inv, J[17]);
        J[18] = fma(t.d_v_b, dx_inv, J[18]); J[19] = fma(t.d_eta_H * ec, dx_inv, J[19]);
    }
    /* left cell sigma_xx into the left u row */
    {
        double eH = ETA(eta, i, j - 1, ny, nx) * CE(H, i, j - 1, ny, nx);
        SigN s = sigma_xx(VF(u, i, j - 1, ny, nx), u_l, HF(v, i, j - 1, ny, nx), HF(v, i + 1, j - 1, ny, nx),
                          eH, dx_inv, i, j - 1, ny, nx);
        r[0] = fma(-s.res, dx_inv, r[0]);
        J[0] = fma(-s.d_u_r, dx_inv, J[0]);
    }
    /* right cell sigma_xx into the right u row */
    {
        double eH = ETA(eta, i, j + 1, ny, nx) * CE(H, i, j + 1, ny, nx);
        SigN s = sigma_xx(u_r, VF(u, i, j + 2, ny, nx), HF(v, i, j + 1, ny, nx), HF(v, i + 1, j + 1, ny, nx),
                          eH, dx_inv, i, j + 1, ny, nx);
        r[1] = fma(s.res, dx_inv, r[1]);
        J[6] = fma(s.d_u_l, dx_inv, J[6]);
    }
    /* top cell sigma_yy into the top v row */
    {
        double eH = ETA(eta, i - 1, j, ny, nx) * CE(H, i - 1, j, ny, nx);
        SigN s = sigma_yy(VF(u, i - 1, j, ny, nx), VF(u, i - 1, j + 1, ny, nx), HF(v, i - 1, j, ny, nx), v_t,
                          eH, dx_inv, i - 1, j, ny, nx);
        r[2] = fma(s.res, dx_inv, r[2]);
        J[12] = fma(s.d_v_b, dx_inv, J[12]);
    }
    /* bottom cell sigma_yy into the bottom v row */
    {
        double eH = ETA(eta, i + 1, j, ny, nx) * CE(H, i + 1, j, ny, nx);
        SigN s = sigma_yy(VF(u, i + 1, j, ny, nx), VF(u, i + 1, j + 1, ny, nx), v_b, HF(v, i + 2, j, ny, nx),
                          eH, dx_inv, i + 1, j, ny, nx);
        r[3] = fma(-s.res, dx_inv, r[3]);
        J[18] = fma(-s.d_v_t, dx_inv, J[18]);
    }
    /* top-left vertex (i, j) */
    {
        double e_tl = ETA(eta, i - 1, j - 1, ny, nx), e_t = ETA(eta, i - 1, j, ny, nx);
        double e_l = ETA(eta, i, j - 1, ny, nx), e_c = ETA(eta, i, j, ny, nx);
        double eH = etaH_vertex(e_tl, e_t, e_l, e_c, CE(H, i - 1, j - 1, ny, nx), CE(H, i - 1, j, ny, nx),
                                CE(H, i, j - 1, ny, nx), H_c);
        double dH_br = 0.25 * e_c;
        SigS s = sigma_xy(VF(u, i - 1, j, ny, nx), u_l, HF(v, i, j - 1, ny, nx), v_t, eH, dx_inv, i, j, ny, nx);
        r[0] = fma(s.res, dx_inv, r[0]);
        J[0] = fma(s.d_u_b, dx_inv, J[0]);
        J[4] = fma(s.d_eta_H * dH_br, dx_inv, J[4]);
        r[2] = fma(-s.res, dx_inv, r[2]);
        J[12] = fma(-s.d_v_r, dx_inv, J[12]);
        J[14] = fma(-(s.d_eta_H * dH_br), dx_inv, J[14]);
        if (g_full_block) { J[2] = fma(s.d_v_r, dx_inv, J[2]); J[10] = fma(-s.d_u_b, dx_inv, J[10]); }
    }
    /* top-right vertex (i, j+1) */
    {
        double e_t = ETA(eta, i - 1, j, ny, nx), e_tr = ETA(eta, i - 1, j + 1, ny, nx);
        double e_c = ETA(eta, i, j, ny, nx), e_r = ETA(eta, i, j + 1, ny, nx);
        double eH = etaH_vertex(e_t, e_tr, e_c, e_r, CE(H, i - 1, j, ny, nx), CE(H, i - 1, j + 1, ny, nx),
                                H_c, CE(H, i, j + 1, ny, nx));
        double dH_bl = 0.25 * e_c;
        SigS s = sigma_xy(VF(u, i - 1, j + 1, ny, nx), u_r, v_t, HF(v, i, j + 1, ny, nx), eH, dx_inv, i, j + 1, ny, nx);
        r[1] = fma(s.res, dx_inv, r[1]);
        J[6] = fma(s.d_u_b, dx_inv, J[6]);
        J[9] = fma(s.d_eta_H * dH_bl, dx_inv, J[9]);
        r[2] = fma(s.res, dx_inv, r[2]);
        J[12] = fma(s.d_v_l, dx_inv, J[12]);
        J[14] = fma(s.d_eta_H * dH_bl, dx_inv, J[14]);
        if (g_full_block) { J[7] = fma(s.d_v_l, dx_inv, J[7]); J[11] = fma(s.d_u_b, dx_inv, J[11]); }
    }
    /* bottom-left vertex (i+1, j) */
    {
        double e_l = ETA(eta, i, j - 1, ny, nx), e_c = ETA(eta, i, j, ny, nx);
        double e_bl = ETA(eta, i + 1, j - 1, ny, nx), e_b = ETA(eta, i + 1, j, ny, nx);
        double eH = etaH_vertex(e_l, e_c, e_bl, e_b, CE(H, i, j - 1, ny, nx), H_c,
                                CE(H, i + 1, j - 1, ny, nx), CE(H, i + 1, j, ny, nx));
        double dH_tr = 0.25 * e_c;
        SigS s = sigma_xy(u_l, VF(u, i + 1, j, ny, nx), HF(v, i + 1, j - 1, ny, nx), v_b, eH, dx_inv, i + 1, j, ny, nx);
        r[0] = fma(-s.res, dx_inv, r[0]);
        J[0] = fma(-s.d_u_t, dx_inv, J[0]);
        J[4] = fma(-(s.d_eta_H * dH_tr), dx_inv, J[4]);
        r[3] = fma(-s.res, dx_inv, r[3]);
        J[18] = fma(-s.d_v_r, dx_inv, J[18]);
        J[19] = fma(-(s.d_eta_H * dH_tr), dx_inv, J[19]);
        if (g_full_block) { J[3] = fma(-s.d_v_r, dx_inv, J[3]); J[15] = fma(-s.d_u_t, dx_inv, J[15]); }
    }
    /* bottom-right vertex (i+1, j+1) */
    {
        double e_c = ETA(eta, i, j, ny, nx), e_r = ETA(eta, i, j + 1, ny, nx);
        double e_b = ETA(eta, i + 1, j, ny, nx), e_br = ETA(eta, i + 1, j + 1, ny, nx);
        double eH = etaH_vertex(e_c, e_r, e_b, e_br, H_c, CE(H, i, j + 1, ny, nx),
                                CE(H, i + 1, j, ny, nx), CE(H, i + 1, j + 1, ny, nx));
        double dH_tl = 0.25 * e_c;
        SigS s = sigma_xy(u_r, VF(u, i + 1, j + 1, ny, nx), v_b, HF(v, i + 1, j + 1, ny, nx), eH, dx_inv, i + 1, j + 1, ny, nx);
        r[1] = fma(-s.res, dx_inv, r[1]);
        J[6] = fma(-s.d_u_t, dx_inv, J[6]);
        J[9] = fma(-(s.d_eta_H * dH_tl), dx_inv, J[9]);
        r[3] = fma(s.res, dx_inv, r[3]);
        J[18] = fma(s.d_v_l, dx_inv, J[18]);
        J[19] = fma(s.d_eta_H * dH_tl, dx_inv, J[19]);
        if (g_full_block) { J[8] = fma(-s.d_v_l, dx_inv, J[8]); J[16] = fma(s.d_u_t, dx_inv, J[16]); }
    }
    /* basal drag */
    {
        TauB t = tau_b(u_l, CE(H, i, j - 1, ny, nx), H_c, CE(bed, i, j - 1, ny, nx), CE(bed, i, j, ny, nx),
                       CE(beta, i, j - 1, ny, nx), CE(beta, i, j, ny, nx), p);
        r[0] += t.res; J[0] += t.d_u; J[4] += t.d_H_r;
    }
    {
        TauB t = tau_b(u_r, H_c, CE(H, i, j + 1, ny, nx), CE(bed, i, j, ny, nx), CE(bed, i, j + 1, ny, nx),
                       CE(beta, i, j, ny, nx), CE(beta, i, j + 1, ny, nx), p);
        r[1] += t.res; J[6] += t.d_u; J[9] += t.d_H_l;
    }
    {
        TauB t = tau_b(v_t, CE(H, i - 1, j, ny, nx), H_c, CE(bed, i - 1, j, ny, nx), CE(bed, i, j, ny, nx),
                       CE(beta, i - 1, j, ny, nx), CE(beta, i, j, ny, nx), p);
        r[2] += t.res; J[12] += t.d_u; J[14] += t.d_H_r;
    }
    {
        TauB t = tau_b(v_b, H_c, CE(H, i + 1, j, ny, nx), CE(bed, i, j, ny, nx), CE(bed, i + 1, j, ny, nx),
                       CE(beta, i, j, ny, nx), CE(beta, i + 1, j, ny, nx), p);
        r[3] += t.res; J[18] += t.d_u; J[19] += t.d_H_l;
    }
    /* driving stress */
    {
        TauD t = tau_dx(CE(H, i, j - 1, ny, nx), H_c, CE(bed, i, j - 1, ny, nx), CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
        r[0] -= t.res; J[4] -= t.d_H_r;
    }
    {
        TauD t = tau_dx(H_c, CE(H, i, j + 1, ny, nx), CE(bed, i, j, ny, nx), CE(bed, i, j + 1, ny, nx), dx_inv, i, j + 1, ny, nx);
        r[1] -= t.res; J[9] -= t.d_H_l;
    }
    {
        TauD t = tau_dy(CE(H, i - 1, j, ny, nx), H_c, CE(bed, i - 1, j, ny, nx), CE(bed, i, j, ny, nx), dx_inv, i, j, ny, nx);
        r[2] -= t.res; J[14] -= t.d_H_r;
    }
    {
        TauD t = tau_dy(H_c, CE(H, i + 1, j, ny, nx), CE(bed, i, j, ny, nx), CE(bed, i + 1, j, ny, nx), dx_inv, i + 1, j, ny, nx);
        r[3] -= t.res; J[19] -= t.d_H_l;
    }
    if (mode == 0) {   /* vanka.cu:406-410 */
        J[0] -= 1.0; J[6] -= 1.0; J[12] -= 1.0; J[18] -= 1.0; J[24] += 1.0;
    }
}

/* Local block (J, r) at the cell's current values, exported for tests. */
void or_vanka_block(const OGeom *g, const OParams *p, const double *U, const double *f,
                    const double *bed, const double *beta, const double *gamma, const double *eta,
                    int i, int j, int mode, int adj_masked, double *J, double *r)
{
    VCtx c = {g, p, U, U + NU(g), U + NU(g) + NV(g), f, f + NU(g), f + NU(g) + NV(g), bed, beta, gamma, eta};
    int ny = g->ny, nx = g->nx;
    vanka_assemble(&c, i, j, VF(c.u, i, j, ny, nx), VF(c.u, i, j + 1, ny, nx), HF(c.v, i, j, ny, nx),
                   HF(c.v, i + 1, j, ny, nx), CE(c.H, i, j, ny, nx), CE(gamma, i, j, ny, nx), mode,
                   adj_masked, J, r);
}

/* One forward Vanka sweep (vanka.cu:3-439 / :844-1295 + grid.py:305-317):
 * per cell 10 damped Newton steps with eta frozen, then U += 0.5 * delta,
 * where delta_face = sum of 0.5*(x - x0) from the (at most two) adjacent
 * smoothed cells and delta_H = H_c - H_c0.  active == NULL smooths all. */
void or_vanka_sweep(const OGeom *g, const OParams *p, double *U, const double *f,
                    const double *bed, const double *B, const double *beta, const double *gamma,
                    const unsigned char *active, int n_newton, double *eta, double *cellnew)
{
    int ny = g->ny, nx = g->nx;
    size_t nu = NU(g), nv = NV(g);
    or_eta(g, p, eta, U, B);
    VCtx c = {g, p, U, U + nu, U + nu + nv, f, f + nu, f + nu + nv, bed, beta, gamma, eta};
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; i++) {
        for (int j = 0; j < nx; j++) {
            size_t ci = (size_t)i * nx + j;
            if (active && !active[ci]) continue;
            double u_l = VF(c.u, i, j, ny, nx), u_r = VF(c.u, i, j + 1, ny, nx);
            double v_t = HF(c.v, i, j, ny, nx), v_b = HF(c.v, i + 1, j, ny, nx);
            double H_c = CE(c.H, i, j, ny, nx), th = CE(gamma, i, j, ny, nx);
            for (int k = 0; k < n_newton; k++) {
                double J[25], r[5], dxl[5];
                vanka_assemble(&c, i, j, u_l, u_r, v_t, v_b, H_c, th, 0, 0, J, r);
                or_lu5(J, r, dxl);
                u_l -= 0.25 * dxl[0];
                u_r -= 0.25 * dxl[1];
                v_t -= 0.25 * dxl[2];
                v_b -= 0.25 * dxl[3];
                H_c -= 0.25 * dxl[4];
                H_c = fmax(H_c, th);
            }
            double *o = cellnew + 5 * ci;
            o[0] = 0.5 * (u_l - VF(c.u, i, j, ny, nx));
            o[1] = 0.5 * (u_r - VF(c.u, i, j + 1, ny, nx));
            o[2] = 0.5 * (v_t - HF(c.v, i, j, ny, nx));
            o[3] = 0.5 * (v_b - HF(c.v, i + 1, j, ny, nx));
            o[4] = H_c - CE(c.H, i, j, ny, nx);
        }
    }
    double *u = U, *v = U + nu, *H = U + nu + nv;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i <= ny; i++) {
        for (int j = 0; j <= nx; j++) {
            if (i < ny) {      /* u face (i, j): right face of (i, j-1), left face of (i, j) */
                double d = 0.0;
                if (j > 0 && (!active || active[(size_t)i * nx + j - 1])) d = d + cellnew[5 * ((size_t)i * nx + j - 1) + 1];
                if (j < nx && (!active || active[(size_t)i * nx + j])) d = d + cellnew[5 * ((size_t)i * nx + j) + 0];
                u[(size_t)i * (nx + 1) + j] += 0.5 * d;
            }
            if (j < nx) {      /* v face (i, j): bottom face of (i-1, j), top face of (i, j) */
                double d = 0.0;
                if (i > 0 && (!active || active[(size_t)(i - 1) * nx + j])) d = d + cellnew[5 * ((size_t)(i - 1) * nx + j) + 3];
                if (i < ny && (!active || active[(size_t)i * nx + j])) d = d + cellnew[5 * ((size_t)i * nx + j) + 2];
                v[(size_t)i * nx + j] += 0.5 * d;
            }
            if (i < ny && j < nx) {
                size_t ci = (size_t)i * nx + j;
                double d = (!active || active[ci]) ? cellnew[5 * ci + 4] : 0.0;
                H[ci] += 0.5 * d;
            }
        }
    }
}

/* Activation of the local smoother (vanka.cu:870-888 with grid.py:182-188
 * tiling): cell (i, j) is smoothed iff any cell of its 16x16 launch-block
 * footprint (block (i/14, j/14), rows 14by-1..14by+14, cols 14bx-1..14bx+14)
 * has error_mask > 0.5. */
void or_local_activation(int ny, int nx, const double *error_mask, unsigned char *active)
{
    int nby = ny / 14 + 1, nbx = nx / 14 + 1;
    unsigned char *blk = (unsigned char *)calloc((size_t)nby * nbx, 1);
    for (int by = 0; by < nby; by++)
        for (int bx = 0; bx < nbx; bx++) {
            int on = 0;
            for (int i = 14 * by - 1; i <= 14 * by + 14 && !on; i++)
                for (int j = 14 * bx - 1; j <= 14 * bx + 14; j++)
                    if (i >= 0 && i < ny && j >= 0 && j < nx && error_mask[(size_t)i * nx + j] > 0.5) { on = 1; break; }
            blk[(size_t)by * nbx + bx] = (unsigned char)on;
        }
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++)
            active[(size_t)i * nx + j] = blk[(size_t)(i / 14) * nbx + j / 14];
    free(blk);
}

/* error_mask = (|r_u right| + |r_u left| + |r_v bottom| + |r_v top| + |FB(r_H, H - gamma)|) > 0.01
 * (solver.py:148-154) */
void or_error_mask(const OGeom *g, const double *r, const double *U, const double *gamma, double *error_mask)
{
    int ny = g->ny, nx = g->nx;
    const double *r_u = r, *r_v = r + NU(g), *r_H = r + NU(g) + NV(g);
    const double *H = U + NU(g) + NV(g);
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            size_t ci = (size_t)i * nx + j;
            double a = r_H[ci], b = H[ci] - gamma[ci];
            double fb = a + b - sqrt(a * a + b * b);
            double te = fabs(r_u[(size_t)i * (nx + 1) + j + 1]) + fabs(r_u[(size_t)i * (nx + 1) + j]) +
                        fabs(r_v[(size_t)(i + 1) * nx + j]) + fabs(r_v[(size_t)i * nx + j]) + fabs(fb);
            error_mask[ci] = te > 0.01 ? 1.0 : 0.0;
        }
}

/* Adjoint red-black Vanka pass (vanka.cu:441-840): for cells of colour
 * (i+j)%2 == color solve J_loc^T dl = r_adj and write lambda_out. */
void or_vanka_adjoint_pass(const OGeom *g, const OParams *p, double *lam_out, const double *lam,
                           const double *mask, const double *r_adj, const double *U,
                           const double *bed, const double *B, const double *beta,
                           const double *gamma, int color, double omega, double *eta)
{
    int ny = g->ny, nx = g->nx;
    size_t nu = NU(g), nv = NV(g);
    or_eta(g, p, eta, U, B);
    VCtx c = {g, p, U, U + nu, U + nu + nv, NULL, NULL, NULL, bed, beta, gamma, eta};
    const double *lu = lam, *lv = lam + nu, *lH = lam + nu + nv;
    const double *ru = r_adj, *rv = r_adj + nu, *rH = r_adj + nu + nv;
    double *ou = lam_out, *ov = lam_out + nu, *oH = lam_out + nu + nv;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; i++) {
        for (int j = 0; j < nx; j++) {
            if ((i + j) % 2 != color) continue;
            double J[25], r[5], JT[25], rhs[5], dl[5];
            int masked = CE(mask, i, j, ny, nx) > 0.5;
            vanka_assemble(&c, i, j, VF(c.u, i, j, ny, nx), VF(c.u, i, j + 1, ny, nx), HF(c.v, i, j, ny, nx),
                           HF(c.v, i + 1, j, ny, nx), CE(c.H, i, j, ny, nx), CE(gamma, i, j, ny, nx), 1, masked, J, r);
            for (int a = 0; a < 5; a++)
                for (int b = 0; b < 5; b++) JT[a * 5 + b] = J[b * 5 + a];
            rhs[0] = VF(ru, i, j, ny, nx);
            rhs[1] = VF(ru, i, j + 1, ny, nx);
            rhs[2] = HF(rv, i, j, ny, nx);
            rhs[3] = HF(rv, i + 1, j, ny, nx);
            rhs[4] = CE(rH, i, j, ny, nx);
            or_lu5(JT, rhs, dl);
            ou[(size_t)i * (nx + 1) + j] = VF(lu, i, j, ny, nx) + 0.5 * omega * dl[0];
            ou[(size_t)i * (nx + 1) + j + 1] = VF(lu, i, j + 1, ny, nx) + 0.5 * omega * dl[1];
            ov[(size_t)i * nx + j] = HF(lv, i, j, ny, nx) + 0.5 * omega * dl[2];
            ov[(size_t)(i + 1) * nx + j] = HF(lv, i + 1, j, ny, nx) + 0.5 * omega * dl[3];
            oH[(size_t)i * nx + j] = CE(lH, i, j, ny, nx) + omega * dl[4];
        }
    }
}

/* grad_beta += lambda . dR/dbeta (grad.cu:1-74).  quirk != 0 reproduces the
 * reference's skipped right/bottom-cell terms at j = nx-1 and i = ny-1
 * (grad.cu:53,71) and drops the ghost-cell terms of boundary faces;
 * quirk == 0 gives the exact transpose (ghost terms go to the clamped cell). */
void or_grad_beta(const OGeom *g, const OParams *p, double *grad, const double *U, const double *lam,
                  const double *bed, const double *beta, int quirk)
{
    int ny = g->ny, nx = g->nx;
    const double *u = U, *v = U + NU(g), *H = U + NU(g) + NV(g);
    const double *lu = lam, *lv = lam + NU(g);
    memset(grad, 0, NC(g) * sizeof(double));
    for (int i = 0; i <= ny; i++)
        for (int j = 0; j <= nx; j++) {
            if (i < ny) {
                TauB t = tau_b(VF(u, i, j, ny, nx), CE(H, i, j - 1, ny, nx), CE(H, i, j, ny, nx),
                               CE(bed, i, j - 1, ny, nx), CE(bed, i, j, ny, nx),
                               CE(beta, i, j - 1, ny, nx), CE(beta, i, j, ny, nx), p);
                double l = VF(lu, i, j, ny, nx);
                if (quirk) {
                    if (j > 0) grad[(size_t)i * nx + j - 1] += l * t.d_beta_l;
                    if (j < nx - 1) grad[(size_t)i * nx + j] += l * t.d_beta_r;
                } else {   /* boundary faces read clamped (ghost) beta: route to the clamped cell */
                    grad[(size_t)i * nx + clampi(j - 1, 0, nx - 1)] += l * t.d_beta_l;
                    grad[(size_t)i * nx + clampi(j, 0, nx - 1)] += l * t.d_beta_r;
                }
            }
            if (j < nx) {
                TauB t = tau_b(HF(v, i, j, ny, nx), CE(H, i - 1, j, ny, nx), CE(H, i, j, ny, nx),
                               CE(bed, i - 1, j, ny, nx), CE(bed, i, j, ny, nx),
                               CE(beta, i - 1, j, ny, nx), CE(beta, i, j, ny, nx), p);
                double l = HF(lv, i, j, ny, nx);
                if (quirk) {
                    if (i > 0) grad[(size_t)(i - 1) * nx + j] += l * t.d_beta_l;
                    if (i < ny - 1) grad[(size_t)i * nx + j] += l * t.d_beta_r;
                } else {
                    grad[(size_t)clampi(i - 1, 0, ny - 1) * nx + j] += l * t.d_beta_l;
                    grad[(size_t)clampi(i, 0, ny - 1) * nx + j] += l * t.d_beta_r;
                }
            }
        }
}

/* dL/dB = lambda . dR/dB (no reference counterpart: glide/physics.py:211-242
 * returns dL/dbeta only).  B enters R only through eta = 0.5 B (...)^p
 * (viscosity.cu:61-112), so d eta_c / d B_c = eta_c / B_c, and eta_c enters
 * only through eta*H in sigma_xx and sigma_yy of cell c (stress.cu:44-102:
 * + on its left u face, - on its right; - on its top v face, + on its
 * bottom) and through the vertex average eta*H_v = 1/4 sum eta_k H_k in
 * sigma_xy of its interior corner vertices (stress.cu:160-183: vertex (a, b)
 * enters r_u(a, b) +, r_u(a-1, b) -, r_v(a, b) -, r_v(a, b-1) +). */
void or_grad_B(const OGeom *g, const OParams *p, double *grad, const double *U, const double *lam,
               const double *B)
{
    int ny = g->ny, nx = g->nx;
    double dx_inv = 1.0 / g->dx;
    const double *u = U, *v = U + NU(g), *H = U + NU(g) + NV(g);
    const double *lu = lam, *lv = lam + NU(g);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            double eta = eta_cell(i, j, u, v, B, p->n, p->eps_reg, g->dx, ny, nx);
            double u_l = VF(u, i, j, ny, nx), u_r = VF(u, i, j + 1, ny, nx);
            double v_t = HF(v, i, j, ny, nx), v_b = HF(v, i + 1, j, ny, nx);
            double exx = fma(2.0 * (u_r - u_l), dx_inv, (v_t - v_b) * dx_inv);
            double eyy = fma(u_r - u_l, dx_inv, 2.0 * (v_t - v_b) * dx_inv);
            double acc = 2.0 * exx * (VF(lu, i, j, ny, nx) - VF(lu, i, j + 1, ny, nx));
            acc = fma(2.0 * eyy, HF(lv, i + 1, j, ny, nx) - HF(lv, i, j, ny, nx), acc);
            for (int k = 0; k < 4; k++) {
                int a = i + (k >> 1), b = j + (k & 1);
                if (a < 1 || a > ny - 1 || b < 1 || b > nx - 1) continue;
                double exy = 0.5 * fma(VF(u, a - 1, b, ny, nx) - VF(u, a, b, ny, nx), dx_inv,
                                       (HF(v, a, b, ny, nx) - HF(v, a, b - 1, ny, nx)) * dx_inv);
                double comb = (VF(lu, a, b, ny, nx) - VF(lu, a - 1, b, ny, nx)) -
                              (HF(lv, a, b, ny, nx) - HF(lv, a, b - 1, ny, nx));
                acc = fma(0.5 * exy, comb, acc);
            }
            grad[(size_t)i * nx + j] = acc * CE(H, i, j, ny, nx) * dx_inv * (eta / CE(B, i, j, ny, nx));
        }
}

/* ------------------------------------------------------------------ */
/* Transfer operators (utility.cu)                                     */
/* ------------------------------------------------------------------ */
/* utility.cu:1-19 */
void or_restrict_u(const double *uf, double *uc, int nyc, int nxc)
{
    for (int J = 0; J < nyc; J++)
        for (int I = 0; I <= nxc; I++)
            uc[(size_t)J * (nxc + 1) + I] = 0.5 * (uf[(size_t)(2 * J) * (2 * nxc + 1) + 2 * I] +
                                                   uf[(size_t)(2 * J + 1) * (2 * nxc + 1) + 2 * I]);
}
/* utility.cu:125-143 */
void or_restrict_v(const double *vf, double *vc, int nyc, int nxc)
{
    for (int J = 0; J <= nyc; J++)
        for (int I = 0; I < nxc; I++)
            vc[(size_t)J * nxc + I] = 0.5 * (vf[(size_t)(2 * J) * (2 * nxc) + 2 * I] +
                                             vf[(size_t)(2 * J) * (2 * nxc) + 2 * I + 1]);
}
/* utility.cu:56-75 */
void or_restrict_cell(const double *hf, double *hc, int nyc, int nxc)
{
    int nxf = 2 * nxc;
    for (int J = 0; J < nyc; J++)
        for (int I = 0; I < nxc; I++)
            hc[(size_t)J * nxc + I] = 0.25 * (hf[(size_t)(2 * J) * nxf + 2 * I] + hf[(size_t)(2 * J) * nxf + 2 * I + 1] +
                                              hf[(size_t)(2 * J + 1) * nxf + 2 * I] + hf[(size_t)(2 * J + 1) * nxf + 2 * I + 1]);
}
/* utility.cu:77-101 */
void or_restrict_max(const double *hf, double *hc, int nyc, int nxc)
{
    int nxf = 2 * nxc;
    for (int J = 0; J < nyc; J++)
        for (int I = 0; I < nxc; I++) {
            double a = fmax(hf[(size_t)(2 * J) * nxf + 2 * I], hf[(size_t)(2 * J) * nxf + 2 * I + 1]);
            double b = fmax(hf[(size_t)(2 * J + 1) * nxf + 2 * I], hf[(size_t)(2 * J + 1) * nxf + 2 * I + 1]);
            hc[(size_t)J * nxc + I] = fmax(a, b);
        }
}
/* utility.cu:104-123 (injection); sign = -1 prolongates -h_coarse */
void or_prolong_inject(const double *hc, double *hf, int nyf, int nxf, double sign)
{
    for (int j = 0; j < nyf; j++)
        for (int i = 0; i < nxf; i++)
            hf[(size_t)j * nxf + i] = sign * hc[(size_t)(j / 2) * (nxf / 2) + i / 2];
}
static inline double bilin(double ty, double tx, double v00, double v01, double v10, double v11)
{
    return (1.0 - ty) * ((1.0 - tx) * v00 + tx * v01) + ty * ((1.0 - tx) * v10 + tx * v11);
}
/* utility.cu:180-232 */
void or_prolong_u(const double *uc, double *uf, int nyf, int nxf)
{
    int nyc = nyf / 2, nxc = nxf / 2;
    for (int j = 0; j < nyf; j++)
        for (int i = 0; i <= nxf; i++) {
            double Jf = (j - 0.5) * 0.5, If = i * 0.5;
            Jf = fmax(0.0, fmin(Jf, (double)(nyc - 1)));
            If = fmax(0.0, fmin(If, (double)nxc));
            int Jl = (int)Jf, Il = (int)If;
            int Jh = Jl + 1 < nyc - 1 ? Jl + 1 : nyc - 1;
            int Ih = Il + 1 < nxc ? Il + 1 : nxc;
            double ty = Jf - Jl, tx = If - Il;
            int s = nxc + 1;
            uf[(size_t)j * (nxf + 1) + i] = bilin(ty, tx, uc[(size_t)Jl * s + Il], uc[(size_t)Jl * s + Ih],
                                                  uc[(size_t)Jh * s + Il], uc[(size_t)Jh * s + Ih]);
        }
}
/* utility.cu:235-287 */
void or_prolong_v(const double *vc, double *vf, int nyf, int nxf)
{
    int nyc = nyf / 2, nxc = nxf / 2;
    for (int j = 0; j <= nyf; j++)
        for (int i = 0; i < nxf; i++) {
            double Jf = j * 0.5, If = (i - 0.5) * 0.5;
            Jf = fmax(0.0, fmin(Jf, (double)nyc));
            If = fmax(0.0, fmin(If, (double)(nxc - 1)));
            int Jl = (int)Jf, Il = (int)If;
            int Jh = Jl + 1 < nyc ? Jl + 1 : nyc;
            int Ih = Il + 1 < nxc - 1 ? Il + 1 : nxc - 1;
            double ty = Jf - Jl, tx = If - Il;
            int s = nxc;
            vf[(size_t)j * nxf + i] = bilin(ty, tx, vc[(size_t)Jl * s + Il], vc[(size_t)Jl * s + Ih],
                                            vc[(size_t)Jh * s + Il], vc[(size_t)Jh * s + Ih]);
        }
}
/* utility.cu:289-337 */
void or_prolong_H(const double *hc, double *hf, int nyf, int nxf)
{
    int nyc = nyf / 2, nxc = nxf / 2;
    for (int j = 0; j < nyf; j++)
        for (int i = 0; i < nxf; i++) {
            double Jf = (j - 0.5) * 0.5, If = (i - 0.5) * 0.5;
            Jf = fmax(0.0, fmin(Jf, (double)(nyc - 1)));
            If = fmax(0.0, fmin(If, (double)(nxc - 1)));
            int Jl = (int)Jf, Il = (int)If;
            int Jh = Jl + 1 < nyc - 1 ? Jl + 1 : nyc - 1;
            int Ih = Il + 1 < nxc - 1 ? Il + 1 : nxc - 1;
            double ty = Jf - Jl, tx = If - Il;
            hf[(size_t)j * nxf + i] = bilin(ty, tx, hc[(size_t)Jl * nxc + Il], hc[(size_t)Jl * nxc + Ih],
                                            hc[(size_t)Jh * nxc + Il], hc[(size_t)Jh * nxc + Ih]);
        }
}

/* ------------------------------------------------------------------ */
/* Hierarchy, FASCD and adjoint V-cycles (grid.py, solver.py,          */
/* physics.py)                                                         */
/* ------------------------------------------------------------------ */
#define OMAXLEV 12
typedef struct {
    OGeom g; OParams p;
    size_t nu, nv, nc, nt;
    double *U, *f, *r, *F, *Z, *w, *y, *z;
    double *Lam, *Lam_out, *f_adj, *r_adj, *l;
    double *H_prev, *bed, *B, *beta, *smb, *gamma, *chi, *phi, *mask, *error_mask;
    double *eta, *cellnew;
    Dual *etad;
    unsigned char *active;
} OLevel;

typedef struct {
    int nlev;
    double thklim;
    OLevel L[OMAXLEV];
    int sweeps_pre, sweeps_post, sweeps_coarse, local_passes, local_sweeps, n_newton;
    int adj_sweeps, adj_sweeps_coarse;
    double bweight;   /* weight of the boundary-face rows of restricted residuals (1 = solver.py) */
} OHier;

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static void level_alloc(OLevel *L, int ny, int nx, double dx, const OParams *p)
{
    memset(L, 0, sizeof(*L));
    L->g.ny = ny; L->g.nx = nx; L->g.dx = dx; L->g.dt = 1.0;
    L->p = *p;
    L->nu = (size_t)ny * (nx + 1); L->nv = (size_t)(ny + 1) * nx; L->nc = (size_t)ny * nx;
    L->nt = L->nu + L->nv + L->nc;
    L->U = zalloc(L->nt); L->f = zalloc(L->nt); L->r = zalloc(L->nt); L->F = zalloc(L->nt);
    L->Z = zalloc(L->nt); L->w = zalloc(L->nt); L->y = zalloc(L->nt); L->z = zalloc(L->nt);
    L->Lam = zalloc(L->nt); L->Lam_out = zalloc(L->nt); L->f_adj = zalloc(L->nt);
    L->r_adj = zalloc(L->nt); L->l = zalloc(L->nt);
    L->H_prev = zalloc(L->nc); L->bed = zalloc(L->nc); L->B = zalloc(L->nc); L->beta = zalloc(L->nc);
    L->smb = zalloc(L->nc); L->gamma = zalloc(L->nc); L->chi = zalloc(L->nc); L->phi = zalloc(L->nc);
    L->mask = zalloc(L->nc); L->error_mask = zalloc(L->nc); L->eta = zalloc(L->nc);
    L->cellnew = zalloc(5 * L->nc);
    L->etad = (Dual *)calloc(L->nc, sizeof(Dual));
    L->active = (unsigned char *)calloc(L->nc, 1);
}

static void level_free(OLevel *L)
{
    double *ptrs[] = {L->U, L->f, L->r, L->F, L->Z, L->w, L->y, L->z, L->Lam, L->Lam_out, L->f_adj,
                      L->r_adj, L->l, L->H_prev, L->bed, L->B, L->beta, L->smb, L->gamma, L->chi,
                      L->phi, L->mask, L->error_mask, L->eta, L->cellnew};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
    free(L->etad);
    free(L->active);
}

/* physics.py:49-80 + grid.py:167-180: children inherit n, eps_reg,
 * water_drag, calving_rate but fall back to gl_sigmoid_c = 0.1 and
 * gl_derivatives = False. */
OHier *oh_create(int ny, int nx, double dx, int nlev, const OParams *p, double thklim)
{
    if (nlev < 1 || nlev > OMAXLEV) return NULL;
    OHier *h = (OHier *)calloc(1, sizeof(OHier));
    h->nlev = nlev; h->thklim = thklim;
    h->sweeps_pre = 10; h->sweeps_post = 10; h->sweeps_coarse = 200;
    h->local_passes = 10; h->local_sweeps = 10; h->n_newton = 10;
    h->adj_sweeps = 10; h->adj_sweeps_coarse = 100;
    h->bweight = 1.0;
    OParams pc = *p;
    pc.gl_sigmoid_c = 0.1; pc.gl_derivatives = 0;
    for (int k = 0; k < nlev; k++)
        level_alloc(&h->L[k], ny >> k, nx >> k, dx * (double)(1 << k), k == 0 ? p : &pc);
    return h;
}

void oh_destroy(OHier *h)
{
    if (!h) return;
    for (int k = 0; k < h->nlev; k++) level_free(&h->L[k]);
    free(h);
}

void oh_set_schedule(OHier *h, int pre, int post, int coarse, int local_passes, int local_sweeps,
                     int n_newton, int adj_sweeps, int adj_coarse)
{
    h->sweeps_pre = pre; h->sweeps_post = post; h->sweeps_coarse = coarse;
    h->local_passes = local_passes; h->local_sweeps = local_sweeps; h->n_newton = n_newton;
    h->adj_sweeps = adj_sweeps; h->adj_sweeps_coarse = adj_coarse;
}

/* Weight of the boundary-face rows (u faces j = 0 and j = nx, v faces
 * i = 0 and i = ny) of every restricted residual.  solver.py:34-44 restricts
 * them like interior rows (weight 1).  A boundary face has one neighbouring
 * cell, so its momentum row is a one-sided stress sigma / dx (scaling as
 * 1/h) where interior rows are stress differences (1/h^2); restricted with
 * weight 1 the coarse correction of those rows is twice too strong.  0.5
 * restores the interior rows' scaling. */
void oh_set_boundary_weight(OHier *h, double w) { h->bweight = w; }

/* dst (a coarse state-sized vector) *= w on its boundary-face rows */
static void weight_boundary_faces(OLevel *C, double *dst, double w)
{
    int ny = C->g.ny, nx = C->g.nx;
    double *u = dst, *v = dst + C->nu;
    for (int i = 0; i < ny; i++) {
        u[(size_t)i * (nx + 1)] *= w;
        u[(size_t)i * (nx + 1) + nx] *= w;
    }
    for (int j = 0; j < nx; j++) {
        v[j] *= w;
        v[(size_t)ny * nx + j] *= w;
    }
}

/* field access for the ctypes wrapper */
double *oh_field(OHier *h, int lev, int which)
{
    OLevel *L = &h->L[lev];
    switch (which) {
    case 0: return L->U; case 1: return L->f; case 2: return L->r; case 3: return L->F;
    case 4: return L->w; case 5: return L->y; case 6: return L->z; case 7: return L->Lam;
    case 8: return L->f_adj; case 9: return L->r_adj; case 10: return L->l; case 11: return L->H_prev;
    case 12: return L->bed; case 13: return L->B; case 14: return L->beta; case 15: return L->smb;
    case 16: return L->gamma; case 17: return L->chi; case 18: return L->phi; case 19: return L->mask;
    case 20: return L->error_mask;
    default: return NULL;
    }
}

static void fill(double *a, size_t n, double v) { for (size_t k = 0; k < n; k++) a[k] = v; }

static void restrict_vec(const OLevel *F, OLevel *C, const double *src, double *dst)
{
    or_restrict_u(src, dst, C->g.ny, C->g.nx);
    or_restrict_v(src + F->nu, dst + C->nu, C->g.ny, C->g.nx);
    or_restrict_cell(src + F->nu + F->nv, dst + C->nu + C->nv, C->g.ny, C->g.nx);
}

/* physics.py:82-99, 131-139 */
void oh_set_geometry(OHier *h, const double *bed, const double *H)
{
    OLevel *L0 = &h->L[0];
    memcpy(L0->bed, bed, L0->nc * sizeof(double));
    memcpy(L0->U + L0->nu + L0->nv, H, L0->nc * sizeof(double));
    memcpy(L0->H_prev, H, L0->nc * sizeof(double));
    fill(L0->gamma, L0->nc, h->thklim);
    for (int k = 0; k + 1 < h->nlev; k++) {
        OLevel *P = &h->L[k], *C = &h->L[k + 1];
        or_restrict_cell(P->bed, C->bed, C->g.ny, C->g.nx);
        or_restrict_cell(P->U + P->nu + P->nv, C->U + C->nu + C->nv, C->g.ny, C->g.nx);
        or_restrict_cell(P->H_prev, C->H_prev, C->g.ny, C->g.nx);
        fill(C->gamma, C->nc, h->thklim);
    }
}

/* solver.py:53-68 */
static void restrict_parameters(OLevel *P, OLevel *C)
{
    or_restrict_cell(P->bed, C->bed, C->g.ny, C->g.nx);
    or_restrict_cell(P->B, C->B, C->g.ny, C->g.nx);
    or_restrict_cell(P->beta, C->beta, C->g.ny, C->g.nx);
    or_restrict_cell(P->H_prev, C->H_prev, C->g.ny, C->g.nx);
    or_restrict_cell(P->smb, C->smb, C->g.ny, C->g.nx);
}

/* physics.py:101-129 */
void oh_set_parameters(OHier *h, const double *B, const double *beta, const double *smb)
{
    OLevel *L0 = &h->L[0];
    if (B) memcpy(L0->B, B, L0->nc * sizeof(double));
    if (beta) memcpy(L0->beta, beta, L0->nc * sizeof(double));
    if (smb) memcpy(L0->smb, smb, L0->nc * sizeof(double));
    for (int k = 0; k + 1 < h->nlev; k++) restrict_parameters(&h->L[k], &h->L[k + 1]);
}

void oh_restrict_parameters(OHier *h)
{
    for (int k = 0; k + 1 < h->nlev; k++) restrict_parameters(&h->L[k], &h->L[k + 1]);
}

static void lev_residual(OLevel *L, double *out, const double *f)
{
    or_residual(&L->g, &L->p, out, L->U, f, L->bed, L->B, L->beta, L->mask, L->gamma, L->eta);
}

static void lev_sweep(OHier *h, OLevel *L, int n, const unsigned char *active)
{
    for (int s = 0; s < n; s++)
        or_vanka_sweep(&L->g, &L->p, L->U, L->f, L->bed, L->B, L->beta, L->gamma, active,
                       h->n_newton, L->eta, L->cellnew);
}

/* solver.py:71-157 */
static void fascd(OHier *h, int k, int finest)
{
    OLevel *L = &h->L[k];
    double *H = L->U + L->nu + L->nv, *wH = L->w + L->nu + L->nv;
    if (finest) {
        memcpy(L->w, L->U, L->nt * sizeof(double));
        for (size_t c = 0; c < L->nc; c++) L->chi[c] = L->gamma[c] - H[c];
    }
    if (k == h->nlev - 1) {
        for (size_t c = 0; c < L->nc; c++) L->gamma[c] = wH[c] + L->chi[c];
        lev_sweep(h, L, h->sweeps_coarse, NULL);
        fill(L->gamma, L->nc, h->thklim);
        return;
    }
    OLevel *C = &h->L[k + 1];
    or_restrict_max(L->chi, C->chi, C->g.ny, C->g.nx);
    or_prolong_inject(C->chi, L->phi, L->g.ny, L->g.nx, -1.0);
    for (size_t c = 0; c < L->nc; c++) L->phi[c] += L->chi[c];
    for (size_t c = 0; c < L->nc; c++) L->gamma[c] = wH[c] + L->phi[c];
    lev_sweep(h, L, h->sweeps_pre, NULL);
    fill(L->gamma, L->nc, h->thklim);
    for (size_t t = 0; t < L->nt; t++) L->y[t] = L->U[t] - L->w[t];
    restrict_vec(L, C, L->U, C->U);
    memcpy(C->w, C->U, C->nt * sizeof(double));
    lev_residual(L, L->r, L->f);
    restrict_vec(L, C, L->r, C->r);
    if (h->bweight != 1.0) weight_boundary_faces(C, C->r, h->bweight);
    lev_residual(C, C->F, C->Z);
    for (size_t t = 0; t < C->nt; t++) C->f[t] = C->F[t] - C->r[t];
    fascd(h, k + 1, 0);
    for (size_t t = 0; t < C->nt; t++) C->z[t] = C->U[t] - C->w[t];
    or_prolong_u(C->z, L->z, L->g.ny, L->g.nx);
    or_prolong_v(C->z + C->nu, L->z + L->nu, L->g.ny, L->g.nx);
    or_prolong_inject(C->z + C->nu + C->nv, L->z + L->nu + L->nv, L->g.ny, L->g.nx, 1.0);
    for (size_t t = 0; t < L->nt; t++) L->z[t] += L->y[t];
    for (size_t t = 0; t < L->nt; t++) L->U[t] = L->w[t] + L->z[t];
    for (size_t c = 0; c < L->nc; c++) L->gamma[c] = wH[c] + L->chi[c];
    lev_sweep(h, L, h->sweeps_post, NULL);
    for (int pass = 0; pass < h->local_passes; pass++) {
        lev_residual(L, L->r, L->f);
        or_error_mask(&L->g, L->r, L->U, L->gamma, L->error_mask);
        or_local_activation(L->g.ny, L->g.nx, L->error_mask, L->active);
        lev_sweep(h, L, h->local_sweeps, L->active);
    }
    fill(L->gamma, L->nc, h->thklim);
}

void oh_vcycle(OHier *h) { fascd(h, 0, 1); }

/* ||(r_u, r_v, FB(r_H, H - gamma))||_2 at the finest level (physics.py:176-203) */
double oh_residual_norm(OHier *h)
{
    OLevel *L = &h->L[0];
    lev_residual(L, L->r, L->f);
    const double *H = L->U + L->nu + L->nv, *rH = L->r + L->nu + L->nv;
    double s = 0.0;
    for (size_t t = 0; t < L->nu + L->nv; t++) s += L->r[t] * L->r[t];
    for (size_t c = 0; c < L->nc; c++) {
        double a = rH[c], b = H[c] - L->gamma[c];
        double fb = a + b - sqrt(a * a + b * b);
        s += fb * fb;
    }
    return sqrt(s);
}

/* physics.py:141-209.  Runs n_vcycles V-cycles (or, with tol > 0, until the
 * relative residual falls to tol, at most max(n_vcycles, 1) cycles).  Writes
 * the relative residual after each cycle to hist (if given) and returns the
 * number of cycles run. */
int oh_forward(OHier *h, double dt, int n_vcycles, double tol, double *hist, int update_geometry)
{
    for (int k = 0; k < h->nlev; k++) h->L[k].g.dt = dt;
    OLevel *L = &h->L[0];
    double *fH = L->f + L->nu + L->nv;
    for (size_t c = 0; c < L->nc; c++) fH[c] = L->H_prev[c] / dt + L->smb[c];
    double r0 = (tol > 0.0 || hist) ? oh_residual_norm(h) : 0.0;
    int done = 0;
    for (int it = 0; it < n_vcycles; it++) {
        oh_vcycle(h);
        done++;
        if (tol > 0.0 || hist) {
            double rel = r0 > 0.0 ? oh_residual_norm(h) / r0 : 0.0;
            if (hist) hist[it] = rel;
            if (tol > 0.0 && rel <= tol) break;
        }
    }
    if (update_geometry) memcpy(L->H_prev, L->U + L->nu + L->nv, L->nc * sizeof(double));
    return done;
}

/* grid.py:332-334 */
static void compute_mask(OLevel *L)
{
    const double *H = L->U + L->nu + L->nv;
    for (size_t c = 0; c < L->nc; c++) L->mask[c] = (H[c] <= L->gamma[c] + 0.1) ? 1.0 : 0.0;
}

static void lev_vjp(OLevel *L)
{
    or_vjp(&L->g, &L->p, L->l, L->U, L->Lam, L->bed, L->B, L->beta, L->mask, L->gamma, L->etad);
}

/* grid.py:286-330 */
static void adjoint_sweeps(OLevel *L, int n, double omega)
{
    for (int s = 0; s < n; s++) {
        for (int color = 0; color < 2; color++) {
            lev_vjp(L);
            for (size_t t = 0; t < L->nt; t++) L->r_adj[t] = L->f_adj[t] - L->l[t];
            memcpy(L->Lam_out, L->Lam, L->nt * sizeof(double));
            or_vanka_adjoint_pass(&L->g, &L->p, L->Lam_out, L->Lam, L->mask, L->r_adj, L->U, L->bed,
                                  L->B, L->beta, L->gamma, color, omega, L->eta);
            memcpy(L->Lam, L->Lam_out, L->nt * sizeof(double));
        }
    }
}

/* solver.py:160-220 */
static void adjoint_vc(OHier *h, int k)
{
    OLevel *L = &h->L[k];
    if (k == h->nlev - 1) {
        compute_mask(L);
        adjoint_sweeps(L, h->adj_sweeps_coarse, 1.0);
        fill(L->mask, L->nc, 0.0);
        return;
    }
    OLevel *C = &h->L[k + 1];
    compute_mask(L);
    adjoint_sweeps(L, h->adj_sweeps, 1.0);
    fill(L->mask, L->nc, 0.0);
    lev_vjp(L);
    for (size_t t = 0; t < L->nt; t++) L->r_adj[t] = L->f_adj[t] - L->l[t];
    restrict_vec(L, C, L->Lam, C->Lam);
    restrict_parameters(L, C);
    fill(C->f_adj, C->nt, 0.0);
    restrict_vec(L, C, L->r_adj, C->f_adj);
    fill(C->Lam, C->nt, 0.0);
    fill(C->Lam_out, C->nt, 0.0);
    adjoint_vc(h, k + 1);
    fill(L->z, L->nt, 0.0);
    or_prolong_u(C->Lam, L->z, L->g.ny, L->g.nx);
    or_prolong_v(C->Lam + C->nu, L->z + L->nu, L->g.ny, L->g.nx);
    or_prolong_H(C->Lam + C->nu + C->nv, L->z + L->nu + L->nv, L->g.ny, L->g.nx);
    for (size_t t = 0; t < L->nt; t++) L->Lam[t] += L->z[t];
    compute_mask(L);
    adjoint_sweeps(L, h->adj_sweeps, 1.0);
    fill(L->mask, L->nc, 0.0);
}

void oh_adjoint_vcycle(OHier *h) { adjoint_vc(h, 0); }

/* physics.py:211-242 (+ grid.py:336-355).  dL arrays are u, v, H fields. */
void oh_adjoint(OHier *h, const double *dLdu, const double *dLdv, const double *dLdH,
                int n_vcycles, double *grad, int quirk)
{
    OLevel *L = &h->L[0];
    for (size_t t = 0; t < L->nu; t++) L->f_adj[t] = -dLdu[t];
    for (size_t t = 0; t < L->nv; t++) L->f_adj[L->nu + t] = -dLdv[t];
    for (size_t t = 0; t < L->nc; t++) L->f_adj[L->nu + L->nv + t] = -dLdH[t];
    fill(L->Lam, L->nt, 0.0);
    fill(L->Lam_out, L->nt, 0.0);
    for (int c = 0; c < n_vcycles; c++) oh_adjoint_vcycle(h);
    compute_mask(L);
    or_grad_beta(&L->g, &L->p, grad, L->U, L->Lam, L->bed, L->beta, quirk);
    fill(L->mask, L->nc, 0.0);
}

/* adjoint residual ||f_adj - J^T lambda|| at the finest level, mask as used by the smoother */
double oh_adjoint_residual_norm(OHier *h)
{
    OLevel *L = &h->L[0];
    compute_mask(L);
    lev_vjp(L);
    fill(L->mask, L->nc, 0.0);
    double s = 0.0;
    for (size_t t = 0; t < L->nt; t++) { double d = L->f_adj[t] - L->l[t]; s += d * d; }
    return sqrt(s);
}

int oh_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oh_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* deterministic math exported for the accuracy test */
double or_exp(double x) { return OEXP(x); }
double or_pow(double x, double p) { return OPOW(x, p); }
#ifndef ORACLE_LIBM
double or_log(double x) { return dm_log(x); }
#else
double or_log(double x) { return log(x); }
#endif
