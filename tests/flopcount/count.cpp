// Counts FP64 operations of the solver's per-cell code by running it on the
// host with a counting scalar type (the headers are included after a textual
// double -> R substitution made by tests/test_flop_count.py).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

static long g_flops = 0;
struct CD {
    double v;
    CD() : v(0) {}
    CD(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
};
#define OP(sym)                                                                        \
    inline CD operator sym(CD a, CD b) { g_flops++; return CD(a.v sym b.v); }          \
    inline CD operator sym(CD a, double b) { g_flops++; return CD(a.v sym b); }        \
    inline CD operator sym(double a, CD b) { g_flops++; return CD(a sym b.v); }
OP(+) OP(-) OP(*) OP(/)
#undef OP
#define AOP(sym, op)                                                                   \
    inline CD &operator sym(CD &a, CD b) { g_flops++; a.v = a.v op b.v; return a; }    \
    inline CD &operator sym(CD &a, double b) { g_flops++; a.v = a.v op b; return a; }
AOP(+=, +) AOP(-=, -) AOP(*=, *) AOP(/=, /)
#undef AOP
inline CD operator-(CD a) { return CD(-a.v); }
#define CMP(sym)                                                                       \
    inline bool operator sym(CD a, CD b) { return a.v sym b.v; }                       \
    inline bool operator sym(CD a, double b) { return a.v sym b; }                     \
    inline bool operator sym(double a, CD b) { return a sym b.v; }
CMP(<) CMP(<=) CMP(>) CMP(>=) CMP(==) CMP(!=)
#undef CMP
inline CD fmin(CD a, CD b) { return CD(std::fmin(a.v, b.v)); }
inline CD fmax(CD a, CD b) { return CD(std::fmax(a.v, b.v)); }
inline CD fmin(CD a, double b) { return CD(std::fmin(a.v, b)); }
inline CD fmax(CD a, double b) { return CD(std::fmax(a.v, b)); }
inline CD fabs(CD a) { return CD(std::fabs(a.v)); }
inline CD copysign(double a, CD b) { return CD(std::copysign(a, b.v)); }
inline CD fma(CD a, CD b, CD c) { g_flops += 2; return CD(std::fma(a.v, b.v, c.v)); }
inline CD count_floor(CD a) { return CD(std::floor(a.v)); }
using R = CD;

#define __host__
#define __device__
#define __forceinline__ inline
#define __restrict__
#include "glide_hip.h"
#include "vanka_device.h"

using namespace glide;

int main()
{
    const int ny = 8, nx = 8, i = 3, j = 4;
    const size_t nu = ny * (nx + 1), nv = (ny + 1) * nx, nc = ny * nx;
    static R U[1000], B[100];
    for (size_t k = 0; k < nu + nv; k++) U[k] = R(50.0 + 13.0 * (double)((k * 7) % 11) - 40.0 * (double)(k % 3));
    for (size_t k = 0; k < nc; k++) { U[nu + nv + k] = R(700.0 + 10.0 * (double)(k % 5)); B[k] = R(23.9); }
    Params p{R(3.0), R(1e-5), R(1e-3), R(1.0), R(0.1), 0};
    const R *u = U, *v = U + nu;
    g_flops = 0;
    R e = eta_cell(i, j, u, v, B, p, R(1500.0), ny, nx);
    long eta_flops = g_flops;
    CellNbhd N;
    R *fields = (R *)&N;
    for (size_t k = 0; k < sizeof(CellNbhd) / sizeof(R); k++) fields[k] = R(1.0 + 0.01 * (double)k);
    N.e_c = N.e_l = N.e_r = N.e_t = N.e_b = N.e_tl = N.e_tr = N.e_bl = N.e_br = e;
    N.H_l = N.H_r = N.H_t = N.H_b = N.H_tl = N.H_tr = N.H_bl = N.H_br = R(700.0);
    N.bed_c = N.bed_l = N.bed_r = N.bed_t = N.bed_b = R(300.0);
    N.beta_c = N.beta_l = N.beta_r = N.beta_t = N.beta_b = R(0.1);
    g_flops = 0;
    nbhd_grounding(N, p);
    long setup_flops = g_flops;
    R J[25], r[5], d[5];
    R ul(40.0), ur(60.0), vt(-20.0), vb(10.0), Hc(700.0), th(0.1);
    g_flops = 0;
    vanka_assemble<0>(N, p, i, j, ny, nx, R(1.0) / R(1500.0), R(10.0), ul, ur, vt, vb, Hc, th, false, J, r);
    lu5_solve(J, r, d);
    g_flops -= 1;   // R(1)/R(1500) above is the caller's dx_inv, computed once per kernel
    ul -= 0.25 * d[0];
    ur -= 0.25 * d[1];
    vt -= 0.25 * d[2];
    vb -= 0.25 * d[3];
    Hc -= 0.25 * d[4];
    Hc = fmax(Hc, th);
    bool constrained = (Hc - R(10.0) * r[4]) <= th;
    printf("%ld %ld %d %ld\n", g_flops, eta_flops, (int)constrained, setup_flops);
    return 0;
}
