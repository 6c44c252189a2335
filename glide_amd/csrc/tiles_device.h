// tiles_device.h - tile bodies shared by the single-level and the batched
// (ensemble) kernels: the cell-viscosity LDS tile, the residual stencil, the
// forward Vanka smoother, its gather-apply pass and the local-smoother block
// activation.  Each body works on one (bx, by) tile of TX x TY points or
// cells with a (TX, TY) thread block, so a kernel may run it once per
// workgroup (one tile per block) or in a loop over many tiles (persistent
// batched launches).  The per-point arithmetic does not depend on the tile
// shape, so every shape gives bitwise-identical results.
#pragma once

#include "vanka_device.h"

namespace glide {

// records msg as glide_last_error() and returns code (host side)
int set_error(int code, const char *msg);

constexpr int RB = 14;  // reference launch-block interior (glide/grid.py:184-186)
constexpr int NXCD = 8; // XCDs per MI355X

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Persistent tile loop.  With at least as many workgroups as tiles each
// workgroup runs tile blockIdx.x (the hardware dispatcher balances them).
// Otherwise workgroups loop over tiles: workgroups are dispatched round-robin
// over the XCDs (workgroup b runs on XCD b % 8), the tiles are cut into chunks
// of up to TCH consecutive tiles (neighbours sharing halo rows; smaller when
// there are few tiles, so every XCD gets at least two chunks) dealt
// round-robin to the XCDs, and each XCD's workgroups walk its chunks' tiles.
// Halos then meet in one L2 while work clustered in space (the local
// smoother's active blocks) still spreads over all XCDs.  The launch must
// have at least 8 workgroups.
constexpr int TCH = 8;

__host__ __device__ inline size_t n_u(int ny, int nx) { return (size_t)ny * (nx + 1); }
__host__ __device__ inline size_t n_v(int ny, int nx) { return (size_t)(ny + 1) * nx; }
__host__ __device__ inline size_t n_c(int ny, int nx) { return (size_t)ny * nx; }
__host__ __device__ inline size_t n_t(int ny, int nx) { return n_u(ny, nx) + n_v(ny, nx) + n_c(ny, nx); }

template <class Body>
__device__ __forceinline__ void tile_loop(int total, const Body &body)
{
    const int G = gridDim.x, b = blockIdx.x;
    if (G >= total) {
        if (b < total) body(b);
        return;
    }
    const int tch = max(1, min(TCH, total / (2 * NXCD)));
    const int x = b % NXCD, lane = b / NXCD;
    const int nwg = (G - x + NXCD - 1) / NXCD;
    const int nchunks = (total + tch - 1) / tch;
    const int my_chunks = nchunks > x ? (nchunks - x + NXCD - 1) / NXCD : 0;
    for (int k = lane; k < my_chunks * tch; k += nwg) {
        const int t = (x + NXCD * (k / tch)) * tch + k % tch;
        if (t < total) body(t);
    }
}

// One tile per workgroup with the tiles dealt to the XCDs in chunks of up to
// TCH consecutive tiles, as tile_loop deals them: workgroup b (on XCD b % 8)
// takes tile (x + 8 c) tch + o for x = b % 8, c = (b / 8) / tch, o = (b / 8)
// % tch.  Over a grid of a multiple of 8 tch workgroups this is a bijection;
// workgroups whose tile is past the last one (t >= total) have no tile.
__device__ __forceinline__ int chunked_tile(int b, int total)
{
    const int tch = max(1, min(TCH, total / (2 * NXCD)));
    const int k = b / NXCD;
    return (b % NXCD + NXCD * (k / tch)) * tch + k % tch;
}
__host__ __device__ inline int chunked_grid(int total)
{
    const int tch = total / (2 * NXCD) < 1 ? 1 : (total / (2 * NXCD) > TCH ? TCH : total / (2 * NXCD));
    const int span = NXCD * tch;
    return (total + span - 1) / span * span;
}

// ---------------------------------------------------------------------------
// viscosity tile: cells [i0-1, i0+TY] x [j0-1, j0+TX] into LDS (0 outside)
// ---------------------------------------------------------------------------
template <int TX, int TY>
__device__ __forceinline__ void fill_eta_tile(double (*eta_s)[TX + 2], int i0, int j0, const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ B,
                                              const Params &p, double dx, int ny, int nx)
{
    int tid = threadIdx.y * TX + threadIdx.x;
    for (int k = tid; k < (TY + 2) * (TX + 2); k += TX * TY) {
        int a = k / (TX + 2), b = k - a * (TX + 2);
        int ci = i0 - 1 + a, cj = j0 - 1 + b;
        eta_s[a][b] = (ci >= 0 && ci < ny && cj >= 0 && cj < nx) ? eta_cell(ci, cj, u, v, B, p, dx, ny, nx) : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Residual r = F(U) - f (- sub)  (glide/cuda/residuals.cu:6-292), LDS-staged:
// the tile's u, v (two-ring halo), H, bed, beta (one-ring halo) are staged
// once; every cell's viscosity, grounding sigmoid, sigma_xx and sigma_yy and
// every vertex's sigma_xy are evaluated once into LDS; each point then sums
// its rows from LDS.  Every quantity is the same expression of the same
// (clamped) values as in the per-point form of the reference and the oracle,
// and each row sums its terms in the same order, so the bits are the same.
// ---------------------------------------------------------------------------
template <int TX, int TY>
struct ResidLDS {
    double u[TY + 4][TX + 4];   // u[i0-2+a][j0-2+b] (clamped)
    double v[TY + 4][TX + 4];   // v[i0-2+a][j0-2+b] (clamped)
    double H[TY + 2][TX + 2];   // cells [i0-1+a][j0-1+b] (clamped)
    double bed[TY + 2][TX + 2];
    double beta[TY + 2][TX + 2];
    double E[TY + 2][TX + 2];   // viscosity (0 outside the domain)
    double g[TY + 2][TX + 2];   // grounding sigmoid and its derivative
    double gd[TY + 2][TX + 2];
    double sxx[TY + 2][TX + 2]; // sigma_xx / sigma_yy residual parts
    double syy[TY + 2][TX + 2];
    double sxy[TY + 1][TX + 1]; // vertices [i0+a][j0+b]
};

// EDGE_FACES: the tiles cover the cells only (cdiv(nx, TX) x cdiv(ny, TY));
// where a tile ends exactly at the last column (row), its last lane (row of
// threads) also computes the u faces at j = nx (v faces at i = ny), so the
// extra face column and row need no tiles of their own.  Without it the
// tiles cover the points [0, ny] x [0, nx].  Same arithmetic per point.
template <int TX, int TY, bool EDGE_FACES = false>
__device__ __forceinline__ void residual_tile_lds(const GlideGeom &g, const Params &p, double *__restrict__ r,
                                                  const double *__restrict__ U, const double *__restrict__ f,
                                                  const double *__restrict__ sub, const double *__restrict__ bed,
                                                  const double *__restrict__ B, const double *__restrict__ beta,
                                                  const double *__restrict__ mask, const double *__restrict__ gamma,
                                                  int bx, int by, ResidLDS<TX, TY> &S)
{
    constexpr int NT = TX * TY;
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv;
    const int i0 = by * TY, j0 = bx * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    // 1. stage the inputs
    for (int k = tid; k < (TY + 4) * (TX + 4); k += NT) {
        const int a = k / (TX + 4), b = k - a * (TX + 4);
        S.u[a][b] = ld_u(u, i0 - 2 + a, j0 - 2 + b, ny, nx);
        S.v[a][b] = ld_v(v, i0 - 2 + a, j0 - 2 + b, ny, nx);
    }
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        const int a = k / (TX + 2), b = k - a * (TX + 2);
        S.H[a][b] = ld_c(H, i0 - 1 + a, j0 - 1 + b, ny, nx);
        S.bed[a][b] = ld_c(bed, i0 - 1 + a, j0 - 1 + b, ny, nx);
        S.beta[a][b] = ld_c(beta, i0 - 1 + a, j0 - 1 + b, ny, nx);
    }
    __syncthreads();
    auto UL = [&](int a, int b) { return S.u[a - i0 + 2][b - j0 + 2]; };   // global (a, b)
    auto VL = [&](int a, int b) { return S.v[a - i0 + 2][b - j0 + 2]; };
    auto HC = [&](int a, int b) { return S.H[a - i0 + 1][b - j0 + 1]; };
    auto EC = [&](int a, int b) { return S.E[a - i0 + 1][b - j0 + 1]; };
    // 2. per cell: viscosity and grounding sigmoid
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        const int a = k / (TX + 2), b = k - a * (TX + 2);
        const int ci = i0 - 1 + a, cj = j0 - 1 + b;
        S.E[a][b] = (ci >= 0 && ci < ny && cj >= 0 && cj < nx)
                        ? eta_cell_acc(ci, cj, UL, VL, B[(size_t)ci * nx + cj], p, g.dx, ny, nx)
                        : 0.0;
        const Ground gr = ground(S.bed[a][b], S.H[a][b], p);
        S.g[a][b] = gr.g;
        S.gd[a][b] = gr.gd;
    }
    __syncthreads();
    // 3. per cell normal stresses, per vertex shear stress
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        const int a = k / (TX + 2), b = k - a * (TX + 2);
        const int ci = i0 - 1 + a, cj = j0 - 1 + b;
        double sx = 0.0, sy = 0.0;
        // cells [i0, i0+TY) x [j0-1, j0+TX), and j0+TX for the faces at j = nx
        if (a >= 1 && a <= TY && b <= TX + (EDGE_FACES ? 1 : 0))
            sx = sigma_xx(UL(ci, cj), UL(ci, cj + 1), VL(ci, cj), VL(ci + 1, cj), S.E[a][b] * S.H[a][b], dx_inv, cj,
                          nx).res;
        // cells [i0-1, i0+TY) x [j0, j0+TX), and i0+TY for the faces at i = ny
        if (a <= TY + (EDGE_FACES ? 1 : 0) && b >= 1 && b <= TX)
            sy = sigma_yy(UL(ci, cj), UL(ci, cj + 1), VL(ci, cj), VL(ci + 1, cj), S.E[a][b] * S.H[a][b], dx_inv, ci,
                          ny).res;
        S.sxx[a][b] = sx;
        S.syy[a][b] = sy;
    }
    for (int k = tid; k < (TY + 1) * (TX + 1); k += NT) {
        const int a = k / (TX + 1), b = k - a * (TX + 1);
        const int vi = i0 + a, vj = j0 + b;
        const double eH = etaH_vertex(EC(vi - 1, vj - 1), EC(vi - 1, vj), EC(vi, vj - 1), EC(vi, vj),
                                      HC(vi - 1, vj - 1), HC(vi - 1, vj), HC(vi, vj - 1), HC(vi, vj));
        S.sxy[a][b] = sigma_xy(UL(vi - 1, vj), UL(vi, vj), VL(vi, vj - 1), VL(vi, vj), eH, dx_inv, vi, vj, ny, nx).res;
    }
    __syncthreads();
    // 4. rows of point (i, j): cell (i < ny, j < nx), u face (i < ny), v face (j < nx)
    auto SXX = [&](int ci, int cj) { return S.sxx[ci - i0 + 1][cj - j0 + 1]; };
    auto SYY = [&](int ci, int cj) { return S.syy[ci - i0 + 1][cj - j0 + 1]; };
    auto SXY = [&](int vi, int vj) { return S.sxy[vi - i0][vj - j0]; };
    auto rows = [&](int i, int j) {
    const int a = i - i0 + 1, b = j - j0 + 1;   // cell (i, j) in the one-ring arrays
    if (i < ny && j < nx) {
        double H_c = S.H[a][b];
        const size_t kc = (size_t)i * nx + j;
        double rH = H_c / dt - (f ? f[nu + nv + kc] : 0.0);
        Calv jc = calving_g(H_c, Ground{S.g[a][b], S.gd[a][b]}, p);
        rH -= jc.res;
        double H_l = S.H[a][b - 1];
        Flux q = flux_x(UL(i, j), H_l, H_c, j, nx);
        rH -= q.res * dx_inv;
        double H_r = S.H[a][b + 1];
        q = flux_x(UL(i, j + 1), H_c, H_r, j + 1, nx);
        rH += q.res * dx_inv;
        double H_t = S.H[a - 1][b];
        q = flux_y(VL(i, j), H_t, H_c, i, ny);
        rH += q.res * dx_inv;
        double H_b = S.H[a + 1][b];
        q = flux_y(VL(i + 1, j), H_c, H_b, i + 1, ny);
        rH -= q.res * dx_inv;
        double m = mask[kc];
        double th = gamma[kc];
        double val = (1.0 - m) * rH + m * (H_c - th);
        size_t k = nu + nv + kc;
        r[k] = sub ? val - sub[k] : val;
    }
    if (i < ny) {
        const size_t k = (size_t)i * (nx + 1) + j;
        double ru = f ? -f[k] : -0.0;
        ru += SXX(i, j) * dx_inv;
        ru -= SXX(i, j - 1) * dx_inv;
        ru += SXY(i, j) * dx_inv;
        ru -= SXY(i + 1, j) * dx_inv;
        {
            TauB t = tau_b_g(UL(i, j), S.g[a][b - 1], S.g[a][b], S.gd[a][b - 1], S.gd[a][b], S.beta[a][b - 1],
                             S.beta[a][b], p);
            ru += t.res;
        }
        {
            TauD t = tau_dx(S.H[a][b - 1], S.H[a][b], S.bed[a][b - 1], S.bed[a][b], dx_inv, j, nx);
            ru -= t.res;
        }
        r[k] = sub ? ru - sub[k] : ru;
    }
    if (j < nx) {
        const size_t k = nu + (size_t)i * nx + j;
        double rv = f ? -f[k] : -0.0;
        rv += SYY(i - 1, j) * dx_inv;
        rv -= SYY(i, j) * dx_inv;
        rv -= SXY(i, j) * dx_inv;
        rv += SXY(i, j + 1) * dx_inv;
        {
            TauB t = tau_b_g(VL(i, j), S.g[a - 1][b], S.g[a][b], S.gd[a - 1][b], S.gd[a][b], S.beta[a - 1][b],
                             S.beta[a][b], p);
            rv += t.res;
        }
        {
            TauD t = tau_dy(S.H[a - 1][b], S.H[a][b], S.bed[a - 1][b], S.bed[a][b], dx_inv, i, ny);
            rv -= t.res;
        }
        r[k] = sub ? rv - sub[k] : rv;
    }
    };
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    if (i > ny || j > nx) return;
    rows(i, j);
    if (EDGE_FACES) {
        if (j == nx - 1 && j0 + TX == nx && i < ny) rows(i, nx);   // u face at j = nx
        if (i == ny - 1 && i0 + TY == ny && j < nx) rows(ny, j);   // v face at i = ny
    }
}

template <class EtaFn>
__device__ __forceinline__ CellNbhd load_nbhd(int i, int j, int ny, int nx, size_t nu, size_t nv,
                                              const double *__restrict__ u, const double *__restrict__ v,
                                              const double *__restrict__ H, const double *__restrict__ f,
                                              const double *__restrict__ bed, const double *__restrict__ beta,
                                              const EtaFn &E)
{
    CellNbhd N;
    N.e_c = E(i, j); N.e_l = E(i, j - 1); N.e_r = E(i, j + 1); N.e_t = E(i - 1, j); N.e_b = E(i + 1, j);
    N.e_tl = E(i - 1, j - 1); N.e_tr = E(i - 1, j + 1); N.e_bl = E(i + 1, j - 1); N.e_br = E(i + 1, j + 1);
    N.H_l = ld_c(H, i, j - 1, ny, nx); N.H_r = ld_c(H, i, j + 1, ny, nx);
    N.H_t = ld_c(H, i - 1, j, ny, nx); N.H_b = ld_c(H, i + 1, j, ny, nx);
    N.H_tl = ld_c(H, i - 1, j - 1, ny, nx); N.H_tr = ld_c(H, i - 1, j + 1, ny, nx);
    N.H_bl = ld_c(H, i + 1, j - 1, ny, nx); N.H_br = ld_c(H, i + 1, j + 1, ny, nx);
    N.bed_c = ld_c(bed, i, j, ny, nx); N.bed_l = ld_c(bed, i, j - 1, ny, nx); N.bed_r = ld_c(bed, i, j + 1, ny, nx);
    N.bed_t = ld_c(bed, i - 1, j, ny, nx); N.bed_b = ld_c(bed, i + 1, j, ny, nx);
    N.beta_c = ld_c(beta, i, j, ny, nx); N.beta_l = ld_c(beta, i, j - 1, ny, nx);
    N.beta_r = ld_c(beta, i, j + 1, ny, nx); N.beta_t = ld_c(beta, i - 1, j, ny, nx);
    N.beta_b = ld_c(beta, i + 1, j, ny, nx);
    N.u_ll = ld_u(u, i, j - 1, ny, nx); N.u_rr = ld_u(u, i, j + 2, ny, nx);
    N.u_tl = ld_u(u, i - 1, j, ny, nx); N.u_tr = ld_u(u, i - 1, j + 1, ny, nx);
    N.u_bl = ld_u(u, i + 1, j, ny, nx); N.u_br = ld_u(u, i + 1, j + 1, ny, nx);
    N.v_lt = ld_v(v, i, j - 1, ny, nx); N.v_lb = ld_v(v, i + 1, j - 1, ny, nx);
    N.v_rt = ld_v(v, i, j + 1, ny, nx); N.v_rb = ld_v(v, i + 1, j + 1, ny, nx);
    N.v_tt = ld_v(v, i - 1, j, ny, nx); N.v_bb = ld_v(v, i + 2, j, ny, nx);
    if (f) {
        N.f0 = ld_u(f, i, j, ny, nx); N.f1 = ld_u(f, i, j + 1, ny, nx);
    N.f2 = ld_v(f + nu, i, j, ny, nx); N.f3 = ld_v(f + nu, i + 1, j, ny, nx);
    N.f4 = ld_v(f + nu + nv, i, j, ny, nx);
    } else {
        N.f0 = N.f1 = N.f2 = N.f3 = N.f4 = 0.0;
    }

    return N;
}

// ---------------------------------------------------------------------------
// One cell's Vanka update: neighbourhood, damped Newton iterations, and the
// per-cell deltas (glide/cuda/vanka.cu:56-439).
// INTERIOR: the cell lies in [1, ny-2] x [1, nx-2].  Every boundary test of
// the assembly compares a cell index in [i-1, i+1], or a face or vertex index
// in [i, i+1] (likewise for j), with 0 and with ny (nx); all of them take the
// interior branch for such a cell, as they do for the cell (1, 1) of a 3 x 3
// grid.  Passing those constants lets the compiler drop the tests; the
// arithmetic, and so every bit of the result, is unchanged.
// ---------------------------------------------------------------------------
template <bool INTERIOR, class EtaFn>
__device__ __forceinline__ void vanka_cell_deltas(const GlideGeom &g, const Params &p, const double *__restrict__ u,
                                                  const double *__restrict__ v, const double *__restrict__ H,
                                                  const double *__restrict__ f, const double *__restrict__ bed,
                                                  const double *__restrict__ beta, const double *__restrict__ gamma,
                                                  int n_newton, int i, int j, const EtaFn &E, double d5[5])
{
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx);
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    CellNbhd N = load_nbhd(i, j, ny, nx, nu, nv, u, v, H, f, bed, beta, E);
    nbhd_grounding(N, p);
    const double u_l0 = ld_u(u, i, j, ny, nx), u_r0 = ld_u(u, i, j + 1, ny, nx);
    const double v_t0 = ld_v(v, i, j, ny, nx), v_b0 = ld_v(v, i + 1, j, ny, nx);
    const double H_c0 = ld_c(H, i, j, ny, nx), th = ld_c(gamma, i, j, ny, nx);
    double u_l = u_l0, u_r = u_r0, v_t = v_t0, v_b = v_b0, H_c = H_c0;
    const int ai = INTERIOR ? 1 : i, aj = INTERIOR ? 1 : j, any = INTERIOR ? 3 : ny, anx = INTERIOR ? 3 : nx;
    for (int k = 0; k < n_newton; k++) {
        double J[25], r[5], d[5];
        vanka_assemble<0>(N, p, ai, aj, any, anx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, false, J, r);
        lu5_solve(J, r, d);
        u_l -= 0.25 * d[0];
        u_r -= 0.25 * d[1];
        v_t -= 0.25 * d[2];
        v_b -= 0.25 * d[3];
        H_c -= 0.25 * d[4];
        H_c = fmax(H_c, th);
    }
    d5[0] = 0.5 * (u_l - u_l0);
    d5[1] = 0.5 * (u_r - u_r0);
    d5[2] = 0.5 * (v_t - v_t0);
    d5[3] = 0.5 * (v_b - v_b0);
    d5[4] = H_c - H_c0;
}

template <bool INTERIOR, class EtaFn>
__device__ __forceinline__ void vanka_cell(const GlideGeom &g, const Params &p, const double *__restrict__ u,
                                           const double *__restrict__ v, const double *__restrict__ H,
                                           const double *__restrict__ f, const double *__restrict__ bed,
                                           const double *__restrict__ beta, const double *__restrict__ gamma,
                                           int n_newton, double *__restrict__ cd, int i, int j, const EtaFn &E)
{
    double d5[5];
    vanka_cell_deltas<INTERIOR>(g, p, u, v, H, f, bed, beta, gamma, n_newton, i, j, E, d5);
    const size_t nc = n_c(g.ny, g.nx), ci = (size_t)i * g.nx + j;
    cd[ci] = d5[0];
    cd[nc + ci] = d5[1];
    cd[2 * nc + ci] = d5[2];
    cd[3 * nc + ci] = d5[3];
    cd[4 * nc + ci] = d5[4];
}

// ---------------------------------------------------------------------------
// Forward Vanka: per-cell damped Newton (glide/cuda/vanka.cu:3-439, 844-1295)
// cd (5 x nc, structure of arrays) receives 0.5*(x - x0) for the four faces
// and (H - H0) for the cell.
// ---------------------------------------------------------------------------
template <int TX, int TY, bool LOCAL>
__device__ __forceinline__ void vanka_tile(const GlideGeom &g, const Params &p, const double *__restrict__ U,
                                           const double *__restrict__ f, const double *__restrict__ bed,
                                           const double *__restrict__ B, const double *__restrict__ beta,
                                           const double *__restrict__ gamma, const unsigned char *__restrict__ blk,
                                           int nbx, int n_newton, double *__restrict__ cd, int bx, int by,
                                           double (*eta_s)[TX + 2])
{
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv;
    const int i0 = by * TY, j0 = bx * TX;
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    const bool mine = (i < ny && j < nx);
    // A local sweep smooths every cell of a tile that holds an active cell:
    // the apply pass reads only the active cells' updates, so the idle
    // cells' chains change nothing, and a wave with part of its lanes idle
    // ran its Newton chain up to 2.3x slower than a full wave (measured on
    // the 64^2 to 512^2 levels).
    if (LOCAL) {
        if (!__syncthreads_or(mine && blk[(i / RB) * nbx + j / RB])) return;   // uniform: the whole tile is idle
    }
    fill_eta_tile<TX, TY>(eta_s, i0, j0, u, v, B, p, g.dx, ny, nx);
    __syncthreads();
    if (!mine) return;
    auto E = [&](int ci, int cj) { return eta_s[ci - i0 + 1][cj - j0 + 1]; };
    // A tile whose cells all lie in [1, ny-2] x [1, nx-2] runs the cell
    // with its boundary tests resolved at compile time (same bits).  The
    // whole per-cell body is instantiated twice so that no value is live
    // across the branch.
    if (i0 >= 1 && i0 + TY <= ny - 1 && j0 >= 1 && j0 + TX <= nx - 1)
        vanka_cell<true>(g, p, u, v, H, f, bed, beta, gamma, n_newton, cd, i, j, E);
    else
        vanka_cell<false>(g, p, u, v, H, f, bed, beta, gamma, n_newton, cd, i, j, E);
}

// Gather the per-cell updates: U += 0.5 * delta (glide/grid.py:305-317)
template <int TX, int TY, bool LOCAL>
__device__ __forceinline__ void apply_tile(int ny, int nx, double *__restrict__ U, const double *__restrict__ cd,
                                           const unsigned char *__restrict__ blk, int nbx, int bx, int by)
{
    const int i = by * TY + threadIdx.y, j = bx * TX + threadIdx.x;
    if (i > ny || j > nx) return;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx);
    auto act = [&](int ci, int cj) { return !LOCAL || blk[(ci / RB) * nbx + cj / RB] != 0; };
    if (i < ny) {
        double d = 0.0;
        if (j > 0 && act(i, j - 1)) d = d + cd[nc + (size_t)i * nx + j - 1];
        if (j < nx && act(i, j)) d = d + cd[(size_t)i * nx + j];
        U[(size_t)i * (nx + 1) + j] += 0.5 * d;
    }
    if (j < nx) {
        double d = 0.0;
        if (i > 0 && act(i - 1, j)) d = d + cd[3 * nc + (size_t)(i - 1) * nx + j];
        if (i < ny && act(i, j)) d = d + cd[2 * nc + (size_t)i * nx + j];
        U[nu + (size_t)i * nx + j] += 0.5 * d;
    }
    if (i < ny && j < nx) {
        size_t ci = (size_t)i * nx + j;
        double d = act(i, j) ? cd[4 * nc + ci] : 0.0;
        U[nu + nv + ci] += 0.5 * d;
    }
}

// One Vanka sweep with its apply pass folded in, from the state U into Uo
// (another buffer).  The workgroup's TX x TY threads run the cells
// [i0-1, i0+TY-2] x [j0-1, j0+TX-2]: the (TX-2) x (TY-2) cells it owns,
// (i0, j0) = (by (TY-2), bx (TX-2)), and the ring around them, whose
// updates the owned faces also receive.  The per-cell updates meet in LDS
// and every point the workgroup owns gets U + 0.5 d with d summed as the
// apply pass sums it: the u faces and H of its cells, their top v faces,
// and the level's last face column and row where its cells reach them.
// Each point of the level is owned once, so Uo is written whole.  Threads
// whose cell lies outside the level run the nearest cell of the level and
// drop the result (a wave with idle lanes ran its chain slower).  With
// LOCAL, idle cells' updates are not read (as in the apply pass), and a
// workgroup whose cells are all idle runs no chain.  Same bits as vanka_tile
// followed by apply_tile.  The grid is cdiv(nx, TX-2) x cdiv(ny, TY-2).
template <int TX, int TY, bool LOCAL>
__device__ __forceinline__ void vanka_halo_tile(const GlideGeom &g, const Params &p, const double *__restrict__ U,
                                                double *__restrict__ Uo, const double *__restrict__ f,
                                                const double *__restrict__ bed, const double *__restrict__ B,
                                                const double *__restrict__ beta, const double *__restrict__ gamma,
                                                const unsigned char *__restrict__ blk, int nbx, int n_newton, int bx,
                                                int by, double (*eta_s)[TX + 2], double (*cd_s)[TY][TX])
{
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv;
    const int i0 = by * (TY - 2), j0 = bx * (TX - 2);   // owned cells [i0, ie) x [j0, je)
    const int ie = min(i0 + TY - 2, ny), je = min(j0 + TX - 2, nx);
    const int a = threadIdx.y, b = threadIdx.x;
    const int i = i0 - 1 + a, j = j0 - 1 + b;
    const bool in = i >= 0 && i < ny && j >= 0 && j < nx;
    auto act = [&](int ci, int cj) { return !LOCAL || blk[(ci / RB) * nbx + cj / RB] != 0; };
    bool run = true;
    if (LOCAL) run = __syncthreads_or(in && act(i, j));   // uniform
    if (run) {
        fill_eta_tile<TX, TY>(eta_s, i0 - 1, j0 - 1, u, v, B, p, g.dx, ny, nx);
        __syncthreads();
        const int ic = min(max(i, 0), ny - 1), jc = min(max(j, 0), nx - 1);
        auto E = [&](int ci, int cj) { return eta_s[ci - i0 + 2][cj - j0 + 2]; };
        double d5[5];
        // every run cell in [1, ny-2] x [1, nx-2]: the boundary tests resolved at compile time
        if (i0 >= 2 && i0 + TY <= ny && j0 >= 2 && j0 + TX <= nx)
            vanka_cell_deltas<true>(g, p, u, v, H, f, bed, beta, gamma, n_newton, ic, jc, E, d5);
        else
            vanka_cell_deltas<false>(g, p, u, v, H, f, bed, beta, gamma, n_newton, ic, jc, E, d5);
#pragma unroll
        for (int q = 0; q < 5; q++) cd_s[q][a][b] = d5[q];
        __syncthreads();
    }
    auto cdv = [&](int q, int ci, int cj) { return cd_s[q][ci - i0 + 1][cj - j0 + 1]; };
    double *uo = Uo, *vo = Uo + nu, *Ho = Uo + nu + nv;
    if (i >= i0 && i < ie && j >= j0 && (j < je || j == nx)) {
        double d = 0.0;
        if (j > 0 && act(i, j - 1)) d = d + cdv(1, i, j - 1);
        if (j < nx && act(i, j)) d = d + cdv(0, i, j);
        const size_t k = (size_t)i * (nx + 1) + j;
        uo[k] = u[k] + 0.5 * d;
    }
    if (j >= j0 && j < je && i >= i0 && (i < ie || i == ny)) {
        double d = 0.0;
        if (i > 0 && act(i - 1, j)) d = d + cdv(3, i - 1, j);
        if (i < ny && act(i, j)) d = d + cdv(2, i, j);
        const size_t k = (size_t)i * nx + j;
        vo[k] = v[k] + 0.5 * d;
    }
    if (i >= i0 && i < ie && j >= j0 && j < je) {
        const size_t k = (size_t)i * nx + j;
        const double d = act(i, j) ? cdv(4, i, j) : 0.0;
        Ho[k] = H[k] + 0.5 * d;
    }
}

// Local-smoother activation: one 16x16 workgroup per 14x14 reference block
// reads its 16x16 footprint (glide/solver.py:148-154, glide/cuda/vanka.cu:870-888).
__device__ __forceinline__ void activation_block(int ny, int nx, const double *__restrict__ r,
                                                 const double *__restrict__ U, const double *__restrict__ gamma,
                                                 unsigned char *__restrict__ blk, int bx, int by, int nbx,
                                                 int *__restrict__ any_active = nullptr)
{
    const int i = by * RB - 1 + (int)threadIdx.y, j = bx * RB - 1 + (int)threadIdx.x;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    int bad = 0;
    if (i >= 0 && i < ny && j >= 0 && j < nx) {
        size_t ci = (size_t)i * nx + j;
        double a = r[nu + nv + ci], b = U[nu + nv + ci] - gamma[ci];
        double fb = a + b - sqrt(a * a + b * b);
        double te = fabs(r[(size_t)i * (nx + 1) + j + 1]) + fabs(r[(size_t)i * (nx + 1) + j]) +
                    fabs(r[nu + (size_t)(i + 1) * nx + j]) + fabs(r[nu + (size_t)i * nx + j]) + fabs(fb);
        bad = te > 0.01;
    }
    int any = __syncthreads_or(bad);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        blk[by * nbx + bx] = (unsigned char)(any != 0);
        if (any && any_active) *any_active = 1;   // every writer stores the same value
    }
}

// ---------------------------------------------------------------------------
// Bilinear prolongation samples (glide/cuda/utility.cu:180-337)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bilin(double ty, double tx, double v00, double v01, double v10, double v11)
{
    return (1.0 - ty) * ((1.0 - tx) * v00 + tx * v01) + ty * ((1.0 - tx) * v10 + tx * v11);
}
// coarse sample helper: value of (a[k] - b[k]) when b != NULL, else a[k]
__device__ __forceinline__ double cdiff(const double *__restrict__ a, const double *__restrict__ b, size_t k)
{
    return b ? a[k] - b[k] : a[k];
}
// bilinear u prolongation of (uc - wc) at fine face (j, i) (utility.cu:180-232)
__device__ __forceinline__ double prol_u_at(const double *uc, const double *wc, int j, int i, int nyc, int nxc)
{
    double Jf = (j - 0.5) * 0.5, If = i * 0.5;
    Jf = fmax(0.0, fmin(Jf, (double)(nyc - 1)));
    If = fmax(0.0, fmin(If, (double)nxc));
    int Jl = (int)Jf, Il = (int)If;
    int Jh = min(Jl + 1, nyc - 1), Ih = min(Il + 1, nxc);
    double ty = Jf - Jl, tx = If - Il;
    size_t s = nxc + 1;
    return bilin(ty, tx, cdiff(uc, wc, Jl * s + Il), cdiff(uc, wc, Jl * s + Ih), cdiff(uc, wc, Jh * s + Il),
                 cdiff(uc, wc, Jh * s + Ih));
}
// utility.cu:235-287
__device__ __forceinline__ double prol_v_at(const double *vc, const double *wc, int j, int i, int nyc, int nxc)
{
    double Jf = j * 0.5, If = (i - 0.5) * 0.5;
    Jf = fmax(0.0, fmin(Jf, (double)nyc));
    If = fmax(0.0, fmin(If, (double)(nxc - 1)));
    int Jl = (int)Jf, Il = (int)If;
    int Jh = min(Jl + 1, nyc), Ih = min(Il + 1, nxc - 1);
    double ty = Jf - Jl, tx = If - Il;
    size_t s = nxc;
    return bilin(ty, tx, cdiff(vc, wc, Jl * s + Il), cdiff(vc, wc, Jl * s + Ih), cdiff(vc, wc, Jh * s + Il),
                 cdiff(vc, wc, Jh * s + Ih));
}
// utility.cu:289-337
__device__ __forceinline__ double prol_H_at(const double *hc, const double *wc, int j, int i, int nyc, int nxc)
{
    double Jf = (j - 0.5) * 0.5, If = (i - 0.5) * 0.5;
    Jf = fmax(0.0, fmin(Jf, (double)(nyc - 1)));
    If = fmax(0.0, fmin(If, (double)(nxc - 1)));
    int Jl = (int)Jf, Il = (int)If;
    int Jh = min(Jl + 1, nyc - 1), Ih = min(Il + 1, nxc - 1);
    double ty = Jf - Jl, tx = If - Il;
    size_t s = nxc;
    return bilin(ty, tx, cdiff(hc, wc, Jl * s + Il), cdiff(hc, wc, Jl * s + Ih), cdiff(hc, wc, Jh * s + Il),
                 cdiff(hc, wc, Jh * s + Ih));
}

// ---------------------------------------------------------------------------
// Linearisation, LDS-staged like the residual: JVP j = J(U) d
// (glide/cuda/residuals.cu:299-579) and the gather-form VJP l = J(U)^T lambda
// (residuals.cu:586-1023; the (u, v) block through its symmetry, as a JVP
// along lambda with a zero thickness direction).  The tile's u, v and the
// direction (two-ring halo) and H, bed, beta and the thickness direction (or
// lambda_H (1 - mask)) (one-ring halo) are staged once; each cell's dual
// viscosity, grounding sigmoid and the directional derivatives of sigma_xx
// and sigma_yy, and each vertex's of sigma_xy, are evaluated once; each point
// then sums its terms from LDS.  Every quantity is the expression of the
// per-point form on the same clamped values and every output sums its terms
// in the same order, so the bits are those of the per-point form.
// Tiles cover the points [0, ny] x [0, nx].
// ---------------------------------------------------------------------------
template <int TX, int TY>
struct LinLDS {
    double u[TY + 4][TX + 4], v[TY + 4][TX + 4];     // state, clamped, global (i0-2+a, j0-2+b)
    double du[TY + 4][TX + 4], dv[TY + 4][TX + 4];   // direction (JVP) or lambda (VJP), clamped
    double H[TY + 2][TX + 2], dH[TY + 2][TX + 2];    // cells (i0-1+a, j0-1+b), clamped; dH: direction or lambda_H (1 - mask)
    double bed[TY + 2][TX + 2], beta[TY + 2][TX + 2];
    double ev[TY + 2][TX + 2], ed[TY + 2][TX + 2];   // dual viscosity (0 outside the domain)
    double g[TY + 2][TX + 2], gd[TY + 2][TX + 2];    // grounding sigmoid and its derivative
    double txx[TY + 2][TX + 2], tyy[TY + 2][TX + 2]; // directional derivatives of sigma_xx, sigma_yy
    double txy[TY + 1][TX + 1];                      // ... of sigma_xy, vertices (i0+a, j0+b)
};

// eta*H as a Dual: cell and vertex average (viscosity.cu:150-217)
__device__ __forceinline__ Dual eH_cell_dual(Dual eta, Dual Hh) { return Dual{Hh.v * eta.v, Hh.v * eta.d + eta.v * Hh.d}; }
__device__ __forceinline__ Dual eH_vertex_dual(Dual a, Dual b, Dual c, Dual d, Dual Ha, Dual Hb, Dual Hc, Dual Hdd)
{
    return Dual{etaH_vertex(a.v, b.v, c.v, d.v, Ha.v, Hb.v, Hc.v, Hdd.v),
                0.25 * Ha.v * a.d + 0.25 * a.v * Ha.d + 0.25 * Hb.v * b.d + 0.25 * b.v * Hb.d + 0.25 * Hc.v * c.d +
                    0.25 * c.v * Hc.d + 0.25 * Hdd.v * d.d + 0.25 * d.v * Hdd.d};
}

// VJP: out = base - l when base != NULL (the adjoint residual f_adj - J^T lambda)
template <int TX, int TY, bool VJP>
__device__ __forceinline__ void lin_tile_lds(const GlideGeom &g, const Params &p, double *__restrict__ out,
                                             const double *__restrict__ base, const double *__restrict__ U,
                                             const double *__restrict__ D, const double *__restrict__ bed,
                                             const double *__restrict__ B, const double *__restrict__ beta,
                                             const double *__restrict__ mask, int bx, int by, LinLDS<TX, TY> &S)
{
    constexpr int NT = TX * TY;
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv;
    const double *du = D, *dv = D + nu, *dH = D + nu + nv;
    const int i0 = by * TY, j0 = bx * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    // 1. stage the inputs
    for (int k = tid; k < (TY + 4) * (TX + 4); k += NT) {
        const int a = k / (TX + 4), b = k - a * (TX + 4);
        const int gi = i0 - 2 + a, gj = j0 - 2 + b;
        S.u[a][b] = ld_u(u, gi, gj, ny, nx);
        S.v[a][b] = ld_v(v, gi, gj, ny, nx);
        S.du[a][b] = ld_u(du, gi, gj, ny, nx);
        S.dv[a][b] = ld_v(dv, gi, gj, ny, nx);
    }
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        const int a = k / (TX + 2), b = k - a * (TX + 2);
        const int ci = i0 - 1 + a, cj = j0 - 1 + b;
        S.H[a][b] = ld_c(H, ci, cj, ny, nx);
        S.bed[a][b] = ld_c(bed, ci, cj, ny, nx);
        S.beta[a][b] = ld_c(beta, ci, cj, ny, nx);
        S.dH[a][b] = VJP ? ld_c(dH, ci, cj, ny, nx) * (1.0 - ld_c(mask, ci, cj, ny, nx)) : ld_c(dH, ci, cj, ny, nx);
    }
    __syncthreads();
    auto UL = [&](int a, int b) { return S.u[a - i0 + 2][b - j0 + 2]; };   // global (a, b)
    auto VL = [&](int a, int b) { return S.v[a - i0 + 2][b - j0 + 2]; };
    auto DU = [&](int a, int b) { return S.du[a - i0 + 2][b - j0 + 2]; };
    auto DV = [&](int a, int b) { return S.dv[a - i0 + 2][b - j0 + 2]; };
    auto Ud = [&](int a, int b) { return Dual{UL(a, b), DU(a, b)}; };
    auto Vd = [&](int a, int b) { return Dual{VL(a, b), DV(a, b)}; };
    auto HC = [&](int a, int b) { return S.H[a - i0 + 1][b - j0 + 1]; };
    auto LM = [&](int a, int b) { return S.dH[a - i0 + 1][b - j0 + 1]; };   // VJP: lambda_H (1 - mask)
    auto Hd = [&](int a, int b) { return Dual{HC(a, b), VJP ? 0.0 : S.dH[a - i0 + 1][b - j0 + 1]}; };
    auto E = [&](int a, int b) { return Dual{S.ev[a - i0 + 1][b - j0 + 1], S.ed[a - i0 + 1][b - j0 + 1]}; };
    auto G = [&](int a, int b) { return Ground{S.g[a - i0 + 1][b - j0 + 1], S.gd[a - i0 + 1][b - j0 + 1]}; };
    auto BED = [&](int a, int b) { return S.bed[a - i0 + 1][b - j0 + 1]; };
    auto BETA = [&](int a, int b) { return S.beta[a - i0 + 1][b - j0 + 1]; };
    // 2. per cell: dual viscosity and grounding sigmoid
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        const int a = k / (TX + 2), b = k - a * (TX + 2);
        const int ci = i0 - 1 + a, cj = j0 - 1 + b;
        const Dual e = (ci >= 0 && ci < ny && cj >= 0 && cj < nx)
                           ? eta_cell_dual_acc(ci, cj, Ud, Vd, B[(size_t)ci * nx + cj], p, g.dx, ny, nx)
                           : Dual{0.0, 0.0};
        S.ev[a][b] = e.v;
        S.ed[a][b] = e.d;
        const Ground gr = ground(S.bed[a][b], S.H[a][b], p);
        S.g[a][b] = gr.g;
        S.gd[a][b] = gr.gd;
    }
    __syncthreads();
    // 3. directional derivatives of the stresses: sigma_xx on cells [i0, i0+TY) x
    //    [j0-1, j0+TX), sigma_yy on [i0-1, i0+TY) x [j0, j0+TX), sigma_xy on the vertices
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        const int a = k / (TX + 2), b = k - a * (TX + 2);
        const int ci = i0 - 1 + a, cj = j0 - 1 + b;
        double tx = 0.0, ty = 0.0;
        if ((a >= 1 && a <= TY && b <= TX) || (a <= TY && b >= 1 && b <= TX)) {
            const Dual eH = eH_cell_dual(E(ci, cj), Hd(ci, cj));
            const Dual ul = Ud(ci, cj), ur = Ud(ci, cj + 1), vt = Vd(ci, cj), vb = Vd(ci + 1, cj);
            if (a >= 1 && a <= TY && b <= TX) {
                const SigN s = sigma_xx(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, cj, nx);
                tx = sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d);
            }
            if (a <= TY && b >= 1 && b <= TX) {
                const SigN s = sigma_yy(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, ci, ny);
                ty = sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d);
            }
        }
        S.txx[a][b] = tx;
        S.tyy[a][b] = ty;
    }
    for (int k = tid; k < (TY + 1) * (TX + 1); k += NT) {
        const int a = k / (TX + 1), b = k - a * (TX + 1);
        const int vi = i0 + a, vj = j0 + b;
        const Dual eH = eH_vertex_dual(E(vi - 1, vj - 1), E(vi - 1, vj), E(vi, vj - 1), E(vi, vj), Hd(vi - 1, vj - 1),
                                       Hd(vi - 1, vj), Hd(vi, vj - 1), Hd(vi, vj));
        const Dual ut = Ud(vi - 1, vj), ub = Ud(vi, vj), vl = Vd(vi, vj - 1), vr = Vd(vi, vj);
        const SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, vi, vj, ny, nx);
        S.txy[a][b] = sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d);
    }
    __syncthreads();
    auto TXX = [&](int ci, int cj) { return S.txx[ci - i0 + 1][cj - j0 + 1]; };
    auto TYY = [&](int ci, int cj) { return S.tyy[ci - i0 + 1][cj - j0 + 1]; };
    auto TXY = [&](int vi, int vj) { return S.txy[vi - i0][vj - j0]; };
    // 2 eps_xy at vertex (vi, vj), 0 on boundary vertices
    auto EXY = [&](int vi, int vj) {
        return sigma_xy(UL(vi - 1, vj), UL(vi, vj), VL(vi, vj - 1), VL(vi, vj), 0.0, dx_inv, vi, vj, ny, nx).d_eta_H;
    };
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    if (i > ny || j > nx) return;
    auto emit = [&](size_t k, double val) { out[k] = (VJP && base) ? base[k] - val : val; };
    // 4a. u row (i, j): J_u d (glide/cuda/residuals.cu:378-499)
    if (i < ny) {
        double d = 0.0;
        d += TXX(i, j) * dx_inv;
        d -= TXX(i, j - 1) * dx_inv;
        d += TXY(i, j) * dx_inv;
        d -= TXY(i + 1, j) * dx_inv;
        const Dual ul = Ud(i, j), Hl = Hd(i, j - 1), Hc = Hd(i, j);
        {
            const Ground gl = G(i, j - 1), gc = G(i, j);
            TauB t = tau_b_g(ul.v, gl.g, gc.g, gl.gd, gc.gd, BETA(i, j - 1), BETA(i, j), p);
            d += t.d_u * ul.d + t.d_H_l * Hl.d + t.d_H_r * Hc.d;
        }
        {
            TauD t = tau_dx(Hl.v, Hc.v, BED(i, j - 1), BED(i, j), dx_inv, j, nx);
            d -= t.d_H_l * Hl.d + t.d_H_r * Hc.d;
        }
        if (VJP) {   // plus the mass rows through the fluxes of this face
            Flux q = flux_x(ul.v, HC(i, j - 1), HC(i, j), j, nx);
            double acc = 0.0;
            if (j > 0) acc += LM(i, j - 1) * q.d_u * dx_inv;
            if (j < nx) acc += -LM(i, j) * q.d_u * dx_inv;
            acc += d;
            d = acc;
        }
        emit((size_t)i * (nx + 1) + j, d);
    }
    // 4b. v row (i, j) (residuals.cu:501-576)
    if (j < nx) {
        double d = 0.0;
        d += TYY(i - 1, j) * dx_inv;
        d -= TYY(i, j) * dx_inv;
        d -= TXY(i, j) * dx_inv;
        d += TXY(i, j + 1) * dx_inv;
        const Dual vt = Vd(i, j), Ht = Hd(i - 1, j), Hc = Hd(i, j);
        {
            const Ground gt = G(i - 1, j), gc = G(i, j);
            TauB t = tau_b_g(vt.v, gt.g, gc.g, gt.gd, gc.gd, BETA(i - 1, j), BETA(i, j), p);
            d += t.d_u * vt.d + t.d_H_l * Ht.d + t.d_H_r * Hc.d;
        }
        {
            TauD t = tau_dy(Ht.v, Hc.v, BED(i - 1, j), BED(i, j), dx_inv, i, ny);
            d -= t.d_H_l * Ht.d + t.d_H_r * Hc.d;
        }
        if (VJP) {
            Flux q = flux_y(vt.v, HC(i - 1, j), HC(i, j), i, ny);
            double acc = 0.0;
            if (i > 0) acc += -LM(i - 1, j) * q.d_u * dx_inv;
            if (i < ny) acc += LM(i, j) * q.d_u * dx_inv;
            acc += d;
            d = acc;
        }
        emit(nu + (size_t)i * nx + j, d);
    }
    if (i >= ny || j >= nx) return;
    const size_t kc = (size_t)i * nx + j;
    if (!VJP) {   // 4c. mass row of cell (i, j) (residuals.cu:343-375)
        const Dual H_c = Hd(i, j);
        double d = H_c.d / dt;
        Calv jc = calving_g(H_c.v, G(i, j), p);
        d -= jc.d_H * H_c.d;
        Flux q = flux_x(UL(i, j), HC(i, j - 1), H_c.v, j, nx);
        d -= (q.d_u * DU(i, j) + q.d_H_l * Hd(i, j - 1).d + q.d_H_r * H_c.d) * dx_inv;
        q = flux_x(UL(i, j + 1), H_c.v, HC(i, j + 1), j + 1, nx);
        d += (q.d_u * DU(i, j + 1) + q.d_H_l * H_c.d + q.d_H_r * Hd(i, j + 1).d) * dx_inv;
        q = flux_y(VL(i, j), HC(i - 1, j), H_c.v, i, ny);
        d += (q.d_u * DV(i, j) + q.d_H_l * Hd(i - 1, j).d + q.d_H_r * H_c.d) * dx_inv;
        q = flux_y(VL(i + 1, j), H_c.v, HC(i + 1, j), i + 1, ny);
        d -= (q.d_u * DV(i + 1, j) + q.d_H_l * H_c.d + q.d_H_r * Hd(i + 1, j).d) * dx_inv;
        const double m = mask[kc];
        out[nu + nv + kc] = (1.0 - m) * d + m * H_c.d;
        return;
    }
    // 4c. VJP, cell (i, j): every row's derivative with respect to H(i, j)
    const double Hc = HC(i, j), eta = E(i, j).v;
    const double Hl = HC(i, j - 1), Hr = HC(i, j + 1), Ht = HC(i - 1, j), Hb = HC(i + 1, j);
    const double lc = LM(i, j);
    Flux qxl = flux_x(UL(i, j), Hl, Hc, j, nx);
    Flux qxr = flux_x(UL(i, j + 1), Hc, Hr, j + 1, nx);
    Flux qyt = flux_y(VL(i, j), Ht, Hc, i, ny);
    Flux qyb = flux_y(VL(i + 1, j), Hc, Hb, i + 1, ny);
    Calv jc = calving_g(Hc, G(i, j), p);
    // own mass row (residuals.cu:648-690)
    double acc = lc / dt;
    acc += -lc * jc.d_H;
    acc += -lc * qxl.d_H_r * dx_inv;
    acc += lc * qxr.d_H_l * dx_inv;
    acc += lc * qyt.d_H_r * dx_inv;
    acc += -lc * qyb.d_H_l * dx_inv;
    acc += dH[kc] * mask[kc];
    // neighbours' mass rows through the shared faces
    if (j > 0) acc += LM(i, j - 1) * qxl.d_H_r * dx_inv;
    if (j < nx - 1) acc += -LM(i, j + 1) * qxr.d_H_l * dx_inv;
    if (i > 0) acc += -LM(i - 1, j) * qyt.d_H_r * dx_inv;
    if (i < ny - 1) acc += LM(i + 1, j) * qyb.d_H_l * dx_inv;
    // momentum rows through eta*H of this cell: normal stresses
    const double ul = UL(i, j), ur = UL(i, j + 1), vt = VL(i, j), vb = VL(i + 1, j);
    SigN sx = sigma_xx(ul, ur, vt, vb, 0.0, dx_inv, j, nx);
    SigN sy = sigma_yy(ul, ur, vt, vb, 0.0, dx_inv, i, ny);
    const double lu_l = DU(i, j), lu_r = DU(i, j + 1), lv_t = DV(i, j), lv_b = DV(i + 1, j);
    acc += lu_l * sx.d_eta_H * eta * dx_inv;
    acc += -lu_r * sx.d_eta_H * eta * dx_inv;
    acc += -lv_t * sy.d_eta_H * eta * dx_inv;
    acc += lv_b * sy.d_eta_H * eta * dx_inv;
    // shear stresses of the four corner vertices
    const double q4 = 0.25 * eta;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int a = i + (c >> 1), b = j + (c & 1);
        const double dv_ = EXY(a, b);
        const double w = DU(a, b) - DU(a - 1, b) - DV(a, b) + DV(a, b - 1);
        acc += w * dv_ * q4 * dx_inv;
    }
    // basal drag (grounding-line derivatives) and driving stress on the four faces
    {
        const Ground ga = G(i, j - 1), gb = G(i, j);
        TauB t = tau_b_g(ul, ga.g, gb.g, ga.gd, gb.gd, BETA(i, j - 1), BETA(i, j), p);
        TauD d = tau_dx(Hl, Hc, BED(i, j - 1), BED(i, j), dx_inv, j, nx);
        acc += lu_l * t.d_H_r;
        acc += -lu_l * d.d_H_r;
    }
    {
        const Ground ga = G(i, j), gb = G(i, j + 1);
        TauB t = tau_b_g(ur, ga.g, gb.g, ga.gd, gb.gd, BETA(i, j), BETA(i, j + 1), p);
        TauD d = tau_dx(Hc, Hr, BED(i, j), BED(i, j + 1), dx_inv, j + 1, nx);
        acc += lu_r * t.d_H_l;
        acc += -lu_r * d.d_H_l;
    }
    {
        const Ground ga = G(i - 1, j), gb = G(i, j);
        TauB t = tau_b_g(vt, ga.g, gb.g, ga.gd, gb.gd, BETA(i - 1, j), BETA(i, j), p);
        TauD d = tau_dy(Ht, Hc, BED(i - 1, j), BED(i, j), dx_inv, i, ny);
        acc += lv_t * t.d_H_r;
        acc += -lv_t * d.d_H_r;
    }
    {
        const Ground ga = G(i, j), gb = G(i + 1, j);
        TauB t = tau_b_g(vb, ga.g, gb.g, ga.gd, gb.gd, BETA(i, j), BETA(i + 1, j), p);
        TauD d = tau_dy(Hc, Hb, BED(i, j), BED(i + 1, j), dx_inv, i + 1, ny);
        acc += lv_b * t.d_H_l;
        acc += -lv_b * d.d_H_l;
    }
    emit(nu + nv + kc, acc);
}

}  // namespace glide
