// glide_hip.hip - MI355X (gfx950) kernels and C ABI of the SSA / FASCD solver.
//
// Layout: FP64 row-major MAC fields, state vector [u | v | H] per level.
// Stencil kernels work on 64 x 4 point tiles (one wave64 per tile row, so
// every row access is a contiguous 512-byte segment); the cell viscosity of
// the tile plus a one-cell ring is computed once per tile into LDS.  The
// Vanka smoother is one lane per cell (the 10-step local Newton stays in
// registers) writing its 5 local updates to a scratch array; a second
// gather pass applies them (each face sums its <= 2 adjacent cells' updates,
// so no atomics are needed and the result is deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "glide_hip.h"
#include "tiles_device.h"

using namespace glide;

namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NT = TX * TY;

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail((int)e, std::string(what) + ": " + hipGetErrorString(e));
    return GLIDE_OK;
}


int check_geom(const GlideGeom *g)
{
    if (!g) return fail(GLIDE_E_NULL, "geometry is NULL");
    if (g->ny < 1 || g->nx < 1 || g->ny > (1 << 16) || g->nx > (1 << 16))
        return fail(GLIDE_E_SIZE, "grid size out of range");
    if (!(g->dx > 0.0) || !(g->dt > 0.0) || !std::isfinite(g->dx) || !std::isfinite(g->dt))
        return fail(GLIDE_E_ARG, "dx and dt must be positive and finite");
    return GLIDE_OK;
}

inline hipStream_t S(void *s) { return (hipStream_t)s; }

// Workgroups per CU of the persistent single-level launches (Vanka sweep,
// residual); set by glide_set_persistence.  The Vanka sweep's 64 gives one
// tile per workgroup at every level size used here (measured fastest: the
// hardware then balances clustered local-smoother work); the residual's 8
// loops each workgroup over several tiles where a level has more tiles than
// that.  Launch shapes only: results do not depend on them.
std::atomic<int> g_wg_per_cu[2] = {{64}, {8}};
// Levels with at most g_small_max_cells cells sweep with the small-level
// kernel: variant 0 = k_vanka compiled for one wave per SIMD (the whole
// register file), 1 = the row split (k_vanka_rows), 2 = k_vanka_halo (the
// apply pass folded into the sweep, the state alternating between U and the
// Vanka scratch); set by glide_set_small_levels.  Measured (one sweep with
// its apply, 10 Newton iterations, HIP events over 10 sweeps): 32^2 22.2 /
// 19.9 us, 64^2 23.4 / 20.5, 128^2 23.7-24.6 / 20.7-21.1 (variant 0 / 2);
// the row split measured 23.2, 24.0 and 24.1 against 22.5, 23.6 and 23.4
// for variant 0 in an earlier build.
std::atomic<long> g_small_max_cells{128 * 128};
std::atomic<int> g_small_variant{2};
// Tile heights (64 x TY points) of the single-level residual (4, 8, 16; 0 =
// by level size) and of the JVP / VJP (0 = the per-point form, 4, 6, 8); set
// by glide_set_tile_height.  Launch shapes only: results are the same bits.
// Measured at 512^2: residual 64 x 16 tiles 11.3 us, 64 x 8 12.0, 64 x 4 13.4;
// at 1024^2: 64 x 8 39.7 us, 64 x 16 42.0, 64 x 4 46.4 (8 workgroups per CU).
// VJP at 512^2: 64 x 4 22.5 us, 64 x 8 26.1, 64 x 6 31.8, per point 39.6.
std::atomic<int> g_res_ty{0};
std::atomic<int> g_lin_ty{4};
int residual_ty(int ny, int nx)
{
    const int ty = g_res_ty;
    if (ty) return ty;
    const long nc = (long)ny * nx;
    return nc > (1L << 18) ? 8 : (nc >= (1L << 16) ? 16 : 4);
}

int n_cu()
{
    static int cu[16] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 256;
    if (!cu[dev]) {
        hipDeviceProp_t prop;
        cu[dev] = hipGetDeviceProperties(&prop, dev) == hipSuccess ? prop.multiProcessorCount : 256;
    }
    return cu[dev];
}
// min(tiles, CUs x per_cu) workgroups, at least 8 and a multiple of 8 (tile_loop)
dim3 persistent_grid(long tiles, int per_cu)
{
    long g = tiles < (long)n_cu() * per_cu ? tiles : (long)n_cu() * per_cu;
    g = g < NXCD ? NXCD : g;
    return dim3((unsigned)((g + NXCD - 1) / NXCD * NXCD));
}
dim3 vanka_grid(int ny, int nx) { return persistent_grid((long)cdiv(nx, TX) * cdiv(ny, TY), g_wg_per_cu[0]); }
dim3 residual_grid(int ny, int nx, int ty) { return persistent_grid((long)cdiv(nx, TX) * cdiv(ny, ty), g_wg_per_cu[1]); }
dim3 point_grid(int ny, int nx) { return dim3((nx + 1 + TX - 1) / TX, (ny + 1 + TY - 1) / TY); }
dim3 cell_grid(int ny, int nx) { return dim3((nx + TX - 1) / TX, (ny + TY - 1) / TY); }
inline unsigned blocks1d(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// error channel shared with the batched entry points (glide_batch.hip)
int glide::set_error(int code, const char *msg) { return fail(code, msg); }

// ---------------------------------------------------------------------------
// Residual r = F(U) - f (- sub)  (glide/cuda/residuals.cu:6-292)
// ---------------------------------------------------------------------------
template <int RTY>
__global__ __launch_bounds__(TX *RTY) void k_residual(GlideGeom g, Params p, double *__restrict__ r,
                                                      const double *__restrict__ U, const double *__restrict__ f,
                                                      const double *__restrict__ sub, const double *__restrict__ bed,
                                                      const double *__restrict__ B, const double *__restrict__ beta,
                                                      const double *__restrict__ mask,
                                                      const double *__restrict__ gamma, int *__restrict__ zero)
{
    __shared__ ResidLDS<TX, RTY> S;
    if (zero && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *zero = 0;   // a flag the next kernel raises
    const int tx = cdiv(g.nx, TX), total = tx * cdiv(g.ny, RTY);   // cell tiles; edge faces by the last tiles
    tile_loop(total, [&](int t) {
        __syncthreads();   // the previous tile's readers of S are done
        residual_tile_lds<TX, RTY, true>(g, p, r, U, f, sub, bed, B, beta, mask, gamma, t % tx, t / tx, S);
    });
}

// ---------------------------------------------------------------------------
// Forward Vanka: per-cell damped Newton (glide/cuda/vanka.cu:3-439, 844-1295)
// cd (5 x nc, structure of arrays) receives 0.5*(x - x0) for the four faces
// and (H - H0) for the cell.
// ---------------------------------------------------------------------------
// WPE waves per SIMD: the register budget both instantiated cell bodies
// (interior and general) must share.  Large levels take 2 (occupancy hides
// the Newton chain's latency); small levels, whose few waves each run alone
// on a SIMD, take 1 (the whole register file: no spills, longer schedules).
template <bool LOCAL, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_vanka(GlideGeom g, Params p,
                                                                                   const double *__restrict__ U,
                                              const double *__restrict__ f, const double *__restrict__ bed,
                                              const double *__restrict__ B, const double *__restrict__ beta,
                                              const double *__restrict__ gamma, const unsigned char *__restrict__ blk,
                                              int nbx, int n_newton, double *__restrict__ cd)
{
    __shared__ double eta_s[TY + 2][TX + 2];
    const int tx = cdiv(g.nx, TX), total = tx * cdiv(g.ny, TY);
    tile_loop(total, [&](int t) {
        __syncthreads();   // the previous tile's readers of eta_s are done
        vanka_tile<TX, TY, LOCAL>(g, p, U, f, bed, B, beta, gamma, blk, nbx, n_newton, cd, t % tx, t / tx, eta_s);
    });
}

// Small-level sweep with the apply pass folded in (vanka_halo_tile): U -> Uo,
// one workgroup per 62 x 2 owned cells, one wave per SIMD.
template <bool LOCAL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1))) void k_vanka_halo(
    GlideGeom g, Params p, const double *__restrict__ U, double *__restrict__ Uo, const double *__restrict__ f,
    const double *__restrict__ bed, const double *__restrict__ B, const double *__restrict__ beta,
    const double *__restrict__ gamma, const unsigned char *__restrict__ blk, int nbx, int n_newton)
{
    __shared__ double eta_s[TY + 2][TX + 2];
    __shared__ double cd_s[5][TY][TX];
    vanka_halo_tile<TX, TY, LOCAL>(g, p, U, Uo, f, bed, B, beta, gamma, blk, nbx, n_newton, blockIdx.x, blockIdx.y,
                                   eta_s, cd_s);
}
dim3 halo_grid(int ny, int nx) { return dim3(cdiv(nx, TX - 2), cdiv(ny, TY - 2)); }


// ---------------------------------------------------------------------------
// Row-split Vanka sweep for small levels, where a sweep takes as long as one
// wave's Newton chain.  A workgroup holds 5 waves x 64 cells of one grid row;
// wave w assembles row w of each cell's 5x5 block (vanka_assemble with a row
// mask: each entry receives the operations it receives in the full
// assembly), the rows meet in LDS (double buffered, one barrier per Newton
// iteration) and every wave solves the block and updates its copy of the
// unknowns.  Same bits as k_vanka.
// ---------------------------------------------------------------------------
template <int ROWS>
__device__ __forceinline__ void rows_to_lds(const CellNbhd &N, const Params &p, int i, int j, int ny, int nx,
                                            double dx_inv, double dt, double u_l, double u_r, double v_t, double v_b,
                                            double H_c, double th, double (*x)[6][64], int lane)
{
    double J[25], r[5];
    vanka_assemble<0, ROWS>(N, p, i, j, ny, nx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, false, J, r);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        if (ROWS & (1 << q)) {
#pragma unroll
            for (int c = 0; c < 5; c++) x[q][c][lane] = J[5 * q + c];
            x[q][5][lane] = r[q];
        }
    }
}

constexpr int ROW_WAVES = 5;

template <bool LOCAL>
__global__ __launch_bounds__(64 * ROW_WAVES) void k_vanka_rows(GlideGeom g, Params p, const double *__restrict__ U,
                                                    const double *__restrict__ f, const double *__restrict__ bed,
                                                    const double *__restrict__ B, const double *__restrict__ beta,
                                                    const double *__restrict__ gamma,
                                                    const unsigned char *__restrict__ blk, int nbx, int n_newton,
                                                    double *__restrict__ cd)
{
    __shared__ double eta_s[3][64 + 2];
    __shared__ double xch[2][5][6][64];
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv;
    const int lane = threadIdx.x, w = threadIdx.y;
    const int i = blockIdx.y, j0 = blockIdx.x * 64, j = j0 + lane;
    bool mine = j < nx;   // i < ny by the grid
    if (LOCAL) {
        mine = mine && blk[(i / RB) * nbx + j / RB];
        if (!__syncthreads_or(mine)) return;   // uniform: the whole tile is idle
    }
    for (int k = w * 64 + lane; k < 3 * 66; k += 64 * ROW_WAVES) {   // viscosity of cells [i-1, i+1] x [j0-1, j0+64]
        const int a = k / 66, b = k - a * 66;
        const int ci = i - 1 + a, cj = j0 - 1 + b;
        eta_s[a][b] = (ci >= 0 && ci < ny && cj >= 0 && cj < nx) ? eta_cell(ci, cj, u, v, B, p, g.dx, ny, nx) : 0.0;
    }
    __syncthreads();
    // lanes past the last column work on the last cell (every wave reaches
    // every barrier of the Newton loop) and discard the result
    const int jj = j < nx ? j : nx - 1;
    auto E = [&](int ci, int cj) { return eta_s[ci - i + 1][cj - j0 + 1]; };
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    CellNbhd N = load_nbhd(i, jj, ny, nx, nu, nv, u, v, H, f, bed, beta, E);
    nbhd_grounding(N, p);
    const double u_l0 = ld_u(u, i, jj, ny, nx), u_r0 = ld_u(u, i, jj + 1, ny, nx);
    const double v_t0 = ld_v(v, i, jj, ny, nx), v_b0 = ld_v(v, i + 1, jj, ny, nx);
    const double H_c0 = ld_c(H, i, jj, ny, nx), th = ld_c(gamma, i, jj, ny, nx);
    double u_l = u_l0, u_r = u_r0, v_t = v_t0, v_b = v_b0, H_c = H_c0;
    for (int k = 0; k < n_newton; k++) {
        double (*x)[6][64] = xch[k & 1];
        if (w == 0)
            rows_to_lds<1>(N, p, i, jj, ny, nx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, x, lane);
        else if (w == 1)
            rows_to_lds<2>(N, p, i, jj, ny, nx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, x, lane);
        else if (w == 2)
            rows_to_lds<4>(N, p, i, jj, ny, nx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, x, lane);
        else if (w == 3)
            rows_to_lds<8>(N, p, i, jj, ny, nx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, x, lane);
        else
            rows_to_lds<16>(N, p, i, jj, ny, nx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, x, lane);
        // one barrier per iteration: buffer k & 1 is written again at k + 2,
        // after every wave has passed barrier k + 1 and so finished reading it
        __syncthreads();
        double J[25], r[5], d[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
#pragma unroll
            for (int c = 0; c < 5; c++) J[5 * q + c] = x[q][c][lane];
            r[q] = x[q][5][lane];
        }
        lu5_solve(J, r, d);
        u_l -= 0.25 * d[0];
        u_r -= 0.25 * d[1];
        v_t -= 0.25 * d[2];
        v_b -= 0.25 * d[3];
        H_c -= 0.25 * d[4];
        H_c = fmax(H_c, th);
    }
    if (w == 0 && mine) {
        const size_t ci = (size_t)i * nx + j;
        cd[ci] = 0.5 * (u_l - u_l0);
        cd[nc + ci] = 0.5 * (u_r - u_r0);
        cd[2 * nc + ci] = 0.5 * (v_t - v_t0);
        cd[3 * nc + ci] = 0.5 * (v_b - v_b0);
        cd[4 * nc + ci] = H_c - H_c0;
    }
}

// Local block (J, r) of every cell after k Newton updates (test hook):
// out[30*ci + 0..24] = J, out[30*ci + 25..29] = r.
__global__ __launch_bounds__(NT) void k_vanka_block(GlideGeom g, Params p, const double *__restrict__ U,
                                                    const double *__restrict__ f, const double *__restrict__ bed,
                                                    const double *__restrict__ B, const double *__restrict__ beta,
                                                    const double *__restrict__ gamma, int k_updates,
                                                    double *__restrict__ out)
{
    __shared__ double eta_s[TY + 2][TX + 2];
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    fill_eta_tile<TX, TY>(eta_s, i0, j0, u, v, B, p, g.dx, ny, nx);
    __syncthreads();
    if (i >= ny || j >= nx) return;
    auto E = [&](int ci, int cj) { return eta_s[ci - i0 + 1][cj - j0 + 1]; };
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    CellNbhd N = load_nbhd(i, j, ny, nx, nu, nv, u, v, H, f, bed, beta, E);
    nbhd_grounding(N, p);
    double u_l = ld_u(u, i, j, ny, nx), u_r = ld_u(u, i, j + 1, ny, nx);
    double v_t = ld_v(v, i, j, ny, nx), v_b = ld_v(v, i + 1, j, ny, nx);
    double H_c = ld_c(H, i, j, ny, nx);
    const double th = ld_c(gamma, i, j, ny, nx);
    double J[25], r[5], d[5];
    for (int k = 0; k <= k_updates; k++) {
        vanka_assemble<0>(N, p, i, j, ny, nx, dx_inv, dt, u_l, u_r, v_t, v_b, H_c, th, false, J, r);
        if (k == k_updates) break;
        lu5_solve(J, r, d);
        u_l -= 0.25 * d[0];
        u_r -= 0.25 * d[1];
        v_t -= 0.25 * d[2];
        v_b -= 0.25 * d[3];
        H_c -= 0.25 * d[4];
        H_c = fmax(H_c, th);
    }
    double *o = out + 30 * ((size_t)i * nx + j);
    for (int q = 0; q < 25; q++) o[q] = J[q];
    for (int q = 0; q < 5; q++) o[25 + q] = r[q];
}

// Gather the per-cell updates: U += 0.5 * delta (glide/grid.py:305-317)
template <bool LOCAL>
__global__ __launch_bounds__(NT) void k_vanka_apply(int ny, int nx, double *__restrict__ U,
                                                    const double *__restrict__ cd,
                                                    const unsigned char *__restrict__ blk, int nbx)
{
    apply_tile<TX, TY, LOCAL>(ny, nx, U, cd, blk, nbx, blockIdx.x, blockIdx.y);
}

// Local-smoother activation: one workgroup per 14x14 reference block reads
// its 16x16 footprint (glide/solver.py:148-154, glide/cuda/vanka.cu:870-888).
__global__ __launch_bounds__(256) void k_local_activation(int ny, int nx, const double *__restrict__ r,
                                                          const double *__restrict__ U,
                                                          const double *__restrict__ gamma,
                                                          unsigned char *__restrict__ blk, int *__restrict__ any_active)
{
    activation_block(ny, nx, r, U, gamma, blk, blockIdx.x, blockIdx.y, gridDim.x, any_active);
}

// ---------------------------------------------------------------------------
// Transfer operators (glide/cuda/utility.cu)
// ---------------------------------------------------------------------------
// bw multiplies the boundary-face rows (I = 0, I = nxc for u; J = 0, J = nyc
// for v): 1.0 for states and the reference's residual restriction.
__global__ void k_restrict_u(const double *__restrict__ uf, double *__restrict__ uc, double *__restrict__ uc2,
                             int nyc, int nxc, double bw)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nyc * (nxc + 1)) return;
    int J = idx / (nxc + 1), I = idx - J * (nxc + 1);
    double val = 0.5 * (uf[(size_t)(2 * J) * (2 * nxc + 1) + 2 * I] + uf[(size_t)(2 * J + 1) * (2 * nxc + 1) + 2 * I]);
    if (I == 0 || I == nxc) val *= bw;
    uc[idx] = val;
    if (uc2) uc2[idx] = val;
}
__global__ void k_restrict_v(const double *__restrict__ vf, double *__restrict__ vc, double *__restrict__ vc2,
                             int nyc, int nxc, double bw)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (nyc + 1) * nxc) return;
    int J = idx / nxc, I = idx - J * nxc;
    double val = 0.5 * (vf[(size_t)(2 * J) * (2 * nxc) + 2 * I] + vf[(size_t)(2 * J) * (2 * nxc) + 2 * I + 1]);
    if (J == 0 || J == nyc) val *= bw;
    vc[idx] = val;
    if (vc2) vc2[idx] = val;
}
__global__ void k_restrict_cell(const double *__restrict__ hf, double *__restrict__ hc, double *__restrict__ hc2,
                                int nyc, int nxc)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nyc * nxc) return;
    int J = idx / nxc, I = idx - J * nxc;
    size_t nxf = 2 * (size_t)nxc;
    double val = 0.25 * (hf[(2 * J) * nxf + 2 * I] + hf[(2 * J) * nxf + 2 * I + 1] + hf[(2 * J + 1) * nxf + 2 * I] +
                         hf[(2 * J + 1) * nxf + 2 * I + 1]);
    hc[idx] = val;
    if (hc2) hc2[idx] = val;
}
__global__ void k_restrict_max(const double *__restrict__ hf, double *__restrict__ hc, int nyc, int nxc)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nyc * nxc) return;
    int J = idx / nxc, I = idx - J * nxc;
    size_t nxf = 2 * (size_t)nxc;
    double a = fmax(hf[(2 * J) * nxf + 2 * I], hf[(2 * J) * nxf + 2 * I + 1]);
    double b = fmax(hf[(2 * J + 1) * nxf + 2 * I], hf[(2 * J + 1) * nxf + 2 * I + 1]);
    hc[idx] = fmax(a, b);
}
// The three restrictions of a state vector [u | v | H] in one launch (the
// arithmetic of k_restrict_u, k_restrict_v and k_restrict_cell per point);
// optionally also a[0, n_fill) = fill_val (the launch covers both ranges).
__global__ void k_restrict_state(const double *__restrict__ src, double *__restrict__ dst, double *__restrict__ dst2,
                                 int nyc, int nxc, double bw, double *__restrict__ fill_a, size_t n_fill,
                                 double fill_val)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n_fill) fill_a[idx] = fill_val;
    const size_t nuc = n_u(nyc, nxc), nvc = n_v(nyc, nxc), ncc = n_c(nyc, nxc);
    const size_t nxf = 2 * (size_t)nxc, nuf = n_u(2 * nyc, 2 * nxc), nvf = n_v(2 * nyc, 2 * nxc);
    double val;
    if (idx < nuc) {
        const int J = (int)(idx / (nxc + 1)), I = (int)(idx - (size_t)J * (nxc + 1));
        val = 0.5 * (src[(size_t)(2 * J) * (nxf + 1) + 2 * I] + src[(size_t)(2 * J + 1) * (nxf + 1) + 2 * I]);
        if (I == 0 || I == nxc) val *= bw;
    } else if (idx < nuc + nvc) {
        const size_t k = idx - nuc;
        const int J = (int)(k / nxc), I = (int)(k - (size_t)J * nxc);
        const double *vf = src + nuf;
        val = 0.5 * (vf[(size_t)(2 * J) * nxf + 2 * I] + vf[(size_t)(2 * J) * nxf + 2 * I + 1]);
        if (J == 0 || J == nyc) val *= bw;
    } else if (idx < nuc + nvc + ncc) {
        const size_t k = idx - nuc - nvc;
        const int J = (int)(k / nxc), I = (int)(k - (size_t)J * nxc);
        const double *hf = src + nuf + nvf;
        val = 0.25 * (hf[(2 * J) * nxf + 2 * I] + hf[(2 * J) * nxf + 2 * I + 1] + hf[(2 * J + 1) * nxf + 2 * I] +
                      hf[(2 * J + 1) * nxf + 2 * I + 1]);
    } else {
        return;
    }
    dst[idx] = val;
    if (dst2) dst2[idx] = val;
}
__global__ void k_prolong_inject(const double *__restrict__ hc, double *__restrict__ hf, int nyf, int nxf, double sign)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nyf * nxf) return;
    int j = idx / nxf, i = idx - j * nxf;
    hf[idx] = sign * hc[(size_t)(j / 2) * (nxf / 2) + i / 2];
}

__global__ void k_prolong_u(const double *__restrict__ uc, double *__restrict__ uf, int nyf, int nxf)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nyf * (nxf + 1)) return;
    int j = idx / (nxf + 1), i = idx - j * (nxf + 1);
    uf[idx] = prol_u_at(uc, nullptr, j, i, nyf / 2, nxf / 2);
}
__global__ void k_prolong_v(const double *__restrict__ vc, double *__restrict__ vf, int nyf, int nxf)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (nyf + 1) * nxf) return;
    int j = idx / nxf, i = idx - j * nxf;
    vf[idx] = prol_v_at(vc, nullptr, j, i, nyf / 2, nxf / 2);
}
__global__ void k_prolong_H(const double *__restrict__ hc, double *__restrict__ hf, int nyf, int nxf)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nyf * nxf) return;
    int j = idx / nxf, i = idx - j * nxf;
    hf[idx] = prol_H_at(hc, nullptr, j, i, nyf / 2, nxf / 2);
}

// ---------------------------------------------------------------------------
// FASCD element-wise steps (glide/solver.py:71-157), fused
// ---------------------------------------------------------------------------
// w = U, chi = gamma - H
__global__ void k_fas_begin(size_t nt, size_t nuv, double *__restrict__ w, const double *__restrict__ U,
                            double *__restrict__ chi, const double *__restrict__ gamma)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    double x = U[k];
    w[k] = x;
    if (k >= nuv) chi[k - nuv] = gamma[k - nuv] - x;
}
__global__ void k_fill(size_t n, double *__restrict__ a, double val)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = val;
}
__global__ void k_copy(size_t n, double *__restrict__ dst, const double *__restrict__ src)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}
// gamma = w_H + chi, or (pre-smooth) gamma = w_H + ((-chi_c[inject]) + chi)
__global__ void k_gamma(int ny, int nx, double *__restrict__ gamma, const double *__restrict__ wH,
                        const double *__restrict__ chi, const double *__restrict__ chi_c)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ny * nx) return;
    double ph;
    if (chi_c) {
        int j = idx / nx, i = idx - j * nx;
        double inj = -1.0 * chi_c[(size_t)(j / 2) * (nx / 2) + i / 2];
        ph = inj + chi[idx];
    } else {
        ph = chi[idx];
    }
    gamma[idx] = wH[idx] + ph;
}
// k_restrict_max into chi_c followed by the pre-smoothing k_gamma, in one
// launch: each fine cell forms its parent's max itself (the same fmax
// order), and the cell at the parent's top-left corner stores it.
__global__ void k_gamma_pre(int ny, int nx, double *__restrict__ gamma, const double *__restrict__ wH,
                            const double *__restrict__ chi, double *__restrict__ chi_c)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ny * nx) return;
    const int j = idx / nx, i = idx - j * nx, J = j / 2, I = i / 2;
    const size_t f0 = (size_t)(2 * J) * nx + 2 * I, f1 = f0 + nx;
    const double a = fmax(chi[f0], chi[f0 + 1]);
    const double b = fmax(chi[f1], chi[f1 + 1]);
    const double c = fmax(a, b);
    if (j == 2 * J && i == 2 * I) chi_c[(size_t)J * (nx / 2) + I] = c;
    const double inj = -1.0 * c;
    gamma[idx] = wH[idx] + (inj + chi[idx]);
}
// U = w + (P(U_c - w_c) + (U - w)): bilinear for u, v and injection for H
// (glide/solver.py:131-140); with gamma, the cells also take the
// post-correction k_gamma, gamma = w_H + chi (it reads no U).
__global__ void k_fas_correct(int ny, int nx, double *__restrict__ U, const double *__restrict__ w,
                              const double *__restrict__ Uc, const double *__restrict__ wc,
                              double *__restrict__ gamma, const double *__restrict__ chi,
                              double *__restrict__ gamma_c, double thklim)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gamma_c && k < n_c(ny / 2, nx / 2)) gamma_c[k] = thklim;   // the child's reset (fewer cells than nt)
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx);
    if (k >= nu + nv + nc) return;
    const int nyc = ny / 2, nxc = nx / 2;
    const size_t nuc = n_u(nyc, nxc), nvc = n_v(nyc, nxc);
    double z;
    if (k < nu) {
        int j = (int)(k / (nx + 1)), i = (int)(k - (size_t)j * (nx + 1));
        z = prol_u_at(Uc, wc, j, i, nyc, nxc);
    } else if (k < nu + nv) {
        size_t kk = k - nu;
        int j = (int)(kk / nx), i = (int)(kk - (size_t)j * nx);
        z = prol_v_at(Uc + nuc, wc + nuc, j, i, nyc, nxc);
    } else {
        size_t kk = k - nu - nv;
        int j = (int)(kk / nx), i = (int)(kk - (size_t)j * nx);
        size_t kc = nuc + nvc + (size_t)(j / 2) * nxc + i / 2;
        z = Uc[kc] - wc[kc];
        if (gamma) gamma[kk] = w[k] + chi[kk];
    }
    double y = U[k] - w[k];
    z += y;
    U[k] = w[k] + z;
}

// f_H = H_prev / dt + smb (glide/physics.py:173)
__global__ void k_forcing(size_t nc, double *__restrict__ fH, const double *__restrict__ H_prev,
                          const double *__restrict__ smb, double dt)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nc) fH[k] = H_prev[k] / dt + smb[k];
}

// Count of non-finite entries of x (NaN or +-Inf) into *count.
__global__ __launch_bounds__(256) void k_count_nonfinite(const double *__restrict__ x, size_t n,
                                                         unsigned long long *__restrict__ count)
{
    unsigned c = 0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const double v = x[p];
        c += !(v - v == 0.0);   // NaN - NaN and Inf - Inf are NaN
    }
    if (c) atomicAdd(count, (unsigned long long)c);
}

__global__ void k_zero_u64(unsigned long long *__restrict__ x) { *x = 0ull; }

// ||(r_u, r_v, FB(r_H, H - gamma))||^2: per-block partial sums (fixed tree order)
__global__ __launch_bounds__(256) void k_norm_partial(int ny, int nx, const double *__restrict__ r,
                                                      const double *__restrict__ U,
                                                      const double *__restrict__ gamma, double *__restrict__ part)
{
    __shared__ double s[256];
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx), nt = nu + nv + nc;
    double acc = 0.0;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < nt; k += (size_t)gridDim.x * 256) {
        double x;
        if (k < nu + nv) {
            x = r[k];
        } else {
            double a = r[k], b = U[k] - gamma[k - nu - nv];
            x = a + b - sqrt(a * a + b * b);
        }
        acc += x * x;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(256) void k_norm_final(int n, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) acc += part[k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(s[0]);
}


// ===========================================================================
// Linearisation: JVP and gather-form VJP
// ===========================================================================
// Dual viscosities and grounding sigmoids of the cells [i0-1, i0+TY] x
// [j0-1, j0+TX] into LDS.  The sigmoid of a cell outside the domain is that
// of its clamped neighbour, as every clamped read of bed and H gives it.
__device__ __forceinline__ void fill_eta_dual_tile(Dual (*eta_s)[TX + 2], Ground (*g_s)[TX + 2], int i0, int j0,
                                                   const double *__restrict__ u, const double *__restrict__ v,
                                                   const double *__restrict__ du, const double *__restrict__ dv,
                                                   const double *__restrict__ H, const double *__restrict__ bed,
                                                   const double *__restrict__ B, const Params &p, double dx, int ny,
                                                   int nx)
{
    int tid = threadIdx.y * TX + threadIdx.x;
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        int a = k / (TX + 2), b = k - a * (TX + 2);
        int ci = i0 - 1 + a, cj = j0 - 1 + b;
        eta_s[a][b] = (ci >= 0 && ci < ny && cj >= 0 && cj < nx) ? eta_cell_dual(ci, cj, u, v, du, dv, B, p, dx, ny, nx)
                                                                 : Dual{0.0, 0.0};
        g_s[a][b] = ground(ld_c(bed, ci, cj, ny, nx), ld_c(H, ci, cj, ny, nx), p);
    }
}

// d/dt of the u-row R_u(i, j) along (du, dv, dH); dH == NULL means 0
// (the u-face part of glide/cuda/residuals.cu:378-499)
template <class EF, class GF>
__device__ __forceinline__ double jvp_u_row(int i, int j, int ny, int nx, double dx_inv, const Params &p, const EF &E,
                                            const GF &G,
                                            const double *u, const double *v, const double *H, const double *du,
                                            const double *dv, const double *dH, const double *bed,
                                            const double *beta)
{
    auto Hd = [&](int a, int b) { return Dual{ld_c(H, a, b, ny, nx), dH ? ld_c(dH, a, b, ny, nx) : 0.0}; };
    auto Ud = [&](int a, int b) { return Dual{ld_u(u, a, b, ny, nx), ld_u(du, a, b, ny, nx)}; };
    auto Vd = [&](int a, int b) { return Dual{ld_v(v, a, b, ny, nx), ld_v(dv, a, b, ny, nx)}; };
    auto eHc = [](Dual eta, Dual Hh) { return Dual{Hh.v * eta.v, Hh.v * eta.d + eta.v * Hh.d}; };
    auto eHv = [](Dual a, Dual b, Dual c, Dual d, Dual Ha, Dual Hb, Dual Hc, Dual Hdd) {
        return Dual{etaH_vertex(a.v, b.v, c.v, d.v, Ha.v, Hb.v, Hc.v, Hdd.v),
                    0.25 * Ha.v * a.d + 0.25 * a.v * Ha.d + 0.25 * Hb.v * b.d + 0.25 * b.v * Hb.d +
                        0.25 * Hc.v * c.d + 0.25 * c.v * Hc.d + 0.25 * Hdd.v * d.d + 0.25 * d.v * Hdd.d};
    };
    double d = 0.0;
    {
        Dual eH = eHc(E(i, j), Hd(i, j));
        Dual ul = Ud(i, j), ur = Ud(i, j + 1), vt = Vd(i, j), vb = Vd(i + 1, j);
        SigN s = sigma_xx(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, j, nx);
        d += sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d) * dx_inv;
    }
    {
        Dual eH = eHc(E(i, j - 1), Hd(i, j - 1));
        Dual ul = Ud(i, j), ull = Ud(i, j - 1), vt = Vd(i, j - 1), vb = Vd(i + 1, j - 1);
        SigN s = sigma_xx(ull.v, ul.v, vt.v, vb.v, eH.v, dx_inv, j - 1, nx);
        d -= sigN_jvp(s, ull.d, ul.d, vt.d, vb.d, eH.d) * dx_inv;
    }
    {
        Dual eH = eHv(E(i - 1, j - 1), E(i - 1, j), E(i, j - 1), E(i, j), Hd(i - 1, j - 1), Hd(i - 1, j), Hd(i, j - 1),
                      Hd(i, j));
        Dual ut = Ud(i - 1, j), ub = Ud(i, j), vl = Vd(i, j - 1), vr = Vd(i, j);
        SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i, j, ny, nx);
        d += sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
    }
    {
        Dual eH = eHv(E(i, j - 1), E(i, j), E(i + 1, j - 1), E(i + 1, j), Hd(i, j - 1), Hd(i, j), Hd(i + 1, j - 1),
                      Hd(i + 1, j));
        Dual ut = Ud(i, j), ub = Ud(i + 1, j), vl = Vd(i + 1, j - 1), vr = Vd(i + 1, j);
        SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i + 1, j, ny, nx);
        d -= sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
    }
    {
        Dual ul = Ud(i, j), Hl = Hd(i, j - 1), Hc = Hd(i, j);
        const Ground gl = G(i, j - 1), gc = G(i, j);
        TauB t = tau_b_g(ul.v, gl.g, gc.g, gl.gd, gc.gd, ld_c(beta, i, j - 1, ny, nx), ld_c(beta, i, j, ny, nx), p);
        d += t.d_u * ul.d + t.d_H_l * Hl.d + t.d_H_r * Hc.d;
    }
    {
        Dual Hl = Hd(i, j - 1), Hc = Hd(i, j);
        TauD t = tau_dx(Hl.v, Hc.v, ld_c(bed, i, j - 1, ny, nx), ld_c(bed, i, j, ny, nx), dx_inv, j, nx);
        d -= t.d_H_l * Hl.d + t.d_H_r * Hc.d;
    }
    return d;
}

// v-row R_v(i, j) (glide/cuda/residuals.cu:501-576)
template <class EF, class GF>
__device__ __forceinline__ double jvp_v_row(int i, int j, int ny, int nx, double dx_inv, const Params &p, const EF &E,
                                            const GF &G,
                                            const double *u, const double *v, const double *H, const double *du,
                                            const double *dv, const double *dH, const double *bed,
                                            const double *beta)
{
    auto Hd = [&](int a, int b) { return Dual{ld_c(H, a, b, ny, nx), dH ? ld_c(dH, a, b, ny, nx) : 0.0}; };
    auto Ud = [&](int a, int b) { return Dual{ld_u(u, a, b, ny, nx), ld_u(du, a, b, ny, nx)}; };
    auto Vd = [&](int a, int b) { return Dual{ld_v(v, a, b, ny, nx), ld_v(dv, a, b, ny, nx)}; };
    auto eHc = [](Dual eta, Dual Hh) { return Dual{Hh.v * eta.v, Hh.v * eta.d + eta.v * Hh.d}; };
    auto eHv = [](Dual a, Dual b, Dual c, Dual d, Dual Ha, Dual Hb, Dual Hc, Dual Hdd) {
        return Dual{etaH_vertex(a.v, b.v, c.v, d.v, Ha.v, Hb.v, Hc.v, Hdd.v),
                    0.25 * Ha.v * a.d + 0.25 * a.v * Ha.d + 0.25 * Hb.v * b.d + 0.25 * b.v * Hb.d +
                        0.25 * Hc.v * c.d + 0.25 * c.v * Hc.d + 0.25 * Hdd.v * d.d + 0.25 * d.v * Hdd.d};
    };
    double d = 0.0;
    {
        Dual eH = eHc(E(i - 1, j), Hd(i - 1, j));
        Dual ul = Ud(i - 1, j), ur = Ud(i - 1, j + 1), vt = Vd(i - 1, j), vb = Vd(i, j);
        SigN s = sigma_yy(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, i - 1, ny);
        d += sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d) * dx_inv;
    }
    {
        Dual eH = eHc(E(i, j), Hd(i, j));
        Dual ul = Ud(i, j), ur = Ud(i, j + 1), vt = Vd(i, j), vb = Vd(i + 1, j);
        SigN s = sigma_yy(ul.v, ur.v, vt.v, vb.v, eH.v, dx_inv, i, ny);
        d -= sigN_jvp(s, ul.d, ur.d, vt.d, vb.d, eH.d) * dx_inv;
    }
    {
        Dual eH = eHv(E(i - 1, j - 1), E(i - 1, j), E(i, j - 1), E(i, j), Hd(i - 1, j - 1), Hd(i - 1, j), Hd(i, j - 1),
                      Hd(i, j));
        Dual ut = Ud(i - 1, j), ub = Ud(i, j), vl = Vd(i, j - 1), vr = Vd(i, j);
        SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i, j, ny, nx);
        d -= sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
    }
    {
        Dual eH = eHv(E(i - 1, j), E(i - 1, j + 1), E(i, j), E(i, j + 1), Hd(i - 1, j), Hd(i - 1, j + 1), Hd(i, j),
                      Hd(i, j + 1));
        Dual ut = Ud(i - 1, j + 1), ub = Ud(i, j + 1), vl = Vd(i, j), vr = Vd(i, j + 1);
        SigS s = sigma_xy(ut.v, ub.v, vl.v, vr.v, eH.v, dx_inv, i, j + 1, ny, nx);
        d += sigS_jvp(s, ut.d, ub.d, vl.d, vr.d, eH.d) * dx_inv;
    }
    {
        Dual vt = Vd(i, j), Ht = Hd(i - 1, j), Hc = Hd(i, j);
        const Ground gt = G(i - 1, j), gc = G(i, j);
        TauB t = tau_b_g(vt.v, gt.g, gc.g, gt.gd, gc.gd, ld_c(beta, i - 1, j, ny, nx), ld_c(beta, i, j, ny, nx), p);
        d += t.d_u * vt.d + t.d_H_l * Ht.d + t.d_H_r * Hc.d;
    }
    {
        Dual Ht = Hd(i - 1, j), Hc = Hd(i, j);
        TauD t = tau_dy(Ht.v, Hc.v, ld_c(bed, i - 1, j, ny, nx), ld_c(bed, i, j, ny, nx), dx_inv, i, ny);
        d -= t.d_H_l * Ht.d + t.d_H_r * Hc.d;
    }
    return d;
}

__global__ __launch_bounds__(NT) void k_jvp_point(GlideGeom g, Params p, double *__restrict__ jo, const double *__restrict__ U,
                                            const double *__restrict__ dU, const double *__restrict__ bed,
                                            const double *__restrict__ B, const double *__restrict__ beta,
                                            const double *__restrict__ mask)
{
    __shared__ Dual eta_s[TY + 2][TX + 2];
    __shared__ Ground g_s[TY + 2][TX + 2];
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv, *du = dU, *dv = dU + nu, *dH = dU + nu + nv;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    fill_eta_dual_tile(eta_s, g_s, i0, j0, u, v, du, dv, H, bed, B, p, g.dx, ny, nx);
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    if (i > ny || j > nx) return;
    auto E = [&](int ci, int cj) { return eta_s[ci - i0 + 1][cj - j0 + 1]; };
    auto G = [&](int ci, int cj) { return g_s[ci - i0 + 1][cj - j0 + 1]; };
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    if (i < ny && j < nx) {   // glide/cuda/residuals.cu:343-375
        Dual H_c = {ld_c(H, i, j, ny, nx), ld_c(dH, i, j, ny, nx)};
        double d = H_c.d / dt;
        Calv jc = calving_g(H_c.v, G(i, j), p);
        d -= jc.d_H * H_c.d;
        double Hl = ld_c(H, i, j - 1, ny, nx), dHl = ld_c(dH, i, j - 1, ny, nx);
        double ul = ld_u(u, i, j, ny, nx), dul = ld_u(du, i, j, ny, nx);
        Flux q = flux_x(ul, Hl, H_c.v, j, nx);
        d -= (q.d_u * dul + q.d_H_l * dHl + q.d_H_r * H_c.d) * dx_inv;
        double Hr = ld_c(H, i, j + 1, ny, nx), dHr = ld_c(dH, i, j + 1, ny, nx);
        double ur = ld_u(u, i, j + 1, ny, nx), dur = ld_u(du, i, j + 1, ny, nx);
        q = flux_x(ur, H_c.v, Hr, j + 1, nx);
        d += (q.d_u * dur + q.d_H_l * H_c.d + q.d_H_r * dHr) * dx_inv;
        double Ht = ld_c(H, i - 1, j, ny, nx), dHt = ld_c(dH, i - 1, j, ny, nx);
        double vt = ld_v(v, i, j, ny, nx), dvt = ld_v(dv, i, j, ny, nx);
        q = flux_y(vt, Ht, H_c.v, i, ny);
        d += (q.d_u * dvt + q.d_H_l * dHt + q.d_H_r * H_c.d) * dx_inv;
        double Hb = ld_c(H, i + 1, j, ny, nx), dHb = ld_c(dH, i + 1, j, ny, nx);
        double vb = ld_v(v, i + 1, j, ny, nx), dvb = ld_v(dv, i + 1, j, ny, nx);
        q = flux_y(vb, H_c.v, Hb, i + 1, ny);
        d -= (q.d_u * dvb + q.d_H_l * H_c.d + q.d_H_r * dHb) * dx_inv;
        double m = ld_c(mask, i, j, ny, nx);
        jo[nu + nv + (size_t)i * nx + j] = (1.0 - m) * d + m * H_c.d;
    }
    if (i < ny) jo[(size_t)i * (nx + 1) + j] = jvp_u_row(i, j, ny, nx, dx_inv, p, E, G, u, v, H, du, dv, dH, bed, beta);
    if (j < nx) jo[nu + (size_t)i * nx + j] = jvp_v_row(i, j, ny, nx, dx_inv, p, E, G, u, v, H, du, dv, dH, bed, beta);
}

// shear-rate derivative factor 2*eps_xy at vertex (a, b), 0 on boundary vertices
__device__ __forceinline__ double vertex_deps(int a, int b, int ny, int nx, double dx_inv, const double *u,
                                              const double *v)
{
    SigS s = sigma_xy(ld_u(u, a - 1, b, ny, nx), ld_u(u, a, b, ny, nx), ld_v(v, a, b - 1, ny, nx), ld_v(v, a, b, ny, nx),
                      0.0, dx_inv, a, b, ny, nx);
    return s.d_eta_H;
}

// l = J^T lambda in gather form; out = base - l when base != NULL
// (the adjoint residual r_adj = f_adj - J^T lambda, glide/grid.py:322-324)
__global__ __launch_bounds__(NT) void k_vjp_point(GlideGeom g, Params p, double *__restrict__ out, const double *__restrict__ base,
                                            const double *__restrict__ U, const double *__restrict__ lam,
                                            const double *__restrict__ bed, const double *__restrict__ B,
                                            const double *__restrict__ beta, const double *__restrict__ mask)
{
    __shared__ Dual eta_s[TY + 2][TX + 2];
    __shared__ Ground g_s[TY + 2][TX + 2];
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv, *lu = lam, *lv = lam + nu, *lH = lam + nu + nv;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    fill_eta_dual_tile(eta_s, g_s, i0, j0, u, v, lu, lv, H, bed, B, p, g.dx, ny, nx);
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    if (i > ny || j > nx) return;
    auto E = [&](int ci, int cj) { return eta_s[ci - i0 + 1][cj - j0 + 1]; };
    auto G = [&](int ci, int cj) { return g_s[ci - i0 + 1][cj - j0 + 1]; };
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    // lambda_H of the mass row of cell (a, b) with the active-set mask applied
    auto lm = [&](int a, int b) { return ld_c(lH, a, b, ny, nx) * (1.0 - ld_c(mask, a, b, ny, nx)); };
    auto emit = [&](size_t k, double val) { out[k] = base ? base[k] - val : val; };

    if (i < ny) {   // u face (i, j)
        Flux q = flux_x(ld_u(u, i, j, ny, nx), ld_c(H, i, j - 1, ny, nx), ld_c(H, i, j, ny, nx), j, nx);
        double acc = 0.0;
        if (j > 0) acc += lm(i, j - 1) * q.d_u * dx_inv;
        if (j < nx) acc += -lm(i, j) * q.d_u * dx_inv;
        acc += jvp_u_row(i, j, ny, nx, dx_inv, p, E, G, u, v, H, lu, lv, (const double *)nullptr, bed, beta);
        emit((size_t)i * (nx + 1) + j, acc);
    }
    if (j < nx) {   // v face (i, j)
        Flux q = flux_y(ld_v(v, i, j, ny, nx), ld_c(H, i - 1, j, ny, nx), ld_c(H, i, j, ny, nx), i, ny);
        double acc = 0.0;
        if (i > 0) acc += -lm(i - 1, j) * q.d_u * dx_inv;
        if (i < ny) acc += lm(i, j) * q.d_u * dx_inv;
        acc += jvp_v_row(i, j, ny, nx, dx_inv, p, E, G, u, v, H, lu, lv, (const double *)nullptr, bed, beta);
        emit(nu + (size_t)i * nx + j, acc);
    }
    if (i < ny && j < nx) {   // cell (i, j)
        const double Hc = ld_c(H, i, j, ny, nx), eta = E(i, j).v;
        const double Hl = ld_c(H, i, j - 1, ny, nx), Hr = ld_c(H, i, j + 1, ny, nx);
        const double Ht = ld_c(H, i - 1, j, ny, nx), Hb = ld_c(H, i + 1, j, ny, nx);
        const double lc = lm(i, j);
        Flux qxl = flux_x(ld_u(u, i, j, ny, nx), Hl, Hc, j, nx);
        Flux qxr = flux_x(ld_u(u, i, j + 1, ny, nx), Hc, Hr, j + 1, nx);
        Flux qyt = flux_y(ld_v(v, i, j, ny, nx), Ht, Hc, i, ny);
        Flux qyb = flux_y(ld_v(v, i + 1, j, ny, nx), Hc, Hb, i + 1, ny);
        Calv jc = calving_g(Hc, G(i, j), p);
        // own mass row (residuals.cu:648-690)
        double acc = lc / dt;
        acc += -lc * jc.d_H;
        acc += -lc * qxl.d_H_r * dx_inv;
        acc += lc * qxr.d_H_l * dx_inv;
        acc += lc * qyt.d_H_r * dx_inv;
        acc += -lc * qyb.d_H_l * dx_inv;
        acc += ld_c(lH, i, j, ny, nx) * ld_c(mask, i, j, ny, nx);
        // neighbours' mass rows through the shared faces
        if (j > 0) acc += lm(i, j - 1) * qxl.d_H_r * dx_inv;
        if (j < nx - 1) acc += -lm(i, j + 1) * qxr.d_H_l * dx_inv;
        if (i > 0) acc += -lm(i - 1, j) * qyt.d_H_r * dx_inv;
        if (i < ny - 1) acc += lm(i + 1, j) * qyb.d_H_l * dx_inv;
        // momentum rows through eta*H of this cell: normal stresses
        const double ul = ld_u(u, i, j, ny, nx), ur = ld_u(u, i, j + 1, ny, nx);
        const double vt = ld_v(v, i, j, ny, nx), vb = ld_v(v, i + 1, j, ny, nx);
        SigN sx = sigma_xx(ul, ur, vt, vb, 0.0, dx_inv, j, nx);
        SigN sy = sigma_yy(ul, ur, vt, vb, 0.0, dx_inv, i, ny);
        const double lu_l = ld_u(lu, i, j, ny, nx), lu_r = ld_u(lu, i, j + 1, ny, nx);
        const double lv_t = ld_v(lv, i, j, ny, nx), lv_b = ld_v(lv, i + 1, j, ny, nx);
        acc += lu_l * sx.d_eta_H * eta * dx_inv;
        acc += -lu_r * sx.d_eta_H * eta * dx_inv;
        acc += -lv_t * sy.d_eta_H * eta * dx_inv;
        acc += lv_b * sy.d_eta_H * eta * dx_inv;
        // shear stresses of the four corner vertices
        const double q4 = 0.25 * eta;
        for (int c = 0; c < 4; c++) {
            const int a = i + (c >> 1), b = j + (c & 1);
            const double dv_ = vertex_deps(a, b, ny, nx, dx_inv, u, v);
            const double w = ld_u(lu, a, b, ny, nx) - ld_u(lu, a - 1, b, ny, nx) - ld_v(lv, a, b, ny, nx) +
                             ld_v(lv, a, b - 1, ny, nx);
            acc += w * dv_ * q4 * dx_inv;
        }
        // basal drag (grounding-line derivatives) and driving stress on the four faces
        {
            const Ground ga = G(i, j - 1), gb = G(i, j);
            TauB t = tau_b_g(ul, ga.g, gb.g, ga.gd, gb.gd, ld_c(beta, i, j - 1, ny, nx), ld_c(beta, i, j, ny, nx), p);
            TauD d = tau_dx(Hl, Hc, ld_c(bed, i, j - 1, ny, nx), ld_c(bed, i, j, ny, nx), dx_inv, j, nx);
            acc += lu_l * t.d_H_r;
            acc += -lu_l * d.d_H_r;
        }
        {
            const Ground ga = G(i, j), gb = G(i, j + 1);
            TauB t = tau_b_g(ur, ga.g, gb.g, ga.gd, gb.gd, ld_c(beta, i, j, ny, nx), ld_c(beta, i, j + 1, ny, nx), p);
            TauD d = tau_dx(Hc, Hr, ld_c(bed, i, j, ny, nx), ld_c(bed, i, j + 1, ny, nx), dx_inv, j + 1, nx);
            acc += lu_r * t.d_H_l;
            acc += -lu_r * d.d_H_l;
        }
        {
            const Ground ga = G(i - 1, j), gb = G(i, j);
            TauB t = tau_b_g(vt, ga.g, gb.g, ga.gd, gb.gd, ld_c(beta, i - 1, j, ny, nx), ld_c(beta, i, j, ny, nx), p);
            TauD d = tau_dy(Ht, Hc, ld_c(bed, i - 1, j, ny, nx), ld_c(bed, i, j, ny, nx), dx_inv, i, ny);
            acc += lv_t * t.d_H_r;
            acc += -lv_t * d.d_H_r;
        }
        {
            const Ground ga = G(i, j), gb = G(i + 1, j);
            TauB t = tau_b_g(vb, ga.g, gb.g, ga.gd, gb.gd, ld_c(beta, i, j, ny, nx), ld_c(beta, i + 1, j, ny, nx), p);
            TauD d = tau_dy(Hc, Hb, ld_c(bed, i, j, ny, nx), ld_c(bed, i + 1, j, ny, nx), dx_inv, i + 1, ny);
            acc += lv_b * t.d_H_l;
            acc += -lv_b * d.d_H_l;
        }
        emit(nu + nv + (size_t)i * nx + j, acc);
    }
}

// LDS-staged JVP and VJP on 64 x LTY point tiles (lin_tile_lds).  Tiles are
// numbered down the columns of tiles and dealt to the XCDs in chunks of
// consecutive tiles (chunked_tile), so a chunk's vertical neighbours, which
// share the larger halos, meet in one L2.
template <int LTY>
__global__ __launch_bounds__(TX *LTY) void k_jvp(GlideGeom g, Params p, double *__restrict__ jo,
                                                 const double *__restrict__ U, const double *__restrict__ dU,
                                                 const double *__restrict__ bed, const double *__restrict__ B,
                                                 const double *__restrict__ beta, const double *__restrict__ mask)
{
    __shared__ LinLDS<TX, LTY> S;
    const int tiles_y = cdiv(g.ny + 1, LTY), total = cdiv(g.nx + 1, TX) * tiles_y;
    const int t = chunked_tile(blockIdx.x, total);
    if (t >= total) return;
    lin_tile_lds<TX, LTY, false>(g, p, jo, nullptr, U, dU, bed, B, beta, mask, t / tiles_y, t % tiles_y, S);
}
template <int LTY>
__global__ __launch_bounds__(TX *LTY) void k_vjp(GlideGeom g, Params p, double *__restrict__ out,
                                                 const double *__restrict__ base, const double *__restrict__ U,
                                                 const double *__restrict__ lam, const double *__restrict__ bed,
                                                 const double *__restrict__ B, const double *__restrict__ beta,
                                                 const double *__restrict__ mask)
{
    __shared__ LinLDS<TX, LTY> S;
    const int tiles_y = cdiv(g.ny + 1, LTY), total = cdiv(g.nx + 1, TX) * tiles_y;
    const int t = chunked_tile(blockIdx.x, total);
    if (t >= total) return;
    lin_tile_lds<TX, LTY, true>(g, p, out, base, U, lam, bed, B, beta, mask, t / tiles_y, t % tiles_y, S);
}
// grid of the staged JVP / VJP
inline dim3 lin_grid(const GlideGeom &g, int ty) { return dim3(chunked_grid(cdiv(g.nx + 1, TX) * cdiv(g.ny + 1, ty))); }

// Adjoint red-black Vanka pass, in place: cells of one colour own disjoint
// dofs, so lam is read and written by one lane per dof
// (glide/cuda/vanka.cu:441-840 with Lambda_out = Lambda, grid.py:291-303)
__global__ __launch_bounds__(NT) void k_vanka_adj(GlideGeom g, Params p, double *lam_out, const double *lam,
                                                  const double *__restrict__ mask, const double *__restrict__ r_adj,
                                                  const double *__restrict__ U, const double *__restrict__ bed,
                                                  const double *__restrict__ B, const double *__restrict__ beta,
                                                  const double *__restrict__ gamma, int color, double omega)
{
    __shared__ double eta_s[TY + 2][TX + 2];
    const int ny = g.ny, nx = g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    fill_eta_tile<TX, TY>(eta_s, i0, j0, u, v, B, p, g.dx, ny, nx);
    __syncthreads();
    if (i >= ny || j >= nx || ((i + j) & 1) != color) return;
    auto E = [&](int ci, int cj) { return eta_s[ci - i0 + 1][cj - j0 + 1]; };
    const double dx_inv = 1.0 / g.dx, dt = g.dt;
    CellNbhd N = load_nbhd(i, j, ny, nx, nu, nv, u, v, H, (const double *)nullptr, bed, beta, E);
    nbhd_grounding(N, p);
    double J[25], r[5], JT[25], rhs[5], d[5];
    const bool masked = ld_c(mask, i, j, ny, nx) > 0.5;
    vanka_assemble<1>(N, p, i, j, ny, nx, dx_inv, dt, ld_u(u, i, j, ny, nx), ld_u(u, i, j + 1, ny, nx),
                      ld_v(v, i, j, ny, nx), ld_v(v, i + 1, j, ny, nx), ld_c(H, i, j, ny, nx), ld_c(gamma, i, j, ny, nx),
                      masked, J, r);
#pragma unroll
    for (int a = 0; a < 5; a++)
#pragma unroll
        for (int b = 0; b < 5; b++) JT[a * 5 + b] = J[b * 5 + a];
    const size_t ku0 = (size_t)i * (nx + 1) + j, ku1 = ku0 + 1;
    const size_t kv0 = nu + (size_t)i * nx + j, kv1 = nu + (size_t)(i + 1) * nx + j;
    const size_t kh = nu + nv + (size_t)i * nx + j;
    rhs[0] = r_adj[ku0];
    rhs[1] = r_adj[ku1];
    rhs[2] = r_adj[kv0];
    rhs[3] = r_adj[kv1];
    rhs[4] = r_adj[kh];
    lu5_solve(JT, rhs, d);
    const double l0 = lam[ku0], l1 = lam[ku1], l2 = lam[kv0], l3 = lam[kv1], l4 = lam[kh];
    lam_out[ku0] = l0 + 0.5 * omega * d[0];
    lam_out[ku1] = l1 + 0.5 * omega * d[1];
    lam_out[kv0] = l2 + 0.5 * omega * d[2];
    lam_out[kv1] = l3 + 0.5 * omega * d[3];
    lam_out[kh] = l4 + omega * d[4];
}

// grad_beta = lambda . dR/dbeta, gather form in the oracle's summation order
// (glide/cuda/grad.cu:1-74); ref_edges reproduces the reference's skipped
// edge terms, otherwise boundary faces' ghost terms go to the clamped cell.
__global__ void k_grad_beta(GlideGeom g, Params p, double *__restrict__ grad, const double *__restrict__ U,
                            const double *__restrict__ lam, const double *__restrict__ bed,
                            const double *__restrict__ beta, int ref_edges)
{
    const int ny = g.ny, nx = g.nx;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ny * nx) return;
    const int i = idx / nx, j = idx - i * nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv, *lu = lam, *lv = lam + nu;
    auto tbx = [&](int a, int b) {   // drag on u face (a, b)
        return tau_b(ld_u(u, a, b, ny, nx), ld_c(H, a, b - 1, ny, nx), ld_c(H, a, b, ny, nx), ld_c(bed, a, b - 1, ny, nx),
                     ld_c(bed, a, b, ny, nx), ld_c(beta, a, b - 1, ny, nx), ld_c(beta, a, b, ny, nx), p);
    };
    auto tby = [&](int a, int b) {   // drag on v face (a, b)
        return tau_b(ld_v(v, a, b, ny, nx), ld_c(H, a - 1, b, ny, nx), ld_c(H, a, b, ny, nx), ld_c(bed, a - 1, b, ny, nx),
                     ld_c(bed, a, b, ny, nx), ld_c(beta, a - 1, b, ny, nx), ld_c(beta, a, b, ny, nx), p);
    };
    double acc = 0.0;
    {   // u face (i, j): this cell is its right cell
        TauB t = tbx(i, j);
        double l = lu[(size_t)i * (nx + 1) + j];
        if (!ref_edges && j == 0) acc += l * t.d_beta_l;
        if (!ref_edges || j < nx - 1) acc += l * t.d_beta_r;
    }
    {   // v face (i, j): this cell is its bottom cell
        TauB t = tby(i, j);
        double l = lv[(size_t)i * nx + j];
        if (!ref_edges && i == 0) acc += l * t.d_beta_l;
        if (!ref_edges || i < ny - 1) acc += l * t.d_beta_r;
    }
    {   // u face (i, j+1): left cell
        TauB t = tbx(i, j + 1);
        double l = lu[(size_t)i * (nx + 1) + j + 1];
        acc += l * t.d_beta_l;
        if (!ref_edges && j + 1 == nx) acc += l * t.d_beta_r;
    }
    {   // v face (i+1, j): top cell
        TauB t = tby(i + 1, j);
        double l = lv[(size_t)(i + 1) * nx + j];
        acc += l * t.d_beta_l;
        if (!ref_edges && i + 1 == ny) acc += l * t.d_beta_r;
    }
    grad[idx] = acc;
}

// mask = (H <= gamma + 0.1) (glide/grid.py:332-334)
__global__ void k_compute_mask(size_t nc, double *__restrict__ mask, const double *__restrict__ H,
                               const double *__restrict__ gamma)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nc) mask[k] = (H[k] <= gamma[k] + 0.1) ? 1.0 : 0.0;
}

// lam += P(lam_c): bilinear for all three fields (glide/solver.py:207-215)
__global__ void k_adj_correct(int ny, int nx, double *__restrict__ lam, const double *__restrict__ lamc)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx);
    if (k >= nu + nv + nc) return;
    const int nyc = ny / 2, nxc = nx / 2;
    const size_t nuc = n_u(nyc, nxc), nvc = n_v(nyc, nxc);
    double z;
    if (k < nu) {
        int j = (int)(k / (nx + 1)), i = (int)(k - (size_t)j * (nx + 1));
        z = prol_u_at(lamc, nullptr, j, i, nyc, nxc);
    } else if (k < nu + nv) {
        size_t kk = k - nu;
        int j = (int)(kk / nx), i = (int)(kk - (size_t)j * nx);
        z = prol_v_at(lamc + nuc, nullptr, j, i, nyc, nxc);
    } else {
        size_t kk = k - nu - nv;
        int j = (int)(kk / nx), i = (int)(kk - (size_t)j * nx);
        z = prol_H_at(lamc + nuc + nvc, nullptr, j, i, nyc, nxc);
    }
    lam[k] = lam[k] + z;
}

// dL/dB = lambda . dR/dB: B enters the residual only through eta = 1/2 B
// (...)^p, so d eta_c / d B_c = eta_c / B_c, and eta_c only through eta*H in
// sigma_xx and sigma_yy of cell c and in the vertex average of sigma_xy at
// its interior corners.  One thread per cell, gather form.
__global__ void k_grad_B(GlideGeom g, Params p, double *__restrict__ grad, const double *__restrict__ U,
                         const double *__restrict__ lam, const double *__restrict__ B)
{
    const int ny = g.ny, nx = g.nx;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ny * nx) return;
    const int i = idx / nx, j = idx - i * nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
    const double *u = U, *v = U + nu, *H = U + nu + nv, *lu = lam, *lv = lam + nu;
    const double dx_inv = 1.0 / g.dx;
    const double eta = eta_cell(i, j, u, v, B, p, g.dx, ny, nx);
    const double u_l = ld_u(u, i, j, ny, nx), u_r = ld_u(u, i, j + 1, ny, nx);
    const double v_t = ld_v(v, i, j, ny, nx), v_b = ld_v(v, i + 1, j, ny, nx);
    const double exx = fma(2.0 * (u_r - u_l), dx_inv, (v_t - v_b) * dx_inv);
    const double eyy = fma(u_r - u_l, dx_inv, 2.0 * (v_t - v_b) * dx_inv);
    double acc = 2.0 * exx * (ld_u(lu, i, j, ny, nx) - ld_u(lu, i, j + 1, ny, nx));
    acc = fma(2.0 * eyy, ld_v(lv, i + 1, j, ny, nx) - ld_v(lv, i, j, ny, nx), acc);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int a = i + (k >> 1), b = j + (k & 1);
        if (a < 1 || a > ny - 1 || b < 1 || b > nx - 1) continue;
        const double exy = 0.5 * fma(ld_u(u, a - 1, b, ny, nx) - ld_u(u, a, b, ny, nx), dx_inv,
                                     (ld_v(v, a, b, ny, nx) - ld_v(v, a, b - 1, ny, nx)) * dx_inv);
        const double comb = (ld_u(lu, a, b, ny, nx) - ld_u(lu, a - 1, b, ny, nx)) -
                            (ld_v(lv, a, b, ny, nx) - ld_v(lv, a, b - 1, ny, nx));
        acc = fma(0.5 * exy, comb, acc);
    }
    grad[idx] = acc * H[idx] * dx_inv * (eta / B[idx]);
}

// ===========================================================================
// Host side
// ===========================================================================
namespace {

// zero (optional): an int the launch also sets to 0
int launch_residual(const GlideGeom &g, const Params &p, double *r, const double *U, const double *f, const double *sub,
                    const double *bed, const double *B, const double *beta, const double *mask, const double *gamma,
                    hipStream_t st, int *zero = nullptr)
{
    const int ty = residual_ty(g.ny, g.nx);
    if (ty == 16)
        hipLaunchKernelGGL(k_residual<16>, residual_grid(g.ny, g.nx, 16), dim3(TX, 16), 0, st, g, p, r, U, f, sub, bed,
                           B, beta, mask, gamma, zero);
    else if (ty == 8)
        hipLaunchKernelGGL(k_residual<8>, residual_grid(g.ny, g.nx, 8), dim3(TX, 8), 0, st, g, p, r, U, f, sub, bed, B,
                           beta, mask, gamma, zero);
    else
        hipLaunchKernelGGL(k_residual<4>, residual_grid(g.ny, g.nx, 4), dim3(TX, 4), 0, st, g, p, r, U, f, sub, bed, B,
                           beta, mask, gamma, zero);
    return check_launch("residual");
}

inline int nbx_of(int nx) { return nx / RB + 1; }
inline int nby_of(int ny) { return ny / RB + 1; }

int launch_sweeps(const GlideGeom &g, const Params &p, double *U, const double *f, const double *bed, const double *B,
                  const double *beta, const double *gamma, const unsigned char *blk, int n_sweeps, int n_newton,
                  double *scratch, hipStream_t st)
{
    const int nbx = nbx_of(g.nx);
    const bool small = (long)g.ny * g.nx <= g_small_max_cells;
    const bool rows = small && g_small_variant == 1;
    const dim3 rgrid(cdiv(g.nx, 64), g.ny), rblock(64, ROW_WAVES);
    if (small && g_small_variant == 2) {
        // The sweeps alternate the state between U and the scratch (5 nc >=
        // nu + nv + nc doubles); an odd count runs its first sweep in place
        // with the apply pass, so that the last sweep writes U.
        int s = 0;
        if (n_sweeps & 1) {
            if (blk) {
                hipLaunchKernelGGL((k_vanka<true, 1>), vanka_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, U, f, bed,
                                   B, beta, gamma, blk, nbx, n_newton, scratch);
                hipLaunchKernelGGL(k_vanka_apply<true>, point_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g.ny, g.nx, U,
                                   scratch, blk, nbx);
            } else {
                hipLaunchKernelGGL((k_vanka<false, 1>), vanka_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, U, f, bed,
                                   B, beta, gamma, blk, nbx, n_newton, scratch);
                hipLaunchKernelGGL(k_vanka_apply<false>, point_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g.ny, g.nx, U,
                                   scratch, blk, nbx);
            }
            s = 1;
        }
        for (; s < n_sweeps; s++) {
            const bool to_u = ((n_sweeps - 1 - s) & 1) == 0;
            const double *src = to_u ? scratch : U;
            double *dst = to_u ? U : scratch;
            if (blk)
                hipLaunchKernelGGL(k_vanka_halo<true>, halo_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, src, dst, f,
                                   bed, B, beta, gamma, blk, nbx, n_newton);
            else
                hipLaunchKernelGGL(k_vanka_halo<false>, halo_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, src, dst,
                                   f, bed, B, beta, gamma, blk, nbx, n_newton);
        }
        return check_launch("vanka sweep");
    }
    for (int s = 0; s < n_sweeps; s++) {
        if (rows) {
            if (blk) {
                hipLaunchKernelGGL(k_vanka_rows<true>, rgrid, rblock, 0, st, g, p, U, f, bed, B, beta, gamma, blk, nbx,
                                   n_newton, scratch);
                hipLaunchKernelGGL(k_vanka_apply<true>, point_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g.ny, g.nx, U,
                                   scratch, blk, nbx);
            } else {
                hipLaunchKernelGGL(k_vanka_rows<false>, rgrid, rblock, 0, st, g, p, U, f, bed, B, beta, gamma, blk,
                                   nbx, n_newton, scratch);
                hipLaunchKernelGGL(k_vanka_apply<false>, point_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g.ny, g.nx, U,
                                   scratch, blk, nbx);
            }
            continue;
        }
        if (blk) {
            if (small)
                hipLaunchKernelGGL((k_vanka<true, 1>), vanka_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, U, f, bed, B,
                                   beta, gamma, blk, nbx, n_newton, scratch);
            else
                hipLaunchKernelGGL((k_vanka<true, 2>), vanka_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, U, f, bed, B,
                                   beta, gamma, blk, nbx, n_newton, scratch);
            hipLaunchKernelGGL(k_vanka_apply<true>, point_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g.ny, g.nx, U,
                               scratch, blk, nbx);
        } else {
            if (small)
                hipLaunchKernelGGL((k_vanka<false, 1>), vanka_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, U, f, bed,
                                   B, beta, gamma, blk, nbx, n_newton, scratch);
            else
                hipLaunchKernelGGL((k_vanka<false, 2>), vanka_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, U, f, bed,
                                   B, beta, gamma, blk, nbx, n_newton, scratch);
            hipLaunchKernelGGL(k_vanka_apply<false>, point_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g.ny, g.nx, U,
                               scratch, blk, nbx);
        }
    }
    return check_launch("vanka sweep");
}

int launch_activation(const GlideGeom &g, const double *r, const double *U, const double *gamma, unsigned char *blk,
                      hipStream_t st, int *any_active = nullptr)
{
    hipLaunchKernelGGL(k_local_activation, dim3(nbx_of(g.nx), nby_of(g.ny)), dim3(16, 16), 0, st, g.ny, g.nx, r, U,
                       gamma, blk, any_active);
    return check_launch("local activation");
}

}  // namespace

// Host side of the single-domain V-cycle's idle-pass check (include/glide_hip.h):
// two pinned slots for the activation flags of consecutive passes, an event
// after each copy, and the pass counters.
struct GlideVcycleWork {
    int device = 0;
    int *slot = nullptr;
    hipEvent_t ev[2] = {};
    bool pending[2] = {false, false};   // copy posted, not yet waited for
    hipStream_t stream = nullptr;       // stream of the last posted copy
    long run = 0, skipped = 0;
};

namespace {

// Make the workspace ready for a V-cycle on stream st: flag copies still in
// flight on another stream are waited for, so none can land in a slot after
// this V-cycle's own copy.
int work_begin(GlideVcycleWork *w, hipStream_t st)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail((int)e, std::string("local pass check: ") + hipGetErrorString(e));
    if (dev != w->device) return fail(GLIDE_E_ARG, "local pass check: the workspace belongs to another device");
    if (st != w->stream) {
        for (int k = 0; k < 2; k++) {
            if (!w->pending[k]) continue;
            e = hipEventSynchronize(w->ev[k]);
            if (e != hipSuccess) return fail((int)e, std::string("local pass check: ") + hipGetErrorString(e));
            w->pending[k] = false;
        }
        w->stream = st;
    }
    return GLIDE_OK;
}

// Enqueue the copy of a pass's activation flag into its pinned slot.
int post_flag(GlideVcycleWork *w, int k, const int *flag_dev, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(&w->slot[k], flag_dev, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(w->ev[k], st);
    if (e != hipSuccess) return fail((int)e, std::string("local pass check: ") + hipGetErrorString(e));
    w->pending[k] = true;
    return GLIDE_OK;
}

// Wait for slot k's copy and report whether that activation marked a block.
int read_flag(GlideVcycleWork *w, int k, bool *any)
{
    hipError_t e = hipEventSynchronize(w->ev[k]);
    if (e != hipSuccess) return fail((int)e, std::string("local pass check: ") + hipGetErrorString(e));
    w->pending[k] = false;
    *any = w->slot[k] != 0;
    return GLIDE_OK;
}

// F is the 2x refinement of C (check_levels); fill_a (optional) receives
// fill_val in its first n_fill entries in the same launch.
int restrict_state(const GlideLevel &F, const GlideLevel &C, const double *src, double *dst, double *dst2,
                   hipStream_t st, double bw = 1.0, double *fill_a = nullptr, size_t n_fill = 0, double fill_val = 0.0)
{
    (void)F;
    const int nyc = C.g.ny, nxc = C.g.nx;
    const size_t n = std::max(n_t(nyc, nxc), fill_a ? n_fill : 0);
    hipLaunchKernelGGL(k_restrict_state, dim3(blocks1d(n)), dim3(256), 0, st, src, dst, dst2, nyc, nxc, bw, fill_a,
                       fill_a ? n_fill : 0, fill_val);
    return check_launch("restrict state");
}

int restrict_cell(const double *hf, double *hc, int nyc, int nxc, hipStream_t st)
{
    hipLaunchKernelGGL(k_restrict_cell, dim3(blocks1d(n_c(nyc, nxc))), dim3(256), 0, st, hf, hc, nullptr, nyc, nxc);
    return check_launch("restrict cell");
}

int fill(double *a, size_t n, double v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill, dim3(blocks1d(n)), dim3(256), 0, st, n, a, v);
    return check_launch("fill");
}

#define TRY(x)                    \
    do {                          \
        int _rc = (x);            \
        if (_rc != GLIDE_OK) return _rc; \
    } while (0)

GlideSchedule default_schedule() { return GlideSchedule{10, 10, 200, 10, 10, 10, 10, 100, 1.0, 0}; }

int check_schedule(const GlideSchedule &sc)
{
    if (sc.pre < 0 || sc.post < 0 || sc.coarse < 0 || sc.local_passes < 0 || sc.local_sweeps < 0 || sc.n_newton < 0 ||
        sc.adj_sweeps < 0 || sc.adj_coarse < 0)
        return fail(GLIDE_E_ARG, "schedule: negative count");
    if (!std::isfinite(sc.boundary_weight)) return fail(GLIDE_E_ARG, "schedule: boundary_weight is not finite");
    return GLIDE_OK;
}

int check_levels(const GlideLevel *L, int n)
{
    if (!L) return fail(GLIDE_E_NULL, "levels is NULL");
    if (n < 1 || n > 16) return fail(GLIDE_E_SIZE, "level count out of range");
    for (int k = 0; k < n; k++) {
        TRY(check_geom(&L[k].g));
        if (!L[k].U || !L[k].f || !L[k].r || !L[k].w || !L[k].z || !L[k].bed || !L[k].B || !L[k].beta ||
            !L[k].gamma || !L[k].chi || !L[k].mask || !L[k].scratch || !L[k].active)
            return fail(GLIDE_E_NULL, "level " + std::to_string(k) + " has a NULL forward buffer");
        if (k > 0 && (L[k].g.ny * 2 != L[k - 1].g.ny || L[k].g.nx * 2 != L[k - 1].g.nx))
            return fail(GLIDE_E_SIZE, "level " + std::to_string(k) + " is not a 2x coarsening of its parent");
    }
    return GLIDE_OK;
}

// glide/solver.py:71-157
int fascd(const GlideLevel *L, int n, int k, bool finest, double thklim, const GlideSchedule &sc,
          GlideVcycleWork *work, hipStream_t st)
{
    const GlideLevel &A = L[k];
    const int ny = A.g.ny, nx = A.g.nx;
    const size_t nu = n_u(ny, nx), nv = n_v(ny, nx), nc = n_c(ny, nx), nt = nu + nv + nc;
    const Params p = to_params(A.p);
    double *wH = A.w + nu + nv;
    if (finest) {
        hipLaunchKernelGGL(k_fas_begin, dim3(blocks1d(nt)), dim3(256), 0, st, nt, nu + nv, A.w, A.U, A.chi, A.gamma);
        TRY(check_launch("fas begin"));
    }
    if (k == n - 1) {
        hipLaunchKernelGGL(k_gamma, dim3(blocks1d(nc)), dim3(256), 0, st, ny, nx, A.gamma, wH, A.chi, nullptr);
        TRY(launch_sweeps(A.g, p, A.U, A.f, A.bed, A.B, A.beta, A.gamma, nullptr, sc.coarse, sc.n_newton, A.scratch, st));
        return k > 0 ? GLIDE_OK : fill(A.gamma, nc, thklim, st);   // a parent resets gamma in its FAS correction
    }
    const GlideLevel &C = L[k + 1];
    const int nyc = C.g.ny, nxc = C.g.nx;
    const Params pc = to_params(C.p);
    hipLaunchKernelGGL(k_gamma_pre, dim3(blocks1d(nc)), dim3(256), 0, st, ny, nx, A.gamma, wH, A.chi, C.chi);
    TRY(check_launch("pre gamma"));
    TRY(launch_sweeps(A.g, p, A.U, A.f, A.bed, A.B, A.beta, A.gamma, nullptr, sc.pre, sc.n_newton, A.scratch, st));
    TRY(restrict_state(A, C, A.U, C.U, C.w, st, 1.0, A.gamma, nc, thklim));   // and gamma = thklim
    TRY(launch_residual(A.g, p, A.r, A.U, A.f, nullptr, A.bed, A.B, A.beta, A.mask, A.gamma, st));
    TRY(restrict_state(A, C, A.r, C.r, nullptr, st, sc.boundary_weight));
    // f_c = F_c(U_c) - r_c
    TRY(launch_residual(C.g, pc, C.f, C.U, nullptr, C.r, C.bed, C.B, C.beta, C.mask, C.gamma, st));
    TRY(fascd(L, n, k + 1, false, thklim, sc, work, st));
    // also the post-correction gamma and the child's gamma = thklim
    hipLaunchKernelGGL(k_fas_correct, dim3(blocks1d(nt)), dim3(256), 0, st, ny, nx, A.U, A.w, C.U, C.w, A.gamma,
                       A.chi, C.gamma, thklim);
    TRY(check_launch("fas correct"));
    TRY(launch_sweeps(A.g, p, A.U, A.f, A.bed, A.B, A.beta, A.gamma, nullptr, sc.post, sc.n_newton, A.scratch, st));
    // A pass whose activation marks no block leaves U unchanged (its apply
    // passes add +0.0), so every later pass of this V-cycle recomputes the
    // same residual and activation and is idle too.  With sc.local_skip
    // the activation also raises an any-block flag (pass k's in int slot
    // k & 1 at the start of the Vanka scratch, which the sweeps rewrite only
    // after the flag has been copied out), and the host stops the level's
    // passes once a flag reads 0:
    // - pass 0's flag is waited for before its sweeps are enqueued (levels
    //   that start idle are the common case and skip everything);
    // - from pass 2 on, the host waits for the previous pass's flag, which
    //   the GPU has produced while this pass's residual and activation were
    //   already queued, so an active level never drains the stream; an
    //   idle pass found this way has had its (idle) sweeps enqueued.
    // The only difference from running every pass is the sign of zero
    // entries of U that idle passes would turn from -0.0 into +0.0.
    GlideVcycleWork *chk = sc.local_skip && sc.local_passes > 0 ? work : nullptr;
    int *flags = reinterpret_cast<int *>(A.scratch);
    for (int pass = 0; pass < sc.local_passes; pass++) {
        // the residual launch also clears this pass's flag slot
        TRY(launch_residual(A.g, p, A.r, A.U, A.f, nullptr, A.bed, A.B, A.beta, A.mask, A.gamma, st,
                            chk ? flags + (pass & 1) : nullptr));
        if (chk) {
            const int k = pass & 1;
            TRY(launch_activation(A.g, A.r, A.U, A.gamma, A.active, st, flags + k));
            TRY(post_flag(chk, k, flags + k, st));
            bool any = true;
            if (pass == 0)
                TRY(read_flag(chk, 0, &any));
            else if (pass >= 2)
                TRY(read_flag(chk, k ^ 1, &any));   // pass - 1 (pass 0 was active)
            if (!any) {
                chk->skipped += sc.local_passes - pass;
                break;
            }
            chk->run++;
        } else {
            TRY(launch_activation(A.g, A.r, A.U, A.gamma, A.active, st));
        }
        TRY(launch_sweeps(A.g, p, A.U, A.f, A.bed, A.B, A.beta, A.gamma, A.active, sc.local_sweeps, sc.n_newton,
                          A.scratch, st));
    }
    return k > 0 ? GLIDE_OK : fill(A.gamma, nc, thklim, st);   // a parent resets gamma in its FAS correction
}


int launch_vjp(const GlideGeom &g, const Params &p, double *out, const double *base, const double *U, const double *lam,
               const double *bed, const double *B, const double *beta, const double *mask, hipStream_t st)
{
    const int ty = g_lin_ty;
    const dim3 grid4 = lin_grid(g, 4), grid8 = lin_grid(g, 8), grid6 = lin_grid(g, 6);
    if (ty == 8)
        hipLaunchKernelGGL(k_vjp<8>, grid8, dim3(TX, 8), 0, st, g, p, out, base, U, lam, bed, B, beta, mask);
    else if (ty == 6)
        hipLaunchKernelGGL(k_vjp<6>, grid6, dim3(TX, 6), 0, st, g, p, out, base, U, lam, bed, B, beta, mask);
    else if (ty == 4)
        hipLaunchKernelGGL(k_vjp<4>, grid4, dim3(TX, 4), 0, st, g, p, out, base, U, lam, bed, B, beta, mask);
    else
        hipLaunchKernelGGL(k_vjp_point, point_grid(g.ny, g.nx), dim3(TX, TY), 0, st, g, p, out, base, U, lam, bed, B,
                           beta, mask);
    return check_launch("vjp");
}

int compute_mask(const GlideLevel &A, hipStream_t st)
{
    const size_t nc = n_c(A.g.ny, A.g.nx), off = n_u(A.g.ny, A.g.nx) + n_v(A.g.ny, A.g.nx);
    hipLaunchKernelGGL(k_compute_mask, dim3(blocks1d(nc)), dim3(256), 0, st, nc, A.mask, A.U + off, A.gamma);
    return check_launch("compute_mask");
}

// glide/grid.py:319-330: per sweep, for each colour: r_adj = f_adj - J^T lam; smooth
int adjoint_sweeps(const GlideLevel &A, int n, double omega, hipStream_t st)
{
    const Params p = to_params(A.p);
    for (int s = 0; s < n; s++) {
        for (int color = 0; color < 2; color++) {
            TRY(launch_vjp(A.g, p, A.r_adj, A.f_adj, A.U, A.Lambda, A.bed, A.B, A.beta, A.mask, st));
            hipLaunchKernelGGL(k_vanka_adj, cell_grid(A.g.ny, A.g.nx), dim3(TX, TY), 0, st, A.g, p, A.Lambda, A.Lambda,
                               A.mask, A.r_adj, A.U, A.bed, A.B, A.beta, A.gamma, color, omega);
            TRY(check_launch("vanka adjoint"));
        }
    }
    return GLIDE_OK;
}

int check_adjoint_levels(const GlideLevel *L, int n)
{
    TRY(check_levels(L, n));
    for (int k = 0; k < n; k++)
        if (!L[k].Lambda || !L[k].f_adj || !L[k].r_adj)
            return fail(GLIDE_E_NULL, "level " + std::to_string(k) + " has a NULL adjoint buffer");
    return GLIDE_OK;
}

// glide/solver.py:160-220
int adjoint_vc(const GlideLevel *L, int n, int k, const GlideSchedule &sc, hipStream_t st)
{
    const GlideLevel &A = L[k];
    const size_t nc = n_c(A.g.ny, A.g.nx), nt = n_t(A.g.ny, A.g.nx);
    if (k == n - 1) {
        TRY(compute_mask(A, st));
        TRY(adjoint_sweeps(A, sc.adj_coarse, 1.0, st));
        return fill(A.mask, nc, 0.0, st);
    }
    const GlideLevel &C = L[k + 1];
    const int nyc = C.g.ny, nxc = C.g.nx;
    TRY(compute_mask(A, st));
    TRY(adjoint_sweeps(A, sc.adj_sweeps, 1.0, st));
    TRY(fill(A.mask, nc, 0.0, st));
    TRY(launch_vjp(A.g, to_params(A.p), A.r_adj, A.f_adj, A.U, A.Lambda, A.bed, A.B, A.beta, A.mask, st));
    // restrict_solution(adjoint=True) writes the child's lambda, which is reset below: skipped
    TRY(restrict_cell(A.bed, C.bed, nyc, nxc, st));
    TRY(restrict_cell(A.B, C.B, nyc, nxc, st));
    TRY(restrict_cell(A.beta, C.beta, nyc, nxc, st));
    if (A.H_prev && C.H_prev) TRY(restrict_cell(A.H_prev, C.H_prev, nyc, nxc, st));
    if (A.smb && C.smb) TRY(restrict_cell(A.smb, C.smb, nyc, nxc, st));
    TRY(restrict_state(A, C, A.r_adj, C.f_adj, nullptr, st));
    TRY(fill(C.Lambda, n_t(nyc, nxc), 0.0, st));
    if (C.Lambda_out) TRY(fill(C.Lambda_out, n_t(nyc, nxc), 0.0, st));
    TRY(adjoint_vc(L, n, k + 1, sc, st));
    hipLaunchKernelGGL(k_adj_correct, dim3(blocks1d(nt)), dim3(256), 0, st, A.g.ny, A.g.nx, A.Lambda, C.Lambda);
    TRY(check_launch("adjoint correction"));
    TRY(compute_mask(A, st));
    TRY(adjoint_sweeps(A, sc.adj_sweeps, 1.0, st));
    return fill(A.mask, nc, 0.0, st);
}
}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free Vanka sweep: the per-cell updates go to a static device
// array of 5 x 2^20 doubles (40 MiB, levels up to 1024 x 1024 cells) that
// every GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel: its device address is looked up
// once per device (hipGetSymbolAddress, with that device current) and kept.
// ---------------------------------------------------------------------------
namespace noalloc {
constexpr long MAX_CELLS = 1L << 20;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[5 * MAX_CELLS];
std::mutex mu;
double *ptr[MAX_DEVICES] = {};

// Device address of this device's copy of scratch.
int address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail((int)e, std::string("vanka_sweep_noalloc: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(GLIDE_E_ARG, "vanka_sweep_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(e != hipSuccess ? (int)e : GLIDE_E_NULL,
                        std::string("vanka_sweep_noalloc: scratch address: ") + hipGetErrorString(e));
        ptr[dev] = static_cast<double *>(p);
    }
    *out = ptr[dev];
    return GLIDE_OK;
}
}  // namespace noalloc

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

size_t glide_last_error(char *buf, size_t cap)
{
    const size_t n = g_last_error.size();
    if (buf && cap > 0) {
        const size_t m = n < cap - 1 ? n : cap - 1;
        std::memcpy(buf, g_last_error.data(), m);
        buf[m] = '\0';
    }
    return n;
}

int glide_vcycle_work_create(GlideVcycleWork **out)
{
    if (!out) return fail(GLIDE_E_NULL, "vcycle_work_create: out is NULL");
    *out = nullptr;
    GlideVcycleWork *w = new GlideVcycleWork();
    hipError_t e = hipGetDevice(&w->device);
    if (e == hipSuccess) e = hipHostMalloc((void **)&w->slot, 2 * sizeof(int), hipHostMallocDefault);
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipEventCreateWithFlags(&w->ev[k], hipEventDisableTiming);
    if (e != hipSuccess) {
        glide_vcycle_work_destroy(w);
        return fail((int)e, std::string("vcycle_work_create: ") + hipGetErrorString(e));
    }
    w->slot[0] = w->slot[1] = 0;
    *out = w;
    return GLIDE_OK;
}

int glide_vcycle_work_destroy(GlideVcycleWork *w)
{
    if (!w) return GLIDE_OK;
    for (int k = 0; k < 2; k++)
        if (w->pending[k] && w->ev[k]) (void)hipEventSynchronize(w->ev[k]);
    for (int k = 0; k < 2; k++)
        if (w->ev[k]) (void)hipEventDestroy(w->ev[k]);
    if (w->slot) (void)hipHostFree(w->slot);
    delete w;
    return GLIDE_OK;
}

int glide_vcycle_work_counts(GlideVcycleWork *w, long *run, long *skipped, int reset)
{
    if (!w || !run || !skipped) return fail(GLIDE_E_NULL, "vcycle_work_counts: NULL argument");
    *run = w->run;
    *skipped = w->skipped;
    if (reset) w->run = w->skipped = 0;
    return GLIDE_OK;
}

int glide_set_small_levels(long max_cells, int variant)
{
    if (max_cells < 0) return fail(GLIDE_E_ARG, "max_cells must be >= 0");
    if (variant < 0 || variant > 2)
        return fail(GLIDE_E_ARG, "small-level variant must be 0 (per-cell), 1 (row split) or 2 (apply folded in)");
    g_small_max_cells = max_cells;
    g_small_variant = variant;
    return GLIDE_OK;
}

int glide_set_tile_height(int kernel, int ty)
{
    if (kernel == 0) {
        if (ty != 0 && ty != 4 && ty != 8 && ty != 16)
            return fail(GLIDE_E_ARG, "residual tile height must be 0 (by level size), 4, 8 or 16");
        g_res_ty = ty;
    } else if (kernel == 1) {
        if (ty != 0 && ty != 4 && ty != 6 && ty != 8) return fail(GLIDE_E_ARG, "JVP/VJP tile height must be 0, 4, 6 or 8");
        g_lin_ty = ty;
    } else {
        return fail(GLIDE_E_ARG, "kernel must be 0 (residual) or 1 (JVP/VJP)");
    }
    return GLIDE_OK;
}

int glide_set_persistence(int kernel, int wg_per_cu)
{
    if (kernel < 0 || kernel > 1) return fail(GLIDE_E_ARG, "kernel must be 0 (Vanka sweep) or 1 (residual)");
    if (wg_per_cu < 1 || wg_per_cu > 64) return fail(GLIDE_E_ARG, "workgroups per CU must be in [1, 64]");
    g_wg_per_cu[kernel] = wg_per_cu;
    return GLIDE_OK;
}

const char *glide_version(void) { return "glide_amd 0.1 (gfx950, fp64)"; }

int glide_residual(const GlideGeom *g, const GlideParams *p, double *r, const double *U, const double *f,
                   const double *sub, const double *bed, const double *B, const double *beta, const double *mask,
                   const double *gamma, void *stream)
{
    TRY(check_geom(g));
    if (!p || !r || !U || !bed || !B || !beta || !mask || !gamma) return fail(GLIDE_E_NULL, "residual: NULL argument");
    return launch_residual(*g, to_params(*p), r, U, f, sub, bed, B, beta, mask, gamma, S(stream));
}

int glide_vanka_sweep(const GlideGeom *g, const GlideParams *p, double *U, const double *f, const double *bed,
                      const double *B, const double *beta, const double *gamma, const unsigned char *active,
                      int n_sweeps, int n_newton, double *scratch, void *stream)
{
    TRY(check_geom(g));
    if (!p || !U || !f || !bed || !B || !beta || !gamma || !scratch) return fail(GLIDE_E_NULL, "vanka: NULL argument");
    if (n_sweeps < 0 || n_newton < 0) return fail(GLIDE_E_ARG, "vanka: negative count");
    return launch_sweeps(*g, to_params(*p), U, f, bed, B, beta, gamma, active, n_sweeps, n_newton, scratch, S(stream));
}


int glide_vanka_sweep_noalloc(const GlideGeom *g, const GlideParams *p, double *U, const double *f, const double *bed,
                              const double *B, const double *beta, const double *gamma, const unsigned char *active,
                              int n_sweeps, int n_newton, void *stream)
{
    TRY(check_geom(g));
    if (!p || !U || !f || !bed || !B || !beta || !gamma) return fail(GLIDE_E_NULL, "vanka_sweep_noalloc: NULL argument");
    if (n_sweeps < 0 || n_newton < 0) return fail(GLIDE_E_ARG, "vanka_sweep_noalloc: negative count");
    if ((long)g->ny * g->nx > noalloc::MAX_CELLS)
        return fail(GLIDE_E_SIZE, "vanka_sweep_noalloc: " + std::to_string(g->ny) + " x " + std::to_string(g->nx) +
                                      " cells exceed the static scratch (" + std::to_string(noalloc::MAX_CELLS) +
                                      " cells)");
    double *scratch = nullptr;
    TRY(noalloc::address(&scratch));
    return launch_sweeps(*g, to_params(*p), U, f, bed, B, beta, gamma, active, n_sweeps, n_newton, scratch, S(stream));
}

int glide_vanka_block(const GlideGeom *g, const GlideParams *p, const double *U, const double *f, const double *bed,
                      const double *B, const double *beta, const double *gamma, int k_updates, double *out,
                      void *stream)
{
    TRY(check_geom(g));
    if (!p || !U || !bed || !B || !beta || !gamma || !out) return fail(GLIDE_E_NULL, "vanka_block: NULL argument");
    if (k_updates < 0) return fail(GLIDE_E_ARG, "vanka_block: negative count");
    hipLaunchKernelGGL(k_vanka_block, cell_grid(g->ny, g->nx), dim3(TX, TY), 0, S(stream), *g, to_params(*p), U, f,
                       bed, B, beta, gamma, k_updates, out);
    return check_launch("vanka_block");
}

int glide_local_activation(const GlideGeom *g, const double *r, const double *U, const double *gamma,
                           unsigned char *active, void *stream)
{
    TRY(check_geom(g));
    if (!r || !U || !gamma || !active) return fail(GLIDE_E_NULL, "activation: NULL argument");
    return launch_activation(*g, r, U, gamma, active, S(stream));
}


int glide_jvp(const GlideGeom *g, const GlideParams *p, double *j, const double *U, const double *dU, const double *bed,
              const double *B, const double *beta, const double *mask, const double *gamma, void *stream)
{
    TRY(check_geom(g));
    if (!p || !j || !U || !dU || !bed || !B || !beta || !mask) return fail(GLIDE_E_NULL, "jvp: NULL argument");
    (void)gamma;
    const int ty = g_lin_ty;
    const dim3 grid4 = lin_grid(*g, 4), grid8 = lin_grid(*g, 8), grid6 = lin_grid(*g, 6);
    if (ty == 8)
        hipLaunchKernelGGL(k_jvp<8>, grid8, dim3(TX, 8), 0, S(stream), *g, to_params(*p), j, U, dU, bed, B, beta, mask);
    else if (ty == 6)
        hipLaunchKernelGGL(k_jvp<6>, grid6, dim3(TX, 6), 0, S(stream), *g, to_params(*p), j, U, dU, bed, B, beta, mask);
    else if (ty == 4)
        hipLaunchKernelGGL(k_jvp<4>, grid4, dim3(TX, 4), 0, S(stream), *g, to_params(*p), j, U, dU, bed, B, beta, mask);
    else
    hipLaunchKernelGGL(k_jvp_point, point_grid(g->ny, g->nx), dim3(TX, TY), 0, S(stream), *g, to_params(*p), j, U, dU, bed,
                       B, beta, mask);
    return check_launch("jvp");
}

int glide_vjp(const GlideGeom *g, const GlideParams *p, double *l, const double *U, const double *lam, const double *bed,
              const double *B, const double *beta, const double *mask, const double *gamma, void *stream)
{
    TRY(check_geom(g));
    if (!p || !l || !U || !lam || !bed || !B || !beta || !mask) return fail(GLIDE_E_NULL, "vjp: NULL argument");
    (void)gamma;
    return launch_vjp(*g, to_params(*p), l, nullptr, U, lam, bed, B, beta, mask, S(stream));
}

int glide_vanka_adjoint_pass(const GlideGeom *g, const GlideParams *p, double *lam_out, const double *lam,
                             const double *mask, const double *r_adj, const double *U, const double *bed,
                             const double *B, const double *beta, const double *gamma, int color, double omega,
                             void *stream)
{
    TRY(check_geom(g));
    if (!p || !lam_out || !lam || !mask || !r_adj || !U || !bed || !B || !beta || !gamma)
        return fail(GLIDE_E_NULL, "vanka_adjoint_pass: NULL argument");
    if (color != 0 && color != 1) return fail(GLIDE_E_ARG, "color must be 0 or 1");
    hipLaunchKernelGGL(k_vanka_adj, cell_grid(g->ny, g->nx), dim3(TX, TY), 0, S(stream), *g, to_params(*p), lam_out,
                       lam, mask, r_adj, U, bed, B, beta, gamma, color, omega);
    return check_launch("vanka_adjoint_pass");
}

int glide_grad_beta(const GlideGeom *g, const GlideParams *p, double *grad, const double *U, const double *lam,
                    const double *bed, const double *beta, int reference_edges, void *stream)
{
    TRY(check_geom(g));
    if (!p || !grad || !U || !lam || !bed || !beta) return fail(GLIDE_E_NULL, "grad_beta: NULL argument");
    hipLaunchKernelGGL(k_grad_beta, dim3(blocks1d(n_c(g->ny, g->nx))), dim3(256), 0, S(stream), *g, to_params(*p), grad,
                       U, lam, bed, beta, reference_edges);
    return check_launch("grad_beta");
}

int glide_grad_B(const GlideGeom *g, const GlideParams *p, double *grad, const double *U, const double *lam,
                 const double *B, void *stream)
{
    TRY(check_geom(g));
    if (!p || !grad || !U || !lam || !B) return fail(GLIDE_E_NULL, "grad_B: NULL argument");
    hipLaunchKernelGGL(k_grad_B, dim3(blocks1d(n_c(g->ny, g->nx))), dim3(256), 0, S(stream), *g, to_params(*p), grad,
                       U, lam, B);
    return check_launch("grad_B");
}

int glide_adjoint_vcycle(const GlideLevel *L, int n, const GlideSchedule *sched, void *stream)
{
    TRY(check_adjoint_levels(L, n));
    GlideSchedule sc = sched ? *sched : default_schedule();
    TRY(check_schedule(sc));
    return adjoint_vc(L, n, 0, sc, S(stream));
}

int glide_compute_mask(const GlideLevel *level, void *stream)
{
    if (!level) return fail(GLIDE_E_NULL, "compute_mask: NULL");
    TRY(check_geom(&level->g));
    if (!level->U || !level->mask || !level->gamma) return fail(GLIDE_E_NULL, "compute_mask: NULL buffer");
    return compute_mask(*level, S(stream));
}

int glide_adjoint_residual(const GlideLevel *level, void *stream)
{
    if (!level) return fail(GLIDE_E_NULL, "adjoint_residual: NULL");
    TRY(check_geom(&level->g));
    const GlideLevel &A = *level;
    if (!A.U || !A.Lambda || !A.f_adj || !A.r_adj || !A.mask) return fail(GLIDE_E_NULL, "adjoint_residual: NULL buffer");
    return launch_vjp(A.g, to_params(A.p), A.r_adj, A.f_adj, A.U, A.Lambda, A.bed, A.B, A.beta, A.mask, S(stream));
}

int glide_restrict_u(const double *uf, double *uc, int nyc, int nxc, void *stream)
{
    if (!uf || !uc) return fail(GLIDE_E_NULL, "restrict_u: NULL");
    if (nyc < 1 || nxc < 1) return fail(GLIDE_E_SIZE, "restrict_u: size");
    hipLaunchKernelGGL(k_restrict_u, dim3(blocks1d(n_u(nyc, nxc))), dim3(256), 0, S(stream), uf, uc, nullptr, nyc, nxc,
                       1.0);
    return check_launch("restrict_u");
}
int glide_restrict_v(const double *vf, double *vc, int nyc, int nxc, void *stream)
{
    if (!vf || !vc) return fail(GLIDE_E_NULL, "restrict_v: NULL");
    if (nyc < 1 || nxc < 1) return fail(GLIDE_E_SIZE, "restrict_v: size");
    hipLaunchKernelGGL(k_restrict_v, dim3(blocks1d(n_v(nyc, nxc))), dim3(256), 0, S(stream), vf, vc, nullptr, nyc, nxc,
                       1.0);
    return check_launch("restrict_v");
}
int glide_restrict_cell(const double *hf, double *hc, int nyc, int nxc, void *stream)
{
    if (!hf || !hc) return fail(GLIDE_E_NULL, "restrict_cell: NULL");
    if (nyc < 1 || nxc < 1) return fail(GLIDE_E_SIZE, "restrict_cell: size");
    return restrict_cell(hf, hc, nyc, nxc, S(stream));
}
int glide_restrict_max(const double *hf, double *hc, int nyc, int nxc, void *stream)
{
    if (!hf || !hc) return fail(GLIDE_E_NULL, "restrict_max: NULL");
    if (nyc < 1 || nxc < 1) return fail(GLIDE_E_SIZE, "restrict_max: size");
    hipLaunchKernelGGL(k_restrict_max, dim3(blocks1d(n_c(nyc, nxc))), dim3(256), 0, S(stream), hf, hc, nyc, nxc);
    return check_launch("restrict_max");
}
int glide_prolong_inject(const double *hc, double *hf, int nyf, int nxf, double sign, void *stream)
{
    if (!hc || !hf) return fail(GLIDE_E_NULL, "prolong_inject: NULL");
    if (nyf < 2 || nxf < 2 || (nyf & 1) || (nxf & 1)) return fail(GLIDE_E_SIZE, "prolong_inject: size");
    hipLaunchKernelGGL(k_prolong_inject, dim3(blocks1d(n_c(nyf, nxf))), dim3(256), 0, S(stream), hc, hf, nyf, nxf, sign);
    return check_launch("prolong_inject");
}
int glide_prolong_u(const double *uc, double *uf, int nyf, int nxf, void *stream)
{
    if (!uc || !uf) return fail(GLIDE_E_NULL, "prolong_u: NULL");
    if (nyf < 2 || nxf < 2 || (nyf & 1) || (nxf & 1)) return fail(GLIDE_E_SIZE, "prolong_u: size");
    hipLaunchKernelGGL(k_prolong_u, dim3(blocks1d(n_u(nyf, nxf))), dim3(256), 0, S(stream), uc, uf, nyf, nxf);
    return check_launch("prolong_u");
}
int glide_prolong_v(const double *vc, double *vf, int nyf, int nxf, void *stream)
{
    if (!vc || !vf) return fail(GLIDE_E_NULL, "prolong_v: NULL");
    if (nyf < 2 || nxf < 2 || (nyf & 1) || (nxf & 1)) return fail(GLIDE_E_SIZE, "prolong_v: size");
    hipLaunchKernelGGL(k_prolong_v, dim3(blocks1d(n_v(nyf, nxf))), dim3(256), 0, S(stream), vc, vf, nyf, nxf);
    return check_launch("prolong_v");
}
int glide_prolong_H(const double *hc, double *hf, int nyf, int nxf, void *stream)
{
    if (!hc || !hf) return fail(GLIDE_E_NULL, "prolong_H: NULL");
    if (nyf < 2 || nxf < 2 || (nyf & 1) || (nxf & 1)) return fail(GLIDE_E_SIZE, "prolong_H: size");
    hipLaunchKernelGGL(k_prolong_H, dim3(blocks1d(n_c(nyf, nxf))), dim3(256), 0, S(stream), hc, hf, nyf, nxf);
    return check_launch("prolong_H");
}

int glide_restrict_parameters(const GlideLevel *L, int n, void *stream)
{
    TRY(check_levels(L, n));
    hipStream_t st = S(stream);
    for (int k = 0; k + 1 < n; k++) {
        int nyc = L[k + 1].g.ny, nxc = L[k + 1].g.nx;
        TRY(restrict_cell(L[k].bed, L[k + 1].bed, nyc, nxc, st));
        TRY(restrict_cell(L[k].B, L[k + 1].B, nyc, nxc, st));
        TRY(restrict_cell(L[k].beta, L[k + 1].beta, nyc, nxc, st));
        if (L[k].H_prev && L[k + 1].H_prev) TRY(restrict_cell(L[k].H_prev, L[k + 1].H_prev, nyc, nxc, st));
        if (L[k].smb && L[k + 1].smb) TRY(restrict_cell(L[k].smb, L[k + 1].smb, nyc, nxc, st));
    }
    return GLIDE_OK;
}

int glide_propagate_geometry(const GlideLevel *L, int n, double thklim, void *stream)
{
    TRY(check_levels(L, n));
    hipStream_t st = S(stream);
    for (int k = 0; k + 1 < n; k++) {
        const GlideLevel &F = L[k], &C = L[k + 1];
        int nyc = C.g.ny, nxc = C.g.nx;
        const size_t offf = n_u(F.g.ny, F.g.nx) + n_v(F.g.ny, F.g.nx), offc = n_u(nyc, nxc) + n_v(nyc, nxc);
        TRY(restrict_cell(F.bed, C.bed, nyc, nxc, st));
        TRY(restrict_cell(F.U + offf, C.U + offc, nyc, nxc, st));
        if (F.H_prev && C.H_prev) TRY(restrict_cell(F.H_prev, C.H_prev, nyc, nxc, st));
        TRY(fill(C.gamma, n_c(nyc, nxc), thklim, st));
    }
    return GLIDE_OK;
}

int glide_fascd_vcycle(const GlideLevel *L, int n, double thklim, const GlideSchedule *sched, GlideVcycleWork *work,
                       void *stream)
{
    TRY(check_levels(L, n));
    GlideSchedule sc = sched ? *sched : default_schedule();
    TRY(check_schedule(sc));
    if (sc.local_skip && sc.local_passes > 0) {
        if (!work) return fail(GLIDE_E_NULL, "fascd_vcycle: local_skip needs a GlideVcycleWork");
        TRY(work_begin(work, S(stream)));
    }
    return fascd(L, n, 0, true, thklim, sc, work, S(stream));
}

int glide_forward_setup(GlideLevel *L, int n, double dt, void *stream)
{
    TRY(check_levels(L, n));
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(GLIDE_E_ARG, "dt must be positive");
    if (!L[0].H_prev || !L[0].smb) return fail(GLIDE_E_NULL, "finest level needs H_prev and smb");
    for (int k = 0; k < n; k++) L[k].g.dt = dt;
    const size_t nc = n_c(L[0].g.ny, L[0].g.nx);
    double *fH = L[0].f + n_u(L[0].g.ny, L[0].g.nx) + n_v(L[0].g.ny, L[0].g.nx);
    hipLaunchKernelGGL(k_forcing, dim3(blocks1d(nc)), dim3(256), 0, S(stream), nc, fH, L[0].H_prev, L[0].smb, dt);
    return check_launch("forcing");
}

int glide_count_nonfinite(const double *x, long n, unsigned long long *counter, unsigned long long *count_host,
                          void *stream)
{
    if (!x || !counter || !count_host) return fail(GLIDE_E_NULL, "count_nonfinite: NULL argument");
    if (n < 0) return fail(GLIDE_E_SIZE, "count_nonfinite: negative size");
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, counter);
    if (n > 0)
        hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)std::min<size_t>(((size_t)n + 255) / 256, 2048)), dim3(256),
                           0, st, x, (size_t)n, counter);
    TRY(check_launch("count_nonfinite"));
    hipError_t e = hipMemcpyAsync(count_host, counter, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail((int)e, std::string("count_nonfinite: ") + hipGetErrorString(e));
    return GLIDE_OK;
}

int glide_residual_norm(const GlideLevel *level, double *out_host, void *stream)
{
    if (!level || !out_host) return fail(GLIDE_E_NULL, "residual_norm: NULL");
    TRY(check_geom(&level->g));
    const GlideLevel &A = *level;
    if (!A.U || !A.f || !A.r || !A.bed || !A.B || !A.beta || !A.mask || !A.gamma || !A.scratch)
        return fail(GLIDE_E_NULL, "residual_norm: the level has a NULL buffer");
    hipStream_t st = S(stream);
    // partial sums and the result go to the scratch buffer (5*nc doubles):
    // 1024 blocks where it holds 1025 values, fewer on levels below 205 cells
    const size_t cap = 5 * n_c(A.g.ny, A.g.nx) - 1;
    const int nblk = (int)(cap < 1024 ? cap : 1024);
    TRY(launch_residual(A.g, to_params(A.p), A.r, A.U, A.f, nullptr, A.bed, A.B, A.beta, A.mask, A.gamma, st));
    hipLaunchKernelGGL(k_norm_partial, dim3(nblk), dim3(256), 0, st, A.g.ny, A.g.nx, A.r, A.U, A.gamma, A.scratch);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, st, nblk, A.scratch, A.scratch + nblk);
    TRY(check_launch("residual_norm"));
    hipError_t e = hipMemcpyAsync(out_host, A.scratch + nblk, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return fail((int)e, hipGetErrorString(e));
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail((int)e, hipGetErrorString(e));
    return GLIDE_OK;
}

}  // extern "C"
