// glide_batch.hip - batched ensembles of uneven grids: every operation of the
// FASCD V-cycle runs as ONE persistent launch over all members' tiles.
//
// Plan layout.  Member m with L_m levels is aligned at its coarsest level:
// its level k sits in phase d = k + (L_max - L_m), so every member's coarse
// solve (the 200 latency-bound coarsest sweeps) shares one set of launches,
// and phase d's launches cover every (member, level) entry of that phase.
// An extra phase holds each member's finest level (forcing, residual norms).
//
// Tile tables.  For each phase and each tiling (points / cells / 14x14
// reference blocks of the entry's level or of its child, and each tile
// height TY) the plan keeps the prefix sum of the entries' tile counts.  A
// workgroup maps a global tile index to its entry by binary search.
//
// Persistent launches.  A launch has min(tiles, CUs x WGs-per-CU) workgroups
// (a multiple of 8) that loop over the tiles (tile_loop, tiles_device.h).
// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X, so
// workgroup b runs on XCD b % 8; the tile list is cut into chunks of up to 8
// consecutive tiles dealt round-robin to the XCDs, so neighbouring tiles
// (shared halo rows) meet in one L2 while spatially clustered work (the
// local smoother's active blocks) still spreads over every XCD.
//
// The per-point arithmetic is the single-level kernels' (tiles_device.h),
// so a member's trajectory in the ensemble is bitwise identical to solving it
// alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "glide_hip.h"
#include "tiles_device.h"

using namespace glide;

namespace {

constexpr int BTX = 64;           // tile width: one wave64 per tile row
constexpr int PTY = 4;            // tile height of the element-wise passes
constexpr int NSHAPE = 5;         // tile heights TY = 1, 2, 4, 8, 16
__host__ __device__ constexpr int ty_of(int s) { return 1 << s; }
constexpr int SHAPE_PTY = 2;      // index of TY = 4

enum TilingKind { TK_POINT = 0, TK_CELL = 1, TK_BLOCK = 2 };
// table index inside a phase: ((child * 3) + kind) * NSHAPE + shape
__host__ __device__ constexpr int table_id(int child, int kind, int shape) { return (child * 3 + kind) * NSHAPE + shape; }
constexpr int NTABLE = 2 * 3 * NSHAPE;

struct BEntry {
    GlideLevel L;   // this level
    GlideLevel C;   // its child (zeroed at a member's coarsest level)
    int member, level, finest, has_child;
};

enum PointOp {
    OP_FAS_BEGIN = 0,   // finest: w = U, chi = gamma - H
    OP_GAMMA_PRE,       // gamma = w_H + ((-chi_c[inject]) + chi)
    OP_GAMMA,           // gamma = w_H + chi
    OP_FILL_GAMMA,      // gamma = thklim
    OP_FAS_CORRECT,     // U = w + (P(U_c - w_c) + (U - w))
    OP_FORCING,         // finest: f_H = H_prev / dt + smb
    OP_RESTRICT_CHI,    // child tiling: chi_c = max-pool(chi)
    OP_RESTRICT_U,      // child tiling: U_c = w_c = R U
    OP_RESTRICT_R,      // child tiling: r_c = R r
};

int bfail(int code, const std::string &msg) { return glide::set_error(code, msg.c_str()); }

}  // namespace

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_entry(const int *__restrict__ pre, int n, int t)
{
    int lo = 0, hi = n;   // pre[0] = 0 <= t < pre[n]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pre[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// batched kernels
// ---------------------------------------------------------------------------
// CHILD = false: r = F(U) - f on each entry's level;
// CHILD = true : f_c = F_c(U_c) - r_c on each entry's child (FAS right-hand side)
template <int TY, bool CHILD>
__global__ __launch_bounds__(BTX *TY) void kb_residual(const BEntry *__restrict__ E, const int *__restrict__ pre, int n,
                                                        const unsigned char *__restrict__ on)
{
    __shared__ ResidLDS<BTX, TY> S;
    tile_loop(pre[n], [&](int t) {
        const int e = find_entry(pre, n, t);
        const BEntry &B = E[e];
        if (!on[B.member]) return;
        const GlideLevel &A = CHILD ? B.C : B.L;
        const int lt = t - pre[e], tx = cdiv(A.g.nx, BTX);   // cell tiles; edge faces by the last tiles
        __syncthreads();   // the previous tile's readers of S are done
        if (CHILD)
            residual_tile_lds<BTX, TY, true>(A.g, to_params(A.p), A.f, A.U, nullptr, A.r, A.bed, A.B, A.beta, A.mask,
                                             A.gamma, lt % tx, lt / tx, S);
        else
            residual_tile_lds<BTX, TY, true>(A.g, to_params(A.p), A.r, A.U, A.f, nullptr, A.bed, A.B, A.beta, A.mask,
                                             A.gamma, lt % tx, lt / tx, S);
    });
}

template <int TY, bool LOCAL>
__global__ __launch_bounds__(BTX *TY) __attribute__((amdgpu_waves_per_eu(2))) void kb_vanka(const BEntry *__restrict__ E, const int *__restrict__ pre, int n,
                                                     const unsigned char *__restrict__ on, int n_newton)
{
    __shared__ double eta_s[TY + 2][BTX + 2];
    tile_loop(pre[n], [&](int t) {
        const int e = find_entry(pre, n, t);
        const GlideLevel &A = E[e].L;
        if (!on[E[e].member]) return;
        const int lt = t - pre[e], tx = cdiv(A.g.nx, BTX);
        __syncthreads();
        vanka_tile<BTX, TY, LOCAL>(A.g, to_params(A.p), A.U, A.f, A.bed, A.B, A.beta, A.gamma, A.active,
                                   A.g.nx / RB + 1, n_newton, A.scratch, lt % tx, lt / tx, eta_s);
    });
}

template <bool LOCAL>
__global__ __launch_bounds__(BTX *PTY) void kb_apply(const BEntry *__restrict__ E, const int *__restrict__ pre, int n,
                                                      const unsigned char *__restrict__ on)
{
    tile_loop(pre[n], [&](int t) {
        const int e = find_entry(pre, n, t);
        const GlideLevel &A = E[e].L;
        if (!on[E[e].member]) return;
        const int lt = t - pre[e], tx = cdiv(A.g.nx + 1, BTX);
        apply_tile<BTX, PTY, LOCAL>(A.g.ny, A.g.nx, A.U, A.scratch, A.active, A.g.nx / RB + 1, lt % tx, lt / tx);
    });
}

__global__ __launch_bounds__(256) void kb_activation(const BEntry *__restrict__ E, const int *__restrict__ pre, int n,
                                                     const unsigned char *__restrict__ on, int *__restrict__ any_active)
{
    tile_loop(pre[n], [&](int t) {
        const int e = find_entry(pre, n, t);
        const GlideLevel &A = E[e].L;
        if (!on[E[e].member]) return;
        const int nbx = A.g.nx / RB + 1, lt = t - pre[e];
        activation_block(A.g.ny, A.g.nx, A.r, A.U, A.gamma, A.active, lt % nbx, lt / nbx, nbx, any_active);
    });
}

__global__ void kb_set_int(int *__restrict__ x, int v) { *x = v; }

// Element-wise passes on point tiles: point (i, j) owns u[i, j] (i < ny),
// v[i, j] (j < nx) and the cell (i, j) (both).  Same expressions as the
// single-level 1-D kernels.  arg is thklim, except for OP_RESTRICT_R where it
// is the weight of the restricted residual's boundary-face rows.
template <int OP>
__global__ __launch_bounds__(BTX *PTY) void kb_point(const BEntry *__restrict__ E, const int *__restrict__ pre, int n,
                                                      const unsigned char *__restrict__ on, double arg)
{
    const double thklim = arg;
    tile_loop(pre[n], [&](int t) {
        const int e = find_entry(pre, n, t);
        const BEntry &B = E[e];
        if (!on[B.member]) return;
        if ((OP == OP_FAS_BEGIN || OP == OP_FORCING) && !B.finest) return;
        constexpr bool ON_CHILD = OP >= OP_RESTRICT_CHI;
        const GlideLevel &A = ON_CHILD ? B.C : B.L;
        const int ny = A.g.ny, nx = A.g.nx;
        const int lt = t - pre[e], tx = cdiv(nx + 1, BTX);
        const int i = (lt / tx) * PTY + threadIdx.y, j = (lt % tx) * BTX + threadIdx.x;
        if (i > ny || j > nx) return;
        const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
        const bool hu = i < ny, hv = j < nx, hc = hu && hv;
        const size_t ku = (size_t)i * (nx + 1) + j, kv = nu + (size_t)i * nx + j, kc = (size_t)i * nx + j;
        if (OP == OP_FAS_BEGIN) {
            if (hu) A.w[ku] = A.U[ku];
            if (hv) A.w[kv] = A.U[kv];
            if (hc) {
                double x = A.U[nu + nv + kc];
                A.w[nu + nv + kc] = x;
                A.chi[kc] = A.gamma[kc] - x;
            }
        } else if (OP == OP_GAMMA_PRE) {
            if (hc) {
                double inj = -1.0 * B.C.chi[(size_t)(i / 2) * (nx / 2) + j / 2];
                double ph = inj + A.chi[kc];
                A.gamma[kc] = A.w[nu + nv + kc] + ph;
            }
        } else if (OP == OP_GAMMA) {
            if (hc) A.gamma[kc] = A.w[nu + nv + kc] + A.chi[kc];
        } else if (OP == OP_FILL_GAMMA) {
            if (hc) A.gamma[kc] = thklim;
        } else if (OP == OP_FORCING) {
            if (hc) A.f[nu + nv + kc] = A.H_prev[kc] / A.g.dt + A.smb[kc];
        } else if (OP == OP_FAS_CORRECT) {
            const GlideLevel &C = B.C;
            const int nyc = ny / 2, nxc = nx / 2;
            const size_t nuc = n_u(nyc, nxc), nvc = n_v(nyc, nxc);
            if (hu) {
                double z = prol_u_at(C.U, C.w, i, j, nyc, nxc);
                double y = A.U[ku] - A.w[ku];
                z += y;
                A.U[ku] = A.w[ku] + z;
            }
            if (hv) {
                double z = prol_v_at(C.U + nuc, C.w + nuc, i, j, nyc, nxc);
                double y = A.U[kv] - A.w[kv];
                z += y;
                A.U[kv] = A.w[kv] + z;
            }
            if (hc) {
                size_t kk = nuc + nvc + (size_t)(i / 2) * nxc + j / 2;
                double z = C.U[kk] - C.w[kk];
                size_t k = nu + nv + kc;
                double y = A.U[k] - A.w[k];
                z += y;
                A.U[k] = A.w[k] + z;
            }
        } else {
            // restrictions: (i, j) is a point of the child; the fine level is B.L
            const GlideLevel &F = B.L;
            const int nxf = F.g.nx;
            const size_t nuf = n_u(F.g.ny, nxf), nvf = n_v(F.g.ny, nxf);
            if (OP == OP_RESTRICT_CHI) {
                if (hc) {
                    const double *hf = F.chi;
                    double a = fmax(hf[(size_t)(2 * i) * nxf + 2 * j], hf[(size_t)(2 * i) * nxf + 2 * j + 1]);
                    double b = fmax(hf[(size_t)(2 * i + 1) * nxf + 2 * j], hf[(size_t)(2 * i + 1) * nxf + 2 * j + 1]);
                    A.chi[kc] = fmax(a, b);
                }
            } else {
                const double *src = OP == OP_RESTRICT_U ? F.U : F.r;
                double *dst = OP == OP_RESTRICT_U ? A.U : A.r;
                double *dst2 = OP == OP_RESTRICT_U ? A.w : nullptr;
                const double bw = OP == OP_RESTRICT_R ? arg : 1.0;
                if (hu) {
                    double val = 0.5 * (src[(size_t)(2 * i) * (2 * nx + 1) + 2 * j] +
                                        src[(size_t)(2 * i + 1) * (2 * nx + 1) + 2 * j]);
                    if (j == 0 || j == nx) val *= bw;
                    dst[ku] = val;
                    if (dst2) dst2[ku] = val;
                }
                if (hv) {
                    const double *vf = src + nuf;
                    double val = 0.5 * (vf[(size_t)(2 * i) * (2 * nx) + 2 * j] + vf[(size_t)(2 * i) * (2 * nx) + 2 * j + 1]);
                    if (i == 0 || i == ny) val *= bw;
                    dst[kv] = val;
                    if (dst2) dst2[kv] = val;
                }
                if (hc) {
                    const double *hf = src + nuf + nvf;
                    size_t s = 2 * (size_t)nx;
                    double val = 0.25 * (hf[(2 * i) * s + 2 * j] + hf[(2 * i) * s + 2 * j + 1] +
                                         hf[(2 * i + 1) * s + 2 * j] + hf[(2 * i + 1) * s + 2 * j + 1]);
                    size_t k = nu + nv + kc;
                    dst[k] = val;
                    if (dst2) dst2[k] = val;
                }
            }
        }
    });
}

// ||(r_u, r_v, FB(r_H, H - gamma))||^2 per point tile (fixed order), then per entry.
__global__ __launch_bounds__(BTX *PTY) void kb_norm_partial(const BEntry *__restrict__ E, const int *__restrict__ pre,
                                                             int n, double *__restrict__ part)
{
    __shared__ double s[BTX * PTY];
    const int tid = threadIdx.y * BTX + threadIdx.x;
    tile_loop(pre[n], [&](int t) {
        const int e = find_entry(pre, n, t);
        const GlideLevel &A = E[e].L;
        const int ny = A.g.ny, nx = A.g.nx;
        const int lt = t - pre[e], tx = cdiv(nx + 1, BTX);
        const int i = (lt / tx) * PTY + threadIdx.y, j = (lt % tx) * BTX + threadIdx.x;
        const size_t nu = n_u(ny, nx), nv = n_v(ny, nx);
        double acc = 0.0;
        if (i <= ny && j <= nx) {
            if (i < ny) {
                double x = A.r[(size_t)i * (nx + 1) + j];
                acc += x * x;
            }
            if (j < nx) {
                double x = A.r[nu + (size_t)i * nx + j];
                acc += x * x;
            }
            if (i < ny && j < nx) {
                size_t kc = (size_t)i * nx + j;
                double a = A.r[nu + nv + kc], b = A.U[nu + nv + kc] - A.gamma[kc];
                double x = a + b - sqrt(a * a + b * b);
                acc += x * x;
            }
        }
        __syncthreads();
        s[tid] = acc;
        __syncthreads();
        for (int w = BTX * PTY / 2; w > 0; w >>= 1) {
            if (tid < w) s[tid] += s[tid + w];
            __syncthreads();
        }
        if (tid == 0) part[t] = s[0];
    });
}
__global__ __launch_bounds__(256) void kb_norm_final(const int *__restrict__ pre, const double *__restrict__ part,
                                                     double *__restrict__ out)
{
    __shared__ double s[256];
    const int e = blockIdx.x;
    double acc = 0.0;
    for (int k = pre[e] + threadIdx.x; k < pre[e + 1]; k += 256) acc += part[k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[e] = sqrt(s[0]);
}

// ---------------------------------------------------------------------------
// host: the plan
// ---------------------------------------------------------------------------
enum BatchKernel { BK_VANKA = 0, BK_VANKA_LOCAL = 1, BK_RESIDUAL = 2, BK_RESIDUAL_CHILD = 3, BK_COUNT = 4 };

struct BPhase {
    int n = 0, off = 0, toff = 0;     // entries, first entry, first table int
    int total[NTABLE] = {};           // tiles per table
    bool any_finest = false;
    int shape[BK_COUNT] = {SHAPE_PTY, SHAPE_PTY, SHAPE_PTY, SHAPE_PTY};
    int occ[BK_COUNT] = {0, 0, 0, 0};  // workgroups per CU (0: kernel default)
};

struct GlideBatch {
    int n_members = 0, n_phases = 0, n_cu = 256;
    std::vector<int> n_levels;
    std::vector<BEntry> h_entries;    // mirror of d_entries
    std::vector<BPhase> ph;           // n_phases + 1 (last: finest levels)
    BEntry *d_entries = nullptr;
    int *d_tables = nullptr;
    unsigned char *d_on = nullptr;
    double *d_part = nullptr, *d_norm = nullptr;
    double *h_norm = nullptr;         // pinned
    unsigned char *h_on = nullptr;    // pinned
    BEntry *h_stage = nullptr;        // pinned upload staging
    int *d_any = nullptr;             // any-block flags of two consecutive local passes
    int *h_any = nullptr;             // pinned copies of d_any
    hipEvent_t any_ev[2] = {};        // recorded after each copy
    std::vector<int> h_tables;
};

namespace {

int bcheck(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail((int)e, std::string(what) + ": " + hipGetErrorString(e));
    return GLIDE_OK;
}

#define BTRY(x)                          \
    do {                                 \
        int _rc = (x);                   \
        if (_rc != GLIDE_OK) return _rc; \
    } while (0)

int tiles_of(const GlideLevel &A, int kind, int shape)
{
    const int ny = A.g.ny, nx = A.g.nx, ty = ty_of(shape);
    if (kind == TK_POINT) return cdiv(nx + 1, BTX) * cdiv(ny + 1, ty);
    if (kind == TK_CELL) return cdiv(nx, BTX) * cdiv(ny, ty);
    return (nx / RB + 1) * (ny / RB + 1);
}

// workgroups of a persistent launch: min(tiles, CUs x occupancy), >= 8 and a multiple of 8
unsigned grid_of(const GlideBatch *b, int tiles, int per_cu)
{
    long g = std::min<long>((long)tiles, (long)b->n_cu * std::max(per_cu, 1));
    g = std::max<long>(g, NXCD);
    return (unsigned)((g + NXCD - 1) / NXCD * NXCD);
}

const int *table(const GlideBatch *b, int p, int tid) { return b->d_tables + b->ph[p].toff + tid * (b->ph[p].n + 1); }
const BEntry *entries(const GlideBatch *b, int p) { return b->d_entries + b->ph[p].off; }

// default workgroups per CU from the register budget of each kernel
int default_occ(int kernel, int ty)
{
    int waves_per_cu = (kernel == BK_VANKA || kernel == BK_VANKA_LOCAL) ? 16 : 32;
    return std::max(1, waves_per_cu / ty);
}

template <int TY>
void launch_residual_ty(const GlideBatch *b, int p, bool child, unsigned grid, hipStream_t st)
{
    const int tid = table_id(child ? 1 : 0, TK_CELL, __builtin_ctz(TY));
    if (child)
        hipLaunchKernelGGL((kb_residual<TY, true>), dim3(grid), dim3(BTX, TY), 0, st, entries(b, p), table(b, p, tid),
                           b->ph[p].n, b->d_on);
    else
        hipLaunchKernelGGL((kb_residual<TY, false>), dim3(grid), dim3(BTX, TY), 0, st, entries(b, p), table(b, p, tid),
                           b->ph[p].n, b->d_on);
}

int b_residual(const GlideBatch *b, int p, bool child, int shape, int occ, hipStream_t st)
{
    const BPhase &P = b->ph[p];
    const int kern = child ? BK_RESIDUAL_CHILD : BK_RESIDUAL;
    const int total = P.total[table_id(child ? 1 : 0, TK_CELL, shape)];
    if (total == 0) return GLIDE_OK;
    const unsigned grid = grid_of(b, total, occ > 0 ? occ : default_occ(kern, ty_of(shape)));
    switch (shape) {
    case 1: launch_residual_ty<2>(b, p, child, grid, st); break;
    case 2: launch_residual_ty<4>(b, p, child, grid, st); break;
    case 3: launch_residual_ty<8>(b, p, child, grid, st); break;
    case 4: launch_residual_ty<16>(b, p, child, grid, st); break;
    default: return bfail(GLIDE_E_ARG, "residual tile height must be 2, 4, 8 or 16");
    }
    return bcheck("batched residual");
}

template <int TY, bool LOCAL>
void launch_vanka_ty(const GlideBatch *b, int p, unsigned grid, int n_newton, hipStream_t st)
{
    const int tid = table_id(0, TK_CELL, __builtin_ctz(TY));
    hipLaunchKernelGGL((kb_vanka<TY, LOCAL>), dim3(grid), dim3(BTX, TY), 0, st, entries(b, p), table(b, p, tid),
                       b->ph[p].n, b->d_on, n_newton);
}

// Vanka smoother only (writes the per-cell updates into scratch; U unchanged)
int b_vanka(const GlideBatch *b, int p, bool local, int shape, int occ, int n_newton, hipStream_t st)
{
    const BPhase &P = b->ph[p];
    const int kern = local ? BK_VANKA_LOCAL : BK_VANKA;
    const int total = P.total[table_id(0, TK_CELL, shape)];
    if (total == 0) return GLIDE_OK;
    const unsigned grid = grid_of(b, total, occ > 0 ? occ : default_occ(kern, ty_of(shape)));
    switch (shape * 2 + (local ? 1 : 0)) {
    case 0: launch_vanka_ty<1, false>(b, p, grid, n_newton, st); break;
    case 1: launch_vanka_ty<1, true>(b, p, grid, n_newton, st); break;
    case 2: launch_vanka_ty<2, false>(b, p, grid, n_newton, st); break;
    case 3: launch_vanka_ty<2, true>(b, p, grid, n_newton, st); break;
    case 4: launch_vanka_ty<4, false>(b, p, grid, n_newton, st); break;
    case 5: launch_vanka_ty<4, true>(b, p, grid, n_newton, st); break;
    case 6: launch_vanka_ty<8, false>(b, p, grid, n_newton, st); break;
    case 7: launch_vanka_ty<8, true>(b, p, grid, n_newton, st); break;
    default: return bfail(GLIDE_E_ARG, "Vanka tile height must be 1, 2, 4 or 8");
    }
    return bcheck("batched vanka");
}

int b_sweeps(const GlideBatch *b, int p, bool local, int n_sweeps, int n_newton, hipStream_t st)
{
    const BPhase &P = b->ph[p];
    const int kern = local ? BK_VANKA_LOCAL : BK_VANKA;
    const int apt = table_id(0, TK_POINT, SHAPE_PTY);
    const unsigned ga = grid_of(b, P.total[apt], 8);
    for (int s = 0; s < n_sweeps; s++) {
        BTRY(b_vanka(b, p, local, P.shape[kern], P.occ[kern], n_newton, st));
        if (local)
            hipLaunchKernelGGL(kb_apply<true>, dim3(ga), dim3(BTX, PTY), 0, st, entries(b, p), table(b, p, apt), P.n,
                               b->d_on);
        else
            hipLaunchKernelGGL(kb_apply<false>, dim3(ga), dim3(BTX, PTY), 0, st, entries(b, p), table(b, p, apt), P.n,
                               b->d_on);
    }
    return bcheck("batched apply");
}

template <int OP>
int b_point(const GlideBatch *b, int p, double thklim, hipStream_t st)
{
    constexpr bool child = OP >= OP_RESTRICT_CHI;
    const int tid = table_id(child ? 1 : 0, TK_POINT, SHAPE_PTY);
    const int total = b->ph[p].total[tid];
    if (total == 0) return GLIDE_OK;
    hipLaunchKernelGGL(kb_point<OP>, dim3(grid_of(b, total, 8)), dim3(BTX, PTY), 0, st, entries(b, p), table(b, p, tid),
                       b->ph[p].n, b->d_on, thklim);
    return bcheck("batched element-wise pass");
}

int b_activation(const GlideBatch *b, int p, hipStream_t st, int *any_active = nullptr)
{
    const int tid = table_id(0, TK_BLOCK, 0);
    hipLaunchKernelGGL(kb_activation, dim3(grid_of(b, b->ph[p].total[tid], 8)), dim3(16, 16), 0, st, entries(b, p),
                       table(b, p, tid), b->ph[p].n, b->d_on, any_active);
    return bcheck("batched activation");
}

// Activation of local pass `pass` with its any-block flag copied to the host
// (slot pass & 1); sets *stop when the host has seen an idle pass: pass 0's
// flag is waited for, from pass 2 on the previous pass's (see the
// single-domain V-cycle for the argument that every later pass is idle).
int b_checked_activation(const GlideBatch *b, int p, int pass, hipStream_t st, bool *stop)
{
    const int k = pass & 1;
    hipLaunchKernelGGL(kb_set_int, dim3(1), dim3(1), 0, st, b->d_any + k, 0);
    BTRY(b_activation(b, p, st, b->d_any + k));
    hipError_t e = hipMemcpyAsync(b->h_any + k, b->d_any + k, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(b->any_ev[k], st);
    int look = pass == 0 ? 0 : (pass >= 2 ? (k ^ 1) : -1);
    if (e == hipSuccess && look >= 0) e = hipEventSynchronize(b->any_ev[look]);
    if (e != hipSuccess) return bfail((int)e, std::string("local pass check: ") + hipGetErrorString(e));
    *stop = look >= 0 && b->h_any[look] == 0;
    return GLIDE_OK;
}

// glide/solver.py:71-157 over every entry of phase d at once
int b_fascd(const GlideBatch *b, int d, double thklim, const GlideSchedule &sc, hipStream_t st)
{
    const BPhase &P = b->ph[d];
    if (P.any_finest) BTRY(b_point<OP_FAS_BEGIN>(b, d, thklim, st));
    if (d == b->n_phases - 1) {
        BTRY(b_point<OP_GAMMA>(b, d, thklim, st));
        BTRY(b_sweeps(b, d, false, sc.coarse, sc.n_newton, st));
        return b_point<OP_FILL_GAMMA>(b, d, thklim, st);
    }
    BTRY(b_point<OP_RESTRICT_CHI>(b, d, thklim, st));
    BTRY(b_point<OP_GAMMA_PRE>(b, d, thklim, st));
    BTRY(b_sweeps(b, d, false, sc.pre, sc.n_newton, st));
    BTRY(b_point<OP_FILL_GAMMA>(b, d, thklim, st));
    BTRY(b_point<OP_RESTRICT_U>(b, d, thklim, st));
    BTRY(b_residual(b, d, false, P.shape[BK_RESIDUAL], P.occ[BK_RESIDUAL], st));
    BTRY(b_point<OP_RESTRICT_R>(b, d, sc.boundary_weight, st));
    BTRY(b_residual(b, d, true, P.shape[BK_RESIDUAL_CHILD], P.occ[BK_RESIDUAL_CHILD], st));
    BTRY(b_fascd(b, d + 1, thklim, sc, st));
    BTRY(b_point<OP_FAS_CORRECT>(b, d, thklim, st));
    BTRY(b_point<OP_GAMMA>(b, d, thklim, st));
    BTRY(b_sweeps(b, d, false, sc.post, sc.n_newton, st));
    const bool skip = sc.local_skip != 0;
    for (int pass = 0; pass < sc.local_passes; pass++) {
        BTRY(b_residual(b, d, false, P.shape[BK_RESIDUAL], P.occ[BK_RESIDUAL], st));
        if (skip) {
            bool stop = false;
            BTRY(b_checked_activation(b, d, pass, st, &stop));
            if (stop) break;
        } else {
            BTRY(b_activation(b, d, st));
        }
        BTRY(b_sweeps(b, d, true, sc.local_sweeps, sc.n_newton, st));
    }
    return b_point<OP_FILL_GAMMA>(b, d, thklim, st);
}

int upload_entries(GlideBatch *b, hipStream_t st)
{
    const size_t bytes = b->h_entries.size() * sizeof(BEntry);
    std::memcpy(b->h_stage, b->h_entries.data(), bytes);
    hipError_t e = hipMemcpyAsync(b->d_entries, b->h_stage, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bfail((int)e, std::string("entry upload: ") + hipGetErrorString(e));
    return GLIDE_OK;
}

bool level_ok(const GlideLevel &A)
{
    return A.U && A.f && A.r && A.w && A.bed && A.B && A.beta && A.gamma && A.chi && A.mask && A.scratch &&
           A.active && A.H_prev && A.smb && A.g.ny >= 1 && A.g.nx >= 1 && A.g.dx > 0.0 && A.g.dt > 0.0;
}

void destroy(GlideBatch *b)
{
    if (!b) return;
    (void)hipFree(b->d_entries);
    (void)hipFree(b->d_tables);
    (void)hipFree(b->d_on);
    (void)hipFree(b->d_part);
    (void)hipFree(b->d_norm);
    (void)hipHostFree(b->h_norm);
    (void)hipHostFree(b->h_on);
    (void)hipHostFree(b->h_stage);
    (void)hipFree(b->d_any);
    (void)hipHostFree(b->h_any);
    for (int k = 0; k < 2; k++)
        if (b->any_ev[k]) (void)hipEventDestroy(b->any_ev[k]);
    delete b;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int glide_batch_create(const GlideLevel *const *members, const int *n_levels, int n_members, GlideBatch **out)
{
    if (!members || !n_levels || !out) return bfail(GLIDE_E_NULL, "members, n_levels and out are required");
    *out = nullptr;
    if (n_members < 1 || n_members > 4096) return bfail(GLIDE_E_SIZE, "member count out of range");
    int Lmax = 0;
    for (int m = 0; m < n_members; m++) {
        if (!members[m]) return bfail(GLIDE_E_NULL, "member " + std::to_string(m) + " has no levels");
        if (n_levels[m] < 1 || n_levels[m] > 16) return bfail(GLIDE_E_SIZE, "level count out of range");
        for (int k = 0; k < n_levels[m]; k++) {
            const GlideLevel &A = members[m][k];
            if (!level_ok(A))
                return bfail(GLIDE_E_NULL, "member " + std::to_string(m) + " level " + std::to_string(k) +
                                               " has a NULL buffer or a bad geometry");
            if (k > 0 && (A.g.ny * 2 != members[m][k - 1].g.ny || A.g.nx * 2 != members[m][k - 1].g.nx))
                return bfail(GLIDE_E_SIZE, "member " + std::to_string(m) + " level " + std::to_string(k) +
                                               " is not a 2x coarsening of its parent");
        }
        Lmax = std::max(Lmax, n_levels[m]);
    }
    GlideBatch *b = new GlideBatch();
    b->n_members = n_members;
    b->n_phases = Lmax;
    b->n_levels.assign(n_levels, n_levels + n_members);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        b->n_cu = prop.multiProcessorCount;
    // entries per phase: phases 0..Lmax-1 (aligned at the coarsest level), then the finest levels
    b->ph.resize(Lmax + 1);
    for (int d = 0; d <= Lmax; d++) {
        BPhase &P = b->ph[d];
        P.off = (int)b->h_entries.size();
        for (int m = 0; m < n_members; m++) {
            const int k = d < Lmax ? d - (Lmax - n_levels[m]) : 0;
            if (k < 0 || k >= n_levels[m]) continue;
            BEntry e;
            std::memset(&e, 0, sizeof(e));
            e.L = members[m][k];
            e.has_child = k + 1 < n_levels[m];
            if (e.has_child) e.C = members[m][k + 1];
            e.member = m;
            e.level = k;
            e.finest = k == 0;
            P.any_finest = P.any_finest || e.finest;
            b->h_entries.push_back(e);
            P.n++;
        }
    }
    // prefix tables
    size_t max_tiles = 1;
    for (int d = 0; d <= Lmax; d++) {
        BPhase &P = b->ph[d];
        P.toff = (int)b->h_tables.size();
        for (int child = 0; child < 2; child++)
            for (int kind = 0; kind < 3; kind++)
                for (int s = 0; s < NSHAPE; s++) {
                    int acc = 0;
                    b->h_tables.push_back(0);
                    for (int q = 0; q < P.n; q++) {
                        const BEntry &e = b->h_entries[P.off + q];
                        if (!child || e.has_child) acc += tiles_of(child ? e.C : e.L, kind, s);
                        b->h_tables.push_back(acc);
                    }
                    P.total[table_id(child, kind, s)] = acc;
                    max_tiles = std::max(max_tiles, (size_t)acc);
                }
    }
    hipError_t e = hipSuccess;
    const size_t eb = b->h_entries.size() * sizeof(BEntry), tb = b->h_tables.size() * sizeof(int);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_entries, eb);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_tables, tb);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_on, n_members);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_part, max_tiles * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_norm, n_members * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_norm, n_members * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_on, n_members, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_stage, std::max(eb, tb), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_any, 2 * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_any, 2 * sizeof(int), hipHostMallocDefault);
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipEventCreateWithFlags(&b->any_ev[k], hipEventDisableTiming);
    if (e == hipSuccess) {
        std::memcpy(b->h_stage, b->h_tables.data(), tb);
        e = hipMemcpy(b->d_tables, b->h_stage, tb, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        std::memset(b->h_on, 1, n_members);
        e = hipMemcpy(b->d_on, b->h_on, n_members, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        std::memcpy(b->h_stage, b->h_entries.data(), eb);
        e = hipMemcpy(b->d_entries, b->h_stage, eb, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        destroy(b);
        return bfail((int)e, std::string("batch allocation: ") + hipGetErrorString(e));
    }
    *out = b;
    return GLIDE_OK;
}

int glide_batch_destroy(GlideBatch *b)
{
    destroy(b);
    return GLIDE_OK;
}

int glide_batch_info(const GlideBatch *b, int *n_phases, int *entries_per_phase, int *cells_per_phase)
{
    if (!b || !n_phases) return bfail(GLIDE_E_NULL, "batch and n_phases are required");
    *n_phases = b->n_phases;
    for (int d = 0; d < b->n_phases; d++) {
        if (entries_per_phase) entries_per_phase[d] = b->ph[d].n;
        if (cells_per_phase) {
            long c = 0;
            for (int q = 0; q < b->ph[d].n; q++) {
                const GlideLevel &A = b->h_entries[b->ph[d].off + q].L;
                c += (long)A.g.ny * A.g.nx;
            }
            cells_per_phase[d] = (int)std::min<long>(c, 0x7fffffff);
        }
    }
    return GLIDE_OK;
}

int glide_batch_set_enabled(GlideBatch *b, const unsigned char *on, void *stream)
{
    if (!b || !on) return bfail(GLIDE_E_NULL, "batch and flags are required");
    hipStream_t st = (hipStream_t)stream;
    std::memcpy(b->h_on, on, b->n_members);
    hipError_t e = hipMemcpyAsync(b->d_on, b->h_on, b->n_members, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bfail((int)e, std::string("set enabled: ") + hipGetErrorString(e));
    return GLIDE_OK;
}

int glide_batch_forward_setup(GlideBatch *b, double dt, void *stream)
{
    if (!b) return bfail(GLIDE_E_NULL, "batch is NULL");
    if (!(dt > 0.0) || !std::isfinite(dt)) return bfail(GLIDE_E_ARG, "dt must be positive and finite");
    hipStream_t st = (hipStream_t)stream;
    for (BEntry &e : b->h_entries) {
        e.L.g.dt = dt;
        if (e.has_child) e.C.g.dt = dt;
    }
    BTRY(upload_entries(b, st));
    return b_point<OP_FORCING>(b, b->n_phases, 0.0, st);
}

int glide_batch_vcycle(GlideBatch *b, double thklim, const GlideSchedule *sched, void *stream)
{
    if (!b) return bfail(GLIDE_E_NULL, "batch is NULL");
    GlideSchedule sc = sched ? *sched : GlideSchedule{10, 10, 200, 10, 10, 10, 10, 100, 1.0, 0};
    if (sc.pre < 0 || sc.post < 0 || sc.coarse < 0 || sc.local_passes < 0 || sc.local_sweeps < 0 || sc.n_newton < 0)
        return bfail(GLIDE_E_ARG, "schedule: negative count");
    if (!std::isfinite(sc.boundary_weight)) return bfail(GLIDE_E_ARG, "schedule: boundary_weight is not finite");
    return b_fascd(b, 0, thklim, sc, (hipStream_t)stream);
}

int glide_batch_residual(GlideBatch *b, int phase, void *stream)
{
    if (!b) return bfail(GLIDE_E_NULL, "batch is NULL");
    if (phase < 0 || phase > b->n_phases) return bfail(GLIDE_E_ARG, "phase out of range");
    const BPhase &P = b->ph[phase];
    return b_residual(b, phase, false, P.shape[BK_RESIDUAL], P.occ[BK_RESIDUAL], (hipStream_t)stream);
}

int glide_batch_residual_norms(GlideBatch *b, double *out_host, void *stream)
{
    if (!b || !out_host) return bfail(GLIDE_E_NULL, "batch and output are required");
    hipStream_t st = (hipStream_t)stream;
    const int F = b->n_phases;
    const BPhase &P = b->ph[F];
    BTRY(b_residual(b, F, false, P.shape[BK_RESIDUAL], P.occ[BK_RESIDUAL], st));
    const int tid = table_id(0, TK_POINT, SHAPE_PTY);
    hipLaunchKernelGGL(kb_norm_partial, dim3(grid_of(b, P.total[tid], 8)), dim3(BTX, PTY), 0, st, entries(b, F),
                       table(b, F, tid), P.n, b->d_part);
    hipLaunchKernelGGL(kb_norm_final, dim3(P.n), dim3(256), 0, st, table(b, F, tid), b->d_part, b->d_norm);
    BTRY(bcheck("batched norm"));
    hipError_t e = hipMemcpyAsync(b->h_norm, b->d_norm, P.n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bfail((int)e, std::string("norm readback: ") + hipGetErrorString(e));
    for (int q = 0; q < P.n; q++) out_host[b->h_entries[P.off + q].member] = b->h_norm[q];
    return GLIDE_OK;
}

int glide_batch_set_tiles(GlideBatch *b, int kernel, int phase, int ty, int wg_per_cu)
{
    if (!b) return bfail(GLIDE_E_NULL, "batch is NULL");
    if (kernel < 0 || kernel >= BK_COUNT) return bfail(GLIDE_E_ARG, "unknown kernel id");
    if (phase < 0 || phase > b->n_phases) return bfail(GLIDE_E_ARG, "phase out of range");
    const bool vanka = kernel == BK_VANKA || kernel == BK_VANKA_LOCAL;
    int s = -1;
    for (int q = 0; q < NSHAPE; q++)
        if (ty_of(q) == ty) s = q;
    if (s < 0 || (vanka && ty > 8) || (!vanka && ty < 2))
        return bfail(GLIDE_E_ARG, vanka ? "Vanka tile height must be 1, 2, 4 or 8"
                                        : "residual tile height must be 2, 4, 8 or 16");
    if (wg_per_cu < 0 || wg_per_cu > 64) return bfail(GLIDE_E_ARG, "workgroups per CU out of range");
    b->ph[phase].shape[kernel] = s;
    b->ph[phase].occ[kernel] = wg_per_cu;
    return GLIDE_OK;
}

int glide_batch_time_kernel(GlideBatch *b, int kernel, int phase, int ty, int wg_per_cu, int reps, double *ms,
                            void *stream)
{
    if (!b || !ms) return bfail(GLIDE_E_NULL, "batch and ms are required");
    if (reps < 1 || reps > 1000) return bfail(GLIDE_E_ARG, "reps out of range");
    if (kernel < 0 || kernel >= BK_COUNT) return bfail(GLIDE_E_ARG, "unknown kernel id");
    if (phase < 0 || phase > b->n_phases) return bfail(GLIDE_E_ARG, "phase out of range");
    if (kernel == BK_RESIDUAL_CHILD && (phase >= b->n_phases - 1))
        return bfail(GLIDE_E_ARG, "the coarsest phase has no child level");
    BPhase &P = b->ph[phase];
    const int keep_s = P.shape[kernel], keep_o = P.occ[kernel];
    BTRY(glide_batch_set_tiles(b, kernel, phase, ty, wg_per_cu));
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess) return bfail((int)he, "event creation failed");
    const int s = P.shape[kernel], o = P.occ[kernel];
    int rc = GLIDE_OK;
    auto run = [&]() {
        switch (kernel) {
        case BK_VANKA: return b_vanka(b, phase, false, s, o, 10, st);
        case BK_VANKA_LOCAL: return b_vanka(b, phase, true, s, o, 10, st);
        case BK_RESIDUAL: return b_residual(b, phase, false, s, o, st);
        default: return b_residual(b, phase, true, s, o, st);
        }
    };
    rc = run();   // warm-up
    if (rc == GLIDE_OK) (void)hipEventRecord(e0, st);
    for (int k = 0; k < reps && rc == GLIDE_OK; k++) rc = run();
    if (rc == GLIDE_OK) {
        (void)hipEventRecord(e1, st);
        he = hipEventSynchronize(e1);
        float t = 0.0f;
        if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
        if (he != hipSuccess) rc = bfail((int)he, "event timing failed");
        *ms = (double)t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P.shape[kernel] = keep_s;
    P.occ[kernel] = keep_o;
    return rc;
}

int glide_batch_time_restriction(GlideBatch *b, int phase, int reps, double *ms, double *bytes, void *stream)
{
    if (!b || !ms || !bytes) return bfail(GLIDE_E_NULL, "batch, ms and bytes are required");
    if (reps < 1 || reps > 1000) return bfail(GLIDE_E_ARG, "reps out of range");
    if (phase < 0 || phase >= b->n_phases) return bfail(GLIDE_E_ARG, "phase out of range");
    const BPhase &P = b->ph[phase];
    double nb = 0.0;
    for (int e = P.off; e < P.off + P.n; e++) {
        const BEntry &E = b->h_entries[e];
        if (!E.has_child) continue;
        const size_t nf = n_u(E.L.g.ny, E.L.g.nx) + n_v(E.L.g.ny, E.L.g.nx) + n_c(E.L.g.ny, E.L.g.nx);
        const size_t nc = n_u(E.C.g.ny, E.C.g.nx) + n_v(E.C.g.ny, E.C.g.nx) + n_c(E.C.g.ny, E.C.g.nx);
        nb += 8.0 * (double)(nf + nc);
    }
    *bytes = nb;
    *ms = 0.0;
    if (nb == 0.0) return GLIDE_OK;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess) return bfail((int)he, "event creation failed");
    int rc = b_point<OP_RESTRICT_R>(b, phase, 1.0, st);   // warm-up
    if (rc == GLIDE_OK) (void)hipEventRecord(e0, st);
    for (int k = 0; k < reps && rc == GLIDE_OK; k++) rc = b_point<OP_RESTRICT_R>(b, phase, 1.0, st);
    if (rc == GLIDE_OK) {
        (void)hipEventRecord(e1, st);
        he = hipEventSynchronize(e1);
        float t = 0.0f;
        if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
        if (he != hipSuccess) rc = bfail((int)he, "event timing failed");
        *ms = (double)t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // extern "C"
