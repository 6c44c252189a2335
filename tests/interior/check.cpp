// Host check of the interior specialisation of the Vanka assembly: for every
// cell (i, j) in [1, ny-2] x [1, nx-2] of several grids, vanka_assemble with
// the cell's own indices and with the constant indices (1, 1) of a 3 x 3 grid
// gives the same bits, for random neighbourhoods, states and both grounding
// modes; a boundary cell must differ (so the check can see a difference).
// Prints "ok <cells checked>" or the first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#define __host__
#define __device__
#define __forceinline__ inline
#define __restrict__
#include "glide_hip.h"
#include "vanka_device.h"

using namespace glide;

static void fill(CellNbhd &N, std::mt19937_64 &rng)
{
    std::uniform_real_distribution<double> U(0.5, 2.0), S(-1.0, 1.0);
    double *d = reinterpret_cast<double *>(&N);
    for (size_t k = 0; k < sizeof(CellNbhd) / sizeof(double); k++) d[k] = U(rng);
    N.bed_c = 300.0 * S(rng); N.bed_l = 300.0 * S(rng); N.bed_r = 300.0 * S(rng);
    N.bed_t = 300.0 * S(rng); N.bed_b = 300.0 * S(rng);
    N.u_ll = 50.0 * S(rng); N.u_rr = 50.0 * S(rng); N.v_tt = 50.0 * S(rng); N.v_bb = 50.0 * S(rng);
    N.H_l *= 400.0; N.H_r *= 400.0; N.H_t *= 400.0; N.H_b *= 400.0;
}

int main()
{
    std::mt19937_64 rng(1234);
    std::uniform_real_distribution<double> U(0.5, 2.0), S(-1.0, 1.0);
    const int sizes[][2] = {{3, 3}, {4, 7}, {16, 16}, {33, 20}};
    long checked = 0;
    bool boundary_differs = false;
    for (int gl = 0; gl < 2; gl++) {
        Params p{3.0, 1e-5, 1e-3, 1.0, 0.1, gl};
        for (auto &sz : sizes) {
            const int ny = sz[0], nx = sz[1];
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++) {
                    for (int rep = 0; rep < 3; rep++) {
                        CellNbhd N;
                        fill(N, rng);
                        nbhd_grounding(N, p);
                        const double ul = 30 * S(rng), ur = 30 * S(rng), vt = 30 * S(rng), vb = 30 * S(rng);
                        const double Hc = 400 * U(rng), dxi = 1.0 / 1500.0, dt = 1.0, th = 0.1;
                        double J1[25], r1[5], J2[25], r2[5];
                        vanka_assemble<0>(N, p, i, j, ny, nx, dxi, dt, ul, ur, vt, vb, Hc, th, false, J1, r1);
                        vanka_assemble<0>(N, p, 1, 1, 3, 3, dxi, dt, ul, ur, vt, vb, Hc, th, false, J2, r2);
                        const bool same = !std::memcmp(J1, J2, sizeof J1) && !std::memcmp(r1, r2, sizeof r1);
                        const bool interior = i >= 1 && i <= ny - 2 && j >= 1 && j <= nx - 2;
                        if (interior && !same) {
                            std::printf("mismatch at cell (%d, %d) of %d x %d\n", i, j, ny, nx);
                            return 1;
                        }
                        if (!interior && !same) boundary_differs = true;
                        checked += interior;
                    }
                }
        }
    }
    if (!boundary_differs) {
        std::printf("no boundary cell differed\n");
        return 1;
    }
    std::printf("ok %ld\n", checked);
    return 0;
}
