// grid_data.hip - input preparation on the GPU: observations interpolated to
// the model grid, and ice thickness recomputed from surface and bed.
//
// glide_interp_grid follows the reference's interpolate_to_grid
// (glide/data.py:386-431): a linear RegularGridInterpolator over
// (y_data, x_data), evaluated at every (y_t[i], x_t[j]), fill outside the
// data's extent, NaN where a target coordinate is NaN.  A descending axis is
// read reversed (the interpolator flips it to ascending).  Each target point
// locates its interval on both axes by binary search over the coordinates
// (L2-resident), then combines the four corners in the interpolator's term
// order: v00 (1-y0)(1-y1) + v01 (1-y0) y1 + v10 y0 (1-y1) + v11 y0 y1, summed
// left to right, with y = (x - g[i]) / (g[i+1] - g[i]).  One thread per
// target point, consecutive threads along x_t.
//
// glide_bedmachine_thickness follows glide/data.py:257-258:
// base = max(bed, surface * (1 - 1 / (1 - 0.917))), H = surface - base + thklim
// (NaN propagates through the max as in numpy.maximum).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "glide_hip.h"
#include "glide_io.h"

namespace glide {
int set_error(int code, const char *msg);
}

namespace {

int dfail(int code, const std::string &msg) { return glide::set_error(code, msg.c_str()); }

struct Axis {
    const double *g;
    int n;
    int desc;   // stored descending: ascending index k is stored at n - 1 - k
    __device__ __forceinline__ double at(int k) const { return g[desc ? n - 1 - k : k]; }
    __device__ __forceinline__ int row(int k) const { return desc ? n - 1 - k : k; }
};

// interval i in [0, n-2] with at(i) <= x < at(i+1) (clipped at the ends) and
// the normalised distance of x in it
__device__ __forceinline__ void locate(const Axis &a, double x, int &i, double &w)
{
    int lo = 0, hi = a.n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.at(mid) <= x)
            lo = mid;
        else
            hi = mid;
    }
    i = lo;
    const double g0 = a.at(lo), g1 = a.at(lo + 1);
    w = (x - g0) / (g1 - g0);
}

__global__ __launch_bounds__(256) void k_interp_grid(const double *__restrict__ data, int nx_d, Axis ay, Axis ax,
                                                     const double *__restrict__ y_t, int ny_t,
                                                     const double *__restrict__ x_t, int nx_t, double fill,
                                                     double *__restrict__ out)
{
    const size_t n = (size_t)ny_t * nx_t;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const int ti = (int)(p / nx_t), tj = (int)(p - (size_t)ti * nx_t);
        const double y = y_t[ti], x = x_t[tj];
        double r;
        if (y != y || x != x) {
            r = __builtin_nan("");
        } else if (y < ay.at(0) || y > ay.at(ay.n - 1) || x < ax.at(0) || x > ax.at(ax.n - 1)) {
            r = fill;
        } else {
            int i0, i1;
            double y0, y1;
            locate(ay, y, i0, y0);
            locate(ax, x, i1, y1);
            const size_t ra = (size_t)ay.row(i0) * nx_d, rb = (size_t)ay.row(i0 + 1) * nx_d;
            const int ca = ax.row(i1), cb = ax.row(i1 + 1);
            const double a0 = 1.0 - y0, a1 = 1.0 - y1;
            const double t00 = data[ra + ca] * a0 * a1;
            const double t01 = data[ra + cb] * a0 * y1;
            const double t10 = data[rb + ca] * y0 * a1;
            const double t11 = data[rb + cb] * y0 * y1;
            r = t00 + t01 + t10 + t11;
        }
        out[p] = r;
    }
}

__global__ __launch_bounds__(256) void k_bedmachine_thickness(const double *__restrict__ bed,
                                                              const double *__restrict__ surface, size_t n,
                                                              double thklim, double *__restrict__ H)
{
    const double c = 1.0 - 1.0 / (1.0 - 0.917);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const double b = bed[p], s = surface[p];
        const double f = s * c;
        const double base = (b != b || f != f) ? __builtin_nan("") : (b >= f ? b : f);
        H[p] = s - base + thklim;
    }
}

unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 8192); }

}  // namespace

extern "C" {

int glide_interp_grid(const double *data, int ny_d, int nx_d, const double *y_data, int y_descending,
                      const double *x_data, int x_descending, const double *y_t, int ny_t, const double *x_t,
                      int nx_t, double fill, double *out, void *stream)
{
    if (!data || !y_data || !x_data || !y_t || !x_t || !out) return dfail(GLIDE_E_NULL, "interp_grid: NULL argument");
    if (ny_d < 2 || nx_d < 2) return dfail(GLIDE_E_SIZE, "interp_grid: the data needs at least 2 x 2 points");
    if (ny_t < 1 || nx_t < 1) return dfail(GLIDE_E_SIZE, "interp_grid: empty target grid");
    const Axis ay{y_data, ny_d, y_descending ? 1 : 0}, ax{x_data, nx_d, x_descending ? 1 : 0};
    const size_t n = (size_t)ny_t * nx_t;
    hipLaunchKernelGGL(k_interp_grid, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, data, nx_d, ay, ax, y_t,
                       ny_t, x_t, nx_t, fill, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return dfail((int)e, std::string("interp_grid: ") + hipGetErrorString(e));
    return GLIDE_OK;
}

int glide_bedmachine_thickness(const double *bed, const double *surface, long n, double thklim, double *thickness,
                               void *stream)
{
    if (!bed || !surface || !thickness) return dfail(GLIDE_E_NULL, "bedmachine_thickness: NULL argument");
    if (n < 0) return dfail(GLIDE_E_SIZE, "bedmachine_thickness: negative size");
    if (n == 0) return GLIDE_OK;
    hipLaunchKernelGGL(k_bedmachine_thickness, dim3(grid_for((size_t)n)), dim3(256), 0, (hipStream_t)stream, bed,
                       surface, (size_t)n, thklim, thickness);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return dfail((int)e, std::string("bedmachine_thickness: ") + hipGetErrorString(e));
    return GLIDE_OK;
}

}  // extern "C"
