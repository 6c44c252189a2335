// vti.hip - appended-data block of a VTK ImageData snapshot, packed on the GPU.
//
// The reference's write_vti (glide/io.py:20-131) converts each field to
// float32 on the host, flips it along y, interleaves vector components and
// writes every array as [UInt32 byte count][float32 values].  Here one launch
// produces that whole block from the FP64 fields in HBM: blockIdx.y selects
// the array, threads stride over its points; each point reads its n_comp
// doubles (coalesced along a row) and writes n_comp contiguous floats.
// FP64 -> FP32 is the hardware conversion (round to nearest even, denormals
// kept), the same rounding as numpy's astype(float32).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "glide_hip.h"
#include "glide_io.h"

namespace glide {
int set_error(int code, const char *msg);
}

namespace {

struct VtiPlan {
    const double *comp[GLIDE_VTI_MAX_ARRAYS][3];
    unsigned long long offset[GLIDE_VTI_MAX_ARRAYS];   // byte offset of the array's header in the block
    int n_comp[GLIDE_VTI_MAX_ARRAYS];
};

__global__ __launch_bounds__(256) void k_vti_pack(VtiPlan P, int ny, int nx, int flip, unsigned char *__restrict__ out)
{
    const int a = blockIdx.y;
    const int nc = P.n_comp[a];
    unsigned char *base = out + P.offset[a];
    const size_t npts = (size_t)ny * nx;
    if (blockIdx.x == 0 && threadIdx.x == 0) *reinterpret_cast<uint32_t *>(base) = (uint32_t)(npts * nc * 4);
    float *dst = reinterpret_cast<float *>(base + 4);
    const double *c0 = P.comp[a][0], *c1 = P.comp[a][1], *c2 = P.comp[a][2];
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < npts; p += (size_t)gridDim.x * blockDim.x) {
        const size_t r = p / nx, c = p - r * nx;
        const size_t src = (flip ? (size_t)(ny - 1) - r : r) * nx + c;
        float *d = dst + p * nc;
        d[0] = (float)c0[src];
        if (nc > 1) d[1] = (float)c1[src];
        if (nc > 2) d[2] = (float)c2[src];
    }
}

int vfail(int code, const std::string &msg) { return glide::set_error(code, msg.c_str()); }

// Offsets of every array's header; returns the block size, 0 if invalid.
size_t layout(const GlideVtiArray *arrays, int n, int ny, int nx, unsigned long long *offset, std::string *why,
              int *code)
{
    *code = GLIDE_E_SIZE;
    if (n < 1 || n > GLIDE_VTI_MAX_ARRAYS) return *why = "vti: 1 to 16 arrays", 0;
    if (ny < 1 || nx < 1) return *why = "vti: ny and nx must be >= 1", 0;
    *code = GLIDE_E_NULL;
    if (!arrays) return *why = "vti: NULL array table", 0;
    const size_t npts = (size_t)ny * nx;
    size_t off = 0;
    for (int a = 0; a < n; a++) {
        const int nc = arrays[a].n_comp;
        *code = GLIDE_E_SIZE;
        if (nc < 1 || nc > 3) return *why = "vti: n_comp must be 1, 2 or 3", 0;
        const size_t bytes = npts * nc * 4;
        if (bytes > 0xFFFFFFFFull) return *why = "vti: array exceeds the UInt32 size header", 0;
        *code = GLIDE_E_NULL;
        for (int q = 0; q < nc; q++)
            if (!arrays[a].comp[q]) return *why = "vti: NULL component", 0;
        if (offset) offset[a] = off;
        off += 4 + bytes;
    }
    return off;
}

}  // namespace

extern "C" {

size_t glide_vti_block_bytes(const GlideVtiArray *arrays, int n_arrays, int ny, int nx)
{
    std::string why;
    int code;
    return layout(arrays, n_arrays, ny, nx, nullptr, &why, &code);
}

int glide_vti_pack(const GlideVtiArray *arrays, int n_arrays, int ny, int nx, int flip_y, unsigned char *out,
                   size_t out_bytes, void *stream)
{
    VtiPlan P = {};
    std::string why;
    int code;
    const size_t need = layout(arrays, n_arrays, ny, nx, P.offset, &why, &code);
    if (!need) return vfail(code, why);
    if (!out) return vfail(GLIDE_E_NULL, "vti: NULL output");
    if (out_bytes < need) return vfail(GLIDE_E_SIZE, "vti: output smaller than the block");
    for (int a = 0; a < n_arrays; a++) {
        P.n_comp[a] = arrays[a].n_comp;
        for (int q = 0; q < 3; q++) P.comp[a][q] = q < arrays[a].n_comp ? arrays[a].comp[q] : nullptr;
    }
    const size_t npts = (size_t)ny * nx;
    const unsigned gx = (unsigned)std::min<size_t>((npts + 255) / 256, 2048);
    hipLaunchKernelGGL(k_vti_pack, dim3(gx, n_arrays), dim3(256), 0, (hipStream_t)stream, P, ny, nx, flip_y ? 1 : 0,
                       out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return vfail((int)e, std::string("vti_pack: ") + hipGetErrorString(e));
    return GLIDE_OK;
}

}  // extern "C"
