/* glide_io.h - data formats either side of the solver (libglide_hip.so): the
 * device-side packing of VTK ImageData snapshots, and input preparation.
 *
 * The reference writes snapshots with glide/io.py:20-131 (write_vti): every
 * field is converted to float32 on the host, flipped along y, vector
 * components interleaved, and the arrays appended raw after the XML header,
 * each preceded by its byte count as a little-endian UInt32.  Here the
 * appended block is produced by one kernel from the FP64 fields in HBM; the
 * host copies it out once and writes header + block.
 *
 * Conventions as in glide_hip.h: 0 = success, < 0 = GLIDE_E_* argument
 * errors, > 0 = hipError_t; glide_last_error() holds the message.
 */
#ifndef GLIDE_IO_H
#define GLIDE_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLIDE_VTI_MAX_ARRAYS 16

/* One data array of a snapshot: n_comp (1..3) row-major ny x nx FP64
 * fields on the device.  A scalar has n_comp = 1. */
typedef struct {
    const double *comp[3];
    int n_comp;
} GlideVtiArray;

/* Bytes of the appended block for these arrays: sum of 4 + 4 * n_comp * ny * nx.
 * Returns 0 on invalid arguments (see glide_vti_pack). */
size_t glide_vti_block_bytes(const GlideVtiArray *arrays, int n_arrays, int ny, int nx);

/* Fill out[0 .. glide_vti_block_bytes) on the device: for each array in
 * order, its byte count (UInt32) then its values as float32, point (r, c)
 * at index r * nx + c (components interleaved), taken from source row
 * ny - 1 - r when flip_y is set, else row r.  Replaces the host loop of
 * write_vti (glide/io.py:47-62, 95-131).  Asynchronous on `stream`.
 * Errors: NULL pointers, n_arrays outside 1..GLIDE_VTI_MAX_ARRAYS, n_comp
 * outside 1..3, ny or nx < 1, an array larger than 4 GiB - 1 bytes (the
 * UInt32 header), out_bytes smaller than the block. */
int glide_vti_pack(const GlideVtiArray *arrays, int n_arrays, int ny, int nx, int flip_y, unsigned char *out,
                   size_t out_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Input preparation
 * ------------------------------------------------------------------------- */

/* Linear interpolation of data (ny_d x nx_d, row-major, rows along y_data,
 * columns along x_data) to every target point (y_t[i], x_t[j]); out is
 * ny_t x nx_t FP64.  Replaces interpolate_to_grid (glide/data.py:386-431, a
 * linear RegularGridInterpolator with bounds_error = False): targets outside
 * [min, max] of either axis get `fill`, a NaN target coordinate gives NaN.
 * Each axis is strictly ascending (flag 0) or strictly descending (flag 1);
 * the caller checks that.  Needs ny_d, nx_d >= 2. */
int glide_interp_grid(const double *data, int ny_d, int nx_d, const double *y_data, int y_descending,
                      const double *x_data, int x_descending, const double *y_t, int ny_t, const double *x_t,
                      int nx_t, double fill, double *out, void *stream);

/* thickness = surface - max(bed, surface * (1 - 1 / (1 - 0.917))) + thklim
 * over n cells (glide/data.py:257-258, BedMachine's thickness recomputed so
 * that floating ice is in hydrostatic balance). */
int glide_bedmachine_thickness(const double *bed, const double *surface, long n, double thklim, double *thickness,
                               void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GLIDE_IO_H */
