// fcmd_common.hpp -- small device helpers shared by the kernels.
#pragma once

#include "fcmd_dev.hpp"

namespace fcmd {

// Fortran INT() of a double, back in double: truncation toward zero.  One
// v_trunc_f64 instead of a convert to int and back; the two agree for
// |v| < 2^31, which every argument here (ratios of lengths to the box side or
// to sigma) satisfies by far.
__device__ __forceinline__ double fint(double v) { return trunc(v); }

// Mixed-path position record of one atom: its offset from the centre of the
// cell it was binned into at the last list build, in FP32 (at most half a cell
// side plus the drift since, so rounding costs ~1e-7 sigma), and the cell's
// coordinates packed CELL_BITS each in w.  A pair separation then needs no
// FP64: (cell_i - cell_j) * side + (off_i - off_j) with the cell difference
// wrapped to a neighbour.
__device__ __forceinline__ float4 make_posf(double x, double y, double z, int cx, int cy, int cz,
                                            const TrajScalars &ts)
{
    return make_float4((float)(x - ((double)cx + 0.5) * ts.cside[0]),
                       (float)(y - ((double)cy + 0.5) * ts.cside[1]),
                       (float)(z - ((double)cz + 0.5) * ts.cside[2]),
                       __int_as_float(cx | (cy << CELL_BITS) | (cz << (2 * CELL_BITS))));
}

__device__ __forceinline__ void unpack_cell(float w, int &cx, int &cy, int &cz)
{
    const int c = __float_as_int(w);
    cx = c & CELL_MAX;
    cy = (c >> CELL_BITS) & CELL_MAX;
    cz = c >> (2 * CELL_BITS);
}

// the cell coordinates of a packed cell as floats (one byte conversion each)
__device__ __forceinline__ float cell_xf(float w) { return (float)(__float_as_uint(w) & 0xffu); }
__device__ __forceinline__ float cell_yf(float w) { return (float)((__float_as_uint(w) >> 8) & 0xffu); }
__device__ __forceinline__ float cell_zf(float w) { return (float)((__float_as_uint(w) >> 16) & 0xffu); }

// One coordinate of a pair separation from cell-offset records, on a grid of
// nc >= 3 cells a side (inc = 1/nc): the cell difference dc = ci - cj of two
// cells that are neighbours modulo nc is -1, 0, 1 or +-(nc - 1); dc/nc rounds
// to the periodic image (0 or +-1), so dc - nc*rint(dc/nc) is the wrapped
// difference, exactly (small integers).  Then dc*side + (off_i - off_j).
// FP32 throughout; the two roundings of the fused multiply-adds are the
// mixed path's.
__device__ __forceinline__ float sep_wrap(float ci, float cj, float nc, float inc, float side, float doff)
{
    const float dc = ci - cj;
    const float dw = __builtin_fmaf(-nc, rintf(dc * inc), dc);
    return __builtin_fmaf(dw, side, doff);
}

// the same on a side of fewer than 3 cells: the cell difference, then the
// remaining image by rounding
__device__ __forceinline__ float sep_small(float ci, float cj, float side, float doff, float box, float ibox)
{
    const float s = __builtin_fmaf(ci - cj, side, doff);
    return __builtin_fmaf(-box, rintf(s * ibox), s);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over a BLK-thread block; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds4)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v, double *lds4)
{
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = fmax(fmax(lds4[0], lds4[1]), fmax(lds4[2], lds4[3]));
    __syncthreads();
    return r;
}

// Several block reductions at the end of a kernel with one barrier between
// them all: each value's wavefront result, then (w0 + w1) + (w2 + w3) (or the
// max) by thread 0, as block_sum / block_max do one at a time, so the results
// are the same bits.  Valid in thread 0 only; lds is not reused after.
template <int N>
__device__ __forceinline__ void block_reduce_end(double (&v)[N], const bool (&is_max)[N], double (*lds)[4])
{
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = is_max[k] ? wave_max(v[k]) : wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) lds[k][w] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k)
            v[k] = is_max[k] ? fmax(fmax(lds[k][0], lds[k][1]), fmax(lds[k][2], lds[k][3]))
                             : (lds[k][0] + lds[k][1]) + (lds[k][2] + lds[k][3]);
    }
}

__device__ __forceinline__ int block_sum_i(int v, int *lds4)
{
    v = wave_sum_i(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    int r = 0;
    if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return r;
}

}  // namespace fcmd
