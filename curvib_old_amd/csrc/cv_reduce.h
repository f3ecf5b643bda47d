// cv_reduce.h -- deterministic block reductions for gfx950 (64-lane waves).
#pragma once
#include <hip/hip_runtime.h>

#define CV_RED_BLOCK 256
#define CV_RED_GRID 1024

__device__ inline double cv_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum over the block; result valid in thread 0.  sh: >= blockDim/64 doubles.
__device__ inline double cv_block_sum(double v, double *sh) {
  v = cv_wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < nw; q++) s += sh[q];
  }
  return s;
}

// max that keeps NaN: a NaN operand wins, so a norm or maximum over a field
// holding a NaN is NaN (fmax would drop it and hide a blown-up step)
__device__ inline double cv_nmax(double a, double b) { return (b > a || b != b) ? b : a; }

__device__ inline double cv_wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = cv_nmax(v, __shfl_xor(v, o, 64));
  return v;
}

// max over the block; result valid in thread 0.  sh: >= blockDim/64 doubles.
__device__ inline double cv_block_max(double v, double *sh) {
  v = cv_wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double m = v;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < nw; q++) m = cv_nmax(m, sh[q]);
  }
  return m;
}

// final pass: out[q] = sum_b part[q*nb + b], one block per q, fixed order
__global__ void cv_k_final_sum(const double *part, int nb, double *out, int nq);
