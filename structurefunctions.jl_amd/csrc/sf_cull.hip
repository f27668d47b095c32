// sf_cull.hip -- device preparation of spatial sort and tile culling (gfx950).
//
// For a slice whose largest distance edge is small against its extent, the binary32 records are
// put in Morton (Z-curve) order of their key-space coordinates (flat: x; sphere: the unit
// vector), so that a tile of consecutive records is spatially compact, and the host drops every
// work item whose two tiles' bounding boxes are farther apart than the largest edge.  This file
// computes, on the device and in the caller's stream:
//   cull_coords   key-space coordinates per point and a bounding box per workgroup
//   cull_reduce   the slice's bounding box, its largest coordinate magnitude, a non-finite flag
//   cull_keys     30-bit Morton keys (10 bits per axis of the bounding box) and the identity order
//   (radix sort)  stable sort of the keys, carrying point indices: the record order
//   cull_tiles    the bounding box of every tile of `rows` consecutive records
// The host reads back the slice box and the tile boxes (a few KB) to build its work items.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include "sf_internal.h"
#include "sfmath.h"

namespace sfk {

namespace {

constexpr int kCullBlock = 256;

// key-space coordinates of point i (x is point-major, W values per point)
__device__ void cull_point(const double *x, int64_t i, int W, int D, int geom, double c[3]) {
  if (geom == 0) {
    for (int k = 0; k < 3; ++k) c[k] = k < D ? x[(size_t)W * i + k] : 0.0;
    return;
  }
  double so, co, sa, ca;
  if (geom == 1) {
    sf_sincosd(x[2 * i], &so, &co);
    sf_sincosd(x[2 * i + 1], &sa, &ca);
  } else {
    sf_sincos(x[2 * i], &so, &co);
    sf_sincos(x[2 * i + 1], &sa, &ca);
  }
  c[0] = ca * co;
  c[1] = ca * so;
  c[2] = sa;
}

// per workgroup: lo[3], hi[3], non-finite count, largest |coordinate| (8 doubles)
__global__ void __launch_bounds__(kCullBlock) cull_coords(const double *x, int64_t N, int W, int D,
                                                          int geom, double *kc, double *part) {
  __shared__ double red[8][kCullBlock];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double bad = 0.0, cmax = 0.0;
  const int64_t i = (int64_t)blockIdx.x * kCullBlock + threadIdx.x;
  if (i < N) {
    double c[3];
    cull_point(x, i, W, D, geom, c);
    for (int k = 0; k < 3; ++k) {
      kc[3 * i + k] = c[k];
      if (!(fabs(c[k]) < INFINITY)) bad = 1.0;
      lo[k] = c[k];
      hi[k] = c[k];
      cmax = fmax(cmax, fabs(c[k]));
    }
  }
  const int t = threadIdx.x;
  for (int k = 0; k < 3; ++k) {
    red[k][t] = lo[k];
    red[3 + k][t] = hi[k];
  }
  red[6][t] = bad;
  red[7][t] = cmax;
  __syncthreads();
  for (int w = kCullBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      for (int k = 0; k < 3; ++k) {
        red[k][t] = fmin(red[k][t], red[k][t + w]);
        red[3 + k][t] = fmax(red[3 + k][t], red[3 + k][t + w]);
      }
      red[6][t] += red[6][t + w];
      red[7][t] = fmax(red[7][t], red[7][t + w]);
    }
    __syncthreads();
  }
  if (t < 8) part[(size_t)blockIdx.x * 8 + t] = red[t][0];
}

// one workgroup: combine nblk partial boxes into glob (same 8-double layout)
__global__ void __launch_bounds__(kCullBlock) cull_reduce(const double *part, int nblk, double *glob) {
  __shared__ double red[8][kCullBlock];
  const int t = threadIdx.x;
  double v[8] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0, 0.0};
  for (int b = t; b < nblk; b += kCullBlock) {
    const double *p = part + (size_t)b * 8;
    for (int k = 0; k < 3; ++k) {
      v[k] = fmin(v[k], p[k]);
      v[3 + k] = fmax(v[3 + k], p[3 + k]);
    }
    v[6] += p[6];
    v[7] = fmax(v[7], p[7]);
  }
  for (int k = 0; k < 8; ++k) red[k][t] = v[k];
  __syncthreads();
  for (int w = kCullBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      for (int k = 0; k < 3; ++k) {
        red[k][t] = fmin(red[k][t], red[k][t + w]);
        red[3 + k][t] = fmax(red[3 + k][t], red[3 + k][t + w]);
      }
      red[6][t] += red[6][t + w];
      red[7][t] = fmax(red[7][t], red[7][t + w]);
    }
    __syncthreads();
  }
  if (t < 8) glob[t] = red[t][0];
}

__device__ uint32_t spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void __launch_bounds__(kCullBlock) cull_keys(const double *kc, int64_t N, const double *glob,
                                                        uint32_t *key, int32_t *idx) {
  const int64_t i = (int64_t)blockIdx.x * kCullBlock + threadIdx.x;
  if (i >= N) return;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const double w = glob[3 + k] - glob[k];
    const double t = w > 0 ? (kc[3 * i + k] - glob[k]) / w : 0.0;
    // a non-finite coordinate (the host then refuses culling) still gets an in-range key
    q[k] = (uint32_t)fmin(1023.0, fmax(0.0, floor(t * 1024.0)));
  }
  key[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  idx[i] = (int32_t)i;
}

// one workgroup per tile of `rows` records: the tile's box (lo[3], hi[3])
__global__ void __launch_bounds__(kCullBlock) cull_tiles(const double *kc, const int32_t *perm, int64_t N,
                                                         int rows, double *boxes) {
  __shared__ double red[6][kCullBlock];
  const int t = threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.x * rows;
  const int64_t k1 = k0 + rows < N ? k0 + rows : N;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t k = k0 + t; k < k1; k += kCullBlock) {
    const int64_t i = perm[k];
    for (int c = 0; c < 3; ++c) {
      v[c] = fmin(v[c], kc[3 * i + c]);
      v[3 + c] = fmax(v[3 + c], kc[3 * i + c]);
    }
  }
  for (int c = 0; c < 6; ++c) red[c][t] = v[c];
  __syncthreads();
  for (int w = kCullBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      for (int c = 0; c < 3; ++c) {
        red[c][t] = fmin(red[c][t], red[c][t + w]);
        red[3 + c][t] = fmax(red[3 + c][t], red[3 + c][t + w]);
      }
    }
    __syncthreads();
  }
  if (t < 6) boxes[(size_t)blockIdx.x * 6 + t] = red[t][0];
}

struct CullWs {
  double *kc, *part;
  uint32_t *key_in, *key_out;
  int32_t *idx;
  void *sort_tmp;
  size_t sort_bytes;
  size_t total;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// carves the workspace at base (sizes only when base is null)
hipError_t cull_ws_layout(int64_t N, char *base, CullWs &w) {
  const int64_t nblk = (N + kCullBlock - 1) / kCullBlock;
  size_t sort_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t *)nullptr,
                                                    (uint32_t *)nullptr, (const int32_t *)nullptr,
                                                    (int32_t *)nullptr, (int)N, 0, 30, (hipStream_t)0);
  if (e != hipSuccess) return e;
  size_t o = 0;
  auto at = [&](size_t bytes) {
    void *q = base ? (void *)(base + o) : nullptr;
    o += align256(bytes);
    return q;
  };
  w.kc = (double *)at(sizeof(double) * 3 * N);
  w.part = (double *)at(sizeof(double) * 8 * nblk);
  w.key_in = (uint32_t *)at(sizeof(uint32_t) * N);
  w.key_out = (uint32_t *)at(sizeof(uint32_t) * N);
  w.idx = (int32_t *)at(sizeof(int32_t) * N);
  w.sort_tmp = at(sort_bytes);
  w.sort_bytes = sort_bytes;
  w.total = o;
  return hipSuccess;
}

}  // namespace

hipError_t cull_workspace_bytes(int64_t N, size_t *bytes) {
  CullWs w;
  const hipError_t e = cull_ws_layout(N, nullptr, w);
  *bytes = w.total;
  return e;
}

// x: the slice's positions (device, point-major, W per point); geom 0 flat, 1 sphere in degrees,
// 2 sphere in radians.  Writes the record order to perm (N entries) and to info the slice box
// (8 doubles: lo[3], hi[3], non-finite count, largest |coordinate|) followed by one box per tile
// (6 doubles each, ceil(N / rows) tiles).  ws: cull_workspace_bytes(N) bytes of device memory.
hipError_t launch_cull(const double *x, int64_t N, int W, int D, int geom, int rows, void *ws,
                       int32_t *perm, double *info, hipStream_t st) {
  if (N < 1 || N > 0x7fffffffLL || rows < 1) return hipErrorInvalidValue;
  CullWs w;
  hipError_t e = cull_ws_layout(N, (char *)ws, w);
  if (e != hipSuccess) return e;
  const int nblk = (int)((N + kCullBlock - 1) / kCullBlock);
  const int ntile = (int)((N + rows - 1) / rows);
  hipLaunchKernelGGL(cull_coords, dim3(nblk), dim3(kCullBlock), 0, st, x, N, W, D, geom, w.kc, w.part);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(cull_reduce, dim3(1), dim3(kCullBlock), 0, st, (const double *)w.part, nblk, info);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(cull_keys, dim3(nblk), dim3(kCullBlock), 0, st, (const double *)w.kc, N,
                     (const double *)info, w.key_in, w.idx);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  size_t sb = w.sort_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, sb, (const uint32_t *)w.key_in, w.key_out,
                                         (const int32_t *)w.idx, perm, (int)N, 0, 30, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(cull_tiles, dim3(ntile), dim3(kCullBlock), 0, st, (const double *)w.kc,
                     (const int32_t *)perm, N, rows, info + 8);
  return hipGetLastError();
}

}  // namespace sfk
