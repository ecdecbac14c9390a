// phem_gen.hip -- synthetic graph generators on the device.
//
// Counter-based streams: draw(seed, stream, i) = mix64(base(seed, stream) + i)
// with base(seed, stream) = mix64(seed * GOLDEN + stream).  Stream use:
// 0/1 uniform endpoints, 3 uniform self-loop redraws (index i*64 + t),
// 2 weights, 8 + a RMAT attempt a (index i*16 + level/2).  Edges are stored
// with u < v; weights are exact float32 k * 2^-24.
#include "phem_dev.h"
#include "phem_internal.h"

namespace phem {

constexpr int GBLK = 256;

__device__ __forceinline__ uint64_t draw(uint64_t base, uint64_t i) {
  return mix64(base + i);
}

__global__ __launch_bounds__(GBLK) void k_gen_uniform(
    uint32_t n, uint64_t m, uint64_t bu, uint64_t bv, uint64_t bre, uint64_t bw,
    uint64_t id0, int positive, uint32_t *u, uint32_t *v, uint32_t *w) {
  for (uint64_t k = (uint64_t)blockIdx.x * GBLK + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * GBLK) {
    const uint64_t i = id0 + k;
    uint32_t a = mulhi_n(draw(bu, i), n);
    uint32_t b = mulhi_n(draw(bv, i), n);
    for (int t = 0; t < 64 && a == b; t++) b = mulhi_n(draw(bre, i * 64 + t), n);
    if (a == b) b = (uint32_t)(((uint64_t)a + 1) % n);
    u[k] = a < b ? a : b;
    v[k] = a < b ? b : a;
    w[k] = weight_bits(draw(bw, i), positive);
  }
}

__global__ __launch_bounds__(GBLK) void k_gen_rmat(
    uint32_t scale, uint64_t m, uint64_t seed, uint32_t t1, uint32_t t2,
    uint32_t t3, uint64_t bw, int positive, uint32_t *u, uint32_t *v,
    uint32_t *w) {
  for (uint64_t i = (uint64_t)blockIdx.x * GBLK + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * GBLK) {
    uint32_t a = 0, b = 0;
    for (int att = 0; att < 64; att++) {
      const uint64_t base = stream_base(seed, 8 + att);
      a = 0;
      b = 0;
      for (uint32_t lvl = 0; lvl < scale; lvl += 2) {
        const uint64_t h = draw(base, i * 16 + lvl / 2);
        uint32_t x = (uint32_t)(h >> 32);
        a = (a << 1) | (x >= t2);
        b = (b << 1) | ((x >= t1 && x < t2) || x >= t3);
        if (lvl + 1 < scale) {
          x = (uint32_t)h;
          a = (a << 1) | (x >= t2);
          b = (b << 1) | ((x >= t1 && x < t2) || x >= t3);
        }
      }
      if (a != b) break;
    }
    if (a == b) b = a ^ 1u;
    u[i] = a < b ? a : b;
    v[i] = a < b ? b : a;
    w[i] = weight_bits(draw(bw, i), positive);
  }
}

__global__ __launch_bounds__(GBLK) void k_pack(const uint32_t *__restrict__ u,
                                               const uint32_t *__restrict__ v,
                                               uint64_t m, ull *__restrict__ e) {
  for (uint64_t i = (uint64_t)blockIdx.x * GBLK + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * GBLK)
    e[i] = ((ull)u[i] << 32) | v[i];
}

}  // namespace phem

using namespace phem;

extern "C" int phem_gen_uniform(uint32_t n, uint64_t m, uint64_t seed,
                                uint64_t id0, int positive_w, uint32_t *d_u,
                                uint32_t *d_v, uint32_t *d_w, void *stream) {
  if (n < 2) return set_err(PHEM_EINVAL, "uniform generator needs n >= 2");
  if (m && (!d_u || !d_v || !d_w)) return set_err(PHEM_EINVAL, "NULL output");
  if (!m) return PHEM_OK;
  {
    PHEM_PROF("k_gen_uniform", (hipStream_t)stream);
    hipLaunchKernelGGL(k_gen_uniform, dim3(grid_for(m, GBLK)), dim3(GBLK), 0,
                       (hipStream_t)stream, n, m, stream_base(seed, 0),
                       stream_base(seed, 1), stream_base(seed, 3),
                       stream_base(seed, 2), id0, positive_w, d_u, d_v, d_w);
  }
  PHEM_LAUNCHED("k_gen_uniform");
  return PHEM_OK;
}

extern "C" int phem_gen_rmat(uint32_t scale, uint64_t m, uint64_t seed,
                             uint32_t t1, uint32_t t2, uint32_t t3,
                             int positive_w, uint32_t *d_u, uint32_t *d_v,
                             uint32_t *d_w, void *stream) {
  if (scale < 1 || scale > 31)
    return set_err(PHEM_EINVAL, "rmat scale %u outside 1..31", scale);
  if (!(t1 <= t2 && t2 <= t3)) return set_err(PHEM_EINVAL, "thresholds unordered");
  if (m && (!d_u || !d_v || !d_w)) return set_err(PHEM_EINVAL, "NULL output");
  if (!m) return PHEM_OK;
  {
    PHEM_PROF("k_gen_rmat", (hipStream_t)stream);
    hipLaunchKernelGGL(k_gen_rmat, dim3(grid_for(m, GBLK)), dim3(GBLK), 0,
                       (hipStream_t)stream, scale, m, seed, t1, t2, t3,
                       stream_base(seed, 2), positive_w, d_u, d_v, d_w);
  }
  PHEM_LAUNCHED("k_gen_rmat");
  return PHEM_OK;
}

extern "C" int phem_pack_edges(const uint32_t *d_u, const uint32_t *d_v,
                               uint64_t m, uint64_t *d_edges, void *stream) {
  if (m && (!d_u || !d_v || !d_edges)) return set_err(PHEM_EINVAL, "NULL array");
  if (!m) return PHEM_OK;
  {
    PHEM_PROF("k_pack", (hipStream_t)stream);
    hipLaunchKernelGGL(k_pack, dim3(grid_for(m, GBLK)), dim3(GBLK), 0,
                       (hipStream_t)stream, d_u, d_v, m, (ull *)d_edges);
  }
  PHEM_LAUNCHED("k_pack");
  return PHEM_OK;
}
