// Fused full leapfrog step (H then E) for gfx950, two field banks.
//
// One launch reads bank A (E, H, CPML memories) and writes bank B, so no tile
// ever reads a value another tile is overwriting.  Each workgroup owns a
// 64 (x) x kTY (y) column of cells and marches along z over a chunk of planes:
//
//   plane k:  H_new(k) at its own cells        -> registers, LDS, bank B
//             H_new(k) on a one-cell low halo  -> LDS only (recomputed, the
//             (row j0-1 and column i0-1)          owning tile writes it)
//             barrier
//             E_new(k) at its own cells from H_new(k) (LDS: own, j-1, i-1) and
//             H_new(k-1) (registers: same (j, i)) -> bank B
//             barrier
//
// Every E/H/coefficient element is read from HBM once per step and every
// output written once (halo re-reads are served by the caches), which is the
// compulsory-traffic bound the roofline is priced on.
//
// Numerical contract: identical to phase.hip (kernels.py:458-626,
// lattice.py:523-575, kernels.py:184-264, _boundary_compile.py:118-160);
// arithmetic is float32 in the same operation order.

#include "common.h"

namespace bzh {
namespace {

constexpr int kTX = 64;
constexpr int kTY = 4;

template <typename T>
__device__ __forceinline__ float ldf(const void* p, int i) {
  return load_f(static_cast<const T*>(p), i);
}

__device__ __forceinline__ float cf(const float* arr, int st, int i) { return arr[i * st]; }

// Packed CPML correction reading psi from bank A and, for owned cells, writing
// bank B.  ``E`` is the psi extent along the term's axis (2L for uniform L).
template <bool kWrite, typename F>
__device__ __forceinline__ float corr(const PhaseParams& p, void* const* psi_out, int t, float d,
                                      int c, F lin) {
  const DevTerm& tm = p.term[t];
  int low, high;
  if (p.uniform_l > 0) {
    low = high = p.uniform_l;
  } else {
    low = p.meta[t * 5 + 2];
    high = p.meta[t * 5 + 3];
  }
  int q = -1;
  if (c < low) q = c;
  else if (c >= tm.n - high) q = low + c - (tm.n - high);
  if (q < 0 || q >= tm.low + tm.high) return d;
  const int idx = lin(q, tm.low + tm.high);
  float psi = tm.psi_bf16 ? load_f(static_cast<const bf16_t*>(tm.psi), idx)
                          : static_cast<const float*>(tm.psi)[idx];
  psi = tm.b[q] * psi + tm.a[q] * d;
  if constexpr (kWrite) {
    if (tm.psi_bf16) store_f(static_cast<bf16_t*>(psi_out[t]), idx, psi);
    else static_cast<float*>(psi_out[t])[idx] = psi;
  }
  return d * tm.ik[q] + psi;
}

// ---- H updates at one cell (reads bank A); bounds are the caller's ----------
template <typename T, bool kCpml, bool kWrite>
__device__ __forceinline__ float h_x(const FusedParams& f, int k, int j, int i) {
  const PhaseParams& p = f.h;
  const int ny = p.ny, nx = p.nx;
  const int ez = (k * (ny + 1) + j) * (nx + 1) + i;
  const int ey = (k * ny + j) * (nx + 1) + i;
  float d0 = (ldf<T>(p.src[2], ez + nx + 1) - ldf<T>(p.src[2], ez)) * p.inv_dx;
  float d1 = (ldf<T>(p.src[1], ey + ny * (nx + 1)) - ldf<T>(p.src[1], ey)) * p.inv_dx;
  if constexpr (kCpml) {
    d0 = corr<kWrite>(p, f.psi_h_out, 0, d0, j, [&](int q, int E) { return (k * E + q) * (nx + 1) + i; });
    d1 = corr<kWrite>(p, f.psi_h_out, 1, d1, k, [&](int q, int) { return (q * ny + j) * (nx + 1) + i; });
  }
  const int o = (k * ny + j) * (nx + 1) + i;
  float h = cf(p.dec[0], p.dec_st[0], o) * ldf<T>(p.tin[0], o) -
            cf(p.scf[0], p.scf_st[0], o) * p.ratio * (d0 - d1);
  if (((p.metallic & kLeft) && i == 0) || ((p.metallic & kRight) && i == nx)) h = 0.0f;
  return h;
}

template <typename T, bool kCpml, bool kWrite>
__device__ __forceinline__ float h_y(const FusedParams& f, int k, int j, int i) {
  const PhaseParams& p = f.h;
  const int ny = p.ny, nx = p.nx;
  const int ex = (k * (ny + 1) + j) * nx + i;
  const int ez = (k * (ny + 1) + j) * (nx + 1) + i;
  float d2 = (ldf<T>(p.src[0], ex + (ny + 1) * nx) - ldf<T>(p.src[0], ex)) * p.inv_dx;
  float d3 = (ldf<T>(p.src[2], ez + 1) - ldf<T>(p.src[2], ez)) * p.inv_dx;
  if constexpr (kCpml) {
    d2 = corr<kWrite>(p, f.psi_h_out, 2, d2, k, [&](int q, int) { return (q * (ny + 1) + j) * nx + i; });
    d3 = corr<kWrite>(p, f.psi_h_out, 3, d3, i, [&](int q, int E) { return (k * (ny + 1) + j) * E + q; });
  }
  const int o = ex;
  float h = cf(p.dec[1], p.dec_st[1], o) * ldf<T>(p.tin[1], o) -
            cf(p.scf[1], p.scf_st[1], o) * p.ratio * (d2 - d3);
  if (((p.metallic & kBottom) && j == 0) || ((p.metallic & kTop) && j == ny)) h = 0.0f;
  return h;
}

template <typename T, bool kCpml, bool kWrite>
__device__ __forceinline__ float h_z(const FusedParams& f, int k, int j, int i) {
  const PhaseParams& p = f.h;
  const int nz = p.nz, ny = p.ny, nx = p.nx;
  const int ey = (k * ny + j) * (nx + 1) + i;
  const int ex = (k * (ny + 1) + j) * nx + i;
  float d4 = (ldf<T>(p.src[1], ey + 1) - ldf<T>(p.src[1], ey)) * p.inv_dx;
  float d5 = (ldf<T>(p.src[0], ex + nx) - ldf<T>(p.src[0], ex)) * p.inv_dx;
  if constexpr (kCpml) {
    d4 = corr<kWrite>(p, f.psi_h_out, 4, d4, i, [&](int q, int E) { return (k * ny + j) * E + q; });
    d5 = corr<kWrite>(p, f.psi_h_out, 5, d5, j, [&](int q, int E) { return (k * E + q) * nx + i; });
  }
  const int o = (k * ny + j) * nx + i;
  float h = cf(p.dec[2], p.dec_st[2], o) * ldf<T>(p.tin[2], o) -
            cf(p.scf[2], p.scf_st[2], o) * p.ratio * (d4 - d5);
  if (((p.metallic & kFront) && k == 0) || ((p.metallic & kBack) && k == nz)) h = 0.0f;
  return h;
}

// Backward difference with the ghost rule from already-loaded neighbours:
// ``cur`` is H at c (valid when c < N), ``prev`` is H at c - 1 (valid when c > 0).
__device__ __forceinline__ float gdiff(float cur, float prev, int c, int N, bool mlo, bool mhi) {
  const float up = (c < N) ? cur : (mhi ? 0.0f : prev);
  const float lo = (c > 0) ? prev : (mlo ? 0.0f : cur);
  return up - lo;
}

template <typename T, bool kCpml>
__global__ void __launch_bounds__(kTX * kTY) fused_step(FusedParams f) {
  __shared__ float sHx[kTY + 1][kTX + 1];
  __shared__ float sHy[kTY + 1][kTX + 1];
  __shared__ float sHz[kTY + 1][kTX + 1];

  const PhaseParams& hp = f.h;
  const PhaseParams& ep = f.e;
  const int nz = hp.nz, ny = hp.ny, nx = hp.nx;
  const int m = hp.metallic;
  const int tx = threadIdx.x, ty = threadIdx.y;

  // tile order: x fastest, then y, then z-chunk
  int b = blockIdx.x;
  const int bx = b % f.tiles_x;
  b /= f.tiles_x;
  const int by = b % f.tiles_y;
  const int bz = b / f.tiles_y;
  const int i0 = bx * kTX, j0 = by * kTY;
  const int i = i0 + tx, j = j0 + ty;
  const int kz0 = bz * f.kz;
  const int kz1 = min(kz0 + f.kz, nz + 1);

  // H_new(k-1) at this cell, carried along z
  float prev_hx = 0.0f, prev_hy = 0.0f;
  if (kz0 > 0) {
    const int k = kz0 - 1;  // < nz
    if (j < ny && i <= nx) prev_hx = h_x<T, kCpml, false>(f, k, j, i);
    if (j <= ny && i < nx) prev_hy = h_y<T, kCpml, false>(f, k, j, i);
  }

  for (int k = kz0; k < kz1; ++k) {
    // ---- H at owned cells ----
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
    const bool vhx = k < nz && j < ny && i <= nx;
    const bool vhy = k < nz && j <= ny && i < nx;
    const bool vhz = j < ny && i < nx;  // k <= nz always here
    if (vhx) {
      hx = h_x<T, kCpml, true>(f, k, j, i);
      store_f(static_cast<T*>(hp.tout[0]), (k * ny + j) * (nx + 1) + i, hx);
    }
    if (vhy) {
      hy = h_y<T, kCpml, true>(f, k, j, i);
      store_f(static_cast<T*>(hp.tout[1]), (k * (ny + 1) + j) * nx + i, hy);
    }
    if (vhz) {
      hz = h_z<T, kCpml, true>(f, k, j, i);
      store_f(static_cast<T*>(hp.tout[2]), (k * ny + j) * nx + i, hz);
    }
    sHx[ty + 1][tx + 1] = hx;
    sHy[ty + 1][tx + 1] = hy;
    sHz[ty + 1][tx + 1] = hz;
    // ---- halo: row j0-1 (Hx, Hz) and column i0-1 (Hy, Hz), recomputed ----
    if (ty == 0 && j0 > 0) {
      const int jh = j0 - 1;
      sHx[0][tx + 1] = (k < nz && i <= nx) ? h_x<T, kCpml, false>(f, k, jh, i) : 0.0f;
      sHz[0][tx + 1] = (i < nx) ? h_z<T, kCpml, false>(f, k, jh, i) : 0.0f;
    }
    if (tx == 0 && i0 > 0) {
      const int ih = i0 - 1;
      sHy[ty + 1][0] = (k < nz && j <= ny) ? h_y<T, kCpml, false>(f, k, j, ih) : 0.0f;
      sHz[ty + 1][0] = (j < ny) ? h_z<T, kCpml, false>(f, k, j, ih) : 0.0f;
    }
    __syncthreads();

    // ---- E at owned cells ----
    // Ex (nz+1, ny+1, nx): d(Hz)/dy (+), d(Hy)/dz (-)
    if (j <= ny && i < nx) {
      const int o = (k * (ny + 1) + j) * nx + i;
      float d0 = gdiff(sHz[ty + 1][tx + 1], sHz[ty][tx + 1], j, ny, m & kBottom, m & kTop) * ep.inv_dx;
      float d1 = gdiff(hy, prev_hy, k, nz, m & kFront, m & kBack) * ep.inv_dx;
      if constexpr (kCpml) {
        d0 = corr<true>(ep, f.psi_e_out, 0, d0, j, [&](int q, int E) { return (k * E + q) * nx + i; });
        d1 = corr<true>(ep, f.psi_e_out, 1, d1, k, [&](int q, int) { return (q * (ny + 1) + j) * nx + i; });
      }
      float e = cf(ep.dec[0], ep.dec_st[0], o) * ldf<T>(ep.tin[0], o) +
                cf(ep.scf[0], ep.scf_st[0], o) * ep.ratio * (d0 - d1);
      if (((m & kFront) && k == 0) || ((m & kBack) && k == nz) ||
          ((m & kBottom) && j == 0) || ((m & kTop) && j == ny))
        e = 0.0f;
      store_f(static_cast<T*>(ep.tout[0]), o, e);
    }
    // Ey (nz+1, ny, nx+1): d(Hx)/dz (+), d(Hz)/dx (-)
    if (j < ny && i <= nx) {
      const int o = (k * ny + j) * (nx + 1) + i;
      float d2 = gdiff(hx, prev_hx, k, nz, m & kFront, m & kBack) * ep.inv_dx;
      float d3 = gdiff(sHz[ty + 1][tx + 1], sHz[ty + 1][tx], i, nx, m & kLeft, m & kRight) * ep.inv_dx;
      if constexpr (kCpml) {
        d2 = corr<true>(ep, f.psi_e_out, 2, d2, k, [&](int q, int) { return (q * ny + j) * (nx + 1) + i; });
        d3 = corr<true>(ep, f.psi_e_out, 3, d3, i, [&](int q, int E) { return (k * ny + j) * E + q; });
      }
      float e = cf(ep.dec[1], ep.dec_st[1], o) * ldf<T>(ep.tin[1], o) +
                cf(ep.scf[1], ep.scf_st[1], o) * ep.ratio * (d2 - d3);
      if (((m & kFront) && k == 0) || ((m & kBack) && k == nz) ||
          ((m & kLeft) && i == 0) || ((m & kRight) && i == nx))
        e = 0.0f;
      store_f(static_cast<T*>(ep.tout[1]), o, e);
    }
    // Ez (nz, ny+1, nx+1): d(Hy)/dx (+), d(Hx)/dy (-)
    if (k < nz && j <= ny && i <= nx) {
      const int o = (k * (ny + 1) + j) * (nx + 1) + i;
      float d4 = gdiff(sHy[ty + 1][tx + 1], sHy[ty + 1][tx], i, nx, m & kLeft, m & kRight) * ep.inv_dx;
      float d5 = gdiff(sHx[ty + 1][tx + 1], sHx[ty][tx + 1], j, ny, m & kBottom, m & kTop) * ep.inv_dx;
      if constexpr (kCpml) {
        d4 = corr<true>(ep, f.psi_e_out, 4, d4, i, [&](int q, int E) { return (k * (ny + 1) + j) * E + q; });
        d5 = corr<true>(ep, f.psi_e_out, 5, d5, j, [&](int q, int E) { return (k * E + q) * (nx + 1) + i; });
      }
      float e = cf(ep.dec[2], ep.dec_st[2], o) * ldf<T>(ep.tin[2], o) +
                cf(ep.scf[2], ep.scf_st[2], o) * ep.ratio * (d4 - d5);
      if (((m & kBottom) && j == 0) || ((m & kTop) && j == ny) ||
          ((m & kLeft) && i == 0) || ((m & kRight) && i == nx))
        e = 0.0f;
      store_f(static_cast<T*>(ep.tout[2]), o, e);
    }
    __syncthreads();
    prev_hx = hx;
    prev_hy = hy;
  }
}

template <typename T>
void dispatch(hipStream_t s, const FusedParams& p) {
  dim3 block(kTX, kTY, 1);
  dim3 grid(p.tiles_x * p.tiles_y * p.nzc, 1, 1);
  if (p.h.nterms > 0) hipLaunchKernelGGL((fused_step<T, true>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((fused_step<T, false>), grid, block, 0, s, p);
}

}  // namespace

void launch_fused3d(hipStream_t s, const FusedParams& p, int dtype) {
  BZH_REQUIRE(p.tiles_x == (p.h.nx + 1 + kTX - 1) / kTX && p.tiles_y == (p.h.ny + 1 + kTY - 1) / kTY,
              "fused tiling does not match the grid");
  BZH_REQUIRE(p.kz >= 1 && p.nzc * p.kz >= p.h.nz + 1, "fused z chunks do not cover the grid");
  switch (dtype) {
    case BZH_F32: dispatch<float>(s, p); break;
    case BZH_BF16: dispatch<bf16_t>(s, p); break;
    case BZH_F16: dispatch<__half>(s, p); break;
    default: throw Failure{"unsupported field dtype"};
  }
  BZH_HIP(hipGetLastError());
}

}  // namespace bzh
