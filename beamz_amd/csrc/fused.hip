// Fused full leapfrog step (H then E) for gfx950, two field banks.
//
// One launch reads bank A (E, H, CPML memories) and writes bank B, so no
// workgroup ever reads a value another is overwriting.
//
// Mapping: one wave64 per workgroup.  A wave owns a (x, y) tile of 62 x kTY
// output cells and marches along z over a chunk of planes.  Lane l sits at
// x = i0 - 1 + l: lanes 1..62 own outputs, lanes 0 and 63 are halo columns.
// Each lane keeps kTY + 2 rows (y = j0 - 1 .. j0 + kTY) of E in registers.
//
//   plane k:  issue every load of the plane (E(k+1), H_A(k), coefficients)
//             H_new(k) on rows j0-1 .. j0+kTY-1, lanes 0..62 (row j0-1 and
//               lane 0 are recomputed halos; only owned cells are stored)
//             E_new(k) on owned cells from H_new(k) (own, y-1 in registers,
//               x-1 by cross-lane permute) and H_new(k-1) (carried along z)
//
// There is no LDS traffic and no barrier: x-neighbours cross lanes through
// ds_bpermute, y-neighbours are in the lane's own registers, z-neighbours
// are carried from the previous plane.  Every E/H/coefficient element is read
// from HBM once per step and every output written once (halo re-reads hit the
// caches): the compulsory-traffic bound the roofline is priced on.
//
// Numerical contract: identical to phase.hip (kernels.py:458-626,
// lattice.py:523-575, kernels.py:184-264 with a uniform slab thickness L,
// _boundary_compile.py:118-160); arithmetic is float32.

#include <cstdlib>

#include "common.h"

namespace bzh {
namespace {

template <typename T>
__device__ __forceinline__ float ldf(const void* p, int i) {
  return load_f(static_cast<const T*>(p), i);
}

// Value of the lane above / below (x + 1 / x - 1).
__device__ __forceinline__ float from_right(float v, int addr_dn) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr_dn, __float_as_int(v)));
}
__device__ __forceinline__ float from_left(float v, int addr_up) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr_up, __float_as_int(v)));
}

// Ghost-rule backward difference from loaded neighbours (lattice.py:523-575).
__device__ __forceinline__ float gdiff(float cur, float prev, int c, int N, bool mlo, bool mhi) {
  const float up = (c < N) ? cur : (mhi ? 0.0f : prev);
  const float lo = (c > 0) ? prev : (mlo ? 0.0f : cur);
  return up - lo;
}

// Packed slab coordinate for a uniform thickness L on an axis of extent n.
__device__ __forceinline__ int slab_q(int c, int n, int L) {
  return c < L ? c : (c >= n - L ? c - (n - 2 * L) : -1);
}

struct Ctx {
  int nz, ny, nx, L, m;
  float idx;
};

// Apply one packed CPML term: psi read from bank A (f.psi[t]), written to
// bank B (psi + delta) when ``own``.  ``lin`` is the psi linear index.
template <int kT>
__device__ __forceinline__ float cpml(const FusedParams& f, float d, int q, int lin, bool own) {
  const int L2 = 2 * f.L;
  const float* tab = f.tab + kT * 3 * L2;
  if (f.psi_bf16) {
    const bf16_t* pin = static_cast<const bf16_t*>(f.psi[kT]);
    float psi = tab[L2 + q] * load_f(pin, lin) + tab[q] * d;
    if (own) store_f(const_cast<bf16_t*>(pin) + f.psi_delta, lin, psi);
    return d * tab[2 * L2 + q] + psi;
  }
  const float* pin = static_cast<const float*>(f.psi[kT]);
  float psi = tab[L2 + q] * pin[lin] + tab[q] * d;
  if (own) const_cast<float*>(pin)[f.psi_delta + lin] = psi;
  return d * tab[2 * L2 + q] + psi;
}

template <typename T, int kTY>
struct Plane {
  // E(k) on rows 0..kTY+1 (y = j0-1+r), this lane's column
  float ex[kTY + 2], ey[kTY + 2], ez[kTY + 2];
};

template <typename T, int kTY>
__device__ __forceinline__ void load_e(const FusedParams& f, const Ctx& c, int k, int j0, int x,
                                       Plane<T, kTY>& P) {
  const int nz = c.nz, ny = c.ny, nx = c.nx;
#pragma unroll
  for (int r = 0; r < kTY + 2; ++r) {
    const int y = j0 - 1 + r;
    const bool okx = k <= nz && y >= 0 && y <= ny && x >= 0 && x < nx;
    const bool oky = k <= nz && y >= 0 && y < ny && x >= 0 && x <= nx;
    const bool okz = k < nz && y >= 0 && y <= ny && x >= 0 && x <= nx;
    // clamp to a valid element so every load is unconditional; the value of an
    // out-of-support sample is replaced by zero
    const int kx = min(max(k, 0), nz), yx = min(max(y, 0), ny), xx = min(max(x, 0), nx - 1);
    const int yy = min(max(y, 0), ny - 1), xy = min(max(x, 0), nx);
    const int kz = min(max(k, 0), nz - 1);
    const float vx = ldf<T>(f.e_in[0], (kx * (ny + 1) + yx) * nx + xx);
    const float vy = ldf<T>(f.e_in[1], (kx * ny + yy) * (nx + 1) + xy);
    const float vz = ldf<T>(f.e_in[2], (kz * (ny + 1) + yx) * (nx + 1) + xy);
    P.ex[r] = okx ? vx : 0.0f;
    P.ey[r] = oky ? vy : 0.0f;
    P.ez[r] = okz ? vz : 0.0f;
  }
}

// H_new(k) rows 0..kTY at this lane.  ``own`` selects the cells written to
// bank B.  kCpml enables the packed CPML terms.
template <typename T, int kTY, bool kCpml>
__device__ __forceinline__ void h_plane(const FusedParams& f, const Ctx& c, int k, int j0, int x,
                                        bool lane_own, bool write, const Plane<T, kTY>& E0,
                                        const Plane<T, kTY>& E1, int dn, float* hx, float* hy,
                                        float* hz, const float hc[6]) {
  const int nz = c.nz, ny = c.ny, nx = c.nx, m = c.m;
  const float idx = c.idx;
  // H_A(k) loads (clamped, unconditional)
  float ha[3][kTY + 1];
#pragma unroll
  for (int r = 0; r <= kTY; ++r) {
    const int y = j0 - 1 + r;
    const int kk = min(k, nz - 1), ky = min(max(y, 0), ny - 1), kyy = min(max(y, 0), ny);
    const int xi = min(max(x, 0), nx), xo = min(max(x, 0), nx - 1);
    ha[0][r] = ldf<T>(f.h_in[0], (kk * ny + ky) * (nx + 1) + xi);
    ha[1][r] = ldf<T>(f.h_in[1], (kk * (ny + 1) + kyy) * nx + xo);
    ha[2][r] = ldf<T>(f.h_in[2], (min(k, nz) * ny + ky) * nx + xo);
  }
  // per-cell H coefficients (magnetic loss) are rare: scalars live in hc[]
#pragma unroll
  for (int r = 0; r <= kTY; ++r) {
    const int y = j0 - 1 + r;
    const bool own = write && lane_own && r >= 1;
    // Hx (nz, ny, nx+1): dEz/dy (+), dEy/dz (-)
    {
      const bool ok = k < nz && y >= 0 && y < ny && x >= 0 && x <= nx;
      float d0 = (E0.ez[r + 1] - E0.ez[r]) * idx;
      float d1 = (E1.ey[r] - E0.ey[r]) * idx;
      if constexpr (kCpml) {
        if (ok) {
          const int q0 = slab_q(y, ny, c.L);
          if (q0 >= 0) d0 = cpml<0>(f, d0, q0, (k * 2 * c.L + q0) * (nx + 1) + x, own);
          const int q1 = slab_q(k, nz, c.L);
          if (q1 >= 0) d1 = cpml<1>(f, d1, q1, (q1 * ny + y) * (nx + 1) + x, own);
        }
      }
      float h = hc[0] * ha[0][r] - hc[3] * (d0 - d1);
      if (((m & kLeft) && x == 0) || ((m & kRight) && x == nx)) h = 0.0f;
      hx[r] = ok ? h : 0.0f;
    }
    // Hy (nz, ny+1, nx): dEx/dz (+), dEz/dx (-)
    {
      const bool ok = k < nz && y >= 0 && y <= ny && x >= 0 && x < nx;
      float d2 = (E1.ex[r] - E0.ex[r]) * idx;
      float d3 = (from_right(E0.ez[r], dn) - E0.ez[r]) * idx;
      if constexpr (kCpml) {
        if (ok) {
          const int q2 = slab_q(k, nz, c.L);
          if (q2 >= 0) d2 = cpml<2>(f, d2, q2, (q2 * (ny + 1) + y) * nx + x, own);
          const int q3 = slab_q(x, nx, c.L);
          if (q3 >= 0) d3 = cpml<3>(f, d3, q3, (k * (ny + 1) + y) * 2 * c.L + q3, own);
        }
      }
      float h = hc[1] * ha[1][r] - hc[4] * (d2 - d3);
      if (((m & kBottom) && y == 0) || ((m & kTop) && y == ny)) h = 0.0f;
      hy[r] = ok ? h : 0.0f;
    }
    // Hz (nz+1, ny, nx): dEy/dx (+), dEx/dy (-)
    {
      const bool ok = k <= nz && y >= 0 && y < ny && x >= 0 && x < nx;
      float d4 = (from_right(E0.ey[r], dn) - E0.ey[r]) * idx;
      float d5 = (E0.ex[r + 1] - E0.ex[r]) * idx;
      if constexpr (kCpml) {
        if (ok) {
          const int q4 = slab_q(x, nx, c.L);
          if (q4 >= 0) d4 = cpml<4>(f, d4, q4, (k * ny + y) * 2 * c.L + q4, own);
          const int q5 = slab_q(y, ny, c.L);
          if (q5 >= 0) d5 = cpml<5>(f, d5, q5, (k * 2 * c.L + q5) * nx + x, own);
        }
      }
      float h = hc[2] * ha[2][r] - hc[5] * (d4 - d5);
      if (((m & kFront) && k == 0) || ((m & kBack) && k == nz)) h = 0.0f;
      hz[r] = ok ? h : 0.0f;
    }
  }
}

template <typename T, int kTY, bool kCpml>
__device__ __forceinline__ void plane_step(const FusedParams& f, const Ctx& c, int k, int j0, int x,
                                           bool lane_own, const Plane<T, kTY>& E0,
                                           const Plane<T, kTY>& E1, int dn, int up,
                                           float* prev_hx, float* prev_hy, const float hc[6],
                                           const float ec[6]) {
  const int nz = c.nz, ny = c.ny, nx = c.nx, m = c.m;
  const float idx = c.idx;
  // E coefficients of this plane's owned rows (per-cell arrays or scalars)
  float es[3][kTY];
#pragma unroll
  for (int r = 0; r < kTY; ++r) {
    const int y = j0 + r;
    const int kk = min(k, nz), yy1 = min(y, ny), yy0 = min(y, ny - 1), kz = min(k, nz - 1);
    const int xo = min(max(x, 0), nx - 1), xi = min(max(x, 0), nx);
    es[0][r] = f.esrc_st[0] ? f.esrc[0][(kk * (ny + 1) + yy1) * nx + xo] : ec[3];
    es[1][r] = f.esrc_st[1] ? f.esrc[1][(kk * ny + yy0) * (nx + 1) + xi] : ec[4];
    es[2][r] = f.esrc_st[2] ? f.esrc[2][(kz * (ny + 1) + yy1) * (nx + 1) + xi] : ec[5];
  }
  float hx[kTY + 1], hy[kTY + 1], hz[kTY + 1];
  h_plane<T, kTY, kCpml>(f, c, k, j0, x, lane_own, true, E0, E1, dn, hx, hy, hz, hc);
  // store owned H
#pragma unroll
  for (int r = 1; r <= kTY; ++r) {
    const int y = j0 - 1 + r;
    if (lane_own) {
      if (k < nz && y < ny && x <= nx) store_f(static_cast<T*>(f.h_out[0]), (k * ny + y) * (nx + 1) + x, hx[r]);
      if (k < nz && y <= ny && x < nx) store_f(static_cast<T*>(f.h_out[1]), (k * (ny + 1) + y) * nx + x, hy[r]);
      if (y < ny && x < nx) store_f(static_cast<T*>(f.h_out[2]), (k * ny + y) * nx + x, hz[r]);
    }
  }
  // E_new(k) on owned rows
#pragma unroll
  for (int r = 1; r <= kTY; ++r) {
    const int y = j0 - 1 + r;
    const float hz_l = from_left(hz[r], up);
    const float hy_l = from_left(hy[r], up);
    // Ex (nz+1, ny+1, nx): dHz/dy (+), dHy/dz (-)
    if (lane_own && y <= ny && x < nx) {
      float d0 = gdiff(hz[r], hz[r - 1], y, ny, m & kBottom, m & kTop) * idx;
      float d1 = gdiff(hy[r], prev_hy[r - 1], k, nz, m & kFront, m & kBack) * idx;
      if constexpr (kCpml) {
        const int q0 = slab_q(y, ny + 1, c.L);
        if (q0 >= 0) d0 = cpml<6>(f, d0, q0, (k * 2 * c.L + q0) * nx + x, true);
        const int q1 = slab_q(k, nz + 1, c.L);
        if (q1 >= 0) d1 = cpml<7>(f, d1, q1, (q1 * (ny + 1) + y) * nx + x, true);
      }
      const int o = (k * (ny + 1) + y) * nx + x;
      float e = (f.edec_st[0] ? f.edec[0][o] : ec[0]) * E0.ex[r] + es[0][r - 1] * f.ratio_e * (d0 - d1);
      if (((m & kFront) && k == 0) || ((m & kBack) && k == nz) || ((m & kBottom) && y == 0) ||
          ((m & kTop) && y == ny))
        e = 0.0f;
      store_f(static_cast<T*>(f.e_out[0]), o, e);
    }
    // Ey (nz+1, ny, nx+1): dHx/dz (+), dHz/dx (-)
    if (lane_own && y < ny && x <= nx) {
      float d2 = gdiff(hx[r], prev_hx[r - 1], k, nz, m & kFront, m & kBack) * idx;
      float d3 = gdiff(hz[r], hz_l, x, nx, m & kLeft, m & kRight) * idx;
      if constexpr (kCpml) {
        const int q2 = slab_q(k, nz + 1, c.L);
        if (q2 >= 0) d2 = cpml<8>(f, d2, q2, (q2 * ny + y) * (nx + 1) + x, true);
        const int q3 = slab_q(x, nx + 1, c.L);
        if (q3 >= 0) d3 = cpml<9>(f, d3, q3, (k * ny + y) * 2 * c.L + q3, true);
      }
      const int o = (k * ny + y) * (nx + 1) + x;
      float e = (f.edec_st[1] ? f.edec[1][o] : ec[1]) * E0.ey[r] + es[1][r - 1] * f.ratio_e * (d2 - d3);
      if (((m & kFront) && k == 0) || ((m & kBack) && k == nz) || ((m & kLeft) && x == 0) ||
          ((m & kRight) && x == nx))
        e = 0.0f;
      store_f(static_cast<T*>(f.e_out[1]), o, e);
    }
    // Ez (nz, ny+1, nx+1): dHy/dx (+), dHx/dy (-)
    if (lane_own && k < nz && y <= ny && x <= nx) {
      float d4 = gdiff(hy[r], hy_l, x, nx, m & kLeft, m & kRight) * idx;
      float d5 = gdiff(hx[r], hx[r - 1], y, ny, m & kBottom, m & kTop) * idx;
      if constexpr (kCpml) {
        const int q4 = slab_q(x, nx + 1, c.L);
        if (q4 >= 0) d4 = cpml<10>(f, d4, q4, (k * (ny + 1) + y) * 2 * c.L + q4, true);
        const int q5 = slab_q(y, ny + 1, c.L);
        if (q5 >= 0) d5 = cpml<11>(f, d5, q5, (k * 2 * c.L + q5) * (nx + 1) + x, true);
      }
      const int o = (k * (ny + 1) + y) * (nx + 1) + x;
      float e = (f.edec_st[2] ? f.edec[2][o] : ec[2]) * E0.ez[r] + es[2][r - 1] * f.ratio_e * (d4 - d5);
      if (((m & kBottom) && y == 0) || ((m & kTop) && y == ny) || ((m & kLeft) && x == 0) ||
          ((m & kRight) && x == nx))
        e = 0.0f;
      store_f(static_cast<T*>(f.e_out[2]), o, e);
    }
  }
#pragma unroll
  for (int r = 1; r <= kTY; ++r) {
    prev_hx[r - 1] = hx[r];
    prev_hy[r - 1] = hy[r];
  }
}

// Workgroups are dispatched to the 8 XCDs round-robin (g % 8); give each XCD
// a contiguous range of logical tiles so that neighbouring tiles, which share
// halo rows and 128-byte lines, run against the same L2.
__device__ __forceinline__ int xcd_remap(int g, int n) {
  constexpr int kXcd = 8;
  const int c = g % kXcd, m = g / kXcd;
  const int base = n / kXcd, rem = n % kXcd;
  return c * base + min(c, rem) + m;
}

template <typename T, int kTY, bool kCpml>
__global__ void __launch_bounds__(64) fused_step(FusedParams f) {
  Ctx c;
  c.nz = f.nz;
  c.ny = f.ny;
  c.nx = f.nx;
  c.L = f.L;
  c.m = f.metallic;
  c.idx = f.inv_dx;
  const int lane = threadIdx.x;
  const int dn = ((lane + 1) & 63) << 2;
  const int up = ((lane - 1) & 63) << 2;

  int bx, by, kz0, kz1;
  const int g = xcd_remap(blockIdx.x, gridDim.x);
  if (f.tiles != nullptr) {
    const int4 t = f.tiles[g];
    bx = t.x;
    by = t.y;
    kz0 = t.z;
    kz1 = t.w;
  } else {
    int b = g;
    bx = f.box_tx0 + b % f.box_ntx;
    b /= f.box_ntx;
    by = f.box_ty0 + b % f.box_nty;
    const int bz = b / f.box_nty;
    kz0 = f.box_z0 + bz * f.box_kz;
    kz1 = min(kz0 + f.box_kz, f.box_z1);
  }
  const int i0 = bx * 62, j0 = by * kTY;
  const int x = i0 - 1 + lane;
  const bool lane_own = lane >= 1 && lane <= 62 && x <= c.nx;

  // uniform coefficients (rank-0 device scalars) are read once
  float hc[6], ec[6];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    hc[q] = f.hdec_st[q] ? 0.0f : f.hdec[q][0];
    hc[3 + q] = (f.hsrc_st[q] ? 0.0f : f.hsrc[q][0]) * f.ratio_h;
    ec[q] = f.edec_st[q] ? 0.0f : f.edec[q][0];
    ec[3 + q] = f.esrc_st[q] ? 0.0f : f.esrc[q][0];
  }

  Plane<T, kTY> E0, E1;
  float prev_hx[kTY], prev_hy[kTY];
#pragma unroll
  for (int r = 0; r < kTY; ++r) prev_hx[r] = prev_hy[r] = 0.0f;

  if (kz0 > 0) {
    // H_new(kz0 - 1) on this lane's rows (not stored: the chunk below owns it)
    load_e<T, kTY>(f, c, kz0 - 1, j0, x, E1);
    load_e<T, kTY>(f, c, kz0, j0, x, E0);
    float hx[kTY + 1], hy[kTY + 1], hz[kTY + 1];
    h_plane<T, kTY, kCpml>(f, c, kz0 - 1, j0, x, lane_own, false, E1, E0, dn, hx, hy, hz, hc);
#pragma unroll
    for (int r = 1; r <= kTY; ++r) {
      prev_hx[r - 1] = hx[r];
      prev_hy[r - 1] = hy[r];
    }
  } else {
    load_e<T, kTY>(f, c, kz0, j0, x, E0);
  }

  for (int k = kz0; k < kz1; ++k) {
    load_e<T, kTY>(f, c, k + 1, j0, x, E1);
    plane_step<T, kTY, kCpml>(f, c, k, j0, x, lane_own, E0, E1, dn, up, prev_hx, prev_hy, hc, ec);
    E0 = E1;
  }
}

// ---------------------------------------------------------------------------
// Fast path: tiles whose x/y footprint (halos included) avoids every domain
// face and the x/y CPML slabs.  All loads and stores are unconditional, the
// z faces and the z CPML slabs are plane-uniform branches, and every address
// is a 32-bit byte offset on a uniform base (global_load ... saddr).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float ldT(const void* b, uint32_t e) {
  return load_f(reinterpret_cast<const T*>(static_cast<const char*>(b) + e * uint32_t(sizeof(T))), 0);
}
template <typename T>
__device__ __forceinline__ void stT(void* b, uint32_t e, float v) {
  store_f(reinterpret_cast<T*>(static_cast<char*>(b) + e * uint32_t(sizeof(T))), 0, v);
}
__device__ __forceinline__ float ldF(const float* b, uint32_t e) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(b) + e * 4u);
}

// z-term CPML (kT: 1 Hx_z, 2 Hy_z, 7 Ex_z, 8 Ey_z) at element offset e of psi
template <int kT>
__device__ __forceinline__ float zcpml(const FusedParams& f, float d, int q, uint32_t e, bool own) {
  const int L2 = 2 * f.L;
  const float* tab = f.tab + kT * 3 * L2;
  const float a = tab[q], b = tab[L2 + q], ik = tab[2 * L2 + q];
  if (f.psi_bf16) {
    const bf16_t* pin = static_cast<const bf16_t*>(f.psi[kT]);
    const float psi = b * load_f(pin, e) + a * d;
    if (own) store_f(const_cast<bf16_t*>(pin) + f.psi_delta, e, psi);
    return d * ik + psi;
  }
  const float* pin = static_cast<const float*>(f.psi[kT]);
  const float psi = b * ldF(pin, e) + a * d;
  if (own) *reinterpret_cast<float*>(reinterpret_cast<char*>(const_cast<float*>(pin)) +
                                     (static_cast<int64_t>(e) + f.psi_delta) * 4) = psi;
  return d * ik + psi;
}

template <typename T, int kTY, bool kEArr>
__global__ void __launch_bounds__(64) fused_fast(FusedParams f) {
  const int lane = threadIdx.x;
  const int g = xcd_remap(blockIdx.x, gridDim.x);
  int b = g;
  const int bx = f.box_tx0 + b % f.box_ntx;
  b /= f.box_ntx;
  const int by = f.box_ty0 + b % f.box_nty;
  const int bz = b / f.box_nty;
  const int kz0 = f.box_z0 + bz * f.box_kz;
  const int kz1 = min(kz0 + f.box_kz, f.box_z1);

  const int nz = f.nz, ny = f.ny, nx = f.nx, L = f.L, m = f.metallic;
  const float idx = f.inv_dx;
  const int j0 = by * kTY;
  const int x = bx * 62 - 1 + lane;
  const bool own = lane >= 1 && lane <= 62;

  // element geometry of the four layout classes (plane stride, row stride)
  //   X: Ex, Hy, e_source_x    (ny+1) x nx
  //   Y: Ey, Hx, e_source_y     ny    x (nx+1)
  //   Z: Ez, e_source_z        (ny+1) x (nx+1)
  //   W: Hz                     ny    x nx
  const uint32_t rX = nx, rY = nx + 1, rZ = nx + 1, rW = nx;
  const uint32_t pX = (ny + 1) * nx, pY = ny * (nx + 1), pZ = (ny + 1) * (nx + 1), pW = ny * nx;
  const uint32_t y0 = j0 - 1;
  const uint32_t bX = y0 * rX + x, bY = y0 * rY + x, bZ = y0 * rZ + x, bW = y0 * rW + x;

  float hc[6], ec[6];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    hc[q] = f.hdec[q][0];
    hc[3 + q] = f.hsrc[q][0] * f.ratio_h;
    ec[q] = f.edec[q][0];
    ec[3 + q] = kEArr ? 0.0f : f.esrc[q][0] * f.ratio_e;
  }
  const int dn = ((lane + 1) & 63) << 2;
  const int up = ((lane - 1) & 63) << 2;

  float ex0[kTY + 2], ey0[kTY + 2], ez0[kTY + 2];
  float ex1[kTY + 2], ey1[kTY + 2], ez1[kTY + 2];
  float phx[kTY], phy[kTY];

  auto load_plane = [&](int k, float* ex, float* ey, float* ez) {
    // Ex/Ey exist on planes 0..nz, Ez on 0..nz-1; a missing plane reads as 0
    const bool vxy = k <= nz, vz = k < nz;
    const uint32_t kx = min(k, nz), kz = min(k, nz - 1);
    const uint32_t oX = kx * pX + bX, oY = kx * pY + bY, oZ = kz * pZ + bZ;
#pragma unroll
    for (int r = 0; r < kTY + 2; ++r) {
      const float vx = ldT<T>(f.e_in[0], oX + r * rX);
      const float vy = ldT<T>(f.e_in[1], oY + r * rY);
      const float vzz = ldT<T>(f.e_in[2], oZ + r * rZ);
      ex[r] = vxy ? vx : 0.0f;
      ey[r] = vxy ? vy : 0.0f;
      ez[r] = vz ? vzz : 0.0f;
    }
  };

  // H_new(k) on rows 0..kTY (row 0 = y0 halo) at lanes 0..62
  auto h_rows = [&](int k, const float* ex, const float* ey, const float* ez, const float* exn,
                    const float* eyn, float* hx, float* hy, float* hz, bool write) {
    const bool vxy = k < nz;  // Hx, Hy exist on planes 0..nz-1
    const uint32_t kk = min(k, nz - 1);
    const uint32_t oY = kk * pY + bY, oX = kk * pX + bX, oW = k * pW + bW;
    float ha[3][kTY + 1];
#pragma unroll
    for (int r = 0; r <= kTY; ++r) {
      ha[0][r] = ldT<T>(f.h_in[0], oY + r * rY);
      ha[1][r] = ldT<T>(f.h_in[1], oX + r * rX);
      ha[2][r] = ldT<T>(f.h_in[2], oW + r * rW);
    }
    const int qz = L > 0 && vxy ? slab_q(k, nz, L) : -1;
    const bool hz_pec = ((m & kFront) && k == 0) || ((m & kBack) && k == nz);
#pragma unroll
    for (int r = 0; r <= kTY; ++r) {
      const bool w = write && own && r >= 1;
      float d0 = (ez[r + 1] - ez[r]) * idx;
      float d1 = (eyn[r] - ey[r]) * idx;
      float d2 = (exn[r] - ex[r]) * idx;
      float d3 = (from_right(ez[r], dn) - ez[r]) * idx;
      float d4 = (from_right(ey[r], dn) - ey[r]) * idx;
      float d5 = (ex[r + 1] - ex[r]) * idx;
      if (qz >= 0) {
        d1 = zcpml<1>(f, d1, qz, qz * pY + bY + r * rY, w);
        d2 = zcpml<2>(f, d2, qz, qz * pX + bX + r * rX, w);
      }
      hx[r] = vxy ? hc[0] * ha[0][r] - hc[3] * (d0 - d1) : 0.0f;
      hy[r] = vxy ? hc[1] * ha[1][r] - hc[4] * (d2 - d3) : 0.0f;
      hz[r] = hz_pec ? 0.0f : hc[2] * ha[2][r] - hc[5] * (d4 - d5);
    }
  };

#pragma unroll
  for (int r = 0; r < kTY; ++r) phx[r] = phy[r] = 0.0f;
  if (kz0 > 0) {
    float hx[kTY + 1], hy[kTY + 1], hz[kTY + 1];
    load_plane(kz0 - 1, ex1, ey1, ez1);
    load_plane(kz0, ex0, ey0, ez0);
    h_rows(kz0 - 1, ex1, ey1, ez1, ex0, ey0, hx, hy, hz, false);
#pragma unroll
    for (int r = 0; r < kTY; ++r) {
      phx[r] = hx[r + 1];
      phy[r] = hy[r + 1];
    }
  } else {
    load_plane(kz0, ex0, ey0, ez0);
  }

  for (int k = kz0; k < kz1; ++k) {
    load_plane(k + 1, ex1, ey1, ez1);
    float es[3][kTY];
    if constexpr (kEArr) {
      const uint32_t kx = min(k, nz), kz = min(k, nz - 1);
      const uint32_t oX = kx * pX + bX + rX, oY = kx * pY + bY + rY, oZ = kz * pZ + bZ + rZ;
#pragma unroll
      for (int r = 0; r < kTY; ++r) {
        es[0][r] = ldF(f.esrc[0], oX + r * rX) * f.ratio_e;
        es[1][r] = ldF(f.esrc[1], oY + r * rY) * f.ratio_e;
        es[2][r] = ldF(f.esrc[2], oZ + r * rZ) * f.ratio_e;
      }
    } else {
#pragma unroll
      for (int r = 0; r < kTY; ++r) {
        es[0][r] = ec[3];
        es[1][r] = ec[4];
        es[2][r] = ec[5];
      }
    }
    float hx[kTY + 1], hy[kTY + 1], hz[kTY + 1];
    h_rows(k, ex0, ey0, ez0, ex1, ey1, hx, hy, hz, true);

    const bool vxy = k < nz;
    const uint32_t kk = min(k, nz - 1);
    const uint32_t hoY = kk * pY + bY, hoX = kk * pX + bX, hoW = k * pW + bW;
    if (own) {
#pragma unroll
      for (int r = 1; r <= kTY; ++r) {
        if (vxy) {
          stT<T>(f.h_out[0], hoY + r * rY, hx[r]);
          stT<T>(f.h_out[1], hoX + r * rX, hy[r]);
        }
        stT<T>(f.h_out[2], hoW + r * rW, hz[r]);
      }
    }
    // E_new(k): z ghost rule and z-face PEC are plane-uniform
    const int qz = L > 0 ? slab_q(k, nz + 1, L) : -1;
    const bool exy_pec = ((m & kFront) && k == 0) || ((m & kBack) && k == nz);
    const bool mlo = m & kFront, mhi = m & kBack;
    const uint32_t kx = min(k, nz), kz = min(k, nz - 1);
    const uint32_t eoX = kx * pX + bX, eoY = kx * pY + bY, eoZ = kz * pZ + bZ;
#pragma unroll
    for (int r = 1; r <= kTY; ++r) {
      const float hz_l = from_left(hz[r], up);
      const float hy_l = from_left(hy[r], up);
      float d0 = (hz[r] - hz[r - 1]) * idx;
      float d1 = gdiff(hy[r], phy[r - 1], k, nz, mlo, mhi) * idx;
      float d2 = gdiff(hx[r], phx[r - 1], k, nz, mlo, mhi) * idx;
      float d3 = (hz[r] - hz_l) * idx;
      float d4 = (hy[r] - hy_l) * idx;
      float d5 = (hx[r] - hx[r - 1]) * idx;
      if (qz >= 0) {
        d1 = zcpml<7>(f, d1, qz, qz * pX + bX + r * rX, own);
        d2 = zcpml<8>(f, d2, qz, qz * pY + bY + r * rY, own);
      }
      const float nex = exy_pec ? 0.0f : ec[0] * ex0[r] + es[0][r - 1] * (d0 - d1);
      const float ney = exy_pec ? 0.0f : ec[1] * ey0[r] + es[1][r - 1] * (d2 - d3);
      const float nez = ec[2] * ez0[r] + es[2][r - 1] * (d4 - d5);
      if (own) {
        stT<T>(f.e_out[0], eoX + r * rX, nex);
        stT<T>(f.e_out[1], eoY + r * rY, ney);
        if (k < nz) stT<T>(f.e_out[2], eoZ + r * rZ, nez);
      }
    }
#pragma unroll
    for (int r = 0; r < kTY; ++r) {
      phx[r] = hx[r + 1];
      phy[r] = hy[r + 1];
    }
#pragma unroll
    for (int r = 0; r < kTY + 2; ++r) {
      ex0[r] = ex1[r];
      ey0[r] = ey1[r];
      ez0[r] = ez1[r];
    }
  }
}

template <typename T, int kR>
void dispatch_rows(hipStream_t s, const FusedParams& p, int nblocks) {
  dim3 block(64, 1, 1);
  dim3 grid(nblocks, 1, 1);
  if (p.tiles == nullptr) {
    bool earr = p.esrc_st[0] && p.esrc_st[1] && p.esrc_st[2];
    if (earr) hipLaunchKernelGGL((fused_fast<T, kR, true>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((fused_fast<T, kR, false>), grid, block, 0, s, p);
  } else if (p.cpml_kernel) {
    hipLaunchKernelGGL((fused_step<T, kR, true>), grid, block, 0, s, p);
  } else {
    hipLaunchKernelGGL((fused_step<T, kR, false>), grid, block, 0, s, p);
  }
}

// Rows per wave (y extent of a tile): 2, 4 or 8; BZH_FUSED_ROWS overrides the
// default for tuning sweeps.
int rows_setting() {
  int r = 4;
  if (const char* env = std::getenv("BZH_FUSED_ROWS")) r = std::atoi(env);
  return (r == 2 || r == 4 || r == 8) ? r : 4;
}

template <typename T>
void dispatch(hipStream_t s, const FusedParams& p, int nblocks) {
  switch (rows_setting()) {
    case 2: dispatch_rows<T, 2>(s, p, nblocks); break;
    case 8: dispatch_rows<T, 8>(s, p, nblocks); break;
    default: dispatch_rows<T, 4>(s, p, nblocks); break;
  }
}

}  // namespace

int fused_rows() { return rows_setting(); }

void launch_fused3d(hipStream_t s, const FusedParams& p, int dtype) {
  int nblocks;
  if (p.tiles != nullptr) {
    nblocks = p.ntiles;
  } else {
    BZH_REQUIRE(p.box_ntx >= 0 && p.box_nty >= 0 && p.box_kz >= 1, "fused box is malformed");
    BZH_REQUIRE(p.box_tx0 >= 0 && (p.box_tx0 + p.box_ntx - 1) * 62 <= p.nx + 62 &&
                    p.box_ty0 >= 0 && p.box_z0 >= 0 && p.box_z1 <= p.nz + 1,
                "fused box exceeds the grid");
    nblocks = p.box_ntx * p.box_nty * p.box_nzc;
  }
  BZH_REQUIRE(p.cpml_kernel || p.L == 0, "CPML tiles need the CPML-capable kernel");
  if (nblocks <= 0) return;
  switch (dtype) {
    case BZH_F32: dispatch<float>(s, p, nblocks); break;
    case BZH_BF16: dispatch<bf16_t>(s, p, nblocks); break;
    case BZH_F16: dispatch<__half>(s, p, nblocks); break;
    default: throw Failure{"unsupported field dtype"};
  }
  BZH_HIP(hipGetLastError());
}

}  // namespace bzh
