// Fused full leapfrog step (H then E) for gfx950, two field banks.
//
// One launch reads bank A (E, H, CPML memories) and writes bank B, so no
// workgroup ever reads a value another is overwriting.
//
// Mapping: a workgroup of kR + 1 waves owns a tile of 62 (x) by kR (y) cells
// and marches along z over a chunk of planes.  Lane l of every wave sits at
// x = i0 - 1 + l: lanes 1..62 own outputs, lanes 0 and 63 are halo columns.
// Wave w sits on row y = j0 - 1 + w: waves 1..kR own rows j0 .. j0+kR-1,
// wave 0 computes the halo row j0-1 of H (Hx, Hz, needed by the E update of
// row j0) and loads the top halo row j0+kR of E (Ez, Ex, needed by the H
// update of row j0+kR-1).  Per plane k each wave
//
//   issues one batch of loads: E(k+1), H_A(k), E coefficients and the CPML
//     memories of every active term of its row,
//   computes H_new(k) on its row (E(k) of the row above comes from LDS, x+1
//     neighbours by cross-lane permute),
//   publishes H_new(k) and E(k+1) of its row to LDS, stores owned H,
//   barrier,
//   computes E_new(k) on its row from H_new(k) (row below from LDS, x-1 by
//     permute) and H_new(k-1) (carried in registers along z), stores owned E.
//
// Rows cross through double-buffered LDS with a single barrier per plane.
// Boundary handling costs nothing where it does not apply: everything that
// depends on z (faces, ghost rule, CPML z-terms) is a plane-uniform branch,
// everything that depends on y a wave-uniform branch, and only the tiles that
// touch an x face or an x CPML slab run the kXEdge variant with per-lane x
// checks.  Addresses are an array base (scalar) plus a 32-bit byte offset
// (global_load v, v_off, s[base]); a group of 4 GiB or more takes the wide
// variants, which form 64-bit offsets.
//
// Numerical contract: identical to phase.hip (kernels.py:458-626,
// lattice.py:523-575, kernels.py:184-264 with per-axis slab extents at the
// low and high faces, _boundary_compile.py:118-160); arithmetic is float32.

#ifndef BZH_FUSED_KERNELS_H
#define BZH_FUSED_KERNELS_H

#include <cstdlib>
#include <string>
#include <type_traits>

#include "common.h"

namespace bzh {
namespace {

// Constant-address-space view: uniform indices become scalar (s_load) reads.
using cfloat = const __attribute__((address_space(4))) float;
__device__ __forceinline__ cfloat* as_const(const float* p) { return (cfloat*)(p); }

// Value of the lane to the right / left (x + 1 / x - 1).
__device__ __forceinline__ float lane_shift(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// Ghost-rule backward difference from loaded neighbours (lattice.py:523-575):
// ``cur`` is the sample at c, ``prev`` at c - 1, N the sample count.
__device__ __forceinline__ float gdiff(float cur, float prev, int c, int N, bool mlo, bool mhi) {
  const float up = (c < N) ? cur : (mhi ? 0.0f : prev);
  const float lo = (c > 0) ? prev : (mlo ? 0.0f : cur);
  return up - lo;
}

// Packed slab coordinate of sample c on an axis of n samples whose CPML
// memory holds the lo samples at the low face and the hi samples at the high
// face (kernels.py:184-222); -1 outside the slabs.  Requires 0 <= c < n.
__device__ __forceinline__ int slab_q(int c, int n, int lo, int hi) {
  return c < lo ? c : (c >= n - hi ? c - (n - lo - hi) : -1);
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

// Work item of this workgroup.  Without bands (box_band == 0) the logical
// index is x tile fastest, then y tile, then z chunk, dealt to the XCDs in
// contiguous ranges.  With bands, the (x, y) tiles are cut into 8 bands of
// box_band tiles and XCD c walks band c chunk by chunk, so at any time all
// XCDs stream the same z chunk of the grid (one contiguous region per array)
// while tiles that share halo rows stay in one L2.  Indices past the last
// tile of a band are padding; their workgroups return at once.
__device__ __forceinline__ bool decode_tile(const FusedParams& f, int& bx, int& by, int& kz0,
                                            int& kz1) {
  const int lam = xcd_remap(blockIdx.x, gridDim.x);
  int t, zc;
  if (f.box_band > 0) {
    const int per = f.box_band * f.box_nzc;
    const int c = lam / per, rem = lam - c * per;
    zc = rem / f.box_band;
    t = c * f.box_band + (rem - zc * f.box_band);
    if (t >= f.box_ntx * f.box_nty) return false;
  } else {
    t = lam % (f.box_ntx * f.box_nty);
    zc = lam / (f.box_ntx * f.box_nty);
  }
  bx = t % f.box_ntx;
  by = f.box_ty0 + t / f.box_ntx;
  if (bx >= f.box_tsplit) bx += f.box_tskip;
  bx += f.box_tx0;
  kz0 = f.box_z0 + zc * f.box_kz;
  kz1 = min(kz0 + f.box_kz, f.box_z1);
  return true;
}

// Work item of block ``blk`` of a pair box (see PairBox).  Indices past the
// last tile of a band, or past the last chunk, are padding: false.
__device__ __forceinline__ bool decode_pair_box(const PairBox& b, int blk, int& bx, int& by,
                                                int& kz0, int& kz1) {
  const int lam = xcd_remap(blk, b.nblk);
  const int tiles = b.ntx * b.nty;
  int t, zc;
  if (b.band > 0) {
    const int per = b.band * b.nzc;
    const int c = lam / per, rem = lam - c * per;
    zc = rem / b.band;
    t = c * b.band + (rem - zc * b.band);
    if (t >= tiles) return false;
  } else {
    t = lam % tiles;
    zc = lam / tiles;
    if (zc >= b.nzc) return false;
  }
  int tx = t % b.ntx, ty = t / b.ntx;
  if (tx >= b.tsplit) tx += b.tskip;
  if (ty >= b.tysplit) ty += b.tyskip;
  bx = b.tx0 + tx;
  by = b.ty0 + ty;
  if (zc < b.nzc_lo) {
    kz0 = b.z0 + zc * b.kz;
    kz1 = min(kz0 + b.kz, b.zsplit);
  } else {
    kz0 = b.zhi + (zc - b.nzc_lo) * b.kz;
    kz1 = min(kz0 + b.kz, b.z1);
  }
  return true;
}

struct Coef {
  float a, b, ik;
};
// Term t's recurrence coefficients at packed coordinate q; tab is [12][3][st].
__device__ __forceinline__ Coef coef_u(cfloat* tab, int t, int st, int q) {
  return {tab[(t * 3 + 0) * st + q], tab[(t * 3 + 1) * st + q], tab[(t * 3 + 2) * st + q]};
}
__device__ __forceinline__ Coef coef_v(const float* tab, int t, int st, int q) {
  return {tab[(t * 3 + 0) * st + q], tab[(t * 3 + 1) * st + q], tab[(t * 3 + 2) * st + q]};
}
// psi' = b psi + a d;  returns d / kappa + psi'  (kernels.py:237-264)
__device__ __forceinline__ float cpml_step(const Coef& c, float d, float psi, float& out) {
  out = c.b * psi + c.a * d;
  return d * c.ik + out;
}

// Global-address-space views: the row base stays a scalar pair and the load
// takes the ``global_load v, v_off, s[base]`` form.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

// Group base (kernel argument, uniform) plus a byte offset formed next to
// the access: ``aoff + row * sizeof(T) + xb`` with the array's offset in its
// group, the uniform row start and the per-lane byte offset xb.  O is the
// offset type: uint32_t when every group spans less than 4 GiB (checked on
// the host), so that the access selects ``global_load v, v_off, s[base]``
// with one scalar pair per group; uint64_t in the wide variants (banks of
// 4 GiB or more, each array under 4 GiB).
template <typename T, typename O>
__device__ __forceinline__ gptr<const char> at(const char* base, O aoff, uint32_t row,
                                               uint32_t xb) {
  const O off = aoff + static_cast<O>(row) * static_cast<O>(sizeof(T)) + static_cast<O>(xb);
  return (gptr<const char>)(base) + off;
}
__device__ __forceinline__ float ld_g16(gptr<const char> p, bool bf) {
  const uint16_t u = *reinterpret_cast<gptr<const uint16_t>>(p);
  return bf ? __uint_as_float(static_cast<uint32_t>(u) << 16) : __half2float(__ushort_as_half(u));
}
template <typename T, typename O>
__device__ __forceinline__ float ldu(const char* base, O aoff, uint32_t row, uint32_t xb) {
  gptr<const char> p = at<T, O>(base, aoff, row, xb);
  if constexpr (sizeof(T) == 4) return *reinterpret_cast<gptr<const float>>(p);
  else return ld_g16(p, std::is_same<T, bf16_t>::value);
}
template <typename T, typename O>
__device__ __forceinline__ void stu(const char* base, O aoff, uint32_t row, uint32_t xb,
                                    float v) {
  gptr<char> p = (gptr<char>)(at<T, O>(base, aoff, row, xb));
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<gptr<float>>(p) = v;
  } else {
    T tmp;
    store_f(&tmp, 0, v);
    uint16_t u;
    __builtin_memcpy(&u, &tmp, 2);
    *reinterpret_cast<gptr<uint16_t>>(p) = u;
  }
}
// Access at ``base + soff + voff``: a uniform byte offset (plane and array)
// plus a per-lane byte offset (row and column) held in a vector register.
template <typename T, typename O>
__device__ __forceinline__ float ldo(const char* base, O soff, uint32_t voff) {
  return ldu<T, O>(base, soff, 0u, voff);
}

// The raw bits of one element at ``base + soff + voff`` (zero-extended), for
// a load whose value is needed only an iteration later: the conversion to
// float32 (``from_raw``) happens at the use, so the load does not wait here.
template <typename T, typename O>
__device__ __forceinline__ uint32_t ldraw(const char* base, O soff, uint32_t voff) {
  gptr<const char> p = at<T, O>(base, soff, 0u, voff);
  if constexpr (sizeof(T) == 4) return *reinterpret_cast<gptr<const uint32_t>>(p);
  else return *reinterpret_cast<gptr<const uint16_t>>(p);
}
template <typename T>
__device__ __forceinline__ float from_raw(uint32_t u) {
  if constexpr (sizeof(T) == 4) return __uint_as_float(u);
  else if constexpr (std::is_same<T, bf16_t>::value) return __uint_as_float(u << 16);
  else return __half2float(__ushort_as_half(static_cast<unsigned short>(u)));
}

// One dword per lane from ``base + off`` (global) into an LDS row (lane l at
// row[l]), asynchronously: the wave goes on, and the compiler waits for the
// transfer before the row is read.
template <typename O>
__device__ __forceinline__ void dma_row(const char* base, O soff, uint32_t voff, float* row) {
  __builtin_amdgcn_global_load_lds(
      (__attribute__((address_space(1))) void*)((gptr<const char>)(base) + (soff + static_cast<O>(voff))),
      (__attribute__((address_space(3))) void*)(row), 4, 0, 0);
}

// Field store: bfloat16 fields take stochastic rounding when ``sr`` (uniform)
// is set, every other storage type the ordinary conversion.
template <typename T, typename O>
__device__ __forceinline__ void stf(const char* base, O aoff, uint32_t row, uint32_t xb,
                                    float v, bool sr, SrState& rng) {
  if constexpr (std::is_same<T, bf16_t>::value) {
    if (sr) {
      gptr<char> p = (gptr<char>)(at<T, O>(base, aoff, row, xb));
      *reinterpret_cast<gptr<uint16_t>>(p) = bf16_bits_sr(v, rng);
      return;
    }
  }
  stu<T, O>(base, aoff, row, xb, v);
}
// The seed of a launch's rounding sequence: the lane's global index and the
// absolute time step (*step0 + kstep, or kstep without a step counter).
__device__ __forceinline__ SrState sr_launch_seed(const FusedParams& f, int step_offset) {
  const int32_t s0 = f.step0 != nullptr ? *(const __attribute__((address_space(4))) int32_t*)(f.step0) : 0;
  return sr_seed(blockIdx.x * blockDim.x + threadIdx.x,
                 static_cast<uint32_t>(s0 + f.kstep + step_offset));
}
// The kernel argument block, re-read at a use: in the x-interior kernel for
// 2-byte fields (8 waves per SIMD, fewer scalar registers) the CPML memory
// offsets and the coefficient table, needed only on slab rows and planes, are
// reloaded through the scalar cache there instead of holding scalar registers
// across the z-march (interior 1.47 -> 1.39 ms at 1024x512x256 f16).  FP32
// and the x-edge kernels keep them resident (reloading measured 2.53 -> 2.56
// ms in FP32 and 0.33 -> 0.36 ms for the f16 x-edge box).
__device__ __forceinline__ __attribute__((address_space(4))) const FusedParams* kargs() {
  auto p = (__attribute__((address_space(4))) const FusedParams*)(
      __builtin_amdgcn_kernarg_segment_ptr());
  asm volatile("" : "+s"(p));
  return p;
}
template <bool kReload>
struct SlabArgs {
  const FusedParams& f;
  __device__ __forceinline__ const char* base() const {
    if constexpr (kReload) return kargs()->pbase;
    else return f.pbase;
  }
  __device__ __forceinline__ uint32_t in(int t) const {
    if constexpr (kReload) return kargs()->poff[t];
    else return f.poff[t];
  }
  __device__ __forceinline__ uint32_t out(int t) const {
    if constexpr (kReload) return kargs()->poff[t] + kargs()->pdelta;
    else return f.poff[t] + f.pdelta;
  }
  __device__ __forceinline__ cfloat* tab() const {
    if constexpr (kReload) return as_const(kargs()->tab);
    else return as_const(f.tab);
  }
};

// The pair kernel's argument block, re-read at a use (see kargs): its CPML
// offsets, coefficient table, source slabs and monitor boxes are needed only
// on some planes and rows, and holding them across the z-march spills scalar
// registers to vector lanes.
__device__ __forceinline__ __attribute__((address_space(4))) const PairParams* pair_kargs() {
  auto p = (__attribute__((address_space(4))) const PairParams*)(
      __builtin_amdgcn_kernarg_segment_ptr());
  asm volatile("" : "+s"(p));
  return p;
}

using cint = const __attribute__((address_space(4))) int32_t;
using kslab = const __attribute__((address_space(4))) MidSlab;

// Add one source slab to plane k of a row (execute.py:374-461): the slab's
// plane and row are uniform, its columns per lane (lane at x).  vx, vy, vz are
// the E or H components of the row, as the slab's ``hfield`` says.
__device__ __forceinline__ void slab_add(kslab& ms, int step, int k, int y, int x, int nz, int ny,
                                         int nx, float& vx, float& vy, float& vz) {
  cint* st = (cint*)(ms.start);
  const int cz = k - st[0], cy = y - st[1], cx = x - st[2];
  const int comp = ms.comp;
  int Dz, Dy, Dx;  // the component's extent
  if (ms.hfield) {
    Dz = comp == 2 ? nz + 1 : nz;
    Dy = comp == 1 ? ny + 1 : ny;
    Dx = comp == 0 ? nx + 1 : nx;
  } else {
    Dz = comp == 2 ? nz : nz + 1;
    Dy = comp == 1 ? ny : ny + 1;
    Dx = comp == 0 ? nx : nx + 1;
  }
  if (cz < 0 || cz >= ms.sz || cy < 0 || cy >= ms.sy || k >= Dz || y < 0 || y >= Dy) return;
  if (ms.post && ((ms.mlo[0] && k == 0) || (ms.mhi[0] && k == Dz - 1) || (ms.mlo[1] && y == 0) ||
                  (ms.mhi[1] && y == Dy - 1)))
    return;
  int widx = step + ms.kstep;
  widx = widx < 0 ? 0 : (widx > ms.T - 1 ? ms.T - 1 : widx);
  const float wv = as_const(ms.wave)[widx];
  const bool in = cx >= 0 && cx < ms.sx && x >= 0 && x < Dx &&
                  !(ms.post && ((ms.mlo[2] && x == 0) || (ms.mhi[2] && x == Dx - 1)));
  float add = 0.0f;
  if (in) add = ((const float*)ms.coeff)[(cz * ms.sy + cy) * ms.sx + cx] * wv;
  // by value (a component picked through the references is indexed in
  // scratch memory); adding 0 leaves the other two unchanged
  vx += comp == 0 ? add : 0.0f;
  vy += comp == 1 ? add : 0.0f;
  vz += comp == 2 ? add : 0.0f;
}

// Does slab ``ms`` cross row y on any lane of this wave (lane at x)?
__device__ __forceinline__ bool slab_hits_wave(kslab& ms, int y, int x) {
  cint* st = (cint*)(ms.start);
  const int cy = y - st[1], cx = x - st[2];
  const bool hit = cy >= 0 && cy < ms.sy && cx >= 0 && cx < ms.sx;
  return __builtin_amdgcn_readfirstlane(__any(hit)) != 0;
}

template <typename T, typename P, typename C, int kR, bool kEArr, bool kXEdge, bool kHSrc,
          bool kWide>
__device__ __forceinline__ void fused_body(const FusedParams& f) {
  const SlabArgs<sizeof(T) == 2 && !kXEdge> sa{f};
  // field and E coefficient offsets in their groups: 32-bit, 64-bit in the
  // wide variant (kWide: a bank of 4 GiB or more)
  auto fo = [&](int c) {
    if constexpr (kWide) return f.foff64[c];
    else return f.foff[c];
  };
  auto co = [&](int c) {
    if constexpr (kWide) return f.coff64[c];
    else return f.coff[c];
  };
  __shared__ float lE[2][kR + 2][2][64];
  __shared__ float lH[2][kR + 1][2][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool halo = w == 0;

  int bx, by, kz0, kz1;
  if (!decode_tile(f, bx, by, kz0, kz1)) return;  // padding workgroup

  const int nz = f.nz, ny = f.ny, nx = f.nx, L = f.L, TS = f.tab_st;
  // CPML memory extents along the slab axis (x and y; z slabs lead)
  const int n2xH = f.clo[kSlabH] + f.chi[kSlabH], n2yH = f.clo[kSlabH + 1] + f.chi[kSlabH + 1];
  const int n2xE = f.clo[kSlabE] + f.chi[kSlabE], n2yE = f.clo[kSlabE + 1] + f.chi[kSlabE + 1];
  const int m = f.metallic;
  const bool mf = m & kFront, mk = m & kBack, mb = m & kBottom, mt = m & kTop;
  const bool ml = m & kLeft, mr = m & kRight;
  const float idx = f.inv_dx;
  const int j0 = by * kR;
  const int x = bx * 62 - 1 + lane;
  const int dn = ((lane + 1) & 63) << 2;  // permute address of lane + 1
  const int up = ((lane - 1) & 63) << 2;  // ... of lane - 1

  // ---- x: per lane ----
  bool vxX = true, vxY = true, xface = false;
  int xX = x, xY = x, qxH = -1, qxE = -1;
  if constexpr (kXEdge) {
    vxX = x >= 0 && x < nx;
    vxY = x >= 0 && x <= nx;
    xX = min(max(x, 0), nx - 1);
    xY = min(max(x, 0), nx);
    qxH = vxX ? slab_q(x, nx, f.clo[kSlabH], f.chi[kSlabH]) : -1;
    qxE = vxY ? slab_q(x, nx + 1, f.clo[kSlabE], f.chi[kSlabE]) : -1;
    xface = (ml && x == 0) || (mr && x == nx);
  }
  const bool own = !halo && lane >= 1 && lane <= 62 && (!kXEdge || x <= nx);
  const uint32_t tX = xX * sizeof(T), tY = xY * sizeof(T);
  const uint32_t pX = xX * sizeof(P), pY = xY * sizeof(P);
  const uint32_t cX = xX * sizeof(C), cY = xY * sizeof(C);
  const uint32_t qH = max(qxH, 0) * sizeof(P), qE = max(qxE, 0) * sizeof(P);
  Coef c3{}, c4{}, c9{}, c10{};
  if constexpr (kXEdge) {
    if (qxH >= 0) {
      c3 = coef_v(f.tab, 3, TS, qxH);
      c4 = coef_v(f.tab, 4, TS, qxH);
    }
    if (qxE >= 0) {
      c9 = coef_v(f.tab, 9, TS, qxE);
      c10 = coef_v(f.tab, 10, TS, qxE);
    }
  }

  // ---- y: this wave's row (uniform) ----
  const int y = j0 - 1 + w;
  const bool vyX = y >= 0 && y <= ny, vyY = y >= 0 && y < ny;
  const int yX = min(max(y, 0), ny), yY = min(max(y, 0), ny - 1);
  const int qyH = vyY ? slab_q(y, ny, f.clo[kSlabH + 1], f.chi[kSlabH + 1]) : -1;      // Hx_y, Hz_y: ny samples
  const int qyE = vyX ? slab_q(y, ny + 1, f.clo[kSlabE + 1], f.chi[kSlabE + 1]) : -1;  // Ex_y, Ez_y: ny + 1 samples
  const bool y_pec_h = (mb && y == 0) || (mt && y == ny);  // Hy normal to a y face
  // top halo row of E, loaded by wave 0
  const int yt = j0 + kR;
  const bool vtX = yt <= ny;
  const int ytX = min(yt, ny);

  float hc[6], ec[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    hc[q] = as_const(f.hdec[q])[0];
    hc[3 + q] = as_const(f.hsrc[q])[0] * f.ratio_h;
    ec[q] = as_const(f.edec[q])[0];
  }
  float es0 = 0.0f, es1 = 0.0f, es2 = 0.0f;
  if constexpr (!kEArr) {
    es0 = as_const(f.esrc[0])[0] * f.ratio_e;
    es1 = as_const(f.esrc[1])[0] * f.ratio_e;
    es2 = as_const(f.esrc[2])[0] * f.ratio_e;
  }

  // E(k) on this wave's row; Ex/Ey exist on planes 0..nz, Ez on 0..nz-1
  auto load_e = [&](int k, float& ex, float& ey, float& ez) {
    const bool vxy = k <= nz, vz = k < nz;
    const uint32_t kx = min(k, nz), kz = min(k, nz - 1);
    const float a = ldu<T>(f.fin, fo(0), (kx * (ny + 1) + yX) * nx, tX);
    const float c = ldu<T>(f.fin, fo(1), (kx * ny + yY) * (nx + 1), tY);
    const float d = ldu<T>(f.fin, fo(2), (kz * (ny + 1) + yX) * (nx + 1), tY);
    ex = (vxy && vyX && vxX) ? a : 0.0f;
    ey = (vxy && vyY && vxY) ? c : 0.0f;
    ez = (vz && vyX && vxY) ? d : 0.0f;
  };
  auto load_top = [&](int k, float& ex, float& ez) {
    const bool vxy = k <= nz, vz = k < nz;
    const uint32_t kx = min(k, nz), kz = min(k, nz - 1);
    const float a = ldu<T>(f.fin, fo(0), (kx * (ny + 1) + ytX) * nx, tX);
    const float d = ldu<T>(f.fin, fo(2), (kz * (ny + 1) + ytX) * (nx + 1), tY);
    ex = (vxy && vtX && vxX) ? a : 0.0f;
    ez = (vz && vtX && vxY) ? d : 0.0f;
  };

  // stochastic rounding of bfloat16 field stores (uniform switch)
  const bool sr = std::is_same<T, bf16_t>::value && f.bf16_sr != 0;
  SrState rng{0u};
  if (sr) rng = sr_launch_seed(f, 0);

  bool hwave = false;  // an H-timing slab crosses this wave's row
  if constexpr (kHSrc)
    for (int s = 0; s < f.hslab_n; ++s) hwave = hwave || slab_hits_wave(kargs()->hslab[s], y, x);

  const int ks = kz0 > 0 ? kz0 - 1 : 0;  // plane kz0-1 only rebuilds H_new(kz0-1)
  float ex, ey, ez;
  load_e(ks, ex, ey, ez);
  lE[ks & 1][w][0][lane] = ez;
  lE[ks & 1][w][1][lane] = ex;
  if (halo) {
    float tx_, tz_;
    load_top(ks, tx_, tz_);
    lE[ks & 1][kR + 1][0][lane] = tz_;
    lE[ks & 1][kR + 1][1][lane] = tx_;
  }
  __syncthreads();

  float phx = 0.0f, phy = 0.0f;  // H_new(k - 1) on this row
  for (int k = ks; k < kz1; ++k) {
    const int bb = k & 1;
    const bool st = k >= kz0;  // this chunk owns plane k
    const bool vxy = k < nz;   // Hx/Hy planes
    const uint32_t kk = min(k, nz - 1);
    const uint32_t kx = min(k, nz), kzc = min(k, nz - 1);
    const int qzH = vxy ? slab_q(k, nz, f.clo[kSlabH + 2], f.chi[kSlabH + 2]) : -1;
    const int qzE = slab_q(k, nz + 1, f.clo[kSlabE + 2], f.chi[kSlabE + 2]);

    // ---- one batch of loads: E(k+1), H_A(k), coefficients, CPML memories ----
    float exn, eyn, ezn, etx = 0.0f, etz = 0.0f;
    load_e(k + 1, exn, eyn, ezn);
    if (halo) load_top(k + 1, etx, etz);
    const float ha0 = ldu<T>(f.fin, fo(3), (kk * ny + yY) * (nx + 1), tY);
    const float ha1 = halo ? 0.0f : ldu<T>(f.fin, fo(4), (kk * (ny + 1) + yX) * nx, tX);
    const float ha2 = ldu<T>(f.fin, fo(5), (k * ny + yY) * nx, tX);
    float p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0;
    if (qyH >= 0) {
      p0 = ldu<P>(sa.base(), sa.in(0), (kk * n2yH + qyH) * (nx + 1), pY);
      p5 = ldu<P>(sa.base(), sa.in(5), (k * n2yH + qyH) * nx, pX);
    }
    if (qzH >= 0) {
      p1 = ldu<P>(sa.base(), sa.in(1), (qzH * ny + yY) * (nx + 1), pY);
      if (!halo) p2 = ldu<P>(sa.base(), sa.in(2), (qzH * (ny + 1) + yX) * nx, pX);
    }
    if constexpr (kXEdge) {
      if (qxH >= 0) {
        if (!halo) p3 = ldu<P>(sa.base(), sa.in(3), (kk * (ny + 1) + yX) * n2xH, qH);
        p4 = ldu<P>(sa.base(), sa.in(4), (k * ny + yY) * n2xH, qH);
      }
    }
    float p6 = 0, p7 = 0, p8 = 0, p9 = 0, p10 = 0, p11 = 0;
    if (!halo && st) {
      if constexpr (kEArr) {
        es0 = ldu<C>(f.cbase, co(0), (kx * (ny + 1) + yX) * nx, cX) * f.ratio_e;
        es1 = ldu<C>(f.cbase, co(1), (kx * ny + yY) * (nx + 1), cY) * f.ratio_e;
        es2 = ldu<C>(f.cbase, co(2), (kzc * (ny + 1) + yX) * (nx + 1), cY) * f.ratio_e;
      }
      if (qyE >= 0) {
        p6 = ldu<P>(sa.base(), sa.in(6), (kx * n2yE + qyE) * nx, pX);
        p11 = ldu<P>(sa.base(), sa.in(11), (kzc * n2yE + qyE) * (nx + 1), pY);
      }
      if (qzE >= 0) {
        p7 = ldu<P>(sa.base(), sa.in(7), (qzE * (ny + 1) + yX) * nx, pX);
        p8 = ldu<P>(sa.base(), sa.in(8), (qzE * ny + yY) * (nx + 1), pY);
      }
      if constexpr (kXEdge) {
        if (qxE >= 0) {
          p9 = ldu<P>(sa.base(), sa.in(9), (kx * ny + yY) * n2xE, qE);
          p10 = ldu<P>(sa.base(), sa.in(10), (kzc * (ny + 1) + yX) * n2xE, qE);
        }
      }
    }
    const float ez_up = lE[bb][w + 1][0][lane];
    const float ex_up = lE[bb][w + 1][1][lane];

    // ---- H_new(k) on this row ----
    float d0 = (ez_up - ez) * idx;                   // dEz/dy
    float d1 = (eyn - ey) * idx;                     // dEy/dz
    float d2 = (exn - ex) * idx;                     // dEx/dz
    float d3 = (lane_shift(ez, dn) - ez) * idx;      // dEz/dx
    float d4 = (lane_shift(ey, dn) - ey) * idx;      // dEy/dx
    float d5 = (ex_up - ex) * idx;                   // dEx/dy
    const bool okhx = vxy && vyY && vxY;
    const bool okhy = vxy && vyX && vxX;
    const bool okhz = vyY && vxX;
    const bool wh = own && st;
    float n;
    if (qyH >= 0) {
      d0 = cpml_step(coef_u(sa.tab(), 0, TS, qyH), d0, p0, n);
      if (wh && okhx) stu<P>(sa.base(), sa.out(0), (k * n2yH + qyH) * (nx + 1), pY, n);
      d5 = cpml_step(coef_u(sa.tab(), 5, TS, qyH), d5, p5, n);
      if (wh && okhz) stu<P>(sa.base(), sa.out(5), (k * n2yH + qyH) * nx, pX, n);
    }
    if (qzH >= 0) {
      d1 = cpml_step(coef_u(sa.tab(), 1, TS, qzH), d1, p1, n);
      if (wh && okhx) stu<P>(sa.base(), sa.out(1), (qzH * ny + yY) * (nx + 1), pY, n);
      d2 = cpml_step(coef_u(sa.tab(), 2, TS, qzH), d2, p2, n);
      if (wh && okhy) stu<P>(sa.base(), sa.out(2), (qzH * (ny + 1) + yX) * nx, pX, n);
    }
    if constexpr (kXEdge) {
      if (qxH >= 0) {
        d3 = cpml_step(c3, d3, p3, n);
        if (wh && okhy) stu<P>(sa.base(), sa.out(3), (k * (ny + 1) + yX) * n2xH, qH, n);
        d4 = cpml_step(c4, d4, p4, n);
        if (wh && okhz) stu<P>(sa.base(), sa.out(4), (k * ny + yY) * n2xH, qH, n);
      }
    }
    float hx = hc[0] * ha0 - hc[3] * (d0 - d1);
    float hy = hc[1] * ha1 - hc[4] * (d2 - d3);
    float hz = hc[2] * ha2 - hc[5] * (d4 - d5);
    if (kXEdge && xface) hx = 0.0f;
    if (y_pec_h) hy = 0.0f;
    if ((mf && k == 0) || (mk && k == nz)) hz = 0.0f;
    hx = okhx ? hx : 0.0f;
    hy = okhy ? hy : 0.0f;
    hz = okhz ? hz : 0.0f;
    // H-timing source slabs of this step (execute.py:467-541: after the H
    // update, before E reads H)
    if (kHSrc && hwave) {
      const int step = *(cint*)(kargs()->step0);
      for (int s = 0; s < kargs()->hslab_n; ++s)
        slab_add(kargs()->hslab[s], step, k, y, x, nz, ny, nx, hx, hy, hz);
    }

    // publish H_new(k) and E(k+1) rows, store owned H
    lH[bb][w][0][lane] = hz;
    lH[bb][w][1][lane] = hx;
    lE[bb ^ 1][w][0][lane] = ezn;
    lE[bb ^ 1][w][1][lane] = exn;
    if (halo) {
      lE[bb ^ 1][kR + 1][0][lane] = etz;
      lE[bb ^ 1][kR + 1][1][lane] = etx;
    }
    if (wh) {
      if (okhx) stf<T>(f.fout, fo(3), (k * ny + yY) * (nx + 1), tY, hx, sr, rng);
      if (okhy) stf<T>(f.fout, fo(4), (k * (ny + 1) + yX) * nx, tX, hy, sr, rng);
      if (okhz) stf<T>(f.fout, fo(5), (k * ny + yY) * nx, tX, hz, sr, rng);
    }
    __syncthreads();

    // ---- E_new(k) on this row (owned waves) ----
    if (!halo && st) {
      const float hz_dn = lH[bb][w - 1][0][lane];
      const float hx_dn = lH[bb][w - 1][1][lane];
      const float hz_l = lane_shift(hz, up);
      const float hy_l = lane_shift(hy, up);
      float e0 = gdiff(hz, hz_dn, y, ny, mb, mt) * idx;  // dHz/dy
      float e1 = gdiff(hy, phy, k, nz, mf, mk) * idx;    // dHy/dz
      float e2 = gdiff(hx, phx, k, nz, mf, mk) * idx;    // dHx/dz
      float e3, e4;                                      // dHz/dx, dHy/dx
      if constexpr (kXEdge) {
        e3 = gdiff(hz, hz_l, x, nx, ml, mr) * idx;
        e4 = gdiff(hy, hy_l, x, nx, ml, mr) * idx;
      } else {
        e3 = (hz - hz_l) * idx;
        e4 = (hy - hy_l) * idx;
      }
      float e5 = gdiff(hx, hx_dn, y, ny, mb, mt) * idx;  // dHx/dy
      const bool okx = own && k <= nz && vyX && vxX;
      const bool oky = own && k <= nz && vyY && vxY;
      const bool okz = own && k < nz && vyX && vxY;
      if (qyE >= 0) {
        e0 = cpml_step(coef_u(sa.tab(), 6, TS, qyE), e0, p6, n);
        if (okx) stu<P>(sa.base(), sa.out(6), (kx * n2yE + qyE) * nx, pX, n);
        e5 = cpml_step(coef_u(sa.tab(), 11, TS, qyE), e5, p11, n);
        if (okz) stu<P>(sa.base(), sa.out(11), (kzc * n2yE + qyE) * (nx + 1), pY, n);
      }
      if (qzE >= 0) {
        e1 = cpml_step(coef_u(sa.tab(), 7, TS, qzE), e1, p7, n);
        if (okx) stu<P>(sa.base(), sa.out(7), (qzE * (ny + 1) + yX) * nx, pX, n);
        e2 = cpml_step(coef_u(sa.tab(), 8, TS, qzE), e2, p8, n);
        if (oky) stu<P>(sa.base(), sa.out(8), (qzE * ny + yY) * (nx + 1), pY, n);
      }
      if constexpr (kXEdge) {
        if (qxE >= 0) {
          e3 = cpml_step(c9, e3, p9, n);
          if (oky) stu<P>(sa.base(), sa.out(9), (kx * ny + yY) * n2xE, qE, n);
          e4 = cpml_step(c10, e4, p10, n);
          if (okz) stu<P>(sa.base(), sa.out(10), (kzc * (ny + 1) + yX) * n2xE, qE, n);
        }
      }
      float nex = ec[0] * ex + es0 * (e0 - e1);
      float ney = ec[1] * ey + es1 * (e2 - e3);
      float nez = ec[2] * ez + es2 * (e4 - e5);
      const bool zf = (mf && k == 0) || (mk && k == nz);
      const bool yf = (mb && y == 0) || (mt && y == ny);
      if (zf || yf) nex = 0.0f;
      if (zf || (kXEdge && xface)) ney = 0.0f;
      if (yf || (kXEdge && xface)) nez = 0.0f;
      if (okx) stf<T>(f.fout, fo(0), (kx * (ny + 1) + yX) * nx, tX, nex, sr, rng);
      if (oky) stf<T>(f.fout, fo(1), (kx * ny + yY) * (nx + 1), tY, ney, sr, rng);
      if (okz) stf<T>(f.fout, fo(2), (kzc * (ny + 1) + yX) * (nx + 1), tY, nez, sr, rng);
    }
    phx = hx;
    phy = hy;
    ex = exn;
    ey = eyn;
    ez = ezn;
  }
}

// x-interior tiles (no per-lane x checks) and x-edge tiles.  With 2-byte
// fields the interior variant is held to 8 waves per SIMD, so that two
// 16-wave workgroups share a compute unit: its loads move half the bytes per
// instruction, and the second workgroup's loads cover the latency (1.80 ms
// against 2.11 ms at 1024x512x256 f16).  FP32 keeps the compiler's 7 waves:
// forced to 8 it spills to scratch and runs slower (2.68 ms against 2.53 ms).
// kHSrc: the variant that applies H-timing source slabs (mode-source runs);
// the others carry none of its code or registers.
template <typename T, typename P, typename C, int kR, bool kEArr, bool kHSrc, bool kWide>
__global__ void __launch_bounds__(64 * (kR + 1)) fused_interior(FusedParams f) {
  fused_body<T, P, C, kR, kEArr, false, kHSrc, kWide>(f);
}
template <typename T, typename P, typename C, int kR, bool kEArr, bool kHSrc>
__global__ void __launch_bounds__(64 * (kR + 1)) __attribute__((amdgpu_waves_per_eu(8, 8)))
fused_interior_half(FusedParams f) {
  fused_body<T, P, C, kR, kEArr, false, kHSrc, false>(f);
}
template <typename T, typename P, typename C, int kR, bool kEArr, bool kHSrc, bool kWide>
__global__ void __launch_bounds__(64 * (kR + 1)) fused_edge(FusedParams f) {
  fused_body<T, P, C, kR, kEArr, true, kHSrc, kWide>(f);
}

template <typename T, typename P, typename C, int kR, bool kHSrc>
void launch_wg_h(hipStream_t s, const FusedParams& p, int nblocks) {
  const dim3 block(64 * (kR + 1), 1, 1), grid(nblocks, 1, 1);
  const bool earr = p.esrc_st[0] != 0;
  if (p.wide) {
    // 64-bit offsets: float32 fields, memories and coefficients, 15-row tiles
    if constexpr (std::is_same<T, float>::value && std::is_same<P, float>::value &&
                  std::is_same<C, float>::value && kR == 15) {
      if (p.xedge) {
        if (earr) hipLaunchKernelGGL((fused_edge<T, P, C, kR, true, kHSrc, true>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((fused_edge<T, P, C, kR, false, kHSrc, true>), grid, block, 0, s, p);
      } else {
        if (earr) hipLaunchKernelGGL((fused_interior<T, P, C, kR, true, kHSrc, true>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((fused_interior<T, P, C, kR, false, kHSrc, true>), grid, block, 0, s, p);
      }
      return;
    }
    throw Failure{"the wide fused step takes float32 fields, CPML memories and coefficients "
                  "and 15-row tiles"};
  }
  if (p.xedge) {
    if (earr) hipLaunchKernelGGL((fused_edge<T, P, C, kR, true, kHSrc, false>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((fused_edge<T, P, float, kR, false, kHSrc, false>), grid, block, 0, s, p);
  } else if constexpr (sizeof(T) == 2) {
    if (earr) hipLaunchKernelGGL((fused_interior_half<T, P, C, kR, true, kHSrc>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((fused_interior_half<T, P, float, kR, false, kHSrc>), grid, block, 0, s, p);
  } else {
    if (earr) hipLaunchKernelGGL((fused_interior<T, P, C, kR, true, kHSrc, false>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((fused_interior<T, P, float, kR, false, kHSrc, false>), grid, block, 0, s, p);
  }
}

template <typename T, typename P, typename C, int kR>
void launch_wg(hipStream_t s, const FusedParams& p, int nblocks) {
  if (p.hslab_n > 0) launch_wg_h<T, P, C, kR, true>(s, p, nblocks);
  else launch_wg_h<T, P, C, kR, false>(s, p, nblocks);
}

// Rows per workgroup: 15 by default, 8 selectable (BZH_FUSED_ROWS) for
// tuning sweeps.  (13 rows, two 14-wave workgroups per CU at 7 waves per
// SIMD, measured within 1 % of 15 in FP32 and 25 % slower in f16.)
int rows_setting() {
  int r = 0;
  if (const char* env = std::getenv("BZH_FUSED_ROWS")) r = std::atoi(env);
  return r == 8 ? 8 : 15;
}

template <typename T, typename P, typename C>
void launch_rows(hipStream_t s, const FusedParams& p, int nblocks) {
  if (p.rows == 8) launch_wg<T, P, C, 8>(s, p, nblocks);
  else launch_wg<T, P, C, 15>(s, p, nblocks);
}

template <typename T, typename P>
void launch_coef(hipStream_t s, const FusedParams& p, int nblocks) {
  if constexpr (std::is_same<T, float>::value) {
    launch_rows<T, P, float>(s, p, nblocks);
  } else {
    // half-precision fields carry float16 (folded) or float32 E coefficients
    if (p.coef_f16) launch_rows<T, P, __half>(s, p, nblocks);
    else launch_rows<T, P, float>(s, p, nblocks);
  }
}

template <typename T>
void launch_field(hipStream_t s, const FusedParams& p, int nblocks) {
  if (p.psi_bf16) launch_coef<T, bf16_t>(s, p, nblocks);
  else launch_coef<T, float>(s, p, nblocks);
}


// ---- two leapfrog steps per launch -------------------------------------------
//
// A workgroup of 16 waves owns 60 (x) by 13 (y) cells of the second step's
// output and marches z over a chunk of planes.  Lane l sits at
// x = i0 - 2 + l, wave w on row y = j0 - 2 + w.  The first step (H', E') is
// computed on the whole 64 x 16 block: H' is complete on lanes 0..62, E' on
// lanes 1..62 and waves 1..15.  The second (H'', E'') runs where its inputs
// are complete: H'' on lanes 1..61 and waves 1..14, E'' on lanes 2..61 and
// waves 2..14, the owned cells.  Along z the second step lags, so rows still
// cross through LDS with one barrier per plane.  Iteration k computes
//   before the barrier: H'(k)   from E(k), E(k+1), E(k) of the row above
//                                (published at k-1);
//   after it:           E'(k)   from H'(k), H'(k-1), H'(k) of the row below
//                                (published before this barrier),
//                       H''(k-1) from E'(k-1), E'(k), E'(k-1) of the row above
//                                (published at k-1),
//                       E''(k-2) from H''(k-2), H''(k-3), H''(k-2) of the
//                                row below (published at k-1).
// A chunk owning planes [kz0, kz1) starts at plane kz0 - 2 and ends at
// kz1 + 1.  The first step's fields stay in registers and LDS; they reach
// HBM only inside the monitor boxes (the intermediate bank read by the DFT of
// the first step).  Between the steps the kernel applies the first step's
// E-timing source slabs (before that sample) and the second step's pre-E
// slabs (after it), in the order the single-step schedule applies them.
// Arithmetic per update is the single-step kernel's; in half precision the
// intermediate fields are not rounded to the storage type.
constexpr int kPairWaves = kPairRows + 3;
// LDS rows of the pair kernel, shared by both of its variants.
struct PairLds {
  float e1[2][kPairWaves + 1][2][64];  // E rows (Ez, Ex) for H'
  float h1[2][kPairWaves][2][64];      // H' rows (Hz, Hx) for E'
  float e2[2][kPairWaves][2][64];      // E' rows for H''
  float h2[2][kPairWaves][2][64];      // H'' rows for E''
  // per wave and DFT component, the monitor capture of its row: x0, x1, z0,
  // z1 (x0 > x1: none), and the slot index base + x * stride + z
  int cap[kPairWaves][6][6];
};
// The core kernel's prefetch rows: each wave's inputs of its next iteration,
// fetched by LDS DMA while the current one computes: E(k+1) x,y,z, H(k)
// x,y,z, E coefficients x,y,z, top-row E(k+1) x,z (wave 0).
using PairPrefetch = float[kPairWaves][11][64];

template <typename T, typename P, typename C, bool kEArr, bool kXEdge, bool kHSrc, bool kCore,
          bool kUni, bool kWide>
__device__ __forceinline__ void pair_body(const PairParams& pp, PairLds& lds, PairPrefetch* pf,
                                          int box, int blk) {
  const FusedParams& f = pp.f;
  // field and E coefficient offsets in their groups: 32-bit, 64-bit in the
  // wide variant (kWide: a bank of 4 GiB or more)
  auto fo = [&](int c) {
    if constexpr (kWide) return f.foff64[c];
    else return f.foff[c];
  };
  auto co = [&](int c) {
    if constexpr (kWide) return f.coff64[c];
    else return f.coff[c];
  };
  const SlabArgs<kXEdge> sa{f};  // x-edge tiles: reloaded (x-slab terms raise the pressure)
  constexpr int kW = kPairWaves;
  auto& lE1 = lds.e1;
  auto& lH1 = lds.h1;
  auto& lE2 = lds.e2;
  auto& lH2 = lds.h2;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool halo = w == 0;  // computes H' (Hx, Hz) only and loads the top E row

  int bx, by, kz0, kz1;
  if (!decode_pair_box(pp.box[box], blk, bx, by, kz0, kz1)) return;

  const int nz = f.nz, ny = f.ny, nx = f.nx, L = f.L;
  // CPML slab extents (lo, hi) per kSlabH/kSlabE + axis, the memories' extent
  // along a slab axis and the coefficient table's row stride.  kUni: one
  // thickness L on every face, so that all of them derive from L (fewer
  // scalar registers held across the z-march: the general layout measured
  // 1.6 % slower in the shell blocks)
  auto clo = [&](int i) { return kUni ? L : f.clo[i]; };
  auto chi = [&](int i) { return kUni ? L : f.chi[i]; };
  auto n2 = [&](int i) { return kUni ? 2 * L : f.clo[i] + f.chi[i]; };
  auto zq = [&](int k, int n, int i) { return slab_q(k, n, clo(i), chi(i)); };
  auto tst = [&]() { return kUni ? 2 * L : f.tab_st; };
  const int m = f.metallic;
  const bool mf = m & kFront, mk = m & kBack, mb = m & kBottom, mt = m & kTop;
  const bool ml = m & kLeft, mr = m & kRight;
  const float idx = f.inv_dx;
  const int j0 = by * kPairRows;
  const int x = bx * kPairCols - 2 + lane;
  const int dn = ((lane + 1) & 63) << 2;
  const int up = ((lane - 1) & 63) << 2;

  bool vxX = true, vxY = true, xface = false;
  int xX = x, xY = x, qxH = -1, qxE = -1;
  if constexpr (kXEdge) {
    vxX = x >= 0 && x < nx;
    vxY = x >= 0 && x <= nx;
    xX = min(max(x, 0), nx - 1);
    xY = min(max(x, 0), nx);
    qxH = vxX ? slab_q(x, nx, clo(kSlabH), chi(kSlabH)) : -1;
    qxE = vxY ? slab_q(x, nx + 1, clo(kSlabE), chi(kSlabE)) : -1;
    xface = (ml && x == 0) || (mr && x == nx);
  }
  const bool own = lane >= 2 && lane <= 61 && w >= 2 && w <= kPairRows + 1;
  const uint32_t tX = xX * sizeof(T), tY = xY * sizeof(T);
  const uint32_t pX = xX * sizeof(P), pY = xY * sizeof(P);
  const uint32_t cX = xX * sizeof(C), cY = xY * sizeof(C);
  const uint32_t qH = max(qxH, 0) * sizeof(P), qE = max(qxE, 0) * sizeof(P);
  Coef c3{}, c4{}, c9{}, c10{};
  if constexpr (kXEdge) {
    if (qxH >= 0) {
      c3 = coef_v(f.tab, 3, tst(), qxH);
      c4 = coef_v(f.tab, 4, tst(), qxH);
    }
    if (qxE >= 0) {
      c9 = coef_v(f.tab, 9, tst(), qxE);
      c10 = coef_v(f.tab, 10, tst(), qxE);
    }
  }

  // core tiles: every row lies in [max(L, 1), ny - L - 1] (checked on the
  // host), so the y terms, faces and validity cuts are compiled out
  const int y = j0 - 2 + w;
  const bool vyX = kCore || (y >= 0 && y <= ny), vyY = kCore || (y >= 0 && y < ny);
  const int yX = kCore ? y : min(max(y, 0), ny), yY = kCore ? y : min(max(y, 0), ny - 1);
  const int qyH = (!kCore && vyY) ? slab_q(y, ny, clo(kSlabH + 1), chi(kSlabH + 1)) : -1;
  const int qyE = (!kCore && vyX) ? slab_q(y, ny + 1, clo(kSlabE + 1), chi(kSlabE + 1)) : -1;
  const bool y_pec_h = !kCore && ((mb && y == 0) || (mt && y == ny));
  const bool row_plain = kCore || (y >= 1 && y < ny);  // no y face, ghost or validity cut on this row
  const int yt = j0 - 2 + kW;
  const bool vtX = yt <= ny;
  const int ytX = min(yt, ny);

  float hc[6], ec[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    hc[q] = as_const(f.hdec[q])[0];
    hc[3 + q] = as_const(f.hsrc[q])[0] * f.ratio_h;
    ec[q] = as_const(f.edec[q])[0];
  }
  const float hci[3] = {hc[3] * idx, hc[4] * idx, hc[5] * idx};
  float es0 = 0.0f, es1 = 0.0f, es2 = 0.0f;
  if constexpr (!kEArr) {
    es0 = as_const(f.esrc[0])[0] * f.ratio_e;
    es1 = as_const(f.esrc[1])[0] * f.ratio_e;
    es2 = as_const(f.esrc[2])[0] * f.ratio_e;
  }

  // monitor boxes: bit b set on lanes whose (y, x) lies in box b; planes are
  // tested per k.  Only owned cells write the intermediate bank.
  const bool capp = pp.cap != nullptr;
  uint32_t mbits = 0;
  int zs0 = 1 << 30, zs1 = -1;
  bool mwave = false;
  if (capp) {
    // the capture: lane c < 6 finds component c's box of the one monitor
    // whose cells this row holds (the host checked there is at most one)
    int e0 = 1, e1 = 0, e2 = 0, e3 = -1, e4 = 0, e5 = 0;
    const int xl = bx * kPairCols, xh = xl + kPairCols - 1;
    if (lane < 6 && w >= 2 && w <= kPairRows + 1) {
      for (int mm = 0; mm < pp.cap_m; ++mm) {
        const int32_t* t = pp.capbox + (mm * 6 + lane) * 8;
        if (t[0] > t[1] || y < t[2] || y > t[3] || t[5] < xl || t[4] > xh) continue;
        const int nzb = t[1] - t[0] + 1, stride = nzb * (t[3] - t[2] + 1);
        e0 = t[4];
        e1 = t[5];
        e2 = t[0];
        e3 = t[1];
        e4 = t[6] - t[4] * stride + (y - t[2]) * nzb - t[0];
        e5 = stride;
        break;
      }
      int* dst = lds.cap[w][lane];
      dst[0] = e0;
      dst[1] = e1;
      dst[2] = e2;
      dst[3] = e3;
      dst[4] = e4;
      dst[5] = e5;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int x0c = __builtin_amdgcn_readlane(e0, c), x1c = __builtin_amdgcn_readlane(e1, c);
      if (x0c > x1c) continue;
      mwave = true;
      zs0 = min(zs0, __builtin_amdgcn_readlane(e2, c));
      zs1 = max(zs1, __builtin_amdgcn_readlane(e3, c));
      if (own && x >= x0c && x <= x1c) mbits = 1u;
    }
  } else {
    for (int b = 0; b < pp.mbox_n; ++b) {
      cint* bx6 = (cint*)(pair_kargs()->mbox) + 6 * b;
      if (y >= bx6[2] && y <= bx6[3] && x >= bx6[4] && x <= bx6[5]) mbits |= 1u << b;
    }
    mbits = own ? mbits : 0u;
    mwave = __builtin_amdgcn_readfirstlane(__any(mbits != 0u)) != 0;
    // the planes this wave samples: the hull of the z ranges of the boxes
    // that hold any of its lanes (one uniform test per plane; lanes in a box
    // whose z range misses a plane of the hull write cells no gather reads,
    // which is harmless: the intermediate bank has the field bank's layout)
    if (mwave) {
      for (int b = 0; b < pp.mbox_n; ++b) {
        if (__builtin_amdgcn_readfirstlane(__any((mbits >> b) & 1u)) == 0) continue;
        cint* bx6 = (cint*)(pair_kargs()->mbox) + 6 * b;
        zs0 = min(zs0, bx6[0]);
        zs1 = max(zs1, bx6[1]);
      }
    }
  }
  const bool mlane = mbits != 0u;
  // the intermediate step is sampled on the chunk's own planes
  zs0 = max(zs0, kz0);
  zs1 = min(zs1, kz1 - 1);
  auto sample_plane = [&](int k) { return k >= zs0 && k <= zs1; };
  // one captured value of component c (DFT order) on plane k
  auto cap_store = [&](int c, int k, float v) {
    const int* e = lds.cap[w][c];
    if (x >= e[0] && x <= e[1] && k >= e[2] && k <= e[3]) pp.cap[e[4] + x * e[5] + k] = v;
  };

  auto load_e = [&](int k, float& ex, float& ey, float& ez) {
    const bool vxy = k <= nz, vz = k < nz;
    const uint32_t kx = min(k, nz), kz = min(k, nz - 1);
    const float a = ldu<T>(f.fin, fo(0), (kx * (ny + 1) + yX) * nx, tX);
    const float c = ldu<T>(f.fin, fo(1), (kx * ny + yY) * (nx + 1), tY);
    const float d = ldu<T>(f.fin, fo(2), (kz * (ny + 1) + yX) * (nx + 1), tY);
    ex = (vxy && vyX && vxX) ? a : 0.0f;
    ey = (vxy && vyY && vxY) ? c : 0.0f;
    ez = (vz && vyX && vxY) ? d : 0.0f;
  };
  auto load_top = [&](int k, float& ex, float& ez) {
    const bool vxy = k <= nz, vz = k < nz;
    const uint32_t kx = min(k, nz), kz = min(k, nz - 1);
    const float a = ldu<T>(f.fin, fo(0), (kx * (ny + 1) + ytX) * nx, tX);
    const float d = ldu<T>(f.fin, fo(2), (kz * (ny + 1) + ytX) * (nx + 1), tY);
    ex = (vxy && vtX && vxX) ? a : 0.0f;
    ez = (vz && vtX && vxY) ? d : 0.0f;
  };

  // CPML memories of plane k: H terms 0..5 at [t], E terms 6..11 at [t - 6]
  auto load_psi_h = [&](int k, float (&p)[6]) {
    const uint32_t kk = min(k, nz - 1), kh = min(k, nz);
    const int qzH = k < nz ? zq(k, nz, kSlabH + 2) : -1;
    if (qyH >= 0) {
      p[0] = ldu<P>(sa.base(), sa.in(0), (kk * n2(kSlabH + 1) + qyH) * (nx + 1), pY);
      p[5] = ldu<P>(sa.base(), sa.in(5), (kh * n2(kSlabH + 1) + qyH) * nx, pX);
    }
    if (qzH >= 0) {
      p[1] = ldu<P>(sa.base(), sa.in(1), (qzH * ny + yY) * (nx + 1), pY);
      if (!halo) p[2] = ldu<P>(sa.base(), sa.in(2), (qzH * (ny + 1) + yX) * nx, pX);
    }
    if constexpr (kXEdge) {
      if (qxH >= 0) {
        if (!halo) p[3] = ldu<P>(sa.base(), sa.in(3), (kk * (ny + 1) + yX) * n2(kSlabH), qH);
        p[4] = ldu<P>(sa.base(), sa.in(4), (kh * ny + yY) * n2(kSlabH), qH);
      }
    }
  };
  auto load_psi_e = [&](int k, float (&p)[6]) {
    const uint32_t kx = min(k, nz), kzc = min(k, nz - 1);
    const int qzE = k <= nz ? zq(k, nz + 1, kSlabE + 2) : -1;
    if (qyE >= 0) {
      p[0] = ldu<P>(sa.base(), sa.in(6), (kx * n2(kSlabE + 1) + qyE) * nx, pX);
      p[5] = ldu<P>(sa.base(), sa.in(11), (kzc * n2(kSlabE + 1) + qyE) * (nx + 1), pY);
    }
    if (qzE >= 0) {
      p[1] = ldu<P>(sa.base(), sa.in(7), (qzE * (ny + 1) + yX) * nx, pX);
      p[2] = ldu<P>(sa.base(), sa.in(8), (qzE * ny + yY) * (nx + 1), pY);
    }
    if constexpr (kXEdge) {
      if (qxE >= 0) {
        p[3] = ldu<P>(sa.base(), sa.in(9), (kx * ny + yY) * n2(kSlabE), qE);
        p[4] = ldu<P>(sa.base(), sa.in(10), (kzc * (ny + 1) + yX) * n2(kSlabE), qE);
      }
    }
  };

  // H update of plane k on this row (kernels.py:458-503), psi terms 0..5
  // from pin to pout; invalid positions come out zero.  On planes and rows
  // away from every face (``plain``, uniform) the face masks and the validity
  // zeroing are no-ops and are skipped.
  auto h_update = [&](int k, float ez, float ey, float ex, float eyn, float exn, float ez_up,
                      float ex_up, float ha0, float ha1, float ha2, const float (&pin)[6],
                      float (&pout)[6], float& hx, float& hy, float& hz) {
    const bool vxy = k < nz;
    const int qzH = vxy ? zq(k, nz, kSlabH + 2) : -1;
    const float ez_r = lane_shift(ez, dn), ey_r = lane_shift(ey, dn);
    if (!kXEdge && k >= 1 && k < nz && row_plain && qyH < 0 && qzH < 0) {
      // no face and no CPML term: 1/dx folded into the coefficients
      hx = hc[0] * ha0 - hci[0] * ((ez_up - ez) - (eyn - ey));
      hy = hc[1] * ha1 - hci[1] * ((exn - ex) - (ez_r - ez));
      hz = hc[2] * ha2 - hci[2] * ((ey_r - ey) - (ex_up - ex));
      return;
    }
    float d0 = (ez_up - ez) * idx;               // dEz/dy
    float d1 = (eyn - ey) * idx;                 // dEy/dz
    float d2 = (exn - ex) * idx;                 // dEx/dz
    float d3 = (ez_r - ez) * idx;                // dEz/dx
    float d4 = (ey_r - ey) * idx;                // dEy/dx
    float d5 = (ex_up - ex) * idx;               // dEx/dy
    if (qyH >= 0) {
      d0 = cpml_step(coef_u(sa.tab(), 0, tst(), qyH), d0, pin[0], pout[0]);
      d5 = cpml_step(coef_u(sa.tab(), 5, tst(), qyH), d5, pin[5], pout[5]);
    }
    if (qzH >= 0) {
      d1 = cpml_step(coef_u(sa.tab(), 1, tst(), qzH), d1, pin[1], pout[1]);
      d2 = cpml_step(coef_u(sa.tab(), 2, tst(), qzH), d2, pin[2], pout[2]);
    }
    if constexpr (kXEdge) {
      if (qxH >= 0) {
        d3 = cpml_step(c3, d3, pin[3], pout[3]);
        d4 = cpml_step(c4, d4, pin[4], pout[4]);
      }
    }
    hx = hc[0] * ha0 - hc[3] * (d0 - d1);
    hy = hc[1] * ha1 - hc[4] * (d2 - d3);
    hz = hc[2] * ha2 - hc[5] * (d4 - d5);
    const bool plain = !kXEdge && k >= 1 && k < nz && row_plain;
    if (!plain) {
      if (kXEdge && xface) hx = 0.0f;
      if (y_pec_h) hy = 0.0f;
      if ((mf && k == 0) || (mk && k == nz)) hz = 0.0f;
      hx = (vxy && vyY && vxY) ? hx : 0.0f;
      hy = (vxy && vyX && vxX) ? hy : 0.0f;
      hz = (k <= nz && vyY && vxX) ? hz : 0.0f;
    }
  };

  // E update of plane k on this row (kernels.py:554-626, ghost rule
  // lattice.py:523-575), psi terms 6..11 from pin[t - 6] to pout[t - 6];
  // invalid positions come out zero.  On ``plain`` planes and rows the ghost
  // rule reduces to plain differences and the masks are no-ops.
  auto e_update = [&](int k, float hx, float hy, float hz, float hz_dn, float hx_dn, float phx,
                      float phy, float ex, float ey, float ez, float s0, float s1, float s2,
                      const float (&pin)[6], float (&pout)[6], float& nex, float& ney,
                      float& nez) {
    const int qzE = k <= nz ? zq(k, nz + 1, kSlabE + 2) : -1;
    const bool plain = !kXEdge && k >= 1 && k < nz && row_plain;
    const float hz_l = lane_shift(hz, up);
    const float hy_l = lane_shift(hy, up);
    if (plain && qyE < 0 && qzE < 0) {
      // no face and no CPML term: 1/dx folded into the coefficients
      nex = ec[0] * ex + (s0 * idx) * ((hz - hz_dn) - (hy - phy));
      ney = ec[1] * ey + (s1 * idx) * ((hx - phx) - (hz - hz_l));
      nez = ec[2] * ez + (s2 * idx) * ((hy - hy_l) - (hx - hx_dn));
      return;
    }
    float e0, e1, e2, e5;
    if (plain) {
      e0 = (hz - hz_dn) * idx;
      e1 = (hy - phy) * idx;
      e2 = (hx - phx) * idx;
      e5 = (hx - hx_dn) * idx;
    } else {
      e0 = gdiff(hz, hz_dn, y, ny, mb, mt) * idx;  // dHz/dy
      e1 = gdiff(hy, phy, k, nz, mf, mk) * idx;    // dHy/dz
      e2 = gdiff(hx, phx, k, nz, mf, mk) * idx;    // dHx/dz
      e5 = gdiff(hx, hx_dn, y, ny, mb, mt) * idx;  // dHx/dy
    }
    float e3, e4;                                  // dHz/dx, dHy/dx
    if constexpr (kXEdge) {
      e3 = gdiff(hz, hz_l, x, nx, ml, mr) * idx;
      e4 = gdiff(hy, hy_l, x, nx, ml, mr) * idx;
    } else {
      e3 = (hz - hz_l) * idx;
      e4 = (hy - hy_l) * idx;
    }
    if (qyE >= 0) {
      e0 = cpml_step(coef_u(sa.tab(), 6, tst(), qyE), e0, pin[0], pout[0]);
      e5 = cpml_step(coef_u(sa.tab(), 11, tst(), qyE), e5, pin[5], pout[5]);
    }
    if (qzE >= 0) {
      e1 = cpml_step(coef_u(sa.tab(), 7, tst(), qzE), e1, pin[1], pout[1]);
      e2 = cpml_step(coef_u(sa.tab(), 8, tst(), qzE), e2, pin[2], pout[2]);
    }
    if constexpr (kXEdge) {
      if (qxE >= 0) {
        e3 = cpml_step(c9, e3, pin[3], pout[3]);
        e4 = cpml_step(c10, e4, pin[4], pout[4]);
      }
    }
    nex = ec[0] * ex + s0 * (e0 - e1);
    ney = ec[1] * ey + s1 * (e2 - e3);
    nez = ec[2] * ez + s2 * (e4 - e5);
    if (!plain) {
      const bool zf = (mf && k == 0) || (mk && k == nz);
      const bool yf = (mb && y == 0) || (mt && y == ny);
      if (zf || yf) nex = 0.0f;
      if (zf || (kXEdge && xface)) ney = 0.0f;
      if (yf || (kXEdge && xface)) nez = 0.0f;
      nex = (k <= nz && vyX && vxX) ? nex : 0.0f;
      ney = (k <= nz && vyY && vxY) ? ney : 0.0f;
      nez = (k < nz && vyX && vxY) ? nez : 0.0f;
    }
  };

  // stores of plane k (fields to ``out``, CPML memories to the written bank)
  // (bfloat16 stores of the second step's fields to bank B round
  // stochastically when f.bf16_sr; the intermediate bank keeps nearest)
  const bool sr = std::is_same<T, bf16_t>::value && f.bf16_sr != 0;
  SrState rng{0u};
  if (sr) rng = sr_launch_seed(f, 1);
  auto store_h = [&](const char* out, int k, float hx, float hy, float hz) {
    const bool s = sr && out == f.fout;
    if (k < nz && vyY && vxY) stf<T>(out, fo(3), (k * ny + yY) * (nx + 1), tY, hx, s, rng);
    if (k < nz && vyX && vxX) stf<T>(out, fo(4), (k * (ny + 1) + yX) * nx, tX, hy, s, rng);
    if (k <= nz && vyY && vxX) stf<T>(out, fo(5), (k * ny + yY) * nx, tX, hz, s, rng);
  };
  auto store_e = [&](const char* out, int k, float ex, float ey, float ez) {
    const bool s = sr && out == f.fout;
    if (k <= nz && vyX && vxX) stf<T>(out, fo(0), (k * (ny + 1) + yX) * nx, tX, ex, s, rng);
    if (k <= nz && vyY && vxY) stf<T>(out, fo(1), (k * ny + yY) * (nx + 1), tY, ey, s, rng);
    if (k < nz && vyX && vxY) stf<T>(out, fo(2), (k * (ny + 1) + yX) * (nx + 1), tY, ez, s, rng);
  };
  auto store_psi_h = [&](int k, const float (&p)[6]) {
    const bool okx = k < nz && vyY && vxY, oky = k < nz && vyX && vxX, okz = k <= nz && vyY && vxX;
    const int qzH = k < nz ? zq(k, nz, kSlabH + 2) : -1;
    if (qyH >= 0) {
      if (okx) stu<P>(sa.base(), sa.out(0), (k * n2(kSlabH + 1) + qyH) * (nx + 1), pY, p[0]);
      if (okz) stu<P>(sa.base(), sa.out(5), (k * n2(kSlabH + 1) + qyH) * nx, pX, p[5]);
    }
    if (qzH >= 0) {
      if (okx) stu<P>(sa.base(), sa.out(1), (qzH * ny + yY) * (nx + 1), pY, p[1]);
      if (oky) stu<P>(sa.base(), sa.out(2), (qzH * (ny + 1) + yX) * nx, pX, p[2]);
    }
    if constexpr (kXEdge) {
      if (qxH >= 0) {
        if (oky) stu<P>(sa.base(), sa.out(3), (k * (ny + 1) + yX) * n2(kSlabH), qH, p[3]);
        if (okz) stu<P>(sa.base(), sa.out(4), (k * ny + yY) * n2(kSlabH), qH, p[4]);
      }
    }
  };
  auto store_psi_e = [&](int k, const float (&p)[6]) {
    const bool okx = k <= nz && vyX && vxX, oky = k <= nz && vyY && vxY, okz = k < nz && vyX && vxY;
    const int qzE = k <= nz ? zq(k, nz + 1, kSlabE + 2) : -1;
    if (qyE >= 0) {
      if (okx) stu<P>(sa.base(), sa.out(6), (k * n2(kSlabE + 1) + qyE) * nx, pX, p[0]);
      if (okz) stu<P>(sa.base(), sa.out(11), (k * n2(kSlabE + 1) + qyE) * (nx + 1), pY, p[5]);
    }
    if (qzE >= 0) {
      if (okx) stu<P>(sa.base(), sa.out(7), (qzE * (ny + 1) + yX) * nx, pX, p[1]);
      if (oky) stu<P>(sa.base(), sa.out(8), (qzE * ny + yY) * (nx + 1), pY, p[2]);
    }
    if constexpr (kXEdge) {
      if (qxE >= 0) {
        if (oky) stu<P>(sa.base(), sa.out(9), (k * ny + yY) * n2(kSlabE), qE, p[3]);
        if (okz) stu<P>(sa.base(), sa.out(10), (k * (ny + 1) + yX) * n2(kSlabE), qE, p[4]);
      }
    }
  };

  // source slabs of the pair's list on plane k of this row: part 0 E-timing of
  // the first step, 1 pre-E of the second, 2 H-timing of the first, 3
  // H-timing of the second (bounds re-read from the argument block, so that
  // they hold no scalar registers across the z-march)
  auto mid_sources = [&](int part, int k, float& vx, float& vy, float& vz) {
    auto* a = pair_kargs();
    const int s0 = part == 0 ? 0 : (part == 1 ? a->msrc_post : (part == 2 ? a->msrc_e : a->msrc_h1));
    const int s1 = part == 0 ? a->msrc_post : (part == 1 ? a->msrc_e : (part == 2 ? a->msrc_h1 : a->msrc_n));
    if (s1 <= s0) return;
    const int step = *(cint*)(a->step0);
    for (int s = s0; s < s1; ++s) slab_add(a->msrc[s], step, k, y, x, nz, ny, nx, vx, vy, vz);
  };

  // does any slab between the steps cross this wave's row and lanes?  (its
  // planes are tested per k inside mid_sources)
  bool swave = false;
  for (int s = 0; s < pair_kargs()->msrc_n; ++s)
    swave = swave || slab_hits_wave(pair_kargs()->msrc[s], y, x);

  const int ks = max(kz0 - 2, 0);
  const int kE = max(kz0 - 1, 0);  // first plane of E' the chunk needs
  float ex, ey, ez;                // E(k) on this row
  load_e(ks, ex, ey, ez);
  lE1[ks & 1][w][0][lane] = ez;
  lE1[ks & 1][w][1][lane] = ex;
  if (halo) {
    float tx_, tz_;
    load_top(ks, tx_, tz_);
    lE1[ks & 1][kW][0][lane] = tz_;
    lE1[ks & 1][kW][1][lane] = tx_;
  }
  __syncthreads();

  float h1x = 0.0f, h1y = 0.0f, h1z = 0.0f;  // H'(k-1)
  float e1x = 0.0f, e1y = 0.0f, e1z = 0.0f;  // E'(k-1)
  float e2x = 0.0f, e2y = 0.0f, e2z = 0.0f;  // E'(k-2)
  float g2x = 0.0f, g2y = 0.0f, g2z = 0.0f;  // H''(k-2)
  float g3x = 0.0f, g3y = 0.0f;              // H''(k-3)
  float sa0 = es0, sa1 = es1, sa2 = es2;     // E source coefficients of plane k-1
  float sb0 = es0, sb1 = es1, sb2 = es2;     // ... of plane k-2
  float mh[6] = {}, me1[6] = {}, me2[6] = {};  // psi': H terms at k-1, E terms at k-1, k-2

  // Lean iterations: on an x-interior row with no y face, ghost or CPML term,
  // iteration k touches planes k-2 .. k+1 only; when all of them are clear of
  // the z faces and z slabs and the chunk owns all four stages' outputs, the
  // iteration runs a branch-free form of the same arithmetic (the ``plain``
  // forms of h_update and e_update, bit for bit) with per-lane byte offsets
  // computed once per row and one uniform plane offset per array family.
  // Families (plane stride, row stride): A (ny+1)*nx, nx: Ex, Hy;  B ny*(nx+1),
  // nx+1: Ey, Hx;  C (ny+1)*(nx+1), nx+1: Ez;  D ny*nx, nx: Hz.
  const bool lean_row = !kXEdge && row_plain && qyH < 0 && qyE < 0;
  const int lean_lo = max(L, 1), lean_hi = nz - max(L, 1) - 1;  // planes clear of faces and slabs
  const int k_lo = lean_lo + 2, k_hi = lean_hi - 1;
  constexpr uint32_t szT = sizeof(T), szC = sizeof(C);
  const uint32_t oA = (yX * nx + xX) * szT, oB = (yY * (nx + 1) + xY) * szT;
  const uint32_t oC = (yX * (nx + 1) + xY) * szT, oD = (yY * nx + xX) * szT;
  const uint32_t oAt = (ytX * nx + xX) * szT, oCt = (ytX * (nx + 1) + xY) * szT;
  const uint32_t cA = (yX * nx + xX) * szC, cB = (yY * (nx + 1) + xY) * szC;
  const uint32_t cC = (yX * (nx + 1) + xY) * szC;
  const uint32_t nA = (ny + 1) * nx, nB = ny * (nx + 1), nC = (ny + 1) * (nx + 1), nD = ny * nx;
  // float32 fields and coefficients: a lean iteration's loads are issued as
  // LDS DMA during the previous lean iteration (after its barrier), so their
  // latency hides behind that iteration's second half
  constexpr bool kPf = kCore && sizeof(T) == 4 && sizeof(C) == 4;
  float (*pfw)[64] = kPf ? (*pf)[w] : nullptr;
  bool pf_ok = false;  // pfw holds this iteration's inputs
  // 2-byte fields: the same prefetch into registers (raw bits, converted at
  // the use), since these kernels have registers to spare
  constexpr bool kRp = kCore && !kPf;
  uint32_t rp[11] = {};

  for (int k = ks; k <= kz1 + 1; ++k) {
    // core launches: every row and iteration is lean by construction; the
    // shell kernel runs the general iteration only
    if constexpr (kCore) {
      const int bb = k & 1;
      const bool g1 = k <= kz1;
      const bool gE1 = g1 && k >= kE && w >= 1;
      const bool gH2 = k - 1 >= kE && k - 1 < kz1 && w >= 1 && w <= kW - 2;
      const bool gE2 = k - 2 >= kz0 && k - 2 < kz1 && w >= 2 && w <= kW - 2;
      const bool sampled = mwave && sample_plane(k);
      // plane k in bytes, per family; plane k + d is z + d * p
      const uint32_t pA = nA * szT, pB = nB * szT, pC = nC * szT, pD = nD * szT;
      const uint32_t zA = k * pA, zB = k * pB, zC = k * pC, zD = k * pD;
      // ---- first step: H'(k) (values used only after every load is issued) ----
      float exn = 0.0f, eyn = 0.0f, ezn = 0.0f, hx = 0.0f, hy = 0.0f, hz = 0.0f;
      float s0 = es0, s1 = es1, s2 = es2;
      if (g1) {
        float etx = 0.0f, etz = 0.0f, ha0, ha1 = 0.0f, ha2;
        if (kRp && pf_ok) {
          exn = from_raw<T>(rp[0]);
          eyn = from_raw<T>(rp[1]);
          ezn = from_raw<T>(rp[2]);
          ha0 = from_raw<T>(rp[3]);
          if (!halo) ha1 = from_raw<T>(rp[4]);
          ha2 = from_raw<T>(rp[5]);
          if constexpr (kEArr) {
            if (gE1) {
              s0 = from_raw<C>(rp[6]);
              s1 = from_raw<C>(rp[7]);
              s2 = from_raw<C>(rp[8]);
            }
          }
          if (halo) {
            etx = from_raw<T>(rp[9]);
            etz = from_raw<T>(rp[10]);
          }
        } else if (kPf && pf_ok) {
          exn = pfw[0][lane];
          eyn = pfw[1][lane];
          ezn = pfw[2][lane];
          ha0 = pfw[3][lane];
          if (!halo) ha1 = pfw[4][lane];
          ha2 = pfw[5][lane];
          if constexpr (kEArr) {
            if (gE1) {
              s0 = pfw[6][lane];
              s1 = pfw[7][lane];
              s2 = pfw[8][lane];
            }
          }
          if (halo) {
            etx = pfw[9][lane];
            etz = pfw[10][lane];
          }
        } else {
          exn = ldo<T>(f.fin, fo(0) + zA + pA, oA);
          eyn = ldo<T>(f.fin, fo(1) + zB + pB, oB);
          ezn = ldo<T>(f.fin, fo(2) + zC + pC, oC);
          if (halo) {
            etx = ldo<T>(f.fin, fo(0) + zA + pA, oAt);
            etz = ldo<T>(f.fin, fo(2) + zC + pC, oCt);
          }
          ha0 = ldo<T>(f.fin, fo(3) + zB, oB);
          if (!halo) ha1 = ldo<T>(f.fin, fo(4) + zA, oA);
          ha2 = ldo<T>(f.fin, fo(5) + zD, oD);
          if constexpr (kEArr) {
            if (gE1) {
              s0 = ldo<C>(f.cbase, co(0) + k * nA * szC, cA);
              s1 = ldo<C>(f.cbase, co(1) + k * nB * szC, cB);
              s2 = ldo<C>(f.cbase, co(2) + k * nC * szC, cC);
            }
          }
        }
        const float ez_up = lE1[bb][w + 1][0][lane];
        const float ex_up = lE1[bb][w + 1][1][lane];
        const float ez_r = lane_shift(ez, dn), ey_r = lane_shift(ey, dn);
        hx = hc[0] * ha0 - hci[0] * ((ez_up - ez) - (eyn - ey));
        hy = hc[1] * ha1 - hci[1] * ((exn - ex) - (ez_r - ez));
        hz = hc[2] * ha2 - hci[2] * ((ey_r - ey) - (ex_up - ex));
        if (kHSrc && swave) mid_sources(2, k, hx, hy, hz);
        lH1[bb][w][0][lane] = hz;
        lH1[bb][w][1][lane] = hx;
        lE1[bb ^ 1][w][0][lane] = ezn;
        lE1[bb ^ 1][w][1][lane] = exn;
        if (halo) {
          lE1[bb ^ 1][kW][0][lane] = (vtX && vxY) ? etz : 0.0f;
          lE1[bb ^ 1][kW][1][lane] = (vtX && vxX) ? etx : 0.0f;
        }
        if (sampled && mlane) {
          if (capp) {
            cap_store(3, k, hx);
            cap_store(4, k, hy);
            cap_store(5, k, hz);
          } else {
            const char* cm = pair_kargs()->cmid;
            stf<T>(cm, fo(3) + zB, 0u, oB, hx, false, rng);
            stf<T>(cm, fo(4) + zA, 0u, oA, hy, false, rng);
            stf<T>(cm, fo(5) + zD, 0u, oD, hz, false, rng);
          }
        }
      }
      __syncthreads();

      // the next iteration's inputs, when it is lean too (its loads would
      // otherwise wait at its top): planes k + 2 (E) and k + 1 (H, E
      // coefficients); the rows of this one were all read before the barrier
      if constexpr (kRp) {
        const int kn = k + 1;
        pf_ok = kn <= k_hi && kn <= kz1;
        if (pf_ok) {
          rp[0] = ldraw<T>(f.fin, fo(0) + zA + 2 * pA, oA);
          rp[1] = ldraw<T>(f.fin, fo(1) + zB + 2 * pB, oB);
          rp[2] = ldraw<T>(f.fin, fo(2) + zC + 2 * pC, oC);
          rp[3] = ldraw<T>(f.fin, fo(3) + zB + pB, oB);
          if (!halo) rp[4] = ldraw<T>(f.fin, fo(4) + zA + pA, oA);
          rp[5] = ldraw<T>(f.fin, fo(5) + zD + pD, oD);
          if constexpr (kEArr) {
            if (w >= 1) {
              rp[6] = ldraw<C>(f.cbase, co(0) + kn * nA * szC, cA);
              rp[7] = ldraw<C>(f.cbase, co(1) + kn * nB * szC, cB);
              rp[8] = ldraw<C>(f.cbase, co(2) + kn * nC * szC, cC);
            }
          }
          if (halo) {
            rp[9] = ldraw<T>(f.fin, fo(0) + zA + 2 * pA, oAt);
            rp[10] = ldraw<T>(f.fin, fo(2) + zC + 2 * pC, oCt);
          }
        }
      }
      if constexpr (kPf) {
        const int kn = k + 1;
        pf_ok = kn <= k_hi && kn <= kz1;  // (lean iterations are contiguous)
        if (pf_ok) {
          dma_row(f.fin, fo(0) + zA + 2 * pA, oA, pfw[0]);
          dma_row(f.fin, fo(1) + zB + 2 * pB, oB, pfw[1]);
          dma_row(f.fin, fo(2) + zC + 2 * pC, oC, pfw[2]);
          dma_row(f.fin, fo(3) + zB + pB, oB, pfw[3]);
          if (!halo) dma_row(f.fin, fo(4) + zA + pA, oA, pfw[4]);
          dma_row(f.fin, fo(5) + zD + pD, oD, pfw[5]);
          if constexpr (kEArr) {
            if (w >= 1) {
              dma_row(f.cbase, co(0) + kn * nA * szC, cA, pfw[6]);
              dma_row(f.cbase, co(1) + kn * nB * szC, cB, pfw[7]);
              dma_row(f.cbase, co(2) + kn * nC * szC, cC, pfw[8]);
            }
          }
          if (halo) {
            dma_row(f.fin, fo(0) + zA + 2 * pA, oAt, pfw[9]);
            dma_row(f.fin, fo(2) + zC + 2 * pC, oCt, pfw[10]);
          }
        }
      }

      // ---- first step: E'(k), then the sources between the steps ----
      float nex = 0.0f, ney = 0.0f, nez = 0.0f;
      if (gE1) {
        const float hz_dn = lH1[bb][w - 1][0][lane];
        const float hx_dn = lH1[bb][w - 1][1][lane];
        const float hz_l = lane_shift(hz, up), hy_l = lane_shift(hy, up);
        nex = ec[0] * ex + (s0 * idx) * ((hz - hz_dn) - (hy - h1y));
        ney = ec[1] * ey + (s1 * idx) * ((hx - h1x) - (hz - hz_l));
        nez = ec[2] * ez + (s2 * idx) * ((hy - hy_l) - (hx - hx_dn));
        if (swave) mid_sources(0, k, nex, ney, nez);
        if (sampled && mlane) {
          if (capp) {
            cap_store(0, k, nex);
            cap_store(1, k, ney);
            cap_store(2, k, nez);
          } else {
            const char* cm = pair_kargs()->cmid;
            stf<T>(cm, fo(0) + zA, 0u, oA, nex, false, rng);
            stf<T>(cm, fo(1) + zB, 0u, oB, ney, false, rng);
            stf<T>(cm, fo(2) + zC, 0u, oC, nez, false, rng);
          }
        }
        if (swave) mid_sources(1, k, nex, ney, nez);
        lE2[bb][w][0][lane] = nez;
        lE2[bb][w][1][lane] = nex;
      }

      // ---- second step: H''(k-1) ----
      float gx = 0.0f, gy = 0.0f, gz = 0.0f;
      if (gH2) {
        const float ez_up2 = lE2[bb ^ 1][w + 1][0][lane];
        const float ex_up2 = lE2[bb ^ 1][w + 1][1][lane];
        const float ez_r2 = lane_shift(e1z, dn), ey_r2 = lane_shift(e1y, dn);
        gx = hc[0] * h1x - hci[0] * ((ez_up2 - e1z) - (ney - e1y));
        gy = hc[1] * h1y - hci[1] * ((nex - e1x) - (ez_r2 - e1z));
        gz = hc[2] * h1z - hci[2] * ((ey_r2 - e1y) - (ex_up2 - e1x));
        if (kHSrc && swave) mid_sources(3, k - 1, gx, gy, gz);
        lH2[bb][w][0][lane] = gz;
        lH2[bb][w][1][lane] = gx;
        if (own && k - 1 >= kz0) {
          stf<T>(f.fout, fo(3) + zB - pB, 0u, oB, gx, sr, rng);
          stf<T>(f.fout, fo(4) + zA - pA, 0u, oA, gy, sr, rng);
          stf<T>(f.fout, fo(5) + zD - pD, 0u, oD, gz, sr, rng);
        }
      }

      // ---- second step: E''(k-2) ----
      if (gE2) {
        const float hz_dn = lH2[bb ^ 1][w - 1][0][lane];
        const float hx_dn = lH2[bb ^ 1][w - 1][1][lane];
        const float hz_l = lane_shift(g2z, up), hy_l = lane_shift(g2y, up);
        const float fx = ec[0] * e2x + (sb0 * idx) * ((g2z - hz_dn) - (g2y - g3y));
        const float fy = ec[1] * e2y + (sb1 * idx) * ((g2x - g3x) - (g2z - hz_l));
        const float fz = ec[2] * e2z + (sb2 * idx) * ((g2y - hy_l) - (g2x - hx_dn));
        if (own) {
          stf<T>(f.fout, fo(0) + zA - 2 * pA, 0u, oA, fx, sr, rng);
          stf<T>(f.fout, fo(1) + zB - 2 * pB, 0u, oB, fy, sr, rng);
          stf<T>(f.fout, fo(2) + zC - 2 * pC, 0u, oC, fz, sr, rng);
        }
      }

      // the CPML carries (mh, me1, me2) are not advanced: the planes this
      // iteration touches have no z term, and the general iterations that
      // follow read them only on planes that have one
      g3x = g2x;
      g3y = g2y;
      g2x = gx;
      g2y = gy;
      g2z = gz;
      e2x = e1x;
      e2y = e1y;
      e2z = e1z;
      e1x = nex;
      e1y = ney;
      e1z = nez;
      h1x = hx;
      h1y = hy;
      h1z = hz;
      sb0 = sa0;
      sb1 = sa1;
      sb2 = sa2;
      sa0 = s0;
      sa1 = s1;
      sa2 = s2;
      ex = exn;
      ey = eyn;
      ez = ezn;
    } else {
    const int bb = k & 1;
    const bool g1 = k <= kz1;
    const bool gE1 = g1 && k >= kE && w >= 1;
    const bool gH2 = k - 1 >= kE && k - 1 < kz1 && w >= 1 && w <= kW - 2;
    const bool gE2 = k - 2 >= kz0 && k - 2 < kz1 && w >= 2 && w <= kW - 2;
    const bool sampled = mwave && sample_plane(k) && mlane;

    // ---- first step: H'(k) ----
    float exn = 0.0f, eyn = 0.0f, ezn = 0.0f;
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
    float q1[6] = {}, pe[6] = {};
    float s0 = es0, s1 = es1, s2 = es2;
    if (g1) {
      float etx = 0.0f, etz = 0.0f;
      load_e(k + 1, exn, eyn, ezn);
      if (halo) load_top(k + 1, etx, etz);
      const uint32_t kk = min(k, nz - 1), kh = min(k, nz);
      const float ha0 = ldu<T>(f.fin, fo(3), (kk * ny + yY) * (nx + 1), tY);
      const float ha1 = halo ? 0.0f : ldu<T>(f.fin, fo(4), (kk * (ny + 1) + yX) * nx, tX);
      const float ha2 = ldu<T>(f.fin, fo(5), (kh * ny + yY) * nx, tX);
      float ph[6] = {};
      load_psi_h(k, ph);
      if (gE1) {
        if constexpr (kEArr) {
          const uint32_t kx = min(k, nz), kzc = min(k, nz - 1);
          s0 = ldu<C>(f.cbase, co(0), (kx * (ny + 1) + yX) * nx, cX);
          s1 = ldu<C>(f.cbase, co(1), (kx * ny + yY) * (nx + 1), cY);
          s2 = ldu<C>(f.cbase, co(2), (kzc * (ny + 1) + yX) * (nx + 1), cY);
        }
        load_psi_e(k, pe);
      }
      const float ez_up = lE1[bb][w + 1][0][lane];
      const float ex_up = lE1[bb][w + 1][1][lane];
      h_update(k, ez, ey, ex, eyn, exn, ez_up, ex_up, ha0, ha1, ha2, ph, q1, hx, hy, hz);
      if (kHSrc && swave) mid_sources(2, k, hx, hy, hz);
      lH1[bb][w][0][lane] = hz;
      lH1[bb][w][1][lane] = hx;
      lE1[bb ^ 1][w][0][lane] = ezn;
      lE1[bb ^ 1][w][1][lane] = exn;
      if (halo) {
        lE1[bb ^ 1][kW][0][lane] = etz;
        lE1[bb ^ 1][kW][1][lane] = etx;
      }
      if (sampled && capp) {
        cap_store(3, k, hx);
        cap_store(4, k, hy);
        cap_store(5, k, hz);
      } else if (sampled) {
        store_h(pair_kargs()->cmid, k, hx, hy, hz);
      }
    }
    __syncthreads();

    // ---- first step: E'(k), then the sources between the steps ----
    float nex = 0.0f, ney = 0.0f, nez = 0.0f;
    float q2[6] = {};
    if (gE1) {
      const float hz_dn = lH1[bb][w - 1][0][lane];
      const float hx_dn = lH1[bb][w - 1][1][lane];
      e_update(k, hx, hy, hz, hz_dn, hx_dn, h1x, h1y, ex, ey, ez, s0, s1, s2, pe, q2, nex, ney,
               nez);
      if (swave) mid_sources(0, k, nex, ney, nez);
      if (sampled && capp) {
        cap_store(0, k, nex);
        cap_store(1, k, ney);
        cap_store(2, k, nez);
      } else if (sampled) {
        store_e(pair_kargs()->cmid, k, nex, ney, nez);
      }
      if (swave) mid_sources(1, k, nex, ney, nez);
      lE2[bb][w][0][lane] = nez;
      lE2[bb][w][1][lane] = nex;
    }

    // ---- second step: H''(k-1) ----
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    if (gH2) {
      const int k1 = k - 1;
      const float ez_up = lE2[bb ^ 1][w + 1][0][lane];
      const float ex_up = lE2[bb ^ 1][w + 1][1][lane];
      float q3[6] = {};
      h_update(k1, e1z, e1y, e1x, ney, nex, ez_up, ex_up, h1x, h1y, h1z, mh, q3, gx, gy, gz);
      if (kHSrc && swave) mid_sources(3, k1, gx, gy, gz);
      lH2[bb][w][0][lane] = gz;
      lH2[bb][w][1][lane] = gx;
      if (own && k1 >= kz0) {
        store_h(f.fout, k1, gx, gy, gz);
        if (L > 0) store_psi_h(k1, q3);
      }
    }

    // ---- second step: E''(k-2) ----
    if (gE2) {
      const int k2 = k - 2;
      const float hz_dn = lH2[bb ^ 1][w - 1][0][lane];
      const float hx_dn = lH2[bb ^ 1][w - 1][1][lane];
      float q4[6] = {};
      float fx, fy, fz;
      e_update(k2, g2x, g2y, g2z, hz_dn, hx_dn, g3x, g3y, e2x, e2y, e2z, sb0, sb1, sb2, me2, q4,
               fx, fy, fz);
      if (own) {
        store_e(f.fout, k2, fx, fy, fz);
        if (L > 0) store_psi_e(k2, q4);
      }
    }

    g3x = g2x;
    g3y = g2y;
    g2x = gx;
    g2y = gy;
    g2z = gz;
    e2x = e1x;
    e2y = e1y;
    e2z = e1z;
    e1x = nex;
    e1y = ney;
    e1z = nez;
    h1x = hx;
    h1y = hy;
    h1z = hz;
    sb0 = sa0;
    sb1 = sa1;
    sb2 = sa2;
    sa0 = s0;
    sa1 = s1;
    sa2 = s2;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      me2[t] = me1[t];
      me1[t] = q2[t];
      mh[t] = q1[t];
    }
    ex = exn;
    ey = eyn;
    ez = ezn;
    }  // general iteration
  }
}

// One launch for both boxes: the x-edge workgroups first (the slower ones,
// so that they do not form the tail), then the x-interior ones.  Both
// variants run one 16-wave workgroup per CU, so merging costs no occupancy and
// the two boxes overlap instead of running back to back.
template <typename T, typename P, typename C, bool kEArr, bool kHSrc, bool kUni, bool kWide>
__global__ void __launch_bounds__(64 * kPairWaves) pair_step(PairParams p) {
  __shared__ PairLds lds;
  __shared__ PairPrefetch pf;
  int b = blockIdx.x, i = 0;
  while (i + 1 < p.nbox && b >= p.box[i].nblk) b -= p.box[i++].nblk;
  const int kind = p.box[i].xedge;  // 0 x-interior, 1 x-edge, 2 core
  if (kind == 1) pair_body<T, P, C, kEArr, true, kHSrc, false, kUni, kWide>(p, lds, nullptr, i, b);
  else if (kind == 2) pair_body<T, P, C, kEArr, false, kHSrc, true, true, kWide>(p, lds, &pf, i, b);
  else pair_body<T, P, C, kEArr, false, kHSrc, false, kUni, kWide>(p, lds, nullptr, i, b);
}

// The core kernel: x-interior tiles whose rows all lie clear of the y faces
// and y slabs, on z chunks whose iterations all lie clear of the z faces and z
// slabs.  Only the lean iteration is compiled, so its registers are not
// sized by the general one.
template <typename T, typename P, typename C, bool kEArr, bool kHSrc, bool kWide>
__global__ void __launch_bounds__(64 * kPairWaves) pair_core(PairParams p) {
  __shared__ PairLds lds;
  __shared__ PairPrefetch pf;
  pair_body<T, P, C, kEArr, false, kHSrc, true, true, kWide>(p, lds, &pf, 0, blockIdx.x);
}

// One CPML thickness on every face (or none): the kUni variant of the shell.
bool cpml_uniform(const FusedParams& f) {
  if (f.L == 0) return true;
  for (int i = 0; i < 6; ++i)
    if (f.clo[i] != f.L || f.chi[i] != f.L) return false;
  return f.tab_st == 2 * f.L;
}

// The wide variants (64-bit offsets) exist for float32 fields, memories and
// coefficients; their shell takes every CPML layout (no kUni variant).
template <typename T, typename P, typename C>
constexpr bool kWideTypes =
    std::is_same<T, float>::value && std::is_same<P, float>::value && std::is_same<C, float>::value;

template <typename T, typename P, typename C, bool kEArr, bool kHSrc>
void launch_pair_shell(hipStream_t s, const PairParams& p, dim3 grid, dim3 block) {
  if (p.f.wide) {
    if constexpr (kWideTypes<T, P, C>) {
      hipLaunchKernelGGL((pair_step<T, P, C, kEArr, kHSrc, false, true>), grid, block, 0, s, p);
      return;
    }
    throw Failure{"the wide pair step takes float32 fields, CPML memories and coefficients"};
  }
  if (cpml_uniform(p.f))
    hipLaunchKernelGGL((pair_step<T, P, C, kEArr, kHSrc, true, false>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((pair_step<T, P, C, kEArr, kHSrc, false, false>), grid, block, 0, s, p);
}

template <typename T, typename P, typename C, bool kEArr, bool kHSrc>
void launch_pair_core(hipStream_t s, const PairParams& p, dim3 grid, dim3 block) {
  if (p.f.wide) {
    if constexpr (kWideTypes<T, P, C>) {
      hipLaunchKernelGGL((pair_core<T, P, C, kEArr, kHSrc, true>), grid, block, 0, s, p);
      return;
    }
    throw Failure{"the wide pair step takes float32 fields, CPML memories and coefficients"};
  }
  hipLaunchKernelGGL((pair_core<T, P, C, kEArr, kHSrc, false>), grid, block, 0, s, p);
}

template <typename T, typename P, typename C>
void launch_pair_wg(hipStream_t s, const PairParams& p, int nblocks) {
  const dim3 block(64 * kPairWaves, 1, 1), grid(nblocks, 1, 1);
  const bool earr = p.f.esrc_st[0] != 0;
  const bool hsrc = p.msrc_n > p.msrc_e;  // H-timing slabs in the list
  if (p.core) {
    if (hsrc) {
      if (earr) launch_pair_core<T, P, C, true, true>(s, p, grid, block);
      else launch_pair_core<T, P, float, false, true>(s, p, grid, block);
    } else {
      if (earr) launch_pair_core<T, P, C, true, false>(s, p, grid, block);
      else launch_pair_core<T, P, float, false, false>(s, p, grid, block);
    }
  } else if (hsrc) {
    if (earr) launch_pair_shell<T, P, C, true, true>(s, p, grid, block);
    else launch_pair_shell<T, P, float, false, true>(s, p, grid, block);
  } else {
    if (earr) launch_pair_shell<T, P, C, true, false>(s, p, grid, block);
    else launch_pair_shell<T, P, float, false, false>(s, p, grid, block);
  }
}

template <typename T, typename P>
void launch_pair_coef(hipStream_t s, const PairParams& p, int nblocks) {
  if constexpr (std::is_same<T, float>::value) {
    launch_pair_wg<T, P, float>(s, p, nblocks);
  } else {
    if (p.f.coef_f16) launch_pair_wg<T, P, __half>(s, p, nblocks);
    else launch_pair_wg<T, P, float>(s, p, nblocks);
  }
}

template <typename T>
void launch_pair_field(hipStream_t s, const PairParams& p, int nblocks) {
  if (p.f.psi_bf16) launch_pair_coef<T, bf16_t>(s, p, nblocks);
  else launch_pair_coef<T, float>(s, p, nblocks);
}


}  // namespace
}  // namespace bzh

#endif  // BZH_FUSED_KERNELS_H
