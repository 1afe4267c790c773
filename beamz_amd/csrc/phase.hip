// Per-phase Yee updates for gfx950: one launch advances the three H (or E)
// components of a 3D grid, or the TMxy pair / Ez of a 2D plane, including the
// packed CPML recurrences and the PEC face rule.
//
// Numerical contract (reference files):
//   3D H: beamz/simulation/kernels.py:458-503, 267-290; lattice.py:467-473
//   3D E: kernels.py:554-626, 326-388; ghost rule lattice.py:523-575
//   TMxy: kernels.py:731-811, 1217-1290
//   CPML: kernels.py:184-264 (packed slabs), PEC: _boundary_compile.py:118-160
//
// Mapping: wave64 along x (fastest storage axis), block 64x4, one thread per
// (k, j, i) of the union of the three component supports; each thread updates
// every component whose support contains it.  All field indices are 32-bit:
// the host rejects arrays of 2^31 elements or more.

#include "common.h"

namespace bzh {
namespace {

constexpr int kBX = 64;
constexpr int kBY = 4;

template <typename T>
__device__ __forceinline__ float ld(const void* p, int i) {
  return load_f(static_cast<const T*>(p), i);
}

__device__ __forceinline__ float metric_at(const PhaseParams& p, int axis, int c) {
  const float* m = p.metric[axis];
  return m == nullptr ? p.inv_dx : m[c * p.met_st[axis]];
}

// Packed CPML correction of one derivative (kernels.py:184-264).  ``c`` is the
// coordinate along the term's axis, ``lin_of`` maps a packed coordinate to the
// psi linear index.  Returns d/kappa + psi inside the slab and d outside.
template <typename F>
__device__ __forceinline__ float cpml_correct(const PhaseParams& p, int t, float d,
                                              int c, F lin_of) {
  const DevTerm& tm = p.term[t];
  int low, high;
  if (p.uniform_l > 0) {
    low = high = p.uniform_l;
  } else {
    low = p.meta[t * 5 + 2];
    high = p.meta[t * 5 + 3];
  }
  int packed = -1;
  if (c < low) {
    packed = c;
  } else if (c >= tm.n - high) {
    packed = low + c - (tm.n - high);
  }
  // tm.low/tm.high carry the psi extent split validated on the host; a packed
  // index outside [0, low + high) is never dereferenced.
  if (packed < 0 || packed >= tm.low + tm.high) return d;
  const int q = lin_of(packed);
  float psi;
  if (tm.psi_bf16) {
    psi = load_f(static_cast<const bf16_t*>(tm.psi), q);
  } else {
    psi = static_cast<const float*>(tm.psi)[q];
  }
  psi = tm.b[packed] * psi + tm.a[packed] * d;
  if (tm.psi_bf16) {
    store_f(static_cast<bf16_t*>(tm.psi), q, psi);
  } else {
    static_cast<float*>(tm.psi)[q] = psi;
  }
  return d * tm.ik[packed] + psi;
}

__device__ __forceinline__ float coef(const float* arr, int st, float s, int i) {
  return arr == nullptr ? s : arr[i * st];
}
// E source coefficient c of cell o: scf is a table indexed by the cell's
// 8-bit ID (four IDs per int32 word, lowest byte first; compile.py:146-168)
// when scf_ids[c] is set, else as coef
__device__ __forceinline__ float scf_at(const PhaseParams& p, int c, int o) {
  if (p.scf_ids[c] != nullptr) {
    const uint32_t word = static_cast<uint32_t>(p.scf_ids[c][o >> 2]);
    return p.scf[c][(word >> ((o & 3) * 8)) & 0xFFu];
  }
  return coef(p.scf[c], p.scf_st[c], p.scf_s[c], o);
}

// ---------------------------------------------------------------------------
// 3D H phase
// ---------------------------------------------------------------------------
template <typename T, bool kCpml>
__global__ void __launch_bounds__(kBX * kBY)
h_phase_3d(PhaseParams p) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int k = blockIdx.z;
  const int nz = p.nz, ny = p.ny, nx = p.nx;
  // source (E) strides
  const int exy = nx, exz = (ny + 1) * nx;             // Ex (nz+1, ny+1, nx)
  const int eyy = nx + 1, eyz = ny * (nx + 1);         // Ey (nz+1, ny, nx+1)
  const int ezy = nx + 1, ezz = (ny + 1) * (nx + 1);   // Ez (nz, ny+1, nx+1)
  const void* Ex = p.src[0];
  const void* Ey = p.src[1];
  const void* Ez = p.src[2];

  // Hx (nz, ny, nx+1): d0 = dEz/dy (+), d1 = dEy/dz (-)
  if (k < nz && j < ny && i <= nx) {
    const int o = (k * ny + j) * (nx + 1) + i;
    const int ez = k * ezz + j * ezy + i;
    const int ey = k * eyz + j * eyy + i;
    float d0 = (ld<T>(Ez, ez + ezy) - ld<T>(Ez, ez)) * metric_at(p, 1, j);
    float d1 = (ld<T>(Ey, ey + eyz) - ld<T>(Ey, ey)) * metric_at(p, 0, k);
    if constexpr (kCpml) {
      d0 = cpml_correct(p, 0, d0, j, [&](int q) { return (k * (p.term[0].low + p.term[0].high) + q) * (nx + 1) + i; });
      d1 = cpml_correct(p, 1, d1, k, [&](int q) { return (q * ny + j) * (nx + 1) + i; });
    }
    const float curl = d0 - d1;
    float h = coef(p.dec[0], p.dec_st[0], p.dec_s[0], o) * ld<T>(p.tin[0], o) -
              coef(p.scf[0], p.scf_st[0], p.scf_s[0], o) * p.ratio * curl;
    if (((p.metallic & kLeft) && i == 0) || ((p.metallic & kRight) && i == nx)) h = 0.0f;
    store_f(static_cast<T*>(p.tout[0]), o, h);
  }
  // Hy (nz, ny+1, nx): d2 = dEx/dz (+), d3 = dEz/dx (-)
  if (k < nz && j <= ny && i < nx) {
    const int o = (k * (ny + 1) + j) * nx + i;
    const int ex = k * exz + j * exy + i;
    const int ez = k * ezz + j * ezy + i;
    float d2 = (ld<T>(Ex, ex + exz) - ld<T>(Ex, ex)) * metric_at(p, 0, k);
    float d3 = (ld<T>(Ez, ez + 1) - ld<T>(Ez, ez)) * metric_at(p, 2, i);
    if constexpr (kCpml) {
      d2 = cpml_correct(p, 2, d2, k, [&](int q) { return (q * (ny + 1) + j) * nx + i; });
      d3 = cpml_correct(p, 3, d3, i, [&](int q) { return (k * (ny + 1) + j) * (p.term[3].low + p.term[3].high) + q; });
    }
    const float curl = d2 - d3;
    float h = coef(p.dec[1], p.dec_st[1], p.dec_s[1], o) * ld<T>(p.tin[1], o) -
              coef(p.scf[1], p.scf_st[1], p.scf_s[1], o) * p.ratio * curl;
    if (((p.metallic & kBottom) && j == 0) || ((p.metallic & kTop) && j == ny)) h = 0.0f;
    store_f(static_cast<T*>(p.tout[1]), o, h);
  }
  // Hz (nz+1, ny, nx): d4 = dEy/dx (+), d5 = dEx/dy (-)
  if (k <= nz && j < ny && i < nx) {
    const int o = (k * ny + j) * nx + i;
    const int ey = k * eyz + j * eyy + i;
    const int ex = k * exz + j * exy + i;
    float d4 = (ld<T>(Ey, ey + 1) - ld<T>(Ey, ey)) * metric_at(p, 2, i);
    float d5 = (ld<T>(Ex, ex + exy) - ld<T>(Ex, ex)) * metric_at(p, 1, j);
    if constexpr (kCpml) {
      d4 = cpml_correct(p, 4, d4, i, [&](int q) { return (k * ny + j) * (p.term[4].low + p.term[4].high) + q; });
      d5 = cpml_correct(p, 5, d5, j, [&](int q) { return (k * (p.term[5].low + p.term[5].high) + q) * nx + i; });
    }
    const float curl = d4 - d5;
    float h = coef(p.dec[2], p.dec_st[2], p.dec_s[2], o) * ld<T>(p.tin[2], o) -
              coef(p.scf[2], p.scf_st[2], p.scf_s[2], o) * p.ratio * curl;
    if (((p.metallic & kFront) && k == 0) || ((p.metallic & kBack) && k == nz)) h = 0.0f;
    store_f(static_cast<T*>(p.tout[2]), o, h);
  }
}

// Backward difference of ghost-padded H along one axis (lattice.py:523-575):
// ``idx`` is the linear index of coordinate ``c`` (valid only when c < N).
template <typename T>
__device__ __forceinline__ float ghost_bdiff(const void* h, int idx, int stride, int c,
                                             int N, bool metal_lo, bool metal_hi) {
  const float up = (c < N) ? ld<T>(h, idx) : (metal_hi ? 0.0f : ld<T>(h, idx - stride));
  const float lo = (c > 0) ? ld<T>(h, idx - stride) : (metal_lo ? 0.0f : ld<T>(h, idx));
  return up - lo;
}

// ---------------------------------------------------------------------------
// 3D E phase
// ---------------------------------------------------------------------------
template <typename T, bool kCpml>
__global__ void __launch_bounds__(kBX * kBY)
e_phase_3d(PhaseParams p) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int k = blockIdx.z;
  const int nz = p.nz, ny = p.ny, nx = p.nx;
  const int m = p.metallic;
  const void* Hx = p.src[0];  // (nz, ny, nx+1)
  const void* Hy = p.src[1];  // (nz, ny+1, nx)
  const void* Hz = p.src[2];  // (nz+1, ny, nx)
  const int hxy = nx + 1, hxz = ny * (nx + 1);
  const int hyy = nx, hyz = (ny + 1) * nx;
  const int hzy = nx, hzz = ny * nx;

  // Ex (nz+1, ny+1, nx): d0 = d(Hz)/dy (+), d1 = d(Hy)/dz (-)
  if (k <= nz && j <= ny && i < nx) {
    const int o = (k * (ny + 1) + j) * nx + i;
    float d0 = ghost_bdiff<T>(Hz, k * hzz + j * hzy + i, hzy, j, ny, m & kBottom, m & kTop) *
               metric_at(p, 1, j);
    float d1 = ghost_bdiff<T>(Hy, k * hyz + j * hyy + i, hyz, k, nz, m & kFront, m & kBack) *
               metric_at(p, 0, k);
    if constexpr (kCpml) {
      d0 = cpml_correct(p, 0, d0, j, [&](int q) { return (k * (p.term[0].low + p.term[0].high) + q) * nx + i; });
      d1 = cpml_correct(p, 1, d1, k, [&](int q) { return (q * (ny + 1) + j) * nx + i; });
    }
    const float curl = d0 - d1;
    float e = coef(p.dec[0], p.dec_st[0], p.dec_s[0], o) * ld<T>(p.tin[0], o) +
              scf_at(p, 0, o) * p.ratio * curl;
    if (((m & kFront) && k == 0) || ((m & kBack) && k == nz) ||
        ((m & kBottom) && j == 0) || ((m & kTop) && j == ny))
      e = 0.0f;
    store_f(static_cast<T*>(p.tout[0]), o, e);
  }
  // Ey (nz+1, ny, nx+1): d2 = d(Hx)/dz (+), d3 = d(Hz)/dx (-)
  if (k <= nz && j < ny && i <= nx) {
    const int o = (k * ny + j) * (nx + 1) + i;
    float d2 = ghost_bdiff<T>(Hx, k * hxz + j * hxy + i, hxz, k, nz, m & kFront, m & kBack) *
               metric_at(p, 0, k);
    float d3 = ghost_bdiff<T>(Hz, k * hzz + j * hzy + i, 1, i, nx, m & kLeft, m & kRight) *
               metric_at(p, 2, i);
    if constexpr (kCpml) {
      d2 = cpml_correct(p, 2, d2, k, [&](int q) { return (q * ny + j) * (nx + 1) + i; });
      d3 = cpml_correct(p, 3, d3, i, [&](int q) { return (k * ny + j) * (p.term[3].low + p.term[3].high) + q; });
    }
    const float curl = d2 - d3;
    float e = coef(p.dec[1], p.dec_st[1], p.dec_s[1], o) * ld<T>(p.tin[1], o) +
              scf_at(p, 1, o) * p.ratio * curl;
    if (((m & kFront) && k == 0) || ((m & kBack) && k == nz) ||
        ((m & kLeft) && i == 0) || ((m & kRight) && i == nx))
      e = 0.0f;
    store_f(static_cast<T*>(p.tout[1]), o, e);
  }
  // Ez (nz, ny+1, nx+1): d4 = d(Hy)/dx (+), d5 = d(Hx)/dy (-)
  if (k < nz && j <= ny && i <= nx) {
    const int o = (k * (ny + 1) + j) * (nx + 1) + i;
    float d4 = ghost_bdiff<T>(Hy, k * hyz + j * hyy + i, 1, i, nx, m & kLeft, m & kRight) *
               metric_at(p, 2, i);
    float d5 = ghost_bdiff<T>(Hx, k * hxz + j * hxy + i, hxy, j, ny, m & kBottom, m & kTop) *
               metric_at(p, 1, j);
    if constexpr (kCpml) {
      d4 = cpml_correct(p, 4, d4, i, [&](int q) { return (k * (ny + 1) + j) * (p.term[4].low + p.term[4].high) + q; });
      d5 = cpml_correct(p, 5, d5, j, [&](int q) { return (k * (p.term[5].low + p.term[5].high) + q) * (nx + 1) + i; });
    }
    const float curl = d4 - d5;
    float e = coef(p.dec[2], p.dec_st[2], p.dec_s[2], o) * ld<T>(p.tin[2], o) +
              scf_at(p, 2, o) * p.ratio * curl;
    if (((m & kBottom) && j == 0) || ((m & kTop) && j == ny) ||
        ((m & kLeft) && i == 0) || ((m & kRight) && i == nx))
      e = 0.0f;
    store_f(static_cast<T*>(p.tout[2]), o, e);
  }
}

// ---------------------------------------------------------------------------
// 2D TMxy (Ez, Hx, Hy) on (ny, nx): kernels.py:731-811, 1217-1290
// ---------------------------------------------------------------------------

// The folded timed source slabs of target ``comp`` at cell (j, i), added to
// the updated value v in declaration order (as successive source launches
// add them, to rounding); float32 fields only.
__device__ __forceinline__ float tm_sources(const PhaseParams& p, int comp, int j, int i, float v) {
  for (int g = 0; g < p.npsrc; ++g) {
    const PhaseSrc& s = p.psrc[g];
    if (s.comp != comp) continue;
    if ((s.mlo[1] && j == 0) || (s.mhi[1] && j == s.fy - 1) || (s.mlo[2] && i == 0) ||
        (s.mhi[2] && i == s.fx - 1))
      continue;
    int widx = *p.step0 + p.k;
    widx = widx < 0 ? 0 : (widx > s.T - 1 ? s.T - 1 : widx);
    for (int q = 0; q < s.n; ++q) {
      const int cz = -s.starts[q * 3 + 0];  // the 2D fields' single plane z = 0
      const int cy = j - s.starts[q * 3 + 1], cx = i - s.starts[q * 3 + 2];
      if (cz < 0 || cz >= s.sz || cy < 0 || cy >= s.sy || cx < 0 || cx >= s.sx) continue;
      const int64_t c = ((static_cast<int64_t>(q) * s.sz + cz) * s.sy + cy) * s.sx + cx;
      v = v + s.coeff[c] * s.wave[static_cast<int64_t>(q) * s.T + widx];
    }
  }
  return v;
}
template <typename T, bool kCpml>
__global__ void __launch_bounds__(kBX * kBY)
h_phase_tm(PhaseParams p) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int ny = p.ny, nx = p.nx;
  const int m = p.metallic;
  const void* Ez = p.src[2];  // (ny+1, nx+1)
  // Hx (ny, nx+1): corrected(+dEz/dy)
  if (j < ny && i <= nx) {
    const int o = j * (nx + 1) + i;
    const int ez = j * (nx + 1) + i;
    float d = (ld<T>(Ez, ez + nx + 1) - ld<T>(Ez, ez)) * metric_at(p, 1, j);
    if constexpr (kCpml) {
      d = cpml_correct(p, 0, d, j, [&](int q) { return q * (nx + 1) + i; });
    }
    float h = coef(p.dec[0], p.dec_st[0], p.dec_s[0], o) * ld<T>(p.tin[0], o) -
              coef(p.scf[0], p.scf_st[0], p.scf_s[0], o) * p.ratio * d;
    if (((m & kLeft) && i == 0) || ((m & kRight) && i == nx)) h = 0.0f;
    if (p.npsrc > 0) h = tm_sources(p, 0, j, i, h);
    store_f(static_cast<T*>(p.tout[0]), o, h);
  }
  // Hy (ny+1, nx): corrected(-dEz/dx)
  if (j <= ny && i < nx) {
    const int o = j * nx + i;
    const int ez = j * (nx + 1) + i;
    float d = (ld<T>(Ez, ez + 1) - ld<T>(Ez, ez)) * metric_at(p, 2, i);
    if constexpr (kCpml) {
      d = cpml_correct(p, 1, d, i, [&](int q) { return j * (p.term[1].low + p.term[1].high) + q; });
    }
    float h = coef(p.dec[1], p.dec_st[1], p.dec_s[1], o) * ld<T>(p.tin[1], o) +
              coef(p.scf[1], p.scf_st[1], p.scf_s[1], o) * p.ratio * d;
    if (((m & kBottom) && j == 0) || ((m & kTop) && j == ny)) h = 0.0f;
    if (p.npsrc > 0) h = tm_sources(p, 1, j, i, h);
    store_f(static_cast<T*>(p.tout[1]), o, h);
  }
}

template <typename T, bool kCpml>
__global__ void __launch_bounds__(kBX * kBY)
e_phase_tm(PhaseParams p) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int ny = p.ny, nx = p.nx;
  const int m = p.metallic;
  if (j > ny || i > nx) return;
  const void* Hx = p.src[0];  // (ny, nx+1)
  const void* Hy = p.src[1];  // (ny+1, nx)
  const int o = j * (nx + 1) + i;
  float d0 = ghost_bdiff<T>(Hy, j * nx + i, 1, i, nx, m & kLeft, m & kRight) * metric_at(p, 2, i);
  float d1 = ghost_bdiff<T>(Hx, j * (nx + 1) + i, nx + 1, j, ny, m & kBottom, m & kTop) *
             metric_at(p, 1, j);
  if constexpr (kCpml) {
    d0 = cpml_correct(p, 0, d0, i, [&](int q) { return j * (p.term[0].low + p.term[0].high) + q; });
    d1 = cpml_correct(p, 1, d1, j, [&](int q) { return q * (nx + 1) + i; });
  }
  const float curl = d0 - d1;
  float e = coef(p.dec[2], p.dec_st[2], p.dec_s[2], o) * ld<T>(p.tin[2], o) +
            scf_at(p, 2, o) * p.ratio * curl;
  if (((m & kBottom) && j == 0) || ((m & kTop) && j == ny) ||
      ((m & kLeft) && i == 0) || ((m & kRight) && i == nx))
    e = 0.0f;
  if (p.npsrc > 0) e = tm_sources(p, 2, j, i, e);
  store_f(static_cast<T*>(p.tout[2]), o, e);
}

template <typename T>
void dispatch3d(hipStream_t s, const PhaseParams& p, int phase) {
  dim3 block(kBX, kBY, 1);
  dim3 grid((p.nx + 1 + kBX - 1) / kBX, (p.ny + 1 + kBY - 1) / kBY, p.nz + 1);
  const bool cpml = p.nterms > 0;
  if (phase == BZH_PHASE_H) {
    if (cpml) hipLaunchKernelGGL((h_phase_3d<T, true>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((h_phase_3d<T, false>), grid, block, 0, s, p);
  } else {
    if (cpml) hipLaunchKernelGGL((e_phase_3d<T, true>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((e_phase_3d<T, false>), grid, block, 0, s, p);
  }
}

template <typename T>
void dispatch_tm(hipStream_t s, const PhaseParams& p, int phase) {
  dim3 block(kBX, kBY, 1);
  dim3 grid((p.nx + 1 + kBX - 1) / kBX, (p.ny + 1 + kBY - 1) / kBY, 1);
  const bool cpml = p.nterms > 0;
  if (phase == BZH_PHASE_H) {
    if (cpml) hipLaunchKernelGGL((h_phase_tm<T, true>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((h_phase_tm<T, false>), grid, block, 0, s, p);
  } else {
    if (cpml) hipLaunchKernelGGL((e_phase_tm<T, true>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((e_phase_tm<T, false>), grid, block, 0, s, p);
  }
}

}  // namespace

void launch_phase3d(hipStream_t s, const PhaseParams& p, int phase, int dtype) {
  BZH_REQUIRE(!p.scf_f16, "f16 coefficient arrays are read only by the fused step");
  BZH_REQUIRE(p.npsrc == 0, "folded source slabs are for the 2D kernels");
  switch (dtype) {
    case BZH_F32: dispatch3d<float>(s, p, phase); break;
    case BZH_BF16: dispatch3d<bf16_t>(s, p, phase); break;
    case BZH_F16: dispatch3d<__half>(s, p, phase); break;
    default: throw Failure{"unsupported field dtype"};
  }
  BZH_HIP(hipGetLastError());
}

void launch_phase_tm(hipStream_t s, const PhaseParams& p, int phase, int dtype) {
  BZH_REQUIRE(!p.scf_f16, "f16 coefficient arrays are read only by the fused step");
  BZH_REQUIRE(p.npsrc >= 0 && p.npsrc <= kMaxPhaseSrc && (p.npsrc == 0 || p.step0 != nullptr),
              "folded source slabs are malformed");
  BZH_REQUIRE(p.npsrc == 0 || dtype == BZH_F32, "folded source slabs need float32 fields");
  switch (dtype) {
    case BZH_F32: dispatch_tm<float>(s, p, phase); break;
    case BZH_BF16: dispatch_tm<bf16_t>(s, p, phase); break;
    case BZH_F16: dispatch_tm<__half>(s, p, phase); break;
    default: throw Failure{"unsupported field dtype"};
  }
  BZH_HIP(hipGetLastError());
}

}  // namespace bzh
