// Source injection and DFT-monitor accumulation kernels for gfx950.
//
//   slab sources: beamz/simulation/execute.py:374-461 (field[start:start+size]
//                 += coeff * waveform[clip(step, 0, T-1)]); post-update timings
//                 skip PEC-constrained cells (equivalent to re-masking,
//                 execute.py:488-521)
//   DFT monitors: beamz/simulation/observe.py:315-357 (gating, window),
//                 observe.py:529-585 (weighted gather and accumulation)
//
// The DFT accumulators and their phase factors are float64: the phase
// 2*pi*f*t_phys is evaluated in float64 from t_phys = t + (k + 1) dt.

#include <climits>

#include "common.h"

namespace bzh {
namespace {

template <typename T>
__global__ void __launch_bounds__(256) source_kernel(SourceParams p) {
  const int cx = blockIdx.x * 64 + threadIdx.x;
  const int cy = blockIdx.y * 4 + threadIdx.y;
  const int cz = blockIdx.z;
  if (cx >= p.sdims[2] || cy >= p.sdims[1] || cz >= p.sdims[0]) return;
  int widx = *p.step0 + p.k;
  widx = widx < 0 ? 0 : (widx > p.T - 1 ? p.T - 1 : widx);
  const int slab_cells = p.sdims[0] * p.sdims[1] * p.sdims[2];
  const int c = (cz * p.sdims[1] + cy) * p.sdims[2] + cx;
  T* f = static_cast<T*>(p.field);
  for (int s = p.slab0; s < p.slab0 + p.nslabs; ++s) {
    const int z = p.starts[s * 3 + 0] + cz;
    const int y = p.starts[s * 3 + 1] + cy;
    const int x = p.starts[s * 3 + 2] + cx;
    if (z < 0 || y < 0 || x < 0 || z >= p.fdims[0] || y >= p.fdims[1] || x >= p.fdims[2]) continue;
    if ((p.mlo[0] && z == 0) || (p.mhi[0] && z == p.fdims[0] - 1) ||
        (p.mlo[1] && y == 0) || (p.mhi[1] && y == p.fdims[1] - 1) ||
        (p.mlo[2] && x == 0) || (p.mhi[2] && x == p.fdims[2] - 1))
      continue;
    const int o = (z * p.fdims[1] + y) * p.fdims[2] + x;
    const float add = p.coeff[static_cast<int64_t>(s) * slab_cells + c] *
                      p.wave[static_cast<int64_t>(s) * p.T + widx] * p.scale;
    store_f(f, o, load_f(f, o) + add);
  }
}

constexpr int kDftBlock = 128;
constexpr int kMaxFreq = 256;

// Weighted gather of one monitor sample (observe.py:360-370, 529-585),
// from the field or, when ``p.cap`` is set, from the capture of the
// monitored cells (the corner's cell mapped into its component's box).
template <typename T>
__device__ __forceinline__ float gather(const DftParams& p, int m, int c, int pt) {
  const T* fld = static_cast<const T*>(p.field[c]);
  const int64_t base = ((static_cast<int64_t>(m) * 6 + c) * p.Pmax + pt) * p.K;
  float v = 0.0f;
  if (p.cap != nullptr) {
    const int32_t* t = p.capbox + (m * 6 + c) * 8;
    const int z0 = t[0], z1 = t[1], y0 = t[2], y1 = t[3], x0 = t[4], x1 = t[5];
    const int sy = p.fdims[c][1], sx = p.fdims[c][2];
    const int nzb = z1 - z0 + 1, nyb = y1 - y0 + 1;
    for (int q = 0; q < p.K; ++q) {
      const int idx = p.indices[base + q];
      const float w = p.weights[base + q];
      if (w == 0.0f || idx < 0 || idx >= p.fsize[c]) continue;
      const int z = idx / (sy * sx), r = idx - z * (sy * sx);
      const int y = r / sx, x = r - y * sx;
      if (z < z0 || z > z1 || y < y0 || y > y1 || x < x0 || x > x1) continue;
      v += w * p.cap[t[6] + ((x - x0) * nyb + (y - y0)) * nzb + (z - z0)];
    }
    return v;
  }
  for (int q = 0; q < p.K; ++q) {
    const int idx = p.indices[base + q];
    const float w = p.weights[base + q];
    if (w != 0.0f && idx >= 0 && idx < p.fsize[c]) v += w * load_f(fld, idx);
  }
  return v;
}

// Gate, window and amplitude factor of step k of this call for monitor m
// (observe.py:315-357).  Returns false when the step is not sampled.
__device__ __forceinline__ bool dft_gate(const DftParams& p, int m, int k, double& t,
                                         double& window, double& scale) {
  const int interval = p.counts[m * 5 + 2] > 1 ? p.counts[m * 5 + 2] : 1;
  const int step = *p.step0 + k;
  const double t0 = p.time_f64 ? *static_cast<const double*>(p.time)
                               : static_cast<double>(*static_cast<const float*>(p.time));
  t = t0 + static_cast<double>(k + 1) * p.dt;
  const double ts = p.windows[m * 3 + 0];
  const double te = p.windows[m * 3 + 1];
  if (!(t >= ts && t <= te) || (step % interval) != 0) return false;
  window = 1.0;
  if (p.codes[m * 2 + 0] == 1 && isfinite(te) && te > ts) {
    double tau = (t - ts) / fmax(te - ts, 1e-30);
    tau = fmin(fmax(tau, 0.0), 1.0);
    window = 0.5 * (1.0 - cos(2.0 * M_PI * tau));
  }
  scale = fmax(window, 0.0);
  if (p.codes[m * 2 + 1] == 1) {
    const double lu = p.windows[m * 3 + 2];
    scale *= (p.dt * interval * 299792458.0 / lu) / sqrt(2.0 * M_PI);
  }
  return true;
}

// Buffered path, step k: stage the gathered sample (f32) in ring slot k % slots.
template <typename T>
__global__ void __launch_bounds__(kDftBlock) dft_stage_kernel(DftParams p) {
  const int m = blockIdx.z / 6;
  const int c = blockIdx.z % 6;
  const int P = p.counts[m * 5 + 1];
  double t, window, scale;
  if (!dft_gate(p, m, p.k, t, window, scale)) return;
  const int pt = blockIdx.x * kDftBlock + threadIdx.x;
  if (p.masks[m * 6 + c] == 0.0f || pt >= P || pt >= p.Pmax || p.field[c] == nullptr) return;
  const int64_t slot = p.k % p.slots;
  p.stage[((slot * p.M + m) * 6 + c) * p.Pmax + pt] = gather<T>(p, m, c, pt);
}

// Buffered path: fold steps k0 .. k0+nk-1 of the ring into the accumulators.
// Each step's term is formed exactly as in the direct path (f32 sample times
// 1/scale, mask and amplitude in f64, times the f64 phase factor); the nk
// terms of one (frequency, point) are summed in f64 and added once.
template <typename A>
__global__ void __launch_bounds__(kDftBlock) dft_flush_kernel(DftParams p) {
  constexpr int kFt = 16;  // frequencies per LDS tile
  const int m = blockIdx.z / 6;
  const int c = blockIdx.z % 6;
  const int P = p.counts[m * 5 + 1];
  const int F = p.counts[m * 5 + 0];
  const int64_t voff = p.counts[m * 5 + 3];
  const int64_t woff = p.counts[m * 5 + 4];
  const int nk = p.nk < kDftMaxSlots ? p.nk : kDftMaxSlots;
  __shared__ double s_t[kDftMaxSlots], s_scale[kDftMaxSlots], s_win[kDftMaxSlots];
  __shared__ int s_on[kDftMaxSlots];
  __shared__ double s_cos[kDftMaxSlots][kFt], s_sin[kDftMaxSlots][kFt];
  if (threadIdx.x < nk) {
    double t = 0.0, window = 0.0, scale = 0.0;
    const bool on = dft_gate(p, m, p.k0 + threadIdx.x, t, window, scale);
    s_t[threadIdx.x] = t;
    s_scale[threadIdx.x] = on ? scale : 0.0;
    s_win[threadIdx.x] = on ? window : 0.0;
    s_on[threadIdx.x] = on ? 1 : 0;
  }
  __syncthreads();
  const int nf = F < p.Fmax ? F : p.Fmax;
  if (c == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    A* w = static_cast<A*>(p.wsum);
    for (int f = 0; f < nf; ++f) {
      double acc = 0.0;
      for (int j = 0; j < nk; ++j) acc += s_win[j];
      if (woff + f < p.w_len) w[woff + f] += static_cast<A>(acc);
    }
  }
  const float cmask = p.masks[m * 6 + c];
  const int pt = blockIdx.x * kDftBlock + threadIdx.x;
  const bool live = cmask != 0.0f && pt < P && pt < p.Pmax && p.field[c] != nullptr;
  double dv[kDftMaxSlots];
#pragma unroll
  for (int j = 0; j < kDftMaxSlots; ++j) {
    dv[j] = 0.0;
    if (live && j < nk && s_on[j]) {
      const int64_t slot = (p.k0 + j) % p.slots;
      const float v = p.stage[((slot * p.M + m) * 6 + c) * p.Pmax + pt];
      dv[j] = static_cast<double>(v) * p.inv_scale[c] * cmask * s_scale[j];
    }
  }
  A* re = static_cast<A*>(p.re);
  A* im = static_cast<A*>(p.im);
  for (int f0 = 0; f0 < nf; f0 += kFt) {
    __syncthreads();
    for (int e = threadIdx.x; e < nk * kFt; e += kDftBlock) {
      const int j = e / kFt, f = f0 + e % kFt;
      double sn = 0.0, cs = 0.0;
      if (f < nf) {
        const double fr = p.freqs_f64 ? static_cast<const double*>(p.freqs)[m * p.Fmax + f]
                                      : static_cast<double>(static_cast<const float*>(p.freqs)[m * p.Fmax + f]);
        sincos(2.0 * M_PI * fr * s_t[j], &sn, &cs);
      }
      s_cos[j][e % kFt] = cs;
      s_sin[j][e % kFt] = sn;
    }
    __syncthreads();
    if (!live) continue;
    for (int ff = 0; ff < kFt && f0 + ff < nf; ++ff) {
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int j = 0; j < kDftMaxSlots; ++j) {
        if (j < nk) {
          ar += dv[j] * s_cos[j][ff];
          ai += dv[j] * s_sin[j][ff];
        }
      }
      const int64_t o = voff + (static_cast<int64_t>(c) * F + f0 + ff) * P + pt;
      if (o < p.re_len) {
        re[o] += static_cast<A>(ar);
        im[o] += static_cast<A>(ai);
      }
    }
  }
}

template <typename T, typename A>
__global__ void __launch_bounds__(kDftBlock) dft_kernel(DftParams p) {
  const int m = blockIdx.z / 6;
  const int c = blockIdx.z % 6;
  const int P = p.counts[m * 5 + 1];
  const int F = p.counts[m * 5 + 0];
  const int interval = p.counts[m * 5 + 2] > 1 ? p.counts[m * 5 + 2] : 1;
  const int64_t voff = p.counts[m * 5 + 3];
  const int64_t woff = p.counts[m * 5 + 4];
  const int step = *p.step0 + p.k;
  const double t0 = p.time_f64 ? *static_cast<const double*>(p.time)
                               : static_cast<double>(*static_cast<const float*>(p.time));
  const double t = t0 + static_cast<double>(p.k + 1) * p.dt;
  const double ts = p.windows[m * 3 + 0];
  const double te = p.windows[m * 3 + 1];
  // observe.py:315-338: gate on the time window and the sampling interval
  if (!(t >= ts && t <= te) || (step % interval) != 0) return;
  double window = 1.0;
  if (p.codes[m * 2 + 0] == 1 && isfinite(te) && te > ts) {
    double tau = (t - ts) / fmax(te - ts, 1e-30);
    tau = fmin(fmax(tau, 0.0), 1.0);
    window = 0.5 * (1.0 - cos(2.0 * M_PI * tau));
  }
  double scale = fmax(window, 0.0);
  if (p.codes[m * 2 + 1] == 1) {
    // observe.py:341-357 physical normalization
    const double lu = p.windows[m * 3 + 2];
    scale *= (p.dt * interval * 299792458.0 / lu) / sqrt(2.0 * M_PI);
  }
  __shared__ double s_cos[kMaxFreq];
  __shared__ double s_sin[kMaxFreq];
  const int nf = F < p.Fmax ? F : p.Fmax;
  for (int f = threadIdx.x; f < nf && f < kMaxFreq; f += kDftBlock) {
    const double fr = p.freqs_f64 ? static_cast<const double*>(p.freqs)[m * p.Fmax + f]
                                  : static_cast<double>(static_cast<const float*>(p.freqs)[m * p.Fmax + f]);
    double sn, cs;
    sincos(2.0 * M_PI * fr * t, &sn, &cs);
    s_cos[f] = cs;
    s_sin[f] = sn;
  }
  __syncthreads();
  // weight sum: one thread per monitor adds the window for each frequency
  if (c == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    A* w = static_cast<A*>(p.wsum);
    for (int f = 0; f < nf; ++f)
      if (woff + f < p.w_len) w[woff + f] += static_cast<A>(window);
  }
  const float cmask = p.masks[m * 6 + c];
  const int pt = blockIdx.x * kDftBlock + threadIdx.x;
  if (cmask == 0.0f || pt >= P || pt >= p.Pmax || p.field[c] == nullptr) return;
  const float v = gather<T>(p, m, c, pt);
  const double dv = static_cast<double>(v) * p.inv_scale[c] * cmask * scale;
  A* re = static_cast<A*>(p.re);
  A* im = static_cast<A*>(p.im);
  for (int f = 0; f < nf && f < kMaxFreq; ++f) {
    const int64_t o = voff + (static_cast<int64_t>(c) * F + f) * P + pt;
    if (o < p.re_len) {
      re[o] += static_cast<A>(dv * s_cos[f]);
      im[o] += static_cast<A>(dv * s_sin[f]);
    }
  }
}

template <typename T>
void dft_dispatch(hipStream_t s, const DftParams& p) {
  dim3 block(kDftBlock, 1, 1);
  dim3 grid((p.Pmax + kDftBlock - 1) / kDftBlock, 1, p.M * 6);
  if (p.mode == 1) {
    hipLaunchKernelGGL((dft_stage_kernel<T>), grid, block, 0, s, p);
  } else if (p.mode == 2) {
    if (p.acc_f64) hipLaunchKernelGGL((dft_flush_kernel<double>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((dft_flush_kernel<float>), grid, block, 0, s, p);
  } else {
    if (p.acc_f64) hipLaunchKernelGGL((dft_kernel<T, double>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((dft_kernel<T, float>), grid, block, 0, s, p);
  }
}


// Bounding boxes of the monitors' gather points, for the pair step's
// intermediate bank: box m = [zmin, zmax, ymin, ymax, xmin, xmax] over every
// point the gather reads (weight != 0, index inside the component).  The
// init kernel empties the boxes; blocks (b, m) fold slices of monitor m's
// points in with atomics.
__global__ void __launch_bounds__(64) monbox_init_kernel(MonBoxParams p) {
  const int i = threadIdx.x;
  if (i < p.M * 6) p.box[i] = (i % 2 == 0) ? INT_MAX : INT_MIN;
}

constexpr int kBoxBlock = 256;
constexpr int kBoxSlices = 64;

__global__ void __launch_bounds__(kBoxBlock) monbox_kernel(MonBoxParams p) {
  __shared__ int red[6][kBoxBlock];
  const int m = blockIdx.y;
  int v[6] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  const int64_t per = static_cast<int64_t>(p.Pmax) * p.K;
  for (int c = 0; c < 6; ++c) {
    const int Sz = p.dims[c][0], Sy = p.dims[c][1], Sx = p.dims[c][2];
    const int64_t n = static_cast<int64_t>(Sz) * Sy * Sx;
    if (n <= 0) continue;
    const int64_t base = (static_cast<int64_t>(m) * 6 + c) * per;
    for (int64_t i = blockIdx.x * int64_t{kBoxBlock} + threadIdx.x; i < per;
         i += int64_t{kBoxBlock} * gridDim.x) {
      const int idx = p.indices[base + i];
      if (p.weights[base + i] == 0.0f || idx < 0 || idx >= n) continue;
      const int z = idx / (Sy * Sx);
      const int r = idx - z * (Sy * Sx);
      const int y = r / Sx;
      const int x = r - y * Sx;
      v[0] = min(v[0], z);
      v[1] = max(v[1], z);
      v[2] = min(v[2], y);
      v[3] = max(v[3], y);
      v[4] = min(v[4], x);
      v[5] = max(v[5], x);
    }
  }
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int h = kBoxBlock / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      for (int q = 0; q < 6; ++q) {
        const int a = red[q][threadIdx.x], b = red[q][threadIdx.x + h];
        red[q][threadIdx.x] = (q % 2 == 0) ? min(a, b) : max(a, b);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int q = 0; q < 6; q += 2) {
      if (red[q][0] != INT_MAX) atomicMin(&p.box[m * 6 + q], red[q][0]);
      if (red[q + 1][0] != INT_MIN) atomicMax(&p.box[m * 6 + q + 1], red[q + 1][0]);
    }
  }
}

}  // namespace

void launch_source(hipStream_t s, const SourceParams& p) {
  if (p.sdims[0] <= 0 || p.sdims[1] <= 0 || p.sdims[2] <= 0 || p.nslabs <= 0) return;
  dim3 block(64, 4, 1);
  dim3 grid((p.sdims[2] + 63) / 64, (p.sdims[1] + 3) / 4, p.sdims[0]);
  switch (p.dtype) {
    case BZH_F32: hipLaunchKernelGGL((source_kernel<float>), grid, block, 0, s, p); break;
    case BZH_BF16: hipLaunchKernelGGL((source_kernel<bf16_t>), grid, block, 0, s, p); break;
    case BZH_F16: hipLaunchKernelGGL((source_kernel<__half>), grid, block, 0, s, p); break;
    default: throw Failure{"unsupported source field dtype"};
  }
  BZH_HIP(hipGetLastError());
}

void launch_dft(hipStream_t s, const DftParams& p) {
  if (p.M <= 0 || p.Pmax <= 0) return;
  BZH_REQUIRE(p.Fmax <= kMaxFreq, "DFT monitors support at most 256 frequencies");
  if (p.mode != 0) {
    BZH_REQUIRE(p.stage != nullptr && p.slots >= 1 && p.slots <= kDftMaxSlots,
                "buffered DFT needs a staging ring of 1..16 slots");
    BZH_REQUIRE(p.mode != 2 || (p.nk >= 1 && p.nk <= p.slots), "DFT flush range exceeds the ring");
  }
  switch (p.dtype) {
    case BZH_F32: dft_dispatch<float>(s, p); break;
    case BZH_BF16: dft_dispatch<bf16_t>(s, p); break;
    case BZH_F16: dft_dispatch<__half>(s, p); break;
    default: throw Failure{"unsupported DFT field dtype"};
  }
  BZH_HIP(hipGetLastError());
}


void launch_monbox(hipStream_t s, const MonBoxParams& p) {
  BZH_REQUIRE(p.M >= 1 && p.M <= kMaxMonBoxes && p.box != nullptr, "monitor boxes are malformed");
  if (p.Pmax <= 0 || p.K <= 0) return;
  hipLaunchKernelGGL(monbox_init_kernel, dim3(1), dim3(64), 0, s, p);
  hipLaunchKernelGGL(monbox_kernel, dim3(kBoxSlices, p.M), dim3(kBoxBlock), 0, s, p);
  BZH_HIP(hipGetLastError());
}

}  // namespace bzh
