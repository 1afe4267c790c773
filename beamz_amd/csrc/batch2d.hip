// Batched TMxy steps: many independent 2D simulations of uneven sizes advance
// together, one persistent launch per time step.
//
// Work: every simulation is cut into tiles of 64 (x) by kTY (y) Ez cells; the
// tiles of all simulations form one queue (``tiles``, host-built, largest
// simulations first).  A fixed grid of workgroups pops tiles from a per-step
// counter with one atomic per tile until the queue is empty, so uneven grid
// sizes load-balance without one launch per simulation.
//
// One tile, one full leapfrog step from bank A to bank B (kernels.py:731-811,
// 1217-1290; ghost rule lattice.py:523-575; packed CPML kernels.py:184-264;
// PEC faces _boundary_compile.py:118-160):
//   1. Ez(A) rows j0-1 .. j0+kTY, columns i0-1 .. i0+64 into LDS, with the
//      pre-E slab sources of the step added on the fly (execute.py:374-461;
//      a cell gets the same sum in every tile that loads it);
//   2. Hx_new on rows j0-1 .. j0+kTY-1 and Hy_new on columns i0-1 .. i0+63
//      (the halo row/column is recomputed, not stored) into LDS;
//   3. Ez_new on the owned cells from LDS.
// CPML memories are read from bank A and written to bank B for owned cells
// only.  Between steps the kernel boundary is the only synchronization.

#include "common.h"

namespace bzh {
namespace {

__device__ __forceinline__ int slab_q2(int c, int n, int L) {
  return L <= 0 ? -1 : (c < L ? c : (c >= n - L ? c - (n - 2 * L) : -1));
}

__device__ __forceinline__ float gdiff2(float cur, float prev, int c, int N, bool mlo, bool mhi) {
  const float up = (c < N) ? cur : (mhi ? 0.0f : prev);
  const float lo = (c > 0) ? prev : (mlo ? 0.0f : cur);
  return up - lo;
}

template <int kTY>
__global__ void __launch_bounds__(256) batch2d_step(Batch2dParams p) {
  constexpr int kTX = 64;
  __shared__ float s_ez[kTY + 2][kTX + 2];
  __shared__ float s_hx[kTY + 1][kTX];
  __shared__ float s_hy[kTY][kTX + 1];
  __shared__ int s_tile;
  const int tid = threadIdx.x;
  const int step = *p.step0 + p.k;
  const int a = p.bank, bnk = p.bank ^ 1;

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(p.counter, 1);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();
    if (t >= p.ntiles) break;
    const int4 td = p.tiles[t];
    const Sim2dDev& S = p.sims[td.x];
    const int ny = S.ny, nx = S.nx, L = S.L, L2 = 2 * S.L, m = S.metallic;
    const bool mb = m & kBottom, mt = m & kTop, ml = m & kLeft, mr = m & kRight;
    const float idx = S.inv_dx;
    const float h_dec = *S.h_dec, h_src = *S.h_src, e_dec = *S.e_dec;
    const int j0 = td.y * kTY, i0 = td.z * kTX;
    const int W = nx + 1;  // row length of Ez and Hx

    // 1. Ez(A) with this step's pre-E sources
    for (int e = tid; e < (kTY + 2) * (kTX + 2); e += 256) {
      const int r = e / (kTX + 2), c = e - r * (kTX + 2);
      const int y = j0 - 1 + r, x = i0 - 1 + c;
      float v = 0.0f;
      if (y >= 0 && y <= ny && x >= 0 && x <= nx) {
        v = S.ez[a][y * W + x];
        for (int q = 0; q < S.nsrc; ++q) {
          const int sy = y - S.src_y0[q], sx = x - S.src_x0[q];
          if (sy >= 0 && sy < S.src_sy[q] && sx >= 0 && sx < S.src_sx[q]) {
            int wi = step;
            wi = wi < 0 ? 0 : (wi > S.src_T[q] - 1 ? S.src_T[q] - 1 : wi);
            v += S.src_coeff[q][sy * S.src_sx[q] + sx] * S.src_wave[q][wi];
          }
        }
      }
      s_ez[r][c] = v;
    }
    __syncthreads();

    // 2a. Hx (ny, nx+1): dH Hx - sH * corrected(dEz/dy); rows j0-1 .. j0+kTY-1
    for (int e = tid; e < (kTY + 1) * kTX; e += 256) {
      const int r = e / kTX, c = e - r * kTX;
      const int y = j0 - 1 + r, x = i0 + c;
      float h = 0.0f;
      if (y >= 0 && y < ny && x <= nx) {
        float d = (s_ez[r + 1][c + 1] - s_ez[r][c + 1]) * idx;
        const int q = slab_q2(y, ny, L);
        if (q >= 0) {
          const int o = q * W + x;
          const float psi = S.tab[(0 * 3 + 1) * L2 + q] * S.psi[a][0][o] + S.tab[(0 * 3 + 0) * L2 + q] * d;
          if (r >= 1) S.psi[bnk][0][o] = psi;
          d = d * S.tab[(0 * 3 + 2) * L2 + q] + psi;
        }
        const int o = y * W + x;
        h = h_dec * S.hx[a][o] - h_src * d;
        if ((ml && x == 0) || (mr && x == nx)) h = 0.0f;
        if (r >= 1) S.hx[bnk][o] = h;
      }
      s_hx[r][c] = h;
    }
    // 2b. Hy (ny+1, nx): dH Hy + sH * corrected(dEz/dx); columns i0-1 .. i0+63
    for (int e = tid; e < kTY * (kTX + 1); e += 256) {
      const int r = e / (kTX + 1), c = e - r * (kTX + 1);
      const int y = j0 + r, x = i0 - 1 + c;
      float h = 0.0f;
      if (y <= ny && x >= 0 && x < nx) {
        float d = (s_ez[r + 1][c + 1] - s_ez[r + 1][c]) * idx;
        const int q = slab_q2(x, nx, L);
        if (q >= 0) {
          const int o = y * L2 + q;
          const float psi = S.tab[(1 * 3 + 1) * L2 + q] * S.psi[a][1][o] + S.tab[(1 * 3 + 0) * L2 + q] * d;
          if (c >= 1) S.psi[bnk][1][o] = psi;
          d = d * S.tab[(1 * 3 + 2) * L2 + q] + psi;
        }
        const int o = y * nx + x;
        h = h_dec * S.hy[a][o] + h_src * d;
        if ((mb && y == 0) || (mt && y == ny)) h = 0.0f;
        if (c >= 1) S.hy[bnk][o] = h;
      }
      s_hy[r][c] = h;
    }
    __syncthreads();

    // 3. Ez (ny+1, nx+1): dE Ez + sE * (corrected dHy/dx - corrected dHx/dy)
    for (int e = tid; e < kTY * kTX; e += 256) {
      const int r = e / kTX, c = e - r * kTX;
      const int y = j0 + r, x = i0 + c;
      if (y > ny || x > nx) continue;
      float d0 = gdiff2(s_hy[r][c + 1], s_hy[r][c], x, nx, ml, mr) * idx;
      float d1 = gdiff2(s_hx[r + 1][c], s_hx[r][c], y, ny, mb, mt) * idx;
      const int qx = slab_q2(x, nx + 1, L);
      if (qx >= 0) {
        const int o = y * L2 + qx;
        const float psi = S.tab[(2 * 3 + 1) * L2 + qx] * S.psi[a][2][o] + S.tab[(2 * 3 + 0) * L2 + qx] * d0;
        S.psi[bnk][2][o] = psi;
        d0 = d0 * S.tab[(2 * 3 + 2) * L2 + qx] + psi;
      }
      const int qy = slab_q2(y, ny + 1, L);
      if (qy >= 0) {
        const int o = qy * W + x;
        const float psi = S.tab[(3 * 3 + 1) * L2 + qy] * S.psi[a][3][o] + S.tab[(3 * 3 + 0) * L2 + qy] * d1;
        S.psi[bnk][3][o] = psi;
        d1 = d1 * S.tab[(3 * 3 + 2) * L2 + qy] + psi;
      }
      const int o = y * W + x;
      const float es = S.e_src[S.e_src_st ? o : 0];
      float ez = e_dec * s_ez[r + 1][c + 1] + es * (d0 - d1);
      if ((mb && y == 0) || (mt && y == ny) || (ml && x == 0) || (mr && x == nx)) ez = 0.0f;
      S.ez[bnk][o] = ez;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) zero_counters(int* c, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) c[i] = 0;
}

}  // namespace

void launch_batch2d(hipStream_t s, const Batch2dParams& p) {
  BZH_REQUIRE(p.workgroups >= 1, "batched launch needs at least one workgroup");
  BZH_REQUIRE(p.sims != nullptr && p.tiles != nullptr && p.counter != nullptr && p.step0 != nullptr,
              "batched launch is missing its descriptors");
  const dim3 block(256, 1, 1), grid(p.workgroups, 1, 1);
  switch (p.tile_rows) {
    case 4: hipLaunchKernelGGL((batch2d_step<4>), grid, block, 0, s, p); break;
    case 8: hipLaunchKernelGGL((batch2d_step<8>), grid, block, 0, s, p); break;
    case 16: hipLaunchKernelGGL((batch2d_step<16>), grid, block, 0, s, p); break;
    case 32: hipLaunchKernelGGL((batch2d_step<32>), grid, block, 0, s, p); break;
    default: throw Failure{"batched tile height must be 4, 8, 16 or 32"};
  }
  BZH_HIP(hipGetLastError());
}

void launch_zero_counters(hipStream_t s, const ZeroParams& p) {
  if (p.n <= 0) return;
  hipLaunchKernelGGL(zero_counters, dim3((p.n + 255) / 256), dim3(256), 0, s, p.counters, p.n);
  BZH_HIP(hipGetLastError());
}

}  // namespace bzh
