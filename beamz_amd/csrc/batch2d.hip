// Batched TMxy steps: many independent 2D simulations of uneven sizes advance
// together, one persistent launch per time step.
//
// Work: every simulation is cut into items of 62 (x) by R (y) Ez cells; the
// items of all simulations form one list (``tiles``, host-built, largest
// simulations first, row-major within one), so uneven grid sizes share one
// launch.  Items are equal but for the simulations' edges, so wave w takes
// items w, w + W, w + 2W, ... of the W waves of the grid: no queue atomics
// (a single shared counter serialized the pops), and the four waves of a
// workgroup take four neighbouring items of one row band, whose shared edge
// lines then sit in one L2.
//
// One item is one wave marching up its R rows, one full leapfrog step from
// bank A to bank B (kernels.py:731-811, 1217-1290; ghost rule
// lattice.py:523-575; packed CPML kernels.py:184-264; PEC faces
// _boundary_compile.py:118-160).  Lane l holds column x = 62 c - 1 + l
// (lanes 1..62 own their column, lanes 0 and 63 are halos); x neighbours come
// from ds_bpermute, and the y neighbours are the rows the wave carries in
// registers, so no LDS and no barrier is involved.  At row j the wave loads
// Ez(j+1), Hx(j), Hy(j) and the E coefficient (two rows ahead of their use,
// so that the loads of rows j+1 and j+2 fly while row j is computed), then
// forms
//   Hx_new(j) from Ez(j+1) - Ez(j),
//   Hy_new(j) from Ez(j, x+1) - Ez(j, x),
//   Ez_new(j) from Hy_new(x) - Hy_new(x-1) and Hx_new(j) - Hx_new(j-1),
// carrying Ez(j+1) and Hx_new(j) to the next row.  The row below the item is
// recomputed (Hx_new only).  CPML memories are read from bank A and written
// to bank B for owned cells only.  Between steps the kernel boundary is the
// only synchronization.  The pre-E slab sources of a step (execute.py:374-461)
// are added to bank A's Ez by a small launch before the march.

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

__device__ __forceinline__ float shift2(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// Eight waves per SIMD: the march waits on HBM most of its cycles, and the
// SGPR budget of eight waves (some scalars spill to VGPR lanes) measured
// 2.2 % faster than the seven its natural 106 SGPRs allow.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
batch2d_march(Batch2dParams p) {
  const int lane = threadIdx.x & 63;
  const int step = *p.step0 + p.k;
  const int a = p.bank, bnk = p.bank ^ 1;
  const int R = p.tile_rows;
  const int dn = ((lane + 1) & 63) << 2;  // permute address of lane + 1
  const int up = ((lane - 1) & 63) << 2;  // ... of lane - 1

  const int nwaves = gridDim.x * (blockDim.x >> 6);
  for (int t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
       t < p.ntiles; t += nwaves) {
    const int4 td = p.tiles[t];
    // the descriptor's fields are read once into registers: the field stores
    // below could otherwise alias it and force reloads on every row (the
    // source slabs are unrolled over their fixed maximum, so that no local
    // array is indexed at run time and nothing goes to scratch)
    const Sim2dDev& S = p.sims[td.x];
    const float* __restrict__ ez_a = S.ez[a];
    const float* __restrict__ hx_a = S.hx[a];
    const float* __restrict__ hy_a = S.hy[a];
    float* __restrict__ ez_b = S.ez[bnk];
    float* __restrict__ hx_b = S.hx[bnk];
    float* __restrict__ hy_b = S.hy[bnk];
    const float* __restrict__ pa0 = S.psi[a][0];
    const float* __restrict__ pa1 = S.psi[a][1];
    const float* __restrict__ pa2 = S.psi[a][2];
    const float* __restrict__ pa3 = S.psi[a][3];
    float* __restrict__ pb0 = S.psi[bnk][0];
    float* __restrict__ pb1 = S.psi[bnk][1];
    float* __restrict__ pb2 = S.psi[bnk][2];
    float* __restrict__ pb3 = S.psi[bnk][3];
    const float* __restrict__ tab = S.tab;
    const float* __restrict__ e_src = S.e_src;
    const int ny = S.ny, nx = S.nx, L = S.L, L2 = 2 * S.L, m = S.metallic;
    const bool mb = m & kBottom, mt = m & kTop, ml = m & kLeft, mr = m & kRight;
    const float idx = S.inv_dx;
    const float h_dec = *S.h_dec, h_src = *S.h_src, e_dec = *S.e_dec;
    const int W = nx + 1;  // row length of Ez and Hx
    const int y0 = td.y * R, y1 = min(y0 + R, ny + 1);
    const int x = 62 * td.z - 1 + lane;
    const bool vez = x >= 0 && x <= nx;   // Ez / Hx column
    const bool vhy = x >= 0 && x < nx;    // Hy column
    const bool own = lane >= 1 && lane <= 62 && vez;
    const bool xpec = (ml && x == 0) || (mr && x == nx);
    // x-slab terms of this lane: Hy_x (term 1, nx samples), Ez_x (term 2, nx+1)
    const int q1 = vhy ? slab_q2(x, nx, L) : -1;
    const int q2 = vez ? slab_q2(x, nx + 1, L) : -1;
    float a1 = 0, b1 = 0, k1 = 0, a2 = 0, b2 = 0, k2 = 0;
    if (q1 >= 0) {
      a1 = tab[(1 * 3 + 0) * L2 + q1];
      b1 = tab[(1 * 3 + 1) * L2 + q1];
      k1 = tab[(1 * 3 + 2) * L2 + q1];
    }
    if (q2 >= 0) {
      a2 = tab[(2 * 3 + 0) * L2 + q2];
      b2 = tab[(2 * 3 + 1) * L2 + q2];
      k2 = tab[(2 * 3 + 2) * L2 + q2];
    }

    // Ez(A) at (j, x), 0 off the grid
    auto ez_row = [&](int j) -> float { return (j <= ny && vez) ? ez_a[j * W + x] : 0.0f; };

    const bool es_arr = S.e_src_st != 0;
    const float es_uni = es_arr ? 0.0f : e_src[0];
    auto hx_row = [&](int j) { return (j < ny && vez) ? hx_a[j * W + x] : 0.0f; };
    auto hy_row = [&](int j) { return (j <= ny && vhy) ? hy_a[j * nx + x] : 0.0f; };
    auto es_row = [&](int j) { return (es_arr && j <= ny && vez) ? e_src[j * W + x] : es_uni; };

    const int js = y0 > 0 ? y0 - 1 : 0;
    // Each row's operands are loaded two rows ahead: the loads of rows j+1
    // and j+2 fly while row j is computed.
    float ez = ez_row(js), ezn = ez_row(js + 1), ezn2 = ez_row(js + 2);
    float hxa = hx_row(js), hya = hy_row(js), esv = es_row(js);
    float hxa2 = hx_row(js + 1), hya2 = hy_row(js + 1), esv2 = es_row(js + 1);
    float hx_prev = 0.0f;  // Hx_new(j - 1)
    for (int j = js; j < y1; ++j) {
      const bool st = j >= y0;  // this item owns row j
      float ezn3 = 0.0f, hxa3 = 0.0f, hya3 = 0.0f, esv3 = es_uni;
      if (j + 2 < y1) {
        ezn3 = ez_row(j + 3);
        hxa3 = hx_row(j + 2);
        hya3 = hy_row(j + 2);
        esv3 = es_row(j + 2);
      }
      // Hx_new(j) (ny rows, nx+1 columns): dH Hx - sH * corrected(dEz/dy)
      float hx = 0.0f;
      if (j < ny && vez) {
        float d = (ezn - ez) * idx;
        const int q = slab_q2(j, ny, L);
        if (q >= 0) {
          const int o = q * W + x;
          const float psi = tab[(0 * 3 + 1) * L2 + q] * pa0[o] + tab[(0 * 3 + 0) * L2 + q] * d;
          if (st && own) pb0[o] = psi;
          d = d * tab[(0 * 3 + 2) * L2 + q] + psi;
        }
        hx = h_dec * hxa - h_src * d;
        if (xpec) hx = 0.0f;
        if (st && own) hx_b[j * W + x] = hx;
      }
      if (st) {
        // Hy_new(j) (ny+1 rows, nx columns): dH Hy + sH * corrected(dEz/dx)
        const float ez_r = shift2(ez, dn);
        float hy = 0.0f;
        if (vhy) {
          float d = (ez_r - ez) * idx;
          if (q1 >= 0) {
            const int o = j * L2 + q1;
            const float psi = b1 * pa1[o] + a1 * d;
            if (own) pb1[o] = psi;
            d = d * k1 + psi;
          }
          hy = h_dec * hya + h_src * d;
          if ((mb && j == 0) || (mt && j == ny)) hy = 0.0f;
          if (own) hy_b[j * nx + x] = hy;
        }
        // Ez_new(j): dE Ez + sE * (corrected dHy/dx - corrected dHx/dy)
        const float hy_l = shift2(hy, up);
        if (own) {
          float d0 = gdiff2(hy, hy_l, x, nx, ml, mr) * idx;
          float d1 = gdiff2(hx, hx_prev, j, ny, mb, mt) * idx;
          if (q2 >= 0) {
            const int o = j * L2 + q2;
            const float psi = b2 * pa2[o] + a2 * d0;
            pb2[o] = psi;
            d0 = d0 * k2 + psi;
          }
          const int qy = slab_q2(j, ny + 1, L);
          if (qy >= 0) {
            const int o = qy * W + x;
            const float psi = tab[(3 * 3 + 1) * L2 + qy] * pa3[o] + tab[(3 * 3 + 0) * L2 + qy] * d1;
            pb3[o] = psi;
            d1 = d1 * tab[(3 * 3 + 2) * L2 + qy] + psi;
          }
          float e = e_dec * ez + esv * (d0 - d1);
          if ((mb && j == 0) || (mt && j == ny) || xpec) e = 0.0f;
          ez_b[j * W + x] = e;
        }
      }
      hx_prev = hx;
      ez = ezn;
      ezn = ezn2;
      ezn2 = ezn3;
      hxa = hxa2;
      hya = hya2;
      esv = esv2;
      hxa2 = hxa3;
      hya2 = hya3;
      esv2 = esv3;
    }
  }
}

// The pre-E slab sources of one step: block (b, q, i) adds slab q of
// simulation i to bank ``bank``'s Ez, 256 cells per block row.
__global__ void __launch_bounds__(256) batch2d_sources(Batch2dParams p) {
  const Sim2dDev& S = p.sims[blockIdx.z];
  const int q = blockIdx.y;
  if (q >= S.nsrc) return;
  const int n = S.src_sy[q] * S.src_sx[q];
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int sy = c / S.src_sx[q], sx = c - sy * S.src_sx[q];
  const int y = S.src_y0[q] + sy, x = S.src_x0[q] + sx;
  if (y < 0 || x < 0 || y > S.ny || x > S.nx) return;
  int wi = *p.step0 + p.k;
  wi = wi < 0 ? 0 : (wi > S.src_T[q] - 1 ? S.src_T[q] - 1 : wi);
  float* ez = S.ez[p.bank];
  // (atomic: two slabs of one simulation may share a cell)
  atomicAdd(&ez[y * (S.nx + 1) + x], S.src_coeff[q][c] * S.src_wave[q][wi]);
}

}  // namespace

void launch_batch2d(hipStream_t s, const Batch2dParams& p) {
  BZH_REQUIRE(p.workgroups >= 1, "batched launch needs at least one workgroup");
  BZH_REQUIRE(p.sims != nullptr && p.tiles != nullptr && p.step0 != nullptr,
              "batched launch is missing its descriptors");
  BZH_REQUIRE(p.tile_rows == 8 || p.tile_rows == 16 || p.tile_rows == 32 || p.tile_rows == 64,
              "batched item height must be 8, 16, 32 or 64 rows");
  const dim3 block(256, 1, 1), grid(p.workgroups, 1, 1);
  hipLaunchKernelGGL(batch2d_march, grid, block, 0, s, p);
  BZH_HIP(hipGetLastError());
}

void launch_batch2d_sources(hipStream_t s, const Batch2dParams& p) {
  if (p.nsims <= 0 || p.max_src_cells <= 0) return;
  const dim3 block(256, 1, 1), grid((p.max_src_cells + 255) / 256, 4, p.nsims);
  hipLaunchKernelGGL(batch2d_sources, grid, block, 0, s, p);
  BZH_HIP(hipGetLastError());
}


}  // namespace bzh
