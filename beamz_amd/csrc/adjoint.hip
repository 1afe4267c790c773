// Reverse-time (adjoint) TMxy steps for gradients with respect to the E
// source coefficient, i.e. the permittivity, over a design region.
//
// The forward step n (kernels.py:731-811, 1217-1290; lattice.py:523-575;
// packed CPML kernels.py:184-264; PEC _boundary_compile.py:118-160) is, with
// Ez^a = Ez_n + pre-E sources,
//   Hx' = dH Hx - sH dy',  dy = (Ez^a[j+1,i] - Ez^a[j,i]) / dx   (Hx_y CPML)
//   Hy' = dH Hy + sH dx',  dx = (Ez^a[j,i+1] - Ez^a[j,i]) / dx   (Hy_x CPML)
//   Ez' = dE Ez^a + sE (d0' - d1'),  d0 = ghost x-diff of Hy', d1 = ghost
//         y-diff of Hx'                                           (Ez_x, Ez_y)
// with d' = d/kappa + psi_new, psi_new = b psi + a d, and PEC outputs forced
// to zero.  Every operation is linear, so the exact discrete adjoint runs the
// transposed operations in reverse order.  For the recurrence the adjoint
// memory m of psi obeys  Lambda = m + a_d',  m <- b Lambda,
// a_d = a_d'/kappa + a Lambda.
//
// One reverse step is three launches over the adjoint state (aEz, aHx, aHy and
// the four adjoint memories):
//   adj_e:  g = mask(aEz + c_n W);  aEz <- dE g;  t = sE g;  E-term adjoint
//           recurrences; A0 = a_d0/dx, A1 = a_d1/dx;  grad += g * curl_n
//   adj_h:  aHy += ghost-x-diff^T(A0), aHx += ghost-y-diff^T(A1); then the H
//           transposes: aH <- dH g_h, B = a_d/dx with the H-term recurrences
//   adj_ez: aEz += forward-diff^T(Bx, By)
// where W is the monitor's adjoint source profile and c_n its time factor for
// step n, and curl_n the forward (d0' - d1') recorded over the design region.

#include "common.h"

namespace bzh {
namespace {

__device__ __forceinline__ int slab_qa(int c, int n, int L) {
  return L <= 0 ? -1 : (c < L ? c : (c >= n - L ? c - (n - 2 * L) : -1));
}

// Adjoint of one packed CPML correction at packed coordinate q of term t:
// returns a_d from a_d' and updates the adjoint memory in place.
__device__ __forceinline__ float cpml_adj(const float* tab, int t, int L2, int q, float ad,
                                          float* m) {
  const float a = tab[(t * 3 + 0) * L2 + q];
  const float b = tab[(t * 3 + 1) * L2 + q];
  const float ik = tab[(t * 3 + 2) * L2 + q];
  const float lam = *m + ad;
  *m = b * lam;
  return ad * ik + a * lam;
}

__global__ void __launch_bounds__(256) adj_e(Adjoint2dParams p) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  const int ny = p.ny, nx = p.nx, W = nx + 1, L = p.L, L2 = 2 * p.L, m = p.metallic;
  if (j > ny || i > nx) return;
  const int o = j * W + i;
  float g = p.a_ez[o];
  if (p.src_profile) g += p.src_time[p.n] * p.src_profile[o];
  if (((m & kBottom) && j == 0) || ((m & kTop) && j == ny) || ((m & kLeft) && i == 0) ||
      ((m & kRight) && i == nx))
    g = 0.0f;
  const float es = p.e_src[p.e_src_st ? o : 0];
  p.a_ez[o] = *p.e_dec * g;
  const float t = es * g;
  float a0 = t, a1 = -t;  // adjoints of d0' and d1'
  const int qx = slab_qa(i, nx + 1, L);
  if (qx >= 0) a0 = cpml_adj(p.tab, 2, L2, qx, a0, &p.a_psi[2][j * L2 + qx]);
  const int qy = slab_qa(j, ny + 1, L);
  if (qy >= 0) a1 = cpml_adj(p.tab, 3, L2, qy, a1, &p.a_psi[3][qy * W + i]);
  p.w_e0[o] = a0 * p.inv_dx;
  p.w_e1[o] = a1 * p.inv_dx;
  // gradient with respect to sE over the design region: dEz'/dsE = curl_n
  const int dy = j - p.design[0], dx = i - p.design[2];
  const int dh = p.design[1] - p.design[0], dw = p.design[3] - p.design[2];
  if (p.grad && dy >= 0 && dy < dh && dx >= 0 && dx < dw) {
    const float c = p.record[(static_cast<int64_t>(p.slot) * dh + dy) * dw + dx];
    p.grad[dy * dw + dx] += static_cast<double>(g) * static_cast<double>(c);
  }
}

__global__ void __launch_bounds__(256) adj_h(Adjoint2dParams p) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  const int ny = p.ny, nx = p.nx, W = nx + 1, L = p.L, L2 = 2 * p.L, m = p.metallic;
  const bool mb = m & kBottom, mt = m & kTop, ml = m & kLeft, mr = m & kRight;
  const float h_dec = *p.h_dec, h_src = *p.h_src;
  // Hy (ny+1, nx): d0(i) = up(i) - lo(i) over Ez nodes i = 0..nx (ghost rule)
  if (j <= ny && i < nx) {
    const float* A0 = p.w_e0 + j * W;
    float a = p.a_hy[j * nx + i] + A0[i] - A0[i + 1];
    if (i == nx - 1 && !mr) a += A0[nx];
    if (i == 0 && !ml) a -= A0[0];
    if ((mb && j == 0) || (mt && j == ny)) a = 0.0f;  // Hy' forced to zero
    p.a_hy[j * nx + i] = h_dec * a;
    float ad = h_src * a;  // adjoint of dx'
    const int q = slab_qa(i, nx, L);
    if (q >= 0) ad = cpml_adj(p.tab, 1, L2, q, ad, &p.a_psi[1][j * L2 + q]);
    p.w_hy[j * nx + i] = ad * p.inv_dx;
  }
  // Hx (ny, nx+1): d1(j) = up(j) - lo(j) over Ez nodes j = 0..ny
  if (j < ny && i <= nx) {
    const float* A1 = p.w_e1;
    float a = p.a_hx[j * W + i] + A1[j * W + i] - A1[(j + 1) * W + i];
    if (j == ny - 1 && !mt) a += A1[ny * W + i];
    if (j == 0 && !mb) a -= A1[i];
    if ((ml && i == 0) || (mr && i == nx)) a = 0.0f;  // Hx' forced to zero
    p.a_hx[j * W + i] = h_dec * a;
    float ad = -h_src * a;  // adjoint of dy'
    const int q = slab_qa(j, ny, L);
    if (q >= 0) ad = cpml_adj(p.tab, 0, L2, q, ad, &p.a_psi[0][q * W + i]);
    p.w_hx[j * W + i] = ad * p.inv_dx;
  }
}

__global__ void __launch_bounds__(256) adj_ez(Adjoint2dParams p) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  const int ny = p.ny, nx = p.nx, W = nx + 1;
  if (j > ny || i > nx) return;
  float a = 0.0f;
  if (i >= 1) a += p.w_hy[j * nx + i - 1];
  if (i < nx) a -= p.w_hy[j * nx + i];
  if (j >= 1) a += p.w_hx[(j - 1) * W + i];
  if (j < ny) a -= p.w_hx[j * W + i];
  p.a_ez[j * W + i] += a;
}

// Forward curl (d0' - d1') of the current H over the design region into
// record[slot]; the design region holds no CPML slab and no PEC face.
__global__ void __launch_bounds__(256) adj_record(Adjoint2dParams p, const float* hx, const float* hy) {
  const int dx = blockIdx.x * 64 + threadIdx.x, dy = blockIdx.y * 4 + threadIdx.y;
  const int dh = p.design[1] - p.design[0], dw = p.design[3] - p.design[2];
  if (dy >= dh || dx >= dw) return;
  const int j = p.design[0] + dy, i = p.design[2] + dx, nx = p.nx, W = nx + 1;
  const float d0 = (hy[j * nx + i] - hy[j * nx + i - 1]) * p.inv_dx;
  const float d1 = (hx[j * W + i] - hx[(j - 1) * W + i]) * p.inv_dx;
  p.record[(static_cast<int64_t>(p.slot) * dh + dy) * dw + dx] = d0 - d1;
}

}  // namespace

void launch_adjoint_step(hipStream_t s, const Adjoint2dParams& p) {
  const dim3 block(64, 4, 1), grid((p.nx + 1 + 63) / 64, (p.ny + 1 + 3) / 4, 1);
  hipLaunchKernelGGL(adj_e, grid, block, 0, s, p);
  hipLaunchKernelGGL(adj_h, grid, block, 0, s, p);
  hipLaunchKernelGGL(adj_ez, grid, block, 0, s, p);
  BZH_HIP(hipGetLastError());
}

void launch_adjoint_record(hipStream_t s, const Adjoint2dParams& p, const float* hx, const float* hy) {
  const int dh = p.design[1] - p.design[0], dw = p.design[3] - p.design[2];
  const dim3 block(64, 4, 1), grid((dw + 63) / 64, (dh + 3) / 4, 1);
  hipLaunchKernelGGL(adj_record, grid, block, 0, s, p, hx, hy);
  BZH_HIP(hipGetLastError());
}

}  // namespace bzh
