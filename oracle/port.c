/*
 * FP32 C/OpenMP restatement of the reference 3D Yee + CPML step, used only as
 * the timed CPU baseline (TEST INFRASTRUCTURE: never linked by the beamz_amd
 * package).
 *
 * Follows beamz/simulation/kernels.py:458-626 (H and E phases with packed CPML
 * terms, kernels.py:184-264) and lattice.py:523-575 (ghost rule).  PEC faces
 * are masked as in _boundary_compile.py:118-160.  Fields use the complete Yee
 * supports of lattice.py:32-45, row-major (z, y, x).
 *
 * Every CPML term has the same uniform thickness L on both ends of its axis
 * (the benchmark workloads), so a term's packed coordinate of c along an axis
 * of extent n is c (c < L) or L + c - (n - L) (c >= n - L).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int nz, ny, nx;   /* material grid */
  float inv_dx;
  float h_src;      /* dt / mu0 (H coefficients are uniform) */
  const float* e_src[3];  /* per-cell E source coefficients (decay == 1) */
  int L;            /* uniform CPML thickness (0: none) */
  const float* a[12];     /* packed 1-D term coefficients, length 2L */
  const float* b[12];
  const float* ik[12];
  float* psi[12];   /* packed recurrence memories */
  int metallic;     /* face bits front,back,bottom,top,left,right */
} port_grid;

static inline int packed(int c, int n, int L) {
  if (c < L) return c;
  if (c >= n - L) return L + c - (n - L);
  return -1;
}

static inline float corr(const port_grid* g, int t, float d, int c, int n, long q_base,
                         long q_stride) {
  if (g->L == 0) return d;
  int p = packed(c, n, g->L);
  if (p < 0) return d;
  float* psi = g->psi[t] + q_base + (long)p * q_stride;
  float v = g->b[t][p] * (*psi) + g->a[t][p] * d;
  *psi = v;
  return d * g->ik[t][p] + v;
}

/* One full step: H phase then E phase (no sources, no monitors). */
void port_step(const port_grid* g, float* hx, float* hy, float* hz, float* ex, float* ey,
               float* ez) {
  const int nz = g->nz, ny = g->ny, nx = g->nx, L2 = 2 * g->L;
  const float idx = g->inv_dx, hs = g->h_src;
  const int m = g->metallic;
  /* ---- H phase ---- */
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= nz; ++k) {
    for (int j = 0; j <= ny; ++j) {
      if (k < nz && j < ny) { /* Hx (nz, ny, nx+1) */
        for (int i = 0; i <= nx; ++i) {
          long o = ((long)k * ny + j) * (nx + 1) + i;
          float d0 = (ez[((long)k * (ny + 1) + j + 1) * (nx + 1) + i] - ez[((long)k * (ny + 1) + j) * (nx + 1) + i]) * idx;
          float d1 = (ey[((long)(k + 1) * ny + j) * (nx + 1) + i] - ey[((long)k * ny + j) * (nx + 1) + i]) * idx;
          d0 = corr(g, 0, d0, j, ny, (long)k * L2 * (nx + 1) + i, nx + 1);
          d1 = corr(g, 1, d1, k, nz, (long)j * (nx + 1) + i, (long)ny * (nx + 1));
          float h = hx[o] - hs * (d0 - d1);
          if (((m & 16) && i == 0) || ((m & 32) && i == nx)) h = 0.0f;
          hx[o] = h;
        }
      }
      if (k < nz) { /* Hy (nz, ny+1, nx) */
        for (int i = 0; i < nx; ++i) {
          long o = ((long)k * (ny + 1) + j) * nx + i;
          float d2 = (ex[((long)(k + 1) * (ny + 1) + j) * nx + i] - ex[o]) * idx;
          long ez0 = ((long)k * (ny + 1) + j) * (nx + 1) + i;
          float d3 = (ez[ez0 + 1] - ez[ez0]) * idx;
          d2 = corr(g, 2, d2, k, nz, (long)j * nx + i, (long)(ny + 1) * nx);
          d3 = corr(g, 3, d3, i, nx, ((long)k * (ny + 1) + j) * L2, 1);
          float h = hy[o] - hs * (d2 - d3);
          if (((m & 4) && j == 0) || ((m & 8) && j == ny)) h = 0.0f;
          hy[o] = h;
        }
      }
      if (j < ny) { /* Hz (nz+1, ny, nx) */
        for (int i = 0; i < nx; ++i) {
          long o = ((long)k * ny + j) * nx + i;
          long ey0 = ((long)k * ny + j) * (nx + 1) + i;
          long ex0 = ((long)k * (ny + 1) + j) * nx + i;
          float d4 = (ey[ey0 + 1] - ey[ey0]) * idx;
          float d5 = (ex[ex0 + nx] - ex[ex0]) * idx;
          d4 = corr(g, 4, d4, i, nx, ((long)k * ny + j) * L2, 1);
          d5 = corr(g, 5, d5, j, ny, (long)k * L2 * nx + i, nx);
          float h = hz[o] - hs * (d4 - d5);
          if (((m & 1) && k == 0) || ((m & 2) && k == nz)) h = 0.0f;
          hz[o] = h;
        }
      }
    }
  }
  /* ---- E phase ---- */
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= nz; ++k) {
    for (int j = 0; j <= ny; ++j) {
      { /* Ex (nz+1, ny+1, nx) */
        for (int i = 0; i < nx; ++i) {
          long o = ((long)k * (ny + 1) + j) * nx + i;
          float up = j < ny ? hz[((long)k * ny + j) * nx + i] : ((m & 8) ? 0.0f : hz[((long)k * ny + ny - 1) * nx + i]);
          float lo = j > 0 ? hz[((long)k * ny + j - 1) * nx + i] : ((m & 4) ? 0.0f : hz[((long)k * ny) * nx + i]);
          float d0 = (up - lo) * idx;
          up = k < nz ? hy[((long)k * (ny + 1) + j) * nx + i] : ((m & 2) ? 0.0f : hy[((long)(nz - 1) * (ny + 1) + j) * nx + i]);
          lo = k > 0 ? hy[((long)(k - 1) * (ny + 1) + j) * nx + i] : ((m & 1) ? 0.0f : hy[(long)j * nx + i]);
          float d1 = (up - lo) * idx;
          d0 = corr(g, 6, d0, j, ny + 1, (long)k * L2 * nx + i, nx);
          d1 = corr(g, 7, d1, k, nz + 1, (long)j * nx + i, (long)(ny + 1) * nx);
          float e = ex[o] + g->e_src[0][o] * (d0 - d1);
          if (((m & 1) && k == 0) || ((m & 2) && k == nz) || ((m & 4) && j == 0) || ((m & 8) && j == ny)) e = 0.0f;
          ex[o] = e;
        }
      }
      if (j < ny) { /* Ey (nz+1, ny, nx+1) */
        for (int i = 0; i <= nx; ++i) {
          long o = ((long)k * ny + j) * (nx + 1) + i;
          float up = k < nz ? hx[((long)k * ny + j) * (nx + 1) + i] : ((m & 2) ? 0.0f : hx[((long)(nz - 1) * ny + j) * (nx + 1) + i]);
          float lo = k > 0 ? hx[((long)(k - 1) * ny + j) * (nx + 1) + i] : ((m & 1) ? 0.0f : hx[(long)j * (nx + 1) + i]);
          float d2 = (up - lo) * idx;
          long hz0 = ((long)k * ny + j) * nx;
          up = i < nx ? hz[hz0 + i] : ((m & 32) ? 0.0f : hz[hz0 + nx - 1]);
          lo = i > 0 ? hz[hz0 + i - 1] : ((m & 16) ? 0.0f : hz[hz0]);
          float d3 = (up - lo) * idx;
          d2 = corr(g, 8, d2, k, nz + 1, (long)j * (nx + 1) + i, (long)ny * (nx + 1));
          d3 = corr(g, 9, d3, i, nx + 1, ((long)k * ny + j) * L2, 1);
          float e = ey[o] + g->e_src[1][o] * (d2 - d3);
          if (((m & 1) && k == 0) || ((m & 2) && k == nz) || ((m & 16) && i == 0) || ((m & 32) && i == nx)) e = 0.0f;
          ey[o] = e;
        }
      }
      if (k < nz) { /* Ez (nz, ny+1, nx+1) */
        for (int i = 0; i <= nx; ++i) {
          long o = ((long)k * (ny + 1) + j) * (nx + 1) + i;
          long hy0 = ((long)k * (ny + 1) + j) * nx;
          float up = i < nx ? hy[hy0 + i] : ((m & 32) ? 0.0f : hy[hy0 + nx - 1]);
          float lo = i > 0 ? hy[hy0 + i - 1] : ((m & 16) ? 0.0f : hy[hy0]);
          float d4 = (up - lo) * idx;
          up = j < ny ? hx[((long)k * ny + j) * (nx + 1) + i] : ((m & 8) ? 0.0f : hx[((long)k * ny + ny - 1) * (nx + 1) + i]);
          lo = j > 0 ? hx[((long)k * ny + j - 1) * (nx + 1) + i] : ((m & 4) ? 0.0f : hx[((long)k * ny) * (nx + 1) + i]);
          float d5 = (up - lo) * idx;
          d4 = corr(g, 10, d4, i, nx + 1, ((long)k * (ny + 1) + j) * L2, 1);
          d5 = corr(g, 11, d5, j, ny + 1, (long)k * L2 * (nx + 1) + i, nx + 1);
          float e = ez[o] + g->e_src[2][o] * (d4 - d5);
          if (((m & 4) && j == 0) || ((m & 8) && j == ny) || ((m & 16) && i == 0) || ((m & 32) && i == nx)) e = 0.0f;
          ez[o] = e;
        }
      }
    }
  }
}

int port_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}
