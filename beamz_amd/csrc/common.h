// Internal definitions shared by the gfx950 kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <string>

#include "beamz_hip.h"

namespace bzh {

// ---- error reporting -------------------------------------------------------
void set_error(const std::string& message);
void clear_error();

struct Failure {
  std::string message;
};

#define BZH_REQUIRE(cond, msg)                                  \
  do {                                                          \
    if (!(cond)) throw ::bzh::Failure{std::string(msg)};        \
  } while (0)

#define BZH_HIP(expr)                                                          \
  do {                                                                         \
    hipError_t bzh_status_ = (expr);                                           \
    if (bzh_status_ != hipSuccess)                                             \
      throw ::bzh::Failure{std::string(#expr) + ": " +                         \
                           hipGetErrorString(bzh_status_)};                    \
  } while (0)

// ---- storage types ---------------------------------------------------------
// Fields may be stored as f32, bf16 or f16; arithmetic is always f32.
struct bf16_t {
  uint16_t bits;
};

__device__ __forceinline__ float load_f(const float* p, int64_t i) { return p[i]; }
__device__ __forceinline__ float load_f(const bf16_t* p, int64_t i) {
  return __uint_as_float(static_cast<uint32_t>(p[i].bits) << 16);
}
__device__ __forceinline__ float load_f(const __half* p, int64_t i) {
  return __half2float(p[i]);
}
__device__ __forceinline__ void store_f(float* p, int64_t i, float v) { p[i] = v; }
__device__ __forceinline__ void store_f(bf16_t* p, int64_t i, float v) {
  // round to nearest even; NaN stays NaN
  uint32_t u = __float_as_uint(v);
  uint32_t r = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) r = (u >> 16) | 0x40u;
  p[i].bits = static_cast<uint16_t>(r);
}
__device__ __forceinline__ void store_f(__half* p, int64_t i, float v) {
  p[i] = __float2half(v);
}

// Stochastic rounding of float32 to bfloat16: add 16 random bits below the
// kept mantissa and truncate, so that E[round(v)] = v and the rounding error
// of a slowly varying field does not repeat from step to step.  The bits come
// from a per-lane Weyl sequence (state += golden-ratio increment) scrambled by
// one odd multiply; the state is seeded from the lane's global index and the
// time step, so a run is deterministic.  NaN stays NaN.
struct SrState {
  uint32_t s;
};
__device__ __forceinline__ SrState sr_seed(uint32_t lane_id, uint32_t step) {
  uint32_t h = lane_id * 0x9E3779B1u ^ (step + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return {h};
}
__device__ __forceinline__ uint16_t bf16_bits_sr(float v, SrState& st) {
  st.s += 0x9E3779B9u;
  const uint32_t r = (st.s * 0x85EBCA6Bu) >> 16;
  const uint32_t u = __float_as_uint(v);
  uint32_t b = (u + r) >> 16;
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) b = (u >> 16) | 0x40u;
  return static_cast<uint16_t>(b);
}

// ---- phase parameters --------------------------------------------------------
// One packed CPML derivative term on the device.
struct DevTerm {
  void* psi;          // packed recurrence memory (f32 or bf16), target shape with
                      // the derivative axis cut to low + high (only the sum
                      // low + high, the psi extent, is used by the kernels)
  const float* a;     // packed along the axis, length low + high
  const float* b;
  const float* ik;
  int32_t low;
  int32_t high;
  int32_t n;          // target extent along the axis
  int32_t psi_bf16;   // psi stored as bf16
};

// Fields of one phase: three targets (updated in place, or written to an out
// bank) and the three fields of the other phase.  Shapes are derived from the
// material grid (nz, ny, nx) by the complete-Yee rule and validated on the host.
// Timed source slabs folded into a 2D (TMxy) phase launch, float32 fields:
// each updated cell adds its slabs' terms in declaration order, as the
// source launches after the phase would (execute.py:374-461; the kernel may
// fuse a term's multiply and add, so the sums agree to rounding), skipping
// the PEC-constrained ends.  comp: the target within the phase (0 Hx, 1 Hy
// for H; 2 Ez for E).
struct PhaseSrc {
  const float* coeff;      // [n, sz, sy, sx]
  const float* wave;       // [n, T]
  const int32_t* starts;   // [n, 3] (z, y, x)
  int32_t n, sz, sy, sx, T;
  int32_t comp;
  int32_t fy, fx;          // the target's shape (rows, columns)
  int32_t mlo[3], mhi[3];  // PEC-constrained ends (z, y, x)
};
constexpr int kMaxPhaseSrc = 2;

struct PhaseParams {
  const void* tin[3];   // targets read
  void* tout[3];        // targets written (== tin for in-place launches)
  const void* src[3];   // other-phase fields
  const float* dec[3];  // decay coefficients (nullptr -> dec_s); element o * dec_st
  const float* scf[3];  // source coefficients (nullptr -> scf_s); element o * scf_st
  // E phase with packed coefficient IDs (compile.py:146-187): scf[c] is then
  // the component's table (<= 256 float32 values) and scf_ids[c] the cells'
  // 8-bit IDs, four per int32 word; the decay is 1 (lossless)
  const int32_t* scf_ids[3];
  int32_t dec_st[3];    // 1 for a per-cell array, 0 for a device scalar
  int32_t scf_st[3];
  float dec_s[3];
  float scf_s[3];
  float ratio;          // target storage scale / source storage scale
  int32_t scf_f16;      // some source coefficient array is f16 (fused step only)
  const float* metric[3];  // (z, y, x) inverse distances (nullptr -> inv_dx)
  int32_t met_st[3];       // 1 per-coordinate array, 0 device scalar
  float inv_dx;
  int32_t metallic;     // bit mask of PEC faces
  int32_t nz, ny, nx;   // material grid
  int32_t nterms;       // 0 or 6 (3D) / 0 or 2 (TMxy)
  int32_t uniform_l;    // uniform CPML thickness (0 -> read metadata)
  const int32_t* meta;  // term metadata s32[nterms, 5] (device)
  DevTerm term[6];
  // timed source slabs of this step folded into the 2D kernels (npsrc 0: the
  // source launches apply them); the waveform index is *step0 + k
  PhaseSrc psrc[kMaxPhaseSrc];
  int32_t npsrc;
  const int32_t* step0;
  int32_t k;
};

// Face bits (runtime.py:38-40): front, back, bottom, top, left, right.
enum : int32_t {
  kFront = 1,
  kBack = 2,
  kBottom = 4,
  kTop = 8,
  kLeft = 16,
  kRight = 32
};

// A source slab applied inside a fused launch (``sources`` of
// execute.py:374-461 on one E or H component): in the pair kernel between
// its two steps (an E-timing slab of the first step, ``post``: before the
// intermediate monitor sample, skipping PEC-constrained cells; a pre-E slab
// of the second) or after an H update (H-timing slabs, ``post``); in the
// single-step kernel after its H update.
struct MidSlab {
  const float* coeff;     // this slab's [sz, sy, sx] coefficients (storage units)
  const float* wave;      // this slab's waveform [T]
  const int32_t* start;   // device int32[3]: (z, y, x) of the slab origin
  int32_t sz, sy, sx, T;
  int32_t hfield;         // 0: E component, 1: H component
  int32_t comp;           // 0 x, 1 y, 2 z
  int32_t kstep;          // waveform index = *step0 + kstep, clipped to [0, T-1]
  int32_t post;
  int32_t mlo[3], mhi[3]; // PEC masks of post slabs (z, y, x)
};
// (16 slabs between the steps of a pair: a 3D TF/SF mode launch emits 4
// H-timed and 4 E-timed slabs per step, 12 over a pair)
constexpr int kMaxMidSlabs = 16;
constexpr int kMaxHSlabs = 8;

// One fused leapfrog step reading bank A and writing bank B (fields and CPML
// memories).  Requires uniform H coefficients, isotropic metrics and, with
// CPML, the 36 term arrays packed as the rows of tab[12][3][tab_st] (H terms
// then E terms; a, b, 1/kappa; the terms on one axis of one phase share their
// slab extents) and each psi of bank B at psi + psi_delta (elements).
constexpr int kSlabH = 0, kSlabE = 3;
struct FusedParams {
  const void* e_in[3];
  const void* h_in[3];
  void* e_out[3];
  void* h_out[3];
  const float* hdec[3];
  const float* hsrc[3];
  const float* edec[3];
  const float* esrc[3];
  int32_t hdec_st[3], hsrc_st[3], edec_st[3], esrc_st[3];
  float inv_dx;
  float ratio_h;   // H storage scale / E storage scale
  float ratio_e;   // E storage scale / H storage scale
  int32_t nz, ny, nx;
  int32_t metallic;
  int32_t L;             // thickest CPML slab (0: no CPML)
  // CPML slab extents, indexed kSlabH/kSlabE + axis (0 x, 1 y, 2 z): the
  // samples at the low face and at the high face that carry a CPML memory
  // (the memory's extent along the axis is lo + hi; 0 and 0 for an axis
  // without CPML).  The H terms on one axis share their extents, as do the E
  // terms.
  int32_t clo[6], chi[6];
  int32_t tab_st;        // floats per (term, coefficient) row of tab
  const float* tab;      // [12 terms][a, b, 1/kappa][tab_st]
  void* psi[12];         // bank A: H terms 0..5, E terms 6..11
  int64_t psi_delta;
  int32_t psi_bf16;
  // work assignment: a box of (x tile, y tile, z chunk) workgroups.  Logical
  // x tile b in [0, box_ntx) maps to box_tx0 + b, plus box_tskip when
  // b >= box_tsplit (the two x-edge ranges of the grid in one launch).
  int32_t box_tx0, box_ntx, box_tsplit, box_tskip;
  int32_t box_ty0, box_nty;
  int32_t box_z0, box_z1, box_kz, box_nzc;     // z range [z0, z1) in chunks of kz
  int32_t box_band;      // tiles per XCD band (0: contiguous ranges per XCD)
  int32_t xedge;         // 1: tiles touching an x face or x CPML slab
  int32_t rows;          // tile height (y rows) the box was cut with
  int32_t coef_f16;      // E source coefficient arrays stored as float16
  // Based addressing, the only form the kernels use (the pointers above are
  // host-side inputs): one scalar base pair serves a group and each access
  // adds a byte offset, 32-bit when every array of the group lies within
  // 4 GiB above the group's base (64-bit when ``wide``; the CPML memories are
  // always within 4 GiB).  The read bank's fields sit at fin + foff[c],
  // the written bank's at fout + foff[c] (c: Ex, Ey, Ez, Hx, Hy, Hz; both
  // banks share the layout); E source coefficient arrays at cbase + coff[c];
  // CPML memory t is read at pbase + poff[t] and written at
  // pbase + poff[t] + pdelta (modulo 2^32).
  const char* fin;
  const char* fout;
  uint32_t foff[6];
  const char* cbase;
  uint32_t coff[3];
  // 1: a group spans 4 GiB or more (each array under 4 GiB), so the kernels
  // form 64-bit offsets (the wide variants: float32 storage only) from
  // foff64 and coff64, the same offsets at full width (the 32-bit ones are
  // then unused)
  int32_t wide;
  uint64_t foff64[6];
  uint64_t coff64[3];
  const char* pbase;
  uint32_t poff[12];
  uint32_t pdelta;
  // H-timing source slabs of this step (single-step launches only), applied
  // after the H update; the waveform index reads *step0
  MidSlab hslab[kMaxHSlabs];
  int32_t hslab_n;
  const int32_t* step0;
  // step of this launch within the call (the first step of a pair launch):
  // with *step0 it keys the stochastic rounding of bfloat16 field stores
  int32_t kstep;
  int32_t bf16_sr;       // 1: bfloat16 fields are stored with stochastic rounding
};
int fused_rows();

constexpr int kMaxMonBoxes = 8;

// A box of pair work items: x tiles tx0 + [0, ntx), the ones at or past
// tsplit shifted by tskip (two x ranges in one box); y tiles ty0 + [0, nty)
// likewise with tysplit and tyskip; z chunks of kz planes, nzc_lo of them
// over [z0, zsplit) and the rest over [zhi, z1).  Tiles are dealt to the 8
// XCDs in bands of ``band`` tiles (0: contiguous ranges).
struct PairBox {
  int32_t tx0, ntx, tsplit, tskip;
  int32_t ty0, nty, tysplit, tyskip;
  int32_t z0, zsplit, zhi, z1, kz, nzc_lo, nzc;
  int32_t band;
  int32_t nblk;   // workgroups of this box, a multiple of 8
  int32_t xedge;  // 0: x-interior tiles, 1: tiles that touch an x face or an
                  // x CPML slab, 2: core tiles (every iteration lean)
};
constexpr int kMaxPairBoxes = 4;

// Two leapfrog steps in one launch (bank A -> bank B).  ``f`` carries the
// single-step layout and work box (tiles of 60 x 13 owned cells); the first
// step's fields are written to the bank at ``cmid`` (same layout, foff) only
// where a monitor box [z0, z1] x [y0, y1] x [x0, x1] (inclusive Yee indices,
// device int32[mbox_n][6]) holds the cell, for the intermediate DFT sample.
struct PairParams {
  FusedParams f;
  // source slabs, in order: [0, msrc_post) E-timing of the first step,
  // [msrc_post, msrc_e) pre-E of the second, [msrc_e, msrc_h1) H-timing of the
  // first, [msrc_h1, msrc_n) H-timing of the second
  MidSlab msrc[kMaxMidSlabs];
  int32_t msrc_n, msrc_post, msrc_e, msrc_h1;
  const int32_t* step0;
  char* cmid;
  const int32_t* mbox;
  int32_t mbox_n;
  // Capture of the monitored cells (when ``cap`` is set, instead of
  // ``cmid``): the first step's fields of every cell a monitor gathers, f32
  // (``cap_slot`` floats).  ``capbox`` is device int32[cap_m][6][8]: per
  // monitor and DFT component (Ex .. Hz) the inclusive box z0, z1, y0, y1,
  // x0, x1, the offset of its cells and 0 (z0 > z1 when the component is not
  // gathered); a cell's index is off + ((x - x0) * ny_b + (y - y0)) * nz_b +
  // (z - z0), z fastest, so that a wave's stores over the z march fill whole
  // lines.  Each tile row holds cells of at most one monitor (checked on the
  // host).
  float* cap;
  const int32_t* capbox;
  int32_t cap_slot;
  int32_t cap_m;
  // The launch's work: box[0 .. nbox) in order, workgroups numbered through
  // them (each box's block count is a multiple of 8, so every box deals its
  // tiles to the XCDs from XCD 0).  ``core`` launches run the core kernel: one
  // box whose every iteration is a lean one (no face, slab, ghost or x-edge
  // work on any of its rows and planes).
  PairBox box[kMaxPairBoxes];
  int32_t nbox;
  int32_t core;
};
constexpr int kPairCols = 60;   // owned columns of a pair tile
constexpr int kPairRows = 13;   // owned rows of a pair tile (16 waves)

// Bounding boxes of the monitors' gather points (one workgroup per monitor).
struct MonBoxParams {
  const int32_t* indices;  // [M, 6, Pmax, K]
  const float* weights;
  int32_t M, Pmax, K;
  int32_t dims[6][3];      // component shapes (z, y, x)
  int32_t* box;            // out: int32[M][6]
};

// Launchers implemented in the .hip files.
void launch_fused3d(hipStream_t s, const FusedParams& p, int dtype);
void launch_pair3d(hipStream_t s, const PairParams& p, int dtype);
void launch_monbox(hipStream_t s, const MonBoxParams& p);
void launch_phase3d(hipStream_t s, const PhaseParams& p, int phase, int dtype);
void launch_phase_tm(hipStream_t s, const PhaseParams& p, int phase, int dtype);

// Slab-source group on one field component.
struct SourceParams {
  void* field;           // the target field (storage dtype)
  int32_t dtype;
  float scale;           // field storage scale
  int32_t fdims[3];      // field shape (z, y, x); 2D uses fdims[0] = 1
  const float* coeff;    // [n, sz, sy, sx]
  const float* wave;     // [n, T]
  const int32_t* starts; // [n, 3] (z, y, x)
  int32_t n;
  int32_t sdims[3];      // slab shape
  int32_t T;
  const int32_t* step0;  // device scalar: step index of step 0 of this call
  int32_t k;             // step within this call
  int32_t slab0;         // first slab handled by this launch
  int32_t nslabs;        // slabs summed per cell by one thread (coincident group)
  int32_t mlo[3];        // component masked at the low end of axis (z, y, x)
  int32_t mhi[3];        // ... at the high end (only for post-update timings)
};
void launch_source(hipStream_t s, const SourceParams& p);

// Packed DFT monitors.
struct DftParams {
  const void* field[6];  // Ex, Ey, Ez, Hx, Hy, Hz (storage dtype; nullptr absent)
  float inv_scale[6];    // 1 / storage scale per component
  int32_t dtype;
  const int32_t* indices;  // [M, 6, Pmax, K]
  const float* weights;    // [M, 6, Pmax, K]
  const void* freqs;       // [M, Fmax]
  int32_t freqs_f64;
  const float* masks;      // [M, 6]
  const int32_t* counts;   // [M, 5]
  const int32_t* codes;    // [M, 2]
  const float* windows;    // [M, 3]
  void* re;
  void* im;
  void* wsum;
  int32_t acc_f64;
  const void* time;        // device scalar: t at the start of the call
  int32_t time_f64;
  const int32_t* step0;
  int32_t k;
  double dt;
  int32_t M, Pmax, K, Fmax;
  int32_t fsize[6];        // element count of each component (gather bound)
  int64_t re_len;          // arena lengths (accumulation bound)
  int64_t w_len;
  // Buffered accumulation: mode 1 stages this step's gathered samples in
  // stage[k % slots][M][6][Pmax] (f32), mode 2 folds steps k0 .. k0+nk-1 of
  // the ring into the accumulators; mode 0 accumulates step k directly.
  int32_t mode;
  float* stage;
  int32_t slots;
  int32_t k0, nk;
  // Gather from a capture of the monitored cells instead of the fields (the
  // pair kernel's, laid out as ``PairParams::capbox`` describes) when set.
  const float* cap;
  const int32_t* capbox;
  int32_t fdims[6][3];   // component shapes (z, y, x), for the capture index
};
constexpr int kDftMaxSlots = 32;
void launch_dft(hipStream_t s, const DftParams& p);

// One TMxy simulation of a batch on the device: two banks of (Ez, Hx, Hy),
// uniform decays and H source, an E source array or scalar, CPML memories of
// both banks (Hx_y, Hy_x, Ez_x, Ez_y), the coefficient table [4][3][2L] and
// up to four pre-E Ez slab sources.
struct Sim2dDev {
  float* ez[2];
  float* hx[2];
  float* hy[2];
  float* psi[2][4];
  const float* e_src;   // (ny+1, nx+1), or a device scalar when e_src_st == 0
  const float* e_dec;   // device scalars
  const float* h_src;
  const float* h_dec;
  const float* tab;
  const float* src_coeff[4];
  const float* src_wave[4];
  float inv_dx;
  int32_t e_src_st;
  int32_t src_y0[4], src_x0[4], src_sy[4], src_sx[4], src_T[4];
  int32_t nsrc;
  int32_t ny, nx, L, metallic;
};

struct Batch2dParams {
  const Sim2dDev* sims;
  const int4* tiles;     // (simulation, item row, item column, 0), queue order
  int32_t ntiles;
  int32_t bank;          // bank read by this step
  const int32_t* step0;  // device scalar: step index of step 0 of this call
  int32_t k;             // step within this call
  int32_t tile_rows;     // rows per item (items are 62 columns wide)
  int32_t workgroups;    // persistent grid size
  int32_t nsims;         // simulations (the source launch's grid)
  int32_t max_src_cells; // largest pre-E slab, in cells (0: no sources)
};
void launch_batch2d(hipStream_t s, const Batch2dParams& p);
void launch_batch2d_sources(hipStream_t s, const Batch2dParams& p);

// Reverse-time TMxy step (adjoint.hip).
struct Adjoint2dParams {
  int32_t ny, nx, L, metallic;
  float inv_dx;
  const float* e_dec;
  const float* e_src;
  int32_t e_src_st;
  const float* h_dec;
  const float* h_src;
  const float* tab;          // forward CPML table [4][3][2L] (Hx_y, Hy_x, Ez_x, Ez_y)
  float* a_ez;               // adjoint state (Ez, Hx, Hy shapes)
  float* a_hx;
  float* a_hy;
  float* a_psi[4];           // adjoint CPML memories (shapes of the forward ones)
  float* w_e0;               // scratch: Ez, Ez, Hy, Hx shapes
  float* w_e1;
  float* w_hy;
  float* w_hx;
  const float* src_profile;  // Ez-shaped adjoint source weights (nullptr: none)
  const float* src_time;     // time factor per forward step
  int32_t n;                 // forward step being reversed
  int32_t design[4];         // Ez-node box [y0, y1) x [x0, x1)
  float* record;             // forward curl per segment step [K][dy][dx]
  int32_t slot;              // record slot of step n
  double* grad;              // sum over steps of g * curl, [dy][dx]
};
void launch_adjoint_step(hipStream_t s, const Adjoint2dParams& p);
void launch_adjoint_record(hipStream_t s, const Adjoint2dParams& p, const float* hx, const float* hy);


}  // namespace bzh
