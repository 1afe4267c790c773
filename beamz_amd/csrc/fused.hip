// Host entry points of the fused single-step and pair launches: argument
// checks, then the launch of the storage type's kernels.  The kernels
// (fused_kernels.h) are instantiated per storage type in fused_f32.hip,
// fused_f16.hip and fused_bf16.hip, so that the three compile in parallel.

#include "fused_kernels.h"

namespace bzh {

void launch_fused_f32(hipStream_t s, const FusedParams& p, int nblocks);
void launch_fused_f16(hipStream_t s, const FusedParams& p, int nblocks);
void launch_fused_bf16(hipStream_t s, const FusedParams& p, int nblocks);
void launch_pair_f32(hipStream_t s, const PairParams& p, int nblocks);
void launch_pair_f16(hipStream_t s, const PairParams& p, int nblocks);
void launch_pair_bf16(hipStream_t s, const PairParams& p, int nblocks);

int fused_rows() { return rows_setting(); }

void launch_fused3d(hipStream_t s, const FusedParams& p, int dtype) {
  BZH_REQUIRE(p.box_ntx >= 0 && p.box_nty >= 0 && p.box_kz >= 1 && p.box_nzc >= 0,
              "fused box is malformed");
  BZH_REQUIRE(p.rows == 8 || p.rows == 15, "fused tile height is not one the kernel is built for");
  BZH_REQUIRE(!p.coef_f16 || dtype != BZH_F32, "float16 coefficients need half-precision fields");
  const int ntx_total = (p.nx + 1 + 61) / 62, nty_total = (p.ny + 1 + p.rows - 1) / p.rows;
  BZH_REQUIRE(p.box_tx0 >= 0 && p.box_tsplit >= 0 && p.box_tskip >= 0 &&
                  p.box_tx0 + p.box_ntx + p.box_tskip <= ntx_total && p.box_ty0 >= 0 &&
                  p.box_ty0 + p.box_nty <= nty_total && p.box_z0 >= 0 &&
                  p.box_z1 <= p.nz + 1 && p.box_z0 + (p.box_nzc - 1) * p.box_kz < p.box_z1,
              "fused box exceeds the grid");
  if (!p.xedge) {
    // the interior variant reads lanes x = i0-1 .. i0+62 unchecked and assumes
    // no x face or x CPML slab under lanes 0..62
    const int M = p.L > 0 ? p.L : 1;
    const int lo = p.box_tx0 * 62 - 1;
    const int hi = (p.box_tx0 + p.box_ntx + p.box_tskip - 1) * 62 + 62;
    BZH_REQUIRE(p.box_ntx == 0 || (lo >= M && hi <= p.nx - M - 1),
                "interior fused tiles overlap an x face or x CPML slab");
  }
  BZH_REQUIRE(p.L == 0 || p.tab != nullptr, "CPML tiles need the coefficient table");
  BZH_REQUIRE(p.fin != nullptr && p.fout != nullptr && (p.L == 0 || p.pbase != nullptr) &&
                  (p.esrc_st[0] == 0 || p.cbase != nullptr),
              "the fused step needs based addressing of its array groups");
  BZH_REQUIRE(p.hslab_n >= 0 && p.hslab_n <= kMaxHSlabs && (p.hslab_n == 0 || p.step0 != nullptr),
              "fused H source slabs are malformed");
  for (int i = 0; i < p.hslab_n; ++i) {
    const MidSlab& m = p.hslab[i];
    BZH_REQUIRE(m.coeff && m.wave && m.start && m.sz >= 0 && m.sy >= 0 && m.sx >= 0 && m.T >= 1 &&
                    m.comp >= 0 && m.comp < 3 && m.hfield == 1,
                "fused H source slab is malformed");
  }
  for (int c = 0; c < 3; ++c) {
    BZH_REQUIRE(p.hdec_st[c] == 0 && p.hsrc_st[c] == 0 && p.edec_st[c] == 0,
                "the fused step needs uniform H coefficients and E decay");
    BZH_REQUIRE(p.esrc_st[c] == p.esrc_st[0], "E source coefficients must be all arrays or all scalars");
  }
  const int64_t tiles = static_cast<int64_t>(p.box_ntx) * p.box_nty;
  BZH_REQUIRE(p.box_band == 0 || (p.box_band >= 1 && 8 * int64_t{p.box_band} >= tiles),
              "fused bands do not cover the tiles");
  const int64_t nblocks = (p.box_band > 0 ? 8 * int64_t{p.box_band} : tiles) * p.box_nzc;
  BZH_REQUIRE(nblocks < (int64_t{1} << 31), "fused grid too large");
  if (nblocks <= 0) return;
  switch (dtype) {
    case BZH_F32: launch_fused_f32(s, p, static_cast<int>(nblocks)); break;
    case BZH_BF16: launch_fused_bf16(s, p, static_cast<int>(nblocks)); break;
    case BZH_F16: launch_fused_f16(s, p, static_cast<int>(nblocks)); break;
    default: throw Failure{"unsupported field dtype"};
  }
  BZH_HIP(hipGetLastError());
}

void launch_pair3d(hipStream_t s, const PairParams& pp, int dtype) {
  const FusedParams& p = pp.f;
  BZH_REQUIRE(p.rows == kPairRows, "pair tiles are cut with the pair tile height");
  BZH_REQUIRE(!p.coef_f16 || dtype != BZH_F32, "float16 coefficients need half-precision fields");
  BZH_REQUIRE(pp.nbox >= 1 && pp.nbox <= kMaxPairBoxes && (!pp.core || pp.nbox == 1),
              "pair launch boxes are malformed");
  const int ntx_total = (p.nx + 1 + kPairCols - 1) / kPairCols;
  const int nty_total = (p.ny + 1 + kPairRows - 1) / kPairRows;
  const int M = p.L > 0 ? p.L : 1;
  int64_t nblocks = 0;
  for (int i = 0; i < pp.nbox; ++i) {
    const PairBox& b = pp.box[i];
    // tiles and chunks inside the grid, block count matching the box
    BZH_REQUIRE(b.ntx >= 0 && b.nty >= 0 && b.tx0 >= 0 && b.tsplit >= 0 && b.tskip >= 0 &&
                    b.tx0 + b.ntx + b.tskip <= ntx_total && b.ty0 >= 0 && b.tysplit >= 0 &&
                    b.tyskip >= 0 && b.ty0 + b.nty + b.tyskip <= nty_total,
                "pair box tiles exceed the grid");
    BZH_REQUIRE(b.kz >= 1 && b.nzc_lo >= 0 && b.nzc >= b.nzc_lo && 0 <= b.z0 && b.z0 <= b.zsplit &&
                    b.zsplit <= b.zhi && b.zhi <= b.z1 && b.z1 <= p.nz + 1 &&
                    (b.nzc_lo == 0 || b.z0 + (b.nzc_lo - 1) * b.kz < b.zsplit) &&
                    (b.nzc == b.nzc_lo || b.zhi + (b.nzc - b.nzc_lo - 1) * b.kz < b.z1),
                "pair box chunks exceed the grid");
    const int64_t tiles = static_cast<int64_t>(b.ntx) * b.nty;
    BZH_REQUIRE(b.band == 0 || (b.band >= 1 && 8 * int64_t{b.band} >= tiles),
                "pair bands do not cover the tiles");
    const int64_t need = (b.band > 0 ? 8 * int64_t{b.band} : tiles) * b.nzc;
    BZH_REQUIRE(b.nblk % 8 == 0 && b.nblk >= need && b.nblk < need + 8,
                "pair block count does not match its box");
    BZH_REQUIRE(tiles == 0 || b.nblk > 0, "pair box without blocks");
    if (b.xedge != 1 && tiles > 0) {
      // the interior variant reads lanes x = i0-2 .. i0+61 unchecked and
      // assumes no x face or x CPML slab under them
      const int lo = b.tx0 * kPairCols - 2;
      const int hi = (b.tx0 + b.ntx + b.tskip - 1) * kPairCols + kPairCols + 1;
      BZH_REQUIRE(lo >= M && hi <= p.nx - M - 1, "interior pair tiles overlap an x face or x CPML slab");
    }
    BZH_REQUIRE(b.xedge >= 0 && b.xedge <= 2 && (!pp.core || b.xedge == 2),
                "pair box kind is malformed");
    if (b.xedge == 2 && tiles > 0) {
      // core tiles: rows j0-2 .. j0+13 in [M, ny-L-1] and iterations
      // kz0-2 .. kz1+1 touching planes in [M, nz-M-1] (lean everywhere)
      BZH_REQUIRE(b.tyskip == 0 && b.nzc == b.nzc_lo, "core box is malformed");
      const int y_lo = b.ty0 * kPairRows - 2, y_hi = (b.ty0 + b.nty - 1) * kPairRows + kPairRows;
      BZH_REQUIRE(y_lo >= M && y_hi <= p.ny - p.L - 1, "core tiles reach a y face or y CPML slab");
      BZH_REQUIRE(b.z0 - 4 >= M && b.zsplit + 2 <= p.nz - M - 1,
                  "core chunks reach a z face or z CPML slab");
    }
    nblocks += b.nblk;
  }
  BZH_REQUIRE(p.L == 0 || p.tab != nullptr, "CPML tiles need the coefficient table");
  BZH_REQUIRE(p.fin != nullptr && p.fout != nullptr && (p.L == 0 || p.pbase != nullptr) &&
                  (p.esrc_st[0] == 0 || p.cbase != nullptr),
              "the pair step needs based addressing of its array groups");
  for (int c = 0; c < 3; ++c) {
    BZH_REQUIRE(p.hdec_st[c] == 0 && p.hsrc_st[c] == 0 && p.edec_st[c] == 0,
                "the pair step needs uniform H coefficients and E decay");
    BZH_REQUIRE(p.esrc_st[c] == p.esrc_st[0], "E source coefficients must be all arrays or all scalars");
  }
  BZH_REQUIRE(pp.msrc_n >= 0 && pp.msrc_n <= kMaxMidSlabs && 0 <= pp.msrc_post &&
                  pp.msrc_post <= pp.msrc_e && pp.msrc_e <= pp.msrc_h1 && pp.msrc_h1 <= pp.msrc_n &&
                  (pp.msrc_n == 0 || pp.step0 != nullptr),
              "pair source slabs are malformed");
  for (int i = 0; i < pp.msrc_n; ++i) {
    const MidSlab& m = pp.msrc[i];
    BZH_REQUIRE(m.coeff && m.wave && m.start && m.sz >= 0 && m.sy >= 0 && m.sx >= 0 && m.T >= 1 &&
                    m.comp >= 0 && m.comp < 3 && m.hfield == (i >= pp.msrc_e ? 1 : 0),
                "pair source slab is malformed");
  }
  BZH_REQUIRE(p.hslab_n == 0, "pair launches carry their H slabs in the pair list");
  BZH_REQUIRE(pp.cap == nullptr || (pp.cmid == nullptr && pp.mbox_n == 0 && pp.capbox != nullptr &&
                                     pp.cap_m >= 1 && pp.cap_m <= kMaxMonBoxes && pp.cap_slot > 0),
              "pair capture is malformed");
  BZH_REQUIRE(pp.mbox_n >= 0 && pp.mbox_n <= kMaxMonBoxes &&
                  (pp.mbox_n == 0 || (pp.mbox != nullptr && pp.cmid != nullptr)),
              "pair monitor boxes are malformed");
  BZH_REQUIRE(nblocks < (int64_t{1} << 31), "pair grid too large");
  if (nblocks <= 0) return;
  switch (dtype) {
    case BZH_F32: launch_pair_f32(s, pp, static_cast<int>(nblocks)); break;
    case BZH_BF16: launch_pair_bf16(s, pp, static_cast<int>(nblocks)); break;
    case BZH_F16: launch_pair_f16(s, pp, static_cast<int>(nblocks)); break;
    default: throw Failure{"unsupported field dtype"};
  }
  BZH_HIP(hipGetLastError());
}
}  // namespace bzh
