// C-ABI layer of libbeamz_hip.so: descriptor validation, positional decoding
// of the reference's FFI layouts, the per-step schedule and the hipGraph
// cache.
//
// Reference counterparts:
//   ffi_handler.cc:46-189, 290-334, 336-708  -> decode_* below
//   program.cu:110-170 (ValidateProgram)     -> validate_* below
//   program.cu:172-335 (schedules)           -> build_schedule
//   graph.cu:84-154 (graph key, cache)       -> GraphCache
// The per-step order is execute.py:467-541: pre_e sources, H, h sources,
// (H masks, folded into the kernels), E, e sources, (E masks), monitors.

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace bzh {

namespace {
thread_local std::string g_error;
}

thread_local int32_t g_last_fused = 0;
thread_local int32_t g_last_pairs = 0;
thread_local int32_t g_last_capture = 0;
void set_error(const std::string& message) { g_error = message; }
void clear_error() { g_error.clear(); }

namespace {

// ---- buffer helpers -----------------------------------------------------------
int64_t numel(const bzh_buffer& b) {
  if (b.rank < 0 || b.rank > 4) return -1;
  int64_t n = 1;
  for (int d = 0; d < b.rank; ++d) n *= b.dims[d];
  return n;
}

bool is_empty(const bzh_buffer& b) { return b.data == nullptr || numel(b) == 0; }

bool dims_equal(const bzh_buffer& b, int rank, const int64_t* dims) {
  if (b.rank != rank) return false;
  for (int d = 0; d < rank; ++d)
    if (b.dims[d] != dims[d]) return false;
  return true;
}

std::string shape_str(const bzh_buffer& b) {
  std::string s = "(";
  for (int d = 0; d < b.rank; ++d) s += (d ? ", " : "") + std::to_string(b.dims[d]);
  return s + ")";
}

bool is_float_storage(int t) { return t == BZH_F32 || t == BZH_BF16 || t == BZH_F16; }

// Index of a component in the (Hx, Hy, Hz, Ex, Ey, Ez) bank order.
enum { kHx = 0, kHy, kHz, kEx, kEy, kEz };

// Complete Yee support of each component (lattice.py:32-45), 3D.
void yee_shape3d(int c, int nz, int ny, int nx, int64_t out[3]) {
  static const int off[6][3] = {
      // (z, y, x) "+1" where the component's offset is 0
      {0, 0, 1},  // Hx (nz, ny, nx+1)
      {0, 1, 0},  // Hy (nz, ny+1, nx)
      {1, 0, 0},  // Hz (nz+1, ny, nx)
      {1, 1, 0},  // Ex (nz+1, ny+1, nx)
      {1, 0, 1},  // Ey (nz+1, ny, nx+1)
      {0, 1, 1},  // Ez (nz, ny+1, nx+1)
  };
  out[0] = nz + off[c][0];
  out[1] = ny + off[c][1];
  out[2] = nx + off[c][2];
}

struct Grid {
  bool is3d = true;
  int nz = 1, ny = 0, nx = 0;
  int dtype = BZH_F32;
};

// Infer and validate the material grid from the six fields (bank order).
Grid infer_grid(const bzh_buffer* f[6]) {
  Grid g;
  const bzh_buffer& ez = *f[kEz];
  if (ez.rank == 2) {
    // TMxy: Ez (ny+1, nx+1), Hx (ny, nx+1), Hy (ny+1, nx)
    g.is3d = false;
    g.ny = static_cast<int>(ez.dims[0]) - 1;
    g.nx = static_cast<int>(ez.dims[1]) - 1;
    BZH_REQUIRE(g.ny >= 1 && g.nx >= 1, "TMxy grid must have at least one cell");
    const int64_t hx[2] = {g.ny, g.nx + 1}, hy[2] = {g.ny + 1, g.nx};
    BZH_REQUIRE(dims_equal(*f[kHx], 2, hx), "Hx shape " + shape_str(*f[kHx]) + " does not match Ez");
    BZH_REQUIRE(dims_equal(*f[kHy], 2, hy), "Hy shape " + shape_str(*f[kHy]) + " does not match Ez");
    g.dtype = ez.element_type;
    for (int c : {kHx, kHy, kEz}) {
      BZH_REQUIRE(f[c]->element_type == g.dtype, "TMxy fields must share one dtype");
      BZH_REQUIRE(f[c]->data != nullptr, "TMxy field pointer is null");
    }
  } else {
    BZH_REQUIRE(f[kHx]->rank == 3, "3D fields must be rank 3");
    g.nz = static_cast<int>(f[kHx]->dims[0]);
    g.ny = static_cast<int>(f[kHx]->dims[1]);
    g.nx = static_cast<int>(f[kHx]->dims[2]) - 1;
    BZH_REQUIRE(g.nz >= 1 && g.ny >= 1 && g.nx >= 1, "3D grid must have at least one cell");
    g.dtype = f[kHx]->element_type;
    for (int c = 0; c < 6; ++c) {
      int64_t s[3];
      yee_shape3d(c, g.nz, g.ny, g.nx, s);
      BZH_REQUIRE(dims_equal(*f[c], 3, s),
                  "field " + std::to_string(c) + " shape " + shape_str(*f[c]) +
                      " is not a complete Yee support of the grid");
      BZH_REQUIRE(f[c]->element_type == g.dtype, "fields must share one dtype");
      BZH_REQUIRE(f[c]->data != nullptr, "field pointer is null");
    }
  }
  BZH_REQUIRE(is_float_storage(g.dtype), "fields must be f32, bf16 or f16");
  const int64_t cells = static_cast<int64_t>(g.nz + 1) * (g.ny + 1) * (g.nx + 1);
  BZH_REQUIRE(cells < (int64_t{1} << 31), "grid exceeds 32-bit indexing");
  return g;
}

// Which ends of each (z, y, x) axis mask a component on PEC faces
// (_boundary_compile.py:118-160).  2D uses axes 1 (y), 2 (x).
void mask_ends(int comp, int metallic, bool is3d, int lo[3], int hi[3]) {
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = 0;
  auto face = [&](int axis) {
    static const int bits[3][2] = {{kFront, kBack}, {kBottom, kTop}, {kLeft, kRight}};
    lo[axis] = (metallic & bits[axis][0]) != 0;
    hi[axis] = (metallic & bits[axis][1]) != 0;
  };
  if (is3d) {
    switch (comp) {
      case kEx: face(0); face(1); break;
      case kEy: face(0); face(2); break;
      case kEz: face(1); face(2); break;
      case kHx: face(2); break;
      case kHy: face(1); break;
      case kHz: face(0); break;
    }
  } else {
    switch (comp) {
      case kEz: face(1); face(2); break;
      case kHx: face(2); break;
      case kHy: face(1); break;
      default: break;
    }
  }
}

// ---- phase decoding ---------------------------------------------------------
struct PhaseIn {
  const bzh_buffer* tin[3];
  const bzh_buffer* tout[3];
  const bzh_buffer* src[3];
  const bzh_buffer* mat[6];
  const bzh_buffer* meta;
  const bzh_buffer* terms[18];
  const bzh_buffer* psi_in[6];
  const bzh_buffer* psi_out[6];
  const bzh_buffer* metric[3];
  int nterms;
};

// Derivative axis of each term in storage order (z, y, x) = (0, 1, 2).
// 3D: update.cu order Hx_y, Hx_z, Hy_z, Hy_x, Hz_x, Hz_y (and E alike).
const int kAxes3d[6] = {1, 0, 0, 2, 2, 1};
// TMxy H terms (Hx_y, Hy_x) and E terms (Ez_x, Ez_y), in (z, y, x) axes.
const int kAxesTmH[2] = {1, 2};
const int kAxesTmE[2] = {2, 1};

// A coefficient is an f32 device scalar or an f32 array of the target's shape;
// with ``f16`` non-null an f16 array is accepted too (the folded E source
// coefficients of half-precision programs, read only by the fused step).
void fill_coef(const bzh_buffer* b, int64_t expect, const float** ptr, int32_t* st,
               float* s, const char* what, int32_t* f16 = nullptr) {
  *ptr = nullptr;
  *st = 0;
  *s = 0.0f;
  if (b == nullptr) return;
  const int64_t n = numel(*b);
  const bool half_array = f16 != nullptr && b->element_type == BZH_F16 && b->rank > 0 && n > 1;
  BZH_REQUIRE(b->element_type == BZH_F32 || half_array,
              std::string(what) + (f16 ? " must be f32 (or an f16 array)" : " must be f32"));
  if (half_array) *f16 = 1;
  if (b->rank == 0 || n == 1) {
    BZH_REQUIRE(b->data != nullptr, std::string(what) + " scalar pointer is null");
    *ptr = static_cast<const float*>(b->data);
    *st = 0;
    return;
  }
  BZH_REQUIRE(n == expect, std::string(what) + " has " + std::to_string(n) +
                               " elements, expected " + std::to_string(expect));
  *ptr = static_cast<const float*>(b->data);
  *st = 1;
}

// The E phase's component c in the packed coefficient-ID form the
// reference's compiler emits for lossless programs whose E source
// coefficients take at most 256 values (compile.py:146-187): the decay slot
// holds the component's f32 table, the source slot its cells' 8-bit IDs,
// four per int32 word.  n: the component's cell count.
bool coef_ids(const PhaseIn& e, int c, int64_t n) {
  const bzh_buffer* tab = e.mat[c];
  const bzh_buffer* ids = e.mat[3 + c];
  return tab != nullptr && ids != nullptr && tab->data != nullptr && ids->data != nullptr &&
         tab->element_type == BZH_F32 && tab->rank == 1 && numel(*tab) >= 1 &&
         numel(*tab) <= 256 && ids->element_type == BZH_S32 && ids->rank == 1 &&
         numel(*ids) == (n + 3) / 4;
}

// Build kernel parameters for one phase.  ``fields`` is the bank-ordered field
// set used for geometry (targets and sources may come from different banks).
PhaseParams make_phase(const PhaseIn& in, int phase, const Grid& g, float dt,
                       float resolution, int metallic, int uniform_l) {
  PhaseParams p;
  std::memset(&p, 0, sizeof(p));
  p.nz = g.nz;
  p.ny = g.ny;
  p.nx = g.nx;
  p.metallic = metallic;
  p.inv_dx = 1.0f / resolution;
  p.ratio = 1.0f;
  const int base = phase == BZH_PHASE_H ? kHx : kEx;
  const int other = phase == BZH_PHASE_H ? kEx : kHx;
  (void)dt;
  for (int c = 0; c < 3; ++c) {
    const bool active = g.is3d || (phase == BZH_PHASE_H ? c < 2 : c == 2);
    const bool src_active = g.is3d || (phase == BZH_PHASE_H ? c == 2 : c < 2);
    int64_t tshape[3] = {1, 1, 1};
    if (g.is3d) yee_shape3d(base + c, g.nz, g.ny, g.nx, tshape);
    if (active) {
      BZH_REQUIRE(in.tin[c]->data != nullptr, "target field pointer is null");
      p.tin[c] = in.tin[c]->data;
      p.tout[c] = (in.tout[c] && in.tout[c]->data) ? in.tout[c]->data : in.tin[c]->data;
      if (in.tout[c] && in.tout[c]->data)
        BZH_REQUIRE(numel(*in.tout[c]) == numel(*in.tin[c]) &&
                        in.tout[c]->element_type == in.tin[c]->element_type,
                    "output field does not match its input");
      const int64_t n = numel(*in.tin[c]);
      if (phase == BZH_PHASE_E && coef_ids(in, c, n)) {
        // lossless: decay 1, source coefficient from the table by ID
        p.dec[c] = nullptr;
        p.dec_s[c] = 1.0f;
        p.scf[c] = static_cast<const float*>(in.mat[c]->data);
        p.scf_ids[c] = static_cast<const int32_t*>(in.mat[3 + c]->data);
      } else {
        fill_coef(in.mat[c], n, &p.dec[c], &p.dec_st[c], &p.dec_s[c], "decay coefficient");
        fill_coef(in.mat[3 + c], n, &p.scf[c], &p.scf_st[c], &p.scf_s[c], "source coefficient",
                  phase == BZH_PHASE_E ? &p.scf_f16 : nullptr);
        BZH_REQUIRE(p.dec[c] != nullptr, "every active component needs a decay coefficient");
      }
      BZH_REQUIRE(p.scf[c] != nullptr, "every active component needs source coefficients");
    }
    if (src_active) {
      BZH_REQUIRE(in.src[c]->data != nullptr, "source field pointer is null");
      p.src[c] = in.src[c]->data;
    }
    (void)other;
  }
  // metrics (z, y, x): empty -> inv_dx, rank 0 -> device scalar, rank 1 -> array
  const int ext[3] = {g.nz, g.ny, g.nx};
  for (int a = 0; a < 3; ++a) {
    const bzh_buffer* m = in.metric[a];
    p.metric[a] = nullptr;
    p.met_st[a] = 0;
    if (m == nullptr || m->data == nullptr || numel(*m) == 0) continue;
    BZH_REQUIRE(m->element_type == BZH_F32, "metrics must be f32");
    if (numel(*m) == 1) {
      p.metric[a] = static_cast<const float*>(m->data);
      continue;
    }
    const int64_t need = phase == BZH_PHASE_H ? ext[a] : ext[a] + 1;
    BZH_REQUIRE(numel(*m) >= need, "metric array too short for its axis");
    p.metric[a] = static_cast<const float*>(m->data);
    p.met_st[a] = 1;
  }
  // CPML terms
  p.nterms = in.nterms;
  p.uniform_l = uniform_l;
  if (in.nterms > 0) {
    const int want = g.is3d ? 6 : 2;
    BZH_REQUIRE(in.nterms == want, g.is3d ? "3D CPML needs six terms per phase"
                                          : "TMxy CPML needs two terms per phase");
    BZH_REQUIRE(in.meta && in.meta->element_type == BZH_S32 && in.meta->rank == 2 &&
                    in.meta->dims[0] == want && in.meta->dims[1] == 5,
                "term metadata must be s32[nterms, 5]");
    p.meta = static_cast<const int32_t*>(in.meta->data);
    for (int t = 0; t < in.nterms; ++t) {
      const int comp = g.is3d ? t / 2 : (phase == BZH_PHASE_H ? t : 2);
      const int axis = g.is3d ? kAxes3d[t] : (phase == BZH_PHASE_H ? kAxesTmH[t] : kAxesTmE[t]);
      const bzh_buffer& tgt = *in.tin[comp];
      const bzh_buffer& psi = *in.psi_in[t];
      const int raxis = g.is3d ? axis : axis - 1;  // axis within the target's rank
      BZH_REQUIRE(psi.rank == tgt.rank, "psi rank must match its target");
      int64_t ext_axis = 0;
      for (int d = 0; d < tgt.rank; ++d) {
        if (d == raxis) {
          ext_axis = psi.dims[d];
        } else {
          BZH_REQUIRE(psi.dims[d] == tgt.dims[d], "psi " + std::to_string(t) + " shape " +
                                                      shape_str(psi) + " does not match its target " +
                                                      shape_str(tgt));
        }
      }
      BZH_REQUIRE(psi.element_type == BZH_F32 || psi.element_type == BZH_BF16,
                  "psi must be f32 or bf16");
      if (uniform_l > 0)
        BZH_REQUIRE(ext_axis == 2 * uniform_l || ext_axis == 0,
                    "psi extent does not match the uniform CPML thickness");
      DevTerm& d = p.term[t];
      d.psi = psi.data;
      if (in.psi_out[t] && in.psi_out[t]->data)
        BZH_REQUIRE(in.psi_out[t]->data == psi.data,
                    "per-phase launches update psi in place: outputs must alias inputs");
      d.low = static_cast<int32_t>(ext_axis);
      d.high = 0;
      d.n = static_cast<int32_t>(tgt.dims[raxis]);
      d.psi_bf16 = psi.element_type == BZH_BF16;
      for (int q = 0; q < 3; ++q) {
        const bzh_buffer* tb = in.terms[t * 3 + q];
        BZH_REQUIRE(tb && tb->element_type == BZH_F32, "CPML term arrays must be f32");
        BZH_REQUIRE(numel(*tb) >= ext_axis, "CPML term array shorter than its slab");
        BZH_REQUIRE(ext_axis == 0 || tb->data != nullptr, "CPML term array pointer is null");
      }
      d.a = static_cast<const float*>(in.terms[t * 3 + 0]->data);
      d.b = static_cast<const float*>(in.terms[t * 3 + 1]->data);
      d.ik = static_cast<const float*>(in.terms[t * 3 + 2]->data);
      if (ext_axis == 0) d.psi = nullptr;
    }
  }
  return p;
}

void run_phase(hipStream_t s, const PhaseParams& p, int phase, const Grid& g) {
  if (g.is3d) launch_phase3d(s, p, phase, g.dtype);
  else launch_phase_tm(s, p, phase, g.dtype);
}

// ---- schedule ------------------------------------------------------------------
enum class Kind : int32_t {
  kPhase3d, kPhaseTm, kSource, kDft, kFused, kBatch2d, kAdjoint, kPair, kMonBox
};

struct Launch {
  Kind kind;
  int32_t phase;
  int32_t dtype;
  std::vector<uint8_t> bytes;
};

template <typename P>
Launch make_launch(Kind kind, int phase, int dtype, const P& p) {
  Launch l;
  l.kind = kind;
  l.phase = phase;
  l.dtype = dtype;
  l.bytes.resize(sizeof(P));
  std::memcpy(l.bytes.data(), &p, sizeof(P));
  return l;
}

void enqueue(hipStream_t s, const Launch& l) {
  switch (l.kind) {
    case Kind::kPhase3d: {
      PhaseParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_phase3d(s, p, l.phase, l.dtype);
      break;
    }
    case Kind::kPhaseTm: {
      PhaseParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_phase_tm(s, p, l.phase, l.dtype);
      break;
    }
    case Kind::kSource: {
      SourceParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_source(s, p);
      break;
    }
    case Kind::kBatch2d: {
      Batch2dParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      if (l.phase == 1) launch_batch2d_sources(s, p);
      else launch_batch2d(s, p);
      break;
    }
    case Kind::kAdjoint: {
      Adjoint2dParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_adjoint_step(s, p);
      break;
    }
    case Kind::kDft: {
      DftParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_dft(s, p);
      break;
    }
    case Kind::kFused: {
      FusedParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_fused3d(s, p, l.dtype);
      break;
    }
    case Kind::kPair: {
      PairParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_pair3d(s, p, l.dtype);
      break;
    }
    case Kind::kMonBox: {
      MonBoxParams p;
      std::memcpy(&p, l.bytes.data(), sizeof(p));
      launch_monbox(s, p);
      break;
    }
  }
}

// ---- graph cache (graph.cu:84-154) -------------------------------------------
class GraphCache {
 public:
  ~GraphCache() {
    // Executables are owned by the process; the runtime may already be torn down
    // at static destruction, so they are intentionally not destroyed here.
  }

  // Launch ``launches`` on ``s``; replay a cached graph when the kernel
  // arguments match a previous call exactly.  A stream the caller is itself
  // capturing gets the kernels directly (they become part of the caller's
  // graph).  A full cache evicts its least recently launched entry once that
  // entry's last launch has completed.
  void run(hipStream_t s, const std::vector<Launch>& launches, bool use_graph) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (use_graph && s != nullptr) BZH_HIP(hipStreamIsCapturing(s, &cap));
    if (!use_graph || s == nullptr || cap != hipStreamCaptureStatusNone) {
      for (const auto& l : launches) enqueue(s, l);
      return;
    }
    std::string key;
    int device = 0;
    BZH_HIP(hipGetDevice(&device));
    key.append(reinterpret_cast<const char*>(&device), sizeof(device));
    for (const auto& l : launches) {
      key.append(reinterpret_cast<const char*>(&l.kind), sizeof(l.kind));
      key.append(reinterpret_cast<const char*>(&l.phase), sizeof(l.phase));
      key.append(reinterpret_cast<const char*>(&l.dtype), sizeof(l.dtype));
      key.append(reinterpret_cast<const char*>(l.bytes.data()), l.bytes.size());
    }
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto it = cache_.find(key);
      if (it != cache_.end()) {
        launch(it->second, s);
        return;
      }
    }
    {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        for (const auto& l : launches) enqueue(s, l);
        return;
      }
      bool ok = true;
      std::string why;
      try {
        for (const auto& l : launches) enqueue(s, l);
      } catch (const Failure& f) {
        ok = false;
        why = f.message;
      }
      hipError_t end = hipStreamEndCapture(s, &graph);
      if (!ok) {
        if (graph) (void)hipGraphDestroy(graph);
        throw Failure{why};
      }
      if (end != hipSuccess || graph == nullptr ||
          hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        if (graph) (void)hipGraphDestroy(graph);
        for (const auto& l : launches) enqueue(s, l);
        return;
      }
      (void)hipGraphDestroy(graph);
      hipEvent_t done = nullptr;
      BZH_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
      std::lock_guard<std::mutex> lock(mu_);
      auto it = cache_.find(key);
      if (it != cache_.end()) {
        // another thread cached the same launch list meanwhile
        (void)hipGraphExecDestroy(exec);
        (void)hipEventDestroy(done);
      } else {
        if (cache_.size() >= kCapacity) evict_oldest();
        it = cache_.emplace(key, Entry{exec, done, 0}).first;
      }
      launch(it->second, s);
    }
  }

  size_t size() {
    std::lock_guard<std::mutex> lock(mu_);
    return cache_.size();
  }

 private:
  struct Entry {
    hipGraphExec_t exec;
    hipEvent_t done;   // recorded after each launch of ``exec``
    uint64_t tick;     // launch order
  };
  // Called with mu_ held.
  void launch(Entry& e, hipStream_t s) {
    e.tick = ++tick_;
    BZH_HIP(hipGraphLaunch(e.exec, s));
    BZH_HIP(hipEventRecord(e.done, s));
  }
  // Called with mu_ held.  Waits for the entry's last launch to complete
  // before destroying its executable.
  void evict_oldest() {
    auto victim = cache_.begin();
    for (auto it = cache_.begin(); it != cache_.end(); ++it)
      if (it->second.tick < victim->second.tick) victim = it;
    BZH_HIP(hipEventSynchronize(victim->second.done));
    (void)hipGraphExecDestroy(victim->second.exec);
    (void)hipEventDestroy(victim->second.done);
    cache_.erase(victim);
  }

  static constexpr size_t kCapacity = 32;
  std::mutex mu_;
  uint64_t tick_ = 0;
  std::map<std::string, Entry> cache_;
};

GraphCache& graph_cache() {
  static GraphCache* cache = new GraphCache();
  return *cache;
}

// ---- source groups ---------------------------------------------------------------
// (timing, component) order of the nine groups: execute.py:48-52
//   pre_e: Ex Ey Ez, h: Hx Hy Hz, e: Ex Ey Ez
struct SourceGroupIn {
  const bzh_buffer* coeffs;
  const bzh_buffer* waves;
  const bzh_buffer* starts;
  int coincident;
};

void add_source_launches(std::vector<Launch>& out, const SourceGroupIn& grp, int timing,
                         int comp_in_phase, const bzh_buffer* const* bank, const Grid& g,
                         int metallic, const int32_t* step0, int k) {
  const bzh_buffer& c = *grp.coeffs;
  if (c.data == nullptr || c.rank < 1 || c.dims[0] == 0) return;
  const int n = static_cast<int>(c.dims[0]);
  const int comp = (timing == 1 ? kHx : kEx) + comp_in_phase;
  const bzh_buffer& field = *bank[comp];
  BZH_REQUIRE(field.data != nullptr, "source targets an absent field component");
  BZH_REQUIRE(c.element_type == BZH_F32 && grp.waves->element_type == BZH_F32 &&
                  grp.starts->element_type == BZH_S32,
              "source groups are f32 coefficients, f32 waveforms and s32 starts");
  BZH_REQUIRE(grp.waves->rank == 2 && grp.waves->dims[0] == n && grp.waves->dims[1] >= 1,
              "source waveforms must be [n, T]");
  const int fr = field.rank;
  BZH_REQUIRE(grp.starts->rank == 2 && grp.starts->dims[0] == n && grp.starts->dims[1] == 3,
              "source starts must be s32[n, 3]");
  BZH_REQUIRE(c.rank == 4, "source coefficients must be [n, sz, sy, sx]");
  SourceParams p;
  std::memset(&p, 0, sizeof(p));
  p.field = field.data;
  p.dtype = field.element_type;
  p.scale = 1.0f;
  if (fr == 3) {
    for (int d = 0; d < 3; ++d) p.fdims[d] = static_cast<int>(field.dims[d]);
  } else {
    BZH_REQUIRE(fr == 2, "source target must be rank 2 or 3");
    p.fdims[0] = 1;
    p.fdims[1] = static_cast<int>(field.dims[0]);
    p.fdims[2] = static_cast<int>(field.dims[1]);
  }
  p.coeff = static_cast<const float*>(c.data);
  p.wave = static_cast<const float*>(grp.waves->data);
  p.starts = static_cast<const int32_t*>(grp.starts->data);
  p.n = n;
  for (int d = 0; d < 3; ++d) p.sdims[d] = static_cast<int>(c.dims[1 + d]);
  p.T = static_cast<int>(grp.waves->dims[1]);
  p.step0 = step0;
  p.k = k;
  if (timing != 0) {
    int lo[3], hi[3];
    mask_ends(comp, metallic, g.is3d, lo, hi);
    for (int d = 0; d < 3; ++d) {
      p.mlo[d] = lo[d];
      p.mhi[d] = hi[d];
    }
  }
  if (grp.coincident) {
    p.slab0 = 0;
    p.nslabs = n;
    out.push_back(make_launch(Kind::kSource, 0, p.dtype, p));
  } else {
    // overlapping slabs are applied in declaration order, one launch each
    for (int s = 0; s < n; ++s) {
      p.slab0 = s;
      p.nslabs = 1;
      out.push_back(make_launch(Kind::kSource, 0, p.dtype, p));
    }
  }
}

// Fold a timed source group into a 2D phase launch (PhaseParams::psrc): the
// checks of add_source_launches, then the group's slabs, summed per cell in
// declaration order inside the kernel.
void fold_phase_sources(PhaseParams& p, const SourceGroupIn& grp, int comp_in_phase, int comp,
                        const bzh_buffer* const* bank, const Grid& g, int metallic,
                        const int32_t* step0, int k) {
  const bzh_buffer& c = *grp.coeffs;
  if (c.data == nullptr || c.rank < 1 || c.dims[0] == 0) return;
  const int n = static_cast<int>(c.dims[0]);
  const bzh_buffer& field = *bank[comp];
  BZH_REQUIRE(field.data != nullptr && field.rank == 2, "source targets an absent field component");
  BZH_REQUIRE(c.element_type == BZH_F32 && grp.waves->element_type == BZH_F32 &&
                  grp.starts->element_type == BZH_S32,
              "source groups are f32 coefficients, f32 waveforms and s32 starts");
  BZH_REQUIRE(grp.waves->rank == 2 && grp.waves->dims[0] == n && grp.waves->dims[1] >= 1,
              "source waveforms must be [n, T]");
  BZH_REQUIRE(grp.starts->rank == 2 && grp.starts->dims[0] == n && grp.starts->dims[1] == 3,
              "source starts must be s32[n, 3]");
  BZH_REQUIRE(c.rank == 4, "source coefficients must be [n, sz, sy, sx]");
  BZH_REQUIRE(p.npsrc < kMaxPhaseSrc, "too many folded source groups");
  PhaseSrc& q = p.psrc[p.npsrc++];
  q.coeff = static_cast<const float*>(c.data);
  q.wave = static_cast<const float*>(grp.waves->data);
  q.starts = static_cast<const int32_t*>(grp.starts->data);
  q.n = n;
  q.sz = static_cast<int>(c.dims[1]);
  q.sy = static_cast<int>(c.dims[2]);
  q.sx = static_cast<int>(c.dims[3]);
  q.T = static_cast<int>(grp.waves->dims[1]);
  q.comp = comp_in_phase;
  q.fy = static_cast<int>(field.dims[0]);
  q.fx = static_cast<int>(field.dims[1]);
  mask_ends(comp, metallic, g.is3d, q.mlo, q.mhi);
  p.step0 = step0;
  p.k = k;
}

// ---- monitors ----------------------------------------------------------------------
struct MonitorIn {
  const bzh_buffer* b[12];
};

DftParams make_dft(const MonitorIn& in, const bzh_buffer* const* bank, const Grid& g,
                   float dt, int monitor_count) {
  DftParams p;
  std::memset(&p, 0, sizeof(p));
  const bzh_buffer& idx = *in.b[0];
  const bzh_buffer& w = *in.b[1];
  const bzh_buffer& fr = *in.b[2];
  BZH_REQUIRE(idx.element_type == BZH_S32 && idx.rank == 4 && idx.dims[1] == 6,
              "monitor indices must be s32[M, 6, P, K]");
  BZH_REQUIRE(w.element_type == BZH_F32 && dims_equal(w, 4, idx.dims),
              "monitor weights must be f32 with the shape of the indices");
  p.M = static_cast<int>(idx.dims[0]);
  BZH_REQUIRE(monitor_count <= 0 || monitor_count == p.M, "monitor_count does not match");
  p.Pmax = static_cast<int>(idx.dims[2]);
  p.K = static_cast<int>(idx.dims[3]);
  BZH_REQUIRE(fr.rank == 2 && fr.dims[0] == p.M, "monitor frequencies must be [M, F]");
  BZH_REQUIRE(fr.element_type == BZH_F32 || fr.element_type == BZH_F64,
              "monitor frequencies must be f32 or f64");
  p.Fmax = static_cast<int>(fr.dims[1]);
  p.freqs = fr.data;
  p.freqs_f64 = fr.element_type == BZH_F64;
  BZH_REQUIRE(in.b[3]->element_type == BZH_F32 && numel(*in.b[3]) == 6 * p.M,
              "component masks must be f32[M, 6]");
  BZH_REQUIRE(in.b[4]->element_type == BZH_S32 && numel(*in.b[4]) == 5 * p.M,
              "monitor counts must be s32[M, 5]");
  BZH_REQUIRE(in.b[5]->element_type == BZH_S32 && numel(*in.b[5]) == 2 * p.M,
              "monitor codes must be s32[M, 2]");
  BZH_REQUIRE(in.b[6]->element_type == BZH_F32 && numel(*in.b[6]) == 3 * p.M,
              "monitor windows must be f32[M, 3]");
  p.indices = static_cast<const int32_t*>(idx.data);
  p.weights = static_cast<const float*>(w.data);
  p.masks = static_cast<const float*>(in.b[3]->data);
  p.counts = static_cast<const int32_t*>(in.b[4]->data);
  p.codes = static_cast<const int32_t*>(in.b[5]->data);
  p.windows = static_cast<const float*>(in.b[6]->data);
  const bzh_buffer& re = *in.b[7];
  const bzh_buffer& im = *in.b[8];
  const bzh_buffer& ws = *in.b[9];
  BZH_REQUIRE(re.element_type == im.element_type && re.element_type == ws.element_type &&
                  (re.element_type == BZH_F64 || re.element_type == BZH_F32),
              "DFT accumulators must all be f64 (or all f32)");
  BZH_REQUIRE(numel(re) == numel(im), "DFT re/im arenas differ in length");
  p.re = re.data;
  p.im = im.data;
  p.wsum = ws.data;
  p.acc_f64 = re.element_type == BZH_F64;
  p.re_len = numel(re);
  p.w_len = numel(ws);
  const bzh_buffer& t = *in.b[10];
  BZH_REQUIRE((t.element_type == BZH_F64 || t.element_type == BZH_F32) && numel(t) == 1,
              "monitor time must be an f32 or f64 scalar");
  p.time = t.data;
  p.time_f64 = t.element_type == BZH_F64;
  BZH_REQUIRE(in.b[11]->element_type == BZH_S32 && numel(*in.b[11]) == 1,
              "current_step must be an s32 scalar");
  p.step0 = static_cast<const int32_t*>(in.b[11]->data);
  p.dt = dt;
  p.dtype = g.dtype;
  // bank order Hx..Ez -> DFT order Ex, Ey, Ez, Hx, Hy, Hz
  const int order[6] = {kEx, kEy, kEz, kHx, kHy, kHz};
  for (int c = 0; c < 6; ++c) {
    const bzh_buffer& f = *bank[order[c]];
    const bool active = g.is3d || order[c] == kEz || order[c] == kHx || order[c] == kHy;
    p.field[c] = active ? f.data : nullptr;
    p.fsize[c] = active ? static_cast<int32_t>(numel(f)) : 0;
    p.inv_scale[c] = 1.0f;
    for (int d = 0; d < 3; ++d)
      p.fdims[c][d] = !active ? 0 : g.is3d ? static_cast<int32_t>(f.dims[d])
                                        : (d == 0 ? 1 : static_cast<int32_t>(f.dims[d - 1]));
  }
  return p;
}

// ---- program ------------------------------------------------------------------
struct ProgramIn {
  const bzh_buffer* bank_a[6];   // Hx Hy Hz Ex Ey Ez
  const bzh_buffer* bank_b[6];   // workspace (temporal layouts) or nullptr
  const bzh_buffer* psi_b[12];   // psi workspace (H terms then E terms) or nullptr
  PhaseIn h;                     // targets/sources bound per step below
  PhaseIn e;
  SourceGroupIn groups[9];
  int ngroups;
  const bzh_buffer* step;        // s32 scalar for the sources (may be null without sources)
  MonitorIn mon;
  bool has_monitors;
  int monitor_count;
  int nsteps;
  float dt, resolution;
  int metallic, uniform_l;
  int flags;
  bool noalloc;                  // bzh_program_noalloc: scratch from the device array, no graphs
  // optional trailing argument of the monitor layouts (no counterpart in the
  // reference): host s32[M, 6, 6] boxes of the cells each monitor component
  // gathers, which lets pair launches capture them densely
  const bzh_buffer* mon_boxes;
};

// Work split of the fused step -- the counterpart of the reference's
// shell/core block queue (update.cu:655-755).  Every (x, y) tile over every
// z chunk belongs to one of two boxes: the x-interior box, whose lanes
// x = i0-1 .. i0+62 all lie in [M, nx - M) with M = max(L, 1) (no x face, x
// ghost or x CPML slab, so the kernel has no per-lane x checks), and the
// x-edge box (the tiles left of and right of it).  y and z boundaries are
// handled by row- and plane-uniform branches in both.
struct FusedBox {
  int tx0, ntx, tsplit, tskip, nty, kz, nzc;
};
struct FusedSplit {
  FusedBox interior, edge;
};

// z chunk length (BZH_FUSED_KZ overrides it for tuning sweeps): 12 planes
// (the recomputed first plane costs 1/12 of a chunk), shortened (down to 4)
// while the launch has fewer than 2048 workgroups.
int chunk_len(int64_t tiles, int planes) {
  if (const char* env = std::getenv("BZH_FUSED_KZ")) return std::max(1, std::atoi(env));
  int kz = 12;
  while (kz > 4 && tiles * ((planes + kz - 1) / kz) < 2048) kz = std::max(4, kz / 2);
  return kz;
}

// Ring length of the buffered DFT: 32 steps (BZH_DFT_SLOTS overrides it;
// 1 accumulates every step directly).  Never longer than the call.
int dft_slots(int nsteps) {
  int n = kDftMaxSlots;
  if (const char* env = std::getenv("BZH_DFT_SLOTS")) n = std::atoi(env);
  n = std::max(1, std::min(n, kDftMaxSlots));
  return std::min(n, std::max(nsteps, 1));
}

// Staging ring of the buffered DFT, one per (device, stream) so that calls on
// different streams never share it; calls on one stream are ordered by the
// stream.  Grown on demand and kept for the life of the process.
float* dft_stage(hipStream_t s, size_t floats) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, std::pair<float*, size_t>> rings;
  int device = 0;
  BZH_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(mu);
  auto& r = rings[{device, s}];
  if (r.second < floats) {
    if (r.first != nullptr) {
      // a graph that still references the old ring may be replayed later:
      // keep it alive (rings only grow, so this happens at most a few times)
      r.first = nullptr;
    }
    void* d = nullptr;
    BZH_HIP(hipMalloc(&d, floats * sizeof(float)));
    r = {static_cast<float*>(d), floats};
  }
  return r.first;
}

// Scratch of bzh_program_noalloc: a file-scope device array that every GPU
// the library runs on holds until the process ends.  In host code the name
// ``scratch`` is only a host-side stand-in; its device address (one per
// device) comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is handed to kernels.  Calls on one
// GPU share the array: calls that overlap in time on one GPU (different
// streams, or host threads) race on it.  The staging ring it holds is written
// before it is read within each call, so nothing needs clearing between calls.
static __device__ __align__(256) char scratch[BZH_NOALLOC_SCRATCH_BYTES];

char* noalloc_scratch() {
  static std::mutex mu;
  static std::map<int, char*> addresses;
  int device = 0;
  BZH_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(mu);
  auto it = addresses.find(device);
  if (it != addresses.end()) return it->second;
  void* p = nullptr;
  BZH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  BZH_REQUIRE(p != nullptr, "no device address for the no-allocation scratch");
  addresses[device] = static_cast<char*>(p);
  return static_cast<char*>(p);
}

// Band ordering of the fused workgroups (BZH_FUSED_ORDER=xcd selects the
// plain contiguous-range order for comparison).
bool fused_banded() {
  const char* env = std::getenv("BZH_FUSED_ORDER");
  return !(env && std::string(env) == "xcd");
}

FusedSplit fused_split(const Grid& g, int L) {
  const int TY = fused_rows();
  const int Tx = (g.nx + 1 + 61) / 62, Ty = (g.ny + 1 + TY - 1) / TY, planes = g.nz + 1;
  const int M = std::max(L, 1);
  int tx0 = (M + 1 + 61) / 62;                                  // 62 tx0 - 1 >= M
  int tx1 = (g.nx - M - 63) >= 0 ? (g.nx - M - 63) / 62 + 1 : 0;  // 62 (tx1-1) + 62 <= nx-M-1
  if (tx1 <= tx0) tx0 = tx1 = 0;
  FusedSplit sp{};
  sp.interior = {tx0, tx1 - tx0, tx1 - tx0, 0, Ty, 1, 0};
  sp.edge = {0, tx0 + (Tx - tx1), tx0, tx1 - tx0, Ty, 1, 0};
  for (FusedBox* b : {&sp.interior, &sp.edge}) {
    const int64_t tiles = static_cast<int64_t>(b->ntx) * b->nty;
    if (tiles == 0) continue;
    b->kz = chunk_len(tiles, planes);
    b->nzc = (planes + b->kz - 1) / b->kz;
  }
  return sp;
}

// Work split of the pair step (two steps per launch): tiles of 60 x 13
// owned cells whose lanes cover x = 60 t - 2 .. 60 t + 61; x-interior tiles
// have all of them in [M, nx - M - 1].  z chunks of 48 planes (32: 3.35 ms,
// 64: 3.32 ms against 3.29 ms per launch at 1024x512x256; BZH_PAIR_KZ
// overrides it); each chunk recomputes two planes of the first step.
//
// The x-interior tiles whose 16 rows all lie clear of the y faces and y CPML
// slabs, over the z range whose chunks' iterations all lie clear of the z
// faces and z slabs, form the core, run by the core variant of the kernel:
// lean iterations only (per-row byte offsets, no face, slab or ghost work,
// next plane prefetched by LDS DMA).  The rest, the shell, runs the general
// kernel: the x-edge tiles, the x-interior tiles of the outer y tiles, and
// the x-interior core-row tiles on the outer z planes.  (A core covering
// every plane, with the general iteration compiled in for the z-slab
// planes, measured 3.06 against 3.02 ms per pair at 1024x512x256: the
// general iteration's registers slow the lean one.)
// By default the core box joins the shell launch (its blocks last), so that
// the two overlap instead of each leaving a tail; BZH_PAIR_CORE=split gives
// the core its own launch and BZH_PAIR_CORE=0 runs it in the general kernel.
struct PairSplit {
  PairBox shell[kMaxPairBoxes];
  int nshell = 0;
  PairBox core{};
  bool has_core = false;
};

int pair_kz() {
  int kz = 48;
  if (const char* env = std::getenv("BZH_PAIR_KZ")) kz = std::max(1, std::atoi(env));
  return kz;
}

// Finish a box: z chunking over [z0, zsplit) and [zhi, z1) with chunks of at
// most kz planes, bands and block count; returns false for an empty box.
bool finish_box(PairBox& b, int kz, bool even) {
  const int64_t tiles = static_cast<int64_t>(b.ntx) * b.nty;
  auto chunks = [&](int n) {
    if (n <= 0) return std::pair<int, int>(0, kz);
    int c = (n + kz - 1) / kz;
    return std::pair<int, int>(c, even ? (n + c - 1) / c : kz);
  };
  const auto lo = chunks(b.zsplit - b.z0), hi = chunks(b.z1 - b.zhi);
  // both ranges share one chunk length
  b.kz = std::max(lo.first > 0 ? lo.second : 1, hi.first > 0 ? hi.second : 1);
  b.nzc_lo = b.zsplit > b.z0 ? (b.zsplit - b.z0 + b.kz - 1) / b.kz : 0;
  b.nzc = b.nzc_lo + (b.z1 > b.zhi ? (b.z1 - b.zhi + b.kz - 1) / b.kz : 0);
  if (tiles == 0 || b.nzc == 0) return false;
  b.band = fused_banded() ? static_cast<int>((tiles + 7) / 8) : 0;
  const int64_t n = (b.band > 0 ? 8 * int64_t{b.band} : tiles) * b.nzc;
  b.nblk = static_cast<int32_t>((n + 7) / 8 * 8);
  return true;
}

PairSplit pair_split(const Grid& g, int L) {
  const int W = kPairCols, TY = kPairRows;
  const int Tx = (g.nx + 1 + W - 1) / W, Ty = (g.ny + 1 + TY - 1) / TY, planes = g.nz + 1;
  const int M = std::max(L, 1);
  int tx0 = (M + 2 + W - 1) / W;                           // W tx0 - 2 >= M
  int tx1 = g.nx - M - 2 >= 0 ? std::min((g.nx - M - 2) / W, Tx) : 0;  // W tx1 <= nx - M - 2
  if (tx1 <= tx0) tx0 = tx1 = 0;
  const int kz = pair_kz();
  // core rows: 13 by - 2 >= M and 13 by + 13 <= ny - L - 1
  const int cy0 = (M + 2 + TY - 1) / TY;
  const int cy1 = g.ny - L - 1 - TY >= 0 ? (g.ny - L - 1 - TY) / TY + 1 : 0;
  // core planes: chunks [kz0, kz1) with kz0 - 4 >= M and kz1 + 2 <= nz - M - 1
  const int cz0 = M + 4, cz1 = g.nz - M - 3;
  const char* env = std::getenv("BZH_PAIR_CORE");
  const bool core = !(env && std::string(env) == "0") && tx1 > tx0 && cy1 > cy0 && cz1 > cz0;
  PairSplit sp;
  auto box = [&](int bx0, int ntx, int tsplit, int tskip, int by0, int nty, int tysplit, int tyskip,
                 int z0, int zsplit, int zhi, int z1, int xedge) {
    PairBox b{};
    b.tx0 = bx0;
    b.ntx = ntx;
    b.tsplit = tsplit;
    b.tskip = tskip;
    b.ty0 = by0;
    b.nty = nty;
    b.tysplit = tysplit;
    b.tyskip = tyskip;
    b.z0 = z0;
    b.zsplit = zsplit;
    b.zhi = zhi;
    b.z1 = z1;
    b.xedge = xedge;
    return b;
  };
  // the x-edge tiles (slower: first, so that they do not form the tail)
  PairBox e = box(0, tx0 + (Tx - tx1), tx0, tx1 - tx0, 0, Ty, Ty, 0, 0, planes, planes, planes, 1);
  if (finish_box(e, kz, false)) sp.shell[sp.nshell++] = e;
  if (!core) {
    PairBox in = box(tx0, tx1 - tx0, tx1 - tx0, 0, 0, Ty, Ty, 0, 0, planes, planes, planes, 0);
    if (finish_box(in, kz, false)) sp.shell[sp.nshell++] = in;
    return sp;
  }
  // x-interior tiles outside the core rows, every plane
  PairBox out = box(tx0, tx1 - tx0, tx1 - tx0, 0, 0, Ty - (cy1 - cy0), cy0, cy1 - cy0, 0, planes,
                    planes, planes, 0);
  if (finish_box(out, kz, false)) sp.shell[sp.nshell++] = out;
  // x-interior core-row tiles on the planes outside the core range
  PairBox rim = box(tx0, tx1 - tx0, tx1 - tx0, 0, cy0, cy1 - cy0, cy1 - cy0, 0, 0, cz0, cz1, planes, 0);
  if (finish_box(rim, kz, false)) sp.shell[sp.nshell++] = rim;
  // the core, in chunks of equal length
  sp.core = box(tx0, tx1 - tx0, tx1 - tx0, 0, cy0, cy1 - cy0, cy1 - cy0, 0, cz0, cz1, cz1, cz1, 2);
  sp.has_core = finish_box(sp.core, kz, true);
  if (sp.has_core && !(env && std::string(env) == "split")) {
    sp.shell[sp.nshell++] = sp.core;
    sp.has_core = false;
  }
  return sp;
}

// Rounding of bfloat16 field stores in the fused kernels: stochastic by
// default (BZH_BF16_ROUNDING=nearest selects round-to-nearest-even).
bool bf16_stochastic_rounding() {
  const char* env = std::getenv("BZH_BF16_ROUNDING");
  return !(env && std::string(env) == "nearest");
}

// Two steps per launch for every storage type (BZH_PAIR=0 disables it).
// With the core kernel, 2-byte fields run faster in pairs than in single
// steps (f16 1.52 against 1.83 ms per step at 1024x512x256).
bool pairs_enabled(int /*dtype*/) {
  const char* env = std::getenv("BZH_PAIR");
  return !(env && std::string(env) == "0");
}

// Intermediate bank and monitor boxes of the pair step, one per (device,
// stream) like the DFT ring: calls on one stream are ordered by it.  The bank
// grows on demand; a replaced bank is freed once the stream has drained.
struct PairWorkspace {
  char* bank = nullptr;
  size_t bytes = 0;
  int32_t* boxes = nullptr;
  float* cap = nullptr;            // dense capture of the monitored cells
  size_t cap_bytes = 0;
  int32_t* capbox = nullptr;       // its box table, int32[kMaxMonBoxes][6][8]
  std::vector<int32_t> capbox_host;  // what capbox holds
};
// This stream's workspace (one per device and stream, kept until exit).
PairWorkspace& pair_space(hipStream_t s) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, PairWorkspace> spaces;
  int device = 0;
  BZH_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(mu);
  return spaces[{device, s}];
}

PairWorkspace pair_workspace(hipStream_t s, size_t bank_bytes) {
  PairWorkspace& w = pair_space(s);
  if (w.boxes == nullptr) {
    void* d = nullptr;
    BZH_HIP(hipMalloc(&d, kMaxMonBoxes * 6 * sizeof(int32_t)));
    w.boxes = static_cast<int32_t*>(d);
  }
  if (w.bytes < bank_bytes) {
    if (w.bank != nullptr) {
      BZH_HIP(hipStreamSynchronize(s));
      BZH_HIP(hipFree(w.bank));
      w.bank = nullptr;
      w.bytes = 0;
    }
    void* d = nullptr;
    BZH_HIP(hipMalloc(&d, bank_bytes));
    w.bank = static_cast<char*>(d);
    w.bytes = bank_bytes;
  }
  return w;
}

// The capture of the monitored cells for pair launches (PairParams::cap),
// planned from the host boxes of the monitor layouts' trailing argument.
// ``ok`` is false (and the intermediate bank is used) without the boxes, or
// when a tile row would hold cells of two monitors: monitors whose y ranges
// overlap must lie in disjoint 60-column tile columns.
struct CapturePlan {
  bool ok = false;
  std::vector<int32_t> table;  // [M][6][8]
  size_t slot = 0;             // floats per step
};
CapturePlan capture_plan(const ProgramIn& in, const DftParams& probe) {
  CapturePlan cp;
  if (in.mon_boxes == nullptr || in.noalloc || probe.M < 1 || probe.M > kMaxMonBoxes) return cp;
  const int M = probe.M;
  const int32_t* hb = static_cast<const int32_t*>(in.mon_boxes->data);
  cp.table.assign(static_cast<size_t>(M) * 48, 0);
  std::vector<int> uy0(M, INT_MAX), uy1(M, INT_MIN), ux0(M, INT_MAX), ux1(M, INT_MIN);
  size_t off = 0;
  for (int m = 0; m < M; ++m) {
    for (int c = 0; c < 6; ++c) {
      const int32_t* b = hb + (m * 6 + c) * 6;
      int32_t* t = &cp.table[static_cast<size_t>(m * 6 + c) * 8];
      t[0] = 1;  // empty: z0 > z1
      t[1] = 0;
      if (b[0] < 0 || probe.field[c] == nullptr) continue;
      for (int d = 0; d < 3; ++d)
        BZH_REQUIRE(b[2 * d] >= 0 && b[2 * d] <= b[2 * d + 1] && b[2 * d + 1] < probe.fdims[c][d],
                    "monitor box " + std::to_string(m) + " lies outside component " +
                        std::to_string(c));
      for (int q = 0; q < 6; ++q) t[q] = b[q];
      t[6] = static_cast<int32_t>(off);
      off += static_cast<size_t>(b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
      BZH_REQUIRE(off < (size_t{1} << 31), "the monitors' capture exceeds 2^31 cells");
      uy0[m] = std::min(uy0[m], b[2]);
      uy1[m] = std::max(uy1[m], b[3]);
      ux0[m] = std::min(ux0[m], b[4]);
      ux1[m] = std::max(ux1[m], b[5]);
    }
  }
  for (int a = 0; a < M; ++a)
    for (int b = a + 1; b < M; ++b) {
      if (uy0[a] > uy1[a] || uy0[b] > uy1[b]) continue;
      const bool yo = uy0[a] <= uy1[b] && uy0[b] <= uy1[a];
      const bool xo = ux0[a] / kPairCols <= ux1[b] / kPairCols && ux0[b] / kPairCols <= ux1[a] / kPairCols;
      if (yo && xo) return cp;
    }
  cp.slot = off;
  cp.ok = off > 0;
  return cp;
}

// The capture buffer and box table of a plan, on this stream's
// workspace; the table is uploaded when it changes.
void capture_workspace(hipStream_t s, PairWorkspace& w, const CapturePlan& cp) {
  const size_t bytes = cp.slot * sizeof(float);
  if (w.cap_bytes < bytes) {
    if (w.cap != nullptr) {
      BZH_HIP(hipStreamSynchronize(s));
      BZH_HIP(hipFree(w.cap));
      w.cap = nullptr;
      w.cap_bytes = 0;
    }
    void* d = nullptr;
    BZH_HIP(hipMalloc(&d, bytes));
    w.cap = static_cast<float*>(d);
    w.cap_bytes = bytes;
  }
  if (w.capbox == nullptr) {
    void* d = nullptr;
    BZH_HIP(hipMalloc(&d, kMaxMonBoxes * 48 * sizeof(int32_t)));
    w.capbox = static_cast<int32_t*>(d);
  }
  if (w.capbox_host != cp.table) {
    // (rare: a new program) the launches of earlier calls may still read it
    BZH_HIP(hipStreamSynchronize(s));
    BZH_HIP(hipMemcpy(w.capbox, cp.table.data(), cp.table.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
    w.capbox_host = cp.table;
  }
}

// Slab count of a source group (0 when absent).
int group_slabs(const SourceGroupIn& grp) {
  const bzh_buffer* c = grp.coeffs;
  return (c && c->data && c->rank >= 1) ? static_cast<int>(c->dims[0]) : 0;
}

// Append the slabs of a source group (E components, or H with ``hfield``) to
// an in-kernel slab list of at most ``cap`` entries.
void add_mid_slabs(MidSlab* list, int32_t& count, int cap, const SourceGroupIn& grp, int comp,
                   int kstep, bool post, bool hfield, int metallic) {
  const int n = group_slabs(grp);
  if (n == 0) return;
  const bzh_buffer& c = *grp.coeffs;
  BZH_REQUIRE(c.element_type == BZH_F32 && c.rank == 4 && grp.waves->element_type == BZH_F32 &&
                  grp.waves->rank == 2 && grp.waves->dims[0] == n && grp.waves->dims[1] >= 1 &&
                  grp.starts->element_type == BZH_S32 && grp.starts->rank == 2 &&
                  grp.starts->dims[0] == n && grp.starts->dims[1] == 3,
              "source groups are f32 [n, sz, sy, sx] coefficients, f32 [n, T] waveforms and s32 [n, 3] starts");
  BZH_REQUIRE(count + n <= cap, "too many source slabs for the fused step");
  const int64_t cells = c.dims[1] * c.dims[2] * c.dims[3];
  const int T = static_cast<int>(grp.waves->dims[1]);
  for (int s = 0; s < n; ++s) {
    MidSlab m;
    std::memset(&m, 0, sizeof(m));
    m.coeff = static_cast<const float*>(c.data) + s * cells;
    m.wave = static_cast<const float*>(grp.waves->data) + static_cast<int64_t>(s) * T;
    m.start = static_cast<const int32_t*>(grp.starts->data) + 3 * s;
    m.sz = static_cast<int>(c.dims[1]);
    m.sy = static_cast<int>(c.dims[2]);
    m.sx = static_cast<int>(c.dims[3]);
    m.T = T;
    m.hfield = hfield ? 1 : 0;
    m.comp = comp;
    m.kstep = kstep;
    m.post = post ? 1 : 0;
    if (post) mask_ends((hfield ? kHx : kEx) + comp, metallic, true, m.mlo, m.mhi);
    list[count++] = m;
  }
}

size_t elem_bytes(int t) { return t == BZH_BF16 || t == BZH_F16 ? 2 : (t == BZH_F64 ? 8 : 4); }

// The CPML term extents of a phase pair from its device metadata rows
// (component, axis, low, high, sign; runtime.py:56-69): ext[2T], ext[2T + 1]
// are term T's low and high extents, H terms 0..5 then E terms 6..11 in the
// kernels' order.  The rows are copied on the caller's stream (so that a
// metadata write queued before this call lands first) and the copy is waited
// for.  Not while the caller's stream captures a graph (no copy may run
// then): false, and the caller takes the per-phase schedule.
bool cpml_extents(hipStream_t s, const PhaseIn& h, const PhaseIn& e, int32_t ext[24]) {
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (s != nullptr) BZH_HIP(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) return false;
  for (const PhaseIn* ph : {&h, &e})
    if (ph->meta == nullptr || ph->meta->data == nullptr || ph->meta->element_type != BZH_S32 ||
        numel(*ph->meta) != 30)
      return false;
  int32_t rows[2][30];
  BZH_HIP(hipMemcpyAsync(rows[0], h.meta->data, sizeof(rows[0]), hipMemcpyDeviceToHost, s));
  BZH_HIP(hipMemcpyAsync(rows[1], e.meta->data, sizeof(rows[1]), hipMemcpyDeviceToHost, s));
  BZH_HIP(hipStreamSynchronize(s));
  for (int T = 0; T < 12; ++T) {
    const int32_t* r = rows[T / 6] + (T % 6) * 5;
    if (r[0] != (T % 6) / 2 || r[1] != kAxes3d[T % 6]) return false;  // not the kernels' order
    ext[2 * T] = r[2];
    ext[2 * T + 1] = r[3];
  }
  return true;
}

// Can the fused kernel take this phase pair?  Fills the CPML part of ``fp``.
// Term extents: ``uniform_l`` on every face when positive, else ``ext`` (from
// cpml_extents; null: unknown, so no).  The terms on one axis must share their
// extents within a phase, the 36 coefficient arrays must be the rows of one
// [12][3][st] table, and the CPML memories of both banks one arena.
bool fused_cpml_layout(const PhaseIn& h, const PhaseIn& e, const bzh_buffer* const* psi_b,
                       int uniform_l, const int32_t* ext, FusedParams& fp) {
  fp.L = 0;
  fp.tab = nullptr;
  fp.tab_st = 0;
  for (int i = 0; i < 6; ++i) fp.clo[i] = fp.chi[i] = 0;
  for (int T = 0; T < 12; ++T) {
    fp.psi[T] = nullptr;
    fp.poff[T] = 0u;
  }
  if (h.nterms == 0 && e.nterms == 0) return true;
  if (h.nterms != 6 || e.nterms != 6) return false;
  if (uniform_l <= 0 && ext == nullptr) return false;
  int n2[12];
  bool seen[6] = {};
  int T0 = -1, L = 0;
  for (int T = 0; T < 12; ++T) {
    const int lo = uniform_l > 0 ? uniform_l : ext[2 * T];
    const int hi = uniform_l > 0 ? uniform_l : ext[2 * T + 1];
    if (lo < 0 || hi < 0) return false;
    const int slot = (T < 6 ? kSlabH : kSlabE) + (2 - kAxes3d[T % 6]);  // axis 0 x, 1 y, 2 z
    if (!seen[slot]) {
      seen[slot] = true;
      fp.clo[slot] = lo;
      fp.chi[slot] = hi;
    } else if (fp.clo[slot] != lo || fp.chi[slot] != hi) {
      return false;
    }
    n2[T] = lo + hi;
    L = std::max(L, std::max(lo, hi));
    if (n2[T] > 0 && T0 < 0) T0 = T;
  }
  if (T0 < 0) return true;  // no term has a slab: no CPML work
  // the coefficient table: term T0's a and b rows give the row stride
  const PhaseIn& p0 = T0 < 6 ? h : e;
  const char* a0 = static_cast<const char*>(p0.terms[(T0 % 6) * 3 + 0]->data);
  const char* b0 = static_cast<const char*>(p0.terms[(T0 % 6) * 3 + 1]->data);
  if (a0 == nullptr || b0 <= a0 || (b0 - a0) % 4 != 0) return false;
  const int64_t st = (b0 - a0) / 4;
  const char* base = a0 - static_cast<int64_t>(T0) * 3 * st * 4;
  for (int T = 0; T < 12; ++T) {
    if (n2[T] == 0) continue;
    if (n2[T] > st) return false;
    const PhaseIn& ph = T < 6 ? h : e;
    for (int q = 0; q < 3; ++q) {
      const bzh_buffer* b = ph.terms[(T % 6) * 3 + q];
      if (b == nullptr || b->element_type != BZH_F32 || numel(*b) != n2[T]) return false;
      if (static_cast<const char*>(b->data) != base + ((T * 3 + q) * st) * 4) return false;
    }
  }
  // CPML memories: one per term with a slab, in both banks at one distance
  int64_t delta = 0;
  int ptype = -1;
  for (int T = 0; T < 12; ++T) {
    if (n2[T] == 0) continue;
    const bzh_buffer* a = T < 6 ? h.psi_in[T] : e.psi_in[T - 6];
    const bzh_buffer* b = psi_b[T];
    if (a == nullptr || b == nullptr || a->data == nullptr || b->data == nullptr) return false;
    if (numel(*a) != numel(*b) || a->element_type != b->element_type) return false;
    if (a->rank != 3 || a->dims[kAxes3d[T % 6]] != n2[T]) return false;
    if (ptype < 0) ptype = a->element_type;
    if (a->element_type != ptype) return false;
    const int64_t diff = static_cast<const char*>(b->data) - static_cast<const char*>(a->data);
    const int64_t sz = static_cast<int64_t>(elem_bytes(ptype));
    if (diff % sz != 0) return false;
    if (T == T0) delta = diff / sz;
    else if (diff / sz != delta) return false;
    fp.psi[T] = a->data;
  }
  // based addressing: all the memories of both banks within 4 GiB above the lowest
  const int64_t sz = static_cast<int64_t>(elem_bytes(ptype));
  uintptr_t lo = UINTPTR_MAX, hi = 0;
  for (int T = 0; T < 12; ++T) {
    if (n2[T] == 0) continue;
    const bzh_buffer* a = T < 6 ? h.psi_in[T] : e.psi_in[T - 6];
    for (const bzh_buffer* x : {a, psi_b[T]}) {
      const uintptr_t p = reinterpret_cast<uintptr_t>(x->data);
      lo = std::min(lo, p);
      hi = std::max(hi, p + static_cast<uintptr_t>(numel(*x) * sz));
    }
  }
  if (hi - lo >= (uintptr_t{1} << 32)) return false;
  fp.pbase = reinterpret_cast<const char*>(lo);
  for (int T = 0; T < 12; ++T)
    if (fp.psi[T] != nullptr)
      fp.poff[T] = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(fp.psi[T]) - lo);
  fp.pdelta = static_cast<uint32_t>(static_cast<uint64_t>(delta * sz));
  fp.L = L;
  fp.tab = reinterpret_cast<const float*>(base);
  fp.tab_st = static_cast<int32_t>(st);
  fp.psi_delta = delta;
  fp.psi_bf16 = ptype == BZH_BF16;
  return true;
}

// One scalar base for a group of arrays, the lowest address, and each array's
// byte offset from it.  ``wide`` reports a group that spans 4 GiB or more:
// the kernels then form 64-bit offsets (their wide variants).  False for a
// null array or an array of 4 GiB or more (the kernels index within an array
// with 32-bit byte offsets).
bool group_base(const bzh_buffer* const* bufs, int n, const char*& base, uint64_t* off,
                bool& wide) {
  constexpr uintptr_t k4G = uintptr_t{1} << 32;
  uintptr_t lo = UINTPTR_MAX, hi = 0;
  for (int i = 0; i < n; ++i) {
    if (bufs[i] == nullptr || bufs[i]->data == nullptr) return false;
    const uintptr_t p = reinterpret_cast<uintptr_t>(bufs[i]->data);
    const uintptr_t bytes = static_cast<uintptr_t>(numel(*bufs[i]) * elem_bytes(bufs[i]->element_type));
    if (bytes >= k4G) return false;
    lo = std::min(lo, p);
    hi = std::max(hi, p + bytes);
  }
  if (n == 0) return false;
  wide = hi - lo >= k4G;
  base = reinterpret_cast<const char*>(lo);
  for (int i = 0; i < n; ++i) off[i] = reinterpret_cast<uintptr_t>(bufs[i]->data) - lo;
  return true;
}

// The fields of a bank in the kernel's order (Ex, Ey, Ez, Hx, Hy, Hz).
bool bank_base(const bzh_buffer* const* bank, const char*& base, uint64_t off[6], bool& wide) {
  const bzh_buffer* order[6] = {bank[kEx], bank[kEy], bank[kEz], bank[kHx], bank[kHy], bank[kHz]};
  return group_base(order, 6, base, off, wide);
}

// The wide (64-bit offset) kernels for every fused launch, also where 32-bit
// offsets would do (BZH_WIDE=1; a test switch for the wide variants).
bool force_wide() {
  const char* env = std::getenv("BZH_WIDE");
  return env && std::string(env) == "1";
}

bool metrics_empty(const PhaseIn& ph) {
  for (int a = 0; a < 3; ++a)
    if (ph.metric[a] && ph.metric[a]->data && numel(*ph.metric[a]) > 0) return false;
  return true;
}

// Returns true when the fused two-bank kernel carried the whole schedule (psi
// then ping-pongs with the fields); false for the per-phase schedule.
bool run_program(hipStream_t s, ProgramIn& in) {
  BZH_REQUIRE(in.nsteps >= 1, "step count must be positive");
  Grid g = infer_grid(in.bank_a);
  const bool two_banks = in.bank_b[0] != nullptr;
  if (two_banks) {
    const Grid gb = infer_grid(in.bank_b);
    BZH_REQUIRE(gb.nz == g.nz && gb.ny == g.ny && gb.nx == g.nx && gb.dtype == g.dtype,
                "workspace bank does not match the field bank");
  }
  // H-timing slabs run inside the fused kernels (up to kMaxHSlabs per step)
  int h_slabs = 0;
  for (int c = 0; c < 3 && 3 + c < in.ngroups; ++c) h_slabs += group_slabs(in.groups[3 + c]);
  // the fused kernel takes scalar H coefficients and E decay, E source
  // coefficients that are all arrays or all scalars, and arrays addressable
  // with 32-bit element offsets; anything else runs the per-phase schedule
  bool coef_ok = true;
  for (int q = 0; q < 6; ++q)
    if (in.h.mat[q] == nullptr || numel(*in.h.mat[q]) != 1) coef_ok = false;
  for (int q = 0; q < 3; ++q)
    if (in.e.mat[q] == nullptr || numel(*in.e.mat[q]) != 1) coef_ok = false;
  int n_e_arrays = 0;
  for (int q = 3; q < 6; ++q)
    if (in.e.mat[q] != nullptr && numel(*in.e.mat[q]) != 1) ++n_e_arrays;
  // packed coefficient IDs run the per-phase kernels (the reference's
  // streamed kernel reads them)
  for (int q = 0; q < 3; ++q)
    if (in.e.mat[3 + q] != nullptr && in.e.mat[3 + q]->element_type == BZH_S32) coef_ok = false;
  coef_ok = coef_ok && (n_e_arrays == 0 || n_e_arrays == 3);
  // based addressing: each bank's fields at offsets from one base, the same
  // offsets in both banks; the E coefficient arrays likewise.  A group that
  // spans 4 GiB or more takes the wide kernels (64-bit offsets)
  FusedParams cpml_probe;
  std::memset(&cpml_probe, 0, sizeof(cpml_probe));
  bool based = g.is3d && two_banks;
  uint64_t foff_a[6] = {}, foff_b[6] = {};
  bool wide_a = false, wide_b = false, wide_c = false;
  const char* base_a = nullptr;
  const char* base_b = nullptr;
  if (based) {
    based = bank_base(in.bank_a, base_a, foff_a, wide_a) &&
            bank_base(in.bank_b, base_b, foff_b, wide_b) &&
            std::memcmp(foff_a, foff_b, sizeof(foff_a)) == 0;
  }
  if (based && n_e_arrays == 3) {
    const bzh_buffer* src[3] = {in.e.mat[3], in.e.mat[4], in.e.mat[5]};
    based = group_base(src, 3, cpml_probe.cbase, cpml_probe.coff64, wide_c);
  }
  cpml_probe.wide = (wide_a || wide_b || wide_c || force_wide()) ? 1 : 0;
  for (int q = 0; q < 6; ++q) {
    cpml_probe.foff64[q] = foff_a[q];
    cpml_probe.foff[q] = static_cast<uint32_t>(foff_a[q]);  // (used when not wide)
  }
  for (int q = 0; q < 3; ++q) cpml_probe.coff[q] = static_cast<uint32_t>(cpml_probe.coff64[q]);
  // CPML extents other than one thickness on every face: from the term rows
  int32_t cpml_ext_rows[24];
  const int32_t* cpml_ext = nullptr;
  if (based && in.uniform_l <= 0 && in.h.nterms == 6 && in.e.nterms == 6 && !in.noalloc &&
      (in.flags & BZH_FLAG_FUSED_STEP) && cpml_extents(s, in.h, in.e, cpml_ext_rows))
    cpml_ext = cpml_ext_rows;
  const bool fused = two_banks && (in.flags & BZH_FLAG_FUSED_STEP) && g.is3d &&
                     h_slabs <= kMaxHSlabs && (h_slabs == 0 || in.step != nullptr) &&
                     coef_ok && based && metrics_empty(in.h) && metrics_empty(in.e) &&
                     fused_cpml_layout(in.h, in.e, in.psi_b, in.uniform_l, cpml_ext, cpml_probe) &&
                     // the wide kernels: float32 storage and 15-row single-step tiles
                     (!cpml_probe.wide ||
                      (g.dtype == BZH_F32 && !cpml_probe.psi_bf16 && fused_rows() == 15));
  const int32_t* step0 = in.step ? static_cast<const int32_t*>(in.step->data) : nullptr;
  int slots = 1;
  float* stage = nullptr;
  if (in.has_monitors) {
    const DftParams probe = make_dft(in.mon, in.bank_a, g, in.dt, in.monitor_count);
    slots = dft_slots(in.nsteps);
    const size_t per_slot = static_cast<size_t>(probe.M) * 6 * probe.Pmax;
    if (in.noalloc) {
      // the ring lives in the fixed scratch: as many slots as fit
      const size_t cap = BZH_NOALLOC_SCRATCH_BYTES / sizeof(float);
      BZH_REQUIRE(per_slot <= cap,
                  "the monitors' samples of one step (" + std::to_string(per_slot * sizeof(float)) +
                      " bytes) exceed the no-allocation scratch (" +
                      std::to_string(BZH_NOALLOC_SCRATCH_BYTES) + " bytes)");
      slots = std::min<int>(slots, static_cast<int>(cap / per_slot));
      if (slots > 1) stage = reinterpret_cast<float*>(noalloc_scratch());
    } else if (slots > 1) {
      stage = dft_stage(s, static_cast<size_t>(slots) * per_slot);
    }
  }
  // Two steps per launch (the pair kernel) where the fused step runs, the
  // source slabs between the steps fit the kernel and the monitors its boxes.
  // Each pair swaps the banks once, so an even number of pairs per call keeps
  // the single-step bank parity (result in the workspace after an odd count).
  int mid_slabs = 2 * h_slabs;  // the H slabs of both steps
  for (int c = 0; c < 3; ++c) {
    if (c < in.ngroups) mid_slabs += group_slabs(in.groups[c]);
    if (6 + c < in.ngroups) mid_slabs += group_slabs(in.groups[6 + c]);
  }
  int npairs = 0;
  int mon_m = 0;
  if (in.has_monitors) mon_m = make_dft(in.mon, in.bank_a, g, in.dt, in.monitor_count).M;
  // (the no-allocation entry runs pairs without monitors only: their
  // intermediate sample needs the workspace bank)
  if (fused && pairs_enabled(g.dtype) && mid_slabs <= kMaxMidSlabs && mon_m <= kMaxMonBoxes &&
      (mid_slabs == 0 || step0 != nullptr) && !(in.noalloc && mon_m > 0)) {
    npairs = in.nsteps / 2;
    if (npairs % 2 == 1) --npairs;
  }
  // the monitors' samples of the pairs: a dense capture of the cells they
  // gather (with the host boxes), else the intermediate bank (same layout as
  // bank A) and the monitor boxes found on the device
  const bzh_buffer* cbank[6] = {};
  bzh_buffer cbufs[6];
  PairWorkspace pws;
  CapturePlan cap;
  if (npairs > 0 && mon_m > 0)
    cap = capture_plan(in, make_dft(in.mon, in.bank_a, g, in.dt, in.monitor_count));
  if (npairs > 0 && mon_m > 0 && cap.ok) {
    PairWorkspace& w = pair_space(s);
    capture_workspace(s, w, cap);
    pws = w;
  } else if (npairs > 0 && mon_m > 0) {
    size_t span = 0;
    const int order[6] = {kEx, kEy, kEz, kHx, kHy, kHz};
    for (int q = 0; q < 6; ++q) {
      const bzh_buffer& b = *in.bank_a[order[q]];
      span = std::max(span, static_cast<size_t>(cpml_probe.foff64[q]) +
                                static_cast<size_t>(numel(b)) * elem_bytes(b.element_type));
    }
    pws = pair_workspace(s, span);
    for (int q = 0; q < 6; ++q) {
      cbufs[order[q]] = *in.bank_a[order[q]];
      cbufs[order[q]].data = pws.bank + cpml_probe.foff64[q];
    }
    for (int q = 0; q < 6; ++q) cbank[q] = &cbufs[q];
  }

  // The fused-step parameters for a launch that reads bank B when cur_b.
  auto fused_params = [&](const bzh_buffer* const* cur, const bzh_buffer* const* nxt, bool cur_b) {
    PhaseIn h = in.h, e = in.e;
    for (int c = 0; c < 3; ++c) {
      h.tin[c] = cur[kHx + c];
      h.tout[c] = nxt[kHx + c];
      h.src[c] = cur[kEx + c];
      e.tin[c] = cur[kEx + c];
      e.tout[c] = nxt[kEx + c];
      e.src[c] = nxt[kHx + c];
    }
    // validate both phases through the generic decoder, then repack
    const PhaseParams hp = make_phase(h, BZH_PHASE_H, g, in.dt, in.resolution, in.metallic, in.uniform_l);
    const PhaseParams ep = make_phase(e, BZH_PHASE_E, g, in.dt, in.resolution, in.metallic, in.uniform_l);
    FusedParams fp = cpml_probe;
    for (int c = 0; c < 3; ++c) {
      fp.h_in[c] = hp.tin[c];
      fp.h_out[c] = hp.tout[c];
      fp.e_in[c] = ep.tin[c];
      fp.e_out[c] = ep.tout[c];
      fp.hdec[c] = hp.dec[c];
      fp.hsrc[c] = hp.scf[c];
      fp.edec[c] = ep.dec[c];
      fp.esrc[c] = ep.scf[c];
      fp.hdec_st[c] = hp.dec_st[c];
      fp.hsrc_st[c] = hp.scf_st[c];
      fp.edec_st[c] = ep.dec_st[c];
      fp.esrc_st[c] = ep.scf_st[c];
    }
    fp.coef_f16 = ep.scf_f16;
    fp.inv_dx = hp.inv_dx;
    fp.ratio_h = 1.0f;
    fp.ratio_e = 1.0f;
    fp.nz = g.nz;
    fp.ny = g.ny;
    fp.nx = g.nx;
    fp.metallic = in.metallic;
    if (fp.L > 0 && cur_b) {
      // bank B is read: psi roles swap
      for (int t = 0; t < 12; ++t) {
        if (fp.psi[t] == nullptr) continue;  // a term without a slab
        fp.psi[t] = static_cast<char*>(fp.psi[t]) + fp.psi_delta * (fp.psi_bf16 ? 2 : 4);
        fp.poff[t] += fp.pdelta;
      }
      fp.psi_delta = -fp.psi_delta;
      fp.pdelta = 0u - fp.pdelta;
    }
    fp.fin = cur_b ? base_b : base_a;
    fp.fout = cur_b ? base_a : base_b;
    fp.box_ty0 = 0;
    fp.box_z0 = 0;
    fp.box_z1 = g.nz + 1;
    fp.step0 = step0;
    fp.bf16_sr = g.dtype == BZH_BF16 && bf16_stochastic_rounding() ? 1 : 0;
    return fp;
  };
  auto set_box = [&](FusedParams& fp, const FusedBox& b, int v) {
    fp.xedge = v;
    fp.box_tx0 = b.tx0;
    fp.box_ntx = b.ntx;
    fp.box_tsplit = b.tsplit;
    fp.box_tskip = b.tskip;
    fp.box_nty = b.nty;
    fp.box_kz = b.kz;
    fp.box_nzc = b.nzc;
    fp.box_band = fused_banded() ? (b.ntx * b.nty + 7) / 8 : 0;
  };

  std::vector<Launch> launches;
  launches.reserve(static_cast<size_t>(in.nsteps) * 8);
  // the DFT of step k, read from ``bank`` (buffered: stage, and fold the ring
  // every ``slots`` steps and at the end of the call)
  auto add_dft = [&](int k, const bzh_buffer* const* bank) {
    DftParams dp = make_dft(in.mon, bank, g, in.dt, in.monitor_count);
    dp.k = k;
    if (slots > 1) {
      dp.mode = 1;
      dp.slots = slots;
      dp.stage = stage;
      launches.push_back(make_launch(Kind::kDft, 0, g.dtype, dp));
      if ((k + 1) % slots == 0 || k == in.nsteps - 1) {
        dp.mode = 2;
        dp.nk = (k % slots) + 1;
        dp.k0 = k + 1 - dp.nk;
        launches.push_back(make_launch(Kind::kDft, 0, g.dtype, dp));
      }
    } else {
      launches.push_back(make_launch(Kind::kDft, 0, g.dtype, dp));
    }
  };
  // the DFT of the pair's first step, from the capture
  auto add_dft_cap = [&](int k) {
    const size_t before = launches.size();
    add_dft(k, in.bank_a);
    for (size_t i = before; i < launches.size(); ++i) {
      DftParams dp;
      std::memcpy(&dp, launches[i].bytes.data(), sizeof(dp));
      dp.cap = pws.cap;
      dp.capbox = pws.capbox;
      launches[i] = make_launch(Kind::kDft, 0, g.dtype, dp);
    }
  };
  if (npairs > 0 && mon_m > 0 && !cap.ok) {
    const DftParams probe = make_dft(in.mon, in.bank_a, g, in.dt, in.monitor_count);
    MonBoxParams mb;
    std::memset(&mb, 0, sizeof(mb));
    mb.indices = probe.indices;
    mb.weights = probe.weights;
    mb.M = probe.M;
    mb.Pmax = probe.Pmax;
    mb.K = probe.K;
    const int order[6] = {kEx, kEy, kEz, kHx, kHy, kHz};
    for (int q = 0; q < 6; ++q)
      for (int d = 0; d < 3; ++d) mb.dims[q][d] = static_cast<int>(in.bank_a[order[q]]->dims[d]);
    mb.box = pws.boxes;
    launches.push_back(make_launch(Kind::kMonBox, 0, g.dtype, mb));
  }

  const Kind pk = g.is3d ? Kind::kPhase3d : Kind::kPhaseTm;
  // 2D float32 programs fold their timed source slabs into the phase launches
  // (three launches per step instead of five: H, E, the DFT sample); the
  // groups that do not fit keep their source launches
  bool fold_tm = !g.is3d && g.dtype == BZH_F32 && step0 != nullptr;
  for (int c = 0; c < 3 && fold_tm; ++c) {
    if (3 + c < in.ngroups && group_slabs(in.groups[3 + c]) > 0 && c == 2) fold_tm = false;  // (no Hz in TMxy)
    if (6 + c < in.ngroups && group_slabs(in.groups[6 + c]) > 0 && c < 2) fold_tm = false;  // (no Ex, Ey)
  }
  bool cur_b = false;  // the current fields live in bank B
  int pairs_left = npairs;
  for (int k = 0; k < in.nsteps;) {
    const bzh_buffer* const* cur = cur_b ? in.bank_b : in.bank_a;
    const bzh_buffer* const* nxt = two_banks ? (cur_b ? in.bank_a : in.bank_b) : cur;
    // pre_e sources act on the E bank read by this step
    for (int c = 0; c < 3 && c < in.ngroups; ++c)
      add_source_launches(launches, in.groups[c], 0, c, cur, g, in.metallic, step0, k);
    if (pairs_left > 0) {
      // steps k and k + 1 in one launch: E sources of step k and pre-E
      // sources of step k + 1 are applied inside, between the steps
      PairParams pp;
      std::memset(&pp, 0, sizeof(pp));
      pp.f = fused_params(cur, nxt, cur_b);
      pp.f.rows = kPairRows;
      pp.f.kstep = k;
      pp.step0 = step0;
      for (int c = 0; c < 3 && 6 + c < in.ngroups; ++c)
        add_mid_slabs(pp.msrc, pp.msrc_n, kMaxMidSlabs, in.groups[6 + c], c, k, true, false, in.metallic);
      pp.msrc_post = pp.msrc_n;
      for (int c = 0; c < 3 && c < in.ngroups; ++c)
        add_mid_slabs(pp.msrc, pp.msrc_n, kMaxMidSlabs, in.groups[c], c, k + 1, false, false, in.metallic);
      pp.msrc_e = pp.msrc_n;
      for (int c = 0; c < 3 && 3 + c < in.ngroups; ++c)
        add_mid_slabs(pp.msrc, pp.msrc_n, kMaxMidSlabs, in.groups[3 + c], c, k, true, true, in.metallic);
      pp.msrc_h1 = pp.msrc_n;
      for (int c = 0; c < 3 && 3 + c < in.ngroups; ++c)
        add_mid_slabs(pp.msrc, pp.msrc_n, kMaxMidSlabs, in.groups[3 + c], c, k + 1, true, true, in.metallic);
      if (mon_m > 0 && cap.ok) {
        pp.cap = pws.cap;
        pp.capbox = pws.capbox;
        pp.cap_slot = static_cast<int32_t>(cap.slot);
        pp.cap_m = mon_m;
      } else if (mon_m > 0) {
        pp.cmid = pws.bank;
        pp.mbox = pws.boxes;
        pp.mbox_n = mon_m;
      }
      // the shell launch (general kernel), then the core launch
      const PairSplit sp = pair_split(g, pp.f.L);
      pp.nbox = sp.nshell;
      for (int b = 0; b < sp.nshell; ++b) pp.box[b] = sp.shell[b];
      pp.core = 0;
      if (pp.nbox > 0) launches.push_back(make_launch(Kind::kPair, 0, g.dtype, pp));
      if (sp.has_core) {
        PairParams pc = pp;
        pc.nbox = 1;
        pc.box[0] = sp.core;
        pc.core = 1;
        launches.push_back(make_launch(Kind::kPair, 0, g.dtype, pc));
      }
      if (in.has_monitors && cap.ok) {
        add_dft_cap(k);
      } else if (in.has_monitors) {
        add_dft(k, cbank);
      }
      for (int c = 0; c < 3 && 6 + c < in.ngroups; ++c)
        add_source_launches(launches, in.groups[6 + c], 2, c, nxt, g, in.metallic, step0, k + 1);
      // (the second step is sampled from the fields: capturing it too costs
      // the kernel more stores than the strided gather costs)
      if (in.has_monitors) add_dft(k + 1, nxt);
      --pairs_left;
      cur_b = two_banks ? !cur_b : cur_b;
      k += 2;
      continue;
    }
    if (fused) {
      FusedParams fp = fused_params(cur, nxt, cur_b);
      fp.kstep = k;
      for (int c = 0; c < 3 && 3 + c < in.ngroups; ++c)
        add_mid_slabs(fp.hslab, fp.hslab_n, kMaxHSlabs, in.groups[3 + c], c, k, true, true, in.metallic);
      const FusedSplit sp = fused_split(g, fp.L);
      fp.rows = fused_rows();
      for (int v = 0; v < 2; ++v) {
        const FusedBox& b = v == 0 ? sp.interior : sp.edge;
        if (b.ntx == 0 || b.nty == 0 || b.nzc == 0) continue;
        set_box(fp, b, v);
        launches.push_back(make_launch(Kind::kFused, 0, g.dtype, fp));
      }
    } else {
      PhaseIn h = in.h, e = in.e;
      for (int c = 0; c < 3; ++c) {
        h.tin[c] = cur[kHx + c];
        h.tout[c] = nxt[kHx + c];
        h.src[c] = cur[kEx + c];
        e.tin[c] = cur[kEx + c];
        e.tout[c] = nxt[kEx + c];
        e.src[c] = nxt[kHx + c];
      }
      PhaseParams hp = make_phase(h, BZH_PHASE_H, g, in.dt, in.resolution, in.metallic, in.uniform_l);
      PhaseParams ep = make_phase(e, BZH_PHASE_E, g, in.dt, in.resolution, in.metallic, in.uniform_l);
      if (fold_tm) {
        // 2D, float32: the H-timed slabs (Hx, Hy) and the E-timed ones (Ez)
        // run inside the phase launches
        for (int c = 0; c < 2 && 3 + c < in.ngroups; ++c)
          fold_phase_sources(hp, in.groups[3 + c], c, kHx + c, nxt, g, in.metallic, step0, k);
        if (8 < in.ngroups)
          fold_phase_sources(ep, in.groups[8], 2, kEz, nxt, g, in.metallic, step0, k);
      }
      launches.push_back(make_launch(pk, BZH_PHASE_H, g.dtype, hp));
      if (!fold_tm)
        for (int c = 0; c < 3 && 3 + c < in.ngroups; ++c)
          add_source_launches(launches, in.groups[3 + c], 1, c, nxt, g, in.metallic, step0, k);
      launches.push_back(make_launch(pk, BZH_PHASE_E, g.dtype, ep));
    }
    if (!fold_tm)
      for (int c = 0; c < 3 && 6 + c < in.ngroups; ++c)
        add_source_launches(launches, in.groups[6 + c], 2, c, nxt, g, in.metallic, step0, k);
    if (in.has_monitors) add_dft(k, nxt);
    cur_b = two_banks ? !cur_b : cur_b;
    ++k;
  }
  // (no graphs for the no-allocation entry: instantiating one may allocate)
  graph_cache().run(s, launches, (in.flags & BZH_FLAG_GRAPH_CACHE) != 0 && !in.noalloc);
  g_last_pairs = npairs;
  g_last_capture = (npairs > 0 && mon_m > 0 && cap.ok) ? 1 : 0;
  g_last_fused = fused ? 1 : 0;
  return fused;
}

// ---- positional decoding (ffi_handler.cc:336-708) ---------------------------------
const bzh_buffer kEmpty = {nullptr, 1, BZH_F32, {0, 0, 0, 0}};

const bzh_buffer* arg(const bzh_buffer* args, int nargs, int i) {
  BZH_REQUIRE(i >= 0 && i < nargs, "positional argument " + std::to_string(i) +
                                       " missing (got " + std::to_string(nargs) + ")");
  return &args[i];
}

void decode_cpml_payload(const bzh_buffer* args, int nargs, int off, PhaseIn& ph,
                         bool cpml) {
  // off: first material of the phase payload
  for (int m = 0; m < 6; ++m) ph.mat[m] = arg(args, nargs, off + m);
  if (!cpml) {
    ph.nterms = 0;
    return;
  }
  // the CPML payload always carries 18 term arrays and 6 psi slots; a TMxy
  // phase uses the first two terms (metadata s32[2, 5])
  ph.meta = arg(args, nargs, off + 6);
  BZH_REQUIRE(ph.meta->rank == 2 && (ph.meta->dims[0] == 6 || ph.meta->dims[0] == 2),
              "term metadata must be s32[6, 5] (3D) or s32[2, 5] (TMxy)");
  ph.nterms = static_cast<int>(ph.meta->dims[0]);
  for (int t = 0; t < 18; ++t) ph.terms[t] = arg(args, nargs, off + 7 + t);
  for (int t = 0; t < 6; ++t) {
    ph.psi_in[t] = arg(args, nargs, off + 25 + t);
    ph.psi_out[t] = nullptr;
  }
}

void clear_phase(PhaseIn& ph) {
  std::memset(&ph, 0, sizeof(ph));
  for (int t = 0; t < 18; ++t) ph.terms[t] = &kEmpty;
  for (int t = 0; t < 6; ++t) ph.psi_in[t] = ph.psi_out[t] = nullptr;
  for (int a = 0; a < 3; ++a) ph.metric[a] = nullptr;
}

int decode_program(const bzh_buffer* args, int nargs, const bzh_program_attrs& at,
                   ProgramIn& in) {
  std::memset(&in, 0, sizeof(in));
  clear_phase(in.h);
  clear_phase(in.e);
  in.nsteps = at.nsteps;
  in.dt = at.dt;
  in.resolution = at.resolution;
  in.metallic = at.boundary_code & 0xFF;
  in.uniform_l = at.boundary_code >> 8;
  in.flags = at.flags;
  in.monitor_count = at.monitor_count;
  const int layout = at.program_layout;
  const bool cpml = at.cpml_enabled != 0;
  const bool temporal = layout == BZH_LAYOUT_YEE_TEMPORAL ||
                        layout == BZH_LAYOUT_SOURCE_TEMPORAL_CPML ||
                        layout == BZH_LAYOUT_MONITOR_TEMPORAL_CPML;
  BZH_REQUIRE(layout >= 0 && layout <= 6, "unknown program layout");
  if (layout == BZH_LAYOUT_CPML_IN_PLACE) BZH_REQUIRE(cpml, "layout 2 requires cpml_enabled");
  if (layout == BZH_LAYOUT_YEE_IN_PLACE || layout == BZH_LAYOUT_YEE_TEMPORAL)
    BZH_REQUIRE(!cpml, "Yee layouts carry no CPML payload");
  int next = 0;
  if (layout == BZH_LAYOUT_YEE_TEMPORAL) {
    // fields [0-5], workspace [6-11], materials [12-23], metrics [24-29]
    for (int c = 0; c < 6; ++c) in.bank_a[c] = arg(args, nargs, c);
    for (int c = 0; c < 6; ++c) in.bank_b[c] = arg(args, nargs, 6 + c);
    for (int m = 0; m < 6; ++m) in.h.mat[m] = arg(args, nargs, 12 + m);
    for (int m = 0; m < 6; ++m) in.e.mat[m] = arg(args, nargs, 18 + m);
    for (int a = 0; a < 3; ++a) in.h.metric[a] = arg(args, nargs, 24 + a);
    for (int a = 0; a < 3; ++a) in.e.metric[a] = arg(args, nargs, 27 + a);
    next = 30;
  } else if (!cpml) {
    // fields [0-5], H materials [6-11], E materials [12-17], metrics [18-23]
    for (int c = 0; c < 6; ++c) in.bank_a[c] = arg(args, nargs, c);
    for (int m = 0; m < 6; ++m) in.h.mat[m] = arg(args, nargs, 6 + m);
    for (int m = 0; m < 6; ++m) in.e.mat[m] = arg(args, nargs, 12 + m);
    for (int a = 0; a < 3; ++a) in.h.metric[a] = arg(args, nargs, 18 + a);
    for (int a = 0; a < 3; ++a) in.e.metric[a] = arg(args, nargs, 21 + a);
    next = BZH_YEE_GRAPH_INPUT_COUNT;
    if (temporal) {
      // extension: layouts 4 and 6 without CPML take the 24-input Yee graph
      // followed by the six-field workspace bank
      for (int c = 0; c < 6; ++c) in.bank_b[c] = arg(args, nargs, next + c);
      next += 6;
    }
  } else {
    for (int c = 0; c < 6; ++c) in.bank_a[c] = arg(args, nargs, c);
    decode_cpml_payload(args, nargs, 6, in.h, true);
    decode_cpml_payload(args, nargs, BZH_CPML_E_INPUT_OFFSET, in.e, true);
    for (int a = 0; a < 3; ++a) in.h.metric[a] = arg(args, nargs, BZH_CPML_H_METRIC_OFFSET + a);
    for (int a = 0; a < 3; ++a) in.e.metric[a] = arg(args, nargs, BZH_CPML_E_METRIC_OFFSET + a);
    next = BZH_CPML_GRAPH_INPUT_COUNT;
    if (temporal) {
      for (int c = 0; c < 6; ++c) in.bank_b[c] = arg(args, nargs, BZH_TEMPORAL_FIELD_WORKSPACE_INPUT + c);
      // psi workspace (args 80-91: six H terms, then six E terms).  The fused
      // schedule ping-pongs psi with the fields; the per-phase schedule
      // updates bank A in place and copies it here after an odd step count.
      for (int t = 0; t < 12; ++t) in.psi_b[t] = arg(args, nargs, BZH_TEMPORAL_PSI_WORKSPACE_INPUT + t);
      next = BZH_TEMPORAL_PSI_WORKSPACE_INPUT + 12;
    }
  }
  if (layout == BZH_LAYOUT_SOURCE_IN_PLACE || layout == BZH_LAYOUT_SOURCE_TEMPORAL_CPML ||
      layout == BZH_LAYOUT_MONITOR_IN_PLACE || layout == BZH_LAYOUT_MONITOR_TEMPORAL_CPML) {
    for (int gi = 0; gi < BZH_SOURCE_GROUP_COUNT; ++gi) {
      in.groups[gi].coeffs = arg(args, nargs, next + 3 * gi + 0);
      in.groups[gi].waves = arg(args, nargs, next + 3 * gi + 1);
      in.groups[gi].starts = arg(args, nargs, next + 3 * gi + 2);
      in.groups[gi].coincident = (at.coincident_source_group_mask >> gi) & 1;
    }
    in.ngroups = BZH_SOURCE_GROUP_COUNT;
    next += 3 * BZH_SOURCE_GROUP_COUNT;
  }
  if (layout == BZH_LAYOUT_SOURCE_IN_PLACE || layout == BZH_LAYOUT_SOURCE_TEMPORAL_CPML) {
    in.step = arg(args, nargs, next++);
  }
  if (layout == BZH_LAYOUT_MONITOR_IN_PLACE || layout == BZH_LAYOUT_MONITOR_TEMPORAL_CPML) {
    for (int q = 0; q < 12; ++q) in.mon.b[q] = arg(args, nargs, next + q);
    in.step = in.mon.b[11];
    in.has_monitors = true;
    next += 12;
    if (nargs == next + 1) {
      const bzh_buffer* hb = arg(args, nargs, next++);
      BZH_REQUIRE(hb->element_type == BZH_S32 && numel(*hb) == 36 * in.monitor_count &&
                      (hb->data != nullptr || in.monitor_count == 0),
                  "monitor boxes must be host s32[M, 6, 6]");
      in.mon_boxes = hb;
    }
  }
  BZH_REQUIRE(next == nargs, "layout " + std::to_string(layout) + " expects " +
                                 std::to_string(next) + " arguments, got " + std::to_string(nargs));
  if (in.ngroups > 0) BZH_REQUIRE(in.step && in.step->element_type == BZH_S32,
                                  "current_step must be s32");
  return temporal ? 1 : 0;
}

void copy_psi_workspace(hipStream_t s, const bzh_buffer* args, int nargs) {
  // result banks of the temporal layouts: psi of an odd step count lives in
  // the workspace (runtime.py:393-440)
  for (int t = 0; t < 12; ++t) {
    const bzh_buffer* src = arg(args, nargs, (t < 6 ? BZH_CPML_H_PSI_INPUT_OFFSET : BZH_CPML_E_PSI_INPUT_OFFSET) + (t % 6));
    const bzh_buffer* dst = arg(args, nargs, BZH_TEMPORAL_PSI_WORKSPACE_INPUT + t);
    BZH_REQUIRE(numel(*src) == numel(*dst) && src->element_type == dst->element_type,
                "psi workspace does not match psi");
    const size_t bytes = static_cast<size_t>(numel(*src)) * (src->element_type == BZH_BF16 ? 2 : 4);
    if (bytes) BZH_HIP(hipMemcpyAsync(dst->data, src->data, bytes, hipMemcpyDeviceToDevice, s));
  }
}

// ---- batched 2D simulations ---------------------------------------------------
// Device copies of the descriptor table and of the item list, keyed on their
// bytes.  They are kept for the life of the process: a captured graph refers
// to them.
template <typename V>
const void* upload_cached(const std::vector<V>& host) {
  static std::mutex mu;
  static std::map<std::string, void*> cache;
  int device = 0;
  BZH_HIP(hipGetDevice(&device));
  std::string key(reinterpret_cast<const char*>(&device), sizeof(device));
  key.append(reinterpret_cast<const char*>(host.data()), host.size() * sizeof(V));
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  void* d = nullptr;
  BZH_HIP(hipMalloc(&d, std::max<size_t>(host.size() * sizeof(V), 16)));
  BZH_HIP(hipMemcpy(d, host.data(), host.size() * sizeof(V), hipMemcpyHostToDevice));
  cache[key] = d;
  return d;
}

float* f32_ptr(const bzh_buffer& b, int64_t expect, const std::string& what) {
  BZH_REQUIRE(b.element_type == BZH_F32 && b.data != nullptr, what + " must be a non-null f32 buffer");
  BZH_REQUIRE(numel(b) == expect, what + " has " + std::to_string(numel(b)) + " elements, expected " +
                                      std::to_string(expect));
  return static_cast<float*>(b.data);
}


Sim2dDev describe_sim2d(const bzh_sim2d& in, int index) {
  const std::string who = "simulation " + std::to_string(index) + ": ";
  Sim2dDev d;
  std::memset(&d, 0, sizeof(d));
  BZH_REQUIRE(in.ez[0].rank == 2, who + "Ez must be rank 2");
  d.ny = static_cast<int>(in.ez[0].dims[0]) - 1;
  d.nx = static_cast<int>(in.ez[0].dims[1]) - 1;
  BZH_REQUIRE(d.ny >= 1 && d.nx >= 1, who + "grid must have at least one cell");
  BZH_REQUIRE(static_cast<int64_t>(d.ny + 1) * (d.nx + 1) < (int64_t{1} << 31), who + "grid too large");
  const int64_t nez = static_cast<int64_t>(d.ny + 1) * (d.nx + 1);
  const int64_t nhx = static_cast<int64_t>(d.ny) * (d.nx + 1);
  const int64_t nhy = static_cast<int64_t>(d.ny + 1) * d.nx;
  for (int b = 0; b < 2; ++b) {
    d.ez[b] = f32_ptr(in.ez[b], nez, who + "Ez");
    d.hx[b] = f32_ptr(in.hx[b], nhx, who + "Hx");
    d.hy[b] = f32_ptr(in.hy[b], nhy, who + "Hy");
  }
  BZH_REQUIRE(d.ez[0] != d.ez[1] && d.hx[0] != d.hx[1] && d.hy[0] != d.hy[1],
              who + "the two field banks must be distinct");
  d.e_dec = f32_ptr(in.e_decay, 1, who + "E decay");
  d.h_dec = f32_ptr(in.h_decay, 1, who + "H decay");
  d.h_src = f32_ptr(in.h_source, 1, who + "H source");
  d.e_src_st = numel(in.e_source) == 1 ? 0 : 1;
  d.e_src = f32_ptr(in.e_source, d.e_src_st ? nez : 1, who + "E source");
  BZH_REQUIRE(in.resolution > 0.0f, who + "resolution must be positive");
  d.inv_dx = 1.0f / in.resolution;
  d.metallic = in.metallic & (kBottom | kTop | kLeft | kRight);
  d.L = in.cpml_cells;
  BZH_REQUIRE(d.L >= 0 && 2 * d.L <= std::min(d.ny, d.nx), who + "CPML thicker than the grid");
  if (d.L > 0) {
    const int64_t L2 = 2 * d.L;
    d.tab = f32_ptr(in.cpml_table, 12 * L2, who + "CPML table");
    const int64_t sizes[4] = {L2 * (d.nx + 1), (d.ny + 1) * L2, (d.ny + 1) * L2, L2 * (d.nx + 1)};
    for (int b = 0; b < 2; ++b)
      for (int t = 0; t < 4; ++t)
        d.psi[b][t] = f32_ptr(in.psi[b][t], sizes[t], who + "CPML memory " + std::to_string(t));
  }
  BZH_REQUIRE(in.nsrc >= 0 && in.nsrc <= 4, who + "at most four sources");
  d.nsrc = in.nsrc;
  for (int q = 0; q < in.nsrc; ++q) {
    const bzh_buffer& c = in.src_coeff[q];
    BZH_REQUIRE(c.rank == 2, who + "source coefficients must be (sy, sx)");
    d.src_sy[q] = static_cast<int>(c.dims[0]);
    d.src_sx[q] = static_cast<int>(c.dims[1]);
    d.src_y0[q] = in.src_start[q][0];
    d.src_x0[q] = in.src_start[q][1];
    BZH_REQUIRE(d.src_y0[q] >= 0 && d.src_x0[q] >= 0 && d.src_y0[q] + d.src_sy[q] <= d.ny + 1 &&
                    d.src_x0[q] + d.src_sx[q] <= d.nx + 1,
                who + "source slab leaves the Ez support");
    d.src_coeff[q] = f32_ptr(c, static_cast<int64_t>(d.src_sy[q]) * d.src_sx[q], who + "source coefficients");
    d.src_T[q] = static_cast<int>(numel(in.src_wave[q]));
    BZH_REQUIRE(d.src_T[q] >= 1, who + "source waveform is empty");
    d.src_wave[q] = f32_ptr(in.src_wave[q], d.src_T[q], who + "source waveform");
  }
  return d;
}

void run_batch2d(hipStream_t s, const bzh_sim2d* sims, int nsims, const bzh_buffer* step,
                 const bzh_batch_attrs& a) {
  BZH_REQUIRE(sims != nullptr && nsims >= 1, "empty batch");
  BZH_REQUIRE(a.nsteps >= 1, "step count must be positive");
  BZH_REQUIRE(a.tile_rows == 8 || a.tile_rows == 16 || a.tile_rows == 32 || a.tile_rows == 64,
              "tile_rows must be 8, 16, 32 or 64");
  BZH_REQUIRE(step != nullptr && step->element_type == BZH_S32 && step->data && numel(*step) == 1,
              "current_step must be an s32 scalar");
  std::vector<Sim2dDev> desc(nsims);
  for (int i = 0; i < nsims; ++i) desc[i] = describe_sim2d(sims[i], i);
  // queue: largest simulations first, row-major items (62 columns by
  // tile_rows rows) within one
  std::vector<int> order(nsims);
  for (int i = 0; i < nsims; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return static_cast<int64_t>(desc[x].ny) * desc[x].nx > static_cast<int64_t>(desc[y].ny) * desc[y].nx;
  });
  std::vector<int4> tiles;
  for (int i : order) {
    const int ty = (desc[i].ny + 1 + a.tile_rows - 1) / a.tile_rows, tx = (desc[i].nx + 1 + 61) / 62;
    for (int r = 0; r < ty; ++r)
      for (int c = 0; c < tx; ++c) tiles.push_back(make_int4(i, r, c, 0));
  }
  BZH_REQUIRE(tiles.size() < (size_t{1} << 31), "too many tiles");
  int workgroups = a.workgroups;
  if (workgroups <= 0) {
    int device = 0, cus = 0;
    BZH_HIP(hipGetDevice(&device));
    BZH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    workgroups = 4 * std::max(cus, 1);
  }
  Batch2dParams p;
  std::memset(&p, 0, sizeof(p));
  p.sims = static_cast<const Sim2dDev*>(upload_cached(desc));
  p.tiles = static_cast<const int4*>(upload_cached(tiles));
  p.ntiles = static_cast<int>(tiles.size());
  p.step0 = static_cast<const int32_t*>(step->data);
  p.tile_rows = a.tile_rows;
  p.workgroups = workgroups;
  p.nsims = nsims;
  for (const Sim2dDev& d : desc)
    for (int q = 0; q < d.nsrc; ++q) p.max_src_cells = std::max(p.max_src_cells, d.src_sy[q] * d.src_sx[q]);
  std::vector<Launch> launches;
  for (int k = 0; k < a.nsteps; ++k) {
    p.k = k;
    p.bank = k & 1;
    // the step's pre-E sources on the bank it reads, then the march
    if (p.max_src_cells > 0) launches.push_back(make_launch(Kind::kBatch2d, 1, BZH_F32, p));
    launches.push_back(make_launch(Kind::kBatch2d, 0, BZH_F32, p));
  }
  graph_cache().run(s, launches, (a.flags & BZH_FLAG_GRAPH_CACHE) != 0);
}

// ---- adjoint TMxy -----------------------------------------------------------
Adjoint2dParams describe_adjoint(const bzh_adjoint2d& in, int64_t* record_slots) {
  Adjoint2dParams p;
  std::memset(&p, 0, sizeof(p));
  BZH_REQUIRE(in.a_ez.rank == 2, "adjoint Ez must be rank 2");
  p.ny = static_cast<int>(in.a_ez.dims[0]) - 1;
  p.nx = static_cast<int>(in.a_ez.dims[1]) - 1;
  BZH_REQUIRE(p.ny >= 2 && p.nx >= 2, "adjoint grid too small");
  const int64_t nez = static_cast<int64_t>(p.ny + 1) * (p.nx + 1);
  const int64_t nhx = static_cast<int64_t>(p.ny) * (p.nx + 1);
  const int64_t nhy = static_cast<int64_t>(p.ny + 1) * p.nx;
  BZH_REQUIRE(nez < (int64_t{1} << 31), "adjoint grid too large");
  p.a_ez = f32_ptr(in.a_ez, nez, "adjoint Ez");
  p.a_hx = f32_ptr(in.a_hx, nhx, "adjoint Hx");
  p.a_hy = f32_ptr(in.a_hy, nhy, "adjoint Hy");
  p.w_e0 = f32_ptr(in.work[0], nez, "adjoint scratch 0");
  p.w_e1 = f32_ptr(in.work[1], nez, "adjoint scratch 1");
  p.w_hy = f32_ptr(in.work[2], nhy, "adjoint scratch 2");
  p.w_hx = f32_ptr(in.work[3], nhx, "adjoint scratch 3");
  p.e_dec = f32_ptr(in.e_decay, 1, "E decay");
  p.h_dec = f32_ptr(in.h_decay, 1, "H decay");
  p.h_src = f32_ptr(in.h_source, 1, "H source");
  p.e_src_st = numel(in.e_source) == 1 ? 0 : 1;
  p.e_src = f32_ptr(in.e_source, p.e_src_st ? nez : 1, "E source");
  BZH_REQUIRE(in.resolution > 0.0f, "resolution must be positive");
  p.inv_dx = 1.0f / in.resolution;
  p.metallic = in.metallic & (kBottom | kTop | kLeft | kRight);
  p.L = in.cpml_cells;
  BZH_REQUIRE(p.L >= 0 && 2 * p.L <= std::min(p.ny, p.nx), "CPML thicker than the grid");
  if (p.L > 0) {
    const int64_t L2 = 2 * p.L;
    p.tab = f32_ptr(in.cpml_table, 12 * L2, "CPML table");
    const int64_t sizes[4] = {L2 * (p.nx + 1), (p.ny + 1) * L2, (p.ny + 1) * L2, L2 * (p.nx + 1)};
    for (int t = 0; t < 4; ++t) p.a_psi[t] = f32_ptr(in.a_psi[t], sizes[t], "adjoint CPML memory " + std::to_string(t));
  }
  if (!is_empty(in.source_profile)) {
    p.src_profile = f32_ptr(in.source_profile, nez, "adjoint source profile");
    BZH_REQUIRE(in.source_time.element_type == BZH_F32 && in.source_time.data && numel(in.source_time) >= 1,
                "adjoint source time factors must be f32");
    p.src_time = static_cast<const float*>(in.source_time.data);
  }
  for (int q = 0; q < 4; ++q) p.design[q] = in.design[q];
  const int dh = p.design[1] - p.design[0], dw = p.design[3] - p.design[2];
  // the recorded curl is the plain one: the box keeps off CPML slabs and faces
  const int M = std::max(p.L, 1);
  BZH_REQUIRE(dh >= 1 && dw >= 1 && p.design[0] >= M && p.design[1] <= p.ny - M &&
                  p.design[2] >= M && p.design[3] <= p.nx - M,
              "design box must be non-empty and clear of CPML slabs and edges");
  BZH_REQUIRE(in.record.element_type == BZH_F32 && in.record.data && in.record.rank == 3 &&
                  in.record.dims[1] == dh && in.record.dims[2] == dw,
              "record must be f32 [K, design rows, design columns]");
  *record_slots = in.record.dims[0];
  p.record = static_cast<float*>(in.record.data);
  BZH_REQUIRE(in.grad.element_type == BZH_F64 && in.grad.data && numel(in.grad) == int64_t{dh} * dw,
              "grad must be f64 [design rows, design columns]");
  p.grad = static_cast<double*>(in.grad.data);
  return p;
}

template <typename F>
int32_t guarded(F&& f) {
  clear_error();
  try {
    f();
    return 0;
  } catch (const Failure& e) {
    set_error(e.message);
    return 1;
  } catch (const std::exception& e) {
    set_error(e.what());
    return 1;
  } catch (...) {
    set_error("unknown native failure");
    return 1;
  }
}

}  // namespace
}  // namespace bzh

using namespace bzh;

extern "C" {

int32_t bzh_abi_version(void) { return BZH_ABI_VERSION; }
const char* bzh_component_version(void) { return BZH_COMPONENT_VERSION; }
const char* bzh_targets(void) { return "beamz_hip_streamed,beamz_hip_program"; }
const char* bzh_last_error(void) { return bzh::g_error.c_str(); }
int32_t bzh_last_program_fused(void) { return bzh::g_last_fused; }

int32_t bzh_last_program_pairs(void) { return bzh::g_last_pairs; }
int32_t bzh_last_program_captured(void) { return bzh::g_last_capture; }

int32_t bzh_launch_phase(void* stream, const bzh_launch* launch) {
  return guarded([&] {
    BZH_REQUIRE(launch != nullptr, "launch descriptor is null");
    BZH_REQUIRE(launch->abi_version == BZH_ABI_VERSION, "ABI version mismatch");
    BZH_REQUIRE(launch->phase == BZH_PHASE_H || launch->phase == BZH_PHASE_E, "phase must be 0 (H) or 1 (E)");
    PhaseIn ph;
    clear_phase(ph);
    const bzh_buffer* bank[6];
    for (int c = 0; c < 3; ++c) {
      ph.tin[c] = &launch->inputs[c];
      ph.tout[c] = &launch->outputs[c];
      ph.src[c] = &launch->inputs[3 + c];
    }
    for (int m = 0; m < 6; ++m) ph.mat[m] = &launch->inputs[6 + m];
    ph.nterms = launch->nterms;
    if (launch->nterms > 0) {
      ph.meta = &launch->inputs[12];
      for (int t = 0; t < 18; ++t) ph.terms[t] = &launch->inputs[13 + t];
      for (int t = 0; t < 6; ++t) {
        ph.psi_in[t] = &launch->inputs[31 + t];
        ph.psi_out[t] = &launch->outputs[3 + t];
      }
    }
    for (int a = 0; a < 3; ++a) ph.metric[a] = &launch->metrics[a];
    const bool hphase = launch->phase == BZH_PHASE_H;
    for (int c = 0; c < 3; ++c) {
      bank[(hphase ? kHx : kEx) + c] = &launch->inputs[c];
      bank[(hphase ? kEx : kHx) + c] = &launch->inputs[3 + c];
    }
    const Grid g = infer_grid(bank);
    const PhaseParams p = make_phase(ph, launch->phase, g, launch->dt, launch->resolution,
                                     launch->metallic_edges, launch->uniform_cpml_thickness);
    run_phase(static_cast<hipStream_t>(stream), p, launch->phase, g);
  });
}

int32_t bzh_streamed(void* stream, const bzh_buffer* args, int32_t nargs,
                     const bzh_phase_attrs* attrs) {
  return guarded([&] {
    BZH_REQUIRE(attrs != nullptr && args != nullptr, "null arguments");
    BZH_REQUIRE(attrs->abi_version == BZH_ABI_VERSION, "ABI version mismatch");
    BZH_REQUIRE(attrs->phase == BZH_PHASE_H || attrs->phase == BZH_PHASE_E, "phase must be 0 (H) or 1 (E)");
    BZH_REQUIRE(attrs->nterms == 0 || attrs->nterms == 6 || attrs->nterms == 2,
                "nterms must be 0, 2 (TMxy) or 6 (3D)");
    PhaseIn ph;
    clear_phase(ph);
    for (int c = 0; c < 3; ++c) {
      ph.tin[c] = arg(args, nargs, c);
      ph.tout[c] = nullptr;
      ph.src[c] = arg(args, nargs, 3 + c);
    }
    for (int m = 0; m < 6; ++m) ph.mat[m] = arg(args, nargs, 6 + m);
    int next = 12;
    ph.nterms = attrs->nterms;
    if (attrs->nterms > 0) {
      ph.meta = arg(args, nargs, next++);
      for (int t = 0; t < 3 * attrs->nterms; ++t) ph.terms[t] = arg(args, nargs, next++);
      for (int t = 0; t < attrs->nterms; ++t) ph.psi_in[t] = arg(args, nargs, next++);
    }
    for (int a = 0; a < 3; ++a) ph.metric[a] = arg(args, nargs, next++);
    BZH_REQUIRE(next == nargs, "streamed call expects " + std::to_string(next) + " arguments");
    const bool hphase = attrs->phase == BZH_PHASE_H;
    const bzh_buffer* bank[6];
    for (int c = 0; c < 3; ++c) {
      bank[(hphase ? kHx : kEx) + c] = ph.tin[c];
      bank[(hphase ? kEx : kHx) + c] = ph.src[c];
    }
    const Grid g = infer_grid(bank);
    const PhaseParams p = make_phase(ph, attrs->phase, g, attrs->dt, attrs->resolution,
                                     attrs->boundary_code & 0xFF, attrs->boundary_code >> 8);
    run_phase(static_cast<hipStream_t>(stream), p, attrs->phase, g);
  });
}

int32_t bzh_program(void* stream, const bzh_buffer* args, int32_t nargs,
                    const bzh_program_attrs* attrs) {
  return guarded([&] {
    BZH_REQUIRE(attrs != nullptr && args != nullptr, "null arguments");
    BZH_REQUIRE(attrs->abi_version == BZH_ABI_VERSION, "ABI version mismatch");
    ProgramIn in;
    const int temporal = decode_program(args, nargs, *attrs, in);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool fused = run_program(s, in);
    if (temporal && attrs->cpml_enabled && !fused && (attrs->nsteps % 2 == 1))
      copy_psi_workspace(s, args, nargs);
  });
}

int32_t bzh_program_noalloc(void* stream, const bzh_buffer* args, int32_t nargs,
                            const bzh_program_attrs* attrs) {
  return guarded([&] {
    BZH_REQUIRE(attrs != nullptr && args != nullptr, "null arguments");
    BZH_REQUIRE(attrs->abi_version == BZH_ABI_VERSION, "ABI version mismatch");
    ProgramIn in;
    const int temporal = decode_program(args, nargs, *attrs, in);
    in.noalloc = true;
    (void)noalloc_scratch();  // set-up: the device address, once per device
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool fused = run_program(s, in);
    if (temporal && attrs->cpml_enabled && !fused && (attrs->nsteps % 2 == 1))
      copy_psi_workspace(s, args, nargs);
  });
}

int32_t bzh_batch2d(void* stream, const bzh_sim2d* sims, int32_t nsims,
                    const bzh_buffer* current_step, const bzh_batch_attrs* attrs) {
  return guarded([&] {
    BZH_REQUIRE(attrs != nullptr, "null attributes");
    BZH_REQUIRE(attrs->abi_version == BZH_ABI_VERSION, "ABI version mismatch");
    run_batch2d(static_cast<hipStream_t>(stream), sims, nsims, current_step, *attrs);
  });
}

int32_t bzh_adjoint2d_record(void* stream, const bzh_adjoint2d* adj, const bzh_buffer* hx,
                             const bzh_buffer* hy, int32_t slot) {
  return guarded([&] {
    BZH_REQUIRE(adj != nullptr && hx != nullptr && hy != nullptr, "null arguments");
    int64_t slots = 0;
    Adjoint2dParams p = describe_adjoint(*adj, &slots);
    BZH_REQUIRE(slot >= 0 && slot < slots, "record slot out of range");
    p.slot = slot;
    const float* fhx = f32_ptr(*hx, static_cast<int64_t>(p.ny) * (p.nx + 1), "forward Hx");
    const float* fhy = f32_ptr(*hy, static_cast<int64_t>(p.ny + 1) * p.nx, "forward Hy");
    launch_adjoint_record(static_cast<hipStream_t>(stream), p, fhx, fhy);
  });
}

int32_t bzh_adjoint2d_reverse(void* stream, const bzh_adjoint2d* adj, int32_t n_last,
                              int32_t count, int32_t slot_last, int32_t flags) {
  return guarded([&] {
    BZH_REQUIRE(adj != nullptr, "null arguments");
    int64_t slots = 0;
    Adjoint2dParams p = describe_adjoint(*adj, &slots);
    BZH_REQUIRE(count >= 1 && n_last - count + 1 >= 0, "reverse range out of bounds");
    BZH_REQUIRE(slot_last < slots && slot_last - count + 1 >= 0, "record slots out of range");
    if (p.src_profile)
      BZH_REQUIRE(n_last < numel(adj->source_time), "source time factors shorter than the run");
    std::vector<Launch> launches;
    for (int c = 0; c < count; ++c) {
      p.n = n_last - c;
      p.slot = slot_last - c;
      launches.push_back(make_launch(Kind::kAdjoint, 0, BZH_F32, p));
    }
    graph_cache().run(static_cast<hipStream_t>(stream), launches, (flags & BZH_FLAG_GRAPH_CACHE) != 0);
  });
}

int32_t bzh_launch_program(void* stream, const bzh_program_launch* program) {
  return guarded([&] {
    BZH_REQUIRE(program != nullptr, "program descriptor is null");
    BZH_REQUIRE(program->field_bank_count == 1 || program->field_bank_count == 2,
                "field_bank_count must be 1 or 2");
    ProgramIn in;
    std::memset(&in, 0, sizeof(in));
    clear_phase(in.h);
    clear_phase(in.e);
    const bzh_launch& h = program->h_ab;
    const bzh_launch& e = program->e_ab;
    BZH_REQUIRE(h.abi_version == BZH_ABI_VERSION && e.abi_version == BZH_ABI_VERSION,
                "ABI version mismatch");
    for (int c = 0; c < 3; ++c) {
      in.bank_a[kHx + c] = &h.inputs[c];
      in.bank_a[kEx + c] = &e.inputs[c];
      if (program->field_bank_count == 2) {
        in.bank_b[kHx + c] = &h.outputs[c];
        in.bank_b[kEx + c] = &e.outputs[c];
      }
    }
    if (program->field_bank_count == 2 && h.nterms > 0) {
      // two banks: outputs[3..8] of the *_ab launches are the psi workspace
      for (int t = 0; t < 6; ++t) {
        in.psi_b[t] = &h.outputs[3 + t];
        in.psi_b[6 + t] = &e.outputs[3 + t];
      }
    }
    auto fill = [](PhaseIn& ph, const bzh_launch& l) {
      for (int m = 0; m < 6; ++m) ph.mat[m] = &l.inputs[6 + m];
      ph.nterms = l.nterms;
      if (l.nterms > 0) {
        ph.meta = &l.inputs[12];
        for (int t = 0; t < 18; ++t) ph.terms[t] = &l.inputs[13 + t];
        for (int t = 0; t < 6; ++t) ph.psi_in[t] = &l.inputs[31 + t];
      }
      for (int a = 0; a < 3; ++a) ph.metric[a] = &l.metrics[a];
    };
    fill(in.h, h);
    fill(in.e, e);
    in.nsteps = program->nsteps;
    in.dt = h.dt;
    in.resolution = h.resolution;
    in.metallic = h.metallic_edges;
    in.uniform_l = h.uniform_cpml_thickness;
    in.flags = h.flags;
    BZH_REQUIRE(program->source_group_count == 0 || program->source_group_count == 9,
                "source_group_count must be 0 or 9");
    in.ngroups = program->source_group_count;
    for (int gi = 0; gi < in.ngroups; ++gi) {
      const bzh_source_group& sg = program->source_groups[gi];
      in.groups[gi].coeffs = &sg.coefficients;
      in.groups[gi].waves = &sg.waveforms;
      in.groups[gi].starts = &sg.starts;
      in.groups[gi].coincident = sg.coincident;
      if (sg.current_step.data) in.step = &sg.current_step;
    }
    if (program->monitors != nullptr) {
      const bzh_dft_group& m = *program->monitors;
      const bzh_buffer* list[12] = {&m.indices, &m.weights, &m.frequencies, &m.component_masks,
                                    &m.counts, &m.codes, &m.windows, &m.dft_re, &m.dft_im,
                                    &m.dft_weight, &m.time, &m.current_step};
      for (int q = 0; q < 12; ++q) in.mon.b[q] = list[q];
      in.has_monitors = true;
      in.monitor_count = m.monitor_count;
      if (in.step == nullptr) in.step = &m.current_step;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool fused = run_program(s, in);
    if (program->field_bank_count == 2 && h.nterms > 0 && !fused && (in.nsteps % 2 == 1)) {
      // per-phase schedule: psi was updated in place; the result convention
      // puts it in the workspace after an odd step count
      for (int t = 0; t < 12; ++t) {
        const bzh_buffer* src = t < 6 ? in.h.psi_in[t] : in.e.psi_in[t - 6];
        const bzh_buffer* dst = in.psi_b[t];
        BZH_REQUIRE(numel(*src) == numel(*dst) && src->element_type == dst->element_type,
                    "psi workspace does not match psi");
        const size_t bytes = static_cast<size_t>(numel(*src)) * (src->element_type == BZH_BF16 ? 2 : 4);
        if (bytes) BZH_HIP(hipMemcpyAsync(dst->data, src->data, bytes, hipMemcpyDeviceToDevice, s));
      }
    }
  });
}

}  // extern "C"
