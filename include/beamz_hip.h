/*
 * beamz_hip.h -- C-ABI of the MI355X (gfx950) Yee/CPML time-stepping library
 * libbeamz_hip.so.
 *
 * Plain pointers and sizes only.  Every buffer is caller-owned device memory;
 * the library keeps non-owning descriptors for the duration of a call and
 * updates outputs in place (outputs alias inputs, as the reference's
 * input_output_aliases do).  All work is enqueued on the caller's HIP stream;
 * a nonzero return means nothing further was enqueued and bzh_last_error()
 * holds a message for the calling thread.
 *
 * Reference interface replaced (paths relative to the BeamZ tree):
 *   cuda/src/launch.h:12-77            descriptors  -> bzh_buffer, bzh_launch, ...
 *   cuda/src/extension.cc:10-23        module query -> bzh_abi_version, ...
 *   cuda/src/ffi_handler.cc:722-735    beamz_cuda_streamed -> bzh_streamed
 *   cuda/src/ffi_handler.cc:737-765    beamz_cuda_program  -> bzh_program
 *   cuda/src/update.cu:919-971         BeamzLaunchStreamed -> bzh_launch_phase
 *   cuda/src/program.cu:339-354        BeamzLaunchProgram  -> bzh_launch_program
 *   cuda/abi_layout.json               layout constants (mirrored below)
 */
#ifndef BEAMZ_HIP_H_
#define BEAMZ_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ABI constants (cuda/abi_layout.json:1-67) ---------------------------- */
#define BZH_ABI_VERSION 12
#define BZH_COMPONENT_VERSION "0.12.0-hip"

#define BZH_FLAG_GRAPH_CACHE (1 << 7)  /* abi_layout.json "graph_cache": 7 */
#define BZH_FLAG_BF16_PSI (1 << 12)    /* abi_layout.json "bf16_psi": 12 (accepted; psi dtype is read from the buffer) */
#define BZH_FLAG_FUSED_STEP (1 << 16)  /* no reference counterpart: allow the fused two-bank full-step kernel */

#define BZH_LAYOUT_YEE_IN_PLACE 0
#define BZH_LAYOUT_YEE_TEMPORAL 1
#define BZH_LAYOUT_CPML_IN_PLACE 2
#define BZH_LAYOUT_SOURCE_IN_PLACE 3
#define BZH_LAYOUT_SOURCE_TEMPORAL_CPML 4
#define BZH_LAYOUT_MONITOR_IN_PLACE 5
#define BZH_LAYOUT_MONITOR_TEMPORAL_CPML 6

#define BZH_FIELD_COUNT 6
#define BZH_CPML_TERM_COUNT 6
#define BZH_SOURCE_GROUP_COUNT 9
#define BZH_SOURCE_GROUP_BUFFER_COUNT 3
#define BZH_MONITOR_INPUT_COUNT 12
#define BZH_YEE_GRAPH_INPUT_COUNT 24
#define BZH_CPML_GRAPH_INPUT_COUNT 74
#define BZH_CPML_PHASE_INPUT_COUNT 31
#define BZH_CPML_E_INPUT_OFFSET 37
#define BZH_CPML_H_METRIC_OFFSET 68
#define BZH_CPML_E_METRIC_OFFSET 71
#define BZH_CPML_H_PSI_INPUT_OFFSET 31
#define BZH_CPML_E_PSI_INPUT_OFFSET 62
#define BZH_TEMPORAL_FIELD_WORKSPACE_INPUT 74
#define BZH_TEMPORAL_PSI_WORKSPACE_INPUT 80

#define BZH_PHASE_H 0
#define BZH_PHASE_E 1

/* Element types (cuda/src/launch.h:6-10; F16 and F64 are additions). */
enum bzh_element_type {
  BZH_F32 = 0,
  BZH_S32 = 1,
  BZH_BF16 = 2,
  BZH_F16 = 3,
  BZH_F64 = 4
};

/* A non-owning device-buffer descriptor (cuda/src/launch.h:12-17).  Row-major,
 * contiguous.  rank 0 is a scalar; an empty buffer has a zero dimension. */
typedef struct bzh_buffer {
  void* data;
  int32_t rank;
  int32_t element_type;
  int64_t dims[4];
} bzh_buffer;

/* One Yee phase (cuda/src/launch.h:19-35).  inputs[] follows the streamed
 * positional order: 3 targets, 3 sources (the other phase's fields), 6
 * materials (decay x,y,z then source x,y,z; rank-0 = uniform), then when
 * nterms == 6: term metadata s32[6,5] (component, axis, low, high, sign),
 * 18 term arrays (a, b, 1/kappa per term, packed along the term's axis) and
 * 6 psi buffers.  metrics[] are (z, y, x) inverse distances (rank 0 or empty
 * for isotropic grids).  outputs[] alias the targets and psi. */
typedef struct bzh_launch {
  int32_t abi_version;
  int32_t flags;
  int32_t phase;
  int32_t nterms;
  int32_t metric_kind;
  float dt;
  float resolution;
  float inv_resolution;
  float dt_over_eps;
  float dt_over_mu;
  int32_t metallic_edges; /* bits: front,back,bottom,top,left,right = 0..5 */
  int32_t uniform_cpml_thickness;
  bzh_buffer inputs[37];
  bzh_buffer metrics[3];
  bzh_buffer outputs[9];
} bzh_launch;

/* One (timing, component) slab-source group (cuda/src/launch.h:37-45).
 * coefficients f32[n, sz, sy, sx]; waveforms f32[n, T]; starts s32[n, 3]. */
typedef struct bzh_source_group {
  bzh_buffer coefficients;
  bzh_buffer waveforms;
  bzh_buffer starts;
  bzh_buffer current_step; /* s32[] device scalar: step index of the first step */
  int32_t component;       /* 0..2 within the phase's field triple */
  int32_t timing;          /* 0 pre_e, 1 h, 2 e */
  int32_t coincident;      /* all slabs share one origin: serial per-cell sum */
} bzh_source_group;

/* Packed DFT monitors (cuda/src/launch.h:47-61).  dft_re/dft_im/dft_weight are
 * f64 (or f32) arenas; time is an f64 (or f32) device scalar holding t at the
 * start of the call; current_step an s32 device scalar. */
typedef struct bzh_dft_group {
  bzh_buffer indices;         /* s32[M, 6, Pmax, K] flat indices into each component */
  bzh_buffer weights;         /* f32[M, 6, Pmax, K] */
  bzh_buffer frequencies;     /* f32 or f64 [M, Fmax] in Hz */
  bzh_buffer component_masks; /* f32[M, 6] */
  bzh_buffer counts;          /* s32[M, 5]: F, P, interval, value_offset, weight_offset */
  bzh_buffer codes;           /* s32[M, 2]: window (0 rect, 1 Hann), normalization */
  bzh_buffer windows;         /* f32[M, 3]: t_start, t_end, length_unit */
  bzh_buffer dft_re;
  bzh_buffer dft_im;
  bzh_buffer dft_weight;
  bzh_buffer time;
  bzh_buffer current_step;
  int32_t monitor_count;
} bzh_dft_group;

/* A complete multi-step program (cuda/src/launch.h:63-77).  With
 * field_bank_count == 2 the *_ab launches read bank A and write bank B and the
 * *_ba launches the reverse; the result of an odd step count is in bank B. */
typedef struct bzh_program_launch {
  bzh_launch h_ab;
  bzh_launch e_ab;
  bzh_launch h_ba;
  bzh_launch e_ba;
  const bzh_source_group* source_groups;
  int32_t source_group_count;
  const bzh_dft_group* monitors;
  int32_t field_bank_count;
  int32_t nsteps;
} bzh_program_launch;

/* Attributes of the positional entry points (ffi_handler.cc:722-765). */
typedef struct bzh_phase_attrs {
  int32_t abi_version;
  int32_t flags;
  int32_t phase;
  int32_t nterms;
  float dt;
  float resolution;
  int32_t boundary_code; /* metallic edge mask | (uniform CPML cells << 8) */
  int32_t metric_kind;   /* 0 isotropic, 1 axis-uniform, 2 rectilinear */
} bzh_phase_attrs;

typedef struct bzh_program_attrs {
  int32_t abi_version;
  int32_t flags;
  int32_t nsteps;
  float dt;
  float resolution;
  int32_t boundary_code;
  int32_t metric_kind;
  int32_t program_layout;
  int32_t cpml_enabled;
  int32_t monitor_count;
  int32_t coincident_source_group_mask;
} bzh_program_attrs;

/* ---- module queries (cuda/src/extension.cc:10-23) ------------------------- */
int32_t bzh_abi_version(void);
const char* bzh_component_version(void);
/* Comma-separated target names ("beamz_hip_streamed,beamz_hip_program"),
 * the counterparts of registrations() (extension.cc:14-21). */
const char* bzh_targets(void);
/* Message of the last failing call on this thread ("" if none). */
const char* bzh_last_error(void);

/* ---- descriptor-level launches ------------------------------------------- */
/* cuda/src/update.cu:919-971 (BeamzLaunchStreamed). */
int32_t bzh_launch_phase(void* stream, const bzh_launch* launch);
/* cuda/src/program.cu:339-354 (BeamzLaunchProgram). */
int32_t bzh_launch_program(void* stream, const bzh_program_launch* program);

/* ---- positional launches (what the reference's FFI handlers decode) ------- */
/* beamz_cuda_streamed, ffi_handler.cc:290-334 and 722-735: args in the order
 * of beamz/simulation/cuda/runtime.py:131-181 (targets, sources, materials,
 * [metadata, 18 term arrays, 6 psi], 3 metrics).  Outputs alias args. */
int32_t bzh_streamed(void* stream, const bzh_buffer* args, int32_t nargs,
                     const bzh_phase_attrs* attrs);
/* beamz_cuda_program, ffi_handler.cc:336-708 and 737-765: args in the order of
 * program layout attrs->program_layout (runtime.py:284-750).  Outputs alias
 * args (field/psi workspaces included). */
int32_t bzh_program(void* stream, const bzh_buffer* args, int32_t nargs,
                    const bzh_program_attrs* attrs);

#ifdef __cplusplus
}
#endif

#endif /* BEAMZ_HIP_H_ */
