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
 * materials (decay x,y,z then source x,y,z; rank-0 = uniform; for the E
 * phase of a lossless program also the packed ID form of
 * beamz/simulation/compile.py:146-187: the decay slot an f32 table of at most
 * 256 values, the source slot s32 words of four 8-bit cell IDs), then when
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
/* Diagnostic with no counterpart in the reference: 1 when the last program
 * call on this thread ran the fused two-bank step kernels, 0 when it ran the
 * per-phase schedule (e.g. more than eight H-timed source slabs, per-cell H
 * coefficients, a field array of 4 GiB or more, or a bank over 4 GiB in
 * 2-byte storage). */
int32_t bzh_last_program_fused(void);
/* Diagnostic with no counterpart in the reference: the number of two-step
 * (pair) launches in the last program call on this thread.  The pair kernel
 * advances the fused step two steps per pass; a call runs an even number of
 * pairs (so the bank parity of the single-step schedule holds) and single
 * steps for the rest.  BZH_PAIR=0 in the environment disables it. */
int32_t bzh_last_program_pairs(void);
/* Diagnostic with no counterpart in the reference: 1 when the last program
 * call's pair launches captured the monitored cells densely (the monitor
 * layouts' optional trailing argument, host s32[M, 6, 6] boxes of the cells
 * each monitor component gathers, was given and no tile row holds cells of
 * two monitors), 0 when they used the intermediate bank or ran no pairs. */
int32_t bzh_last_program_captured(void);

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
 * args (field/psi workspaces included).  The monitor layouts (5 and 6) take
 * one optional trailing argument with no counterpart in the reference: HOST
 * s32[M, 6, 6] inclusive boxes (z0, z1, y0, y1, x0, x1; -1 when unused) of
 * the cells each monitor gathers per component (Ex, Ey, Ez, Hx, Hy, Hz).
 * With it, two-step launches capture those cells densely instead of writing
 * an intermediate bank; results are the same. */
int32_t bzh_program(void* stream, const bzh_buffer* args, int32_t nargs,
                    const bzh_program_attrs* attrs);

/* No counterpart in the reference.  bzh_program for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, layouts and results, but
 * it calls no hipMalloc, hipFree or graph instantiation.  The buffered DFT's
 * staging ring lives in a fixed device array of BZH_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends; the ring takes as many steps (up to 32)
 * as fit, and a call whose monitors sample more than that array holds in one
 * step fails with an error.  Pairs of steps run only without monitors.  Calls
 * share the array per GPU: calls that overlap in time on one GPU race. */
#define BZH_NOALLOC_SCRATCH_BYTES (64u << 20)
int32_t bzh_program_noalloc(void* stream, const bzh_buffer* args, int32_t nargs,
                            const bzh_program_attrs* attrs);

/* ---- batched 2D simulations ------------------------------------------------
 * No counterpart in the reference, which runs one simulation per call (its
 * FFI is vmapped sequentially, cuda/runtime.py:85).  Many independent TMxy
 * (Ez, Hx, Hy) simulations of uneven sizes advance together: each step is
 * one source launch (the pre-E slabs) and one march launch whose waves take
 * the work items (62 columns by tile_rows rows) of every simulation
 * round-robin and march up their rows; a call of nsteps steps is captured
 * once as a hipGraph and replayed.  Two field banks per simulation: bank 0 is
 * read by the first step, and after an odd step count the result is in bank
 * 1 (the caller swaps the roles).  All buffers f32. */
typedef struct bzh_sim2d {
  bzh_buffer ez[2];        /* (ny+1, nx+1) */
  bzh_buffer hx[2];        /* (ny, nx+1) */
  bzh_buffer hy[2];        /* (ny+1, nx) */
  bzh_buffer e_decay;      /* scalar */
  bzh_buffer e_source;     /* scalar or (ny+1, nx+1) */
  bzh_buffer h_decay;      /* scalar */
  bzh_buffer h_source;     /* scalar */
  bzh_buffer cpml_table;   /* [4, 3, 2L]: Hx_y, Hy_x, Ez_x, Ez_y x (a, b, 1/kappa); empty without CPML */
  bzh_buffer psi[2][4];    /* per bank: Hx_y (2L, nx+1), Hy_x (ny+1, 2L), Ez_x (ny+1, 2L), Ez_y (2L, nx+1) */
  bzh_buffer src_coeff[4]; /* pre-E Ez slab coefficients (sy, sx) */
  bzh_buffer src_wave[4];  /* waveform [T], sampled at clip(step, 0, T-1) */
  int32_t src_start[4][2]; /* slab origin (y, x) in the Ez support */
  int32_t nsrc;            /* 0..4 */
  int32_t cpml_cells;      /* CPML thickness on all four edges (0: none) */
  int32_t metallic;        /* PEC edge bits: bottom 4, top 8, left 16, right 32 */
  float resolution;        /* cell size (the inverse enters the curls) */
} bzh_sim2d;

typedef struct bzh_batch_attrs {
  int32_t abi_version;
  int32_t flags;           /* BZH_FLAG_GRAPH_CACHE */
  int32_t nsteps;
  int32_t tile_rows;       /* rows per work item (62 columns wide): 8, 16, 32 or 64 */
  int32_t workgroups;      /* persistent grid size; 0 picks 4 per compute unit */
} bzh_batch_attrs;

/* current_step: s32 device scalar, the step index of the first step. */
int32_t bzh_batch2d(void* stream, const bzh_sim2d* sims, int32_t nsims,
                    const bzh_buffer* current_step, const bzh_batch_attrs* attrs);

/* ---- adjoint (reverse-time) TMxy steps ------------------------------------
 * No counterpart in the reference, whose gradient is a host-side overlap of
 * two forward runs (beamz/optimization/topology.py:314-342).  These entry
 * points run the exact discrete adjoint of the TMxy step (CPML and PEC
 * included) backward in time and accumulate d(objective)/d(E source
 * coefficient) over an interior design box.  The forward problem is
 * described by the same coefficients, CPML table, edge mask and resolution
 * as the forward program.  All buffers f32 except grad (f64). */
typedef struct bzh_adjoint2d {
  bzh_buffer e_decay;        /* scalar */
  bzh_buffer e_source;       /* scalar or (ny+1, nx+1) */
  bzh_buffer h_decay;        /* scalar */
  bzh_buffer h_source;       /* scalar */
  bzh_buffer cpml_table;     /* [4, 3, 2L] or empty */
  int32_t cpml_cells;
  int32_t metallic;          /* PEC edge bits: bottom 4, top 8, left 16, right 32 */
  float resolution;
  bzh_buffer a_ez, a_hx, a_hy;  /* adjoint state: shapes of Ez, Hx, Hy */
  bzh_buffer a_psi[4];          /* adjoint CPML memories: Hx_y, Hy_x, Ez_x, Ez_y */
  bzh_buffer work[4];           /* scratch with the shapes of Ez, Ez, Hy, Hx */
  bzh_buffer source_profile;    /* Ez-shaped adjoint source weights W, or empty */
  bzh_buffer source_time;       /* [N]: the adjoint source of forward step n is c_n W */
  int32_t design[4];            /* Ez-node box [y0, y1) x [x0, x1), inside CPML and PEC */
  bzh_buffer record;            /* [K, y1-y0, x1-x0]: forward curl per step of a segment */
  bzh_buffer grad;              /* f64 [y1-y0, x1-x0]: += sum_n g_n curl_n */
} bzh_adjoint2d;

/* Record the forward curl of the current (Hx, Hy) over the design box into
 * record[slot]; call after each forward step of a recomputed segment. */
int32_t bzh_adjoint2d_record(void* stream, const bzh_adjoint2d* adj, const bzh_buffer* hx,
                             const bzh_buffer* hy, int32_t slot);
/* Reverse forward steps n_last, n_last-1, ..., n_last-count+1, whose curls
 * sit in record slots slot_last, slot_last-1, ...  flags: BZH_FLAG_GRAPH_CACHE. */
int32_t bzh_adjoint2d_reverse(void* stream, const bzh_adjoint2d* adj, int32_t n_last,
                              int32_t count, int32_t slot_last, int32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* BEAMZ_HIP_H_ */
