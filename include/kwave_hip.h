/*
 * kwave_hip.h -- C ABI of the MI355X (gfx950) k-space pseudospectral solver.
 *
 * The reference reaches its accelerated time loop through a process
 * boundary: kspaceFirstOrder3DC.m:536-566 runs
 *   system('... ./kspaceFirstOrder-{OMP,CUDA} -i in.h5 -o out.h5 <flags>')
 * after kspaceFirstOrder_saveToDisk.m writes the input file, and reads the
 * results back with h5read (kspaceFirstOrder3DC.m:568-778).  This header is
 * the in-process form of that contract: a simulation is described by the
 * datasets of that input file (as plain host arrays), run, and its outputs
 * fetched by the dataset names of the output file.  The file-level form of
 * the same contract (HDF5 in, HDF5 out) is in kwave_io.h and the
 * kspaceFirstOrder-HIP executable built on it.
 *
 * Conventions: status codes, never exceptions; host buffers are caller-owned
 * and copied in; device buffers are library-owned; handles are independent
 * (one HIP stream each); kw_last_error() returns a thread-local message.
 * Arrays are in HDF5 C order: a grid field is (Nz, Ny, Nx) with x fastest,
 * which is MATLAB's column-major [Nx, Ny, Nz].
 */
#ifndef KWAVE_HIP_H
#define KWAVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KW_ABI_VERSION 1

typedef enum kw_status {
  KW_OK = 0,
  KW_ERR_INVALID = 1,     /* malformed description or argument */
  KW_ERR_UNSUPPORTED = 2, /* valid k-Wave input this build does not run */
  KW_ERR_DEVICE = 3,      /* HIP runtime error */
  KW_ERR_OOM = 4,         /* device allocation failed */
  KW_ERR_IO = 5           /* file I/O (kwave_io.h) */
} kw_status;

/* A float32 dataset that is a scalar (count == 1) or a full grid field
 * (count == Nx*Ny*Nz).  Mirrors writeMatrix.m's scalar-or-matrix datasets
 * (c0, rho0, rho0_sgx/y/z, BonA, alpha_coeff: saveToDisk.m:62-73, 155-215). */
typedef struct kw_real_array {
  const float* data;
  uint64_t count;
} kw_real_array;

/* Output selection: the binary's --p_raw ... --u_final flags
 * (kspaceFirstOrder3DC.m:466-519). */
enum kw_record_flag {
  KW_REC_P = 1u << 0,
  KW_REC_P_MAX = 1u << 1,
  KW_REC_P_MIN = 1u << 2,
  KW_REC_P_RMS = 1u << 3,
  KW_REC_P_MAX_ALL = 1u << 4,
  KW_REC_P_MIN_ALL = 1u << 5,
  KW_REC_P_FINAL = 1u << 6,
  KW_REC_U = 1u << 7,
  KW_REC_U_MAX = 1u << 8,
  KW_REC_U_MIN = 1u << 9,
  KW_REC_U_RMS = 1u << 10,
  KW_REC_U_MAX_ALL = 1u << 11,
  KW_REC_U_MIN_ALL = 1u << 12,
  KW_REC_U_FINAL = 1u << 13,
  /* --u_non_staggered_raw: u shifted to the unstaggered grid points,
   * real(ifft(x_shift_neg .* fft(u, [], axis))) (extractSensorData.m:39-55) */
  KW_REC_U_NON_STAGGERED = 1u << 14
};

/* Output datasets (names of the output file read at kspaceFirstOrder3DC.m:568-778).
 * Velocity outputs come in x, y, z order: KW_OUT_UX + axis, etc. */
typedef enum kw_output_id {
  KW_OUT_P = 0,         /* (Nt_rec, Ns)  float32 */
  KW_OUT_P_MAX = 1,     /* (Ns) */
  KW_OUT_P_MIN = 2,
  KW_OUT_P_RMS = 3,
  KW_OUT_P_MAX_ALL = 4, /* (Nz, Ny, Nx) */
  KW_OUT_P_MIN_ALL = 5,
  KW_OUT_P_FINAL = 6,
  KW_OUT_UX = 7, /* +axis: (Nt_rec, Ns) */
  KW_OUT_UX_MAX = 10,
  KW_OUT_UX_MIN = 13,
  KW_OUT_UX_RMS = 16,
  KW_OUT_UX_MAX_ALL = 19,
  KW_OUT_UX_MIN_ALL = 22,
  KW_OUT_UX_FINAL = 25,
  KW_OUT_UX_NON_STAGGERED = 28, /* +axis: (Nt_rec, Ns) */
  KW_OUT_COUNT = 31
} kw_output_id;

/* One simulation: the dataset set of the HDF5 input file
 * (kspaceFirstOrder_saveToDisk.m:44-414, writeFlags.m, writeGrid.m).
 * 2-D inputs have Nz == 1 and no z datasets (saveToDisk.m:420-427).
 * Indices are 1-based linear indices as written by the front end (find()). */
typedef struct kw_sim_desc {
  uint32_t abi_version; /* KW_ABI_VERSION */
  int32_t device;       /* HIP device ordinal (the binary's -g) */
  uint64_t Nx, Ny, Nz, Nt;
  double dx, dy, dz, dt, c_ref;
  uint64_t pml_x_size, pml_y_size, pml_z_size;
  double pml_x_alpha, pml_y_alpha, pml_z_alpha;
  kw_real_array c0, rho0, rho0_sgx, rho0_sgy, rho0_sgz;
  uint64_t nonlinear_flag;
  kw_real_array BonA;
  uint64_t absorbing_flag; /* 0 lossless, 1 power law, 2 stokes */
  kw_real_array alpha_coeff; /* dB/(MHz^y cm), converted by db2neper.m:50 inside */
  double alpha_power;
  /* initial pressure (p0_source_flag / p0_source_input); NULL if none */
  const float* p0_source_input;
  /* pressure source: flag = number of signal samples (0 = none) */
  uint64_t p_source_flag, p_source_mode, p_source_many, p_source_npoints;
  const uint64_t* p_source_index; /* p_source_npoints */
  const float* p_source_input;    /* (p_source_flag, many ? npoints : 1) */
  /* velocity sources share u_source_index */
  uint64_t ux_source_flag, uy_source_flag, uz_source_flag;
  uint64_t u_source_mode, u_source_many, u_source_npoints;
  const uint64_t* u_source_index;
  const float* ux_source_input;
  const float* uy_source_input;
  const float* uz_source_input;
  /* sensor (sensor_mask_type 0: binary-mask indices) */
  uint64_t sensor_npoints;
  const uint64_t* sensor_mask_index;
  /* outputs */
  uint64_t record_flags;       /* kw_record_flag bits */
  uint64_t record_start_index; /* the binary's -s, 1-based (0 treated as 1) */
  /* transducer source (kspaceFirstOrder3D.m:889-902; 3-D only): while
   * t <= transducer_source_flag, ux(u_source_index(j)) +=
   * transducer_source_input(delay_mask(j) + t - 1) (1-based), the signal
   * pre-scaled by 2 c0 dt / dx (scaleSourceTerms.m:376-386).  Shares
   * u_source_index / u_source_npoints and excludes ux/uy/uz sources. */
  uint64_t transducer_source_flag;
  uint64_t transducer_source_length;    /* samples in transducer_source_input */
  const float* transducer_source_input;
  const uint64_t* delay_mask;           /* u_source_npoints, 1-based (inputChecking.m:1213) */
  /* axisymmetric_flag (kspaceFirstOrderASC.m, saveToDisk.m:56,125): a 2-D grid
   * with x axial and y radial from the axis, run with the WSWA radial
   * symmetry of kspaceFirstOrderAS.m (inputChecking.m:1400-1411); the radial
   * PML sits at the outer edge only (getPML.m:73-77); lossless or stokes
   * absorption (inputChecking.m:165-200). */
  uint64_t axisymmetric_flag;
} kw_sim_desc;

typedef struct kw_sim kw_sim;

/* Copies the description to the device, runs the FP64 operator precompute
 * (kspaceFirstOrder3D.m:583-659) and allocates state.
 * Replaces: the binary's start-up after system() at kspaceFirstOrder3DC.m:544. */
kw_status kw_sim_create(const kw_sim_desc* desc, kw_sim** out);

/* Advances nsteps time steps (clamped to Nt) on `stream` (a hipStream_t;
 * NULL uses the handle's own stream).  Asynchronous with respect to the host.
 * Replaces: the time loop inside the binary (kspaceFirstOrder3D.m:767-1206). */
kw_status kw_sim_run(kw_sim* sim, uint64_t nsteps, void* stream);

/* Blocks until all work queued on the handle's streams has finished. */
kw_status kw_sim_sync(kw_sim* sim);

/* Number of time steps completed so far. */
uint64_t kw_sim_steps_done(const kw_sim* sim);

/* The handle's own stream (a hipStream_t), used by kw_sim_run(..., NULL). */
void* kw_sim_stream(kw_sim* sim);

/* Byte size of an output (0 if not recorded). */
size_t kw_sim_output_bytes(const kw_sim* sim, int output_id);

/* Copies an output to host memory (synchronises first).
 * Replaces: h5read of the output file (kspaceFirstOrder3DC.m:568-778). */
kw_status kw_sim_fetch(kw_sim* sim, int output_id, void* host_dst, size_t bytes);

/* Frees device memory and the handle. */
kw_status kw_sim_destroy(kw_sim* sim);

/* Checkpoint-restart (the binaries' --checkpoint_file, file_type "checkpoint",
 * private/getH5Literals.m:48; release notes: "command line inputs for
 * checkpoint-restart").  The state is the time index, the fields carried
 * between steps and every running output; a handle created from the same
 * description restores it with kw_sim_checkpoint_load and continues exactly as
 * the uninterrupted run would. */
size_t kw_sim_checkpoint_bytes(const kw_sim* sim);
kw_status kw_sim_checkpoint_save(kw_sim* sim, void* host_dst, size_t bytes);
kw_status kw_sim_checkpoint_load(kw_sim* sim, const void* host_src, size_t bytes);

/* Thread-local description of the last failure on this thread. */
const char* kw_last_error(void);

/* Per-kernel timing: when enabled, every launch of kw_sim_run is bracketed
 * by HIP events on its stream and its duration accumulated per kernel kind.
 * kw_sim_profile_get returns the number of kinds and fills up to `cap`
 * entries of names (static strings), total milliseconds and launch counts. */
kw_status kw_sim_set_profiling(kw_sim* sim, int enable);
int kw_sim_profile_get(kw_sim* sim, const char** names, double* total_ms, uint64_t* launches, int cap);

/* Launch-geometry autotuner: times each candidate set of workgroup sizes of
 * the row, pencil and plane kernels and each radix plan of the pencils
 * (coordinate descent from the defaults) over `trial_steps` steps of this
 * simulation, keeps the fastest (cached per grid shape and physics branch for
 * the rest of the process) and resets the simulation to step 0.  The full
 * chosen configuration is reported by kw_sim_launch_config.  Call before
 * kw_sim_run.  The chosen sizes are returned (0: kernel kind not used;
 * pencil_nt is the y pass's, or the last axis's when there is no y pass).
 * No counterpart in the reference (its binaries have fixed launch shapes). */
kw_status kw_sim_autotune(kw_sim* sim, int trial_steps, int* row_nt, int* pencil_nt, int* plane_nt);

/* The launch configuration a handle runs with (after kw_sim_autotune, the
 * tuned one): workgroup sizes of the x-row kernels, the y pass, the
 * last-axis pass and the plane pass (0: kind not used by this grid), and the
 * radix plans of the two pencil passes (0: radices 16, 8, 4, 2, 3, 5, 7
 * greedily; 1: radix 8 at most).  No counterpart in the reference. */
typedef struct kw_launch_config {
  int32_t row_nt, ypass_nt, last_nt, plane_nt;
  int32_t ypass_plan, last_plan;
  int32_t kx_major; /* 1: kx-major spectra and the plane pass */
  int32_t row_wg_target; /* x-row kernels: workgroups per launch the row pairs per workgroup aim at */
} kw_launch_config;
kw_status kw_sim_launch_config(const kw_sim* sim, kw_launch_config* config);

/* Device memory (bytes) a description would need at its peak during
 * kw_sim_create and after it; 0 on invalid input.  Follows kw_sim_create's
 * allocation sequence.  Used for HBM-footprint admission by the serving loop. */
uint64_t kw_sim_footprint_bytes(const kw_sim_desc* desc);

/* Device bytes a handle holds now and held at most since kw_sim_create
 * (the figures kw_sim_footprint_bytes predicts).  No counterpart in the
 * reference. */
kw_status kw_sim_memory(const kw_sim* sim, uint64_t* allocated_bytes, uint64_t* peak_bytes);

/* Diagnostic with no counterpart in the reference: plain device copies of
 * `bytes` bytes (one untimed, then ten timed with HIP events; *ms is the mean
 * per copy).  Gives the FETCH_SIZE/WRITE_SIZE counters a kernel of known
 * traffic for calibration, and the copy bandwidth a practical ceiling. */
kw_status kw_diag_copy(uint64_t bytes, int device, double* ms);

/* The same with a choice of copy kernel, for the measured copy ceiling:
 * variant 0 is kw_diag_copy's kernel (one 16-byte load per thread per
 * iteration); 1 and 2 keep 4 and 8 16-byte loads in flight per thread; 3 and
 * 4 are 1 and 2 with non-temporal loads and stores.  blocks: number of
 * 256-thread workgroups (0: 2048). */
kw_status kw_diag_copy_ex(uint64_t bytes, int device, int variant, int blocks, double* ms);

/* ---------------------------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one whole simulation -- kw_sim_create, kw_sim_run(Nt),
 * kw_sim_fetch of each listed output, kw_sim_destroy -- with every device
 * buffer (operators, state, work spectra, recorded outputs) inside a fixed
 * scratch region that the library holds statically, KW_NOALLOC_SCRATCH_BYTES
 * on each GPU it runs on, until the process ends.  It calls no hipMalloc,
 * hipFree or other allocating routine; the work runs on `stream` (a
 * hipStream_t; NULL is the null stream) and the call returns when it is done.
 * Results equal kw_sim_create + kw_sim_run + kw_sim_fetch bit for bit.
 * A description needing more than the region (the arena figure of its
 * allocation sequence, with no early releases) fails with KW_ERR_UNSUPPORTED
 * before any device work.  Calls on one GPU share its region: within a
 * process they are serialised; the region keeps the last call's contents
 * (everything that must start at zero is cleared by the call itself).
 * ------------------------------------------------------------------------- */
#define KW_NOALLOC_SCRATCH_BYTES (128ull << 20)

typedef struct kw_output_buffer {
  int32_t output_id; /* kw_output_id */
  int32_t reserved;
  void* host_dst;    /* kw_sim_output_bytes of that output */
  size_t bytes;
} kw_output_buffer;

kw_status kw_sim_run_noalloc(const kw_sim_desc* desc, void* stream, const kw_output_buffer* outputs,
                             int noutputs);

/* ---------------------------------------------------------------------------
 * Serving loop.  The reference runs one simulation per process launch
 * (kspaceFirstOrder3DC.m:544, one system() call per kspaceFirstOrder3DC
 * call; kWaveTester.m:1936-1950 loops over many such calls).  A server keeps
 * many simulations resident on one device instead: requests are copied in at
 * submit, admitted in arrival order when their device footprint
 * (kw_sim_footprint_bytes) fits the HBM budget and a slot is free, advanced
 * concurrently (each on its own stream), and their outputs are copied to
 * host memory when they finish, after which their device memory is freed and
 * the next request can be admitted.  A request's results are the same as
 * those of kw_sim_create + kw_sim_run(Nt) + kw_sim_fetch on its own, bit for
 * bit -- except with kw_server_config.autotune set: a tuned request runs the
 * workgroup sizes kw_sim_autotune picked for its shape (the same for every
 * request of that shape in the process), which may round differently from
 * the default sizes (FP32 rounding level).
 * ------------------------------------------------------------------------- */
#define KW_ERR_TIMEOUT ((kw_status)6) /* kw_server_wait: not finished in time */

typedef struct kw_server_config {
  uint32_t abi_version;       /* KW_ABI_VERSION */
  int32_t device;             /* HIP device ordinal */
  uint64_t hbm_budget_bytes;  /* 0: 90% of the device memory free at create */
  uint32_t max_active;        /* simulations resident at once; 0: 8 */
  uint32_t steps_per_quantum; /* time steps queued per submission; 0: 16 */
  uint32_t autotune;          /* 1: kw_sim_autotune each admitted simulation (cached per shape) */
  /* 1: batched mode.  Admitted 3-D simulations of one grid shape, physics
   * branch and launch configuration advance together: each kernel of the time
   * step is launched once for the whole group (one simulation per
   * blockIdx.z, arguments in device tables, per-simulation time index kept
   * on the device), and a quantum of steps is captured in a hipGraph and
   * replayed while the group's membership is unchanged.  Simulations join
   * and leave at quantum boundaries.  Results are the same as alone, bit for
   * bit.  Simulations that cannot be batched (2-D, axisymmetric,
   * u_non_staggered, additive sources with one signal per point, lengths
   * without specialised kernels) run as in the default mode. */
  uint32_t batching;
} kw_server_config;

typedef enum kw_request_state {
  KW_REQ_QUEUED = 0,
  KW_REQ_RUNNING = 1,
  KW_REQ_DONE = 2,
  KW_REQ_FAILED = 3,
  KW_REQ_UNKNOWN = 4 /* no such request (never submitted, or released) */
} kw_request_state;

/* Seconds are measured from kw_server_create on a monotonic clock. */
typedef struct kw_request_info {
  int32_t state;   /* kw_request_state */
  int32_t status;  /* kw_status of a failed request, else KW_OK */
  uint64_t footprint_bytes;
  uint64_t point_updates; /* Nx*Ny*Nz*Nt */
  double t_submit, t_admit, t_done;
} kw_request_info;

typedef struct kw_server_stats {
  uint64_t submitted, completed, failed, queued, active;
  uint64_t hbm_budget_bytes, hbm_reserved_bytes, hbm_peak_reserved_bytes;
  uint64_t point_updates_done;
  double elapsed_s;
} kw_server_stats;

typedef struct kw_server kw_server;

kw_status kw_server_create(const kw_server_config* cfg, kw_server** out);
/* Copies the description and every array it points to (the caller may free
 * them on return) and queues it.  Fails at once (no id) for a description
 * kw_sim_footprint_bytes rejects or whose footprint exceeds the budget. */
kw_status kw_server_submit(kw_server* srv, const kw_sim_desc* desc, uint64_t* id);
int kw_server_poll(kw_server* srv, uint64_t id); /* kw_request_state */
/* Waits until the request is DONE or FAILED; timeout_s < 0 waits forever.
 * KW_OK when finished (check the state), KW_ERR_TIMEOUT otherwise. */
kw_status kw_server_wait(kw_server* srv, uint64_t id, double timeout_s);
kw_status kw_server_request_info(kw_server* srv, uint64_t id, kw_request_info* info);
/* Message of a FAILED request (empty otherwise); valid until release. */
const char* kw_server_request_error(kw_server* srv, uint64_t id);
size_t kw_server_output_bytes(kw_server* srv, uint64_t id, int output_id);
kw_status kw_server_fetch(kw_server* srv, uint64_t id, int output_id, void* host_dst, size_t bytes);
/* Drops a finished request and its host copies of the outputs. */
kw_status kw_server_release(kw_server* srv, uint64_t id);
kw_status kw_server_stats_get(kw_server* srv, kw_server_stats* stats);
/* Fails queued requests, lets running ones finish, joins and frees. */
kw_status kw_server_destroy(kw_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* KWAVE_HIP_H */
