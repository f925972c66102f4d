/*
 * kwave_io.h -- the file-level contract of the k-Wave external binaries.
 *
 * The MATLAB front end writes an HDF5 input file (kspaceFirstOrder_saveToDisk.m
 * via writeMatrix.m / writeFlags.m / writeGrid.m / writeAttributes.m), runs
 *   ./<binary> -i <in.h5> -o <out.h5> <options>      (kspaceFirstOrder3DC.m:544)
 * and reads the output datasets back with h5read (kspaceFirstOrder3DC.m:568-778).
 * kw_run_file is that binary's work as a function; kspaceFirstOrder-HIP is a
 * thin executable over it.  The kwio_* calls read and write datasets in the
 * same layout (float32 "float" / uint64 "long" with data_type and domain_type
 * attributes, file format version 1.2: private/getH5Literals.m:26-55) so
 * harnesses without an HDF5 binding can build inputs and read outputs.
 */
#ifndef KWAVE_IO_H
#define KWAVE_IO_H

#include <stddef.h>
#include <stdint.h>

#include "kwave_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Options of the binary's command line (kspaceFirstOrder3DC.m:308-524). */
typedef struct kw_run_options {
  uint64_t record_flags;       /* kw_record_flag bits (--p_raw ... --u_final) */
  uint64_t record_start_index; /* -s, 1-based */
  int32_t device;              /* -g */
  int32_t verbose;             /* --verbose 0..2 */
  int32_t compression_level;   /* -c, deflate level of the output datasets */
  int32_t reserved;
  /* filled in by kw_run_file: wall-clock seconds of each phase */
  double t_load, t_precompute, t_loop, t_save;
  /* checkpoint-restart (--checkpoint_file, --checkpoint_interval,
   * --checkpoint_timesteps; file_type "checkpoint", getH5Literals.m:48).
   * If checkpoint_file names an existing checkpoint of this input, the run
   * resumes from it.  Once checkpoint_interval seconds (> 0) of time loop or
   * checkpoint_timesteps steps (> 0) have passed in this call, the state is
   * saved there, checkpointed is set to 1 and no output file is written; the
   * same call again continues.  A run that finishes writes the output file
   * and deletes the checkpoint. */
  const char* checkpoint_file;
  double checkpoint_interval;
  uint64_t checkpoint_timesteps;
  int32_t checkpointed; /* out */
  int32_t reserved2;
} kw_run_options;

/* Reads the input file, runs the simulation, writes the output file.
 * Replaces: system('./kspaceFirstOrder-CUDA -i in -o out ...') at
 * kspaceFirstOrder3DC.m:544 (and kspaceFirstOrder2DG.m:76, 3DG.m:69). */
kw_status kw_run_file(const char* input_path, const char* output_path, kw_run_options* opts);

/* ---- dataset access ------------------------------------------------------ */
typedef struct kwio_file kwio_file;

enum kwio_dtype { KWIO_FLOAT32 = 0, KWIO_UINT64 = 1 };

/* mode 0: open read-only, 1: create/truncate for writing.
 * Replaces: h5create (writeMatrix.m:189-204) / h5read (3DC.m:568). */
kw_status kwio_open(const char* path, int mode, kwio_file** out);
kw_status kwio_close(kwio_file* f);

/* Writes a real dataset with C-order dims (d0, d1, d2) == MATLAB [d2, d1, d0],
 * chunked as writeMatrix.m:56-110 does, with data_type/domain_type attributes.
 * Replaces: writeMatrix(filename, matrix, name, compression) (writeMatrix.m:1). */
kw_status kwio_write(kwio_file* f, const char* name, int dtype, const uint64_t dims[3], const void* data,
                     int compression_level);

/* Root attributes of writeAttributes.m:56-70 (file_type "input"/"output"). */
kw_status kwio_write_root_attributes(kwio_file* f, const char* file_type, const char* description);

/* Dataset shape (C order, padded to 3 dims with leading 1s) and type. */
kw_status kwio_info(kwio_file* f, const char* name, int* dtype, uint64_t dims[3]);
/* Rank (up to 8) and C-order dimensions of a dataset (cuboid-corner outputs
 * are 4-D: (Nt, cz, cy, cx) == MATLAB [cx, cy, cz, Nt]). */
kw_status kwio_info_nd(kwio_file* f, const char* name, int* dtype, int* rank, uint64_t dims[8]);
kw_status kwio_read(kwio_file* f, const char* name, void* dst, size_t bytes);
int kwio_exists(kwio_file* f, const char* name);

/* Names of every dataset with its group path ("p", "p/1", ...),
 * '\n'-separated; returns the byte count needed (call with cap 0 to size the
 * buffer). */
size_t kwio_list(kwio_file* f, char* buf, size_t cap);

/* Thread-local message of the last kwio_* / kw_run_file failure. */
const char* kwio_last_error(void);

/* String attribute of the root group ("/") or a dataset. */
kw_status kwio_read_string_attribute(kwio_file* f, const char* object, const char* name, char* buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif /* KWAVE_IO_H */
