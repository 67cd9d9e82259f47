/* dpmd_h5 -- optional HDF5 I/O in the reference's file layout (libdpmd_h5.so).
 *
 * Built only where an HDF5 C library is present; it links libdpmd_amd.so for the particle
 * loader.  Files follow the reference's layout:
 *
 *   <dir>/meta.h5        /static  class_name, neighbor_method (UTF-8 strings), n_particles,
 *                                 n_systems, n_vertices (int scalars), static fields
 *                        /init, /final   state fields;  /restart  step (int) + fields
 *   <dir>/trajectory.h5  /timestep (T, 1) int;  /<field> (T, N) or (T, N, 2), chunked along T;
 *                        /ragged/<field>/step_<k> for frames whose length changed
 *
 * with per-particle fields as (N,) or (N, 2) datasets of native double / int, in canonical
 * particle order (utils/h5_io.hpp:234-364, utils/output_manager.hpp:169-311, :704-1040).
 * All functions return DPMD_OK (0) or a nonzero status; dpmd_h5_last_error() gives the
 * message of the calling thread's last failure. */
#ifndef DPMD_H5_H
#define DPMD_H5_H

#include <stdint.h>

#include "dpmd.h"

#ifdef __cplusplus
extern "C" {
#endif

enum dpmd_h5_type { DPMD_H5_F64 = 0, DPMD_H5_I32 = 1 };
enum dpmd_h5_mode { DPMD_H5_CREATE = 0 /* truncate */, DPMD_H5_READ_WRITE = 1 /* create if missing */,
                    DPMD_H5_READ_ONLY = 2 };

const char* dpmd_h5_last_error(void);

/* File handles (H5Fcreate / H5Fopen, as the reference's OutputManager::initialize,
 * output_manager.hpp:169-240). */
int dpmd_h5_open(const char* path, int mode, int64_t* file);
int dpmd_h5_close(int64_t file);
int dpmd_h5_flush(int64_t file);

/* Links: h5_link_exists (h5_io.hpp:88), H5Ldelete, H5Lmove (output_manager.hpp:1000-1011). */
int dpmd_h5_exists(int64_t file, const char* path, int* exists);
int dpmd_h5_delete(int64_t file, const char* path);
int dpmd_h5_move(int64_t file, const char* from, const char* to);

/* Datasets, replacing any existing link; missing parent groups are created.
 * cols = 0 writes (rows,), cols = 2 writes (rows, 2) from row-major x0 y0 x1 y1 ...
 * (write_vector / write_vector_2d / write_scalar, h5_io.hpp:234-364). */
int dpmd_h5_write_array(int64_t file, const char* path, int type, int64_t rows, int cols, const void* data);
int dpmd_h5_write_int(int64_t file, const char* path, int value);
int dpmd_h5_write_string(int64_t file, const char* path, const char* value);

/* Reads (read_scalar / read_vector / read_vector_2d, h5_io.hpp:144-341).  shape: ndims and up
 * to 3 dims; type is the dataset's class mapped to dpmd_h5_type (-1 for others).  read_array
 * converts to the requested type; count must equal the dataset's element count. */
int dpmd_h5_shape(int64_t file, const char* path, int* ndims, int64_t* dims3, int* type);
int dpmd_h5_read_array(int64_t file, const char* path, int type, void* out, int64_t count);
int dpmd_h5_read_int(int64_t file, const char* path, int* value);
int dpmd_h5_read_string(int64_t file, const char* path, char* buf, int64_t capacity);
/* Members of a group, one "<kind> <name>" line each (kind: group, dataset or other), in name
 * order, NUL-terminated; DPMD_ERR_OVERFLOW when capacity is too small. */
int dpmd_h5_list(int64_t file, const char* group, char* buf, int64_t capacity);

/* One trajectory frame (ensure_traj_dataset_locked / write_trajectory,
 * output_manager.hpp:728-932): the dataset at path is
 * (T, n) for cols = 0 or (T, n, 2) for cols = 2, chunked and unlimited along T; it is created
 * on first use and extended to cover `row`.  A frame whose n differs from the dataset's is an
 * error (DPMD_ERR_ARGUMENT); the caller then writes it as a ragged frame. */
int dpmd_h5_write_frame(int64_t file, const char* path, int type, int64_t n, int cols, const void* data, int64_t row);
/* Shrinks every unlimited dataset under the root to `rows` frames (shrink_to_cursor_locked and
 * trim_trajectories_to_restart, output_manager.hpp:791-795, :934-990). */
int dpmd_h5_trim_frames(int64_t file, int64_t rows);

/* BaseParticle::load_from_hdf5(path, location) (base_particle.hpp:120-160): reads
 * <run_dir>/meta.h5, allocates the context from /static, loads the static fields and then
 * those present in /<location> ("init", "final" or "restart"), syncs and builds neighbors.
 * Fields the context derives (system_id, system_size, particle_offset, vertex ids) are
 * checked against the file rather than loaded. */
int dpmd_h5_load(dpmd_ctx* ctx, const char* run_dir, const char* location);

#ifdef __cplusplus
}
#endif

#endif
