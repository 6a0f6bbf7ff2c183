/* mara_checkpoint — Mara checkpoint files (HDF5) for the device path, C ABI.
 *
 * Built as its own library (libmara_h5.so, linked against the system's HDF5 and against
 * libmaragpu.so), so that the solver library carries no HDF5 dependency.
 *
 * The layout is the one Mara's Lua drivers write and read:
 *   /prim/<field>       one 3-D double dataset per primitive field (rho, pre, vx, vy, vz, Bx,
 *                       By, Bz), shape (Nx, Ny, Nz), interior zones only, contiguous
 *                       (DataManagerHDF5:write_sequential / read, modules/unigrid.lua:132-212;
 *                       the hyperslabs of :243-292)
 *   /<name>             a scalar fixed-length string dataset of exactly the text's length
 *                       (Indexable:__newindex__ for strings, lua-hdf5/LuaHDF5.lua:174-191):
 *                       drvtrb's runargs, measure, program, version and driving
 *                       (ctf-core/drvtrb.lua:509-549); run-mara's measure_log, status, git_sha
 *                       (ctf-core/run-mara.lua:180-204)
 *   /status/<key>       a scalar double (numbers) or string dataset per status entry
 *                       (ctf-core/drvtrb.lua:544-548, LuaHDF5.lua:192-201)
 *
 * Return codes are those of maragpu.h (0 success, < 0 error). A failed HDF5 call returns
 * MARAGPU_E_IO; mara_h5_last_error() names the step that failed.
 */
#ifndef MARA_CHECKPOINT_H
#define MARA_CHECKPOINT_H

#include <stddef.h>

#include "maragpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { MARA_H5_TRUNCATE = 0, MARA_H5_UPDATE = 1 };       /* file modes 'w' and 'r+' */
enum { MARA_H5_MISSING = 0, MARA_H5_NUMBER = 1, MARA_H5_STRING = 2, MARA_H5_OTHER = 3 };

const char *mara_h5_last_error(void);

/* DataManagerHDF5:write (modules/unigrid.lua:85-177) of a host primitive array in Mara's AoS
 * layout [i][j][k][q] with Ng ghosts: the interior of field q goes to <group>/<names[q]>,
 * shape (N0, N1, N2). mode MARA_H5_TRUNCATE creates the file, as the first writer does;
 * MARA_H5_UPDATE opens an existing one and replaces the datasets. */
int mara_h5_write_prim(const char *path, const char *group, const double *P, const int N[3], int Ng,
                       int Nq, const char *const *names, int mode);

/* DataManagerHDF5:read (modules/unigrid.lua:181-212): the datasets <group>/<names[q]> into the
 * interior of field q of the AoS array P; ghosts are left as they are. Each dataset must
 * have shape (N0, N1, N2). */
int mara_h5_read_prim(const char *path, const char *group, double *P, const int N[3], int Ng, int Nq,
                      const char *const *names);

/* The extent of a dataset, as MyMara:initialize_solver reads prim/rho's
 * (ctf-core/run-mara.lua:103-107): *ndim and dims[0 .. *ndim - 1], at most 8 dimensions. */
int mara_h5_extent(const char *path, const char *dset, int *ndim, long long dims[8]);

/* Scalar datasets under group ("" or NULL: the root), created if missing, replaced if
 * present (clobber_if_H5Lexists, LuaHDF5.lua:75-83). Strings are written as a fixed-length
 * C string type of exactly n bytes. The file must exist. */
int mara_h5_write_string(const char *path, const char *group, const char *name, const char *bytes, size_t n);
int mara_h5_write_number(const char *path, const char *group, const char *name, double value);

/* What group/name holds (MARA_H5_*). */
int mara_h5_kind(const char *path, const char *group, const char *name);
/* A string dataset's bytes: *n = its length; copies min(n, cap) bytes. */
int mara_h5_read_string(const char *path, const char *group, const char *name, char *buf, size_t cap,
                        size_t *n);
int mara_h5_read_number(const char *path, const char *group, const char *name, double *value);
/* The link names of a group, '\n'-separated, in HDF5's name order; *n = bytes needed. */
int mara_h5_keys(const char *path, const char *group, char *buf, size_t cap, size_t *n);

/* The device forms: maragpu_download_field / maragpu_upload_field per field, to and from
 * <group>/<rho .. Bz>. Reading sets the committed state, periodic ghosts included. */
int maragpu_checkpoint_write(maragpu_ctx *ctx, const char *path, const char *group, int mode);
int maragpu_checkpoint_read(maragpu_ctx *ctx, const char *path, const char *group);

/* JSON strings as the drivers' json module (dkjson 2.1, modules/json.lua) writes and reads
 * them, for byte strings such as the driving blob (ctf-core/drvtrb.lua:542,740).
 * quote: the pure-Lua quotestring (:289-304), quotes included; *n = bytes needed.
 * unquote: scanstring (:529-590) on a quoted string, from its opening quote. */
int mara_json_quote(const char *in, size_t len, char *out, size_t cap, size_t *n);
int mara_json_unquote(const char *in, size_t len, char *out, size_t cap, size_t *n);

#ifdef __cplusplus
}
#endif

#endif
