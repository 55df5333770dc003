/* libfs2io: HDF5 particle output of the MI355X fs2 path (serial HDF5).
 *
 * Built only where an HDF5 C library is found; everything else of the
 * path lives in libfs2hip.so (include/fs2hip.h), which this library uses
 * through its C ABI.
 */
#ifndef FS2IO_H
#define FS2IO_H

#include "fs2hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* py/hdf5_write.cpp:28-139 / py_hdf5_io.cpp _hdf5_write_particles:
 * writes into `filename` (opened read-write if it exists, else created) the
 * group /parameters {boxsize, omega_m, ax, av (f64), nc (i32) =
 * round(np_total^(1/3))} and, for each letter of `var` (a subsequence of
 * "ixvcf12" in this order): i -> "id" (np,) u64; x -> "x", v -> "v" (raw),
 * c -> "v" (COLA velocity), f -> "f", 1 -> "dx1", 2 -> "dx2", each (np, 3)
 * f32 little endian.  FS2_ERR_IO if the file or group cannot be created,
 * FS2_ERR_VALUE for letters out of order or unknown. */
int fs2_write_hdf5(fs2_particles* p, const char* filename, const char* var);

#ifdef __cplusplus
}
#endif

#endif /* FS2IO_H */
