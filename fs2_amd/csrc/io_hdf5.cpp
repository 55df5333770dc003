// HDF5 particle output (py/hdf5_write.cpp:28-246), one rank, serial HDF5.
// Host-side only: the data come from libfs2hip through its C ABI.
#include <hdf5.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fs2_internal.h"
#include "fs2io.h"

namespace {

using fs2::fail;

struct H5Handle {   // closes an HDF5 id with the matching close call
  hid_t id;
  herr_t (*close)(hid_t);
  H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Handle() {
    if (id >= 0) close(id);
  }
  H5Handle(const H5Handle&) = delete;
  H5Handle& operator=(const H5Handle&) = delete;
};

void check(fs2_particles* p, int rc) {
  (void)p;
  if (rc) fail(rc, "%s", fs2_last_error());
}

void write_scalar(hid_t loc, const char* name, hid_t file_type, hid_t mem_type, const void* v) {
  H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
  H5Handle d(H5Dcreate2(loc, name, file_type, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
             H5Dclose);
  if (d.id < 0) fail(FS2_ERR_IO, "unable to create data %s", name);
  if (H5Dwrite(d.id, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
    fail(FS2_ERR_IO, "unable to write data %s", name);
}

// hdf5_write.cpp:231-280: (nrow,) for one column, else (nrow, ncol)
void write_table(hid_t loc, const char* name, hsize_t nrow, hsize_t ncol, hid_t mem_type,
                 hid_t file_type, const void* data) {
  if (nrow == 0) return;   // the reference warns and writes nothing
  const int rank = ncol == 1 ? 1 : 2;
  const hsize_t dims[2] = {nrow, ncol};
  H5Handle sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  H5Handle d(H5Dcreate2(loc, name, file_type, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
             H5Dclose);
  if (d.id < 0) fail(FS2_ERR_IO, "unable to create dataset %s", name);
  if (H5Dwrite(d.id, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    fail(FS2_ERR_IO, "unable to write dataset %s", name);
}

}  // namespace

extern "C" int fs2_write_hdf5(fs2_particles* p, const char* filename, const char* var) {
  return fs2::guard([&] {
    if (!p) fail(FS2_ERR_VALUE, "null particles handle");
    if (!filename || !var) fail(FS2_ERR_VALUE, "null argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    hid_t f = H5Fopen(filename, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) f = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) fail(FS2_ERR_IO, "unable to create HDF5 file, %s", filename);
    H5Handle file(f, H5Fclose);

    uint64_t np = 0, np_total = 0;
    double boxsize = 0, om = 0, ax = 0, av = 0, af = 0;
    check(p, fs2_particles_len(p, &np));
    check(p, fs2_particles_np_total(p, &np_total));
    check(p, fs2_particles_boxsize(p, &boxsize));
    check(p, fs2_cosmology_omega_m(&om));
    check(p, fs2_particles_get_times(p, &ax, &av, &af));
    {   // write_header, hdf5_write.cpp:148-172
      H5Handle g(H5Gcreate2(file.id, "parameters", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                 H5Gclose);
      if (g.id < 0) fail(FS2_ERR_IO, "unable to open group, parameters");
      write_scalar(g.id, "boxsize", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &boxsize);
      write_scalar(g.id, "omega_m", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &om);
      write_scalar(g.id, "ax", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &ax);
      write_scalar(g.id, "av", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &av);
      const int nc = (int)std::round(std::pow((double)np_total, 1.0 / 3.0));
      write_scalar(g.id, "nc", H5T_STD_I32LE, H5T_NATIVE_INT, &nc);
    }

    std::vector<float> buf(3 * np);
    const char* v = var;
    if (*v == 'i') {
      std::vector<uint64_t> id(np);
      check(p, fs2_particles_get_id(p, id.data()));
      write_table(file.id, "id", np, 1, H5T_NATIVE_UINT64, H5T_STD_U64LE, id.data());
      ++v;
    }
    struct Col {
      char letter;
      const char* name;
      int field;   // FS2_FIELD_*, or -1 for the COLA velocity
    };
    const Col cols[] = {{'x', "x", FS2_FIELD_X},    {'v', "v", FS2_FIELD_V},
                        {'c', "v", -1},             {'f', "f", FS2_FIELD_FORCE},
                        {'1', "dx1", FS2_FIELD_DX1}, {'2', "dx2", FS2_FIELD_DX2}};
    for (const Col& c : cols) {
      if (*v != c.letter) continue;
      if (c.field < 0)
        check(p, fs2_cola_velocity(p, buf.data()));
      else
        check(p, fs2_particles_get_field(p, c.field, buf.data()));
      write_table(file.id, c.name, np, 3, H5T_NATIVE_FLOAT, H5T_IEEE_F32LE, buf.data());
      ++v;
    }
    if (*v != '\0') fail(FS2_ERR_VALUE, "unknown option for hdf5_write, %s", v);
  });
}
