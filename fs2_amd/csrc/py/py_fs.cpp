// fs._fs -- CPython module over the libfs2hip C ABI.
//
// Same module name, method names, argument formats and capsule names as the
// reference's method table (py/py_package.cpp:28-157, py/py_pm.cpp,
// py/py_cola.cpp, py/py_lpt.cpp, py/py_particles.cpp, py/py_fft.cpp,
// py/py_power.cpp, py/py_cosmology.cpp), so fs/*.py wrappers and their
// callers run unchanged.  C-ABI status codes become the Python exceptions the
// reference raises, now with a message (py/error.h:4-14 mapped as in the
// py_*.cpp catch blocks).  Entries that are new here (no reference
// counterpart) are marked "new" in their docstrings.
//
// Out of scope and not registered: leapfrog, FoF, kd-tree, HDF5 and Gadget
// writers, _mem_* (py/py_package.cpp:122-155).
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>

#include <dlfcn.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fs2hip.h"

namespace {

fs2_pm* g_pm = nullptr;
bool g_check_total_density = false;
bool g_domain_initialised = false;

PyObject* raise(int rc) {
  PyObject* exc = PyExc_RuntimeError;
  switch (rc) {
    case FS2_ERR_ASSERTION: exc = PyExc_AssertionError; break;
    case FS2_ERR_MEMORY: exc = PyExc_MemoryError; break;
    case FS2_ERR_IO: exc = PyExc_IOError; break;
    case FS2_ERR_VALUE: exc = PyExc_ValueError; break;
    case FS2_ERR_INDEX: exc = PyExc_IndexError; break;
    default: break;
  }
  PyErr_SetString(exc, fs2_last_error());
  return nullptr;
}

#define CALL(expr)                 \
  do {                             \
    const int rc_ = (expr);        \
    if (rc_) return raise(rc_);    \
  } while (0)

// py/py_assert.h:11-12
#define PY_ASSERT_PTR(e)                                             \
  if (!(e)) {                                                        \
    PyErr_SetString(PyExc_TypeError, "py_assersion error");          \
    return nullptr;                                                  \
  }

fs2_particles* as_particles(PyObject* o) {
  void* p = PyCapsule_GetPointer(o, "_Particles");
  if (!p) PyErr_Clear();
  return (fs2_particles*)p;
}
fs2_grid* as_grid(PyObject* o) {
  void* p = PyCapsule_GetPointer(o, "_FFT");
  if (!p) PyErr_Clear();
  return (fs2_grid*)p;
}
fs2_power* as_power(PyObject* o) {
  void* p = PyCapsule_GetPointer(o, "_PowerSpectrum");
  if (!p) PyErr_Clear();
  return (fs2_power*)p;
}

void particles_free(PyObject* o) {
  fs2_particles* p = (fs2_particles*)PyCapsule_GetPointer(o, "_Particles");
  if (p) fs2_particles_free(p);
}
void grid_free(PyObject* o) {
  fs2_grid* g = (fs2_grid*)PyCapsule_GetPointer(o, "_FFT");
  if (g) fs2_grid_free(g);
}
void power_free(PyObject* o) {
  fs2_power* ps = (fs2_power*)PyCapsule_GetPointer(o, "_PowerSpectrum");
  if (ps) fs2_power_free(ps);
}

PyObject* new_float_array(int nd, npy_intp* dims) { return PyArray_SimpleNew(nd, dims, NPY_FLOAT); }

// ---------------------------------------------------------------- msg / comm / config
PyObject* py_set_loglevel(PyObject*, PyObject* args) {
  int level;
  if (!PyArg_ParseTuple(args, "i", &level)) return nullptr;
  fs2_set_loglevel(level);
  Py_RETURN_NONE;
}
PyObject* py_comm_mpi_init(PyObject*, PyObject*) { Py_RETURN_NONE; }
PyObject* py_comm_mpi_finalise(PyObject*, PyObject*) { Py_RETURN_NONE; }
PyObject* py_comm_hello(PyObject*, PyObject*) {
  std::printf("Hello from node %d / %d\n", 0, 1);
  Py_RETURN_NONE;
}
PyObject* py_comm_this_node(PyObject*, PyObject*) { return Py_BuildValue("i", 0); }
PyObject* py_comm_n_nodes(PyObject*, PyObject*) { return Py_BuildValue("i", 1); }
PyObject* py_config_precision(PyObject*, PyObject*) { return Py_BuildValue("s", "single"); }

// ---------------------------------------------------------------- cosmology
PyObject* py_cosmology_init(PyObject*, PyObject* args) {
  double om;
  if (!PyArg_ParseTuple(args, "d", &om)) return nullptr;
  CALL(fs2_cosmology_init(om));
  Py_RETURN_NONE;
}
PyObject* py_cosmology_D_growth(PyObject*, PyObject* args) {
  double a, out;
  if (!PyArg_ParseTuple(args, "d", &a)) return nullptr;
  CALL(fs2_cosmology_D_growth(a, &out));
  return Py_BuildValue("d", out);
}
PyObject* py_cosmology_D2_growth(PyObject*, PyObject* args) {
  double a, out;
  if (!PyArg_ParseTuple(args, "d", &a)) return nullptr;
  CALL(fs2_cosmology_D2_growth(a, &out));
  return Py_BuildValue("d", out);
}
PyObject* py_cosmology_f_growth_rate(PyObject*, PyObject* args) {
  double a, out;
  if (!PyArg_ParseTuple(args, "d", &a)) return nullptr;
  CALL(fs2_cosmology_f_growth_rate(a, &out));
  return Py_BuildValue("d", out);
}
PyObject* py_cosmology_Dv_growth(PyObject*, PyObject* args) {
  double a, out;
  if (!PyArg_ParseTuple(args, "d", &a)) return nullptr;
  CALL(fs2_cosmology_Dv_growth(a, &out));
  return Py_BuildValue("d", out);
}
PyObject* py_cola_kick_coefficients(PyObject*, PyObject* args) {
  double av, ax, avel, kf, q1, q2;
  if (!PyArg_ParseTuple(args, "ddd", &av, &ax, &avel)) return nullptr;
  CALL(fs2_cola_kick_coefficients(av, ax, avel, &kf, &q1, &q2));
  return Py_BuildValue("ddd", kf, q1, q2);
}
PyObject* py_cola_drift_coefficients(PyObject*, PyObject* args) {
  double ax, av, apos, dt, d1, d2;
  if (!PyArg_ParseTuple(args, "ddd", &ax, &av, &apos)) return nullptr;
  CALL(fs2_cola_drift_coefficients(ax, av, apos, &dt, &d1, &d2));
  return Py_BuildValue("ddd", dt, d1, d2);
}

// ---------------------------------------------------------------- power
PyObject* py_power_alloc(PyObject*, PyObject* args) {
  PyObject* bytes;
  if (!PyArg_ParseTuple(args, "O&", PyUnicode_FSConverter, &bytes)) return nullptr;
  char* filename;
  Py_ssize_t len;
  PyBytes_AsStringAndSize(bytes, &filename, &len);
  fs2_power* ps = nullptr;
  const int rc = fs2_power_alloc(filename, &ps);
  Py_DECREF(bytes);
  if (rc) {
    // py/py_power.cpp:33-35
    PyErr_Format(PyExc_IOError, "PowerSpectrum File Error: %s", fs2_last_error());
    return nullptr;
  }
  return PyCapsule_New(ps, "_PowerSpectrum", power_free);
}
PyObject* py_power_eh98(PyObject*, PyObject* args) {
  double om, ob, h, ns, s8;
  if (!PyArg_ParseTuple(args, "ddddd", &om, &ob, &h, &ns, &s8)) return nullptr;
  fs2_power* ps = nullptr;
  CALL(fs2_power_alloc_eh98(om, ob, h, ns, s8, &ps));
  return PyCapsule_New(ps, "_PowerSpectrum", power_free);
}
PyObject* py_power_n(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_power* ps = as_power(o);
  PY_ASSERT_PTR(ps);
  int n;
  CALL(fs2_power_n(ps, &n));
  return Py_BuildValue("i", n);
}
PyObject* py_power_i(PyObject*, PyObject* args) {
  PyObject* o;
  int i;
  if (!PyArg_ParseTuple(args, "Oi", &o, &i)) return nullptr;
  fs2_power* ps = as_power(o);
  PY_ASSERT_PTR(ps);
  double k, P;
  CALL(fs2_power_i(ps, i, &k, &P));
  return Py_BuildValue("dd", k, P);
}
PyObject* py_power_P(PyObject*, PyObject* args) {
  PyObject* o;
  double k, P;
  if (!PyArg_ParseTuple(args, "Od", &o, &k)) return nullptr;
  fs2_power* ps = as_power(o);
  PY_ASSERT_PTR(ps);
  CALL(fs2_power_P(ps, k, &P));
  return Py_BuildValue("d", P);
}
PyObject* py_power_sigma(PyObject*, PyObject* args) {
  PyObject* o;
  double R, s;
  if (!PyArg_ParseTuple(args, "Od", &o, &R)) return nullptr;
  fs2_power* ps = as_power(o);
  PY_ASSERT_PTR(ps);
  CALL(fs2_power_sigma(ps, R, &s));
  return Py_BuildValue("d", s);
}

// ---------------------------------------------------------------- particles
PyObject* py_particles_alloc(PyObject*, PyObject* args) {
  unsigned long long np;
  double boxsize;
  if (!PyArg_ParseTuple(args, "Kd", &np, &boxsize)) return nullptr;
  fs2_particles* p = nullptr;
  CALL(fs2_particles_alloc(np, boxsize, &p));
  return PyCapsule_New(p, "_Particles", particles_free);
}

PyObject* py_particles_len(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  CALL(fs2_particles_len(p, &n));
  return Py_BuildValue("k", (unsigned long)n);
}

PyObject* py_particles_np_total(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  CALL(fs2_particles_np_total(p, &n));
  return Py_BuildValue("k", (unsigned long)n);
}

PyObject* py_particles_update_np_total(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_particles_update_np_total(p));
  Py_RETURN_NONE;
}

PyObject* records_array(fs2_particles* p, uint64_t start, uint64_t count) {
  npy_intp dims[2] = {(npy_intp)count, 14};
  PyObject* arr = new_float_array(2, dims);
  if (!arr) return nullptr;
  if (count) {
    const int rc = fs2_particles_get_records(p, start, count, PyArray_DATA((PyArrayObject*)arr));
    if (rc) {
      Py_DECREF(arr);
      return raise(rc);
    }
  }
  return arr;
}

// py/py_particles.cpp:217-270
PyObject* py_particles_getitem(PyObject*, PyObject* args) {
  PyObject *o, *index;
  if (!PyArg_ParseTuple(args, "OO", &o, &index)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  CALL(fs2_particles_len(p, &n));
  if (PyNumber_Check(index) && !PySlice_Check(index)) {
    long i = PyLong_AsLong(index);
    if (i == -1 && PyErr_Occurred()) return nullptr;
    if (i < 0) i += (long)n;
    if (i < 0 || i >= (long)n) {
      PyErr_SetNone(PyExc_IndexError);
      return nullptr;
    }
    unsigned char rec[56];
    CALL(fs2_particles_get_records(p, (uint64_t)i, 1, rec));
    float x[3];
    std::memcpy(x, rec, 12);
    return Py_BuildValue("(ddd)", (double)x[0], (double)x[1], (double)x[2]);
  }
  if (PySlice_Check(index)) {
    Py_ssize_t start, stop, step, length;
    if (PySlice_GetIndicesEx(index, (Py_ssize_t)n, &start, &stop, &step, &length)) return nullptr;
    if (step == 1) return records_array(p, (uint64_t)start, (uint64_t)length);
    std::vector<unsigned char> all(56 * n);
    if (n) CALL(fs2_particles_get_records(p, 0, n, all.data()));
    npy_intp dims[2] = {(npy_intp)length, 14};
    PyObject* arr = new_float_array(2, dims);
    if (!arr) return nullptr;
    unsigned char* dst = (unsigned char*)PyArray_DATA((PyArrayObject*)arr);
    Py_ssize_t j = 0;
    for (Py_ssize_t i = start; j < length; i += step, ++j) std::memcpy(dst + 56 * j, &all[56 * i], 56);
    return arr;
  }
  PyErr_SetString(PyExc_TypeError, "index must be an integer or a slice");
  return nullptr;
}

// py/py_particles.cpp:180-215: wrapped particles with x[2] < frac*boxsize
PyObject* py_particles_slice(PyObject*, PyObject* args) {
  PyObject* o;
  double frac;
  if (!PyArg_ParseTuple(args, "Od", &o, &frac)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  double boxsize;
  CALL(fs2_particles_len(p, &n));
  CALL(fs2_particles_boxsize(p, &boxsize));
  std::vector<unsigned char> all(56 * n);
  if (n) CALL(fs2_particles_get_records(p, 0, n, all.data()));
  const float L = (float)boxsize;
  const float xmax = (float)(frac * boxsize);
  std::vector<unsigned char> sel;
  for (uint64_t i = 0; i < n; ++i) {
    float rec[12];
    std::memcpy(rec, &all[56 * i], 48);
    for (int k = 0; k < 3; ++k) {
      if (rec[k] < 0) rec[k] += L;
      if (rec[k] >= L) rec[k] -= L;
    }
    if (rec[2] < xmax) {
      const size_t off = sel.size();
      sel.resize(off + 56);
      std::memcpy(&sel[off], rec, 48);
      std::memcpy(&sel[off + 48], &all[56 * i + 48], 8);
    }
  }
  npy_intp dims[2] = {(npy_intp)(sel.size() / 56), 14};
  PyObject* arr = new_float_array(2, dims);
  if (!arr) return nullptr;
  if (!sel.empty()) std::memcpy(PyArray_DATA((PyArrayObject*)arr), sel.data(), sel.size());
  return arr;
}

PyObject* field_array(PyObject* args, int field) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  CALL(fs2_particles_len(p, &n));
  npy_intp dims[2] = {(npy_intp)n, 3};
  PyObject* arr = new_float_array(2, dims);
  if (!arr) return nullptr;
  const int rc = fs2_particles_get_field(p, field, (float*)PyArray_DATA((PyArrayObject*)arr));
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

// py/py_particles.cpp:306-331: x_type 0 x, 1 dx1, 2 dx2
PyObject* py_particles_x_asarray(PyObject*, PyObject* args) {
  PyObject* o;
  int x_type;
  if (!PyArg_ParseTuple(args, "Oi", &o, &x_type)) return nullptr;
  PY_ASSERT_PTR(x_type >= 0 && x_type <= 2);
  PyObject* a = Py_BuildValue("(O)", o);
  PyObject* r = field_array(a, x_type == 0 ? FS2_FIELD_X : (x_type == 1 ? FS2_FIELD_DX1 : FS2_FIELD_DX2));
  Py_DECREF(a);
  return r;
}
PyObject* py_particles_force_asarray(PyObject*, PyObject* args) { return field_array(args, FS2_FIELD_FORCE); }
PyObject* py_particles_v_asarray(PyObject*, PyObject* args) { return field_array(args, FS2_FIELD_V); }

PyObject* py_particles_id_asarray(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  CALL(fs2_particles_len(p, &n));
  npy_intp dims[1] = {(npy_intp)n};
  PyObject* arr = PyArray_SimpleNew(1, dims, NPY_UINT64);
  if (!arr) return nullptr;
  const int rc = fs2_particles_get_id(p, (uint64_t*)PyArray_DATA((PyArrayObject*)arr));
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

PyObject* py_particles_set_field(PyObject*, PyObject* args) {
  PyObject *o, *src;
  int field;
  if (!PyArg_ParseTuple(args, "OiO", &o, &field, &src)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  CALL(fs2_particles_len(p, &n));
  PyArrayObject* a = (PyArrayObject*)PyArray_FROM_OTF(src, NPY_FLOAT, NPY_ARRAY_IN_ARRAY);
  if (!a) return nullptr;
  if (PyArray_NDIM(a) != 2 || PyArray_DIM(a, 0) != (npy_intp)n || PyArray_DIM(a, 1) != 3) {
    Py_DECREF(a);
    PyErr_SetString(PyExc_ValueError, "expected an (np_local, 3) array");
    return nullptr;
  }
  const int rc = fs2_particles_set_field(p, field, (const float*)PyArray_DATA(a));
  Py_DECREF(a);
  if (rc) return raise(rc);
  Py_RETURN_NONE;
}

PyObject* py_particles_periodic_wrapup(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_particles_periodic_wrapup(p));
  Py_RETURN_NONE;
}

// ---- friends-of-friends (py/py_fof.cpp, py/py_kdtree.cpp) -------------------------------
PyObject* int32_array_from(int (*get)(int32_t*, uint64_t*)) {
  uint64_t n = 0;
  CALL(get(nullptr, &n));
  if (n == 0) Py_RETURN_NONE;
  npy_intp dims[1] = {(npy_intp)n};
  PyObject* arr = PyArray_SimpleNew(1, dims, NPY_INT32);
  if (!arr) return nullptr;
  const int rc = get((int32_t*)PyArray_DATA((PyArrayObject*)arr), &n);
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

// py_fof.cpp:14-54: _fof_find_groups(_particles, linking_length, boxsize3,
// quota, return_nfof) -> nfof array or None
PyObject* py_fof_find_groups(PyObject*, PyObject* args) {
  PyObject *o, *py_box;
  double ll;
  int quota, return_nfof;
  if (!PyArg_ParseTuple(args, "OdOii", &o, &ll, &py_box, &quota, &return_nfof)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  double box[3];
  const double* pbox = nullptr;
  if (py_box != Py_None) {
    PY_ASSERT_PTR(PySequence_Check(py_box));
    for (int k = 0; k < 3; ++k) {
      PyObject* e = PySequence_GetItem(py_box, k);
      PY_ASSERT_PTR(e);
      const bool is_float = PyFloat_Check(e);
      box[k] = is_float ? PyFloat_AsDouble(e) : 0.0;
      Py_DECREF(e);
      PY_ASSERT_PTR(is_float);
    }
    pbox = box;
  }
  CALL(fs2_fof_find_groups(p, ll, pbox, quota, nullptr));
  if (return_nfof) return int32_array_from(fs2_fof_nfof);
  Py_RETURN_NONE;
}

// py_fof.cpp:56-64: _fof_grp() -> grp array, or None before any run
PyObject* py_fof_grp(PyObject*, PyObject*) { return int32_array_from(fs2_fof_grp); }

// py_kdtree.cpp:6-36: list of (i, k, left, right, ibegin, iend)
PyObject* py_kdtree_create_copy(PyObject*, PyObject*) {
  uint64_t n = 0;
  CALL(fs2_kdtree_nodes(nullptr, &n));
  std::vector<fs2_kdtree_node> nodes(n);
  CALL(fs2_kdtree_nodes(nodes.data(), &n));
  PyObject* list = PyList_New(0);
  if (!list) return nullptr;
  for (uint64_t i = 0; i < n; ++i) {
    const fs2_kdtree_node& t = nodes[i];
    PyObject* item = Py_BuildValue("(KiddKK)", (unsigned long long)i, (int)t.k, (double)t.left,
                                   (double)t.right, (unsigned long long)t.ibegin,
                                   (unsigned long long)t.iend);
    if (!item || PyList_Append(list, item) != 0) {
      Py_XDECREF(item);
      Py_DECREF(list);
      return nullptr;
    }
    Py_DECREF(item);
  }
  return list;
}

PyObject* py_kdtree_get_height(PyObject*, PyObject*) {
  uint64_t h = 0;
  CALL(fs2_kdtree_height(&h));
  return Py_BuildValue("K", (unsigned long long)h);
}

// new: _particles_resize(_particles, np_local)
PyObject* py_particles_resize(PyObject*, PyObject* args) {
  PyObject* o;
  unsigned long long n;
  if (!PyArg_ParseTuple(args, "OK", &o, &n)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_particles_resize(p, n));
  Py_RETURN_NONE;
}

PyObject* py_particles_clear(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_particles_clear(p));
  Py_RETURN_NONE;
}

// py/py_particles.cpp:403-462
PyObject* py_particles_append(PyObject*, PyObject* args) {
  PyObject *o, *arr;
  if (!PyArg_ParseTuple(args, "OO", &o, &arr)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  if (arr != Py_None) {
    Py_buffer view;
    if (PyObject_GetBuffer(arr, &view, PyBUF_ANY_CONTIGUOUS | PyBUF_FORMAT) == -1) return nullptr;
    if (view.ndim != 2) {
      PyErr_SetString(PyExc_TypeError, "Expected a 2-dimensional array");
      PyBuffer_Release(&view);
      return nullptr;
    }
    if (std::strcmp(view.format, "d") != 0) {
      PyErr_SetString(PyExc_TypeError, "Expected an array of doubles");
      PyBuffer_Release(&view);
      return nullptr;
    }
    const Py_ssize_t n = view.shape[0], ncol = view.shape[1];
    if (ncol < 3) {
      PyErr_SetString(PyExc_TypeError, "Expected at least 3 columns for x y z");
      PyBuffer_Release(&view);
      return nullptr;
    }
    const int64_t rs = view.strides[0] / (Py_ssize_t)sizeof(double);
    const int64_t cs = view.strides[1] / (Py_ssize_t)sizeof(double);
    const int rc = fs2_particles_append(p, (const double*)view.buf, (uint64_t)n, rs, cs);
    PyBuffer_Release(&view);
    if (rc) return raise(rc);
  } else {
    CALL(fs2_particles_update_np_total(p));
  }
  Py_RETURN_NONE;
}

PyObject* py_particles_times(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  double ax, av, af;
  CALL(fs2_particles_get_times(p, &ax, &av, &af));
  return Py_BuildValue("ddd", ax, av, af);
}

PyObject* py_particles_set_times(PyObject*, PyObject* args) {
  PyObject* o;
  double ax, av, af;
  if (!PyArg_ParseTuple(args, "Oddd", &o, &ax, &av, &af)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_particles_set_times(p, ax, av, af));
  Py_RETURN_NONE;
}

PyObject* py_particles_copy(PyObject*, PyObject* args) {
  PyObject *d, *s;
  if (!PyArg_ParseTuple(args, "OO", &d, &s)) return nullptr;
  fs2_particles* dst = as_particles(d);
  fs2_particles* src = as_particles(s);
  PY_ASSERT_PTR(dst && src);
  CALL(fs2_particles_copy(dst, src));
  Py_RETURN_NONE;
}

PyObject* py_particles_sync(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_particles_sync(p));
  Py_RETURN_NONE;
}

PyObject* py_particles_device_fields(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  float *x, *v, *dx1, *dx2, *f;
  uint64_t* id;
  uint64_t stride;
  CALL(fs2_particles_device_fields(p, &x, &v, &dx1, &dx2, &f, &id, &stride));
  return Py_BuildValue("{s:K,s:K,s:K,s:K,s:K,s:K,s:K}", "x", (unsigned long long)(uintptr_t)x, "v",
                       (unsigned long long)(uintptr_t)v, "dx1", (unsigned long long)(uintptr_t)dx1,
                       "dx2", (unsigned long long)(uintptr_t)dx2, "force",
                       (unsigned long long)(uintptr_t)f, "id", (unsigned long long)(uintptr_t)id,
                       "stride", (unsigned long long)stride);
}

// ---------------------------------------------------------------- lpt
PyObject* py_lpt(PyObject*, PyObject* args) {
  PyObject* o;
  int nc;
  double a, boxsize;
  unsigned long seed;
  const char* kind;
  if (!PyArg_ParseTuple(args, "iddkOs", &nc, &boxsize, &a, &seed, &o, &kind)) return nullptr;
  fs2_power* ps = as_power(o);
  PY_ASSERT_PTR(ps);
  fs2_particles* p = nullptr;
  CALL(fs2_lpt(nc, boxsize, a, (uint64_t)seed, ps, kind, &p));
  return PyCapsule_New(p, "_Particles", particles_free);
}

PyObject* py_lpt_set_offset(PyObject*, PyObject* args) {
  double off;
  if (!PyArg_ParseTuple(args, "d", &off)) return nullptr;
  CALL(fs2_lpt_set_offset(off));
  Py_RETURN_NONE;
}

PyObject* py_lpt_set_zeldovich_force(PyObject*, PyObject* args) {
  PyObject* o;
  double a;
  if (!PyArg_ParseTuple(args, "Od", &o, &a)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_lpt_set_zeldovich_force(p, a));
  Py_RETURN_NONE;
}

PyObject* py_lpt_white_noise(PyObject*, PyObject* args) {
  int nc;
  unsigned long long seed;
  if (!PyArg_ParseTuple(args, "iK", &nc, &seed)) return nullptr;
  npy_intp dims[3] = {nc, nc, nc};
  PyObject* arr = new_float_array(3, dims);
  if (!arr) return nullptr;
  const int rc = fs2_lpt_white_noise(nc, seed, (float*)PyArray_DATA((PyArrayObject*)arr));
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

// ---------------------------------------------------------------- pm
PyObject* py_pm_init(PyObject*, PyObject* args) {
  int nc;
  double pm_factor, boxsize;
  if (!PyArg_ParseTuple(args, "idd", &nc, &pm_factor, &boxsize)) return nullptr;
  // py/pm.cpp:233-238: same parameters -> no-op
  if (g_pm) {
    int n0;
    double f0, b0;
    CALL(fs2_pm_params(g_pm, &n0, &f0, &b0));
    if (n0 == nc && f0 == pm_factor && (float)b0 == (float)boxsize) Py_RETURN_NONE;
    fs2_pm_destroy(g_pm);
    g_pm = nullptr;
  }
  CALL(fs2_pm_create(nc, pm_factor, boxsize, &g_pm));
  Py_RETURN_NONE;
}

#define REQUIRE_PM()                                                              \
  if (!g_pm) {                                                                    \
    PyErr_SetString(PyExc_RuntimeError, "PM not initialised; call pm_init().");   \
    return nullptr;                                                               \
  }

PyObject* py_pm_compute_density(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  REQUIRE_PM();
  fs2_grid* g = nullptr;
  CALL(fs2_pm_compute_density(g_pm, p, &g));
  return PyCapsule_New(g, "_FFT", nullptr);   // borrowed, py_pm.cpp:55
}

PyObject* py_pm_compute_force(PyObject*, PyObject* args) {
  REQUIRE_PM();
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_pm_compute_force(g_pm, p));
  Py_RETURN_NONE;
}

PyObject* py_pm_send_positions(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  g_domain_initialised = true;
  CALL(fs2_pm_send_positions(g_pm, p));
  Py_RETURN_NONE;
}

PyObject* py_pm_check_total_density(PyObject*, PyObject*) {
  // A no-op unless enabled, as the reference build without -DCHECK
  // (py/pm.cpp:357-390).
  if (!g_check_total_density) Py_RETURN_NONE;
  REQUIRE_PM();
  double s;
  CALL(fs2_pm_check_total_density(g_pm, &s));
  Py_RETURN_NONE;
}

PyObject* py_pm_set_check_total_density(PyObject*, PyObject* args) {
  int on;
  if (!PyArg_ParseTuple(args, "p", &on)) return nullptr;
  g_check_total_density = on != 0;
  Py_RETURN_NONE;
}

PyObject* py_pm_total_density(PyObject*, PyObject*) {
  REQUIRE_PM();
  double s = 0.0;
  const int rc = fs2_pm_check_total_density(g_pm, &s);
  if (rc && rc != FS2_ERR_ASSERTION) return raise(rc);
  return Py_BuildValue("d", s);
}

PyObject* py_pm_get_forces(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  if (g_pm) CALL(fs2_pm_get_forces(g_pm, p));
  Py_RETURN_NONE;
}

PyObject* py_pm_domain_init(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  g_domain_initialised = true;
  Py_RETURN_NONE;
}

// py/pm_domain.cpp:393-406
PyObject* py_pm_set_packet_size(PyObject*, PyObject* args) {
  int packet_size;
  if (!PyArg_ParseTuple(args, "i", &packet_size)) return nullptr;
  if (g_domain_initialised) {
    PyErr_SetString(PyExc_RuntimeError,
                    "pm_domain already initialised. packet_size must be set earlier");
    return nullptr;
  }
  Py_RETURN_NONE;
}

PyObject* py_pm_set_paint_variant(PyObject*, PyObject* args) {
  int v;
  if (!PyArg_ParseTuple(args, "i", &v)) return nullptr;
  REQUIRE_PM();
  CALL(fs2_pm_set_paint_variant(g_pm, v));
  Py_RETURN_NONE;
}

PyObject* py_pm_force_mesh(PyObject*, PyObject* args) {
  int axis;
  if (!PyArg_ParseTuple(args, "i", &axis)) return nullptr;
  REQUIRE_PM();
  int nc;
  double f, b;
  CALL(fs2_pm_params(g_pm, &nc, &f, &b));
  npy_intp dims[3] = {nc, nc, nc};
  PyObject* arr = new_float_array(3, dims);
  if (!arr) return nullptr;
  const int rc = fs2_pm_get_force_mesh(g_pm, axis, (float*)PyArray_DATA((PyArrayObject*)arr));
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

PyObject* py_pm_power_spectrum(PyObject*, PyObject* args) {
  int nbins;
  double dk_bin;
  if (!PyArg_ParseTuple(args, "id", &nbins, &dk_bin)) return nullptr;
  REQUIRE_PM();
  if (nbins <= 0) {
    PyErr_SetString(PyExc_ValueError, "nbins must be positive");
    return nullptr;
  }
  npy_intp dims[1] = {nbins};
  PyObject* k = PyArray_SimpleNew(1, dims, NPY_DOUBLE);
  PyObject* P = PyArray_SimpleNew(1, dims, NPY_DOUBLE);
  PyObject* nm = PyArray_SimpleNew(1, dims, NPY_DOUBLE);
  if (!k || !P || !nm) return nullptr;
  const int rc = fs2_pm_power_spectrum(g_pm, nbins, dk_bin, (double*)PyArray_DATA((PyArrayObject*)k),
                                       (double*)PyArray_DATA((PyArrayObject*)P),
                                       (double*)PyArray_DATA((PyArrayObject*)nm));
  if (rc) {
    Py_DECREF(k);
    Py_DECREF(P);
    Py_DECREF(nm);
    return raise(rc);
  }
  return Py_BuildValue("NNN", k, P, nm);
}

// ---------------------------------------------------------------- cola
PyObject* py_cola_kick(PyObject*, PyObject* args) {
  PyObject* o;
  double a_vel;
  if (!PyArg_ParseTuple(args, "Od", &o, &a_vel)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_cola_kick(p, a_vel));
  Py_RETURN_NONE;
}

PyObject* py_cola_drift(PyObject*, PyObject* args) {
  PyObject* o;
  double a_pos;
  if (!PyArg_ParseTuple(args, "Od", &o, &a_pos)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_cola_drift(p, a_pos));
  Py_RETURN_NONE;
}

PyObject* py_cola_velocity(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  uint64_t n;
  CALL(fs2_particles_len(p, &n));
  npy_intp dims[2] = {(npy_intp)n, 3};
  PyObject* arr = new_float_array(2, dims);
  if (!arr) return nullptr;
  const int rc = fs2_cola_velocity(p, (float*)PyArray_DATA((PyArrayObject*)arr));
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

// ---------------------------------------------------------------- output
// py/py_write.cpp: _write_gadget_binary(_particles, filename, use_long_id)
PyObject* py_write_gadget_binary(PyObject*, PyObject* args) {
  PyObject *o, *bytes, *longid;
  if (!PyArg_ParseTuple(args, "OO&O", &o, PyUnicode_FSConverter, &bytes, &longid)) return nullptr;
  fs2_particles* p = as_particles(o);
  if (!p) {
    Py_DECREF(bytes);
    PyErr_SetString(PyExc_TypeError, "py_assersion error");
    return nullptr;
  }
  const int use_long_id = PyObject_IsTrue(longid);
  const int rc = use_long_id < 0 ? -1 : fs2_write_gadget(p, PyBytes_AsString(bytes), use_long_id, 1.0);
  Py_DECREF(bytes);
  if (rc < 0) return nullptr;
  if (rc) return raise(rc);
  Py_RETURN_NONE;
}

// The HDF5 writer is in libfs2io.so next to libfs2hip.so, opened on first
// use so that the module loads where no HDF5 library is installed.
typedef int (*write_hdf5_fn)(fs2_particles*, const char*, const char*);
write_hdf5_fn hdf5_writer(std::string* why) {
  static write_hdf5_fn fn = nullptr;
  static std::string err;
  if (fn || !err.empty()) {
    *why = err;
    return fn;
  }
  Dl_info info;
  std::string dir = ".";
  if (dladdr((void*)&hdf5_writer, &info) && info.dli_fname) {
    dir = info.dli_fname;
    dir = dir.substr(0, dir.find_last_of('/'));
  }
  const std::string path = dir + "/../libfs2io.so";
  void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    const char* e = dlerror();
    err = std::string("HDF5 output is not available (") + (e ? e : path) + ")";
  } else {
    fn = (write_hdf5_fn)dlsym(h, "fs2_write_hdf5");
    if (!fn) err = "libfs2io.so does not export fs2_write_hdf5";
  }
  *why = err;
  return fn;
}

// py/py_hdf5_io.cpp: _hdf5_write_particles(_particles, filename, var)
PyObject* py_hdf5_write_particles(PyObject*, PyObject* args) {
  PyObject *o, *bytes, *vbytes;
  if (!PyArg_ParseTuple(args, "OO&O&", &o, PyUnicode_FSConverter, &bytes, PyUnicode_FSConverter,
                        &vbytes))
    return nullptr;
  fs2_particles* p = as_particles(o);
  std::string why;
  write_hdf5_fn fn = p ? hdf5_writer(&why) : nullptr;
  int rc = 0;
  if (p && fn) rc = fn(p, PyBytes_AsString(bytes), PyBytes_AsString(vbytes));
  Py_DECREF(bytes);
  Py_DECREF(vbytes);
  if (!p) {
    PyErr_SetString(PyExc_TypeError, "py_assersion error");
    return nullptr;
  }
  if (!fn) {
    PyErr_SetString(PyExc_IOError, why.c_str());
    return nullptr;
  }
  if (rc) return raise(rc);
  Py_RETURN_NONE;
}

// ---------------------------------------------------------------- leapfrog
// py/py_leapfrog.cpp:4-65
PyObject* py_leapfrog_initial_velocity(PyObject*, PyObject* args) {
  PyObject* o;
  double a;
  if (!PyArg_ParseTuple(args, "Od", &o, &a)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_leapfrog_initial_velocity(p, a));
  Py_RETURN_NONE;
}

PyObject* py_leapfrog_kick(PyObject*, PyObject* args) {
  PyObject* o;
  double a_vel;
  if (!PyArg_ParseTuple(args, "Od", &o, &a_vel)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_leapfrog_kick(p, a_vel));
  Py_RETURN_NONE;
}

PyObject* py_leapfrog_drift(PyObject*, PyObject* args) {
  PyObject* o;
  double a_pos;
  if (!PyArg_ParseTuple(args, "Od", &o, &a_pos)) return nullptr;
  fs2_particles* p = as_particles(o);
  PY_ASSERT_PTR(p);
  CALL(fs2_leapfrog_drift(p, a_pos));
  Py_RETURN_NONE;
}

PyObject* py_leapfrog_factors(PyObject*, PyObject* args) {
  double ai, af, kf, dt;
  if (!PyArg_ParseTuple(args, "dd", &ai, &af)) return nullptr;
  CALL(fs2_leapfrog_factors(ai, af, &kf, &dt));
  return Py_BuildValue("dd", kf, dt);
}

// ---------------------------------------------------------------- fft
PyObject* py_fft_alloc(PyObject*, PyObject* args) {
  int nc;
  if (!PyArg_ParseTuple(args, "i", &nc)) return nullptr;
  fs2_grid* g = nullptr;
  CALL(fs2_grid_alloc(nc, &g));
  return PyCapsule_New(g, "_FFT", grid_free);
}

PyObject* py_fft_set_test_data(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_grid* g = as_grid(o);
  PY_ASSERT_PTR(g);
  CALL(fs2_grid_set_test_data(g));
  Py_RETURN_NONE;
}

PyObject* py_fft_fx_global_as_array(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_grid* g = as_grid(o);
  PY_ASSERT_PTR(g);
  int nc;
  CALL(fs2_grid_nc(g, &nc));
  npy_intp dims[3] = {nc, nc, nc};
  PyObject* arr = new_float_array(3, dims);
  if (!arr) return nullptr;
  const int rc = fs2_grid_get_real(g, (float*)PyArray_DATA((PyArrayObject*)arr));
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

PyObject* py_fft_fk_as_array(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_grid* g = as_grid(o);
  PY_ASSERT_PTR(g);
  int nc;
  CALL(fs2_grid_nc(g, &nc));
  npy_intp dims[3] = {nc, nc, nc / 2 + 1};
  PyObject* arr = PyArray_SimpleNew(3, dims, NPY_COMPLEX64);
  if (!arr) return nullptr;
  const int rc = fs2_grid_get_complex(g, (float*)PyArray_DATA((PyArrayObject*)arr));
  if (rc) {
    Py_DECREF(arr);
    return raise(rc);
  }
  return arr;
}

PyObject* py_fft_set_fx(PyObject*, PyObject* args) {
  PyObject *o, *src;
  if (!PyArg_ParseTuple(args, "OO", &o, &src)) return nullptr;
  fs2_grid* g = as_grid(o);
  PY_ASSERT_PTR(g);
  int nc;
  CALL(fs2_grid_nc(g, &nc));
  PyArrayObject* a = (PyArrayObject*)PyArray_FROM_OTF(src, NPY_FLOAT, NPY_ARRAY_IN_ARRAY);
  if (!a) return nullptr;
  if (PyArray_SIZE(a) != (npy_intp)nc * nc * nc) {
    Py_DECREF(a);
    PyErr_SetString(PyExc_ValueError, "expected nc^3 values");
    return nullptr;
  }
  const int rc = fs2_grid_set_real(g, (const float*)PyArray_DATA(a));
  Py_DECREF(a);
  if (rc) return raise(rc);
  Py_RETURN_NONE;
}

PyObject* py_fft_set_fk(PyObject*, PyObject* args) {
  PyObject *o, *src;
  if (!PyArg_ParseTuple(args, "OO", &o, &src)) return nullptr;
  fs2_grid* g = as_grid(o);
  PY_ASSERT_PTR(g);
  int nc;
  CALL(fs2_grid_nc(g, &nc));
  PyArrayObject* a = (PyArrayObject*)PyArray_FROM_OTF(src, NPY_COMPLEX64, NPY_ARRAY_IN_ARRAY);
  if (!a) return nullptr;
  if (PyArray_SIZE(a) != (npy_intp)nc * nc * (nc / 2 + 1)) {
    Py_DECREF(a);
    PyErr_SetString(PyExc_ValueError, "expected nc*nc*(nc/2+1) values");
    return nullptr;
  }
  const int rc = fs2_grid_set_complex(g, (const float*)PyArray_DATA(a));
  Py_DECREF(a);
  if (rc) return raise(rc);
  Py_RETURN_NONE;
}

PyObject* py_fft_r2c(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_grid* g = as_grid(o);
  PY_ASSERT_PTR(g);
  CALL(fs2_grid_r2c(g));
  Py_RETURN_NONE;
}

PyObject* py_fft_c2r(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_grid* g = as_grid(o);
  PY_ASSERT_PTR(g);
  CALL(fs2_grid_c2r(g));
  Py_RETURN_NONE;
}

// ---------------------------------------------------------------- runtime
PyObject* py_device_init(PyObject*, PyObject*) {
  CALL(fs2_device_init());
  Py_RETURN_NONE;
}
PyObject* py_synchronize(PyObject*, PyObject*) {
  CALL(fs2_synchronize());
  Py_RETURN_NONE;
}
PyObject* py_profile_enable(PyObject*, PyObject* args) {
  int on;
  if (!PyArg_ParseTuple(args, "p", &on)) return nullptr;
  CALL(fs2_profile_enable(on));
  Py_RETURN_NONE;
}
PyObject* py_profile_reset(PyObject*, PyObject*) {
  CALL(fs2_profile_reset());
  Py_RETURN_NONE;
}
// ---- py_mem.cpp: _mem_alloc(name, size) -> "_Mem" capsule, _mem(m) -> (alloc, using)
void mem_free(PyObject* cap) {
  fs2_mem_free((fs2_mem*)PyCapsule_GetPointer(cap, "_Mem"));
}

PyObject* py_mem_alloc(PyObject*, PyObject* args) {
  const char* name;
  unsigned long long size;
  if (!PyArg_ParseTuple(args, "sK", &name, &size)) return nullptr;
  fs2_mem* m = nullptr;
  CALL(fs2_mem_alloc(name, size, &m));
  return PyCapsule_New(m, "_Mem", mem_free);
}

PyObject* py_mem(PyObject*, PyObject* args) {
  PyObject* o;
  if (!PyArg_ParseTuple(args, "O", &o)) return nullptr;
  fs2_mem* m = (fs2_mem*)PyCapsule_GetPointer(o, "_Mem");
  PY_ASSERT_PTR(m);
  uint64_t a = 0, u = 0;
  CALL(fs2_mem_info(m, &a, &u));
  return Py_BuildValue("(KK)", (unsigned long long)a, (unsigned long long)u);
}

// new: _copy_bandwidth(bytes, grid, reps) -> best ms of a streaming float4 copy
PyObject* py_copy_bandwidth(PyObject*, PyObject* args) {
  unsigned long long bytes;
  int grid, reps;
  if (!PyArg_ParseTuple(args, "Kii", &bytes, &grid, &reps)) return nullptr;
  double ms = 0.0;
  CALL(fs2_copy_bandwidth(bytes, grid, reps, &ms));
  return PyFloat_FromDouble(ms);
}

PyObject* py_profile_get(PyObject*, PyObject*) {
  int n;
  CALL(fs2_profile_count(&n));
  PyObject* list = PyList_New(0);
  for (int i = 0; i < n; ++i) {
    const char* name;
    uint64_t launches;
    double ms, bytes;
    const int rc = fs2_profile_get(i, &name, &launches, &ms, &bytes);
    if (rc) {
      Py_DECREF(list);
      return raise(rc);
    }
    PyObject* t = Py_BuildValue("(sKdd)", name, (unsigned long long)launches, ms, bytes);
    PyList_Append(list, t);
    Py_DECREF(t);
  }
  return list;
}

#define M(name, fn, doc) {name, (PyCFunction)fn, METH_VARARGS, doc}

PyMethodDef methods[] = {
    M("set_loglevel", py_set_loglevel, "set loglevel: 0=debug, 1=verbose, ..."),
    M("comm_mpi_init", py_comm_mpi_init, "one rank: no-op"),
    M("comm_mpi_finalise", py_comm_mpi_finalise, "one rank: no-op"),
    M("comm_hello", py_comm_hello, "print node / n_nodes"),
    M("comm_this_node", py_comm_this_node, "rank (always 0)"),
    M("comm_n_nodes", py_comm_n_nodes, "number of ranks (always 1)"),
    M("_cosmology_init", py_cosmology_init, "_cosmology_init(omega_m0)"),
    M("_cosmology_D_growth", py_cosmology_D_growth, "_cosmology_D_growth(a)"),
    M("_cosmology_D2_growth", py_cosmology_D2_growth, "_cosmology_D2_growth(a)"),
    M("_cosmology_f_growth_rate", py_cosmology_f_growth_rate, "new: f(a)"),
    M("_cosmology_Dv_growth", py_cosmology_Dv_growth, "new: Dv(a, D1(a))"),
    M("_cola_kick_coefficients", py_cola_kick_coefficients, "new: (kick_factor, q1, q2)"),
    M("_cola_drift_coefficients", py_cola_drift_coefficients, "new: (dt, da1, da2)"),
    M("_power_alloc", py_power_alloc, "allocate a new _ps object from a k P table"),
    M("_power_eh98", py_power_eh98, "new: EH98 no-wiggle P(k) (om, ob, h, ns, sigma8)"),
    M("_power_n", py_power_n, "_power_n(_ps); number of P(k) data"),
    M("_power_i", py_power_i, "_power_i(_ps, i); (k[i], P[i])"),
    M("_power_P", py_power_P, "new: P(k)"),
    M("_power_sigma", py_power_sigma, "new: sigma(R)"),
    M("_particles_alloc", py_particles_alloc, "allocate a new _particles object"),
    M("_particles_slice", py_particles_slice, "particles with x[2] < frac*boxsize as np.array"),
    M("_particles_len", py_particles_len, "local number of particles"),
    M("_particles_np_total", py_particles_np_total, "total number of particles"),
    M("_particles_update_np_total", py_particles_update_np_total, "update np_total"),
    M("_particles_getitem", py_particles_getitem, "_particles_getitem(_particles, index)"),
    M("_particles_id_asarray", py_particles_id_asarray, "_particles_id_asarray(_particles)"),
    M("_particles_x_asarray", py_particles_x_asarray, "_particles_x_asarray(_particles, type)"),
    M("_particles_force_asarray", py_particles_force_asarray, "_particles_force_asarray(_particles)"),
    M("_particles_v_asarray", py_particles_v_asarray, "new: velocities (np_local, 3)"),
    M("_particles_set_field", py_particles_set_field, "new: upload one (np_local,3) field"),
    M("_particles_periodic_wrapup", py_particles_periodic_wrapup, "periodic wrap of x"),
    M("_particles_resize", py_particles_resize, "new: set np_local"),
    M("_fof_find_groups", py_fof_find_groups,
      "_fof_find_groups(_particles, linking_length, boxsize3, quota, return_nfof)"),
    M("_fof_grp", py_fof_grp, "_fof_grp()"),
    M("_kdtree_create_copy", py_kdtree_create_copy, "_kdtree_create_copy()"),
    M("_kdtree_get_height", py_kdtree_get_height, "_kdtree_get_height()"),
    M("_write_gadget_binary", py_write_gadget_binary,
      "_write_gadget_binary(_particles, filename, use_long_id)"),
    M("_hdf5_write_particles", py_hdf5_write_particles,
      "_hdf5_write_particles(_particles, filename, var)"),
    M("_particles_clear", py_particles_clear, "_particles_clear(_particles)"),
    M("_particles_append", py_particles_append, "_particles_append(_particles, x)"),
    M("_particles_times", py_particles_times, "new: (a_x, a_v, a_f)"),
    M("_particles_set_times", py_particles_set_times, "new: set (a_x, a_v, a_f)"),
    M("_particles_copy", py_particles_copy, "new: device copy dst <- src"),
    M("_particles_sync", py_particles_sync, "new: apply the deferred COLA kick and drift"),
    M("_particles_device_fields", py_particles_device_fields, "new: device pointers"),
    M("_lpt", py_lpt, "_lpt(nc, boxsize, a, seed, _ps, kind); 2LPT particles"),
    M("_lpt_set_offset", py_lpt_set_offset, "_lpt_set_offset(offset)"),
    M("_lpt_set_zeldovich_force", py_lpt_set_zeldovich_force, "_lpt_set_zeldovich_force(_particles, a)"),
    M("_lpt_white_noise", py_lpt_white_noise, "new: the white-noise field of _lpt(nc, seed)"),
    M("_pm_init", py_pm_init, "_pm_init(nc_pm, pm_factor, boxsize); initialise pm module"),
    M("_pm_compute_force", py_pm_compute_force, "_pm_compute_force(_particles)"),
    M("_pm_compute_density", py_pm_compute_density, "_pm_compute_density(_particles)"),
    M("_pm_domain_init", py_pm_domain_init, "_pm_domain_init(_particles)"),
    M("_pm_send_positions", py_pm_send_positions, "_pm_send_positions(_particles)"),
    M("_pm_check_total_density", py_pm_check_total_density, "_pm_check_total_density"),
    M("_pm_set_check_total_density", py_pm_set_check_total_density, "new: enable the check"),
    M("_pm_total_density", py_pm_total_density, "new: FP64 sum of delta"),
    M("_pm_get_forces", py_pm_get_forces, "_pm_get_forces(_particles)"),
    M("_pm_set_packet_size", py_pm_set_packet_size, "_pm_set_packet_size(packet_size)"),
    M("_pm_set_paint_variant", py_pm_set_paint_variant, "new: 0 atomic, 1 sorted"),
    M("_pm_force_mesh", py_pm_force_mesh, "new: real force mesh of one axis"),
    M("_pm_power_spectrum", py_pm_power_spectrum, "new: (k, P, nmodes) of the density"),
    M("_cola_kick", py_cola_kick, "_cola_kick(_particles, a_vel)"),
    M("_leapfrog_initial_velocity", py_leapfrog_initial_velocity,
      "_leapfrog_initial_velocity(_particles, a)"),
    M("_leapfrog_kick", py_leapfrog_kick, "_leapfrog_kick(_particles, a_vel)"),
    M("_leapfrog_drift", py_leapfrog_drift, "_leapfrog_drift(_particles, a_pos)"),
    M("_leapfrog_factors", py_leapfrog_factors, "new: (kick_factor, dt) of SphiStd/SqStd"),
    M("_cola_drift", py_cola_drift, "_cola_drift(_particles, a_pos)"),
    M("_cola_velocity", py_cola_velocity, "new: physical COLA velocity"),
    M("_fft_alloc", py_fft_alloc, "_fft_alloc(nc)"),
    M("_fft_set_test_data", py_fft_set_test_data, "_fft_set_test_data(_fft)"),
    M("_fft_fx_global_as_array", py_fft_fx_global_as_array, "fft->fx as nc^3 np.array"),
    M("_fft_fk_as_array", py_fft_fk_as_array, "new: half-complex view"),
    M("_fft_set_fx", py_fft_set_fx, "new: upload a real grid"),
    M("_fft_set_fk", py_fft_set_fk, "new: upload a half-complex grid"),
    M("_fft_r2c", py_fft_r2c, "new: forward r2c in place"),
    M("_fft_c2r", py_fft_c2r, "new: inverse c2r in place"),
    M("config_precision", py_config_precision, "get 'single' or 'double'"),
    M("_device_init", py_device_init, "new: initialise the GPU"),
    M("_synchronize", py_synchronize, "new: wait for device work"),
    M("_profile_enable", py_profile_enable, "new: per-kernel HIP-event timing"),
    M("_profile_reset", py_profile_reset, "new"),
    M("_copy_bandwidth", py_copy_bandwidth, "new: best ms of a streaming float4 copy"),
    M("_mem_alloc", py_mem_alloc, "create a new Mem object"),
    M("_mem", py_mem, "return mem allocated and using"),
    M("_profile_get", py_profile_get, "new: [(name, launches, total_ms, bytes_per_launch)]"),
    {nullptr, nullptr, 0, nullptr}};

PyModuleDef module = {PyModuleDef_HEAD_INIT, "_fs", "PM + COLA hot path on MI355X", -1, methods,
                      nullptr, nullptr, nullptr, nullptr};

}  // namespace

PyMODINIT_FUNC PyInit__fs(void) {
  import_array();
  return PyModule_Create(&module);
}
