/*
 * fs2hip.h -- C ABI of the MI355X (gfx950) PM + COLA hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No
 * C++ exception and no torch type crosses this boundary.  Every entry point
 * returns FS2_OK (0) or an FS2_ERR_* code; fs2_last_error() gives the message
 * of the last failure on the calling thread.
 *
 * Each declaration names the fs2 reference interface it replaces (paths are
 * relative to the fs2 source tree).  The reference keeps PM, LPT and
 * cosmology state in module globals (py/pm.cpp:22-30, py/lpt.cpp:22-40,
 * py/cosmology.cpp:15-21); here PM state lives in an fs2_pm handle and the
 * CPython layer keeps the process-wide one, as py/py_pm.cpp does.
 *
 * Device layout: particle fields live in HBM as planes [3][np_alloc] of FP32
 * (x, v, dx1, dx2, force) plus a u64 id array.  Host views use the
 * reference's 56-byte AoS record (py/particle.h:12-18) -- see
 * fs2_particles_get_records.  Meshes are padded in place r2c grids
 * [nc][nc][2(nc/2+1)] FP32 like py/fft.cpp:19-50.
 *
 * All device work is enqueued on one library-owned HIP stream (default
 * priority); calls that hand data back to the host synchronise it.
 */
#ifndef FS2HIP_H
#define FS2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mapped by the CPython layer onto the exception types the
 * reference raises (py/error.h:4-14; py/py_pm.cpp, py/py_power.cpp). */
enum {
  FS2_OK = 0,
  FS2_ERR_RUNTIME = 1,    /* RuntimeError   */
  FS2_ERR_ASSERTION = 2,  /* AssertionError */
  FS2_ERR_MEMORY = 3,     /* MemoryError    */
  FS2_ERR_IO = 4,         /* IOError        */
  FS2_ERR_VALUE = 5,      /* ValueError     */
  FS2_ERR_DEVICE = 6,     /* RuntimeError: no GPU, launch or HIP failure */
  FS2_ERR_INDEX = 7       /* IndexError     */
};

/* Paint variants for fs2_pm_set_paint_variant.  All replace the scatter
 * loop of py/pm.cpp:57-139 and give the same density to FP32 rounding.
 * The two tiled variants share the binning: particles are counting-sorted
 * by the tile of their lower CIC cell (wavefront-aggregated atomics), and
 * one workgroup per tile accumulates its cells in LDS and stores each cell
 * once (no mesh clear, no global float atomics). */
enum {
  FS2_PAINT_ATOMIC = 0,   /* tiled, LDS FP32 atomics; summation order varies
                             run to run                                     */
  FS2_PAINT_SORTED = 1,   /* default: tiled, per-tile sums in int64 fixed
                             point (unit 2^-k, fac 2^k <= 2^30: 2^-27 at
                             pm_factor 2); bitwise reproducible and, on
                             gfx950, faster than FP32 LDS atomics           */
  FS2_PAINT_GLOBAL = 2    /* one thread per particle, 8 global FP32 atomics
                             onto a mesh cleared to -1 (pm.cpp's loop as is) */
};

/* Particle fields for fs2_particles_get_field / set_field. */
enum {
  FS2_FIELD_X = 0, FS2_FIELD_DX1 = 1, FS2_FIELD_DX2 = 2,
  FS2_FIELD_V = 3, FS2_FIELD_FORCE = 4
};

typedef struct fs2_particles fs2_particles;
typedef struct fs2_pm fs2_pm;
typedef struct fs2_grid fs2_grid;
typedef struct fs2_power fs2_power;
typedef struct fs2_diff fs2_diff;
typedef struct fs2_batch fs2_batch;

/* ---- library ---------------------------------------------------------- */
/* Message of the last failed call on this thread (no reference counterpart:
 * the reference raises a Python exception, py/error.h:4-14). */
const char* fs2_last_error(void);
/* Library version (no reference counterpart). */
int fs2_version(void);
/* py/py_msg.cpp set_loglevel: 0 debug .. 4 fatal (py/msg.h:4). */
int fs2_set_loglevel(int level);
/* Selects device 0 and creates the library stream (no reference
 * counterpart).  FS2_ERR_DEVICE when no GPU is present.  Called implicitly by
 * every entry point that needs it. */
int fs2_device_init(void);
/* Waits for the library stream (no reference counterpart). */
int fs2_synchronize(void);
/* The library stream as a hipStream_t cast to void* (no reference
 * counterpart). */
int fs2_get_stream(void** stream);

/* ---- cosmology (FP64, host) ------------------------------------------- */
/* py/cosmology.cpp:24-30 cosmology_init */
int fs2_cosmology_init(double omega_m);
/* py/cosmology.cpp:33-38 cosmology_D_growth */
int fs2_cosmology_D_growth(double a, double* out);
/* py/py_cosmology.cpp:29-38: D2(a, D1(a)), py/cosmology.cpp:40-46 */
int fs2_cosmology_D2_growth(double a, double* out);
/* py/cosmology.cpp:71-82 cosmology_f_growth_rate */
int fs2_cosmology_f_growth_rate(double a, double* out);
/* py/cosmology.cpp:48-56 cosmology_Dv_growth(a, D1(a)) */
int fs2_cosmology_Dv_growth(double a, double* out);
/* py/cosmology.cpp:104-108 cosmology_hubble_function */
int fs2_cosmology_hubble(double a, double* out);
/* Omega_m of fs2_cosmology_init (py/cosmology.cpp cosmology_omega_m) */
int fs2_cosmology_omega_m(double* out);
/* py/cola.cpp:57-66: FP64 kick factor, q1 = D1(a_x), q2 = D2 - D1^2. */
int fs2_cola_kick_coefficients(double a_v, double a_x, double a_vel,
                               double* kick_factor, double* q1, double* q2);
/* py/cola.cpp:102-109: dt = Sq(a_x, a_pos, a_v), da1, da2. */
int fs2_cola_drift_coefficients(double a_x, double a_v, double a_pos,
                                double* dt, double* da1, double* da2);

/* ---- linear power spectrum (FP64, host) -------------------------------- */
/* py/power.cpp:15-40 PowerSpectrum(filename); FS2_ERR_IO on a bad file. */
int fs2_power_alloc(const char* filename, fs2_power** out);
/* No reference counterpart: Eisenstein & Hu 1998 no-wiggle P(k) normalised
 * to sigma8 (top hat R = 8 Mpc/h), k in h/Mpc. */
int fs2_power_alloc_eh98(double omega_m, double omega_b, double h, double ns,
                         double sigma8, fs2_power** out);
/* py/py_power.cpp:42-50 py_power_free (the capsule destructor). */
void fs2_power_free(fs2_power* ps);
/* py/py_power.cpp:66-77 _power_n */
int fs2_power_n(const fs2_power* ps, int* n);
/* py/py_power.cpp:120-138 _power_i; FS2_ERR_INDEX out of range. */
int fs2_power_i(const fs2_power* ps, int i, double* k, double* P);
/* py/power.cpp:50-54 PowerSpectrum::P */
int fs2_power_P(const fs2_power* ps, double k, double* P);
/* py/power.cpp:146-171 compute_sigma (tables); top-hat integral (EH98). */
int fs2_power_sigma(const fs2_power* ps, double R, double* sigma);

/* ---- particles ---------------------------------------------------------- */
/* py/particle.cpp:13-27 Particles(np_alloc, boxsize) / py_particles_alloc */
int fs2_particles_alloc(uint64_t np_alloc, double boxsize, fs2_particles** out);
/* py/py_particles.cpp:157-164 py_particles_free */
void fs2_particles_free(fs2_particles* p);
/* py/py_particles.cpp:166-178 _particles_len */
int fs2_particles_len(const fs2_particles* p, uint64_t* np_local);
/* py/py_particles.cpp:358-369 _particles_np_total */
int fs2_particles_np_total(const fs2_particles* p, uint64_t* np_total);
/* py/particle.cpp:35-41 update_np_total (one rank: np_total = np_local) */
int fs2_particles_update_np_total(fs2_particles* p);
/* Particles::np_allocated (py/particle.h:36) */
int fs2_particles_np_allocated(const fs2_particles* p, uint64_t* np_alloc);
/* Particles::boxsize (py/particle.h:38) */
int fs2_particles_boxsize(const fs2_particles* p, double* boxsize);
/* py/py_particles.cpp:403-462 _particles_append: n rows of FP64 x,y,z at
 * x[i*row_stride + c*col_stride] (strides in elements).  Other fields of the
 * new particles are zero.  FS2_ERR_MEMORY past np_alloc. */
int fs2_particles_append(fs2_particles* p, const double* x, uint64_t n,
                         int64_t row_stride, int64_t col_stride);
/* py/py_particles.cpp:386-401 _particles_clear */
int fs2_particles_clear(fs2_particles* p);
/* py/fs/particles.py:61-63 Particles.resize (its _particles_reseize is not
 * registered in the reference): np_local = n <= np_alloc; particles past the
 * old np_local keep whatever the planes hold. */
int fs2_particles_resize(fs2_particles* p, uint64_t n);
/* py/py_particles.cpp:306-355 _particles_x_asarray (x, dx1, dx2) and
 * _particles_force_asarray: (np_local, 3) FP32 row-major into out. */
int fs2_particles_get_field(fs2_particles* p, int field, float* out);
/* Host -> device upload of one field for np_local particles (no reference
 * counterpart; used to stage test states). */
int fs2_particles_set_field(fs2_particles* p, int field, const float* in);
/* py/py_particles.cpp:291-304 _particles_id_asarray */
int fs2_particles_get_id(fs2_particles* p, uint64_t* out);
/* py/py_particles.cpp:217-270 _particles_getitem / _slice rows: 56-byte
 * Particle records (py/particle.h:12-18) for [start, start+count). */
int fs2_particles_get_records(fs2_particles* p, uint64_t start, uint64_t count,
                              void* out);
/* py/py_particles.cpp:371-384 _particles_periodic_wrapup (py/util.h:14-24) */
int fs2_particles_periodic_wrapup(fs2_particles* p);
/* Particles::a_x, a_v, a_f (py/particle.h:33) */
int fs2_particles_get_times(const fs2_particles* p, double* a_x, double* a_v,
                            double* a_f);
/* Sets Particles::a_x, a_v, a_f (py/particle.h:33), as py/lpt.cpp and
 * py/cola.cpp assign them. */
int fs2_particles_set_times(fs2_particles* p, double a_x, double a_v, double a_f);
/* Device planes [3][stride] of each field and the id array (no reference
 * counterpart; for callers that launch their own work, e.g. autograd).  Applies pending wraps,
 * kicks and drifts; from then on fs2_cola_kick and fs2_cola_drift on these particles run at
 * once. */
int fs2_particles_device_fields(fs2_particles* p, float** x, float** v,
                                float** dx1, float** dx2, float** force,
                                uint64_t** id, uint64_t* stride);
/* Device-to-device copy of all state of src into dst (same np_alloc; no
 * reference counterpart). */
int fs2_particles_copy(fs2_particles* dst, const fs2_particles* src);
/* Runs the COLA kick and drift that fs2_cola_kick / fs2_cola_drift recorded
 * and that no later call has applied yet (no reference counterpart: the
 * reference updates the particles inside cola.cpp:48-127 at once; here the
 * next density binning applies them in its own pass).  For timing loops
 * that must end with the particles' last update done. */
int fs2_particles_sync(fs2_particles* p);

/* ---- FFT grids ---------------------------------------------------------- */
/* py/py_fft.cpp:10-30 _fft_alloc(nc): owning padded real/complex grid */
int fs2_grid_alloc(int nc, fs2_grid** out);
/* py/fft.cpp:68-78 FFT::~FFT (the _FFT capsule destructor). */
void fs2_grid_free(fs2_grid* g);
/* FFT::nc (py/fft.h:27) */
int fs2_grid_nc(const fs2_grid* g, int* nc);
/* py/py_fft.cpp:57-85 _fft_set_test_data: fx[ix,iy,iz] = index + 1 */
int fs2_grid_set_test_data(fs2_grid* g);
/* py/py_fft.cpp:87-188 _fft_fx_global_as_array: (nc,nc,nc) FP32 copy */
int fs2_grid_get_real(fs2_grid* g, float* out);
/* Host -> device real grid (no reference counterpart; stages test data). */
int fs2_grid_set_real(fs2_grid* g, const float* in);
/* Half-complex view nc*nc*(nc/2+1) complex64 in [kx][ky][kz] order, the
 * layout of FFT::fk (py/fft.h:29); no reference accessor. */
int fs2_grid_get_complex(fs2_grid* g, float* out);
/* Host -> device half-complex grid (no reference counterpart). */
int fs2_grid_set_complex(fs2_grid* g, const float* in);
/* py/fft.cpp:80-89 execute_forward (r2c, unnormalised, sign -1) */
int fs2_grid_r2c(fs2_grid* g);
/* py/fft.cpp:91-100 execute_inverse (c2r, unnormalised) */
int fs2_grid_c2r(fs2_grid* g);

/* ---- PM ----------------------------------------------------------------- */
/* py/pm.cpp:225-268 pm_init: FS2_ERR_RUNTIME for nc_pm <= 1 or boxsize <= 0 */
int fs2_pm_create(int nc_pm, double pm_factor, double boxsize, fs2_pm** out);
/* py/pm.cpp:270-274 pm_free */
void fs2_pm_destroy(fs2_pm* pm);
/* The parameters of fs2_pm_create (py/pm.cpp:233-238 keeps them for the
 * same-parameters check of pm_init). */
int fs2_pm_params(const fs2_pm* pm, int* nc, double* pm_factor, double* boxsize);
/* Paint variant FS2_PAINT_* (no reference counterpart: the reference has
 * one paint loop, py/pm.cpp:57-139, which FS2_PAINT_GLOBAL follows). */
int fs2_pm_set_paint_variant(fs2_pm* pm, int variant);
/* py/pm_domain.cpp:94-126 pm_domain_send_positions: the periodic wrap;
 * fused into the next paint, or run on its own before any other use of x. */
int fs2_pm_send_positions(fs2_pm* pm, fs2_particles* p);
/* py/pm.cpp:341-355 pm_compute_density; *density is borrowed (valid until
 * the next call on pm), like the NULL-destructor capsule of py_pm.cpp:55. */
int fs2_pm_compute_density(fs2_pm* pm, fs2_particles* p, fs2_grid** density);
/* py/pm.cpp:357-390 pm_check_total_density: FP64 sum of delta; returns
 * FS2_ERR_ASSERTION when |sum| > 10 FLT_EPSILON nc^3. */
int fs2_pm_check_total_density(fs2_pm* pm, double* sum);
/* py/pm.cpp:306-338 pm_compute_force: r2c, Green x gradient fused into the
 * first c2r pass, 3 c2r, CIC readout of all three axes in one pass.
 * Returns early when a_f == a_x; FS2_ERR_RUNTIME when no density is ready. */
int fs2_pm_compute_force(fs2_pm* pm, fs2_particles* p);
/* py/pm_domain.cpp:129-161 pm_domain_get_forces (one rank: no-op) */
int fs2_pm_get_forces(fs2_pm* pm, fs2_particles* p);
/* Real-space force mesh of one axis from the last compute_force, (nc)^3
 * FP32 (no reference counterpart; exposes the fft_pm->fx of pm.cpp:507). */
int fs2_pm_get_force_mesh(fs2_pm* pm, int axis, float* out);
/* No reference counterpart: binned P(k) of the current density mesh (FP64
 * sums).  Must follow compute_density.  Bin b holds b <= |k|/(dk_bin k_f)
 * < b+1; outputs nbins values each (k in h/Mpc). */
int fs2_pm_power_spectrum(fs2_pm* pm, int nbins, double dk_bin, double* k,
                          double* P, double* nmodes);
/* No counterpart in the reference (which has no P(k) estimator at all):
 * fs2_pm_power_spectrum without any allocation, for callers that cannot
 * allocate -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  The FP64 bin sums live in one fixed per-GPU array
 * of 3 x FS2_PK_NOALLOC_MAX_BINS doubles (48 KiB), held until the process
 * ends; nbins above FS2_PK_NOALLOC_MAX_BINS gives FS2_ERR_VALUE.  Each call
 * clears the bins it uses first.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other host threads, other streams) race. */
#define FS2_PK_NOALLOC_MAX_BINS 2048
int fs2_pm_power_spectrum_noalloc(fs2_pm* pm, int nbins, double dk_bin, double* k,
                                  double* P, double* nmodes);

/* ---- COLA ---------------------------------------------------------------- */
/* py/cola.cpp:48-92 cola_kick.  The coefficients are computed (and checked)
 * here; the velocity update is deferred and fused with the next
 * fs2_cola_drift into one pass (bitwise the same result).  Any other call
 * that reads or writes these particles applies a deferred kick first. */
int fs2_cola_kick(fs2_particles* p, double a_vel);
/* py/cola.cpp:94-127 cola_drift (runs a deferred kick in the same pass) */
int fs2_cola_drift(fs2_particles* p, double a_pos);
/* py/cola.cpp:129-148 cola_velocity: (np_local, 3) FP32 into out */
int fs2_cola_velocity(fs2_particles* p, float* out);

/* ---- named allocations ------------------------------------------------- */
typedef struct fs2_mem fs2_mem;
/* py/py_mem.cpp:8-31 _mem_alloc / py/mem.cpp:28-75 Mem(name, size): `size`
 * bytes (rounded up to 16) of device memory, nothing for size 0;
 * FS2_ERR_MEMORY when it does not fit. */
int fs2_mem_alloc(const char* name, uint64_t size, fs2_mem** out);
/* py/py_mem.cpp:42-53 _mem: (size_alloc, size_using) */
int fs2_mem_info(const fs2_mem* m, uint64_t* size_alloc, uint64_t* size_using);
/* py/py_mem.cpp:35-41 py_mem_free (the _Mem capsule destructor). */
void fs2_mem_free(fs2_mem* m);

/* ---- measurement --------------------------------------------------------- */
/* No reference counterpart (roofline harness): best time in ms over reps
 * launches of a grid-stride float4 copy of `bytes` between two fresh device
 * buffers, `grid` blocks of 256 threads, on the library stream. */
int fs2_copy_bandwidth(uint64_t bytes, int grid, int reps, double* ms);

/* ---- friends-of-friends halo finder (CPU, fed by a pinned D2H) -------- */
/* py/py_fof.cpp:14-54 _fof_find_groups / py/fof.cpp:45-99 fof_find_groups:
 * copies the np_local positions to pinned host memory, builds the balanced
 * kd-tree (py/kdtree.cpp:119-190) and links pairs closer than
 * linking_length (periodic, FP32).  boxsize3 (3 doubles, or NULL for the
 * bounding box) only steers the cut directions.  As in the reference the
 * tree build reorders the particles: x, v, dx1, dx2 and id are permuted on
 * the device, the force array is not.  ngroups (optional) receives the
 * number of groups.  The result stays in library state for the calls below. */
int fs2_fof_find_groups(fs2_particles* p, double linking_length,
                        const double* boxsize3, int quota, uint64_t* ngroups);
/* py/py_fof.cpp:56-64 _fof_grp (py/fof.cpp:121-124): grp[i] = lowest index
 * of particle i's group.  *n = np_local of the last run (0: none); out may
 * be NULL to query n. */
int fs2_fof_grp(int32_t* out, uint64_t* n);
/* py/fof.cpp:101-113 fof_compute_nfof: nfof[i] = members of the group whose
 * lowest index is i, else 0.  Same n convention as fs2_fof_grp. */
int fs2_fof_nfof(int32_t* out, uint64_t* n);
/* py/kdtree.h:9-15 KdTree node; py/py_kdtree.cpp:6-36 _kdtree_create_copy
 * and _kdtree_get_height (py/kdtree.cpp:212-220).  Node i has children
 * 2i+1 and 2i+2; there are 2^(height+1) - 1 nodes, sized from the
 * allocation, and the ones below the leaves are zero. */
typedef struct fs2_kdtree_node {
  int32_t k;
  float left, right;
  int32_t ibegin, iend;
} fs2_kdtree_node;
/* py/py_kdtree.cpp:36-39 _kdtree_get_height (py/kdtree.cpp kdtree_get_height). */
int fs2_kdtree_height(uint64_t* height);
/* py/py_kdtree.cpp:6-34 _kdtree_create_copy: the nodes of the last tree. */
int fs2_kdtree_nodes(fs2_kdtree_node* out, uint64_t* n);

/* ---- output ------------------------------------------------------------------ */
/* py/gadget_file.cpp:15-114 / py_write.cpp _write_gadget_binary: Gadget-1
 * binary snapshot (256-byte header, positions, velocities / sqrt(a),
 * 4- or 8-byte ids) of one rank's particles.  h = hubble_param of the
 * header (1.0 in the reference's call).  FS2_ERR_IO if the file cannot be
 * written.  The HDF5 writer lives in libfs2io.so (include/fs2io.h). */
int fs2_write_gadget(fs2_particles* p, const char* filename, int use_long_id, double h);

/* ---- standard leapfrog ----------------------------------------------------- */
/* py/leapfrog.cpp:96-141 SphiStd(ai, af) and SqStd(ai, af): the kick factor
 * int da/(a^2 H) and the drift factor int da/(a^3 H), FP64. */
int fs2_leapfrog_factors(double ai, double af, double* kick_factor, double* dt);
/* py/leapfrog.cpp:18-38 / py_leapfrog.cpp:4-22 _leapfrog_initial_velocity:
 * v = Dv(a) dx1 + D2v(a) dx2, a_v = a */
int fs2_leapfrog_initial_velocity(fs2_particles* p, double a);
/* py/leapfrog.cpp:40-68 / py_leapfrog.cpp:24-44 _leapfrog_kick:
 * v += -1.5 Om f SphiStd(a_v, a_vel) */
int fs2_leapfrog_kick(fs2_particles* p, double a_vel);
/* py/leapfrog.cpp:70-94 / py_leapfrog.cpp:43-61 _leapfrog_drift:
 * x += v SqStd(a_x, a_pos) */
int fs2_leapfrog_drift(fs2_particles* p, double a_pos);

/* ---- 2LPT ---------------------------------------------------------------- */
/* py/py_lpt.cpp:9-42 _lpt(nc, boxsize, a, seed, ps, kind) with kind one of
 * "zeldovich", "2lpt", "cola".  The Gaussian field is seeded white noise
 * (Philox4x32-10 + Box-Muller) instead of N-GenIC's GSL stream. */
int fs2_lpt(int nc, double boxsize, double a, uint64_t seed, const fs2_power* ps,
            const char* kind, fs2_particles** out);
/* py/lpt.cpp:227-230 lpt_set_offset */
int fs2_lpt_set_offset(double offset);
/* py/py_lpt.cpp:57-83 _lpt_set_zeldovich_force: f = D1(a) dx1 */
int fs2_lpt_set_zeldovich_force(fs2_particles* p, double a);
/* The white-noise field of fs2_lpt for (nc, seed), (nc)^3 FP32 on the host
 * (no reference counterpart; exposes the field for checks). */
int fs2_lpt_white_noise(int nc, uint64_t seed, float* out);

/* ---- differentiable COLA run ---------------------------------------------- */
/* No reference counterpart: fs2 has no adjoint.  The forward is the
 * reference's 'cola' pipeline (py/lpt.cpp 2LPT with v0 = 0, then nstep
 * kick/drift steps of py/cola.cpp:48-127 on the schedule of
 * script/cola.py:37-44, PM forces of py/pm.cpp, final CIC density of
 * py/pm.cpp:57-139); the backward is its vector-Jacobian product, which
 * the reference lacks.  All buffers are DEVICE pointers on the library
 * stream (fs2_get_stream); both calls synchronize that stream on return.
 *   noise:     nc_lpt^3 FP32, [ix][iy][iz] contiguous (the white-noise field)
 *   delta:     nc_pm^3 FP32, the final density contrast
 * The handle keeps the checkpoints of its last forward; fs2_diff_backward
 * differentiates that forward (any number of times) and fails with
 * FS2_ERR_RUNTIME if none ran.  Requires fs2_cosmology_init. */
enum {
  FS2_DIFF_AUTO = 0,       /* store when it fits in half of the free device memory */
  FS2_DIFF_RECOMPUTE = 1,  /* keep positions only; re-bin and re-transform per step */
  FS2_DIFF_STORE = 2       /* also keep every step's force meshes and particle bins */
};
int fs2_diff_create(int nc_lpt, int nc_pm, double boxsize, double a0, double a1, int nstep,
                    const fs2_power* ps, int checkpoint, fs2_diff** out);
void fs2_diff_destroy(fs2_diff* d);
int fs2_diff_params(const fs2_diff* d, int* nc_lpt, int* nc_pm, int* nstep, uint64_t* np,
                    int* stores_forces);
/* FS2_PAINT_SORTED (default: tiled deterministic paint and tiled adjoint
 * kernels) or FS2_PAINT_GLOBAL (global atomics and gather kernels). */
int fs2_diff_set_paint_variant(fs2_diff* d, int variant);
int fs2_diff_forward(fs2_diff* d, const float* noise, float* delta);
/* delta_bar: d loss / d delta (nc_pm^3); noise_bar: d loss / d noise (nc_lpt^3). */
int fs2_diff_backward(fs2_diff* d, const float* delta_bar, float* noise_bar);

/* ---- batched ensemble ------------------------------------------------------- */
/* No reference counterpart: the reference runs one box per process
 * (script/cola.py:28-48).  A batch holds nclass classes; class k has
 * nbox[k] independent boxes of nc_lpt[k]^3 particles on nc_pm[k]^3 meshes
 * of side boxsize[k].  Every kernel of a step covers a whole class; each
 * class runs on its own stream, and the nstep-step run of all classes is
 * captured once into a hipGraph and replayed with one launch.  Per box the
 * results are bit-identical to fs2_lpt + fs2_pm_* + fs2_cola_* on that box. */
int fs2_batch_create(int nclass, const int* nc_lpt, const int* nc_pm, const int* nbox,
                     const double* boxsize, fs2_batch** out);
void fs2_batch_destroy(fs2_batch* b);
int fs2_batch_info(const fs2_batch* b, int* nclass, uint64_t* np_total);
/* 2LPT ('cola' kind, v = 0) at scale factor a for every box; seeds holds one
 * seed per box, class-major.  Kept as the state fs2_batch_reset restores. */
int fs2_batch_init_lpt(fs2_batch* b, const fs2_power* ps, double a, const uint64_t* seeds);
/* Paint / readout tile extents of a class (powers of two dividing nc_pm). */
int fs2_batch_set_tile(fs2_batch* b, int cls, int tx, int ty, int tz);
int fs2_batch_get_tile(const fs2_batch* b, int cls, int* tx, int* ty, int* tz);
/* Times every candidate tile shape of every class on its own data and keeps
 * the fastest; best_ms (nclass entries, may be null) receives the time of
 * one bin + paint + readout pass with the chosen shape.  Leaves the state
 * at the initial conditions. */
int fs2_batch_autotune(fs2_batch* b, int reps, double* best_ms);
/* Steps a0 -> a1 in nstep COLA steps (schedule of script/cola.py:37-44). */
int fs2_batch_set_schedule(fs2_batch* b, double a0, double a1, int nstep);
int fs2_batch_capture(fs2_batch* b);
/* Runs the schedule from the current state: use_graph = 1 replays the
 * captured graph (capturing first if needed), 0 launches eagerly.
 * Asynchronous on the library stream. */
int fs2_batch_run(fs2_batch* b, int use_graph);
int fs2_batch_reset(fs2_batch* b);
/* One box's field (FS2_FIELD_*) as (nc_lpt^3, 3) FP32 rows on the host. */
int fs2_batch_get_field(fs2_batch* b, int cls, int box, int field, float* out);

/* ---- kernel timing --------------------------------------------------------- */
/* No reference counterpart (the reference has no timers).  When enabled,
 * every kernel launch is bracketed by HIP events on the library stream and
 * accumulated per kernel name with its algorithmic HBM bytes per launch. */
int fs2_profile_enable(int on);
/* No reference counterpart. */
int fs2_profile_reset(void);
/* No reference counterpart. */
int fs2_profile_count(int* n);
/* No reference counterpart. */
int fs2_profile_get(int i, const char** name, uint64_t* launches,
                    double* total_ms, double* bytes_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* FS2HIP_H */
