// Internal declarations shared by the HIP translation units of libfs2hip.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "fs2hip.h"

// Linear power spectrum: a "k P" table (natural cubic spline of ln P in
// ln k, py/power.cpp:15-54) or the Eisenstein-Hu no-wiggle form.
struct fs2_power {
  int kind = 0;   // 0 table, 1 EH98 no-wiggle
  std::vector<double> logk, logP, m2;
  double om = 0, ob = 0, h = 0, ns = 0, A = 1.0, tcmb = 2.7255;
};

namespace fs2 {

// ---- errors ---------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const char* fmt, ...);
void set_error(int code, const std::string& msg);

#define FS2_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      ::fs2::fail(FS2_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,          \
                  hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

// Runs f, turning exceptions into a status code and the fs2_last_error text.
template <typename F>
int guard(F&& f) {
  try {
    f();
    return FS2_OK;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(FS2_ERR_MEMORY, "host allocation failed");
    return FS2_ERR_MEMORY;
  } catch (const std::exception& e) {
    set_error(FS2_ERR_RUNTIME, e.what());
    return FS2_ERR_RUNTIME;
  }
}

// ---- runtime --------------------------------------------------------------
void device_init();                 // throws Error(FS2_ERR_DEVICE) without a GPU
// The stream every launch and copy of this thread goes to: the library
// stream, or the stream of the innermost live StreamScope.
hipStream_t stream();
hipStream_t library_stream();
int device_cus();                   // compute units of the device
class StreamScope {
 public:
  explicit StreamScope(hipStream_t s);
  ~StreamScope();
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;

 private:
  hipStream_t prev_;
};
void* dmalloc(size_t bytes);        // throws Error(FS2_ERR_MEMORY) on OOM
void dfree(void* p);
void sync();
int loglevel();
void logf(int level, const char* fmt, ...);

// ---- kernel timing ----------------------------------------------------------
// Every launch goes through prof_begin/prof_end; both are no-ops unless
// fs2_profile_enable(1) was called.
int prof_begin(const char* name, double bytes);
void prof_end(int token);
bool profiling_enabled();
void set_profiling(bool on);

template <typename K, typename... Args>
inline void launch(const char* name, double bytes, K kernel, dim3 grid, dim3 block,
                   size_t shmem, Args... args) {
  // a grid sized from an empty input (no particles) has no work to do
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  int tok = prof_begin(name, bytes);
  hipLaunchKernelGGL(kernel, grid, block, shmem, stream(), args...);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    fail(FS2_ERR_DEVICE, "launch of %s failed: %s", name, hipGetErrorString(e));
  prof_end(tok);
}

// ---- mesh layout -----------------------------------------------------------
// In-place r2c grids: n*n rows; a row holds n reals or n/2+1 complex values,
// and rows start every grid_cpitch(n) float2 = grid_rpitch(n) floats, a
// multiple of 128 bytes so no row segment straddles a cache line.
inline int grid_cpitch(int n) { return ((n / 2 + 1) + 15) & ~15; }
inline int grid_rpitch(int n) { return 2 * grid_cpitch(n); }
inline size_t grid_floats(int n) { return (size_t)n * n * grid_rpitch(n); }

// ---- FFT (fft.hip) ------------------------------------------------------------
// Complex values are in natural [kx][ky][kz] order (kz = 0..n/2).  nb > 1
// transforms nb meshes stored back to back (grid_floats(n) apart) in one
// launch per pass.
bool fft_supported(int n);
void fft_r2c_3d(float* grid, int n, int nb = 1);
void fft_c2r_3d(float* grid, int n, int nb = 1);
// Green x gradient fused into the first inverse pass (x pencils):
//   dst = IFFT_x( i f2(k) delta_k ),  f2 = f1 k_axis / |k|^2, Nyquist -> +n/2
// then the y pass and the z c2r pass in place on dst.
void fft_green_c2r_3d(const float* delta_k, float* dst, int n, int axis, float f1);
// The whole PM mesh transform (pm.cpp:431-508): r2c of the density in place
// (z and y passes), then one fused x pass that writes the three Green x
// gradient spectra, then y and z c2r passes of each force mesh.  The density
// mesh is left holding its y-transformed intermediate.
void fft_force_meshes(float* mesh, float* const out[3], int n, float f1, int nb = 1);
// fft_force_meshes in two parts (tuned sizes): the spectra of the three force
// meshes after the x and y inverse passes (half-complex [x][y][kz] rows),
// and the three z c2r passes that turn them into the real meshes.
// z_done: the mesh already holds the z r2c of the density (paint_tiled_r2c).
void fft_force_spectra(float* mesh, float* const out[3], int n, float f1, bool z_done = false);
void fft_force_finish(float* const out[3], int n);
// The FP32 twiddle table e^{-2 pi i k/N}, k < N, of the transforms.
const float2* fft_twiddle_table(int N);
// Transpose of the force-mesh operator: p[0] <- -sum_a M_a p[a] (real meshes,
// M_a = c2r . (i f1 k_a/|k|^2) . r2c); p[1] and p[2] are clobbered.
void fft_adjoint_green(float* const p[3], int n, float f1);

// ---- PM kernels (pm_kernels.hip) -----------------------------------------------
struct PosArgs {
  float* x;            // planes [3][stride]
  int64_t stride;
  int64_t np;
};
void fill_grid(float* grid, int n, float value);
void wrap_positions(PosArgs p, float boxsize);
void paint_atomic(PosArgs p, float* grid, int n, float boxsize, float fac, bool wrap);
void readout3(PosArgs p, const float* g0, const float* g1, const float* g2, float* f,
              int n, float boxsize);
double grid_sum(const float* grid, int n);   // FP64 sum over the n^3 cells

// Tiled paint (paint_tiled.hip): device scratch sized for np particles and
// nbins = paint_tiled_bins(n) bins.
struct PaintScratch {
  unsigned* counts = nullptr;    // [nbins]
  unsigned* offsets = nullptr;   // [nbins + 1]
  unsigned* cursor = nullptr;    // [nbins]
  unsigned* block_sums = nullptr;   // [4096] scan block totals
  float4* rec = nullptr;         // [np] (x, y, z, particle index) in bin order
  unsigned* amax = nullptr;      // [4] per-field max |value| of the adjoint paint
  size_t np = 0, nbins = 0;
};
// Tile extents and batching of the tiled kernels.  Zero extents pick the
// default 8 x 8 x 64 (clamped to n); extents are powers of two dividing n.
// nbox > 1: the particle planes hold nbox boxes of np / nbox particles each,
// box-major, and the meshes are nbox grids of grid_floats(n) floats back to
// back.
struct TileOpts {
  int tx = 0, ty = 0, tz = 0;
  int nbox = 1;
};
size_t paint_tiled_bins(int n, const TileOpts& o = TileOpts());
void tile_extents(int n, const TileOpts& o, int ext[3]);   // resolved tile extents
void ensure_paint_scratch(PaintScratch& s, size_t np, size_t nbins);   // grows on demand
void free_paint_scratch(PaintScratch& s);
void paint_tiled(PosArgs p, float* grid, int n, float boxsize, float fac, bool wrap,
                 bool deterministic, const PaintScratch& s, const TileOpts& o = TileOpts());
// paint_tiled in two parts: the binning (bin counts, scan, scatter of the
// positions into bin-ordered records) and the tile kernel that writes the
// cells from those records.
void paint_tiled_bin(PosArgs p, int n, float boxsize, bool wrap, const PaintScratch& s,
                     const TileOpts& o);
void paint_tiled_cells(int64_t np, float* grid, int n, float boxsize, float fac,
                       bool deterministic, const PaintScratch& s, const TileOpts& o);
// The deterministic tile kernel fused with the z r2c pass of the force
// transform (tiles of full z rows, one row per wavefront): grid receives the
// z-transformed density, exactly what paint_tiled_cells followed by the z
// r2c pass writes, without the real density.
bool paint_r2c_supported(int n, const TileOpts& o);
void paint_tiled_r2c(int64_t np, float* grid, int n, float boxsize, float fac,
                     const PaintScratch& s, const TileOpts& o);
// Readout of the three force meshes at the particles binned by the last
// paint_tiled call (positions unchanged since); forces land at f[i + a*fstride].
void readout_tiled(const float* g0, const float* g1, const float* g2, float* f, int64_t fstride,
                   int64_t np, int n, float boxsize, const PaintScratch& s,
                   const TileOpts& o = TileOpts());
// Fused z c2r + readout for meshes where column_readout_supported(n): the
// paint bins particles into column_tile_opts(n) tiles, and readout_columns
// finishes the force spectra s[a] (fft_force_spectra) row by row in LDS and
// reads the forces out of them, without writing the real force meshes.
bool column_readout_supported(int n);
TileOpts column_tile_opts(int n);
void readout_columns(const float* s0, const float* s1, const float* s2, float* f,
                     int64_t fstride, int64_t np, int n, float boxsize, const PaintScratch& s,
                     const TileOpts& o);
// Adjoint kernels over the bins of the last paint_tiled call (positions
// unchanged since).  paint_values3_tiled: out_a(c) = sum_p W(x_p - c) v_a(p)
// for the value planes v_a = vals[i + a*vstride], int64 fixed-point sums, every
// cell written.  grad_readout_tiled: xbar_p += sum_c grad W(x_p - c)
// [fac rhobar(c) + sum_a fbar_a(p) F_a(c)] (F may be null: rhobar term only).
void paint_values3_tiled(const float* vals, int64_t vstride, int64_t np, float* const out[3],
                         int n, float boxsize, const PaintScratch& s,
                         const TileOpts& o = TileOpts());
void grad_readout_tiled(const float* rhobar, float fac, const float* const F[3],
                        const float* fbar, float* xbar, int64_t stride, int64_t np, int n,
                        float boxsize, const PaintScratch& s, const TileOpts& o = TileOpts());
void power_spectrum(const float* delta_k, int n, double boxsize, int nbins,
                    double dk_bin, double* k, double* P, double* nmodes);
// The same without any allocation: the sums go to the per-GPU `scratch`
// array of pm_kernels.hip (nbins <= FS2_PK_NOALLOC_MAX_BINS).
void power_spectrum_noalloc(const float* delta_k, int n, double boxsize, int nbins,
                            double dk_bin, double* k, double* P, double* nmodes);
// Records the device address of that array for the current device; called
// once per device at library set-up, with the device current.
void pk_scratch_init();
// Builds the twiddle tables of an n^3 mesh's passes, so later transforms of
// that size allocate nothing.
void fft_prepare(int n);

// ---- COLA kernels (cola_kernels.hip) --------------------------------------------
// Coefficients of a COLA kick (kick = 1; cola.cpp:57-66, Float factors and
// -1.5 Omega_m in double) followed by a drift (cola.cpp:106-109, dt double,
// da1 and da2 Float); kick = 0 is the drift alone.
struct KickDrift {
  int kick;
  double m15om;
  float kf, q1, q2;
  double dt;
  float da1, da2;
};
// paint_tiled_bin with the COLA step before it in the same pass over the
// particles (paint_tiled.hip): kick (kd.kick) and drift of p.x / v as
// cola_kick_drift or cola_drift would apply them, the wrap, then the bin
// counts, scan and scatter of the updated positions.
void paint_tiled_bin_kick_drift(PosArgs p, float* v, const float* f, const float* dx1,
                                const float* dx2, const KickDrift& kd, int n, float boxsize,
                                bool wrap, const PaintScratch& s, const TileOpts& o);
void cola_kick(float* v, const float* f, const float* dx1, const float* dx2,
               int64_t stride, int64_t np, double m15om, float kf, float q1, float q2);
void cola_drift(float* x, const float* v, const float* dx1, const float* dx2,
                int64_t stride, int64_t np, double dt, float da1, float da2);
// Best time (ms) of a grid-stride float4 copy of `bytes` (probe.hip).
double copy_probe_ms(uint64_t bytes, int grid, int reps);

// ---- friends-of-friends (host, fof.cpp) -------------------------------------
// One kd-tree node (py/kdtree.h:9-15): children of node i are 2i+1, 2i+2.
struct KdNode {
  int k;               // cut axis of the parent (root: 0)
  float left, right;   // extent of the node's particles along k
  int32_t ibegin, iend;
};
struct FofResult {
  std::vector<KdNode> tree;
  int height = 0, quota = 0;
  std::vector<int32_t> grp;    // lowest index of each particle's group
  std::vector<int32_t> perm;   // index before the tree build, per final index
};
// Builds the tree over positions xyz[a][i] (i < n), reorders nothing in
// place but returns the reference's particle order in perm, and links groups.
void fof_groups(const float* const xyz[3], int64_t n, int64_t np_alloc, double boxsize,
                double linking_length, const double* boxsize3, int quota, FofResult& out);
void fof_counts(const std::vector<int32_t>& grp, std::vector<int32_t>& nfof);
// dst[i] = src[perm[i]] for n elements of 4 or 8 bytes
void gather_u32(uint32_t* dst, const uint32_t* src, const int32_t* perm, int64_t n);
void gather_u64(uint64_t* dst, const uint64_t* src, const int32_t* perm, int64_t n);

// kick then drift in one pass (bitwise equal to the two calls in sequence);
// the drift reads positions from xin and writes them to x (may be equal)
void cola_kick_drift(float* x, const float* xin, float* v, const float* f, const float* dx1,
                     const float* dx2, int64_t stride, int64_t np, double m15om, float kf,
                     float q1, float q2, double dt, float da1, float da2);
void cola_velocity(float* out, const float* v, const float* dx1, const float* dx2,
                   int64_t stride, int64_t np, float Dv, float D2v);
void scale_field(float* out, const float* in, int64_t stride, int64_t np, float s);
// standard (non-COLA) leapfrog, py/leapfrog.cpp:18-94
void leapfrog_initial_velocity(float* v, const float* dx1, const float* dx2, int64_t stride,
                               int64_t np, float Dv, float D2v);
void leapfrog_kick(float* v, const float* f, int64_t stride, int64_t np, double m15om, float kf);
void leapfrog_drift(float* x, const float* v, int64_t stride, int64_t np, double dt);

// ---- LPT kernels (lpt_kernels.hip) --------------------------------------------------
void white_noise(float* grid, int n, uint64_t seed);
void lpt_psi_ij(float* dk, float* const out[6], int n, double boxsize,
                const double* amp /* device, indexed by integer |k|^2 */);
void lpt_div2(float* out, const float* const psi[6], int n);
void lpt_psi12(float* psi1_z_inout, const float* div_k, float* const out5[5], int n,
               double boxsize, const double* amp);
struct LatticeArgs {
  float *x, *v, *dx1, *dx2;
  uint64_t* id;
  int64_t stride;
};
// out.id may be null (no ids written).
void lpt_lattice(LatticeArgs out, const float* const psi1[3], const float* const psi2[3],
                 int n, float dx, float offset, float D1, float D2, float Dv, float D2v,
                 double nmesh3_inv);

// Host orchestration of the 2LPT displacement fields (capi.cpp).
// lpt_amp_table: sqrt(P(k)/V)/n^1.5 indexed by integer |k|^2 for the
// non-Nyquist modes (py/lpt.cpp:328-331).
std::vector<double> lpt_amp_table(int n, double boxsize, const fs2_power* ps);
// S holds the real white noise on entry.  On return psi1 = {G[1], G[2], S}
// and psi2 = {G[3], G[4], G[5]} hold the real first- and (n^3-scaled)
// second-order displacements; G[0] holds the spectrum of div Psi2.
void lpt_displacements(float* S, float* const G[6], int n, double boxsize, const double* d_amp);
// As lpt_displacements up to the six real Psi_ij fields (G[0..5]); S is
// left holding delta_k.
void lpt_psi_fields(float* S, float* const G[6], int n, double boxsize, const double* d_amp);
float lpt_offset_value();   // lattice offset in cells (fs2_lpt_set_offset)
double power_P(const fs2_power* ps, double k);

// ---- adjoint kernels (adjoint_kernels.hip) -------------------------------------------
// Vector-Jacobian products of the differentiable COLA run.  Particle arrays
// are planes [3][stride]; meshes are padded grids of size n.
void adj_paint_values3(const float* x, const float* val, int64_t stride, int64_t np,
                       float* const out[3], int n, float boxsize);
void adj_grad_readout(const float* x, int64_t stride, int64_t np, const float* rhobar, float fac,
                      const float* const F[3], const float* fbar, float* xbar, int n,
                      float boxsize);
void adj_cola_step(const float* xbar, float* vbar, float* dx1bar, float* dx2bar, float* fbar,
                   int64_t n3, float dt, float da1, float da2, float c, float q1, float q2);
void adj_axpy(float* y, const float* x, float a, int64_t n);
void plane_to_grid(const float* plane, float* grid, int n);
void adj_lpt_div2(const float* divbar, float* const psi[6], int n);
void adj_lpt_divbar_k(float* const g[3], int n, double boxsize);
void adj_lpt_noisebar_k(float* const h[3], float* const q[6], int n, double boxsize,
                        const double* amp);

// ---- host numerics (host_numerics.cpp) ------------------------------------------------
namespace host {
void cosmology_init(double omega_m);
bool cosmology_ready();
double omega_m();
double D_growth(double a);
double D2_growth(double a, double D);
double f_growth_rate(double a);
double hubble(double a);
double Dv_growth(double a, double D);
double D2v_growth(double a, double D2);
double drift_Sq(double ai, double af, double av);
double leapfrog_Sphi(double ai, double af);
double leapfrog_Sq(double ai, double af);
}  // namespace host

}  // namespace fs2
