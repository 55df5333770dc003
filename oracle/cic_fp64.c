/*
 * cic_fp64.c -- FP64 CIC mass assignment and readout (test oracle only).
 *
 * Test infrastructure: only tests/, the smoke check and the bench's CPU
 * baseline load this library, to compare against; the product path never
 * does.
 *
 * The same arithmetic as oracle/cic.py (wdtype float64), in C so that the
 * large-mesh parity tests (16.8M and 134M particles) finish in seconds:
 *
 *   paint    follows pm_assign_cic_density, py/pm.cpp:57-139, with the -1
 *            of clear_density (py/pm.cpp:414-426) applied afterwards;
 *   readout  follows force_at_particle_locations, py/pm.cpp:142-220, for
 *            three meshes at once.
 *
 * Index rules of the reference: the cell is the C (int) cast of x * nc/L
 * (truncation toward zero, pm.cpp:95-97), an index equal to nc becomes 0
 * (pm.cpp:109-111), and the upper neighbour wraps (pm.cpp:113-115).  The
 * particle loop is sequential, so sums are reproducible; tests pin this file
 * against the numpy restatement, which is itself pinned by known answers.
 *
 * Meshes are unpadded (nc, nc, nc) C-order FP64 arrays.
 */
#include <stdint.h>

static inline void cell(double X, int nc, int* i0, int* i1, double* w0, double* w1)
{
  int i = (int)X;
  *w1 = X - (double)i;
  *w0 = 1.0 - *w1;
  if (i >= nc) i = 0;
  int j = i + 1;
  if (j >= nc) j -= nc;
  /* positions more than a cell outside [0, L) (never the case for wrapped
   * particles) index from the far end, as numpy's negative indices do, so
   * no access leaves the mesh */
  i = ((i % nc) + nc) % nc;
  j = ((j % nc) + nc) % nc;
  *i0 = i;
  *i1 = j;
}

/* rho[(ix*nc + iy)*nc + iz] += fac wx wy wz for np particles x[3*i + a];
 * rho is accumulated into (the caller zeroes it and subtracts 1). */
void oracle_cic_paint(const double* x, int64_t np, int nc, double boxsize, double fac,
                      double* rho)
{
  const double dx_inv = (double)nc / boxsize;
  const int64_t n = nc;
  for (int64_t p = 0; p < np; ++p) {
    int ix[2], iy[2], iz[2];
    double wx[2], wy[2], wz[2];
    cell(x[3 * p + 0] * dx_inv, nc, &ix[0], &ix[1], &wx[0], &wx[1]);
    cell(x[3 * p + 1] * dx_inv, nc, &iy[0], &iy[1], &wy[0], &wy[1]);
    cell(x[3 * p + 2] * dx_inv, nc, &iz[0], &iz[1], &wz[0], &wz[1]);
    for (int cx = 0; cx < 2; ++cx)
      for (int cy = 0; cy < 2; ++cy)
        for (int cz = 0; cz < 2; ++cz)
          rho[((int64_t)ix[cx] * n + iy[cy]) * n + iz[cz]] += fac * wx[cx] * wy[cy] * wz[cz];
  }
}

/* out[3*i + a] = sum over the 8 corners of mesh_a * wx wy wz (pm.cpp:142-220)
 * for the three meshes m0, m1, m2. */
void oracle_cic_readout3(const double* x, int64_t np, int nc, double boxsize,
                         const double* m0, const double* m1, const double* m2, double* out)
{
  const double dx_inv = (double)nc / boxsize;
  const int64_t n = nc;
  const double* m[3] = {m0, m1, m2};
  /* particles are independent: threads change no result bit */
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; ++p) {
    int ix[2], iy[2], iz[2];
    double wx[2], wy[2], wz[2];
    cell(x[3 * p + 0] * dx_inv, nc, &ix[0], &ix[1], &wx[0], &wx[1]);
    cell(x[3 * p + 1] * dx_inv, nc, &iy[0], &iy[1], &wy[0], &wy[1]);
    cell(x[3 * p + 2] * dx_inv, nc, &iz[0], &iz[1], &wz[0], &wz[1]);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int cx = 0; cx < 2; ++cx)
      for (int cy = 0; cy < 2; ++cy)
        for (int cz = 0; cz < 2; ++cz) {
          const int64_t c = ((int64_t)ix[cx] * n + iy[cy]) * n + iz[cz];
          const double w = wx[cx] * wy[cy] * wz[cz];
          for (int a = 0; a < 3; ++a)
            if (m[a]) acc[a] += m[a][c] * w;
        }
    for (int a = 0; a < 3; ++a) out[3 * p + a] = acc[a];
  }
}
