// Host-side geometry preparation of the CDA solve engine: near list,
// regular-grid detection and Ewald records. Plain C++ (no HIP), so the CPU
// tests compile and check it with g++ directly.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/lumina_hip.h"

namespace lh {

// One dipole geometry (positions + near list), shared by every solve on it.
struct Geometry {
  int n = 0;
  int use_fcd = 0;
  double cell = 0.0;
  float rskip2f = 0.0f;        // exclusion radius^2 on FP32 positions (0: no near list)
  std::vector<double> x, y, z;
  std::vector<double> cls;     // [n][3] positions for the FCD decision, or empty
  int kw = 0;                  // ELL width (largest near count of a dipole)
  std::vector<int> ell_col;    // [kw][n] source index j
  std::vector<int> ell_tid;    // [kw][n] displacement-table id, -1 = pad (row end)
  std::vector<double> disp;    // [ntab][4] distinct r_i - r_j and the FCD flag
  int ntab = 0;
  long long nnz = 0;
  // regular grid (lattice solves): x = x0 + ix hx etc. for every dipole
  int grid = 0;                // 1 when detect_grid found one
  int gn[3] = {1, 1, 1};       // grid points per axis
  double g0[3] = {0, 0, 0}, gh[3] = {1, 1, 1};
  std::vector<int> gpack;      // ix * NYd * NZd + iy * NZd + iz
  std::vector<int> gizj2;      // 2 iz - (nz - 1)
};

// Finds a regular grid carrying every dipole (each coordinate within 1e-12
// of x0 + integer * h, relative to the coordinate span's magnitude) and
// fills the grid fields; leaves grid = 0 otherwise.
void detect_grid(Geometry& g);

// Free-space solves with at least this many dipoles on a full grid use the
// FFT block-Toeplitz matvec.
constexpr int kFftMinDipoles = 4096;

// For a geometry whose grid (detect_grid) has every node occupied exactly
// once (fft_matvec.rs:20-75: a full box of gn[0] x gn[1] x gn[2] points),
// the grid index (ix, iy, iz) of every dipole, [n][3]; false otherwise.
bool full_grid_indices(const Geometry& g, std::vector<int>& idx);

// On a full grid (idx from full_grid_indices): true when every near pair at
// one grid offset (ix_i - ix_j, iy_i - iy_j, iz_i - iz_j) carries the same
// FCD flag, so that the pair tensor is a function of the offset alone; offs
// then lists the FCD offsets, [m][3], in increasing lexicographic order.
bool fcd_offsets(const Geometry& g, const std::vector<int>& idx, std::vector<int>& offs);

// Builds the near list: every pair j != i whose FP32-rounded positions
// satisfy r2f_exact <= rskip2f, i.e. exactly the pairs the far-field kernels
// skip. Each pair points to a table of distinct displacements r_i - r_j
// (FP64), flagged FCD when sqrt(|r_i - r_j|^2) <= 2 cell_size, the
// reference's own test (greens.rs:102-106), evaluated on g.cls when given.
void build_near_list(Geometry& g);

// Second derivatives y2 of the natural cubic spline through (xs, ys), n >= 2
// points with strictly increasing xs (spline.rs:32-66: the tridiagonal
// sweep of Numerical Recipes with y2[0] = y2[n-1] = 0).
void natural_spline(const double* xs, const double* ys, int n, double* y2);

// What the device needs to set up every wavelength of a sweep
// (runner.rs:150-186): the optical classes of the dipoles (one class = one
// material, volume and depolarisation hint), the permittivity splines and
// the substrate material.
struct SweepSetup {
  std::vector<double> lam;          // [S] wavelength of each solve, nm
  double n_env = 1.0;
  std::vector<int> cls_of;          // [n] class of each dipole; empty: one class
  std::vector<int> cls_material;    // [C] material of the class
  std::vector<int> cls_hint;        // [C] ellipsoid hint, -1: isotropic
  std::vector<double> cls_volume;   // [C] spacing^3
  std::vector<double> hint_depol;   // [H][3] depolarisation factors
  std::vector<double> hint_rot;     // [H][9] rotation, row-major
  std::vector<int> mat_off, mat_n;  // knots of material m: [mat_off[m], mat_off[m] + mat_n[m])
  std::vector<double> xs, yr, yi, y2r, y2i;   // knots, Re/Im eps and their spline y2
  int sub_material = -1;            // substrate material, -1: none
  int sub_retarded = 1;             // retarded (1) or quasi-static (0) reflection factor
};

struct Problem {
  int geom = 0;
  double k = 0.0;
  std::vector<lh_c64> ainv;     // 9 (uniform) or 9n
  int ainv_uniform = 1;
  int has_sub = 0;
  double z_interface = 0.0;
  lh_c64 delta{0.0, 0.0};
  // 2D lattice (LatticeSpec, lattice.rs:29-44) and Bloch vector
  int has_lat = 0;
  double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
  double eta = 0.0;            // <= 0: sqrt(pi / A)
  int n_real = 5, n_recip = 5;
  double kb[3] = {0, 0, 0};
};

// The hardware v_sin_f32 / v_cos_f32 take the phase in revolutions and are
// valid for |x| <= 256 revolutions. The largest far-field phase of a solve
// is k times its geometry's bounding-box diagonal; with a substrate the
// image of a source lies at most 2 diag + 2 |z_interface| + 2 max |z| away.
// FP32 far fields use the hardware functions only while every solve's
// phase stays below kMaxHwRevolutions.
constexpr double kMaxHwRevolutions = 200.0;
double max_phase_revolutions(const Geometry& g, const Problem& p);

// Ewald records of a lattice problem at its k: the real-space images
// (EwReal, m1 outer, m2 inner over [-n_real, n_real]) followed by the
// spectral terms (EwRecip, n1 outer, n2 inner, |q| < 1e-10 left out), as
// doubles; returns the counts and the splitting parameter eta.
void ewald_records(const Problem& p, std::vector<double>& out, int& nreal, int& nrecip,
                   double& eta);

}  // namespace lh
