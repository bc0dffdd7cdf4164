// Shared host/device data layout of the batched CDA solve engine.
//
// A batch holds S solves (one solve = one dipole set at one wavelength and
// one incident field). Solves that share a geometry share its positions and
// near list; everything else is per solve.
//
// HBM layout (all arrays device-resident for the life of the batch):
//   pos_x/pos_y/pos_z   [sum over geometries of n]    FP64 SoA positions
//   posf_x/posf_y/posf_z[sum over geometries of n]    the same rounded to FP32
//                                                     (exclusion test of every kernel)
//   ell_col / ell_tid   [sum over geometries of kw*n] ELL near list: source j and
//                                                     displacement-table id (-1 = pad),
//                                                     slot-major so lanes read consecutively
//   near_tab            [sum over solves of 12*U]     FP64 near tensors (xx yy zz xy xz yz,
//                                                     re/im) per unique displacement
//   tab_disp            [sum over geometries of 4*U]  unique displacement (dx dy dz, fcd flag)
//   ainv                [sum over solves of 9*n or 9] FP64 complex alpha^-1 blocks
//   vec                 [sum over solves of NSLOT*3n] FP64 complex Krylov storage, slot-major
//   src64 / src32       [sum over solves of n]        per-matvec source records
//   part                [sum over solves of nt*3n]    far-field partial sums, one slice per
//                                                     256-dipole tile
//   gpack / gizj2       [sum over geometries of n]    lattice solves on a regular grid:
//                                                     packed grid coordinates of each dipole
//   ewl                 [sum over lattice solves]     Ewald real-space image records and
//                                                     spectral records (EwReal, EwRecip)
//   ewt64 / ewt32       [sum over lattice solves of   periodic tensors per grid offset
//                        (1 + sub) * ewn * 12]        (xx yy zz xy xz yz, re/im): direct
//                                                     table, then the Fresnel image table
#pragma once
#include <stdint.h>

#define LH_MMAX 30             // GMRES restart length (iterative.rs:40)
#define LH_SLOT_W (LH_MMAX + 1)
#define LH_SLOT_X (LH_MMAX + 2)
#define LH_SLOT_R (LH_MMAX + 3)
#define LH_SLOT_RHS (LH_MMAX + 4)
#define LH_SLOT_E (LH_MMAX + 5)   // incident field (for cross sections)
#define LH_NSLOT (LH_MMAX + 6)

#define LH_VB 256              // threads per vector-op block; 256 dipoles = 768 complex per item
#define LH_TILE 256            // far-field tile: 256 targets x 256 sources
#define LH_NDOT 32             // stride of the per-item dot partials

// Participation codes of a launch (which solves a kernel works on): solves
// in the Arnoldi loop, starting a cycle, normalising a new basis vector, all
// of them.
#define kWhichInner 0
#define kWhichStarted 1
#define kWhichNormalise 2
#define kWhichAll 3

struct SolveDev {
  int n;            // dipoles
  int m;            // restart length min(30, 3n)
  int pos_off;      // first dipole in the position arrays
  int kw;           // ELL width of the near list
  long long ell_off;   // first ELL entry (kw*n entries, slot-major)
  long long tab_off;   // first near tensor of this solve (12 doubles each)
  long long dtab_off;  // first unique displacement of the geometry (4 doubles each)
  int ntab;            // unique displacements of the geometry
  int nt;              // 256-dipole tiles = far partial slices
  long long ainv_off;
  int ainv_uniform;
  int has_sub;
  long long vec_off;   // complex offset of slot 0
  long long src_off;   // first source record
  long long part_off;  // complex offset of tile 0 partials
  int vi_off;          // first vector item of this solve
  int nvi;             // vector items of this solve
  int max_iter;
  float rskip2f;       // far-field exclusion radius^2, FP32 (0 without FCD)
  double k;
  double inv_k;
  double c0;           // k^2 / (4 pi)
  double z2;           // 2 * z_interface
  double dre, dim;     // Fresnel factor Delta
  double cell;         // FCD cell size
  double tol;
  // 2D periodic lattice (Ewald) solves; lat = 0 for free-space solves
  int lat;             // 0 none, 1 grid tables, 2 per-pair evaluation
  int ew_nreal, ew_nrecip;   // real-space images, spectral terms
  int ew_nz1;          // grid points along z minus one
  long long ew_off;    // first table entry of this solve (direct; image at + ewn)
  long long ewn;       // entries per table = NXd * NYd * NZd
  long long ewl_off;   // first EwReal record (EwRecip records follow)
  int ew_nyd, ew_nzd;  // table extents along y and z (2 n - 1)
  int ew_center;       // table index of the zero offset
  int fft;             // 1: G x by the FFT block-Toeplitz matvec (one partial slice)
  double ew_eta;
  double ew_h[3];      // grid spacing
  double ew_zimg0;     // 2 z_min - 2 z_interface (image z at iz_i + iz_j = 0)
};

// Real-space image R of the Ewald sum (ewald.rs:119-174) with its Bloch
// phase e^{i k_par . R}; zero = 1 for the R = 0 image.
struct EwReal {
  double x, y, z, zero, br, bi;
};

// Spectral term (ewald.rs:186-273) at Q = k_par + G: (Qx, Qy), q =
// csqrt+(|Q|^2 - k^2) and c = (i / 2A) / q. Terms with |q| < 1e-10 are
// left out of the list.
struct EwRecip {
  double qx, qy, qr, qi, cr, ci;
};

// Far-field work item: targets [t0, t0 + 256) against sources [c0, c1) of
// solve s. Row sums go to partial slice c0/256, column sums (symmetric
// kernels, off-diagonal tiles) to slice t0/256; except in the FP32
// one-tensor-per-pair kernel, whose items are runs of source tiles S <= T
// (T = t0/256): there the row sums of the whole run go to slot fs, and the
// column sums of source tile S to slot T - S - 1 (slot layout of a tile X:
// column sums from the tiles T > X first, then its own runs' row sums).
struct FarItem {
  int s;
  int t0;
  int c0;
  int c1;
  int fs;
};

// Vector work item: dipoles [t0, t0 + LH_VB) of solve s (3 complex each).
struct VecItem {
  int s;
  int t0;
  int b;    // index of this item within its solve
  int pad;
};

// Per-solve GMRES state (iterative.rs:33-160), advanced by the engine's
// per-solve kernels.
struct GState {
  int status;       // 0 running, 1 converged, 2 max iterations, 3 zero rhs
  int inner;        // in the Arnoldi loop of the current cycle
  int j;            // Arnoldi steps completed in this cycle
  int total_iters;
  int cycles;
  int started;      // a cycle was started by the last check (normalise r -> v0)
  int nflag;        // normalise w -> v_{j+1} after this step
  int matvecs;      // matvecs performed for this solve
  int matvecs32;    // of which with the FP32 far field
  int pad;
  double rhs_norm;
  double abs_tol;
  double res_norm;  // last explicit residual norm
  double nscale;    // 1 / h_{j+1,j} (0 on breakdown) or 1 / beta
  double cs[LH_MMAX];
  double sn_re[LH_MMAX], sn_im[LH_MMAX];
  double g_re[LH_MMAX + 1], g_im[LH_MMAX + 1];
  double y_re[LH_MMAX], y_im[LH_MMAX];
  double hp_re[LH_MMAX + 1], hp_im[LH_MMAX + 1];   // this Gram-Schmidt pass
  double hc_re[LH_MMAX + 1], hc_im[LH_MMAX + 1];   // column accumulated over passes
  double H_re[(LH_MMAX + 1) * LH_MMAX], H_im[(LH_MMAX + 1) * LH_MMAX];
};

// FP64 source record of one dipole for one matvec, read by every lane of a
// wave through the scalar cache.
struct Src64 {
  typedef double value_type;
  double x, y, z;       // position
  float xf, yf, zf, pf; // position rounded to FP32 (exclusion test)
  double v[6];          // x_j (re, im) x 3
  double w[6];          // (-D, -D, +D) o x_j for the image term
  double pad;
};

// Source record of one solved dipole for the field maps (position, moment).
struct FieldSrc {
  double x, y, z, pad;
  double v[6];          // p_j (re, im) x 3
};

// FP32 source record.
struct Src32 {
  typedef float value_type;
  float x, y, z, pf;    // position rounded to FP32
  float v[6];
  float w[6];
};
