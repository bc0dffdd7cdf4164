// Shared host/device data layout of the batched CDA solve engine.
//
// A batch holds S solves (one solve = one dipole set at one wavelength and
// one incident field). Solves that share a geometry share its positions and
// FCD near-pair list; everything else is per solve.
//
// HBM layout (all arrays device-resident for the life of the batch):
//   pos_x/pos_y/pos_z [sum over geometries of n]       FP64 SoA positions
//   near_rp           [sum over geometries of (n+1)]   CSR row pointers (global entry ids)
//   near_col          [sum of nnz]                     source index j (local to the geometry)
//   near_fcd          [sum of nnz]                     1: FCD 27-node average, 0: point G
//   near_t            [6][sum over solves of nnz]      FP64 complex near tensors (xx yy zz xy xz yz)
//   ainv              [sum over solves of 9*n or 9]    FP64 complex alpha^-1 blocks
//   vec               [sum over solves of NSLOT*3n]    FP64 complex Krylov storage, slot-major
//   src64 / src32     [sum over solves of n]           per-matvec source records
//   part              [sum over solves of nsc*3n]      far-field partial sums per source chunk
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
#define LH_NDOT 32             // stride of the per-item dot partials

struct SolveDev {
  int n;            // dipoles
  int m;            // restart length min(30, 3n)
  int pos_off;      // first dipole in pos_x/y/z
  int rp_off;       // first row pointer in near_rp (n+1 entries)
  long long e0;     // near_rp[rp_off] (global id of the first near entry)
  long long t_off;  // first tensor of this solve in near_t (per component)
  long long ainv_off;
  int ainv_uniform;
  int has_sub;
  long long vec_off;   // complex offset of slot 0
  long long src_off;   // first source record
  long long part_off;  // complex offset of chunk 0 partials
  int nsc;             // source chunks (partials per target)
  int vi_off;          // first vector item of this solve
  int nvi;             // vector items of this solve
  int max_iter;
  double k;
  double inv_k;
  double c0;           // k^2 / (4 pi)
  double z2;           // 2 * z_interface
  double dre, dim;     // Fresnel factor Delta
  double rskip2;       // far-field exclusion radius^2 (0 without FCD)
  double cell;         // FCD cell size
  double tol;
};

// Far-field work item: targets [t0, t0 + tile) of solve s against sources
// [c0, c1), written to partial chunk `chunk`.
struct FarItem {
  int s;
  int t0;
  int c0;
  int c1;
  int chunk;
  int pad;
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

// FP64 source record: position and the (optionally image-scaled) input
// vector of one dipole, read by every lane of a wave through the scalar
// cache.
struct Src64 {
  typedef double value_type;
  double x, y, z, pad;
  double v[6];          // x_j (re, im) x 3
  double w[6];          // (-D, -D, +D) o x_j for the image term
};

struct Src32 {
  typedef float value_type;
  double x, y, z;       // FP64: exclusion test and displacement
  float v[6];
  float w[6];
  float pad[2];
};
