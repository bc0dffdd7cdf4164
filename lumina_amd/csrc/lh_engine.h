// Host side of the batched CDA solve engine.
#pragma once
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "lh_types.h"
#include "lh_host.h"
#include "../../include/lumina_hip.h"

namespace lh {

struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

void check_hip(hipError_t e, const char* what);
#define LH_HIP(x) ::lh::check_hip((x), #x)

// Device allocation owned by one engine; freed with it.
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem();
  void alloc(size_t b);
  void release();
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Far-field launch options. block[] is the autotuner's search space: 64, 128
// or 256 targets per work item of the row kernels, kSymBlock for the kernels
// that evaluate one tensor per unordered pair with LDS-staged sources, or
// (FP64 only) kSymRotBlock for the unordered-pair kernel whose sources rotate
// around the wave.
constexpr int kSymBlock = 512;
constexpr int kSymRotBlock = 1024;
struct Tiling {
  int hw_sincos = 1;          // FP32 far field: hardware v_sin/v_cos instead of sincos_f32
  int graph = 1;              // replay the restart cycle as a captured HIP graph
  int block[2] = {0, 0};      // far-field form per precision [FP64, FP32]:
                              // 64, 128, 256, kSymBlock or (FP64) kSymRotBlock;
                              // 0 = timed on the batch at first use
  int ewald_pairs = 0;        // lattice solves: per-pair evaluation even on a grid
  int fft = 1;                // free-space solves filling a regular grid: FFT block-Toeplitz
                              // matvec instead of the pair sum; 1 = from kFftMinDipoles
                              // dipoles, 2 = from 2, 0 = never
  int cgs2_fp32 = 0;          // Arnoldi steps with an FP32 matvec: classical Gram-Schmidt
                              // with reorthogonalisation (1) or one classical pass (0)
  int mgs = 0;                // FP64 Arnoldi steps: modified Gram-Schmidt in the reference's
                              // order (1) instead of CGS2 (0)
  int part32 = 1;             // FP32 matvecs store their far partial slices in FP32
  int run_tiles = 2;          // FP32 one-tensor-per-pair form: source tiles per work item
                              // (a run of one class, near or clean, for one target tile)
  int graph_cycles = 0;       // restart cycles per graph launch: 0 = 1, 2, 4, ... up to
                              // kAutoGraphCycles; N > 0 = N; -1 = enough cycles to reach
                              // max_iter in one launch (at most kMaxGraphCycles)
  int groups = 0;             // solve groups: 1, 2, or 0 = two when the batch has at least
                              // two solves and kGroupMinPairs pair evaluations per matvec
};
constexpr double kGroupMinPairs = 1e8;
constexpr int kAutoGraphCycles = 8;
constexpr int kMaxGraphCycles = 64;

// Device arrays of a sweep's per-wavelength setup (lh_sweep.hip).
struct SweepDevArgs {
  int S, C;
  double n_env;
  const double* lam;
  const int *cmat, *chint, *cls_of;
  const double *cvol, *hdepol, *hrot;
  const int *moff, *mcnt;
  const double *xs, *yr, *yi, *y2r, *y2i;
  int sub_material, sub_retarded;
  double2* cls_ainv;
  double2* delta;
  int* bad;
};

struct SolveResult {
  int status = 0;
  int group = 0;
  int iterations = 0, cycles = 0, matvecs = 0, matvecs32 = 0;
  double residual = 0.0;
  double pair_evals = 0.0;
};

class Engine {
 public:
  Engine(hipStream_t stream) : stream_(stream) {}
  ~Engine();
  void build(std::vector<Geometry>& geoms, std::vector<Problem>& probs, const Tiling& t);
  // Re-targets problem s (same geometry): k, alpha^-1, substrate factor.
  void set_problem(int s, const Problem& p);
  void compute_near_tensors();   // and the Ewald tables of lattice solves
  int num_solves() const { return (int)probs_.size(); }
  int n_of(int s) const { return geoms_[probs_[s].geom].n; }

  // Host <-> slot copies (3n complex).
  void upload_slot(int s, int slot, const lh_c64* h);
  void download_slot(int s, int slot, lh_c64* h);
  void fill_incident(int s, const lh_incident& inc);   // RHS and E slots

  // One y = A x for every solve: x from `in_slot`, y to W.
  void matvec_all(int in_slot, bool fp32);
  // GMRES on every solve, RHS slot -> X slot.
  void solve(double tol, long long max_iter, bool mixed, bool use_graph);
  std::vector<SolveResult> results();
  // C_ext, C_abs, C_sca per solve from the X (moments) and E slots.
  std::vector<double> cross_sections(const std::vector<double>& amplitude);

  double pairs_per_matvec(int s) const;

  // Adjoint C_ext gradients of every solve (lh_adjoint.hip), with the forward
  // moments in slot p_slot, the adjoint solution in X and the incident field
  // in E: 5 doubles per dipole (dC/dRe eps, dC/dIm eps, dC/dr x y z), in
  // solve order. dainv: d alpha^-1 / d eps blocks (9 or 9n complex) of
  // solve 0's layout. positions = false leaves dC/dr at zero.
  void gradients(int p_slot, const double dir[3], double amplitude,
                 const std::vector<lh_c64>& dainv, int dainv_uniform, bool positions,
                 std::vector<double>& out);
  // fields.rs:18-48: the total field at m observation points (host, 3m
  // doubles) from the moments in slot `slot` of solve s; field (3m complex,
  // may be null) and |E|^2 (m, may be null) to the host.
  void field_at_points(int s, int slot, const lh_incident& inc, long long m, const double* pts,
                       lh_c64* field, double* intensity);
  // field_at_points on device buffers (points 3m doubles, field 3m complex
  // and intensity m, each may be null) from the moments mom (3n complex,
  // device; null: slot `slot`), with `scratch` (device, bytes long) as the
  // only working memory: allocates nothing.
  void field_at_points_scratch(int s, int slot, const double2* mom, const lh_incident& inc,
                               long long m, const double* pts, double2* field, double* intensity,
                               char* scratch, size_t bytes);
  const SolveDev& solve_dev(int s) const { return sd_[s]; }
  // fields.rs:156-221: |(I - rhat rhat) F|^2 for m directions (host, 3m
  // doubles) from the moments in slot `slot` of solve s.
  void far_field(int s, int slot, long long m, const double* rhat, double* intensity);
  // The per-wavelength setup of a sweep on the device (lh_sweep.hip): every
  // solve's alpha^-1 blocks, substrate reflection factor and incident field
  // (RHS and E slots). Throws LH_INVALID_ARGUMENT on a singular block.
  void sweep_setup(const SweepSetup& su, const lh_incident& inc);
  hipStream_t stream() const { return stream_; }

  // Far-field launch timing: HIP events around every far-field launch on
  // the engine's stream, recorded into the captured graphs as event nodes
  // when the cycle is replayed from a graph.
  void set_timing(bool on);
  // Sums over the far launches since the last solve() started:
  // {far64_ms, far32_ms, far64_launches, far32_launches}.
  void far_times(double out[4]);
  double last_run_ms = 0.0;
  double tune_ms = 0.0;
  // host round trips (graph launches or host-issued cycles, each followed by
  // one counter read) and restart cycles issued by the last solve()
  int last_launches = 0;
  long long last_cycles_issued = 0;
  int block(int fp32) const { return tiling_.block[fp32]; }
  int groups() const { return ng_; }
  // restart cycles per graph launch (Tiling::graph_cycles)
  void set_graph_cycles(int c) { tiling_.graph_cycles = c; }
  int fft_solves() const { return (int)fft_.size(); }

 private:
  struct EvPair {
    hipEvent_t a, b;
    bool fp32;
  };
  std::vector<EvPair> evs_;                 // host-issued launches' events not yet read
  std::vector<EvPair>* cap_evs_ = nullptr;  // events of the graph being captured
  double far_ms_[2] = {0.0, 0.0};
  long long far_n_[2] = {0, 0};
  bool timing_ = false;
  void clear_events();
  void record(hipEvent_t e, hipStream_t st);
  // adds the elapsed times of evs to the far-field sums (destroy: and frees them)
  void harvest(std::vector<EvPair>& evs, bool destroy);
  // g = -1: every solve on stream_; g = 0, 1: the solves of group g on its
  // stream
  void far_field(int p, int which, int in_slot, int g = -1, hipEvent_t wait_ev = nullptr);
  int part32(bool fp32) const { return (fp32 && tiling_.part32) ? 1 : 0; }
  void matvec(int which, int in_slot, int out_slot, int mode, bool fp32, int g = -1);
  void reduce(int which, int in_slot, int out_slot, int mode, bool fp32, int g);
  // an Arnoldi step after its far field: reduction, Gram-Schmidt, Givens,
  // normalisation
  void step_rest(int j, bool fp32, int g);
  void arnoldi_step(int j, bool fp32, int g);
  void enqueue_cycle(bool mixed, bool early_exit);
  void enqueue_cycle_groups(bool mixed, bool early_exit);

  // Solve groups. With two, the solves are split into two groups of about
  // equal pair work, whose restart cycles are issued on two streams
  // (stream_ and gst_[1]); the far fields of the groups take turns (group 0
  // step j, group 1 step j, group 0 step j + 1, ...), so that one group's
  // reduction, Gram-Schmidt and normalisation run beside the other group's
  // far field. Every solve performs exactly the arithmetic of the one-group
  // engine, in the same order: results are identical.
  int ng_ = 1;
  hipStream_t gst_[2] = {nullptr, nullptr};
  std::vector<int> order_;          // solves in group order
  std::vector<int> group_of_;
  int glo_[3] = {0, 0, 0};          // group g: order_[glo_[g] .. glo_[g + 1])
  int vlo_[3] = {0, 0, 0};          // group g's vector items: vit_[vlo_[g] .. vlo_[g + 1])
  int flo_[2][5][3] = {};           // far lists [precision][free, clean, sub, ewt, ewp]:
                                    // group g's items [flo_[p][l][g], flo_[p][l][g + 1])
  int subn_[2][3] = {};             // one-tensor-per-pair form: group g's near items lead
                                    // its range of the substrate list
  DevMem d_order_;
  hipEvent_t gev_[3] = {nullptr, nullptr, nullptr};   // far field of group 0 / 1 done; join
  hipStream_t gstream(int g) const { return g <= 0 ? stream_ : gst_[1]; }
  int vlo(int g) const { return g < 0 ? 0 : vlo_[g]; }
  int vcount(int g) const { return g < 0 ? (int)vit_.size() : vlo_[g + 1] - vlo_[g]; }
  const int* slist(int g) const { return g < 0 ? nullptr : d_order_.as<int>() + glo_[g]; }
  int scount(int g) const { return g < 0 ? (int)sd_.size() : glo_[g + 1] - glo_[g]; }
  int* counter(int g) const { return d_counter_.as<int>() + (g < 0 ? 0 : g); }
  // stream b waits for what stream a has issued so far (event e)
  void wait_for(hipStream_t b, hipStream_t a, hipEvent_t e);
  void join_groups();
  // K restart cycles captured once as a HIP graph and replayed, keyed by
  // (K, timing events, precision)
  struct CycleGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<EvPair> evs;
  };
  std::map<int, CycleGraph> graphs_;
  CycleGraph& cycle_graph(bool mixed, int cycles);
  void drop_graphs();
  // far work items of precision p for the block size tiling_.block[p]
  void build_far_items(int p);
  // times one all-solve matvec per candidate block size, keeps the fastest
  void autotune(int p);
  int read_counter();
  void upload_solves();

  hipStream_t stream_;
  std::vector<Geometry> geoms_;
  std::vector<Problem> probs_;
  std::vector<SolveDev> sd_;
  std::vector<FarItem> far_free_[2], far_sub_[2];   // [FP64, FP32]
  // free-space row items with no excluded pair (k_far_row<.., NEAR = false>)
  std::vector<FarItem> far_clean_[2];
  DevMem d_far_clean_[2];
  // FP32 one-tensor-per-pair form in runs of source tiles: partial slots per
  // 256-dipole tile (indexed like vit_), read by k_reduce in FP32 matvecs
  bool runs32_ = false;
  std::vector<int> nslot32_;
  DevMem d_nslot32_;
  // bounding boxes of the FP32-rounded positions of every 64-dipole block of
  // each geometry: lo xyz, hi xyz
  std::vector<std::vector<double>> box64_;
  std::vector<VecItem> vit_;
  int n_titems_ = 0;
  Tiling tiling_;

  DevMem d_px_, d_py_, d_pz_, d_ell_col_, d_ell_tid_, d_disp_, d_tab_, d_ainv_, d_vec_,
      d_src64_, d_src32_, d_part_, d_dot_, d_norm_, d_cross_, d_solves_, d_state_, d_counter_,
      d_vit_, d_tis_, d_tiu_;
  DevMem d_far_free_[2], d_far_sub_[2];
  // lattice solves: grid-table and per-pair far items, Ewald records, tables
  std::vector<FarItem> far_ewt_[2], far_ewp_[2];
  DevMem d_far_ewt_[2], d_far_ewp_[2];
  DevMem d_gpack_, d_gizj2_, d_ewl_, d_ewt64_, d_ewt32_, d_ewi_s_, d_ewi_e_;
  DevMem d_gtab_, d_dainv_, d_gout_, d_far_grad_;   // adjoint: near gradient tensors, D
                                                     // blocks, output, far items
  DevMem d_fsrc_, d_fpts_, d_fpart_, d_fout_, d_fint_;   // field maps
  void field_sources(int s, int slot);
  // FFT block-Toeplitz matvec of one full-grid solve (lh_fft.hip)
  struct FftSolve {
    int s = 0;
    int P[3] = {1, 1, 1};      // circulant period per axis
    int nn[3] = {1, 1, 1};     // grid points per axis
    double h[3] = {1, 1, 1};   // grid spacing
    long long len = 0;         // P0 P1 P2
    hipfftHandle plan64 = nullptr, plan32 = nullptr;   // 3D, batch 3, in place
    int n_fcd = 0;             // FCD grid offsets folded into K (fft = 2)
    DevMem K64, K32, X64, X32, pt2dip, dip2pt, fcd_offs;
    ~FftSolve();
  };
  std::vector<std::unique_ptr<FftSolve>> fft_;
  void setup_fft();                       // plans and index maps
  void fft_kernels();                     // spectra of the tensor for the current k
  void fft_matvec(int in_slot, bool fp32, int g, hipStream_t st);
  bool has_pair_work() const;             // any solve on the pair-sum kernels
  long long n_tab_total_ = 0;
  int n_ewitems_ = 0;
  std::vector<int> ewl_cap_;   // records reserved per solve
  std::vector<double> ewl_stage_;   // host staging of a set_problem record update
};

}  // namespace lh
