// Host side of the batched CDA solve engine: batch layout, near lists, the
// GMRES cycle driver and the cross-section reduction.
#include "lh_engine.h"
#include "lh_device.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace lh {

// kernel launchers (lh_pairsum.hip / lh_krylov.hip)
void launch_near_tensors(hipStream_t, int, const SolveDev*, const int*, const int*,
                         const double*, double*);
void launch_prep64(hipStream_t, int, const VecItem*, const SolveDev*, const GState*, int,
                   const double*, const double*, const double*, const double2*, int, Src64*);
void launch_prep32(hipStream_t, int, const VecItem*, const SolveDev*, const GState*, int,
                   const double*, const double*, const double*, const double2*, int, Src32*);
void launch_far64(hipStream_t, bool, int, const FarItem*, const SolveDev*, const GState*, int,
                  const Src64*, double2*, int);
void launch_far32(hipStream_t, bool, int, const FarItem*, const SolveDev*, const GState*, int,
                  const Src32*, double2*, bool, int, int);
void launch_far_clean64(hipStream_t, int, const FarItem*, const SolveDev*, const GState*, int,
                        const Src64*, double2*, int);
void launch_far_clean32(hipStream_t, int, const FarItem*, const SolveDev*, const GState*, int,
                        const Src32*, double2*, bool, int, int);
void launch_far_sym32(hipStream_t, bool, bool, int, const FarItem*, const SolveDev*, const GState*,
                      int, const Src32*, double2*, int);
void launch_far_sym64(hipStream_t, bool, bool, int, const FarItem*, const SolveDev*, const GState*,
                      int, const Src64*, double2*);
void launch_far_symrot64(hipStream_t, int, const FarItem*, const SolveDev*, const GState*, int,
                         const Src64*, double2*);
void launch_reduce(hipStream_t, int, const VecItem*, const SolveDev*, const GState*, int,
                   const int*, const int*, const double*, const double2*, double2*, int, int,
                   const double2*, int, const int*, int, double*);
void launch_vnorm(hipStream_t, int, const VecItem*, const SolveDev*, const GState*, int,
                  const double2*, int, double*);
void launch_scale(hipStream_t, int, const VecItem*, const SolveDev*, const GState*, int,
                  double2*, int, int);
void launch_dots(hipStream_t, int, const VecItem*, const SolveDev*, const GState*,
                 const double2*, int, int, double2*);
void launch_update(hipStream_t, int, const VecItem*, const SolveDev*, const GState*, double2*,
                   int, int, double*);
void launch_xupdate(hipStream_t, int, const VecItem*, const SolveDev*, const GState*, double2*);
void launch_cross(hipStream_t, int, const VecItem*, const SolveDev*, const double2*,
                  const double2*, double*);
void launch_copy_slot(hipStream_t, int, const VecItem*, const SolveDev*, double2*, int, int);
void launch_zero(hipStream_t, int*);
void launch_fin_dots(hipStream_t, int, const SolveDev*, GState*, int, int, const double2*, int,
                     int*, const int*);
void launch_fin_step(hipStream_t, int, const SolveDev*, GState*, int, const double*, int, int*,
                     const int*);
void launch_backsub(hipStream_t, int, GState*, const int*);
void launch_check(hipStream_t, int, const SolveDev*, GState*, const double*, int, int, int*,
                  const int*);

void launch_ewald_table(hipStream_t, int, const SolveDev*, const int*, const long long*,
                        const double*, double*, float*);
void launch_ewp(hipStream_t, int, const FarItem*, const SolveDev*, const GState*, int,
                const Src64*, const double*, double2*, int, int);
void launch_ewt_shift32(hipStream_t, int, const FarItem*, const SolveDev*, const GState*, int,
                        const Src32*, const int*, const int*, const float*, double2*, int);
void launch_ewt_shift64(hipStream_t, int, const FarItem*, const SolveDev*, const GState*, int,
                        const Src64*, const int*, const int*, const double*, double2*);

void launch_near_grad_tensors(hipStream_t, int, const SolveDev*, const int*, const int*,
                              const double*, double*);
void launch_prep_grad(hipStream_t, int, const VecItem*, const SolveDev*, const double*,
                      const double*, const double*, const double2*, int, Src64*);
void launch_grad_far(hipStream_t, int, const FarItem*, const SolveDev*, const Src64*, double2*,
                     int);
void launch_fft_kernel(hipStream_t, long long, const int[3], const int[3], const double[3], double,
                       double2*);
void launch_fft_to32(hipStream_t, long long, const double2*, float2*);
void launch_fft_fcd(hipStream_t, int, const int*, long long, const int[3], const double[3], double,
                    double, double2*);
void launch_fft_scatter64(hipStream_t, long long, const int*, const double2*, double2*);
void launch_fft_scatter32(hipStream_t, long long, const int*, const double2*, float2*);
void launch_fft_mul64(hipStream_t, long long, const double2*, double2*);
void launch_fft_mul32(hipStream_t, long long, const float2*, float2*);
void launch_fft_gather64(hipStream_t, int, long long, const int*, const double2*, double, double2*,
                         long long);
void launch_fft_gather32(hipStream_t, int, long long, const int*, const float2*, double, double2*,
                         long long, int);
void launch_sweep_setup(hipStream_t, const SweepDevArgs&, const SolveDev*, int, double2*);
void launch_sweep_incident(hipStream_t, const SolveDev*, int, int, const double*, const double*,
                           const double*, const double[3], double, const double*, double2*);
void launch_field_pack(hipStream_t, int, const double*, const double*, const double*,
                       const double2*, FieldSrc*);
void launch_field_part(hipStream_t, int, const double*, const FieldSrc*, int, int, double,
                       double2*);
void launch_field_reduce(hipStream_t, int, int, const double*, const double2*, double,
                         const double[3], double, const double*, double2*, double*);
void launch_ff_part(hipStream_t, int, const double*, const FieldSrc*, int, int, double, double2*);
void launch_ff_reduce(hipStream_t, int, int, const double*, const double2*, double*);
void launch_grad_reduce(hipStream_t, int, const VecItem*, const SolveDev*, const int*,
                        const int*, const double*, const double2*, int, const double2*,
                        const double2*, int, long long, const double[3], double, int, double*);

void check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int st = (e == hipErrorOutOfMemory) ? LH_OUT_OF_MEMORY : LH_DEVICE_ERROR;
  throw Error(st, std::string(what) + ": " + hipGetErrorString(e));
}

DevMem::~DevMem() { release(); }

void DevMem::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void DevMem::alloc(size_t b) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
  bytes = b;
  if (b == 0) return;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    throw Error(LH_OUT_OF_MEMORY, "OutOfMemory: requested " + std::to_string(b) +
                                      " bytes, available " + std::to_string(fr));
  }
}

// Copies v to device memory m, keeping m's allocation (and address) when it
// is already large enough -- captured HIP graphs hold these addresses.
template <class T>
static void upload(DevMem& m, const std::vector<T>& v, hipStream_t s) {
  const size_t need = std::max<size_t>(v.size() * sizeof(T), 16);
  if (m.p == nullptr || m.bytes < need) m.alloc(need);
  if (!v.empty())
    LH_HIP(hipMemcpyAsync(m.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// Teardown order: every kernel queued on the engine's stream finishes first;
// then the graph executables and timing events that refer to the engine's
// memory, then the hipFFT plans (and their buffers), then, as members, the
// device memory. The owner destroys the stream last.
Engine::~Engine() {
  (void)hipStreamSynchronize(stream_);
  if (gst_[1]) (void)hipStreamSynchronize(gst_[1]);
  clear_events();
  drop_graphs();
  fft_.clear();
  for (auto& e : gev_)
    if (e) (void)hipEventDestroy(e);
  if (gst_[1]) (void)hipStreamDestroy(gst_[1]);
}

namespace {
// Smallest 2^a 3^b 5^c 7^d >= m: the circulant period of an axis of n grid
// points is smooth7(2 n - 1).
int smooth7(int m) {
  for (int v = std::max(m, 1);; ++v) {
    int r = v;
    for (int f : {2, 3, 5, 7})
      while (r % f == 0) r /= f;
    if (r == 1) return v;
  }
}

long long fft_len_of(const Geometry& g) {
  long long len = 1;
  for (int a = 0; a < 3; ++a) len *= smooth7(2 * g.gn[a] - 1);
  return len;
}

// hipFFT batch distances are 32-bit; three component buffers per batch
constexpr long long kMaxFftLen = 1LL << 28;
}  // namespace

void Engine::build(std::vector<Geometry>& geoms, std::vector<Problem>& probs, const Tiling& t) {
  geoms_ = std::move(geoms);
  probs_ = std::move(probs);
  tiling_ = t;
  // hardware sin/cos in the FP32 far field only within their phase range
  double max_rev = 0.0;
  for (const auto& p : probs_) max_rev = std::max(max_rev, max_phase_revolutions(geoms_[p.geom], p));
  if (max_rev > kMaxHwRevolutions) tiling_.hw_sincos = 0;
  // the one-tensor-per-pair FP32 kernel uses the hardware sin/cos
  if (!tiling_.hw_sincos && tiling_.block[1] == kSymBlock) tiling_.block[1] = LH_TILE;
  if (tiling_.block[1] == kSymRotBlock) tiling_.block[1] = LH_TILE;   // FP64 only
  const int S = (int)probs_.size();
  if (S == 0) throw Error(LH_INVALID_ARGUMENT, "empty batch");

  box64_.assign(geoms_.size(), {});
  for (size_t gi = 0; gi < geoms_.size(); ++gi) {
    const Geometry& g = geoms_[gi];
    auto& b = box64_[gi];
    b.assign(6 * (size_t)((g.n + 63) / 64), 0.0);
    for (int i = 0; i < g.n; ++i) {
      double* e = &b[6 * (size_t)(i / 64)];
      const double c[3] = {(double)(float)g.x[i], (double)(float)g.y[i], (double)(float)g.z[i]};
      for (int a = 0; a < 3; ++a) {
        if (i % 64 == 0 || c[a] < e[a]) e[a] = c[a];
        if (i % 64 == 0 || c[a] > e[3 + a]) e[3 + a] = c[a];
      }
    }
  }

  std::vector<double> px, py, pz, disp;
  std::vector<int> ecol, etid, gpack, gizj2;
  std::vector<int> g_pos(geoms_.size());
  std::vector<long long> g_ell(geoms_.size()), g_disp(geoms_.size());
  for (size_t gi = 0; gi < geoms_.size(); ++gi) {
    Geometry& g = geoms_[gi];
    g_pos[gi] = (int)px.size();
    g_ell[gi] = (long long)ecol.size();
    g_disp[gi] = (long long)disp.size() / 4;
    px.insert(px.end(), g.x.begin(), g.x.end());
    py.insert(py.end(), g.y.begin(), g.y.end());
    pz.insert(pz.end(), g.z.begin(), g.z.end());
    ecol.insert(ecol.end(), g.ell_col.begin(), g.ell_col.end());
    etid.insert(etid.end(), g.ell_tid.begin(), g.ell_tid.end());
    disp.insert(disp.end(), g.disp.begin(), g.disp.end());
    if (g.grid) {
      gpack.insert(gpack.end(), g.gpack.begin(), g.gpack.end());
      gizj2.insert(gizj2.end(), g.gizj2.begin(), g.gizj2.end());
    } else {
      gpack.insert(gpack.end(), g.n, 0);
      gizj2.insert(gizj2.end(), g.n, 0);
    }
  }

  // solve groups: two of about equal pair work (largest solves first, each
  // to the lighter group), solves of a group in index order
  {
    std::vector<double> cost(S);
    double tot = 0.0;
    for (int s = 0; s < S; ++s) {
      const double n = (double)geoms_[probs_[s].geom].n;
      cost[s] = n * n * (probs_[s].has_sub ? 2.0 : 1.0);
      tot += cost[s];
    }
    const int want = tiling_.groups > 0 ? tiling_.groups : (tot >= kGroupMinPairs ? 2 : 1);
    ng_ = (want >= 2 && S >= 2) ? 2 : 1;
    group_of_.assign(S, 0);
    if (ng_ == 2) {
      std::vector<int> by(S);
      for (int s = 0; s < S; ++s) by[s] = s;
      std::stable_sort(by.begin(), by.end(), [&](int a, int b) { return cost[a] > cost[b]; });
      double load[2] = {0.0, 0.0};
      for (int s : by) {
        const int g = load[1] < load[0] ? 1 : 0;
        group_of_[s] = g;
        load[g] += cost[s];
      }
    }
    order_.clear();
    for (int g = 0; g < ng_; ++g) {
      glo_[g] = (int)order_.size();
      for (int s = 0; s < S; ++s)
        if (group_of_[s] == g) order_.push_back(s);
    }
    glo_[ng_] = (int)order_.size();
    if (ng_ == 1) glo_[2] = glo_[1];
    if (ng_ == 2 && !gst_[1]) {
      LH_HIP(hipStreamCreateWithFlags(&gst_[1], hipStreamNonBlocking));
      for (auto& e : gev_) LH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    gst_[0] = stream_;
  }

  sd_.assign(S, SolveDev{});
  vit_.clear();
  long long vec_off = 0, src_off = 0, part_off = 0, tab_off = 0, ainv_off = 0;
  int vi_off = 0;
  std::vector<lh_c64> ainv_all;
  std::vector<int> tis, tiu, ewis;
  std::vector<long long> ewie;
  std::vector<double> ewl;
  long long ew_off = 0;
  ewl_cap_.assign(S, 0);
  // grid tables beyond these sizes fall back to per-pair evaluation
  const long long kMaxEntries = 1LL << 24, kMaxTotalEntries = 1LL << 28;
  for (int s = 0; s < S; ++s) {
    const Problem& p = probs_[s];
    const Geometry& g = geoms_[p.geom];
    SolveDev& d = sd_[s];
    d.n = g.n;
    d.m = std::min(LH_MMAX, 3 * g.n);
    d.pos_off = g_pos[p.geom];
    d.kw = g.kw;
    d.ell_off = g_ell[p.geom];
    d.tab_off = tab_off;
    d.dtab_off = g_disp[p.geom];
    d.ntab = g.ntab;
    d.nt = (g.n + LH_TILE - 1) / LH_TILE;
    d.ainv_off = ainv_off;
    d.ainv_uniform = p.ainv_uniform;
    d.has_sub = p.has_sub;
    d.vec_off = vec_off;
    d.src_off = src_off;
    d.part_off = part_off;
    d.nvi = (g.n + LH_VB - 1) / LH_VB;
    d.max_iter = 0;
    d.rskip2f = g.rskip2f;
    d.k = p.k;
    d.inv_k = 1.0 / p.k;
    d.c0 = p.k * p.k / (4.0 * M_PI);
    d.z2 = 2.0 * p.z_interface;
    d.dre = p.delta.re;
    d.dim = p.delta.im;
    d.cell = g.cell;
    d.tol = 0.0;
    d.fft = 0;
    std::vector<int> gidx;
    if (tiling_.fft && !p.has_lat && !p.has_sub &&
        g.n >= (tiling_.fft == 2 ? 2 : kFftMinDipoles) &&
        full_grid_indices(g, gidx) && fft_len_of(g) <= kMaxFftLen) {
      // FCD flags that depend on the grid offset alone go into the FFT
      // kernel; otherwise the near list keeps the FCD correction
      std::vector<int> offs;
      d.fft = fcd_offsets(g, gidx, offs) ? 2 : 1;
    }
    for (int u = 0; u < g.ntab; u += 256) {
      tis.push_back(s);
      tiu.push_back(u);
    }
    if (p.has_lat) {
      std::vector<double> rec;
      int nr = 0, nq = 0;
      double eta = 0.0;
      ewald_records(p, rec, nr, nq, eta);
      const int cap = nr + (2 * p.n_recip + 1) * (2 * p.n_recip + 1);
      d.ew_nreal = nr;
      d.ew_nrecip = nq;
      d.ew_eta = eta;
      d.ewl_off = (long long)ewl.size() / 6;
      ewl_cap_[s] = cap;
      rec.resize((size_t)cap * 6, 0.0);
      ewl.insert(ewl.end(), rec.begin(), rec.end());
      const long long nyd = 2LL * g.gn[1] - 1, nzd = 2LL * g.gn[2] - 1;
      const long long ewn = (2LL * g.gn[0] - 1) * nyd * nzd;
      const long long need = ewn * (p.has_sub ? 2 : 1);
      if (g.grid && !tiling_.ewald_pairs && ewn <= kMaxEntries &&
          ew_off + need <= kMaxTotalEntries) {
        d.lat = 1;
        d.ew_off = ew_off;
        d.ewn = ewn;
        d.ew_nyd = (int)nyd;
        d.ew_nzd = (int)nzd;
        d.ew_nz1 = g.gn[2] - 1;
        d.ew_center = (int)((g.gn[0] - 1) * nyd * nzd + (g.gn[1] - 1) * nzd + (g.gn[2] - 1));
        for (int c = 0; c < 3; ++c) d.ew_h[c] = g.gh[c];
        d.ew_zimg0 = 2.0 * g.g0[2] - 2.0 * p.z_interface;
        for (long long e = 0; e < need; e += 256) {
          ewis.push_back(s);
          ewie.push_back(e);
        }
        ew_off += need;
      } else {
        d.lat = 2;
      }
    }
    ainv_all.insert(ainv_all.end(), p.ainv.begin(), p.ainv.end());
    ainv_off += (long long)p.ainv.size();
    tab_off += g.ntab;
    vec_off += (long long)LH_NSLOT * 3 * g.n;
    src_off += g.n;
    part_off += (long long)d.nt * 3 * g.n;
  }
  // vector items in group order: each group's items are one range of vit_
  for (int g = 0; g < ng_; ++g) {
    vlo_[g] = vi_off;
    for (int oi = glo_[g]; oi < glo_[g + 1]; ++oi) {
      const int s = order_[oi];
      SolveDev& d = sd_[s];
      d.vi_off = vi_off;
      for (int b = 0; b < d.nvi; ++b) vit_.push_back(VecItem{s, b * LH_VB, b, 0});
      vi_off += d.nvi;
    }
  }
  vlo_[ng_] = vi_off;
  if (ng_ == 1) vlo_[2] = vlo_[1];
  n_titems_ = (int)tis.size();
  n_ewitems_ = (int)ewis.size();
  n_tab_total_ = tab_off;

  hipStream_t st = stream_;
  upload(d_px_, px, st);
  upload(d_py_, py, st);
  upload(d_pz_, pz, st);
  upload(d_ell_col_, ecol, st);
  upload(d_ell_tid_, etid, st);
  upload(d_disp_, disp, st);
  upload(d_ainv_, ainv_all, st);
  upload(d_order_, order_, st);
  for (int q = 0; q < 2; ++q) build_far_items(q);
  upload(d_vit_, vit_, st);
  upload(d_tis_, tis, st);
  upload(d_tiu_, tiu, st);
  upload(d_gpack_, gpack, st);
  upload(d_gizj2_, gizj2, st);
  upload(d_ewl_, ewl, st);
  upload(d_ewi_s_, ewis, st);
  upload(d_ewi_e_, ewie, st);
  d_ewt64_.alloc(sizeof(double) * 12 * (size_t)std::max<long long>(ew_off, 1));
  d_ewt32_.alloc(sizeof(float) * 12 * (size_t)std::max<long long>(ew_off, 1));
  d_tab_.alloc(sizeof(double) * 12 * (size_t)std::max<long long>(tab_off, 1));
  d_vec_.alloc(sizeof(double2) * (size_t)std::max<long long>(vec_off, 1));
  LH_HIP(hipMemsetAsync(d_vec_.p, 0, d_vec_.bytes, st));
  d_src64_.alloc(sizeof(Src64) * (size_t)src_off);
  d_src32_.alloc(sizeof(Src32) * (size_t)src_off);
  d_part_.alloc(sizeof(double2) * (size_t)std::max<long long>(part_off, 1));
  d_dot_.alloc(sizeof(double2) * (size_t)vi_off * LH_NDOT);
  d_norm_.alloc(sizeof(double) * (size_t)vi_off);
  d_cross_.alloc(sizeof(double) * 3 * (size_t)vi_off);
  d_state_.alloc(sizeof(GState) * (size_t)S);
  LH_HIP(hipMemsetAsync(d_state_.p, 0, d_state_.bytes, st));
  d_counter_.alloc(sizeof(int) * 4);
  setup_fft();
  upload_solves();
  compute_near_tensors();
  LH_HIP(hipStreamSynchronize(st));
}

namespace {
// Squared distance between the boxes of the 64-blocks [b0, b1) and [c0, c1).
double box_gap2(const std::vector<double>& box, int b0, int b1, int c0, int c1) {
  double lo[2][3], hi[2][3];
  const int r0[2] = {b0, c0}, r1[2] = {b1, c1};
  for (int t = 0; t < 2; ++t)
    for (int a = 0; a < 3; ++a) {
      lo[t][a] = box[6 * (size_t)r0[t] + a];
      hi[t][a] = box[6 * (size_t)r0[t] + 3 + a];
      for (int q = r0[t] + 1; q < r1[t]; ++q) {
        lo[t][a] = std::min(lo[t][a], box[6 * (size_t)q + a]);
        hi[t][a] = std::max(hi[t][a], box[6 * (size_t)q + 3 + a]);
      }
    }
  double d2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double gap = std::max({0.0, lo[1][a] - hi[0][a], lo[0][a] - hi[1][a]});
    d2 += gap * gap;
  }
  return d2;
}
}  // namespace

void Engine::build_far_items(int p) {
  far_free_[p].clear();
  far_clean_[p].clear();
  far_sub_[p].clear();
  far_ewt_[p].clear();
  far_ewp_[p].clear();
  // before tuning, any block size serves (the far items are rebuilt then);
  // kSymBlock and kSymRotBlock select the one-tensor-per-unordered-pair
  // kernels for the free-space solves (items T >= S of 256 x 256 tiles; the
  // LDS form splits off the items beyond the exclusion radius, the rotating
  // form takes every item with the test), the others keep the row kernel
  // with 256 targets per item
  const bool rot = tiling_.block[p] == kSymRotBlock;
  const bool sym = tiling_.block[p] == kSymBlock || rot;
  const int tb = (tiling_.block[p] > 0 && !sym) ? tiling_.block[p] : LH_TILE;
  // FP32 one-tensor-per-pair form: items are runs of source tiles
  const bool runs = p == 1 && sym && !rot;
  // free-space solves on a substrate in the one-tensor-per-pair form (the
  // LDS-staged kernels, not the rotating one): their items T >= S go to
  // far_sub_, within each group the near items (subn_) before the clean ones
  const bool subsym = sym && !rot;
  std::vector<FarItem> subn, subc;
  auto flush_sub = [&](int g) {
    if (runs)
      for (auto* v : {&subn, &subc})
        std::stable_sort(v->begin(), v->end(), [](const FarItem& a, const FarItem& b) {
          return a.c1 - a.c0 > b.c1 - b.c0;
        });
    subn_[p][g] = (int)subn.size();
    far_sub_[p].insert(far_sub_[p].end(), subn.begin(), subn.end());
    far_sub_[p].insert(far_sub_[p].end(), subc.begin(), subc.end());
    subn.clear();
    subc.clear();
  };
  std::vector<int>& nslot = nslot32_;
  if (runs) {
    nslot.assign(vit_.size(), 0);
    for (const SolveDev& d : sd_)
      std::fill(nslot.begin() + d.vi_off, nslot.begin() + d.vi_off + d.nvi, d.nt);
  }
  const int run_max = std::max(1, tiling_.run_tiles);
  std::vector<FarItem>* lists[5] = {&far_free_[p], &far_clean_[p], &far_sub_[p], &far_ewt_[p],
                                    &far_ewp_[p]};
  // solves in group order: each group's items are one range of every list
  for (int oi = 0; oi < (int)order_.size(); ++oi) {
    for (int g = 0; g < ng_; ++g)
      if (oi == glo_[g]) {
        if (g > 0) flush_sub(g - 1);
        for (int l = 0; l < 5; ++l) flo_[p][l][g] = (int)lists[l]->size();
      }
    const size_t s = (size_t)order_[oi];
    const SolveDev& d = sd_[s];
    if (d.fft) continue;   // G x by FFT
    const bool free_space = d.lat == 0 && !d.has_sub;
    const bool sub_pairs = subsym && d.lat == 0 && d.has_sub;
    // tile pairs whose boxes lie beyond the exclusion radius (with a margin
    // for FP32 rounding of |d|^2) have no excluded pair and no i == j;
    // decided per 256 x 256 tile pair, so that every block size sends the
    // same pairs through the same arithmetic (the one-tensor-per-pair kernels
    // split their items T >= S the same way)
    auto clean = [&](int ti, int tj) {
      const Geometry& g = geoms_[probs_[s].geom];
      const double lim = (double)g.rskip2f * (1.0 + 1e-5);
      const int ta = ti * LH_TILE, tz = std::min(d.n, ta + LH_TILE);
      const int ca = tj * LH_TILE, cz = std::min(d.n, ca + LH_TILE);
      const double gap2 =
          box_gap2(box64_[probs_[s].geom], ta / 64, (tz + 63) / 64, ca / 64, (cz + 63) / 64);
      return gap2 > lim && gap2 > 0.0;
    };
    if (runs && (free_space || sub_pairs)) {
      // target tile T: runs of up to run_max consecutive source tiles S <= T
      // of one class; slots of T's dipoles: column sums from the tiles
      // T' > T (slot T' - T - 1), then one row-sum slot per run
      for (int ti = 0; ti < d.nt; ++ti) {
        int r = 0;
        for (int tj = 0; tj <= ti; ++r) {
          const bool c = clean(ti, tj);
          int tk = tj + 1;
          while (tk <= ti && tk - tj < run_max && clean(ti, tk) == c) ++tk;
          const FarItem f{(int)s, ti * LH_TILE, tj * LH_TILE, std::min(d.n, tk * LH_TILE),
                          d.nt - 1 - ti + r};
          if (sub_pairs)
            (c ? subc : subn).push_back(f);
          else
            (c ? far_clean_[p] : far_free_[p]).push_back(f);
          tj = tk;
        }
        nslot[(size_t)d.vi_off + ti] = d.nt - 1 - ti + r;
      }
      continue;
    }
    // lattice tables: one wave of 64 targets per item (k_ewt_shift)
    const int tbs = d.lat == 1 ? 64 : tb;
    for (int ti = 0; ti < (d.n + tbs - 1) / tbs; ++ti)
      for (int tj = 0; tj < d.nt; ++tj) {
        const FarItem f{(int)s, ti * tbs, tj * LH_TILE, std::min(d.n, (tj + 1) * LH_TILE), tj};
        if (d.lat == 1) far_ewt_[p].push_back(f);
        else if (d.lat == 2) far_ewp_[p].push_back(f);
        else if (sub_pairs) {
          if (ti >= tj) (clean(ti, tj) ? subc : subn).push_back(f);
        } else if (d.has_sub) {
          far_sub_[p].push_back(f);
        }
        else if (!(sym && free_space) || (ti >= tj && !rot))
          (clean(f.t0 / LH_TILE, tj) ? far_clean_[p] : far_free_[p]).push_back(f);
        else if (ti >= tj)
          far_free_[p].push_back(f);
      }
  }
  flush_sub(ng_ - 1);
  for (int l = 0; l < 5; ++l)
    for (int g = ng_; g < 3; ++g) flo_[p][l][g] = (int)lists[l]->size();
  if (runs) {
    // longest runs first within each group, so that the short ones fill the
    // end of its launch
    for (int l = 0; l < 2; ++l)
      for (int g = 0; g < ng_; ++g)
        std::stable_sort(lists[l]->begin() + flo_[p][l][g], lists[l]->begin() + flo_[p][l][g + 1],
                         [](const FarItem& a, const FarItem& b) {
                           return a.c1 - a.c0 > b.c1 - b.c0;
                         });
    upload(d_nslot32_, nslot, stream_);
  }
  if (p == 1) runs32_ = runs;
  drop_graphs();
  upload(d_far_free_[p], far_free_[p], stream_);
  upload(d_far_clean_[p], far_clean_[p], stream_);
  upload(d_far_sub_[p], far_sub_[p], stream_);
  upload(d_far_ewt_[p], far_ewt_[p], stream_);
  upload(d_far_ewp_[p], far_ewp_[p], stream_);
}

void Engine::drop_graphs() {
  for (auto& kv : graphs_) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto& e : kv.second.evs) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
  }
  graphs_.clear();
}

bool Engine::has_pair_work() const {
  for (const auto& d : sd_)
    if (!d.fft) return true;
  return false;
}

void Engine::autotune(int p) {
  if (!has_pair_work()) {
    tiling_.block[p] = LH_TILE;
    return;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const bool was_timing = timing_;
  timing_ = false;
  hipEvent_t a, b;
  LH_HIP(hipEventCreate(&a));
  LH_HIP(hipEventCreate(&b));
  int best = LH_TILE;
  float best_ms = 0.0f;
  std::vector<int> cands = {64, 128, 256};
  if (p == 0 || tiling_.hw_sincos) cands.push_back(kSymBlock);
  if (p == 0) cands.push_back(kSymRotBlock);
  for (int cand : cands) {
    tiling_.block[p] = cand;
    build_far_items(p);
    matvec(3, LH_SLOT_E, LH_SLOT_W, 0, p != 0);   // warm-up
    // the fastest of three timed matvecs: candidates within a few per cent
    // of each other are told apart by their best case, not by noise
    float tot = 0.0f;
    for (int rep = 0; rep < 3; ++rep) {
      LH_HIP(hipEventRecord(a, stream_));
      matvec(3, LH_SLOT_E, LH_SLOT_W, 0, p != 0);
      LH_HIP(hipEventRecord(b, stream_));
      LH_HIP(hipEventSynchronize(b));
      float ms = 0.0f;
      LH_HIP(hipEventElapsedTime(&ms, a, b));
      tot = rep == 0 ? ms : std::min(tot, ms);
    }
    if (cand == cands.front() || tot < best_ms) {
      best = cand;
      best_ms = tot;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  timing_ = was_timing;
  tiling_.block[p] = best;
  build_far_items(p);
  LH_HIP(hipStreamSynchronize(stream_));
  tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Engine::upload_solves() {
  upload(d_solves_, sd_, stream_);
}

void Engine::set_problem(int s, const Problem& p) {
  if (s < 0 || s >= (int)probs_.size()) throw Error(LH_INVALID_ARGUMENT, "bad solve index");
  Problem& q = probs_[s];
  const Geometry& g = geoms_[q.geom];
  if ((int)p.ainv.size() != (p.ainv_uniform ? 9 : 9 * g.n) ||
      p.ainv_uniform != q.ainv_uniform)
    throw Error(LH_INVALID_ARGUMENT, "polarisability layout differs from the session's");
  q.k = p.k;
  q.ainv = p.ainv;
  q.delta = p.delta;
  const bool hw_off = tiling_.hw_sincos && max_phase_revolutions(g, q) > kMaxHwRevolutions;
  SolveDev& d = sd_[s];
  d.k = p.k;
  d.inv_k = 1.0 / p.k;
  d.c0 = p.k * p.k / (4.0 * M_PI);
  d.dre = p.delta.re;
  d.dim = p.delta.im;
  if (q.has_lat) {
    std::vector<double> rec;
    int nr = 0, nq = 0;
    double eta = 0.0;
    ewald_records(q, rec, nr, nq, eta);
    if (nr + nq > ewl_cap_[s]) throw Error(LH_DEVICE_ERROR, "Ewald record overflow");
    d.ew_nreal = nr;
    d.ew_nrecip = nq;
    d.ew_eta = eta;
    ewl_stage_ = rec;
    LH_HIP(hipMemcpyAsync(d_ewl_.as<double>() + 6 * d.ewl_off, ewl_stage_.data(),
                          ewl_stage_.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
  }
  LH_HIP(hipMemcpyAsync(d_ainv_.as<double2>() + d.ainv_off, q.ainv.data(),
                        q.ainv.size() * sizeof(lh_c64), hipMemcpyHostToDevice, stream_));
  LH_HIP(hipMemcpyAsync(d_solves_.as<SolveDev>() + s, &d, sizeof(SolveDev),
                        hipMemcpyHostToDevice, stream_));
  if (hw_off) {
    // the new k takes the FP32 far field beyond the hardware sin/cos range:
    // software sincos_f32 from now on, and off the unordered-pair FP32 form,
    // which uses the hardware functions (rebuilding the items drops the
    // captured graphs, which hold the old launch arguments)
    tiling_.hw_sincos = 0;
    if (tiling_.block[1] == kSymBlock) tiling_.block[1] = LH_TILE;
    build_far_items(1);
  }
  compute_near_tensors();
  LH_HIP(hipStreamSynchronize(stream_));
}

void Engine::compute_near_tensors() {
  launch_near_tensors(stream_, n_titems_, d_solves_.as<SolveDev>(), d_tis_.as<int>(),
                      d_tiu_.as<int>(), d_disp_.as<double>(), d_tab_.as<double>());
  launch_ewald_table(stream_, n_ewitems_, d_solves_.as<SolveDev>(), d_ewi_s_.as<int>(),
                     d_ewi_e_.as<long long>(), d_ewl_.as<double>(), d_ewt64_.as<double>(),
                     d_ewt32_.as<float>());
  LH_HIP(hipGetLastError());
  fft_kernels();
}

void Engine::upload_slot(int s, int slot, const lh_c64* h) {
  const SolveDev& d = sd_[s];
  LH_HIP(hipMemcpyAsync(d_vec_.as<double2>() + d.vec_off + (long long)slot * 3 * d.n, h,
                        sizeof(lh_c64) * 3 * d.n, hipMemcpyHostToDevice, stream_));
}

void Engine::download_slot(int s, int slot, lh_c64* h) {
  const SolveDev& d = sd_[s];
  LH_HIP(hipMemcpyAsync(h, d_vec_.as<double2>() + d.vec_off + (long long)slot * 3 * d.n,
                        sizeof(lh_c64) * 3 * d.n, hipMemcpyDeviceToHost, stream_));
  LH_HIP(hipStreamSynchronize(stream_));
}

// types.rs:60-74 IncidentField::at_position, E0 e exp(i k d.r), for every
// dipole; written to the RHS and E slots.
void Engine::fill_incident(int s, const lh_incident& inc) {
  const Problem& p = probs_[s];
  const Geometry& g = geoms_[p.geom];
  std::vector<lh_c64> e(3 * (size_t)g.n);
  for (int i = 0; i < g.n; ++i) {
    const double kdotr =
        p.k * (inc.direction[0] * g.x[i] + inc.direction[1] * g.y[i] + inc.direction[2] * g.z[i]);
    const double pr = std::cos(kdotr), pi = std::sin(kdotr);
    for (int c = 0; c < 3; ++c) {
      const double er = inc.amplitude * inc.polarisation[c].re;
      const double ei = inc.amplitude * inc.polarisation[c].im;
      e[3 * i + c] = lh_c64{er * pr - ei * pi, er * pi + ei * pr};
    }
  }
  upload_slot(s, LH_SLOT_RHS, e.data());
  upload_slot(s, LH_SLOT_E, e.data());
  LH_HIP(hipStreamSynchronize(stream_));
}

double Engine::pairs_per_matvec(int s) const {
  if (sd_[s].fft) return 0.0;   // no pair evaluations in the FFT matvec
  const double n = geoms_[probs_[s].geom].n;
  // with a lattice the self block is a lattice sum too (assembly.rs:241-243)
  const double direct = probs_[s].has_lat ? n * n : n * (n - 1.0);
  return direct + (probs_[s].has_sub ? n * n : 0.0);
}

// ---------------------------------------------------------------------------
// Matvec and GMRES cycle.
// ---------------------------------------------------------------------------

// The far field of precision p (0 FP64, 1 FP32) for the solves `which`
// selects (of group g, on its stream; g = -1: every solve, on stream_):
// source records, then (after wait_ev, when given) every far-field kernel of
// that precision (and the FFT solves' block-Toeplitz products) into the
// partial slices, bracketed by HIP events when timing.
void Engine::far_field(int p, int which, int in_slot, int g, hipEvent_t wait_ev) {
  hipStream_t st = gstream(g);
  const bool fp32 = p == 1;
  const int nv = vcount(g);
  const VecItem* vi = d_vit_.as<VecItem>() + vlo(g);
  // group g's range of far list l
  auto lo = [&](int l) { return g < 0 ? 0 : flo_[p][l][g]; };
  auto cnt = [&](int l, size_t all) { return g < 0 ? (int)all : flo_[p][l][g + 1] - flo_[p][l][g]; };
  const int nfree = cnt(0, far_free_[p].size()), nclean = cnt(1, far_clean_[p].size());
  const int nsub = cnt(2, far_sub_[p].size());
  const int newt = cnt(3, far_ewt_[p].size()), newp = cnt(4, far_ewp_[p].size());
  const FarItem* ifree = d_far_free_[p].as<FarItem>() + lo(0);
  const FarItem* iclean = d_far_clean_[p].as<FarItem>() + lo(1);
  const FarItem* isub = d_far_sub_[p].as<FarItem>() + lo(2);
  const FarItem* iewt = d_far_ewt_[p].as<FarItem>() + lo(3);
  const FarItem* iewp = d_far_ewp_[p].as<FarItem>() + lo(4);
  const bool rot = tiling_.block[p] == kSymRotBlock;
  const bool sym = tiling_.block[p] == kSymBlock || rot;
  const int blk = (tiling_.block[p] > 0 && !sym) ? tiling_.block[p] : LH_TILE;
  // substrate solves in the one-tensor-per-pair form: per group, the near
  // items of far_sub_ and then the clean ones (fn(near, count, items))
  const bool subsym = sym && !rot;
  auto sub_pair_launches = [&](auto&& fn) {
    for (int gg = g < 0 ? 0 : g; gg <= (g < 0 ? ng_ - 1 : g); ++gg) {
      const int a = flo_[p][2][gg], m = subn_[p][gg], e = flo_[p][2][gg + 1];
      const FarItem* base = d_far_sub_[p].as<FarItem>();
      fn(true, m, base + a);
      fn(false, e - a - m, base + a + m);
    }
  };
  const SolveDev* sd = d_solves_.as<SolveDev>();
  const GState* gs = d_state_.as<GState>();
  double2* vec = d_vec_.as<double2>();
  double2* part = d_part_.as<double2>();
  EvPair ev{nullptr, nullptr, fp32};
  if (timing_) {
    LH_HIP(hipEventCreate(&ev.a));
    LH_HIP(hipEventCreate(&ev.b));
    (cap_evs_ ? *cap_evs_ : evs_).push_back(ev);
  }
  // an FP32 matvec keeps its partial slices in FP32
  const int p32 = part32(fp32);
  if (fp32) {
    launch_prep32(st, nv, vi, sd, gs, which, d_px_.as<double>(), d_py_.as<double>(),
                  d_pz_.as<double>(), vec, in_slot, d_src32_.as<Src32>());
    // the per-pair lattice form is FP64 in either precision
    if (newp > 0)
      launch_prep64(st, nv, vi, sd, gs, which, d_px_.as<double>(), d_py_.as<double>(),
                    d_pz_.as<double>(), vec, in_slot, d_src64_.as<Src64>());
    if (wait_ev) LH_HIP(hipStreamWaitEvent(st, wait_ev, 0));
    if (timing_) record(ev.a, st);
    if (sym) {
      launch_far_sym32(st, true, false, nfree, ifree, sd, gs, which, d_src32_.as<Src32>(), part,
                       p32);
      launch_far_sym32(st, false, false, nclean, iclean, sd, gs, which, d_src32_.as<Src32>(),
                       part, p32);
    } else {
      launch_far32(st, false, nfree, ifree, sd, gs, which, d_src32_.as<Src32>(), part,
                   tiling_.hw_sincos != 0, blk, p32);
      launch_far_clean32(st, nclean, iclean, sd, gs, which, d_src32_.as<Src32>(), part,
                         tiling_.hw_sincos != 0, blk, p32);
    }
    if (subsym)
      sub_pair_launches([&](bool near, int n, const FarItem* f) {
        launch_far_sym32(st, near, true, n, f, sd, gs, which, d_src32_.as<Src32>(), part, p32);
      });
    else
      launch_far32(st, true, nsub, isub, sd, gs, which, d_src32_.as<Src32>(), part,
                   tiling_.hw_sincos != 0, blk, p32);
    launch_ewt_shift32(st, newt, iewt, sd, gs, which, d_src32_.as<Src32>(), d_gpack_.as<int>(),
                       d_gizj2_.as<int>(), d_ewt32_.as<float>(), part, p32);
  } else {
    launch_prep64(st, nv, vi, sd, gs, which, d_px_.as<double>(), d_py_.as<double>(),
                  d_pz_.as<double>(), vec, in_slot, d_src64_.as<Src64>());
    if (wait_ev) LH_HIP(hipStreamWaitEvent(st, wait_ev, 0));
    if (timing_) record(ev.a, st);
    if (rot) {
      launch_far_symrot64(st, nfree, ifree, sd, gs, which, d_src64_.as<Src64>(), part);
    } else if (sym) {
      launch_far_sym64(st, true, false, nfree, ifree, sd, gs, which, d_src64_.as<Src64>(), part);
      launch_far_sym64(st, false, false, nclean, iclean, sd, gs, which, d_src64_.as<Src64>(),
                       part);
    } else {
      launch_far64(st, false, nfree, ifree, sd, gs, which, d_src64_.as<Src64>(), part, blk);
      launch_far_clean64(st, nclean, iclean, sd, gs, which, d_src64_.as<Src64>(), part, blk);
    }
    if (subsym)
      sub_pair_launches([&](bool near, int n, const FarItem* f) {
        launch_far_sym64(st, near, true, n, f, sd, gs, which, d_src64_.as<Src64>(), part);
      });
    else
      launch_far64(st, true, nsub, isub, sd, gs, which, d_src64_.as<Src64>(), part, blk);
    launch_ewt_shift64(st, newt, iewt, sd, gs, which, d_src64_.as<Src64>(), d_gpack_.as<int>(),
                       d_gizj2_.as<int>(), d_ewt64_.as<double>(), part);
  }
  launch_ewp(st, newp, iewp, sd, gs, which, d_src64_.as<Src64>(), d_ewl_.as<double>(), part, blk,
             p32);
  fft_matvec(in_slot, fp32, g, st);
  if (timing_) record(ev.b, st);
}

// y = A x for the solves `which` selects (of group g).
void Engine::reduce(int which, int in_slot, int out_slot, int mode, bool fp32, int g) {
  launch_reduce(gstream(g), vcount(g), d_vit_.as<VecItem>() + vlo(g), d_solves_.as<SolveDev>(),
                d_state_.as<GState>(), which, d_ell_col_.as<int>(), d_ell_tid_.as<int>(),
                d_tab_.as<double>(), d_ainv_.as<double2>(), d_vec_.as<double2>(), in_slot,
                out_slot, d_part_.as<double2>(), part32(fp32),
                fp32 && runs32_ ? d_nslot32_.as<int>() : nullptr, mode, d_norm_.as<double>());
}

void Engine::matvec(int which, int in_slot, int out_slot, int mode, bool fp32, int g) {
  far_field(fp32 ? 1 : 0, which, in_slot, g);
  reduce(which, in_slot, out_slot, mode, fp32, g);
}

void Engine::matvec_all(int in_slot, bool fp32) {
  matvec(3, in_slot, LH_SLOT_W, 0, fp32);
  LH_HIP(hipGetLastError());
  LH_HIP(hipStreamSynchronize(stream_));
}

// Stream b waits for the work stream a has issued so far.
void Engine::wait_for(hipStream_t b, hipStream_t a, hipEvent_t e) {
  LH_HIP(hipEventRecord(e, a));
  LH_HIP(hipStreamWaitEvent(b, e, 0));
}

// stream_ waits for group 1's stream (a no-op with one group).
void Engine::join_groups() {
  if (ng_ == 2) wait_for(stream_, gst_[1], gev_[2]);
}

// Solves counted by the last check or Arnoldi step, over the groups.
int Engine::read_counter() {
  join_groups();
  int c[2] = {0, 0};
  LH_HIP(hipMemcpyAsync(c, d_counter_.p, sizeof(int) * ng_, hipMemcpyDeviceToHost, stream_));
  LH_HIP(hipStreamSynchronize(stream_));
  return c[0] + c[1];
}

// Arnoldi step j of group g after its far field: reduction into W,
// Gram-Schmidt, Givens and the exit test, normalisation of v_{j+1}.
void Engine::step_rest(int j, bool fp32, int g) {
  hipStream_t st = gstream(g);
  const int nv = vcount(g), S = scount(g);
  const VecItem* vi = d_vit_.as<VecItem>() + vlo(g);
  const int* sl = slist(g);
  const SolveDev* sd = d_solves_.as<SolveDev>();
  GState* gs = d_state_.as<GState>();
  double2* vec = d_vec_.as<double2>();
  reduce(0, j, LH_SLOT_W, 0, fp32, g);
  if (tiling_.mgs && !fp32) {
    // modified Gram-Schmidt as the reference (iterative.rs:92-96): for
    // i = 0..j, h_ij = <v_i, w> of the current w, then w -= h_ij v_i, one
    // (dot, sum, update) triple of launches per basis vector
    for (int i = 0; i <= j; ++i) {
      launch_dots(st, nv, vi, sd, gs, vec, i, i, d_dot_.as<double2>());
      launch_fin_dots(st, S, sd, gs, i, i, d_dot_.as<double2>(), 0, counter(g), sl);
      launch_update(st, nv, vi, sd, gs, vec, i, i, d_norm_.as<double>());
    }
  } else {
    // FP64 steps: classical Gram-Schmidt with one reorthogonalisation pass
    // (CGS2), orthogonal to working precision like the reference's MGS
    // (iterative.rs:92-96), with all dots of a pass in one reduction. Steps
    // with an FP32 matvec take one classical pass (FP64 arithmetic): the
    // basis only has to be as good as the FP32 products it spans, and every
    // cycle restarts from an FP64 residual. On the 1,024-solve batch this
    // left the iteration counts unchanged (623.11 mean, 607 converged, C_ext
    // within 8.9e-8 of FP64) and the helix's forward and adjoint counts
    // identical, at half the Gram-Schmidt traffic.
    const int passes = (fp32 && !tiling_.cgs2_fp32) ? 1 : 2;
    for (int pass = 0; pass < passes; ++pass) {
      launch_dots(st, nv, vi, sd, gs, vec, 0, j, d_dot_.as<double2>());
      // pass 0 also clears the counter k_fin_step counts into
      launch_fin_dots(st, S, sd, gs, 0, j, d_dot_.as<double2>(), pass, counter(g), sl);
      launch_update(st, nv, vi, sd, gs, vec, 0, j, d_norm_.as<double>());
    }
  }
  launch_fin_step(st, S, sd, gs, j, d_norm_.as<double>(), fp32 ? 1 : 0, counter(g), sl);
  launch_scale(st, nv, vi, sd, gs, 2, vec, LH_SLOT_W, -1);
}

void Engine::arnoldi_step(int j, bool fp32, int g) {
  far_field(fp32 ? 1 : 0, 0, j, g);
  step_rest(j, fp32, g);
}

void Engine::clear_events() {
  for (auto& e : evs_) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  evs_.clear();
}

void Engine::set_timing(bool on) {
  if (on != timing_) drop_graphs();   // captured graphs hold (or lack) event nodes
  timing_ = on;
}

// An event on stream st; during graph capture an event record node of the
// graph (hipEventRecordExternal), re-recorded at every replay.
void Engine::record(hipEvent_t e, hipStream_t st) {
  if (cap_evs_) LH_HIP(hipEventRecordWithFlags(e, st, hipEventRecordExternal));
  else LH_HIP(hipEventRecord(e, st));
}

void Engine::harvest(std::vector<EvPair>& evs, bool destroy) {
  for (auto& e : evs) {
    float ms = 0.0f;
    LH_HIP(hipEventElapsedTime(&ms, e.a, e.b));
    far_ms_[e.fp32 ? 1 : 0] += ms;
    far_n_[e.fp32 ? 1 : 0] += 1;
    if (destroy) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
  }
  if (destroy) evs.clear();
}

void Engine::far_times(double out[4]) {
  LH_HIP(hipStreamSynchronize(stream_));
  harvest(evs_, true);
  out[0] = far_ms_[0];
  out[1] = far_ms_[1];
  out[2] = (double)far_n_[0];
  out[3] = (double)far_n_[1];
}

// `cycles` restart cycles, every kernel of them, captured once and replayed;
// solves that converge (or leave the Arnoldi loop) inside the graph skip the
// rest of it at the cost of one flag read per work item.
Engine::CycleGraph& Engine::cycle_graph(bool mixed, int cycles) {
  const int key = (cycles << 2) | (timing_ ? 2 : 0) | (mixed ? 1 : 0);
  auto it = graphs_.find(key);
  if (it != graphs_.end()) return it->second;
  CycleGraph cg;
  hipStream_t st = stream_;
  hipGraph_t g = nullptr;
  LH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  cap_evs_ = &cg.evs;
  try {
    // group 1's stream joins the capture (gev_[1] then marks its start),
    // and the capture ends joined
    if (ng_ == 2) {
      wait_for(gst_[1], st, gev_[2]);
      LH_HIP(hipEventRecord(gev_[1], gst_[1]));
    }
    for (int c = 0; c < cycles; ++c) enqueue_cycle(mixed, false);
    join_groups();
  } catch (...) {
    cap_evs_ = nullptr;
    if (ng_ == 2 && hipEventRecord(gev_[2], gst_[1]) == hipSuccess)
      (void)hipStreamWaitEvent(st, gev_[2], 0);
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    for (auto& e : cg.evs) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    throw;
  }
  cap_evs_ = nullptr;
  LH_HIP(hipStreamEndCapture(st, &g));
  const hipError_t e = hipGraphInstantiate(&cg.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    for (auto& ev : cg.evs) {
      (void)hipEventDestroy(ev.a);
      (void)hipEventDestroy(ev.b);
    }
    LH_HIP(e);
  }
  return graphs_.emplace(key, std::move(cg)).first->second;
}

void Engine::solve(double tol, long long max_iter, bool mixed, bool use_graph) {
  clear_events();
  far_ms_[0] = far_ms_[1] = 0.0;
  far_n_[0] = far_n_[1] = 0;
  hipStream_t st = stream_;
  const int nv = (int)vit_.size(), S = (int)sd_.size();
  const VecItem* vi = d_vit_.as<VecItem>();
  SolveDev* sdd = d_solves_.as<SolveDev>();
  GState* gs = d_state_.as<GState>();
  double2* vec = d_vec_.as<double2>();
  for (auto& d : sd_) {
    d.tol = tol;
    d.max_iter = (int)std::min<long long>(max_iter, 1LL << 30);
  }
  upload_solves();
  sdd = d_solves_.as<SolveDev>();
  if (tiling_.block[0] == 0) autotune(0);
  if (mixed && tiling_.block[1] == 0) autotune(1);
  // x = 0, r = rhs - A 0 = rhs (the matvec of the zero vector is exactly 0)
  launch_copy_slot(st, nv, vi, sdd, vec, -1, LH_SLOT_X);
  launch_copy_slot(st, nv, vi, sdd, vec, LH_SLOT_RHS, LH_SLOT_R);
  launch_vnorm(st, nv, vi, sdd, gs, 3, vec, LH_SLOT_R, d_norm_.as<double>());
  launch_zero(st, counter(0));
  if (ng_ == 2) launch_zero(st, counter(1));
  launch_check(st, S, sdd, gs, d_norm_.as<double>(), 1, 0, counter(-1), nullptr);
  LH_HIP(hipGetLastError());
  int running = read_counter();
  // Every cycle performs at least one Arnoldi step of every running solve,
  // so at most max_iter + 1 cycles can start. With graphs, each launch
  // replays K cycles and is followed by one counter read: K doubles from 1
  // up to kAutoGraphCycles by default (a batch that converges early wastes
  // at most that many empty cycles), or is fixed, or covers max_iter.
  long long cycles = 0;
  int launches = 0;
  const bool graph = use_graph && tiling_.graph;
  // host-issued cycles: group 1's stream starts after the set-up above (a
  // captured graph forks it inside)
  if (ng_ == 2 && !graph) {
    wait_for(gst_[1], st, gev_[2]);
    LH_HIP(hipEventRecord(gev_[1], gst_[1]));
  }
  const int whole = (int)std::min<long long>(kMaxGraphCycles, max_iter / LH_MMAX + 1);
  int next = 1;
  while (running > 0 && cycles <= max_iter + 1) {
    if (graph) {
      const int k = tiling_.graph_cycles > 0 ? std::min(tiling_.graph_cycles, kMaxGraphCycles)
                    : tiling_.graph_cycles < 0 ? std::max(whole, 1)
                                               : next;
      CycleGraph& g = cycle_graph(mixed, k);
      LH_HIP(hipGraphLaunch(g.exec, st));
      LH_HIP(hipGetLastError());
      cycles += k;
      next = std::min(2 * next, kAutoGraphCycles);
      running = read_counter();
      if (timing_) harvest(g.evs, false);
    } else {
      enqueue_cycle(mixed, true);
      LH_HIP(hipGetLastError());
      ++cycles;
      running = read_counter();
      if (timing_) harvest(evs_, true);
    }
    ++launches;
  }
  LH_HIP(hipStreamSynchronize(st));
  last_launches = launches;
  last_cycles_issued = cycles;
}

// One restart cycle for every running solve (iterative.rs:69-158): v_0 =
// r / beta, up to 30 Arnoldi steps, back substitution, x += V y, the explicit
// residual r = rhs - A x (FP64 matvec) and its check, which also opens the
// next cycle. With early_exit the host stops issuing Arnoldi steps once no
// solve is left in the loop (one counter read per step); a captured cycle
// issues all 30.
void Engine::enqueue_cycle(bool mixed, bool early_exit) {
  if (ng_ == 2) {
    enqueue_cycle_groups(mixed, early_exit);
    return;
  }
  hipStream_t st = stream_;
  const int nv = (int)vit_.size(), S = (int)sd_.size();
  const VecItem* vi = d_vit_.as<VecItem>();
  const SolveDev* sdd = d_solves_.as<SolveDev>();
  GState* gs = d_state_.as<GState>();
  double2* vec = d_vec_.as<double2>();
  launch_scale(st, nv, vi, sdd, gs, 1, vec, LH_SLOT_R, 0);
  for (int j = 0; j < LH_MMAX; ++j) {
    arnoldi_step(j, mixed, -1);
    if (early_exit) {
      LH_HIP(hipGetLastError());
      if (read_counter() == 0) break;
    }
  }
  launch_backsub(st, S, gs, nullptr);
  launch_xupdate(st, nv, vi, sdd, gs, vec);
  matvec(1, LH_SLOT_X, LH_SLOT_R, 1, false);
  launch_zero(st, counter(0));
  launch_check(st, S, sdd, gs, d_norm_.as<double>(), 0, 1, counter(0), nullptr);
}

// The same cycle with two solve groups on two streams. The far fields take
// turns: group 1's far field of step j waits for group 0's, and group 0's
// far field of step j + 1 for group 1's of step j, so exactly one far field
// runs at a time, while the other group reduces, orthogonalises and
// normalises beside it. Each group's kernels keep their order on its stream.
void Engine::enqueue_cycle_groups(bool mixed, bool early_exit) {
  const SolveDev* sdd = d_solves_.as<SolveDev>();
  GState* gs = d_state_.as<GState>();
  double2* vec = d_vec_.as<double2>();
  hipStream_t s0 = gst_[0], s1 = gst_[1];
  for (int g = 0; g < 2; ++g)
    launch_scale(gstream(g), vcount(g), d_vit_.as<VecItem>() + vlo(g), sdd, gs, 1, vec,
                 LH_SLOT_R, 0);
  // The far fields of group 0 then group 1, each group's remaining work
  // after its own far field. Each group's source records are prepared before
  // it waits for its turn (gev_[1]: group 1's last far field, first recorded
  // when group 1's stream joined; gev_[0]: group 0's far field of this
  // step). Without the turns (the two groups' chains independent inside the
  // graph) the batch workload's step measured 21.5-22.0 s against 18.54 s
  // for one group and 18.06 s with the turns: the far fields then run at
  // once and slow each other.
  auto turns = [&](int p, int which, int in_slot, auto&& rest) {
    far_field(p, which, in_slot, 0, gev_[1]);
    LH_HIP(hipEventRecord(gev_[0], s0));
    far_field(p, which, in_slot, 1, gev_[0]);
    LH_HIP(hipEventRecord(gev_[1], s1));
    rest(0);
    rest(1);
  };
  for (int j = 0; j < LH_MMAX; ++j) {
    turns(mixed ? 1 : 0, 0, j, [&](int g) { step_rest(j, mixed, g); });
    if (early_exit) {
      LH_HIP(hipGetLastError());
      if (read_counter() == 0) break;
    }
  }
  for (int g = 0; g < 2; ++g) {
    launch_backsub(gstream(g), scount(g), gs, slist(g));
    launch_xupdate(gstream(g), vcount(g), d_vit_.as<VecItem>() + vlo(g), sdd, gs, vec);
  }
  turns(0, 1, LH_SLOT_X, [&](int g) {
    reduce(1, LH_SLOT_X, LH_SLOT_R, 1, false, g);
    launch_zero(gstream(g), counter(g));
    launch_check(gstream(g), scount(g), sdd, gs, d_norm_.as<double>(), 0, 1, counter(g),
                 slist(g));
  });
}

void Engine::gradients(int p_slot, const double dir[3], double amplitude,
                       const std::vector<lh_c64>& dainv, int dainv_uniform, bool positions,
                       std::vector<double>& out) {
  hipStream_t st = stream_;
  const int nv = (int)vit_.size();
  long long ntot = 0;
  for (const auto& d : sd_) ntot += d.n;
  upload(d_dainv_, dainv, st);
  const SolveDev* sd = d_solves_.as<SolveDev>();
  if (positions) {
    const size_t gbytes = sizeof(double) * 36 * (size_t)std::max<long long>(n_tab_total_, 1);
    if (d_gtab_.bytes < gbytes) d_gtab_.alloc(gbytes);
    launch_near_grad_tensors(st, n_titems_, sd, d_tis_.as<int>(), d_tiu_.as<int>(),
                             d_disp_.as<double>(), d_gtab_.as<double>());
    launch_prep_grad(st, nv, d_vit_.as<VecItem>(), sd, d_px_.as<double>(), d_py_.as<double>(),
                     d_pz_.as<double>(), d_vec_.as<double2>(), p_slot, d_src64_.as<Src64>());
    // every ordered (256 targets, 256 sources) tile pair of the free-space solves
    std::vector<FarItem> items;
    for (size_t s = 0; s < sd_.size(); ++s) {
      const SolveDev& d = sd_[s];
      if (d.lat != 0 || d.has_sub) continue;
      for (int ti = 0; ti < d.nt; ++ti)
        for (int tj = 0; tj < d.nt; ++tj)
          items.push_back(FarItem{(int)s, ti * LH_TILE, tj * LH_TILE,
                                  std::min(d.n, (tj + 1) * LH_TILE)});
    }
    upload(d_far_grad_, items, st);
    launch_grad_far(st, (int)items.size(), d_far_grad_.as<FarItem>(), sd, d_src64_.as<Src64>(),
                    d_part_.as<double2>(), LH_TILE);
  }
  const size_t obytes = sizeof(double) * 5 * (size_t)std::max<long long>(ntot, 1);
  if (d_gout_.bytes < obytes) d_gout_.alloc(obytes);
  const double K = sd_[0].k / (amplitude * amplitude);
  launch_grad_reduce(st, nv, d_vit_.as<VecItem>(), sd, d_ell_col_.as<int>(),
                     d_ell_tid_.as<int>(), d_gtab_.p ? d_gtab_.as<double>() : nullptr,
                     d_vec_.as<double2>(), p_slot, d_part_.as<double2>(),
                     d_dainv_.as<double2>(), dainv_uniform, 0, dir, K, positions ? 1 : 0,
                     d_gout_.as<double>());
  LH_HIP(hipGetLastError());
  out.resize(5 * (size_t)ntot);
  LH_HIP(hipMemcpyAsync(out.data(), d_gout_.p, sizeof(double) * out.size(),
                        hipMemcpyDeviceToHost, st));
  LH_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Sweep setup (lh_sweep.hip).
// ---------------------------------------------------------------------------

void Engine::sweep_setup(const SweepSetup& su, const lh_incident& inc) {
  hipStream_t st = stream_;
  const int S = (int)sd_.size();
  if ((int)su.lam.size() != S) throw Error(LH_INVALID_ARGUMENT, "one wavelength per solve");
  const int C = (int)su.cls_material.size();
  int max_n = 1;
  for (const auto& d : sd_) max_n = std::max(max_n, d.n);
  // set-up inputs: a few KB, uploaded into scratch that lives for this call
  DevMem lam, cmat, chint, cvol, hdepol, hrot, cls_of, moff, mcnt, xs, yr, yi, y2r, y2i, pol;
  DevMem cls_ainv, delta, bad;
  upload(lam, su.lam, st);
  upload(cmat, su.cls_material, st);
  upload(chint, su.cls_hint, st);
  upload(cvol, su.cls_volume, st);
  upload(hdepol, su.hint_depol, st);
  upload(hrot, su.hint_rot, st);
  upload(cls_of, su.cls_of, st);
  upload(moff, su.mat_off, st);
  upload(mcnt, su.mat_n, st);
  upload(xs, su.xs, st);
  upload(yr, su.yr, st);
  upload(yi, su.yi, st);
  upload(y2r, su.y2r, st);
  upload(y2i, su.y2i, st);
  std::vector<double> polv(6);
  for (int c = 0; c < 3; ++c) {
    polv[2 * c] = inc.polarisation[c].re;
    polv[2 * c + 1] = inc.polarisation[c].im;
  }
  upload(pol, polv, st);
  cls_ainv.alloc(sizeof(double2) * 9 * (size_t)S * (size_t)std::max(C, 1));
  delta.alloc(sizeof(double2) * (size_t)S);
  bad.alloc(sizeof(int) * 4);
  LH_HIP(hipMemsetAsync(bad.p, 0, bad.bytes, st));
  SweepDevArgs a{S, C, su.n_env, lam.as<double>(), cmat.as<int>(), chint.as<int>(),
                 su.cls_of.empty() ? nullptr : cls_of.as<int>(), cvol.as<double>(),
                 hdepol.as<double>(), hrot.as<double>(), moff.as<int>(), mcnt.as<int>(),
                 xs.as<double>(), yr.as<double>(), yi.as<double>(), y2r.as<double>(),
                 y2i.as<double>(), su.sub_material, su.sub_retarded, cls_ainv.as<double2>(),
                 delta.as<double2>(), bad.as<int>()};
  launch_sweep_setup(st, a, d_solves_.as<SolveDev>(), max_n, d_ainv_.as<double2>());
  launch_sweep_incident(st, d_solves_.as<SolveDev>(), S, max_n, d_px_.as<double>(),
                        d_py_.as<double>(), d_pz_.as<double>(), inc.direction, inc.amplitude,
                        pol.as<double>(), d_vec_.as<double2>());
  LH_HIP(hipGetLastError());
  int nbad = 0;
  LH_HIP(hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  std::vector<lh_c64> dl(su.sub_material >= 0 ? S : 0);
  if (!dl.empty())
    LH_HIP(hipMemcpyAsync(dl.data(), delta.p, sizeof(lh_c64) * dl.size(), hipMemcpyDeviceToHost,
                          st));
  LH_HIP(hipStreamSynchronize(st));
  if (nbad > 0) throw Error(LH_INVALID_ARGUMENT, "Singular polarisability tensor");
  // the reflection factors enter the source records through the solve
  // descriptors
  for (int s = 0; s < (int)dl.size(); ++s) {
    probs_[s].delta = dl[s];
    sd_[s].dre = dl[s].re;
    sd_[s].dim = dl[s].im;
  }
  if (!dl.empty()) {
    upload_solves();
    LH_HIP(hipStreamSynchronize(st));
  }
}

// ---------------------------------------------------------------------------
// FFT block-Toeplitz matvec (lh_fft.hip).
// ---------------------------------------------------------------------------

namespace {
void check_fft(hipfftResult r, const char* what) {
  if (r != HIPFFT_SUCCESS)
    throw Error(LH_DEVICE_ERROR, std::string("hipFFT: ") + what + " failed (" +
                                     std::to_string((int)r) + ")");
}

}  // namespace

Engine::FftSolve::~FftSolve() {
  if (plan64) (void)hipfftDestroy(plan64);
  if (plan32) (void)hipfftDestroy(plan32);
}

void Engine::setup_fft() {
  fft_.clear();
  hipStream_t st = stream_;
  for (size_t s = 0; s < sd_.size(); ++s) {
    if (!sd_[s].fft) continue;
    const Geometry& g = geoms_[probs_[s].geom];
    std::vector<int> idx;
    if (!full_grid_indices(g, idx)) throw Error(LH_DEVICE_ERROR, "FFT grid lost");
    auto f = std::make_unique<FftSolve>();
    f->s = (int)s;
    long long len = 1;
    for (int a = 0; a < 3; ++a) {
      f->nn[a] = g.gn[a];
      f->P[a] = smooth7(2 * g.gn[a] - 1);
      f->h[a] = g.gh[a];
      len *= f->P[a];
    }
    f->len = len;
    std::vector<int> pt2dip((size_t)len, -1), dip2pt((size_t)g.n);
    for (int i = 0; i < g.n; ++i) {
      const long long q = ((long long)idx[3 * i] * f->P[1] + idx[3 * i + 1]) * f->P[2] +
                          idx[3 * i + 2];
      pt2dip[(size_t)q] = i;
      dip2pt[(size_t)i] = (int)q;
    }
    upload(f->pt2dip, pt2dip, st);
    upload(f->dip2pt, dip2pt, st);
    if (sd_[s].fft == 2) {
      std::vector<int> offs;
      if (!fcd_offsets(g, idx, offs)) throw Error(LH_DEVICE_ERROR, "FFT FCD offsets lost");
      f->n_fcd = (int)offs.size() / 3;
      upload(f->fcd_offs, offs, st);
    }
    f->K64.alloc(sizeof(double2) * 6 * (size_t)len);
    f->K32.alloc(sizeof(float2) * 6 * (size_t)len);
    f->X64.alloc(sizeof(double2) * 3 * (size_t)len);
    f->X32.alloc(sizeof(float2) * 3 * (size_t)len);
    LH_HIP(hipMemsetAsync(f->X64.p, 0, f->X64.bytes, st));
    LH_HIP(hipMemsetAsync(f->X32.p, 0, f->X32.bytes, st));
    // axes of one point drop out of the transform (same flat layout)
    int dims[3], rank = 0;
    for (int a = 0; a < 3; ++a)
      if (f->P[a] > 1) dims[rank++] = f->P[a];
    check_fft(hipfftPlanMany(&f->plan64, rank, dims, nullptr, 1, (int)len, nullptr, 1, (int)len,
                             HIPFFT_Z2Z, 3),
              "plan Z2Z");
    check_fft(hipfftPlanMany(&f->plan32, rank, dims, nullptr, 1, (int)len, nullptr, 1, (int)len,
                             HIPFFT_C2C, 3),
              "plan C2C");
    check_fft(hipfftSetStream(f->plan64, st), "set stream");
    check_fft(hipfftSetStream(f->plan32, st), "set stream");
    // one transform of each plan outside any graph capture, so that work
    // buffers and kernels are in place before the first captured cycle
    check_fft(hipfftExecC2C(f->plan32, f->X32.as<hipfftComplex>(), f->X32.as<hipfftComplex>(),
                            HIPFFT_FORWARD),
              "exec C2C");
    fft_.push_back(std::move(f));
  }
  LH_HIP(hipStreamSynchronize(st));
}

// The tensor at every circulant offset and its spectrum (six components,
// two batches of three), for the solve's current k; FP32 copy for mixed
// matvecs.
void Engine::fft_kernels() {
  hipStream_t st = stream_;
  for (auto& fp : fft_) {
    FftSolve& f = *fp;
    const SolveDev& d = sd_[f.s];
    double2* K = f.K64.as<double2>();
    launch_fft_kernel(st, f.len, f.P, f.nn, f.h, d.k, K);
    launch_fft_fcd(st, f.n_fcd, f.fcd_offs.as<int>(), f.len, f.P, f.h, d.cell, d.k, K);
    LH_HIP(hipGetLastError());
    check_fft(hipfftSetStream(f.plan64, st), "set stream");
    for (int half = 0; half < 2; ++half) {
      hipfftDoubleComplex* k3 = reinterpret_cast<hipfftDoubleComplex*>(K + 3 * half * f.len);
      check_fft(hipfftExecZ2Z(f.plan64, k3, k3, HIPFFT_FORWARD), "exec Z2Z");
    }
    launch_fft_to32(st, 6 * f.len, K, f.K32.as<float2>());
    LH_HIP(hipGetLastError());
  }
}

// G x of every FFT solve (of group g; g = -1: all) into its partial slice 0
// (k_reduce subtracts it), on stream st.
void Engine::fft_matvec(int in_slot, bool fp32, int g, hipStream_t st) {
  for (auto& fp : fft_) {
    FftSolve& f = *fp;
    if (g >= 0 && group_of_[f.s] != g) continue;
    const SolveDev& d = sd_[f.s];
    check_fft(hipfftSetStream(fp32 ? f.plan32 : f.plan64, st), "set stream");
    const double2* x = d_vec_.as<double2>() + d.vec_off + (long long)in_slot * 3 * d.n;
    double2* part = d_part_.as<double2>();
    const double scale = 1.0 / (double)f.len;
    if (fp32) {
      hipfftComplex* X = f.X32.as<hipfftComplex>();
      launch_fft_scatter32(st, f.len, f.pt2dip.as<int>(), x, f.X32.as<float2>());
      check_fft(hipfftExecC2C(f.plan32, X, X, HIPFFT_FORWARD), "exec C2C");
      launch_fft_mul32(st, f.len, f.K32.as<float2>(), f.X32.as<float2>());
      check_fft(hipfftExecC2C(f.plan32, X, X, HIPFFT_BACKWARD), "exec C2C");
      launch_fft_gather32(st, d.n, f.len, f.dip2pt.as<int>(), f.X32.as<float2>(), scale, part,
                          d.part_off, part32(true));
    } else {
      hipfftDoubleComplex* X = f.X64.as<hipfftDoubleComplex>();
      launch_fft_scatter64(st, f.len, f.pt2dip.as<int>(), x, f.X64.as<double2>());
      check_fft(hipfftExecZ2Z(f.plan64, X, X, HIPFFT_FORWARD), "exec Z2Z");
      launch_fft_mul64(st, f.len, f.K64.as<double2>(), f.X64.as<double2>());
      check_fft(hipfftExecZ2Z(f.plan64, X, X, HIPFFT_BACKWARD), "exec Z2Z");
      launch_fft_gather64(st, d.n, f.len, f.dip2pt.as<int>(), f.X64.as<double2>(), scale, part,
                          d.part_off);
    }
  }
}

// ---------------------------------------------------------------------------
// Field maps (lh_fields.hip).
// ---------------------------------------------------------------------------

namespace {
constexpr long long kFieldBatch = 1LL << 22;   // observation points per launch

// Dipoles per chunk: enough (256-point block, chunk) work groups to fill the
// chip, at least 64 dipoles per chunk.
int field_chunk(long long m, int n) {
  const long long pblocks = (m + 255) / 256;
  long long nchunk = std::max<long long>(1, (2048 + pblocks - 1) / pblocks);
  nchunk = std::min<long long>(nchunk, (n + 63) / 64);
  return (int)((n + nchunk - 1) / nchunk);
}

void ensure(DevMem& m, size_t bytes) {
  if (m.p == nullptr || m.bytes < bytes) m.alloc(std::max<size_t>(bytes, 16));
}
}  // namespace

void Engine::field_sources(int s, int slot) {
  const SolveDev& d = sd_[s];
  ensure(d_fsrc_, sizeof(FieldSrc) * (size_t)d.n);
  launch_field_pack(stream_, d.n, d_px_.as<double>() + d.pos_off, d_py_.as<double>() + d.pos_off,
                    d_pz_.as<double>() + d.pos_off,
                    d_vec_.as<double2>() + d.vec_off + (long long)slot * 3 * d.n,
                    d_fsrc_.as<FieldSrc>());
}

void Engine::field_at_points(int s, int slot, const lh_incident& inc, long long m,
                             const double* pts, lh_c64* field, double* intensity) {
  const SolveDev& d = sd_[s];
  hipStream_t st = stream_;
  field_sources(s, slot);
  double pol[6];
  for (int c = 0; c < 3; ++c) {
    pol[2 * c] = inc.polarisation[c].re;
    pol[2 * c + 1] = inc.polarisation[c].im;
  }
  for (long long q0 = 0; q0 < m; q0 += kFieldBatch) {
    const int mb = (int)std::min(kFieldBatch, m - q0);
    const int chunk = field_chunk(mb, d.n);
    const int nchunk = (d.n + chunk - 1) / chunk;
    ensure(d_fpts_, sizeof(double) * 3 * (size_t)mb);
    ensure(d_fpart_, sizeof(double2) * 3 * (size_t)mb * nchunk);
    ensure(d_fout_, sizeof(double2) * 3 * (size_t)mb);
    ensure(d_fint_, sizeof(double) * (size_t)mb);
    LH_HIP(hipMemcpyAsync(d_fpts_.p, pts + 3 * q0, sizeof(double) * 3 * (size_t)mb,
                          hipMemcpyHostToDevice, st));
    launch_field_part(st, mb, d_fpts_.as<double>(), d_fsrc_.as<FieldSrc>(), d.n, chunk, d.k,
                      d_fpart_.as<double2>());
    launch_field_reduce(st, mb, nchunk, d_fpts_.as<double>(), d_fpart_.as<double2>(), d.k,
                        inc.direction, inc.amplitude, pol, d_fout_.as<double2>(),
                        d_fint_.as<double>());
    LH_HIP(hipGetLastError());
    if (field)
      LH_HIP(hipMemcpyAsync(field + 3 * q0, d_fout_.p, sizeof(double2) * 3 * (size_t)mb,
                            hipMemcpyDeviceToHost, st));
    if (intensity)
      LH_HIP(hipMemcpyAsync(intensity + q0, d_fint_.p, sizeof(double) * (size_t)mb,
                            hipMemcpyDeviceToHost, st));
    LH_HIP(hipStreamSynchronize(st));
  }
}

// field_at_points with every buffer on the device and the scratch memory
// given: the source records and the chunk partials live in `scratch`
// (`bytes` long, 256-byte aligned); nothing is allocated. mom: the moments
// (3n complex, device) or null for slot `slot`. The point batches are sized
// so that their partials fit the scratch; the caller has checked that the
// source records do.
void Engine::field_at_points_scratch(int s, int slot, const double2* mom, const lh_incident& inc,
                                     long long m, const double* pts, double2* field,
                                     double* intensity, char* scratch, size_t bytes) {
  const SolveDev& d = sd_[s];
  hipStream_t st = stream_;
  const size_t src_bytes = (sizeof(FieldSrc) * (size_t)d.n + 255) / 256 * 256;
  FieldSrc* src = reinterpret_cast<FieldSrc*>(scratch);
  double2* part = reinterpret_cast<double2*>(scratch + src_bytes);
  const size_t part_cap = (bytes - src_bytes) / (3 * sizeof(double2));   // points x chunks
  launch_field_pack(st, d.n, d_px_.as<double>() + d.pos_off, d_py_.as<double>() + d.pos_off,
                    d_pz_.as<double>() + d.pos_off,
                    mom ? mom : d_vec_.as<double2>() + d.vec_off + (long long)slot * 3 * d.n,
                    src);
  double pol[6];
  for (int c = 0; c < 3; ++c) {
    pol[2 * c] = inc.polarisation[c].re;
    pol[2 * c + 1] = inc.polarisation[c].im;
  }
  for (long long q0 = 0; q0 < m;) {
    // the batch of field_at_points, halved until its partials fit
    int mb = (int)std::min(kFieldBatch, m - q0);
    int chunk = field_chunk(mb, d.n), nchunk = (d.n + chunk - 1) / chunk;
    while ((size_t)mb * nchunk > part_cap && mb > 256) {
      mb = std::max(256, mb / 2);
      chunk = field_chunk(mb, d.n);
      nchunk = (d.n + chunk - 1) / chunk;
    }
    if ((size_t)mb * nchunk > part_cap)
      throw Error(LH_INVALID_ARGUMENT, "field map scratch too small for this dipole count");
    launch_field_part(st, mb, pts + 3 * q0, src, d.n, chunk, d.k, part);
    launch_field_reduce(st, mb, nchunk, pts + 3 * q0, part, d.k, inc.direction, inc.amplitude,
                        pol, field ? field + 3 * q0 : nullptr,
                        intensity ? intensity + q0 : nullptr);
    LH_HIP(hipGetLastError());
    q0 += mb;
  }
  LH_HIP(hipStreamSynchronize(st));
}

void Engine::far_field(int s, int slot, long long m, const double* rhat, double* intensity) {
  const SolveDev& d = sd_[s];
  hipStream_t st = stream_;
  field_sources(s, slot);
  for (long long q0 = 0; q0 < m; q0 += kFieldBatch) {
    const int mb = (int)std::min(kFieldBatch, m - q0);
    const int chunk = field_chunk(mb, d.n);
    const int nchunk = (d.n + chunk - 1) / chunk;
    ensure(d_fpts_, sizeof(double) * 3 * (size_t)mb);
    ensure(d_fpart_, sizeof(double2) * 3 * (size_t)mb * nchunk);
    ensure(d_fint_, sizeof(double) * (size_t)mb);
    LH_HIP(hipMemcpyAsync(d_fpts_.p, rhat + 3 * q0, sizeof(double) * 3 * (size_t)mb,
                          hipMemcpyHostToDevice, st));
    launch_ff_part(st, mb, d_fpts_.as<double>(), d_fsrc_.as<FieldSrc>(), d.n, chunk, d.k,
                   d_fpart_.as<double2>());
    launch_ff_reduce(st, mb, nchunk, d_fpts_.as<double>(), d_fpart_.as<double2>(),
                     d_fint_.as<double>());
    LH_HIP(hipGetLastError());
    LH_HIP(hipMemcpyAsync(intensity + q0, d_fint_.p, sizeof(double) * (size_t)mb,
                          hipMemcpyDeviceToHost, st));
    LH_HIP(hipStreamSynchronize(st));
  }
}

std::vector<SolveResult> Engine::results() {
  std::vector<GState> h(sd_.size());
  LH_HIP(hipMemcpyAsync(h.data(), d_state_.p, sizeof(GState) * h.size(), hipMemcpyDeviceToHost,
                        stream_));
  LH_HIP(hipStreamSynchronize(stream_));
  std::vector<SolveResult> r(sd_.size());
  for (size_t s = 0; s < h.size(); ++s) {
    const GState& g = h[s];
    r[s].status = (g.status == 2) ? LH_CONVERGENCE_FAILURE : LH_OK;
    r[s].iterations = g.total_iters;
    r[s].cycles = g.cycles;
    r[s].matvecs = g.matvecs;
    r[s].matvecs32 = g.matvecs32;
    r[s].residual = (g.rhs_norm > 0.0) ? g.res_norm / g.rhs_norm : 0.0;
    r[s].pair_evals = pairs_per_matvec((int)s) * g.matvecs;
    r[s].group = group_of_[s];
  }
  return r;
}

// cda/mod.rs:187-235: C_ext = k/E0^2 sum Im(conj(E) p);
// C_abs = k/E0^2 sum [Im(p . conj(ainv) . conj(p)) - 2/3 k^3 |p|^2];
// C_sca = max(C_ext - C_abs, 0).
std::vector<double> Engine::cross_sections(const std::vector<double>& amplitude) {
  const int nv = (int)vit_.size();
  launch_cross(stream_, nv, d_vit_.as<VecItem>(), d_solves_.as<SolveDev>(), d_vec_.as<double2>(),
               d_ainv_.as<double2>(), d_cross_.as<double>());
  LH_HIP(hipGetLastError());
  std::vector<double> part(3 * (size_t)nv);
  LH_HIP(hipMemcpyAsync(part.data(), d_cross_.p, sizeof(double) * part.size(),
                        hipMemcpyDeviceToHost, stream_));
  LH_HIP(hipStreamSynchronize(stream_));
  std::vector<double> out(3 * sd_.size());
  for (size_t s = 0; s < sd_.size(); ++s) {
    const SolveDev& d = sd_[s];
    double ce = 0.0, ca = 0.0, ps = 0.0;
    for (int b = 0; b < d.nvi; ++b) {
      ce += part[3 * (size_t)(d.vi_off + b) + 0];
      ca += part[3 * (size_t)(d.vi_off + b) + 1];
      ps += part[3 * (size_t)(d.vi_off + b) + 2];
    }
    const double k = d.k;
    const double e0sq = amplitude[s] * amplitude[s];
    const double cext = ce * (k / e0sq);
    const double cabs = (ca - (2.0 / 3.0) * (k * k * k) * ps) * (k / e0sq);
    out[3 * s + 0] = cext;
    out[3 * s + 1] = cabs;
    out[3 * s + 2] = std::max(cext - cabs, 0.0);
  }
  return out;
}

}  // namespace lh
