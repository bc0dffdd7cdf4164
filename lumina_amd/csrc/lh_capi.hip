// C ABI (include/lumina_hip.h): backend, stored-matrix session, on-the-fly
// pair-sum sessions and batched solves.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "lh_engine.h"

namespace {

thread_local std::string g_last_error;

int fail(int st, const std::string& m) {
  g_last_error = m;
  return st;
}

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    return f();
  } catch (const lh::Error& e) {
    return fail(e.status, e.what());
  } catch (const std::bad_alloc&) {
    return fail(LH_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(LH_DEVICE_ERROR, e.what());
  }
}

// assembly.rs:335-363 invert_3x3: cofactor inverse, |det| > 1e-30 required
// (the reference panics on a singular block; here it is an argument error).
void invert3(const lh_c64* m, lh_c64* out) {
  auto mul = [](lh_c64 a, lh_c64 b) {
    return lh_c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  };
  auto sub = [](lh_c64 a, lh_c64 b) { return lh_c64{a.re - b.re, a.im - b.im}; };
  auto add = [](lh_c64 a, lh_c64 b) { return lh_c64{a.re + b.re, a.im + b.im}; };
  const lh_c64 a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
               k = m[8];
  const lh_c64 det = add(sub(mul(a, sub(mul(e, k), mul(f, h))), mul(b, sub(mul(d, k), mul(f, g)))),
                         mul(c, sub(mul(d, h), mul(e, g))));
  const double dn = std::hypot(det.re, det.im);
  if (!(dn > 1e-30))
    throw lh::Error(LH_INVALID_ARGUMENT,
                    "Singular polarisability tensor (det = " + std::to_string(dn) + ")");
  const double nsq = det.re * det.re + det.im * det.im;
  const lh_c64 inv{det.re / nsq, -det.im / nsq};
  out[0] = mul(inv, sub(mul(e, k), mul(f, h)));
  out[1] = mul(inv, sub(mul(c, h), mul(b, k)));
  out[2] = mul(inv, sub(mul(b, f), mul(c, e)));
  out[3] = mul(inv, sub(mul(f, g), mul(d, k)));
  out[4] = mul(inv, sub(mul(a, k), mul(c, g)));
  out[5] = mul(inv, sub(mul(c, d), mul(a, f)));
  out[6] = mul(inv, sub(mul(d, h), mul(e, g)));
  out[7] = mul(inv, sub(mul(b, g), mul(a, h)));
  out[8] = mul(inv, sub(mul(a, e), mul(b, d)));
}

std::vector<lh_c64> inverse_blocks(const lh_c64* alpha, int uniform, size_t n) {
  const size_t nb = uniform ? 1 : n;
  std::vector<lh_c64> out(9 * nb);
  for (size_t i = 0; i < nb; ++i) invert3(alpha + 9 * i, out.data() + 9 * i);
  return out;
}

void check_spec(const lh_pairsum_spec* s) {
  if (!s) throw lh::Error(LH_INVALID_ARGUMENT, "null spec");
  if (s->n_dipoles == 0) throw lh::Error(LH_INVALID_ARGUMENT, "No dipoles provided");
  if (s->n_dipoles > (size_t)(1 << 26)) throw lh::Error(LH_INVALID_ARGUMENT, "too many dipoles");
  if (!s->positions || !s->polarisability)
    throw lh::Error(LH_INVALID_ARGUMENT, "null positions or polarisability");
  if (!(s->k > 0.0) || !std::isfinite(s->k))
    throw lh::Error(LH_INVALID_ARGUMENT, "k must be positive and finite");
  if (s->use_fcd && !(s->cell_size > 0.0))
    throw lh::Error(LH_INVALID_ARGUMENT, "cell_size must be positive with FCD");
  if (s->has_lattice) {
    const double cz = s->a1[0] * s->a2[1] - s->a1[1] * s->a2[0];
    for (int c = 0; c < 3; ++c)
      if (!std::isfinite(s->a1[c]) || !std::isfinite(s->a2[c]) || !std::isfinite(s->k_bloch[c]))
        throw lh::Error(LH_INVALID_ARGUMENT, "non-finite lattice or Bloch vector");
    // a 2D lattice in the xy plane: its spectral sum uses the in-plane
    // reciprocal vectors (lattice.rs:113-147)
    if (s->a1[2] != 0.0 || s->a2[2] != 0.0 || !(std::fabs(cz) > 0.0))
      throw lh::Error(LH_INVALID_ARGUMENT, "lattice vectors must span the xy plane");
    if (s->truncation_real < 0 || s->truncation_recip < 0 || s->truncation_real > 50 ||
        s->truncation_recip > 50)
      throw lh::Error(LH_INVALID_ARGUMENT, "lattice truncations must lie in [0, 50]");
    if (!std::isfinite(s->ewald_eta)) throw lh::Error(LH_INVALID_ARGUMENT, "non-finite eta");
  }
  if (s->precision != LH_PREC_FP64 && s->precision != LH_PREC_MIXED)
    throw lh::Error(LH_INVALID_ARGUMENT, "unknown precision");
}

lh::Geometry make_geometry(const lh_pairsum_spec* s) {
  lh::Geometry g;
  g.n = (int)s->n_dipoles;
  g.use_fcd = s->use_fcd ? 1 : 0;
  g.cell = s->cell_size;
  // Far-field exclusion radius: 5% beyond the FCD radius 2 cell_size, tested
  // on FP32-rounded positions; every FCD pair is inside it as long as FP32
  // rounding of the positions moves a displacement by far less than 5% of
  // 2 cell_size, which the coordinate bound below guarantees.
  const double R = 2.0 * s->cell_size * 1.05;
  // with a lattice every pair is an Ewald pair: no FCD, no near list
  // (assembly.rs:244-251)
  g.rskip2f = (g.use_fcd && !s->has_lattice) ? (float)(R * R) : 0.0f;
  g.x.resize(g.n);
  g.y.resize(g.n);
  g.z.resize(g.n);
  double amax = 0.0;
  for (int i = 0; i < g.n; ++i) {
    g.x[i] = s->positions[3 * i];
    g.y[i] = s->positions[3 * i + 1];
    g.z[i] = s->positions[3 * i + 2];
    if (!std::isfinite(g.x[i]) || !std::isfinite(g.y[i]) || !std::isfinite(g.z[i]))
      throw lh::Error(LH_INVALID_ARGUMENT, "non-finite dipole position");
    amax = std::max({amax, std::fabs(g.x[i]), std::fabs(g.y[i]), std::fabs(g.z[i])});
  }
  if (s->classify_positions) {
    g.cls.assign(s->classify_positions, s->classify_positions + 3 * (size_t)g.n);
    for (double v : g.cls)
      if (!std::isfinite(v)) throw lh::Error(LH_INVALID_ARGUMENT, "non-finite classify position");
  }
  if (g.use_fcd && !s->has_lattice && amax > 1e4 * s->cell_size)
    throw lh::Error(LH_INVALID_ARGUMENT,
                    "positions span more than 1e4 cell sizes from the origin; centre the object");
  if (s->has_lattice) {
    lh::detect_grid(g);
  } else {
    lh::build_near_list(g);
    lh::detect_grid(g);   // a full grid qualifies for the FFT matvec
  }
  return g;
}

lh::Problem make_problem(const lh_pairsum_spec* s, int geom) {
  lh::Problem p;
  p.geom = geom;
  p.k = s->k;
  p.ainv_uniform = s->polarisability_uniform ? 1 : 0;
  p.ainv = inverse_blocks(s->polarisability, p.ainv_uniform, s->n_dipoles);
  p.has_sub = s->has_substrate ? 1 : 0;
  p.z_interface = s->z_interface;
  p.delta = s->has_substrate ? s->delta_eps : lh_c64{0.0, 0.0};
  if (s->has_lattice) {
    p.has_lat = 1;
    for (int c = 0; c < 3; ++c) {
      p.a1[c] = s->a1[c];
      p.a2[c] = s->a2[c];
      p.kb[c] = s->k_bloch[c];
    }
    p.eta = s->ewald_eta;
    p.n_real = s->truncation_real;
    p.n_recip = s->truncation_recip;
  }
  return p;
}

// Launch options. LUMINA_FP32_SINCOS=sw selects the software sin/cos of the
// FP32 far field (default: the hardware instructions where their range
// allows); LUMINA_GRAPH=0 issues every cycle's kernels from the host instead
// of replaying a captured graph; LUMINA_FAR_BLOCK=64|128|256 fixes the
// far-field targets per work item, 512 the FP32 one-tensor-per-pair kernel
// (FP64 then uses 256), sym that kernel in both precisions (default: timed
// on the batch);
// LUMINA_EWALD=pairs evaluates the Ewald tensor per pair even where the
// dipoles sit on a grid (default: grid tables); LUMINA_FFT=0 keeps the pair
// sum for full-grid solves, LUMINA_FFT=force takes the FFT matvec for any
// full grid of two or more dipoles (default: from 4,096 dipoles);
// LUMINA_GROUPS=1|2 fixes the number of solve groups whose cycles run on
// separate streams (default: two for batches of 1e8 pair evaluations per
// matvec or more).
lh::Tiling tiling_from_env() {
  lh::Tiling t;
  const char* sc = std::getenv("LUMINA_FP32_SINCOS");
  if (sc && std::string(sc) == "sw") t.hw_sincos = 0;
  const char* gr = std::getenv("LUMINA_GRAPH");
  if (gr && std::string(gr) == "0") t.graph = 0;
  const char* ew = std::getenv("LUMINA_EWALD");
  if (ew && std::string(ew) == "pairs") t.ewald_pairs = 1;
  const char* ff = std::getenv("LUMINA_FFT");
  if (ff && std::string(ff) == "0") t.fft = 0;
  if (ff && std::string(ff) == "force") t.fft = 2;
  const char* og = std::getenv("LUMINA_ORTHO");
  if (og && std::string(og) == "mgs") t.mgs = 1;
  const char* pp = std::getenv("LUMINA_PART64");
  if (pp && std::string(pp) == "1") t.part32 = 0;
  const char* rt = std::getenv("LUMINA_RUN_TILES");
  if (rt && std::atoi(rt) > 0) t.run_tiles = std::atoi(rt);
  const char* gc = std::getenv("LUMINA_GRAPH_CYCLES");
  if (gc && std::string(gc) == "all") t.graph_cycles = -1;
  else if (gc && std::atoi(gc) > 0) t.graph_cycles = std::atoi(gc);
  const char* gp = std::getenv("LUMINA_GROUPS");
  if (gp && (std::atoi(gp) == 1 || std::atoi(gp) == 2)) t.groups = std::atoi(gp);
  const char* cg = std::getenv("LUMINA_CGS2");
  if (cg && std::string(cg) == "1") t.cgs2_fp32 = 1;
  const char* bl = std::getenv("LUMINA_FAR_BLOCK");
  if (bl && std::string(bl) == "sym") t.block[0] = t.block[1] = lh::kSymBlock;
  if (bl) {
    const int b = std::atoi(bl);
    if (b == 64 || b == 128 || b == 256) t.block[0] = t.block[1] = b;
    if (b == lh::kSymBlock) {
      t.block[0] = 256;
      t.block[1] = lh::kSymBlock;
    }
    if (b == lh::kSymRotBlock) {
      t.block[0] = lh::kSymRotBlock;
      t.block[1] = 256;
    }
  }
  return t;
}

void fill_info(const lh::SolveResult& r, lh_solve_info* info) {
  if (!info) return;
  info->iterations = (uint64_t)r.iterations;
  info->cycles = (uint64_t)r.cycles;
  info->matvecs = (uint64_t)r.matvecs;
  info->matvecs_fp32 = (uint64_t)r.matvecs32;
  info->residual = r.residual;
  info->pair_evals = r.pair_evals;
  info->status = r.status;
  info->solve_group = r.group;
}

}  // namespace

// Working memory of lh_pairsum_field_at_points_noalloc: a fixed array in
// every GPU the library runs on, held until the process ends: the source
// records of up to kNoAllocMaxDipoles dipoles and the chunk partials of the
// point batches. Its name is a host-side stand-in, not a device address;
// host code uses the address hipGetSymbolAddress returns on each device
// (lh_backend::scratch), taken once per device when its backend is created.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr size_t kNoAllocMaxDipoles = 262144;   // 20 MiB of 80-byte source records
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
// Calls on one device share its scratch array: they are serialised per
// device (ordinals beyond the table share a lock).
constexpr int kLockedDevices = 64;
std::mutex g_scratch_lock[kLockedDevices];
}  // namespace

struct lh_backend {
  int device = 0;
  hipDeviceProp_t prop;
  void* scratch = nullptr;   // this device's copy of `scratch`
};

struct lh_session {
  lh_backend* b = nullptr;
  size_t n = 0;
  hipStream_t stream = nullptr;
  lh::DevMem A, x, y;
  bool uploaded = false;
};

struct lh_pairsum {
  lh_backend* b = nullptr;
  hipStream_t stream = nullptr;
  std::unique_ptr<lh::Engine> eng;
  size_t n = 0;
  int precision = 0;
  lh_c64 delta{0.0, 0.0};
  int has_sub = 0;
  int has_lat = 0;
};

struct lh_batch {
  lh_backend* b = nullptr;
  hipStream_t stream = nullptr;
  std::unique_ptr<lh::Engine> eng;
  std::vector<double> amp;
  bool ran = false;
};

namespace lh {
void launch_dense_matvec(hipStream_t st, int n, const double2* A, const double2* x, double2* y);
void launch_assemble(hipStream_t st, int n, const double* pos, double k, const double2* diag,
                     double2* A);
}

extern "C" {

const char* lh_last_error(void) { return g_last_error.c_str(); }

int lh_backend_create(int device_ordinal, lh_backend** out) {
  return guarded([&]() {
    if (!out) return fail(LH_INVALID_ARGUMENT, "null out");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
      (void)hipGetLastError();
      return fail(LH_UNAVAILABLE, "Backend not available: no HIP device");
    }
    if (device_ordinal < 0 || device_ordinal >= count)
      return fail(LH_INVALID_ARGUMENT, "device ordinal out of range");
    auto b = std::make_unique<lh_backend>();
    b->device = device_ordinal;
    LH_HIP(hipSetDevice(device_ordinal));
    LH_HIP(hipGetDeviceProperties(&b->prop, device_ordinal));
    // the device address of this device's scratch array (with the device current)
    LH_HIP(hipGetSymbolAddress(&b->scratch, HIP_SYMBOL(scratch)));
    *out = b.release();
    return (int)LH_OK;
  });
}

void lh_backend_destroy(lh_backend* backend) { delete backend; }

int lh_backend_device_info(const lh_backend* b, lh_device_info* out) {
  return guarded([&]() {
    if (!b || !out) return fail(LH_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, b->prop.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, b->prop.gcnArchName, sizeof(out->arch) - 1);
    if (out->name[0] == '\0') {
      // some drivers report no marketing name: name the device by its ISA
      const std::string isa(b->prop.gcnArchName);
      const std::string nm = "AMD GPU " + isa.substr(0, isa.find(':'));
      std::strncpy(out->name, nm.c_str(), sizeof(out->name) - 1);
    }
    out->backend_type = 1;
    out->memory_bytes = (uint64_t)b->prop.totalGlobalMem;
    out->compute_units = (uint32_t)b->prop.multiProcessorCount;
    return (int)LH_OK;
  });
}

// ---- stored-matrix session ------------------------------------------------

int lh_session_create(lh_backend* b, size_t n, lh_session** out) {
  return guarded([&]() {
    if (!b || !out || n == 0) return fail(LH_INVALID_ARGUMENT, "bad argument");
    *out = nullptr;
    LH_HIP(hipSetDevice(b->device));
    auto s = std::make_unique<lh_session>();
    s->b = b;
    s->n = n;
    const size_t bytes = n * n * sizeof(lh_c64);
    size_t fr = 0, tot = 0;
    LH_HIP(hipMemGetInfo(&fr, &tot));
    if (bytes > fr)
      return fail(LH_OUT_OF_MEMORY, "OutOfMemory: requested " + std::to_string(bytes) +
                                        " bytes, available " + std::to_string(fr));
    LH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->A.alloc(bytes);
    s->x.alloc(n * sizeof(lh_c64));
    s->y.alloc(n * sizeof(lh_c64));
    *out = s.release();
    return (int)LH_OK;
  });
}

int lh_session_upload_matrix(lh_session* s, const lh_c64* m, size_t n) {
  return guarded([&]() {
    if (!s || !m) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (n != s->n)
      return fail(LH_DEVICE_ERROR, "Matrix dimension mismatch: session " + std::to_string(s->n) +
                                       ", matrix " + std::to_string(n));
    LH_HIP(hipSetDevice(s->b->device));
    LH_HIP(hipMemcpyAsync(s->A.p, m, n * n * sizeof(lh_c64), hipMemcpyHostToDevice, s->stream));
    LH_HIP(hipStreamSynchronize(s->stream));
    s->uploaded = true;
    return (int)LH_OK;
  });
}

int lh_session_matvec(lh_session* s, const lh_c64* x, lh_c64* y, size_t n) {
  return guarded([&]() {
    if (!s || !x || !y) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (n != s->n)
      return fail(LH_DEVICE_ERROR, "Vector dimension mismatch: session " + std::to_string(s->n) +
                                       ", vector " + std::to_string(n));
    if (!s->uploaded) return fail(LH_DEVICE_ERROR, "matrix not uploaded");
    LH_HIP(hipSetDevice(s->b->device));
    LH_HIP(hipMemcpyAsync(s->x.p, x, n * sizeof(lh_c64), hipMemcpyHostToDevice, s->stream));
    lh::launch_dense_matvec(s->stream, (int)n, s->A.as<double2>(), s->x.as<double2>(),
                            s->y.as<double2>());
    LH_HIP(hipGetLastError());
    LH_HIP(hipMemcpyAsync(y, s->y.p, n * sizeof(lh_c64), hipMemcpyDeviceToHost, s->stream));
    LH_HIP(hipStreamSynchronize(s->stream));
    return (int)LH_OK;
  });
}

int lh_session_assemble(lh_backend* b, const double* positions, size_t n_dipoles, double k,
                        const lh_c64* diagonal_blocks, lh_session** out) {
  return guarded([&]() {
    if (!b || !out) return fail(LH_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (n_dipoles == 0) return fail(LH_INVALID_ARGUMENT, "No dipoles provided");
    if (!positions || !diagonal_blocks) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (!(std::isfinite(k) && k > 0.0))
      return fail(LH_INVALID_ARGUMENT, "k must be positive and finite");
    // one grid row per block row i (gridDim.y); the dense matrix of 65,535
    // dipoles is 619 GB, beyond any device memory
    if (n_dipoles > 65535)
      return fail(LH_INVALID_ARGUMENT, "dense assembly is limited to 65535 dipoles");
    for (size_t q = 0; q < 3 * n_dipoles; ++q)
      if (!std::isfinite(positions[q])) return fail(LH_INVALID_ARGUMENT, "non-finite dipole position");
    for (size_t q = 0; q < 9 * n_dipoles; ++q)
      if (!std::isfinite(diagonal_blocks[q].re) || !std::isfinite(diagonal_blocks[q].im))
        return fail(LH_INVALID_ARGUMENT, "non-finite diagonal block");
    lh_session* raw = nullptr;
    const int st = lh_session_create(b, 3 * n_dipoles, &raw);
    if (st != LH_OK) return st;
    std::unique_ptr<lh_session, void (*)(lh_session*)> s(raw, lh_session_destroy);
    lh::DevMem pos, diag;
    pos.alloc(3 * n_dipoles * sizeof(double));
    diag.alloc(9 * n_dipoles * sizeof(lh_c64));
    LH_HIP(hipMemcpyAsync(pos.p, positions, pos.bytes, hipMemcpyHostToDevice, s->stream));
    LH_HIP(hipMemcpyAsync(diag.p, diagonal_blocks, diag.bytes, hipMemcpyHostToDevice, s->stream));
    lh::launch_assemble(s->stream, (int)n_dipoles, pos.as<double>(), k, diag.as<double2>(),
                        s->A.as<double2>());
    LH_HIP(hipGetLastError());
    LH_HIP(hipStreamSynchronize(s->stream));
    s->uploaded = true;
    *out = s.release();
    return (int)LH_OK;
  });
}

void lh_session_destroy(lh_session* s) {
  if (!s) return;
  (void)hipSetDevice(s->b->device);
  // queued work first, then the memory it uses, then the stream
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  s->A.release();
  s->x.release();
  s->y.release();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// ---- pair-sum session -----------------------------------------------------

int lh_pairsum_create(lh_backend* b, const lh_pairsum_spec* spec, lh_pairsum** out) {
  return guarded([&]() {
    if (!b || !out) return fail(LH_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    check_spec(spec);
    LH_HIP(hipSetDevice(b->device));
    auto s = std::make_unique<lh_pairsum>();
    s->b = b;
    s->n = spec->n_dipoles;
    s->precision = spec->precision;
    s->has_sub = spec->has_substrate ? 1 : 0;
    s->delta = spec->delta_eps;
    s->has_lat = spec->has_lattice ? 1 : 0;
    LH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    std::vector<lh::Geometry> geoms;
    geoms.push_back(make_geometry(spec));
    std::vector<lh::Problem> probs;
    probs.push_back(make_problem(spec, 0));
    s->eng = std::make_unique<lh::Engine>(s->stream);
    s->eng->build(geoms, probs, tiling_from_env());
    *out = s.release();
    return (int)LH_OK;
  });
}

int lh_pairsum_set_wavelength(lh_pairsum* s, double k, const lh_c64* alpha, int uniform,
                              lh_c64 delta_eps) {
  return guarded([&]() {
    if (!s || !alpha) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (!(k > 0.0) || !std::isfinite(k)) return fail(LH_INVALID_ARGUMENT, "k must be positive");
    LH_HIP(hipSetDevice(s->b->device));
    lh::Problem p;
    p.k = k;
    p.ainv_uniform = uniform ? 1 : 0;
    p.ainv = inverse_blocks(alpha, p.ainv_uniform, s->n);
    p.delta = s->has_sub ? delta_eps : lh_c64{0.0, 0.0};
    s->eng->set_problem(0, p);
    return (int)LH_OK;
  });
}

int lh_pairsum_matvec(lh_pairsum* s, const lh_c64* x, lh_c64* y, int precision) {
  return guarded([&]() {
    if (!s || !x || !y) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (precision != LH_PREC_FP64 && precision != LH_PREC_MIXED)
      return fail(LH_INVALID_ARGUMENT, "unknown precision");
    LH_HIP(hipSetDevice(s->b->device));
    s->eng->upload_slot(0, 0, x);
    s->eng->matvec_all(0, precision == LH_PREC_MIXED);
    s->eng->download_slot(0, LH_SLOT_W, y);
    return (int)LH_OK;
  });
}

int lh_pairsum_solve(lh_pairsum* s, const lh_c64* rhs, double tol, size_t max_iter, lh_c64* x,
                     lh_solve_info* info) {
  return guarded([&]() {
    if (!s || !rhs || !x) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (!(tol >= 0.0)) return fail(LH_INVALID_ARGUMENT, "tolerance must be >= 0");
    LH_HIP(hipSetDevice(s->b->device));
    s->eng->upload_slot(0, LH_SLOT_RHS, rhs);
    s->eng->solve(tol, (long long)max_iter, s->precision == LH_PREC_MIXED, true);
    s->eng->download_slot(0, LH_SLOT_X, x);
    const auto r = s->eng->results();
    fill_info(r[0], info);
    if (r[0].status != LH_OK)
      return fail(LH_CONVERGENCE_FAILURE,
                  "GMRES did not converge within " + std::to_string(max_iter) +
                      " iterations (residual " + std::to_string(r[0].residual) + ")");
    return (int)LH_OK;
  });
}

int lh_pairsum_incident(lh_pairsum* s, const lh_incident* inc, lh_c64* e_inc) {
  return guarded([&]() {
    if (!s || !inc || !e_inc) return fail(LH_INVALID_ARGUMENT, "null argument");
    LH_HIP(hipSetDevice(s->b->device));
    s->eng->fill_incident(0, *inc);
    s->eng->download_slot(0, LH_SLOT_E, e_inc);
    return (int)LH_OK;
  });
}

int lh_pairsum_cross_sections(lh_pairsum* s, const lh_c64* p, const lh_c64* e_inc,
                              double amplitude, double out[3]) {
  return guarded([&]() {
    if (!s || !p || !e_inc || !out) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (!(amplitude != 0.0)) return fail(LH_INVALID_ARGUMENT, "amplitude must be non-zero");
    LH_HIP(hipSetDevice(s->b->device));
    s->eng->upload_slot(0, LH_SLOT_X, p);
    s->eng->upload_slot(0, LH_SLOT_E, e_inc);
    const auto cs = s->eng->cross_sections({amplitude});
    out[0] = cs[0];
    out[1] = cs[1];
    out[2] = cs[2];
    return (int)LH_OK;
  });
}

int lh_pairsum_compute_cross_sections(lh_pairsum* s, const lh_incident* inc, double tol,
                                      size_t max_iter, double out[3], lh_c64* p_out,
                                      lh_solve_info* info) {
  return guarded([&]() {
    if (!s || !inc || !out) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (!(tol >= 0.0)) return fail(LH_INVALID_ARGUMENT, "tolerance must be >= 0");
    if (!(inc->amplitude != 0.0)) return fail(LH_INVALID_ARGUMENT, "amplitude must be non-zero");
    LH_HIP(hipSetDevice(s->b->device));
    s->eng->fill_incident(0, *inc);
    s->eng->solve(tol, (long long)max_iter, s->precision == LH_PREC_MIXED, true);
    const auto r = s->eng->results();
    fill_info(r[0], info);
    const auto cs = s->eng->cross_sections({inc->amplitude});
    out[0] = cs[0];
    out[1] = cs[1];
    out[2] = cs[2];
    if (p_out) s->eng->download_slot(0, LH_SLOT_X, p_out);
    if (r[0].status != LH_OK)
      return fail(LH_CONVERGENCE_FAILURE,
                  "GMRES did not converge within " + std::to_string(max_iter) +
                      " iterations (residual " + std::to_string(r[0].residual) + ")");
    return (int)LH_OK;
  });
}

namespace {
// Observation points / directions per C-ABI call: kernel indexing is 32-bit.
constexpr size_t kMaxFieldPoints = (size_t)1 << 28;

int moments_slot(lh_pairsum* s, const lh_c64* moments) {
  if (!moments) return LH_SLOT_X;
  s->eng->upload_slot(0, LH_SLOT_W, moments);
  return LH_SLOT_W;
}
}  // namespace

int lh_pairsum_field_at_points(lh_pairsum* s, const lh_c64* moments, const lh_incident* inc,
                               size_t m, const double* points, lh_c64* field,
                               double* intensity) {
  return guarded([&]() {
    if (!s || !inc || (m > 0 && !points)) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (m > kMaxFieldPoints) return fail(LH_INVALID_ARGUMENT, "too many observation points");
    if (m == 0 || (!field && !intensity)) return (int)LH_OK;
    LH_HIP(hipSetDevice(s->b->device));
    const int slot = moments_slot(s, moments);
    s->eng->field_at_points(0, slot, *inc, (long long)m, points, field, intensity);
    return (int)LH_OK;
  });
}

int lh_pairsum_field_at_points_noalloc(lh_pairsum* s, const lh_c64* moments,
                                       const lh_incident* inc, size_t m, const double* points,
                                       lh_c64* field, double* intensity) {
  return guarded([&]() {
    if (!s || !inc || (m > 0 && !points)) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (m > kMaxFieldPoints) return fail(LH_INVALID_ARGUMENT, "too many observation points");
    if (s->n > kNoAllocMaxDipoles)
      return fail(LH_INVALID_ARGUMENT,
                  "lh_pairsum_field_at_points_noalloc: " + std::to_string(s->n) +
                      " dipoles exceed the fixed scratch's " +
                      std::to_string(kNoAllocMaxDipoles));
    if (m == 0 || (!field && !intensity)) return (int)LH_OK;
    if (!s->b->scratch) return fail(LH_UNAVAILABLE, "no scratch address on this device");
    LH_HIP(hipSetDevice(s->b->device));
    std::lock_guard<std::mutex> lock(g_scratch_lock[s->b->device % kLockedDevices]);
    s->eng->field_at_points_scratch(0, LH_SLOT_X, reinterpret_cast<const double2*>(moments), *inc,
                                    (long long)m, points, reinterpret_cast<double2*>(field),
                                    intensity, static_cast<char*>(s->b->scratch),
                                    kNoAllocScratchBytes);
    return (int)LH_OK;
  });
}

int lh_pairsum_far_field(lh_pairsum* s, const lh_c64* moments, size_t n_theta, size_t n_phi,
                         double* theta, double* phi, double* intensity) {
  return guarded([&]() {
    if (!s || !intensity) return fail(LH_INVALID_ARGUMENT, "null argument");
    if (n_phi > 0 && n_theta > kMaxFieldPoints / n_phi)
      return fail(LH_INVALID_ARGUMENT, "too many directions");
    const size_t m = n_theta * n_phi;
    if (m == 0) return (int)LH_OK;
    std::vector<double> rhat(3 * m);
    size_t t = 0;
    for (size_t it = 0; it < n_theta; ++it) {
      const double th = M_PI * (double)it / (double)std::max<size_t>(n_theta - 1, 1);
      const double st = std::sin(th), ct = std::cos(th);
      for (size_t ip = 0; ip < n_phi; ++ip, ++t) {
        const double ph = 2.0 * M_PI * (double)ip / (double)n_phi;
        rhat[3 * t] = st * std::cos(ph);
        rhat[3 * t + 1] = st * std::sin(ph);
        rhat[3 * t + 2] = ct;
        if (theta) theta[t] = th;
        if (phi) phi[t] = ph;
      }
    }
    LH_HIP(hipSetDevice(s->b->device));
    const int slot = moments_slot(s, moments);
    s->eng->far_field(0, slot, (long long)m, rhat.data(), intensity);
    return (int)LH_OK;
  });
}

int lh_pairsum_adjoint_grad(lh_pairsum* s, const lh_incident* inc, double tol, size_t max_iter,
                            const lh_c64* dainv_deps, int dainv_uniform, double cross_sections[3],
                            double* grad_eps, double* grad_pos, lh_solve_info* fwd_info,
                            lh_solve_info* adj_info) {
  return guarded([&]() {
    if (!s || !inc || !dainv_deps || !cross_sections || !grad_eps)
      return fail(LH_INVALID_ARGUMENT, "null argument");
    if (!(tol >= 0.0)) return fail(LH_INVALID_ARGUMENT, "tolerance must be >= 0");
    if (!(inc->amplitude != 0.0)) return fail(LH_INVALID_ARGUMENT, "amplitude must be non-zero");
    // with a lattice, G_per(-r) != G_per(r) for pairs at equal height
    // (sgn 0 = +1 in the spectral sum), so A is not symmetric
    if (s->has_lat)
      return fail(LH_UNAVAILABLE,
                  "the adjoint solve needs a complex-symmetric operator (no lattice)");
    if (grad_pos && (s->has_sub || s->has_lat))
      return fail(LH_UNAVAILABLE, "position gradients are built for free space only");
    LH_HIP(hipSetDevice(s->b->device));
    const bool mixed = s->precision == LH_PREC_MIXED;
    const size_t n3 = 3 * s->n;
    // forward: A p = E
    s->eng->fill_incident(0, *inc);
    s->eng->solve(tol, (long long)max_iter, mixed, true);
    const auto r1 = s->eng->results();
    fill_info(r1[0], fwd_info);
    const auto cs = s->eng->cross_sections({inc->amplitude});
    for (int q = 0; q < 3; ++q) cross_sections[q] = cs[q];
    std::vector<lh_c64> p(n3), e(n3);
    s->eng->download_slot(0, LH_SLOT_X, p.data());
    s->eng->download_slot(0, LH_SLOT_E, e.data());
    // adjoint: A lam = conj(E) (A = A^T)
    std::vector<lh_c64> ec(n3);
    for (size_t q = 0; q < n3; ++q) ec[q] = lh_c64{e[q].re, -e[q].im};
    s->eng->upload_slot(0, LH_SLOT_RHS, ec.data());
    s->eng->solve(tol, (long long)max_iter, mixed, true);
    const auto r2 = s->eng->results();
    fill_info(r2[0], adj_info);
    // gradients with p in Krylov slot 0 (free after the solves)
    s->eng->upload_slot(0, 0, p.data());
    std::vector<lh_c64> dv(dainv_deps, dainv_deps + (dainv_uniform ? 9 : 9 * s->n));
    std::vector<double> out;
    s->eng->gradients(0, inc->direction, inc->amplitude, dv, dainv_uniform ? 1 : 0,
                      grad_pos != nullptr, out);
    for (size_t i = 0; i < s->n; ++i) {
      grad_eps[2 * i] = out[5 * i];
      grad_eps[2 * i + 1] = out[5 * i + 1];
      if (grad_pos)
        for (int c = 0; c < 3; ++c) grad_pos[3 * i + c] = out[5 * i + 2 + c];
    }
    // leave the session as after a forward solve
    s->eng->upload_slot(0, LH_SLOT_X, p.data());
    s->eng->upload_slot(0, LH_SLOT_RHS, e.data());
    if (r1[0].status != LH_OK || r2[0].status != LH_OK)
      return fail(LH_CONVERGENCE_FAILURE,
                  std::string("GMRES did not converge within ") + std::to_string(max_iter) +
                      " iterations (" + (r1[0].status != LH_OK ? "forward" : "adjoint") +
                      " solve)");
    return (int)LH_OK;
  });
}

void lh_pairsum_destroy(lh_pairsum* s) {
  if (!s) return;
  (void)hipSetDevice(s->b->device);
  s->eng.reset();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// ---- batch ----------------------------------------------------------------

int lh_batch_create(lh_backend* b, const lh_batch_problem* problems, size_t count,
                    lh_batch** out) {
  return guarded([&]() {
    if (!b || !problems || count == 0 || !out) return fail(LH_INVALID_ARGUMENT, "bad argument");
    *out = nullptr;
    LH_HIP(hipSetDevice(b->device));
    struct Key {
      const double* p;
      size_t n;
      int fcd;
      double cell;
      int lat;
      const double* cls;
    };
    std::vector<Key> keys;
    std::vector<lh::Geometry> geoms;
    std::vector<lh::Problem> probs;
    auto bt = std::make_unique<lh_batch>();
    bt->b = b;
    for (size_t q = 0; q < count; ++q) {
      lh_pairsum_spec sp = problems[q].spec;
      sp.precision = LH_PREC_FP64;
      check_spec(&sp);
      if (!(problems[q].incident.amplitude != 0.0))
        return fail(LH_INVALID_ARGUMENT, "amplitude must be non-zero");
      int gi = -1;
      for (size_t t = 0; t < keys.size(); ++t)
        if (keys[t].p == sp.positions && keys[t].n == sp.n_dipoles &&
            keys[t].fcd == (sp.use_fcd ? 1 : 0) && keys[t].cell == sp.cell_size &&
            keys[t].lat == (sp.has_lattice ? 1 : 0) && keys[t].cls == sp.classify_positions) {
          gi = (int)t;
          break;
        }
      if (gi < 0) {
        gi = (int)geoms.size();
        keys.push_back(
            Key{sp.positions, sp.n_dipoles, sp.use_fcd ? 1 : 0, sp.cell_size,
                sp.has_lattice ? 1 : 0, sp.classify_positions});
        geoms.push_back(make_geometry(&sp));
      }
      probs.push_back(make_problem(&sp, gi));
      bt->amp.push_back(problems[q].incident.amplitude);
    }
    LH_HIP(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
    bt->eng = std::make_unique<lh::Engine>(bt->stream);
    bt->eng->build(geoms, probs, tiling_from_env());
    for (size_t q = 0; q < count; ++q) bt->eng->fill_incident((int)q, problems[q].incident);
    *out = bt.release();
    return (int)LH_OK;
  });
}

int lh_batch_run(lh_batch* bt, double tol, size_t max_iter, int precision) {
  return guarded([&]() {
    if (!bt) return fail(LH_INVALID_ARGUMENT, "null batch");
    if (precision != LH_PREC_FP64 && precision != LH_PREC_MIXED)
      return fail(LH_INVALID_ARGUMENT, "unknown precision");
    if (!(tol >= 0.0)) return fail(LH_INVALID_ARGUMENT, "tolerance must be >= 0");
    LH_HIP(hipSetDevice(bt->b->device));
    const auto t0 = std::chrono::steady_clock::now();
    bt->eng->solve(tol, (long long)max_iter, precision == LH_PREC_MIXED, true);
    const auto t1 = std::chrono::steady_clock::now();
    bt->eng->last_run_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    bt->ran = true;
    return (int)LH_OK;
  });
}

int lh_batch_results(lh_batch* bt, double* cross_sections, lh_c64* const* moments,
                     lh_solve_info* infos) {
  return guarded([&]() {
    if (!bt || !cross_sections) return fail(LH_INVALID_ARGUMENT, "bad argument");
    if (!bt->ran) return fail(LH_INVALID_ARGUMENT, "lh_batch_run has not been called");
    LH_HIP(hipSetDevice(bt->b->device));
    const auto res = bt->eng->results();
    const auto cs = bt->eng->cross_sections(bt->amp);
    std::memcpy(cross_sections, cs.data(), sizeof(double) * cs.size());
    bool all_ok = true;
    for (size_t q = 0; q < res.size(); ++q) {
      if (infos) fill_info(res[q], &infos[q]);
      if (moments && moments[q]) bt->eng->download_slot((int)q, LH_SLOT_X, moments[q]);
      if (res[q].status != LH_OK) all_ok = false;
    }
    if (!all_ok) return fail(LH_CONVERGENCE_FAILURE, "at least one solve did not converge");
    return (int)LH_OK;
  });
}

void lh_batch_destroy(lh_batch* bt) {
  if (!bt) return;
  (void)hipSetDevice(bt->b->device);
  bt->eng.reset();
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  delete bt;
}

int lh_batch_set_timing(lh_batch* bt, int enable) {
  return guarded([&]() {
    if (!bt) return fail(LH_INVALID_ARGUMENT, "null batch");
    bt->eng->set_timing(enable != 0);
    return (int)LH_OK;
  });
}

int lh_batch_set_graph_cycles(lh_batch* bt, int cycles) {
  return guarded([&]() {
    if (!bt) return fail(LH_INVALID_ARGUMENT, "null batch");
    if (cycles < -1) return fail(LH_INVALID_ARGUMENT, "cycles must be -1, 0 or positive");
    bt->eng->set_graph_cycles(cycles);
    return (int)LH_OK;
  });
}

int lh_batch_kernel_times(lh_batch* bt, lh_kernel_times* out) {
  return guarded([&]() {
    if (!bt || !out) return fail(LH_INVALID_ARGUMENT, "bad argument");
    LH_HIP(hipSetDevice(bt->b->device));
    double t[4];
    bt->eng->far_times(t);
    std::memset(out, 0, sizeof(*out));
    out->far64_ms = t[0];
    out->far32_ms = t[1];
    out->far64_launches = (uint64_t)t[2];
    out->far32_launches = (uint64_t)t[3];
    const auto res = bt->eng->results();
    for (size_t q = 0; q < res.size(); ++q) {
      const double p = bt->eng->pairs_per_matvec((int)q);
      out->far32_pairs += p * res[q].matvecs32;
      out->far64_pairs += p * (res[q].matvecs - res[q].matvecs32);
    }
    out->run_ms = bt->eng->last_run_ms;
    out->far64_block = bt->eng->block(0);
    out->far32_block = bt->eng->block(1);
    out->tune_ms = bt->eng->tune_ms;
    out->host_round_trips = bt->eng->last_launches;
    out->solve_groups = bt->eng->groups();
    out->cycles_issued = (uint64_t)bt->eng->last_cycles_issued;
    return (int)LH_OK;
  });
}

int lh_sweep_solve(lh_backend* b, const lh_sweep_spec* spec, const lh_material* materials,
                   size_t n_materials, const double* wavelengths_nm, size_t n_wl,
                   const lh_incident* inc, double tol, size_t max_iter, double* cross_sections,
                   lh_solve_info* infos) {
  lh_batch* bt = nullptr;
  int rc = guarded([&]() {
    if (!b || !spec || !materials || n_materials == 0 || !wavelengths_nm || n_wl == 0 || !inc ||
        !cross_sections)
      return fail(LH_INVALID_ARGUMENT, "bad argument");
    if (!(tol >= 0.0)) return fail(LH_INVALID_ARGUMENT, "tolerance must be >= 0");
    if (!(inc->amplitude != 0.0)) return fail(LH_INVALID_ARGUMENT, "amplitude must be non-zero");
    if (!(spec->n_env > 0.0) || !std::isfinite(spec->n_env))
      return fail(LH_INVALID_ARGUMENT, "n_env must be positive and finite");
    // the geometry checks of a session, with placeholder optics
    lh_pairsum_spec g = spec->geometry;
    const lh_c64 unit[9] = {{1, 0}, {0, 0}, {0, 0}, {0, 0}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {1, 0}};
    g.polarisability = unit;
    g.polarisability_uniform = 1;
    g.k = 1.0;
    g.delta_eps = lh_c64{0.0, 0.0};
    check_spec(&g);
    const size_t n = g.n_dipoles;
    lh::SweepSetup su;
    su.n_env = spec->n_env;
    // materials: knots and spline second derivatives (spline.rs:32-66)
    for (size_t m = 0; m < n_materials; ++m) {
      const lh_material& mt = materials[m];
      if (mt.n_points < 2 || !mt.wavelength_nm || !mt.epsilon)
        return fail(LH_INVALID_ARGUMENT, "material " + std::to_string(m) + ": need >= 2 points");
      const int np = (int)mt.n_points;
      std::vector<double> re(np), im(np), y2(np);
      for (int q = 0; q < np; ++q) {
        if (!std::isfinite(mt.wavelength_nm[q]) || !std::isfinite(mt.epsilon[q].re) ||
            !std::isfinite(mt.epsilon[q].im) || (q > 0 && !(mt.wavelength_nm[q] > mt.wavelength_nm[q - 1])))
          return fail(LH_INVALID_ARGUMENT,
                      "material " + std::to_string(m) + ": wavelengths must increase strictly");
        re[q] = mt.epsilon[q].re;
        im[q] = mt.epsilon[q].im;
      }
      su.mat_off.push_back((int)su.xs.size());
      su.mat_n.push_back(np);
      su.xs.insert(su.xs.end(), mt.wavelength_nm, mt.wavelength_nm + np);
      su.yr.insert(su.yr.end(), re.begin(), re.end());
      su.yi.insert(su.yi.end(), im.begin(), im.end());
      lh::natural_spline(mt.wavelength_nm, re.data(), np, y2.data());
      su.y2r.insert(su.y2r.end(), y2.begin(), y2.end());
      lh::natural_spline(mt.wavelength_nm, im.data(), np, y2.data());
      su.y2i.insert(su.y2i.end(), y2.begin(), y2.end());
    }
    // optical classes: (material, volume, hint), volume = spacing^3 as powi(3)
    std::vector<char> used(n_materials, 0);
    std::map<std::tuple<int, uint64_t, int>, int> classes;
    std::vector<int> cls_of(n);
    for (size_t i = 0; i < n; ++i) {
      const int m = spec->material ? spec->material[i] : 0;
      if (m < 0 || (size_t)m >= n_materials)
        return fail(LH_INVALID_ARGUMENT, "material index out of range at dipole " + std::to_string(i));
      const double d = spec->spacing ? spec->spacing[i] : spec->uniform_spacing;
      if (!(d > 0.0) || !std::isfinite(d))
        return fail(LH_INVALID_ARGUMENT, "spacing must be positive at dipole " + std::to_string(i));
      const int h = spec->hint ? spec->hint[i] : -1;
      if (h >= 0 && ((size_t)h >= spec->n_hints || !spec->hint_depol || !spec->hint_rotation))
        return fail(LH_INVALID_ARGUMENT, "hint index out of range at dipole " + std::to_string(i));
      const double vol = d * d * d;
      const auto key = std::make_tuple(m, __builtin_bit_cast(uint64_t, vol), h < 0 ? -1 : h);
      auto it = classes.find(key);
      if (it == classes.end()) {
        it = classes.emplace(key, (int)su.cls_material.size()).first;
        su.cls_material.push_back(m);
        su.cls_hint.push_back(h < 0 ? -1 : h);
        su.cls_volume.push_back(vol);
      }
      cls_of[i] = it->second;
      used[m] = 1;
    }
    if (su.cls_material.size() > 1) su.cls_of = cls_of;
    if (spec->hint && spec->n_hints > 0) {
      su.hint_depol.assign(spec->hint_depol, spec->hint_depol + 3 * spec->n_hints);
      su.hint_rot.assign(spec->hint_rotation, spec->hint_rotation + 9 * spec->n_hints);
    }
    if (g.has_substrate) {
      const int m = spec->substrate_material;
      if (m < 0 || (size_t)m >= n_materials)
        return fail(LH_INVALID_ARGUMENT, "substrate material index out of range");
      su.sub_material = m;
      su.sub_retarded = spec->substrate_retarded ? 1 : 0;
      used[m] = 1;
    }
    // MaterialError::OutOfRange (johnson_christy.rs:266-274)
    for (size_t w = 0; w < n_wl; ++w) {
      const double L = wavelengths_nm[w];
      if (!(L > 0.0) || !std::isfinite(L))
        return fail(LH_INVALID_ARGUMENT, "wavelengths must be positive and finite");
      for (size_t m = 0; m < n_materials; ++m) {
        if (!used[m]) continue;
        const double lo = materials[m].wavelength_nm[0];
        const double hi = materials[m].wavelength_nm[materials[m].n_points - 1];
        if (L < lo || L > hi)
          return fail(LH_INVALID_ARGUMENT, "material " + std::to_string(m) + ": wavelength " +
                                               std::to_string(L) + " nm outside [" +
                                               std::to_string(lo) + ", " + std::to_string(hi) + "]");
      }
      su.lam.push_back(L);
    }
    // one geometry, one problem per wavelength; optics filled on the device
    LH_HIP(hipSetDevice(b->device));
    std::vector<lh::Geometry> geoms;
    geoms.push_back(make_geometry(&g));
    std::vector<lh::Problem> probs;
    const bool uniform = su.cls_material.size() == 1;
    for (size_t w = 0; w < n_wl; ++w) {
      lh::Problem p = make_problem(&g, 0);
      p.k = 2.0 * M_PI * spec->n_env / wavelengths_nm[w];   // cda/mod.rs:169-171
      p.ainv_uniform = uniform ? 1 : 0;
      p.ainv.assign(uniform ? 9 : 9 * n, lh_c64{0.0, 0.0});
      p.delta = lh_c64{0.0, 0.0};
      probs.push_back(std::move(p));
    }
    auto owned = std::make_unique<lh_batch>();
    owned->b = b;
    owned->amp.assign(n_wl, inc->amplitude);
    LH_HIP(hipStreamCreateWithFlags(&owned->stream, hipStreamNonBlocking));
    bt = owned.release();
    bt->eng = std::make_unique<lh::Engine>(bt->stream);
    bt->eng->build(geoms, probs, tiling_from_env());
    bt->eng->sweep_setup(su, *inc);
    bt->eng->solve(tol, (long long)max_iter, g.precision == LH_PREC_MIXED, true);
    const auto res = bt->eng->results();
    const auto cs = bt->eng->cross_sections(bt->amp);
    int failed = 0;
    for (size_t w = 0; w < n_wl; ++w) {
      const bool ok = res[w].status == LH_OK;
      failed += ok ? 0 : 1;
      for (int c = 0; c < 3; ++c)
        cross_sections[3 * w + c] = ok ? cs[3 * w + c] : std::nan("");
      if (infos) fill_info(res[w], &infos[w]);
    }
    if (failed)
      return fail(LH_CONVERGENCE_FAILURE, std::to_string(failed) + " of " + std::to_string(n_wl) +
                                              " wavelengths did not converge within " +
                                              std::to_string(max_iter) + " iterations");
    return (int)LH_OK;
  });
  std::string msg = g_last_error;
  lh_batch_destroy(bt);
  g_last_error = msg;
  return rc;
}

int lh_batch_solve(lh_backend* b, const lh_batch_problem* problems, size_t count, double tol,
                   size_t max_iter, int precision, double* cross_sections,
                   lh_c64* const* moments, lh_solve_info* infos) {
  lh_batch* bt = nullptr;
  int st = lh_batch_create(b, problems, count, &bt);
  if (st != LH_OK) return st;
  st = lh_batch_run(bt, tol, max_iter, precision);
  if (st == LH_OK) st = lh_batch_results(bt, cross_sections, moments, infos);
  std::string msg = g_last_error;
  lh_batch_destroy(bt);
  g_last_error = msg;
  return st;
}

}  // extern "C"
