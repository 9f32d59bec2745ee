// Host orchestration and C ABI of libdfhip.so.
//
// The call sequence follows contract_coulomb_densities_{2d,4d}
// (diabatization/intor.f90:28-270): shell tables, Schwarz norms, metric,
// Cholesky, screened 3c2e contracted with the densities, B = L^-1 d, and the
// 8-fold Coulomb-tensor assembly.  All arithmetic runs on the GPU; the host
// parses the basis, sorts shell pairs / auxiliary shells by their Schwarz
// norms and derives the screened prefix lengths.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/dfhip.h"
#include "gen3c_classes.h"
#include "grid.h"
#include "int3c2e_impl.h"
#include "int3c2e_mm_impl.h"
#include "linalg.h"
#include "schwarz4c_impl.h"
#include "schwarz4m_impl.h"
#include "schwarz4s_impl.h"

namespace dfhip {

void register_int3c2e_l0(Int3Table&);
void register_int3c2e_l1(Int3Table&);
void register_int3c2e_l2(Int3Table&);
void register_int3c2e_l3(Int3Table&);
void register_int3c2e_l4(Int3Table&);
void register_int3c2e_l5(Int3Table&);
void register_int3c2e_mm_l0(Int3MTable&, EtTable&);
void register_int3c2e_mm_l1(Int3MTable&, EtTable&);
void register_int3c2e_mm_l2(Int3MTable&, EtTable&);
void register_int3c2e_mm_l3(Int3MTable&, EtTable&);
void register_int3c2e_mm_l4(Int3MTable&, EtTable&);
void register_schwarz4c(Schwarz4Table&);
void register_schwarz4c_cached(Schwarz4CachedTable&);
void register_schwarz4m(Schwarz4MTable&);
void register_schwarz4s(Schwarz4STable&);

namespace {

struct Error {
  int code;
  std::string msg;
};

thread_local int t_status = DFHIP_OK;
thread_local std::string t_msg;

#define HC(x)                                                                                     \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw Error{DFHIP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// Cache of device blocks in power-of-two size classes (256 B .. 256 MiB):
// blocks a call frees are kept for the next call, which then pays neither
// hipMalloc nor hipFree (the latter synchronizes the device).  Larger blocks
// are allocated and freed exactly.  All device work runs on the library's
// one stream, so a block handed out again is only touched by work queued
// after the work of its previous owner.  dfhip_release_workspace empties it.
class DevPool {
 public:
  static constexpr int KMIN = 8, KMAX = 28;
  void* get(size_t bytes, size_t& cap) {
    std::lock_guard<std::mutex> g(m_);
    void* p = nullptr;
    if (bytes > ((size_t)1 << KMAX)) {
      cap = bytes;
    } else {
      int k = KMIN;
      while (((size_t)1 << k) < bytes) k++;
      cap = (size_t)1 << k;
      if (!free_[k].empty()) {
        p = free_[k].back();
        free_[k].pop_back();
        return p;
      }
    }
    if (hipMalloc(&p, cap) != hipSuccess) {
      // give the cached blocks back to the runtime and try once more
      release_locked();
      if (hipMalloc(&p, cap) != hipSuccess) return nullptr;
    }
    return p;
  }
  void put(void* p, size_t cap) {
    std::lock_guard<std::mutex> g(m_);
    if (cap > ((size_t)1 << KMAX)) {
      (void)hipFree(p);
      return;
    }
    int k = KMIN;
    while (((size_t)1 << k) < cap) k++;
    free_[k].push_back(p);
  }
  void release() {
    std::lock_guard<std::mutex> g(m_);
    release_locked();
  }

 private:
  void release_locked() {
    for (auto& v : free_) {
      for (void* p : v) (void)hipFree(p);
      v.clear();
    }
  }
  std::mutex m_;
  std::vector<void*> free_[KMAX + 1];
};

DevPool& dev_pool() {
  static DevPool* pool = new DevPool();  // lives until process end
  return *pool;
}

// Pool of the calling thread's lane (see Lane); the default lane uses dev_pool().
thread_local DevPool* t_pool = nullptr;
DevPool& cur_pool() { return t_pool ? *t_pool : dev_pool(); }

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t cap = 0;            // bytes of the pool block
  DevPool* pool = nullptr;   // the pool the block came from
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    pool = &cur_pool();
    p = (T*)pool->get(count * sizeof(T), cap);
    if (!p) throw Error{DFHIP_ERR_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
    n = count;
  }
  void release() {
    if (p) pool->put(p, cap);
    p = nullptr;
    n = 0;
    cap = 0;
  }
  ~DBuf() { release(); }
};

template <class T>
void upload(DBuf<T>& d, const std::vector<T>& h, hipStream_t st) {
  d.alloc(std::max<size_t>(h.size(), 1));
  if (!h.empty()) HC(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}


// ---------------------------------------------------------------------------
// Tables

double dfact(int n) {
  double r = 1.0;
  while (n > 1) {
    r *= n;
    n -= 2;
  }
  return r;
}
double fact(int n) {
  double r = 1.0;
  while (n > 1) r *= n--;
  return r;
}
double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return r;
}

// Real solid harmonics S_lm (m = -l..l) as Cartesian polynomials, normalized
// like x^l and divided by sqrt((2l-1)!!) (Helgaker, Jorgensen, Olsen,
// Molecular Electronic-Structure Theory, 6.4.47-6.4.50).
void build_sph_table(SphTable& t) {
  int n = 0;
  for (int L = 0; L <= LMAX_SPH; L++) {
    const double scale = 1.0 / std::sqrt(dfact(2 * L - 1));
    for (int m = -L; m <= L; m++) {
      t.start[L][m + L] = n;
      const int am = std::abs(m), vm2 = m < 0 ? 1 : 0;
      const double N = 1.0 / (std::pow(2.0, am) * fact(L)) *
                       std::sqrt(2.0 * fact(L + am) * fact(L - am) / (m == 0 ? 2.0 : 1.0));
      // accumulate per monomial
      std::vector<double> coef((L + 1) * (L + 1), 0.0);  // index lx*(L+1)+ly
      for (int tt = 0; tt <= (L - am) / 2; tt++)
        for (int u = 0; u <= tt; u++)
          for (int v2 = vm2; v2 <= am; v2 += 2) {
            const int lx = 2 * tt + am - 2 * u - v2, ly = 2 * u + v2;
            const double sgn = ((tt + (v2 - vm2) / 2) & 1) ? -1.0 : 1.0;
            coef[lx * (L + 1) + ly] +=
                sgn * std::pow(0.25, tt) * binom(L, tt) * binom(L - tt, am + tt) * binom(tt, u) * binom(am, v2);
          }
      for (int lx = L; lx >= 0; lx--)
        for (int ly = L - lx; ly >= 0; ly--) {
          const double c = coef[lx * (L + 1) + ly];
          if (c == 0.0) continue;
          if (n >= SPH_MAX_TERMS) throw Error{DFHIP_ERR_ARG, "solid-harmonic table overflow"};
          t.terms[n++] = SphTerm{N * c * scale, lx, ly, L - lx - ly};
        }
    }
    t.start[L][2 * L + 1] = n;
  }
}

// F_n(k h) by the convergent series at the highest order and downward
// recursion.
void build_boys_table(std::vector<double>& tab) {
  tab.assign((size_t)BOYS_NGRID * BOYS_NORD, 0.0);
  const int nmax = BOYS_NORD - 1;
  for (int k = 0; k < BOYS_NGRID; k++) {
    const double T = k * BOYS_H;
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int j = 1; j < 4000; j++) {
      term *= 2.0 * T / (2 * nmax + 2 * j + 1);
      sum += term;
      if (term < 1e-18 * sum) break;
    }
    const double eT = std::exp(-T);
    double* f = &tab[(size_t)k * BOYS_NORD];
    f[nmax] = eT * sum;
    for (int n = nmax - 1; n >= 0; n--) f[n] = (2.0 * T * f[n + 1] + eT) / (2 * n + 1);
  }
}

// A stream with its own workspace and block pool.  Calls run on the default
// lane; a batch runs its molecules concurrently on several lanes, one host
// thread each.  Blocks are only reused within their lane, i.e. by work queued
// on the same stream after the work of their previous owner.
// Pinned host buffer, grown on demand: a device-to-host copy into it returns
// at once, one into pageable memory waits for the stream to reach it.
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T>
  T* get(size_t n) {
    const size_t need = std::max<size_t>(n, 1) * sizeof(T);
    if (need > bytes) {
      release();
      HC(hipHostMalloc(&p, need, hipHostMallocDefault));
      bytes = need;
    }
    return (T*)p;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Bump allocator over pinned host memory for the uploads queued behind
// running kernels.  A pageable upload goes through the runtime's staging
// buffers, which are reused only once the copies holding them have run, so a
// series of pageable uploads queued behind long kernels stalls the host;
// copies from pinned memory return at once.  reset() only when no copy from
// the arena is in flight (at the start of a call: every call ends with its
// streams drained).
struct PinnedArena {
  std::vector<PinnedBuf> chunks;
  size_t used = 0;   // bytes used in the last chunk
  size_t total = 0;  // bytes handed out since the last reset
  void reset() {
    if (chunks.size() > 1) {
      // one chunk for what the last call needed
      const size_t need = total;
      for (auto& c : chunks) c.release();
      chunks.assign(1, PinnedBuf{});
      chunks[0].get<char>(need);
    }
    used = 0;
    total = 0;
  }
  template <class T>
  T* alloc(size_t n) {
    const size_t b = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
    if (chunks.empty() || used + b > chunks.back().bytes) {
      chunks.emplace_back();
      chunks.back().get<char>(std::max<size_t>(b, (size_t)4 << 20));
      used = 0;
    }
    T* p = (T*)((char*)chunks.back().p + used);
    used += b;
    total += b;
    return p;
  }
  void release() {
    for (auto& c : chunks) c.release();
    chunks.clear();
    used = total = 0;
  }
};

struct Lane {
  hipStream_t st = nullptr;
  // side stream of the lane: the metric factorization runs there while the
  // integrals run on st (created on first use)
  hipStream_t st2 = nullptr;
  // (and the density packing before it, concurrent with the Schwarz norms)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_pack = nullptr;
  int* hinfo = nullptr;  // pinned host copy of the factorization's status
  void ensure_side() {
    if (st2) return;
    HC(hipHostMalloc((void**)&hinfo, sizeof(int), hipHostMallocDefault));
    HC(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
    HC(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HC(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HC(hipEventCreateWithFlags(&ev_pack, hipEventDisableTiming));
  }
  // the Schwarz partial sums and the aux norms come back through pinned
  // buffers, each signalled by its own event, so that the host can build the
  // pair part of the plan while the Schwarz kernels of later classes and the
  // metric still run
  hipEvent_t ev_met = nullptr;
  PinnedBuf hpart, hqc, hqvl;
  PinnedArena arena;
  // copy stream: the metric's inputs go up on it while the Schwarz kernels
  // run (ev_mprep); the Schwarz sums come down on it in NSHIP shipments,
  // each as soon as its classes' kernels end (ev_ship[i], on the host at
  // ev_shipd[i]) while st goes on with the next classes: all classes but the
  // last NSHIP - 1, then those one by one; the sorted pairs go up and their
  // bra records are built on it
  // while the metric runs (ev_plan).  ev_pairs: the device pair list is
  // ready.
  hipStream_t st3 = nullptr;
  static constexpr int NSHIP = 3;
  hipEvent_t ev_mprep = nullptr, ev_pairs = nullptr, ev_plan = nullptr, ev_ship[NSHIP] = {}, ev_shipd[NSHIP] = {};
  void ensure_events() {
    if (ev_met) return;
    HC(hipEventCreateWithFlags(&ev_met, hipEventDisableTiming));
    HC(hipStreamCreateWithFlags(&st3, hipStreamNonBlocking));
    HC(hipEventCreateWithFlags(&ev_mprep, hipEventDisableTiming));
    HC(hipEventCreateWithFlags(&ev_pairs, hipEventDisableTiming));
    HC(hipEventCreateWithFlags(&ev_plan, hipEventDisableTiming));
    for (int i = 0; i < NSHIP; i++) {
      HC(hipEventCreateWithFlags(&ev_ship[i], hipEventDisableTiming));
      HC(hipEventCreateWithFlags(&ev_shipd[i], hipEventDisableTiming));
    }
  }
  DevPool pool;
  // device workspace reused across calls (grown on demand, freed by
  // dfhip_release_workspace)
  DBuf<double> wsT, wsA, wsDp, wsPart, wsDraw, wsDT;
  void release() {
    wsDT.release();
    arena.release();
    hpart.release();
    hqc.release();
    hqvl.release();
    wsT.release();
    wsA.release();
    wsDp.release();
    wsPart.release();
    wsDraw.release();
    pool.release();
  }
};

// Device scratch of the allocation-free entry point
// (dfhip_int_schwarz_aux_noalloc): every GPU the library runs on holds its own
// copy of this array, at its own address, until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is only
// a stand-in for the symbol): it uses the device address that
// hipGetSymbolAddress returns for the current device, kept per device in
// Ctx::scratch_dev.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;  // 16 MiB per device
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct Ctx {
  bool ready = false;
  int device = -1;
  unsigned char* scratch_dev[NOALLOC_MAX_DEVICES] = {};  // device address of `scratch`, per device
  Lane lane0;                                 // default lane (its blocks come from dev_pool())
  std::vector<std::unique_ptr<Lane>> lanes;   // batch lanes 1..
  DBuf<double> boys;
  DBuf<SphTable> sph;
  Int3Table int3{};
  Int3MTable int3m{};
  EtTable et{};
  Schwarz4Table s4{};
  Schwarz4CachedTable s4c{};
  Schwarz4MTable s4m{};
  Schwarz4STable s4s{};
  Gen3Table gen3{};
  rocblas_handle blas = nullptr;  // rocBLAS / rocSOLVER handle of the eigen route (created on first use)
};

template <class T>
T* reserve(DBuf<T>& b, size_t n) {
  n = std::max<size_t>(n, 1);
  if (b.n < n) b.alloc(n);
  return b.p;
}

Ctx& ctx() {
  static Ctx* c = new Ctx();  // lives until process end
  return *c;
}
std::mutex& ctx_mutex() {
  static std::mutex m;
  return m;
}

thread_local Lane* t_lane = nullptr;
Lane& cur_lane() { return t_lane ? *t_lane : ctx().lane0; }

// A batch runs several lanes at once: their copy-stream work goes on each
// lane's own stream instead, so that the streams of all lanes stay within
// the hardware queues (work of streams that share a queue serializes).
thread_local bool t_few_streams = false;
// lanes of the batch call running beside this one: the automatic HBM
// budgets of a call (staging matrix, Et tables) are shares of the free memory
// divided by it, so that concurrent lanes do not over-commit the device
thread_local int t_lanes = 1;
hipStream_t copy_stream(Lane& l) { return t_few_streams ? l.st : l.st3; }

// Uploads through the lane's pinned arena, for uploads queued behind running
// kernels (see PinnedArena).
template <class T>
void h2d_pinned(T* d, const T* h, size_t n, hipStream_t st) {
  if (!n) return;
  T* b = cur_lane().arena.alloc<T>(n);
  std::memcpy(b, h, n * sizeof(T));
  HC(hipMemcpyAsync(d, b, n * sizeof(T), hipMemcpyHostToDevice, st));
}

template <class T>
void upload_pinned(DBuf<T>& d, const std::vector<T>& h, hipStream_t st) {
  d.alloc(std::max<size_t>(h.size(), 1));
  h2d_pinned(d.p, h.data(), h.size(), st);
}

// Several arrays in one device block, staged in one pinned block and sent
// with one copy (each copy costs the stream a few microseconds whatever its
// size): add() reserves an aligned range and returns its offset; after
// allocate(), host<T>(off) is where the caller writes the array and
// dev<T>(off) where the device will find it once commit() has queued the
// copy.
struct UploadBatch {
  size_t size = 0;
  char* h = nullptr;
  DBuf<char> d;
  template <class T>
  size_t add(size_t n) {
    const size_t off = size;
    size += (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
    return off;
  }
  void allocate() {
    h = cur_lane().arena.alloc<char>(size);
    d.alloc(std::max<size_t>(size, 1));
  }
  template <class T>
  T* host(size_t off) {
    return (T*)(h + off);
  }
  template <class T>
  T* dev(size_t off) {
    return (T*)(d.p + off);
  }
  void commit(hipStream_t st) {
    if (size) HC(hipMemcpyAsync(d.p, h, size, hipMemcpyHostToDevice, st));
  }
};

// A device address with the .p of a DBuf, for arrays inside an UploadBatch.
template <class T>
struct DView {
  T* p = nullptr;
};

// Device address of the allocation-free scratch array on device dev (the
// current device), looked up once per device.
unsigned char* resolve_scratch(Ctx& c, int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw Error{DFHIP_ERR_UNSUPPORTED, "device index out of range"};
  if (!c.scratch_dev[dev]) {
    void* p = nullptr;
    HC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error{DFHIP_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array"};
    c.scratch_dev[dev] = (unsigned char*)p;
  }
  return c.scratch_dev[dev];
}

void load_env_tuning();  // the table named by DFHIP_TUNING, if any

void ensure_ctx() {
  Ctx& c = ctx();
  if (c.ready) return;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Error{DFHIP_ERR_HIP, "no HIP device available"};
  HC(hipGetDevice(&c.device));
  resolve_scratch(c, c.device);
  load_env_tuning();
  HC(hipStreamCreateWithFlags(&c.lane0.st, hipStreamNonBlocking));
  const hipStream_t st0 = c.lane0.st;
  std::vector<double> tab;
  build_boys_table(tab);
  upload(c.boys, tab, st0);
  SphTable* h = new SphTable();
  std::memset(h, 0, sizeof(SphTable));
  build_sph_table(*h);
  c.sph.alloc(1);
  HC(hipMemcpyAsync(c.sph.p, h, sizeof(SphTable), hipMemcpyHostToDevice, st0));
  HC(hipStreamSynchronize(st0));
  delete h;
  register_int3c2e_mm_l0(c.int3m, c.et);
  register_int3c2e_mm_l1(c.int3m, c.et);
  register_int3c2e_mm_l2(c.int3m, c.et);
  register_int3c2e_mm_l3(c.int3m, c.et);
  register_int3c2e_mm_l4(c.int3m, c.et);
  register_int3c2e_l0(c.int3);
  register_int3c2e_l1(c.int3);
  register_int3c2e_l2(c.int3);
  register_int3c2e_l3(c.int3);
  register_int3c2e_l4(c.int3);
  register_int3c2e_l5(c.int3);
  register_schwarz4c(c.s4);
  register_schwarz4c_cached(c.s4c);
  register_schwarz4m(c.s4m);
  register_schwarz4s(c.s4s);
  register_gen3_all(c.gen3);
  for (int la = 0; la <= LMAX_AUX; la++)
    for (int lb = 0; lb <= (la <= LMAX_ORB ? la : 0); lb++)
      for (int lc = 0; lc <= LMAX_AUX; lc++)
        if (!c.int3[la][lb][lc].fn) throw Error{DFHIP_ERR_UNSUPPORTED, "missing 3c2e class kernel"};
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++)
      if (!c.s4[la][lb].fn) throw Error{DFHIP_ERR_UNSUPPORTED, "missing 4c2e class kernel"};
  c.ready = true;
}

// ---------------------------------------------------------------------------
// Shells

struct HostShells {
  int n = 0, nbf = 0, lmax = -1, maxprim = 0;
  std::vector<int> L, nprim, pptr, f0;
  std::vector<double> xyz, exps, coefs;
};

// common/normalization.f90:10-38
void norm_cgto_coeffs(double* c, const double* e, int n, int L) {
  double N = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      N += c[i] * c[j] / std::pow(e[i] + e[j], L + 1.5) * std::pow(std::sqrt(e[i] * e[j]), L + 1.5);
  for (int i = 0; i < n; i++) c[i] *= std::sqrt(std::pow(e[i], L + 1.5) / (std::pow(PI, 1.5) / std::pow(2.0, L) * N));
}

// common/shells.f90:97-187
HostShells parse_shells(const dfhip_basis& b, int lmax_allowed, const char* what) {
  HostShells s;
  if (b.nshells < 1 || !b.bas_centers || !b.bas_spec || !b.bas_data)
    throw Error{DFHIP_ERR_ARG, std::string(what) + ": empty basis"};
  const int n = b.nshells;
  s.n = n;
  s.L.resize(n);
  s.nprim.resize(n);
  s.pptr.resize(n);
  s.f0.resize(n);
  s.xyz.resize(3 * n);
  for (int i = 0; i < n; i++) {
    const int cptr = b.bas_centers[i + 2 * n] - 1;
    const int L = b.bas_spec[i + n];
    const int np = b.bas_spec[i + 2 * n];
    const int pp = b.bas_spec[i + 4 * n] - 1;
    if (L < 0 || L > lmax_allowed)
      throw Error{L < 0 ? DFHIP_ERR_ARG : DFHIP_ERR_UNSUPPORTED,
                  std::string(what) + ": shell " + std::to_string(i) + " has L=" + std::to_string(L) +
                      " (supported 0.." + std::to_string(lmax_allowed) + ")"};
    if (np < 1) throw Error{DFHIP_ERR_ARG, std::string(what) + ": shell without primitives"};
    if (cptr < 0 || cptr + 3 > b.ndata || pp < 0 || pp + 2 * np > b.ndata)
      throw Error{DFHIP_ERR_ARG, std::string(what) + ": pointer outside bas_data"};
    s.L[i] = L;
    s.nprim[i] = np;
    s.pptr[i] = (int)s.exps.size();
    s.f0[i] = s.nbf;
    s.nbf += 2 * L + 1;
    s.lmax = std::max(s.lmax, L);
    s.maxprim = std::max(s.maxprim, np);
    for (int d = 0; d < 3; d++) s.xyz[3 * i + d] = b.bas_data[cptr + d];
    std::vector<double> c(b.bas_data + pp, b.bas_data + pp + np), e(b.bas_data + pp + np, b.bas_data + pp + 2 * np);
    for (double x : e)
      if (!(x > 0.0)) throw Error{DFHIP_ERR_ARG, std::string(what) + ": non-positive exponent"};
    norm_cgto_coeffs(c.data(), e.data(), np, L);
    s.exps.insert(s.exps.end(), e.begin(), e.end());
    s.coefs.insert(s.coefs.end(), c.begin(), c.end());
  }
  return s;
}

struct DevShellBufs {
  DBuf<int> L, nprim, pptr, f0;
  DBuf<double> xyz, exps, coefs;
  DevShells view{};
  void upload_from(const HostShells& h, hipStream_t st) {
    upload(L, h.L, st);
    upload(nprim, h.nprim, st);
    upload(pptr, h.pptr, st);
    upload(f0, h.f0, st);
    upload(xyz, h.xyz, st);
    upload(exps, h.exps, st);
    upload(coefs, h.coefs, st);
    view = DevShells{h.n, L.p, nprim.p, pptr.p, f0.p, xyz.p, exps.p, coefs.p};
  }
};

// ---------------------------------------------------------------------------
// Timing

struct Timer {
  hipStream_t st;
  std::vector<hipEvent_t> ev;
  explicit Timer(hipStream_t s) : st(s) {}
  int mark(hipStream_t on = nullptr) {
    hipEvent_t e;
    HC(hipEventCreate(&e));
    HC(hipEventRecord(e, on ? on : st));
    ev.push_back(e);
    return (int)ev.size() - 1;
  }
  double ms(int a, int b) {
    float t = 0.f;
    HC(hipEventSynchronize(ev[b]));
    HC(hipEventElapsedTime(&t, ev[a], ev[b]));
    return t;
  }
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// ---------------------------------------------------------------------------
// Pairs and classes

constexpr int NPCLS = (LMAX_ORB + 1) * (LMAX_ORB + 2) / 2;
inline int pcls(int la, int lb) { return la * (la + 1) / 2 + lb; }

struct PairSet {
  // global pair order: class by class; inside a class as given by the caller
  std::vector<int> s1, s2, packed;
  std::vector<int> cls_begin, cls_end;  // per class, into the global order
  int n() const { return (int)s1.size(); }
};

// All a >= b shell pairs, oriented with L(s1) >= L(s2) and grouped by class
// (la, lb); the order inside a class is that of PairClassTable, which
// pair_classes fills and gen_pairs follows on the device.
struct PairClasses {
  PairClassTable t;
  std::vector<int> shflat;  // shells of each L, ascending, L by L
};

PairClasses pair_classes(const HostShells& o) {
  PairClasses pc;
  std::memset(&pc.t, 0, sizeof(pc.t));
  std::vector<int> loff(LMAX_ORB + 2, 0), nl(LMAX_ORB + 1, 0);
  for (int a = 0; a < o.n; a++) nl[o.L[a]]++;
  for (int L = 0; L <= LMAX_ORB; L++) loff[L + 1] = loff[L] + nl[L];
  pc.shflat.resize(o.n);
  std::vector<int> pos(loff.begin(), loff.end() - 1);
  for (int a = 0; a < o.n; a++) pc.shflat[pos[o.L[a]]++] = a;
  PairClassTable& t = pc.t;
  t.ncls = NPCLS;
  int p = 0;
  int64_t row = 0;
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++) {
      const int k = pcls(la, lb);
      t.begin[k] = p;
      t.la[k] = la;
      t.lb[k] = lb;
      t.off_a[k] = loff[la];
      t.off_b[k] = loff[lb];
      t.n_b[k] = nl[lb];
      t.w[k] = (2 * la + 1) * (2 * lb + 1);
      t.row0[k] = row;
      const int n = la != lb ? nl[la] * nl[lb] : nl[la] * (nl[la] + 1) / 2;
      p += n;
      row += (int64_t)n * t.w[k];
    }
  t.begin[NPCLS] = p;
  return pc;
}

// The host copy of the pair list in the order of pc (gen_pairs' order).
PairSet make_pairs(const HostShells& o, const PairClasses& pc) {
  const PairClassTable& t = pc.t;
  PairSet ps;
  const int np = t.begin[NPCLS];
  ps.s1.resize(np);
  ps.s2.resize(np);
  ps.packed.resize(np);
  ps.cls_begin.assign(t.begin, t.begin + NPCLS);
  ps.cls_end.assign(t.begin + 1, t.begin + NPCLS + 1);
  for (int k = 0; k < NPCLS; k++) {
    int p = t.begin[k];
    if (t.la[k] != t.lb[k]) {
      const int na = t.n_b[k] ? (t.begin[k + 1] - t.begin[k]) / t.n_b[k] : 0;
      for (int ia = 0; ia < na; ia++)
        for (int ib = 0; ib < t.n_b[k]; ib++, p++) {
          ps.s1[p] = pc.shflat[t.off_a[k] + ia];
          ps.s2[p] = pc.shflat[t.off_b[k] + ib];
        }
    } else {
      for (int ia = 0; p < t.begin[k + 1]; ia++)
        for (int ib = 0; ib <= ia; ib++, p++) {
          ps.s1[p] = pc.shflat[t.off_a[k] + ia];
          ps.s2[p] = pc.shflat[t.off_b[k] + ib];
        }
    }
  }
  for (int p = 0; p < np; p++) {
    const int a = std::max(ps.s1[p], ps.s2[p]), b = std::min(ps.s1[p], ps.s2[p]);
    ps.packed[p] = a * (a + 1) / 2 + b;
  }
  return ps;
}

// Schwarz norms of all pairs, in the PairSet's order.  Classes whose pairs all
// fit the primitive-pair cache use the cached kernel (one partial per pair),
// the others the chunked kernel.
// Schwarz norms in two parts, so that the caller can queue other work
// behind the kernels: schwarz_bra_launch enqueues them on the lane's stream
// and, class by class, the copies of their partial sums to the host on the
// lane's copy stream; schwarz_class_finish waits for one class's sums and
// reduces them to Q_ab.
struct SchwarzPending {
  DBuf<int> s1, s2, shflat;
  DBuf<int64_t> prow0;
  DBuf<double> part, scr;
  std::vector<size_t> poff;
  std::vector<int> nch;
  std::vector<int> ship;  // per class: the shipment that brings its sums
  bool shipped[Lane::NSHIP] = {};
};

// With sk.thresh > 0 the pairs whose bound on Q_ab keeps no triplet get a
// zero sum without their quartets being evaluated.
void schwarz_bra_launch(Ctx& c, const HostShells& o, const DevShells& dev, const PairClasses& pc, SchwarzPending& sp,
                        const PairSkip& sk = PairSkip{}) {
  hipStream_t st = cur_lane().st;
  // the pair list (and each pair's first row) generated on the device
  const int npairs = pc.t.begin[NPCLS];
  DBuf<int>& s1 = sp.s1;
  DBuf<int>& s2 = sp.s2;
  s1.alloc(std::max(npairs, 1));
  s2.alloc(std::max(npairs, 1));
  sp.prow0.alloc(std::max(npairs, 1));
  Lane& lane = cur_lane();
  lane.ensure_events();
  upload_pinned(sp.shflat, pc.shflat, st);
  gen_pairs(pc.t, sp.shflat.p, s1.p, s2.p, sp.prow0.p, st);
  HC(hipGetLastError());
  HC(hipEventRecord(lane.ev_pairs, st));  // other streams may read the list after this
  // shells of each L with the most primitives
  int maxprim[LMAX_ORB + 1] = {};
  for (int a = 0; a < o.n; a++) maxprim[o.L[a]] = std::max(maxprim[o.L[a]], o.nprim[a]);
  PairSet ps;  // class ranges only
  ps.cls_begin.assign(pc.t.begin, pc.t.begin + NPCLS);
  ps.cls_end.assign(pc.t.begin + 1, pc.t.begin + NPCLS + 1);
  std::vector<size_t>& poff = sp.poff;
  poff.assign(NPCLS + 1, 0);
  // kernel per class: 3 = matrix cores with the tables in LDS, 2 = matrix
  // cores streaming the tables from global scratch, 1 = cached VALU,
  // 0 = chunked VALU (DFHIP_SCHWARZ_KERNEL caps the choice, for tests)
  std::vector<int>& nch = sp.nch;
  nch.assign(NPCLS, 1);
  std::vector<int> kind(NPCLS, 0);
  const char* cap_env = std::getenv("DFHIP_SCHWARZ_KERNEL");
  const int cap = (cap_env && *cap_env) ? std::atoi(cap_env) : 3;
  size_t scr_need = 0;
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++) {
      const int k = pcls(la, lb);
      const int np = ps.cls_end[k] - ps.cls_begin[k];
      const int maxnpp = np > 0 ? maxprim[la] * maxprim[lb] : 0;
      if (cap >= 3 && c.s4m[la][lb].fn != nullptr && maxnpp <= c.s4m[la][lb].nppm) {
        kind[k] = 3;
      } else if (cap >= 2 && c.s4s[la][lb].fn != nullptr && maxnpp <= c.s4s[la][lb].npps) {
        kind[k] = 2;
        scr_need = std::max(scr_need, (size_t)np * (size_t)c.s4s[la][lb].perpair);
      } else if (cap >= 1 && c.s4c[la][lb].fn != nullptr && maxnpp <= c.s4c[la][lb].nppc) {
        kind[k] = 1;
      }
      nch[k] = kind[k] ? 1 : c.s4[la][lb].nch;
      poff[k + 1] = (size_t)(ps.cls_end[k] - ps.cls_begin[k]) * nch[k];
    }
  for (int k = 0; k < NPCLS; k++) poff[k + 1] += poff[k];
  DBuf<double>& part = sp.part;
  part.alloc(std::max<size_t>(poff[NPCLS], 1));
  // scratch of the streamed kernels: all pairs of the largest class, at most 4 GiB
  DBuf<double>& scr = sp.scr;
  scr.alloc(std::min<size_t>(scr_need, (size_t)1 << 29));
  double* hp = lane.hpart.get<double>(poff[NPCLS]);
  // shipments: the last NSHIP - 1 non-empty classes one each, the classes
  // before them in the first (shipment i ends after class shipend[i])
  constexpr int NS = Lane::NSHIP;
  std::vector<int> nonempty;
  for (int k = 0; k < NPCLS; k++)
    if (ps.cls_end[k] > ps.cls_begin[k]) nonempty.push_back(k);
  int shipend[NS];
  for (int i = 0; i < NS; i++) {
    const int j = (int)nonempty.size() - NS + i;  // shipment i ends after nonempty[j]
    shipend[i] = j >= 0 ? nonempty[j] : -1;
  }
  sp.ship.assign(NPCLS, NS - 1);
  for (int k = NPCLS - 1; k >= 0; k--)
    for (int i = 0; i < NS; i++)
      if (k <= shipend[i]) {
        sp.ship[k] = i;
        break;
      }
  auto ship = [&](int i, size_t b, size_t e) {
    HC(hipEventRecord(lane.ev_ship[i], st));
    const hipStream_t s3 = copy_stream(lane);
    HC(hipStreamWaitEvent(s3, lane.ev_ship[i], 0));
    if (e > b) HC(hipMemcpyAsync(hp + b, part.p + b, (e - b) * sizeof(double), hipMemcpyDeviceToHost, s3));
    HC(hipEventRecord(lane.ev_shipd[i], s3));
    sp.shipped[i] = true;
  };
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++) {
      const int k = pcls(la, lb), b0 = ps.cls_begin[k], np = ps.cls_end[k] - b0;
      if (np == 0) continue;
      if (kind[k] == 3)
        c.s4m[la][lb].fn(dev, s1.p + b0, s2.p + b0, np, c.sph.p, c.boys.p, part.p + poff[k], sk, st);
      else if (kind[k] == 2)
        c.s4s[la][lb].fn(dev, s1.p + b0, s2.p + b0, np, c.sph.p, c.boys.p, part.p + poff[k], scr.p, scr.n, sk, st);
      else if (kind[k] == 1)
        c.s4c[la][lb].fn(dev, s1.p + b0, s2.p + b0, np, c.sph.p, c.boys.p, part.p + poff[k], st);
      else
        c.s4[la][lb].fn(dev, s1.p + b0, s2.p + b0, np, c.sph.p, c.boys.p, part.p + poff[k], sk, st);
      HC(hipGetLastError());
      for (int i = 0; i < NS - 1; i++)
        if (k == shipend[i]) ship(i, i > 0 && shipend[i - 1] >= 0 ? poff[shipend[i - 1] + 1] : 0, poff[k + 1]);
    }
  ship(NS - 1, NS > 1 && shipend[NS - 2] >= 0 ? poff[shipend[NS - 2] + 1] : 0, poff[NPCLS]);
}

// Q_ab of the pairs of class k (make_pairs order) into q[cls_begin[k] ..];
// waits for that class's sums only (work queued behind them may still run).
void schwarz_class_finish(const PairSet& ps, const SchwarzPending& sp, int k, double* q) {
  Lane& lane = cur_lane();
  const std::vector<size_t>& poff = sp.poff;
  const int nch = sp.nch[k];
  if (sp.shipped[sp.ship[k]]) HC(hipEventSynchronize(lane.ev_shipd[sp.ship[k]]));
  const double* hp = (const double*)lane.hpart.p;
  for (int p = ps.cls_begin[k]; p < ps.cls_end[k]; p++) {
    double s = 0.0;
    for (int z = 0; z < nch; z++) s += hp[poff[k] + (size_t)(p - ps.cls_begin[k]) * nch + z];
    q[p] = std::sqrt(std::sqrt(s));
    if (!std::isfinite(q[p])) throw Error{DFHIP_ERR_HIP, "non-finite Schwarz norm"};
  }
}

std::vector<double> schwarz_bra_finish(const PairSet& ps, SchwarzPending& sp) {
  std::vector<double> q(ps.n());
  for (int k = 0; k < NPCLS; k++) schwarz_class_finish(ps, sp, k, q.data());
  return q;
}

std::vector<double> schwarz_bra_dev(Ctx& c, const HostShells& o, const DevShells& dev, const PairClasses& pc,
                                    const PairSet& ps) {
  SchwarzPending sp;
  schwarz_bra_launch(c, o, dev, pc, sp);
  return schwarz_bra_finish(ps, sp);
}

bool use_matrix_cores_3c() {
  const char* v = std::getenv("DFHIP_NO_MM");
  return !(v && *v && *v != '0');
}

bool use_generated() {
  const char* v = std::getenv("DFHIP_NO_GEN");
  return !(v && *v && *v != '0');
}

inline bool gen3_has(int la, int lb, int lc) {
  return la <= GEN3_LA_MAX && lb <= la && lc <= GEN3_LC_MAX;
}

// Classes with a generated kernel where the matrix-core kernel is faster on
// MI355X (measured per class on the 60-atom s-d / s-f workload: (dd|d) 577
// against 776 us, (dd|f) 388 against 557 us); in mixed mode their FP64
// window goes to the matrix-core kernel, the FP32 window stays generated.
inline bool prefer_mm(int la, int lb, int lc) { return la == 2 && lb == 2 && lc >= 2; }

// Kernel choices measured by dfhip_autotune or loaded from its file.
// fam[k][L][w]: for window w (0 = FP64, 1 = the FP32 window of the mixed
// path) of class (pair class k, aux L) 1 = generated VALU kernel, 2 =
// matrix-core kernel, 0 = not tuned (the rule above decides).  contract:
// per density count, the densities per thread of the contraction's output
// tile and the scale of its split-K depth.
struct Tuning {
  signed char fam[NPCLS][LMAX_AUX + 1][2] = {};
  double ms[NPCLS][LMAX_AUX + 1][2][2] = {};  // measured [w][gen, mm] (for the file)
  struct Contract {
    int ndens, ipt;
    double nscale, ms;
  };
  std::vector<Contract> contract;
  const Contract* contract_for(int ndens) const {
    for (const auto& c : contract)
      if (c.ndens == ndens) return &c;
    return nullptr;
  }
};
Tuning& tuning() {
  static Tuning* t = new Tuning();
  return *t;
}
std::mutex& tuning_mutex() {
  static std::mutex m;
  return m;
}

// Set by dfhip_autotune on its thread for the duration of one measured run:
// forced choices (0 = the table / rule) and the measured times.
struct TuneProbe {
  int fam64 = 0, fam32 = 0, ipt = 0;
  double nscale = 0.0;
  double ms[NPCLS][LMAX_AUX + 1][2] = {};
  bool seen[NPCLS][LMAX_AUX + 1][2] = {};
};
thread_local TuneProbe* t_probe = nullptr;

// family of window w of class (k, L): the probe's, else the table's, else 0
int tuned_family(int k, int L, int w) {
  if (t_probe && (w == 0 ? t_probe->fam64 : t_probe->fam32)) return w == 0 ? t_probe->fam64 : t_probe->fam32;
  std::lock_guard<std::mutex> g(tuning_mutex());
  return tuning().fam[k][L][w];
}

void launch_bra_records(int la, int lb, DevShells bra, const int* s1, const int* s2, const int64_t* recoff, int rs,
                        int npairs, double* recs, hipStream_t st) {
  if (npairs <= 0) return;
#define DFHIP_BR(A, B)                                                                                          \
  if (la == A && lb == B) {                                                                                    \
    if (rs != 4 + BraRecMap<A, B>::N) throw Error{DFHIP_ERR_HIP, "bra record size mismatch"};                    \
    hipLaunchKernelGGL((bra_records_kernel<A, B>), dim3(npairs), dim3(BRA_THREADS), 0, st, bra, s1, s2, recoff, rs, \
                       recs);                                                                                  \
    return;                                                                                                    \
  }
  DFHIP_BR(0, 0) DFHIP_BR(1, 0) DFHIP_BR(1, 1) DFHIP_BR(2, 0) DFHIP_BR(2, 1) DFHIP_BR(2, 2)
#undef DFHIP_BR
  throw Error{DFHIP_ERR_UNSUPPORTED, "no bra-record kernel for this class"};
}

// Bra records of the generated kernels for one pair class: offsets (in
// doubles, per pair), primitive-pair counts and the records themselves.  off
// and npp point into off_buf / npp_buf or into a caller's upload batch.
struct BraRecords {
  DBuf<int64_t> off_buf;
  DBuf<int> npp_buf;
  const int64_t* off = nullptr;
  const int* npp = nullptr;
  DBuf<double> recs;
  DBuf<float> recs32;  // mixed precision only
};

// Record layout of n pairs: off[p], npp[p] (host arrays); returns the total
// number of doubles.
int64_t bra_record_layout(int rs, const HostShells& bra, const int* s1, const int* s2, int n, int64_t* off,
                          int* npp) {
  int64_t o = 0;
  for (int p = 0; p < n; p++) {
    off[p] = o;
    npp[p] = bra.nprim[s1[p]] * (s2[p] >= 0 ? bra.nprim[s2[p]] : 1);
    o += (int64_t)npp[p] * rs;
  }
  return o;
}

// FP32 copy of the bra records for the FP32 class kernels.
__global__ void records_to_f32_kernel(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// The records of n pairs whose layout (br.off, br.npp, total o doubles) is
// on the device already.
void fill_bra_records(Ctx& c, int la, int lb, const DevShells& dev, const int* ds1, const int* ds2, int n, int64_t o,
                      BraRecords& br, bool fp32, hipStream_t s) {
  const int rs = c.gen3[la][lb][0].rs;
  br.recs.alloc(std::max<int64_t>(o, 1));
  launch_bra_records(la, lb, dev, ds1, ds2, br.off, rs, n, br.recs.p, s);
  HC(hipGetLastError());
  if (fp32 && o > 0) {
    br.recs32.alloc(o);
    hipLaunchKernelGGL(records_to_f32_kernel, dim3((unsigned)((o + 255) / 256)), dim3(256), 0, s, br.recs.p,
                       br.recs32.p, o);
    HC(hipGetLastError());
  }
}

// Layout, upload and records of n pairs (host s1/s2, device ds1/ds2).
void build_bra_records(Ctx& c, int la, int lb, const HostShells& bra, const DevShells& dev, const int* ds1,
                       const int* ds2, const int* s1, const int* s2, int n, BraRecords& br, bool fp32 = false,
                       hipStream_t s = nullptr) {
  if (!s) s = cur_lane().st;
  std::vector<int64_t> off(n);
  std::vector<int> npp(n);
  const int64_t o = bra_record_layout(c.gen3[la][lb][0].rs, bra, s1, s2, n, off.data(), npp.data());
  upload_pinned(br.off_buf, off, s);
  upload_pinned(br.npp_buf, npp, s);
  br.off = br.off_buf.p;
  br.npp = br.npp_buf.p;
  fill_bra_records(c, la, lb, dev, ds1, ds2, n, o, br, fp32, s);
}

void launch_gen3(Ctx& c, int la, int lb, int lc, const BraRecords& br, DevShells ket, PairArgs pa, AuxArgs xa,
                 double* out, int64_t ldo, bool fp32 = false, hipStream_t st = nullptr) {
  if (!st) st = cur_lane().st;
  const Gen3Launch& g = c.gen3[la][lb][lc];
  const int nchunks = (xa.hi - xa.lo + 63) / 64;
  if (pa.npairs <= 0 || nchunks <= 0) return;
  if (fp32 && !br.recs32.p) throw Error{DFHIP_ERR_HIP, "FP32 bra records missing"};
  // x: the m-groups of a pair interleaved by XCD (8 (G (pidx / 8) + group) +
  // pidx % 8, the pairs padded to a multiple of 8; the kernel decodes it)
  const int64_t gx = (int64_t)(fp32 ? g.groups32 : g.groups) * ((pa.npairs + 7) / 8 * 8);
  if (gx > INT32_MAX) throw Error{DFHIP_ERR_ARG, "too many shell pairs in one class"};
  hipLaunchKernelGGL(fp32 ? g.kernel32 : g.kernel, dim3((unsigned)gx, nchunks), dim3(64), 0, st, br.recs.p,
                     br.recs32.p, br.off, br.npp, ket, pa, xa, c.boys.p, out, ldo);
  HC(hipGetLastError());
}

// Full metric (P|Q) (naux x naux) into M via the 2c2e form of the 3c2e
// kernels, in two parts: metric_prepare uploads the aux "pair" lists and
// builds their bra records on stream s (a side stream may run it while other
// work runs), metric_launch queues the kernels on the lane's stream once the
// caller has ordered that stream behind the preparation.
struct MetricPrep {
  DBuf<int> ds1, ds2, dns, dlsh, diota, dfL;
  DBuf<int64_t> drow, dlcol;
  int64_t nbf = 0;
  std::vector<BraRecords> brs;
  std::vector<int> lbeg, kmax;
  bool gen = false;
  MetricPrep() : brs(LMAX_AUX + 1) {}
};

void metric_prepare(Ctx& c, const HostShells& x, const DevShells& dev, MetricPrep& mp, hipStream_t s) {
  std::vector<std::vector<int>> byL(LMAX_AUX + 1);
  for (int i = 0; i < x.n; i++) byL[x.L[i]].push_back(i);
  std::vector<int> s1, s2, nsurv;
  std::vector<int64_t> row;
  std::vector<int> lsh;
  std::vector<int64_t> lcol;
  mp.lbeg.assign(LMAX_AUX + 2, 0);
  mp.kmax.assign(LMAX_AUX + 1, 1);
  for (int L = 0; L <= LMAX_AUX; L++) {
    mp.lbeg[L] = (int)s1.size();
    for (int i : byL[L]) {
      s1.push_back(i);
      s2.push_back(-1);
      row.push_back(x.f0[i]);
      nsurv.push_back(std::numeric_limits<int>::max());
      lsh.push_back(i);
      lcol.push_back(x.f0[i]);
      mp.kmax[L] = std::max(mp.kmax[L], x.nprim[i]);
    }
  }
  mp.lbeg[LMAX_AUX + 1] = (int)s1.size();
  // the window [i, n) of bra i in its own L's list, and the L of each function
  std::vector<int> iota(std::max<size_t>(s1.size(), 1)), fL(std::max(x.nbf, 1));
  std::iota(iota.begin(), iota.end(), 0);
  for (int i = 0; i < x.n; i++)
    for (int m = 0; m < 2 * x.L[i] + 1; m++) fL[x.f0[i] + m] = x.L[i];
  mp.nbf = x.nbf;
  upload_pinned(mp.diota, iota, s);
  upload_pinned(mp.dfL, fL, s);
  upload_pinned(mp.ds1, s1, s);
  upload_pinned(mp.ds2, s2, s);
  upload_pinned(mp.dns, nsurv, s);
  upload_pinned(mp.drow, row, s);
  upload_pinned(mp.dlsh, lsh, s);
  upload_pinned(mp.dlcol, lcol, s);
  mp.gen = use_generated();
  for (int LP = 0; LP <= LMAX_AUX; LP++) {
    const int np = mp.lbeg[LP + 1] - mp.lbeg[LP];
    if (np > 0 && mp.gen && gen3_has(LP, 0, 0))
      build_bra_records(c, LP, 0, x, dev, mp.ds1.p + mp.lbeg[LP], mp.ds2.p + mp.lbeg[LP], s1.data() + mp.lbeg[LP],
                        s2.data() + mp.lbeg[LP], np, mp.brs[LP], false, s);
  }
}

// One side of the symmetric metric: the blocks with bra L < ket L, and for
// equal L the ket shells from the bra's own on (the window [i, n) of the
// generated kernels; the other kernels take the whole list), then the other
// side by symmetry.
void metric_launch(Ctx& c, const DevShells& dev, const MetricPrep& mp, double* M, int64_t ldm,
                   hipStream_t st = nullptr) {
  if (!st) st = cur_lane().st;
  for (int LP = 0; LP <= LMAX_AUX; LP++) {
    const int np = mp.lbeg[LP + 1] - mp.lbeg[LP];
    if (np == 0) continue;
    PairArgs pa{mp.ds1.p + mp.lbeg[LP], mp.ds2.p + mp.lbeg[LP], mp.drow.p + mp.lbeg[LP], mp.dns.p + mp.lbeg[LP], np};
    for (int LQ = LP; LQ <= LMAX_AUX; LQ++) {
      const int nq = mp.lbeg[LQ + 1] - mp.lbeg[LQ];
      if (nq == 0) continue;
      AuxArgs xa{mp.dlsh.p + mp.lbeg[LQ], mp.dlcol.p + mp.lbeg[LQ], 0, nq, mp.kmax[LQ]};
      if (mp.gen && gen3_has(LP, 0, LQ)) {
        PairArgs pw = pa;
        if (LQ == LP) pw.nbeg = mp.diota.p;
        launch_gen3(c, LP, 0, LQ, mp.brs[LP], dev, pw, xa, M, ldm, false, st);
      } else {
        c.int3[LP][0][LQ].fn(dev, dev, pa, xa, c.sph.p, c.boys.p, M, ldm, st);
        HC(hipGetLastError());
      }
    }
  }
  symmetrize_metric(M, ldm, mp.nbf, mp.dfL.p, st);
  HC(hipGetLastError());
}

// Both parts on the lane's stream.  No wait: the buffers go back to the
// lane's pool, whose blocks are only reused by work queued behind these
// kernels.
void metric_dev(Ctx& c, const HostShells& x, const DevShells& dev, double* M, int64_t ldm) {
  MetricPrep mp;
  metric_prepare(c, x, dev, mp, cur_lane().st);
  metric_launch(c, dev, mp, M, ldm);
}

// Q_C = sqrt(||(C|C)||_F) of every aux shell from the diagonal blocks alone
// (common/integrals/intor.f90:97-129 evaluates int_2c2e(L, L, c, c) only):
// per L, the 2c2e kernels with bra shell i limited to its own list entry
// (window [i, i + 1)), writing into a compact buffer whose row is the shell's
// function and whose column is m.  O(naux) work, against O(naux^2) for the
// full metric.
void schwarz_aux_dev(Ctx& c, const HostShells& x, const DevShells& dev, double* qc, hipStream_t st) {
  constexpr int LDC = 2 * LMAX_AUX + 1;
  std::vector<std::vector<int>> byL(LMAX_AUX + 1);
  for (int i = 0; i < x.n; i++) byL[x.L[i]].push_back(i);
  std::vector<int> s1, s2, nbeg, nend, lbeg(LMAX_AUX + 2, 0), kmax(LMAX_AUX + 1, 1);
  std::vector<int64_t> row, col;
  for (int L = 0; L <= LMAX_AUX; L++) {
    lbeg[L] = (int)s1.size();
    for (size_t k = 0; k < byL[L].size(); k++) {
      const int i = byL[L][k];
      s1.push_back(i);
      s2.push_back(-1);
      row.push_back(x.f0[i]);
      col.push_back(0);
      nbeg.push_back((int)k);
      nend.push_back((int)k + 1);
      kmax[L] = std::max(kmax[L], x.nprim[i]);
    }
  }
  lbeg[LMAX_AUX + 1] = (int)s1.size();
  DBuf<int> ds1, ds2, dnb, dne;
  DBuf<int64_t> drow, dcol;
  upload_pinned(ds1, s1, st);
  upload_pinned(ds2, s2, st);
  upload_pinned(dnb, nbeg, st);
  upload_pinned(dne, nend, st);
  upload_pinned(drow, row, st);
  upload_pinned(dcol, col, st);
  DBuf<double> blk((size_t)std::max(x.nbf, 1) * LDC);
  const bool gen = use_generated();
  std::vector<BraRecords> brs(LMAX_AUX + 1);
  for (int L = 0; L <= LMAX_AUX; L++) {
    const int b0 = lbeg[L], np = lbeg[L + 1] - b0;
    if (np == 0) continue;
    PairArgs pa{ds1.p + b0, ds2.p + b0, drow.p + b0, dne.p + b0, np};
    pa.nbeg = dnb.p + b0;
    AuxArgs xa{ds1.p + b0, dcol.p + b0, 0, np, kmax[L]};
    if (gen && gen3_has(L, 0, L)) {
      build_bra_records(c, L, 0, x, dev, ds1.p + b0, ds2.p + b0, s1.data() + b0, s2.data() + b0, np, brs[L], false,
                        st);
      launch_gen3(c, L, 0, L, brs[L], dev, pa, xa, blk.p, LDC, false, st);
    } else {
      c.int3[L][0][L].fn(dev, dev, pa, xa, c.sph.p, c.boys.p, blk.p, LDC, st);
      HC(hipGetLastError());
    }
  }
  schwarz_aux_from_metric(blk.p, LDC, dev.L, dev.f0, x.n, qc, st, true);
  HC(hipGetLastError());
  // the pool blocks above go back before the caller's copy of qc has run
  HC(hipStreamSynchronize(st));
}

dfhip_options defaults() {
  dfhip_options o;
  std::memset(&o, 0, sizeof(o));
  o.precision = DFHIP_FP64;
  o.screening = DFHIP_SCREEN_SCHWARZ;
  o.thresh = (double)1e-10f;  // diabatization/intor.f90:140 (real(4) literal)
  o.dens_layout = DFHIP_DENS_FORTRAN;
  const char* v = std::getenv("DFHIP_VERBOSE");
  o.verbose = (v && *v) ? std::max(0, std::atoi(v)) : 0;
  o.max_tile_bytes = 0;
  o.fp32_below = 1e-2;
  return o;
}

// ---------------------------------------------------------------------------
// The contraction path

struct Plan {
  HostShells orb, aux;
  PairSet ps;              // sorted by Q_ab descending inside each class
  std::vector<double> qab, qc;
  std::vector<int64_t> prow;            // per pair, row of its first function
  std::vector<int> from;                // per pair, its index in the make_pairs order
  int64_t npf = 0;
  // aux lists per L, sorted by Q_C descending; permuted column of each list entry
  std::vector<std::vector<int>> alist;
  std::vector<std::vector<int64_t>> acol;
  std::vector<int> orig;                // permuted column -> original aux function
  std::vector<std::vector<int>> nsurv;  // [L][pair]
  std::vector<std::vector<int>> n64;    // [L][pair] FP64 part of the prefix (mixed precision)
  int64_t ncomputed = 0, ntriplets = 0, ncomputed32 = 0;
  int64_t class_trip[NPCLS * 6] = {}, class_prim[NPCLS * 6] = {};
};

// Sort key of a non-negative finite double for descending order: the IEEE
// bits of x >= 0 increase with x, so their complement decreases.
inline uint64_t desc_key(double x) {
  uint64_t b;
  std::memcpy(&b, &x, sizeof b);
  return ~b;
}

// idx (indices into q) reordered so that q[idx[]] is non-increasing; equal
// values keep their order (stable).  An LSD radix sort on the upper 32 bits
// of desc_key (8 bits per pass, the counts of all passes from one read,
// passes with a constant digit skipped), then runs with equal upper bits
// are ordered by the full key with a stable insertion sort.  The result is
// that of a stable sort by the full 64-bit key.
void sort_desc(const double* q, std::vector<int>& idx) {
  const size_t n = idx.size();
  if (n < 2) return;
  std::vector<uint32_t> key(n), k2(n);
  std::vector<int> v2(n);
  uint32_t cnt[4][256] = {};
  for (size_t i = 0; i < n; i++) {
    const uint32_t k = (uint32_t)(desc_key(q[idx[i]]) >> 32);
    key[i] = k;
    for (int d = 0; d < 4; d++) cnt[d][(k >> (8 * d)) & 255]++;
  }
  for (int d = 0; d < 4; d++) {
    uint32_t* c = cnt[d];
    const int shift = 8 * d;
    if (c[(key[0] >> shift) & 255] == n) continue;
    uint32_t sum = 0;
    for (int b = 0; b < 256; b++) {
      const uint32_t t = c[b];
      c[b] = sum;
      sum += t;
    }
    for (size_t i = 0; i < n; i++) {
      const uint32_t pos = c[(key[i] >> shift) & 255]++;
      k2[pos] = key[i];
      v2[pos] = idx[i];
    }
    key.swap(k2);
    idx.swap(v2);
  }
  for (size_t a = 0; a < n;) {
    size_t b = a + 1;
    while (b < n && key[b] == key[a]) b++;
    for (size_t i = a + 1; i < b; i++) {
      const int v = idx[i];
      const uint64_t kv = desc_key(q[v]);
      size_t j = i;
      while (j > a && desc_key(q[idx[j - 1]]) > kv) {
        idx[j] = idx[j - 1];
        j--;
      }
      idx[j] = v;
    }
    a = b;
  }
}

// Runs fn(t) for t = 0..n-1; the tasks with big[t] on separate host threads
// (each task is independent and writes only its own outputs), the rest on
// the calling thread.
template <class F>
void run_tasks(int n, const std::vector<char>& big, F fn) {
  std::vector<std::thread> th;
  for (int t = 0; t < n; t++)
    if (big[t]) th.emplace_back(fn, t);
  for (int t = 0; t < n; t++)
    if (!big[t]) fn(t);
  for (auto& t : th) t.join();
}

// Pair classes of at least 4096 pairs are sorted, and on such systems the aux
// L swept, on their own host threads.
bool big_system(const PairSet& ps, std::vector<char>* big_cls) {
  std::vector<char> b(NPCLS, 0);
  int nbig = 0;
  for (int k = 0; k < NPCLS; k++) nbig += b[k] = ps.cls_end[k] - ps.cls_begin[k] >= 4096;
  if (nbig <= 1) std::fill(b.begin(), b.end(), 0);
  if (big_cls) *big_cls = b;
  return nbig > 1;
}

// Pair part of the plan for class k (needs that class's Q_ab only): its
// pairs sorted by Q_ab descending, ties in make_pairs order, into pl.ps,
// pl.qab, pl.from and pl.prow at the class's range (these have their full
// sizes).  ps0 / q0: the pairs in make_pairs order and their Q_ab; row0: the
// class's first row (all pairs of a class have the same number of functions).
// compact: pairs with Q_ab = 0 (they keep no triplet) get no rows of the
// staging matrix; their prow is the class's row end.  Returns that end.
int64_t plan_pair_class(Plan& pl, const PairSet& ps0, const std::vector<double>& q0, int k, int64_t row0,
                        bool compact) {
  const HostShells& o = pl.orb;
  const int p0 = ps0.cls_begin[k], p1 = ps0.cls_end[k];
  if (p1 == p0) return row0;
  std::vector<int> idx(p1 - p0);
  for (int p = p0; p < p1; p++) idx[p - p0] = p;
  sort_desc(q0.data(), idx);
  const int64_t w = (int64_t)(2 * o.L[ps0.s1[p0]] + 1) * (2 * o.L[ps0.s2[p0]] + 1);
  for (int r = 0; r < p1 - p0; r++) {
    const int p = idx[r];
    pl.ps.s1[p0 + r] = ps0.s1[p];
    pl.ps.s2[p0 + r] = ps0.s2[p];
    pl.ps.packed[p0 + r] = ps0.packed[p];
    pl.from[p0 + r] = p;
    pl.qab[p0 + r] = q0[p];
  }
  int64_t row = row0;
  for (int r = 0; r < p1 - p0; r++) {
    pl.prow[p0 + r] = row;
    if (!compact || pl.qab[p0 + r] > 0.0) row += w;
  }
  if (compact)  // the zero-Q pairs sort last: they all start at the end
    for (int r = 0; r < p1 - p0; r++)
      if (!(pl.qab[p0 + r] > 0.0)) pl.prow[p0 + r] = row;
  return row;
}

// Aux lists of the plan (needs Q_C): per L sorted by Q_C descending, their
// columns in the staging matrix and the original function of each column.
void build_aux_lists(Plan& pl) {
  const HostShells& o = pl.orb;
  const HostShells& x = pl.aux;
  pl.alist.assign(LMAX_AUX + 1, {});
  pl.acol.assign(LMAX_AUX + 1, {});
  for (int i = 0; i < x.n; i++) pl.alist[x.L[i]].push_back(i);
  pl.orig.assign(x.nbf, 0);
  int64_t col = 0;
  for (int L = 0; L <= LMAX_AUX; L++) {
    auto& v = pl.alist[L];
    sort_desc(pl.qc.data(), v);
    for (int sh : v) {
      pl.acol[L].push_back(col);
      for (int m = 0; m < 2 * L + 1; m++) pl.orig[col + m] = x.f0[sh] + m;
      col += 2 * L + 1;
    }
  }
  pl.ntriplets = (int64_t)o.n * (o.n + 1) / 2 * x.n;
}

// Screened prefixes of the plan (needs the aux lists and the sorted pairs)
// and the triplet counts; under QVL they come from apply_qvl instead.
void screen_sweep(Plan& pl, const dfhip_options& opt) {
  const HostShells& o = pl.orb;
  const HostShells& x = pl.aux;
  const int np = pl.ps.n();
  std::vector<char> big_l(LMAX_AUX + 1, 0);
  if (big_system(pl.ps, nullptr)) {
    std::vector<int> nl(LMAX_AUX + 1, 0);
    for (int i = 0; i < x.n; i++) nl[x.L[i]]++;
    for (int L = 0; L <= LMAX_AUX; L++) big_l[L] = nl[L] > 0;
  }
  pl.nsurv.assign(LMAX_AUX + 1, std::vector<int>(np, 0));
  pl.ncomputed = 0;
  if (opt.screening == DFHIP_SCREEN_QVL) return;  // prefixes and counts from apply_qvl
  // Per aux L: the screened prefix of every pair (skip when Q_ab * Q_C <
  // thresh; inside a pair class Q_ab is non-increasing, so the prefix length
  // is too and one downward sweep per class finds all of them), then the
  // per-class counts of screened triplets and primitive triplets.
  int64_t ncomp[LMAX_AUX + 1] = {};
  std::vector<int> npp(np);
  for (int p = 0; p < np; p++) npp[p] = o.nprim[pl.ps.s1[p]] * o.nprim[pl.ps.s2[p]];
  run_tasks(LMAX_AUX + 1, big_l, [&](int L) {
    const auto& v = pl.alist[L];
    const int nv = (int)v.size();
    std::vector<double> qcs(nv);
    std::vector<int64_t> pref(nv + 1, 0);
    for (int i = 0; i < nv; i++) {
      qcs[i] = pl.qc[v[i]];
      pref[i + 1] = pref[i] + x.nprim[v[i]];
    }
    int* ns = pl.nsurv[L].data();
    for (int k = 0; k < NPCLS; k++) {
      int n = nv;
      int64_t trip = 0, prim = 0;
      for (int p = pl.ps.cls_begin[k]; p < pl.ps.cls_end[k]; p++) {
        if (opt.screening != DFHIP_SCREEN_NONE) {
          const double qa = pl.qab[p];
          // the product is monotone in Q_C, so the kept set is a prefix
          while (n > 0 && qa * qcs[n - 1] < opt.thresh) n--;
        }
        ns[p] = n;
        trip += n;
        prim += (int64_t)npp[p] * pref[n];
      }
      pl.class_trip[k * 6 + L] += trip;
      pl.class_prim[k * 6 + L] += prim;
      ncomp[L] += trip;
    }
  });
  for (int L = 0; L <= LMAX_AUX; L++) pl.ncomputed += ncomp[L];
}

// The FP32 windows of the generated kernels (taken only when the tuning
// table or DFHIP_GEN32=1 asks for them) stay below this Schwarz product
// whatever fp32_below says: at 1e-2 one molecule of the 256-molecule batch
// moves J by 4.4e-8 through them, at 1e-3 the worst is 6e-11.
constexpr double GEN32_TAU_MAX = 1e-3;

// Mixed precision: the FP64 part of each screened prefix, the list entries
// with Q_ab Q_C >= tau (a prefix again, non-increasing along a class), for the
// classes with FP32 kernels; the rest of the prefix runs in FP32.
void split_precision(Plan& pl, double tau, const std::vector<std::vector<char>>& lowp) {
  pl.n64 = pl.nsurv;
  for (int L = 0; L <= LMAX_AUX; L++) {
    const auto& v = pl.alist[L];
    for (int k = 0; k < NPCLS; k++) {
      if (!lowp[k][L]) continue;
      const double t = lowp[k][L] == 1 ? std::min(tau, GEN32_TAU_MAX) : tau;
      int n = (int)v.size();
      for (int p = pl.ps.cls_begin[k]; p < pl.ps.cls_end[k]; p++) {
        const double qa = pl.qab[p];
        while (n > 0 && qa * pl.qc[v[n - 1]] < t) n--;
        pl.n64[L][p] = std::min(n, pl.nsurv[L][p]);
      }
    }
  }
}

// Class k, aux L back to one FP64 window (its low-precision kernel cannot run).
void unsplit_precision(Plan& pl, int k, int L) {
  for (int p = pl.ps.cls_begin[k]; p < pl.ps.cls_end[k]; p++) pl.n64[L][p] = pl.nsurv[L][p];
}

// Verbose progress line, flushed at once so that a caller redirecting stdout
// sees each phase as it finishes.
void vlog(const dfhip_options& opt, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void vlog(const dfhip_options& opt, const char* fmt, ...) {
  if (!opt.verbose) return;
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  std::fflush(stdout);
}

// erfinv(y), 0 <= y < 1, by the rational Chebyshev approximations of Blair,
// Edwards and Johnson (Math. Comp. 30 (1976) 827-830), the method of
// common/erfinv.f90:88-127: x R1(x^2 - 0.5625) for y <= 0.75,
// x R2(x^2 - 0.87890625) for y <= 0.9375, and R3(t)/t with
// t = 1/sqrt(-log(1 - y)) above.  The coefficients are the paper's published
// tables; each R = P(z)/Q(z) is evaluated by Horner's rule.
double erfinv_pos(double y) {
  static const double p1[7] = {0.160304955844066229311e2, -0.90784959262960326650e2, 0.18644914861620987391e3,
                               -0.16900142734642382420e3, 0.6545466284794487048e2,  -0.864213011587247794e1,
                               0.176058782139590e0};
  static const double q1[7] = {0.147806470715138316110e2, -0.91374167024260313936e2, 0.21015790486205317714e3,
                               -0.22210254121855132366e3, 0.10760453916055123830e3,  -0.206010730328265443e2,
                               1.0};
  static const double p2[8] = {-0.152389263440726128e-1, 0.3444556924136125216e0,  -0.29344398672542478687e1,
                               0.11763505705217627302e2, -0.22655292823101104193e2, 0.19121334396580330163e2,
                               -0.5478927619598318769e1, 0.237516689024448000e0};
  static const double q2[8] = {-0.108465169602059954e-1, 0.2610628885843078511e0,  -0.24068318104393757995e1,
                               0.10695129973387014469e2, -0.23716715521596581025e2, 0.24640158943917284883e2,
                               -0.10014376349783070835e2, 1.0};
  static const double p3[9] = {0.3100808562552958465e-4,  0.40974876030119399241e-2, 0.1214902662897276161664e0,
                               0.11091676946390282026357e1, 0.32283798556639239484387e1, 0.2881691815651598826211e1,
                               0.2047972087262996049729e1, 0.854592208197214832806e0,  0.3551095884622383139e-2};
  static const double q3[9] = {0.3100809298564522487e-4,  0.40975286786639145041e-2, 0.1215907800748757143068e0,
                               0.11186271676316964210428e1, 0.34323639843052901805054e1, 0.4140284677116202150554e1,
                               0.4119797271272204121285e1, 0.2162961962641434560920e1, 1.0};
  auto ratio = [](const double* p, const double* q, int n, double z) {
    double a = p[n - 1], b = q[n - 1];
    for (int i = n - 2; i >= 0; i--) {
      a = a * z + p[i];
      b = b * z + q[i];
    }
    return a / b;
  };
  if (y <= 0.75) return y * ratio(p1, q1, 7, y * y - 0.5625);
  if (y <= 0.9375) return y * ratio(p2, q2, 8, y * y - 0.87890625);
  const double t = 1.0 / std::sqrt(-std::log(1.0 - y));
  return ratio(p3, q3, 9, t) / t;
}

// QVL screening (common/screening.f90:101-163 set-up, :267-299 estimate):
// aux records on the host, pair records and the screened-in prefixes of every
// (pair, aux L) on the device; fills pl.nsurv and the counts.  The returned
// buffers are the records the 3c2e kernels test each triplet against.
struct QvlRecords {
  DBuf<double> pair, aux;
  // per aux L and pair: the prefix end, the kept count and the kept
  // primitives ([L][3][npairs]); the prefixes feed the kernels' masks
  DBuf<int> counts;
  size_t per = 0;
  // per aux L: the kept entries of every pair's prefix (qvl_compact), for the
  // matrix-core kernels
  DBuf<int64_t> koff[LMAX_AUX + 1];
  DBuf<int> kidx[LMAX_AUX + 1];
};

// Aux records [extent, beta, Q_C, 0] per aux shell (common/screening.f90:107-115):
// auxiliary_extent(exps, 1d-8) (:168-196), beta = min exp^(-(2L+3)/4) sqrt((2L-1)!!).
std::vector<double> aux_qvl_records(const HostShells& x, const std::vector<double>& qc) {
  std::vector<double> xr((size_t)4 * x.n, 0.0);
  for (int sh = 0; sh < x.n; sh++) {
    const int L = x.L[sh];
    double ext = 0.0, emin = x.exps[x.pptr[sh]];
    const double pt = -std::log(1e-8);  // auxiliary_extent(exps, 1d-8), :168-196
    for (int i = 0; i < x.nprim[sh]; i++) {
      const double e = x.exps[x.pptr[sh] + i];
      const double arg = (pt + 0.5 * std::log(e)) / e;
      ext = std::max(ext, arg < 0 ? 0.01 : std::sqrt(arg));
      emin = std::min(emin, e);
    }
    double df = 1.0;  // (2L-1)!!
    for (int k = 2 * L - 1; k > 1; k -= 2) df *= k;
    xr[4 * sh + 0] = ext;
    xr[4 * sh + 1] = std::pow(emin, -(2.0 * L + 3.0) / 4.0) * std::sqrt(df);
    xr[4 * sh + 2] = qc[sh];
  }
  return xr;
}

// S_l of pair_bound.h: per degree the largest sum of |coefficients| of one
// real solid harmonic of the table.
const double* sph_abs_sums() {
  static const std::array<double, LMAX_SPH + 1> sl = [] {
    std::array<double, LMAX_SPH + 1> r{};
    std::unique_ptr<SphTable> t(new SphTable());
    build_sph_table(*t);
    for (int L = 0; L <= LMAX_SPH; L++)
      for (int m = 0; m < 2 * L + 1; m++) {
        double s = 0.0;
        for (int e = t->start[L][m]; e < t->start[L][m + 1]; e++) s += std::fabs(t->terms[e].c);
        r[L] = std::max(r[L], s);
      }
    return r;
  }();
  return sl.data();
}

DevShells host_view(const HostShells& s) {
  return DevShells{s.n, s.L.data(), s.nprim.data(), s.pptr.data(), s.f0.data(), s.xyz.data(), s.exps.data(),
                   s.coefs.data()};
}

// Upper bound of Q_C of aux shell sh (pair_bound.h, one centre).  A single
// normalized primitive has it in closed form: its components are orthonormal
// and, on one centre with one exponent, Coulomb-orthogonal with the same
// self-energy, which the Fourier transform of r^l Y_lm exp(-a r^2) gives as
// 2 pi / (a (l + 1/2)), so Q_C = (2l+1)^(1/4) sqrt(2 pi / (a (l + 1/2)))
// (1 % added for rounding).
double aux_schwarz_bound(const HostShells& x, int sh) {
  const int L = x.L[sh];
  if (x.nprim[sh] == 1)
    return 1.01 * std::pow(2.0 * L + 1.0, 0.25) * std::sqrt(2.0 * PI / (x.exps[x.pptr[sh]] * (L + 0.5)));
  double u = 0.0;
  for (int i = 0; i < x.nprim[sh]; i++) {
    const double b = x.exps[x.pptr[sh] + i], q = 0.5 * b;
    const double C = L > 0 ? std::pow(L / b, 0.5 * L) * std::exp(-0.5 * L) : 1.0;
    u += std::fabs(x.coefs[x.pptr[sh] + i]) * C * std::sqrt(2.0 * std::pow(PI, 2.5) / (q * q * std::sqrt(2.0 * q)));
  }
  return std::sqrt(2.0 * L + 1.0) * sph_abs_sums()[L] * u;
}

// Drop-test parameters of a screened contraction (pair_bound.h): the bound
// on max Q_C and, under QVL, the far-field factor from the aux extents and
// betas the screener uses.  DFHIP_NO_PAIR_SKIP=1 turns the test off.
PairSkip pair_skip_params(const HostShells& x, const dfhip_options& opt) {
  PairSkip sk;
  const char* v = std::getenv("DFHIP_NO_PAIR_SKIP");
  if (opt.screening == DFHIP_SCREEN_NONE || (v && *v && *v != '0') || !(opt.thresh > 0.0)) return sk;
  for (int l = 0; l <= LMAX_SPH; l++) sk.sl[l] = sph_abs_sums()[l];
  for (int sh = 0; sh < x.n; sh++) sk.qc = std::max(sk.qc, aux_schwarz_bound(x, sh));
  if (opt.screening == DFHIP_SCREEN_QVL) {
    const std::vector<double> xr = aux_qvl_records(x, std::vector<double>(x.n, 0.0));
    double g = 0.0;
    for (int sh = 0; sh < x.n; sh++) g = std::max(g, xr[4 * sh + 1] / std::pow(xr[4 * sh], x.L[sh] + 1));
    sk.qvl = PI * std::sqrt(2.0) * g;
  }
  if (!std::isfinite(sk.qc) || !std::isfinite(sk.qvl)) return PairSkip{};
  sk.thresh = opt.thresh;
  return sk;
}

// erfinv(1 - thresh_ws) of the pair extents, thresh_ws = 1d-4
// (t_df_screener(shells, shells_aux, 1d-4), diabatization/intor.f90:84)
double qvl_erfi() { return erfinv_pos(1.0 - 1e-4); }

void apply_qvl(Plan& pl, const DevShells& dorb, const DevShells& daux, const int* ps1, const int* ps2,
               const std::vector<DView<int>>& dalist, double thresh, QvlRecords& qr, hipStream_t st) {
  const HostShells& o = pl.orb;
  const int np = pl.ps.n();
  upload(qr.aux, aux_qvl_records(pl.aux, pl.qc), st);
  DBuf<double> dq;
  upload(dq, pl.qab, st);
  qr.pair.alloc((size_t)8 * std::max(np, 1));
  qvl_pair_records(dorb, ps1, ps2, dq.p, np, qvl_erfi(), qr.pair.p, st);
  HC(hipGetLastError());
  // the screens of all aux L first, one copy of their counts into pinned
  // memory and one wait; then per L the kept-list offsets go up through the
  // pinned arena and the compaction is queued, without further waits.  Host
  // loops cover the pairs with Q_ab > 0 (a leading run of each class; the
  // others keep nothing)
  const size_t per = (size_t)3 * std::max(np, 1);
  qr.per = per;
  qr.counts.alloc(per * (LMAX_AUX + 1));
  DBuf<int>& ns = qr.counts;
  pl.ncomputed = 0;
  if (np == 0) return;
  for (int L = 0; L <= LMAX_AUX; L++) {
    const int nl = (int)pl.alist[L].size();
    int* n3 = ns.p + per * L;
    if (nl == 0) {
      HC(hipMemsetAsync(n3, 0, per * sizeof(int), st));
      continue;
    }
    qvl_screen(qr.pair.p, np, daux, dalist[L].p, nl, L, qr.aux.p, thresh, n3, n3 + np, n3 + 2 * np, st);
    HC(hipGetLastError());
  }
  std::vector<int> live(NPCLS);
  for (int k = 0; k < NPCLS; k++)
    live[k] = thresh > 0.0 ? (int)(std::partition_point(pl.qab.begin() + pl.ps.cls_begin[k],
                                                        pl.qab.begin() + pl.ps.cls_end[k],
                                                        [](double q) { return q > 0.0; }) -
                                   pl.qab.begin())
                           : pl.ps.cls_end[k];
  int* hall = cur_lane().hqvl.get<int>(per * (LMAX_AUX + 1));
  HC(hipMemcpyAsync(hall, ns.p, per * (LMAX_AUX + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
  for (int L = 0; L <= LMAX_AUX; L++) {
    if (pl.alist[L].empty()) continue;
    const int* h = hall + per * L;
    int* n3 = ns.p + per * L;
    int* nsv = pl.nsurv[L].data();
    int64_t* ko = cur_lane().arena.alloc<int64_t>((size_t)np + 1);
    ko[0] = 0;
    for (int la = 0; la <= LMAX_ORB; la++)
      for (int lb = 0; lb <= la; lb++) {
        const int k = pcls(la, lb), p0 = pl.ps.cls_begin[k], p1 = pl.ps.cls_end[k];
        for (int p = p0; p < live[k]; p++) {
          nsv[p] = h[p];
          ko[p + 1] = ko[p] + h[np + p];
          pl.class_trip[k * 6 + L] += h[np + p];
          pl.class_prim[k * 6 + L] += (int64_t)o.nprim[pl.ps.s1[p]] * o.nprim[pl.ps.s2[p]] * h[2 * np + p];
        }
        for (int p = std::max(p0, live[k]); p < p1; p++) {
          nsv[p] = 0;
          ko[p + 1] = ko[p];
        }
      }
    pl.ncomputed += ko[np];
    qr.koff[L].alloc((size_t)np + 1);
    HC(hipMemcpyAsync(qr.koff[L].p, ko, ((size_t)np + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    qr.kidx[L].alloc((size_t)std::max<int64_t>(ko[np], 1));
    qvl_compact(qr.pair.p, np, daux, dalist[L].p, L, qr.aux.p, thresh, n3, qr.koff[L].p, qr.kidx[L].p, st);
    HC(hipGetLastError());
  }
}

struct Outputs {
  double* df = nullptr;         // host, column-major naux x ndens
  double* J = nullptr;          // host, column-major nstates^4
  int nstates = 0;
  double* dense3c = nullptr;    // host, (nbf, nbf, naux) row-major
  double* gram = nullptr;       // host, column-major ndens x ndens: B^T B
  double* blocknorm = nullptr;  // host, [a(a+1)/2 + b][aux shell c]: ||(ab|C)||_F of the computed blocks
};

// Frobenius norm of every (pair, aux shell) block of one column block of the
// staging matrix (rows of pair p from prow[p], the block's aux shells at
// columns bcol[q] - c0): norms[p nshx + bsh[q]].  One thread per block.
__global__ void block_norm_kernel(const double* __restrict__ T, int64_t ldt, int64_t c0,
                                  const int64_t* __restrict__ prow, const int* __restrict__ s1,
                                  const int* __restrict__ s2, const int* __restrict__ Lorb, int npairs,
                                  const int* __restrict__ bsh, const int64_t* __restrict__ bcol,
                                  const int* __restrict__ bL, int nb, int nshx, double* __restrict__ norms) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)npairs * nb) return;
  const int p = (int)(e / nb), q = (int)(e - (int64_t)p * nb);
  const int nab = (2 * Lorb[s1[p]] + 1) * (2 * Lorb[s2[p]] + 1), nc = 2 * bL[q] + 1;
  const double* t = T + prow[p] * ldt + (bcol[q] - c0);
  double s = 0.0;
  for (int r = 0; r < nab; r++)
    for (int m = 0; m < nc; m++) {
      const double v = t[(int64_t)r * ldt + m];
      s += v * v;
    }
  norms[(int64_t)p * nshx + bsh[q]] = sqrt(s);
}

// take_lock = false: the caller holds the library mutex and has set up the
// context (a batch worker on its own lane).
void run(const dfhip_basis& B, const dfhip_basis& X, int ndens, int nbfs, const double* dens,
         const dfhip_options& opt, Outputs out, dfhip_stats* stats, bool take_lock = true) {
  const auto entry_t = std::chrono::steady_clock::now();
  if (opt.precision != DFHIP_FP64 && opt.precision != DFHIP_MIXED)
    throw Error{DFHIP_ERR_UNSUPPORTED, "unknown precision"};
  const bool mixed = opt.precision == DFHIP_MIXED;
  if (mixed && !(opt.fp32_below >= 0.0)) throw Error{DFHIP_ERR_ARG, "fp32_below must be >= 0"};
  if (opt.screening != DFHIP_SCREEN_SCHWARZ && opt.screening != DFHIP_SCREEN_NONE &&
      opt.screening != DFHIP_SCREEN_QVL)
    throw Error{DFHIP_ERR_UNSUPPORTED, "unknown screening mode"};
  Plan pl;
  pl.orb = parse_shells(B, LMAX_ORB, "orbital basis");
  pl.aux = parse_shells(X, LMAX_AUX, "auxiliary basis");
  const HostShells& o = pl.orb;
  const HostShells& x = pl.aux;
  const bool need_dens = out.df || out.J || out.gram;
  if (out.blocknorm && opt.screening != DFHIP_SCREEN_NONE)
    throw Error{DFHIP_ERR_ARG, "block norms need the unscreened integrals"};
  if (need_dens) {
    if (ndens < 1) throw Error{DFHIP_ERR_ARG, "ndens must be >= 1"};
    if (nbfs != o.nbf)
      throw Error{DFHIP_ERR_ARG, "nbfs=" + std::to_string(nbfs) + " but the basis has " + std::to_string(o.nbf) +
                                     " spherical functions"};
    if (!dens) throw Error{DFHIP_ERR_ARG, "densities is NULL"};
  }
  if (out.J && out.nstates * (out.nstates + 1) / 2 != ndens)
    throw Error{DFHIP_ERR_ARG, "ndens is not nstates(nstates+1)/2"};
  const int64_t naux = x.nbf;

  std::unique_lock<std::mutex> lock(ctx_mutex(), std::defer_lock);
  if (take_lock) {
    lock.lock();
    ensure_ctx();
  }
  Ctx& c = ctx();
  hipStream_t st = cur_lane().st;
  Timer tm(st);
  const int e0 = tm.mark();
  // verbose 2: wall time of the host steps up to the integrals (each step
  // synchronized first, so the times add up); verbose 3: no waits, each step
  // logs when the host got there and when the stream did (printed at the end)
  const auto run_t0 = std::chrono::steady_clock::now();
  auto step_t0 = run_t0;
  std::vector<std::pair<const char*, std::pair<double, int>>> vmarks;
  auto vstep = [&](const char* what) {
    if (opt.verbose < 2) return;
    const auto t = std::chrono::steady_clock::now();
    if (opt.verbose >= 3) {
      vmarks.push_back({what, {std::chrono::duration<double, std::milli>(t - run_t0).count(), tm.mark()}});
      return;
    }
    HC(hipStreamSynchronize(st));
    const auto t2 = std::chrono::steady_clock::now();
    vlog(opt, "  %-28s %8.3f ms\n", what, std::chrono::duration<double, std::milli>(t2 - step_t0).count());
    step_t0 = t2;
  };
  DevShellBufs dorb, daux;
  dorb.upload_from(o, st);
  daux.upload_from(x, st);
  const int e1 = tm.mark();

  Lane& lane = cur_lane();
  lane.arena.reset();
  if (need_dens) lane.ensure_side();
  DBuf<int> info(1);
  // on every way out of run() (errors included) the lane's streams are
  // drained before Ap, info, the pinned buffers or the pool blocks can be
  // reused
  struct StreamDrain {
    Lane& l;
    bool side = false;
    ~StreamDrain() {
      (void)hipStreamSynchronize(l.st);
      if (side) (void)hipStreamSynchronize(l.st2);
      if (l.st3) (void)hipStreamSynchronize(l.st3);
    }
  } drain{lane, need_dens};

  // Schwarz norms of the shell pairs (common/integrals/intor.f90:131-160);
  // the metric kernels queue behind them
  vstep("shells");
  const PairClasses pcl = pair_classes(o);
  SchwarzPending spend;
  schwarz_bra_launch(c, o, dorb.view, pcl, spend, pair_skip_params(x, opt));
  const int e2 = tm.mark();
  vstep("Schwarz queued");
  pl.ps = make_pairs(o, pcl);
  vstep("pairs");

  // the metric's inputs go up on the side stream (the metric follows them
  // there) while the Schwarz kernels run; not on the copy stream, whose
  // shipments wait for the Schwarz classes
  MetricPrep mprep;
  const hipStream_t sm = need_dens ? lane.st2 : st;
  HC(hipStreamWaitEvent(sm, tm.ev[e1], 0));
  metric_prepare(c, x, daux.view, mprep, sm);
  HC(hipEventRecord(lane.ev_mprep, sm));
  vstep("metric inputs queued");


  // metric (common/integrals/intor.f90:40-95) and Q_C from its diagonal
  // blocks: first on the side stream, beside the Schwarz kernels (without
  // densities on the lane's stream).
  // A = [[M], [I], [D^T]]: naux columns, rows 0..naux-1 the metric, then the
  // identity border whose factorization gives L^-T, then d^T (ndens rows)
  const int64_t lda = need_dens ? 2 * naux + ndens : naux;
  double* Ap = reserve(cur_lane().wsA, (size_t)lda * naux);
  HC(hipMemsetAsync(Ap, 0, (size_t)lda * naux * sizeof(double), sm));
  HC(hipStreamWaitEvent(sm, lane.ev_mprep, 0));
  const int e2m = tm.mark(sm);  // the metric's kernels start here
  metric_launch(c, daux.view, mprep, Ap, lda, sm);
  DBuf<double> dqc(x.n);
  schwarz_aux_from_metric(Ap, lda, daux.L.p, daux.f0.p, x.n, dqc.p, sm);
  const int e3 = tm.mark(sm);
  double* hqc = lane.hqc.get<double>(x.n);
  HC(hipMemcpyAsync(hqc, dqc.p, x.n * sizeof(double), hipMemcpyDeviceToHost, sm));
  HC(hipEventRecord(lane.ev_met, sm));
  vstep("metric queued");

  // densities: upload and pair-pack (diabatization/intor.f90:155-176
  // symmetry) on the side stream behind the metric, concurrent with the
  // Schwarz kernels.  The
  // packed rows follow the make_pairs order; the contraction reads them
  // through rowperm once the pairs are sorted.
  double* Dp = nullptr;
  int h2d0 = -1, h2d1 = -1;
  std::vector<int64_t> prow0(pl.ps.n());
  int64_t npf0 = 0;
  for (int p = 0; p < pl.ps.n(); p++) {
    prow0[p] = npf0;
    npf0 += (int64_t)(2 * o.L[pl.ps.s1[p]] + 1) * (2 * o.L[pl.ps.s2[p]] + 1);
  }
  if (npf0 >= ((int64_t)1 << 31)) throw Error{DFHIP_ERR_UNSUPPORTED, "more than 2^31 pair functions"};
  // device arrays of the side stream's packing: they live until run() ends,
  // so that no work on st gets their blocks while the packing reads them
  DBuf<int> rowpair0, dpairof, dfshell;
  if (need_dens) {
    const hipStream_t sp = lane.st2;
    HC(hipStreamWaitEvent(sp, lane.ev_pairs, 0));  // the shell arrays and the pair list are on the device
    const double* Dsrc = dens;
    if (!opt.dens_on_device) {
      const size_t nd = (size_t)ndens * nbfs * nbfs;
      double* Draw = reserve(cur_lane().wsDraw, nd);
      h2d0 = tm.mark(sp);
      HC(hipMemcpyAsync(Draw, dens, nd * sizeof(double), hipMemcpyHostToDevice, sp));
      h2d1 = tm.mark(sp);
      Dsrc = Draw;
    }
    int64_t sI, si, sj;
    if (opt.dens_layout == DFHIP_DENS_C) {
      sI = (int64_t)nbfs * nbfs;
      si = nbfs;
      sj = 1;
    } else {
      sI = 1;
      si = ndens;
      sj = (int64_t)ndens * nbfs;
    }
    Dp = reserve(cur_lane().wsDp, (size_t)npf0 * ndens);
    if (sj == 1 && si == nbfs && ndens > 48) {
      // density index slowest and many densities: tiled transpose (pair lookup
      // by shell pair, shell by function).  With few densities the packing
      // by rows below is faster.
      std::vector<int> pairof((size_t)o.n * o.n), fshell(o.nbf);
      for (int p = 0; p < pl.ps.n(); p++) {
        const int a = pl.ps.s1[p], b = pl.ps.s2[p];
        pairof[(size_t)a * o.n + b] = p | (1 << 30);
        if (a != b) pairof[(size_t)b * o.n + a] = p;
      }
      for (int sh = 0; sh < o.n; sh++)
        for (int m = 0; m < 2 * o.L[sh] + 1; m++) fshell[o.f0[sh] + m] = sh;
      upload(dpairof, pairof, sp);
      upload(dfshell, fshell, sp);
      pack_density_t(dpairof.p, dfshell.p, spend.prow0.p, dorb.L.p, dorb.f0.p, o.n, o.nbf, Dsrc, sI, ndens, Dp, ndens,
                     sp);
    } else {
      if (sj == 1 && si == nbfs && (size_t)ndens * nbfs * nbfs * sizeof(double) <= ((size_t)1 << 32)) {
        // density index slowest, few densities: turn D into (i, j, I) first,
        // so that the packing reads runs of densities (the direct reads along
        // I of the C layout each take their own cache line, about ten times
        // the bytes of D; the turn moves D twice)
        const int64_t nn = (int64_t)nbfs * nbfs;
        double* DT = reserve(cur_lane().wsDT, (size_t)ndens * nn);
        transpose(Dsrc, ndens, nn, DT, sp);
        Dsrc = DT;
        sI = 1;
        si = (int64_t)nbfs * ndens;
        sj = ndens;
      }
      rowpair0.alloc((size_t)std::max<int64_t>(npf0, 1));
      fill_rowpair(spend.prow0.p, pl.ps.n(), npf0, rowpair0.p, nullptr, nullptr, sp);
      pack_density(rowpair0.p, spend.s1.p, spend.s2.p, spend.prow0.p, npf0, dorb.L.p, dorb.f0.p, Dsrc, sI, si, sj, ndens, Dp,
                   ndens, sp);
    }
    HC(hipGetLastError());
    HC(hipEventRecord(lane.ev_pack, sp));
  }
  vstep("packing queued");

  // Cholesky of the metric bordered by the identity (common/linalg.f90:78 and
  // the dtrtri of diabatization/intor.f90:198) behind the metric and the
  // packing on the side stream, concurrent with the Schwarz kernels and the
  // integrals: it touches only rows < 2 naux, the contraction writes d^T
  // below them (after waiting for ev_met: the memset before the metric
  // clears those rows too).  Ap and info outlive the join.
  if (need_dens) {
    // not before the Schwarz kernels are done: its panel workgroups hold LDS
    // that the Schwarz workgroups need, the 3c2e kernels mostly do not
    HC(hipStreamWaitEvent(sm, tm.ev[e2], 0));
    set_identity_rows(Ap, lda, naux, sm);
    HC(hipMemsetAsync(info.p, 0, sizeof(int), sm));
    cholesky_inverse_rows(Ap, lda, naux, info.p, sm);
    HC(hipGetLastError());
    HC(hipEventRecord(lane.ev_join, sm));
  }

  // pair part of the plan class by class as the Schwarz sums arrive (the
  // sorts overlap the Schwarz kernels of the later classes and the metric);
  // the sorted pair arrays and the bra-record layouts go up in one copy
  const PairSet ps0 = pl.ps;
  const int npairs = ps0.n();
  std::vector<double> q0(npairs);
  pl.qab.assign(npairs, 0.0);
  pl.from.assign(npairs, 0);
  pl.prow.assign(npairs, 0);
  // the staging matrix holds rows of the pairs with Q_ab > 0 only (the others
  // keep no triplet under a positive threshold), so its column blocks are
  // wider; the dense and block-norm dumps read every pair's rows
  const bool compact_rows = opt.screening != DFHIP_SCREEN_NONE && opt.thresh > 0.0 && !out.dense3c &&
                            !out.blocknorm;
  int64_t rows_next = 0;
  const bool gen = use_generated();
  UploadBatch pb;
  const size_t o_s1 = pb.add<int>(npairs), o_s2 = pb.add<int>(npairs), o_prow = pb.add<int64_t>(npairs);
  const size_t o_qab = pb.add<double>(npairs), o_dprow = pb.add<int64_t>(need_dens ? npairs : 0);
  std::vector<size_t> o_boff(NPCLS, 0), o_bnpp(NPCLS, 0);
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++) {
      const int k = pcls(la, lb), n = ps0.cls_end[k] - ps0.cls_begin[k];
      if (n > 0 && gen && gen3_has(la, lb, 0)) {
        o_boff[k] = pb.add<int64_t>(n);
        o_bnpp[k] = pb.add<int>(n);
      }
    }
  pb.allocate();
  DView<int> ps1{pb.dev<int>(o_s1)}, ps2{pb.dev<int>(o_s2)};
  DView<int64_t> prow{pb.dev<int64_t>(o_prow)}, ddprow{pb.dev<int64_t>(o_dprow)};
  DView<double> dqab{pb.dev<double>(o_qab)};
  std::vector<BraRecords> brs(NPCLS);
  std::vector<int64_t> brtot(NPCLS, 0);
  double ms_plan = 0.0;
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++) {
      const int k = pcls(la, lb), p0 = ps0.cls_begin[k], n = ps0.cls_end[k] - p0;
      schwarz_class_finish(ps0, spend, k, q0.data());
      if (n == 0) continue;
      static const char* const wname[] = {"class 0 sums", "class 1 sums", "class 2 sums", "class 3 sums",
                                          "class 4 sums", "class 5 sums", "class 6+ sums"};
      static const char* const sname[] = {"class 0 sorted", "class 1 sorted", "class 2 sorted", "class 3 sorted",
                                          "class 4 sorted", "class 5 sorted", "class 6+ sorted"};
      if (opt.verbose >= 3) vstep(wname[std::min(k, 6)]);
      const auto hp0 = std::chrono::steady_clock::now();
      rows_next = plan_pair_class(pl, ps0, q0, k, rows_next, compact_rows);
      ms_plan += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - hp0).count();
      if (opt.verbose >= 3) vstep(sname[std::min(k, 6)]);
      std::memcpy(pb.host<int>(o_s1) + p0, pl.ps.s1.data() + p0, n * sizeof(int));
      std::memcpy(pb.host<int>(o_s2) + p0, pl.ps.s2.data() + p0, n * sizeof(int));
      std::memcpy(pb.host<int64_t>(o_prow) + p0, pl.prow.data() + p0, n * sizeof(int64_t));
      std::memcpy(pb.host<double>(o_qab) + p0, pl.qab.data() + p0, n * sizeof(double));
      if (need_dens) {
        // packed density row of each sorted pair's first function
        int64_t* dprow = pb.host<int64_t>(o_dprow);
        for (int r = p0; r < p0 + n; r++) dprow[r] = prow0[pl.from[r]];
      }
      if (gen && gen3_has(la, lb, 0)) {
        brtot[k] = bra_record_layout(c.gen3[la][lb][0].rs, o, pl.ps.s1.data() + p0, pl.ps.s2.data() + p0, n,
                                     pb.host<int64_t>(o_boff[k]), pb.host<int>(o_bnpp[k]));
        brs[k].off = pb.dev<int64_t>(o_boff[k]);
        brs[k].npp = pb.dev<int>(o_bnpp[k]);
      }
    }
  // the upload, the bra records and the row maps run on the copy stream,
  // beside the metric; the lane's stream waits for them before its first use
  // of the sorted pairs
  const hipStream_t s3 = copy_stream(lane);
  pb.commit(s3);
  // FP32 copies of the records only for classes that may get an FP32 window
  // of the generated kernels (the tuning table's choice or DFHIP_GEN32=1)
  const char* g32env = std::getenv("DFHIP_GEN32");
  const bool gen32_all = g32env && *g32env && *g32env != '0';
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++) {
      const int k = pcls(la, lb), p0 = ps0.cls_begin[k], n = ps0.cls_end[k] - p0;
      if (n > 0 && gen && gen3_has(la, lb, 0)) {
        bool rec32 = false;
        for (int L = 0; L <= LMAX_AUX && mixed; L++)
          rec32 = rec32 || (gen3_has(la, lb, L) && (gen32_all || tuned_family(k, L, 1) == 1));
        fill_bra_records(c, la, lb, dorb.view, ps1.p + p0, ps2.p + p0, n, brtot[k], brs[k], rec32, s3);
      }
    }
  pl.npf = rows_next;
  // per staging-matrix row (sorted pairs): its pair, for the contraction's
  // mask, and the packed density row it reads
  DBuf<int> drowpair, drowperm;
  if (need_dens) {
    drowpair.alloc((size_t)std::max<int64_t>(pl.npf, 1));
    drowperm.alloc((size_t)std::max<int64_t>(pl.npf, 1));
    fill_rowpair(prow.p, npairs, pl.npf, drowpair.p, ddprow.p, drowperm.p, s3);
  }
  HC(hipGetLastError());
  vstep("bra records queued");

  // Et tables of the classes run on the matrix-core 3c2e kernels, built once
  // for all column blocks after the aux plan, for the pairs with screened-in
  // triplets in any such aux L (a class whose tables would take more than a
  // sixth of the free HBM, the staging matrix this lane holds counted as
  // free, keeps the LDS kernels).  need(k, L, p): pair p (global index) of
  // class k is to be tabled for aux L; defer: leave a class whose table is
  // too large for a later call instead.
  std::vector<DBuf<double>> etg(NPCLS);
  std::vector<DBuf<int64_t>> etoff(NPCLS);
  std::vector<DBuf<int>> etlist(NPCLS);
  std::vector<std::vector<char>> use_mm(NPCLS, std::vector<char>(LMAX_AUX + 1, 0));
  size_t et_cap = 0;
  std::vector<char> haveL(LMAX_AUX + 1, 0);  // aux L present in the basis
  for (int i = 0; i < x.n; i++) haveL[x.L[i]] = 1;
  if (use_matrix_cores_3c()) {
    for (int la = 0; la <= LMAX_ORB; la++)
      for (int lb = 0; lb <= la; lb++) {
        const int k = pcls(la, lb);
        if (pl.ps.cls_end[k] == pl.ps.cls_begin[k] || c.et[la][lb].fn == nullptr) continue;
        for (int L = 0; L <= LMAX_AUX; L++) {
          // where both kernels exist: the tuned family, else the measured rule
          const int f = tuned_family(k, L, 0);
          const bool mm_first = f ? f == 2 : prefer_mm(la, lb, L);
          use_mm[k][L] = haveL[L] && (!(gen && gen3_has(la, lb, L)) || mm_first) && c.int3m[la][lb][L].fn != nullptr;
        }
      }
  }
  auto build_et = [&](auto need, hipStream_t es, bool defer) {
    if (et_cap == 0) {
      size_t fr = 0, tot = 0;
      HC(hipMemGetInfo(&fr, &tot));
      et_cap = std::max<size_t>((fr + cur_lane().wsT.n * sizeof(double)) / 6 / t_lanes, 1);
    }
    for (int la = 0; la <= LMAX_ORB; la++)
      for (int lb = 0; lb <= la; lb++) {
        const int k = pcls(la, lb), p0 = pl.ps.cls_begin[k], np = pl.ps.cls_end[k] - p0;
        if (np == 0 || etg[k].p) continue;
        bool any = false;
        for (int L = 0; L <= LMAX_AUX; L++) any = any || use_mm[k][L];
        if (!any) continue;
        std::vector<int> list;
        std::vector<int64_t> off(np, 0);
        int64_t totd = 0;
        for (int p = 0; p < np; p++) {
          bool nd = false;
          for (int L = 0; L <= LMAX_AUX && !nd; L++) nd = use_mm[k][L] && need(k, L, p0 + p);
          if (!nd) continue;
          list.push_back(p);
          off[p] = totd;
          totd += (int64_t)o.nprim[pl.ps.s1[p0 + p]] * o.nprim[pl.ps.s2[p0 + p]] * c.et[la][lb].etpp;
        }
        if ((size_t)totd * sizeof(double) > et_cap) {
          if (!defer)
            for (int L = 0; L <= LMAX_AUX; L++) use_mm[k][L] = 0;
          continue;
        }
        if (list.empty()) continue;
        upload_pinned(etlist[k], list, es);
        upload_pinned(etoff[k], off, es);
        etg[k].alloc((size_t)totd);
        c.et[la][lb].fn(dorb.view, ps1.p + p0, ps2.p + p0, etlist[k].p, etoff[k].p, (int)list.size(), c.sph.p,
                        etg[k].p, es);
        HC(hipGetLastError());
      }
  };

  HC(hipEventRecord(lane.ev_plan, s3));
  HC(hipStreamWaitEvent(st, lane.ev_plan, 0));
  vstep("early Et queued");
  // Et tables of all pairs of the classes that run on the matrix cores for
  // an aux L the generated kernels cover too ((dd|d), (dd|f): most of their
  // pairs are needed), queued while the host waits for Q_C anyway; the rest
  // after the aux plan, for the pairs with screened-in triplets only (for the
  // f and g classes, or the g and h aux shells, tables of all pairs cost
  // several times the needed ones)
  if (opt.screening != DFHIP_SCREEN_QVL) {
    std::vector<char> early(NPCLS, 0);
    for (int la = 0; la <= LMAX_ORB; la++)
      for (int lb = 0; lb <= la; lb++)
        for (int L = 0; L <= LMAX_AUX; L++)
          if (haveL[L] && gen && gen3_has(la, lb, L) && use_mm[pcls(la, lb)][L]) early[pcls(la, lb)] = 1;
    build_et([&](int k, int, int) { return early[k] != 0; }, st, true);
  }
  vstep("pair plan, bra records");

  // aux part of the plan once Q_C is back (the packing may still run)
  HC(hipEventSynchronize(lane.ev_met));
  pl.qc.assign(hqc, hqc + x.n);
  for (const double q : pl.qc)
    if (!std::isfinite(q)) throw Error{DFHIP_ERR_HIP, "non-finite auxiliary Schwarz norm"};
  vstep("Q_C on the host");

  // the rest of the Et tables: the pairs with screened-in triplets (under
  // Schwarz screening, or none, nsurv[L][p] > 0 exactly when Q_ab times the
  // largest Q_C of aux L passes, the test of the plan's sweep)
  if (opt.screening != DFHIP_SCREEN_QVL) {
    std::vector<double> qcmax(LMAX_AUX + 1, -1.0);
    for (int i = 0; i < x.n; i++) qcmax[x.L[i]] = std::max(qcmax[x.L[i]], pl.qc[i]);
    const bool all = opt.screening == DFHIP_SCREEN_NONE;
    build_et(
        [&](int k, int L, int p) { return qcmax[L] >= 0.0 && (all || !(pl.qab[p] * qcmax[L] < opt.thresh)); }, st,
        false);
  }
  vstep("Et tables queued");
  if (opt.verbose) {
    vlog(opt, "Schwarz norms of %d shell pairs done (%.1f ms).\n", pl.ps.n(), tm.ms(e1, e2));
    vlog(opt, "Metric (P|Q) of %d aux functions done (%.1f ms).\n", x.nbf, tm.ms(e2m, e3));
  }
  const auto hp1 = std::chrono::steady_clock::now();
  build_aux_lists(pl);
  // mixed precision: the low-precision kernel of each (class, aux L), 1 the
  // generated FP32 kernels, 2 the FP32 matrix-core kernels (0: FP64 only);
  // the split itself is final once the tables of the FP32 matrix-core
  // kernels are built, below.  A class of the generated kernels keeps one
  // FP64 window unless the tuning table (or DFHIP_GEN32=1) picks its FP32
  // kernel: those run at the FP64 rate (FP32 VALU without packed math) and
  // the split only adds a launch (60-atom s-d workload: 8.0 ms mixed with
  // the FP32 windows, 6.9 ms FP64)
  std::vector<std::vector<char>> lowp(NPCLS, std::vector<char>(LMAX_AUX + 1, 0));
  if (mixed) {
    for (int la = 0; la <= LMAX_ORB; la++)
      for (int lb = 0; lb <= la; lb++)
        for (int L = 0; L <= LMAX_AUX; L++) {
          const int k = pcls(la, lb);
          const bool mm32 = use_mm[k][L] && c.int3m[la][lb][L].fn32 != nullptr;
          const int tf = tuned_family(k, L, 1);
          if (gen && gen3_has(la, lb, L))
            lowp[k][L] = (mm32 && tf == 2) ? 2 : (tf == 1 || gen32_all) ? 1 : 0;
          else if (mm32) lowp[k][L] = 2;
        }
  }
  bool any_low = false;
  for (const auto& v : lowp)
    for (char w : v) any_low = any_low || w != 0;
  // Schwarz screening (or none) without FP32 windows: the prefixes come from
  // the device, and the host's sweep (the counts, and the prefixes that the
  // contraction's row segments use) runs once the integrals are queued
  const bool dev_prefix =
      !any_low && (opt.screening == DFHIP_SCREEN_SCHWARZ || opt.screening == DFHIP_SCREEN_NONE);
  if (!dev_prefix) screen_sweep(pl, opt);
  ms_plan += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - hp1).count();
  vlog(opt, "Main basis: %6d shells, %6d basis functions\n", o.n, o.nbf);
  vlog(opt, "Aux. basis: %6d shells, %6d basis functions\n", x.n, x.nbf);
  vlog(opt, " Densities: %6d\n", ndens);
  vstep("aux plan");

  // device plan arrays of the aux side in one copy: the lists and their
  // columns per L, the Q_C of the lists (for the device prefixes), the
  // original aux function of each column and, for the contraction's mask,
  // the L and list index of each column
  UploadBatch ab;
  std::vector<size_t> o_al(LMAX_AUX + 1), o_ac(LMAX_AUX + 1);
  size_t qtot = 0;
  for (int L = 0; L <= LMAX_AUX; L++) {
    o_al[L] = ab.add<int>(pl.alist[L].size());
    o_ac[L] = ab.add<int64_t>(pl.acol[L].size());
    qtot += pl.alist[L].size();
  }
  const size_t o_qcs = ab.add<double>(qtot), o_lists = ab.add<int>(2 * (LMAX_AUX + 1));
  const size_t o_orig = ab.add<int>(pl.orig.size()), o_colL = ab.add<int>(naux), o_colia = ab.add<int>(naux);
  ab.allocate();
  std::vector<DView<int>> dalist(LMAX_AUX + 1);
  std::vector<DView<int64_t>> dacol(LMAX_AUX + 1);
  std::vector<int> kmax(LMAX_AUX + 1, 1);
  {
    int* hl = ab.host<int>(o_lists);
    double* hq = ab.host<double>(o_qcs);
    int* colL = ab.host<int>(o_colL);
    int* colia = ab.host<int>(o_colia);
    size_t q = 0;
    for (int L = 0; L <= LMAX_AUX; L++) {
      const size_t nl = pl.alist[L].size();
      std::copy(pl.alist[L].begin(), pl.alist[L].end(), ab.host<int>(o_al[L]));
      std::copy(pl.acol[L].begin(), pl.acol[L].end(), ab.host<int64_t>(o_ac[L]));
      hl[L] = (int)q;
      hl[LMAX_AUX + 1 + L] = (int)nl;
      for (size_t i = 0; i < nl; i++) {
        hq[q + i] = pl.qc[pl.alist[L][i]];
        for (int m = 0; m < 2 * L + 1; m++) {
          colL[pl.acol[L][i] + m] = L;
          colia[pl.acol[L][i] + m] = (int)i;
        }
      }
      q += nl;
      dalist[L].p = ab.dev<int>(o_al[L]);
      dacol[L].p = ab.dev<int64_t>(o_ac[L]);
      for (int sh : pl.alist[L]) kmax[L] = std::max(kmax[L], x.nprim[sh]);
    }
    std::copy(pl.orig.begin(), pl.orig.end(), ab.host<int>(o_orig));
  }
  ab.commit(st);
  QvlRecords qvl;
  const bool use_qvl = opt.screening == DFHIP_SCREEN_QVL;
  if (use_qvl) {
    apply_qvl(pl, dorb.view, daux.view, ps1.p, ps2.p, dalist, opt.thresh, qvl, st);
    if (opt.verbose >= 1) {
      int64_t npre = 0;
      for (int L = 0; L <= LMAX_AUX; L++)
        for (int p = 0; p < pl.ps.n(); p++) npre += pl.nsurv[L][p];
      vlog(opt, "QVL: %lld triplets kept, %lld in the screened prefixes.\n", (long long)pl.ncomputed,
           (long long)npre);
    }
  }
  // screened prefixes, one flat array [L][pair] for the kernels and the
  // contraction's mask
  DBuf<int> dnsflat;
  {
    const size_t nf = (size_t)(LMAX_AUX + 1) * std::max(pl.ps.n(), 1);
    dnsflat.alloc(nf);
    if (dev_prefix) {
      const int* dl = ab.dev<int>(o_lists);
      schwarz_prefixes(dqab.p, pl.ps.n(), ab.dev<double>(o_qcs), dl, dl + LMAX_AUX + 1, LMAX_AUX + 1, opt.thresh,
                       opt.screening == DFHIP_SCREEN_NONE, dnsflat.p, st);
      HC(hipGetLastError());
    } else if (use_qvl && pl.ps.n() > 0) {
      // the prefix ends of the QVL screen, already on the device
      for (int L = 0; L <= LMAX_AUX; L++)
        HC(hipMemcpyAsync(dnsflat.p + (size_t)L * pl.ps.n(), qvl.counts.p + qvl.per * L,
                          (size_t)pl.ps.n() * sizeof(int), hipMemcpyDeviceToDevice, st));
    } else {
      int* nsflat = lane.arena.alloc<int>(nf);
      for (int L = 0; L <= LMAX_AUX; L++)
        std::copy(pl.nsurv[L].begin(), pl.nsurv[L].end(), nsflat + (size_t)L * pl.ps.n());
      HC(hipMemcpyAsync(dnsflat.p, nsflat, nf * sizeof(int), hipMemcpyHostToDevice, st));
    }
  }
  auto dns = [&](int L) { return dnsflat.p + (size_t)L * pl.ps.n(); };
  DBuf<int> dn64flat;  // FP64 windows of the prefixes, [L][pair] like dnsflat
  DBuf<unsigned long long> dn32count;  // QVL: kept triplets of the FP32 windows
  if (any_low) split_precision(pl, opt.fp32_below, lowp);
  const DView<int> dorig{ab.dev<int>(o_orig)};
  vstep("plan uploads");
  const int e5 = tm.mark();

  // column blocks of the (pair function x aux function) staging matrix
  size_t cap = (size_t)opt.max_tile_bytes;
  if (cap == 0) {
    size_t fr = 0, tot = 0;
    HC(hipMemGetInfo(&fr, &tot));
    // the staging matrix this lane already holds counts as free: a repeated
    // call gets the same column blocks as the first one
    fr += cur_lane().wsT.n * sizeof(double);
    // 60 % of it (the Et tables take at most a sixth more): fewer column
    // blocks mean wider aux windows per class launch (200 atoms: 6 blocks
    // instead of 7 at half of it, the step 2.53 -> 2.49 s)
    cap = std::max<size_t>((size_t)(0.6 * fr / t_lanes), (size_t)1 << 28);
  }
  const int64_t rowbytes = std::max<int64_t>(pl.npf, 1) * (int64_t)sizeof(double);
  int64_t maxcols = std::max<int64_t>(1, (int64_t)(cap / rowbytes));
  // a whole number of the contraction's column tiles per block: the GEMM runs
  // every tile of a block in full, so a block of 287 columns (200 atoms)
  // paid for 3 tiles of 128 where 2.24 hold columns
  if (maxcols >= contract_tile()) maxcols = maxcols / contract_tile() * contract_tile();
  // block boundaries at aux-shell granularity over the permuted column order
  std::vector<int64_t> bounds{0};
  {
    int64_t cur = 0;
    for (int L = 0; L <= LMAX_AUX; L++)
      for (size_t k = 0; k < pl.alist[L].size(); k++) {
        const int64_t w = 2 * L + 1;
        if (cur + w - bounds.back() > maxcols && cur > bounds.back()) bounds.push_back(cur);
        cur += w;
      }
    if (bounds.back() != cur) bounds.push_back(cur);
  }
  const int nblocks = (int)bounds.size() - 1;
  int64_t widest = 0;
  for (int b = 0; b < nblocks; b++) widest = std::max(widest, bounds[b + 1] - bounds[b]);
  double* Tp = reserve(cur_lane().wsT, (size_t)std::max<int64_t>(pl.npf, 1) * std::max<int64_t>(widest, 1));
  double* partp = nullptr;
  int nsplit = 1;
  // screened-prefix mask of T (the 3c2e kernels write only screened-in blocks)
  ScreenMask mk{};
  if (need_dens) mk = ScreenMask{drowpair.p, drowperm.p, ab.dev<int>(o_colL), ab.dev<int>(o_colia), dnsflat.p, pl.ps.n()};
  vstep("column blocks, mask");
  // density-tile width and split-K depth of the contraction: the probe's, the
  // tuned ones for this density count, or the defaults
  int cipt = 0;
  double cscale = 1.0;
  if (t_probe && t_probe->ipt) {
    cipt = t_probe->ipt;
    if (t_probe->nscale > 0.0) cscale = t_probe->nscale;
  } else {
    std::lock_guard<std::mutex> g(tuning_mutex());
    if (const Tuning::Contract* tc = tuning().contract_for(ndens)) {
      cipt = tc->ipt;
      cscale = tc->nscale;
    }
  }
  if (need_dens) {
    const int64_t tiles = contract_ntiles(widest, ndens, cipt);
    nsplit = (int)std::max<int64_t>(
        1, std::min<int64_t>(512, (int64_t)(cscale * 2048.0 / (double)std::max<int64_t>(tiles, 1))));
    nsplit = (int)std::min<int64_t>(nsplit, std::max<int64_t>(1, (pl.npf + 15) / 16));
    partp = reserve(cur_lane().wsPart, (size_t)nsplit * widest * ndens);
  }
  // Non-zero row segments of T per column tile: for each pair class, the
  // rows from the class start to the last pair whose screened prefix reaches
  // the tile (under Schwarz screening these pairs are exactly a prefix of
  // the class: pairs are sorted by Q_ab and the prefixes are non-increasing).
  if (opt.screening == DFHIP_SCREEN_QVL)
    build_et([&](int k, int L, int p) { return (lowp[k][L] ? pl.n64[L][p] : pl.nsurv[L][p]) > 0; }, st, false);
  // FP32 Et tables of the FP32 matrix-core windows [n64, nsurv); a class whose
  // table does not fit (or that lost its FP64 table) stays FP64
  std::vector<DBuf<float>> etg32(NPCLS);
  std::vector<DBuf<int64_t>> etoff32(NPCLS);
  std::vector<DBuf<int>> etlist32(NPCLS);
  if (any_low) {
    size_t fr = 0, tot = 0;
    HC(hipMemGetInfo(&fr, &tot));
    size_t cap32 = std::max<size_t>(fr / 6 / t_lanes, 1);
    // DFHIP_ET32_CAP (bytes) replaces the budget of one class's FP32 table:
    // a small value sends the FP32 windows back to FP64 (for tests)
    if (const char* e = std::getenv("DFHIP_ET32_CAP"))
      if (*e) cap32 = (size_t)std::strtoull(e, nullptr, 10);
    bool refit64 = false;
    for (int la = 0; la <= LMAX_ORB; la++)
      for (int lb = 0; lb <= la; lb++) {
        const int k = pcls(la, lb), p0 = pl.ps.cls_begin[k], np = pl.ps.cls_end[k] - p0;
        bool any = false;
        for (int L = 0; L <= LMAX_AUX; L++) {
          if (lowp[k][L] == 2 && !(use_mm[k][L] && etg[k].p)) {
            lowp[k][L] = 0;
            unsplit_precision(pl, k, L);
          }
          any = any || lowp[k][L] == 2;
        }
        if (!any || np == 0) continue;
        std::vector<int> list;
        std::vector<int64_t> off(np, 0);
        int64_t tote = 0;
        for (int p = 0; p < np; p++) {
          bool nd = false;
          for (int L = 0; L <= LMAX_AUX && !nd; L++) nd = lowp[k][L] == 2 && pl.nsurv[L][p0 + p] > pl.n64[L][p0 + p];
          if (!nd) continue;
          list.push_back(p);
          off[p] = tote;
          tote += (int64_t)o.nprim[pl.ps.s1[p0 + p]] * o.nprim[pl.ps.s2[p0 + p]] * c.et[la][lb].etpp;
        }
        if ((size_t)tote * sizeof(float) > cap32) {
          for (int L = 0; L <= LMAX_AUX; L++)
            if (lowp[k][L] == 2) {
              lowp[k][L] = 0;
              unsplit_precision(pl, k, L);
            }
          // the FP64 table holds only the pairs with an FP64 window: built
          // again below for the whole prefixes (or the class leaves the
          // matrix cores)
          if (etg[k].p) {
            etg[k].release();
            refit64 = true;
          }
          continue;
        }
        if (list.empty()) continue;
        upload_pinned(etlist32[k], list, st);
        upload_pinned(etoff32[k], off, st);
        etg32[k].alloc((size_t)tote);
        c.et[la][lb].fn32(dorb.view, ps1.p + p0, ps2.p + p0, etlist32[k].p, etoff32[k].p, (int)list.size(),
                          c.sph.p, etg32[k].p, st);
        HC(hipGetLastError());
      }
    if (refit64)
      build_et([&](int k, int L, int p) { return (lowp[k][L] ? pl.n64[L][p] : pl.nsurv[L][p]) > 0; }, st, false);
    pl.ncomputed32 = 0;
    for (int L = 0; L <= LMAX_AUX; L++)
      for (int p = 0; p < pl.ps.n(); p++) pl.ncomputed32 += pl.nsurv[L][p] - pl.n64[L][p];
    std::vector<int> f((size_t)(LMAX_AUX + 1) * std::max(pl.ps.n(), 1));
    for (int L = 0; L <= LMAX_AUX; L++)
      std::copy(pl.n64[L].begin(), pl.n64[L].end(), f.begin() + (size_t)L * pl.ps.n());
    upload_pinned(dn64flat, f, st);
    // under QVL the windows [n64, nsurv) hold dropped entries too: count the
    // kept ones on the device (read with the other results at the end)
    if (use_qvl && pl.ps.n() > 0) {
      dn32count.alloc(1);
      HC(hipMemsetAsync(dn32count.p, 0, sizeof(unsigned long long), st));
      for (int L = 0; L <= LMAX_AUX; L++)
        if (!pl.alist[L].empty())
          qvl_count_from(qvl.koff[L].p, qvl.kidx[L].p, dn64flat.p + (size_t)L * pl.ps.n(), pl.ps.n(), dn32count.p, st);
      HC(hipGetLastError());
    }
  }
  vstep("Et tables");
  const int GT = contract_tile();
  // suffix maxima of the prefixes inside each class: under Schwarz screening
  // (or none) the prefixes are non-increasing and are their own suffix
  // maxima; under QVL an upper envelope
  std::vector<std::vector<int>> nsmax_qvl;
  if (use_qvl) {
    nsmax_qvl.assign(LMAX_AUX + 1, {});
    for (int L = 0; L <= LMAX_AUX; L++) {
      nsmax_qvl[L] = pl.nsurv[L];
      for (int k = 0; k < NPCLS; k++)
        for (int p = pl.ps.cls_end[k] - 2; p >= pl.ps.cls_begin[k]; p--)
          nsmax_qvl[L][p] = std::max(nsmax_qvl[L][p], nsmax_qvl[L][p + 1]);
    }
  }
  const std::vector<std::vector<int>>& nsmax = use_qvl ? nsmax_qvl : pl.nsurv;
  auto tile_segments = [&](int64_t cA, int64_t w, std::vector<int64_t>& seg) {
    const int ntile = (int)((w + GT - 1) / GT);
    seg.assign((size_t)ntile * (2 * NPCLS + 1), 0);
    for (int t = 0; t < ntile; t++) {
      const int64_t c0 = cA + (int64_t)t * GT, c1 = std::min(cA + w, c0 + GT);
      int64_t* sg = &seg[(size_t)t * (2 * NPCLS + 1)];
      int64_t tot = 0;
      for (int k = 0; k < NPCLS; k++) {
        const int p0 = pl.ps.cls_begin[k], p1 = pl.ps.cls_end[k];
        int count = 0;
        if (p1 > p0) {
          for (int L = 0; L <= LMAX_AUX; L++) {
            const auto& cols = pl.acol[L];
            if (cols.empty()) continue;
            // first list entry whose columns reach past c0, and whether it starts before c1
            int ia = (int)(std::upper_bound(cols.begin(), cols.end(), c0) - cols.begin()) - 1;
            if (ia < 0) ia = 0;
            if (cols[ia] + 2 * L + 1 <= c0) ia++;
            if (ia >= (int)cols.size() || cols[ia] >= c1) continue;
            // pairs up to the last one with nsurv[L][p] > ia
            const auto& ns = nsmax[L];
            int lo = p0, hi = p1;
            while (lo < hi) {
              const int mid = (lo + hi) / 2;
              if (ns[mid] > ia) lo = mid + 1;
              else hi = mid;
            }
            count = std::max(count, lo - p0);
          }
        }
        const int64_t rlo = p1 > p0 ? pl.prow[p0] : 0;
        const int64_t rhi = count > 0 ? (p0 + count < pl.ps.n() ? pl.prow[p0 + count] : pl.npf) : rlo;
        sg[1 + 2 * k] = rlo;
        sg[2 + 2 * k] = rhi;
        tot += rhi - rlo;
      }
      sg[0] = tot;
    }
  };
  // Pairs of class k launched for aux L (a leading run: the pairs are sorted
  // by Q_ab descending, so the Schwarz prefixes, and under QVL their suffix
  // maxima, are non-increasing along a class) and the end of the longest
  // prefix among them; the whole class without screening.
  auto live_window = [&](int k, int L, int& np, int& hi_live) {
    const int p0 = pl.ps.cls_begin[k], p1 = pl.ps.cls_end[k];
    np = p1 - p0;
    if (opt.screening == DFHIP_SCREEN_NONE || np == 0) return;
    const auto& v = pl.alist[L];
    if (v.empty()) {
      np = 0;
      return;
    }
    if (use_qvl) {
      const std::vector<int>& ns = nsmax_qvl[L];
      np = (int)(std::partition_point(ns.begin() + p0, ns.begin() + p1, [](int n) { return n > 0; }) -
                 (ns.begin() + p0));
      if (np > 0) hi_live = std::min(hi_live, ns[p0]);
    } else {
      const double qm = pl.qc[v[0]];
      np = (int)(std::partition_point(pl.qab.begin() + p0, pl.qab.begin() + p1,
                                      [&](double q) { return !(q * qm < opt.thresh); }) -
                 (pl.qab.begin() + p0));
      if (np > 0) {
        const double qa = pl.qab[p0];
        const int n = (int)(std::partition_point(v.begin(), v.end(),
                                                 [&](int sh) { return !(qa * pl.qc[sh] < opt.thresh); }) -
                            v.begin());
        hi_live = std::min(hi_live, n);
      }
    }
  };
  std::vector<int64_t> hseg;
  DBuf<int64_t> dseg;
  DBuf<int> dkjab;  // QVL: kept-list windows per pair, [window kind][ja, jb][pair]
  if (use_qvl) dkjab.alloc((size_t)6 * std::max(pl.ps.n(), 1));
  double ms_int = 0.0, ms_con = 0.0;
  std::vector<double> dense_stage;
  vstep("integral set-up");
  DBuf<double> dnorm;  // block norms, [sorted pair][aux shell]
  if (out.blocknorm) dnorm.alloc((size_t)std::max(pl.ps.n(), 1) * x.n);
  // under dfhip_autotune: events around each class's launches, read after
  // each column block
  struct ProbeEvents {
    std::vector<hipEvent_t> ev;
    std::vector<int> tag;
    size_t used = 0;
    ~ProbeEvents() {
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
  } pev;
  TuneProbe* const probe = t_probe;
  auto probe_mark = [&](int tagv) {
    if (!probe) return;
    if (pev.used == pev.ev.size()) {
      hipEvent_t e;
      HC(hipEventCreate(&e));
      pev.ev.push_back(e);
      pev.tag.push_back(0);
    }
    pev.tag[pev.used] = tagv;
    HC(hipEventRecord(pev.ev[pev.used++], st));
  };
  for (int blk = 0; blk < nblocks; blk++) {
    const int64_t cA = bounds[blk], cB = bounds[blk + 1], w = cB - cA;
    // only the dense dump reads screened-out entries; the contraction masks them
    if (out.dense3c) HC(hipMemsetAsync(Tp, 0, (size_t)pl.npf * w * sizeof(double), st));
    const int b0 = tm.mark();
    for (int L = 0; L <= LMAX_AUX; L++) {
      const auto& cols = pl.acol[L];
      // list window with columns inside [cA, cB)
      const int lo = (int)(std::lower_bound(cols.begin(), cols.end(), cA) - cols.begin());
      const int hi = (int)(std::lower_bound(cols.begin(), cols.end(), cB) - cols.begin());
      if (lo >= hi) continue;
      // QVL: each pair's kept-list window for the matrix-core launches of this
      // L and column block (the whole prefix, and with FP32 windows [0, n64)
      // and [n64, nsurv)), so that a workgroup reads it instead of searching
      const int* kja[3] = {}, *kjb[3] = {};
      if (use_qvl && pl.ps.n() > 0) {
        const int npall = pl.ps.n();
        const int* n64v = any_low ? dn64flat.p + (size_t)L * npall : nullptr;
        for (int wk = 0; wk < (any_low ? 3 : 1); wk++) {
          int* a0 = dkjab.p + (size_t)(2 * wk) * npall;
          qvl_windows(qvl.koff[L].p, qvl.kidx[L].p, npall, lo, hi, wk == 2 ? n64v : nullptr, wk == 1 ? n64v : dns(L),
                      a0, a0 + npall, st);
          kja[wk] = a0;
          kjb[wk] = a0 + npall;
        }
        HC(hipGetLastError());
      }
      for (int la = 0; la <= LMAX_ORB; la++)
        for (int lb = 0; lb <= la; lb++) {
          const int k = pcls(la, lb), p0 = pl.ps.cls_begin[k];
          // no workgroups for the pairs past the class's last one with a kept
          // entry of L, nor for list entries past the longest prefix
          int np = 0, hi_live = hi;
          live_window(k, L, np, hi_live);
          if (np == 0 || hi_live <= lo) continue;
          PairArgs pa{ps1.p + p0, ps2.p + p0, prow.p + p0, dns(L) + p0, np};
          AuxArgs xa{dalist[L].p, dacol[L].p, lo, hi_live, kmax[L]};
          if (use_qvl) {
            pa.qvl = qvl.pair.p + (size_t)8 * p0;
            xa.qvl = qvl.aux.p;
            xa.thresh = opt.thresh;
            pa.koff = qvl.koff[L].p + p0;  // read by the matrix-core kernels only
            pa.kidx = qvl.kidx[L].p;
            pa.kja = kja[0] + p0;
            pa.kjb = kjb[0] + p0;
          }
          const bool mmk = use_mm[k][L] && etg[k].p;
          const Int3MLaunch& mml = c.int3m[la][lb][L];
          auto launch_mm = [&](const PairArgs& q) {
            mml.fn(dorb.view, daux.view, q, xa, etg[k].p, etoff[k].p, c.sph.p, c.boys.p, Tp - cA, w, st);
            HC(hipGetLastError());
          };
          auto launch_mm32 = [&](const PairArgs& q) {
            mml.fn32(dorb.view, daux.view, q, xa, etg32[k].p, etoff32[k].p, c.sph.p, c.boys.p, Tp - cA, w, st);
            HC(hipGetLastError());
          };
          const int tag0 = (k * (LMAX_AUX + 1) + L) * 2;  // probe tags: window 0 / 1
          if (mixed && lowp[k][L]) {
            PairArgs p64 = pa, p32 = pa;
            p64.nsurv = dn64flat.p + (size_t)L * pl.ps.n() + p0;  // [0, n64): FP64
            p32.nbeg = dn64flat.p + (size_t)L * pl.ps.n() + p0;   // [n64, nsurv): FP32
            if (use_qvl) {
              p64.kja = kja[1] + p0;
              p64.kjb = kjb[1] + p0;
              p32.kja = kja[2] + p0;
              p32.kjb = kjb[2] + p0;
            }
            probe_mark(tag0);
            if (mmk) launch_mm(p64);
            else launch_gen3(c, la, lb, L, brs[k], daux.view, p64, xa, Tp - cA, w);
            probe_mark(-1);
            probe_mark(tag0 + 1);
            if (lowp[k][L] == 1) {
              launch_gen3(c, la, lb, L, brs[k], daux.view, p32, xa, Tp - cA, w, true);
            } else {
              launch_mm32(p32);
            }
            probe_mark(-1);
          } else {
            probe_mark(tag0);
            if (mmk) {
              launch_mm(pa);
            } else if (gen && gen3_has(la, lb, L)) {
              launch_gen3(c, la, lb, L, brs[k], daux.view, pa, xa, Tp - cA, w);
            } else {
              c.int3[la][lb][L].fn(dorb.view, daux.view, pa, xa, c.sph.p, c.boys.p, Tp - cA, w, st);
              HC(hipGetLastError());
            }
            probe_mark(-1);
          }
        }
    }
    if (blk == 0) {
      vstep("integrals queued");
      if (dev_prefix) {
        const auto t = std::chrono::steady_clock::now();
        screen_sweep(pl, opt);
        ms_plan += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
      }
    }
    const int b1 = tm.mark();
    if (out.blocknorm) {
      std::vector<int> bsh, bL;
      std::vector<int64_t> bcol;
      for (int L = 0; L <= LMAX_AUX; L++)
        for (size_t i = 0; i < pl.acol[L].size(); i++)
          if (pl.acol[L][i] >= cA && pl.acol[L][i] < cB) {
            bsh.push_back(pl.alist[L][i]);
            bcol.push_back(pl.acol[L][i]);
            bL.push_back(L);
          }
      DBuf<int> dbsh, dbL;
      DBuf<int64_t> dbcol;
      upload(dbsh, bsh, st);
      upload(dbL, bL, st);
      upload(dbcol, bcol, st);
      const int64_t nt = (int64_t)pl.ps.n() * (int64_t)bsh.size();
      if (nt > 0)
        hipLaunchKernelGGL(block_norm_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, Tp, w, cA,
                           prow.p, ps1.p, ps2.p, dorb.L.p, pl.ps.n(), dbsh.p, dbcol.p, dbL.p, (int)bsh.size(), x.n,
                           dnorm.p);
      HC(hipGetLastError());
      HC(hipStreamSynchronize(st));  // the small uploads above go back to the pool
    }
    if (need_dens) {
      // the packed densities come from the side stream, and d^T's rows were
      // cleared there before the metric
      if (blk == 0) {
        HC(hipStreamWaitEvent(st, lane.ev_pack, 0));
        HC(hipStreamWaitEvent(st, lane.ev_met, 0));
      }
      // d^T into the border rows: A[(naux + I) + lda * P]
      tile_segments(cA, w, hseg);
      upload_pinned(dseg, hseg, st);
      ScreenMask mb = mk;
      mb.colL = mk.colL + cA;
      mb.colia = mk.colia + cA;
      contract_T(Tp, w, Dp, ndens, (int)w, ndens, NPCLS, dseg.p, mb, partp, nsplit, dorig.p + cA, Ap + 2 * naux, lda, 1,
                 st, cipt);
      HC(hipGetLastError());
    }
    const int b2 = tm.mark();
    if (out.dense3c) {
      dense_stage.resize((size_t)pl.npf * w);
      HC(hipMemcpyAsync(dense_stage.data(), Tp, dense_stage.size() * sizeof(double), hipMemcpyDeviceToHost, st));
      HC(hipStreamSynchronize(st));
      const int64_t nbf = o.nbf;
      for (int p = 0; p < pl.ps.n(); p++) {
        const int s1 = pl.ps.s1[p], s2 = pl.ps.s2[p];
        const int n1 = 2 * o.L[s1] + 1, n2 = 2 * o.L[s2] + 1;
        for (int mu = 0; mu < n1; mu++)
          for (int nu = 0; nu < n2; nu++) {
            const double* src = &dense_stage[(size_t)(pl.prow[p] + mu * n2 + nu) * w];
            const int64_t i = o.f0[s1] + mu, j = o.f0[s2] + nu;
            for (int64_t col = 0; col < w; col++) {
              const int64_t P = pl.orig[cA + col];
              out.dense3c[(i * nbf + j) * naux + P] = src[col];
              out.dense3c[(j * nbf + i) * naux + P] = src[col];
            }
          }
      }
    }
    HC(hipStreamSynchronize(st));
    ms_int += tm.ms(b0, b1);
    ms_con += tm.ms(b1, b2);
    // the probe's class times of this block: pairs (tag, -1) of marks
    for (size_t i = 0; probe && i + 1 < pev.used; i += 2) {
      float ms = 0.0f;
      HC(hipEventElapsedTime(&ms, pev.ev[i], pev.ev[i + 1]));
      const int t = pev.tag[i], kk = t / 2 / (LMAX_AUX + 1), LL = t / 2 % (LMAX_AUX + 1), ww = t % 2;
      probe->ms[kk][LL][ww] += ms;
      probe->seen[kk][LL][ww] = true;
    }
    pev.used = 0;
    vlog(opt, "Column block %d of %d: integrals %.1f ms, contraction %.1f ms.\n", blk + 1, nblocks, tm.ms(b0, b1),
         tm.ms(b1, b2));
  }
  const int e6 = tm.mark();

  // B^T = d^T L^-T (the dtrmm of diabatization/intor.f90:198) and the packed
  // Coulomb tensor -B^T B (:241-269) once the factorization has joined
  int e7 = e6, e8 = e6;
  DBuf<double> dB, dJ, dBT, dG, dGp;
  if (need_dens) {
    HC(hipStreamWaitEvent(st, lane.ev_join, 0));
    e7 = tm.mark();
    // the status is read after the final wait: no stop here between the join
    // and the products
    HC(hipMemcpyAsync(lane.hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
    dBT.alloc((size_t)naux * ndens);
    DBuf<double> dLp((size_t)apply_linv_part_size(naux, ndens));
    apply_linv(Ap + 2 * naux, lda, Ap + naux, lda, naux, ndens, dLp.p, dBT.p, st);
    HC(hipGetLastError());
    if (out.df) {
      dB.alloc((size_t)naux * ndens);
      bt_to_b(dBT.p, naux, ndens, dB.p, st);
      HC(hipGetLastError());
    }
    if (out.J || out.gram) {
      dGp.alloc((size_t)gram_chunks(naux, ndens) * ndens * ndens);
      dG.alloc((size_t)ndens * ndens);
      gram_neg(dBT.p, naux, ndens, dGp.p, dG.p, st);
      HC(hipGetLastError());
    }
    if (out.J) {
      const size_t nj = (size_t)out.nstates * out.nstates * out.nstates * out.nstates;
      dJ.alloc(nj);
      HC(hipMemsetAsync(dJ.p, 0, nj * sizeof(double), st));
      assemble_j(dG.p, ndens, 0, ndens, out.nstates, dJ.p, st);
      HC(hipGetLastError());
    }
    e8 = tm.mark();
  }
  if (out.df) HC(hipMemcpyAsync(out.df, dB.p, (size_t)naux * ndens * sizeof(double), hipMemcpyDeviceToHost, st));
  if (out.J) HC(hipMemcpyAsync(out.J, dJ.p, dJ.n * sizeof(double), hipMemcpyDeviceToHost, st));
  if (out.gram)
    HC(hipMemcpyAsync(out.gram, dG.p, (size_t)ndens * ndens * sizeof(double), hipMemcpyDeviceToHost, st));
  const int e9 = tm.mark();
  HC(hipStreamSynchronize(st));
  if (out.gram)  // the Gram kernels give -B^T B (the Coulomb tensor's sign)
    for (size_t i = 0; i < (size_t)ndens * ndens; i++) out.gram[i] = -out.gram[i];
  if (out.blocknorm) {
    std::vector<double> hn((size_t)pl.ps.n() * x.n);
    HC(hipMemcpy(hn.data(), dnorm.p, hn.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int p = 0; p < pl.ps.n(); p++) {
      const int a = std::max(pl.ps.s1[p], pl.ps.s2[p]), b = std::min(pl.ps.s1[p], pl.ps.s2[p]);
      std::memcpy(out.blocknorm + ((size_t)a * (a + 1) / 2 + b) * x.n, &hn[(size_t)p * x.n], x.n * sizeof(double));
    }
  }
  if (need_dens && lane.hinfo[0] != 0)
    throw Error{DFHIP_ERR_NOT_SPD, "metric (P|Q) is not positive definite (leading minor " +
                                       std::to_string(lane.hinfo[0]) + ")"};

  if (dn32count.p) {
    unsigned long long n32 = 0;
    HC(hipMemcpy(&n32, dn32count.p, sizeof(n32), hipMemcpyDeviceToHost));
    pl.ncomputed32 = (int64_t)n32;
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->ntriplets = pl.ntriplets;
    stats->ncomputed = pl.ncomputed;
    stats->ncomputed_fp32 = pl.ncomputed32;
    stats->npairs_zero_q = std::count(pl.qab.begin(), pl.qab.end(), 0.0);
    stats->nskipped = pl.ntriplets - pl.ncomputed;
    stats->nbf = o.nbf;
    stats->naux = naux;
    stats->npair_functions = pl.npf;
    stats->ncolumn_blocks = nblocks;
    stats->ms_upload = tm.ms(e0, e1);
    stats->ms_schwarz = tm.ms(e1, e2);
    stats->ms_metric = tm.ms(e2m, e3);
    stats->ms_cholesky = tm.ms(e6, e7);
    stats->ms_pack = tm.ms(e2, e5);
    stats->ms_int3c2e = ms_int;
    stats->ms_contract = ms_con;
    stats->ms_solve = 0.0;
    stats->ms_assemble = tm.ms(e7, e8);
    stats->ms_download = tm.ms(e8, e9);
    stats->ms_total = tm.ms(e0, e9);
    stats->ms_h2d_densities = h2d0 >= 0 ? tm.ms(h2d0, h2d1) : 0.0;
    stats->ms_host_plan = ms_plan;
    for (int k = 0; k < NPCLS * 6; k++) {
      stats->class_triplets[k] = pl.class_trip[k];
      stats->class_prim_triplets[k] = pl.class_prim[k];
    }
  }
  if (opt.verbose >= 3) {
    vlog(opt, "  %-28s host %8.3f ms before the first event\n", "call entry",
         std::chrono::duration<double, std::milli>(run_t0 - entry_t).count());
    vlog(opt, "  %-28s host %8.3f ms\n", "results on the host",
         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - run_t0).count());
  }
  if (opt.verbose >= 3)
    for (const auto& m : vmarks)
      vlog(opt, "  %-28s host %8.3f ms, stream %8.3f ms\n", m.first, m.second.first, tm.ms(e0, m.second.second));
  if (opt.verbose) {
    std::fflush(stdout);
    std::printf("Schwarz norms took %8.4f s.\n", tm.ms(e1, e2) * 1e-3);
    std::printf("Calculation of density fitting metric (P|Q) took %8.4f s.\n", tm.ms(e2m, e3) * 1e-3);
    std::printf("Cholesky factorization of (P|Q) bordered by d took %8.4f s.\n", tm.ms(e6, e7) * 1e-3);
    std::printf("Density contraction took %8.4f s.\n", (ms_int + ms_con) * 1e-3);
    const int64_t skipped = pl.ntriplets - pl.ncomputed;
    std::printf("Screened out %12lld of %12lld integral triplets (%8.2f%%)\n", (long long)skipped,
                (long long)pl.ntriplets, 100.0 * (double)skipped / (double)std::max<int64_t>(pl.ntriplets, 1));
    std::fflush(stdout);
  }
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    t_status = DFHIP_OK;
    t_msg.clear();
  } catch (const Error& e) {
    t_status = e.code;
    t_msg = e.msg;
  } catch (const std::exception& e) {
    t_status = DFHIP_ERR_ARG;
    t_msg = e.what();
  }
  if (t_status != DFHIP_OK && std::getenv("DFHIP_QUIET") == nullptr)
    std::fprintf(stderr, "dfhip error %d: %s\n", t_status, t_msg.c_str());
  return t_status;
}

dfhip_basis mkbasis(int32_t n, int32_t nd, const int32_t* bc, const int32_t* bs, const double* bd) {
  return dfhip_basis{n, nd, bc, bs, bd};
}

std::vector<double> schwarz_bra_host(const dfhip_basis& b) {
  std::lock_guard<std::mutex> lock(ctx_mutex());
  ensure_ctx();
  Ctx& c = ctx();
  HostShells o = parse_shells(b, LMAX_ORB, "basis");
  DevShellBufs d;
  d.upload_from(o, cur_lane().st);
  cur_lane().arena.reset();
  const PairClasses pc = pair_classes(o);
  PairSet ps = make_pairs(o, pc);
  std::vector<double> q = schwarz_bra_dev(c, o, d.view, pc, ps);
  std::vector<double> outp(ps.n());
  for (int p = 0; p < ps.n(); p++) outp[ps.packed[p]] = q[p];
  return outp;
}

std::vector<double> schwarz_aux_host(const dfhip_basis& b) {
  std::lock_guard<std::mutex> lock(ctx_mutex());
  ensure_ctx();
  HostShells x = parse_shells(b, LMAX_AUX, "basis");
  cur_lane().arena.reset();
  const hipStream_t st = cur_lane().st;
  DevShellBufs d;
  d.upload_from(x, st);
  DBuf<double> dq(x.n);
  schwarz_aux_dev(ctx(), x, d.view, dq.p, st);
  std::vector<double> q(x.n);
  HC(hipMemcpyAsync(q.data(), dq.p, x.n * sizeof(double), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
  return q;
}

// The allocation-free form of schwarz_aux_host for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops):
// the same diagonal blocks (C|C) through the LDS 2c2e kernels, with every
// device buffer (shell arrays, aux lists, the compact blocks, the norms)
// carved out of the static scratch array; no hipMalloc, hipFree or pinned
// allocation (host-side vectors only).  Needs f_init() on this device first,
// which builds the tables and looks up the scratch array's address.  The
// library's lock serializes calls in this process and each call waits for its
// work before returning, so the scratch array is never used by two calls at
// once here; the array keeps the last call's contents (nothing in it needs to
// start at zero: every element a kernel reads is written earlier in the same
// call).
void schwarz_aux_noalloc(const dfhip_basis& b, double* integrals) {
  std::lock_guard<std::mutex> lock(ctx_mutex());
  Ctx& c = ctx();
  if (!c.ready) throw Error{DFHIP_ERR_ARG, "call f_init() before the allocation-free entry point"};
  int dev = -1;
  HC(hipGetDevice(&dev));
  if (dev != c.device)
    throw Error{DFHIP_ERR_UNSUPPORTED, "the library was set up on device " + std::to_string(c.device) +
                                           ", the current device is " + std::to_string(dev)};
  unsigned char* base = resolve_scratch(c, dev);
  const HostShells x = parse_shells(b, LMAX_AUX, "basis");
  constexpr int LDC = 2 * LMAX_AUX + 1;
  const int n = x.n, nexp = (int)x.exps.size();
  // layout of the scratch array, each region 256-byte aligned
  size_t top = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = top;
    top = (top + bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t oL = carve(n * sizeof(int)), onp = carve(n * sizeof(int)), opp = carve(n * sizeof(int)),
               of0 = carve(n * sizeof(int)), oxyz = carve(3 * n * sizeof(double)),
               oex = carve(nexp * sizeof(double)), oco = carve(nexp * sizeof(double)),
               os1 = carve(n * sizeof(int)), os2 = carve(n * sizeof(int)), onb = carve(n * sizeof(int)),
               one = carve(n * sizeof(int)), orow = carve(n * sizeof(int64_t)), ocol = carve(n * sizeof(int64_t));
  const size_t nup = top;
  const size_t oblk = carve((size_t)x.nbf * LDC * sizeof(double)), oq = carve(n * sizeof(double));
  if (top > NOALLOC_SCRATCH_BYTES)
    throw Error{DFHIP_ERR_ARG, "the basis needs " + std::to_string(top) +
                                   " bytes of device scratch; the allocation-free entry point holds at most " +
                                   std::to_string(NOALLOC_SCRATCH_BYTES) +
                                   " (88 bytes per auxiliary function plus about 100 per shell)"};
  // host image of everything that goes up, in the same layout
  std::vector<unsigned char> img(nup, 0);
  auto put = [&](size_t o, const void* src, size_t bytes) {
    if (bytes) std::memcpy(img.data() + o, src, bytes);
  };
  put(oL, x.L.data(), n * sizeof(int));
  put(onp, x.nprim.data(), n * sizeof(int));
  put(opp, x.pptr.data(), n * sizeof(int));
  put(of0, x.f0.data(), n * sizeof(int));
  put(oxyz, x.xyz.data(), 3 * n * sizeof(double));
  put(oex, x.exps.data(), nexp * sizeof(double));
  put(oco, x.coefs.data(), nexp * sizeof(double));
  // per L: bra shell i of the list meets only its own list entry (window [k, k + 1))
  std::vector<int> lbeg(LMAX_AUX + 2, 0), kmax(LMAX_AUX + 1, 1);
  {
    std::vector<int> s1, s2, nb, ne;
    std::vector<int64_t> row, col;
    for (int L = 0; L <= LMAX_AUX; L++) {
      lbeg[L] = (int)s1.size();
      int k = 0;
      for (int i = 0; i < n; i++) {
        if (x.L[i] != L) continue;
        s1.push_back(i);
        s2.push_back(-1);
        nb.push_back(k);
        ne.push_back(k + 1);
        row.push_back(x.f0[i]);
        col.push_back(0);
        kmax[L] = std::max(kmax[L], x.nprim[i]);
        k++;
      }
    }
    lbeg[LMAX_AUX + 1] = (int)s1.size();
    put(os1, s1.data(), n * sizeof(int));
    put(os2, s2.data(), n * sizeof(int));
    put(onb, nb.data(), n * sizeof(int));
    put(one, ne.data(), n * sizeof(int));
    put(orow, row.data(), n * sizeof(int64_t));
    put(ocol, col.data(), n * sizeof(int64_t));
  }
  const hipStream_t st = c.lane0.st;
  HC(hipMemcpyAsync(base, img.data(), nup, hipMemcpyHostToDevice, st));
  auto at = [&](size_t o) { return base + o; };
  const DevShells dv{n,
                     (const int*)at(oL),
                     (const int*)at(onp),
                     (const int*)at(opp),
                     (const int*)at(of0),
                     (const double*)at(oxyz),
                     (const double*)at(oex),
                     (const double*)at(oco)};
  const int* ds1 = (const int*)at(os1);
  const int* ds2 = (const int*)at(os2);
  const int* dnb = (const int*)at(onb);
  const int* dne = (const int*)at(one);
  const int64_t* drow = (const int64_t*)at(orow);
  const int64_t* dcol = (const int64_t*)at(ocol);
  double* blk = (double*)at(oblk);
  double* q = (double*)at(oq);
  for (int L = 0; L <= LMAX_AUX; L++) {
    const int b0 = lbeg[L], np = lbeg[L + 1] - b0;
    if (np == 0) continue;
    PairArgs pa{ds1 + b0, ds2 + b0, drow + b0, dne + b0, np};
    pa.nbeg = dnb + b0;
    const AuxArgs xa{ds1 + b0, dcol + b0, 0, np, kmax[L]};
    c.int3[L][0][L].fn(dv, dv, pa, xa, c.sph.p, c.boys.p, blk, LDC, st);
    HC(hipGetLastError());
  }
  schwarz_aux_from_metric(blk, LDC, dv.L, dv.f0, n, q, st, true);
  HC(hipGetLastError());
  HC(hipMemcpyAsync(integrals, q, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
}

// The QVL screener records of the device path: per packed pair
// k = a(a+1)/2 + b (a >= b) [centre x, y, z, extent, (min a + min b)^(1/4),
// Q_ab, 0, 0] from qvl_pair_kernel, per aux shell [extent, beta, Q_C, 0].
void qvl_records_host(const dfhip_basis& b, const dfhip_basis& xb, double* pair_out, double* aux_out) {
  std::lock_guard<std::mutex> lock(ctx_mutex());
  ensure_ctx();
  Ctx& c = ctx();
  HostShells o = parse_shells(b, LMAX_ORB, "orbital basis");
  HostShells x = parse_shells(xb, LMAX_AUX, "auxiliary basis");
  cur_lane().arena.reset();
  const hipStream_t st = cur_lane().st;
  DevShellBufs dorb, daux;
  dorb.upload_from(o, st);
  daux.upload_from(x, st);
  const PairClasses pc = pair_classes(o);
  PairSet ps = make_pairs(o, pc);
  const std::vector<double> qab = schwarz_bra_dev(c, o, dorb.view, pc, ps);
  DBuf<double> dqc(x.n);
  schwarz_aux_dev(c, x, daux.view, dqc.p, st);
  std::vector<double> qc(x.n);
  HC(hipMemcpyAsync(qc.data(), dqc.p, x.n * sizeof(double), hipMemcpyDeviceToHost, st));
  const int np = ps.n();
  DBuf<int> ds1, ds2;
  DBuf<double> dq, rec((size_t)8 * std::max(np, 1));
  upload(ds1, ps.s1, st);
  upload(ds2, ps.s2, st);
  upload(dq, qab, st);
  qvl_pair_records(dorb.view, ds1.p, ds2.p, dq.p, np, qvl_erfi(), rec.p, st);
  HC(hipGetLastError());
  std::vector<double> h((size_t)8 * np);
  HC(hipMemcpyAsync(h.data(), rec.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
  for (int p = 0; p < np; p++) std::memcpy(pair_out + (size_t)8 * ps.packed[p], &h[(size_t)8 * p], 8 * sizeof(double));
  const std::vector<double> xr = aux_qvl_records(x, qc);
  std::memcpy(aux_out, xr.data(), xr.size() * sizeof(double));
}

std::vector<double> metric_host(const dfhip_basis& b, std::vector<double>* qc) {
  std::lock_guard<std::mutex> lock(ctx_mutex());
  ensure_ctx();
  Ctx& c = ctx();
  HostShells x = parse_shells(b, LMAX_AUX, "basis");
  cur_lane().arena.reset();
  DevShellBufs d;
  d.upload_from(x, cur_lane().st);
  DBuf<double> M((size_t)x.nbf * x.nbf);
  metric_dev(c, x, d.view, M.p, x.nbf);
  std::vector<double> h((size_t)x.nbf * x.nbf);
  if (qc) {
    DBuf<double> dq(x.n);
    schwarz_aux_from_metric(M.p, x.nbf, d.L.p, d.f0.p, x.n, dq.p, cur_lane().st);
    qc->resize(x.n);
    HC(hipMemcpyAsync(qc->data(), dq.p, x.n * sizeof(double), hipMemcpyDeviceToHost, cur_lane().st));
  }
  HC(hipMemcpyAsync(h.data(), M.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, cur_lane().st));
  HC(hipStreamSynchronize(cur_lane().st));
  return h;
}

// ---------------------------------------------------------------------------
// (P|Q)^power by the eigen route of common/linalg.f90:36-58 (matrix_powerh):
// A = V diag(lambda) V^T with rocSOLVER's divide-and-conquer dsyevd (the
// reference's LAPACK dsyevd, upper triangle), then V diag(lambda^power) V^T
// with one rocBLAS dgemm.  A (n x n, column-major, device) is overwritten by
// V; out receives the power.  A negative eigenvalue with a fractional power
// gives NaN, as in the reference.

__global__ void scale_columns_pow_kernel(const double* __restrict__ V, const double* __restrict__ lam, int n,
                                         double power, double* __restrict__ W) {
  const int64_t tot = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x)
    W[e] = V[e] * pow(lam[e / n], power);
}

#define RBC(x)                                                                                          \
  do {                                                                                                  \
    rocblas_status s_ = (x);                                                                            \
    if (s_ != rocblas_status_success)                                                                   \
      throw Error{DFHIP_ERR_HIP, std::string(#x) + ": " + rocblas_status_to_string(s_)};               \
  } while (0)

void matrix_powerh_dev(Ctx& c, double* A, int n, double power, double* out, hipStream_t st) {
  if (!c.blas) RBC(rocblas_create_handle(&c.blas));
  RBC(rocblas_set_stream(c.blas, st));
  DBuf<double> lam(n), E(n), W((size_t)n * n);
  DBuf<int> info(1);
  RBC(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, n, A, n, lam.p, E.p, info.p));
  hipLaunchKernelGGL(scale_columns_pow_kernel, dim3(1024), dim3(256), 0, st, A, lam.p, n, power, W.p);
  HC(hipGetLastError());
  const double one = 1.0, zero = 0.0;
  RBC(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one, W.p, n, A, n, &zero,
                    out, n));
  int hinfo = 0;
  HC(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
  if (hinfo != 0) throw Error{DFHIP_ERR_NOT_SPD, "dsyevd did not converge (info " + std::to_string(hinfo) + ")"};
}

// (P|Q)^power of an aux basis into host out (naux x naux).
void metric_powerh_host(const dfhip_basis& b, double power, double* out) {
  std::lock_guard<std::mutex> lock(ctx_mutex());
  ensure_ctx();
  Ctx& c = ctx();
  HostShells x = parse_shells(b, LMAX_AUX, "basis");
  cur_lane().arena.reset();
  const hipStream_t st = cur_lane().st;
  DevShellBufs d;
  d.upload_from(x, st);
  const int n = x.nbf;
  DBuf<double> M((size_t)n * n), P((size_t)n * n);
  metric_dev(c, x, d.view, M.p, n);
  matrix_powerh_dev(c, M.p, n, power, P.p, st);
  HC(hipMemcpyAsync(out, P.p, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// The screening-estimate records of intor/benchmark.f90:15-132
// (benchmark_int3c2e): for a = 1..n, b = 1..a (1-based, the reference's
// loops), pairs that are neither concentric (|A - B| <= 1e-6) nor small
// (Q_ab <= 1d-10), and every aux shell c: [Schwarz estimate Q_ab Q_C, QVL
// estimate, ||(ab|C)||_F of the unscreened block, R = |contracted centre - C|]
// (screening.f90:315-326).  Also the counts and phase times the reference
// logs.
struct Estimates {
  std::vector<double> rec;  // 4 per record
  int64_t ntriplets = 0, nskipped = 0;
  int nbf = 0, naux = 0, nshells = 0, nshells_aux = 0;
  double t_screener = 0, t_metric = 0, t_powerh = 0, t_ints = 0;  // seconds
};

Estimates int3c2e_estimates(const dfhip_basis& b, const dfhip_basis& xb) {
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point z) { return std::chrono::duration<double>(z - a).count(); };
  Estimates es;
  const HostShells o = parse_shells(b, LMAX_ORB, "orbital basis");
  const HostShells x = parse_shells(xb, LMAX_AUX, "auxiliary basis");
  es.nbf = o.nbf;
  es.naux = x.nbf;
  es.nshells = o.n;
  es.nshells_aux = x.n;
  const size_t npk = (size_t)o.n * (o.n + 1) / 2;
  auto t0 = clk::now();
  std::vector<double> pr(8 * npk), xr(4 * (size_t)x.n);
  qvl_records_host(b, xb, pr.data(), xr.data());
  auto t1 = clk::now();
  es.t_screener = secs(t0, t1);
  {
    // the metric and its inverse square root (timed, not used further: the
    // reference computes them the same way)
    std::vector<double> P((size_t)x.nbf * x.nbf);
    const auto m0 = clk::now();
    metric_host(xb, nullptr);
    const auto m1 = clk::now();
    metric_powerh_host(xb, -0.5, P.data());
    es.t_metric = secs(m0, m1);
    es.t_powerh = secs(m1, clk::now());
  }
  std::vector<double> norms(npk * (size_t)x.n, 0.0);
  const auto i0 = clk::now();
  {
    dfhip_options opt = defaults();
    opt.screening = DFHIP_SCREEN_NONE;
    Outputs out;
    out.blocknorm = norms.data();
    run(b, xb, 0, 0, nullptr, opt, out, nullptr);
  }
  es.ntriplets = (int64_t)npk * x.n;
  const double f32_1e10 = (double)1e-10f;  // intor/benchmark.f90:100, a real(4) literal
  for (int a = 0; a < o.n; a++)
    for (int bb = 0; bb <= a; bb++) {
      const size_t k = (size_t)a * (a + 1) / 2 + bb;
      const double dx = o.xyz[3 * a] - o.xyz[3 * bb], dy = o.xyz[3 * a + 1] - o.xyz[3 * bb + 1],
                   dz = o.xyz[3 * a + 2] - o.xyz[3 * bb + 2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) <= 1e-6 || pr[8 * k + 5] <= 1e-10) continue;
      for (int c = 0; c < x.n; c++) {
        const double* C = &x.xyz[3 * c];
        const double qvl = qvl_estimate(&pr[8 * k], &xr[4 * c], C[0], C[1], C[2], x.L[c]);
        const double rx = pr[8 * k] - C[0], ry = pr[8 * k + 1] - C[1], rz = pr[8 * k + 2] - C[2];
        if (qvl < f32_1e10) es.nskipped++;
        es.rec.push_back(pr[8 * k + 5] * xr[4 * c + 2]);
        es.rec.push_back(qvl);
        es.rec.push_back(norms[k * x.n + c]);
        es.rec.push_back(std::sqrt(rx * rx + ry * ry + rz * rz));
      }
    }
  es.t_ints = secs(i0, clk::now());
  return es;
}

// ---------------------------------------------------------------------------
// The std/ DF tensors (std/intor.f90:15-292) through the Coulomb path: each
// is d(P, I) = sum_{mu nu} (mu nu|P) D_I(mu, nu) for a set of model densities,
// B = L^-1 d, under the reference's Schwarz rule Q_ab Q_C < 1e-10 (the
// single-precision literal, std/intor.f90:78,223).
//   df1c(mu, P)  = B(P, mu) with D_mu = e_mu e_mu^T: the diagonal bra
//                  elements (mu mu|Q) of the (aa|C) blocks (:67-103), times L^-T;
//   eri2c        = df1c df1c^T = B^T B (:116-130);
//   df2c_mo(ia, P) = B(P, ia) with D_ia = (C_i C_a^T + C_a C_i^T) / 2 for the
//                  packed MO pairs i >= a (:176-274): sum_{mu nu} C_mu,i C_nu,a
//                  (mu nu|Q) for the symmetric (mu nu|Q).
// The reference also drops single integrals and half-transformed elements of
// magnitude <= 1e-10 inside df2c_mo (:244-266); this evaluation keeps them.

// D[(mu nbf + mu) nbf + mu] = 1 (D cleared before)
__global__ void unit_diag_densities_kernel(double* __restrict__ D, int nbf) {
  const int64_t mu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mu < nbf) D[(mu * nbf + mu) * nbf + mu] = 1.0;
}

// D[pack][mu][nu] = (C[mu, i] C[nu, a] + C[nu, i] C[mu, a]) / 2, pack = i(i+1)/2 + a,
// C column-major (nbf, nmo)
__global__ void mo_pair_densities_kernel(const double* __restrict__ C, int nbf, int nmo, double* __restrict__ D) {
  const int64_t npack = (int64_t)nmo * (nmo + 1) / 2, nn = (int64_t)nbf * nbf, tot = npack * nn;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pk = e / nn, r = e - pk * nn;
    const int mu = (int)(r / nbf), nu = (int)(r - (int64_t)mu * nbf);
    int i = (int)((std::sqrt(8.0 * (double)pk + 1.0) - 1.0) * 0.5);
    while ((int64_t)i * (i + 1) / 2 > pk) i--;
    while ((int64_t)(i + 1) * (i + 2) / 2 <= pk) i++;
    const int a = (int)(pk - (int64_t)i * (i + 1) / 2);
    D[e] = 0.5 * (C[mu + (int64_t)nbf * i] * C[nu + (int64_t)nbf * a] + C[nu + (int64_t)nbf * i] * C[mu + (int64_t)nbf * a]);
  }
}

int basis_nsph(const dfhip_basis& b) {
  int n = 0;
  for (int i = 0; i < b.nshells; i++) n += 2 * b.bas_spec[i + b.nshells] + 1;
  return n;
}

// The model densities of kind 0 (df1c / eri2c) or 1 (df2c_mo) on the device,
// then the Coulomb path into out (df and/or gram).
void std_tensors(const dfhip_basis& b, const dfhip_basis& x, int kind, int nmo, const double* C, Outputs out) {
  if (b.nshells < 1 || x.nshells < 1 || !b.bas_spec || !x.bas_spec)
    throw Error{DFHIP_ERR_ARG, "empty basis"};
  const int nbf = basis_nsph(b);
  std::lock_guard<std::mutex> lock(ctx_mutex());
  ensure_ctx();
  const hipStream_t st = cur_lane().st;
  cur_lane().arena.reset();
  int ndens;
  DBuf<double> D, dC;
  if (kind == 0) {
    ndens = nbf;
    D.alloc((size_t)nbf * nbf * nbf);
    HC(hipMemsetAsync(D.p, 0, D.n * sizeof(double), st));
    hipLaunchKernelGGL(unit_diag_densities_kernel, dim3((nbf + 255) / 256), dim3(256), 0, st, D.p, nbf);
  } else {
    if (nmo < 1 || !C) throw Error{DFHIP_ERR_ARG, "mo_tensor is empty"};
    ndens = nmo * (nmo + 1) / 2;
    std::vector<double> hc(C, C + (size_t)nbf * nmo);
    upload(dC, hc, st);
    D.alloc((size_t)ndens * nbf * nbf);
    hipLaunchKernelGGL(mo_pair_densities_kernel, dim3(2048), dim3(256), 0, st, dC.p, nbf, nmo, D.p);
  }
  HC(hipGetLastError());
  HC(hipStreamSynchronize(st));
  dfhip_options o = defaults();
  o.dens_layout = DFHIP_DENS_C;
  o.dens_on_device = 1;
  run(b, x, ndens, nbf, D.p, o, out, nullptr, false);
}

// ---- tuning table: file, measurement ------------------------------------

void tuning_save(const Tuning& t, const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) throw Error{DFHIP_ERR_ARG, std::string("cannot write ") + path};
  std::fprintf(f, "# libdfhip tuning table (dfhip_autotune)\n");
  std::fprintf(f, "# int3c2e la lb lc window family(1 generated, 2 matrix cores) ms_generated ms_matrix_cores\n");
  std::fprintf(f, "# contract ndens densities_per_thread split_scale ms\n");
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++)
      for (int L = 0; L <= LMAX_AUX; L++)
        for (int w = 0; w < 2; w++) {
          const int k = pcls(la, lb);
          if (t.fam[k][L][w])
            std::fprintf(f, "int3c2e %d %d %d %d %d %.6f %.6f\n", la, lb, L, w, t.fam[k][L][w], t.ms[k][L][w][0],
                         t.ms[k][L][w][1]);
        }
  for (const auto& c : t.contract) std::fprintf(f, "contract %d %d %.3f %.6f\n", c.ndens, c.ipt, c.nscale, c.ms);
  const bool ok = std::fclose(f) == 0;
  if (!ok) throw Error{DFHIP_ERR_ARG, std::string("cannot write ") + path};
}

Tuning tuning_load(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) throw Error{DFHIP_ERR_ARG, std::string("cannot read ") + path};
  Tuning t;
  char line[512];
  int lineno = 0;
  std::string bad;
  while (std::fgets(line, sizeof(line), f)) {
    lineno++;
    char kind[32] = {};
    if (line[0] == '#' || std::sscanf(line, "%31s", kind) != 1) continue;
    if (std::strcmp(kind, "int3c2e") == 0) {
      int la, lb, L, w, fam;
      double m0 = 0.0, m1 = 0.0;
      if (std::sscanf(line, "%*s %d %d %d %d %d %lf %lf", &la, &lb, &L, &w, &fam, &m0, &m1) < 5 || la < 0 ||
          la > LMAX_ORB || lb < 0 || lb > la || L < 0 || L > LMAX_AUX || w < 0 || w > 1 || fam < 1 || fam > 2) {
        bad = "line " + std::to_string(lineno);
        break;
      }
      const int k = pcls(la, lb);
      t.fam[k][L][w] = (signed char)fam;
      t.ms[k][L][w][0] = m0;
      t.ms[k][L][w][1] = m1;
    } else if (std::strcmp(kind, "contract") == 0) {
      Tuning::Contract c{};
      if (std::sscanf(line, "%*s %d %d %lf %lf", &c.ndens, &c.ipt, &c.nscale, &c.ms) < 3 || c.ndens < 1 ||
          !contract_ipt_ok(c.ipt) || !(c.nscale > 0.0) || c.nscale > 64.0) {
        bad = "line " + std::to_string(lineno);
        break;
      }
      t.contract.push_back(c);
    } else {
      bad = "line " + std::to_string(lineno);
      break;
    }
  }
  std::fclose(f);
  if (!bad.empty()) throw Error{DFHIP_ERR_ARG, std::string(path) + ": malformed " + bad};
  return t;
}

// a tuning table named by DFHIP_TUNING (dfhip_autotune's file), read when
// the library sets up its device context
void load_env_tuning() {
  const char* tp = std::getenv("DFHIP_TUNING");
  if (!tp || !*tp) return;
  Tuning t = tuning_load(tp);
  std::lock_guard<std::mutex> g(tuning_mutex());
  tuning() = t;
}

// dfhip_autotune: measured runs of the contraction on the caller's workload
// (densities as for dfhip_contract_coulomb_densities; the Gram matrix is the
// only output), one per candidate:
//   1. every class with both kernels on the generated one, then on the
//      matrix-core one (mixed: also the FP32 window on the matrix cores);
//      per (la, lb, lc) and window the faster family is kept, the rule's
//      choice unless the other is at least 2 % faster;
//   2. with those families, the contraction's densities per thread over the
//      supported widths, then its split-K depth at 1/2 and 2 times the default.
// A first run with the current choices warms the tables and the workspaces.
Tuning autotune_host(const dfhip_basis& B, const dfhip_basis& X, int ndens, int nbfs, const double* dens,
                     const dfhip_options& opt) {
  const bool mixed = opt.precision == DFHIP_MIXED;
  std::vector<double> gram((size_t)ndens * ndens);
  Outputs out;
  out.gram = gram.data();
  struct Reset {
    ~Reset() { t_probe = nullptr; }
  } reset;
  auto measure = [&](TuneProbe& p) {
    dfhip_stats st{};
    t_probe = &p;
    run(B, X, ndens, nbfs, dens, opt, out, &st);
    t_probe = nullptr;
    return st.ms_contract;
  };
  {
    TuneProbe warm;
    measure(warm);
  }
  TuneProbe gen64, mm64, mm32;
  gen64.fam64 = 1;
  gen64.fam32 = 1;
  mm64.fam64 = 2;
  mm64.fam32 = 1;
  mm32.fam64 = 2;
  mm32.fam32 = 2;
  measure(gen64);
  measure(mm64);
  if (mixed) measure(mm32);
  Tuning t;
  const bool gen = use_generated();
  Ctx& c = ctx();
  for (int la = 0; la <= LMAX_ORB; la++)
    for (int lb = 0; lb <= la; lb++)
      for (int L = 0; L <= LMAX_AUX; L++) {
        const int k = pcls(la, lb);
        if (!(gen && gen3_has(la, lb, L)) || c.int3m[la][lb][L].fn == nullptr) continue;
        if (gen64.seen[k][L][0] && mm64.seen[k][L][0]) {
          const double tg = gen64.ms[k][L][0], tm = mm64.ms[k][L][0];
          const bool rule_mm = prefer_mm(la, lb, L);
          const bool pick_mm = rule_mm ? !(tg < 0.98 * tm) : tm < 0.98 * tg;
          t.fam[k][L][0] = pick_mm ? 2 : 1;
          t.ms[k][L][0][0] = tg;
          t.ms[k][L][0][1] = tm;
        }
        // the FP32 window on the matrix cores needs the FP64 window there
        if (mixed && t.fam[k][L][0] == 2 && c.int3m[la][lb][L].fn32 != nullptr && mm64.seen[k][L][1] &&
            mm32.seen[k][L][1]) {
          const double tg = mm64.ms[k][L][1], tm = mm32.ms[k][L][1];
          t.fam[k][L][1] = tm < 0.98 * tg ? 2 : 1;
          t.ms[k][L][1][0] = tg;
          t.ms[k][L][1][1] = tm;
        }
      }
  // the contraction with the families above in place
  Tuning saved;
  {
    std::lock_guard<std::mutex> g(tuning_mutex());
    saved = tuning();
    std::memcpy(tuning().fam, t.fam, sizeof(t.fam));
    tuning().contract.clear();
  }
  struct Restore {
    Tuning& s;
    bool on = true;
    ~Restore() {
      if (!on) return;
      std::lock_guard<std::mutex> g(tuning_mutex());
      tuning() = s;
    }
  } restore{saved};
  // the default width is kept unless another one is at least 2 % faster
  const int ipt0 = contract_ipt(ndens);
  std::vector<std::pair<int, double>> cand;
  for (int ipt : {1, 2, 3, 4, 5, 7, 102, 103, 104, 105, 106, 107}) {
    // tiles from a quarter to twice the densities (the default always): VALU
    // codes are 16 densities per unit, matrix-core codes 100 + 32-density units
    const int wd = ipt > 100 ? 32 * (ipt - 100) : 16 * ipt;
    if (ipt != ipt0 && (wd > 2 * std::max(ndens, 16) || 4 * wd < ndens)) continue;
    TuneProbe p;
    p.ipt = ipt;
    cand.push_back({ipt, measure(p)});
  }
  int best_ipt = ipt0;
  double best = 0.0, best_scale = 1.0;
  for (const auto& cd : cand)
    if (cd.first == ipt0) best = cd.second;
  for (const auto& cd : cand)
    if (cd.second < 0.98 * best) {
      best = cd.second;
      best_ipt = cd.first;
    }
  for (double sc : {0.5, 2.0}) {
    TuneProbe p;
    p.ipt = best_ipt;
    p.nscale = sc;
    const double ms = measure(p);
    if (ms < 0.98 * best) {
      best = ms;
      best_scale = sc;
    }
  }
  t.contract.push_back(Tuning::Contract{ndens, best_ipt, best_scale, best});
  restore.on = false;
  return t;
}

// grid/grid.f90:112-286 (eval_densities): out[j + ndens p] = density j at
// point p (grid3d[3 p + d]) from the Cartesian density matrices dens
// (Fortran (ndens, nbfs, nbfs), element (j, mu, nu) read for mu >= nu) in
// blocks of bs points, pairs screened per block by thresh.  The points are
// walked in batches of whole blocks whose AO values fit a fixed budget.
void eval_densities_host(const dfhip_basis& b, int64_t npoints, const double* grid3d, int ndens, int nbfs,
                         const double* dens, double* out, int bs, double thresh) {
  if (!out) throw Error{DFHIP_ERR_ARG, "grid_densities is NULL"};
  if (bs < 1) throw Error{DFHIP_ERR_ARG, "blk_size must be >= 1"};
  if (npoints < 0 || ndens < 0) throw Error{DFHIP_ERR_ARG, "npoints and ndens must be >= 0"};
  // cart_gto3d_rel (grid.f90:17-110) covers L <= 4 and returns NaN above
  HostShells s = parse_shells(b, 4, "basis");
  std::vector<int> c0(s.n);
  int ncart = 0;
  for (int i = 0; i < s.n; i++) {
    c0[i] = ncart;
    ncart += (s.L[i] + 1) * (s.L[i] + 2) / 2;
  }
  if (nbfs != ncart)
    throw Error{DFHIP_ERR_ARG, "nbfs=" + std::to_string(nbfs) + " but the basis has " + std::to_string(ncart) +
                                   " Cartesian functions"};
  if (npoints == 0 || ndens == 0) return;
  if (!grid3d || !dens) throw Error{DFHIP_ERR_ARG, "grid3d or densities is NULL"};
  std::lock_guard<std::mutex> lock(ctx_mutex());
  ensure_ctx();
  cur_lane().arena.reset();
  const hipStream_t st = cur_lane().st;
  DBuf<double> dxyz, dexp, dcoef, dpts, dD, dDd, drho;
  DBuf<int> dL, dnp, dpp, dc0;
  upload(dxyz, s.xyz, st);
  upload(dexp, s.exps, st);
  upload(dcoef, s.coefs, st);
  upload(dL, s.L, st);
  upload(dnp, s.nprim, st);
  upload(dpp, s.pptr, st);
  upload(dc0, c0, st);
  dpts.alloc((size_t)3 * npoints);
  HC(hipMemcpyAsync(dpts.p, grid3d, dpts.n * sizeof(double), hipMemcpyHostToDevice, st));
  const size_t nd = (size_t)ndens * ncart * ncart;
  dD.alloc(nd);
  dDd.alloc(nd);
  HC(hipMemcpyAsync(dD.p, dens, nd * sizeof(double), hipMemcpyHostToDevice, st));
  grid_transpose_densities(dD.p, ndens, ncart, dDd.p, st);
  HC(hipGetLastError());
  HC(hipStreamSynchronize(st));
  dD.release();
  drho.alloc((size_t)npoints * ndens);
  const GridShells gs{dxyz.p, dexp.p, dcoef.p, dL.p, dnp.p, dpp.p, dc0.p, s.n, ncart};
  const int64_t nblk = (npoints + bs - 1) / bs;
  const size_t per_blk = (size_t)ncart * bs;
  const int64_t bmax = std::max<int64_t>(1, std::min<int64_t>(nblk, ((size_t)1 << 28) / per_blk));
  DBuf<double> dX((size_t)bmax * per_blk), dm((size_t)bmax * ncart);
  for (int64_t b0 = 0; b0 < nblk; b0 += bmax) {
    const int nb = (int)std::min(bmax, nblk - b0);
    const int64_t p0 = b0 * bs, np = std::min<int64_t>((int64_t)nb * bs, npoints - p0);
    grid_ao_values(gs, dpts.p, p0, np, bs, dX.p, st);
    grid_block_means(dX.p, nb, ncart, bs, np, dm.p, st);
    grid_block_densities(dX.p, dm.p, dDd.p, ndens, ncart, nb, bs, p0, np, thresh, drho.p, st);
    HC(hipGetLastError());
  }
  HC(hipMemcpyAsync(out, drho.p, drho.n * sizeof(double), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
}

}  // namespace
}  // namespace dfhip

using namespace dfhip;

extern "C" {

void f_init(void) {
  guarded([&] {
    std::lock_guard<std::mutex> lock(ctx_mutex());
    ensure_ctx();
  });
}

int dfhip_last_status(void) { return t_status; }

int dfhip_release_workspace(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lock(ctx_mutex());
    Ctx& c = ctx();
    c.lane0.release();
    for (auto& ln : c.lanes) ln->release();
    dev_pool().release();
  });
}
const char* dfhip_last_error(void) { return t_msg.c_str(); }

void dfhip_default_options(dfhip_options* opt) {
  if (opt) *opt = defaults();
}

int dfhip_contract_coulomb_densities(const dfhip_basis* bas, const dfhip_basis* aux, int32_t ndens, int32_t nbfs,
                                     const double* densities, const dfhip_options* opt, double* df_tensor,
                                     int32_t nstates, double* coulomb_tensor, dfhip_stats* stats) {
  return guarded([&] {
    if (!bas || !aux) throw Error{DFHIP_ERR_ARG, "basis is NULL"};
    const dfhip_options o = opt ? *opt : defaults();
    Outputs out;
    out.df = df_tensor;
    out.J = coulomb_tensor;
    out.nstates = nstates;
    run(*bas, *aux, ndens, nbfs, densities, o, out, stats);
  });
}

int dfhip_contract_coulomb_densities_batch(int32_t nmol, const dfhip_basis* bas, const dfhip_basis* aux,
                                           const int32_t* ndens, const int32_t* nbfs,
                                           const double* const* densities, const dfhip_options* opt,
                                           const int32_t* nstates, double* const* coulomb_tensors,
                                           dfhip_stats* stats, int32_t nlanes) {
  return guarded([&] {
    if (nmol < 0) throw Error{DFHIP_ERR_ARG, "nmol < 0"};
    if (nmol == 0) return;
    if (!bas || !aux || !ndens || !nbfs || !densities || !nstates || !coulomb_tensors)
      throw Error{DFHIP_ERR_ARG, "NULL argument"};
    const dfhip_options o = opt ? *opt : defaults();
    std::lock_guard<std::mutex> lock(ctx_mutex());
    ensure_ctx();
    Ctx& c = ctx();
    const int nl = std::max(1, std::min({nlanes > 0 ? (int)nlanes : 4, 8, (int)nmol}));
    while ((int)c.lanes.size() < nl - 1) {
      auto ln = std::make_unique<Lane>();
      HC(hipStreamCreateWithFlags(&ln->st, hipStreamNonBlocking));
      c.lanes.push_back(std::move(ln));
    }
    std::atomic<int> next(0);
    std::mutex emu;
    int ecode = DFHIP_OK;
    std::string emsg;
    auto worker = [&](int w) {
      t_lane = w == 0 ? &c.lane0 : c.lanes[w - 1].get();
      t_pool = w == 0 ? nullptr : &t_lane->pool;
      t_few_streams = nl > 1;
      t_lanes = nl;
      struct Reset {
        ~Reset() {
          t_few_streams = false;
          t_lanes = 1;
        }
      } reset_streams;
      if (hipSetDevice(c.device) != hipSuccess) {
        std::lock_guard<std::mutex> g(emu);
        if (ecode == DFHIP_OK) {
          ecode = DFHIP_ERR_HIP;
          emsg = "hipSetDevice failed in a batch worker";
        }
        return;
      }
      for (int k = next++; k < nmol; k = next++) {
        try {
          Outputs out;
          out.J = coulomb_tensors[k];
          out.nstates = nstates[k];
          run(bas[k], aux[k], ndens[k], nbfs[k], densities[k], o, out, stats ? &stats[k] : nullptr, false);
        } catch (const Error& e) {
          std::lock_guard<std::mutex> g(emu);
          if (ecode == DFHIP_OK) {
            ecode = e.code;
            emsg = "molecule " + std::to_string(k) + ": " + e.msg;
          }
          next = nmol;
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> g(emu);
          if (ecode == DFHIP_OK) {
            ecode = DFHIP_ERR_HIP;
            emsg = "molecule " + std::to_string(k) + ": " + e.what();
          }
          next = nmol;
        }
      }
      t_lane = nullptr;
      t_pool = nullptr;
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nl; w++) th.emplace_back(worker, w);
    worker(0);
    for (auto& t : th) t.join();
    if (ecode != DFHIP_OK) throw Error{ecode, emsg};
  });
}

int dfhip_int3c2e_dense(const dfhip_basis* bas, const dfhip_basis* aux, const dfhip_options* opt, double* out3,
                        dfhip_stats* stats) {
  return guarded([&] {
    if (!bas || !aux || !out3) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    const dfhip_options o = opt ? *opt : defaults();
    Outputs out;
    out.dense3c = out3;
    run(*bas, *aux, 0, 0, nullptr, o, out, stats);
  });
}

int dfhip_int_schwarz_aux_noalloc(const dfhip_basis* aux, double* integrals) {
  return guarded([&] {
    if (!aux || !integrals) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    schwarz_aux_noalloc(*aux, integrals);
  });
}

int dfhip_metric_powerh(const dfhip_basis* aux, double power, double* out) {
  return guarded([&] {
    if (!aux || !out) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    metric_powerh_host(*aux, power, out);
  });
}

int dfhip_int3c2e_estimates(const dfhip_basis* bas, const dfhip_basis* aux, double* records, int64_t capacity,
                            int64_t* nrecords) {
  return guarded([&] {
    if (!bas || !aux || !nrecords) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    const Estimates es = int3c2e_estimates(*bas, *aux);
    const int64_t n = (int64_t)es.rec.size() / 4;
    *nrecords = n;
    if (records) {
      if (capacity < n)
        throw Error{DFHIP_ERR_ARG, "records holds " + std::to_string(capacity) + " records, " + std::to_string(n) +
                                       " are needed"};
      std::memcpy(records, es.rec.data(), es.rec.size() * sizeof(double));
    }
  });
}

void f_benchmark_int3c2e(int32_t nshells, int32_t ndata, const int32_t* bas_centers, const int32_t* bas_spec,
                         const double* bas_data, int32_t nshells_aux, int32_t ndata_aux,
                         const int32_t* bas_centers_aux, const int32_t* bas_spec_aux, const double* bas_data_aux) {
  guarded([&] {
    const dfhip_basis b = mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data);
    const dfhip_basis x = mkbasis(nshells_aux, ndata_aux, bas_centers_aux, bas_spec_aux, bas_data_aux);
    const Estimates es = int3c2e_estimates(b, x);
    std::printf("Main basis:%6d shells, %6d basis functions\n", es.nshells, es.nbf);
    std::printf("Aux. basis:%6d shells, %6d basis functions\n", es.nshells_aux, es.naux);
    std::printf("Initializing the DF-integral screener took %6.2f s\n", es.t_screener);
    std::printf("Calculation of density fitting metric (P|Q) took%8.4f s.\n", es.t_metric);
    std::printf("Calculation of (P|Q)**-0.5 took%8.4f s.\n", es.t_powerh);
    std::printf("Density contraction took%8.4f s.\n", es.t_ints);
    std::printf("Skipped %12lld of %12lld integral triplets (%8.2f%%)\n", (long long)es.nskipped,
                (long long)es.ntriplets, 100.0 * (double)es.nskipped / (double)std::max<int64_t>(es.ntriplets, 1));
    std::fflush(stdout);
    // intor/benchmark.f90:43,117: a stream of float64 records in the working directory
    FILE* f = std::fopen("int_estimates.bin", "wb");
    if (!f) throw Error{DFHIP_ERR_ARG, "cannot open int_estimates.bin for writing"};
    const size_t nw = std::fwrite(es.rec.data(), sizeof(double), es.rec.size(), f);
    std::fclose(f);
    if (nw != es.rec.size()) throw Error{DFHIP_ERR_ARG, "short write to int_estimates.bin"};
  });
}

int dfhip_qvl_records(const dfhip_basis* bas, const dfhip_basis* aux, double* pair_records, double* aux_records) {
  return guarded([&] {
    if (!bas || !aux || !pair_records || !aux_records) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    qvl_records_host(*bas, *aux, pair_records, aux_records);
  });
}

int dfhip_schwarz_bounds(const dfhip_basis* bas, const dfhip_basis* aux, double* pair_bounds, double* aux_bounds) {
  return guarded([&] {
    if (!bas || !aux || !pair_bounds || !aux_bounds) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    const HostShells o = parse_shells(*bas, LMAX_ORB, "orbital basis");
    const HostShells x = parse_shells(*aux, LMAX_AUX, "auxiliary basis");
    const DevShells v = host_view(o);
    for (int a = 0; a < o.n; a++)
      for (int b = 0; b <= a; b++) pair_bounds[(size_t)a * (a + 1) / 2 + b] = schwarz_pair_bound(v, a, b, sph_abs_sums());
    for (int sh = 0; sh < x.n; sh++) aux_bounds[sh] = aux_schwarz_bound(x, sh);
  });
}

int dfhip_metric(const dfhip_basis* aux, double* out) {
  return guarded([&] {
    if (!aux || !out) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    std::vector<double> M = metric_host(*aux, nullptr);
    std::memcpy(out, M.data(), M.size() * sizeof(double));
  });
}

int dfhip_boys(int64_t npts, const int32_t* n, const double* T, double* F) {
  return guarded([&] {
    if (npts < 0 || (npts > 0 && (!n || !T || !F))) throw Error{DFHIP_ERR_ARG, "bad arguments"};
    for (int64_t i = 0; i < npts; i++)
      if (n[i] < 0 || n[i] > BOYS_NORD - BOYS_NTAYLOR || !(T[i] >= 0.0))
        throw Error{DFHIP_ERR_ARG, "order or argument out of range"};
    std::lock_guard<std::mutex> lock(ctx_mutex());
    ensure_ctx();
    Ctx& c = ctx();
    DBuf<int> dn(std::max<int64_t>(npts, 1));
    DBuf<double> dT(std::max<int64_t>(npts, 1)), dF(std::max<int64_t>(npts, 1));
    if (npts > 0) {
      HC(hipMemcpyAsync(dn.p, n, npts * sizeof(int), hipMemcpyHostToDevice, cur_lane().st));
      HC(hipMemcpyAsync(dT.p, T, npts * sizeof(double), hipMemcpyHostToDevice, cur_lane().st));
      boys_probe(dn.p, dT.p, npts, c.boys.p, dF.p, cur_lane().st);
      HC(hipGetLastError());
      HC(hipMemcpyAsync(F, dF.p, npts * sizeof(double), hipMemcpyDeviceToHost, cur_lane().st));
    }
    HC(hipStreamSynchronize(cur_lane().st));
  });
}

void f_contract_coulomb_densities_2d(int32_t nshells, int32_t ndata, const int32_t* bas_centers,
                                     const int32_t* bas_spec, const double* bas_data, int32_t nshells_aux,
                                     int32_t ndata_aux, const int32_t* bas_centers_aux, const int32_t* bas_spec_aux,
                                     const double* bas_data_aux, int32_t ndens, int32_t nbfs, const double* densities,
                                     int32_t naux, double* df_tensor) {
  guarded([&] {
    const dfhip_basis b = mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data);
    const dfhip_basis x = mkbasis(nshells_aux, ndata_aux, bas_centers_aux, bas_spec_aux, bas_data_aux);
    int nsph = 0;
    for (int i = 0; i < nshells_aux; i++) nsph += 2 * bas_spec_aux[i + nshells_aux] + 1;
    if (naux != nsph)
      throw Error{DFHIP_ERR_ARG, "naux=" + std::to_string(naux) + " but the auxiliary basis has " +
                                     std::to_string(nsph) + " spherical functions"};
    const dfhip_options o = defaults();
    Outputs out;
    out.df = df_tensor;
    run(b, x, ndens, nbfs, densities, o, out, nullptr);
  });
}

void f_contract_coulomb_densities_4d(int32_t nshells, int32_t ndata, const int32_t* bas_centers,
                                     const int32_t* bas_spec, const double* bas_data, int32_t nshells_aux,
                                     int32_t ndata_aux, const int32_t* bas_centers_aux, const int32_t* bas_spec_aux,
                                     const double* bas_data_aux, int32_t ndens, int32_t nbfs, const double* densities,
                                     int32_t nstates, double* coulomb_tensor) {
  guarded([&] {
    const dfhip_basis b = mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data);
    const dfhip_basis x = mkbasis(nshells_aux, ndata_aux, bas_centers_aux, bas_spec_aux, bas_data_aux);
    const dfhip_options o = defaults();
    Outputs out;
    out.J = coulomb_tensor;
    out.nstates = nstates;
    run(b, x, ndens, nbfs, densities, o, out, nullptr);
  });
}

void f_intor_df1c(int32_t nshells, int32_t ndata, const int32_t* bas_centers, const int32_t* bas_spec,
                  const double* bas_data, int32_t nshells_aux, int32_t ndata_aux, const int32_t* bas_centers_aux,
                  const int32_t* bas_spec_aux, const double* bas_data_aux, int32_t nbfs, int32_t nbfs_aux,
                  double* df1c_tensor) {
  guarded([&] {
    const dfhip_basis b = mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data);
    const dfhip_basis x = mkbasis(nshells_aux, ndata_aux, bas_centers_aux, bas_spec_aux, bas_data_aux);
    if (!df1c_tensor) throw Error{DFHIP_ERR_ARG, "df1c_tensor is NULL"};
    if (nbfs != basis_nsph(b) || nbfs_aux != basis_nsph(x))
      throw Error{DFHIP_ERR_ARG, "nbfs / nbfs_aux do not match the spherical functions of the bases"};
    std::vector<double> B((size_t)nbfs_aux * nbfs);
    Outputs out;
    out.df = B.data();
    std_tensors(b, x, 0, 0, nullptr, out);
    for (int64_t mu = 0; mu < nbfs; mu++)
      for (int64_t P = 0; P < nbfs_aux; P++) df1c_tensor[mu + (int64_t)nbfs * P] = B[P + (int64_t)nbfs_aux * mu];
  });
}

void f_intor_eri2c(int32_t nshells, int32_t ndata, const int32_t* bas_centers, const int32_t* bas_spec,
                   const double* bas_data, int32_t nshells_aux, int32_t ndata_aux, const int32_t* bas_centers_aux,
                   const int32_t* bas_spec_aux, const double* bas_data_aux, int32_t nbfs, double* eri2c_tensor) {
  guarded([&] {
    const dfhip_basis b = mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data);
    const dfhip_basis x = mkbasis(nshells_aux, ndata_aux, bas_centers_aux, bas_spec_aux, bas_data_aux);
    if (!eri2c_tensor) throw Error{DFHIP_ERR_ARG, "eri2c_tensor is NULL"};
    if (nbfs != basis_nsph(b)) throw Error{DFHIP_ERR_ARG, "nbfs does not match the spherical functions of the basis"};
    Outputs out;
    out.gram = eri2c_tensor;
    std_tensors(b, x, 0, 0, nullptr, out);
  });
}

void f_intor_df2c_mo(int32_t nshells, int32_t ndata, const int32_t* bas_centers, const int32_t* bas_spec,
                     const double* bas_data, int32_t nshells_aux, int32_t ndata_aux, const int32_t* bas_centers_aux,
                     const int32_t* bas_spec_aux, const double* bas_data_aux, int32_t nbfs, int32_t nmos,
                     const double* mo_tensor, int32_t nbfs_aux, double* df2c_mo_tensor) {
  guarded([&] {
    const dfhip_basis b = mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data);
    const dfhip_basis x = mkbasis(nshells_aux, ndata_aux, bas_centers_aux, bas_spec_aux, bas_data_aux);
    if (!df2c_mo_tensor || !mo_tensor) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    if (nbfs != basis_nsph(b) || nbfs_aux != basis_nsph(x))
      throw Error{DFHIP_ERR_ARG, "nbfs / nbfs_aux do not match the spherical functions of the bases"};
    const int64_t npack = (int64_t)nmos * (nmos + 1) / 2;
    std::vector<double> B((size_t)nbfs_aux * npack);
    Outputs out;
    out.df = B.data();
    std_tensors(b, x, 1, nmos, mo_tensor, out);
    for (int64_t pk = 0; pk < npack; pk++)
      for (int64_t P = 0; P < nbfs_aux; P++) df2c_mo_tensor[pk + npack * P] = B[P + (int64_t)nbfs_aux * pk];
  });
}

void f_int_schwarz_aux(int32_t nshells, int32_t ndata, const int32_t* bas_centers, const int32_t* bas_spec,
                       const double* bas_data, double* integrals) {
  guarded([&] {
    if (!integrals) throw Error{DFHIP_ERR_ARG, "integrals is NULL"};
    std::vector<double> qc = schwarz_aux_host(mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data));
    std::memcpy(integrals, qc.data(), qc.size() * sizeof(double));
  });
}

void f_int_schwarz_bra(int32_t nshells, int32_t ndata, const int32_t* bas_centers, const int32_t* bas_spec,
                       const double* bas_data, double* integrals) {
  guarded([&] {
    std::vector<double> q = schwarz_bra_host(mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data));
    std::memcpy(integrals, q.data(), q.size() * sizeof(double));
  });
}

int dfhip_autotune(const dfhip_basis* bas, const dfhip_basis* aux, int32_t ndens, int32_t nbfs,
                   const double* densities, const dfhip_options* opt, const char* path) {
  return guarded([&] {
    if (!bas || !aux || !densities) throw Error{DFHIP_ERR_ARG, "NULL argument"};
    const dfhip_options o = opt ? *opt : defaults();
    Tuning t = autotune_host(*bas, *aux, ndens, nbfs, densities, o);
    if (path && *path) tuning_save(t, path);
    std::lock_guard<std::mutex> g(tuning_mutex());
    tuning() = t;
  });
}

int dfhip_load_tuning(const char* path) {
  return guarded([&] {
    Tuning t = (path && *path) ? tuning_load(path) : Tuning();
    std::lock_guard<std::mutex> g(tuning_mutex());
    tuning() = t;
  });
}

int32_t dfhip_tuning_choice(int32_t la, int32_t lb, int32_t lc, int32_t window) {
  if (la < 0 || la > LMAX_ORB || lb < 0 || lb > la || lc < 0 || lc > LMAX_AUX || window < 0 || window > 1) return -1;
  std::lock_guard<std::mutex> g(tuning_mutex());
  return tuning().fam[pcls(la, lb)][lc][window];
}

int32_t dfhip_tuning_contract(int32_t ndens, int32_t* ipt, double* split_scale) {
  std::lock_guard<std::mutex> g(tuning_mutex());
  const Tuning::Contract* c = tuning().contract_for(ndens);
  if (!c) return 0;
  if (ipt) *ipt = c->ipt;
  if (split_scale) *split_scale = c->nscale;
  return 1;
}

void f_eval_densities(int32_t nshells, int32_t ndata, const int32_t* bas_centers, const int32_t* bas_spec,
                      const double* bas_data, int32_t npoints, const double* grid3d, int32_t ndens, int32_t nbfs,
                      const double* densities, double* grid_densities, int32_t blk_size, double thresh) {
  guarded([&] {
    eval_densities_host(mkbasis(nshells, ndata, bas_centers, bas_spec, bas_data), npoints, grid3d, ndens, nbfs,
                        densities, grid_densities, blk_size, thresh);
  });
}

}  // extern "C"
