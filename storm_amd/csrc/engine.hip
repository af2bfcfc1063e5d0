// Host orchestration of one IMC step on one MI355X and the C ABI
// (include/storm_gpu.h).  One engine = one HIP stream; not re-entrant,
// like the reference's serial manager (MonteCarloManagerSerial.hpp:269-270
// keeps function-static buffers).
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <type_traits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "../../include/storm_gpu.h"
#include "../../include/storm_mesh.h"
#include "kernels.h"
#include "launch_policy.h"
#include "kernels_unstructured.h"
#include "common/umesh_tables.h"
#include "voronoi_build.h"

namespace storm {
namespace gpu {

thread_local std::string g_last_error;

struct EngineError : std::runtime_error {
  int code;
  EngineError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) throw EngineError(2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// While an engine steps, the buffers it outgrows are kept (RetiredBuffers)
// instead of freed: hipFree waits for all work on the device and unmapping a
// large buffer costs ~0.3 ms, so a census growing through a power of two
// would stall the step for milliseconds.  Buffers grow by doubling, so what
// is kept is less than what is in use.
struct Retired {
  void *p;
  size_t bytes;
};
thread_local std::vector<Retired> *t_retired = nullptr;

// Scratch of storm_gpu_step_noalloc: the step's transient buffers (packets
// in flight, the comb's keys and prefix sums, rocPRIM's temporary storage,
// reduction partials, counters and read-back mailboxes) are carved out of the
// file-scope device array `scratch` below instead of device allocations.
// While a ScratchArena is installed on the thread, DevBuf::ensure of a
// transient buffer takes memory from it, and a buffer of the engine's state
// that would have to grow throws instead of allocating.
struct ScratchArena {
  unsigned char *base;
  size_t size;
  size_t used;
  void *take(size_t bytes) {
    const size_t at = (used + 255u) & ~size_t(255);
    if (at > size || bytes > size - at)
      throw EngineError(4, "storm_gpu_step_noalloc: this step needs more than the library's fixed scratch of " +
                               std::to_string(size >> 20) + " MiB (too many packets in one step); use storm_gpu_step");
    used = at + bytes;
    return base + at;
  }
};
thread_local ScratchArena *t_arena = nullptr;

// The scratch array itself: one copy per GPU, held by every process that
// loads this library until it ends.  Host code never passes `scratch` to a
// kernel (on the host it only names the array); it uses the device address
// that hipGetSymbolAddress returns for the current device, looked up once per
// device when an engine is created there.  All engines on one GPU share the
// array: storm_gpu_step_noalloc calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxScratchDevices] = {};

void *noalloc_scratch_address() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) throw EngineError(2, "device index beyond the scratch address table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw EngineError(2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = ptr;
  }
  return g_scratch_dev[dev];
}

// A buffer's allocation while it is lent to a ScratchArena.
struct SavedBuf {
  void *p;
  size_t n;
};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  bool view = false;  // points into a ScratchArena (never freed here)
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && !view) {
      if (t_retired) t_retired->push_back(Retired{p, n * sizeof(T)});
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  // Hand the allocation to `saved` and become an empty view into the
  // installed ScratchArena; give_back restores the allocation.
  SavedBuf lend() {
    SavedBuf sb{p, n};
    p = nullptr;
    n = 0;
    view = true;
    return sb;
  }
  void give_back(const SavedBuf &sb) {
    p = static_cast<T *>(sb.p);
    n = sb.n;
    view = false;
  }
  // Grow without keeping the contents, to a power of two at least twice the
  // old capacity: a census that grows step by step then reallocates a few
  // times only (hipFree waits for the device and costs about a millisecond).
  void ensure(size_t m) {
    if (m <= n && p) return;
    if (t_arena) {
      if (!view)
        throw EngineError(4, "storm_gpu_step_noalloc: a buffer of the engine's state must grow to " +
                                 std::to_string(m) + " elements (capacity " + std::to_string(n) +
                                 "; the census outgrew what earlier steps allocated); use storm_gpu_step");
      p = static_cast<T *>(t_arena->take(m * sizeof(T)));
      n = m;
      return;
    }
    size_t cap = 1;
    while (cap < m || cap < 2 * n) cap <<= 1;
    release();
    HIPCHECK(hipMalloc(reinterpret_cast<void **>(&p), cap * sizeof(T)));
    n = cap;
  }
};

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

constexpr uint64_t kMaxSegmentsPerStep = uint64_t(1) << 24;

// The random-walk tables in double, built once per process (the engine
// converts them to its precision on upload).
static const std::vector<double> &rw_host_tables() {
  static const std::vector<double> tb = [] {
    std::vector<double> t(rw::kAllSize);
    rw::build_all_tables(t.data());
    return t;
  }();
  return tb;
}

template <typename Real>
struct PacketBufs {
  DevBuf<typename V4<Real>::T> pos, dir;
  DevBuf<double> w;
  DevBuf<uint4> meta;
  DevBuf<Real> nu;
  bool with_nu = false;
  void ensure(size_t n) {
    pos.ensure(n);
    dir.ensure(n);
    w.ensure(n);
    meta.ensure(n);
    if (with_nu) nu.ensure(n);
  }
  PacketSoA<Real> soa() { return PacketSoA<Real>{pos.p, dir.p, w.p, meta.p, with_nu ? nu.p : nullptr}; }
  size_t cap() const { return meta.n; }
};

template <typename Real>
struct CensusBufs {
  DevBuf<typename V4<Real>::T> pos, dir;
  DevBuf<double> w;
  DevBuf<uint2> meta;
  void ensure(size_t n) {
    pos.ensure(n);
    dir.ensure(n);
    w.ensure(n);
    meta.ensure(n);
  }
  CensusSoA<Real> soa() { return CensusSoA<Real>{pos.p, dir.p, w.p, meta.p}; }
  size_t cap() const { return meta.n; }
};

// ---------------------------------------------------------------------------
struct EngineBase : LaunchPolicy {
  virtual ~EngineBase() = default;
  virtual void step(double dt) = 0;
  virtual void step_noalloc(double dt) = 0;
  virtual void get_cells(double *, double *, double *, double *, double *, double *) = 0;
  virtual void get_group_erad(double *) = 0;
  virtual void get_segments(uint64_t *) = 0;
  virtual void get_begin(uint64_t *) = 0;
  virtual void get_ledger(storm_gpu_ledger *) = 0;
  virtual void get_instance_ledger(uint32_t, storm_gpu_ledger *) = 0;
  virtual void set_cell_state(const double *, const double *, const double *, const double *, const double *) = 0;
  virtual uint64_t census_size() = 0;
  virtual void export_census(storm_packet_record *, uint64_t, uint64_t *) = 0;
  virtual void import_census(const storm_packet_record *, uint64_t) = 0;
  virtual void queue_source_packets(const storm_packet_record *, const double *, uint64_t) = 0;
  virtual void debug_inputs(uint64_t *, void *, void *, int64_t *) = 0;
  virtual void debug_tallies(uint64_t *, uint64_t *, uint64_t *, uint64_t *, uint32_t *) = 0;
  virtual void debug_accel(void *, uint64_t *) = 0;
  virtual void debug_spectrum(int32_t *, double *, void *, void *, void *) = 0;
  storm_gpu_counters counters{};
  double source_T = 0.0;
  int32_t seg_per_pass = 1 << 20;
  bool keep_inputs = false;
};

template <typename Real>
struct Engine final : EngineBase {
  // problem (host copies)
  int32_t nx, ny, nz;
  uint32_t N;
  Morton morton{};  // Z-order of the cells (the combed census order)
  std::vector<double> xe, ye, ze, Vh, sa0h, ss0h, rhoh;
  storm_material_desc matd;
  storm_boundary_desc bnd;
  storm_imc_params imc;
  uint64_t seed;
  int32_t precision, material_mode;
  bool frozen = false;
  double frozenT4s[2] = {0.0, 0.0};
  double source_T_hi = 0.0;
  uint32_t step_index = 0;
  int32_t eq_prev = 0;  // quantum exponent of the resident census weights
  double census_total_prev = 0.0;  // erg
  uint64_t n_census = 0;           // census resident in `cnext`

  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // device: mesh and cells
  DevBuf<Real> widths;  // MeshView::width
  DevBuf<double> V, sa0, ss0, rho, T, e_mat, e_sync, fleck, sig_a, sig_s, Eemit, wemit, erad_avg, census_erad,
      clamp, partials, host_sa, host_ss, host_cv, wface, dsc;
  bool have_host_sa = false, have_host_ss = false, have_host_cv = false;
  DevBuf<CellXS<Real>> xs;
  DevBuf<uint32_t> nphot, nphot_z, err, run_start, run_end, ncopy, keys, keys_out, counters_u32;
  DevBuf<uint64_t> offsets, offs64, nphot64;
  DevBuf<uint32_t> vals, vals_out;  // comb: census indices sorted with the keys
  DevBuf<unsigned long long> tally_u64, emit_limbs, requant, leaving;
  DevBuf<double> tally_f64;
  // per-XCD replicas 1..kTallyReplicas-1 of the per-visit tallies
  DevBuf<unsigned long long> rep_u64;
  DevBuf<double> rep_f64;
  DevBuf<uint32_t> rep_u32;
  DevBuf<uint64_t> limbs_sorted, cum;  // sorted census weights in quanta, their prefix sums
  DevBuf<double> wsorted;              // sorted census weights
  DevBuf<CombCell> comb_cells;
  DevBuf<int32_t> qexp;
  // scalars read back by the host: gathered on the device, one copy into
  // pinned memory, one synchronisation
  DevBuf<unsigned long long> mail;
  unsigned long long *mail_h = nullptr;
  // pinned staging of the boundary face weights (both x sides, sized at creation)
  double *wface_h = nullptr;
  // host staging of the boundary sources' group CDFs (capacity reserved at creation)
  std::vector<Real> bcdf_h;
  std::vector<double> planck_acc;
  DevBuf<CellAccel<Real>> accel;
  DevBuf<double> Dtil;
  DevBuf<Real> rw_tables;  // rw::tables_at's layout
  bool accel_on = false;
  // multigroup (storm_group_desc)
  int32_t groups = 0;
  uint32_t strat_groups = 1, strat_gbits = 0;  // stratified comb: groups per cell and their key bits
  bool eg_on = false;  // withEgTimeAvg in a multigroup problem
  DevBuf<unsigned long long> eg_u64;  // per-group path integral, mixed engine
  DevBuf<double> eg_f64;              // FP64 engine
  double nu_pow = 0.0, nu_ref = 1.0;
  std::vector<double> bounds_h;  // erg
  DevBuf<double> bounds_erg;
  DevBuf<Real> bounds_kev, cdf, sig_s_cell, bcdf;
  DevBuf<unsigned char> temp;
  Tallies tal{};

  PacketBufs<Real> pa, pb;
  CensusBufs<Real> cpre, cnext;
  CensusBufs<Real> cstage;  // k_transport_refill: census staged per block (its input range)

  // polyhedral meshes (storm_unstructured_mesh): host tables and their device copies
  bool unstructured = false;
  double box_lo[3] = {0, 0, 0}, box_hi[3] = {1, 1, 1};
  umesh::Tables<Real> ut;
  std::vector<double> gen_h;  // mesh points, 3 per cell
  std::vector<double> chord_h;  // mean chord 4 V / S per cell
  DevBuf<double> chord;
  DevBuf<uint32_t> u_face_start, u_tet_start;
  DevBuf<UFace<Real>> u_face;
  DevBuf<Real> u_face_centroid, u_ddmc_rate;
  DevBuf<double2> u_face_geo;
  DevBuf<UTet<Real>> u_tets;
  DevBuf<USource> u_src;
  DevBuf<UTri<Real>> u_tris;
  // batches of independent instances (cell_instance)
  std::vector<storm_instance_desc> inst;
  std::vector<uint32_t> cell_inst_h;
  std::vector<double> inst_T4;  // sources' T^4 per instance (frozen at the first step if asked)
  std::vector<uint32_t> inst_first_h;  // first cell of each instance, and N
  DevBuf<uint32_t> d_cell_inst, d_inst_first, d_inst_ncells;
  DevBuf<uint64_t> d_inst_nphot, d_inst_nmin;
  DevBuf<double> d_inst_factor, d_inst_etot, d_inst_wtot;
  // per instance: boundary source total of the step, census energy after the
  // last step, material energy before this one (erg); unit exponents of this
  // step and the last (device, and host copies for the source weights, the
  // census export and the ledgers)
  DevBuf<double> d_inst_src, d_inst_cen, d_inst_mat0;
  DevBuf<int32_t> d_inst_shift, d_inst_shift_prev;
  double *inst_src_h = nullptr;    // pinned staging, K
  int32_t *inst_shift_h = nullptr; // pinned staging, K
  std::vector<int32_t> inst_shift_now, inst_shift_last;
  InstanceView instview() const {
    InstanceView v{};
    if (inst.empty()) return v;
    v.cell = d_cell_inst.p;
    v.ncells = d_inst_ncells.p;
    v.nphot = d_inst_nphot.p;
    v.nmin = d_inst_nmin.p;
    v.factor = d_inst_factor.p;
    v.etot = d_inst_etot.p;
    v.wtot = d_inst_wtot.p;
    v.shift = d_inst_shift.p;
    v.shift_prev = d_inst_shift_prev.p;
    return v;
  }
  uint32_t last_nfaces_ = 0;    // source faces of the last step
  // host-made source packets queued for the next step (weights in erg) and,
  // for the ledger, the last step's ones on the step's quantum
  std::vector<HostPacket> host_queue_;
  std::vector<std::pair<uint32_t, double>> host_last_;  // (cell, weight on the quantum)
  DevBuf<HostPacket> host_pk;

  // kept transport inputs of the last step (debug / replay)
  std::vector<Packet<Real>> kept_packets;
  std::vector<CellXS<Real>> kept_xs;
  StepConsts kept_consts{};

  Engine(const storm_problem *p, int32_t prec, int32_t mmode) : precision(prec), material_mode(mmode) {
    if (p->unstructured) {
      // polyhedral cells, numbered in Z order already: the per-cell kernels
      // run over a flat nx = N x 1 x 1 layout and packets carry cell indices
      const storm_unstructured_mesh &um = *p->unstructured;
      unstructured = true;
      if (um.ncells < 1 || um.ncells > (1u << 30)) throw EngineError(1, "unstructured mesh: 1 .. 2^30 cells");
      nx = static_cast<int32_t>(um.ncells);
      ny = nz = 1;
      N = um.ncells;
      uint32_t b = 0;
      while ((1ull << b) < N) ++b;
      morton.bits[0] = b;
      morton.bits[1] = morton.bits[2] = 0;
      morton.total = b;
      morton.flat = 1;
      if (morton.total > 31) throw EngineError(1, "mesh too large for 32-bit Z-order cell codes");
      // axis "edges" only for the engine's Cartesian bookkeeping (unused)
      xe = {0.0, 1.0};
      ye = {0.0, 1.0};
      ze = {0.0, 1.0};
      for (int a = 0; a < 3; ++a) {
        box_lo[a] = um.box_lo[a];
        box_hi[a] = um.box_hi[a];
      }
    } else {
      nx = p->mesh.nx;
      ny = p->mesh.ny;
      nz = p->mesh.nz;
      if (nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 1024 || nz > 1024)
        throw EngineError(1, "mesh dimensions must be 1..4096 x 1..1024 x 1..1024");
      // cell indices stay below 2^31 in every tally offset (2 c + 1)
      if (static_cast<uint64_t>(nx) * ny * nz > (1ull << 30)) throw EngineError(1, "more than 2^30 cells");
      N = static_cast<uint32_t>(nx) * ny * nz;
      {
        const int32_t n3[3] = {nx, ny, nz};
        morton.total = 0;
        morton.flat = 0;
        for (int a = 0; a < 3; ++a) {
          uint32_t b = 0;
          while ((1ll << b) < n3[a]) ++b;
          morton.bits[a] = b;
          morton.total += b;
        }
        // the comb's keys hold the code and one more value (dropped packets)
        if (morton.total > 31) throw EngineError(1, "mesh too large for 32-bit Z-order cell codes");
      }
      xe.assign(p->mesh.x_edges, p->mesh.x_edges + nx + 1);
      ye.assign(p->mesh.y_edges, p->mesh.y_edges + ny + 1);
      ze.assign(p->mesh.z_edges, p->mesh.z_edges + nz + 1);
      for (auto *e : {&xe, &ye, &ze})
        for (size_t q = 1; q < e->size(); ++q)
          if (!((*e)[q] > (*e)[q - 1])) throw EngineError(1, "mesh edges must be strictly increasing");
    }
    matd = p->material;
    bnd = p->boundary;
    imc = p->imc;
    seed = p->seed;
    if (imc.new_photons_per_cell == 0) throw EngineError(1, "RadiationIMC requires newPhotonsPerCell > 0");
    if (imc.comb < 0 || imc.comb > 2) throw EngineError(1, "imc.comb must be 0, 1 or 2");
    accel_on = imc.with_random_walk || imc.with_ddmc;
    if (p->groups.num_groups >= 2) {
      groups = p->groups.num_groups;
      if (groups > 4096) throw EngineError(1, "at most 4096 groups");
      if (!p->groups.bounds) throw EngineError(1, "groups.bounds is required");
      bounds_h.assign(p->groups.bounds, p->groups.bounds + groups + 1);
      for (int32_t g = 0; g < groups; ++g)
        if (!(bounds_h[g + 1] > bounds_h[g]) || !(bounds_h[0] > 0.0))
          throw EngineError(1, "group bounds must be positive and strictly increasing");
      nu_pow = p->groups.nu_pow;
      nu_ref = p->groups.nu_ref;
      if (!(nu_ref > 0.0)) throw EngineError(1, "groups.nu_ref must be positive");
    }
    // The starting schedule (the watch corrects it as the problem evolves),
    // measured over the first steps at 256 threads: lanes refill from ranges
    // of 8 chunks per wavefront (the 64^3 cube, steps 3-12: 25.8 against 18.0
    // G segments/s for pass in the mixed engine, 13.5 against 8.2 in FP64;
    // the step slab 19.2 against 16.4), 4 for the mixed engine on polyhedral
    // meshes (the batch of slabs: 36.0 against 26.5); the mixed multigroup
    // problem on a Voronoi mesh runs passes (10.9 against 9.8).
    const bool mg_poly = unstructured && precision != STORM_GPU_FP64 && p->groups.num_groups >= 2;
    schedule = mg_poly ? STORM_GPU_SCHEDULE_PASS : STORM_GPU_SCHEDULE_REFILL;
    if (unstructured && precision != STORM_GPU_FP64) chunks_of[1] = 4;
    // its passes from 1 chunk per wavefront (11.1 against 10.9 at 2)
    if (mg_poly) chunks_of[0] = 1;
    // the FP64 engine refills from 16 (its 2-3 waves per SIMD leave each
    // block a larger LDS table): the cube as stated 15.4 against 14.9 G
    // segments/s over 50 steps, the step slab even
    if (precision == STORM_GPU_FP64) chunks_of[1] = 16;
    // DDMC on a Cartesian mesh: 512-thread blocks refilling from 16 chunks
    // per wavefront (the step slab, steps 3-8, mixed engine: 24.2 G
    // segments/s against 19.3 at 256 threads and 8 chunks, 22.0 at 512 and 8,
    // 18.7 at 256 and 16; FP64: 14.5 against 12.7)
    if (!unstructured && imc.with_ddmc) {
      block = 512;
      chunks_of[1] = 16;
    }
    if (bnd.source_face != -1 && bnd.source_face != STORM_XLO)
      throw EngineError(1, "source_face must be -1 or STORM_XLO (an x+ source is source_xhi)");
    if (bnd.source_xhi && !(bnd.source_temperature_xhi >= 0.0)) throw EngineError(1, "invalid x+ source temperature");
    source_T_hi = bnd.source_temperature_xhi;
    if (!p->material.sigma_a0 && material_mode == STORM_GPU_MATERIAL_DEVICE)
      throw EngineError(1, "material.sigma_a0 is required");
    source_T = bnd.source_temperature;
    if (p->num_instances && !unstructured)
      throw EngineError(1, "batched instances need an unstructured mesh with cell_instance");
    Vh.resize(N);
    chord_h.resize(N);
    if (unstructured) {
      const storm_unstructured_mesh &um = *p->unstructured;
      if (p->num_instances) {
        if (!um.cell_instance || !p->instances) throw EngineError(1, "instances need cell_instance and instances");
        if (groups >= 2) throw EngineError(1, "batched instances are gray");
        inst.assign(p->instances, p->instances + p->num_instances);
        inst_T4.assign(inst.size(), 0.0);
        cell_inst_h.assign(um.cell_instance, um.cell_instance + N);
        for (uint32_t c = 0; c < N; ++c) {
          if (cell_inst_h[c] >= p->num_instances) throw EngineError(1, "cell_instance out of range");
          if (c && cell_inst_h[c] < cell_inst_h[c - 1])
            throw EngineError(1, "the cells of each instance must be contiguous, instances in order");
        }
        inst_first_h.assign(p->num_instances + 1, 0);
        for (uint32_t k = 0, c = 0; k <= p->num_instances; ++k) {
          while (c < N && cell_inst_h[c] < k) ++c;
          inst_first_h[k] = c;
        }
        for (uint32_t k = 0; k < p->num_instances; ++k)
          if (inst_first_h[k + 1] == inst_first_h[k]) throw EngineError(1, "an instance has no cells");
        for (const storm_instance_desc &d : inst)
          if (d.new_photons_per_cell == 0) throw EngineError(1, "every instance needs new_photons_per_cell > 0");
      } else if (um.cell_instance) {
        throw EngineError(1, "cell_instance given without instances");
      }
      try {
        ut = umesh::build<Real>(um, bnd.source_face == STORM_XLO, bnd.source_xhi != 0, [&](uint32_t c) -> uint64_t {
          return inst.empty() ? bnd.source_packets_per_face : inst[cell_inst_h[c]].source_packets_per_face;
        });
      } catch (const std::exception &e) {
        throw EngineError(1, e.what());
      }
      if (!inst.empty())
        for (uint32_t c = 0; c < N; ++c)
          for (uint32_t f = ut.face_start[c]; f < ut.face_start[c + 1]; ++f) {
            const int32_t nb = ut.face[f].nbr;
            if (nb >= 0 && cell_inst_h[static_cast<uint32_t>(nb)] != cell_inst_h[c])
              throw EngineError(1, "cells of different instances share a face");
          }
      for (uint32_t c = 0; c < N; ++c) {
        Vh[c] = um.volume[c];
        chord_h[c] = ut.chord[c];
      }
      gen_h.assign(um.mesh_point, um.mesh_point + 3 * static_cast<size_t>(N));
    } else {
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
          for (int k = 0; k < nz; ++k) {
            const size_t c = (static_cast<size_t>(i) * ny + j) * nz + k;
            const double dx = xe[i + 1] - xe[i], dy = ye[j + 1] - ye[j], dz = ze[k + 1] - ze[k];
            Vh[c] = dx * dy * dz;
            chord_h[c] = 4.0 * Vh[c] / (2.0 * (dy * dz + dx * dz + dx * dy));
          }
    }
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&ev0));
    HIPCHECK(hipEventCreate(&ev1));
    // mesh widths and their reciprocals, interleaved (MeshView::width)
    {
      std::vector<Real> wp;
      for (const std::vector<double> *e : {&xe, &ye, &ze}) {
        std::vector<Real> w(e->size() - 1);
        for (size_t q = 0; q + 1 < e->size(); ++q) w[q] = static_cast<Real>((*e)[q + 1] - (*e)[q]);
        const size_t at = wp.size();
        wp.resize(at + 2 * w.size());
        width_pairs<Real>(w.data(), static_cast<int>(w.size()), wp.data() + at);
      }
      widths.ensure(wp.size());
      HIPCHECK(hipMemcpy(widths.p, wp.data(), wp.size() * sizeof(Real), hipMemcpyHostToDevice));
    }
    auto up = [&](DevBuf<double> &d, const double *src, size_t n) {
      d.ensure(n);
      HIPCHECK(hipMemcpy(d.p, src, n * sizeof(double), hipMemcpyHostToDevice));
    };
    up(V, Vh.data(), N);
    up(chord, chord_h.data(), N);
    if (unstructured) {
      auto upv = [&](auto &d, const auto &h) {
        d.ensure(std::max<size_t>(h.size(), 1));
        if (!h.empty()) HIPCHECK(hipMemcpy(d.p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
      };
      upv(u_face_start, ut.face_start);
      upv(u_face, ut.face);
      upv(u_face_centroid, ut.face_centroid);
      std::vector<double2> geo(ut.face.size());
      for (size_t f = 0; f < geo.size(); ++f) geo[f] = make_double2(ut.face_geo[2 * f], ut.face_geo[2 * f + 1]);
      upv(u_face_geo, geo);
      upv(u_tet_start, ut.tet_start);
      upv(u_tets, ut.tets);
      upv(u_src, ut.src);
      upv(u_tris, ut.tris);
      if (imc.with_ddmc) u_ddmc_rate.ensure(std::max<size_t>(ut.face.size(), 1));
      if (!inst.empty()) {
        const size_t K = inst.size();
        std::vector<uint32_t> nc(K);
        std::vector<uint64_t> nph(K), nmn(K);
        std::vector<double> fac(K);
        for (size_t k = 0; k < K; ++k) {
          nc[k] = inst_first_h[k + 1] - inst_first_h[k];
          nph[k] = inst[k].new_photons_per_cell;
          nmn[k] = std::max<uint64_t>(inst[k].comb_nmin, 1);
          fac[k] = inst[k].comb_factor;
        }
        upv(d_cell_inst, cell_inst_h);
        upv(d_inst_first, inst_first_h);
        upv(d_inst_ncells, nc);
        upv(d_inst_nphot, nph);
        upv(d_inst_nmin, nmn);
        upv(d_inst_factor, fac);
        d_inst_etot.ensure(K);
        d_inst_wtot.ensure(K);
        const std::vector<double> zd(K, 0.0);
        const std::vector<int32_t> zi(K, 0);
        upv(d_inst_src, zd);
        upv(d_inst_cen, zd);
        upv(d_inst_mat0, zd);
        upv(d_inst_shift, zi);
        upv(d_inst_shift_prev, zi);
        inst_shift_now.assign(K, 0);
        inst_shift_last.assign(K, 0);
        HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&inst_src_h), K * sizeof(double)));
        HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&inst_shift_h), K * sizeof(int32_t)));
      }
    }
    if (p->material.sigma_a0) up(sa0, p->material.sigma_a0, N);
    if (p->material.sigma_s0) up(ss0, p->material.sigma_s0, N);
    if (p->material.density) up(rho, p->material.density, N);
    // initial temperature and internal energy (e = EOS(T0) * V)
    std::vector<double> e0(N);
    for (uint32_t c = 0; c < N; ++c) {
      double r = p->material.density ? p->material.density[c] : 1.0;
      e0[c] = matd.eos_f * std::pow(p->T0[c], matd.eos_beta) * std::pow(r, 1.0 - matd.eos_mu) * Vh[c];
    }
    up(T, p->T0, N);
    up(e_mat, e0.data(), N);
    up(e_sync, e0.data(), N);
    for (DevBuf<double> *b : {&fleck, &sig_a, &sig_s, &Eemit, &wemit, &erad_avg, &census_erad, &clamp})
      b->ensure(N);
    partials.ensure(1024);
    xs.ensure(N);
    nphot.ensure(N);
    nphot64.ensure(N + 1);
    nphot_z.ensure((size_t(1) << morton.total) + 1);
    offsets.ensure((size_t(1) << morton.total) + 1);
    err.ensure(1);
    // the stratified comb (a multigroup problem with imc.comb = 2) works on
    // (cell, group) bins; with one group it is the plain comb
    strat_groups = (imc.comb == 2 && groups >= 2) ? static_cast<uint32_t>(groups) : 1u;
    if (strat_groups > 1) {
      if (strat_groups > kMaxStratGroups)
        throw EngineError(1, "the stratified comb takes at most 64 groups");
      strat_gbits = static_cast<uint32_t>(fx::ceil_log2(static_cast<double>(strat_groups)));
      if (morton.total + strat_gbits > 31u) throw EngineError(1, "too many cells and groups for the stratified comb");
    }
    run_start.ensure(static_cast<size_t>(N) * strat_groups);
    run_end.ensure(static_cast<size_t>(N) * strat_groups);
    comb_cells.ensure(static_cast<size_t>(N) * strat_groups);
    emit_limbs.ensure(2 * static_cast<size_t>(N));
    requant.ensure(2 * static_cast<size_t>(N));
    leaving.ensure(5);  // escapes, random-walk steps, DDMC steps, segments, longest history
    dsc.ensure(3);      // material energy before the step, census energy after it, emission total
    qexp.ensure(2);     // energy quantum exponents of the step
    mail.ensure(8);
    HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&mail_h), 8 * sizeof(unsigned long long)));
    // boundary source faces: at most both x sides of ny * nz faces
    {
      const size_t max_faces = std::max<size_t>(unstructured ? ut.src.size() : 2 * static_cast<size_t>(ny) * nz, 1);
      wface.ensure(max_faces);
      HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&wface_h), max_faces * sizeof(double)));
    }
    scratch_ptr = noalloc_scratch_address();
    // The census buffer is the engine's state (the step's input and output);
    // reserve it for twice the census the population control aims at (or
    // four steps of sources without it), so that storm_gpu_step_noalloc,
    // which never allocates, can run from the first step.
    {
      const double per_step = static_cast<double>(N) * static_cast<double>(imc.new_photons_per_cell) +
                              2.0 * ny * nz * static_cast<double>(bnd.source_packets_per_face);
      const double factor = imc.comb_factor >= 1.0 ? std::min(imc.comb_factor, 64.0) : 1.0;
      const double want = imc.comb ? 2.0 * N * static_cast<double>(std::max<uint64_t>(imc.comb_nmin, 1)) * factor +
                                         per_step
                                   : 4.0 * per_step;
      cnext.ensure(static_cast<size_t>(std::min(std::max(want, 1.0), 16777216.0)));
    }
    counters_u32.ensure(4);
    if (groups >= 2) {
      up(bounds_erg, bounds_h.data(), bounds_h.size());
      std::vector<Real> bk(bounds_h.size());
      for (size_t g = 0; g < bk.size(); ++g) bk[g] = static_cast<Real>(bounds_h[g] / units::kev);
      bounds_kev.ensure(bk.size());
      HIPCHECK(hipMemcpy(bounds_kev.p, bk.data(), bk.size() * sizeof(Real), hipMemcpyHostToDevice));
      cdf.ensure(static_cast<size_t>(N) * static_cast<size_t>(groups + 1));
      sig_s_cell.ensure(N);
      bcdf.ensure(2 * static_cast<size_t>(groups + 1));
      bcdf_h.reserve(2 * static_cast<size_t>(groups + 1));
      planck_acc.reserve(static_cast<size_t>(groups + 1));
      pa.with_nu = pb.with_nu = true;
    }
    if (accel_on) {
      accel.ensure(N);
      Dtil.ensure(N);
      if (imc.with_random_walk) {
        // random-walk tables built on the host (once per process), like the
        // RandomWalk constructor the reference calls from RadiationIMC's
        // constructor
        const std::vector<double> &tb = rw_host_tables();
        std::vector<Real> h(tb.begin(), tb.end());
        rw_tables.ensure(h.size());
        HIPCHECK(hipMemcpy(rw_tables.p, h.data(), h.size() * sizeof(Real), hipMemcpyHostToDevice));
      }
    }
    tally_u64.ensure(9 * 2 * static_cast<size_t>(N));
    tally_f64.ensure(3 * static_cast<size_t>(N));
    DevBuf<uint32_t> *segs_begin = new DevBuf<uint32_t>();  // owned below
    segs_begin_.reset(segs_begin);
    segs_begin->ensure(2 * static_cast<size_t>(N));
    unsigned long long *tu = tally_u64.p;
    const size_t n2 = 2 * static_cast<size_t>(N);
    tal.dep = tu;
    tal.pint = tu + n2;
    tal.esc = tu + 2 * n2;
    tal.cin = tu + 3 * n2;
    tal.cpre = tu + 4 * n2;
    tal.cout = tu + 5 * n2;
    tal.rem = tu + 6 * n2;
    tal.rdep = tu + 7 * n2;
    tal.cdrop = tu + 8 * n2;
    tal.dep_fp = tally_f64.p;
    tal.pint_fp = tally_f64.p + N;
    tal.esc_fp = tally_f64.p + 2 * static_cast<size_t>(N);
    tal.segs = segs_begin->p;
    tal.begin = segs_begin->p + N;
    if (precision != STORM_GPU_FP64) {
      tal.dep_fp = nullptr;
      tal.pint_fp = nullptr;
      tal.esc_fp = nullptr;
    }
    {
      const size_t R1 = kTallyReplicas - 1;
      tal.ncell = N;
      rep_u32.ensure(R1 * N);
      tal.segs_rep = rep_u32.p;
      if (precision == STORM_GPU_FP64) {
        rep_f64.ensure(R1 * 2 * static_cast<size_t>(N));
        tal.dep_fp_rep = rep_f64.p;
        tal.pint_fp_rep = rep_f64.p + R1 * N;
        tal.dep_rep = tal.pint_rep = nullptr;
      } else {
        rep_u64.ensure(R1 * 2 * n2);
        tal.dep_rep = rep_u64.p;
        tal.pint_rep = rep_u64.p + R1 * n2;
        tal.dep_fp_rep = tal.pint_fp_rep = nullptr;
      }
    }
    // per-group time-averaged radiation energy (withEgTimeAvg, multigroup)
    eg_on = groups >= 2 && p->groups.with_eg_time_avg != 0;
    tal.eg = nullptr;
    tal.eg_fp = nullptr;
    tal.eg_groups = eg_on ? static_cast<uint32_t>(groups) : 0u;
    if (eg_on) {
      const size_t n = static_cast<size_t>(N) * static_cast<size_t>(groups);
      if (precision == STORM_GPU_FP64) {
        eg_f64.ensure(n);
        zero_f64(eg_f64.p, n);
        tal.eg_fp = eg_f64.p;
      } else {
        eg_u64.ensure(n);
        zero_u64(eg_u64.p, n);
        tal.eg = eg_u64.p;
      }
    }
    zero_u64(tally_u64.p, tally_u64.n);
    zero_f64(tally_f64.p, tally_f64.n);
    zero_u32(segs_begin->p, segs_begin->n);
    zero_u64(requant.p, requant.n);
    HIPCHECK(hipStreamSynchronize(stream));
  }

  ~Engine() override {
    free_retired();
    if (stream) (void)hipStreamSynchronize(stream);
    if (mail_h) (void)hipHostFree(mail_h);
    if (wface_h) (void)hipHostFree(wface_h);
    if (inst_src_h) (void)hipHostFree(inst_src_h);
    if (inst_shift_h) (void)hipHostFree(inst_shift_h);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  std::unique_ptr<DevBuf<uint32_t>> segs_begin_;

  void zero_u64(unsigned long long *p, size_t n) {
    if (n) k_fill_u64<<<blocks_for(n, 256), 256, 0, stream>>>(p, n, 0ull);
  }
  void zero_u32(uint32_t *p, size_t n) {
    if (n) k_fill_u32<<<blocks_for(n, 256), 256, 0, stream>>>(p, n, 0u);
  }
  void zero_f64(double *p, size_t n) {
    if (n) k_fill_f64<<<blocks_for(n, 256), 256, 0, stream>>>(p, n, 0.0);
  }
  // Buffers to zero with one k_zero_ranges launch (a full batch is launched
  // as the next one is added; zero_flush launches the rest).
  struct ZeroBatch {
    ZeroRanges r{};
    uint32_t k = 0;
    uint64_t most = 0;  // largest range, in words
  };
  void zero_add(ZeroBatch &z, void *p, size_t bytes) {
    if (!p || bytes < 4) return;
    if (z.k == static_cast<uint32_t>(kZeroRanges)) zero_flush(z);
    z.r.p[z.k] = static_cast<uint32_t *>(p);
    z.r.n[z.k] = bytes / 4;  // every buffer here holds 4- or 8-byte words
    z.most = std::max<uint64_t>(z.most, bytes / 4);
    ++z.k;
  }
  void zero_flush(ZeroBatch &z) {
    if (!z.k) return;
    const uint32_t bx = static_cast<uint32_t>(std::min<uint64_t>((z.most + 255) / 256, 2048));
    k_zero_ranges<<<dim3(bx, z.k), 256, 0, stream>>>(z.r);
    z = ZeroBatch{};
  }

  MatDev matdev() const {
    MatDev m;
    m.sa0 = sa0.p;
    m.ss0 = ss0.p;
    m.rho = rho.p;
    m.Tref = matd.opacity_T_ref;
    m.Tfloor = matd.opacity_T_floor;
    m.apow = matd.a_temp_pow;
    m.arho = matd.a_rho_pow;
    m.spow = matd.s_temp_pow;
    m.srho = matd.s_rho_pow;
    m.eos_f = matd.eos_f;
    m.eos_beta = matd.eos_beta;
    m.eos_mu = matd.eos_mu;
    m.cvTfloor = matd.eos_cv_T_floor;
    return m;
  }
  AccelView<Real> accelview() const {
    AccelView<Real> a;
    a.cell = accel_on ? accel.p : nullptr;
    a.rw = rw_tables.p;
    // a particle threshold of +inf switches a method off
    a.rw_particle_tau = imc.with_random_walk ? static_cast<Real>(imc.rw_min_particle_optical_depth)
                                             : MathPolicy<Real>::inf();
    a.ddmc_particle_tau = imc.with_ddmc ? static_cast<Real>(imc.ddmc_min_particle_optical_depth)
                                        : MathPolicy<Real>::inf();
    return a;
  }
  MeshDims dims() const {
    return MeshDims{static_cast<uint32_t>(nx), static_cast<uint32_t>(ny), static_cast<uint32_t>(nz),
                    unstructured ? 1u : 0u};
  }
  MeshView<Real> meshview() const {
    MeshView<Real> m;
    m.nx = nx;
    m.ny = ny;
    m.nz = nz;
    m.escape_mask = 0u;
    for (int f = 0; f < 6; ++f)
      if (bnd.face_kind[f] == STORM_FACE_ESCAPE) m.escape_mask |= 1u << f;
    m.width = widths.p;
    m.spec = spectrum();
    return m;
  }
  SpectrumView<Real> spectrum() const {
    SpectrumView<Real> s{};
    s.groups = groups;
    if (groups >= 2) {
      s.nu_pow = static_cast<Real>(nu_pow);
      s.e_lo = static_cast<Real>(bounds_h.front() / units::kev);
      s.e_hi = static_cast<Real>(bounds_h.back() / units::kev);
      s.bounds = bounds_kev.p;
      s.cdf = cdf.p;
      s.sig_s = sig_s_cell.p;
      s.eg = eg_on ? 1 : 0;
    }
    return s;
  }

  UMeshView<Real> umeshview() const {
    UMeshView<Real> m;
    m.ncells = N;
    m.escape_mask = 0u;
    for (int f = 0; f < 6; ++f)
      if (bnd.face_kind[f] == STORM_FACE_ESCAPE) m.escape_mask |= 1u << f;
    m.face_start = u_face_start.p;
    m.face = u_face.p;
    m.face_centroid = u_face_centroid.p;
    m.ddmc_rate = imc.with_ddmc ? u_ddmc_rate.p : nullptr;
    m.spec = spectrum();
    return m;
  }
  // Per transport kernel and block size: the blocks its registers let run
  // on one CU, and the tally-cache slots (the share of the CU's LDS of each
  // such block) for each chunking.
  static int chunk_index(int32_t k) { return k == 1 ? 0 : (k == 2 ? 1 : (k == 4 ? 2 : (k == 8 ? 3 : 4))); }
  static int block_index(int32_t b) { return b == 64 ? 0 : (b == 128 ? 1 : (b == 256 ? 2 : 3)); }
  struct KernelFit {
    uint32_t per_cu = 0;
    uint32_t slots[5] = {};
  };
  KernelFit fit_cache[2][2][4][4];  // [refill][polyhedral][F][block]
  int n_cu = 0;
  template <typename K>
  KernelFit &fit_for(K kernel, bool refill, uint32_t F, int32_t b) {
    KernelFit &f = fit_cache[refill ? 1 : 0][unstructured ? 1 : 0][F][block_index(b)];
    if (!f.per_cu) {
      int per_cu = 0;
      HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, b, 0));
      f.per_cu = static_cast<uint32_t>(std::max(per_cu, 1));
      if (!n_cu) {
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
        n_cu = std::max(n_cu, 1);
      }
    }
    return f;
  }
  // The chunks per wavefront of a launch over n_in packets: the schedule's,
  // halved while the launch would have fewer blocks than the device holds at
  // once (a small step in large block ranges leaves CUs idle: the 100-cell
  // slab ran 5 times slower in ranges of 8 chunks).
  int32_t launch_chunks(const KernelFit &f, uint64_t n_in) const {
    int32_t ch = chunks_per_wave();
    const uint64_t resident = static_cast<uint64_t>(f.per_cu) * static_cast<uint64_t>(n_cu);
    while (ch > 1 && (n_in + static_cast<uint64_t>(block) * ch - 1) / (static_cast<uint64_t>(block) * ch) < resident)
      ch >>= 1;
    return ch;
  }
  uint32_t slots_of(KernelFit &f, int32_t ch, uint32_t nf) {
    uint32_t &r = f.slots[chunk_index(ch)];
    if (!r) r = cache_slots(static_cast<uint32_t>(block) * static_cast<uint32_t>(ch), nf, f.per_cu);
    return r;
  }

  // One transport launch on the kernel variant compiled for this problem's
  // features (random walk / DDMC, multigroup) and mesh, in the schedule in use.
  template <uint32_t F, typename MeshT>
  void launch_transport_m(uint64_t n_in, PacketBufs<Real> &src, PacketBufs<Real> &dst, const MeshT &mesh,
                          const AccelView<Real> &av, const StepConsts &kc) {
    const uint32_t nf = (precision == STORM_GPU_MIXED && groups < 2) ? 1u : 2u;
    const bool refill = schedule == STORM_GPU_SCHEDULE_REFILL;
    KernelFit &f = refill ? fit_for(k_transport_refill<Real, F, MeshT>, true, F, block)
                          : fit_for(k_transport<Real, F, MeshT>, false, F, block);
    const int32_t ch = launch_chunks(f, n_in);
    const uint32_t slots = slots_of(f, ch, nf);
    const uint32_t chunks = static_cast<uint32_t>(ch) * static_cast<uint32_t>(block / 64);
    const uint32_t grid = blocks_for(n_in, 64u * chunks);
    if (refill) {
      k_transport_refill<Real, F, MeshT><<<grid, block, cache_bytes(slots, nf), stream>>>(
          static_cast<uint32_t>(n_in), src.soa(), dst.soa(), counters_u32.p, cpre.soa(), counters_u32.p + 1, mesh,
          xs.p, av, kc, seg_per_pass, tal, leaving.p, slots, nf, chunks, cstage.soa());
    } else {
      k_transport<Real, F, MeshT><<<grid, block, cache_bytes(slots, nf), stream>>>(
          static_cast<uint32_t>(n_in), src.soa(), dst.soa(), counters_u32.p, cpre.soa(), counters_u32.p + 1, mesh,
          xs.p, av, kc, seg_per_pass, tal, leaving.p, slots, nf, chunks);
    }
  }
  template <uint32_t F>
  void launch_transport_f(uint64_t n_in, PacketBufs<Real> &src, PacketBufs<Real> &dst, const MeshView<Real> &mv,
                          const AccelView<Real> &av, const StepConsts &kc) {
    if (unstructured) launch_transport_m<F>(n_in, src, dst, umeshview(), av, kc);
    else launch_transport_m<F>(n_in, src, dst, mv, av, kc);
  }
  void launch_transport(uint64_t n_in, PacketBufs<Real> &src, PacketBufs<Real> &dst, const MeshView<Real> &mv,
                        const AccelView<Real> &av, const StepConsts &kc) {
    const uint32_t f = (accel_on ? kFeatAccel : 0u) | (groups >= 2 ? kFeatMultigroup : 0u);
    switch (f) {
      case 0u: launch_transport_f<0u>(n_in, src, dst, mv, av, kc); break;
      case kFeatAccel: launch_transport_f<kFeatAccel>(n_in, src, dst, mv, av, kc); break;
      case kFeatMultigroup: launch_transport_f<kFeatMultigroup>(n_in, src, dst, mv, av, kc); break;
      default: launch_transport_f<kFeatAll>(n_in, src, dst, mv, av, kc); break;
    }
  }

  // Deterministic FP64 sum (fixed grid and tree) of d[0..n) into *out (device).
  static constexpr int kSumGrid = 512;
  void sum_async(const double *d, uint64_t n, double *out) {
    k_sum_partials<256><<<kSumGrid, 256, 0, stream>>>(d, n, partials.p);
    k_sum_partials<256><<<1, 256, 0, stream>>>(partials.p, kSumGrid, out);
  }

  // Reads up to 8 device scalars (4 or 8 bytes each) with one copy into
  // pinned memory and one synchronisation.
  const unsigned long long *read_scalars(std::initializer_list<std::pair<const void *, uint32_t>> refs) {
    ScalarRefs r{};
    uint32_t n = 0;
    for (const auto &x : refs) {
      r.p[n] = x.first;
      r.bytes[n] = x.second;
      ++n;
    }
    k_gather_scalars<<<1, 64, 0, stream>>>(r, n, mail.p);
    HIPCHECK(hipMemcpyAsync(mail_h, mail.p, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return mail_h;
  }
  static double as_double(unsigned long long v) {
    double d;
    std::memcpy(&d, &v, sizeof d);
    return d;
  }

  template <typename F>
  void ensure_temp(F &&query) {
    size_t bytes = 0;
    query(nullptr, bytes);
    temp.ensure(bytes + 16);
    query(temp.p, bytes);
  }

  // ----------------------------------------------------------------- step
  // Buffers outgrown during steps, freed with the engine, or at the end of a
  // step once they hold more than kRetiredLimit bytes (also when the step
  // throws).
  static constexpr size_t kRetiredLimit = size_t(16) << 30;
  std::vector<Retired> retired;
  void free_retired() {
    if (retired.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (const Retired &r : retired) (void)hipFree(r.p);
    retired.clear();
  }
  struct StepRetire {
    Engine &e;
    explicit StepRetire(Engine &en) : e(en) { t_retired = &e.retired; }
    ~StepRetire() {
      t_retired = nullptr;
      size_t bytes = 0;
      for (const Retired &r : e.retired) bytes += r.bytes;
      if (bytes > kRetiredLimit) e.free_retired();
    }
  };

  bool failed_ = false;  // a step stopped part-way with status 5
  void step(double dt) override {
    if (failed_)
      throw EngineError(5, "an earlier step stopped part-way (status 5); the engine's state is not consistent: "
                           "destroy it");
    if (!(dt > 0.0)) throw EngineError(1, "dt must be positive");
    auto t_start = std::chrono::steady_clock::now();
    StepRetire retire(*this);
    ++step_index;
    counters = storm_gpu_counters{};
    const MatDev md = matdev();
    const MeshDims d = dims();
    const Real s0 = static_cast<Real>(units::clight * dt);
    // the step's tallies, counters and error flags start at zero (one launch)
    {
      ZeroBatch z;
      zero_add(z, tally_u64.p, tally_u64.n * sizeof(unsigned long long));
      if (precision == STORM_GPU_FP64) zero_add(z, tally_f64.p, tally_f64.n * sizeof(double));
      zero_add(z, tal.segs, 2 * static_cast<size_t>(N) * sizeof(uint32_t));
      zero_add(z, tal.eg, static_cast<size_t>(N) * tal.eg_groups * sizeof(unsigned long long));
      zero_add(z, tal.eg_fp, static_cast<size_t>(N) * tal.eg_groups * sizeof(double));
      zero_add(z, rep_u32.p, rep_u32.n * sizeof(uint32_t));
      if (precision == STORM_GPU_FP64) zero_add(z, rep_f64.p, rep_f64.n * sizeof(double));
      else zero_add(z, rep_u64.p, rep_u64.n * sizeof(unsigned long long));
      zero_add(z, requant.p, requant.n * sizeof(unsigned long long));
      zero_add(z, leaving.p, 5 * sizeof(unsigned long long));
      zero_add(z, counters_u32.p, 2 * sizeof(uint32_t));
      zero_add(z, err.p, sizeof(uint32_t));
      zero_flush(z);
    }
    eq_prev_ledger_ = eq_prev;
    sum_async(e_mat.p, N, dsc.p);  // material energy before the step, read by the ledger
    if (!inst.empty())
      k_inst_sum<<<static_cast<uint32_t>(inst.size()), 256, 0, stream>>>(d_inst_first.p, e_mat.p, d_inst_mat0.p);

    // preStep: opacities, Fleck factor, emission energies
    if (groups >= 2) {
      k_prestep_mg<Real><<<blocks_for(N, 256), 256, 0, stream>>>(
          N, T.p, md, have_host_sa ? host_sa.p : nullptr, have_host_ss ? host_ss.p : nullptr,
          have_host_cv ? host_cv.p : nullptr, V.p, dt, groups, bounds_erg.p, nu_pow, nu_ref, xs.p, cdf.p,
          sig_s_cell.p, fleck.p, sig_a.p, sig_s.p, Eemit.p, err.p, chord.p,
          imc.with_random_walk ? imc.rw_min_cell_optical_depth : HUGE_VAL, accel_on ? accel.p : nullptr,
          (imc.with_ddmc && imc.ddmc_multigroup_pgrw) ? imc.ddmc_min_cell_optical_depth : -1.0,
          accel_on ? Dtil.p : nullptr);
    } else {
      k_prestep<Real><<<blocks_for(N, 256), 256, 0, stream>>>(
          N, T.p, md, have_host_sa ? host_sa.p : nullptr, have_host_ss ? host_ss.p : nullptr,
          have_host_cv ? host_cv.p : nullptr, V.p, dt, xs.p, fleck.p, sig_a.p, sig_s.p, Eemit.p, err.p);
    }
    if (accel_on && (groups < 2 || imc.with_ddmc)) {
      k_accel_cells<Real><<<blocks_for(N, 256), 256, 0, stream>>>(
          N, chord.p, sig_a.p, sig_s.p, fleck.p, imc.with_random_walk, imc.rw_min_cell_optical_depth, imc.with_ddmc,
          imc.ddmc_min_cell_optical_depth, groups >= 2 ? 1 : 0, imc.ddmc_multigroup_pgrw ? 1 : 0, Dtil.p,
          accel.p);
      if (unstructured) {
        if (imc.with_ddmc)
          k_accel_rates_u<Real><<<blocks_for(N, 256), 256, 0, stream>>>(N, u_face_start.p, u_face.p, u_face_geo.p, V.p,
                                                                         Dtil.p, accel.p, u_ddmc_rate.p);
      } else {
        k_accel_rates<Real><<<blocks_for(N, 256), 256, 0, stream>>>(N, d, meshview(), Dtil.p, accel.p);
      }
    }
    sum_async(Eemit.p, N, dsc.p + 2);  // emission total, read by k_quantum and k_alloc

    // boundary sources: SideTemperature / TwoSidesTemperature energy per face
    // (T^4 frozen at the first call if asked), sides x- then x+
    const uint32_t per_side = static_cast<uint32_t>(ny) * nz;
    double side_T[2];
    uint32_t nsides = 0;
    if (bnd.source_face == STORM_XLO) side_T[nsides++] = source_T;
    const uint32_t xhi_from = nsides;
    if (bnd.source_xhi) side_T[nsides++] = source_T_hi;
    const uint32_t npc = nsides ? static_cast<uint32_t>(bnd.source_packets_per_face) : 0u;
    const uint32_t nfaces = unstructured ? static_cast<uint32_t>(ut.src.size()) : (npc ? per_side * nsides : 0u);
    double side_T4[2] = {0.0, 0.0};
    for (uint32_t s = 0; s < nsides; ++s) side_T4[s] = side_T[s] * side_T[s] * side_T[s] * side_T[s];
    if (bnd.freeze_source_T4) {
      if (!frozen) {
        for (uint32_t s = 0; s < nsides; ++s) frozenT4s[s] = side_T4[s];
        for (size_t q = 0; q < inst.size(); ++q) {
          const double Ti = inst[q].source_temperature;
          inst_T4[q] = Ti * Ti * Ti * Ti;
        }
        frozen = true;
      }
      for (uint32_t s = 0; s < nsides; ++s) side_T4[s] = frozenT4s[s];
    } else {
      for (size_t q = 0; q < inst.size(); ++q) {
        const double Ti = inst[q].source_temperature;
        inst_T4[q] = Ti * Ti * Ti * Ti;
      }
    }
    // face weights, computed straight into the pinned staging buffer (sized
    // at creation; the stream is idle: the last step ended with a
    // synchronisation, so its copy from this buffer is done)
    double *wface_e = wface_h;
    double source_total = 0.0;
    uint64_t n_bc = 0;
    if (unstructured) {
      // each source face: its side's (or its instance's) T^4, its own area
      // and packets per step
      for (uint32_t f = 0; f < nfaces; ++f) {
        const USource &S = ut.src[f];
        const double T4 = inst.empty() ? side_T4[S.xhi ? xhi_from : 0u] : inst_T4[ut.src_instance[f]];
        wface_e[f] = units::sigma_sb * T4 * ut.src_area[f] * dt / S.count;
        source_total += wface_e[f] * S.count;
        n_bc += S.count;
      }
    } else {
      for (uint32_t f = 0; f < nfaces; ++f) {
        const uint32_t s = f / per_side, jk = f % per_side;
        uint32_t j = jk / nz, k = jk % nz;
        double A = (ye[j + 1] - ye[j]) * (ze[k + 1] - ze[k]);
        wface_e[f] = units::sigma_sb * side_T4[s] * A * dt / npc;
        source_total += wface_e[f] * npc;
      }
      n_bc = static_cast<uint64_t>(nfaces) * npc;
    }
    last_nfaces_ = nfaces;
    // host-made source packets join the step's boundary source
    const uint64_t n_host = host_queue_.size();
    for (const HostPacket &h : host_queue_) source_total += h.w;

    // energy quantum of this step (every energy below 2^62 quanta; in a batch
    // every instance's, in its own unit), photon allocation and offsets; one
    // read-back for the quantum and the count
    if (inst.empty()) {
      k_quantum<<<1, 64, 0, stream>>>(dsc.p + 2, census_total_prev, source_total, units::clight, dt, qexp.p);
    } else {
      const size_t K = inst.size();
      std::fill(inst_src_h, inst_src_h + K, 0.0);
      for (uint32_t f = 0; f < nfaces; ++f) inst_src_h[ut.src_instance[f]] += wface_e[f] * ut.src[f].count;
      for (const HostPacket &h : host_queue_) inst_src_h[cell_inst_h[h.cell]] += h.w;
      HIPCHECK(hipMemcpyAsync(d_inst_src.p, inst_src_h, K * sizeof(double), hipMemcpyHostToDevice, stream));
      k_inst_sum<<<static_cast<uint32_t>(K), 256, 0, stream>>>(d_inst_first.p, Eemit.p, d_inst_etot.p);
      k_inst_quantum<<<1, 256, 0, stream>>>(static_cast<uint32_t>(K), d_inst_cen.p, d_inst_etot.p, d_inst_src.p,
                                             units::clight * dt, d_inst_shift.p, d_inst_shift_prev.p, qexp.p);
    }
    k_alloc<<<blocks_for(N, 256), 256, 0, stream>>>(N, Eemit.p, dsc.p + 2, imc.new_photons_per_cell, qexp.p,
                                                    nphot.p, wemit.p, emit_limbs.p, instview(), step_index);
    // photon counts in the cells' Z order, so emission is laid out like the
    // combed census
    const uint32_t M = 1u << morton.total;
    k_morton_counts<<<blocks_for(M + 1, 256), 256, 0, stream>>>(M, morton, d, nphot.p, nphot_z.p);
    {
      auto q = [&](void *tp, size_t &bytes) {
        HIPCHECK(rocprim::exclusive_scan(tp, bytes, nphot_z.p, offsets.p, uint64_t(0), static_cast<size_t>(M) + 1,
                                         rocprim::plus<uint64_t>(), stream));
      };
      ensure_temp(q);
    }
    if (!inst.empty())
      HIPCHECK(hipMemcpyAsync(inst_shift_h, d_inst_shift.p, inst.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                              stream));
    const unsigned long long *hs =
        read_scalars({{offsets.p + M, 8u}, {qexp.p, 4u}, {qexp.p + 1, 4u}});
    const uint64_t n_emit = hs[0];
    const int32_t eq = static_cast<int32_t>(static_cast<uint32_t>(hs[1]));
    const int32_t et = static_cast<int32_t>(static_cast<uint32_t>(hs[2]));
    counters.energy_quantum_exp = eq;
    if (inst.empty()) {
      for (uint32_t f = 0; f < nfaces; ++f) wface_e[f] = fx::quantize_down(wface_e[f], eq);
    } else {
      // read_scalars synchronised the stream: the shifts are on the host
      inst_shift_last = inst_shift_now;
      inst_shift_now.assign(inst_shift_h, inst_shift_h + inst.size());
      for (uint32_t f = 0; f < nfaces; ++f)
        wface_e[f] = fx::quantize_down(std::ldexp(wface_e[f], inst_shift_now[ut.src_instance[f]]), eq);
    }
    if (groups >= 2 && nfaces) {
      // Planck fractions of each source temperature per group
      // (examples/densmore2012/DensmoreBoundary.hpp:96-104), into host
      // vectors whose capacity was reserved at creation
      bcdf_h.clear();
      for (uint32_t s = 0; s < nsides; ++s) {
        const double kT = units::k_boltz * side_T[s];
        planck_acc.assign(static_cast<size_t>(groups) + 1, 0.0);
        for (int32_t g = 0; g < groups; ++g)
          planck_acc[g + 1] = planck_acc[g] + planck::integral(bounds_h[g] / kT, bounds_h[g + 1] / kT);
        for (int32_t g = 0; g <= groups; ++g)
          bcdf_h.push_back(planck_acc[groups] > 0.0 ? static_cast<Real>(planck_acc[g] / planck_acc[groups]) : Real(0));
      }
      bcdf.ensure(bcdf_h.size());
      HIPCHECK(hipMemcpyAsync(bcdf.p, bcdf_h.data(), bcdf_h.size() * sizeof(Real), hipMemcpyHostToDevice, stream));
      HIPCHECK(hipStreamSynchronize(stream));
    }
    if (nfaces) HIPCHECK(hipMemcpyAsync(wface.p, wface_h, nfaces * sizeof(double), hipMemcpyHostToDevice, stream));
    const uint64_t n_total = n_census + n_emit + n_bc + n_host;
    // 32-bit packet indices; the refill queue head may run past the count by
    // one wavefront of slots per resident wavefront
    if (n_total >= (1ull << 32) - (1ull << 24)) throw EngineError(4, "more than 2^32 - 2^24 packets in one step");
    counters.initial_particles = n_census;
    counters.emitted_particles = n_emit;
    counters.boundary_particles = n_bc + n_host;
    counters.prestep_particles = n_emit + n_bc + n_host;
    // host-made packets on the step's quantum (in a batch, in their
    // instance's unit), their path left within the step
    host_last_.clear();
    for (HostPacket &h : host_queue_) {
      const double w = inst.empty() ? h.w : std::ldexp(h.w, inst_shift_now[cell_inst_h[h.cell]]);
      h.w = fx::quantize_down(w, eq);
      h.s = std::min(h.s, static_cast<double>(s0));
      host_last_.emplace_back(h.cell, h.w);
    }
    counters.start_particles = n_total;


    // packets: [census | emission | boundary]
    pa.ensure(n_total);
    pb.ensure(n_total);
    cpre.ensure(n_total);
    if (n_census)
      k_import<Real><<<blocks_for(n_census, 256), 256, 0, stream>>>(
          static_cast<uint32_t>(n_census), cnext.soa(), pa.soa(), s0, eq_prev, eq, d, tal, requant.p, instview());
    if (n_emit && unstructured)
      k_emit_u<Real><<<blocks_for(n_emit, 256), 256, 0, stream>>>(
          static_cast<uint32_t>(n_emit), M, offsets.p, wemit.p, pa.soa(), static_cast<uint32_t>(n_census), s0, seed,
          step_index, tal.begin, spectrum(), u_tets.p, u_tet_start.p);
    else if (n_emit)
      k_emit<Real><<<blocks_for(n_emit, 256), 256, 0, stream>>>(
          static_cast<uint32_t>(n_emit), M, offsets.p, wemit.p, pa.soa(), static_cast<uint32_t>(n_census), s0, d, morton,
          seed, step_index, tal.begin, spectrum());
    if (n_bc && unstructured)
      k_boundary_u<Real><<<blocks_for(n_bc, 256), 256, 0, stream>>>(
          static_cast<uint32_t>(n_bc), nfaces, u_src.p, u_tris.p, wface.p, pa.soa(),
          static_cast<uint32_t>(n_census + n_emit), s0, seed, step_index, tal.begin, spectrum(),
          groups >= 2 ? bcdf.p : nullptr, xhi_from);
    else if (n_bc)
      k_boundary<Real><<<blocks_for(n_bc, 256), 256, 0, stream>>>(
          static_cast<uint32_t>(n_bc), npc, wface.p, pa.soa(), static_cast<uint32_t>(n_census + n_emit), s0, d, seed,
          step_index, tal.begin, spectrum(), groups >= 2 ? bcdf.p : nullptr, xhi_from);

    if (n_host) {
      host_pk.ensure(n_host);
      HIPCHECK(hipMemcpyAsync(host_pk.p, host_queue_.data(), n_host * sizeof(HostPacket), hipMemcpyHostToDevice,
                              stream));
      k_host_packets<Real><<<blocks_for(n_host, 256), 256, 0, stream>>>(
          static_cast<uint32_t>(n_host), host_pk.p, pa.soa(), static_cast<uint32_t>(n_census + n_emit + n_bc),
          tal.begin);
      // the copy reads host_queue_: wait for it before the queue is cleared
      HIPCHECK(hipStreamSynchronize(stream));
      host_queue_.clear();
    }

    StepConsts kc;
    kc.seed = seed;
    kc.step = step_index;
    kc.eq = eq;
    kc.et = et;
    if (keep_inputs) {
      HIPCHECK(hipStreamSynchronize(stream));
      keep_transport_inputs(n_total, kc);
    }

    // transport passes until every packet is in census, absorbed or gone
    tune_before_step();
    if (schedule == STORM_GPU_SCHEDULE_REFILL) cstage.ensure(n_total);
    counters.transport_block = block;
    counters.transport_chunks = chunks_per_wave();
    counters.transport_schedule = schedule;
    const MeshView<Real> mv = meshview();
    const AccelView<Real> av = accelview();
    uint64_t n_in = n_total;
    bool ping = true;
    double t_transport = 0.0, t_max = 0.0;
    uint32_t n_cen_total = 0;  // both counters were zeroed with the tallies
    bool first_pass = true;
    // A packet may take at most kMaxSegmentsPerStep segments in one step (a
    // history that long is a problem set-up the random walk should cover, and
    // one lane running it would keep the GPU busy for minutes).
    const uint64_t max_passes = (kMaxSegmentsPerStep + static_cast<uint64_t>(seg_per_pass) - 1) /
                                    static_cast<uint64_t>(seg_per_pass) + 1;
    while (n_in > 0) {
      if (counters.transport_passes >= max_passes) {
        // the packets still in flight, for storm_gpu_debug_transport_inputs
        keep_packets(ping ? pa : pb, std::min<uint64_t>(n_in, 4096), kc);
        throw EngineError(6, "transport did not finish: a packet exceeded 2^24 segments in one step "
                             "(optically thick cells without the random walk?)");
      }
      PacketBufs<Real> &src = ping ? pa : pb;
      PacketBufs<Real> &dst = ping ? pb : pa;
      if (!first_pass) zero_u32(counters_u32.p, 1);  // n_out
      first_pass = false;
      HIPCHECK(hipEventRecord(ev0, stream));
      launch_transport(n_in, src, dst, mv, av, kc);
      HIPCHECK(hipEventRecord(ev1, stream));
      HIPCHECK(hipMemcpyAsync(mail_h, counters_u32.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      uint32_t h[2];
      std::memcpy(h, mail_h, sizeof h);
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
      t_transport += ms * 1e-3;
      t_max = std::max(t_max, static_cast<double>(ms) * 1e-3);
      counters.transport_launches++;
      counters.transport_passes++;
      n_in = h[0];
      n_cen_total = h[1];
      ping = !ping;
    }
    k_fold_replicas<<<blocks_for(N, 256), 256, 0, stream>>>(tal, kTallyReplicas, err.p);
    counters.seconds_transport = t_transport;
    counters.seconds_transport_max_launch = t_max;
    counters.census_before_comb = n_cen_total;

    // population control (comb) or a deterministic pass-through
    uint64_t n_out = comb(n_cen_total, eq);
    n_census = n_out;
    counters.end_particles = n_out;

    // postStep
    k_post<Real><<<blocks_for(N, 256), 256, 0, stream>>>(
        N, tal, eq, et, eq_prev, emit_limbs.p, requant.p, dt, V.p, md,
        material_mode == STORM_GPU_MATERIAL_DEVICE ? 1 : 0, T.p, e_mat.p, e_sync.p, erad_avg.p, census_erad.p,
        clamp.p, err.p, precision == STORM_GPU_FP64 ? 1 : 0, xs.p, groups >= 2 ? 0 : 1, instview());
    // census energy for the next step's quantum, total segments, error flags
    // and counters: one synchronisation
    sum_async(census_erad.p, N, dsc.p + 1);
    if (!inst.empty())
      k_inst_sum<<<static_cast<uint32_t>(inst.size()), 256, 0, stream>>>(d_inst_first.p, census_erad.p, d_inst_cen.p);
    k_total_u32<<<std::min<uint32_t>(blocks_for(N, 256), 1024u), 256, 0, stream>>>(N, tal.segs, leaving.p + 3);
    const unsigned long long *hp = read_scalars({{dsc.p + 1, 8u}, {err.p, 4u}, {leaving.p, 8u}, {leaving.p + 1, 8u},
                                                 {leaving.p + 2, 8u}, {leaving.p + 3, 8u}, {leaving.p + 4, 8u}});
    const double hcensus = as_double(hp[0]);
    const uint32_t herr = static_cast<uint32_t>(hp[1]);
    unsigned long long hstats[5];
    for (int k = 0; k < 5; ++k) hstats[k] = hp[2 + k];
    census_total_prev = hcensus;
    counters.segments = hstats[3];
    counters.max_packet_blocks = hstats[4];
    counters.leaving_particles = hstats[0];
    counters.random_walk_steps = hstats[1];
    counters.ddmc_steps = hstats[2];
    eq_prev = eq;
    last_eq_ = eq;
    last_et_ = et;
    last_dt_ = dt;
    // a step that stopped here has replaced the census: the engine takes no
    // further step (status 5 again) and is to be destroyed
    if (herr & (kErrFleck | kErrNegativeEnergy | kErrTallyOverflow)) failed_ = true;
    if (herr & kErrFleck) throw EngineError(5, "Invalid factor fleck in RadiationIMC::preStep");
    if (herr & kErrNegativeEnergy) throw EngineError(5, "Negative material internal energy in RadiationIMC");
    if (herr & kErrTallyOverflow)
      throw EngineError(5, "a cell received 2^28 tally additions in one step; the material state was not updated");
    tune_after_step(static_cast<double>(counters.segments), counters.seconds_transport);
    auto t_end = std::chrono::steady_clock::now();
    counters.seconds_step = std::chrono::duration<double>(t_end - t_start).count();
  }

  int32_t last_eq_ = 0, last_et_ = 0;
  double last_dt_ = 0.0;

  // ---------------------------------------------------------- step_noalloc
  // The same step with every transient buffer carved out of the library's
  // fixed device scratch (storm_gpu_step_noalloc).  Transient buffers carry
  // nothing from one step to the next: small fixed-size ones (reduction
  // partials, counters, mailboxes, error flags) and ones sized by the step's
  // packet count (packets in flight, the census before the comb, the comb's
  // keys and prefix sums, rocPRIM's temporary storage).  The step clears
  // (with fill kernels) every counter that must start at zero, so what an
  // earlier call left in the scratch does not matter.  Buffers of the
  // engine's state are never allocated here: the census buffer must already
  // hold this step's census (EngineError 4 otherwise, raised before the
  // comb writes the census and before the material update; the step counter
  // is restored so that a following step repeats this one exactly).
  void *scratch_ptr = nullptr;
  template <typename F>
  void for_fixed_transient(F &&f) {
    f(partials);
    f(counters_u32);
    f(err);
    f(qexp);
    f(mail);
    f(leaving);
  }
  template <typename F>
  void for_sized_transient(F &&f) {
    for (PacketBufs<Real> *b : {&pa, &pb}) {
      f(b->pos);
      f(b->dir);
      f(b->w);
      f(b->meta);
      f(b->nu);
    }
    f(cpre.pos);
    f(cpre.dir);
    f(cpre.w);
    f(cpre.meta);
    f(cstage.pos);
    f(cstage.dir);
    f(cstage.w);
    f(cstage.meta);
    f(keys);
    f(keys_out);
    f(vals);
    f(vals_out);
    f(limbs_sorted);
    f(wsorted);
    f(cum);
    f(ncopy);
    f(offs64);
    f(temp);
  }
  static constexpr size_t kTransientBufs = 6 + 2 * 5 + 4 + 4 + 10;

  void step_noalloc(double dt) override {
    if (keep_inputs)
      throw EngineError(1, "storm_gpu_step_noalloc: the debug copy of the transport inputs allocates host memory; "
                           "switch it off first (storm_gpu_debug_keep_transport_inputs)");
    if (!scratch_ptr) throw EngineError(2, "storm_gpu_step_noalloc: no scratch address for this device");
    if (!host_queue_.empty())
      throw EngineError(1, "storm_gpu_step_noalloc: host-made source packets are queued (their upload allocates); "
                           "use storm_gpu_step");
    ScratchArena arena{static_cast<unsigned char *>(scratch_ptr), kNoAllocScratchBytes, 0};
    SavedBuf saved[kTransientBufs];
    size_t k = 0;
    for_fixed_transient([&](auto &b) { saved[k++] = b.lend(); });
    for_sized_transient([&](auto &b) { saved[k++] = b.lend(); });
    auto restore = [&] {
      t_arena = nullptr;
      size_t q = 0;
      for_fixed_transient([&](auto &b) { b.give_back(saved[q++]); });
      for_sized_transient([&](auto &b) { b.give_back(saved[q++]); });
    };
    const uint32_t step_before = step_index;
    t_arena = &arena;
    try {
      // the fixed ones are used without an ensure() in the step
      k = 0;
      for_fixed_transient([&](auto &b) { b.ensure(saved[k++].n); });
      step(dt);
    } catch (...) {
      // nothing of this step may still run on the scratch
      (void)hipStreamSynchronize(stream);
      restore();
      step_index = step_before;
      throw;
    }
    restore();
  }

  // Comb (or pass-through) from `cpre` into `cnext`; returns the census size.
  uint64_t comb(uint32_t n, int32_t eq) {
    if (n == 0) return 0;
    const MeshDims d = dims();
    keys.ensure(n);
    keys_out.ensure(n);
    vals.ensure(n);
    vals_out.ensure(n);
    limbs_sorted.ensure(n);
    wsorted.ensure(n);
    cum.ensure(n);
    ncopy.ensure(n);
    offs64.ensure(n);
    const uint32_t salt = static_cast<uint32_t>(seed) * 0x2545F491u + step_index * 0x9E3779B9u;
    // 32-bit keys: the cell's Z-order code (morton.total bits, and 2^total
    // for dropped packets), the photon group for the stratified comb
    // (gbits), then the top pbits bits of the rank; ties by k_comb_ties.
    // (24-bit keys save one radix pass but leave so few rank bits that the
    // ties in the crowded cells of a cold start took 0.8 ms to order.)
    const uint32_t G = strat_groups, gbits = strat_gbits;
    const uint32_t pbits = 31u - morton.total - gbits;
    const CombUnits cu{pbits, gbits, G};
    k_comb_keys<Real><<<blocks_for(n, 256), 256, 0, stream>>>(n, cpre.soa(), morton, salt, pbits, keys.p, vals.p,
                                                              gbits, G > 1 ? bounds_kev.p : nullptr,
                                                              static_cast<int32_t>(G));
    {
      auto q = [&](void *tp, size_t &bytes) {
        HIPCHECK(rocprim::radix_sort_pairs(tp, bytes, keys.p, keys_out.p, vals.p, vals_out.p, static_cast<size_t>(n),
                                           0, 32, stream));
      };
      ensure_temp(q);
    }
    k_comb_ties<<<blocks_for(n, 256), 256, 0, stream>>>(n, keys_out.p, vals_out.p,
                                                        ((1u << morton.total) << gbits) << pbits, cpre.meta.p, salt);
    // runs of cells (bins) over the sorted census, weight limbs and their
    // prefix sums
    zero_u32(run_end.p, static_cast<size_t>(N) * G);
    k_comb_sorted<Real><<<blocks_for(n, 256), 256, 0, stream>>>(n, keys_out.p, vals_out.p, cu, cpre.soa(), eq, morton, d,
                                                                wsorted.p, limbs_sorted.p, run_start.p, run_end.p);
    {
      auto q = [&](void *tp, size_t &bytes) {
        HIPCHECK(rocprim::inclusive_scan(tp, bytes, limbs_sorted.p, cum.p, static_cast<size_t>(n),
                                         rocprim::plus<uint64_t>(), stream));
      };
      ensure_temp(q);
    }
    // per-cell targets; W_c into the ledger's census_pre tally
    if (!inst.empty())
      k_inst_census<<<static_cast<uint32_t>(inst.size()), 256, 0, stream>>>(d_inst_first.p, run_start.p, run_end.p,
                                                                           cum.p, eq, d_inst_wtot.p);
    if (G > 1)
      k_comb_strat<<<blocks_for(N, 256), 256, 0, stream>>>(N, G, run_start.p, run_end.p, cum.p, n, imc.comb_nmin,
                                                           imc.comb_factor, imc.comb_min_per_group, eq, seed,
                                                           step_index, comb_cells.p, tal.cpre, tal.rem, tal.cdrop);
    else
      k_comb_cells<<<blocks_for(N, 256), 256, 0, stream>>>(N, run_start.p, run_end.p, cum.p, n, imc.comb_nmin,
                                                           imc.comb_factor, eq, imc.comb ? 1 : 0, seed, step_index,
                                                           comb_cells.p, tal.cpre, tal.rem, instview());
    k_comb_count<<<blocks_for(n, 256), 256, 0, stream>>>(n, keys_out.p, cu, wsorted.p, comb_cells.p, run_start.p,
                                                         run_end.p, cum.p, eq, morton, d, ncopy.p);
    {
      auto q = [&](void *tp, size_t &bytes) {
        HIPCHECK(rocprim::exclusive_scan(tp, bytes, ncopy.p, offs64.p, uint64_t(0), static_cast<size_t>(n),
                                         rocprim::plus<uint64_t>(), stream));
      };
      ensure_temp(q);
    }
    const unsigned long long *hc = read_scalars({{offs64.p + (n - 1), 8u}, {ncopy.p + (n - 1), 4u}});
    const uint64_t n_out = hc[0] + hc[1];
    if (n_out >= (1ull << 32) - 1) throw EngineError(4, "census larger than 2^32 packets");
    if (n_out == 0) return 0;
    cnext.ensure(n_out);
    k_comb_write<Real><<<blocks_for(n, 256), 256, 0, stream>>>(n, keys_out.p, vals_out.p, cu, cpre.soa(), wsorted.p,
                                                               comb_cells.p, ncopy.p, offs64.p, cnext.soa(), eq, morton,
                                                               d, tal);
    return n_out;
  }

  void keep_transport_inputs(uint64_t n, const StepConsts &kc) { keep_packets(pa, n, kc); }
  // Host copy of the first n packets of `src` with the cell data and step
  // constants they are tracked with.
  void keep_packets(PacketBufs<Real> &src, uint64_t n, const StepConsts &kc) {
    kept_packets.resize(n);
    std::vector<typename V4<Real>::T> hp(n), hd(n);
    std::vector<double> hw(n);
    std::vector<uint4> hm(n);
    std::vector<Real> hn(groups >= 2 ? n : 0);
    if (n) {
      if (groups >= 2) HIPCHECK(hipMemcpy(hn.data(), src.nu.p, n * sizeof(Real), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hp.data(), src.pos.p, n * sizeof(hp[0]), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hd.data(), src.dir.p, n * sizeof(hd[0]), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hw.data(), src.w.p, n * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hm.data(), src.meta.p, n * sizeof(uint4), hipMemcpyDeviceToHost));
    }
    for (uint64_t i = 0; i < n; ++i) {
      Packet<Real> &p = kept_packets[i];
      p.u[0] = hp[i].x;
      p.u[1] = hp[i].y;
      p.u[2] = hp[i].z;
      p.s = hp[i].w;
      p.o[0] = hd[i].x;
      p.o[1] = hd[i].y;
      p.o[2] = hd[i].z;
      p.wmin = hd[i].w;
      p.w = hw[i];
      p.ijk = hm[i].x;
      p.id = hm[i].y;
      p.ctr = hm[i].z;
      p.flags = hm[i].w;
      p.nu = groups >= 2 ? hn[i] : Real(0);
    }
    kept_xs.resize(N);
    HIPCHECK(hipMemcpy(kept_xs.data(), xs.p, N * sizeof(CellXS<Real>), hipMemcpyDeviceToHost));
    kept_consts = kc;
  }

  void debug_inputs(uint64_t *n, void *packets, void *cell_xs, int64_t *consts4) override {
    *n = kept_packets.size();
    if (packets) std::memcpy(packets, kept_packets.data(), kept_packets.size() * sizeof(Packet<Real>));
    if (cell_xs) {
      // two numbers per cell (sig_abs, sig_ev); the reciprocal is derived
      Real *o = static_cast<Real *>(cell_xs);
      for (size_t c = 0; c < kept_xs.size(); ++c) {
        o[2 * c] = kept_xs[c].sig_abs;
        o[2 * c + 1] = kept_xs[c].sig_ev;
      }
    }
    if (consts4) {
      consts4[0] = static_cast<int64_t>(kept_consts.seed);
      consts4[1] = kept_consts.step;
      consts4[2] = kept_consts.eq;
      consts4[3] = kept_consts.et;
    }
  }

  void debug_spectrum(int32_t *g, double *npow, void *bounds, void *cdf_out, void *sig_s_out) override {
    *g = groups;
    if (npow) *npow = nu_pow;
    if (groups < 2) return;
    const size_t G1 = static_cast<size_t>(groups) + 1;
    if (bounds) HIPCHECK(hipMemcpy(bounds, bounds_kev.p, G1 * sizeof(Real), hipMemcpyDeviceToHost));
    if (cdf_out) HIPCHECK(hipMemcpy(cdf_out, cdf.p, G1 * N * sizeof(Real), hipMemcpyDeviceToHost));
    if (sig_s_out) HIPCHECK(hipMemcpy(sig_s_out, sig_s_cell.p, N * sizeof(Real), hipMemcpyDeviceToHost));
  }

  void debug_accel(void *out, uint64_t *bytes) override {
    const uint64_t nb = accel_on ? static_cast<uint64_t>(N) * sizeof(CellAccel<Real>) : 0;
    *bytes = nb;
    if (out && nb) HIPCHECK(hipMemcpy(out, accel.p, nb, hipMemcpyDeviceToHost));
  }

  void debug_tallies(uint64_t *dep, uint64_t *rdep, uint64_t *pint, uint64_t *esc, uint32_t *segs) override {
    const size_t n2 = 2 * static_cast<size_t>(N);
    if (dep) HIPCHECK(hipMemcpy(dep, tal.dep, n2 * 8, hipMemcpyDeviceToHost));
    if (rdep) HIPCHECK(hipMemcpy(rdep, tal.rdep, n2 * 8, hipMemcpyDeviceToHost));
    if (pint) HIPCHECK(hipMemcpy(pint, tal.pint, n2 * 8, hipMemcpyDeviceToHost));
    if (esc) HIPCHECK(hipMemcpy(esc, tal.esc, n2 * 8, hipMemcpyDeviceToHost));
    if (segs) HIPCHECK(hipMemcpy(segs, tal.segs, N * 4, hipMemcpyDeviceToHost));
  }

  void get_cells(double *To, double *eo, double *erad, double *cen, double *fl, double *pl) override {
    HIPCHECK(hipStreamSynchronize(stream));
    auto dl = [&](double *dst, const DevBuf<double> &b) {
      if (dst) HIPCHECK(hipMemcpy(dst, b.p, N * sizeof(double), hipMemcpyDeviceToHost));
    };
    dl(To, T);
    dl(eo, e_mat);
    dl(erad, erad_avg);
    dl(cen, census_erad);
    dl(fl, fleck);
    dl(pl, sig_a);
  }

  // getEgTimeAvg (RadiationIMC.hpp:433-434, normalised as :1896-1903): the
  // per-group path integral of the last step over c dt V, [cell * G + g]
  void get_group_erad(double *out) override {
    if (!eg_on) throw EngineError(1, "the per-group tally is off (storm_group_desc.with_eg_time_avg)");
    HIPCHECK(hipStreamSynchronize(stream));
    const size_t G = static_cast<size_t>(groups), n = static_cast<size_t>(N) * G;
    std::vector<double> v(n);
    if (precision == STORM_GPU_FP64) {
      HIPCHECK(hipMemcpy(v.data(), eg_f64.p, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<unsigned long long> q(n);
      HIPCHECK(hipMemcpy(q.data(), eg_u64.p, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) v[i] = std::ldexp(static_cast<double>(q[i]), last_et_);
    }
    for (uint32_t c = 0; c < N; ++c)
      for (size_t g = 0; g < G; ++g) out[c * G + g] = v[c * G + g] / (units::clight * last_dt_ * Vh[c]);
  }

  void get_segments(uint64_t *out) override {
    std::vector<uint32_t> h(N);
    HIPCHECK(hipMemcpy(h.data(), tal.segs, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < N; ++c) out[c] = h[c];
  }
  void get_begin(uint64_t *out) override {
    std::vector<uint32_t> h(N);
    HIPCHECK(hipMemcpy(h.data(), tal.begin, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < N; ++c) out[c] = h[c];
  }

  static storm_gpu_u128 u128(unsigned __int128 v) {
    storm_gpu_u128 r;
    r.lo = static_cast<uint64_t>(v);
    r.hi = static_cast<uint64_t>(v >> 64);
    return r;
  }

  void get_ledger(storm_gpu_ledger *L) override { ledger_range(L, -1); }

  // The ledger over all cells (k < 0) or over the cells and source faces of
  // instance k of a batch.  Integer fields count quanta of 2^eq erg; over a
  // whole batch they count the step's quantum in the instances' units (each
  // balance is exact, the erg figures are per instance).
  void ledger_range(storm_gpu_ledger *L, int64_t k) {
    const uint32_t c0 = k < 0 ? 0u : inst_first_h[k], c1 = k < 0 ? N : inst_first_h[k + 1];
    const int32_t sh = k < 0 ? 0 : inst_shift_now[k], shp = k < 0 ? 0 : inst_shift_last[k];
    const size_t n2 = 2 * static_cast<size_t>(N);
    std::vector<unsigned long long> h(tally_u64.n), em(n2), rq(n2);
    HIPCHECK(hipMemcpy(h.data(), tally_u64.p, h.size() * 8, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(em.data(), emit_limbs.p, n2 * 8, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(rq.data(), requant.p, n2 * 8, hipMemcpyDeviceToHost));
    auto sum = [&](const unsigned long long *a) {
      unsigned __int128 s = 0;
      for (uint32_t c = c0; c < c1; ++c) s += (static_cast<unsigned __int128>(a[2 * c + 1]) << fx::kLoBits) + a[2 * c];
      return s;
    };
    const unsigned long long *b = h.data();
    std::memset(L, 0, sizeof(*L));
    L->census_in = u128(sum(b + 3 * n2));
    L->emitted = u128(sum(em.data()));
    L->deposited = u128(sum(b) + sum(b + 7 * n2));
    L->escaped = u128(sum(b + 2 * n2));
    if (precision == STORM_GPU_FP64) {
      // FP64 engine: deposits and escapes are FP64 sums, reported rounded to quanta
      std::vector<double> f(tally_f64.n);
      HIPCHECK(hipMemcpy(f.data(), tally_f64.p, f.size() * sizeof(double), hipMemcpyDeviceToHost));
      long double dsum = 0.0L, esum = 0.0L;
      for (uint32_t c = c0; c < c1; ++c) {
        dsum += f[c];
        esum += f[2 * static_cast<size_t>(N) + c];
      }
      const long double Q = std::ldexp(1.0L, last_eq_);
      auto to_u128 = [](long double x) {
        if (!(x > 0.0L)) return static_cast<unsigned __int128>(0);
        const long double two64 = std::ldexp(1.0L, 64);
        long double hi = std::floor(x / two64);
        long double lo = std::floor(x - hi * two64 + 0.5L);
        return (static_cast<unsigned __int128>(static_cast<uint64_t>(hi)) << 64) +
               static_cast<unsigned __int128>(static_cast<uint64_t>(lo));
      };
      L->deposited = u128(to_u128(dsum / Q));
      L->escaped = u128(to_u128(esum / Q));
    }
    L->census_pre = u128(sum(b + 4 * n2));
    L->census_out = u128(sum(b + 5 * n2));
    L->comb_rem = u128(sum(b + 6 * n2));
    L->comb_dropped = u128(sum(b + 8 * n2));
    // boundary source: Npc * w_face per face, exact
    unsigned __int128 src = 0;
    if (last_nfaces_) {
      std::vector<double> wf(last_nfaces_);
      HIPCHECK(hipMemcpy(wf.data(), wface.p, wf.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (uint32_t f = 0; f < last_nfaces_; ++f) {
        if (k >= 0 && ut.src_instance[f] != static_cast<uint32_t>(k)) continue;
        const uint64_t npc = unstructured ? ut.src[f].count : bnd.source_packets_per_face;
        fx::Limbs l = fx::limbs_of(wf[f], last_eq_);
        src += ((static_cast<unsigned __int128>(l.hi) << fx::kLoBits) + l.lo) * npc;
      }
    }
    for (const auto &hp : host_last_) {
      if (hp.first < c0 || hp.first >= c1) continue;
      fx::Limbs l = fx::limbs_of(hp.second, last_eq_);
      src += (static_cast<unsigned __int128>(l.hi) << fx::kLoBits) + l.lo;
    }
    L->source = u128(src);
    L->eq = last_eq_ - sh;
    L->eq_prev = eq_prev_ledger_ - shp;
    double rq_sum = 0.0;
    for (uint32_t c = c0; c < c1; ++c) rq_sum += fx::to_double(rq[2 * c], rq[2 * c + 1], L->eq_prev);
    L->requant_rem = rq_sum;
    std::vector<double> cl(N), em2(N);
    HIPCHECK(hipMemcpy(cl.data(), clamp.p, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(em2.data(), e_mat.p, N * sizeof(double), hipMemcpyDeviceToHost));
    double cls = 0.0, ms = 0.0;
    for (uint32_t c = c0; c < c1; ++c) {
      cls += cl[c];
      ms += em2[c];
    }
    L->clamp = cls;
    L->material_after = ms;
    if (k < 0) HIPCHECK(hipMemcpy(&L->material_before, dsc.p, sizeof(double), hipMemcpyDeviceToHost));
    else HIPCHECK(hipMemcpy(&L->material_before, d_inst_mat0.p + k, sizeof(double), hipMemcpyDeviceToHost));
  }
  void get_instance_ledger(uint32_t k, storm_gpu_ledger *L) override {
    if (k >= inst.size()) throw EngineError(1, "no such instance");
    ledger_range(L, static_cast<int64_t>(k));
  }
  int32_t eq_prev_ledger_ = 0;

  void set_cell_state(const double *Ti, const double *sa, const double *ss, const double *cv,
                      const double *e) override {
    auto up = [&](DevBuf<double> &d, const double *src) {
      d.ensure(N);
      HIPCHECK(hipMemcpy(d.p, src, N * sizeof(double), hipMemcpyHostToDevice));
    };
    if (Ti) up(T, Ti);
    if (sa) {
      up(host_sa, sa);
      have_host_sa = true;
    }
    if (ss) {
      up(host_ss, ss);
      have_host_ss = true;
    }
    if (cv) {
      up(host_cv, cv);
      have_host_cv = true;
    }
    if (e) {
      up(e_mat, e);
      up(e_sync, e);
    }
  }

  uint64_t census_size() override { return n_census; }

  // global position of a census packet (cell lower corner + u * width)
  void export_census(storm_packet_record *out, uint64_t cap, uint64_t *nw) override {
    const uint64_t n = std::min<uint64_t>(cap, n_census);
    std::vector<typename V4<Real>::T> hp(n), hd(n);
    std::vector<double> hw(n);
    std::vector<uint2> hm(n);
    if (n) {
      HIPCHECK(hipMemcpy(hp.data(), cnext.pos.p, n * sizeof(hp[0]), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hd.data(), cnext.dir.p, n * sizeof(hd[0]), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hw.data(), cnext.w.p, n * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hm.data(), cnext.meta.p, n * sizeof(uint2), hipMemcpyDeviceToHost));
    }
    for (uint64_t q = 0; q < n; ++q) {
      storm_packet_record &r = out[q];
      if (unstructured) {
        // position relative to the cell's mesh point
        const uint32_t c = hm[q].x;
        r.x = gen_h[3 * static_cast<size_t>(c)] + static_cast<double>(hp[q].x);
        r.y = gen_h[3 * static_cast<size_t>(c) + 1] + static_cast<double>(hp[q].y);
        r.z = gen_h[3 * static_cast<size_t>(c) + 2] + static_cast<double>(hp[q].z);
        r.vx = static_cast<double>(hd[q].x) * units::clight;
        r.vy = static_cast<double>(hd[q].y) * units::clight;
        r.vz = static_cast<double>(hd[q].z) * units::clight;
        // batches: from the instance's unit to erg
        r.weight = inst.empty() ? hw[q] : std::ldexp(hw[q], -inst_shift_now[cell_inst_h[c]]);
        r.frequency = groups >= 2 ? static_cast<double>(hp[q].w) * units::kev : 0.0;
        r.cell = c;
        r.id = ~0ull;
        continue;
      }
      uint32_t i = ijk_i(hm[q].x), j = ijk_j(hm[q].x), k = ijk_k(hm[q].x);
      r.x = xe[i] + static_cast<double>(hp[q].x) * (xe[i + 1] - xe[i]);
      r.y = ye[j] + static_cast<double>(hp[q].y) * (ye[j + 1] - ye[j]);
      r.z = ze[k] + static_cast<double>(hp[q].z) * (ze[k + 1] - ze[k]);
      r.vx = static_cast<double>(hd[q].x) * units::clight;
      r.vy = static_cast<double>(hd[q].y) * units::clight;
      r.vz = static_cast<double>(hd[q].z) * units::clight;
      r.weight = hw[q];
      r.frequency = groups >= 2 ? static_cast<double>(hp[q].w) * units::kev : 0.0;
      r.cell = (static_cast<uint64_t>(i) * ny + j) * nz + k;
      r.id = ~0ull;
    }
    *nw = n;
  }

  // Host-made new packets for the next step (BoundaryCondition::
  // generateNewBoundaryParticles, RadiationIMC.hpp:1683-1692): each with its
  // own time left; weights in erg, put on the step's quantum then.
  void queue_source_packets(const storm_packet_record *in, const double *time_left, uint64_t n) override {
    if (host_queue_.size() + n >= (1ull << 31)) throw EngineError(4, "more than 2^31 queued source packets");
    auto locate = [](const std::vector<double> &e, double v) {
      auto it = std::upper_bound(e.begin(), e.end(), v);
      int idx = static_cast<int>(it - e.begin()) - 1;
      return std::clamp(idx, 0, static_cast<int>(e.size()) - 2);
    };
    host_queue_.reserve(host_queue_.size() + n);
    for (uint64_t q = 0; q < n; ++q) {
      const storm_packet_record &r = in[q];
      if (!std::isfinite(r.weight) || r.weight < 0.0) throw EngineError(1, "source packet weight is negative or not finite");
      if (!(time_left[q] >= 0.0) || !std::isfinite(time_left[q])) throw EngineError(1, "source packet time left is invalid");
      const double vn = std::sqrt(r.vx * r.vx + r.vy * r.vy + r.vz * r.vz);
      if (!(vn > 0.0)) throw EngineError(1, "source packet with zero velocity");
      HostPacket h{};
      h.o[0] = r.vx / vn;
      h.o[1] = r.vy / vn;
      h.o[2] = r.vz / vn;
      h.w = r.weight;
      h.s = time_left[q] * units::clight;
      h.nu = 0.0f;
      if (groups >= 2) h.nu = static_cast<float>(std::min(std::max(r.frequency, bounds_h.front()), bounds_h.back()) / units::kev);
      if (unstructured) {
        if (r.cell >= N) throw EngineError(1, "source packet on a polyhedral mesh needs its cell index");
        const uint32_t c = static_cast<uint32_t>(r.cell);
        h.u[0] = r.x - gen_h[3 * static_cast<size_t>(c)];
        h.u[1] = r.y - gen_h[3 * static_cast<size_t>(c) + 1];
        h.u[2] = r.z - gen_h[3 * static_cast<size_t>(c) + 2];
        h.cell = c;
        h.code = c;
      } else {
        const int i = r.cell < N ? static_cast<int>(r.cell / (static_cast<uint64_t>(ny) * nz)) : locate(xe, r.x);
        const int j = r.cell < N ? static_cast<int>((r.cell / nz) % ny) : locate(ye, r.y);
        const int k = r.cell < N ? static_cast<int>(r.cell % nz) : locate(ze, r.z);
        auto frac = [](double v, double lo, double hi) {
          const double f = (v - lo) / (hi - lo);
          return f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
        };
        h.u[0] = frac(r.x, xe[i], xe[i + 1]);
        h.u[1] = frac(r.y, ye[j], ye[j + 1]);
        h.u[2] = frac(r.z, ze[k], ze[k + 1]);
        h.cell = static_cast<uint32_t>((static_cast<uint64_t>(i) * ny + j) * nz + k);
        h.code = pack_ijk(i, j, k);
      }
      host_queue_.push_back(h);
    }
  }

  void import_census(const storm_packet_record *in, uint64_t n) override {
    if (n >= (1ull << 32) - 1) throw EngineError(4, "census larger than 2^32 packets");
    std::vector<typename V4<Real>::T> hp(n), hd(n);
    std::vector<double> hw(n);
    std::vector<uint2> hm(n);
    auto locate = [](const std::vector<double> &e, double v) {
      auto it = std::upper_bound(e.begin(), e.end(), v);
      int idx = static_cast<int>(it - e.begin()) - 1;
      return std::clamp(idx, 0, static_cast<int>(e.size()) - 2);
    };
    std::vector<uint32_t> hcell(n);
    double tot = 0.0;
    for (uint64_t q = 0; q < n; ++q) {
      const storm_packet_record &r = in[q];
      if (unstructured) {
        if (r.cell >= N) throw EngineError(1, "census packet on a polyhedral mesh needs its cell index");
        const uint32_t c = static_cast<uint32_t>(r.cell);
        hp[q].x = static_cast<Real>(r.x - gen_h[3 * static_cast<size_t>(c)]);
        hp[q].y = static_cast<Real>(r.y - gen_h[3 * static_cast<size_t>(c) + 1]);
        hp[q].z = static_cast<Real>(r.z - gen_h[3 * static_cast<size_t>(c) + 2]);
        hp[q].w = Real(0);
        if (groups >= 2) {
          const double E = std::min(std::max(r.frequency, bounds_h.front()), bounds_h.back());
          hp[q].w = static_cast<Real>(E / units::kev);
        }
        double vn = std::sqrt(r.vx * r.vx + r.vy * r.vy + r.vz * r.vz);
        if (!(vn > 0.0)) throw EngineError(1, "census packet with zero velocity");
        hd[q].x = static_cast<Real>(r.vx / vn);
        hd[q].y = static_cast<Real>(r.vy / vn);
        hd[q].z = static_cast<Real>(r.vz / vn);
        hd[q].w = Real(0);
        if (!std::isfinite(r.weight) || r.weight < 0.0) throw EngineError(1, "census packet weight is negative or not finite");
        hw[q] = r.weight;
        tot += r.weight;
        hcell[q] = c;
        hm[q] = make_uint2(c, static_cast<uint32_t>(q));
        continue;
      }
      int i = r.cell < N ? static_cast<int>(r.cell / (static_cast<uint64_t>(ny) * nz)) : locate(xe, r.x);
      int j = r.cell < N ? static_cast<int>((r.cell / nz) % ny) : locate(ye, r.y);
      int k = r.cell < N ? static_cast<int>(r.cell % nz) : locate(ze, r.z);
      auto frac = [](double v, double lo, double hi) {
        double f = (v - lo) / (hi - lo);
        return f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
      };
      hp[q].x = static_cast<Real>(frac(r.x, xe[i], xe[i + 1]));
      hp[q].y = static_cast<Real>(frac(r.y, ye[j], ye[j + 1]));
      hp[q].z = static_cast<Real>(frac(r.z, ze[k], ze[k + 1]));
      hp[q].w = Real(0);
      if (groups >= 2) {
        // photon energy in keV, clamped to the group range (clampFrequencyToBounds)
        const double E = std::min(std::max(r.frequency, bounds_h.front()), bounds_h.back());
        hp[q].w = static_cast<Real>(E / units::kev);
      }
      double vn = std::sqrt(r.vx * r.vx + r.vy * r.vy + r.vz * r.vz);
      if (!(vn > 0.0)) throw EngineError(1, "census packet with zero velocity");
      hd[q].x = static_cast<Real>(r.vx / vn);
      hd[q].y = static_cast<Real>(r.vy / vn);
      hd[q].z = static_cast<Real>(r.vz / vn);
      hd[q].w = Real(0);
      if (!std::isfinite(r.weight) || r.weight < 0.0) throw EngineError(1, "census packet weight is negative or not finite");
      hw[q] = r.weight;
      tot += r.weight;
      hcell[q] = static_cast<uint32_t>((static_cast<uint64_t>(i) * ny + j) * nz + k);
      hm[q] = make_uint2(pack_ijk(i, j, k), static_cast<uint32_t>(q));
    }
    // imported weights are arbitrary doubles: put them on a quantum of their
    // own total; the sub-quantum remainders go to the material of their cells
    const int32_t eq_imp = tot > 0.0 ? fx::ceil_log2(tot) - fx::kHeadroomBits : 0;
    std::vector<double> rem(N, 0.0);
    bool any_rem = false;
    for (uint64_t q = 0; q < n; ++q) {
      double wq = fx::quantize_down(hw[q], eq_imp);
      if (wq != hw[q]) {
        rem[hcell[q]] += hw[q] - wq;
        any_rem = true;
      }
      hw[q] = wq;
    }
    if (any_rem) {
      std::vector<double> e(N);
      HIPCHECK(hipMemcpy(e.data(), e_mat.p, N * sizeof(double), hipMemcpyDeviceToHost));
      for (uint32_t c = 0; c < N; ++c) e[c] += rem[c];
      HIPCHECK(hipMemcpy(e_mat.p, e.data(), N * sizeof(double), hipMemcpyHostToDevice));
    }
    cnext.ensure(std::max<uint64_t>(n, 1));
    if (n) {
      HIPCHECK(hipMemcpy(cnext.pos.p, hp.data(), n * sizeof(hp[0]), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(cnext.dir.p, hd.data(), n * sizeof(hd[0]), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(cnext.w.p, hw.data(), n * sizeof(double), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(cnext.meta.p, hm.data(), n * sizeof(uint2), hipMemcpyHostToDevice));
    }
    n_census = n;
    double totq = 0.0;
    for (double w : hw) totq += w;
    census_total_prev = totq;
    eq_prev = eq_imp;
    if (!inst.empty()) {
      // imported weights are in erg: every instance's unit is 1 erg until the
      // next step picks the new ones
      const size_t K = inst.size();
      std::vector<double> ck(K, 0.0);
      for (uint64_t q = 0; q < n; ++q) ck[cell_inst_h[hcell[q]]] += hw[q];
      inst_shift_now.assign(K, 0);
      HIPCHECK(hipMemcpy(d_inst_shift.p, inst_shift_now.data(), K * sizeof(int32_t), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(d_inst_cen.p, ck.data(), K * sizeof(double), hipMemcpyHostToDevice));
    }
  }
};

EngineBase *as_engine(storm_gpu_engine *e) { return reinterpret_cast<EngineBase *>(e); }

template <typename F>
int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const EngineError &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return 9;
  }
}

}  // namespace gpu
}  // namespace storm

using namespace storm::gpu;

extern "C" {

const char *storm_gpu_last_error(void) { return g_last_error.c_str(); }
int storm_gpu_abi_version(void) { return STORM_GPU_ABI_VERSION; }

int storm_gpu_device_info(char *buf, size_t len) {
  return guarded([&] {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, len, "%s | %s | CUs=%d | clock=%d kHz | mem=%zu MiB", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount, prop.clockRate, static_cast<size_t>(prop.totalGlobalMem >> 20));
  });
}

int storm_gpu_calibration_copy(uint64_t bytes, double *ms) {
  return guarded([&] {
    if (bytes == 0 || bytes % 16u != 0u) throw EngineError(1, "copy size must be a positive multiple of 16 bytes");
    if (bytes > (uint64_t(16) << 30)) throw EngineError(1, "copy size above 16 GiB");
    const uint64_t n = bytes / 16u;
    void *a = nullptr, *b = nullptr;
    HIPCHECK(hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) {
      (void)hipFree(a);
      throw EngineError(2, "device allocation failed");
    }
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    const uint32_t block = 256;
    const uint64_t grid = (n + block - 1) / block;
    k_fill_u32<<<static_cast<uint32_t>((bytes / 4 + block - 1) / block), block>>>(static_cast<uint32_t *>(a),
                                                                                   bytes / 4, 0x5a5a5a5au);
    HIPCHECK(hipEventRecord(e0, nullptr));
    k_copy_calibration<<<static_cast<uint32_t>(grid), block>>>(static_cast<const uint4 *>(a), static_cast<uint4 *>(b),
                                                               n);
    HIPCHECK(hipEventRecord(e1, nullptr));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    HIPCHECK(hipFree(a));
    HIPCHECK(hipFree(b));
  });
}

int storm_gpu_create(const storm_problem *prob, int32_t precision, int32_t material_mode, storm_gpu_engine **out) {
  return guarded([&] {
    if (!prob || !out) throw EngineError(1, "null argument");
    if (!prob->T0) throw EngineError(1, "T0 is required");
    if (material_mode != STORM_GPU_MATERIAL_DEVICE && material_mode != STORM_GPU_MATERIAL_HOST)
      throw EngineError(1, "unknown material mode");
    EngineBase *e = nullptr;
    if (precision == STORM_GPU_MIXED)
      e = new Engine<float>(prob, precision, material_mode);
    else if (precision == STORM_GPU_FP64)
      e = new Engine<double>(prob, precision, material_mode);
    else
      throw EngineError(1, "unknown precision");
    *out = reinterpret_cast<storm_gpu_engine *>(e);
  });
}

int storm_gpu_destroy(storm_gpu_engine *eng) {
  return guarded([&] { delete as_engine(eng); });
}

int storm_gpu_step(storm_gpu_engine *eng, double dt) {
  return guarded([&] { as_engine(eng)->step(dt); });
}

int storm_gpu_step_noalloc(storm_gpu_engine *eng, double dt) {
  return guarded([&] { as_engine(eng)->step_noalloc(dt); });
}

int storm_gpu_set_source_temperature(storm_gpu_engine *eng, double T) {
  return guarded([&] { as_engine(eng)->source_T = T; });
}

int storm_gpu_set_cell_state(storm_gpu_engine *eng, const double *T, const double *sa, const double *ss,
                             const double *cv, const double *e) {
  return guarded([&] { as_engine(eng)->set_cell_state(T, sa, ss, cv, e); });
}

int storm_gpu_get_group_erad(storm_gpu_engine *eng, double *out) {
  return guarded([&] { as_engine(eng)->get_group_erad(out); });
}

int storm_gpu_get_cells(storm_gpu_engine *eng, double *T, double *e, double *erad, double *cen, double *fl,
                        double *pl) {
  return guarded([&] { as_engine(eng)->get_cells(T, e, erad, cen, fl, pl); });
}

int storm_gpu_get_segments(storm_gpu_engine *eng, uint64_t *per_cell) {
  return guarded([&] { as_engine(eng)->get_segments(per_cell); });
}

int storm_gpu_get_beginning_counts(storm_gpu_engine *eng, uint64_t *per_cell) {
  return guarded([&] { as_engine(eng)->get_begin(per_cell); });
}

int storm_gpu_get_counters(storm_gpu_engine *eng, storm_gpu_counters *out) {
  return guarded([&] { *out = as_engine(eng)->counters; });
}

int storm_gpu_get_ledger(storm_gpu_engine *eng, storm_gpu_ledger *out) {
  return guarded([&] { as_engine(eng)->get_ledger(out); });
}

int storm_gpu_get_instance_ledger(storm_gpu_engine *eng, uint32_t k, storm_gpu_ledger *out) {
  return guarded([&] { as_engine(eng)->get_instance_ledger(k, out); });
}

int storm_gpu_census_size(storm_gpu_engine *eng, uint64_t *n) {
  return guarded([&] { *n = as_engine(eng)->census_size(); });
}

int storm_gpu_export_census(storm_gpu_engine *eng, storm_packet_record *out, uint64_t cap, uint64_t *nw) {
  return guarded([&] { as_engine(eng)->export_census(out, cap, nw); });
}

int storm_gpu_import_census(storm_gpu_engine *eng, const storm_packet_record *in, uint64_t n) {
  return guarded([&] { as_engine(eng)->import_census(in, n); });
}

int storm_gpu_queue_source_packets(storm_gpu_engine *eng, const storm_packet_record *in, const double *time_left,
                                   uint64_t n) {
  return guarded([&] { as_engine(eng)->queue_source_packets(in, time_left, n); });
}

int storm_gpu_set_tuning(storm_gpu_engine *eng, int32_t block_size, int32_t segments_per_pass, int32_t schedule) {
  return guarded([&] {
    if (block_size != 0 && block_size != 64 && block_size != 128 && block_size != 256 && block_size != 512)
      throw EngineError(1, "block size must be 0 (chosen on line), 64, 128, 256 or 512");
    if (segments_per_pass < 1) throw EngineError(1, "segments per pass must be >= 1");
    if (schedule != STORM_GPU_SCHEDULE_PASS && schedule != STORM_GPU_SCHEDULE_REFILL && schedule != -1)
      throw EngineError(1, "schedule must be STORM_GPU_SCHEDULE_PASS, STORM_GPU_SCHEDULE_REFILL or -1");
    EngineBase *e = as_engine(eng);
    e->seg_per_pass = segments_per_pass;
    if (schedule != -1) e->schedule = schedule;
    e->watch_reset();
    if (block_size == 0) {
      e->tune_start();
      e->watch = true;
    } else {
      e->tune_phase = -1;
      e->block = block_size;
      e->watch = false;
    }
  });
}

int storm_gpu_set_chunks_per_wave(storm_gpu_engine *eng, int32_t chunks) {
  return guarded([&] {
    if (chunks != 1 && chunks != 2 && chunks != 4 && chunks != 8 && chunks != 16)
      throw EngineError(1, "chunks per wavefront must be 1, 2, 4, 8 or 16");
    as_engine(eng)->chunks_ref() = chunks;
  });
}

int storm_gpu_get_launch(storm_gpu_engine *eng, int32_t *block_size, int32_t *chunks_per_wave, int32_t *schedule,
                         int32_t *tuning) {
  return guarded([&] {
    EngineBase *e = as_engine(eng);
    *block_size = e->block;
    *chunks_per_wave = e->chunks_per_wave();
    *schedule = e->schedule;
    *tuning = e->tune_phase >= 0 ? 1 : 0;
  });
}

int storm_gpu_debug_keep_transport_inputs(storm_gpu_engine *eng, int32_t enable) {
  return guarded([&] { as_engine(eng)->keep_inputs = enable != 0; });
}

int storm_gpu_debug_transport_inputs(storm_gpu_engine *eng, uint64_t *n, void *packets, void *cell_xs,
                                     int64_t *consts4) {
  return guarded([&] { as_engine(eng)->debug_inputs(n, packets, cell_xs, consts4); });
}

int storm_gpu_debug_accel(storm_gpu_engine *eng, void *out, uint64_t *bytes) {
  return guarded([&] { as_engine(eng)->debug_accel(out, bytes); });
}

int storm_gpu_debug_spectrum(storm_gpu_engine *eng, int32_t *groups, double *nu_pow, void *bounds_kev, void *cdf,
                             void *sig_s) {
  return guarded([&] {
    if (!groups) throw EngineError(1, "groups must not be NULL");
    as_engine(eng)->debug_spectrum(groups, nu_pow, bounds_kev, cdf, sig_s);
  });
}

int storm_gpu_debug_tallies(storm_gpu_engine *eng, uint64_t *dep, uint64_t *rdep, uint64_t *pint, uint64_t *esc,
                            uint32_t *segs) {
  return guarded([&] { as_engine(eng)->debug_tallies(dep, rdep, pint, esc, segs); });
}

int storm_voronoi_build(const double *points, uint32_t n, const double lo[3], const double hi[3],
                        storm_voronoi **out) {
  return guarded([&] {
    if (!points || !lo || !hi || !out) throw EngineError(1, "null argument");
    try {
      auto *m = new storm::voronoi::Mesh(storm::voronoi::build(points, n, lo, hi));
      *out = reinterpret_cast<storm_voronoi *>(m);
    } catch (const std::invalid_argument &e) {
      throw EngineError(1, e.what());
    }
  });
}

int storm_voronoi_mesh(const storm_voronoi *v, storm_unstructured_mesh *out) {
  return guarded([&] {
    if (!v || !out) throw EngineError(1, "null argument");
    const auto *m = reinterpret_cast<const storm::voronoi::Mesh *>(v);
    storm_unstructured_mesh u{};
    u.ncells = m->ncells;
    u.nfaces = m->nfaces();
    for (int a = 0; a < 3; ++a) {
      u.box_lo[a] = m->lo[a];
      u.box_hi[a] = m->hi[a];
    }
    u.volume = m->volume.data();
    u.mesh_point = m->generator.data();
    u.face_start = m->face_start.data();
    u.face_neighbor = m->face_neighbor.data();
    u.face_normal = m->face_normal.data();
    u.face_centroid = m->face_centroid.data();
    u.face_area = m->face_area.data();
    u.face_vertex_start = m->face_vstart.data();
    u.vertices = m->vertex.data();
    u.cell_instance = nullptr;
    *out = u;
  });
}

int storm_voronoi_free(storm_voronoi *v) {
  return guarded([&] { delete reinterpret_cast<storm::voronoi::Mesh *>(v); });
}

}  // extern "C"
