// pm_api.cpp -- host side of the C ABI: contexts of systems that can come and
// go, their device memory (optionally carved from one HBM reservation), the
// launch tables of system subsets, launch sequencing and result export.
// Compiled by hipcc for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <atomic>
#include <exception>
#include <thread>
#include <string>
#include <utility>
#include <vector>

#include "../../include/protomol_amd.h"
#include "pm_gb.h"
#include "pm_kernels.h"

using pm::ForceDev;
using pm::Launch;
using pm::SwitchDev;
using pm::SysDev;

namespace {

thread_local std::string g_thread_err;
constexpr double kInvTimeFactor = 0.02045482828087295;  // PMConstants.cpp:54
constexpr double kBoltzmann = 0.001987191;              // PMConstants.cpp:70
constexpr size_t kAlign = 256;

struct Fail {
  int code;
  std::string msg;
};

#define HIPCHECK(expr)                                                                           \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
      throw Fail{PM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)};                 \
  } while (0)

inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// One device reservation handed out first-fit in kAlign units, with
// neighbouring free ranges merged on release.  Admission by HBM footprint
// asks it for room; without a reservation blocks come from hipMalloc.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  // admission leaves this much of the reservation free: the pair lists of
  // admitted systems grow into it when an MD step overflows them
  size_t keep = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes

  void init(size_t bytes) {
    bytes = align_up(bytes);
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) throw Fail{PM_ENOMEM, "hipMalloc of the reservation failed"};
    base = static_cast<char *>(p);
    cap = bytes;
    used = 0;
    keep = bytes / 16;
    free_.clear();
    free_[0] = bytes;
  }
  // admit: an admission, which must leave `keep` bytes free
  void *alloc(size_t bytes, bool admit = false) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    if (admit && used + bytes + keep > cap) return nullptr;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used += bytes;
      return base + off;
    }
    return nullptr;
  }
  void release(void *p, size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    size_t off = static_cast<char *>(p) - base;
    used -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
  void destroy() {
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = used = 0;
    free_.clear();
  }
};

// Arrays carved out of one block at kAlign boundaries; with base == nullptr
// it only measures (and every pointer it returns is null).
struct Carver {
  char *base = nullptr;
  size_t off = 0;
  template <class T>
  T *take(size_t n) {
    off = align_up(off);
    T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
    off += std::max<size_t>(n, 1) * sizeof(T);
    return p;
  }
};


// Host-side image of one system: validated inputs and the derived tables,
// ready to be copied into its device block.
struct SysPrep {
  int n = 0;
  std::vector<int> excl_start, excl_partner, excl_class, corr, loc_start, loc_item;
  std::vector<pm::LocRec> loc_rec;  // rows of local-term records (SysDev::loc_rec)
  int loc_rows[pm::kLocShares] = {}, loc_base[pm::kLocShares] = {};
  std::vector<unsigned long long> excl_win;  // per atom: FULL, MODIFIED partner bits (b - a + 31), far flag
  std::vector<double> invm;
  std::vector<float> sig_half, eps_sqrt;
  std::vector<float4> lj, lj12;
  int n_corr = 0, n_bonds = 0, n_angles = 0, ncol = 1, ncx = 1, ncy = 1, rx = 0, ry = 0;
  int cl_cap = 0, slot_cap = 0, sc_cap = 0, list_cap = 0, n_gb_tiles = 0;
  double ax = 0, ay = 0;
  double coul_const = 0, vir_const = 0;
  // Ewald reciprocal space
  std::vector<int4> ew_rows, ew_tiles;
  std::vector<double4> ew_kd;
  std::vector<short2> ew_pts;
  std::vector<int> ew_kmap, ew_fb_rows;
  int ew_nk = 0, ew_xrows = 0, ew_yrows = 0, ew_zrows = 0, ew_chunk = 0, ew_nchunks = 0;
  double ew_c = 0, ew_a8 = 0, ew_fac = 0;
  // molecules (molecular virial): CSR of their atoms in index order
  int n_mol = 0;
  std::vector<int> mol_start, mol_atoms;
  std::vector<int> gb_ident;  // GB: slot = atom
  // the periodic cell of the cutoff path: the system's own box, or for a
  // vacuum system one that encloses its region (vac: the region's bounds)
  double box[3] = {0, 0, 0}, origin[3] = {0, 0, 0};
  bool vac = false;
  double vlo[3] = {0, 0, 0}, vhi[3] = {0, 0, 0};
};

struct SysHost {
  int id = -1;
  int n = 0, ncol = 0, cl_cap = 0, slot_cap = 0, sc_cap = 0, list_cap = 0, n_corr = 0, n_bonded = 0, n_local = 0;
  int n_gb_tiles = 0, n_gb_pairs = 0;  // GB: 64-atom tiles, tile pairs (it <= jt)
  pm::EwLaunch ew{};        // reciprocal-space launch shape (sys filled per launch set)
  double coul_const = 0.0;  // Ewald point-self + charged-system energy (FP64, host)
  double vir_const = 0.0;   // charged-system virial diagonal
  SysDev dev{};             // host copy of the device record
  void *block = nullptr;
  size_t block_bytes = 0;
  void *lists = nullptr;
  size_t lists_bytes = 0;
  int *ctl_snap = nullptr;  // control words saved at the start of an MD advance
  bool md_ready = false;
};

// PM_SWITCH_* -> device record; returns the switch's cutoffSquared() in FP64
double build_switch(const pm_switch_desc &d, SwitchDev &s, bool ewald_hard, double rc) {
  std::memset(&s, 0, sizeof s);
  if (ewald_hard) {
    s.kind = PM_SWITCH_CUTOFF;
    double c2 = rc * rc;
    s.cut2 = (float)c2;
    return c2;
  }
  s.kind = d.kind;
  double cut2 = 0.0;
  switch (d.kind) {
  case PM_SWITCH_CUTOFF:
    if (!(d.cutoff > 0)) throw Fail{PM_EINVAL, "Cutoff switching function needs cutoff > 0"};
    cut2 = d.cutoff * d.cutoff;
    break;
  case PM_SWITCH_C1:  // C1SwitchingFunction.cpp:12-18
    if (!(d.cutoff > 0)) throw Fail{PM_EINVAL, "C1 switching function needs cutoff > 0"};
    cut2 = d.cutoff * d.cutoff;
    s.c15rc1 = (float)(1.5 / d.cutoff);
    s.c05rc3 = (float)(0.5 / (d.cutoff * d.cutoff * d.cutoff));
    s.c15rc3 = (float)(1.5 / (d.cutoff * d.cutoff * d.cutoff));
    break;
  case PM_SWITCH_C2: {  // C2SwitchingFunction.cpp:12-18, make() checks at :33-42
    if (!(d.switchon >= 0.0 && d.cutoff > 0.0 && d.switchon < d.cutoff))
      throw Fail{PM_EINVAL, "C2 switching function: 0 <= switchon < cutoff"};
    double dd = d.cutoff * d.cutoff - d.switchon * d.switchon;
    cut2 = d.cutoff * d.cutoff;
    s.on2 = (float)(d.switchon * d.switchon);
    s.s1 = (float)(1.0 / (dd * dd * dd));
    s.s2 = (float)(d.cutoff * d.cutoff - 3.0 * d.switchon * d.switchon);
    s.s3 = (float)(4.0 / (dd * dd * dd));
    s.c2d = (float)dd;
    s.c2hd = (float)(0.5 * dd);
    s.c2a1 = (float)(2.0 / (dd * dd * dd));
    s.c2sc = (float)(12.0 / (dd * dd * dd));
    break;
  }
  case PM_SWITCH_CN: {  // CnSwitchingFunction.cpp:41-60, make() checks at :83-110
    if (!(d.switchon >= 0.0 && d.cutoff > 0.0 && d.switchon < d.cutoff &&
          !(d.switchoff > 0.0 && d.switchoff != d.cutoff) && d.order >= 2.0 && d.order <= 6.0))
      throw Fail{PM_EINVAL, "Cn switching function: 0 <= switchon < cutoff, order 2..6"};
    s.dcut = d.switchoff != 0.0 ? d.switchoff : d.cutoff;
    s.dcut2 = s.dcut * s.dcut;
    s.don = d.switchon;
    s.don2 = d.switchon * d.switchon;
    s.ord_int = (int)d.order;
    s.ord_idx = s.ord_int - 2;
    if (s.ord_int == 5) s.ord_int = 2;
    s.irange[0] = std::pow(d.switchon - d.cutoff, -(s.ord_int + 1));
    for (int i = 1; i < s.ord_int + 1; ++i) s.irange[i] = s.irange[i - 1] / (d.switchon - d.cutoff);
    cut2 = s.dcut2;
    break;
  }
  case PM_SWITCH_CMPCNCN: {  // CmpCnCnSwitchingFunction.cpp:12-28, make() checks at :47-60
    if (!(d.switchon >= 0.0 && d.cutoff > 0.0 && d.switchoff > 0.0 && d.switchon < d.switchoff &&
          d.switchoff <= d.cutoff && d.order >= 2.0 && d.order <= 6.0))
      throw Fail{PM_EINVAL, "CmpCnCn switching function: 0 <= switchon < switchoff <= cutoff, order 2..6"};
    double ord = d.order == 5.0 ? 2.0 : d.order;
    s.dcut = d.cutoff;
    s.dcut2 = d.cutoff * d.cutoff;
    s.don = d.switchon;
    s.don2 = d.switchon * d.switchon;
    s.doff = d.switchoff;
    s.doff2 = d.switchoff * d.switchoff;
    s.ord_int = (int)ord;
    s.ord_idx = s.ord_int - 2;
    s.irange[0] = 1.0 / std::pow(d.switchon - d.switchoff, s.ord_int + 1);
    s.itrange[0] = 1.0 / std::pow(d.switchoff - d.cutoff, s.ord_int + 1);
    for (int i = 1; i < s.ord_int + 1; ++i) {
      s.irange[i] = s.irange[i - 1] / (d.switchon - d.switchoff);
      s.itrange[i] = s.itrange[i - 1] / (d.switchoff - d.cutoff);
    }
    cut2 = s.dcut2;
    break;
  }
  default:
    throw Fail{PM_EINVAL, "the cell-list pair path needs a switching function with a cutoff"};
  }
  s.cut2 = (float)cut2;
  return cut2;
}


// entries of one super-cluster's list: j clusters whose box comes within
// rlist of the box of kScCl z-consecutive clusters of a column (ax x ay
// cross-section, height kScCl * hz), halved for the half shell, with margin
int estimate_list_cap(double rho, double rlist, double ax, double ay, int cl_cap) {
  const double hz = pm::kCluster / std::max(rho * ax * ay, 1e-9);  // height of one cluster
  const double a = 2.0 * ax, b = 2.0 * ay, c = (pm::kScCl + 1) * hz, r = rlist;
  const double vol = a * b * c + 2.0 * r * (a * b + b * c + c * a) + M_PI * r * r * (a + b + c) +
                     4.0 / 3.0 * M_PI * r * r * r;
  const double half = 0.5 * vol * rho / pm::kCluster;
  int cap = (int)(1.4 * half) + 32;
  cap = (cap + 31) / 32 * 32;
  return std::min(cap, 27 * cl_cap / 2 + 32);
}

// The launch table of a subset of a context's systems: copies of their device
// records, prefix arrays over them, and a buffer for their packed scalars.
struct LaunchSet {
  std::vector<int> ids;
  Launch L;
  std::vector<pm::EwLaunch> ew;  // L.ew points here
  void *mem = nullptr;
  long long *d_scal = nullptr;
  long long *h_scal = nullptr;  // pinned
  // captured MD steps for time step g_h and these Langevin constants: [0]
  // forces only, [1] with the energy pass on the last of the steps
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  double g_h[2] = {0.0, 0.0};
  pm::LangevinDev g_lv[2]{};
  ~LaunchSet() {
    for (hipGraphExec_t g : gexec)
      if (g) (void)hipGraphExecDestroy(g);
    if (h_scal) (void)hipHostFree(h_scal);
    if (mem) (void)hipFree(mem);
  }
};

}  // namespace

struct pm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  pm_force_desc desc{};
  ForceDev P{};
  bool gb = false;
  bool recip = false;       // PM_FLAG_EWALD_RECIPROCAL
  bool molvir = false;      // PM_FLAG_MOLVIRIAL
  // LangevinLeapfrogIntegrator (gamma_ps = 0: plain leapfrog)
  double lv_temperature = 0.0, lv_gamma_ps = 0.0;
  unsigned lv_seed = 1234;
  double rc = 0.0, skin_md = 0.0, skin_eval = 0.0;
  std::vector<std::unique_ptr<SysHost>> sys;  // by id; null once removed
  Arena arena;
  std::map<std::vector<int>, std::unique_ptr<LaunchSet>> sets;
  std::string err;
  double stats[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evp0 = nullptr, evp1 = nullptr, evr0 = nullptr, evr1 = nullptr;
  std::vector<hipEvent_t> evs;  // PM_FLAG_TIMING: start/stop per pair launch of an MD call
  int evs_used = 0;
  // pinned host staging of large transfers (the tables of systems being added,
  // positions and velocities of many systems); pin_done is recorded after the
  // copies that read it
  char *pin = nullptr;
  size_t pin_cap = 0;
  hipEvent_t pin_done = nullptr;
};

namespace {

void set_err(pm_ctx *c, const std::string &m) {
  if (c) c->err = m;
  g_thread_err = m;
}

// admit: the allocation admits a system (it leaves the reservation's growth
// share free); otherwise it grows an admitted one
void *dev_alloc(pm_ctx *c, size_t bytes, bool admit = false) {
  if (c->arena.base) {
    void *p = c->arena.alloc(bytes, admit);
    if (!p) throw Fail{PM_EFULL, "the device reservation has no room for " + std::to_string(bytes) + " bytes"};
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 1)) != hipSuccess) throw Fail{PM_ENOMEM, "hipMalloc failed"};
  return p;
}

void dev_free(pm_ctx *c, void *p, size_t bytes) {
  if (!p) return;
  if (c->arena.base) c->arena.release(p, bytes);
  else (void)hipFree(p);
}

void drop_sets(pm_ctx *c) { c->sets.clear(); }

// forget the launch tables that hold a copy of system `id`
void drop_sets_with(pm_ctx *c, int id) {
  for (auto it = c->sets.begin(); it != c->sets.end();) {
    if (std::find(it->first.begin(), it->first.end(), id) != it->first.end()) it = c->sets.erase(it);
    else ++it;
  }
}

std::vector<int> live_ids(const pm_ctx *c) {
  std::vector<int> v;
  for (size_t i = 0; i < c->sys.size(); ++i)
    if (c->sys[i]) v.push_back((int)i);
  return v;
}

SysHost &sys_of(pm_ctx *c, int id) {
  if (id < 0 || id >= (int)c->sys.size() || !c->sys[id]) throw Fail{PM_EINVAL, "no system with id " + std::to_string(id)};
  return *c->sys[id];
}

std::vector<int> check_ids(pm_ctx *c, int32_t n, const int32_t *ids) {
  if (n < 1 || !ids) throw Fail{PM_EINVAL, "need >= 1 system id"};
  std::vector<int> v(ids, ids + n);
  std::vector<int> u = v;
  std::sort(u.begin(), u.end());
  if (std::adjacent_find(u.begin(), u.end()) != u.end()) throw Fail{PM_EINVAL, "repeated system id"};
  for (int id : v) sys_of(c, id);
  return v;
}

// ------------------------------------------------------------ system set-up

// Reciprocal-space tables of NonbondedFullEwaldSystemForce::initialize
// (NonbondedFullEwaldSystemForce.h:604-667): the k-vectors of the h >= 0 half
// space with |k| < kc in the reference's (h, k, l) order and with its FP64
// expressions, grouped into rows of l at fixed (h, k); the structure-factor
// tiles (h blocks of 8 x kEwSFPoints (k, l) points) and the force pass's row
// blocks.
void ewald_setup(const pm_ctx *c, const pm_system_desc &d, SysPrep &p) {
  constexpr int HB = 8, SFT = pm::kEwSFPoints;
  const double LX = d.box[0], LY = d.box[1], LZ = d.box[2];
  double a1 = LY * LZ - 0.0 * 0.0, d1 = LX * a1 + 0.0 * 0.0 + 0.0 * 0.0;
  double a2 = LZ * LX - 0.0 * 0.0, d2 = 0.0 * 0.0 + LY * a2 + 0.0 * 0.0;
  double a3 = LX * LY - 0.0 * 0.0, d3 = 0.0 * 0.0 + 0.0 * 0.0 + LZ * a3;
  const double Lr[3] = {a1 / d1, a2 / d2, a3 / d3};
  const double alpha = c->desc.ewald_alpha;
  const double pp = -std::log(c->desc.ewald_accuracy);
  const double kc = 2.0 * alpha * std::sqrt(pp);
  const int hmax = (int)std::floor(kc / (2.0 * M_PI) * LX);
  const int kmax = (int)std::floor(kc / (2.0 * M_PI) * LY);
  const int lmax = (int)std::floor(kc / (2.0 * M_PI) * LZ);
  p.ew_xrows = (hmax + 1 + HB - 1) / HB * HB;
  p.ew_yrows = std::max(kmax, 1) + 1;
  p.ew_zrows = std::max(lmax, 1) + 1;
  const double kc2 = kc * kc;
  const double ar = 1.0 / (alpha * alpha);
  const int KW = 2 * kmax + 1, LW = 2 * lmax + 1;
  std::vector<int> kidx((size_t)(hmax + 1) * KW * LW, -1);
  p.ew_rows.clear();
  p.ew_kd.clear();
  for (int h = 0; h <= hmax; h++) {
    const double kx = 2.0 * M_PI * h * Lr[0];
    for (int k = (h == 0 ? 0 : -kmax); k <= kmax; k++) {
      const double ky = 2.0 * M_PI * k * Lr[1];
      bool row = false;
      for (int l = (h == 0 && k == 0 ? 1 : -lmax); l <= lmax; l++) {
        const double kz = 2.0 * M_PI * l * Lr[2];
        if (kx * kx + ky * ky + kz * kz < kc2) {
          const double ksq = kx * kx + ky * ky + kz * kz;
          const double b = 1.0 / ksq * std::exp(-ksq * ar / 4.0);
          const int q = (int)p.ew_kd.size();
          if (!row) p.ew_rows.push_back(make_int4(h, k, l, q));
          row = true;
          kidx[((size_t)h * KW + (k + kmax)) * LW + (l + lmax)] = q;
          p.ew_kd.push_back(make_double4(kx, ky, kz, b));
        }
      }
    }
  }
  p.ew_nk = (int)p.ew_kd.size();
  const int nrows = (int)p.ew_rows.size();
  p.ew_rows.push_back(make_int4(0, 0, 0, p.ew_nk));
  // structure-factor points and tiles per h block
  p.ew_pts.clear();
  p.ew_kmap.clear();
  p.ew_tiles.clear();
  for (int h0 = 0; h0 <= hmax; h0 += HB) {
    const int nh = std::min(HB, hmax + 1 - h0);
    const int first = (int)p.ew_pts.size();
    for (int k = -kmax; k <= kmax; k++)
      for (int l = -lmax; l <= lmax; l++) {
        int m[HB];
        bool any = false;
        for (int t = 0; t < HB; ++t) {
          m[t] = t < nh ? kidx[((size_t)(h0 + t) * KW + (k + kmax)) * LW + (l + lmax)] : -1;
          any |= m[t] >= 0;
        }
        if (!any) continue;
        p.ew_pts.push_back(make_short2((short)k, (short)l));
        p.ew_kmap.insert(p.ew_kmap.end(), m, m + HB);
      }
    const int end = (int)p.ew_pts.size();
    for (int t0 = first; t0 < end; t0 += SFT) p.ew_tiles.push_back(make_int4(h0, t0, end, nh));
  }
  // atom chunks (multiples of 4 atoms): enough workgroups to fill the device
  const int ntiles = (int)p.ew_tiles.size();
  const int n = d.n_atoms;
  constexpr int TA = 4;
  int nch = std::max(1, std::min((4096 + ntiles - 1) / std::max(ntiles, 1), (n + 63) / 64));
  p.ew_chunk = ((n + nch - 1) / nch + TA - 1) / TA * TA;
  p.ew_nchunks = (n + p.ew_chunk - 1) / p.ew_chunk;
  // force pass: row blocks of about equal k count, enough waves to hide the
  // latency of the per-k weights' scalar loads
  const int waves = (n + 127) / 128;  // two atoms per thread
  const int nfb = std::max(1, std::min(nrows, (12288 + waves - 1) / waves));
  p.ew_fb_rows.assign(1, 0);
  long long accum = 0;
  for (int r = 0; r < nrows; ++r) {
    accum += p.ew_rows[r + 1].w - p.ew_rows[r].w;
    const int b = (int)p.ew_fb_rows.size();
    if (b < nfb && r + 1 < nrows && accum * nfb >= (long long)p.ew_nk * b) p.ew_fb_rows.push_back(r + 1);
  }
  p.ew_fb_rows.push_back(nrows);
  const double V = LX * LY * LZ;
  p.ew_c = 4.0 * M_PI * (1.0 / V);
  p.ew_a8 = 8.0 * M_PI * (1.0 / V);
  p.ew_fac = 1.0 / (4.0 * alpha * alpha);
}

// Validate a system description and build its host-side tables
// (exclusions, Ewald correction pairs, per-atom bonded/correction items,
// masses, LJ parameters, cluster grid and capacities).
SysPrep prepare(const pm_ctx *c, const pm_system_desc &d) {
  SysPrep p;
  const int n = d.n_atoms;
  if (n < 1) throw Fail{PM_EINVAL, "n_atoms must be >= 1"};
  if (!d.scaled_charge || !d.mass || !d.molecule) throw Fail{PM_EINVAL, "charge, mass and molecule are required"};
  p.n = n;
  if (d.periodic) {
    for (int k = 0; k < 3; ++k)
      if (!(d.box[k] > 0)) throw Fail{PM_EINVAL, "box edges must be positive"};
  }
  for (int k = 0; k < 3; ++k) {
    p.box[k] = d.box[k];
    p.origin[k] = d.origin[k];
  }
  // exclusion table: last add() of a pair wins (ExclusionTable.cpp:97-120);
  // the Ewald correction walks the table's pairs with exact duplicates dropped
  struct Ex {
    int a, b, cls, order;
  };
  std::vector<Ex> ex;
  ex.reserve(d.n_exclusions);
  for (int k = 0; k < d.n_exclusions; ++k) {
    int a = d.exclusions[3 * k], b = d.exclusions[3 * k + 1], cls = d.exclusions[3 * k + 2];
    if (a < 0 || b < 0 || a >= n || b >= n) throw Fail{PM_EINVAL, "exclusion atom out of range"};
    if (cls != PM_EXCL_FULL && cls != PM_EXCL_MODIFIED) throw Fail{PM_EINVAL, "bad exclusion class"};
    if (a > b) std::swap(a, b);
    if (a == b) continue;
    ex.push_back({a, b, cls, k});
  }
  {
    // correction pairs: first occurrence of each (a, b, class), in table order
    std::vector<Ex> byk = ex;
    std::stable_sort(byk.begin(), byk.end(), [](const Ex &x, const Ex &y) {
      return x.a != y.a ? x.a < y.a : (x.b != y.b ? x.b < y.b : x.cls < y.cls);
    });
    std::vector<Ex> firsts;
    for (size_t i = 0; i < byk.size(); ++i)
      if (i == 0 || byk[i].a != byk[i - 1].a || byk[i].b != byk[i - 1].b || byk[i].cls != byk[i - 1].cls)
        firsts.push_back(byk[i]);
    std::sort(firsts.begin(), firsts.end(), [](const Ex &x, const Ex &y) { return x.order < y.order; });
    const bool ewald = c->desc.term_b == PM_TERM_EWALD_REAL && c->desc.ewald_correction && !c->gb;
    if (ewald)
      for (const Ex &e : firsts) {
        p.corr.push_back(e.a);
        p.corr.push_back(e.b);
        p.corr.push_back(e.cls);
      }
    p.n_corr = (int)p.corr.size() / 3;
  }
  // pair -> class of its last add
  std::stable_sort(ex.begin(), ex.end(), [](const Ex &x, const Ex &y) { return x.a != y.a ? x.a < y.a : x.b < y.b; });
  // the pair terms look exclusions up only for atoms of one molecule
  // (OneAtomPair.h:64-68), so a pair of two molecules is never excluded
  std::vector<Ex> last;
  for (size_t i = 0; i < ex.size(); ++i)
    if ((i + 1 == ex.size() || ex[i + 1].a != ex[i].a || ex[i + 1].b != ex[i].b) &&
        d.molecule[ex[i].a] == d.molecule[ex[i].b])
      last.push_back(ex[i]);
  p.excl_start.assign(n + 1, 0);
  for (const Ex &e : last) {
    p.excl_start[e.a + 1]++;
    p.excl_start[e.b + 1]++;
  }
  for (int i = 0; i < n; ++i) p.excl_start[i + 1] += p.excl_start[i];
  p.excl_partner.resize(p.excl_start[n]);
  p.excl_class.resize(p.excl_start[n]);
  {
    std::vector<int> fill(p.excl_start.begin(), p.excl_start.end() - 1);
    for (const Ex &e : last) {
      p.excl_partner[fill[e.a]] = e.b;
      p.excl_class[fill[e.a]++] = e.cls;
      p.excl_partner[fill[e.b]] = e.a;
      p.excl_class[fill[e.b]++] = e.cls;
    }
  }
  // the same exclusions as bit masks over the atom-index window a - 31 .. a + 31
  // (bit b - a + 31 of the FULL and MODIFIED words); bit 63 of the FULL word
  // marks an atom with a partner outside the window, whose lookups use the CSR
  p.excl_win.assign(2 * (size_t)n, 0ull);
  for (int a = 0; a < n; ++a)
    for (int k = p.excl_start[a]; k < p.excl_start[a + 1]; ++k) {
      const int dd = p.excl_partner[k] - a;
      if (dd < -31 || dd > 31) p.excl_win[2 * (size_t)a] |= 1ull << 63;
      else if (p.excl_class[k] == PM_EXCL_FULL) p.excl_win[2 * (size_t)a] |= 1ull << (dd + 31);
      else if (p.excl_class[k] == PM_EXCL_MODIFIED) p.excl_win[2 * (size_t)a + 1] |= 1ull << (dd + 31);
    }
  // bonded terms
  p.n_bonds = (c->desc.flags & PM_FLAG_BONDED) ? d.n_bonds : 0;
  p.n_angles = (c->desc.flags & PM_FLAG_BONDED) ? d.n_angles : 0;
  if (p.n_bonds && (!d.bonds || !d.bond_k || !d.bond_r0)) throw Fail{PM_EINVAL, "bonds need atoms, k and r0"};
  if (p.n_angles && (!d.angles || !d.angle_k || !d.angle_theta0)) throw Fail{PM_EINVAL, "angles need atoms, k and theta0"};
  for (int k = 0; k < 2 * p.n_bonds; ++k)
    if (d.bonds[k] < 0 || d.bonds[k] >= n) throw Fail{PM_EINVAL, "bond atom out of range"};
  for (int k = 0; k < 3 * p.n_angles; ++k)
    if (d.angles[k] < 0 || d.angles[k] >= n) throw Fail{PM_EINVAL, "angle atom out of range"};
  // per-atom items of the owner-computes kernel: (index << 2) | role, in a CSR
  // keyed by (atom, kind) -- entry 3 * atom + kind -- so the kernel walks
  // one kind at a time
  {
    std::vector<int> cnt(3 * (size_t)n + 1, 0);
    for (int k = 0; k < p.n_bonds; ++k) for (int r = 0; r < 2; ++r) cnt[3 * d.bonds[2 * k + r] + pm::kLocBond + 1]++;
    for (int k = 0; k < p.n_angles; ++k) for (int r = 0; r < 3; ++r) cnt[3 * d.angles[3 * k + r] + pm::kLocAngle + 1]++;
    for (int k = 0; k < p.n_corr; ++k) for (int r = 0; r < 2; ++r) cnt[3 * p.corr[3 * k + r] + pm::kLocCorr + 1]++;
    for (size_t i = 0; i + 1 < cnt.size(); ++i) cnt[i + 1] += cnt[i];
    p.loc_start = cnt;
    p.loc_item.resize(cnt.back());
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < p.n_bonds; ++k)
      for (int r = 0; r < 2; ++r) p.loc_item[fill[3 * d.bonds[2 * k + r] + pm::kLocBond]++] = (k << 2) | r;
    for (int k = 0; k < p.n_angles; ++k)
      for (int r = 0; r < 3; ++r) p.loc_item[fill[3 * d.angles[3 * k + r] + pm::kLocAngle]++] = (k << 2) | r;
    for (int k = 0; k < p.n_corr; ++k)
      for (int r = 0; r < 2; ++r) p.loc_item[fill[3 * p.corr[3 * k + r] + pm::kLocCorr]++] = (k << 2) | r;
    // the records: for share w, row loc_base[w] + k holds item slot + k nslots
    // of the atom's list of that kind (role -1 where the atom has fewer);
    // rows per share from the longest list of each kind
    int maxcnt[3] = {0, 0, 0};
    for (int a = 0; a < n; ++a)
      for (int kind = 0; kind < 3; ++kind)
        maxcnt[kind] = std::max(maxcnt[kind], p.loc_start[3 * a + kind + 1] - p.loc_start[3 * a + kind]);
    int rows = 0;
    for (int w = 0; w < pm::kLocShares; ++w) {
      int kind, slot, ns;
      pm::loc_share(w, kind, slot, ns);
      p.loc_rows[w] = maxcnt[kind] > slot ? (maxcnt[kind] - slot + ns - 1) / ns : 0;
      p.loc_base[w] = rows;
      rows += p.loc_rows[w];
    }
    const double mscale = 1.0 - (double)(float)d.coulomb_scaling;  // as the device's float csf
    p.loc_rec.resize((size_t)rows * n);
    // atoms [a0, a1) of every row; large systems fill their rows on a few threads
    auto fill_rows = [&](int a0, int a1) {
      for (int w = 0; w < pm::kLocShares; ++w) {
        int kind, slot, ns;
        pm::loc_share(w, kind, slot, ns);
        for (int k = 0; k < p.loc_rows[w]; ++k) {
          pm::LocRec *row = p.loc_rec.data() + (size_t)(p.loc_base[w] + k) * n;
          for (int a = a0; a < a1; ++a) {
            const int it = p.loc_start[3 * a + kind] + slot + k * ns;
            pm::LocRec r{-1, -1, -1, -1, 0.0, 0.0};
            if (it < p.loc_start[3 * a + kind + 1]) {
              const int item = p.loc_item[it], t = item >> 2;
              r.role = item & 3;
              if (kind == pm::kLocBond) {
                r.a0 = d.bonds[2 * t];
                r.a1 = d.bonds[2 * t + 1];
                r.p0 = d.bond_k[t];
                r.p1 = d.bond_r0[t];
              } else if (kind == pm::kLocAngle) {
                r.a0 = d.angles[3 * t];
                r.a1 = d.angles[3 * t + 1];
                r.a2 = d.angles[3 * t + 2];
                r.p0 = d.angle_k[t];
                r.p1 = d.angle_theta0[t];
              } else {
                r.a0 = p.corr[3 * t];
                r.a1 = p.corr[3 * t + 1];
                double qq = d.scaled_charge[r.a0] * d.scaled_charge[r.a1];
                if (p.corr[3 * t + 2] == PM_EXCL_MODIFIED) qq *= mscale;
                r.p0 = qq;
              }
            }
            row[a] = r;
          }
        }
      }
    };
    const int nt = std::min(8, 1 + n / 16384);
    if (nt == 1) {
      fill_rows(0, n);
    } else {
      std::vector<std::thread> th;
      for (int t = 0; t < nt; ++t) th.emplace_back(fill_rows, (int)((long long)n * t / nt), (int)((long long)n * (t + 1) / nt));
      for (auto &t : th) t.join();
    }
  }
  // per-atom data
  p.invm.resize(n);
  for (int i = 0; i < n; ++i) {
    if (!(d.mass[i] > 0)) throw Fail{PM_EINVAL, "masses must be positive"};
    p.invm[i] = 1.0 / d.mass[i];
  }
  if (c->molvir) {
    // molecule ids index the molecules (GenericTopology::molecules)
    for (int i = 0; i < n; ++i) {
      if (d.molecule[i] < 0) throw Fail{PM_EINVAL, "molecule ids must be >= 0"};
      p.n_mol = std::max(p.n_mol, d.molecule[i] + 1);
    }
    if (p.n_mol > n) throw Fail{PM_EINVAL, "molecule ids must be < n_atoms"};
    p.mol_start.assign(p.n_mol + 1, 0);
    for (int i = 0; i < n; ++i) p.mol_start[d.molecule[i] + 1]++;
    for (int m = 0; m < p.n_mol; ++m) p.mol_start[m + 1] += p.mol_start[m];
    p.mol_atoms.resize(n);
    std::vector<int> fill(p.mol_start.begin(), p.mol_start.end() - 1);
    for (int i = 0; i < n; ++i) p.mol_atoms[fill[d.molecule[i]]++] = i;
  }
  if (d.lj_mode == PM_LJ_PERATOM) {
    if (!d.lj_sigma || !d.lj_epsilon) throw Fail{PM_EINVAL, "per-atom LJ needs sigma and epsilon"};
    p.sig_half.resize(n);
    p.eps_sqrt.resize(n);
    for (int i = 0; i < n; ++i) {
      p.sig_half[i] = (float)(0.5 * d.lj_sigma[i]);
      p.eps_sqrt[i] = (float)std::sqrt(d.lj_epsilon[i]);
    }
  } else {
    if (d.n_types < 1 || !d.lj_table || !d.type) throw Fail{PM_EINVAL, "LJ table mode needs n_types, type and lj_table"};
    if (d.n_types > 4096) throw Fail{PM_EINVAL, "at most 4096 LJ types"};
    for (int i = 0; i < n; ++i)
      if (d.type[i] < 0 || d.type[i] >= d.n_types) throw Fail{PM_EINVAL, "atom type out of range"};
    p.lj.resize((size_t)d.n_types * d.n_types);
    p.lj12.resize(p.lj.size());
    for (size_t k = 0; k < p.lj.size(); ++k) {
      p.lj[k] = make_float4((float)d.lj_table[4 * k], (float)d.lj_table[4 * k + 1], (float)d.lj_table[4 * k + 2],
                            (float)d.lj_table[4 * k + 3]);
      // the pair kernel's row format: A, B, 12 A, 6 B
      p.lj12[k] = make_float4(p.lj[k].x, p.lj[k].y, 12.0f * p.lj[k].x, 6.0f * p.lj[k].y);
    }
  }
  if (c->gb) {
    if (c->P.gb_on && (!d.gb_radius || !d.gb_scaling)) throw Fail{PM_EINVAL, "GB needs gb_radius and gb_scaling"};
    p.n_gb_tiles = pm::gb_tile_count(n);
    p.cl_cap = 0;
    p.slot_cap = n;  // forces are kept per atom: slot = atom
    p.gb_ident.resize(n);
    for (int i = 0; i < n; ++i) p.gb_ident[i] = i;
    return p;
  }
  const double rlist = c->rc + std::max(c->skin_md, c->skin_eval);
  if (!d.periodic) {
    // VacuumBoundaryConditions (minimalDifference = x_j - x_i) on the cell-list
    // path: box and origin give the region (origin +- box / 2) the atoms stay
    // in; the device works in a periodic cell larger than the region by more
    // than rlist, so that no image of an atom comes within rlist of another
    // atom, and every pair within the cutoff is found at its true separation
    // (minimal image of a vector shorter than half the cell).  An atom that
    // leaves the region stops the run with PM_ECAPACITY.
    if (c->desc.term_b == PM_TERM_EWALD_REAL || c->recip)
      throw Fail{PM_EINVAL, "Ewald summation needs periodic boundary conditions"};
    for (int k = 0; k < 3; ++k) {
      if (!(d.box[k] > 0))
        throw Fail{PM_EINVAL, "a vacuum system on the cutoff path needs box (the extent of the region its atoms "
                              "stay in) and origin (the region's centre)"};
      p.vlo[k] = d.origin[k] - 0.5 * d.box[k];
      p.vhi[k] = d.origin[k] + 0.5 * d.box[k];
      p.box[k] = std::max(d.box[k] + rlist + 1.0, 2.0 * rlist + 1.0);
      p.origin[k] = p.vlo[k] + 0.5 * p.box[k];
    }
    p.vac = true;
  }
  // Ewald point-self and charged-system constants (NonbondedFullEwaldSystemForce.h:694-714)
  if (c->desc.term_b == PM_TERM_EWALD_REAL && c->desc.ewald_correction) {
    double q2 = 0.0, q = 0.0;
    for (int i = 0; i < n; ++i) q2 += d.scaled_charge[i] * d.scaled_charge[i];
    for (int i = 0; i < n; ++i) q += d.scaled_charge[i];
    double alpha = c->desc.ewald_alpha;
    double point = -q2 * alpha / std::sqrt(M_PI);
    double V = std::fabs(d.box[0] * d.box[1] * d.box[2]);
    double charged = 0.0;
    if (std::fabs(q * 0.00268283) > 1.0e-5) charged = -M_PI / (2.0 * V * (alpha * alpha)) * q * q;
    p.coul_const = point + charged;
    p.vir_const = charged;
  }
  // x-y column grid: columns about (8/rho)^(1/3) wide so 8-atom clusters are compact
  // density over the region the atoms occupy (the system's box, or a vacuum
  // system's region)
  const double V = d.box[0] * d.box[1] * d.box[2];
  const double rho = n / V;
  const double scale = c->desc.cell_scale > 0 ? c->desc.cell_scale : 1.0;
  const double a = scale * std::cbrt(pm::kCluster / rho);
  for (int k = 0; k < 3; ++k)
    if (p.box[k] < 2.0 * rlist)
      throw Fail{PM_EINVAL, "box edge " + std::to_string(p.box[k]) + " is shorter than twice the pair-list radius " +
                                std::to_string(rlist)};
  p.ncx = std::max(1, (int)std::floor(p.box[0] / a));
  p.ncy = std::max(1, (int)std::floor(p.box[1] / a));
  p.ax = p.box[0] / p.ncx;
  p.ay = p.box[1] / p.ncy;
  p.ncol = p.ncx * p.ncy;
  p.rx = (int)std::ceil(rlist / p.ax);
  p.ry = (int)std::ceil(rlist / p.ay);
  p.cl_cap = (n + pm::kCluster - 1) / pm::kCluster + p.ncol;
  p.slot_cap = p.cl_cap * pm::kCluster;
  p.sc_cap = p.cl_cap / pm::kScCl + p.ncol + 1;
  p.list_cap = estimate_list_cap(rho, rlist, p.ax, p.ay, p.cl_cap);
  // test knob: a smaller first list capacity (entries per super-cluster), so
  // tests can drive the overflow-and-regrow path; the lists grow as usual
  if (const char *cap = std::getenv("PM_DEBUG_LIST_CAP")) p.list_cap = std::max(1, std::atoi(cap));
  if (c->recip) ewald_setup(c, d, p);
  return p;
}

// Lay out every device array of a system in one block (Carver measures with
// a null base).  Returns the block size.
size_t carve(const pm_ctx *c, const SysPrep &p, const pm_system_desc &d, Carver &cv, SysDev &S, int **ctl_snap) {
  const int n = p.n;
  S.excl_start = cv.take<int>(p.excl_start.size());
  S.excl_partner = cv.take<int>(p.excl_partner.size());
  S.excl_class = cv.take<int>(p.excl_class.size());
  S.excl_win = cv.take<unsigned long long>(p.excl_win.size());
  S.loc_rec = cv.take<pm::LocRec>(p.loc_rec.size());
  for (int w = 0; w < pm::kLocShares; ++w) {
    S.loc_rows[w] = p.loc_rows[w];
    S.loc_base[w] = p.loc_base[w];
  }
  S.charge = cv.take<double>(n);
  S.inv_mass = cv.take<double>(n);
  S.molecule = cv.take<int>(n);
  S.type = cv.take<int>(n);
  if (d.lj_mode == PM_LJ_PERATOM) {
    S.sig_half = cv.take<float>(n);
    S.eps_sqrt = cv.take<float>(n);
  } else {
    S.lj_table = cv.take<float4>(p.lj.size());
    S.lj_table12 = cv.take<float4>(p.lj12.size());
  }
  if (c->molvir) {
    S.mol_start = cv.take<int>(p.mol_start.size());
    S.mol_atoms = cv.take<int>(p.mol_atoms.size());
    S.mass = cv.take<double>(n);
  }
  if (c->gb) {
    S.gb_rho = cv.take<double>(n);
    S.gb_S = cv.take<double>(n);
  }
  if (c->recip) {
    S.ew_rows = cv.take<int4>(p.ew_rows.size());
    S.ew_kd = cv.take<double4>(p.ew_kd.size());
    S.ew_tiles = cv.take<int4>(p.ew_tiles.size());
    S.ew_pts = cv.take<short2>(p.ew_pts.size());
    S.ew_kmap = cv.take<int>(p.ew_kmap.size());
    S.ew_fb_rows = cv.take<int>(p.ew_fb_rows.size());
  }
  // (the tables above are uploaded; everything below starts at zero)
  S.x = cv.take<double>(3 * (size_t)n);
  S.v = cv.take<double>(3 * (size_t)n);
  S.xref = cv.take<double>(3 * (size_t)n);
  S.xaos = cv.take<double>(3 * (size_t)n);
  S.vaos = cv.take<double>(3 * (size_t)n);
  S.faos = cv.take<double>(3 * (size_t)n);
  S.slot = cv.take<int>(n);
  S.lbuild = cv.take<float4>(n);
  S.n_clusters = cv.take<int>(1);
  S.acc = cv.take<long long>((size_t)pm::kAccN * pm::kAccCopies);
  S.census = cv.take<unsigned long long>(2);
  S.err = cv.take<unsigned>(1);
  S.maxdisp2 = cv.take<unsigned>(pm::kDispSlots);
  S.maxdisp_run = cv.take<unsigned>(1);
  S.ctl = cv.take<int>(pm::kCtlN);
  *ctl_snap = cv.take<int>(pm::kCtlN);
  S.xq = cv.take<float4>(p.slot_cap);
  // x, y, z components of the fixed-point forces at a fixed stride, so a lane
  // can address its component as fx + component * f_stride
  S.f_stride = (int)(align_up((size_t)p.slot_cap * sizeof(long long)) / sizeof(long long));
  S.fx = cv.take<long long>(3 * (size_t)S.f_stride);
  S.fy = S.fx ? S.fx + S.f_stride : nullptr;
  S.fz = S.fx ? S.fx + 2 * (size_t)S.f_stride : nullptr;
  S.fl = cv.take<long long>(3 * (size_t)n);
  if (c->molvir) {
    S.n_mol = p.n_mol;
    S.com = cv.take<double>(3 * (size_t)p.n_mol);
  }
  if (c->gb) {
    S.gb_born = cv.take<double>(n);
    S.gb_drfac = cv.take<double>(n);
    S.gb_partial = cv.take<double>(n);
    S.gb_spare = cv.take<double>(n);
    // per tile pair (it <= jt): the i-side and j-side FP64 force partials of
    // its 64 + 64 atoms, summed per atom in a fixed order by k_gb_fsum
    S.gb_nt = p.n_gb_tiles;
    S.gb_fpart = cv.take<double>((size_t)pm::gb_pair_count(p.n_gb_tiles) * pm::kGBPairDoubles);
    return cv.off;
  }
  S.col_count = cv.take<int>(p.ncol);
  S.bucket = cv.take<unsigned long long>((size_t)p.ncol * pm::kSortCap);
  S.col_cl_start = cv.take<int>(p.ncol);
  S.col_ncl = cv.take<int>(p.ncol);
  S.ljt = cv.take<int>(p.slot_cap);
  S.ljp = cv.take<float2>(p.slot_cap);
  S.orig = cv.take<int>(p.slot_cap);
  S.smol = cv.take<int>(p.slot_cap);
  S.cl_origin = cv.take<double>(3 * (size_t)p.cl_cap);
  S.cl_lo = cv.take<float4>(p.cl_cap);
  S.cl_hi = cv.take<float4>(p.cl_cap);
  S.cl_col = cv.take<int>(p.cl_cap);
  S.cl_nvalid = cv.take<int>(p.cl_cap);
  S.col_sc_start = cv.take<int>(p.ncol);
  S.col_nsc = cv.take<int>(p.ncol);
  S.sc_col = cv.take<int>(p.sc_cap);
  S.n_sc = cv.take<int>(1);
  S.list_len = cv.take<int>(p.sc_cap);
  if (c->recip) {
    S.ew_tab = cv.take<double2>((size_t)(p.ew_xrows + p.ew_yrows + p.ew_zrows) * n);
    S.ew_psum = cv.take<double2>((size_t)p.ew_nchunks * p.ew_nk);
    S.ew_sk = cv.take<double4>(p.ew_kd.size());
    S.ew_pf = cv.take<double>((size_t)(p.ew_fb_rows.size() - 1) * 3 * n);
  }
  return cv.off;
}

// exclusion-mask slots per super-cluster: grows with the list capacity
int mask_cap_for(int list_cap) { return std::max(pm::kMaskSlots, list_cap / 8); }

size_t lists_bytes(int sc_cap, int list_cap) {
  Carver cv;
  cv.take<uint4>(2 * (size_t)sc_cap * list_cap);
  cv.take<unsigned long long>((size_t)sc_cap * mask_cap_for(list_cap) * 2 * pm::kScCl);
  return cv.off;
}

// the pair lists of capacity h.list_cap at mem (lists_bytes of them)
void place_lists(SysHost &h, void *mem) {
  Carver cv;
  cv.base = static_cast<char *>(mem);
  h.dev.list_cap = h.list_cap;
  h.dev.mask_cap = mask_cap_for(h.list_cap);
  h.dev.list = cv.take<uint4>(2 * (size_t)h.sc_cap * h.list_cap);
  h.dev.list_mask = cv.take<unsigned long long>((size_t)h.sc_cap * h.dev.mask_cap * 2 * pm::kScCl);
}

void alloc_lists(pm_ctx *c, SysHost &h, bool admit = false) {
  if (h.cl_cap == 0) return;
  const size_t bytes = lists_bytes(h.sc_cap, h.list_cap);
  void *mem = dev_alloc(c, bytes, admit);
  dev_free(c, h.lists, h.lists_bytes);
  h.lists = mem;
  h.lists_bytes = bytes;
  place_lists(h, mem);
}

uint64_t footprint(const pm_ctx *c, const pm_system_desc &d) {
  SysPrep p = prepare(c, d);
  Carver cv;
  SysDev S{};
  int *snap = nullptr;
  size_t b = align_up(carve(c, p, d, cv, S, &snap));
  if (!c->gb) b += align_up(lists_bytes(p.sc_cap, p.list_cap));
  return b;
}

// Table uploads of systems being added: gathered, then copied from one pinned
// staging buffer, one copy per run of adjacent tables (a system's uploaded
// tables lie first in its block, so one run each).
struct Uploads {
  struct Seg {
    char *dst;
    const void *src;
    size_t bytes;
  };
  std::vector<Seg> segs;
};

template <class T>
void upload(Uploads *up, T *dst, const T *src, size_t n) {
  if (!n || !src) return;
  if (up) up->segs.push_back({reinterpret_cast<char *>(dst), src, n * sizeof(T)});
  else HIPCHECK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
}

// f(0) .. f(count - 1) on up to 16 host threads
void parallel_for(int count, const std::function<void(int)> &f) {
  const int nt = std::max(1, std::min({count, 16, (int)std::max(1u, std::thread::hardware_concurrency())}));
  if (nt == 1) {
    for (int i = 0; i < count; ++i) f(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (int i = next++; i < count; i = next++) f(i);
    });
  for (auto &t : th) t.join();
}

// transfers at least this large go through the pinned staging buffer
constexpr size_t kPinnedTransfer = 16u << 20;

// The pinned staging buffer with room for `bytes`, free to be written (the
// copies of its previous use have finished).
char *pinned(pm_ctx *c, size_t bytes) {
  if (!c->pin_done) HIPCHECK(hipEventCreateWithFlags(&c->pin_done, hipEventDisableTiming));
  HIPCHECK(hipEventSynchronize(c->pin_done));
  if (bytes > c->pin_cap) {
    if (c->pin) (void)hipHostFree(c->pin);
    c->pin = nullptr;
    c->pin_cap = 0;
    const size_t cap = std::max(bytes + bytes / 2, (size_t)(4u << 20));
    if (hipHostMalloc(reinterpret_cast<void **>(&c->pin), cap, hipHostMallocDefault) != hipSuccess)
      throw Fail{PM_ENOMEM, "hipHostMalloc of the transfer staging buffer failed"};
    c->pin_cap = cap;
  }
  return c->pin;
}

// Issue the gathered uploads on the context's stream (asynchronous: the
// caller synchronizes the stream before the systems are used; the next use of
// the staging buffer waits for these copies).
void flush_uploads(pm_ctx *c, Uploads &up) {
  if (up.segs.empty()) return;
  std::sort(up.segs.begin(), up.segs.end(), [](const Uploads::Seg &a, const Uploads::Seg &b) { return a.dst < b.dst; });
  // runs: segments whose gaps are alignment padding (< 4 KiB) travel together
  struct Run {
    char *dst;
    size_t bytes, stage;
  };
  std::vector<Run> runs;
  size_t total = 0;
  for (const auto &g : up.segs) {
    if (!runs.empty() && g.dst >= runs.back().dst + runs.back().bytes &&
        g.dst - (runs.back().dst + runs.back().bytes) < 4096) {
      runs.back().bytes = (size_t)(g.dst + g.bytes - runs.back().dst);
    } else {
      total = align_up(total);
      runs.push_back({g.dst, g.bytes, total});
    }
    total = runs.back().stage + runs.back().bytes;
  }
  char *pin = pinned(c, total);
  // segment k's run; each segment also clears the padding before it in its run
  std::vector<int> run_of(up.segs.size());
  for (size_t k = 0, r = 0; k < up.segs.size(); ++k) {
    const auto &g = up.segs[k];
    while (!(g.dst >= runs[r].dst && g.dst + g.bytes <= runs[r].dst + runs[r].bytes)) ++r;
    run_of[k] = (int)r;
  }
  parallel_for((int)up.segs.size(), [&](int k) {
    const auto &g = up.segs[k];
    const Run &R = runs[run_of[k]];
    char *gap0 = (k > 0 && run_of[k - 1] == run_of[k]) ? up.segs[k - 1].dst + up.segs[k - 1].bytes : R.dst;
    if (g.dst > gap0) std::memset(pin + R.stage + (gap0 - R.dst), 0, (size_t)(g.dst - gap0));
    std::memcpy(pin + R.stage + (g.dst - R.dst), g.src, g.bytes);
  });
  for (const Run &R : runs)
    HIPCHECK(hipMemcpyAsync(R.dst, pin + R.stage, R.bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipEventRecord(c->pin_done, c->stream));
  up.segs.clear();
}

// The host record of a system and the scalars of its device record, from
// its prepared tables (no device memory yet).
void init_system(const pm_ctx *c, const pm_system_desc &d, const SysPrep &p, SysHost &h) {
  SysDev &S = h.dev;
  const int n = p.n;
  h.n = n;
  S.n_atoms = n;
  S.lv_seed = c->lv_seed;
  // the cutoff path works in a periodic cell (a vacuum system's encloses its region)
  S.periodic = (d.periodic || p.vac) ? 1 : 0;
  S.vac = p.vac ? 1 : 0;
  for (int k = 0; k < 3; ++k) {
    S.vac_lo[k] = p.vlo[k];
    S.vac_hi[k] = p.vhi[k];
  }
  S.lj_mode = d.lj_mode;
  S.n_types = d.n_types;
  S.csf = (float)d.coulomb_scaling;
  S.lj_scaling = (float)d.lj_scaling;
  if (S.periodic) {
    // PeriodicBoundaryConditions::set (PeriodicBoundaryConditions.cpp:32-77), orthogonal
    const double Lx = p.box[0], Ly = p.box[1], Lz = p.box[2];
    double a1 = Ly * Lz - 0.0 * 0.0, d1 = Lx * a1 + 0.0 * 0.0 + 0.0 * 0.0;
    double a2 = Lz * Lx - 0.0 * 0.0, d2 = 0.0 * 0.0 + Ly * a2 + 0.0 * 0.0;
    double a3 = Lx * Ly - 0.0 * 0.0, d3 = 0.0 * 0.0 + 0.0 * 0.0 + Lz * a3;
    S.L[0] = Lx; S.L[1] = Ly; S.L[2] = Lz;
    S.Lr[0] = a1 / d1; S.Lr[1] = a2 / d2; S.Lr[2] = a3 / d3;
    for (int k = 0; k < 3; ++k) {
      double lo = p.origin[k] - p.box[k] * 0.5, hi = p.origin[k] + p.box[k] * 0.5;
      double mn = std::min(lo, hi), mx = std::max(lo, hi);
      S.H[k] = mx - mn;
      S.H2[k] = S.H[k] * 0.5;
      S.omin[k] = mn;
    }
    double dx = (Lx * 0.5) * (Lx * 0.5), dy = (Ly * 0.5) * (Ly * 0.5), dz = (Lz * 0.5) * (Lz * 0.5);
    S.D = std::min(dx, std::min(dy, dz));
  }
  S.n_corr = h.n_corr = p.n_corr;
  S.n_bonds = p.n_bonds;
  S.n_angles = p.n_angles;
  h.n_bonded = p.n_bonds + p.n_angles;
  h.n_local = (int)p.loc_item.size();
  h.n_gb_tiles = p.n_gb_tiles;
  h.n_gb_pairs = c->gb ? pm::gb_pair_count(p.n_gb_tiles) : 0;
  h.coul_const = p.coul_const;
  h.vir_const = p.vir_const;
  S.ncx = p.ncx;
  S.ncy = p.ncy;
  S.ncol = h.ncol = p.ncol;
  S.ax = p.ax;
  S.ay = p.ay;
  S.rx = p.rx;
  S.ry = p.ry;
  S.cl_cap = h.cl_cap = p.cl_cap;
  S.sc_cap = h.sc_cap = p.sc_cap;
  S.slot_cap = h.slot_cap = p.slot_cap;
  h.list_cap = p.list_cap;
  if (c->recip) {
    S.ew_nk = p.ew_nk;
    S.ew_nrows = (int)p.ew_rows.size() - 1;
    S.ew_xrows = p.ew_xrows;
    S.ew_yrows = p.ew_yrows;
    S.ew_zrows = p.ew_zrows;
    S.ew_chunk = p.ew_chunk;
    S.ew_nchunks = p.ew_nchunks;
    S.ew_ntiles = (int)p.ew_tiles.size();
    S.ew_rl = p.ew_xrows + p.ew_yrows + p.ew_zrows;
    S.ew_nfb = (int)p.ew_fb_rows.size() - 1;
    for (int k = 0; k < 3; ++k) S.ew_org[k] = d.origin[k];
    S.ew_c = p.ew_c;
    S.ew_a8 = p.ew_a8;
    S.ew_fac = p.ew_fac;
    h.ew.n = n;
    h.ew.nk = p.ew_nk;
    h.ew.ntiles = S.ew_ntiles;
    h.ew.nchunks = S.ew_nchunks;
    h.ew.nfb = S.ew_nfb;
    h.ew.lds_bytes = 0;
  }
  Carver dry;
  SysDev tmp{};
  int *snap = nullptr;
  h.block_bytes = carve(c, p, d, dry, tmp, &snap);
}

// Lay the system's arrays out in h.block, clear them (a fill kernel on the
// context's stream when `fill`, else a memset) and upload its tables: into
// `up` when given (copied by flush_uploads, on the same stream after the
// fill), else at once.  `p` must outlive the flush.
void fill_block(pm_ctx *c, const pm_system_desc &d, const SysPrep &p, SysHost &h, bool fill, Uploads *up = nullptr) {
  SysDev &S = h.dev;
  const int n = p.n;
  Carver cv;
  cv.base = static_cast<char *>(h.block);
  carve(c, p, d, cv, S, &h.ctl_snap);
  // (the fill covers block_bytes rounded up to 8; the static layout keeps the
  // block's end padded to 256 bytes)
  if (fill) pm::launch_fill_zero(h.block, (h.block_bytes + 7) / 8 * 8, c->stream);
  else HIPCHECK(hipMemset(h.block, 0, h.block_bytes));
  upload(up, const_cast<int *>(S.excl_start), p.excl_start.data(), p.excl_start.size());
  upload(up, const_cast<int *>(S.excl_partner), p.excl_partner.data(), p.excl_partner.size());
  upload(up, const_cast<int *>(S.excl_class), p.excl_class.data(), p.excl_class.size());
  upload(up, const_cast<unsigned long long *>(S.excl_win), p.excl_win.data(), p.excl_win.size());
  upload(up, const_cast<pm::LocRec *>(S.loc_rec), p.loc_rec.data(), p.loc_rec.size());
  upload(up, const_cast<double *>(S.charge), d.scaled_charge, n);
  upload(up, const_cast<double *>(S.inv_mass), p.invm.data(), n);
  upload(up, const_cast<int *>(S.molecule), d.molecule, n);
  if (c->molvir) {
    upload(up, const_cast<int *>(S.mol_start), p.mol_start.data(), p.mol_start.size());
    upload(up, const_cast<int *>(S.mol_atoms), p.mol_atoms.data(), p.mol_atoms.size());
    upload(up, const_cast<double *>(S.mass), d.mass, n);
  }
  if (d.lj_mode == PM_LJ_PERATOM) {
    upload(up, const_cast<float *>(S.sig_half), p.sig_half.data(), n);
    upload(up, const_cast<float *>(S.eps_sqrt), p.eps_sqrt.data(), n);
  } else {
    upload(up, const_cast<int *>(S.type), d.type, n);
    upload(up, const_cast<float4 *>(S.lj_table), p.lj.data(), p.lj.size());
    upload(up, const_cast<float4 *>(S.lj_table12), p.lj12.data(), p.lj12.size());
  }
  if (c->gb) {
    if (c->P.gb_on) {
      upload(up, const_cast<double *>(S.gb_rho), d.gb_radius, n);
      upload(up, const_cast<double *>(S.gb_S), d.gb_scaling, n);
    }
    upload(up, S.slot, p.gb_ident.data(), n);
    S.ncx = S.ncy = S.ncol = 1;
  }
  if (c->recip) {
    upload(up, const_cast<int4 *>(S.ew_rows), p.ew_rows.data(), p.ew_rows.size());
    upload(up, const_cast<double4 *>(S.ew_kd), p.ew_kd.data(), p.ew_kd.size());
    upload(up, const_cast<int4 *>(S.ew_tiles), p.ew_tiles.data(), p.ew_tiles.size());
    upload(up, const_cast<short2 *>(S.ew_pts), p.ew_pts.data(), p.ew_pts.size());
    upload(up, const_cast<int *>(S.ew_kmap), p.ew_kmap.data(), p.ew_kmap.size());
    upload(up, const_cast<int *>(S.ew_fb_rows), p.ew_fb_rows.data(), p.ew_fb_rows.size());
  }
}

// Place a prepared system in the context: its device block (and pair
// lists), cleared, with its tables queued in `up`.  Returns its id, or -1 when
// the reservation has no room for it (nothing kept).
int commit_system(pm_ctx *c, const pm_system_desc &d, const SysPrep &p, Uploads &up) {
  auto hp = std::make_unique<SysHost>();
  SysHost &h = *hp;
  init_system(c, d, p, h);
  try {
    h.block = dev_alloc(c, h.block_bytes, true);
    fill_block(c, d, p, h, true, &up);
    if (!c->gb) alloc_lists(c, h, true);
  } catch (const Fail &f) {
    if (h.block) {
      // drop this system's queued uploads (they point into its block)
      char *b0 = static_cast<char *>(h.block), *b1 = b0 + h.block_bytes;
      up.segs.erase(std::remove_if(up.segs.begin(), up.segs.end(),
                                   [&](const Uploads::Seg &g) { return g.dst >= b0 && g.dst < b1; }),
                    up.segs.end());
    }
    dev_free(c, h.lists, h.lists_bytes);
    dev_free(c, h.block, h.block_bytes);
    if (f.code == PM_EFULL) return -1;
    throw;
  }
  int id = -1;
  for (size_t i = 0; i < c->sys.size(); ++i)
    if (!c->sys[i]) { id = (int)i; break; }
  if (id < 0) {
    id = (int)c->sys.size();
    c->sys.emplace_back();
  }
  h.id = id;
  c->sys[id] = std::move(hp);
  drop_sets_with(c, id);  // a reused id must not find a stale table
  return id;
}

// Add systems in order: ids[k] = the new id, or -1 when the reservation has no
// room for system k; after max_full such systems the rest are not tried (-1).
// The host tables are prepared on worker threads, a chunk at a time, and each
// chunk's uploads travel in one staged copy per system.  Any other failure
// removes the systems this call added and is raised.
void add_many(pm_ctx *c, int n, const pm_system_desc *ds, int max_full, int32_t *ids) {
  for (int k = 0; k < n; ++k) ids[k] = -1;
  std::vector<int> added;
  int full = 0;
  try {
    constexpr int kChunk = 32;
    for (int k0 = 0; k0 < n && full < max_full; k0 += kChunk) {
      const int m = std::min(kChunk, n - k0);
      std::vector<SysPrep> preps(m);
      std::vector<std::exception_ptr> errs(m);
      parallel_for(m, [&](int i) {
        try {
          preps[i] = prepare(c, ds[k0 + i]);
        } catch (...) {
          errs[i] = std::current_exception();
        }
      });
      Uploads up;
      for (int i = 0; i < m && full < max_full; ++i) {
        if (errs[i]) std::rethrow_exception(errs[i]);
        const int id = commit_system(c, ds[k0 + i], preps[i], up);
        if (id < 0) {
          ++full;
          continue;
        }
        ids[k0 + i] = id;
        added.push_back(id);
      }
      flush_uploads(c, up);  // overlaps the next chunk's preparation
    }
    HIPCHECK(hipStreamSynchronize(c->stream));  // fills and uploads done
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    for (int id : added) {
      SysHost &h = *c->sys[id];
      dev_free(c, h.lists, h.lists_bytes);
      dev_free(c, h.block, h.block_bytes);
      c->sys[id].reset();
      drop_sets_with(c, id);
    }
    for (int k = 0; k < n; ++k) ids[k] = -1;
    throw;
  }
}

// add a system; returns its id
int add_system(pm_ctx *c, const pm_system_desc &d) {
  int32_t id = -1;
  add_many(c, 1, &d, 1, &id);
  if (id < 0) {
    const size_t need = footprint(c, d);
    throw Fail{PM_EFULL, "the device reservation has no room for " + std::to_string(need) + " bytes"};
  }
  return id;
}

void remove_system(pm_ctx *c, int id) {
  SysHost &h = sys_of(c, id);
  HIPCHECK(hipStreamSynchronize(c->stream));
  drop_sets_with(c, id);
  dev_free(c, h.lists, h.lists_bytes);
  dev_free(c, h.block, h.block_bytes);
  c->sys[id].reset();
}

// ------------------------------------------------------------ launch sets

LaunchSet &get_set(pm_ctx *c, const std::vector<int> &ids) {
  auto it = c->sets.find(ids);
  if (it != c->sets.end()) return *it->second;
  if (c->sets.size() >= 64) {
    HIPCHECK(hipStreamSynchronize(c->stream));  // tables may be in use by queued work
    c->sets.clear();
  }
  auto ls = std::make_unique<LaunchSet>();
  ls->ids = ids;
  const int ns = (int)ids.size();
  std::vector<SysDev> table(ns);
  constexpr int kPre = 10;
  std::vector<std::vector<int>> pre(kPre, std::vector<int>(1, 0));
  for (int k = 0; k < ns; ++k) {
    const SysHost &h = *c->sys[ids[k]];
    table[k] = h.dev;
    const int cnt[kPre] = {h.n, h.cl_cap, h.ncol, h.slot_cap, h.n_corr, h.n_bonded, h.n_gb_tiles, h.n_local, h.sc_cap,
                           h.n_gb_pairs};
    for (int q = 0; q < kPre; ++q) pre[q].push_back(pre[q].back() + cnt[q]);
  }
  // one allocation: [SysDev x ns][kPre prefix arrays of ns + 1][kScalN int64 x ns][rebuild flag]
  const size_t o_pre = align_up((size_t)ns * sizeof(SysDev));
  const size_t o_scal = align_up(o_pre + kPre * (size_t)(ns + 1) * sizeof(int));
  const size_t o_flag = align_up(o_scal + (size_t)ns * pm::kScalN * sizeof(long long));
  const size_t bytes = o_flag + sizeof(int);
  HIPCHECK(hipMalloc(&ls->mem, bytes));
  char *base = static_cast<char *>(ls->mem);
  SysDev *d_tab = reinterpret_cast<SysDev *>(base);
  int *d_pre = reinterpret_cast<int *>(base + o_pre);
  ls->d_scal = reinterpret_cast<long long *>(base + o_scal);
  HIPCHECK(hipMemset(base + o_flag, 0, sizeof(int)));
  HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&ls->h_scal), (size_t)ns * pm::kScalN * sizeof(long long), 0));
  std::vector<int> flat;
  for (auto &v : pre) flat.insert(flat.end(), v.begin(), v.end());
  HIPCHECK(hipMemcpy(d_tab, table.data(), ns * sizeof(SysDev), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_pre, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice));
  Launch &L = ls->L;
  const int st = ns + 1;
  L.stream = c->stream;
  L.sys = d_tab;
  L.nsys = ns;
  L.atom_prefix = d_pre;
  L.cl_prefix = d_pre + st;
  L.col_prefix = d_pre + 2 * st;
  L.slot_prefix = d_pre + 3 * st;
  L.corr_prefix = d_pre + 4 * st;
  L.bonded_prefix = d_pre + 5 * st;
  L.gb_prefix = d_pre + 6 * st;
  L.gbp_prefix = d_pre + 9 * st;
  L.sc_prefix = d_pre + 8 * st;
  L.rb_any = reinterpret_cast<int *>(base + o_flag);
  // test knob: a lower limit on the list build's LDS excluded pairs, so
  // tests can drive the per-entry partner walk
  if (const char *hc = std::getenv("PM_DEBUG_HIT_CAP")) L.hit_cap = std::max(0, std::atoi(hc));
  L.n_atoms_total = pre[0].back();
  L.n_cl_total = pre[1].back();
  L.n_cols_total = pre[2].back();
  L.n_slots_total = pre[3].back();
  L.n_corr_total = pre[4].back();
  L.n_bonded_total = pre[5].back();
  L.n_gb_tiles_total = pre[6].back();
  L.n_gb_pairs_total = pre[9].back();
  L.n_local_total = pre[7].back();
  L.n_sc_total = pre[8].back();
  for (int k = 0; k < ns; ++k) L.max_mol = std::max(L.max_mol, c->sys[ids[k]]->dev.n_mol);
  if (c->recip) {
    for (int k = 0; k < ns; ++k) {
      pm::EwLaunch e = c->sys[ids[k]]->ew;
      e.sys = d_tab + k;
      ls->ew.push_back(e);
    }
    L.ew = ls->ew.data();
    L.n_ew = (int)ls->ew.size();
  }
  LaunchSet &ref = *ls;
  c->sets[ids] = std::move(ls);
  return ref;
}

// ------------------------------------------------------------ force evaluation

// Rebuild policy of MD: the list (radius rc + skin) is rebuilt when an atom
// has moved more than skin/2 from where it was built, or after
// rebuild_interval steps when that is > 0.
struct StepPolicy {
  float rlist = 0.f, lim2 = 0.f;
  int kmax = 0;
};

StepPolicy md_policy(const pm_ctx *c) {
  StepPolicy p;
  p.rlist = (float)(c->rc + c->skin_md);
  const double lim = std::max(0.0, 0.5 * c->skin_md - 1e-3);
  p.lim2 = (float)(lim * lim);
  p.kmax = std::max(0, c->desc.rebuild_interval);
  return p;
}

StepPolicy eval_policy(const pm_ctx *c) {
  StepPolicy p;
  p.rlist = (float)(c->rc + c->skin_eval);
  return p;
}

hipEvent_t next_event(pm_ctx *c) {
  if (c->evs_used == (int)c->evs.size()) {
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    c->evs.push_back(e);
  }
  return c->evs[c->evs_used++];
}

// One force evaluation of the systems of L at the positions in S.x.
// mode kStepForce rebuilds the list and zeroes the forces first; kStepAuto
// (MD steps, whose kick has consumed and zeroed the forces) leaves the rebuild
// decision to the device.  kind: pm::kPair* (what the pair kernel adds).
// gb_energies: the GB passes also sum energies (GB has no separate energy pass).
// reciprocal-space energy-pass flags: the virial bits the context asks for
int ewald_energy_mode(const pm_ctx *c) {
  return pm::kEwEnergies | (c->P.virial ? pm::kEwVirial : 0) | (c->molvir ? pm::kEwMolVirial : 0);
}

void force_eval(pm_ctx *c, const Launch &L, int mode, const StepPolicy &pol, int kind, bool gb_energies, bool timed) {
  // the cutoff path's list check and bookkeeping ride on the rebuild kernels
  if (c->gb) pm::launch_step_begin(L, mode, pol.lim2, pol.kmax);
  const bool pool = !timed && (c->desc.flags & PM_FLAG_TIMING);
  // (the cutoff path's list check clears the pair forces itself)
  if (c->gb && (mode & pm::kStepForce)) pm::launch_zero_forces(L);
  const bool en = kind == pm::kPairForcesEnergies;
  const bool local_en = en || (c->gb && gb_energies);
  const int local = pm::kLocalForces | (local_en ? pm::kLocalEnergies : 0);
  // the molecular virial reads the molecules' centres at these positions
  if (c->molvir && local_en) pm::launch_mol_com(L);
  if (c->gb) {
    if (timed) HIPCHECK(hipEventRecord(c->evp0, L.stream));
    if (pool) HIPCHECK(hipEventRecord(next_event(c), L.stream));
    pm::launch_gb(L, c->P, en || gb_energies);
    if (pool) HIPCHECK(hipEventRecord(next_event(c), L.stream));
    if (timed) HIPCHECK(hipEventRecord(c->evp1, L.stream));
    pm::launch_local(L, c->P, local);
    return;
  }
  // a forces-only pass computes the local terms' forces in the list check's
  // launch (they read the same positions)
  const bool local_done =
      pm::launch_rebuild(L, pol.rlist * pol.rlist, mode, pol.lim2, pol.kmax, kind == pm::kPairForces ? &c->P : nullptr);
  if (timed) HIPCHECK(hipEventRecord(c->evp0, L.stream));
  if (pool) HIPCHECK(hipEventRecord(next_event(c), L.stream));
  if (!pm::launch_pair(L, c->P, en && c->molvir ? pm::kPairForcesEnergiesMV : kind))
    throw Fail{PM_EINVAL, "unsupported term combination"};
  if (pool) HIPCHECK(hipEventRecord(next_event(c), L.stream));
  if (timed) HIPCHECK(hipEventRecord(c->evp1, L.stream));
  if (c->recip && kind != pm::kPairCensus) {
    if (timed) HIPCHECK(hipEventRecord(c->evr0, L.stream));
    pm::launch_ewald(L, pm::kEwForces | (en ? ewald_energy_mode(c) : 0));
    if (timed) HIPCHECK(hipEventRecord(c->evr1, L.stream));
  }
  if (!local_done) pm::launch_local(L, c->P, local);
}

// The pair-kernel kind of an MD step's force pass: a step that sums energies
// does it in the same pass as the forces (GB sums its energies inside its own
// passes).  Every kind rounds the forces alike, so the trajectory does not
// depend on which steps sum energies.
int step_kind(const pm_ctx *c, bool energies) {
  return energies && !c->gb ? pm::kPairForcesEnergies : pm::kPairForces;
}

// copy the packed scalars of the set's systems to the host (synchronizes)
void read_scalars(pm_ctx *c, LaunchSet &ls) {
  pm::launch_gather_scalars(ls.L, ls.d_scal);
  HIPCHECK(hipMemcpyAsync(ls.h_scal, ls.d_scal, ls.ids.size() * pm::kScalN * sizeof(long long),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
}

void add_energies(const pm_ctx *c, const SysHost &h, const long long *acc, double *e) {
  const double k = pm::kEScaleInv;
  e[PM_E_COULOMB] += (double)acc[pm::kAccCoul] * k + h.coul_const;
  e[PM_E_LENNARDJONES] += (double)acc[pm::kAccLJ] * k;
  e[PM_E_BOND] += (double)acc[pm::kAccBond] * k;
  e[PM_E_ANGLE] += (double)acc[pm::kAccAngle] * k;
  // the atomic virial slots: everything with the virial on; with only the
  // molecular virial on, the pairs of two molecules (the reference's
  // three-argument addVirial writes both)
  const bool vir = (c->desc.flags & PM_FLAG_VIRIAL) != 0;
  if (vir || c->molvir) {
    const double vc = vir ? h.vir_const : 0.0;
    const double xx = (double)acc[pm::kAccVxx] * k + vc, xy = (double)acc[pm::kAccVxy] * k,
                 xz = (double)acc[pm::kAccVxz] * k, yy = (double)acc[pm::kAccVyy] * k + vc,
                 yz = (double)acc[pm::kAccVyz] * k, zz = (double)acc[pm::kAccVzz] * k + vc;
    double *v = e + PM_E_VIRIALXX;
    v[0] += xx; v[1] += xy; v[2] += xz;
    v[3] += xy; v[4] += yy; v[5] += yz;
    v[6] += xz; v[7] += yz; v[8] += zz;
  }
  if (c->molvir) {
    const double xx = (double)acc[pm::kAccMVxx] * k, xy = (double)acc[pm::kAccMVxy] * k,
                 xz = (double)acc[pm::kAccMVxz] * k, yy = (double)acc[pm::kAccMVyy] * k,
                 yz = (double)acc[pm::kAccMVyz] * k, zz = (double)acc[pm::kAccMVzz] * k;
    double *v = e + PM_E_MOLVIRIALXX;
    v[0] += xx; v[1] += xy; v[2] += xz;
    v[3] += xy; v[4] += yy; v[5] += yz;
    v[6] += xz; v[7] += yz; v[8] += zz;
  }
}

// Raise on device status bits; grow the pair lists that overflowed.
// Returns the ids whose lists grew (their results must be recomputed).
// A list overflow is handled first: a system whose list overflowed may have
// lost pairs (or exclusion masks) in the steps after it, so a non-finite
// position it reports at the same time is not yet an error -- the grown,
// re-run call decides.
std::vector<int> check_errors(pm_ctx *c, LaunchSet &ls) {
  std::vector<int> grown;
  bool any = false;
  for (size_t k = 0; k < ls.ids.size(); ++k) {
    const unsigned e = (unsigned)ls.h_scal[k * pm::kScalN + pm::kScalErr];
    if (!e) continue;
    any = true;
    const int id = ls.ids[k];
    if (e & pm::kErrListCap) {
      grown.push_back(id);
      continue;
    }
    if (e & pm::kErrNonFinite) { pm::launch_clear_err(ls.L); throw Fail{PM_EINVAL, "system " + std::to_string(id) + ": non-finite position"}; }
    if (e & pm::kErrColumnCap) {
      pm::launch_clear_err(ls.L);
      throw Fail{PM_ECAPACITY, "system " + std::to_string(id) + ": more than " + std::to_string(pm::kSortCap) +
                                   " atoms in one cell column"};
    }
    if (e & pm::kErrClusterCap) { pm::launch_clear_err(ls.L); throw Fail{PM_ECAPACITY, "system " + std::to_string(id) + ": cluster capacity"}; }
    if (e & pm::kErrVacuumRegion) {
      pm::launch_clear_err(ls.L);
      throw Fail{PM_ECAPACITY, "system " + std::to_string(id) +
                                   ": an atom left the region (origin +- box / 2) of the vacuum system's cell grid"};
    }
  }
  if (!any) return grown;
  pm::launch_clear_err(ls.L);
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int id : grown) {
    SysHost &h = *c->sys[id];
    h.list_cap *= 2;
    alloc_lists(c, h);
  }
  drop_sets(c);  // the device records changed
  return grown;
}

// H2D of concatenated AoS host arrays into the systems' staging buffers
void put_staging(pm_ctx *c, const std::vector<int> &ids, const double *x, const double *v) {
  size_t atoms = 0;
  for (int id : ids) atoms += c->sys[id]->n;
  const int na = (x ? 1 : 0) + (v ? 1 : 0);
  const size_t arr = 3 * atoms * sizeof(double);
  if (na * arr >= kPinnedTransfer) {
    // large: through the pinned buffer (host copies on worker threads, DMA per array)
    char *pin = pinned(c, na * arr);
    const double *src[2] = {x ? x : v, x ? v : nullptr};
    parallel_for(16 * na, [&](int k) {
      const int a = k / 16, part = k % 16;
      const size_t b0 = arr * part / 16, b1 = arr * (part + 1) / 16;
      std::memcpy(pin + a * arr + b0, reinterpret_cast<const char *>(src[a]) + b0, b1 - b0);
    });
    size_t off = 0;
    for (int id : ids) {
      const SysHost &h = *c->sys[id];
      const size_t b = 3 * (size_t)h.n * sizeof(double);
      const char *px = pin + 3 * off * sizeof(double);
      if (x) HIPCHECK(hipMemcpyAsync(h.dev.xaos, px, b, hipMemcpyHostToDevice, c->stream));
      if (v) HIPCHECK(hipMemcpyAsync(h.dev.vaos, px + (x ? arr : 0), b, hipMemcpyHostToDevice, c->stream));
      off += h.n;
    }
    HIPCHECK(hipEventRecord(c->pin_done, c->stream));
    return;
  }
  size_t off = 0;
  for (int id : ids) {
    const SysHost &h = *c->sys[id];
    const size_t b = 3 * (size_t)h.n * sizeof(double);
    if (x) HIPCHECK(hipMemcpyAsync(h.dev.xaos, x + 3 * off, b, hipMemcpyHostToDevice, c->stream));
    if (v) HIPCHECK(hipMemcpyAsync(h.dev.vaos, v + 3 * off, b, hipMemcpyHostToDevice, c->stream));
    off += h.n;
  }
}

// D2H of the systems' staging buffers into concatenated AoS host arrays; the
// copies have finished when it returns
void get_staging(pm_ctx *c, const std::vector<int> &ids, double *x, double *v, double *f) {
  size_t atoms = 0;
  for (int id : ids) atoms += c->sys[id]->n;
  double *dst[3] = {x, v, f};
  int na = 0;
  for (double *d : dst) na += d ? 1 : 0;
  const size_t arr = 3 * atoms * sizeof(double);
  if (na * arr >= kPinnedTransfer) {
    char *pin = pinned(c, na * arr);
    size_t off = 0;
    for (int id : ids) {
      const SysHost &h = *c->sys[id];
      const size_t b = 3 * (size_t)h.n * sizeof(double);
      int a = 0;
      const double *srcs[3] = {h.dev.xaos, h.dev.vaos, h.dev.faos};
      for (int k = 0; k < 3; ++k)
        if (dst[k]) {
          HIPCHECK(hipMemcpyAsync(pin + a * arr + 3 * off * sizeof(double), srcs[k], b, hipMemcpyDeviceToHost,
                                  c->stream));
          ++a;
        }
      off += h.n;
    }
    HIPCHECK(hipEventRecord(c->pin_done, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    std::vector<double *> outs;
    for (double *d : dst)
      if (d) outs.push_back(d);
    parallel_for(16 * na, [&](int k) {
      const int a = k / 16, part = k % 16;
      const size_t b0 = arr * part / 16, b1 = arr * (part + 1) / 16;
      std::memcpy(reinterpret_cast<char *>(outs[a]) + b0, pin + a * arr + b0, b1 - b0);
    });
    return;
  }
  size_t off = 0;
  for (int id : ids) {
    const SysHost &h = *c->sys[id];
    const size_t b = 3 * (size_t)h.n * sizeof(double);
    if (x) HIPCHECK(hipMemcpyAsync(x + 3 * off, h.dev.xaos, b, hipMemcpyDeviceToHost, c->stream));
    if (v) HIPCHECK(hipMemcpyAsync(v + 3 * off, h.dev.vaos, b, hipMemcpyDeviceToHost, c->stream));
    if (f) HIPCHECK(hipMemcpyAsync(f + 3 * off, h.dev.faos, b, hipMemcpyDeviceToHost, c->stream));
    off += h.n;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
}

size_t atoms_of(const pm_ctx *c, const std::vector<int> &ids) {
  size_t n = 0;
  for (int id : ids) n += c->sys[id]->n;
  return n;
}

// SystemForce::evaluate for the systems `ids` at host positions x: ADDS into f, e
void evaluate_ids(pm_ctx *c, const std::vector<int> &ids, const double *x, double *f, double *e, bool census,
                  uint64_t *count, uint64_t *hash) {
  for (int id : ids) c->sys[id]->md_ready = false;  // the positions are overwritten
  put_staging(c, ids, x, nullptr);
  for (int attempt = 0;; ++attempt) {
    LaunchSet &ls = get_set(c, ids);
    HIPCHECK(hipEventRecord(c->ev0, c->stream));
    pm::launch_import(ls.L, pm::kXferX);
    force_eval(c, ls.L, pm::kStepForce, eval_policy(c), census ? pm::kPairCensus : pm::kPairForcesEnergies, false,
               true);
    HIPCHECK(hipEventRecord(c->ev1, c->stream));
    read_scalars(c, ls);
    if (!check_errors(c, ls).empty()) {
      if (attempt > 4) throw Fail{PM_ECAPACITY, "pair list capacity kept overflowing"};
      continue;
    }
    pm::launch_export(ls.L, pm::kXferF);
    const size_t n = atoms_of(c, ids);
    std::vector<double> fo(f ? 3 * n : 0);
    get_staging(c, ids, nullptr, nullptr, f ? fo.data() : nullptr);
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (f)
      for (size_t k = 0; k < fo.size(); ++k) f[k] += fo[k];
    if (e)
      for (size_t k = 0; k < ids.size(); ++k) add_energies(c, *c->sys[ids[k]], &ls.h_scal[k * pm::kScalN], e + k * PM_E_NSLOTS);
    if (census) {
      *count = (uint64_t)ls.h_scal[pm::kScalCensus];
      *hash = (uint64_t)ls.h_scal[pm::kScalCensus + 1];
    }
    float ms = 0.f, pms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    HIPCHECK(hipEventElapsedTime(&pms, c->evp0, c->evp1));
    c->stats[0] = 1 + attempt;
    c->stats[2] = 0;
    for (size_t k = 0; k < ids.size(); ++k) c->stats[2] += (double)ls.h_scal[k * pm::kScalN + pm::kScalCtl + pm::kCtlEntries];
    c->stats[3] = ms;
    c->stats[4] = pms;
    c->stats[5] = 1;
    c->stats[8] = 0.0;
    c->stats[9] = 0.0;
    if (c->recip && !census) {
      float rms = 0.f;
      HIPCHECK(hipEventElapsedTime(&rms, c->evr0, c->evr1));
      c->stats[8] = rms;
      for (int id : ids) c->stats[9] += c->sys[id]->ew.nk;
    }
    return;
  }
}

// StandardIntegrator::initializeForces: positions/velocities in, F(x0) computed
void set_state_ids(pm_ctx *c, const std::vector<int> &ids, const double *x, const double *v) {
  put_staging(c, ids, x, v);
  for (int attempt = 0;; ++attempt) {
    LaunchSet &ls = get_set(c, ids);
    pm::launch_import(ls.L, pm::kXferX | pm::kXferV);
    // a new state starts its Langevin noise stream at counter 0 (the force
    // evaluation below advances it to 1)
    for (int id : ids)
      HIPCHECK(hipMemsetAsync(c->sys[id]->dev.ctl + pm::kCtlNoise, 0, sizeof(int), c->stream));
    force_eval(c, ls.L, pm::kStepForce, md_policy(c), pm::kPairForces, false, false);
    read_scalars(c, ls);
    if (!check_errors(c, ls).empty()) {
      if (attempt > 4) throw Fail{PM_ECAPACITY, "pair list capacity kept overflowing"};
      continue;
    }
    break;
  }
  for (int id : ids) c->sys[id]->md_ready = true;
}

void require_md(pm_ctx *c, const std::vector<int> &ids) {
  for (int id : ids)
    if (!c->sys[id]->md_ready) throw Fail{PM_ENOTREADY, "system " + std::to_string(id) + ": set the MD state first"};
}

constexpr int kGraphSteps = 10;

// kGraphSteps MD steps of set ls for time step h, captured on first use.
// Kernel arguments are the set's own device tables, which live as long as the
// set does, so the graph stays valid with it.
// LangevinLeapfrogIntegrator constants for time step dt_fs
// (LangevinLeapfrogIntegrator.cpp:24-26, 46-54)
pm::LangevinDev lv_params(const pm_ctx *c, double dt_fs) {
  pm::LangevinDev lv;
  if (!(c->lv_gamma_ps > 0.0)) return lv;
  const double g = c->lv_gamma_ps / (1000 * kInvTimeFactor);  // myGamma
  const double dt = dt_fs * kInvTimeFactor;
  lv.on = 1;
  lv.fdt = (1.0 - std::exp(-0.5 * g * dt)) / g;
  lv.vdt = std::exp(-0.5 * g * dt);
  lv.ndt = std::sqrt((1.0 - std::exp(-g * dt)) / (2.0 * g));
  lv.fc = 2 * kBoltzmann * c->lv_temperature * g;
  return lv;
}

bool same_lv(const pm::LangevinDev &a, const pm::LangevinDev &b) {
  return a.on == b.on && a.vdt == b.vdt && a.fdt == b.fdt && a.ndt == b.ndt && a.fc == b.fc;
}

hipGraphExec_t md_graph(pm_ctx *c, LaunchSet &ls, double h, const StepPolicy &pol, bool energies) {
  const pm::LangevinDev lv = lv_params(c, h / kInvTimeFactor);
  const int g_id = energies ? 1 : 0;
  hipGraphExec_t &gx = ls.gexec[g_id];
  if (gx && ls.g_h[g_id] == h && same_lv(ls.g_lv[g_id], lv)) return gx;
  if (gx) {
    HIPCHECK(hipGraphExecDestroy(gx));
    gx = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < kGraphSteps; ++k) {
      const bool en = energies && k == kGraphSteps - 1;
      pm::launch_kick_drift(ls.L, h, lv, c->gb);
      force_eval(c, ls.L, pm::kStepAuto, pol, step_kind(c, en), en, false);
    }
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHECK(hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    gx = nullptr;
    throw Fail{PM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
  }
  ls.g_h[g_id] = h;
  ls.g_lv[g_id] = lv;
  return gx;
}

// LeapfrogIntegrator::run (LeapfrogIntegrator.cpp:89-104) in three parts:
// md_begin arms the first kick as a half kick, md_advance runs n steps of
// (kick + drift, forces), md_end applies the closing half kick.
void md_begin(pm_ctx *c, const std::vector<int> &ids) {
  require_md(c, ids);
  pm::launch_md_arm(get_set(c, ids).L);
}

// every > 0: the energy pass also runs on every every-th step of the advance
// (an output frequency); energies_last: on the last step.  The energies of
// the last step that ran the pass are what pm_md_get_state reads when that
// step is the last one.
void md_advance(pm_ctx *c, const std::vector<int> &ids, int n_steps, double dt_fs, bool energies_last, int every = 0,
                int depth = 0) {
  require_md(c, ids);
  if (n_steps < 1) return;
  if (!(dt_fs > 0.0)) throw Fail{PM_EINVAL, "dt must be > 0"};
  const double h = dt_fs * kInvTimeFactor;
  const StepPolicy pol = md_policy(c);
  const bool timing = (c->desc.flags & PM_FLAG_TIMING) != 0;
  {
    LaunchSet &ls = get_set(c, ids);
    // state at the start: a pair list that overflows during the advance
    // (capacity grown) sends its system back here to run the advance again
    pm::launch_export(ls.L, pm::kXferX | pm::kXferV);
    for (int id : ids)
      HIPCHECK(hipMemcpyAsync(c->sys[id]->ctl_snap, c->sys[id]->dev.ctl, pm::kCtlN * sizeof(int),
                              hipMemcpyDeviceToDevice, c->stream));
    if (depth == 0) {
      HIPCHECK(hipEventRecord(c->ev0, c->stream));
      c->evs_used = 0;
    }
    if (every >= n_steps) every = 0;  // no energy step before the last
    auto energy_step = [&](int i) { return (energies_last && i == n_steps - 1) || (every > 0 && (i + 1) % every == 0); };
    int i = 0;
    // whole graphs of kGraphSteps steps while the energy steps fall where a
    // graph has them: nowhere, or (every = kGraphSteps) on its last step
    const bool graphs = !timing && (every == 0 || every == kGraphSteps);
    if (graphs)
      while (n_steps - i - (energies_last ? 1 : 0) >= kGraphSteps) {
        HIPCHECK(hipGraphLaunch(md_graph(c, ls, h, pol, every == kGraphSteps), c->stream));
        i += kGraphSteps;
      }
    for (; i < n_steps; ++i) {
      const bool en = energy_step(i);
      pm::launch_kick_drift(ls.L, h, lv_params(c, dt_fs), c->gb);
      force_eval(c, ls.L, pm::kStepAuto, pol, step_kind(c, en), en, false);
    }
    if (depth == 0) HIPCHECK(hipEventRecord(c->ev1, c->stream));
    read_scalars(c, ls);
  }
  LaunchSet &ls = get_set(c, ids);
  std::vector<long long> scal(ls.h_scal, ls.h_scal + ids.size() * pm::kScalN);
  const std::vector<int> grown = check_errors(c, ls);
  if (!grown.empty()) {
    if (depth > 4) throw Fail{PM_ECAPACITY, "pair list capacity kept overflowing"};
    // restore the grown systems to the start of the advance and run it again:
    // positions and velocities, forces at them from a fresh list, and the
    // control words that evaluation advanced -- the run's rebuild count, the
    // first-kick flag and the Langevin noise counter -- so the re-run draws
    // the noise the first attempt drew.  The steps-since-rebuild word stays
    // at the 0 of the fresh list.
    LaunchSet &rs = get_set(c, grown);
    pm::launch_import(rs.L, pm::kXferX | pm::kXferV);
    force_eval(c, rs.L, pm::kStepForce, pol, pm::kPairForces, false, false);
    for (int id : grown) {
      static_assert(pm::kCtlFirst == pm::kCtlCount + 1, "count and first-kick words are adjacent");
      HIPCHECK(hipMemcpyAsync(c->sys[id]->dev.ctl + pm::kCtlCount, c->sys[id]->ctl_snap + pm::kCtlCount,
                              2 * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
      HIPCHECK(hipMemcpyAsync(c->sys[id]->dev.ctl + pm::kCtlNoise, c->sys[id]->ctl_snap + pm::kCtlNoise,
                              sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    }
    md_advance(c, grown, n_steps, dt_fs, energies_last, every, depth + 1);
  }
  if (depth == 0) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    LaunchSet &fs = get_set(c, ids);
    read_scalars(c, fs);
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    double pair_ms = 0.0;
    for (int k = 0; k + 1 < c->evs_used; k += 2) {
      float t = 0.f;
      HIPCHECK(hipEventElapsedTime(&t, c->evs[k], c->evs[k + 1]));
      pair_ms += t;
    }
    unsigned worst = 0;
    double rebuilds = 0, entries = 0;
    for (size_t k = 0; k < ids.size(); ++k) {
      worst = std::max(worst, (unsigned)fs.h_scal[k * pm::kScalN + pm::kScalMaxdispRun]);
      rebuilds += (double)fs.h_scal[k * pm::kScalN + pm::kScalCtl + pm::kCtlCount];
      entries += (double)fs.h_scal[k * pm::kScalN + pm::kScalCtl + pm::kCtlEntries];
    }
    c->stats[2] = entries;
    float w;
    std::memcpy(&w, &worst, sizeof w);
    c->stats[0] = rebuilds;
    c->stats[1] = grown.empty() ? 0 : 1;
    c->stats[3] = ms;
    c->stats[4] = pair_ms;
    c->stats[5] = c->evs_used / 2;
    c->stats[6] = std::sqrt((double)w);
    c->stats[7] = c->skin_md;
  }
  (void)scal;
}

void md_end(pm_ctx *c, const std::vector<int> &ids, double dt_fs) {
  require_md(c, ids);
  // StandardIntegrator::doHalfKick (StandardIntegrator.cpp:80-92), or the
  // Langevin half kick
  pm::launch_half_kick(get_set(c, ids).L, 0.5 * dt_fs * kInvTimeFactor, lv_params(c, dt_fs));
}

void get_state_ids(pm_ctx *c, const std::vector<int> &ids, double *x, double *v, double *f, double *e) {
  require_md(c, ids);
  LaunchSet &ls = get_set(c, ids);
  pm::launch_export(ls.L, pm::kXferX | pm::kXferV | pm::kXferF);
  get_staging(c, ids, x, v, f);
  read_scalars(c, ls);
  if (e) {
    std::memset(e, 0, sizeof(double) * PM_E_NSLOTS * ids.size());
    for (size_t k = 0; k < ids.size(); ++k) add_energies(c, *c->sys[ids[k]], &ls.h_scal[k * pm::kScalN], e + k * PM_E_NSLOTS);
  }
}

// The force description's device record and the context's cutoff and list
// radii, validated as ProtoMol's make() of each force does; n_systems /
// systems (may be 0) bound the pair-list buffer by the smallest box.
void setup_force(pm_ctx *c, const pm_force_desc *force, int32_t n_systems, const pm_system_desc *systems) {
  ForceDev &P = c->P;
  std::memset(&P, 0, sizeof P);
  P.flags = force->flags;
  P.finish_step = 1;
  P.pair_waves = force->pair_waves ? force->pair_waves : 4;
  P.pair_split = force->pair_split ? force->pair_split : pm::kPairSplitDefault;
  if (P.pair_split < 1 || P.pair_split > 64) throw Fail{PM_EINVAL, "pair_split must be 1 .. 64"};
  if (P.pair_waves != 1 && P.pair_waves != 2 && P.pair_waves != 4 && P.pair_waves != 8)
    throw Fail{PM_EINVAL, "pair_waves must be 1, 2, 4 or 8"};
  if (force->cell_scale != 0.0 && !(force->cell_scale >= 0.6 && force->cell_scale <= 2.0))
    throw Fail{PM_EINVAL, "cell_scale must be 0 (default) or within 0.6 .. 2"};
  c->gb = (force->flags & (PM_FLAG_GB | PM_FLAG_ALL_PAIRS)) != 0;  // the all-pairs (vacuum) path
  c->recip = (force->flags & PM_FLAG_EWALD_RECIPROCAL) != 0;
  c->molvir = (force->flags & PM_FLAG_MOLVIRIAL) != 0;
  P.virial = (force->flags & PM_FLAG_VIRIAL) ? 1 : 0;
  P.molvirial = c->molvir ? 1 : 0;
  if ((force->flags & PM_FLAG_GB_ACE) && !(force->flags & PM_FLAG_GB))
    throw Fail{PM_EINVAL, "PM_FLAG_GB_ACE needs PM_FLAG_GB"};
  if (c->recip && (c->gb || force->term_b != PM_TERM_EWALD_REAL))
    throw Fail{PM_EINVAL, "PM_FLAG_EWALD_RECIPROCAL needs term_b = PM_TERM_EWALD_REAL (and no GB)"};
  if (c->gb) {
    pm::gb_setup_force(P, *force);
  } else {
    const int ta = force->term_a, tb = force->term_b;
    const bool ok = (ta == PM_TERM_LJ && (tb == PM_TERM_NONE || tb == PM_TERM_COULOMB || tb == PM_TERM_EWALD_REAL)) ||
                    (ta == PM_TERM_COULOMB && tb == PM_TERM_NONE) || (ta == PM_TERM_NONE && tb == PM_TERM_EWALD_REAL);
    if (!ok) throw Fail{PM_EINVAL, "unsupported term combination"};
    P.term_a = ta;
    P.term_b = tb;
    double cA2 = 0, cB2 = 0;
    if (ta != PM_TERM_NONE) cA2 = build_switch(force->switch_a, P.sw_a, false, 0);
    if (tb == PM_TERM_EWALD_REAL) {
      // NonbondedFullEwaldSystemForce::initialize (NonbondedFullEwaldSystemForce.h:610-619)
      if (!(force->ewald_alpha > 0 && force->ewald_accuracy > 0 && force->ewald_accuracy < 1))
        throw Fail{PM_EINVAL, "Ewald needs alpha > 0 and 0 < accuracy < 1"};
      double p = -std::log(force->ewald_accuracy);
      double rc = std::sqrt(p) / force->ewald_alpha;
      cB2 = build_switch(force->switch_b, P.sw_b, true, rc);
      cB2 = rc * rc;
      P.alpha = (float)force->ewald_alpha;
      P.alpha_p = (float)(0.3275911 * force->ewald_alpha);
      P.alpha2_log2e = (float)(-force->ewald_alpha * force->ewald_alpha / std::log(2.0));
      P.two_alpha_pi = (float)(2.0 * force->ewald_alpha / std::sqrt(M_PI));
      P.dalpha = force->ewald_alpha;
      P.dtwo_alpha_pi = 2.0 * force->ewald_alpha / std::sqrt(M_PI);
    } else if (tb == PM_TERM_COULOMB) {
      cB2 = build_switch(force->switch_b, P.sw_b, false, 0);
      // OneAtomPairTwo includes a pair within the larger of the two switch
      // cutoffs (OneAtomPairTwo.h:25-27, 41-44) and applies both switches when
      // either modifies the values (:66-76).  A CutoffSwitchingFunction
      // (MODIFY = 0) then zeroes its own term beyond its cutoff, a step that
      // must fall where the FP64 test puts it: that term keeps its own
      // cutoff as its inclusion cutoff (FP64-checked like the other).
      // Otherwise a term is included out to the larger cutoff and its
      // switch takes it to zero.
      const double cmax = std::max(cA2, cB2);
      const bool modA = P.sw_a.kind != PM_SWITCH_CUTOFF, modB = P.sw_b.kind != PM_SWITCH_CUTOFF;
      if (modA || modB) {
        if (modA) cA2 = cmax;
        if (modB) cB2 = cmax;
      } else {
        cA2 = cB2 = cmax;
      }
    }
    P.cutA2d = cA2;
    P.cutB2d = cB2;
    P.cutA2 = (float)cA2;
    P.cutB2 = (float)cB2;
    const double cmax = std::max(cA2, cB2);
    P.guard = (float)(2e-3 + 1e-5 * cmax);
    P.cutA2lo = (float)(cA2 - P.guard);
    P.cutA2hi = (float)(cA2 + P.guard);
    P.cutB2lo = (float)(cB2 - P.guard);
    P.cutB2hi = (float)(cB2 + P.guard);
    c->rc = std::sqrt(cmax);
    c->skin_md = force->skin > 0 ? force->skin : 2.0;
    c->skin_eval = 0.1;
    if (force->rebuild_interval < 0) throw Fail{PM_EINVAL, "rebuild_interval must be >= 0"};
    // the pair list radius rc + skin must stay below half the smallest box
    // edge (one image per pair); a larger requested skin is reduced to fit
    // the systems given here
    if (n_systems > 0) {
      double minL = 1e300;
      for (int s = 0; s < n_systems; ++s)
        for (int k = 0; k < 3; ++k) minL = std::min(minL, systems[s].box[k]);
      const double room = 0.5 * minL - c->rc - 0.01;
      if (room < 0.05)
        throw Fail{PM_EINVAL, "box edge " + std::to_string(minL) + " leaves no room for a pair list beyond rc " +
                                  std::to_string(c->rc) + " (needs box > 2 rc)"};
      c->skin_md = std::min(c->skin_md, room);
      c->skin_eval = std::min(c->skin_eval, room);
    }
  }
}

// ------------------------------------------------ allocation-free evaluation

// The device memory of pm_evaluate_static: one file-scope array.  Each GPU
// the library runs on holds its own copy until the process ends.  Host code
// reaches a device's copy only through the address hipGetSymbolAddress gives
// on that device (scratch_ptr); the array's name on the host is a stand-in,
// not a device address.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

void *scratch_ptr() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Fail{PM_EDEVICE, "device id out of range"};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Fail{PM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

// SystemForce::evaluate of one system with every device array -- the
// system's block, its pair lists, the launch table and the packed scalars --
// carved out of the device's scratch array: no context, no device or pinned
// host allocation, no events; the null stream; synchronous on return.
void evaluate_static(const pm_force_desc &force, const pm_system_desc &d, const double *x, double *f, double *e) {
  pm_ctx c;
  c.desc = force;
  c.desc.flags &= ~PM_FLAG_TIMING;
  setup_force(&c, &c.desc, 1, &d);
  if (c.recip) throw Fail{PM_EINVAL, "pm_evaluate_static: no reciprocal-space term (use a context)"};
  char *const base = static_cast<char *>(scratch_ptr());
  const SysPrep p = prepare(&c, d);
  SysHost h;
  init_system(&c, d, p, h);
  for (int attempt = 0;; ++attempt) {
    // [system block | pair lists | SysDev | 10 prefix arrays of 2 ints | scalars]
    constexpr int kPre = 10;
    const size_t o_lists = align_up(h.block_bytes);
    const size_t lb = c.gb ? 0 : lists_bytes(h.sc_cap, h.list_cap);
    const size_t o_tab = align_up(o_lists + lb);
    const size_t o_pre = align_up(o_tab + sizeof(SysDev));
    const size_t o_scal = align_up(o_pre + kPre * 2 * sizeof(int));
    const size_t total = o_scal + pm::kScalN * sizeof(long long);
    if (total > kScratchBytes)
      throw Fail{PM_ECAPACITY, "pm_evaluate_static: the system needs " + std::to_string(total) +
                                   " bytes of device memory, more than the " + std::to_string(kScratchBytes) +
                                   " of the static scratch array (use a context)"};
    h.block = base;
    fill_block(&c, d, p, h, true);
    if (!c.gb) place_lists(h, base + o_lists);
    const int cnt[kPre] = {h.n, h.cl_cap, h.ncol, h.slot_cap, h.n_corr, h.n_bonded, h.n_gb_tiles, h.n_local, h.sc_cap,
                           h.n_gb_pairs};
    int pre[2 * kPre];
    for (int q = 0; q < kPre; ++q) {
      pre[2 * q] = 0;
      pre[2 * q + 1] = cnt[q];
    }
    SysDev *d_tab = reinterpret_cast<SysDev *>(base + o_tab);
    int *d_pre = reinterpret_cast<int *>(base + o_pre);
    long long *d_scal = reinterpret_cast<long long *>(base + o_scal);
    HIPCHECK(hipMemcpy(d_tab, &h.dev, sizeof(SysDev), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_pre, pre, sizeof pre, hipMemcpyHostToDevice));
    Launch L;
    L.stream = nullptr;
    L.sys = d_tab;
    L.nsys = 1;
    L.atom_prefix = d_pre;
    L.cl_prefix = d_pre + 2;
    L.col_prefix = d_pre + 4;
    L.slot_prefix = d_pre + 6;
    L.corr_prefix = d_pre + 8;
    L.bonded_prefix = d_pre + 10;
    L.gb_prefix = d_pre + 12;
    L.sc_prefix = d_pre + 16;
    L.gbp_prefix = d_pre + 18;
    L.n_atoms_total = h.n;
    L.n_cl_total = h.cl_cap;
    L.n_cols_total = h.ncol;
    L.n_slots_total = h.slot_cap;
    L.n_corr_total = h.n_corr;
    L.n_bonded_total = h.n_bonded;
    L.n_gb_tiles_total = h.n_gb_tiles;
    L.n_gb_pairs_total = h.n_gb_pairs;
    L.n_local_total = h.n_local;
    L.n_sc_total = h.sc_cap;
    L.max_mol = h.dev.n_mol;
    const size_t xb = 3 * (size_t)h.n * sizeof(double);
    HIPCHECK(hipMemcpy(h.dev.xaos, x, xb, hipMemcpyHostToDevice));
    pm::launch_import(L, pm::kXferX);
    force_eval(&c, L, pm::kStepForce, eval_policy(&c), pm::kPairForcesEnergies, false, false);
    pm::launch_gather_scalars(L, d_scal);
    long long scal[pm::kScalN];
    HIPCHECK(hipMemcpy(scal, d_scal, sizeof scal, hipMemcpyDeviceToHost));
    const unsigned err = (unsigned)scal[pm::kScalErr];
    if (err & pm::kErrListCap) {
      // a longer list, if it still fits: the whole evaluation again
      if (attempt > 4) throw Fail{PM_ECAPACITY, "pair list capacity kept overflowing"};
      h.list_cap *= 2;
      continue;
    }
    if (err & pm::kErrNonFinite) throw Fail{PM_EINVAL, "non-finite position"};
    if (err & pm::kErrColumnCap)
      throw Fail{PM_ECAPACITY, "more than " + std::to_string(pm::kSortCap) + " atoms in one cell column"};
    if (err & pm::kErrClusterCap) throw Fail{PM_ECAPACITY, "cluster capacity"};
    if (err & pm::kErrVacuumRegion)
      throw Fail{PM_ECAPACITY, "an atom left the region (origin +- box / 2) of the vacuum system's cell grid"};
    if (f) {
      pm::launch_export(L, pm::kXferF);
      std::vector<double> fo(3 * (size_t)h.n);
      HIPCHECK(hipMemcpy(fo.data(), h.dev.faos, xb, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < fo.size(); ++k) f[k] += fo[k];
    }
    if (e) add_energies(&c, h, scal, e);
    HIPCHECK(hipStreamSynchronize(nullptr));
    return;
  }
}

// randomNumber (MathUtilities.cpp:153-187) with both of its generators.
// randomtype 0: the POSIX drand48 sequence -- X <- (0x5DEECE66D X + 11) mod 2^48
// from X = seed << 16 | 0x330E (srand48), each draw X / 2^48 -- which the
// reference calls on Linux; randomtype 1: ProtoMol's Random, the Park-Miller
// minimal standard generator (Schrage's method) behind a Bays-Durham shuffle
// table of 32 (Random.cpp:33-68).  gaussian() is randomGaussianNumber
// (MathUtilities.cpp:206-236): polar Box-Muller, the second deviate of each
// pair returned by the next call.
class RefRandom {
 public:
  RefRandom(uint32_t seed, int randomtype) : type_(randomtype ? 1 : 0) {
    x48_ = ((uint64_t)seed << 16) | 0x330Eu;
    int32_t idum = (int32_t)seed;  // Random::seed keeps an unsigned seed in an int
    if (idum < 1) idum = 1;
    for (int j = kTab + 7; j >= 0; j--) {  // 8 warm-ups, then the table
      idum = step(idum);
      if (j < kTab) tab_[j] = idum;
    }
    last_ = tab_[0];
    idum_ = idum;
  }
  double uniform() {
    if (type_ == 0) {
      x48_ = (0x5DEECE66Dull * x48_ + 11u) & ((1ull << 48) - 1);
      return std::ldexp((double)x48_, -48);
    }
    idum_ = step(idum_);
    const int j = last_ / (1 + (kMod - 1) / kTab);
    last_ = tab_[j];
    tab_[j] = idum_;
    const double t = (1.0 / kMod) * last_, less_than_one = 1.0 - 3.0e-16;
    return t > less_than_one ? less_than_one : t;
  }
  double gaussian() {
    if (spare_ok_) {
      spare_ok_ = false;
      return spare_;
    }
    double r = 2.0, v1 = 0.0, v2 = 0.0;
    while (r >= 1.0 || r < 1.523e-8) {
      v1 = 2.0 * uniform() - 1.0;
      v2 = 2.0 * uniform() - 1.0;
      r = v1 * v1 + v2 * v2;
    }
    const double fac = std::sqrt(-2.0 * std::log(r) / r);
    spare_ = v1 * fac;
    spare_ok_ = true;
    return v2 * fac;
  }

 private:
  static constexpr int kTab = 32;
  static constexpr int32_t kMod = 2147483647;
  static int32_t step(int32_t x) {  // 16807 x mod (2^31 - 1) without overflow
    const int32_t k = x / 127773;
    x = 16807 * (x - k * 127773) - 2836 * k;
    return x < 0 ? x + kMod : x;
  }
  int type_;
  uint64_t x48_;
  int32_t tab_[kTab];
  int32_t last_ = 0, idum_ = 0;
  bool spare_ok_ = false;
  double spare_ = 0.0;
};

int run_guarded(pm_ctx *c, const std::function<void()> &fn) {
  try {
    fn();
    return PM_OK;
  } catch (const Fail &f) {
    set_err(c, f.msg);
    return f.code;
  } catch (const std::exception &e) {
    set_err(c, e.what());
    return PM_EINVAL;
  } catch (...) {
    set_err(c, "unknown error");
    return PM_EINVAL;
  }
}

}  // namespace

extern "C" {

int pm_abi_version(void) { return PM_ABI_VERSION; }

const char *pm_last_error(const pm_ctx *ctx) { return ctx ? ctx->err.c_str() : g_thread_err.c_str(); }

int pm_device_info(char *name, int32_t name_len, int32_t *n_cu, double *hbm_gib) {
  return run_guarded(nullptr, [&] {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHECK(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    if (n_cu) *n_cu = p.multiProcessorCount;
    if (hbm_gib) *hbm_gib = (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
  });
}

int pm_ctx_create(const pm_force_desc *force, int32_t n_systems, const pm_system_desc *systems, pm_ctx **out) {
  if (!out) return PM_EINVAL;
  *out = nullptr;
  pm_ctx *c = new pm_ctx();
  int rc = run_guarded(c, [&] {
    if (!force || n_systems < 0 || (n_systems > 0 && !systems)) throw Fail{PM_EINVAL, "need a force description"};
    c->desc = *force;
    HIPCHECK(hipGetDevice(&c->device));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&c->ev0));
    HIPCHECK(hipEventCreate(&c->ev1));
    HIPCHECK(hipEventCreate(&c->evp0));
    HIPCHECK(hipEventCreate(&c->evp1));
    HIPCHECK(hipEventCreate(&c->evr0));
    HIPCHECK(hipEventCreate(&c->evr1));
    setup_force(c, force, n_systems, systems);
    std::vector<int32_t> ids(std::max(n_systems, 1));
    add_many(c, n_systems, systems, n_systems, ids.data());
    for (int s = 0; s < n_systems; ++s)
      if (ids[s] != s) throw Fail{PM_EINVAL, "system " + std::to_string(s) + " was not placed"};
    HIPCHECK(hipDeviceSynchronize());
  });
  if (rc != PM_OK) {
    g_thread_err = c->err;
    pm_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return PM_OK;
}

int pm_ctx_destroy(pm_ctx *c) {
  if (!c) return PM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->sets.clear();
  for (auto &h : c->sys)
    if (h) {
      dev_free(c, h->lists, h->lists_bytes);
      dev_free(c, h->block, h->block_bytes);
    }
  c->sys.clear();
  c->arena.destroy();
  for (hipEvent_t e : {c->ev0, c->ev1, c->evp0, c->evp1, c->evr0, c->evr1})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->pin) (void)hipHostFree(c->pin);
  if (c->pin_done) (void)hipEventDestroy(c->pin_done);
  delete c;
  return PM_OK;
}

int pm_ctx_reserve(pm_ctx *c, uint64_t bytes) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!live_ids(c).empty() || c->arena.base) throw Fail{PM_EINVAL, "reserve before adding systems, once"};
    if (bytes == 0) throw Fail{PM_EINVAL, "reservation of 0 bytes"};
    c->arena.init(bytes);
  });
}

int pm_ctx_memory(pm_ctx *c, uint64_t *reserved, uint64_t *used) {
  if (!c) return PM_EINVAL;
  if (reserved) *reserved = c->arena.cap;
  if (used) *used = c->arena.used;
  return PM_OK;
}

int pm_system_footprint(pm_ctx *c, const pm_system_desc *system, uint64_t *bytes) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!system || !bytes) throw Fail{PM_EINVAL, "system and bytes required"};
    *bytes = footprint(c, *system);
  });
}

int pm_system_add(pm_ctx *c, const pm_system_desc *system, int32_t *id) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!system || !id) throw Fail{PM_EINVAL, "system and id required"};
    *id = add_system(c, *system);
  });
}

int pm_system_add_many(pm_ctx *c, int32_t n, const pm_system_desc *systems, int32_t max_full, int32_t *ids) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (n < 0 || (n > 0 && (!systems || !ids))) throw Fail{PM_EINVAL, "systems and ids required"};
    add_many(c, n, systems, std::max(max_full, 1), ids);
  });
}

int pm_system_remove(pm_ctx *c, int32_t id) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] { remove_system(c, id); });
}

int pm_evaluate(pm_ctx *c, int32_t system, const double *x, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!x) throw Fail{PM_EINVAL, "positions required"};
    sys_of(c, system);
    evaluate_ids(c, {system}, x, f, e, false, nullptr, nullptr);
  });
}

int pm_evaluate_batch(pm_ctx *c, const double *x, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!x) throw Fail{PM_EINVAL, "positions required"};
    const std::vector<int> ids = live_ids(c);
    if (ids.empty()) throw Fail{PM_EINVAL, "no systems"};
    evaluate_ids(c, ids, x, f, e, false, nullptr, nullptr);
  });
}

int pm_evaluate_systems(pm_ctx *c, int32_t n, const int32_t *ids, const double *x, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!x) throw Fail{PM_EINVAL, "positions required"};
    evaluate_ids(c, check_ids(c, n, ids), x, f, e, false, nullptr, nullptr);
  });
}

int pm_pair_census(pm_ctx *c, int32_t system, const double *x, uint64_t *count, uint64_t *hash) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!count || !hash) throw Fail{PM_EINVAL, "bad census arguments"};
    sys_of(c, system);
    if (c->gb) throw Fail{PM_EINVAL, "the all-pairs GB path evaluates every pair; no census"};
    if (x) {
      evaluate_ids(c, {system}, x, nullptr, nullptr, true, count, hash);
      return;
    }
    // the list the MD state holds, at its positions: no rebuild
    require_md(c, {system});
    LaunchSet &ls = get_set(c, {system});
    HIPCHECK(hipMemsetAsync(c->sys[system]->dev.census, 0, 2 * sizeof(unsigned long long), c->stream));
    pm::ForceDev P = c->P;
    P.finish_step = 0;  // not a force pass: the step counters stay as they are
    if (!pm::launch_pair(ls.L, P, pm::kPairCensus)) throw Fail{PM_EINVAL, "unsupported term combination"};
    read_scalars(c, ls);
    *count = (uint64_t)ls.h_scal[pm::kScalCensus];
    *hash = (uint64_t)ls.h_scal[pm::kScalCensus + 1];
  });
}

int pm_md_set_state(pm_ctx *c, const double *x, const double *v) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!x || !v) throw Fail{PM_EINVAL, "positions and velocities required"};
    const std::vector<int> ids = live_ids(c);
    if (ids.empty()) throw Fail{PM_EINVAL, "no systems"};
    set_state_ids(c, ids, x, v);
  });
}

int pm_md_set_state_systems(pm_ctx *c, int32_t n, const int32_t *ids, const double *x, const double *v) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!x || !v) throw Fail{PM_EINVAL, "positions and velocities required"};
    set_state_ids(c, check_ids(c, n, ids), x, v);
  });
}

int pm_md_run(pm_ctx *c, int32_t n_steps, double dt_fs, int32_t energies_every) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    const std::vector<int> ids = live_ids(c);
    if (ids.empty()) throw Fail{PM_EINVAL, "no systems"};
    require_md(c, ids);
    if (n_steps < 1) return;
    if (!(dt_fs > 0.0)) throw Fail{PM_EINVAL, "dt must be > 0"};
    md_begin(c, ids);
    md_advance(c, ids, n_steps, dt_fs, energies_every > 0, std::max(0, energies_every));
    md_end(c, ids, dt_fs);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int pm_md_begin(pm_ctx *c, int32_t n, const int32_t *ids) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] { md_begin(c, check_ids(c, n, ids)); });
}

int pm_md_advance(pm_ctx *c, int32_t n, const int32_t *ids, int32_t n_steps, double dt_fs, int32_t energies_last) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] { md_advance(c, check_ids(c, n, ids), n_steps, dt_fs, energies_last != 0); });
}

int pm_md_end(pm_ctx *c, int32_t n, const int32_t *ids, double dt_fs) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    md_end(c, check_ids(c, n, ids), dt_fs);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int pm_md_get_state(pm_ctx *c, double *x, double *v, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    const std::vector<int> ids = live_ids(c);
    if (ids.empty()) throw Fail{PM_EINVAL, "no systems"};
    get_state_ids(c, ids, x, v, f, e);
  });
}

int pm_md_get_state_systems(pm_ctx *c, int32_t n, const int32_t *ids, double *x, double *v, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] { get_state_ids(c, check_ids(c, n, ids), x, v, f, e); });
}

int pm_md_set_langevin(pm_ctx *c, double temperature, double gamma_per_ps, uint32_t seed) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!(temperature >= 0.0) || !(gamma_per_ps >= 0.0)) throw Fail{PM_EINVAL, "temperature and gamma must be >= 0"};
    c->lv_temperature = temperature;
    c->lv_gamma_ps = gamma_per_ps;
    c->lv_seed = seed;
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (auto &h : c->sys)
      if (h) h->dev.lv_seed = seed;
    drop_sets(c);
  });
}

int pm_system_set_seed(pm_ctx *c, int32_t id, uint32_t seed) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    SysHost &h = sys_of(c, id);
    HIPCHECK(hipStreamSynchronize(c->stream));
    h.dev.lv_seed = seed;
    drop_sets_with(c, id);
  });
}

int pm_random_velocity(int32_t n, const double *mass, double temperature, uint32_t seed, int32_t randomtype,
                       double *v) {
  return run_guarded(nullptr, [&] {
    if (n < 0 || (n > 0 && (!mass || !v))) throw Fail{PM_EINVAL, "masses and velocities required"};
    if (!(temperature >= 0.0)) throw Fail{PM_EINVAL, "Kelvin temperature cannot be negative"};
    if (randomtype != 0 && randomtype != 1) throw Fail{PM_EINVAL, "randomtype must be 0 or 1"};
    RefRandom rng(seed, randomtype);
    (void)rng.uniform();  // the number randomNumber(seed, randomtype) returns while seeding
    const double kbt = temperature * kBoltzmann;
    for (int32_t i = 0; i < n; ++i) {
      if (!(mass[i] > 0.0)) throw Fail{PM_EINVAL, "masses must be positive"};
      const double s = std::sqrt(kbt / mass[i]);
      for (int d = 0; d < 3; ++d) v[3 * (size_t)i + d] = s * rng.gaussian();
    }
  });
}

int pm_evaluate_static(const pm_force_desc *force, const pm_system_desc *system, const double *x, double *f,
                       double *e) {
  return run_guarded(nullptr, [&] {
    if (!force || !system || !x) throw Fail{PM_EINVAL, "force, system and positions required"};
    evaluate_static(*force, *system, x, f, e);
  });
}

int pm_stats(pm_ctx *c, double *stats, int32_t n) {
  if (!c || !stats) return PM_EINVAL;
  for (int k = 0; k < n && k < 10; ++k) stats[k] = c->stats[k];
  return PM_OK;
}

}  // extern "C"
