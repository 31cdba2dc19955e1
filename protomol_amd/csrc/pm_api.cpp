// pm_api.cpp -- host side of the C ABI: context setup, device memory,
// launch sequencing and result export.  Compiled by hipcc for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
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

struct Allocs {
  std::vector<void *> ptrs;
  template <class T>
  T *get(size_t n) {
    if (n == 0) n = 1;
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw Fail{PM_ENOMEM, "hipMalloc failed"};
    ptrs.push_back(p);
    if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) throw Fail{PM_EDEVICE, "hipMemset failed"};
    return static_cast<T *>(p);
  }
  template <class T>
  T *put(const T *h, size_t n) {
    T *d = get<T>(n);
    if (n && h && hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw Fail{PM_EDEVICE, "hipMemcpy failed"};
    return d;
  }
  void release() {
    for (void *p : ptrs) (void)hipFree(p);
    ptrs.clear();
  }
};

struct SysHost {
  int n = 0, ncol = 0, cl_cap = 0, slot_cap = 0, list_cap = 0, n_corr = 0, n_bonded = 0;
  int n_gb_tiles = 0;
  double coul_const = 0.0;  // Ewald point-self + charged-system energy (FP64, host)
  double vir_const = 0.0;   // charged-system virial diagonal
  SysDev dev{};
  Allocs main, lists;
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

// per-system prefix record on the device: [0, count] for single-system launches
struct Prefixes {
  std::vector<int> atom, cl, col, slot, corr, bonded, gb;
};

}  // namespace

struct pm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  pm_force_desc desc{};
  ForceDev P{};
  bool gb = false;
  double rc = 0.0, skin_md = 0.0, skin_eval = 0.0;
  std::vector<SysHost> sys;
  SysDev *d_sys = nullptr;
  int *d_pref = nullptr;     // 7 prefix arrays of nsys+1 over all systems
  int *d_single = nullptr;   // per system: 7 x [0, count]
  Launch L;
  std::vector<Launch> Ls;    // single-system launches
  double *d_xaos = nullptr, *d_vaos = nullptr, *d_faos = nullptr;
  int n_total = 0;
  std::string err;
  bool md_ready = false;
  double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evp0 = nullptr, evp1 = nullptr;
  std::vector<hipEvent_t> evs;  // PM_FLAG_TIMING: start/stop per pair launch of an MD call
  int evs_used = 0;
  // captured MD steps (kick + drift, rebuild decision, conditional rebuild,
  // forces) for a given time step; replayed between the first and last step
  hipGraphExec_t gexec = nullptr;
  double g_h = 0.0;
  Allocs common;
};

namespace {

void set_err(pm_ctx *c, const std::string &m) {
  if (c) c->err = m;
  g_thread_err = m;
}

void upload_sys(pm_ctx *c, int s) {
  HIPCHECK(hipMemcpy(c->d_sys + s, &c->sys[s].dev, sizeof(SysDev), hipMemcpyHostToDevice));
}

int estimate_list_cap(double rho, double rlist, int cl_cap) {
  // clusters whose bounding box lies within rlist of an i-cluster box, halved
  double b = std::cbrt(pm::kCluster / std::max(rho, 1e-6));
  double w = 2.0 * b, r = rlist;
  double vol = w * w * w + 6.0 * w * w * r + 3.0 * M_PI * w * r * r + 4.0 / 3.0 * M_PI * r * r * r;
  double half = 0.5 * vol * rho / pm::kCluster;
  int cap = (int)(1.4 * half) + 32;
  cap = (cap + 31) / 32 * 32;
  return std::min(cap, 27 * cl_cap / 2 + 32);
}

void alloc_lists(SysHost &h) {
  h.lists.release();
  h.dev.list_cap = h.list_cap;
  h.dev.list = h.lists.get<uint4>((size_t)h.cl_cap * h.list_cap);
  h.dev.list_mask = h.lists.get<unsigned long long>(2 * (size_t)h.cl_cap * h.list_cap);
}

void setup_system(pm_ctx *c, SysHost &h, const pm_system_desc &d) {
  const int n = d.n_atoms;
  if (n < 1) throw Fail{PM_EINVAL, "n_atoms must be >= 1"};
  if (!d.scaled_charge || !d.mass || !d.molecule) throw Fail{PM_EINVAL, "charge, mass and molecule are required"};
  h.n = n;
  SysDev &S = h.dev;
  S.n_atoms = n;
  S.periodic = d.periodic ? 1 : 0;
  S.lj_mode = d.lj_mode;
  S.n_types = d.n_types;
  S.csf = (float)d.coulomb_scaling;
  S.lj_scaling = (float)d.lj_scaling;
  Allocs &A = h.main;

  if (S.periodic) {
    // PeriodicBoundaryConditions::set (PeriodicBoundaryConditions.cpp:32-77), orthogonal
    const double Lx = d.box[0], Ly = d.box[1], Lz = d.box[2];
    if (!(Lx > 0 && Ly > 0 && Lz > 0)) throw Fail{PM_EINVAL, "box edges must be positive"};
    double a1 = Ly * Lz - 0.0 * 0.0, d1 = Lx * a1 + 0.0 * 0.0 + 0.0 * 0.0;
    double a2 = Lz * Lx - 0.0 * 0.0, d2 = 0.0 * 0.0 + Ly * a2 + 0.0 * 0.0;
    double a3 = Lx * Ly - 0.0 * 0.0, d3 = 0.0 * 0.0 + 0.0 * 0.0 + Lz * a3;
    S.L[0] = Lx; S.L[1] = Ly; S.L[2] = Lz;
    S.Lr[0] = a1 / d1; S.Lr[1] = a2 / d2; S.Lr[2] = a3 / d3;
    for (int k = 0; k < 3; ++k) {
      double lo = d.origin[k] - d.box[k] * 0.5, hi = d.origin[k] + d.box[k] * 0.5;
      double mn = std::min(lo, hi), mx = std::max(lo, hi);
      S.H[k] = mx - mn;
      S.H2[k] = S.H[k] * 0.5;
      S.omin[k] = mn;
    }
    double dx = (Lx * 0.5) * (Lx * 0.5), dy = (Ly * 0.5) * (Ly * 0.5), dz = (Lz * 0.5) * (Lz * 0.5);
    S.D = std::min(dx, std::min(dy, dz));
  }

  // exclusion table: last add() of a pair wins (ExclusionTable.cpp:97-120)
  std::map<std::pair<int, int>, int> ex;
  std::vector<int> corr;
  {
    std::vector<std::pair<std::pair<int, int>, int>> seen;
    for (int k = 0; k < d.n_exclusions; ++k) {
      int a = d.exclusions[3 * k], b = d.exclusions[3 * k + 1], cls = d.exclusions[3 * k + 2];
      if (a < 0 || b < 0 || a >= n || b >= n) throw Fail{PM_EINVAL, "exclusion atom out of range"};
      if (cls != PM_EXCL_FULL && cls != PM_EXCL_MODIFIED) throw Fail{PM_EINVAL, "bad exclusion class"};
      if (a > b) std::swap(a, b);
      if (a == b) continue;
      ex[{a, b}] = cls;
      // the table of pairs the Ewald correction walks drops exact duplicates
      bool dup = false;
      for (auto &p : seen)
        if (p.first.first == a && p.first.second == b && p.second == cls) { dup = true; break; }
      if (!dup) {
        seen.push_back({{a, b}, cls});
        corr.push_back(a);
        corr.push_back(b);
        corr.push_back(cls);
      }
    }
  }
  std::vector<int> cnt(n + 1, 0);
  for (auto &kv : ex) { cnt[kv.first.first + 1]++; cnt[kv.first.second + 1]++; }
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> partner(cnt[n]), pcls(cnt[n]), fill(cnt.begin(), cnt.end() - 1);
  for (auto &kv : ex) {
    int a = kv.first.first, b = kv.first.second;
    partner[fill[a]] = b; pcls[fill[a]++] = kv.second;
    partner[fill[b]] = a; pcls[fill[b]++] = kv.second;
  }
  S.excl_start = A.put(cnt.data(), cnt.size());
  S.excl_partner = A.put(partner.data(), partner.size());
  S.excl_class = A.put(pcls.data(), pcls.size());
  const bool ewald = c->desc.term_b == PM_TERM_EWALD_REAL;
  h.n_corr = (ewald && c->desc.ewald_correction) ? (int)corr.size() / 3 : 0;
  S.n_corr = h.n_corr;
  S.corr = A.put(corr.data(), corr.size());

  // per-atom data
  std::vector<double> invm(n);
  for (int i = 0; i < n; ++i) {
    if (!(d.mass[i] > 0)) throw Fail{PM_EINVAL, "masses must be positive"};
    invm[i] = 1.0 / d.mass[i];
  }
  S.charge = A.put(d.scaled_charge, n);
  S.inv_mass = A.put(invm.data(), n);
  S.molecule = A.put(d.molecule, n);
  if (d.lj_mode == PM_LJ_PERATOM) {
    if (!d.lj_sigma || !d.lj_epsilon) throw Fail{PM_EINVAL, "per-atom LJ needs sigma and epsilon"};
    std::vector<float> sh(n), es(n);
    for (int i = 0; i < n; ++i) {
      sh[i] = (float)(0.5 * d.lj_sigma[i]);
      es[i] = (float)std::sqrt(d.lj_epsilon[i]);
    }
    S.sig_half = A.put(sh.data(), n);
    S.eps_sqrt = A.put(es.data(), n);
    S.type = A.get<int>(n);
  } else {
    if (d.n_types < 1 || !d.lj_table || !d.type) throw Fail{PM_EINVAL, "LJ table mode needs n_types, type and lj_table"};
    if (d.n_types > 16 && !c->gb)
      throw Fail{PM_EINVAL, "the cell-list pair kernel keeps the LJ table in LDS: at most 16 types (use PM_LJ_PERATOM)"};
    for (int i = 0; i < n; ++i)
      if (d.type[i] < 0 || d.type[i] >= d.n_types) throw Fail{PM_EINVAL, "atom type out of range"};
    std::vector<float4> t((size_t)d.n_types * d.n_types);
    for (size_t k = 0; k < t.size(); ++k)
      t[k] = make_float4((float)d.lj_table[4 * k], (float)d.lj_table[4 * k + 1], (float)d.lj_table[4 * k + 2],
                         (float)d.lj_table[4 * k + 3]);
    S.lj_table = A.put(t.data(), t.size());
    S.type = A.put(d.type, n);
  }
  S.x = A.get<double>(3 * (size_t)n);
  S.v = A.get<double>(3 * (size_t)n);
  S.xref = A.get<double>(3 * (size_t)n);
  S.slot = A.get<int>(n);
  S.lbuild = A.get<float4>(n);
  S.bin_col = A.get<int>(n);
  S.bin_z = A.get<float>(n);
  S.colbuf = A.get<unsigned long long>(n);

  // bonded terms
  S.n_bonds = (c->desc.flags & PM_FLAG_BONDED) ? d.n_bonds : 0;
  S.n_angles = (c->desc.flags & PM_FLAG_BONDED) ? d.n_angles : 0;
  for (int k = 0; k < 2 * S.n_bonds; ++k)
    if (d.bonds[k] < 0 || d.bonds[k] >= n) throw Fail{PM_EINVAL, "bond atom out of range"};
  for (int k = 0; k < 3 * S.n_angles; ++k)
    if (d.angles[k] < 0 || d.angles[k] >= n) throw Fail{PM_EINVAL, "angle atom out of range"};
  S.bonds = A.put(d.bonds, 2 * (size_t)S.n_bonds);
  S.bond_k = A.put(d.bond_k, S.n_bonds);
  S.bond_r0 = A.put(d.bond_r0, S.n_bonds);
  S.angles = A.put(d.angles, 3 * (size_t)S.n_angles);
  S.angle_k = A.put(d.angle_k, S.n_angles);
  S.angle_t0 = A.put(d.angle_theta0, S.n_angles);
  h.n_bonded = S.n_bonds + S.n_angles;

  // per-system scalars
  S.n_clusters = A.get<int>(1);
  S.acc = A.get<long long>(pm::kAccN);
  S.census = A.get<unsigned long long>(2);
  S.err = A.get<unsigned>(1);
  S.maxdisp2 = A.get<unsigned>(1);
  S.maxdisp_run = A.get<unsigned>(1);
  S.ctl = A.get<int>(pm::kCtlN);

  if (c->gb) {
    if (!d.gb_radius || !d.gb_scaling) throw Fail{PM_EINVAL, "GB needs gb_radius and gb_scaling"};
    pm::gb_setup_system(S, A.put(d.gb_radius, n), A.put(d.gb_scaling, n), A.get<double>(n), A.get<double>(n),
                        A.get<double>(n), A.get<double>(n));
    h.n_gb_tiles = pm::gb_tile_count(n);
    // the cluster machinery is not used in vacuum; keep the arrays minimal
    S.ncx = S.ncy = S.ncol = 1;
    h.ncol = 1;
    h.cl_cap = 0;
    h.slot_cap = n;  // forces are kept per atom: slot = atom
    S.cl_cap = 0;
    S.slot_cap = n;
    S.xq = A.get<float4>(n);
    S.fx = A.get<long long>(n);
    S.fy = A.get<long long>(n);
    S.fz = A.get<long long>(n);
    std::vector<int> ident(n);
    for (int i = 0; i < n; ++i) ident[i] = i;
    HIPCHECK(hipMemcpy(S.slot, ident.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return;
  }

  if (!S.periodic) throw Fail{PM_EINVAL, "the cell-list pair path needs periodic boundary conditions"};
  // Ewald point-self and charged-system constants (NonbondedFullEwaldSystemForce.h:694-714)
  if (ewald && c->desc.ewald_correction) {
    double q2 = 0.0, q = 0.0;
    for (int i = 0; i < n; ++i) q2 += d.scaled_charge[i] * d.scaled_charge[i];
    for (int i = 0; i < n; ++i) q += d.scaled_charge[i];
    double alpha = c->desc.ewald_alpha;
    double point = -q2 * alpha / std::sqrt(M_PI);
    double V = std::fabs(d.box[0] * d.box[1] * d.box[2]);
    double charged = 0.0;
    if (std::fabs(q * 0.00268283) > 1.0e-5) charged = -M_PI / (2.0 * V * (alpha * alpha)) * q * q;
    h.coul_const = point + charged;
    h.vir_const = charged;
  }

  // x-y column grid: columns about (8/rho)^(1/3) wide so 8-atom clusters are compact
  const double V = d.box[0] * d.box[1] * d.box[2];
  const double rho = n / V;
  const double a = std::cbrt(pm::kCluster / rho);
  double rlist = c->rc + std::max(c->skin_md, c->skin_eval);
  for (int k = 0; k < 3; ++k)
    if (d.box[k] < 2.0 * rlist)
      throw Fail{PM_EINVAL, "box edge " + std::to_string(d.box[k]) + " is shorter than twice the pair-list radius " +
                                std::to_string(rlist)};
  S.ncx = std::max(1, (int)std::floor(d.box[0] / a));
  S.ncy = std::max(1, (int)std::floor(d.box[1] / a));
  S.ax = d.box[0] / S.ncx;
  S.ay = d.box[1] / S.ncy;
  S.ncol = S.ncx * S.ncy;
  S.rx = (int)std::ceil(rlist / S.ax);
  S.ry = (int)std::ceil(rlist / S.ay);
  h.ncol = S.ncol;
  h.cl_cap = (n + pm::kCluster - 1) / pm::kCluster + S.ncol;
  h.slot_cap = h.cl_cap * pm::kCluster;
  h.list_cap = estimate_list_cap(rho, rlist, h.cl_cap);
  S.cl_cap = h.cl_cap;
  S.slot_cap = h.slot_cap;
  S.col_count = A.get<int>(S.ncol);
  S.col_cursor = A.get<int>(S.ncol);
  S.col_atom_start = A.get<int>(S.ncol);
  S.col_cl_start = A.get<int>(S.ncol);
  S.col_ncl = A.get<int>(S.ncol);
  S.xq = A.get<float4>(h.slot_cap);
  S.ljt = A.get<int>(h.slot_cap);
  S.ljp = A.get<float2>(h.slot_cap);
  S.orig = A.get<int>(h.slot_cap);
  S.smol = A.get<int>(h.slot_cap);
  S.fx = A.get<long long>(h.slot_cap);
  S.fy = A.get<long long>(h.slot_cap);
  S.fz = A.get<long long>(h.slot_cap);
  S.cl_origin = A.get<double>(3 * (size_t)h.cl_cap);
  S.cl_lo = A.get<float4>(h.cl_cap);
  S.cl_hi = A.get<float4>(h.cl_cap);
  S.cl_col = A.get<int>(h.cl_cap);
  S.cl_nvalid = A.get<int>(h.cl_cap);
  S.list_len = A.get<int>(h.cl_cap);
  alloc_lists(h);
}

void build_launches(pm_ctx *c) {
  const int ns = (int)c->sys.size();
  Prefixes p;
  auto push = [](std::vector<int> &v, int x) { v.push_back(v.empty() ? x : v.back() + x); };
  p.atom = p.cl = p.col = p.slot = p.corr = p.bonded = p.gb = {0};
  for (auto &h : c->sys) {
    push(p.atom, h.n);
    push(p.cl, h.cl_cap);
    push(p.col, h.ncol);
    push(p.slot, h.slot_cap);
    push(p.corr, h.n_corr);
    push(p.bonded, h.n_bonded);
    push(p.gb, h.n_gb_tiles);
  }
  std::vector<int> all;
  for (auto *v : {&p.atom, &p.cl, &p.col, &p.slot, &p.corr, &p.bonded, &p.gb}) all.insert(all.end(), v->begin(), v->end());
  if (!c->d_pref) c->d_pref = c->common.get<int>(all.size());
  HIPCHECK(hipMemcpy(c->d_pref, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> single;
  for (auto &h : c->sys)
    for (int v : {h.n, h.cl_cap, h.ncol, h.slot_cap, h.n_corr, h.n_bonded, h.n_gb_tiles}) {
      single.push_back(0);
      single.push_back(v);
    }
  if (!c->d_single) c->d_single = c->common.get<int>(single.size());
  HIPCHECK(hipMemcpy(c->d_single, single.data(), single.size() * sizeof(int), hipMemcpyHostToDevice));
  Launch &L = c->L;
  const int stride = ns + 1;
  L.stream = c->stream;
  L.sys = c->d_sys;
  L.nsys = ns;
  L.atom_prefix = c->d_pref;
  L.cl_prefix = c->d_pref + stride;
  L.col_prefix = c->d_pref + 2 * stride;
  L.slot_prefix = c->d_pref + 3 * stride;
  L.corr_prefix = c->d_pref + 4 * stride;
  L.bonded_prefix = c->d_pref + 5 * stride;
  L.gb_prefix = c->d_pref + 6 * stride;
  L.n_atoms_total = p.atom.back();
  L.n_cl_total = p.cl.back();
  L.n_cols_total = p.col.back();
  L.n_slots_total = p.slot.back();
  L.n_corr_total = p.corr.back();
  L.n_bonded_total = p.bonded.back();
  L.n_gb_tiles_total = p.gb.back();
  c->Ls.resize(ns);
  for (int s = 0; s < ns; ++s) {
    Launch &l = c->Ls[s];
    const int *b = c->d_single + 14 * s;
    const SysHost &h = c->sys[s];
    l.stream = c->stream;
    l.sys = c->d_sys + s;
    l.nsys = 1;
    l.atom_prefix = b;
    l.cl_prefix = b + 2;
    l.col_prefix = b + 4;
    l.slot_prefix = b + 6;
    l.corr_prefix = b + 8;
    l.bonded_prefix = b + 10;
    l.gb_prefix = b + 12;
    l.n_atoms_total = h.n;
    l.n_cl_total = h.cl_cap;
    l.n_cols_total = h.ncol;
    l.n_slots_total = h.slot_cap;
    l.n_corr_total = h.n_corr;
    l.n_bonded_total = h.n_bonded;
    l.n_gb_tiles_total = h.n_gb_tiles;
  }
}

// one force evaluation of the systems in `L` at the positions already in S.x
hipEvent_t next_event(pm_ctx *c) {
  if (c->evs_used == (int)c->evs.size()) {
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    c->evs.push_back(e);
  }
  return c->evs[c->evs_used++];
}

// Rebuild policy of an MD run: the list (radius rc + skin) is rebuilt when
// an atom has moved more than skin/2 from where it was built, or after
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

// one force evaluation of the systems in `L` at the positions already in S.x.
// mode kStepForce rebuilds the list and zeroes the forces first; kStepAuto
// (MD steps, whose kick has consumed and zeroed the forces) leaves the
// rebuild decision to the device.
void force_eval(pm_ctx *c, const Launch &L, int mode, const StepPolicy &pol, bool energies, bool census, bool timed) {
  pm::launch_step_begin(L, mode, pol.lim2, pol.kmax);
  const bool pool = !timed && (c->desc.flags & PM_FLAG_TIMING);
  if (mode & pm::kStepForce) pm::launch_zero_forces(L);
  if (c->gb) {
    if (timed) HIPCHECK(hipEventRecord(c->evp0, c->stream));
    if (pool) HIPCHECK(hipEventRecord(next_event(c), c->stream));
    pm::launch_gb(L, c->P, energies);
    if (pool) HIPCHECK(hipEventRecord(next_event(c), c->stream));
    if (timed) HIPCHECK(hipEventRecord(c->evp1, c->stream));
    pm::launch_bonded(L, energies);
    return;
  }
  pm::launch_rebuild(L, pol.rlist * pol.rlist);
  if (timed) HIPCHECK(hipEventRecord(c->evp0, c->stream));
  if (pool) HIPCHECK(hipEventRecord(next_event(c), c->stream));
  if (!pm::launch_pair(L, c->P, energies, census)) throw Fail{PM_EINVAL, "unsupported term combination"};
  if (pool) HIPCHECK(hipEventRecord(next_event(c), c->stream));
  if (timed) HIPCHECK(hipEventRecord(c->evp1, c->stream));
  pm::launch_ewald_corr(L, c->P, energies);
  pm::launch_bonded(L, energies);
}

struct Readback {
  std::vector<long long> acc;
  std::vector<unsigned long long> census;
  std::vector<unsigned> err;
};

Readback read_scalars(pm_ctx *c, int s0, int s1) {
  Readback r;
  r.acc.resize((size_t)(s1 - s0) * pm::kAccN);
  r.census.resize(2 * (size_t)(s1 - s0));
  r.err.resize(s1 - s0);
  for (int s = s0; s < s1; ++s) {
    const SysDev &S = c->sys[s].dev;
    HIPCHECK(hipMemcpyAsync(&r.acc[(size_t)(s - s0) * pm::kAccN], S.acc, pm::kAccN * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(&r.census[2 * (size_t)(s - s0)], S.census, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(&r.err[s - s0], S.err, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return r;
}

void add_energies(pm_ctx *c, int s, const long long *acc, double *e) {
  const SysHost &h = c->sys[s];
  const double k = pm::kEScaleInv;
  e[PM_E_COULOMB] += (double)acc[pm::kAccCoul] * k + h.coul_const;
  e[PM_E_LENNARDJONES] += (double)acc[pm::kAccLJ] * k;
  e[PM_E_BOND] += (double)acc[pm::kAccBond] * k;
  e[PM_E_ANGLE] += (double)acc[pm::kAccAngle] * k;
  if (c->desc.flags & PM_FLAG_VIRIAL) {
    const double xx = (double)acc[pm::kAccVxx] * k + h.vir_const, xy = (double)acc[pm::kAccVxy] * k,
                 xz = (double)acc[pm::kAccVxz] * k, yy = (double)acc[pm::kAccVyy] * k + h.vir_const,
                 yz = (double)acc[pm::kAccVyz] * k, zz = (double)acc[pm::kAccVzz] * k + h.vir_const;
    double *v = e + PM_E_VIRIALXX;
    v[0] += xx; v[1] += xy; v[2] += xz;
    v[3] += xy; v[4] += yy; v[5] += yz;
    v[6] += xz; v[7] += yz; v[8] += zz;
  }
}

// raise on device status bits; grows a full pair list and reports it
bool check_errors(pm_ctx *c, int s0, const std::vector<unsigned> &err) {
  bool regrow = false;
  for (size_t k = 0; k < err.size(); ++k) {
    const unsigned e = err[k];
    if (!e) continue;
    const int s = s0 + (int)k;
    HIPCHECK(hipMemset(c->sys[s].dev.err, 0, sizeof(unsigned)));
    if (e & pm::kErrNonFinite) throw Fail{PM_EINVAL, "system " + std::to_string(s) + ": non-finite position"};
    if (e & pm::kErrColumnCap)
      throw Fail{PM_ECAPACITY, "system " + std::to_string(s) + ": more than " + std::to_string(pm::kSortCap) +
                                   " atoms in one cell column"};
    if (e & pm::kErrClusterCap) throw Fail{PM_ECAPACITY, "system " + std::to_string(s) + ": cluster capacity"};
    if (e & pm::kErrListCap) {
      SysHost &h = c->sys[s];
      h.list_cap *= 2;
      alloc_lists(h);
      upload_sys(c, s);
      regrow = true;
    }
  }
  return regrow;
}

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

// evaluate systems [s0, s1) through launch L at host positions x; add into f, e
void evaluate_range(pm_ctx *c, const Launch &L, int s0, int s1, const double *x, double *f, double *e, bool census,
                    uint64_t *count, uint64_t *hash) {
  // an evaluation overwrites the device positions: any MD state is gone
  c->md_ready = false;
  int off = 0;
  for (int s = 0; s < s0; ++s) off += c->sys[s].n;
  int n = 0;
  for (int s = s0; s < s1; ++s) n += c->sys[s].n;
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  HIPCHECK(hipMemcpyAsync(c->d_xaos + 3 * (size_t)off, x, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // the kernels index the staging buffer from the launch's first atom
  pm::launch_import_x(L, c->d_xaos + 3 * (size_t)off);
  for (int attempt = 0;; ++attempt) {
    StepPolicy pol;
    pol.rlist = (float)(c->rc + c->skin_eval);
    force_eval(c, L, pm::kStepForce, pol, true, census, true);
    HIPCHECK(hipEventRecord(c->ev1, c->stream));
    Readback r = read_scalars(c, s0, s1);
    if (check_errors(c, s0, r.err)) {
      if (attempt > 4) throw Fail{PM_ECAPACITY, "pair list capacity kept overflowing"};
      continue;
    }
    pm::launch_export_f(L, c->d_faos + 3 * (size_t)off);
    std::vector<double> fo(3 * (size_t)n);
    HIPCHECK(hipMemcpyAsync(fo.data(), c->d_faos + 3 * (size_t)off, fo.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (f)
      for (size_t k = 0; k < fo.size(); ++k) f[k] += fo[k];
    if (e)
      for (int s = s0; s < s1; ++s) add_energies(c, s, &r.acc[(size_t)(s - s0) * pm::kAccN], e + (size_t)(s - s0) * PM_E_NSLOTS);
    if (census) {
      *count = r.census[0];
      *hash = r.census[1];
    }
    float ms = 0.f, pms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    HIPCHECK(hipEventElapsedTime(&pms, c->evp0, c->evp1));
    c->stats[3] = ms;
    c->stats[4] = pms;
    c->stats[5] = 1;
    c->stats[0] = 1 + attempt;
    return;
  }
}

constexpr int kGraphSteps = 10;

// The graph of kGraphSteps MD steps (kick + drift, then forces with the
// device's own rebuild decision) for time step h, captured on first use.
// Kernel arguments are the context's fixed device tables, so a graph stays
// valid when pair lists are regrown (only the tables' contents change).
template <class Step>
hipGraphExec_t md_graph(pm_ctx *c, double h, const StepPolicy &, Step &step) {
  if (c->gexec && c->g_h == h) return c->gexec;
  if (c->gexec) {
    HIPCHECK(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < kGraphSteps; ++k) step(1, false);
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHECK(hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    throw Fail{PM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
  }
  c->g_h = h;
  return c->gexec;
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
    if (name && name_len > 0) {
      std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (n_cu) *n_cu = p.multiProcessorCount;
    if (hbm_gib) *hbm_gib = (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
  });
}

int pm_ctx_create(const pm_force_desc *force, int32_t n_systems, const pm_system_desc *systems, pm_ctx **out) {
  if (!out) return PM_EINVAL;
  *out = nullptr;
  pm_ctx *c = new pm_ctx();
  int rc = run_guarded(c, [&] {
    if (!force || n_systems < 1 || !systems) throw Fail{PM_EINVAL, "need a force description and >= 1 system"};
    c->desc = *force;
    HIPCHECK(hipGetDevice(&c->device));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&c->ev0));
    HIPCHECK(hipEventCreate(&c->ev1));
    HIPCHECK(hipEventCreate(&c->evp0));
    HIPCHECK(hipEventCreate(&c->evp1));
    ForceDev &P = c->P;
    std::memset(&P, 0, sizeof P);
    P.flags = force->flags;
    c->gb = (force->flags & PM_FLAG_GB) != 0;
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
        P.two_alpha_pi = (float)(2.0 * force->ewald_alpha / std::sqrt(M_PI));
        P.dalpha = force->ewald_alpha;
        P.dtwo_alpha_pi = 2.0 * force->ewald_alpha / std::sqrt(M_PI);
      } else if (tb == PM_TERM_COULOMB) {
        cB2 = build_switch(force->switch_b, P.sw_b, false, 0);
        // OneAtomPairTwo tests the larger of the two switch cutoffs (OneAtomPairTwo.h:25-27)
        cA2 = cB2 = std::max(cA2, cB2);
      }
      P.cutA2d = cA2;
      P.cutB2d = cB2;
      P.cutA2 = (float)cA2;
      P.cutB2 = (float)cB2;
      const double cmax = std::max(cA2, cB2);
      P.guard = (float)(2e-3 + 1e-5 * cmax);
      c->rc = std::sqrt(cmax);
      c->skin_md = force->skin > 0 ? force->skin : 2.0;
      c->skin_eval = 0.1;
    }
    if (!c->gb) {
      // the pair list radius rc + skin must stay below half the smallest box
      // edge (one image per pair); a larger requested skin is reduced to fit
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
    c->sys.resize(n_systems);
    for (int s = 0; s < n_systems; ++s) setup_system(c, c->sys[s], systems[s]);
    c->d_sys = c->common.get<SysDev>(n_systems);
    for (int s = 0; s < n_systems; ++s) upload_sys(c, s);
    for (auto &h : c->sys) c->n_total += h.n;
    c->d_xaos = c->common.get<double>(3 * (size_t)c->n_total);
    c->d_vaos = c->common.get<double>(3 * (size_t)c->n_total);
    c->d_faos = c->common.get<double>(3 * (size_t)c->n_total);
    build_launches(c);
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
  for (auto &h : c->sys) {
    h.lists.release();
    h.main.release();
  }
  c->common.release();
  for (hipEvent_t e : {c->ev0, c->ev1, c->evp0, c->evp1})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PM_OK;
}

int pm_evaluate(pm_ctx *c, int32_t system, const double *x, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (system < 0 || system >= (int)c->sys.size() || !x) throw Fail{PM_EINVAL, "bad system index or positions"};
    evaluate_range(c, c->Ls[system], system, system + 1, x, f, e, false, nullptr, nullptr);
  });
}

int pm_evaluate_batch(pm_ctx *c, const double *x, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!x) throw Fail{PM_EINVAL, "positions required"};
    evaluate_range(c, c->L, 0, (int)c->sys.size(), x, f, e, false, nullptr, nullptr);
  });
}

int pm_pair_census(pm_ctx *c, int32_t system, const double *x, uint64_t *count, uint64_t *hash) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (system < 0 || system >= (int)c->sys.size() || !x || !count || !hash)
      throw Fail{PM_EINVAL, "bad census arguments"};
    if (c->gb) throw Fail{PM_EINVAL, "the all-pairs GB path evaluates every pair; no census"};
    evaluate_range(c, c->Ls[system], system, system + 1, x, nullptr, nullptr, true, count, hash);
  });
}

int pm_md_set_state(pm_ctx *c, const double *x, const double *v) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!x || !v) throw Fail{PM_EINVAL, "positions and velocities required"};
    const size_t bytes = 3 * (size_t)c->n_total * sizeof(double);
    HIPCHECK(hipMemcpyAsync(c->d_xaos, x, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_vaos, v, bytes, hipMemcpyHostToDevice, c->stream));
    pm::launch_import_x(c->L, c->d_xaos);
    pm::launch_import_v(c->L, c->d_vaos);
    // StandardIntegrator::initializeForces: F(x0) before the first step
    for (int attempt = 0;; ++attempt) {
      force_eval(c, c->L, pm::kStepForce, md_policy(c), false, false, false);
      Readback r = read_scalars(c, 0, (int)c->sys.size());
      if (check_errors(c, 0, r.err)) {
        if (attempt > 4) throw Fail{PM_ECAPACITY, "pair list capacity kept overflowing"};
        continue;
      }
      break;
    }
    c->md_ready = true;
  });
}

int pm_md_run(pm_ctx *c, int32_t n_steps, double dt_fs, int32_t energies_every) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!c->md_ready) throw Fail{PM_ENOTREADY, "pm_md_set_state first"};
    if (n_steps < 1) return;
    if (!(dt_fs > 0.0)) throw Fail{PM_EINVAL, "dt must be > 0"};
    // LeapfrogIntegrator::run (LeapfrogIntegrator.cpp:89-104): half kick + drift,
    // forces, then (kick + drift, forces) per step, closing half kick
    const double h = dt_fs * kInvTimeFactor;
    const StepPolicy pol = md_policy(c);
    const bool timing = (c->desc.flags & PM_FLAG_TIMING) != 0;
    const bool en = energies_every > 0;
    const int ns = (int)c->sys.size();
    // snapshot of x, v: a pair-list overflow (capacity grown) re-runs the call
    pm::launch_export_xv(c->L, c->d_xaos, c->d_vaos);
    for (int attempt = 0;; ++attempt) {
      HIPCHECK(hipEventRecord(c->ev0, c->stream));
      c->evs_used = 0;
      auto step = [&](int i, bool energies) {
        pm::launch_kick_drift(c->L, h, i == 0 ? 0 : 1);
        force_eval(c, c->L, pm::kStepAuto | (i == 0 ? pm::kStepResetRun : 0), pol, energies, false, false);
      };
      int i = 0;
      step(i++, en && n_steps == 1);
      if (!timing) {
        // the steps between the first and the last replay a captured graph
        while (n_steps - 1 - i >= kGraphSteps) {
          HIPCHECK(hipGraphLaunch(md_graph(c, h, pol, step), c->stream));
          i += kGraphSteps;
        }
      }
      for (; i < n_steps; ++i) step(i, en && i == n_steps - 1);
      pm::launch_half_kick(c->L, 0.5 * dt_fs * kInvTimeFactor);
      HIPCHECK(hipEventRecord(c->ev1, c->stream));
      std::vector<unsigned> md(ns);
      std::vector<int> ctl(pm::kCtlN * (size_t)ns);
      for (int s = 0; s < ns; ++s) {
        HIPCHECK(hipMemcpyAsync(&md[s], c->sys[s].dev.maxdisp_run, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipMemcpyAsync(&ctl[pm::kCtlN * (size_t)s], c->sys[s].dev.ctl, pm::kCtlN * sizeof(int),
                                hipMemcpyDeviceToHost, c->stream));
      }
      Readback r = read_scalars(c, 0, ns);
      if (check_errors(c, 0, r.err)) {
        if (attempt > 4) throw Fail{PM_ECAPACITY, "pair list capacity kept overflowing"};
        // lists regrown: restore the state of the call's start and run it again
        pm::launch_import_x(c->L, c->d_xaos);
        pm::launch_import_v(c->L, c->d_vaos);
        force_eval(c, c->L, pm::kStepForce, pol, false, false, false);
        continue;
      }
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      double pair_ms = 0.0;
      for (int k = 0; k + 1 < c->evs_used; k += 2) {
        float t = 0.f;
        HIPCHECK(hipEventElapsedTime(&t, c->evs[k], c->evs[k + 1]));
        pair_ms += t;
      }
      unsigned worst = 0;
      double rebuilds = 0;
      for (int s = 0; s < ns; ++s) {
        worst = std::max(worst, md[s]);
        rebuilds += ctl[pm::kCtlN * (size_t)s + pm::kCtlCount];
      }
      float w;
      std::memcpy(&w, &worst, sizeof w);
      c->stats[0] = rebuilds;
      c->stats[1] = attempt;
      c->stats[3] = ms;
      c->stats[4] = pair_ms;
      c->stats[5] = c->evs_used / 2;
      c->stats[6] = std::sqrt((double)w);
      c->stats[7] = c->skin_md;
      break;
    }
  });
}

int pm_md_get_state(pm_ctx *c, double *x, double *v, double *f, double *e) {
  if (!c) return PM_EINVAL;
  return run_guarded(c, [&] {
    if (!c->md_ready) throw Fail{PM_ENOTREADY, "pm_md_set_state first"};
    const size_t bytes = 3 * (size_t)c->n_total * sizeof(double);
    pm::launch_export_xv(c->L, c->d_xaos, c->d_vaos);
    pm::launch_export_f(c->L, c->d_faos);
    if (x) HIPCHECK(hipMemcpyAsync(x, c->d_xaos, bytes, hipMemcpyDeviceToHost, c->stream));
    if (v) HIPCHECK(hipMemcpyAsync(v, c->d_vaos, bytes, hipMemcpyDeviceToHost, c->stream));
    if (f) HIPCHECK(hipMemcpyAsync(f, c->d_faos, bytes, hipMemcpyDeviceToHost, c->stream));
    Readback r = read_scalars(c, 0, (int)c->sys.size());
    if (e) {
      std::memset(e, 0, sizeof(double) * PM_E_NSLOTS * c->sys.size());
      for (size_t s = 0; s < c->sys.size(); ++s) add_energies(c, (int)s, &r.acc[s * pm::kAccN], e + s * PM_E_NSLOTS);
    }
  });
}

int pm_stats(pm_ctx *c, double *stats, int32_t n) {
  if (!c || !stats) return PM_EINVAL;
  for (int k = 0; k < n && k < 8; ++k) stats[k] = c->stats[k];
  return PM_OK;
}

}  // extern "C"
