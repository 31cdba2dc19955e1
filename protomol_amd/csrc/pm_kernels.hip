// pm_kernels.hip -- gfx950 kernels of the nonbonded force path.
//
//   rebuild:  k_bin -> k_col_sort -> k_pairlist
//   force:    k_local (bonded + Ewald correction) lives in pm_local.hip, k_pair in pm_pair.h,
//             instantiated by pm_pair_*.hip
//   step:     k_kick_drift, k_half_kick, k_md_arm; state in and out: k_import, k_export
//
// All kernels take the device table of SysDev records plus a prefix array
// over the systems, so one launch covers every system of a batch.
#include "pm_kernels.h"

#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "../../include/protomol_amd.h"
#include "pm_device.h"
#include "pm_rebuild.h"

namespace pm {

// ------------------------------------------------------------- rebuild

// host staging (AoS FP64, per system) -> device state; what: kXfer* bits
__global__ void k_import(SysDev *sys, const int *ap, int nsys, int what) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ap[nsys]) return;
  const int s = find_sys(ap, nsys, g);
  SysDev &S = sys[s];
  const int a = g - ap[s], n = S.n_atoms;
  for (int d = 0; d < 3; ++d) {
    if (what & kXferX) S.x[d * n + a] = S.xaos[3 * (size_t)a + d];
    if (what & kXferV) S.v[d * n + a] = S.vaos[3 * (size_t)a + d];
  }
}

// one system's step bookkeeping (one thread): counters, the census, and the
// rebuild decision from the largest displacement since the last rebuild
__device__ __forceinline__ void step_begin_ctl(const SysDev &S, int mode, float lim2, int kmax) {
  S.census[0] = 0;
  S.census[1] = 0;
  int *ctl = S.ctl;
  ctl[kCtlFirst] = 0;  // the step's kick has run
  ctl[kCtlNoise] += 1;
  const unsigned d2 = max_disp2(S);
  const bool rb = rebuild_due(S, mode, lim2, kmax);
  *S.maxdisp_run = max(*S.maxdisp_run, d2);
  clear_disp2(S);
  ctl[kCtlRebuild] = rb ? 1 : 0;
  if (rb) {
    ctl[kCtlSince] = 0;
    ctl[kCtlCount] += 1;
    ctl[kCtlEntries] = 0;
  } else {
    ctl[kCtlSince] += 1;
  }
}

__global__ void k_bin(const SysDev *__restrict__ sys, const int *__restrict__ ap, int nsys, int mode, float lim2,
                      int kmax, int *rb_any) {
  bin_thread(sys, ap, nsys, mode, lim2, kmax, blockIdx.x * blockDim.x + threadIdx.x, rb_any);
}

// Rebuild, step 2 of 3, one block per column: the column's first atom,
// cluster and super-cluster from the counts of the columns before it, a
// bitonic sort of its bucket by (z, atom), slot assignment, and each
// cluster's geometry (8 lanes per cluster): bounding box, origin, the atoms'
// FP32 local coordinates.  rb_any (when given): no system of the launch
// rebuilds, every wave returns after one load.
__global__ __launch_bounds__(1024) void k_col_sort(const SysDev *__restrict__ sys, const int *__restrict__ cp, int nsys,
                                                   const int *__restrict__ rb_any) {
  if (rb_any && !*rb_any) return;
  __shared__ unsigned long long keys[kSortCap];
  __shared__ int red[3][1024 / kWave];
  const int gcol = blockIdx.x;
  if (gcol >= cp[nsys]) return;
  const int s = __builtin_amdgcn_readfirstlane(find_sys(cp, nsys, gcol));
  const SysDev &S = sys[s];
  const int col = gcol - cp[s];
  const int t = threadIdx.x, nt = blockDim.x, lane = t & 63, wv = t >> 6;
  if (!S.ctl[kCtlRebuild]) return;
  // clusters and super-clusters of the columns before this one
  int lc = 0, lsc = 0;
  for (int c = t; c < col; c += nt) {
    const int k = (S.col_count[c] + kCluster - 1) / kCluster;
    lc += k;
    lsc += (k + kScCl - 1) / kScCl;
  }
  lc = (int)wave_sum_ll(lc);
  lsc = (int)wave_sum_ll(lsc);
  if (lane == 0) {
    red[0][wv] = lc;
    red[1][wv] = lsc;
  }
  __syncthreads();
  int cl0 = 0, sc0 = 0;
  for (int w = 0; w < nt / kWave; ++w) {
    cl0 += red[0][w];
    sc0 += red[1][w];
  }
  const int n = S.col_count[col];
  if (n > kSortCap) {
    if (t == 0) atomicOr(S.err, (unsigned)kErrColumnCap);
    return;
  }
  const int ncl = (n + kCluster - 1) / kCluster, nsc = (ncl + kScCl - 1) / kScCl;
  if (cl0 + ncl > S.cl_cap || sc0 + nsc > S.sc_cap) {
    if (t == 0) atomicOr(S.err, (unsigned)kErrClusterCap);
    return;
  }
  if (t == 0) {
    S.col_cl_start[col] = cl0;
    S.col_ncl[col] = ncl;
    S.col_sc_start[col] = sc0;
    S.col_nsc[col] = nsc;
    if (col == S.ncol - 1) {  // the system's totals
      *S.n_clusters = cl0 + ncl;
      *S.n_sc = sc0 + nsc;
    }
  }
  for (int j = t; j < nsc; j += nt) S.sc_col[sc0 + j] = col;
  int P = 1;
  while (P < n) P <<= 1;
  const unsigned long long *bk = S.bucket + (size_t)col * kSortCap;
  for (int i = t; i < P; i += nt) keys[i] = i < n ? bk[i] : ~0ull;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P; i += nt) {
        int l = i ^ j;
        if (l > i) {
          unsigned long long a = keys[i], b = keys[l];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int k = t; k < ncl * kCluster; k += nt) {
    int slot = cl0 * kCluster + k;
    if (k < n) {
      int a = (int)(keys[k] & 0xffffffffull);
      S.slot[a] = slot;
      S.orig[slot] = a;
      S.smol[slot] = S.molecule[a];
      if (S.lj_mode == PM_LJ_PERATOM) S.ljp[slot] = make_float2(S.sig_half[a], S.eps_sqrt[a]);
      else S.ljt[slot] = S.type[a];
    } else {
      S.orig[slot] = -1;
      S.smol[slot] = -1;
      if (S.lj_mode == PM_LJ_PERATOM) S.ljp[slot] = make_float2(0.f, 0.f);
      else S.ljt[slot] = 0;
      S.xq[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  for (int c = t; c < ncl; c += nt) {
    S.cl_col[cl0 + c] = col;
    S.cl_nvalid[cl0 + c] = min(kCluster, n - c * kCluster);
  }
  // cluster geometry: lane k of each aligned group of 8 holds atom k of a cluster
  const int na = S.n_atoms;
  for (int i0 = 0; i0 < ncl * kCluster; i0 += nt) {  // block-uniform trip count
    const int i = i0 + t, c = i >> 3;
    const bool has = i < n;
    int a = 0;
    double rel[3] = {0.0, 0.0, 0.0}, lo[3], hi[3];
    if (has) a = (int)(keys[i] & 0xffffffffull);
    for (int d = 0; d < 3; ++d) {
      if (has) rel[d] = wrap_rel(S.x[d * na + a], S.omin[d], S.L[d]);
      lo[d] = has ? rel[d] : 1e300;
      hi[d] = has ? rel[d] : -1e300;
      for (int o = 1; o < kCluster; o <<= 1) {
        lo[d] = fmin(lo[d], __shfl_xor(lo[d], o));
        hi[d] = fmax(hi[d], __shfl_xor(hi[d], o));
      }
    }
    if (i < ncl * kCluster) {
      const int cg = cl0 + c;
      double org[3];
      for (int d = 0; d < 3; ++d) org[d] = 0.5 * (lo[d] + hi[d]);
      if ((i & (kCluster - 1)) == 0) {
        for (int d = 0; d < 3; ++d) S.cl_origin[d * S.cl_cap + cg] = org[d];
        S.cl_lo[cg] = make_float4((float)lo[0] - 1e-3f, (float)lo[1] - 1e-3f, (float)lo[2] - 1e-3f, 0.f);
        S.cl_hi[cg] = make_float4((float)hi[0] + 1e-3f, (float)hi[1] + 1e-3f, (float)hi[2] + 1e-3f, 0.f);
      }
      if (has) {
        float l[3];
        for (int d = 0; d < 3; ++d) {
          const double xa = S.x[d * na + a];
          const double ref = S.omin[d] + org[d] + (xa - S.omin[d] - rel[d]);
          S.xref[d * na + a] = ref;
          l[d] = (float)(xa - ref);
        }
        S.xq[cg * kCluster + (i & (kCluster - 1))] = make_float4(l[0], l[1], l[2], (float)S.charge[a]);
        S.lbuild[a] = make_float4(l[0], l[1], l[2], 0.f);
      }
    }
  }
}

__device__ __forceinline__ int floordiv(int a, int b) {
  int q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// Super-cluster pair list within rlist, half shell over columns.  A
// super-cluster is kScCl z-consecutive clusters of one column; its list holds
// the j clusters within rlist of any of them, with an 8-bit mask of which.
// One wave per super-cluster.  The candidates are the clusters of the
// neighbour columns (periodic x/y image applied) whose z extent, in one of the
// three z images, can come within rlist of the super-cluster's z range: the
// clusters of a column are z-ordered, so two binary searches bound each
// (column, z image) segment.  The wave walks the segments in a fixed order, so
// the list order is the same on every rebuild of the same positions.  Each
// entry is written whole as it is emitted: its FP32 image offset and the
// FULL/MODIFIED lane masks of cluster pairs that share a molecule, from the
// super-cluster's own excluded pairs, gathered in LDS before the walk.
constexpr int kListWaves = 4;                     // waves (super-clusters) per workgroup
constexpr int kListThreads = kListWaves * kWave;
constexpr int kWinChunk = kWave;                  // neighbour columns per segment pass
constexpr int kSegCap = 3 * kWinChunk;            // (column, z image) segments per pass

// per wave, pass B: each lane's entry's FULL and MODIFIED words per i cluster,
// word-major (word w of lane l at mask[w][l]), so a wave's accesses spread
// over the LDS banks
struct ListScratch {
  unsigned long long mask[2 * kScCl][kWave];
};
// per wave, pass B: the excluded pairs of the super-cluster's atoms as (the
// partner's cluster, bits (partner lane << 8) | (i cluster << 4) |
// (i lane << 1) | MODIFIED), chained per bucket of the partner's cluster
// index (head, next; -1 ends a chain); more than kHitCap send the
// super-cluster to the per-entry walk of the j atoms' partners
constexpr int kHitCap = 256, kHitBuckets = 256;
struct ListHits {
  int2 h[kHitCap];
  int next[kHitCap];
  int head[kHitBuckets];
};

// first k in [0, m) with key(k) >= v (m if none), keys ascending
__device__ __forceinline__ int lower_bound_hi_z(const float4 *hi, int c0, int m, float v) {
  int lo = 0, hi_i = m;
  while (lo < hi_i) {
    const int mid = (lo + hi_i) >> 1;
    if (hi[c0 + mid].z > v) hi_i = mid;
    else lo = mid + 1;
  }
  return lo;
}
// first k in [0, m) with lo(k).z >= v (m if none)
__device__ __forceinline__ int lower_bound_lo_z(const float4 *lo_b, int c0, int m, float v) {
  int lo = 0, hi_i = m;
  while (lo < hi_i) {
    const int mid = (lo + hi_i) >> 1;
    if (lo_b[c0 + mid].z >= v) hi_i = mid;
    else lo = mid + 1;
  }
  return lo;
}

__global__ __launch_bounds__(kListThreads) void k_pairlist(const SysDev *__restrict__ sys, const int *__restrict__ scp,
                                                           int nsys, float rlist2, const int *__restrict__ rb_any,
                                                           int hit_cap) {
  if (rb_any && !*rb_any) return;
  __shared__ float4 ib_lo[kListWaves][kScCl], ib_hi[kListWaves][kScCl];
  __shared__ int seg_start[kListWaves][kSegCap + 1], seg_cl[kListWaves][kSegCap], seg_sh[kListWaves][kSegCap];
  __shared__ ListScratch scratch[kListWaves];
  __shared__ ListHits hits[kListWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int scg = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kListWaves + wv));
  if (scg >= scp[nsys]) return;
  const int s = __builtin_amdgcn_readfirstlane(find_sys(scp, nsys, scg));
  const SysDev &S = sys[s];
  if (!S.ctl[kCtlRebuild]) return;
  const int sc = scg - scp[s];
  if (sc >= *S.n_sc) return;
  const int coli = S.sc_col[sc];
  const int kz = sc - S.col_sc_start[coli];
  const int cl0 = S.col_cl_start[coli] + kz * kScCl;
  const int nic = min(kScCl, S.col_ncl[coli] - kz * kScCl);
  const int cxi = coli % S.ncx, cyi = coli / S.ncx;
  // the column's bucket count has been read by every block of k_col_sort:
  // clear it for the next rebuild (a column without atoms kept its 0)
  if (kz == 0 && lane == 0) S.col_count[coli] = 0;
  const float Lx = (float)S.L[0], Ly = (float)S.L[1], Lz = (float)S.L[2];
  const float ax = (float)S.ax, ay = (float)S.ay;
  const float ix0 = cxi * ax, ix1 = (cxi + 1) * ax, iy0 = cyi * ay, iy1 = (cyi + 1) * ay;
  const float rl = sqrtf(rlist2);
  const int cap = S.list_cap;
  const int wx = 2 * S.rx + 1, nwin_all = wx * (2 * S.ry + 1);
  if (lane < kScCl) {
    const bool ok = lane < nic && S.cl_nvalid[cl0 + lane] > 0;
    ib_lo[wv][lane] = ok ? S.cl_lo[cl0 + lane] : make_float4(1e30f, 1e30f, 1e30f, 0.f);
    ib_hi[wv][lane] = ok ? S.cl_hi[cl0 + lane] : make_float4(1e30f, 1e30f, 1e30f, 0.f);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float zlo = 1e30f, zhi = -1e30f;
  for (int c = 0; c < nic; ++c) {
    zlo = fminf(zlo, ib_lo[wv][c].z);
    zhi = fmaxf(zhi, ib_hi[wv][c].z);
  }
  const double ox = S.cl_origin[cl0], oy = S.cl_origin[S.cl_cap + cl0], oz = S.cl_origin[2 * S.cl_cap + cl0];
  // the super-cluster's excluded pairs, from its own atoms' side: lane = i
  // atom (i cluster lane >> 3, i lane lane & 7), one partner per round, each
  // kept when it is in the same molecule -- every excluded pair of an entry's
  // tiles is one of these, found in LDS by the partner's cluster
  int nhit = 0;
  {
    for (int b = lane; b < kHitBuckets; b += kWave) hits[wv].head[b] = -1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int c = lane >> 3, ii = lane & 7;
    int bi = -1, mi = -1, k = 0, kend = 0;
    unsigned long long wf = 0, bits = 0;
    if (c < nic) {
      bi = S.orig[(cl0 + c) * kCluster + ii];
      if (bi >= 0) {
        mi = S.molecule[bi];
        wf = S.excl_win[2 * (size_t)bi];
        if ((wf >> 63) != 0) {
          k = S.excl_start[bi];
          kend = S.excl_start[bi + 1];
        } else {
          bits = (wf | S.excl_win[2 * (size_t)bi + 1]) & ~(1ull << 63);
        }
      }
    }
    while (true) {
      int b = -1, cls = 0;
      if (bits) {
        const int d = __ffsll((long long)bits) - 1;
        bits &= bits - 1;
        b = bi + d - 31;
        cls = ((wf >> d) & 1ull) ? PM_EXCL_FULL : PM_EXCL_MODIFIED;
      } else if (k < kend) {
        b = S.excl_partner[k];
        cls = S.excl_class[k];
        ++k;
      }
      if (wave_ballot(b >= 0) == 0) break;
      bool keep = false;
      int2 hv = make_int2(0, 0);
      if (b >= 0 && (cls == PM_EXCL_FULL || cls == PM_EXCL_MODIFIED) && S.molecule[b] == mi) {
        const int sb = S.slot[b];
        keep = true;
        hv = make_int2(sb >> 3, ((sb & (kCluster - 1)) << 8) | (c << 4) | (ii << 1) |
                                     (cls == PM_EXCL_MODIFIED ? 1 : 0));
      }
      const unsigned long long kb = wave_ballot(keep);
      const int pos = nhit + __popcll(kb & ((1ull << lane) - 1ull));
      if (keep && pos < kHitCap) {
        hits[wv].h[pos] = hv;
        hits[wv].next[pos] = atomicExch(&hits[wv].head[hv.x & (kHitBuckets - 1)], pos);
      }
      nhit += __popcll(kb);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  long long pairs8 = 0;  // 8x8 cluster pairs, for the statistics
  int mslot = 0;
  unsigned long long *mbase = S.list_mask + (size_t)sc * S.mask_cap * (2 * kScCl);
  // this lane's entry: FULL, MODIFIED word per i cluster, word w at M[w * kWave]
  unsigned long long *M = &scratch[wv].mask[0][lane];
  uint4 *out = S.list + 2 * (size_t)sc * cap;
  int count = 0;
  // ---- pass A: entries (j cluster, image, i-cluster mask), segment by segment
  for (int q0 = 0; q0 < nwin_all; q0 += kWinChunk) {
    // lane q: one neighbour column of the window and its three z-image segments
    const int q = q0 + lane;
    int colj = 0, sxy = 0, cs = 0, ncj = 0;
    bool col_ok = false;
    if (q < nwin_all) {
      const int dy = q / wx - S.ry, dx = q % wx - S.rx;
      if (!(dy < 0 || (dy == 0 && dx < 0))) {  // half shell over columns
        const int cyu = cyi + dy, cxu = cxi + dx;
        const float gy = fmaxf(0.f, fmaxf(cyu * ay - iy1, iy0 - (cyu + 1) * ay) - 0.01f);
        const float gx = fmaxf(0.f, fmaxf(cxu * ax - ix1, ix0 - (cxu + 1) * ax) - 0.01f);
        if (gx * gx + gy * gy < rlist2) {
          const int sy = floordiv(cyu, S.ncy), sx = floordiv(cxu, S.ncx);
          colj = (cyu - sy * S.ncy) * S.ncx + (cxu - sx * S.ncx);
          sxy = (sx + 1) | ((sy + 1) << 2) | ((dx == 0 && dy == 0) ? 16 : 0);
          cs = S.col_cl_start[colj];
          ncj = S.col_ncl[colj];
          col_ok = true;
        }
      }
    }
    // segments in (z image, window column) order
    int nseg = 0, ntot = 0;
    for (int sz = -1; sz <= 1; ++sz) {
      const float szL = sz * Lz;
      int k0 = 0, k1 = 0;
      if (col_ok) {
        // a candidate's box in this image must reach above zlo - rl and below zhi + rl
        // (0.01 A of slack: the boxes are z-ordered up to FP32 rounding of the sort keys)
        k0 = lower_bound_hi_z(S.cl_hi, cs, ncj, zlo - rl - szL - 0.01f);
        k1 = lower_bound_lo_z(S.cl_lo, cs, ncj, zhi + rl - szL + 0.01f);
        if (k1 < k0) k1 = k0;
      }
      const int cnt = k1 - k0;
      const unsigned long long has = wave_ballot(cnt > 0);
      // inclusive prefix of cnt over the lanes (wave scan)
      int pre = cnt;
      for (int o = 1; o < kWave; o <<= 1) {
        const int v = __shfl_up(pre, o);
        if (lane >= o) pre += v;
      }
      const int tot = __shfl(pre, kWave - 1);
      if (cnt > 0) {
        const int slot = nseg + __popcll(has & ((1ull << lane) - 1ull));
        seg_start[wv][slot] = ntot + pre - cnt;
        seg_cl[wv][slot] = cs + k0;
        seg_sh[wv][slot] = sxy | ((sz + 1) << 5);
      }
      nseg += __popcll(has);
      ntot += tot;
    }
    if (lane == 0) seg_start[wv][nseg] = ntot;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int base = 0; base < ntot; base += kWave) {
      const int k = base + lane;
      unsigned im = 0;
      int cj = 0, shw = 0;
      if (k < ntot) {
        int lo = 0, hi = nseg;  // seg_start[lo] <= k < seg_start[hi]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (seg_start[wv][mid] <= k) lo = mid;
          else hi = mid;
        }
        cj = seg_cl[wv][lo] + (k - seg_start[wv][lo]);
        shw = seg_sh[wv][lo];
        const int sz = ((shw >> 5) & 3) - 1;
        const bool samecol = (shw & 16) != 0;
        const float sxL = ((shw & 3) - 1) * Lx, syL = (((shw >> 2) & 3) - 1) * Ly, szL = sz * Lz;
        float4 loj = S.cl_lo[cj], hij = S.cl_hi[cj];
        loj.x += sxL;
        hij.x += sxL;
        loj.y += syL;
        hij.y += syL;
        loj.z += szL;
        hij.z += szL;
        if (S.cl_nvalid[cj] > 0)
          for (int c = 0; c < nic; ++c) {
            // each cluster pair once: within a column, j at or above i (z image 0)
            if (samecol && !(sz > 0 || (sz == 0 && cj >= cl0 + c))) continue;
            const float4 loi = ib_lo[wv][c], hii = ib_hi[wv][c];
            const float ex = fmaxf(0.f, fmaxf(loj.x - hii.x, loi.x - hij.x));
            const float ey = fmaxf(0.f, fmaxf(loj.y - hii.y, loi.y - hij.y));
            const float ez = fmaxf(0.f, fmaxf(loj.z - hii.z, loi.z - hij.z));
            if (ex * ex + ey * ey + ez * ez < rlist2) im |= 1u << c;
          }
      }
      const bool hit = im != 0;
      const unsigned long long b = wave_ballot(hit);
      if (b == 0) continue;
      const int pos = count + __popcll(b & ((1ull << lane) - 1ull));
      const bool put = hit && pos < cap;
      // the entry whole: j cluster, image, valid j atoms, the j cluster's
      // origin relative to the super-cluster's (FP32), and its exclusion masks
      unsigned share = 0, w0 = 0;
      float offx = 0.f, offy = 0.f, offz = 0.f;
      for (int w = 0; w < 2 * kScCl; ++w) M[w * kWave] = 0ull;
      if (put) {
        const int sx = (shw & 3) - 1, sy = ((shw >> 2) & 3) - 1, sz = ((shw >> 5) & 3) - 1;
        const int sh = (sx + 1) * 9 + (sy + 1) * 3 + (sz + 1);
        const int nvj = S.cl_nvalid[cj];
        w0 = (unsigned)cj | ((unsigned)sh << 23) | ((unsigned)max(nvj - 1, 0) << 28);
        offx = (float)(S.cl_origin[cj] + (double)sx * S.L[0] - ox);
        offy = (float)(S.cl_origin[S.cl_cap + cj] + (double)sy * S.L[1] - oy);
        offz = (float)(S.cl_origin[2 * S.cl_cap + cj] + (double)sz * S.L[2] - oz);
        pairs8 += __popc(im);
        // the excluded pairs of the entry's tiles, lane bit jj * 8 + ii of the
        // tile's FULL or MODIFIED word: the super-cluster's pairs whose partner
        // is in cluster cj, or, past kHitCap of those, each j atom's exclusion
        // partners (its index window, or the symmetric CSR for an atom with a
        // partner outside it), located by slot -- the super-cluster's i
        // clusters are the contiguous cl0 .. cl0 + nic - 1 -- and kept when
        // the partner is in a tile of the entry and in the same molecule
        if (nhit <= min(hit_cap, kHitCap)) {
          for (int q = hits[wv].head[cj & (kHitBuckets - 1)]; q >= 0; q = hits[wv].next[q]) {
            const int2 hv = hits[wv].h[q];
            const int c = (hv.y >> 4) & 7;
            if (hv.x == cj && ((im >> c) & 1u)) {
              share |= 1u << c;
              M[(2 * c + (hv.y & 1)) * kWave] |= 1ull << (((hv.y >> 8) & 7) * kCluster + ((hv.y >> 1) & 7));
            }
          }
        } else
        for (int jj = 0; jj < kCluster; ++jj) {
          const int sj = cj * kCluster + jj;
          const int bj = S.orig[sj];
          if (bj < 0) continue;
          const int mj = S.smol[sj];
          const unsigned long long wf = S.excl_win[2 * (size_t)bj], wm = S.excl_win[2 * (size_t)bj + 1];
          auto partner = [&](int b, int cls) {
            const int sb = S.slot[b];
            const int c = (sb >> 3) - cl0;
            if (c >= 0 && c < nic && ((im >> c) & 1u) && S.smol[sb] == mj &&
                (cls == PM_EXCL_FULL || cls == PM_EXCL_MODIFIED)) {
              share |= 1u << c;
              M[(2 * c + (cls == PM_EXCL_MODIFIED ? 1 : 0)) * kWave] |=
                  1ull << (jj * kCluster + (sb & (kCluster - 1)));
            }
          };
          if ((wf >> 63) != 0) {
            for (int k = S.excl_start[bj]; k < S.excl_start[bj + 1]; ++k) partner(S.excl_partner[k], S.excl_class[k]);
          } else {
            unsigned long long bits = (wf | wm) & ~(1ull << 63);
            while (bits) {
              const int d = __ffsll((long long)bits) - 1;
              bits &= bits - 1;
              partner(bj + d - 31, ((wf >> d) & 1ull) ? PM_EXCL_FULL : PM_EXCL_MODIFIED);
            }
          }
        }
      }
      // mask slots in entry order
      const unsigned long long fl = wave_ballot(share != 0);
      const int my = mslot + __popcll(fl & ((1ull << lane) - 1ull));
      if (put) {
        unsigned e1 = im;
        if (share != 0) {
          if (my < S.mask_cap) {
            for (int w = 0; w < 2 * kScCl; ++w) mbase[(size_t)my * (2 * kScCl) + w] = M[w * kWave];
            w0 |= 0x80000000u;
            e1 |= (unsigned)my << 8;
          } else {
            atomicOr(S.err, (unsigned)kErrListCap);
          }
        }
        out[2 * pos] = make_uint4(w0, __float_as_uint(offx), __float_as_uint(offy), __float_as_uint(offz));
        out[2 * pos + 1] = make_uint4(e1, 0u, 0u, 0u);
      }
      mslot += __popcll(fl);
      count += __popcll(b);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (count > cap) {
    if (lane == 0) atomicOr(S.err, (unsigned)kErrListCap);
    count = cap;
  }
  if (lane == 0) S.list_len[sc] = count;
  pairs8 = wave_sum_ll(pairs8);
  if (lane == 0 && pairs8) atomicAdd(&S.ctl[kCtlEntries], (int)pairs8);
}

// ------------------------------------------------------------ integrator

// First step of a run (the system's kCtlFirst word is set):
//   doHalfKickdoDrift (LeapfrogIntegrator.cpp:33-56)  v += F*h*0.5/m; x += v*h
// later steps:
//   doKickdoDrift     (LeapfrogIntegrator.cpp:58-87)  v += F*h/m;     x += v*h
// The FP32 cluster-local coordinates for the next force evaluation are
// written; the consumed pair forces are cleared here on the GB path and by
// the next list check (in slot order) on the cutoff path.
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): a counter-based generator, so atom a's noise
// at kick k of its system depends on (seed, a, k) only
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
  for (int r = 0; r < 10; ++r) {
    const unsigned hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const unsigned hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// three N(0, 1) deviates for atom a at noise counter ctr (Box-Muller on the
// four 32-bit outputs, u = (x + 1/2) 2^-32 in (0, 1))
__device__ __forceinline__ void lv_gauss(unsigned seed, int a, unsigned ctr, double g[3]) {
  const uint4 r = philox4x32_10(make_uint4((unsigned)a, ctr, 0x4C414E47u, 0u), make_uint2(seed, 0x50524F54u));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = ((double)r.x + 0.5) * s, u1 = ((double)r.y + 0.5) * s;
  const double u2 = ((double)r.z + 0.5) * s, u3 = ((double)r.w + 0.5) * s;
  const double m0 = sqrt(-2.0 * log(u0)), m1 = sqrt(-2.0 * log(u2));
  double s0, c0, s1, c1;
  sincospi(2.0 * u1, &s0, &c0);
  sincospi(2.0 * u3, &s1, &c1);
  g[0] = m0 * c0;
  g[1] = m0 * s0;
  g[2] = m1 * c1;
}

// one Langevin half kick (LangevinLeapfrogIntegrator.cpp:46-70):
// v = v e^{-g dt/2} + F (1 - e^{-g dt/2}) / (g m) + N(0,1) sqrt(2 kB T g / m) ndt
__device__ __forceinline__ void lv_half(double v[3], const double F[3], double m, const LangevinDev &lv,
                                        unsigned seed, int a, unsigned ctr) {
  double g[3];
  lv_gauss(seed, a, ctr, g);
  const double sf = sqrt(lv.fc / m);
  for (int d = 0; d < 3; ++d) v[d] = v[d] * lv.vdt + F[d] * lv.fdt / m + g[d] * sf * lv.ndt;
}

// One atom's kick + drift (see k_kick_drift); returns the squared distance
// of its cluster-local position from where the pair list was built.
__device__ __forceinline__ float kick_drift_atom(const SysDev &S, int a, double h, const LangevinDev &lv, bool zero_f) {
  const int n = S.n_atoms;
  // every load first (global address space, so none waits on a store), then
  // the arithmetic, then the stores
  PM_GLOBAL long long *fx = gp(S.fx), *fy = gp(S.fy), *fz = gp(S.fz);
  const PM_GLOBAL long long *fl = gp(S.fl);
  PM_GLOBAL double *V = gp(S.v), *X = gp(S.x);
  const PM_GLOBAL double *XR = gp(S.xref);
  PM_GLOBAL float4 *xq = gp(S.xq);
  const int sl = gp(S.slot)[a];
  const int mode = gp(S.ctl)[kCtlFirst] ? 0 : 1;
  const unsigned c = (unsigned)gp(S.ctl)[kCtlNoise];
  const double m = 1.0 / gp(S.inv_mass)[a];
  const long long l0 = fl[a], l1 = fl[n + a], l2 = fl[2 * n + a];
  double v[3] = {V[a], V[n + a], V[2 * n + a]};
  const double x[3] = {X[a], X[n + a], X[2 * n + a]};
  const double xr[3] = {XR[a], XR[n + a], XR[2 * n + a]};
  const float4 b = gp(S.lbuild)[a];
  const long long p0 = fx[sl], p1 = fy[sl], p2 = fz[sl];
  const float qw = xq[sl].w;
  const double F[3] = {(double)(p0 + l0) * kFScaleInv, (double)(p1 + l1) * kFScaleInv, (double)(p2 + l2) * kFScaleInv};
  if (lv.on) {
    // the closing half kick of the previous step (noise 2c) unless this is
    // the first step of a run, then the opening half kick of this one (2c + 1)
    if (mode == 1) lv_half(v, F, m, lv, S.lv_seed, a, 2u * c);
    lv_half(v, F, m, lv, S.lv_seed, a, 2u * c + 1u);
  } else {
    for (int d = 0; d < 3; ++d) {
      if (mode == 0) v[d] += F[d] * h * 0.5 / m;
      else v[d] += F[d] * h / m;
    }
  }
  double xx[3];
  float l[3];
  for (int d = 0; d < 3; ++d) {
    xx[d] = x[d] + v[d] * h;
    l[d] = (float)(xx[d] - xr[d]);
  }
  if (zero_f) {  // (the cutoff path's list check clears them in slot order instead)
    fx[sl] = 0;
    fy[sl] = 0;
    fz[sl] = 0;
  }
  for (int d = 0; d < 3; ++d) {
    V[d * n + a] = v[d];
    X[d * n + a] = xx[d];
  }
  xq[sl] = make_float4(l[0], l[1], l[2], qw);
  const float ex = l[0] - b.x, ey = l[1] - b.y, ez = l[2] - b.z;
  float d2 = ex * ex + ey * ey + ez * ez;
  if (!isfinite(d2)) d2 = 3.0e38f;
  return d2;
}

__global__ __launch_bounds__(256) void k_kick_drift(const SysDev *__restrict__ sys, const int *__restrict__ ap,
                                                    int nsys, double h, LangevinDev lv, int *rb_any, int zero_f) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0 && rb_any) *rb_any = 0;  // the next list check raises it again
  const int total = ap[nsys];
  const bool ok = g < total;
  float d2 = 0.f;
  int s = wave_sys(ap, nsys, g, total);
  if (s >= 0) {
    // the wave's atoms are one system's: its record is read with scalar loads
    const SysDev *__restrict__ S = sys + s;
    if (ok) d2 = kick_drift_atom(*S, g - ap[s], h, lv, zero_f != 0);
  } else {
    s = ok ? find_sys(ap, nsys, g) : 0;
    if (ok) d2 = kick_drift_atom(sys[s], g - ap[s], h, lv, zero_f != 0);
  }
  // max per system: a block whose atoms all belong to one system reduces
  // through LDS to one atomic; a block straddling systems uses one per lane
  __shared__ float red[4];
  __shared__ int bs[2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int glast = min((int)(blockIdx.x * blockDim.x + blockDim.x), total) - 1;
  if (threadIdx.x == 0) bs[0] = s;
  if (g == glast) bs[1] = s;
  const float mx = wave_max_f(d2);
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  if (bs[0] == bs[1]) {
    if (threadIdx.x == 0) {
      float m = red[0];
      for (int k = 1; k < nw; ++k) m = fmaxf(m, red[k]);
      atomicMax(sys[s].maxdisp2 + (blockIdx.x & (kDispSlots - 1)), __float_as_uint(m));
    }
  } else if (ok) {
    atomicMax(sys[s].maxdisp2 + (blockIdx.x & (kDispSlots - 1)), __float_as_uint(d2));
  }
}

// StandardIntegrator::doHalfKick (StandardIntegrator.cpp:80-92): h = 0.5*dt*ITF, v += F*h/m
__device__ __forceinline__ void half_kick_atom(const SysDev &S, int a, double hh, const LangevinDev &lv) {
  const int n = S.n_atoms;
  const PM_GLOBAL long long *fl = gp(S.fl);
  PM_GLOBAL double *V = gp(S.v);
  const int sl = gp(S.slot)[a];
  const unsigned c = (unsigned)gp(S.ctl)[kCtlNoise];
  const double m = 1.0 / gp(S.inv_mass)[a];
  const long long l0 = fl[a], l1 = fl[n + a], l2 = fl[2 * n + a];
  double v[3] = {V[a], V[n + a], V[2 * n + a]};
  const long long p0 = gp(S.fx)[sl], p1 = gp(S.fy)[sl], p2 = gp(S.fz)[sl];
  const double F[3] = {(double)(p0 + l0) * kFScaleInv, (double)(p1 + l1) * kFScaleInv, (double)(p2 + l2) * kFScaleInv};
  if (lv.on) {
    // the closing half kick of the last step: noise counter 2c
    lv_half(v, F, m, lv, S.lv_seed, a, 2u * c);
  } else {
    for (int d = 0; d < 3; ++d) v[d] += F[d] * hh / m;
  }
  for (int d = 0; d < 3; ++d) V[d * n + a] = v[d];
}

__global__ void k_half_kick(const SysDev *__restrict__ sys, const int *__restrict__ ap, int nsys, double hh,
                            LangevinDev lv) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = ap[nsys];
  const int s = wave_sys(ap, nsys, g, total);
  if (g >= total) return;
  if (s >= 0) {
    const SysDev *__restrict__ S = sys + s;
    half_kick_atom(*S, g - ap[s], hh, lv);
  } else {
    const int t = find_sys(ap, nsys, g);
    half_kick_atom(sys[t], g - ap[t], hh, lv);
  }
}

// zero the fixed-point forces of every slot (a fill kernel, not a memset)
__global__ void k_zero_forces(SysDev *sys, const int *slp, int nsys) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= slp[nsys]) return;
  const int s = find_sys(slp, nsys, g);
  SysDev &S = sys[s];
  const int k = g - slp[s];
  S.fx[k] = 0;
  S.fy[k] = 0;
  S.fz[k] = 0;
}

// Start of a force evaluation, one block per system: clear the energy/virial
// accumulators and the census, and decide whether the pair list is rebuilt
// (kStepForce: always; kStepAuto: when an atom moved more than lim from where
// the list was built -- the Verlet criterion, with lim = skin/2 -- or kmax
// steps have passed).  The decision is read by every rebuild kernel.
__global__ void k_step_begin(SysDev *sys, int nsys, int mode, float lim2, int kmax) {
  const int s = blockIdx.x;
  if (s >= nsys) return;
  SysDev &S = sys[s];
  for (int k = threadIdx.x; k < kAccN * kAccCopies; k += blockDim.x) S.acc[k] = 0;
  if (threadIdx.x == 0) step_begin_ctl(S, mode, lim2, kmax);
}

// device state -> host staging (AoS FP64, per system); what: kXfer* bits
__global__ void k_export(SysDev *sys, const int *ap, int nsys, int what) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ap[nsys]) return;
  const int s = find_sys(ap, nsys, g);
  const SysDev &S = sys[s];
  const int a = g - ap[s], n = S.n_atoms;
  const int sl = S.slot[a];
  for (int d = 0; d < 3; ++d) {
    if (what & kXferX) S.xaos[3 * (size_t)a + d] = S.x[d * n + a];
    if (what & kXferV) S.vaos[3 * (size_t)a + d] = S.v[d * n + a];
  }
  if (what & kXferF) {
    S.faos[3 * (size_t)a] = (double)(S.fx[sl] + S.fl[a]) * kFScaleInv;
    S.faos[3 * (size_t)a + 1] = (double)(S.fy[sl] + S.fl[n + a]) * kFScaleInv;
    S.faos[3 * (size_t)a + 2] = (double)(S.fz[sl] + S.fl[2 * n + a]) * kFScaleInv;
  }
}

// per-system scalars of a launch, packed for one device-to-host copy:
// kScalN int64 per system (see kScal*)
__global__ void k_gather_scalars(const SysDev *sys, int nsys, long long *out) {
  const int s = blockIdx.x;
  if (s >= nsys) return;
  const SysDev &S = sys[s];
  long long *o = out + (size_t)s * kScalN;
  for (int k = threadIdx.x; k < kScalN; k += blockDim.x) {
    long long v = 0;
    if (k < kAccN)
      for (int c = 0; c < kAccCopies; ++c) v += S.acc[(size_t)c * kAccN + k];
    else if (k < kScalCensus + 2) v = (long long)S.census[k - kScalCensus];
    else if (k == kScalErr) v = *S.err;
    else if (k >= kScalCtl && k < kScalCtl + kCtlN) v = S.ctl[k - kScalCtl];
    else if (k == kScalMaxdispRun) v = *S.maxdisp_run;
    o[k] = v;
  }
}

// clear the error bits of the systems of a launch
__global__ void k_clear_err(SysDev *sys, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) *sys[s].err = 0u;
}

// ------------------------------------------------------------- launchers


void launch_import(const Launch &L, int what) {
  if (L.n_atoms_total) hipLaunchKernelGGL(k_import, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, what);
}
void launch_export(const Launch &L, int what) {
  if (L.n_atoms_total) hipLaunchKernelGGL(k_export, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, what);
}
void launch_gather_scalars(const Launch &L, long long *out) {
  hipLaunchKernelGGL(k_gather_scalars, dim3(L.nsys), dim3(64), 0, L.stream, L.sys, L.nsys, out);
}
void launch_clear_err(const Launch &L) {
  hipLaunchKernelGGL(k_clear_err, dim3(nblk(L.nsys, 64)), dim3(64), 0, L.stream, L.sys, L.nsys);
}

bool launch_rebuild(const Launch &L, float rlist2, int mode, float lim2, int kmax, const ForceDev *local) {
  if (!L.n_atoms_total) {
    launch_step_begin(L, mode, lim2, kmax);
    return false;
  }
  if (local && L.n_local_total)
    launch_bin_local(L, *local, mode, lim2, kmax);
  else
    hipLaunchKernelGGL(k_bin, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys,
                       mode, lim2, kmax, L.rb_any);
  hipLaunchKernelGGL(k_col_sort, dim3(L.n_cols_total), dim3(1024), 0, L.stream, L.sys, L.col_prefix, L.nsys, L.rb_any);
  hipLaunchKernelGGL(k_pairlist, dim3(nblk(L.n_sc_total, kListWaves)), dim3(kListThreads), 0, L.stream, L.sys,
                     L.sc_prefix, L.nsys, rlist2, L.rb_any, L.hit_cap);
  return local && L.n_local_total;
}

bool launch_pair(const Launch &L, const ForceDev &P, int kind) {
  if (!L.n_sc_total) return true;
  const int a = P.term_a, b = P.term_b;
  if (a == PM_TERM_LJ && b == PM_TERM_NONE) launch_pair_lj(L, P, kind);
  else if (a == PM_TERM_COULOMB && b == PM_TERM_NONE) launch_pair_coulomb(L, P, kind);
  else if (a == PM_TERM_LJ && b == PM_TERM_COULOMB) launch_pair_lj_coulomb(L, P, kind);
  else if (a == PM_TERM_LJ && b == PM_TERM_EWALD_REAL) launch_pair_lj_ewald(L, P, kind);
  else if (a == PM_TERM_NONE && b == PM_TERM_EWALD_REAL) launch_pair_ewald(L, P, kind);
  else return false;
  return true;
}

void launch_kick_drift(const Launch &L, double h, const LangevinDev &lv, bool zero_forces) {
  if (L.n_atoms_total)
    hipLaunchKernelGGL(k_kick_drift, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix,
                       L.nsys, h, lv, L.rb_any, zero_forces ? 1 : 0);
}
void launch_half_kick(const Launch &L, double hh, const LangevinDev &lv) {
  if (L.n_atoms_total) hipLaunchKernelGGL(k_half_kick, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, hh, lv);
}
// buildMolecularCenterOfMass (TopologyUtilities.cpp:677-684) with
// molecularCenterOfMass (:645-674): one thread per molecule, the atoms in
// index order, Kahan-compensated mass-weighted sums, FP64
__global__ __launch_bounds__(256) void k_mol_com(SysDev *sys) {
#pragma clang fp contract(off)
  const SysDev &S = sys[blockIdx.y];
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= S.n_mol) return;
  const int n = S.n_atoms, a0 = S.mol_start[m], a1 = S.mol_start[m + 1];
  double c[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0}, sm = 0.0;
  for (int k = a0; k < a1; ++k) {
    const int a = S.mol_atoms[k];
    const double mass = S.mass[a];
    if (k == a0) {
      sm = mass;
      for (int q = 0; q < 3; ++q) c[q] = S.x[q * n + a] * mass;
      continue;
    }
    sm += mass;
    for (int q = 0; q < 3; ++q) {
      const double tx = S.x[q * n + a] * mass;
      const double ty = tx - t[q];
      const double tt = c[q] + ty;
      t[q] = (tt - c[q]) - ty;
      c[q] = tt;
    }
  }
  for (int q = 0; q < 3; ++q) S.com[q * S.n_mol + m] = a1 > a0 ? c[q] / sm : 0.0;
}

void launch_mol_com(const Launch &L) {
  if (L.max_mol > 0)
    hipLaunchKernelGGL(k_mol_com, dim3(nblk(L.max_mol, 256), L.nsys), dim3(256), 0, L.stream, L.sys);
}

// zero `bytes` (a multiple of 8) at p: a fill kernel, not a memset
__global__ void k_fill_zero(unsigned long long *p, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0ull;
}

void launch_fill_zero(void *p, size_t bytes, hipStream_t stream) {
  const size_t n = bytes / 8;
  if (!n) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, stream, static_cast<unsigned long long *>(p), n);
}

void launch_zero_forces(const Launch &L) {
  if (L.n_slots_total) hipLaunchKernelGGL(k_zero_forces, dim3(nblk(L.n_slots_total, 256)), dim3(256), 0, L.stream, L.sys, L.slot_prefix, L.nsys);
}
// start of an MD run for the systems of the launch: the first kick is a half
// kick, and the run's rebuild count and displacement maximum restart
__global__ void k_md_arm(SysDev *sys, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  SysDev &S = sys[s];
  S.ctl[kCtlFirst] = 1;
  S.ctl[kCtlCount] = 0;
  *S.maxdisp_run = 0u;
}

void launch_md_arm(const Launch &L) {
  hipLaunchKernelGGL(k_md_arm, dim3(nblk(L.nsys, 64)), dim3(64), 0, L.stream, L.sys, L.nsys);
}

void launch_step_begin(const Launch &L, int mode, float lim2, int kmax) {
  hipLaunchKernelGGL(k_step_begin, dim3(L.nsys), dim3(64), 0, L.stream, L.sys, L.nsys, mode, lim2, kmax);
}
}  // namespace pm
