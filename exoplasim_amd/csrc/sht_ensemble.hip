// sht_ensemble.hip -- round trips of an ensemble of members with different
// truncations (T21 / T42 / T63) in two persistent kernels captured in one
// hipGraph.
//
// The reference runs one model per process; an ensemble of uneven
// truncations there is many processes.  Here every member's round trip is a
// list of work items (the blocks of the per-truncation kernels of
// sht_transforms.hip): one persistent kernel drains all inverse items
// (sp2fc/sp2fcdmu/dv2uv fused with fc2gp) of all members, the next drains all
// direct items (gp2fc fused with fc2sp).  Items are handed out by an atomic
// counter, most expensive truncation first, so the blocks that finish the
// small T21 items early take on the remaining work instead of idling.  The
// tile variant of each truncation's items is chosen by exo_sht_ens_autotune.
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "../../include/exo_sht.h"
#include "sht_bodies.h"

namespace exo {

constexpr int ENS_MAXGROUP = 8;
constexpr int ENS_MAXJOB = 3 * ENS_MAXGROUP;

// one launch's worth of items of one group (what a per-truncation launch
// would have been)
struct EnsInvJob {
  InvArgs a;
  const void* A0;
  const void* A1;
  const void* tv;
  const void* tw;
  long plane;
  int ntru, var;
};
struct EnsFwdJob {
  FwdArgs a;
  const void* tv;
  const void* tw;
  int ntru, var;
};
// first item of each job (start[njob] = total); host-side copy
struct EnsIndex {
  int njob;
  int start[ENS_MAXJOB + 1];
};

// Job tables are read through the constant address space: the job index is
// uniform, so every field becomes a scalar load at its point of use instead
// of a value held in registers across the whole item.
template <class S>
using ConstPtr = __attribute__((address_space(4))) const S*;

// job of item: the last job whose first item is <= item (uniform loop)
__device__ __forceinline__ int job_of(ConstPtr<int> start, int njob, int item) {
  int j = 0;
  for (int k = 1; k < njob; ++k)
    if (item >= start[k]) j = k;
  return j;
}

// next item for this block; every block leaves once the counter passes total
__device__ __forceinline__ int next_item(int* counter, int* s_item) {
  if (threadIdx.x == 0) s_item[0] = atomicAdd(counter, 1);
  __syncthreads();
  const int item = __builtin_amdgcn_readfirstlane(s_item[0]);
  __syncthreads();
  return item;
}

template <int T, int MODE, int NT>
__device__ __forceinline__ void ens_inv_item(const EnsInvJob& jb, int blk, char* lds) {
  if (NT == 512) {
    if (jb.a.kind == INV_UV) {
      if (jb.var == 3)
        inv_body<T, MODE, INV_UV, true, 3>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw, lds);
      else
        inv_body<T, MODE, INV_UV, true, 2>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw, lds);
    } else {
      if (jb.var == 3)
        inv_body<T, MODE, INV_SCALAR, true, 3>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw,
                                               lds);
      else
        inv_body<T, MODE, INV_SCALAR, true, 2>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw,
                                               lds);
    }
    return;
  }
  if (jb.a.kind == INV_UV) {
    if (jb.var)
      inv_body<T, MODE, INV_UV, true, 1>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw, lds);
    else
      inv_body<T, MODE, INV_UV, true, 0>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw, lds);
  } else {
    if (jb.var)
      inv_body<T, MODE, INV_SCALAR, true, 1>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw,
                                             lds);
    else
      inv_body<T, MODE, INV_SCALAR, true, 0>(jb.a, blk, jb.A0, jb.A1, jb.plane, jb.tv, jb.tw,
                                             lds);
  }
}

template <int MODE, int NT>
__global__ __launch_bounds__(NT) void k_ens_inv(const EnsInvJob* jobs_, const int* start_,
                                                int njob, int* counter) {
  extern __shared__ __align__(16) char lds[];
  __shared__ __align__(16) int s_item[4];  // keeps the dynamic LDS 16-byte aligned
  const ConstPtr<EnsInvJob> jobs = (ConstPtr<EnsInvJob>)jobs_;
  const ConstPtr<int> start = (ConstPtr<int>)start_;
  const int total = start[njob];
  for (;;) {
    const int item = next_item(counter, s_item);
    if (item >= total) break;
    const int j = __builtin_amdgcn_readfirstlane(job_of(start, njob, item));
    const EnsInvJob& jb = *(const EnsInvJob*)(jobs + j);
    const int blk = item - start[j];
    if (jb.ntru == 63)
      ens_inv_item<63, MODE, NT>(jb, blk, lds);
    else if (jb.ntru == 42)
      ens_inv_item<42, MODE, NT>(jb, blk, lds);
    else
      ens_inv_item<21, MODE, NT>(jb, blk, lds);
    __syncthreads();  // the next item reuses the LDS
  }
}

template <int T, int MODE, int NT>
__device__ __forceinline__ void ens_fwd_item(const EnsFwdJob& jb, int blk, char* lds) {
  if (NT == 512) {
    if (jb.var == 3)
      fwd_body<T, MODE, true, 3>(jb.a, blk, jb.tv, jb.tw, nullptr, lds);
    else
      fwd_body<T, MODE, true, 2>(jb.a, blk, jb.tv, jb.tw, nullptr, lds);
  } else if (jb.var)
    fwd_body<T, MODE, true, 1>(jb.a, blk, jb.tv, jb.tw, nullptr, lds);
  else
    fwd_body<T, MODE, true, 0>(jb.a, blk, jb.tv, jb.tw, nullptr, lds);
}

template <int MODE, int NT>
__global__ __launch_bounds__(NT) void k_ens_fwd(const EnsFwdJob* jobs_, const int* start_,
                                                int njob, int* counter) {
  extern __shared__ __align__(16) char lds[];
  __shared__ __align__(16) int s_item[4];  // keeps the dynamic LDS 16-byte aligned
  const ConstPtr<EnsFwdJob> jobs = (ConstPtr<EnsFwdJob>)jobs_;
  const ConstPtr<int> start = (ConstPtr<int>)start_;
  const int total = start[njob];
  for (;;) {
    const int item = next_item(counter, s_item);
    if (item >= total) break;
    const int j = __builtin_amdgcn_readfirstlane(job_of(start, njob, item));
    const EnsFwdJob& jb = *(const EnsFwdJob*)(jobs + j);
    const int blk = item - start[j];
    if (jb.ntru == 63)
      ens_fwd_item<63, MODE, NT>(jb, blk, lds);
    else if (jb.ntru == 42)
      ens_fwd_item<42, MODE, NT>(jb, blk, lds);
    else
      ens_fwd_item<21, MODE, NT>(jb, blk, lds);
    __syncthreads();
  }
}

// the work counters start at zero for every run (a kernel, not a memset,
// so that the run can be captured in a graph)
__global__ void k_ens_reset(int* counter) {
  if (threadIdx.x < 2) counter[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// host-side sizes per (truncation, mode, variant)
// ---------------------------------------------------------------------------
template <int MODE>
static bool ens_sizes(int ntru, int var, const InvArgs* ia, const FwdArgs* fa, long* items,
                      size_t* lds) {
#define EXO_ENS_CASE(TT)                                                           \
  if (ntru == TT) {                                                                \
    if (ia) {                                                                      \
      *items = var == 3 ? inv_items<TT, MODE, 3>(*ia)                              \
               : var == 2 ? inv_items<TT, MODE, 2>(*ia)                            \
               : var  ? inv_items<TT, MODE, 1>(*ia) : inv_items<TT, MODE, 0>(*ia); \
      *lds = var == 3 ? InvLds<TT, MODE, true, 3>::BYTES                           \
             : var == 2 ? InvLds<TT, MODE, true, 2>::BYTES                         \
             : var  ? InvLds<TT, MODE, true, 1>::BYTES : InvLds<TT, MODE, true, 0>::BYTES; \
    } else {                                                                       \
      if (FwdCfg<TT, MODE, 0>::NPB != 1 || FwdCfg<TT, MODE, 1>::NPB != 1 ||        \
          FwdCfg<TT, MODE, 2>::NPB != 1 || FwdCfg<TT, MODE, 3>::NPB != 1)          \
        return false;                                                              \
      *items = var == 3 ? fwd_items<TT, MODE, 3>(*fa)                              \
               : var == 2 ? fwd_items<TT, MODE, 2>(*fa)                            \
               : var  ? fwd_items<TT, MODE, 1>(*fa) : fwd_items<TT, MODE, 0>(*fa); \
      *lds = var == 3 ? FwdLds<TT, MODE, true, 3>::BYTES                           \
             : var == 2 ? FwdLds<TT, MODE, true, 2>::BYTES                         \
             : var  ? FwdLds<TT, MODE, true, 1>::BYTES : FwdLds<TT, MODE, true, 0>::BYTES; \
    }                                                                              \
    return true;                                                                   \
  }
  EXO_ENS_CASE(21)
  EXO_ENS_CASE(42)
  EXO_ENS_CASE(63)
#undef EXO_ENS_CASE
  return false;
}

static bool ens_sizes_mode(int mode, int ntru, int var, const InvArgs* ia, const FwdArgs* fa,
                           long* items, size_t* lds) {
  switch (mode) {
    case MODE_F32: return ens_sizes<MODE_F32>(ntru, var, ia, fa, items, lds);
    case MODE_F32MFMA: return ens_sizes<MODE_F32MFMA>(ntru, var, ia, fa, items, lds);
    case MODE_BF16X3: return ens_sizes<MODE_BF16X3>(ntru, var, ia, fa, items, lds);
    default: return false;
  }
}

template <int MODE>
static const void* ens_kernel_t(bool inv, bool wide) {
  if (inv)
    return wide ? (const void*)k_ens_inv<MODE, 512> : (const void*)k_ens_inv<MODE, 256>;
  return wide ? (const void*)k_ens_fwd<MODE, 512> : (const void*)k_ens_fwd<MODE, 256>;
}

static const void* ens_kernel(int mode, bool inv, bool wide) {
  switch (mode) {
    case MODE_F32: return ens_kernel_t<MODE_F32>(inv, wide);
    case MODE_F32MFMA: return ens_kernel_t<MODE_F32MFMA>(inv, wide);
    case MODE_BF16X3: return ens_kernel_t<MODE_BF16X3>(inv, wide);
    default: return nullptr;
  }
}

}  // namespace exo

using namespace exo;

struct exo_sht_ens_s {
  int mode = 0;
  int ngroup = 0;
  exo_sht_ens_group g[ENS_MAXGROUP];
  exo_sht_plan_t plan[ENS_MAXGROUP] = {};  // one per group (own tables)
  float* pk[ENS_MAXGROUP] = {};            // BF16X3: parity-ordered spectral input per group
  int inv_var[ENS_MAXGROUP] = {}, fwd_var[ENS_MAXGROUP] = {};
  // tile variants (inverse, inverse_uv, direct) of each group's batched
  // round trip; -1: those of the persistent kernels
  int bvar[ENS_MAXGROUP][3];
  // device state
  EnsInvJob* dinv = nullptr;
  EnsFwdJob* dfwd = nullptr;
  int* dstart = nullptr;  // [2][ENS_MAXJOB + 1]: first item of each job
  int* counter = nullptr;
  EnsIndex iinv{}, ifwd{};
  size_t lds_inv = 0, lds_fwd = 0;
  int grid_inv = 0, grid_fwd = 0;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // what the graph holds: the two persistent kernels (EXO_ENS_PERSISTENT)
  // or one batched round trip per group (EXO_ENS_BATCHED)
  int sched = EXO_ENS_PERSISTENT;
};

// the per-truncation kernels run the same variants
static int plan_var(int v) { return v; }

// whether the inverse (or direct) kernel runs 512-thread blocks (tile
// variants 2 and 3) for every group
static bool ens_wide(const exo_sht_ens_s* e, bool inv) {
  return e->ngroup > 0 && (inv ? e->inv_var[0] : e->fwd_var[0]) >= 2;
}

#define ENS_HIPCK(x)                                 \
  do {                                               \
    if ((x) != hipSuccess) return EXO_ERR_HIP;       \
  } while (0)

static int ens_grid(const void* kern, int nt, size_t lds, int total, int* grid) {
  int dev = 0, ncu = 0, nb = 0;
  ENS_HIPCK(hipGetDevice(&dev));
  ENS_HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  ENS_HIPCK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  ENS_HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, nt, lds));
  if (nb < 1) return EXO_ERR_UNSUPPORTED;
  *grid = std::max(1, std::min(total, nb * ncu));
  return EXO_OK;
}

// job tables, item index, LDS and grid sizes for the current variants
static int ens_prepare(exo_sht_ens_s* e) {
  struct Pending {
    long cost;
    EnsInvJob j;
    long items;
  };
  std::vector<Pending> inv;
  std::vector<std::pair<long, std::pair<EnsFwdJob, long>>> fwd;
  size_t lds_inv = 0, lds_fwd = 0;
  for (int k = 0; k < e->ngroup; ++k) {
    const exo_sht_ens_group& g = e->g[k];
    const Tables& tb = plan_tables(e->plan[k]);
    const int nlev = g.nlev, nfl = 3 * nlev + 1, ngp = nfl + 1 + 2 * nlev;
    const long sps = (long)nfl * 2 * tb.ncsp, gps = (long)ngp * tb.nlat * tb.nlon;
    const void* tv = tb.filt32;
    const void* tw = tb.tw32;
    for (int kind = 0; kind < 3; ++kind) {
      EnsInvJob j{};
      j.a.sp = g.sp_in;
      j.a.out = g.gp;
      j.a.sp_stride = sps;
      j.a.out_stride = gps;
      j.a.nmem = g.nmem;
      j.a.plavor = tb.plavor;
      j.a.var = e->inv_var[k];
      if (e->mode == MODE_BF16X3) {
        j.a.bpk = e->pk[k];
        j.a.bpk_stride = (long)nfl * pk_field(tb.ntru);
        j.a.bpk_f0 = 0;
      }
      int t0 = TQI;
      if (kind == 2) {
        j.a.kind = INV_UV;
        j.a.vor0 = 0;
        j.a.div0 = nlev;
        j.a.u0 = nfl + 1;
        j.a.v0 = nfl + 1 + nlev;
        j.a.nlev = nlev;
        t0 = TQV;
      } else {
        j.a.kind = INV_SCALAR;
        j.a.table = kind == 0 ? TQI : TQJ;
        j.a.src0 = kind == 0 ? 0 : nfl - 1;
        j.a.dst0 = kind == 0 ? 0 : nfl;
        j.a.nfield = kind == 0 ? nfl : 1;
        t0 = j.a.table;
      }
      if (e->mode == MODE_BF16X3) {
        j.A0 = tb.inv_bf[t0];
        j.A1 = tb.inv_bf[TQU];
      } else if (e->mode == MODE_F32MFMA) {
        j.A0 = tb.inv_f32[t0];
        j.A1 = tb.inv_f32[TQU];
      }
      j.tv = tv;
      j.tw = tw;
      j.plane = tb.inv_elems;
      j.ntru = tb.ntru;
      j.var = e->inv_var[k];
      long items = 0;
      size_t lds = 0;
      if (!ens_sizes_mode(e->mode, tb.ntru, j.var, &j.a, nullptr, &items, &lds))
        return EXO_ERR_UNSUPPORTED;
      lds_inv = std::max(lds_inv, lds);
      // cost of one item ~ Legendre terms x coefficients: expensive first
      const long cost = (long)tb.ncsp * (kind == 2 ? 2 : 1);
      if (items > 0) inv.push_back({cost, j, items});
    }
    EnsFwdJob f{};
    f.a.in = g.gp;
    f.a.sp = g.sp_out;
    f.a.in_stride = gps;
    f.a.sp_stride = sps;
    f.a.nmem = g.nmem;
    f.a.src0 = 0;
    f.a.dst0 = 0;
    f.a.nfield = nfl;
    f.a.table = TQC;
    f.a.var = e->fwd_var[k];
    f.tv = tb.fwd4;  // the direct sums' FP32 tables, 4 pairs side by side
    f.tw = tw;
    f.ntru = tb.ntru;
    f.var = e->fwd_var[k];
    long items = 0;
    size_t lds = 0;
    if (!ens_sizes_mode(e->mode, tb.ntru, f.var, nullptr, &f.a, &items, &lds))
      return EXO_ERR_UNSUPPORTED;
    lds_fwd = std::max(lds_fwd, lds);
    if (items > 0) fwd.push_back({(long)tb.ncsp, {f, items}});
  }
  std::stable_sort(inv.begin(), inv.end(),
                   [](const Pending& x, const Pending& y) { return x.cost > y.cost; });
  std::stable_sort(fwd.begin(), fwd.end(),
                   [](const auto& x, const auto& y) { return x.first > y.first; });
  std::vector<EnsInvJob> hi;
  std::vector<EnsFwdJob> hf;
  EnsIndex ii{}, fi{};
  long at = 0;
  for (auto& p : inv) {
    ii.start[ii.njob++] = (int)at;
    hi.push_back(p.j);
    at += p.items;
  }
  ii.start[ii.njob] = (int)at;
  if (at > 0x7fffffffL) return EXO_ERR_ARG;
  at = 0;
  for (auto& p : fwd) {
    fi.start[fi.njob++] = (int)at;
    hf.push_back(p.second.first);
    at += p.second.second;
  }
  fi.start[fi.njob] = (int)at;
  if (at > 0x7fffffffL) return EXO_ERR_ARG;
  if (!hi.empty())
    ENS_HIPCK(hipMemcpy(e->dinv, hi.data(), sizeof(EnsInvJob) * hi.size(), hipMemcpyHostToDevice));
  if (!hf.empty())
    ENS_HIPCK(hipMemcpy(e->dfwd, hf.data(), sizeof(EnsFwdJob) * hf.size(), hipMemcpyHostToDevice));
  ENS_HIPCK(hipMemcpy(e->dstart, ii.start, sizeof(int) * (ii.njob + 1), hipMemcpyHostToDevice));
  ENS_HIPCK(hipMemcpy(e->dstart + ENS_MAXJOB + 1, fi.start, sizeof(int) * (fi.njob + 1),
                      hipMemcpyHostToDevice));
  e->iinv = ii;
  e->ifwd = fi;
  e->lds_inv = lds_inv;
  e->lds_fwd = lds_fwd;
  int rc = EXO_OK;
  if (ii.start[ii.njob] > 0)
    rc = ens_grid(ens_kernel(e->mode, true, ens_wide(e, true)), ens_wide(e, true) ? 512 : 256,
                  lds_inv, ii.start[ii.njob], &e->grid_inv);
  if (rc == EXO_OK && fi.start[fi.njob] > 0)
    rc = ens_grid(ens_kernel(e->mode, false, ens_wide(e, false)), ens_wide(e, false) ? 512 : 256,
                  lds_fwd, fi.start[fi.njob], &e->grid_fwd);
  return rc;
}

template <int MODE>
static void ens_launch_t(const exo_sht_ens_s* e, bool inv, hipStream_t s) {
  const bool wide = ens_wide(e, inv);
  if (inv) {
    if (wide)
      k_ens_inv<MODE, 512><<<e->grid_inv, 512, e->lds_inv, s>>>(e->dinv, e->dstart, e->iinv.njob,
                                                                e->counter);
    else
      k_ens_inv<MODE, 256><<<e->grid_inv, 256, e->lds_inv, s>>>(e->dinv, e->dstart, e->iinv.njob,
                                                                e->counter);
  } else {
    const int* st = e->dstart + ENS_MAXJOB + 1;
    if (wide)
      k_ens_fwd<MODE, 512><<<e->grid_fwd, 512, e->lds_fwd, s>>>(e->dfwd, st, e->ifwd.njob,
                                                                e->counter + 1);
    else
      k_ens_fwd<MODE, 256><<<e->grid_fwd, 256, e->lds_fwd, s>>>(e->dfwd, st, e->ifwd.njob,
                                                                e->counter + 1);
  }
}

static int ens_launch(exo_sht_ens_s* e, bool inv, hipStream_t s) {
  const EnsIndex& ix = inv ? e->iinv : e->ifwd;
  if (ix.njob == 0 || ix.start[ix.njob] == 0) return EXO_OK;
  switch (e->mode) {
    case MODE_F32: ens_launch_t<MODE_F32>(e, inv, s); break;
    case MODE_F32MFMA: ens_launch_t<MODE_F32MFMA>(e, inv, s); break;
    case MODE_BF16X3: ens_launch_t<MODE_BF16X3>(e, inv, s); break;
    default: return EXO_ERR_UNSUPPORTED;
  }
  ENS_HIPCK(hipGetLastError());
  return EXO_OK;
}

static int ens_launch_inv(exo_sht_ens_s* e, hipStream_t s) { return ens_launch(e, true, s); }
static int ens_launch_fwd(exo_sht_ens_s* e, hipStream_t s) { return ens_launch(e, false, s); }

// BF16X3: the pre-split spectral input of every group
static int ens_pack(exo_sht_ens_s* e, hipStream_t s) {
  if (e->mode != MODE_BF16X3) return EXO_OK;
  for (int k = 0; k < e->ngroup; ++k) {
    const exo_sht_ens_group& g = e->g[k];
    const Tables& tb = plan_tables(e->plan[k]);
    const int nfl = 3 * g.nlev + 1;
    if (launch_pack(tb, (const float*)g.sp_in, (long)nfl * 2 * tb.ncsp, 0, nfl, g.nmem, e->pk[k],
                    s) != 0)
      return EXO_ERR_LAUNCH;
  }
  return EXO_OK;
}

// the tile variants of group k's per-truncation kernels: its own batched
// variants once the autotuner has chosen them, else the persistent kernels'
static int ens_plan_tiles(exo_sht_ens_s* e, int k) {
  if (e->bvar[k][0] >= 0)
    return exo_sht_set_tiles(e->plan[k], e->mode, e->bvar[k][0], e->bvar[k][1], e->bvar[k][2]);
  return exo_sht_set_tiles(e->plan[k], e->mode, plan_var(e->inv_var[k]), plan_var(e->inv_var[k]),
                           plan_var(e->fwd_var[k]));
}

// one batched round trip per group, with the group's tile variants
static int ens_groups(exo_sht_ens_s* e, hipStream_t s) {
  for (int k = 0; k < e->ngroup; ++k) {
    const exo_sht_ens_group& g = e->g[k];
    int rc = ens_plan_tiles(e, k);
    if (rc == EXO_OK)
      rc = exo_sht_roundtrip(e->plan[k], e->mode, g.nmem, g.nlev, g.sp_in, g.gp, g.sp_out, s);
    if (rc != EXO_OK) return rc;
  }
  return EXO_OK;
}

// the whole ensemble round trip in the chosen schedule: persistent, reset
// the counters, pre-split the inputs, all inverse items, all direct items;
// batched, one round trip per group
static int ens_enqueue(exo_sht_ens_s* e, hipStream_t s) {
  if (e->sched == EXO_ENS_BATCHED) return ens_groups(e, s);
  k_ens_reset<<<1, 64, 0, s>>>(e->counter);
  ENS_HIPCK(hipGetLastError());
  int rc = ens_pack(e, s);
  if (rc == EXO_OK) rc = ens_launch_inv(e, s);
  if (rc == EXO_OK) rc = ens_launch_fwd(e, s);
  return rc;
}

static void ens_drop_graph(exo_sht_ens_s* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

static int ens_capture(exo_sht_ens_s* e) {
  ens_drop_graph(e);
  if (e->sched == EXO_ENS_BATCHED) {
    // one eager pass first: the plans size their scratch on first use,
    // which must not happen inside the capture
    const int rc0 = ens_groups(e, e->cap);
    if (rc0 != EXO_OK) return rc0;
    ENS_HIPCK(hipStreamSynchronize(e->cap));
  }
  ENS_HIPCK(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
  const int rc = ens_enqueue(e, e->cap);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->cap, &g);
  if (rc != EXO_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  ENS_HIPCK(ec);
  e->graph = g;
  ENS_HIPCK(hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
  return EXO_OK;
}

static void ens_free(exo_sht_ens_s* e) {
  ens_drop_graph(e);
  if (e->cap) (void)hipStreamDestroy(e->cap);
  if (e->dinv) (void)hipFree(e->dinv);
  if (e->dfwd) (void)hipFree(e->dfwd);
  if (e->dstart) (void)hipFree(e->dstart);
  for (int k = 0; k < e->ngroup; ++k)
    if (e->pk[k]) (void)hipFree(e->pk[k]);
  if (e->counter) (void)hipFree(e->counter);
  for (int k = 0; k < e->ngroup; ++k)
    if (e->plan[k]) exo_sht_plan_destroy(e->plan[k]);
  delete e;
}

extern "C" int exo_sht_ens_create(exo_sht_ens_t* ens, int mode, int ngroup,
                                  const exo_sht_ens_group* groups) {
  if (!ens) return EXO_ERR_ARG;
  *ens = nullptr;
  if (ngroup < 1 || ngroup > ENS_MAXGROUP || !groups) return EXO_ERR_ARG;
  if (mode != EXO_F32 && mode != EXO_F32MFMA && mode != EXO_BF16X3) return EXO_ERR_UNSUPPORTED;
  for (int k = 0; k < ngroup; ++k) {
    const exo_sht_ens_group& g = groups[k];
    if (g.ntru != 21 && g.ntru != 42 && g.ntru != 63) return EXO_ERR_UNSUPPORTED;
    if (g.nlev < 1 || g.nmem < 0 || (g.nmem > 0 && (!g.sp_in || !g.gp || !g.sp_out)))
      return EXO_ERR_ARG;
  }
  exo_sht_ens_s* e = new exo_sht_ens_s();
  e->mode = mode;
  e->ngroup = ngroup;
  for (int k = 0; k < ENS_MAXGROUP; ++k) e->bvar[k][0] = e->bvar[k][1] = e->bvar[k][2] = -1;
  int rc = EXO_OK;
  for (int k = 0; k < ngroup && rc == EXO_OK; ++k) {
    e->g[k] = groups[k];
    rc = exo_sht_plan_create(&e->plan[k], groups[k].ntru, nullptr, nullptr, nullptr, nullptr,
                             1.63299310207);
    if (rc == EXO_OK && mode == EXO_BF16X3 && groups[k].nmem > 0 &&
        hipMalloc(&e->pk[k], pack_bytes(plan_tables(e->plan[k]), groups[k].nmem,
                                        3 * groups[k].nlev + 1)) != hipSuccess)
      rc = EXO_ERR_HIP;
  }
  if (rc == EXO_OK &&
      (hipMalloc(&e->dinv, sizeof(EnsInvJob) * ENS_MAXJOB) != hipSuccess ||
       hipMalloc(&e->dfwd, sizeof(EnsFwdJob) * ENS_MAXGROUP) != hipSuccess ||
       hipMalloc(&e->dstart, sizeof(int) * 2 * (ENS_MAXJOB + 1)) != hipSuccess ||
       hipMalloc(&e->counter, 2 * sizeof(int)) != hipSuccess ||
       hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking) != hipSuccess))
    rc = EXO_ERR_HIP;
  if (rc == EXO_OK) rc = ens_prepare(e);
  if (rc == EXO_OK) rc = ens_capture(e);
  if (rc != EXO_OK) {
    (void)hipDeviceSynchronize();
    ens_free(e);
    return rc;
  }
  *ens = e;
  return EXO_OK;
}

extern "C" int exo_sht_ens_destroy(exo_sht_ens_t e) {
  if (!e) return EXO_ERR_ARG;
  (void)hipDeviceSynchronize();
  ens_free(e);
  return EXO_OK;
}

extern "C" int exo_sht_ens_run(exo_sht_ens_t e, void* stream) {
  if (!e || !e->exec) return EXO_ERR_ARG;
  ENS_HIPCK(hipGraphLaunch(e->exec, (hipStream_t)stream));
  return EXO_OK;
}

extern "C" int exo_sht_ens_run_eager(exo_sht_ens_t e, void* stream) {
  if (!e) return EXO_ERR_ARG;
  return ens_enqueue(e, (hipStream_t)stream);
}

// the same round trips as one batched round trip per group (4 launches each)
extern "C" int exo_sht_ens_run_groups(exo_sht_ens_t e, void* stream) {
  if (!e) return EXO_ERR_ARG;
  return ens_groups(e, (hipStream_t)stream);
}

extern "C" int exo_sht_ens_get_schedule(exo_sht_ens_t e, int* sched) {
  if (!e || !sched) return EXO_ERR_ARG;
  *sched = e->sched;
  return EXO_OK;
}

extern "C" int exo_sht_ens_set_schedule(exo_sht_ens_t e, int sched) {
  if (!e || (sched != EXO_ENS_PERSISTENT && sched != EXO_ENS_BATCHED)) return EXO_ERR_ARG;
  (void)hipDeviceSynchronize();
  e->sched = sched;
  return ens_capture(e);
}

// the same round trips one member at a time (4 launches per member)
extern "C" int exo_sht_ens_run_members(exo_sht_ens_t e, void* stream) {
  if (!e) return EXO_ERR_ARG;
  for (int k = 0; k < e->ngroup; ++k) {
    const exo_sht_ens_group& g = e->g[k];
    int dims[4];
    int rc = exo_sht_plan_dims(e->plan[k], dims);
    if (rc == EXO_OK) rc = ens_plan_tiles(e, k);
    if (rc != EXO_OK) return rc;
    const int nfl = 3 * g.nlev + 1, ngp = nfl + 1 + 2 * g.nlev;
    const long sps = (long)nfl * 2 * dims[3], gps = (long)ngp * dims[1] * dims[2];
    for (int m = 0; m < g.nmem; ++m) {
      const float* si = (const float*)g.sp_in + m * sps;
      float* go = (float*)g.gp + m * gps;
      float* so = (float*)g.sp_out + m * sps;
      rc = exo_sht_roundtrip(e->plan[k], e->mode, 1, g.nlev, si, go, so, stream);
      if (rc != EXO_OK) return rc;
    }
  }
  return EXO_OK;
}

extern "C" int exo_sht_ens_get_tiles(exo_sht_ens_t e, int* vars) {
  if (!e || !vars) return EXO_ERR_ARG;
  for (int k = 0; k < e->ngroup; ++k) {
    vars[2 * k] = e->inv_var[k];
    vars[2 * k + 1] = e->fwd_var[k];
  }
  return EXO_OK;
}

extern "C" int exo_sht_ens_set_tiles(exo_sht_ens_t e, const int* vars) {
  if (!e || !vars) return EXO_ERR_ARG;
  for (int k = 0; k < 2 * e->ngroup; ++k)
    if (vars[k] < 0 || vars[k] > 3) return EXO_ERR_ARG;
  // variants 2 and 3 (512-thread blocks) hold for all groups of a kernel or
  // for none
  for (int w = 0; w < 2; ++w) {
    int n2 = 0;
    for (int k = 0; k < e->ngroup; ++k) n2 += vars[2 * k + w] >= 2;
    if (n2 != 0 && n2 != e->ngroup) return EXO_ERR_ARG;
  }
  (void)hipDeviceSynchronize();
  ens_drop_graph(e);  // stale once the job tables change
  for (int k = 0; k < e->ngroup; ++k) {
    e->inv_var[k] = vars[2 * k];
    e->fwd_var[k] = vars[2 * k + 1];
    e->bvar[k][0] = e->bvar[k][1] = e->bvar[k][2] = -1;  // both schedules run these
  }
  int rc = ens_prepare(e);
  if (rc == EXO_OK) rc = ens_capture(e);
  return rc;
}

// Times each combination of tile variants per truncation (the groups of one
// truncation share a variant: 0 or 1 with 256-thread blocks, 2 or 3 with
// 512-thread blocks) for the inverse and the direct kernel of the
// persistent schedule, keeps the fastest, lets each group's plan choose the
// variants of its own batched round trip (exo_sht_autotune), then times the
// two schedules and captures the faster in the graph.
extern "C" int exo_sht_ens_autotune(exo_sht_ens_t e, void* stream, int* vars) {
  if (!e) return EXO_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  (void)hipStreamSynchronize(s);
  ens_drop_graph(e);  // stale once the job tables change
  int truncs[3] = {21, 42, 63};
  hipEvent_t e0, e1;
  ENS_HIPCK(hipEventCreate(&e0));
  ENS_HIPCK(hipEventCreate(&e1));
  int rc = EXO_OK;
  float best[2] = {1e30f, 1e30f};
  int pick[2] = {0, 0};
  int rc0 = ens_pack(e, s);  // the inverse timings read the pre-split inputs
  if (rc0 != EXO_OK) rc = rc0;
  // combinations 0-7: variant 0 or 1 per truncation (256-thread blocks);
  // 8-15: variant 2 or 3 per truncation (512-thread blocks)
  for (int combo = 0; combo < 16 && rc == EXO_OK; ++combo) {
    (void)hipStreamSynchronize(s);
    for (int k = 0; k < e->ngroup; ++k) {
      int bit = 0;
      for (int t = 0; t < 3; ++t)
        if (e->g[k].ntru == truncs[t]) bit = (combo >> t) & 1;
      e->inv_var[k] = e->fwd_var[k] = (combo >> 3) * 2 + bit;
    }
    rc = ens_prepare(e);
    for (int which = 0; which < 2 && rc == EXO_OK; ++which) {
      float tmin = 1e30f;
      for (int rep = 0; rep < 4 && rc == EXO_OK; ++rep) {
        k_ens_reset<<<1, 64, 0, s>>>(e->counter);
        (void)hipEventRecord(e0, s);
        rc = which == 0 ? ens_launch_inv(e, s) : ens_launch_fwd(e, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) tmin = std::min(tmin, ms);
      }
      if (tmin < best[which]) {
        best[which] = tmin;
        pick[which] = combo;
      }
    }
  }
  for (int k = 0; k < e->ngroup; ++k) {
    int bi = 0, bf = 0;
    for (int t = 0; t < 3; ++t)
      if (e->g[k].ntru == truncs[t]) {
        bi = (pick[0] >> t) & 1;
        bf = (pick[1] >> t) & 1;
      }
    e->inv_var[k] = (pick[0] >> 3) * 2 + bi;
    e->fwd_var[k] = (pick[1] >> 3) * 2 + bf;
  }
  if (rc == EXO_OK) rc = ens_prepare(e);
  for (int k = 0; k < e->ngroup && rc == EXO_OK; ++k)
    rc = exo_sht_autotune(e->plan[k], e->mode, e->g[k].nmem, e->g[k].nlev, stream, e->bvar[k]);
  // the whole round trip in each schedule (the first pass of each sizes the
  // plans' scratch and is not timed)
  float tsched[2] = {1e30f, 1e30f};
  for (int sc = 0; sc < 2 && rc == EXO_OK; ++sc) {
    e->sched = sc;
    for (int rep = 0; rep < 4 && rc == EXO_OK; ++rep) {
      (void)hipEventRecord(e0, s);
      rc = ens_enqueue(e, s);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0) tsched[sc] = std::min(tsched[sc], ms);
    }
  }
  e->sched = tsched[EXO_ENS_BATCHED] < tsched[EXO_ENS_PERSISTENT] ? EXO_ENS_BATCHED
                                                                   : EXO_ENS_PERSISTENT;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == EXO_OK) rc = ens_capture(e);
  if (rc == EXO_OK && vars) exo_sht_ens_get_tiles(e, vars);
  return rc;
}
