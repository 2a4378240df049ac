// sht_capi.hip -- C ABI of libexosht.so (declared in include/exo_sht.h).
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/exo_sht.h"
#include "sht_plan.h"

using namespace exo;

#define HIPCK(x)                        \
  do {                                  \
    if ((x) != hipSuccess) return EXO_ERR_HIP; \
  } while (0)

const Tables& exo::plan_tables(const exo_sht_plan_s* p) { return p->tb; }

static int nlat_of(int ntru) {
  switch (ntru) {
    case 21: return 32;
    case 42: return 64;
    case 63: return 96;
    case 127: return 192;
    default: return -1;
  }
}

extern "C" const char* exo_strerror(int code) {
  switch (code) {
    case EXO_OK: return "ok";
    case EXO_ERR_ARG: return "invalid argument";
    case EXO_ERR_UNSUPPORTED: return "unsupported truncation or length";
    case EXO_ERR_LAUNCH: return "kernel launch failed";
    case EXO_ERR_SCRATCH: return "scratch buffer too small";
    case EXO_ERR_HIP: return "HIP runtime error";
    case EXO_ERR_NOINIT: return "exo_legini has not been called";
    default: return "unknown error";
  }
}

extern "C" int exo_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

static void free_tables(Tables& t) {
  void* ps[] = {t.sid, t.gwd, t.tab, t.filt64, t.filt32, t.fwd4, t.skgpsp, t.skspgp, t.tw64, t.tw32};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  for (int i = 0; i < NTAB; ++i) {
    if (t.inv_bf[i]) (void)hipFree(t.inv_bf[i]);
    if (t.inv_f32[i]) (void)hipFree(t.inv_f32[i]);
    if (t.inv_f64[i]) (void)hipFree(t.inv_f64[i]);
    if (t.dir_f32[i]) (void)hipFree(t.dir_f32[i]);
  }
  t = Tables();
}

static int build_tables(Tables& t, int ntru, const double* sid, const double* gwd,
                        const double* skgpsp, const double* skspgp, double plavor) {
  const int nlat = nlat_of(ntru);
  if (nlat < 0) return EXO_ERR_UNSUPPORTED;
  t.ntru = ntru;
  t.nlat = nlat;
  t.nlon = 2 * nlat;
  t.npair = nlat / 2;
  t.ncsp = (ntru + 1) * (ntru + 2) / 2;
  t.kbmax = (ntru + 1 + 31) / 32;
  t.plavor = plavor;
  const long tsz = (long)t.npair * t.ncsp;
  HIPCK(hipMalloc(&t.sid, sizeof(double) * nlat));
  HIPCK(hipMalloc(&t.gwd, sizeof(double) * nlat));
  HIPCK(hipMalloc(&t.tab, sizeof(double) * NTAB * tsz));
  HIPCK(hipMalloc(&t.filt64, sizeof(double) * NTAB * tsz));
  HIPCK(hipMalloc(&t.filt32, sizeof(float) * NTAB * tsz));
  HIPCK(hipMalloc(&t.fwd4, sizeof(float) * NTAB * tsz));
  HIPCK(hipMalloc(&t.skgpsp, sizeof(double) * (ntru + 1)));
  HIPCK(hipMalloc(&t.skspgp, sizeof(double) * (ntru + 1)));
  HIPCK(hipMalloc(&t.tw64, sizeof(double) * 2 * t.nlon));
  HIPCK(hipMalloc(&t.tw32, sizeof(float) * 2 * t.nlon));
  if (sid && gwd) {
    HIPCK(hipMemcpy(t.sid, sid, sizeof(double) * nlat, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(t.gwd, gwd, sizeof(double) * nlat, hipMemcpyHostToDevice));
  } else {
    k_inigau<<<(nlat / 2 + 63) / 64, 64>>>(nlat, t.sid, t.gwd);
    HIPCK(hipGetLastError());
  }
  std::vector<double> ones(ntru + 1, 1.0);
  HIPCK(hipMemcpy(t.skgpsp, skgpsp ? skgpsp : ones.data(), sizeof(double) * (ntru + 1),
                  hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(t.skspgp, skspgp ? skspgp : ones.data(), sizeof(double) * (ntru + 1),
                  hipMemcpyHostToDevice));
  dim3 lg((t.npair + 63) / 64, ntru + 1);
  k_legini<<<lg, 64>>>(ntru, t.npair, t.sid, t.gwd, t.tab);
  HIPCK(hipGetLastError());
  const long nf = NTAB * tsz;
  k_filtered<<<(unsigned)((nf + 255) / 256), 256>>>(ntru, t.npair, t.tab, t.skgpsp, t.skspgp,
                                                    t.filt64, t.filt32);
  HIPCK(hipGetLastError());
  k_pack_fwd4<<<(unsigned)((nf + 255) / 256), 256>>>(t.npair, t.ncsp, t.filt32, t.fwd4);
  HIPCK(hipGetLastError());
  t.inv_elems = (long)(t.npair / 8) * (ntru + 1) * t.kbmax * 512;
  const int inv_tabs[] = {TQI, TQJ, TQU, TQV};
  for (int tt : inv_tabs) {
    const long nbf = (long)(t.npair / 8) * tab_pb(ntru);
    HIPCK(hipMalloc(&t.inv_bf[tt], sizeof(uint16_t) * nbf));
    HIPCK(hipMalloc(&t.inv_f32[tt], sizeof(float) * t.inv_elems));
    HIPCK(hipMalloc(&t.inv_f64[tt], sizeof(double) * t.inv_elems));
    k_pack_inv<<<(unsigned)((t.inv_elems + 255) / 256), 256>>>(
        ntru, t.npair, t.kbmax, tt, t.filt64, t.inv_f32[tt], t.inv_f64[tt]);
    HIPCK(hipGetLastError());
    const long nth = nbf / 3;
    k_pack_inv_bf<<<(unsigned)((nth + 255) / 256), 256>>>(ntru, t.npair, tt, t.filt64,
                                                          t.inv_bf[tt]);
    HIPCK(hipGetLastError());
  }
  t.dir_rb = ((ntru + 2) / 2 + 15) / 16;
  const int dir_tabs[] = {TQC, TQE, TQM, TQQ};
  for (int tt : dir_tabs) {
    const long nd = (long)(ntru + 1) * 2 * t.dir_rb * t.npair * 16;
    HIPCK(hipMalloc(&t.dir_f32[tt], sizeof(float) * nd));
    k_pack_dir<<<(unsigned)((nd + 255) / 256), 256>>>(ntru, t.npair, t.dir_rb, tt, t.filt64,
                                                      t.dir_f32[tt]);
    HIPCK(hipGetLastError());
  }
  k_twiddle<<<(t.nlon + 255) / 256, 256>>>(t.nlon, t.tw64, t.tw32);
  HIPCK(hipGetLastError());
  HIPCK(hipDeviceSynchronize());
  return EXO_OK;
}

// full [8][nlat][ncsp] host copy of the tables, southern half by symmetry
static int copy_tables(const Tables& t, double* out) {
  const long tsz = (long)t.npair * t.ncsp;
  std::vector<double> h(NTAB * tsz);
  HIPCK(hipMemcpy(h.data(), t.tab, sizeof(double) * NTAB * tsz, hipMemcpyDeviceToHost));
  for (int tt = 0; tt < NTAB; ++tt)
    for (int lat = 0; lat < t.nlat; ++lat) {
      const bool south = lat >= t.npair;
      const int pl = south ? t.nlat - 1 - lat : lat;
      for (int m = 0, w = 0; m <= t.ntru; ++m)
        for (int n = m; n <= t.ntru; ++n, ++w) {
          double v = h[tt * tsz + (long)pl * t.ncsp + w];
          if (south) {
            const bool odd = ((n - m) & 1) != 0;
            if (odd != table_is_q(tt)) v = -v;
          }
          out[((long)tt * t.nlat + lat) * t.ncsp + w] = v;
        }
    }
  return EXO_OK;
}

int plan_grow(void** p, size_t* have, size_t need) {
  if (*have >= need) return EXO_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  HIPCK(hipMalloc(p, need));
  *have = need;
  return EXO_OK;
}

// ---------------------------------------------------------------------------
// the no-allocation round trip's scratch
// ---------------------------------------------------------------------------
// Device memory of exo_sht_roundtrip_noalloc: per chunk of members, the
// pre-split operand of the split-BF16 inverse, then the partial direct sums
// of truncations whose direct transform runs in several latitude-pair
// blocks.  Each device holds its own copy of the array, from the loading of
// this library's code on it until the process ends.
static __device__ __align__(256) char scratch[EXO_NOALLOC_SCRATCH_BYTES];

// The array's address on each device, taken once per device with that
// device current.  In host code the name `scratch` is only a host-side
// stand-in for the symbol; kernels and copies get these addresses.
static constexpr int MAX_DEVICES = 64;
static char* g_scratch_dev[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

static int scratch_address(char** out) {
  *out = nullptr;
  int dev = -1;
  HIPCK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return EXO_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    HIPCK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return EXO_ERR_HIP;
    g_scratch_dev[dev] = (char*)ptr;
  }
  *out = g_scratch_dev[dev];
  return EXO_OK;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
extern "C" int exo_sht_plan_create(exo_sht_plan_t* plan, int ntru, const double* sid,
                                   const double* gwd, const double* skgpsp,
                                   const double* skspgp, double plavor) {
  if (!plan) return EXO_ERR_ARG;
  *plan = nullptr;
  exo_sht_plan_s* p = new exo_sht_plan_s();
  int rc = build_tables(p->tb, ntru, sid, gwd, skgpsp, skspgp, plavor);
  // the no-allocation scratch of this device (the plan's tables live here);
  // a failure is reported by exo_sht_roundtrip_noalloc, which needs it
  if (rc == EXO_OK) (void)scratch_address(&p->na_scratch);
  if (rc != EXO_OK) {
    free_tables(p->tb);
    delete p;
    return rc;
  }
  *plan = p;
  return EXO_OK;
}

extern "C" int exo_sht_plan_destroy(exo_sht_plan_t p) {
  if (!p) return EXO_ERR_ARG;
  (void)hipDeviceSynchronize();
  free_tables(p->tb);
  if (p->scratch) (void)hipFree(p->scratch);
  if (p->stage) (void)hipFree(p->stage);
  if (p->pk) (void)hipFree(p->pk);
  delete p;
  return EXO_OK;
}

extern "C" int exo_sht_plan_dims(exo_sht_plan_t p, int* dims) {
  if (!p || !dims) return EXO_ERR_ARG;
  dims[0] = p->tb.ntru;
  dims[1] = p->tb.nlat;
  dims[2] = p->tb.nlon;
  dims[3] = p->tb.ncsp;
  return EXO_OK;
}

extern "C" int exo_sht_plan_grid(exo_sht_plan_t p, double* sid, double* gwd) {
  if (!p) return EXO_ERR_ARG;
  if (sid) HIPCK(hipMemcpy(sid, p->tb.sid, sizeof(double) * p->tb.nlat, hipMemcpyDeviceToHost));
  if (gwd) HIPCK(hipMemcpy(gwd, p->tb.gwd, sizeof(double) * p->tb.nlat, hipMemcpyDeviceToHost));
  return EXO_OK;
}

extern "C" int exo_sht_plan_tables(exo_sht_plan_t p, double* tables) {
  if (!p || !tables) return EXO_ERR_ARG;
  return copy_tables(p->tb, tables);
}

static bool mode_ok(int mode) { return mode >= EXO_F64 && mode <= EXO_BF16X3; }

// MODE_BF16X3: pre-split the fp32 spectral fields [f0, f0 + nf) of every
// member into the plan's pack buffer and point a at it
// (buf: a caller-given buffer of at least pack_bytes(nmem, nf) bytes, or
// null for the plan's own, grown as needed)
static int pack_for(exo_sht_plan_t p, int mode, int nmem, const void* sp, long sps, int f0,
                    int nf, hipStream_t s, InvArgs& a, void* buf = nullptr) {
  a.bpk = nullptr;
  a.bpk_stride = 0;
  a.bpk_f0 = 0;
  if (mode != MODE_BF16X3 || nmem == 0 || nf == 0) return EXO_OK;
  if (!buf) {
    int rc = plan_grow(&p->pk, &p->pk_bytes, pack_bytes(p->tb, nmem, nf));
    if (rc != EXO_OK) return rc;
    buf = p->pk;
  }
  if (launch_pack(p->tb, (const float*)sp, sps, f0, nf, nmem, (float*)buf, s) != 0)
    return EXO_ERR_LAUNCH;
  a.bpk = buf;
  a.bpk_stride = (long)nf * pk_field(p->tb.ntru);
  a.bpk_f0 = f0;
  return EXO_OK;
}

// pk: an existing pack of the fields (round trip), or null to pack here
static int inverse_impl(exo_sht_plan_t p, int mode, int table, int nmem, const void* sp,
                        long sp_stride, int src0, void* gp, long gp_stride, int dst0, int nfield,
                        hipStream_t s, const InvArgs* pk, bool fused = true) {
  InvArgs a{};
  a.sp = sp;
  a.out = gp;
  a.sp_stride = sp_stride;
  a.out_stride = gp_stride;
  a.nmem = nmem;
  a.kind = INV_SCALAR;
  a.table = table;
  a.src0 = src0;
  a.dst0 = dst0;
  a.nfield = nfield;
  a.plavor = p->tb.plavor;
  a.var = p->inv_var[mode];
  if (pk) {
    a.bpk = pk->bpk;
    a.bpk_stride = pk->bpk_stride;
    a.bpk_f0 = pk->bpk_f0;
  } else {
    const int rc = pack_for(p, mode, nmem, sp, sp_stride, src0, nfield, s, a);
    if (rc != EXO_OK) return rc;
  }
  return launch_inverse(p->tb, mode, fused, a, s);
}

static int inverse_uv_impl(exo_sht_plan_t p, int mode, int nmem, const void* sp, long sp_stride,
                           int vor0, int div0, void* gp, long gp_stride, int u0, int v0, int nlev,
                           hipStream_t s, const InvArgs* pk, bool fused = true) {
  InvArgs a{};
  a.sp = sp;
  a.out = gp;
  a.sp_stride = sp_stride;
  a.out_stride = gp_stride;
  a.nmem = nmem;
  a.kind = INV_UV;
  a.vor0 = vor0;
  a.div0 = div0;
  a.u0 = u0;
  a.v0 = v0;
  a.nlev = nlev;
  a.plavor = p->tb.plavor;
  a.var = p->uv_var[mode];
  if (pk) {
    a.bpk = pk->bpk;
    a.bpk_stride = pk->bpk_stride;
    a.bpk_f0 = pk->bpk_f0;
  } else {
    // one pack covering both runs of levels
    const int lo = vor0 < div0 ? vor0 : div0;
    const int hi = (vor0 > div0 ? vor0 : div0) + nlev;
    const int rc = pack_for(p, mode, nmem, sp, sp_stride, lo, hi - lo, s, a);
    if (rc != EXO_OK) return rc;
  }
  return launch_inverse(p->tb, mode, fused, a, s);
}

extern "C" int exo_sht_inverse(exo_sht_plan_t p, int mode, int table, int nmem, const void* sp,
                               long sp_stride, int src0, void* gp, long gp_stride, int dst0,
                               int nfield, void* stream) {
  if (!p || !mode_ok(mode) || (table != EXO_QI && table != EXO_QJ) || nmem < 0 || nfield < 0 ||
      src0 < 0 || dst0 < 0 || (nmem > 0 && nfield > 0 && (!sp || !gp)))
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  return inverse_impl(p, mode, table, nmem, sp, sp_stride, src0, gp, gp_stride, dst0, nfield,
                      (hipStream_t)stream, nullptr);
}

extern "C" int exo_sht_inverse_uv(exo_sht_plan_t p, int mode, int nmem, const void* sp,
                                  long sp_stride, int vor0, int div0, void* gp, long gp_stride,
                                  int u0, int v0, int nlev, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 0 || vor0 < 0 || div0 < 0 ||
      (nmem > 0 && nlev > 0 && (!sp || !gp)))
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  return inverse_uv_impl(p, mode, nmem, sp, sp_stride, vor0, div0, gp, gp_stride, u0, v0, nlev,
                         (hipStream_t)stream, nullptr);
}

// (scr, sbytes: a caller-given buffer for the partial sums of multi-block
// truncations, or null for the plan's own, grown as needed)
static int direct_impl(exo_sht_plan_t p, int mode, bool fused, int nmem, const void* in,
                       long in_stride, int src0, void* sp, long sp_stride, int dst0, int nfield,
                       hipStream_t s, void* scr = nullptr, size_t sbytes = 0) {
  FwdArgs a{};
  a.in = in;
  a.sp = sp;
  a.in_stride = in_stride;
  a.sp_stride = sp_stride;
  a.nmem = nmem;
  a.src0 = src0;
  a.dst0 = dst0;
  a.nfield = nfield;
  a.table = TQC;
  a.var = fused ? p->fwd_var[mode] : 0;
  if (scr) return launch_forward(p->tb, mode, fused, a, scr, sbytes, s);
  const size_t need = forward_scratch_bytes(p->tb, mode, nmem, nfield);
  if (need > 0) {
    int rc = plan_grow(&p->scratch, &p->scratch_bytes, need);
    if (rc != EXO_OK) return rc;
  }
  return launch_forward(p->tb, mode, fused, a, p->scratch, p->scratch_bytes, s);
}

extern "C" int exo_sht_direct(exo_sht_plan_t p, int mode, int nmem, const void* gp,
                              long gp_stride, int src0, void* sp, long sp_stride, int dst0,
                              int nfield, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nfield < 0 || (nmem > 0 && nfield > 0 && (!sp || !gp)))
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  return direct_impl(p, mode, true, nmem, gp, gp_stride, src0, sp, sp_stride, dst0, nfield,
                     (hipStream_t)stream);
}

static int inverse_all_impl(exo_sht_plan_t p, int mode, int nmem, int nlev, const void* sp_in,
                            void* gp, hipStream_t s, void* pkbuf = nullptr);

extern "C" int exo_sht_roundtrip(exo_sht_plan_t p, int mode, int nmem, int nlev, const void* sp_in,
                                 void* gp, void* sp_out, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 1) return EXO_ERR_ARG;
  if (nmem > 0 && (!sp_in || !gp || !sp_out)) return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  const Tables& t = p->tb;
  hipStream_t s = (hipStream_t)stream;
  const int nfl = 3 * nlev + 1;
  const int ngp = nfl + 1 + 2 * nlev;
  const long sps = (long)nfl * 2 * t.ncsp;
  const long gps = (long)ngp * t.nlat * t.nlon;
  // one pre-split of all nfl field-levels serves the three inverse transforms,
  // which run as one launch
  const int rc = inverse_all_impl(p, mode, nmem, nlev, sp_in, gp, s);
  if (rc) return rc;
  return direct_impl(p, mode, true, nmem, gp, gps, 0, sp_out, sps, 0, nfl, s);
}

// Bytes of scratch one chunk of c members needs: the pre-split operand
// (split-BF16) at offset 0, the partial direct sums from pk_off(c) on.
static size_t na_pk_off(const Tables& t, int mode, int nfl, int c) {
  const size_t pk = (mode == MODE_BF16X3) ? pack_bytes(t, c, nfl) : 0;
  return (pk + 255) & ~(size_t)255;
}
static size_t na_bytes(const Tables& t, int mode, int nfl, int c) {
  return na_pk_off(t, mode, nfl, c) + forward_scratch_bytes(t, mode, c, nfl);
}

extern "C" int exo_sht_roundtrip_noalloc(exo_sht_plan_t p, int mode, int nmem, int nlev,
                                         const void* sp_in, void* gp, void* sp_out,
                                         void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 1) return EXO_ERR_ARG;
  if (nmem > 0 && (!sp_in || !gp || !sp_out)) return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  const Tables& t = p->tb;
  if (!p->na_scratch) return EXO_ERR_HIP;
  // the plan's device must be the current one: its scratch address is
  // that device's
  char* dev_scratch = nullptr;
  const int rc0 = scratch_address(&dev_scratch);
  if (rc0 != EXO_OK) return rc0;
  if (dev_scratch != p->na_scratch) return EXO_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int nfl = 3 * nlev + 1;
  const int ngp = nfl + 1 + 2 * nlev;
  const long sps = (long)nfl * 2 * t.ncsp;
  const long gps = (long)ngp * t.nlat * t.nlon;
  // the largest chunk of members whose scratch fits the array
  if (na_bytes(t, mode, nfl, 1) > EXO_NOALLOC_SCRATCH_BYTES) return EXO_ERR_SCRATCH;
  int chunk = 1;
  while (chunk < nmem && na_bytes(t, mode, nfl, 2 * chunk) <= EXO_NOALLOC_SCRATCH_BYTES)
    chunk *= 2;
  {
    int lo = chunk, hi = 2 * chunk;  // na_bytes(lo) fits, na_bytes(hi) does not (or hi > nmem)
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (na_bytes(t, mode, nfl, mid) <= EXO_NOALLOC_SCRATCH_BYTES)
        lo = mid;
      else
        hi = mid;
    }
    chunk = lo < nmem ? lo : (nmem > 0 ? nmem : 1);
  }
  const size_t es = (mode == MODE_F64) ? 8 : 4;
  for (int m0 = 0; m0 < nmem; m0 += chunk) {
    const int c = (nmem - m0 < chunk) ? nmem - m0 : chunk;
    const char* spi = (const char*)sp_in + es * sps * m0;
    char* g = (char*)gp + es * gps * m0;
    char* spo = (char*)sp_out + es * sps * m0;
    char* pk = p->na_scratch;
    char* fwd = p->na_scratch + na_pk_off(t, mode, nfl, c);
    const size_t fwd_bytes = EXO_NOALLOC_SCRATCH_BYTES - na_pk_off(t, mode, nfl, c);
    int rc = inverse_all_impl(p, mode, c, nlev, spi, g, s, pk);
    if (rc) return rc;
    rc = direct_impl(p, mode, true, c, g, gps, 0, spo, sps, 0, nfl, s, fwd, fwd_bytes);
    if (rc) return rc;
  }
  return EXO_OK;
}

extern "C" int exo_sht_set_tiles(exo_sht_plan_t p, int mode, int inv_var, int uv_var,
                                 int fwd_var) {
  if (!p || !mode_ok(mode) || inv_var < 0 || inv_var > 3 || uv_var < 0 || uv_var > 3 ||
      fwd_var < 0 || fwd_var > 3)
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  p->inv_var[mode] = inv_var;
  p->uv_var[mode] = uv_var;
  p->fwd_var[mode] = fwd_var;
  return EXO_OK;
}

extern "C" int exo_sht_get_tiles(exo_sht_plan_t p, int mode, int* vars) {
  if (!p || !mode_ok(mode) || !vars) return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  vars[0] = p->inv_var[mode];
  vars[1] = p->uv_var[mode];
  vars[2] = p->fwd_var[mode];
  return EXO_OK;
}

// the inverse half of exo_sht_roundtrip: pre-split (split-BF16) and the
// three inverse transforms in one launch
static int inverse_all_impl(exo_sht_plan_t p, int mode, int nmem, int nlev, const void* sp_in,
                            void* gp, hipStream_t s, void* pkbuf) {
  const Tables& t = p->tb;
  const int nfl = 3 * nlev + 1;
  const long sps = (long)nfl * 2 * t.ncsp;
  const long gps = (long)(nfl + 1 + 2 * nlev) * t.nlat * t.nlon;
  InvArgs pk{};
  int rc = pack_for(p, mode, nmem, sp_in, sps, 0, nfl, s, pk, pkbuf);
  if (rc) return rc;
  InvArgs aq{};
  aq.sp = sp_in;
  aq.out = gp;
  aq.sp_stride = sps;
  aq.out_stride = gps;
  aq.nmem = nmem;
  aq.plavor = t.plavor;
  aq.var = p->inv_var[mode];
  aq.bpk = pk.bpk;
  aq.bpk_stride = pk.bpk_stride;
  aq.bpk_f0 = pk.bpk_f0;
  InvArgs aj = aq, au = aq;
  aq.kind = INV_SCALAR;
  aq.table = TQI;
  aq.nfield = nfl;
  aj.kind = INV_SCALAR;
  aj.table = TQJ;
  aj.src0 = nfl - 1;
  aj.dst0 = nfl;
  aj.nfield = 1;
  au.kind = INV_UV;
  au.div0 = nlev;
  au.u0 = nfl + 1;
  au.v0 = nfl + 1 + nlev;
  au.nlev = nlev;
  rc = launch_inverse_rt(t, mode, aq, aj, au, s);
  return rc == 0 ? EXO_OK : EXO_ERR_LAUNCH;
}

extern "C" int exo_sht_inverse_all(exo_sht_plan_t p, int mode, int nmem, int nlev,
                                   const void* sp_in, void* gp, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 1) return EXO_ERR_ARG;
  if (nmem > 0 && (!sp_in || !gp)) return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  return inverse_all_impl(p, mode, nmem, nlev, sp_in, gp, (hipStream_t)stream);
}

// device buffers and events of one autotune run, released on every return
struct TuneScratch {
  void *sp = nullptr, *gp = nullptr, *so = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~TuneScratch() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sp) (void)hipFree(sp);
    if (gp) (void)hipFree(gp);
    if (so) (void)hipFree(so);
  }
};

// time each tile variant of the three fused kernels on a round trip of
// nmem members x nlev levels and keep the fastest; the plan's lock is held
// for the whole run, so no other call sees an intermediate tile choice
extern "C" int exo_sht_autotune(exo_sht_plan_t p, int mode, int nmem, int nlev, void* stream,
                                int* vars) {
  if (!p || !mode_ok(mode) || nmem < 1 || nlev < 1) return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  const Tables& t = p->tb;
  hipStream_t s = (hipStream_t)stream;
  const int nfl = 3 * nlev + 1, ngp = nfl + 1 + 2 * nlev;
  const long sps = (long)nfl * 2 * t.ncsp, gps = (long)ngp * t.nlat * t.nlon;
  const size_t es = (mode == MODE_F64) ? 8 : 4;
  TuneScratch b;
  HIPCK(hipMalloc(&b.sp, es * sps * nmem));
  HIPCK(hipMalloc(&b.gp, es * gps * nmem));
  HIPCK(hipMalloc(&b.so, es * sps * nmem));
  HIPCK(hipMemsetAsync(b.sp, 0x3f, es * sps * nmem, s));  // finite, nonzero data
  HIPCK(hipEventCreate(&b.e0));
  HIPCK(hipEventCreate(&b.e1));
  const int keep[3] = {p->inv_var[mode], p->uv_var[mode], p->fwd_var[mode]};
  int rc = EXO_OK;
  float best[3] = {1e30f, 1e30f, 1e30f};
  int pick[3] = {0, 0, 0};
  // each variant of each kernel: 2 untimed calls, then the average of 8
  // back-to-back calls (steady state, no host round trip between them)
  constexpr int WARM = 2, REPS = 8;
  for (int v = 0; v < 4 && rc == EXO_OK; ++v) {
    p->inv_var[mode] = p->uv_var[mode] = p->fwd_var[mode] = v;
    for (int k = 0; k < 3 && rc == EXO_OK; ++k) {
      for (int rep = 0; rep < WARM + REPS && rc == EXO_OK; ++rep) {
        if (rep == WARM && hipEventRecord(b.e0, s) != hipSuccess) rc = EXO_ERR_HIP;
        if (rc != EXO_OK) break;
        if (k == 0)
          rc = inverse_all_impl(p, mode, nmem, nlev, b.sp, b.gp, s);
        else if (k == 1)
          rc = inverse_uv_impl(p, mode, nmem, b.sp, sps, 0, nlev, b.gp, gps, nfl + 1,
                               nfl + 1 + nlev, nlev, s, nullptr);
        else
          rc = direct_impl(p, mode, true, nmem, b.gp, gps, 0, b.so, sps, 0, nfl, s);
      }
      if (rc != EXO_OK) break;
      float ms = 0.f;
      if (hipEventRecord(b.e1, s) != hipSuccess || hipEventSynchronize(b.e1) != hipSuccess ||
          hipEventElapsedTime(&ms, b.e0, b.e1) != hipSuccess) {
        rc = EXO_ERR_HIP;
        break;
      }
      ms /= REPS;
      if (ms < best[k]) {
        best[k] = ms;
        pick[k] = v;
      }
    }
  }
  if (rc != EXO_OK) {
    // a failed run leaves the tile choice as it was
    p->inv_var[mode] = keep[0];
    p->uv_var[mode] = keep[1];
    p->fwd_var[mode] = keep[2];
    (void)hipStreamSynchronize(s);
    return rc;
  }
  p->inv_var[mode] = pick[0];
  p->uv_var[mode] = pick[1];
  p->fwd_var[mode] = pick[2];
  if (vars)
    for (int k = 0; k < 3; ++k) vars[k] = pick[k];
  return EXO_OK;
}

// ---------------------------------------------------------------------------
// drop-in entry points on the current (exo_legini) tables
// ---------------------------------------------------------------------------
static exo_sht_plan_s* g_plan = nullptr;
static std::mutex g_mu;

extern "C" int exo_inigau(int klat, double* pz0, double* pzw) {
  if (klat < 2 || (klat & 1) || !pz0 || !pzw) return EXO_ERR_ARG;
  double* d = nullptr;
  HIPCK(hipMalloc(&d, sizeof(double) * 2 * klat));
  k_inigau<<<(klat / 2 + 63) / 64, 64>>>(klat, d, d + klat);
  int rc = (hipGetLastError() == hipSuccess) ? EXO_OK : EXO_ERR_LAUNCH;
  if (rc == EXO_OK && hipMemcpy(pz0, d, sizeof(double) * klat, hipMemcpyDeviceToHost) != hipSuccess)
    rc = EXO_ERR_HIP;
  if (rc == EXO_OK &&
      hipMemcpy(pzw, d + klat, sizeof(double) * klat, hipMemcpyDeviceToHost) != hipSuccess)
    rc = EXO_ERR_HIP;
  (void)hipFree(d);
  return rc;
}

extern "C" int exo_legini(int ntru, int nlat, int nlev, const double* sid, const double* gwd,
                          const double* skgpsp, const double* skspgp, double plavor) {
  if (nlat_of(ntru) != nlat || nlev < 1) return EXO_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  exo_sht_plan_t p = nullptr;
  int rc = exo_sht_plan_create(&p, ntru, sid, gwd, skgpsp, skspgp, plavor);
  if (rc != EXO_OK) return rc;
  p->nlev = nlev;
  if (g_plan) exo_sht_plan_destroy(g_plan);
  g_plan = p;
  return EXO_OK;
}

extern "C" int exo_legini_tables(double* tables) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_plan) return EXO_ERR_NOINIT;
  return copy_tables(g_plan->tb, tables);
}

// run fn on the current drop-in plan with g_mu held: every read of the plan
// (its sizes and tables) happens under the lock, so a concurrent exo_legini
// cannot replace the plan in between
template <typename Fn>
static int with_plan(Fn fn) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_plan) return EXO_ERR_NOINIT;
  return fn(g_plan);
}

// stage host arrays through the plan's device buffer, run, copy back
// (synchronous); the caller holds g_mu
template <typename R, typename Fn>
static int staged(exo_sht_plan_s* p, size_t in_elems, const R* in, size_t out_elems, R* out,
                  Fn run) {
  const size_t need = sizeof(R) * (in_elems + out_elems);
  int rc = plan_grow(&p->stage, &p->stage_bytes, need);
  if (rc) return rc;
  R* din = (R*)p->stage;
  R* dout = din + in_elems;
  if (in_elems) HIPCK(hipMemcpy(din, in, sizeof(R) * in_elems, hipMemcpyHostToDevice));
  rc = run(p, din, dout);
  if (rc) return rc;
  HIPCK(hipDeviceSynchronize());
  if (out_elems) HIPCK(hipMemcpy(out, dout, sizeof(R) * out_elems, hipMemcpyDeviceToHost));
  return EXO_OK;
}

template <typename R>
static int sp2fc_impl(const R* sp, R* fc, int nfl, int table) {
  if (!sp || !fc || nfl < 0) return EXO_ERR_ARG;
  const int mode = sizeof(R) == 8 ? MODE_F64 : MODE_F32;
  return with_plan([&](exo_sht_plan_s* pl) {
    const Tables& t = pl->tb;
    const long nsp = 2L * t.ncsp, nhor = (long)t.nlat * t.nlon;
    return staged<R>(pl, nsp * nfl, sp, nhor * nfl, fc, [&](exo_sht_plan_s* p, R* din, R* dout) {
      InvArgs a{};
      a.sp = din;
      a.out = dout;
      a.sp_stride = 0;
      a.out_stride = 0;
      a.nmem = 1;
      a.kind = INV_SCALAR;
      a.table = table;
      a.src0 = 0;
      a.dst0 = 0;
      a.nfield = nfl;
      a.plavor = p->tb.plavor;
      return launch_inverse(p->tb, mode, false, a, 0);
    });
  });
}

extern "C" int exo_sp2fc_f64(const double* sp, double* fc, int nfl) {
  return sp2fc_impl<double>(sp, fc, nfl, TQI);
}
extern "C" int exo_sp2fc_f32(const float* sp, float* fc, int nfl) {
  return sp2fc_impl<float>(sp, fc, nfl, TQI);
}
extern "C" int exo_sp2fcdmu_f64(const double* sp, double* fc, int nfl) {
  return sp2fc_impl<double>(sp, fc, nfl, TQJ);
}
extern "C" int exo_sp2fcdmu_f32(const float* sp, float* fc, int nfl) {
  return sp2fc_impl<float>(sp, fc, nfl, TQJ);
}

template <typename R>
static int fc2sp_impl(const R* fc, R* sp, int nfl) {
  if (!sp || !fc || nfl < 0) return EXO_ERR_ARG;
  const int mode = sizeof(R) == 8 ? MODE_F64 : MODE_F32;
  return with_plan([&](exo_sht_plan_s* pl) {
    const Tables& t = pl->tb;
    const long nsp = 2L * t.ncsp, nhor = (long)t.nlat * t.nlon;
    return staged<R>(pl, nhor * nfl, fc, nsp * nfl, sp, [&](exo_sht_plan_s* p, R* din, R* dout) {
      return direct_impl(p, mode, false, 1, din, 0, 0, dout, 0, 0, nfl, 0);
    });
  });
}

extern "C" int exo_fc2sp_f64(const double* fc, double* sp, int nfl) {
  return fc2sp_impl<double>(fc, sp, nfl);
}
extern "C" int exo_fc2sp_f32(const float* fc, float* sp, int nfl) {
  return fc2sp_impl<float>(fc, sp, nfl);
}

template <typename R>
static int dv2uv_impl(const R* pd, const R* pz, R* pu, R* pv) {
  if (!pd || !pz || !pu || !pv) return EXO_ERR_ARG;
  const int mode = sizeof(R) == 8 ? MODE_F64 : MODE_F32;
  return with_plan([&](exo_sht_plan_s* pl) {
    const Tables& t = pl->tb;
    const int nlev = pl->nlev;
    const long nsp = 2L * t.ncsp, nhor = (long)t.nlat * t.nlon;
    std::vector<R> in(2 * nsp * nlev), out(2 * nhor * nlev);
    memcpy(in.data(), pz, sizeof(R) * nsp * nlev);
    memcpy(in.data() + nsp * nlev, pd, sizeof(R) * nsp * nlev);
    int rc = staged<R>(pl, in.size(), in.data(), out.size(), out.data(),
                       [&](exo_sht_plan_s* p, R* din, R* dout) {
                         InvArgs a{};
                         a.sp = din;
                         a.out = dout;
                         a.nmem = 1;
                         a.kind = INV_UV;
                         a.vor0 = 0;
                         a.div0 = nlev;
                         a.u0 = 0;
                         a.v0 = nlev;
                         a.nlev = nlev;
                         a.plavor = p->tb.plavor;
                         return launch_inverse(p->tb, mode, false, a, 0);
                       });
    if (rc) return rc;
    memcpy(pu, out.data(), sizeof(R) * nhor * nlev);
    memcpy(pv, out.data() + nhor * nlev, sizeof(R) * nhor * nlev);
    return EXO_OK;
  });
}

extern "C" int exo_dv2uv_f64(const double* pd, const double* pz, double* pu, double* pv) {
  return dv2uv_impl<double>(pd, pz, pu, pv);
}
extern "C" int exo_dv2uv_f32(const float* pd, const float* pz, float* pu, float* pv) {
  return dv2uv_impl<float>(pd, pz, pu, pv);
}

template <typename R>
static int fft_impl(R* a, int n, int lot, int dir) {
  if (!a || lot < 0) return EXO_ERR_ARG;
  const size_t el = (size_t)n * lot;
  const int mode = sizeof(R) == 8 ? MODE_F64 : MODE_F32;
  return with_plan([&](exo_sht_plan_s* pl) {
    if (n != pl->tb.nlon) return EXO_ERR_UNSUPPORTED;
    return staged<R>(pl, el, a, el, a, [&](exo_sht_plan_s* p, R* din, R* dout) {
      HIPCK(hipMemcpyAsync(dout, din, sizeof(R) * el, hipMemcpyDeviceToDevice, 0));
      return launch_fft(p->tb, mode, dir, dout, n, lot, 0);
    });
  });
}

extern "C" int exo_gp2fc_f64(double* a, int n, int lot) { return fft_impl<double>(a, n, lot, -1); }
extern "C" int exo_gp2fc_f32(float* a, int n, int lot) { return fft_impl<float>(a, n, lot, -1); }
extern "C" int exo_fc2gp_f64(double* a, int n, int lot) { return fft_impl<double>(a, n, lot, +1); }
extern "C" int exo_fc2gp_f32(float* a, int n, int lot) { return fft_impl<float>(a, n, lot, +1); }

// ---------------------------------------------------------------------------
// multi-input direct transforms (uv2dv, qtend, mktend)
// ---------------------------------------------------------------------------
// run one output field of a multi-input direct transform over nlev levels;
// din[i] are device copies of the inputs
template <typename R>
static int dirx_one(exo_sht_plan_s* p, const R* const* din, int nin, R* dout,
                    std::initializer_list<DirxTerm> terms) {
  const int mode = sizeof(R) == 8 ? MODE_F64 : MODE_F32;
  DirxArgs a{};
  for (int i = 0; i < nin; ++i) a.in[i] = din[i];
  a.out = dout;
  a.nin = nin;
  a.nlev = p->nlev;
  a.nmem = 1;
  a.nterm = 0;
  for (const DirxTerm& t : terms) a.terms[a.nterm++] = t;
  const size_t need = dirx_scratch_bytes(p->tb, mode, p->nlev);
  if (need > 0) {
    int rc = plan_grow(&p->scratch, &p->scratch_bytes, need);
    if (rc) return rc;
  }
  return launch_dirx(p->tb, mode, a, p->scratch, p->scratch_bytes, 0);
}

// stage nin Fourier inputs and nout spectral outputs (all nlev levels)
template <typename R, typename Fn>
static int dirx_staged(const R* const* in, int nin, R* const* out, int nout, Fn run) {
  for (int i = 0; i < nin; ++i)
    if (!in[i]) return EXO_ERR_ARG;
  for (int o = 0; o < nout; ++o)
    if (!out[o]) return EXO_ERR_ARG;
  return with_plan([&](exo_sht_plan_s* pl) {
    const Tables& t = pl->tb;
    const long nhor = (long)t.nlat * t.nlon * pl->nlev;
    const long nsp = 2L * t.ncsp * pl->nlev;
    std::vector<R> hin(nhor * nin), hout(nsp * nout);
    for (int i = 0; i < nin; ++i) memcpy(hin.data() + i * nhor, in[i], sizeof(R) * nhor);
    int rc = staged<R>(pl, hin.size(), hin.data(), hout.size(), hout.data(),
                       [&](exo_sht_plan_s* p, R* din, R* dout) {
                         const R* ins[3] = {din, din + nhor, din + 2 * nhor};
                         R* outs[3] = {dout, dout + nsp, dout + 2 * nsp};
                         return run(p, ins, outs);
                       });
    if (rc) return rc;
    for (int o = 0; o < nout; ++o) memcpy(out[o], hout.data() + o * nsp, sizeof(R) * nsp);
    return EXO_OK;
  });
}

// uv2dv(pu, pv, pd, pz), src/legmod.f90:356 (folded branch)
template <typename R>
static int uv2dv_impl(const R* pu, const R* pv, R* pd, R* pz) {
  const R* in[2] = {pu, pv};
  R* out[2] = {pd, pz};
  return dirx_staged<R>(in, 2, out, 2, [&](exo_sht_plan_s* p, const R** din, R** dout) {
    // pd = -qe.v + i qm.u ; pz = qe.u + i qm.v   (inputs 0 = u, 1 = v)
    int rc = dirx_one<R>(p, din, 2, dout[0],
                         {{1, 0, 0, -1, TQE}, {0, 1, 0, -1, TQM},
                          {1, 1, 1, -1, TQE}, {0, 0, 1, +1, TQM}});
    if (rc) return rc;
    return dirx_one<R>(p, din, 2, dout[1],
                       {{0, 0, 0, +1, TQE}, {1, 1, 0, -1, TQM},
                        {0, 1, 1, +1, TQE}, {1, 0, 1, +1, TQM}});
  });
}
extern "C" int exo_uv2dv_f64(const double* pu, const double* pv, double* pd, double* pz) {
  return uv2dv_impl<double>(pu, pv, pd, pz);
}
extern "C" int exo_uv2dv_f32(const float* pu, const float* pv, float* pd, float* pz) {
  return uv2dv_impl<float>(pu, pv, pd, pz);
}

// qtend(q, qn, uq, vq), src/legmod.f90:433 (folded branch)
template <typename R>
static int qtend_impl(R* q, const R* qn, const R* uq, const R* vq) {
  const R* in[3] = {qn, uq, vq};
  R* out[1] = {q};
  return dirx_staged<R>(in, 3, out, 1, [&](exo_sht_plan_s* p, const R** din, R** dout) {
    // q = qe.vq + qc.qn - i qm.uq   (inputs 0 = qn, 1 = uq, 2 = vq)
    return dirx_one<R>(p, din, 3, dout[0],
                       {{2, 0, 0, +1, TQE}, {0, 0, 0, +1, TQC}, {1, 1, 0, +1, TQM},
                        {2, 1, 1, +1, TQE}, {0, 1, 1, +1, TQC}, {1, 0, 1, -1, TQM}});
  });
}
extern "C" int exo_qtend_f64(double* q, const double* qn, const double* uq, const double* vq) {
  return qtend_impl<double>(q, qn, uq, vq);
}
extern "C" int exo_qtend_f32(float* q, const float* qn, const float* uq, const float* vq) {
  return qtend_impl<float>(q, qn, uq, vq);
}

// mktend(d, t, z, tn, fu, fv, ke, ut, vt), src/legmod.f90:505 (folded)
template <typename R>
static int mktend_impl(R* d, R* t, R* z, const R* tn, const R* fu, const R* fv, const R* ke,
                       const R* ut, const R* vt) {
  const R* in_d[3] = {ke, fv, fu};
  const R* in_t[3] = {vt, tn, ut};
  const R* in_z[2] = {fu, fv};
  int rc;
  {
    R* out[1] = {d};
    // D = qq.ke - qe.fv + i qm.fu   (0 = ke, 1 = fv, 2 = fu)
    rc = dirx_staged<R>(in_d, 3, out, 1, [&](exo_sht_plan_s* p, const R** din, R** dout) {
      return dirx_one<R>(p, din, 3, dout[0],
                         {{0, 0, 0, +1, TQQ}, {1, 0, 0, -1, TQE}, {2, 1, 0, -1, TQM},
                          {0, 1, 1, +1, TQQ}, {1, 1, 1, -1, TQE}, {2, 0, 1, +1, TQM}});
    });
    if (rc) return rc;
  }
  {
    R* out[1] = {t};
    // T = qe.vt + qc.tn - i qm.ut   (0 = vt, 1 = tn, 2 = ut)
    rc = dirx_staged<R>(in_t, 3, out, 1, [&](exo_sht_plan_s* p, const R** din, R** dout) {
      return dirx_one<R>(p, din, 3, dout[0],
                         {{0, 0, 0, +1, TQE}, {1, 0, 0, +1, TQC}, {2, 1, 0, +1, TQM},
                          {0, 1, 1, +1, TQE}, {1, 1, 1, +1, TQC}, {2, 0, 1, -1, TQM}});
    });
    if (rc) return rc;
  }
  R* out[1] = {z};
  // Z = qe.fu + i qm.fv   (0 = fu, 1 = fv)
  return dirx_staged<R>(in_z, 2, out, 1, [&](exo_sht_plan_s* p, const R** din, R** dout) {
    return dirx_one<R>(p, din, 2, dout[0],
                       {{0, 0, 0, +1, TQE}, {1, 1, 0, -1, TQM},
                        {0, 1, 1, +1, TQE}, {1, 0, 1, +1, TQM}});
  });
}
extern "C" int exo_mktend_f64(double* d, double* t, double* z, const double* tn,
                              const double* fu, const double* fv, const double* ke,
                              const double* ut, const double* vt) {
  return mktend_impl<double>(d, t, z, tn, fu, fv, ke, ut, vt);
}
extern "C" int exo_mktend_f32(float* d, float* t, float* z, const float* tn, const float* fu,
                              const float* fv, const float* ke, const float* ut,
                              const float* vt) {
  return mktend_impl<float>(d, t, z, tn, fu, fv, ke, ut, vt);
}

// ---------------------------------------------------------------------------
// device-resident Fourier-space transforms on plans (no fused FFT)
// ---------------------------------------------------------------------------
extern "C" int exo_sht_sp2fc(exo_sht_plan_t p, int mode, int table, int nmem, const void* sp,
                             long sp_stride, int src0, void* fc, long fc_stride, int dst0,
                             int nfield, void* stream) {
  if (!p || !mode_ok(mode) || (table != EXO_QI && table != EXO_QJ) || nmem < 0 || nfield < 0 ||
      src0 < 0 || dst0 < 0 || (nmem > 0 && nfield > 0 && (!sp || !fc)))
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  return inverse_impl(p, mode, table, nmem, sp, sp_stride, src0, fc, fc_stride, dst0, nfield,
                      (hipStream_t)stream, nullptr, false);
}

extern "C" int exo_sht_dv2uv(exo_sht_plan_t p, int mode, int nmem, const void* sp,
                             long sp_stride, int vor0, int div0, void* fc, long fc_stride, int u0,
                             int v0, int nlev, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 0 || vor0 < 0 || div0 < 0 || u0 < 0 ||
      v0 < 0 || (nmem > 0 && nlev > 0 && (!sp || !fc)))
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  return inverse_uv_impl(p, mode, nmem, sp, sp_stride, vor0, div0, fc, fc_stride, u0, v0, nlev,
                         (hipStream_t)stream, nullptr, false);
}

extern "C" int exo_sht_fc2sp(exo_sht_plan_t p, int mode, int nmem, const void* fc,
                             long fc_stride, int src0, void* sp, long sp_stride, int dst0,
                             int nfield, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nfield < 0 || src0 < 0 || dst0 < 0 ||
      (nmem > 0 && nfield > 0 && (!sp || !fc)))
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  return direct_impl(p, mode, false, nmem, fc, fc_stride, src0, sp, sp_stride, dst0, nfield,
                     (hipStream_t)stream);
}

extern "C" int exo_sht_fft(exo_sht_plan_t p, int mode, int dir, void* a, long lot,
                           void* stream) {
  if (!p || !mode_ok(mode) || (dir != 1 && dir != -1) || lot < 0 || (lot > 0 && !a))
    return EXO_ERR_ARG;
  if (lot == 0) return EXO_OK;
  std::lock_guard<std::mutex> lk(p->mu);
  const int rc = launch_fft(p->tb, mode, dir, a, p->tb.nlon, lot, (hipStream_t)stream);
  return rc == 0 ? EXO_OK : (rc == -2 ? EXO_ERR_UNSUPPORTED : EXO_ERR_LAUNCH);
}

// one output field of a multi-term direct transform over nmem members x
// nlev levels of device Fourier arrays
static int dirx_dev(exo_sht_plan_t p, int mode, int nmem, int nlev, const void* const* in,
                    int nin, long in_stride, void* out, long out_stride,
                    std::initializer_list<DirxTerm> terms, hipStream_t s) {
  DirxArgs a{};
  for (int i = 0; i < nin; ++i) a.in[i] = in[i];
  a.out = out;
  a.nin = nin;
  a.nlev = nlev;
  a.nmem = nmem;
  a.in_stride = in_stride;
  a.out_stride = out_stride;
  a.nterm = 0;
  for (const DirxTerm& t : terms) a.terms[a.nterm++] = t;
  const size_t need = dirx_scratch_bytes(p->tb, mode, nmem * nlev);
  if (need > 0) {
    const int rc = plan_grow(&p->scratch, &p->scratch_bytes, need);
    if (rc) return rc;
  }
  const int rc = launch_dirx(p->tb, mode, a, p->scratch, p->scratch_bytes, s);
  return rc == 0 ? EXO_OK : (rc == -4 ? EXO_ERR_SCRATCH : EXO_ERR_LAUNCH);
}

static bool all_set(std::initializer_list<const void*> ps) {
  for (const void* q : ps)
    if (!q) return false;
  return true;
}

extern "C" int exo_sht_uv2dv(exo_sht_plan_t p, int mode, int nmem, int nlev, const void* pu,
                             const void* pv, long f_stride, void* pd, void* pz, long s_stride,
                             void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 0 ||
      (nmem > 0 && nlev > 0 && !all_set({pu, pv, pd, pz})))
    return EXO_ERR_ARG;
  if (nmem == 0 || nlev == 0) return EXO_OK;
  std::lock_guard<std::mutex> lk(p->mu);
  hipStream_t s = (hipStream_t)stream;
  const void* in[2] = {pu, pv};
  // pd = -qe.v + i qm.u ; pz = qe.u + i qm.v   (inputs 0 = u, 1 = v)
  int rc = dirx_dev(p, mode, nmem, nlev, in, 2, f_stride, pd, s_stride,
                    {{1, 0, 0, -1, TQE}, {0, 1, 0, -1, TQM}, {1, 1, 1, -1, TQE}, {0, 0, 1, +1, TQM}},
                    s);
  if (rc) return rc;
  return dirx_dev(p, mode, nmem, nlev, in, 2, f_stride, pz, s_stride,
                  {{0, 0, 0, +1, TQE}, {1, 1, 0, -1, TQM}, {0, 1, 1, +1, TQE}, {1, 0, 1, +1, TQM}},
                  s);
}

extern "C" int exo_sht_qtend(exo_sht_plan_t p, int mode, int nmem, int nlev, void* q,
                             long s_stride, const void* qn, const void* uq, const void* vq,
                             long f_stride, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 0 ||
      (nmem > 0 && nlev > 0 && !all_set({q, qn, uq, vq})))
    return EXO_ERR_ARG;
  if (nmem == 0 || nlev == 0) return EXO_OK;
  std::lock_guard<std::mutex> lk(p->mu);
  const void* in[3] = {qn, uq, vq};
  // q = qe.vq + qc.qn - i qm.uq   (inputs 0 = qn, 1 = uq, 2 = vq)
  return dirx_dev(p, mode, nmem, nlev, in, 3, f_stride, q, s_stride,
                  {{2, 0, 0, +1, TQE}, {0, 0, 0, +1, TQC}, {1, 1, 0, +1, TQM},
                   {2, 1, 1, +1, TQE}, {0, 1, 1, +1, TQC}, {1, 0, 1, -1, TQM}},
                  (hipStream_t)stream);
}

extern "C" int exo_sht_mktend(exo_sht_plan_t p, int mode, int nmem, int nlev, void* d, void* t,
                              void* z, long s_stride, const void* tn, const void* fu,
                              const void* fv, const void* ke, const void* ut, const void* vt,
                              long f_stride, void* stream) {
  if (!p || !mode_ok(mode) || nmem < 0 || nlev < 0 ||
      (nmem > 0 && nlev > 0 && !all_set({d, t, z, tn, fu, fv, ke, ut, vt})))
    return EXO_ERR_ARG;
  if (nmem == 0 || nlev == 0) return EXO_OK;
  std::lock_guard<std::mutex> lk(p->mu);
  hipStream_t s = (hipStream_t)stream;
  // D = qq.ke - qe.fv + i qm.fu   (0 = ke, 1 = fv, 2 = fu)
  const void* in_d[3] = {ke, fv, fu};
  int rc = dirx_dev(p, mode, nmem, nlev, in_d, 3, f_stride, d, s_stride,
                    {{0, 0, 0, +1, TQQ}, {1, 0, 0, -1, TQE}, {2, 1, 0, -1, TQM},
                     {0, 1, 1, +1, TQQ}, {1, 1, 1, -1, TQE}, {2, 0, 1, +1, TQM}},
                    s);
  if (rc) return rc;
  // T = qe.vt + qc.tn - i qm.ut   (0 = vt, 1 = tn, 2 = ut)
  const void* in_t[3] = {vt, tn, ut};
  rc = dirx_dev(p, mode, nmem, nlev, in_t, 3, f_stride, t, s_stride,
                {{0, 0, 0, +1, TQE}, {1, 0, 0, +1, TQC}, {2, 1, 0, +1, TQM},
                 {0, 1, 1, +1, TQE}, {1, 1, 1, +1, TQC}, {2, 0, 1, -1, TQM}},
                s);
  if (rc) return rc;
  // Z = qe.fu + i qm.fv   (0 = fu, 1 = fv)
  const void* in_z[2] = {fu, fv};
  return dirx_dev(p, mode, nmem, nlev, in_z, 2, f_stride, z, s_stride,
                  {{0, 0, 0, +1, TQE}, {1, 1, 0, -1, TQM}, {0, 1, 1, +1, TQE}, {1, 0, 1, +1, TQM}},
                  s);
}
