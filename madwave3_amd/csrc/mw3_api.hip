// mw3_api.hip -- C ABI of libmw3 (include/mw3.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "mw3_internal.h"

using namespace mw3;

static thread_local std::string g_create_err;

// Device array of mw3_cheby_spectrum_noalloc: the Chebyshev coefficients of
// one call (kmax + 1 doubles).  Every device the library runs on holds its
// own copy (1 MiB) until the process ends; calls that overlap on one device
// share it and race.  In host code `scratch` names a host-side stand-in, so
// host code uses only the per-device address from hipGetSymbolAddress.
constexpr int kSpecScratch = 1 << 17;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kSpecScratch];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

namespace mw3 {

void timer_begin(Ctx* c, int id, hipStream_t s) {
  if (!c->timing || !c->timing_step) return;
  if (!s) s = c->stream;
  Timer t;
  if (!c->pool.empty()) {
    t = c->pool.back();
    c->pool.pop_back();
  } else {
    // timing only: elapsed times are read after the streams synchronize, so
    // the records need no system-scope fence (which would write back and
    // invalidate the caches at every record, inside the timed work)
    (void)hipEventCreateWithFlags(&t.a, hipEventDisableSystemFence);
    (void)hipEventCreateWithFlags(&t.b, hipEventDisableSystemFence);
  }
  (void)hipEventRecord(t.a, s);
  c->pending.push_back({id, t});
}

void timer_end(Ctx* c, int id, hipStream_t s) {
  if (!c->timing || !c->timing_step) return;
  if (!s) s = c->stream;
  for (auto it = c->pending.rbegin(); it != c->pending.rend(); ++it)
    if (it->first == id) {
      (void)hipEventRecord(it->second.b, s);
      return;
    }
}

static void timer_collect(Ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  if (c->aux) (void)hipStreamSynchronize(c->aux);
  if (c->side) (void)hipStreamSynchronize(c->side);
  for (auto& p : c->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.second.a, p.second.b) == hipSuccess) {
      c->t_ms[p.first] += ms;
      c->t_n[p.first] += 1;
    }
    c->pool.push_back(p.second);
  }
  c->pending.clear();
}

}  // namespace mw3

static int seterr(Ctx* c, int code, const char* what, hipError_t e = hipSuccess) {
  char buf[512];
  if (e != hipSuccess)
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  else
    snprintf(buf, sizeof buf, "%s", what);
  if (c)
    c->err = buf;
  else
    g_create_err = buf;
  return code;
}

#define HIPCHK(ctx, call, what)                                      \
  do {                                                               \
    hipError_t _e = (call);                                          \
    if (_e != hipSuccess) return seterr(ctx, MW3_EDEVICE, what, _e); \
  } while (0)

template <typename T>
static hipError_t upload(void** dst, const std::vector<double>& v) {
  std::vector<T> tmp(v.begin(), v.end());
  size_t bytes = sizeof(T) * std::max<size_t>(tmp.size(), 1);
  hipError_t e = hipMalloc(dst, bytes);
  if (e != hipSuccess) return e;
  if (tmp.empty()) return hipMemset(*dst, 0, bytes);
  return hipMemcpy(*dst, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice);
}

template <typename T>
static hipError_t upload_raw(T** dst, const T* src, size_t n) {
  size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
  hipError_t e = hipMalloc((void**)dst, bytes);
  if (e != hipSuccess) return e;
  if (n == 0) return hipMemset(*dst, 0, bytes);
  return hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice);
}

static hipError_t alloc_zero(void** p, size_t bytes) {
  bytes = std::max<size_t>(bytes, 8);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return e;
  return hipMemset(*p, 0, bytes);
}

static void free_all(Ctx* c) {
  void* ptrs[] = {c->d_pt_dense, c->d_mask, c->d_v, c->d_absfr1, c->d_absfr2, c->d_psi,
                  c->d_psim1, c->d_psi0, c->d_a0, c->d_a0b, c->d_t1, c->d_t2, c->d_fbr,
                  c->d_ang, c->d_hout, c->d_nn1, c->d_p2r1, c->d_p2r2, c->d_divi,
                  c->d_dst_tables, c->d_Dmat, c->d_acoef, c->d_bcoef, c->d_nbr,
                  c->d_flux_dense, c->d_flux_col, c->d_zar1flux, c->d_ring_zr, c->d_ring_zdr,
                  c->d_ring_cvj, c->d_zcr, c->d_zcdr, c->d_zs2, c->d_etot, c->d_fdv, c->d_dcvj,
                  c->d_cvj_skip, c->d_norm_part, c->d_ring_norm, c->d_efe, c->d_zfe, c->d_phi,
                  c->d_wmat, c->d_zero_skip, c->d_ch_elec, c->d_ch_omega, c->d_ch_cpl, c->d_potiflux,
                  c->d_flux_elec, c->d_scratch, c->d_fmap, c->d_ir1flux, c->d_slice,
                  c->d_psi32, c->d_psim132, c->d_scr32, c->d_uw_row, c->d_uw_word,
                  c->d_uw_item, c->d_uw_gfirst, c->d_uw_part, c->d_tprod, c->d_ppart,
                  c->d_dtrans, c->d_ring_prod, c->d_prodwf, c->d_prod_part,
                  c->d_ring_cvjp, c->d_cvjp_last, c->d_zs2prod, c->d_rad_items, c->d_fold_D, c->d_fold_a, c->d_fold_b,
                  c->d_lowT, c->d_mid_items, c->d_mid_sizes, c->d_mid_items2, c->d_mid_wts,
                  c->d_ang_tiles, c->d_fold_jmap, c->d_fold_bm, c->d_fold_bp,
                  c->d_ang_groups};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  free_plans(c);
  for (auto& p : c->pending) c->pool.push_back(p.second);
  c->pending.clear();
  for (auto& t : c->pool) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  c->pool.clear();
  if (c->aux) {
    (void)hipStreamSynchronize(c->aux);
    (void)hipStreamDestroy(c->aux);
  }
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  for (auto& g : c->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  if (c->d_slot) (void)hipFree(c->d_slot);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_slice) (void)hipEventDestroy(c->ev_slice);
  if (c->ev_cvj) (void)hipEventDestroy(c->ev_cvj);
  if (c->ev_psi) (void)hipEventDestroy(c->ev_psi);
  if (c->ev_ang) (void)hipEventDestroy(c->ev_ang);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

extern "C" {

int32_t mw3_abi_version(void) { return MW3_ABI_VERSION; }

const char* mw3_last_error(const mw3_ctx* ctx) {
  if (!ctx) return g_create_err.c_str();
  return reinterpret_cast<const Ctx*>(ctx)->err.c_str();
}

const char* mw3_kernel_name(int32_t id) {
  static const char* names[MW3_NKERNELS] = {"angular_fwd", "angular_bwd", "dst_R_fwd",
                                            "dst_r_scale", "dst_R_inv",   "update_fused",
                                            "projections", "energy_accum", "coriolis",
                                            "prod_projection", "update_tail"};
  if (id < 0 || id >= MW3_NKERNELS) return "";
  return names[id];
}

int32_t mw3_create(const mw3_setup* setup, int32_t precision, mw3_ctx** out) {
  if (!out) return seterr(nullptr, MW3_EINVAL, "mw3_create: out is NULL");
  *out = nullptr;
  if (precision != MW3_FP64 && precision != MW3_MIXED)
    return seterr(nullptr, MW3_EINVAL, "mw3_create: precision must be MW3_FP64 or MW3_MIXED");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return seterr(nullptr, MW3_EDEVICE, "mw3_create: no HIP device");
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return seterr(nullptr, MW3_ENOMEM, "mw3_create: out of host memory");
  std::string err;
  int rc = build_tables(setup, &c->t, &err);
  if (rc != MW3_OK) {
    delete c;
    return seterr(nullptr, rc, err.c_str());
  }
  if (c->t.ne > 8) {
    delete c;
    return seterr(nullptr, MW3_EUNSUP, "mw3_create: more than 8 electronic states");
  }
  c->s = *setup;
  c->precision = precision;
  // the angular term beside the radial passes on its own stream in mixed
  // precision (measured faster there; the FP64 pair ran slower side by side)
  c->opt_side = precision == MW3_MIXED ? 1 : 0;
  (void)hipGetDevice(&c->device);
  const HostTables& t = c->t;
  const size_t nd = (size_t)t.nc * t.nslab;
  const size_t es = precision == MW3_MIXED ? sizeof(float) : sizeof(double);
  hipError_t e;
#define CK(call, what)                                            \
  do {                                                            \
    if ((e = (call)) != hipSuccess) {                             \
      free_all(c);                                                \
      delete c;                                                   \
      return seterr(nullptr, e == hipErrorOutOfMemory ? MW3_ENOMEM : MW3_EDEVICE, what, e); \
    }                                                             \
  } while (0)
  CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "mw3_create: stream");
  CK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking), "mw3_create: stream");
  // the stream-ordering events only order device work between the library's
  // streams (the host never queries them; host reads follow a stream
  // synchronize and a copy on the main stream), so they carry no system-scope
  // fence: with one, every record and wait wrote back and invalidated the
  // caches, about 10 us of idle GPU at each of the step's two crossings
  const unsigned evf = hipEventDisableTiming | hipEventDisableSystemFence;
  CK(hipEventCreateWithFlags(&c->ev_slice, evf), "mw3_create: event");
  CK(hipEventCreateWithFlags(&c->ev_cvj, evf), "mw3_create: event");
  CK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking), "mw3_create: stream");
  CK(hipEventCreateWithFlags(&c->ev_psi, evf), "mw3_create: event");
  CK(hipEventCreateWithFlags(&c->ev_ang, evf), "mw3_create: event");
  CK(hipEventCreateWithFlags(&c->ev_fork, evf), "mw3_create: event");
  CK(alloc_zero((void**)&c->d_slot, sizeof(int32_t)), "mw3_create: ring slot");
  CK(upload_raw(&c->d_pt_dense, t.pt_dense.data(), t.pt_dense.size()), "mw3_create: index map");
  CK(upload_raw(&c->d_mask, t.mask.data(), t.mask.size()), "mw3_create: mask");

  if (precision == MW3_MIXED)
    CK(upload<float>(&c->d_v, t.vdense), "mw3_create: V");
  else
    CK(upload<double>(&c->d_v, t.vdense), "mw3_create: V");
  CK(upload_raw(&c->d_absfr1, setup->absfr1, (size_t)t.n1), "mw3_create: absfr1");
  CK(upload_raw(&c->d_absfr2, setup->absfr2, (size_t)t.n2), "mw3_create: absfr2");
  CK(alloc_zero((void**)&c->d_psi, nd * sizeof(double)), "mw3_create: psi");
  c->graph_psi0 = c->d_psi;
  CK(alloc_zero((void**)&c->d_psim1, nd * sizeof(double)), "mw3_create: psi_m1");
  CK(alloc_zero((void**)&c->d_psi0, nd * sizeof(double)), "mw3_create: psi0");
  CK(alloc_zero((void**)&c->d_hout, nd * sizeof(double)), "mw3_create: hout");
  CK(alloc_zero((void**)&c->d_scratch, nd * sizeof(double)), "mw3_create: scratch");
  if (precision == MW3_MIXED) {
    CK(alloc_zero((void**)&c->d_psi32, nd * sizeof(float)), "mw3_create: psi FP32");
    CK(alloc_zero((void**)&c->d_psim132, nd * sizeof(float)), "mw3_create: psi FP32");
    CK(alloc_zero((void**)&c->d_scr32, nd * sizeof(float)), "mw3_create: psi FP32");
  }
  if (setup->netot > 0) {
    c->reacfct_host.assign(setup->reacfct, setup->reacfct + setup->netot);
    c->etot_host.assign(setup->etots2, setup->etots2 + setup->netot);
  }
  CK(alloc_zero((void**)&c->d_a0, (size_t)t.ntot * sizeof(double)), "mw3_create: a0");
  CK(alloc_zero((void**)&c->d_a0b, (size_t)t.ntot * sizeof(double)), "mw3_create: a0b");
  CK(alloc_zero(&c->d_t1, nd * es), "mw3_create: work slab");
  CK(alloc_zero(&c->d_t2, nd * es), "mw3_create: work slab");
  CK(alloc_zero(&c->d_ang, nd * es), "mw3_create: angular slab");
  CK(alloc_zero(&c->d_fbr, (size_t)t.nc * t.nJ * t.nplane * es), "mw3_create: FBR slab");
  CK(alloc_zero(&c->d_phi, (size_t)t.nc * t.nJ * t.nplane * es), "mw3_create: FBR slab");
  CK(upload_raw(&c->d_nn1, t.nn1.data(), t.nn1.size()), "mw3_create: tables");
  CK(upload_raw(&c->d_p2r1, t.p2r1.data(), t.p2r1.size()), "mw3_create: tables");
  CK(upload_raw(&c->d_p2r2, t.p2r2.data(), t.p2r2.size()), "mw3_create: tables");
  CK(upload_raw(&c->d_divi, t.divi.data(), t.divi.size()), "mw3_create: tables");
  if (precision == MW3_MIXED) {
    CK(upload<float>(&c->d_dst_tables, t.dst_tables), "mw3_create: DST tables");
    CK(upload<float>(&c->d_Dmat, t.Dmat), "mw3_create: D");
  } else {
    CK(upload<double>(&c->d_dst_tables, t.dst_tables), "mw3_create: DST tables");
    CK(upload<double>(&c->d_Dmat, t.Dmat), "mw3_create: D");
  }
  CK(upload<double>(&c->d_acoef, t.acoef), "mw3_create: coefficients");
  CK(upload<double>(&c->d_bcoef, t.bcoef), "mw3_create: coefficients");
  CK(upload_raw(&c->d_nbr, t.nbr.data(), t.nbr.size()), "mw3_create: tables");
  CK(upload_raw(&c->d_ch_elec, t.ch_elec.data(), t.ch_elec.size()), "mw3_create: tables");
  CK(upload_raw(&c->d_ch_omega, t.ch_omega.data(), t.ch_omega.size()), "mw3_create: tables");
  {
    std::vector<int32_t> cpl((size_t)4 * t.nc, -1);
    bool fits = true;
    for (int ch = 0; ch < t.nc && fits; ch++) {
      int k = 0;
      for (int s2 = 0; s2 < t.nc; s2++)
        if (t.ch_omega[s2] == t.ch_omega[ch]) {
          if (k == 4) {
            fits = false;
            break;
          }
          cpl[(size_t)4 * ch + k++] = s2;
        }
    }
    if (fits) CK(upload_raw(&c->d_ch_cpl, cpl.data(), cpl.size()), "mw3_create: tables");
  }
  // projections
  int netot = std::max(setup->netot, 0);
  CK(upload_raw(&c->d_flux_dense, t.flux_dense.data(), t.flux_dense.size()), "mw3_create: flux");
  CK(upload_raw(&c->d_flux_col, t.flux_col.data(), t.flux_col.size()), "mw3_create: flux");
  {
    // zAR1flux is read by the update whenever projections run (zeros if the
    // flux line has no point)
    std::vector<double> zar(t.zar1flux.begin(), t.zar1flux.end());
    if (zar.size() < 2 * (size_t)t.n1) zar.resize(2 * (size_t)t.n1, 0.0);
    CK(upload_raw(&c->d_zar1flux, zar.data(), zar.size()), "mw3_create: flux");
  }
  CK(upload_raw(&c->d_potiflux, t.potiflux.data(), t.potiflux.size()), "mw3_create: flux");
  CK(upload_raw(&c->d_flux_elec, t.flux_elec.data(), t.flux_elec.size()), "mw3_create: flux");
  {
    std::vector<uint8_t> ones((size_t)std::max<int64_t>(t.nflux, 1), 1);
    CK(upload_raw(&c->d_zero_skip, ones.data(), ones.size()), "mw3_create: flux");
  }
  CK(alloc_zero((void**)&c->d_ring_zr, sizeof(double) * kRing * t.nflux), "mw3_create: ring");
  CK(alloc_zero((void**)&c->d_ring_zdr, sizeof(double) * 2 * kRing * t.nflux), "mw3_create: ring");
  CK(alloc_zero((void**)&c->d_ring_cvj, sizeof(double) * kRing * t.ncvj), "mw3_create: ring");
  CK(alloc_zero((void**)&c->d_zcr, sizeof(double) * 2 * t.nflux * netot), "mw3_create: zCR");
  CK(alloc_zero((void**)&c->d_zcdr, sizeof(double) * 2 * t.nflux * netot), "mw3_create: zCdR");
  CK(alloc_zero((void**)&c->d_zs2, sizeof(double) * 2 * (size_t)t.ncvj * netot), "mw3_create: zS2");
  CK(alloc_zero((void**)&c->d_wmat, sizeof(double) * 2 * kRing * netot), "mw3_create: W");
  if (netot > 0)
    CK(upload_raw(&c->d_etot, setup->etots2, (size_t)netot), "mw3_create: energies");
  CK(upload_raw(&c->d_fdv, t.fdv.data(), t.fdv.size()), "mw3_create: fd");
  CK(upload_raw(&c->d_dcvj, t.dcvj.data(), t.dcvj.size()), "mw3_create: D");
  CK(upload_raw(&c->d_cvj_skip, t.cvj_isref.data(), t.cvj_isref.size()), "mw3_create: flags");
  {
    // update work items: for each (channel, angle, 64-R chunk), the r rows with
    // a point of the compressed grid in that chunk (static), cut into items of
    // at most kUpdRows rows, one per wave of k_update
    const int nrc = (t.n2 + 63) / 64;
    std::vector<int32_t> rows, items, gfirst;
    std::vector<uint64_t> words;
    std::vector<int32_t> list_off((size_t)t.na * nrc), list_cnt((size_t)t.na * nrc);
    for (int ia = 0; ia < t.na; ia++)
      for (int rc = 0; rc < nrc; rc++) {
        list_off[(size_t)ia * nrc + rc] = (int32_t)rows.size();
        for (int r = 0; r < t.n1; r++) {
          uint64_t wd = 0;
          for (int l = 0; l < 64 && rc * 64 + l < t.n2; l++)
            if (t.mask[(size_t)ia * t.nplane + (size_t)r * t.n2 + rc * 64 + l]) wd |= 1ull << l;
          if (wd) {
            rows.push_back(r);
            words.push_back(wd);
          }
        }
        list_cnt[(size_t)ia * nrc + rc] = (int32_t)rows.size() - list_off[(size_t)ia * nrc + rc];
      }
    for (int ch = 0; ch < t.nc; ch++)
      for (int ia = 0; ia < t.na; ia++)
        for (int rc = 0; rc < nrc; rc++) {
          const int gi = (ch * t.na + ia) * nrc + rc;
          gfirst.push_back((int32_t)(items.size() / 4));
          const int lo = list_off[(size_t)ia * nrc + rc], n = list_cnt[(size_t)ia * nrc + rc];
          for (int k = 0; k < n; k += kUpdRows) {
            items.push_back(gi);
            items.push_back(lo + k);
            items.push_back(std::min(kUpdRows, n - k));
            items.push_back(0);
          }
        }
    gfirst.push_back((int32_t)(items.size() / 4));
    if (rows.empty()) {  // keep the arrays non-empty
      rows.push_back(0);
      words.push_back(0);
    }
    c->nnorm_blocks = (int)(items.size() / 4);
    CK(upload_raw(&c->d_uw_row, rows.data(), rows.size()), "mw3_create: update rows");
    CK(upload_raw(&c->d_uw_word, words.data(), words.size()), "mw3_create: update rows");
    if (!items.empty())
      CK(upload_raw(&c->d_uw_item, items.data(), items.size()), "mw3_create: update items");
    CK(upload_raw(&c->d_uw_gfirst, gfirst.data(), gfirst.size()), "mw3_create: update items");
    CK(alloc_zero((void**)&c->d_uw_part, sizeof(double) * 3 * 64 * std::max(1, c->nnorm_blocks)),
       "mw3_create: update partials");
  }
  {
    std::vector<int32_t> fmap((size_t)t.nc * t.na * t.n2, -1);
    for (int64_t f = 0; f < t.nflux; f++) {
      // flux_col is the dense index of (c, ia, 0, R): c*slab + ia*plane + R
      uint32_t col = t.flux_col[f];
      int64_t cia = col / t.nplane;
      int64_t R = col % t.n2;
      fmap[(size_t)cia * t.n2 + R] = (int32_t)f;
    }
    CK(upload_raw(&c->d_fmap, fmap.data(), fmap.size()), "mw3_create: flux map");
    std::vector<int32_t> i1f(t.ir1flux.begin(), t.ir1flux.end());
    if (i1f.empty()) i1f.assign(t.na, 1);
    CK(upload_raw(&c->d_ir1flux, i1f.data(), i1f.size()), "mw3_create: flux map");
    CK(alloc_zero((void**)&c->d_slice, sizeof(double) * (size_t)t.nc * t.n1 * t.na),
       "mw3_create: slice");
  }
  CK(alloc_zero((void**)&c->d_norm_part, sizeof(double) * 2 * std::max(1, c->nnorm_blocks)),
     "mw3_create: norms");
  CK(alloc_zero((void**)&c->d_ring_norm, sizeof(double) * 2 * kRing), "mw3_create: norms");
  if (setup->nofe > 0) {
    if (!setup->efe) {
      free_all(c);
      delete c;
      return seterr(nullptr, MW3_EINVAL, "mw3_create: nofe > 0 needs efe");
    }
    CK(upload_raw(&c->d_efe, setup->efe, (size_t)setup->nofe), "mw3_create: efe");
    CK(alloc_zero((void**)&c->d_zfe, sizeof(double) * 2 * setup->nofe * nd), "mw3_create: zfe");
  }
  CK(build_plans(c), "mw3_create: plans");
  CK(hipStreamSynchronize(c->stream), "mw3_create: sync");
#undef CK
  *out = reinterpret_cast<mw3_ctx*>(c);
  return MW3_OK;
}

void mw3_destroy(mw3_ctx* ctx) {
  if (!ctx) return;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
}

int64_t mw3_ntot(const mw3_ctx* ctx) { return ctx ? reinterpret_cast<const Ctx*>(ctx)->t.ntot : -1; }
int64_t mw3_nflux(const mw3_ctx* ctx) {
  return ctx ? reinterpret_cast<const Ctx*>(ctx)->t.nflux : -1;
}
int32_t mw3_cheby_k(const mw3_ctx* ctx) { return ctx ? reinterpret_cast<const Ctx*>(ctx)->k : -1; }

int32_t mw3_hphi(mw3_ctx* ctx, const double* psi, double* hpsi) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !psi || !hpsi) return seterr(c, MW3_EINVAL, "mw3_hphi: NULL argument");
  size_t n = (size_t)c->t.ntot;
  HIPCHK(c, hipMemcpyAsync(c->d_a0, psi, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
         "mw3_hphi: upload");
  HIPCHK(c, launch_scatter_a0(c, c->d_a0, c->d_scratch), "mw3_hphi: scatter");
  HIPCHK(c, launch_to_f32(c, c->d_scratch, c->d_scr32), "mw3_hphi: FP32 copy");
  HIPCHK(c, launch_hpsi(c, c->d_scratch), "mw3_hphi: H psi");
  HIPCHK(c, launch_update(c, 0, c->d_scratch), "mw3_hphi: H psi");
  HIPCHK(c, launch_gather_a0(c, c->d_hout, c->d_a0b), "mw3_hphi: gather");
  HIPCHK(c, hipMemcpyAsync(hpsi, c->d_a0b, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
         "mw3_hphi: download");
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_hphi: sync");
  return MW3_OK;
}

static int clear_accumulators(Ctx* c) {
  const HostTables& t = c->t;
  int netot = std::max(c->s.netot, 0);
  HIPCHK(c, launch_zero(c, c->d_zcr, sizeof(double) * 2 * t.nflux * netot), "clear");
  HIPCHK(c, launch_zero(c, c->d_zcdr, sizeof(double) * 2 * t.nflux * netot), "clear");
  HIPCHK(c, launch_zero(c, c->d_zs2, sizeof(double) * 2 * (size_t)t.ncvj * netot), "clear");
  if (c->s.nofe > 0)
    HIPCHK(c, launch_zero(c, c->d_zfe, sizeof(double) * 2 * c->s.nofe * (size_t)t.nc * t.nslab),
           "clear");
  if (c->prod_on) {
    const size_t nx = (size_t)c->pr_nomb * c->pr_nvj;
    HIPCHK(c, launch_zero(c, c->d_zs2prod, sizeof(double) * 2 * nx * netot), "clear");
    HIPCHK(c, launch_zero(c, c->d_cvjp_last, sizeof(double) * nx), "clear");
  }
  c->ring_fill = 0;
  return MW3_OK;
}

int32_t mw3_cheby_begin(mw3_ctx* ctx, const double* psi0, double* stats3) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !psi0) return seterr(c, MW3_EINVAL, "mw3_cheby_begin: NULL argument");
  size_t n = (size_t)c->t.ntot;
  HIPCHK(c, hipMemcpyAsync(c->d_a0, psi0, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
         "mw3_cheby_begin: upload");
  HIPCHK(c, launch_scatter_a0(c, c->d_a0, c->d_psi), "mw3_cheby_begin: scatter");
  HIPCHK(c, launch_to_f32(c, c->d_psi, c->d_psi32), "mw3_cheby_begin: FP32 copy");
  size_t bytes = sizeof(double) * (size_t)c->t.nc * c->t.nslab;
  HIPCHK(c, hipMemcpyAsync(c->d_psi0, c->d_psi, bytes, hipMemcpyDeviceToDevice, c->stream),
         "mw3_cheby_begin: copy");
  // psi_1 = (H psi_0 - emindlt psi_0)/delta2 is written over psim1; the swap
  // then leaves psi_k = psi_1 and psi_{k-1} = psi_0
  HIPCHK(c, launch_hpsi(c, c->d_psi), "mw3_cheby_begin: H psi0");
  if (stats3) {
    HIPCHK(c, launch_update(c, 0, c->d_psi), "mw3_cheby_begin: H psi0");
    std::vector<double> h(n), p(psi0, psi0 + n);
    HIPCHK(c, launch_gather_a0(c, c->d_hout, c->d_a0b), "mw3_cheby_begin: gather");
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_a0b, n * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream),
           "mw3_cheby_begin: download");
    HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_cheby_begin: sync");
    double em = 0, x1 = 0, x2 = 0;
    for (size_t i = 0; i < n; i++) {
      em += p[i] * h[i];
      x1 += p[i] * p[i];
      x2 += h[i] * h[i];
    }
    stats3[0] = em;
    stats3[1] = x1;
    stats3[2] = x2;
  }
  c->ring_fill = 0;
  c->rec_it.clear();
  c->rec_cvj.clear();
  c->rec_cvjp.clear();
  HIPCHK(c, launch_update(c, 1, c->d_psi), "mw3_cheby_begin: first step");
  std::swap(c->d_psi, c->d_psim1);
  std::swap(c->d_psi32, c->d_psim132);
  c->k = 1;
  c->begun = true;
  int rc = clear_accumulators(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_cheby_begin: sync");
  return MW3_OK;
}

static int run_deferred(Ctx* c);

// Values of one step's Cvjprod record: the product projection's rows when one
// is attached and the energy sums run (netot > 0), else none.
static size_t rec_nprod(const Ctx* c) {
  return (c->prod_on && c->s.netot > 0) ? (size_t)c->pr_nomb * c->pr_nvj : 0;
}

static int flush_ring(Ctx* c, double* norms_dst, double* auto_dst) {
  int rd = run_deferred(c);
  if (rd) return rd;
  if (c->cvj_pending) {  // the reactant-state sums of the ring are done first
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_cvj, 0), "ring flush: wait");
    c->cvj_pending = false;
  }
  if (c->ring_fill == 0) return MW3_OK;
  HIPCHK(c, launch_energy_accum(c), "energy accumulation");
  std::vector<double> nb(2 * (size_t)c->ring_fill);
  HIPCHK(c, hipMemcpyAsync(nb.data(), c->d_ring_norm, nb.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream),
         "norms");
  if (c->rec_on) {  // the ring's Cvj (and Cvjprod, filled by the accumulation above) rows
    const size_t ns = (size_t)c->ring_fill, ncv = (size_t)c->t.ncvj;
    const size_t ncp = rec_nprod(c);
    const size_t o1 = c->rec_cvj.size(), o2 = c->rec_cvjp.size();
    c->rec_cvj.resize(o1 + ns * ncv);
    c->rec_cvjp.resize(o2 + ns * ncp);
    if (ncv)
      HIPCHK(c, hipMemcpyAsync(c->rec_cvj.data() + o1, c->d_ring_cvj, ns * ncv * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream),
             "records");
    if (ncp)
      HIPCHK(c, hipMemcpyAsync(c->rec_cvjp.data() + o2, c->d_ring_cvjp, ns * ncp * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream),
             "records");
    for (size_t i = 0; i < ns; i++) c->rec_it.push_back(c->ring_k0 + (int32_t)i);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
  for (int i = 0; i < c->ring_fill; i++) {
    if (norms_dst) norms_dst[i] = nb[2 * i];
    if (auto_dst) auto_dst[i] = nb[2 * i + 1];
  }
  c->ring_k0 += c->ring_fill;
  c->ring_fill = 0;
  if (c->timing) timer_collect(c);
  return MW3_OK;
}

// Steps run as captured graphs when the context allows it: graph mode on,
// not a step whose kernels are timed (timer events stay out of graphs; with
// sampled timing the other steps still run as graphs) and no psi(E)
// accumulation (its kernel takes the step index k as an argument).
static bool graphs_usable(const Ctx* c) {
  return c->graph_mode && !(c->timing && c->timing_step) && c->s.nofe <= 0;
}

// The captured step body of c on its own streams: on the second stream the
// projection sums of the step before (ring slot from the device counter,
// which they advance), beside this step's H psi; then, once both are done,
// the update (it rewrites the partials those sums read).  Every stream
// forked from c's main stream joins it again before the update.
static hipError_t capture_body(Ctx* c) {
  c->graph_slot_dev = true;
  c->slice_fresh = false;
  auto body = [&]() -> hipError_t {
    hipError_t r;
    if ((r = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess) return r;
    if ((r = hipStreamWaitEvent(c->aux, c->ev_fork, 0)) != hipSuccess) return r;
    if ((r = launch_tail_sums(c, 0, c->aux)) != hipSuccess) return r;
    if ((r = launch_slot_next(c, c->aux)) != hipSuccess) return r;
    if ((r = hipEventRecord(c->ev_cvj, c->aux)) != hipSuccess) return r;
    c->cvj_pending = true;
    if ((r = launch_hpsi(c, c->d_psi)) != hipSuccess) return r;
    if (!c->cvj_joined && (r = hipStreamWaitEvent(c->stream, c->ev_cvj, 0)) != hipSuccess)
      return r;
    return launch_update(c, 2, c->d_psi, true);
  };
  hipError_t e = body();
  c->graph_slot_dev = false;
  c->cvj_pending = false;  // the graph's second-stream work ends inside the graph
  c->cvj_joined = false;
  c->slice_fresh = false;
  return e;
}

static hipError_t end_capture(hipStream_t s, hipError_t eb, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (eb != hipSuccess) e = eb;
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  return e;
}

// Captures one step of c as graph q (psi in buffer q).
static hipError_t capture_step(Ctx* c, int q) {
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  return end_capture(c->stream, capture_body(c), &c->graph[q]);
}

void drop_graphs(Ctx* c) {
  for (auto& g : c->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// the projection sums of the last step, when a captured step left them
// pending, enqueued the ordinary way
static int run_deferred(Ctx* c) {
  if (c->deferred_slot < 0) return MW3_OK;
  const int slot = c->deferred_slot;
  c->deferred_slot = -1;
  HIPCHK(c, launch_projections(c, slot), "mw3_cheby_steps: projections");
  return MW3_OK;
}

// Step of c, first part: ring bookkeeping and whether the step may run as a
// graph.
static bool step_pre(Ctx* c) {
  if (c->ring_fill == 0) c->ring_k0 = c->k + 1;
  c->timing_step = (c->k % c->timing_every) == 0;
  return graphs_usable(c);
}

// Step of c, second part: the kernels, as a graph or one by one.
static int step_body(Ctx* c, bool graphs) {
  if (graphs && c->deferred_slot >= 0) {
    const int q = c->d_psi == c->graph_psi0 ? 0 : 1;
    if (!c->graph[q]) HIPCHK(c, capture_step(c, q), "mw3_cheby_steps: step capture");
    HIPCHK(c, hipGraphLaunch(c->graph[q], c->stream), "mw3_cheby_steps: step graph");
    return MW3_OK;
  }
  int rd = run_deferred(c);  // a graph step before this one left its sums pending
  if (rd) return rd;
  HIPCHK(c, launch_hpsi(c, c->d_psi), "mw3_cheby_steps: H psi");
  // the update rewrites the slice the last k_cvj reads (already waited for
  // through ev_ang when the angular term ran on the third stream)
  if (c->cvj_pending && !c->cvj_joined)
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_cvj, 0), "mw3_cheby_steps: wait");
  HIPCHK(c, launch_update(c, 2, c->d_psi, true), "mw3_cheby_steps: update");
  if (graphs) {  // the next step's graph takes this step's sums (slot counter set here)
    HIPCHK(c, launch_slot_set(c, c->ring_fill, c->stream), "mw3_cheby_steps: slot");
  } else {
    HIPCHK(c, launch_projections(c, c->ring_fill), "mw3_cheby_steps: projections");
    c->slice_fresh = true;  // nothing below writes psi before the next H psi
  }
  return MW3_OK;
}

// Step of c, last part: k, the psi buffers, and the ring flush when the ring
// is full or when this is step nsteps of the call; *done counts the call's
// steps (output index).
static int step_post(Ctx* c, bool graphs, int32_t* done, int32_t nsteps, double* norms,
                     double* autocorr) {
  const int base = *done - c->ring_fill;  // output index of ring slot 0
  if (graphs) c->deferred_slot = c->ring_fill;
  c->k += 1;
  if (!c->trial) HIPCHK(c, launch_psie(c, c->k), "mw3_cheby_steps: psi(E)");
  std::swap(c->d_psi, c->d_psim1);
  std::swap(c->d_psi32, c->d_psim132);
  c->ring_fill += 1;
  *done += 1;
  if (c->trial && (c->ring_fill == kRing || *done == nsteps)) {
    // autotuner trial: the step's sums end, the ring is discarded unread
    int rc = run_deferred(c);
    if (rc) return rc;
    if (c->cvj_pending) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_cvj, 0), "trial: wait");
    c->cvj_pending = false;
    c->ring_fill = 0;
    return MW3_OK;
  }
  if (c->ring_fill == kRing || *done == nsteps) {
    int rc = run_deferred(c);
    if (rc) return rc;
    c->timing_step = true;  // the energy sums of every ring are timed
    rc = flush_ring(c, norms ? norms + base : nullptr, autocorr ? autocorr + base : nullptr);
    if (rc) return rc;
  }
  c->timing_step = true;
  return MW3_OK;
}

// One Chebyshev step of c (H psi_k, update into psim1 = psi_{k+1},
// projections, swap).  In graph mode a step's projection sums run inside the
// next step's graph (or before the flush), on the second stream as before.
static int step_enqueue(Ctx* c, int32_t* done, int32_t nsteps, double* norms, double* autocorr) {
  const bool graphs = step_pre(c);
  int rc = step_body(c, graphs);
  if (rc) return rc;
  return step_post(c, graphs, done, nsteps, norms, autocorr);
}

int32_t mw3_cheby_steps(mw3_ctx* ctx, int32_t nsteps, double* norms, double* autocorr) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  if (!c->begun) return seterr(c, MW3_ESTATE, "mw3_cheby_steps: call mw3_cheby_begin first");
  if (nsteps < 0) return seterr(c, MW3_EINVAL, "mw3_cheby_steps: nsteps < 0");
  int32_t done = 0;
  int rc = MW3_OK;
  while (done < nsteps && rc == MW3_OK) rc = step_enqueue(c, &done, nsteps, norms, autocorr);
  c->slice_fresh = false;  // the caller may write psi before the next H psi
  return rc;
}

int32_t mw3_cheby_steps_batch(mw3_ctx** ctxs, int32_t n, int32_t nsteps, double* norms) {
  if (!ctxs || n < 1 || nsteps < 0) return seterr(nullptr, MW3_EINVAL, "mw3_cheby_steps_batch: bad arguments");
  int dev = -1;
  for (int i = 0; i < n; i++) {
    Ctx* c = reinterpret_cast<Ctx*>(ctxs[i]);
    if (!c) return seterr(nullptr, MW3_EINVAL, "mw3_cheby_steps_batch: NULL context");
    if (!c->begun) return seterr(c, MW3_ESTATE, "mw3_cheby_steps_batch: call mw3_cheby_begin first");
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return seterr(c, MW3_EINVAL, "mw3_cheby_steps_batch: repeated context");
    if (dev < 0) dev = c->device;
    if (c->device != dev)
      return seterr(c, MW3_EINVAL, "mw3_cheby_steps_batch: contexts on different devices");
  }
  std::vector<int32_t> done(n, 0);
  // step-major order: every context's step s is enqueued before any step s+1,
  // so the streams of all contexts hold work at the same time (each member
  // runs its steps as its own captured graphs where it can)
  int rc = MW3_OK;
  for (int32_t st = 0; st < nsteps && rc == MW3_OK; st++)
    for (int i = 0; i < n && rc == MW3_OK; i++) {
      Ctx* c = reinterpret_cast<Ctx*>(ctxs[i]);
      rc = step_enqueue(c, &done[i], nsteps, norms ? norms + (size_t)i * nsteps : nullptr,
                        nullptr);
    }
  for (int i = 0; i < n; i++)  // the caller may write psi before the next H psi
    reinterpret_cast<Ctx*>(ctxs[i])->slice_fresh = false;
  return rc;
}

int32_t mw3_flux_result(mw3_ctx* ctx, double* S2prodtot, double* S2prodelec) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  int netot = c->s.netot, ne = c->t.ne;
  if (netot <= 0) return MW3_OK;
  std::vector<double> sums((size_t)(1 + ne) * netot, 0.0);
  if (c->t.nflux > 0) {
    double* d_out = nullptr;
    HIPCHK(c, hipMalloc(&d_out, sizeof(double) * sums.size()), "mw3_flux_result: alloc");
    hipError_t e = launch_flux_result(c, d_out);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sums.data(), d_out, sizeof(double) * sums.size(), hipMemcpyDeviceToHost,
                         c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_out);
    HIPCHK(c, e, "mw3_flux_result");
  }
  for (int ie = 0; ie < netot; ie++) {
    double f = c->reacfct_host[ie] / c->s.ah1;
    if (S2prodtot) S2prodtot[ie] = sums[ie] * f;
    if (S2prodelec)
      for (int q = 0; q < ne; q++)
        S2prodelec[(size_t)q * netot + ie] = sums[(size_t)(1 + q) * netot + ie] * f;
  }
  return MW3_OK;
}

int32_t mw3_flux_energy_test(mw3_ctx* ctx, int32_t on) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  if (on != 0 && on != 1) return seterr(c, MW3_EINVAL, "mw3_flux_energy_test: on must be 0 or 1");
  c->flux_etest = on;
  return MW3_OK;
}

int32_t mw3_records(mw3_ctx* ctx, int32_t on) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  if (on != 0 && on != 1) return seterr(c, MW3_EINVAL, "mw3_records: on must be 0 or 1");
  c->rec_on = on;
  if (!on) {
    c->rec_it.clear();
    c->rec_cvj.clear();
    c->rec_cvjp.clear();
  }
  return MW3_OK;
}

int32_t mw3_records_get(mw3_ctx* ctx, int32_t maxrec, int32_t* it, double* cvj, double* cvjprod,
                        int32_t* nrec) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  if (maxrec < 0 || !nrec) return seterr(c, MW3_EINVAL, "mw3_records_get: maxrec < 0 or nrec NULL");
  const size_t n = std::min((size_t)maxrec, c->rec_it.size());
  const size_t ncv = (size_t)c->t.ncvj, ncp = rec_nprod(c);
  if (it) std::copy(c->rec_it.begin(), c->rec_it.begin() + n, it);
  if (cvj && ncv) std::copy(c->rec_cvj.begin(), c->rec_cvj.begin() + n * ncv, cvj);
  if (cvjprod && ncp) std::copy(c->rec_cvjp.begin(), c->rec_cvjp.begin() + n * ncp, cvjprod);
  c->rec_it.erase(c->rec_it.begin(), c->rec_it.begin() + n);
  c->rec_cvj.erase(c->rec_cvj.begin(), c->rec_cvj.begin() + n * ncv);
  c->rec_cvjp.erase(c->rec_cvjp.begin(), c->rec_cvjp.begin() + n * ncp);
  *nrec = (int32_t)n;
  return MW3_OK;
}

int32_t mw3_zs2_get(mw3_ctx* ctx, double* zS2) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !zS2) return MW3_EINVAL;
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  size_t n = 2 * (size_t)c->t.ncvj * std::max(c->s.netot, 0);
  HIPCHK(c, hipMemcpyAsync(zS2, c->d_zs2, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
         "mw3_zs2_get");
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_zs2_get");
  return MW3_OK;
}

int32_t mw3_flux_accum_get(mw3_ctx* ctx, double* zCR, double* zCdR) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  int netot = std::max(c->s.netot, 0);
  size_t n = 2 * (size_t)c->t.nflux * netot;
  if (zCR)
    HIPCHK(c, hipMemcpyAsync(zCR, c->d_zcr, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
           "mw3_flux_accum_get");
  if (zCdR)
    HIPCHK(c, hipMemcpyAsync(zCdR, c->d_zcdr, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
           "mw3_flux_accum_get");
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_flux_accum_get");
  // the reference accumulates only where E > V(flux point)
  for (int64_t f = 0; f < (c->flux_etest ? c->t.nflux : 0); f++)
    for (int e = 0; e < netot; e++)
      if (!(c->etot_host[e] > c->t.potiflux[f])) {
        if (zCR) zCR[2 * (f * netot + e)] = zCR[2 * (f * netot + e) + 1] = 0.0;
        if (zCdR) zCdR[2 * (f * netot + e)] = zCdR[2 * (f * netot + e) + 1] = 0.0;
      }
  return MW3_OK;
}

int32_t mw3_flux_accum_put(mw3_ctx* ctx, const double* zCR, const double* zCdR,
                           const double* zS2) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  int netot = std::max(c->s.netot, 0);
  size_t n = 2 * (size_t)c->t.nflux * netot;
  if (zCR)
    HIPCHK(c, hipMemcpyAsync(c->d_zcr, zCR, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
           "mw3_flux_accum_put");
  if (zCdR)
    HIPCHK(c, hipMemcpyAsync(c->d_zcdr, zCdR, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
           "mw3_flux_accum_put");
  if (zS2)
    HIPCHK(c, hipMemcpyAsync(c->d_zs2, zS2, 2 * (size_t)c->t.ncvj * netot * sizeof(double),
                             hipMemcpyHostToDevice, c->stream),
           "mw3_flux_accum_put");
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_flux_accum_put");
  return MW3_OK;
}

int32_t mw3_psie_get(mw3_ctx* ctx, double* zfe) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !zfe) return MW3_EINVAL;
  if (c->s.nofe <= 0) return MW3_OK;
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  size_t nd = (size_t)c->t.nc * c->t.nslab;
  std::vector<double> dense(2 * nd);
  size_t n = (size_t)c->t.ntot;
  for (int e = 0; e < c->s.nofe; e++) {
    HIPCHK(c, hipMemcpyAsync(dense.data(), c->d_zfe + 2 * e * nd, 2 * nd * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream),
           "mw3_psie_get");
    HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_psie_get");
    for (size_t i = 0; i < n; i++) {
      zfe[2 * (e * n + i)] = dense[2 * (size_t)c->t.pt_dense[i]];
      zfe[2 * (e * n + i) + 1] = dense[2 * (size_t)c->t.pt_dense[i] + 1];
    }
  }
  return MW3_OK;
}

static void free_prod(Ctx* c) {
  void** ptrs[] = {&c->d_tprod,
                   (void**)&c->d_ppart,
                   (void**)&c->d_dtrans,
                   (void**)&c->d_ring_prod,
                   (void**)&c->d_prodwf,
                   (void**)&c->d_prod_part,
                   (void**)&c->d_ring_cvjp,
                   (void**)&c->d_cvjp_last,
                   (void**)&c->d_zs2prod};
  for (void** p : ptrs)
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  if (c->p_prod.d_descs) {
    (void)hipFree(c->p_prod.d_descs);
    c->p_prod.d_descs = nullptr;
  }
  c->prod_on = false;
}

int32_t mw3_prod_attach(mw3_ctx* ctx, const mw3_prod_setup* ps) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !ps) return MW3_EINVAL;
  const HostTables& t = c->t;
  if (c->s.iprod != 2)
    return seterr(c, MW3_EINVAL, "mw3_prod_attach: the context was created with iprod != 2");
  if (c->s.netot <= 0)
    return seterr(c, MW3_EINVAL, "mw3_prod_attach: the context has no energy grid (netot = 0)");
  if (ps->n2prod0 < 1 || ps->n2prod1 > t.n2 || ps->n2prod0 > ps->n2prod1 ||
      ps->nangproj0 < 1 || ps->nangproj1 > t.na || ps->nangproj0 > ps->nangproj1 ||
      ps->iomminprod < 0 || ps->iomminprod > ps->iommaxprod || ps->iommaxprod > c->s.Jtot ||
      ps->nviniprod > ps->nvmaxprod || ps->jiniprod > ps->jmaxprod || !ps->noBCprod ||
      !ps->transRgjac || !ps->prodwf || !ps->dtrans)
    return seterr(c, MW3_EINVAL, "mw3_prod_attach: inconsistent ranges or NULL table");
  const int njp = ps->jmaxprod - ps->jiniprod + 1, nvp = ps->nvmaxprod - ps->nviniprod + 1;
  for (int j = 0; j < njp; j++)
    if (ps->noBCprod[j] > ps->nvmaxprod)
      return seterr(c, MW3_EINVAL, "mw3_prod_attach: noBCprod > nvmaxprod");
  if (t.nc > kProdMaxCh)
    return seterr(c, MW3_EUNSUP, "mw3_prod_attach: more than 64 channels");
  // steps already in the ring are accumulated without the product projection
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  drop_graphs(c);  // captured steps hold the kernels of the old set-up
  free_prod(c);
  c->rec_it.clear();  // held records have the old Cvjprod shape
  c->rec_cvj.clear();
  c->rec_cvjp.clear();
  c->pr_n2p0 = ps->n2prod0 - 1;
  c->pr_nap0 = ps->nangproj0 - 1;
  c->pr_n2p = ps->n2prod1 - ps->n2prod0 + 1;
  c->pr_nap = ps->nangproj1 - ps->nangproj0 + 1;
  c->pr_nomb = ps->iommaxprod - ps->iomminprod + 1;
  c->pr_nvj = njp * nvp;
  c->pr_L = (int64_t)c->pr_nap * c->pr_n2p;
  c->pr_shape = {ps->iomminprod, ps->iommaxprod, ps->nviniprod, ps->nvmaxprod, ps->jiniprod,
                 ps->jmaxprod};
  const int64_t L = c->pr_L, K = (int64_t)t.ne * L;
  // K chunks of the flush GEMM: about 512 blocks of 64 x 64 outputs
  const int ntn = (c->pr_nvj + 63) / 64;
  int nchunk = std::max(1, 512 / std::max(1, c->pr_nomb * ntn));
  int64_t kc = std::max<int64_t>(256, (((K + nchunk - 1) / nchunk) + 15) / 16 * 16);
  c->pr_kc = (int)kc;
  c->pr_nchunk = (int)((K + kc - 1) / kc);
  // transRgjac on the dense grid, zero outside the projection range; one
  // table when every channel holds the same values (the reference fills them
  // alike, SRC/mod_coortrans01_02.f:356-366)
  std::vector<double> T0((size_t)t.nslab, 0.0);
  bool same = true;
  std::vector<uint8_t> seen((size_t)t.nslab, 0);
  for (int64_t i = 0; i < t.ntot; i++) {
    const int64_t d = t.pt_dense[i], cc = d / t.nslab, pl = d % t.nslab;
    const int ia = (int)(pl / t.nplane), R = (int)(pl % t.n2);
    double v = ps->transRgjac[i];
    if (ia < c->pr_nap0 || ia >= c->pr_nap0 + c->pr_nap || R < c->pr_n2p0 ||
        R >= c->pr_n2p0 + c->pr_n2p)
      v = 0.0;
    if (!seen[pl]) {
      T0[pl] = v;
      seen[pl] = 1;
    } else if (T0[pl] != v) {
      same = false;
    }
    (void)cc;
  }
  std::vector<double> Tall;
  const std::vector<double>* Tsrc = &T0;
  if (!same) {
    Tall.assign((size_t)t.nc * t.nslab, 0.0);
    for (int64_t i = 0; i < t.ntot; i++) {
      const int64_t d = t.pt_dense[i], pl = d % t.nslab;
      const int ia = (int)(pl / t.nplane), R = (int)(pl % t.n2);
      if (ia < c->pr_nap0 || ia >= c->pr_nap0 + c->pr_nap || R < c->pr_n2p0 ||
          R >= c->pr_n2p0 + c->pr_n2p)
        continue;
      Tall[(size_t)d] = ps->transRgjac[i];
    }
    Tsrc = &Tall;
  }
  c->pr_tstride = same ? 0 : t.nslab;
  // prodwf with the states the reference leaves out zeroed
  std::vector<double> wf((size_t)c->pr_nomb * c->pr_nvj * K);
  for (int o = 0; o < c->pr_nomb; o++)
    for (int jj = 0; jj < njp; jj++)
      for (int iv = 0; iv < nvp; iv++) {
        const int j = ps->jiniprod + jj, v = ps->nviniprod + iv, om = ps->iomminprod + o;
        const bool valid = j >= om && v <= ps->noBCprod[jj];
        const size_t off = (((size_t)o * njp + jj) * nvp + iv) * (size_t)K;
        for (int64_t k = 0; k < K; k++) wf[off + k] = valid ? ps->prodwf[off + k] : 0.0;
      }
  const size_t nx = (size_t)c->pr_nomb * c->pr_nvj;
  const int netot = c->s.netot;
  hipError_t e = c->precision == MW3_MIXED ? upload<float>(&c->d_tprod, *Tsrc)
                                           : upload<double>(&c->d_tprod, *Tsrc);
  if (e == hipSuccess)
    e = upload_raw(&c->d_dtrans, ps->dtrans, (size_t)t.nc * c->pr_nomb * L);
  if (e == hipSuccess) e = upload_raw(&c->d_prodwf, wf.data(), wf.size());
  if (e == hipSuccess)
    e = alloc_zero((void**)&c->d_ppart, sizeof(double) * 64 * std::max(1, c->nnorm_blocks));
  if (e == hipSuccess)
    e = alloc_zero((void**)&c->d_ring_prod, sizeof(double) * kRing * (size_t)c->pr_nomb * K);
  if (e == hipSuccess)
    e = alloc_zero((void**)&c->d_prod_part,
                   sizeof(double) * (size_t)c->pr_nomb * c->pr_nchunk * kRing * c->pr_nvj);
  if (e == hipSuccess) e = alloc_zero((void**)&c->d_ring_cvjp, sizeof(double) * kRing * nx);
  if (e == hipSuccess) e = alloc_zero((void**)&c->d_cvjp_last, sizeof(double) * nx);
  if (e == hipSuccess) e = alloc_zero((void**)&c->d_zs2prod, sizeof(double) * 2 * nx * netot);
  if (e == hipSuccess) e = build_prod_plan(c);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    free_prod(c);
    return seterr(c, e == hipErrorOutOfMemory ? MW3_ENOMEM : MW3_EDEVICE, "mw3_prod_attach", e);
  }
  c->prod_on = true;
  return MW3_OK;
}

static int prod_copy(Ctx* c, double* host, const double* host_in, double* dev, size_t n,
                     const char* what) {
  if (!c->prod_on) return seterr(c, MW3_ESTATE, (std::string(what) + ": no product tables (mw3_prod_attach)").c_str());
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  if (host)
    HIPCHK(c, hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), what);
  else
    HIPCHK(c, hipMemcpyAsync(dev, host_in, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
           what);
  HIPCHK(c, hipStreamSynchronize(c->stream), what);
  return MW3_OK;
}

int32_t mw3_zs2prod_get(mw3_ctx* ctx, double* zS2prod) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !zS2prod) return MW3_EINVAL;
  size_t n = 2 * (size_t)c->pr_nomb * c->pr_nvj * std::max(c->s.netot, 0);
  return prod_copy(c, zS2prod, nullptr, c->d_zs2prod, n, "mw3_zs2prod_get");
}

int32_t mw3_zs2prod_put(mw3_ctx* ctx, const double* zS2prod) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !zS2prod) return MW3_EINVAL;
  size_t n = 2 * (size_t)c->pr_nomb * c->pr_nvj * std::max(c->s.netot, 0);
  return prod_copy(c, nullptr, zS2prod, c->d_zs2prod, n, "mw3_zs2prod_put");
}

int32_t mw3_cvjprod_get(mw3_ctx* ctx, double* cvjprod) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !cvjprod) return MW3_EINVAL;
  size_t n = (size_t)c->pr_nomb * c->pr_nvj;
  return prod_copy(c, cvjprod, nullptr, c->d_cvjp_last, n, "mw3_cvjprod_get");
}

int32_t mw3_state_get(mw3_ctx* ctx, double* psi_km1, double* psi_k, int32_t* k) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  size_t n = (size_t)c->t.ntot;
  if (psi_km1) {
    HIPCHK(c, launch_gather_a0(c, c->d_psim1, c->d_a0), "mw3_state_get");
    HIPCHK(c, hipMemcpyAsync(psi_km1, c->d_a0, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
           "mw3_state_get");
  }
  if (psi_k) {
    HIPCHK(c, launch_gather_a0(c, c->d_psi, c->d_a0b), "mw3_state_get");
    HIPCHK(c, hipMemcpyAsync(psi_k, c->d_a0b, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
           "mw3_state_get");
  }
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_state_get");
  if (k) *k = c->k;
  return MW3_OK;
}

int32_t mw3_state_put(mw3_ctx* ctx, const double* psi_km1, const double* psi_k, int32_t k,
                      const double* psi0) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !psi_km1 || !psi_k) return seterr(c, MW3_EINVAL, "mw3_state_put: NULL argument");
  if (k < 1) return seterr(c, MW3_EINVAL, "mw3_state_put: k must be >= 1");
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  size_t n = (size_t)c->t.ntot;
  HIPCHK(c, hipMemcpyAsync(c->d_a0, psi_km1, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
         "mw3_state_put");
  HIPCHK(c, launch_scatter_a0(c, c->d_a0, c->d_psim1), "mw3_state_put");
  HIPCHK(c, hipMemcpyAsync(c->d_a0b, psi_k, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
         "mw3_state_put");
  HIPCHK(c, launch_scatter_a0(c, c->d_a0b, c->d_psi), "mw3_state_put");
  HIPCHK(c, launch_to_f32(c, c->d_psi, c->d_psi32), "mw3_state_put: FP32 copy");
  HIPCHK(c, launch_to_f32(c, c->d_psim1, c->d_psim132), "mw3_state_put: FP32 copy");
  if (psi0) {
    HIPCHK(c, hipMemcpyAsync(c->d_a0, psi0, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
           "mw3_state_put");
    HIPCHK(c, launch_scatter_a0(c, c->d_a0, c->d_psi0), "mw3_state_put");
  }
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_state_put");
  c->k = k;
  c->begun = true;
  return MW3_OK;
}

int32_t mw3_set_option(mw3_ctx* ctx, int32_t option, int32_t value) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  int rc = flush_ring(c, nullptr, nullptr);  // no step of the old setting left pending
  if (rc) return rc;
  switch (option) {
    case MW3_OPT_GRAPHS:
      if (value != 0 && value != 1) return seterr(c, MW3_EINVAL, "mw3_set_option: graphs 0 or 1");
      c->graph_mode = value;
      return MW3_OK;
    case MW3_OPT_RADIAL:
      if (value != 0 && value != 1) return seterr(c, MW3_EINVAL, "mw3_set_option: radial 0 or 1");
      if (value != c->opt_radial) drop_graphs(c);
      c->opt_radial = value;
      return MW3_OK;
    case MW3_OPT_ANGULAR:
      if (value < 0 || value > 2) return seterr(c, MW3_EINVAL, "mw3_set_option: angular 0..2");
      if (value != c->opt_angular) drop_graphs(c);
      c->opt_angular = value;
      return MW3_OK;
    case MW3_OPT_SIDE_STREAM:
      if (value != 0 && value != 1)
        return seterr(c, MW3_EINVAL, "mw3_set_option: side stream 0 or 1");
      if (value != c->opt_side) drop_graphs(c);
      c->opt_side = value;
      return MW3_OK;
    case MW3_OPT_ANGULAR_GRID:
      if (value != 1 && value != 2)
        return seterr(c, MW3_EINVAL, "mw3_set_option: angular grid divisor 1 or 2");
      if (value != c->opt_ang_grid) drop_graphs(c);
      c->opt_ang_grid = value;
      return MW3_OK;
    case MW3_OPT_RADIAL_GRID:
      if (value != 1 && value != 2 && value != 4)
        return seterr(c, MW3_EINVAL, "mw3_set_option: radial grid divisor 1, 2 or 4");
      if (value != c->opt_rad_grid) drop_graphs(c);
      c->opt_rad_grid = value;
      return MW3_OK;
    case MW3_OPT_FP16_PRODUCTS:
      if (value != 0 && value != 1)
        return seterr(c, MW3_EINVAL, "mw3_set_option: fp16 products 0 or 1");
      if (value != c->opt_fp16) drop_graphs(c);
      c->opt_fp16 = value;
      return MW3_OK;
    default:
      return seterr(c, MW3_EINVAL, "mw3_set_option: unknown option");
  }
}

// the variant each option value selects in this context (the fallback where
// the kernel of that value is not built): radial 1 needs the persistent R
// passes or the fused r passes; angular 0 the folded kernel, 1 a one-pass one
static void effective_options(const Ctx* c, int32_t* out) {
  out[MW3_OPT_GRAPHS] = c->graph_mode;
  out[MW3_OPT_RADIAL] = c->opt_radial && (c->rad_nitems || c->mid_nitems) ? 1 : 0;
  int a = c->opt_angular;
  const bool cor_fold = c->ang_fold_cor && c->precision == MW3_MIXED && c->opt_fp16;
  if (a == 0 && !c->ang_fold && !cor_fold) a = 1;
  if (a == 1 && !c->ang_nt && !c->ang_fold) a = 2;
  if (a == 1 && !c->ang_nt) a = 0;
  out[MW3_OPT_ANGULAR] = a;
  out[MW3_OPT_SIDE_STREAM] = c->opt_side;
  // the grid divisor applies to the staged R-pass kernels only: the FP32
  // passes, and the FP64 ones with their table split over block pairs, run
  // with the table in LDS on a grid of their own
  const bool staged = c->precision != MW3_MIXED && !c->rad_lds64;
  out[MW3_OPT_RADIAL_GRID] = c->rad_nitems && c->opt_radial && staged ? c->opt_rad_grid : 1;
  out[MW3_OPT_ANGULAR_GRID] =
      (c->ang_fold || cor_fold) && out[MW3_OPT_ANGULAR] == 0 ? c->opt_ang_grid : 1;
  out[MW3_OPT_FP16_PRODUCTS] = c->precision == MW3_MIXED ? c->opt_fp16 : 0;
}

int32_t mw3_get_option(const mw3_ctx* ctx, int32_t option, int32_t* value) {
  const Ctx* c = reinterpret_cast<const Ctx*>(ctx);
  if (!c || !value) return MW3_EINVAL;
  if (option < 0 || option >= MW3_NOPTIONS) return MW3_EINVAL;
  int32_t all[MW3_NOPTIONS];
  effective_options(c, all);
  *value = all[option];
  return MW3_OK;
}

// One autotuner trial: the context's current options over 2 + n steps from
// the current state, timed over the last n with events on the main stream
// once every stream has joined it.  No ring flush, no psi(E) sum.
static int trial_ms(Ctx* c, int n, hipEvent_t e0, hipEvent_t e1, double* ms) {
  int32_t done = 0;
  c->trial = true;
  int rc = MW3_OK;
  while (done < 2 && rc == MW3_OK) rc = step_enqueue(c, &done, 2, nullptr, nullptr);
  if (rc == MW3_OK) {
    HIPCHK(c, hipEventRecord(e0, c->stream), "mw3_autotune: event");
    done = 0;
    while (done < n && rc == MW3_OK) rc = step_enqueue(c, &done, n, nullptr, nullptr);
  }
  c->trial = false;
  if (rc) return rc;
  HIPCHK(c, hipEventRecord(e1, c->stream), "mw3_autotune: event");
  HIPCHK(c, hipEventSynchronize(e1), "mw3_autotune: sync");
  float t = 0.f;
  HIPCHK(c, hipEventElapsedTime(&t, e0, e1), "mw3_autotune: elapsed");
  *ms = (double)t / n;
  return MW3_OK;
}

int32_t mw3_autotune(mw3_ctx* ctx, int32_t trial_steps, int32_t* chosen, double* ms_per_step) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  if (trial_steps < 2 || trial_steps > kRing - 2)
    return seterr(c, MW3_EINVAL, "mw3_autotune: trial_steps must be in 2..254");
  const int n = trial_steps & ~1;  // even: the psi buffers end where they began
  int rc = flush_ring(c, nullptr, nullptr);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_autotune: sync");
  const bool timing = c->timing;
  c->timing = false;
  // the state is saved in the two spare dense slabs and restored at the end
  const size_t bytes = sizeof(double) * (size_t)c->t.nc * c->t.nslab;
  double* const p0 = c->d_psi;
  double* const pm0 = c->d_psim1;
  HIPCHK(c, hipMemcpyAsync(c->d_hout, c->d_psi, bytes, hipMemcpyDeviceToDevice, c->stream),
         "mw3_autotune: save");
  HIPCHK(c, hipMemcpyAsync(c->d_scratch, c->d_psim1, bytes, hipMemcpyDeviceToDevice, c->stream),
         "mw3_autotune: save");
  const int k0 = c->k, rk0 = c->ring_k0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(c, hipEventCreate(&e0), "mw3_autotune: event");
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return seterr(c, MW3_EDEVICE, "mw3_autotune: event", he);
  }
  // candidates: every combination of the options, skipping values that fall
  // back to a combination already listed
  typedef std::array<int32_t, MW3_NOPTIONS> Key;
  auto apply = [&](const Key& k) {
    c->graph_mode = k[MW3_OPT_GRAPHS];
    c->opt_radial = k[MW3_OPT_RADIAL];
    c->opt_angular = k[MW3_OPT_ANGULAR];
    c->opt_side = k[MW3_OPT_SIDE_STREAM];
    c->opt_rad_grid = k[MW3_OPT_RADIAL_GRID];
    c->opt_ang_grid = k[MW3_OPT_ANGULAR_GRID];
    c->opt_fp16 = k[MW3_OPT_FP16_PRODUCTS];
  };
  Key best_key;
  effective_options(c, best_key.data());
  std::vector<Key> cands;
  for (int g = 0; g <= 1; g++)
    for (int r = 0; r <= 1; r++)
      for (int a = 0; a <= 2; a++)
        for (int sd = 0; sd <= 1; sd++)
          for (int gd = 1; gd <= 4; gd *= 2)
            for (int ag = 1; ag <= 2; ag++) {
              Key k;
              k[MW3_OPT_GRAPHS] = g;
              k[MW3_OPT_RADIAL] = r;
              k[MW3_OPT_ANGULAR] = a;
              k[MW3_OPT_SIDE_STREAM] = sd;
              k[MW3_OPT_RADIAL_GRID] = gd;
              k[MW3_OPT_ANGULAR_GRID] = ag;
              k[MW3_OPT_FP16_PRODUCTS] = best_key[MW3_OPT_FP16_PRODUCTS];  // kept as set
              apply(k);
              Key eff;
              effective_options(c, eff.data());
              if (std::find(cands.begin(), cands.end(), eff) == cands.end()) cands.push_back(eff);
            }
  double best = 1e30;
  for (const auto& key : cands) {
    drop_graphs(c);
    apply(key);
    double ms = 0;
    rc = trial_ms(c, n, e0, e1, &ms);
    if (rc) break;
    if (ms < best) {
      best = ms;
      best_key = key;
    }
  }
  drop_graphs(c);
  apply(best_key);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->timing = timing;
  // restore the state (an even number of trial steps leaves the buffers in place)
  c->d_psi = p0;
  c->d_psim1 = pm0;
  c->k = k0;
  c->ring_k0 = rk0;
  c->ring_fill = 0;
  c->deferred_slot = -1;
  c->slice_fresh = false;
  HIPCHK(c, hipMemcpyAsync(c->d_psi, c->d_hout, bytes, hipMemcpyDeviceToDevice, c->stream),
         "mw3_autotune: restore");
  HIPCHK(c, hipMemcpyAsync(c->d_psim1, c->d_scratch, bytes, hipMemcpyDeviceToDevice, c->stream),
         "mw3_autotune: restore");
  HIPCHK(c, launch_to_f32(c, c->d_psi, c->d_psi32), "mw3_autotune: FP32 copy");
  HIPCHK(c, launch_to_f32(c, c->d_psim1, c->d_psim132), "mw3_autotune: FP32 copy");
  HIPCHK(c, hipStreamSynchronize(c->stream), "mw3_autotune: sync");
  if (rc) return rc;
  if (chosen) effective_options(c, chosen);
  if (ms_per_step) *ms_per_step = best;
  return MW3_OK;
}

int32_t mw3_timing_enable(mw3_ctx* ctx, int32_t on) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return MW3_EINVAL;
  timer_collect(c);
  if (on < 0) return seterr(c, MW3_EINVAL, "mw3_timing_enable: on < 0");
  c->timing = on != 0;
  c->timing_every = on > 1 ? on : 1;
  c->timing_step = true;
  for (int i = 0; i < MW3_NKERNELS; i++) {
    c->t_ms[i] = 0;
    c->t_n[i] = 0;
  }
  return MW3_OK;
}

int32_t mw3_timing_get(mw3_ctx* ctx, int32_t id, double* total_ms, int64_t* launches) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || id < 0 || id >= MW3_NKERNELS) return MW3_EINVAL;
  timer_collect(c);
  if (total_ms) *total_ms = c->t_ms[id];
  if (launches) *launches = c->t_n[id];
  return MW3_OK;
}

}  // extern "C"

// Device address of this device's coefficient array, looked up once per
// device with that device current.
static hipError_t scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

int32_t mw3_cheby_spectrum_noalloc(const double* d_autoc, int32_t kmax, double emindlt,
                                   double delta2, double hbr, double xkgam, const double* d_e,
                                   int32_t ne, double* d_spectrum, void* stream) {
  if (!d_autoc || !d_e || !d_spectrum || kmax < 0 || ne < 0)
    return seterr(nullptr, MW3_EINVAL, "mw3_cheby_spectrum_noalloc: bad arguments");
  if (!(delta2 > 0.0))
    return seterr(nullptr, MW3_EINVAL, "mw3_cheby_spectrum_noalloc: delta2 must be positive");
  if (kmax >= kSpecScratch) {
    char msg[160];
    snprintf(msg, sizeof msg,
             "mw3_cheby_spectrum_noalloc: kmax = %d exceeds the largest supported, %d",
             (int)kmax, kSpecScratch - 1);
    return seterr(nullptr, MW3_EINVAL, msg);
  }
  if (ne == 0) return MW3_OK;
  int dev = -1;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess)
    return seterr(nullptr, MW3_EDEVICE, "mw3_cheby_spectrum_noalloc: no HIP device", err);
  if (dev < 0 || dev >= kMaxDevices)
    return seterr(nullptr, MW3_EUNSUP, "mw3_cheby_spectrum_noalloc: device index above 63");
  double* coef = nullptr;
  err = scratch_address(dev, &coef);
  if (err != hipSuccess)
    return seterr(nullptr, MW3_EDEVICE, "mw3_cheby_spectrum_noalloc: coefficient array", err);
  // the coefficient kernel writes all kmax + 1 entries before the sum reads
  // them (one stream), so nothing of an earlier call needs clearing
  err = mw3::launch_spectrum(static_cast<hipStream_t>(stream), d_autoc, coef, kmax, hbr, xkgam,
                             d_e, ne, emindlt, delta2, d_spectrum);
  if (err != hipSuccess)
    return seterr(nullptr, MW3_EDEVICE, "mw3_cheby_spectrum_noalloc: launch", err);
  return MW3_OK;
}

int32_t mw3_cheby_spectrum(const double* autoc, int32_t kmax, double emindlt, double delta2,
                           double hbr, double xkgam, const double* e, int32_t ne,
                           double* spectrum) {
  if (!autoc || !e || !spectrum || kmax < 0 || ne < 0)
    return seterr(nullptr, MW3_EINVAL, "mw3_cheby_spectrum: bad arguments");
  if (!(delta2 > 0.0)) return seterr(nullptr, MW3_EINVAL, "mw3_cheby_spectrum: delta2 must be positive");
  if (ne == 0) return MW3_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return seterr(nullptr, MW3_EDEVICE, "mw3_cheby_spectrum: no HIP device");
  double *d_auto = nullptr, *d_a = nullptr, *d_e = nullptr, *d_s = nullptr;
  hipStream_t st = nullptr;
  const size_t nk = (size_t)kmax + 1;
  hipError_t err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (err == hipSuccess) err = hipMalloc((void**)&d_auto, nk * sizeof(double));
  if (err == hipSuccess) err = hipMalloc((void**)&d_a, nk * sizeof(double));
  if (err == hipSuccess) err = hipMalloc((void**)&d_e, (size_t)ne * sizeof(double));
  if (err == hipSuccess) err = hipMalloc((void**)&d_s, (size_t)ne * sizeof(double));
  if (err == hipSuccess)
    err = hipMemcpyAsync(d_auto, autoc, nk * sizeof(double), hipMemcpyHostToDevice, st);
  if (err == hipSuccess)
    err = hipMemcpyAsync(d_e, e, (size_t)ne * sizeof(double), hipMemcpyHostToDevice, st);
  if (err == hipSuccess)
    err = mw3::launch_spectrum(st, d_auto, d_a, kmax, hbr, xkgam, d_e, ne, emindlt, delta2, d_s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(spectrum, d_s, (size_t)ne * sizeof(double), hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (d_auto) (void)hipFree(d_auto);
  if (d_a) (void)hipFree(d_a);
  if (d_e) (void)hipFree(d_e);
  if (d_s) (void)hipFree(d_s);
  if (st) (void)hipStreamDestroy(st);
  if (err != hipSuccess)
    return seterr(nullptr, err == hipErrorOutOfMemory ? MW3_ENOMEM : MW3_EDEVICE,
                  "mw3_cheby_spectrum", err);
  return MW3_OK;
}
