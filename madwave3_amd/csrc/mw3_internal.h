// mw3_internal.h -- internal types of libmw3 (not part of the C ABI).
//
// Device data model (one context = one propagation on one GPU):
//   dense slabs  X[ch][iang][ir1][ir2]   ir2 (R) fastest, iang = global angle-1
//   mask         uint8 [iang][ir1][ir2]  1 on the kept (compressed) points
//   V            [et][es][iang][ir1][ir2] (0 off the mask)
//   FBR          Xt[ch][jj][ir1][ir2]    angular functions index jj
// The compressed A0 vector of the ABI maps to the slabs through pt_dense.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mw3.h"

namespace mw3 {

// Chebyshev steps buffered between energy sums: the FP64 accumulators (about
// 1 GB at 500 energies and 61,440 flux points) are read and written once per
// ring, so a longer ring moves fewer accumulator bytes per step
constexpr int kRing = 256;

// Host-side operator tables built from mw3_setup (mw3_setup.cpp).
struct HostTables {
  // sizes
  int n1 = 0, n2 = 0, na = 0, nc = 0, ne = 0, nJ = 0, nJp = 0;
  int64_t ntot = 0, nplane = 0, nslab = 0;  // nplane = n1*n2, nslab = na*nplane
  // A0 <-> dense
  std::vector<uint32_t> pt_dense;  // [ntot] dense index (ch,iang,ir1,ir2)
  std::vector<uint8_t> mask;       // [na][n1][n2]
  std::vector<double> vdense;      // [ne][ne][na][n1][n2]  (et, es)
  std::vector<int32_t> ch_omega, ch_elec;
  // Tradial: per global angle DST size in r and kinetic tables
  std::vector<int32_t> nn1;     // [na]
  std::vector<double> p2r1;     // [na][n1]
  std::vector<double> p2r2;     // [n2]
  std::vector<double> divi;     // [na]
  double radcut = 0;
  // RODFT00 tables for the folded DST: for size n, h = n/2, ke = n - h,
  // S_e[j][k'] = S[j][2k'] (ke x ke) then S_o[j][k'] = S[j][2k'+1] (h x h,
  // stored in a ke x ke block), S[j][k] = 2 sin(pi (j+1)(k+1)/(n+1))
  std::vector<int32_t> dst_sizes;        // distinct DST lengths (r sizes + n2)
  std::vector<int64_t> dst_table_off;    // offset of each size's S_e block
  std::vector<int> dst_ld;               // leading dimension of each size's blocks
  std::vector<int64_t> low_off;          // offset of each size's low-mode table L[k][j] = S(j,k)
  std::vector<int> low_ld;               // its row stride
  std::vector<int32_t> k1lo, k2lo;       // [na] r / R modes with p2r1 + p2r2 < radcut possible
  // tables of the one-wave-per-row-tile R passes (n2 a multiple of 64): with
  // h = n2 / 2 and k2p = 32 rad_nkb columns k2 (zero for k2 >= n2),
  // Ti[R][k2] = S(R, k2) (R < h) at rad_ti_off and Tt[k2][j] = S(j, k2)
  // (j < h) at rad_tt_off; rad_nkb = 0 when they are not built
  int rad_nkb = 0;
  int64_t rad_ti_off = 0, rad_tt_off = 0;
  std::vector<double> dst_tables;
  // angular FBR
  std::vector<int32_t> jlist;   // [nJ] j values of the FBR axis
  std::vector<double> Dmat;     // [nc][na][nJp]  D_c(iang, j), j stride nJp = nJ rounded to 4
  std::vector<double> acoef;    // [nc][nJ][n1]   j^2 term (angkinj2 FBR)
  std::vector<double> bcoef;    // [nc][3][nJ][n2] l^2 term, q = Omega_s-Omega_t+1
  std::vector<int32_t> nbr;     // [nc][3] source channel for q, -1 if none
  // flux line
  int ir4flux = 0;
  std::vector<int32_t> ir1flux;          // [na]
  std::vector<double> zar1flux;          // [n1][2]
  int64_t nflux = 0;
  std::vector<uint32_t> flux_dense;      // [nflux] dense index of the point
  std::vector<uint32_t> flux_col;        // [nflux] dense index of (ch,iang,0,ir2)
  std::vector<double> potiflux;          // [nflux]
  std::vector<int32_t> flux_elec;        // [nflux] electronic state (0-based)
  // reactant-state projection (Cvjflux_k)
  int ncvj = 0;                          // nc * nj * nv
  std::vector<double> fdv;               // [nc][nj][nv][n1] fd of the channel's state, 0 if unused
  std::vector<double> dcvj;              // [nc][nj][na] Djmmp(iang, j, e, Omega)
  std::vector<uint8_t> cvj_valid;        // [nc][nj][nv]
  std::vector<uint8_t> cvj_isref;        // [nc][nj][nv]
};

// batched GEMM pass: C[m][n] = sum_k A[m][k] B[k][n] per descriptor.
// Folded DST-I: with fold_op = 1 (A) or 2 (B) the operand's k-th element is
// x(k) + fold_sgn x(fold_n-1-k) for k < fold_n/2 (x(k) at the middle of an
// odd fold_n).  kin_ia >= 0 scales output row m (DST index rmul*m + radd) by
// the radial kinetic factor of angle kin_ia.
struct GemmDesc {
  int64_t a_off, b_off, c_off;
  int M, N, K, lda_m;  // A element (m,k) at a_off + m*lda_m + k*lda_k
  int lda_k, fold_op, fold_n, fold_sgn;
  int kin_ia, rmul, radd, kin_mode;  // kin_mode 0: min(E, c) divi; 1: (c - E)_+ divi
};

struct GemmStrides {
  int64_t ldb_k, ldb_n;
  int64_t ldc_m, ldc_n;
};

// one pass of the step as a batch of GEMMs; vec: the operands meet the
// alignment rules of the vectorised kernel; group / share_b: block order
// (consecutive descriptors sharing the A panel / blocks sharing the B panel
// kept on one XCD)
struct GemmPlan {
  GemmDesc* d_descs = nullptr;
  int nb = 0, maxM = 0, maxN = 0, vec = 0, group = 1, share_b = 0;
  GemmStrides st{};
};

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
};

struct Ctx {
  mw3_setup s{};  // scalar copy (pointers not kept)
  int precision = MW3_FP64;
  HostTables t;
  std::string err;
  hipStream_t stream = nullptr;
  // second stream (default priority) for the reactant-state sums, which then
  // overlap the next step's H psi: ev_slice orders them after the update that
  // wrote the slice, ev_cvj orders the next update and the ring flush after them
  hipStream_t aux = nullptr;
  hipEvent_t ev_slice = nullptr, ev_cvj = nullptr;
  // third stream: the angular term runs beside the radial passes; ev_psi
  // orders it after the update that wrote psi, ev_ang the next update after it
  hipStream_t side = nullptr;
  hipEvent_t ev_psi = nullptr, ev_ang = nullptr;
  bool cvj_pending = false;
  // step loop only: ev_slice follows main's last write of psi (the update),
  // so the next H psi's third stream waits on it instead of a new record
  bool slice_fresh = false;
  bool cvj_joined = false;  // the third stream waited ev_cvj before ev_ang
  int device = 0;

  // device buffers
  uint32_t* d_pt_dense = nullptr;
  uint8_t* d_mask = nullptr;
  // update work items (see mw3_create): rows and mask words of each
  // (angle, 64-R chunk), items [nitems][4], item range per (channel, angle,
  // chunk), flux partials [nitems][3][64]
  int32_t* d_uw_row = nullptr;
  uint64_t* d_uw_word = nullptr;
  int32_t* d_uw_item = nullptr;
  int32_t* d_uw_gfirst = nullptr;
  double* d_uw_part = nullptr;
  void* d_v = nullptr;  // double or float dense V
  double* d_absfr1 = nullptr;
  double* d_absfr2 = nullptr;
  double* d_psi = nullptr;    // psi_k dense
  double* d_psim1 = nullptr;  // psi_{k-1} dense
  double* d_psi0 = nullptr;   // rpaq0 dense
  double* d_a0 = nullptr;     // A0 staging [ntot]
  double* d_a0b = nullptr;
  void* d_t1 = nullptr;       // radial work slab (T or float)
  void* d_t2 = nullptr;       // radial work slab
  void* d_fbr = nullptr;      // FBR slab [nc][nJ][n1][n2]
  void* d_ang = nullptr;      // angular result slab [nc][na][n1][n2]
  double* d_hout = nullptr;   // H psi dense (hphi mode)
  // FP32 copies of the psi slabs (mixed precision only): the operator passes
  // read psi rounded to FP32, so the update writes that rounding once next to
  // the FP64 state; they swap with d_psi / d_psim1
  float* d_psi32 = nullptr;
  float* d_psim132 = nullptr;
  float* d_scr32 = nullptr;
  double* d_scratch = nullptr; // input slab of mw3_hphi
  std::vector<double> reacfct_host, etot_host;
  // tables
  int32_t* d_nn1 = nullptr;
  double* d_p2r1 = nullptr;
  double* d_p2r2 = nullptr;
  double* d_divi = nullptr;
  void* d_dst_tables = nullptr;
  void* d_Dmat = nullptr;
  void* d_acoef = nullptr;
  void* d_bcoef = nullptr;
  int32_t* d_nbr = nullptr;
  // projections
  uint32_t* d_flux_dense = nullptr;
  uint32_t* d_flux_col = nullptr;
  double* d_zar1flux = nullptr;
  double* d_ring_zr = nullptr;   // [kRing][nflux]
  double* d_ring_zdr = nullptr;  // [kRing][nflux][2]
  double* d_ring_cvj = nullptr;  // [kRing][ncvj]
  double* d_zcr = nullptr;       // [nflux][netot][2]
  double* d_zcdr = nullptr;
  double* d_zs2 = nullptr;       // [ncvj][netot][2]
  double* d_etot = nullptr;
  double* d_fdv = nullptr;
  double* d_dcvj = nullptr;
  uint8_t* d_cvj_skip = nullptr; // [ncvj] 1 if the reference state (kminelastic)
  double* d_norm_part = nullptr; // partial sums [nblocks][2]
  double* d_ring_norm = nullptr; // [kRing][2] (norm, autocorrelation)
  double* d_efe = nullptr;
  double* d_zfe = nullptr;       // [nofe][nc*nslab][2] dense
  int nnorm_blocks = 0;  // update work items (norm partials)
  int ring_fill = 0;             // steps buffered in the ring
  int ring_k0 = 0;               // k of ring slot 0
  std::vector<double> host_norms;
  // per-step records (mw3_records): every ring flush appends its steps
  int rec_on = 0;
  std::vector<int32_t> rec_it;   // [nrec] k of each record
  std::vector<double> rec_cvj;   // [nrec][ncvj]
  std::vector<double> rec_cvjp;  // [nrec][nomb * nvj] (product projection attached)

  // GEMM passes: R rows, r columns, angular forward, angular back
  GemmPlan p_dstR, p_dstr, p_dstr2, p_dstRi, p_angF, p_angB;
  bool ang_fused = false;  // FBR coefficients applied in the forward GEMM (no Coriolis)
  int ang_nt = 0;          // 16-row tiles of the one-pass angular kernel (0: two GEMMs)
  int32_t* d_rad_items = nullptr;  // R-pass work items {channel-angle, 16-row tile, nn1, k2lo}
  int rad_nitems = 0;              // 0: the R passes run as GEMMs
  int rad_grid[2] = {0, 0};        // persistent blocks of the forward / inverse R pass
  void* d_lowT = nullptr;          // transposed low-mode r tables of k_rad_mid
  int32_t* d_mid_items = nullptr;  // its work items (int4)
  int32_t* d_mid_sizes = nullptr;  // its per-size table records (int4)
  int32_t* d_mid_items2 = nullptr; // packed items of the FP32 kernel (2 int4 each)
  float* d_mid_wts = nullptr;      // its weights (radcut - E)_+ divi [angle][k2][k1]
  int mid_rbw = 0;                 // r blocks per wave of the FP32 kernel
  int mid_nitems = 0;              // 0: the r passes run as two GEMMs
  int mid_nk1 = 0;                 // 16-row tiles of the low r modes
  int rad_ncu = 1;                 // compute units (the LDS-table R passes)
  bool rad_lds64 = false;          // FP64 R passes with the table split over block pairs
  bool rad_inv_h = false;          // FP32 inverse R pass with compensated FP16 products
  bool ang_fold_h = false;         // FP32 folded angular kernel with compensated FP16 products
  int ang_fold_h_nbc = 0;          // its blocks per channel
  int ang_fold_cor = 0;            // tiles of the folded forward / back parts with Coriolis (mixed)
  int ang_cor_nbc[3] = {0, 0, 0};  // their blocks per channel (index part - 1: parts 1 and 3)
  int32_t* d_fold_jmap = nullptr;  // [nc][32 tiles] function j at each class-order position, or -1
  void* d_fold_bm = nullptr;       // [nc][32 tiles][n2] b_j(R) of the Omega - 1 neighbour, class order
  void* d_fold_bp = nullptr;       // the same for Omega + 1
  int opt_fp16 = 1;                // MW3_OPT_FP16_PRODUCTS (mixed only)
  int ang_grid = 0;        // its persistent blocks
  int32_t* d_ang_tiles = nullptr;  // tiles of the one-pass kernels with compressed-grid points
  int ang_ntiles = 0;
  int32_t* d_ang_groups = nullptr;  // 16-column groups with compressed-grid points (folded kernel)
  int ang_ngroups = 0;
  int ang_fold_nbc = 0;             // folded kernel: blocks per channel
  int ang_fold = 0;        // row tiles of the parity-folded one-pass kernel (0: not used)
  bool fbr_all = false;    // Coriolis coefficients by k_fbr_coef_all (2..8 channels)
  void* d_fold_D = nullptr;    // [nc][16 ang_fold][32 ang_fold] D(ia < na/2, j by parity class)
  // operator precision (FP32 in mixed mode, like the D tables)
  void* d_fold_a = nullptr;  // [nc][32 ang_fold][n1] a_j(r1) in the same j order
  void* d_fold_b = nullptr;  // [nc][32 ang_fold][n2] b_j(R)
  void* d_phi = nullptr;        // FBR after the Omega stencil
  double* d_wmat = nullptr;     // [kRing][netot][2] Chebyshev energy factors
  uint8_t* d_zero_skip = nullptr;  // [nflux] all ones (kminelastic zeroing)
  int32_t* d_ch_elec = nullptr;
  int32_t* d_ch_omega = nullptr;
  // [nc][4] the channels of each channel's Omega (itself included) in
  // ascending order, -1 past their count; null when some Omega has more than 4
  int32_t* d_ch_cpl = nullptr;
  double* d_potiflux = nullptr;
  int32_t* d_flux_elec = nullptr;

  int32_t* d_fmap = nullptr;     // [nc][na][n2] flux point index or -1
  int32_t* d_ir1flux = nullptr;  // [na]
  double* d_slice = nullptr;     // [nc][na][n1] psi at R = ir2abs / sqrt(ah2)

  // product-state projection (iprod = 2, mw3_prod_attach).  Per step, the
  // update sums T psi_new over r into d_ppart (T = transRgjac on the dense
  // grid); k_prod_fun reduces them to funprod(c, l), l = (angle, R) of the
  // projection range, rotates by dtrans and sums the channels of each
  // electronic state into the ring d_ring_prod[slot][Omega'][e][l].  At each
  // ring flush one split-K GEMM per Omega' projects the ring on prodwf:
  // Cvjprod[s][Omega'][j'v'] (partials [Omega'][chunk][kRing][nvj]), then the
  // energy accumulation adds C_k(E) Cvjprod to zS2prod.
  bool prod_on = false;
  int pr_n2p0 = 0, pr_nap0 = 0, pr_n2p = 0, pr_nap = 0, pr_nomb = 0, pr_nvj = 0;
  int pr_nchunk = 0, pr_kc = 0;
  int64_t pr_L = 0;
  int64_t pr_tstride = 0;        // channel stride of d_tprod (0: one table for all)
  std::vector<int32_t> pr_shape; // iomminprod, iommaxprod, nviniprod, nvmaxprod, jiniprod, jmaxprod
  void* d_tprod = nullptr;       // T or float dense [nct][na][n1][n2]
  double* d_ppart = nullptr;     // [nitems][64]
  double* d_dtrans = nullptr;    // [nc][nomb][L]
  double* d_ring_prod = nullptr; // [kRing][nomb][ne][L]
  double* d_prodwf = nullptr;    // [nomb][nvj][ne][L]
  double* d_prod_part = nullptr; // [nomb][nchunk][kRing][nvj]
  double* d_ring_cvjp = nullptr; // [kRing][nomb * nvj]
  double* d_cvjp_last = nullptr; // [nomb * nvj]
  double* d_zs2prod = nullptr;   // [nomb * nvj][netot][2]
  GemmPlan p_prod;

  // recursion state
  int k = 0;
  bool begun = false;

  // captured steps (hipGraph): graph[q] is one step -- the projection sums of
  // the step before on the second stream beside this step's H psi, then the
  // update -- for psi in buffer q (0: the buffer d_psi held at creation);
  // d_slot is the ring slot of those projection sums, advanced by the graph
  int graph_mode = 1;             // 0: off; 1: on when the context allows it
  // execution options (include/mw3.h MW3_OPT_*): radial passes as the
  // persistent low-mode kernels (1) or as GEMMs (0); angular term by the
  // parity-folded one-pass kernel (0), the unfolded one-pass kernel (1) or
  // GEMMs (2), each falling back to the next where not built; the angular
  // term on the third stream (1) or the main one (0)
  int opt_radial = 1, opt_angular = 0, opt_side = 0;
  int opt_rad_grid = 1;           // persistent radial grids: resident blocks / this
  int opt_ang_grid = 1;           // folded angular kernel: blocks per channel / this
  int flux_etest = 1;             // E > V(flux point) test when the flux sums are read
  bool trial = false;             // autotuner trial steps: no ring flush, no psi(E)
  bool graph_slot_dev = false;    // kernels being enqueued read their slot from d_slot
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  double* graph_psi0 = nullptr;   // d_psi at creation (buffer 0)
  int32_t* d_slot = nullptr;
  int deferred_slot = -1;         // ring slot whose projection sums are not yet enqueued
  hipEvent_t ev_fork = nullptr;

  // timing
  bool timing = false;
  int timing_every = 1;     // time every N-th step
  bool timing_step = true;  // the current step is timed
  double t_ms[MW3_NKERNELS] = {};
  int64_t t_n[MW3_NKERNELS] = {};
  std::vector<std::pair<int, Timer>> pending;
  std::vector<Timer> pool;
};

// mw3_setup.cpp
int build_tables(const mw3_setup* s, HostTables* t, std::string* err);

// mw3_kernels.hip: launchers, return hipError_t
hipError_t launch_scatter_a0(Ctx* c, const double* a0, double* dense);
hipError_t launch_gather_a0(Ctx* c, const double* dense, double* a0);
hipError_t launch_hpsi(Ctx* c, const double* psi);
float* psi_shadow(Ctx* c, const double* psi);  // FP32 copy of a psi slab (mixed), or null
hipError_t launch_to_f32(Ctx* c, const double* src, float* dst);
hipError_t launch_spectrum(hipStream_t s, const double* d_auto, double* d_a, int kmax,
                           double hbr, double xkgam, const double* d_e, int ne, double emindlt,
                           double delta2, double* d_spec);  // T, V and angular terms into work slabs
// 0: H psi into d_hout, 1: first step, 2: damped step (psi_new into d_psim1)
hipError_t launch_update(Ctx* c, int mode, const double* psi, bool proj = false);
hipError_t launch_projections(Ctx* c, int slot);
hipError_t launch_tail_sums(Ctx* c, int slot, hipStream_t st);
hipError_t launch_energy_accum(Ctx* c);
hipError_t launch_psie(Ctx* c, int k);
hipError_t launch_zero(Ctx* c, void* p, size_t bytes);
hipError_t build_plans(Ctx* c);
void free_plans(Ctx* c);
hipError_t launch_flux_result(Ctx* c, double* d_out /* [(1+ne)][netot] */);
// product projection: plan of the flush GEMM, per-step reduction into the ring
hipError_t build_prod_plan(Ctx* c);
hipError_t launch_prod_step(Ctx* c, int slot, hipStream_t s);
hipError_t launch_prod_flush(Ctx* c);

constexpr int kUpdRows = 32;  // rows per update work item
// most channels of one product-state context (LDS of the per-step product
// reduction: 64 doubles per channel)
constexpr int kProdMaxCh = 64;

// GEMM passes of the step (template tag: one kernel symbol per pass)
enum GemmRole {
  G_ANG_FWD = 0,
  G_ANG_BWD,
  G_DST_R,
  G_DST_r_KIN,
  G_DST_r_INV,
  G_DST_R_INV,
  G_PROD  // product projection of the ring (flush)
};

enum KernelId {
  K_ANG_FWD = 0,
  K_ANG_BWD = 1,
  K_DST_R = 2,
  K_DST_r = 3,
  K_DST_RINV = 4,
  K_UPDATE = 5,
  K_PROJ = 6,
  K_ENERGY = 7,
  K_CORIOLIS = 8,  // FBR coefficients with the Omega +- 1 Coriolis stencil (k_fbr_coef)
  K_PROD = 9,      // product projection: per-step reduction (the flush GEMM is in K_ENERGY)
  K_UPDATE_TAIL = 10,  // flux columns and norm from the update partials (second stream)
};

hipError_t launch_update_tail(Ctx* c, int mode, int slot, bool proj, hipStream_t st);
hipError_t launch_slot_set(Ctx* c, int v, hipStream_t st);
hipError_t launch_slot_next(Ctx* c, hipStream_t st);
void timer_begin(Ctx* c, int id, hipStream_t s = nullptr);
void timer_end(Ctx* c, int id, hipStream_t s = nullptr);

}  // namespace mw3
