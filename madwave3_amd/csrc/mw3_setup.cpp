// mw3_setup.cpp -- operator tables of libmw3 built from the reference's module
// state (struct mw3_setup).
//
//   Tradial      SRC/mod_Hphi_01y2.f:487-620: per-angle r DST size
//                (noptFFT, SRC/liboctdyn.f:1033-1117, floored at npun1min)
//                and DST momenta (sinmom, SRC/liboctdyn.f:1004-1031)
//   angkinj2     SRC/mod_Hphi_01y2.f:166-311 and
//   bfgridl2mat  SRC/mod_Hphi_01y2.f:315-484: the reference stores the
//                angular operator on the gamma grid as
//                  xj2(r1,g,g')   = sum_j D(g,j) D(g',j) min(hb^2(j(j+1)-w^2)/(2 mu1 r1^2), rotcut)
//                  xl2p(R,g,g',q) = sum_j D_t(g,j) D_s(g',j) l2_j(Om_t,Om_s)/r2eff_j(R)^2.
//                libmw3 keeps the factors instead: the matrices D_c(g,j) and
//                the coefficients acoef(j,r1) and bcoef(q,j,R), and applies
//                D diag(coef) D^T on the device.
//   ini_flux     SRC/mod_flux_01y2.f:99-257: flux line r index, d/dr vector
//                zAR1flux (fftmom, SRC/liboctdyn.f:1119-1151), flux points.
//   Cvjflux_k    SRC/mod_flux_01y2.f:663-775: reactant-state tables.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>

#include "mw3_internal.h"

namespace mw3 {

namespace {

// noptFFT: smallest 2^a 3^b 5^c 7^d 11^e >= nin (each exponent bounded by
// ntot, distance below ntot), capped at ntot; 1 if no candidate.
int nopt_fft(int nin, int ntot) {
  const int primes[5] = {2, 3, 5, 7, 11};
  int nexp[5];
  for (int q = 0; q < 5; q++) {
    long v = 1;
    int n = 0;
    for (n = 0; n <= 100; n++) {
      v *= primes[q];
      if (v > ntot) break;
    }
    nexp[q] = n;
  }
  int best = 1;
  long dist = ntot;
  std::vector<long> cand{1};
  for (int q = 0; q < 5; q++) {
    std::vector<long> next;
    for (long c : cand) {
      long v = c;
      for (int e = 0; e <= nexp[q]; e++) {
        next.push_back(v);
        v *= primes[q];
      }
    }
    cand.swap(next);
  }
  for (long i : cand) {
    long d = i - nin;
    if (d >= 0 && d < dist) {
      best = (int)i;
      dist = d;
    }
  }
  return best > ntot ? ntot : best;
}

// sinmom: hb^2 p_k^2/(2 m), p_k = 2 pi k / (2 (n+1) ah), with the reference's
// sequence of roundings (box -> ah -> box').
void sin_mom(int n, double ah_in, double xm, double hbr, double* p2) {
  double pi = std::acos(-1.0);
  double box = (double)(n - 1) * ah_in;
  double hbrxm = 0.5 * hbr * hbr / xm;
  double ah = box / (double)(n - 1);
  double b2 = 2.0 * ((double)n + 1.0) * ah;
  for (int k = 1; k <= n; k++) {
    double p = 2.0 * pi * (double)k / b2;
    p2[k - 1] = p * p * hbrxm;
  }
}

// Smallest and largest eigenvalue of a small symmetric matrix (the reference
// uses DIAGON, SRC/liboctdyn.f:2068-2080, and keeps the two extremes).
void eig_minmax(std::vector<double> a, int n, double* lo, double* hi) {
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += std::fabs(a[p * n + q]);
    if (off == 0.0) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a[p * n + q];
        if (apq == 0.0) continue;
        double th = 0.5 * (a[q * n + q] - a[p * n + p]) / apq;
        double t = 1.0 / (std::fabs(th) + std::sqrt(1.0 + th * th));
        if (th < 0) t = -t;
        double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int r = 0; r < n; r++) {
          double x = a[r * n + p], y = a[r * n + q];
          a[r * n + p] = c * x - s * y;
          a[r * n + q] = s * x + c * y;
        }
        for (int r = 0; r < n; r++) {
          double x = a[p * n + r], y = a[q * n + r];
          a[p * n + r] = c * x - s * y;
          a[q * n + r] = s * x + c * y;
        }
      }
  }
  *lo = *hi = a[0];
  for (int p = 1; p < n; p++) {
    *lo = std::min(*lo, a[p * n + p]);
    *hi = std::max(*hi, a[p * n + p]);
  }
}

bool fail(std::string* err, const std::string& m) {
  *err = m;
  return false;
}

}  // namespace

int build_tables(const mw3_setup* s, HostTables* t, std::string* err) {
  if (!s) return (*err = "mw3_create: setup is NULL", MW3_EINVAL);
  if (s->abi_version != MW3_ABI_VERSION)
    return (*err = "mw3_create: setup.abi_version does not match the library", MW3_EINVAL);
  if (s->proj_j2_up != 0)
    return (*err = "mw3_create: Proj_j2_Up != 0 (SRC/mod_Hphi_01y2.f:274-298) is not implemented",
            MW3_EUNSUP);
  if (s->npun1 < 2 || s->npun2 < 2 || s->nangu < 2)
    return (*err = "mw3_create: npun1, npun2 and nangu must exceed 1 (Hdiatom / rigid-rotor paths are not built)",
            MW3_EUNSUP);
  if (s->npun1 > 4096)
    return (*err = "mw3_create: npun1 above 4096 (the update kernel keeps a mask word and a row index per r row in LDS)",
            MW3_EUNSUP);
  if (s->ncan < 1 || s->nelecmax < 1 || s->npuntot < 1 || s->inc < 1)
    return (*err = "mw3_create: ncan, nelecmax, npuntot and inc must be positive", MW3_EINVAL);
  if (!s->indangreal || !s->npunreal || !s->indr1 || !s->indr2 || !s->VVV || !s->iombas ||
      !s->nelebas || !s->nojbas || !s->jbas || !s->Djmmp || !s->absfr1 || !s->absfr2 ||
      !s->iomdiat || !s->iomatom || !s->sigdiat || !s->sigatom || !s->j00)
    return (*err = "mw3_create: a required setup array is NULL", MW3_EINVAL);
  if (s->netot > 0 && (!s->etots2 || !s->reacfct))
    return (*err = "mw3_create: netot > 0 needs etots2 and reacfct", MW3_EINVAL);
  if (s->delta2 <= 0.0) return (*err = "mw3_create: delta2 must be positive", MW3_EINVAL);

  const int n1 = s->npun1, n2 = s->npun2, na = s->nangu, nc = s->ncan, ne = s->nelecmax;
  const int np = s->npuntot;
  t->n1 = n1;
  t->n2 = n2;
  t->na = na;
  t->nc = nc;
  t->ne = ne;
  t->nplane = (int64_t)n1 * n2;
  t->nslab = (int64_t)na * t->nplane;
  if ((double)nc * (double)t->nslab >= 4294967295.0)
    return (*err = "mw3_create: dense grid exceeds 2^32 values", MW3_EUNSUP);

  for (int i = 0; i < nc; i++) {
    if (s->nelebas[i] < 1 || s->nelebas[i] > ne || s->iombas[i] < 0 || s->iombas[i] > s->Jtot)
      return (*err = "mw3_create: channel with electronic state or Omega out of range", MW3_EINVAL);
    t->ch_omega.push_back(s->iombas[i]);
    t->ch_elec.push_back(s->nelebas[i] - 1);
  }
  std::vector<int> seen(na + 1, 0);
  for (int sl = 0; sl < na; sl++) {
    int ia = s->indangreal[sl];
    if (ia < 1 || ia > na || seen[ia])
      return (*err = "mw3_create: indangreal is not a permutation of 1..nangu", MW3_EINVAL);
    seen[ia] = 1;
    if (s->npunreal[ia - 1] < 0 || s->npunreal[ia - 1] > np)
      return (*err = "mw3_create: npunreal out of [0, npuntot]", MW3_EINVAL);
  }

  // A0 <-> dense maps, mask, dense V (SRC/mod_pot_01y2.f:435-498)
  t->mask.assign((size_t)t->nslab, 0);
  t->vdense.assign((size_t)ne * ne * t->nslab, 0.0);
  int64_t ntot = 0;
  for (int sl = 0; sl < na; sl++) ntot += (int64_t)nc * s->npunreal[s->indangreal[sl] - 1];
  t->ntot = ntot;
  t->pt_dense.resize((size_t)ntot);
  int64_t pos = 0;
  for (int sl = 0; sl < na; sl++) {
    int ia = s->indangreal[sl] - 1;
    int npr = s->npunreal[ia];
    const int32_t* i1 = s->indr1 + (size_t)ia * np;
    const int32_t* i2 = s->indr2 + (size_t)ia * np;
    for (int ir = 0; ir < npr; ir++) {
      if (i1[ir] < 1 || i1[ir] > n1 || i2[ir] < 1 || i2[ir] > n2)
        return (*err = "mw3_create: indr1/indr2 out of the grid", MW3_EINVAL);
      size_t pl = (size_t)ia * t->nplane + (size_t)(i1[ir] - 1) * n2 + (i2[ir] - 1);
      if (t->mask[pl]) return (*err = "mw3_create: repeated grid point in indr1/indr2", MW3_EINVAL);
      t->mask[pl] = 1;
      for (int et = 0; et < ne; et++)
        for (int es = 0; es < ne; es++)
          t->vdense[((size_t)et * ne + es) * t->nslab + pl] =
              s->VVV[(((size_t)et * ne + es) * na + sl) * np + ir];
    }
    for (int ic = 0; ic < nc; ic++)
      for (int ir = 0; ir < npr; ir++) {
        size_t pl = (size_t)ia * t->nplane + (size_t)(i1[ir] - 1) * n2 + (i2[ir] - 1);
        t->pt_dense[pos++] = (uint32_t)((size_t)ic * t->nslab + pl);
      }
  }

  // Tradial (SRC/mod_Hphi_01y2.f:508-599)
  t->nn1.assign(na, 0);
  t->p2r1.assign((size_t)na * n1, 0.0);
  t->p2r2.assign(n2, 0.0);
  t->divi.assign(na, 0.0);
  t->radcut = s->radcutmax;
  std::set<int> sizes;
  for (int ia = 0; ia < na; ia++) {
    int lo = n1, hi = 0;
    for (int ir = 0; ir < s->npunreal[ia]; ir++) {
      int r = s->indr1[(size_t)ia * np + ir];
      lo = std::min(lo, r);
      hi = std::max(hi, r);
    }
    int nn = nopt_fft(hi - lo + 1, n1);
    if (nn < s->npun1min) nn = s->npun1min;
    if (nn < 2 || nn > n1)
      return (*err = "mw3_create: r DST size outside [2, npun1] (check npun1min)", MW3_EINVAL);
    t->nn1[ia] = nn;
    sin_mom(nn, s->ah1, s->xm1red, s->hbr, &t->p2r1[(size_t)ia * n1]);
    t->divi[ia] = 1.0 / (4.0 * (double)(nn + 1) * (double)(n2 + 1));
    sizes.insert(nn);
  }
  sin_mom(n2, s->ah2, s->xm2red, s->hbr, t->p2r2.data());
  sizes.insert(n2);
  {
    double pi = std::acos(-1.0);
    int64_t off = 0;
    for (int n : sizes) {
      t->dst_sizes.push_back(n);
      t->dst_table_off.push_back(off);
      int h = n / 2, ke = n - h, kp = (ke + 3) & ~3;
      t->dst_ld.push_back(kp);
      long m2 = 2L * (n + 1);
      auto S = [&](int j, int k) {
        long m = ((long)(j + 1) * (k + 1)) % m2;
        if (m == 0 || m == n + 1) return 0.0;  // exact zeros of sin(pi m/(n+1))
        return 2.0 * std::sin(pi * (double)m / (double)(n + 1));
      };
      // S_e[j][k] = S(j, 2k) and S_o[j][k] = S(j, 2k+1), each a kp x kp block
      // (leading dimension kp, a multiple of 4, zero padded)
      for (int j = 0; j < kp; j++)
        for (int k = 0; k < kp; k++)
          t->dst_tables.push_back((j < ke && k < ke) ? S(j, 2 * k) : 0.0);
      for (int j = 0; j < kp; j++)
        for (int k = 0; k < kp; k++)
          t->dst_tables.push_back((j < h && k < h) ? S(j, 2 * k + 1) : 0.0);
      off += 2 * (int64_t)kp * kp;
    }
    // Low-mode counts.  The kinetic factor is min(E, radcut) with
    // E = p2r1(k1) + p2r2(k2); min(E, c) = c - (c - E)_+ and the unnormalised
    // DST-I squares to 2(n+1) I, so T psi = c psi - S (c - E)_+ divi S psi,
    // and (c - E)_+ vanishes unless E < c: only modes k1 < k1lo(ia),
    // k2 < k2lo(ia) take part.
    t->k1lo.assign(na, 0);
    t->k2lo.assign(na, 0);
    double p2min = *std::min_element(t->p2r2.begin(), t->p2r2.end());
    for (int ia = 0; ia < na; ia++) {
      const double* q1 = &t->p2r1[(size_t)ia * n1];
      const int nn = t->nn1[ia];
      double p1min = *std::min_element(q1, q1 + nn);
      for (int k = 0; k < nn; k++)
        if (q1[k] + p2min < t->radcut) t->k1lo[ia] = k + 1;
      for (int k = 0; k < n2; k++)
        if (t->p2r2[k] + p1min < t->radcut) t->k2lo[ia] = k + 1;
    }
    // low-mode tables L[k][j] = S(j, k), k < kcap (largest count of that size),
    // row stride a multiple of 4
    for (size_t q = 0; q < t->dst_sizes.size(); q++) {
      const int n = t->dst_sizes[q];
      int kcap = 0;
      for (int ia = 0; ia < na; ia++) {
        if (t->nn1[ia] == n) kcap = std::max(kcap, t->k1lo[ia]);
        if (n == n2) kcap = std::max(kcap, t->k2lo[ia]);
      }
      const int ld = (n + 3) & ~3;
      t->low_off.push_back(off);
      t->low_ld.push_back(ld);
      long m2 = 2L * (n + 1);
      for (int k = 0; k < kcap; k++)
        for (int j = 0; j < ld; j++) {
          long m = ((long)(j + 1) * (k + 1)) % m2;
          double v = 0.0;
          if (j < n && m != 0 && m != n + 1) v = 2.0 * std::sin(pi * (double)m / (double)(n + 1));
          t->dst_tables.push_back(v);
        }
      off += (int64_t)kcap * ld;
    }
    // transposed n2 tables for the one-wave-per-row-tile R passes
    int kmax = 0;
    for (int ia = 0; ia < na; ia++) kmax = std::max(kmax, t->k2lo[ia]);
    const int nkb = ((kmax + 1) / 2 + 15) / 16;
    if (n2 % 64 == 0 && nkb >= 1 && nkb <= 4) {
      const int h = n2 / 2, k2p = 32 * nkb;
      const long m2 = 2L * (n2 + 1);
      auto S = [&](int j, int k) {
        if (k >= n2) return 0.0;
        long m = ((long)(j + 1) * (k + 1)) % m2;
        if (m == 0 || m == n2 + 1) return 0.0;
        return 2.0 * std::sin(pi * (double)m / (double)(n2 + 1));
      };
      t->rad_nkb = nkb;
      t->rad_ti_off = off;
      for (int R = 0; R < h; R++)
        for (int k = 0; k < k2p; k++) t->dst_tables.push_back(S(R, k));
      off += (int64_t)h * k2p;
      t->rad_tt_off = off;
      for (int k = 0; k < k2p; k++)
        for (int j = 0; j < h; j++) t->dst_tables.push_back(S(j, k));
      off += (int64_t)k2p * h;
    }
  }

  // angular FBR tables (angkinj2 / bfgridl2mat restated in factor form)
  const int jtop = (na - 1) * s->inc;
  t->nJ = jtop + 1;
  const int nJ = t->nJ;
  for (int j = 0; j <= jtop; j++) t->jlist.push_back(j);
  if (s->ndj < jtop + 1)
    return (*err = "mw3_create: Djmmp j extent (ndj) below (nangu-1)*inc+1", MW3_EINVAL);
  auto DJ = [&](int iom, int e, int j, int iang) {
    return s->Djmmp[(((size_t)iom * ne + e) * s->ndj + j) * na + (iang - 1)];
  };
  t->nJp = (nJ + 3) & ~3;
  t->Dmat.assign((size_t)nc * na * t->nJp, 0.0);
  for (int c = 0; c < nc; c++)
    for (int ia = 1; ia <= na; ia++)
      for (int jj = 0; jj < nJ; jj++)
        t->Dmat[((size_t)c * na + (ia - 1)) * t->nJp + jj] =
            DJ(t->ch_omega[c], t->ch_elec[c], jj, ia);
  // j^2 (SRC/mod_Hphi_01y2.f:235-272)
  t->acoef.assign((size_t)nc * nJ * n1, 0.0);
  for (int c = 0; c < nc; c++) {
    int iomdi = s->iomdiat[t->ch_elec[c]];
    for (int k = 0; k < s->nojbas[c]; k++) {
      int jk = s->jbas[(size_t)c * na + k];
      if (jk > jtop || jk < 0) continue;
      double xj = (double)jk, xom = (double)iomdi;
      double xkin = (xj * (xj + 1.0) - xom * xom) * s->hbr * s->hbr * 0.5;
      for (int ir1 = 1; ir1 <= n1; ir1++) {
        double r1 = s->rmis1 + (double)(ir1 - 1) * s->ah1;
        double y = xkin / (s->xm1red * r1 * r1);
        if (y > s->rotcutmax) y = s->rotcutmax;
        t->acoef[((size_t)c * nJ + jk) * n1 + (ir1 - 1)] += y;
      }
    }
  }
  // l^2 blocks (SRC/mod_Hphi_01y2.f:343-475)
  t->bcoef.assign((size_t)nc * 3 * nJ * n2, 0.0);
  t->nbr.assign((size_t)nc * 3, -1);
  auto chan = [&](int e, int om) {
    for (int c = 0; c < nc; c++)
      if (t->ch_elec[c] == e && t->ch_omega[c] == om) return c;
    return -1;
  };
  for (int c = 0; c < nc; c++)
    for (int q = -1; q <= 1; q++) t->nbr[c * 3 + q + 1] = chan(t->ch_elec[c], t->ch_omega[c] + q);
  {
    double facmass = s->hbr * s->hbr * 0.5 / s->xm2red;
    int parJ = (s->Jtot % 2 == 0) ? 1 : -1;
    for (int e = 0; e < ne; e++) {
      int sigelec = (s->sigdiat[e] == -1.0 || s->sigatom[e] == -1.0) ? -1 : 1;
      for (int j = s->j0; j <= jtop; j += s->inc) {
        int iom0 = s->iommin;
        if (s->iparity != sigelec * parJ && s->iommin == 0) iom0 = 1;
        if (iom0 > j) continue;
        int iom1 = std::min(j, s->iommax);
        int nnn = iom1 - iom0 + 1;
        if (nnn <= 0) continue;
        std::vector<double> bfl((size_t)nnn * nnn, 0.0);
        double xj = (double)j;
        for (int a = 0; a < nnn; a++)
          for (int b = 0; b < nnn; b++) {
            int iom = iom0 + a, jom = iom0 + b;
            double xk = 0.0;
            if (iom == jom) {
              xk = (xj * (xj + 1.0) + (double)(s->Jtot * (s->Jtot + 1) - 2 * iom * jom)) * facmass;
            } else if (std::abs(iom - jom) == 1) {
              double x1 = (double)(s->Jtot * (s->Jtot + 1) - iom * jom);
              double x2 = xj * (xj + 1.0) - (double)(iom * jom);
              xk = -std::sqrt(x1 * x2) * facmass;
              if (iom * jom == 0) xk *= std::sqrt(2.0);
            }
            bfl[(size_t)a * nnn + b] = xk;
          }
        double ylo, yhi;
        if (nnn > 1)
          eig_minmax(bfl, nnn, &ylo, &yhi);
        else
          ylo = yhi = bfl[0];
        double ymean = (nnn > 1) ? 0.5 * (ylo + yhi) : ylo;
        for (int a = 0; a < nnn; a++)
          for (int b = 0; b < nnn; b++) {
            int iom = iom0 + a, jom = iom0 + b;  // target iom, source jom
            if (std::abs(jom - iom) > 1) continue;
            int tc = chan(e, iom), sc = chan(e, jom);
            if (tc < 0 || sc < 0) continue;
            int q = jom - iom + 1;
            double bv = bfl[(size_t)a * nnn + b];
            for (int ir2 = 1; ir2 <= n2; ir2++) {
              double r2 = s->rmis2 + (double)(ir2 - 1) * s->ah2;
              double r2eff = r2;
              if (ymean / (r2 * r2) > s->rotcutmax) r2eff = std::sqrt(std::fabs(ymean / s->rotcutmax));
              double xl = 0.0;
              if (std::fabs(bv) > 1e-8 && r2eff > 1e-3) xl = bv / (r2eff * r2eff);
              t->bcoef[(((size_t)tc * 3 + q) * nJ + j) * n2 + (ir2 - 1)] += xl;
            }
          }
      }
    }
  }

  // flux line (SRC/mod_flux_01y2.f:108-257)
  if (s->netot > 0) {
    t->ir4flux = 0;
    for (int ir1 = 1; ir1 <= n1; ir1++) {
      double r1 = s->rmis1 + (double)(ir1 - 1) * s->ah1;
      if (r1 <= s->r1flux && ir1 > t->ir4flux) t->ir4flux = ir1;
    }
    double pi = std::acos(-1.0);
    double box = (double)(n1 - 1) * s->ah1;
    double ah = box / (double)(n1 - 1);
    double b2 = ah * (double)n1;
    std::vector<double> xk(n1);
    for (int ir = 1; ir <= n1; ir++) {
      int iii = (ir <= n1 / 2) ? ir - 1 : ir - n1 - 1;
      xk[ir - 1] = 2.0 * pi * (double)iii / b2;
    }
    t->zar1flux.assign((size_t)2 * n1, 0.0);
    for (int m = 1; m <= n1; m++) {
      double d2r = (double)(t->ir4flux - m) * s->ah1;
      double zr = 0, zi = 0;
      for (int k = 0; k < n1; k++) {
        zr += -xk[k] * std::sin(xk[k] * d2r);
        zi += xk[k] * std::cos(xk[k] * d2r);
      }
      t->zar1flux[2 * (m - 1)] = zr / (double)(2 * n1);
      t->zar1flux[2 * (m - 1) + 1] = zi / (double)(2 * n1);
    }
    t->ir1flux.assign(na, 0);
    for (int ia = 0; ia < na; ia++)
      for (int ir = 0; ir < s->npunreal[ia]; ir++) {
        int r = s->indr1[(size_t)ia * np + ir];
        double r1 = (double)(r - 1) * s->ah1 + s->rmis1;
        if (r1 <= s->r1flux && r > t->ir1flux[ia]) t->ir1flux[ia] = r;
      }
    for (int sl = 0; sl < na; sl++) {
      int ia = s->indangreal[sl] - 1;
      for (int c = 0; c < nc; c++) {
        int e = t->ch_elec[c];
        for (int ir = 0; ir < s->npunreal[ia]; ir++) {
          int r1 = s->indr1[(size_t)ia * np + ir];
          if (r1 != t->ir1flux[ia]) continue;
          int r2 = s->indr2[(size_t)ia * np + ir];
          size_t col = (size_t)c * t->nslab + (size_t)ia * t->nplane + (r2 - 1);
          t->flux_dense.push_back((uint32_t)(col + (size_t)(r1 - 1) * n2));
          t->flux_col.push_back((uint32_t)col);
          t->potiflux.push_back(s->VVV[(((size_t)e * ne + e) * na + sl) * np + ir]);
          t->flux_elec.push_back(e);
        }
      }
    }
    t->nflux = (int64_t)t->flux_dense.size();
  }

  // reactant-state projection tables (SRC/mod_flux_01y2.f:678-702)
  if (s->netot > 0 && s->jmax >= s->jini && s->nvmax >= s->nvini) {
    int nj = s->jmax - s->jini + 1, nv = s->nvmax - s->nvini + 1;
    if (!s->noBCstates || !s->fd) return (*err = "mw3_create: noBCstates/fd are NULL", MW3_EINVAL);
    if (s->ir2abs < 1 || s->ir2abs > n2) return (*err = "mw3_create: ir2abs outside 1..npun2", MW3_EINVAL);
    if (s->jmax > jtop) return (*err = "mw3_create: jmax > (nangu-1)*inc", MW3_EINVAL);
    t->ncvj = nc * nj * nv;
    if (t->ncvj > 65535)
      return (*err = "mw3_create: (channels x j x v) of the reactant projection exceeds 65535",
              MW3_EUNSUP);
    t->fdv.assign((size_t)t->ncvj * n1, 0.0);
    t->dcvj.assign((size_t)nc * nj * na, 0.0);
    t->cvj_valid.assign(t->ncvj, 0);
    t->cvj_isref.assign(t->ncvj, 0);
    for (int c = 0; c < nc; c++) {
      int e = t->ch_elec[c], om = t->ch_omega[c];
      for (int j = s->j00[c]; j <= s->jmax; j += s->inc) {
        if (j < s->jini) continue;
        int jj = j - s->jini;
        for (int ia = 1; ia <= na; ia++) t->dcvj[((size_t)c * nj + jj) * na + (ia - 1)] = DJ(om, e, j, ia);
        int nbc = s->noBCstates[(size_t)e * nj + jj];
        int maxnv = std::min(s->nvmax, nbc);
        for (int iv = s->nvini; iv <= maxnv; iv++) {
          int v = iv - s->nvini;
          size_t idx = ((size_t)c * nj + jj) * nv + v;
          t->cvj_valid[idx] = 1;
          for (int r = 0; r < n1; r++)
            t->fdv[idx * n1 + r] = s->fd[(((size_t)e * nj + jj) * nv + v) * n1 + r];
          bool isref = (s->iomref >= 0) ? (c + 1 == s->icanref && iv == s->nvref && j == s->jref)
                                        : (iv == s->nvref && j == s->jref);
          t->cvj_isref[idx] = isref ? 1 : 0;
        }
      }
    }
  }
  return MW3_OK;
}

}  // namespace mw3
