/*
 * mlau_oracle_kernels.c -- point kernels of the CPU restatement (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see mlau_oracle.h).  Restates, operation by
 * operation and in the same evaluation order:
 *   include/funcs.h:23-43      max/min with >= semantics, minmod, minmod3
 *   include/interp.h:30-72     MUSCL minmod / MC / Koren face values
 *   include/interp.h:106-146   WCNS3 (ESWENO3 weights)
 *   include/interp.h:186-238   WCNS4
 *   include/mhd_eigen.h:24-192 Athena-style eigensystem (mhd_eigen_st)
 *   include/mhd_eigen.h:326-453 Alfven/entropy eigensystem (mhd_eigen_ae)
 *   include/mhd_eigen.h:455-530,609-684 characteristic reconstruction
 *   include/mhd_flux.h:26-213  MLAU,  :215-454 Roe,  :456-582 HLLD,  :584-719 LHLLD
 * Compiled with -O2 -ffp-contract=off and no fast-math, like the reference.
 */
#include <math.h>
#include "mlau_oracle.h"

#define EPSF (1e-8) /* include/mhd_flux.h:24 */

static inline double mx_(double a, double b) { return (a >= b) ? a : b; }
static inline double mn_(double a, double b) { return (a >= b) ? b : a; }

/* include/funcs.h:33-37 */
double oc_minmod(double a, double b) { return +mn_(mx_(a, b), 0) + mx_(mn_(a, b), 0); }
/* include/funcs.h:39-43 */
double oc_minmod3(double a, double b, double c)
{
  return +mn_(mx_(mx_(a, b), c), 0) + mx_(mn_(mn_(a, b), c), 0);
}

/* include/interp.h:30-37 */
void oc_muscl_mm(double f1, double f2, double f3, double *fl, double *fr)
{
  double df = 0.5 * oc_minmod(f2 - f1, f3 - f2);
  *fl = f2 + df;
  *fr = f2 - df;
}
/* include/interp.h:45-53 */
void oc_muscl_mc(double f1, double f2, double f3, double *fl, double *fr)
{
  double alpha = 2.0;
  double df = 0.5 * oc_minmod(alpha * oc_minmod(f2 - f1, f3 - f2), 0.5 * (f3 - f1));
  *fl = f2 + df;
  *fr = f2 - df;
}
/* include/interp.h:62-72 */
void oc_muscl_kr(double f1, double f2, double f3, double *fl, double *fr)
{
  double d1 = 2.0 * (f2 - f1);
  double d2 = 2.0 * (f3 - f2);
  double dl = 0.5 * oc_minmod3(d1, d2, (0.5 * d1 + d2) / 3.0);
  double dr = 0.5 * oc_minmod3(d1, d2, (d1 + 0.5 * d2) / 3.0);
  *fl = f2 + dl;
  *fr = f2 - dr;
}

/* include/interp.h:106-146 */
void oc_wcns3(double f0, double f1, double f2, double *fl, double *fr)
{
  const double eps = 1e-40;
  double a, s0, s1, t, w0, w1;
  a = f0 - f1;
  s0 = a * a;
  a = f2 - f1;
  s1 = a * a;
  t = f2 - 2.0 * f1 + f0;
  t = t * t;
  s0 = (1.0 + t / (s0 + eps));
  s1 = (1.0 + t / (s1 + eps));
  w0 = 0.25 * s0;
  w1 = 0.75 * s1;
  w0 /= (w0 + w1);
  w1 = 1.0 - w0;
  *fl = 0.5 * (w0 * (-f0 + 3.0 * f1) + w1 * (f1 + f2));
  w0 = 0.75 * s0;
  w1 = 0.25 * s1;
  w0 /= (w0 + w1);
  w1 = 1.0 - w0;
  *fr = 0.5 * (w1 * (-f2 + 3.0 * f1) + w0 * (f1 + f0));
}

/* include/interp.h:186-238 */
void oc_wcns4(double f0, double f1, double f2, double f3, double f4, double *fl, double *fr)
{
  const double eps = 1e-40;
  double c1, c2, s0, s1, s2, w0, w1, w2, d;
  c1 = f0 - 4.0 * f1 + 3.0 * f2;
  c2 = f0 - 2.0 * f1 + f2;
  s0 = (13. / 12.) * c2 * c2 + 0.25 * c1 * c1;
  c1 = f1 - f3;
  c2 = f1 - 2.0 * f2 + f3;
  s1 = (13. / 12.) * c2 * c2 + 0.25 * c1 * c1;
  c1 = f4 - 4.0 * f3 + 3.0 * f2;
  c2 = f4 - 2.0 * f3 + f2;
  s2 = (13. / 12.) * c2 * c2 + 0.25 * c1 * c1;
  s0 = 1.0 / (s0 * s0 + eps);
  s1 = 1.0 / (s1 * s1 + eps);
  s2 = 1.0 / (s2 * s2 + eps);
  w0 = 0.075 * s0;
  w1 = 0.650 * s1;
  w2 = 0.275 * s2;
  d = 1.0 / (w0 + w1 + w2);
  w0 *= d;
  w1 *= d;
  w2 = 1.0 - w0 - w1;
  *fl = 0.125 * (+w0 * (3.0 * f0 - 10.0 * f1 + 15.0 * f2) + w1 * (-f1 + 6.0 * f2 + 3.0 * f3) +
                 w2 * (+3.0 * f2 + 6.0 * f3 - f4));
  w0 = 0.275 * s0;
  w1 = 0.650 * s1;
  w2 = 0.075 * s2;
  d = 1.0 / (w0 + w1 + w2);
  w0 *= d;
  w1 *= d;
  w2 = 1.0 - w0 - w1;
  *fr = 0.125 * (+w2 * (3.0 * f4 - 10.0 * f3 + 15.0 * f2) + w1 * (-f3 + 6.0 * f2 + 3.0 * f1) +
                 w0 * (+3.0 * f2 + 6.0 * f1 - f0));
}

/* Mirror of eigenvector k into 6-k with the reference's sign pattern
 * (+,-,-,-,+,+,+), include/mhd_eigen.h:107-113 and the r7/r6/r5 blocks. */
static void mirror7(const double *a, double *b)
{
  b[0] = +a[0];
  b[1] = -a[1];
  b[2] = -a[2];
  b[3] = -a[3];
  b[4] = +a[4];
  b[5] = +a[5];
  b[6] = +a[6];
}

/* include/mhd_eigen.h:24-192.  p = (ro,vx,vy,vz,bx,by,bz,pr).
 * L + 7*k is the k-th left eigenvector (row), R + 7*k the k-th right one. */
void oc_eigen_st(const double *p, double gamma, double *al, double *L, double *R)
{
  const double ro = p[0], vx = p[1], bx = p[4], by = p[5], bz = p[6], pr = p[7];
  const double eps = 1e-12;
  const double sqrt2 = sqrt(2.0);
  double iro = 1.0 / ro;
  double sro = sqrt(ro);
  double isro = 1.0 / sro;
  double bxsr = bx * isro, bysr = by * isro, bzsr = bz * isro;
  double bx2 = bxsr * bxsr;
  double bt2 = bysr * bysr + bzsr * bzsr;
  double b2 = bx2 + bt2;
  double a2 = fabs(gamma * pr / ro);
  double aa = sqrt(a2);
  double ca2 = bx2;
  double ca = sqrt(ca2);
  double junk = sqrt(fabs((a2 + b2) * (a2 + b2) - 4.0 * a2 * ca2));
  double cf2 = fabs(0.5 * (a2 + b2 + junk));
  double cf = sqrt(cf2);
  double cs2 = fabs(0.5 * (a2 + b2 - junk));
  double cs = sqrt(cs2);
  bx2 = bx * bx;
  bt2 = by * by + bz * bz;
  b2 = bx2 + bt2;
  double denom = 1.0 / sqrt(bt2);
  int flag = (bt2 > eps * b2);
  double betay = flag ? (by * denom) : (0.5 * sqrt2);
  double betaz = flag ? (bz * denom) : (0.5 * sqrt2);
  denom = 1.0 / sqrt(fabs(cf2 - cs2));
  flag = ((bt2 > eps * b2) || (fabs(gamma * pr - bx2) > eps * gamma * pr));
  double alphaf = flag ? (sqrt(fabs(a2 - cs2)) * denom) : 1.0;
  double alphas = flag ? (sqrt(fabs(cf2 - a2)) * denom) : 0.0;
  double sgnbx = (bx > 0) ? (1.0) : (-1.0);
  double cfalf = cf * alphaf, csals = cs * alphas, aaalf = aa * alphaf, aaals = aa * alphas;
  denom = 0.5 / a2;
  if (al) {
    al[0] = vx - cf; al[1] = vx - ca; al[2] = vx - cs; al[3] = vx;
    al[4] = vx + cs; al[5] = vx + ca; al[6] = vx + cf;
  }
  double *l1 = L, *l2 = L + 7, *l3 = L + 14, *l4 = L + 21, *l5 = L + 28, *l6 = L + 35, *l7 = L + 42;
  l1[0] = 0;
  l1[1] = -cfalf * denom;
  l1[2] = +csals * denom * sgnbx * betay;
  l1[3] = +csals * denom * sgnbx * betaz;
  l1[4] = +aaals * denom * isro * betay;
  l1[5] = +aaals * denom * isro * betaz;
  l1[6] = alphaf * denom * iro;
  mirror7(l1, l7);
  l2[0] = 0;
  l2[1] = 0;
  l2[2] = -0.5 * betaz;
  l2[3] = +0.5 * betay;
  l2[4] = -0.5 * isro * betaz * sgnbx;
  l2[5] = +0.5 * isro * betay * sgnbx;
  l2[6] = 0;
  mirror7(l2, l6);
  l3[0] = 0;
  l3[1] = -csals * denom;
  l3[2] = -cfalf * denom * sgnbx * betay;
  l3[3] = -cfalf * denom * sgnbx * betaz;
  l3[4] = -aaalf * denom * isro * betay;
  l3[5] = -aaalf * denom * isro * betaz;
  l3[6] = alphas * denom * iro;
  mirror7(l3, l5);
  l4[0] = 1.0; l4[1] = 0; l4[2] = 0; l4[3] = 0; l4[4] = 0; l4[5] = 0;
  l4[6] = -1.0 / a2;

  double *r1 = R, *r2 = R + 7, *r3 = R + 14, *r4 = R + 21, *r5 = R + 28, *r6 = R + 35, *r7 = R + 42;
  r1[0] = ro * alphaf;
  r1[1] = -cfalf;
  r1[2] = +csals * sgnbx * betay;
  r1[3] = +csals * sgnbx * betaz;
  r1[4] = +aaals * sro * betay;
  r1[5] = +aaals * sro * betaz;
  r1[6] = alphaf * gamma * pr;
  mirror7(r1, r7);
  r2[0] = 0;
  r2[1] = 0;
  r2[2] = -betaz;
  r2[3] = +betay;
  r2[4] = -sro * betaz * sgnbx;
  r2[5] = +sro * betay * sgnbx;
  r2[6] = 0;
  mirror7(r2, r6);
  r3[0] = ro * alphas;
  r3[1] = -csals;
  r3[2] = -cfalf * sgnbx * betay;
  r3[3] = -cfalf * sgnbx * betaz;
  r3[4] = -aaalf * sro * betay;
  r3[5] = -aaalf * sro * betaz;
  r3[6] = alphas * gamma * pr;
  mirror7(r3, r5);
  r4[0] = 1.0; r4[1] = 0; r4[2] = 0; r4[3] = 0; r4[4] = 0; r4[5] = 0; r4[6] = 0;
}

/* include/mhd_eigen.h:326-453 (eigenvalues are not produced by the reference) */
void oc_eigen_ae(const double *p, double gamma, double *L, double *R)
{
  const double ro = p[0], bx = p[4], by = p[5], bz = p[6], pr = p[7];
  double iro = 1.0 / ro;
  double cs2 = gamma * pr * iro;
  double ca2 = bx * bx * iro;
  double ct2 = (by * by + bz * bz) * iro;
  double cb2 = cs2 + ca2 + ct2;
  double cf2 = 0.5 * (cb2 + sqrt(fabs(cb2 * cb2 - 4.0 * cs2 * ca2)));
  double sro = sqrt(ro);
  double isroh = 0.5 / sro;
  double icf2 = 1.0 / cf2;
  double byh = 0.5 * by, bzh = 0.5 * bz;
  int n;
  for (n = 0; n < 49; n++) { L[n] = 0; R[n] = 0; }
  /* rows l1..l7 */
  L[0 * 7 + 1] = +1.0;
  L[1 * 7 + 0] = cf2; L[1 * 7 + 4] = -by; L[1 * 7 + 5] = -bz; L[1 * 7 + 6] = -1.0;
  L[2 * 7 + 4] = +by; L[2 * 7 + 5] = +bz; L[2 * 7 + 6] = +1.0;
  L[3 * 7 + 2] = +sro; L[3 * 7 + 4] = +1.0;
  L[4 * 7 + 2] = +sro; L[4 * 7 + 4] = -1.0;
  L[5 * 7 + 3] = +sro; L[5 * 7 + 5] = +1.0;
  L[6 * 7 + 3] = +sro; L[6 * 7 + 5] = -1.0;
  /* columns r1..r7 */
  R[0 * 7 + 1] = +1.0;
  R[1 * 7 + 0] = +icf2;
  R[2 * 7 + 0] = +icf2; R[2 * 7 + 6] = +1.0;
  R[3 * 7 + 2] = +isroh; R[3 * 7 + 4] = +0.5; R[3 * 7 + 6] = -byh;
  R[4 * 7 + 2] = +isroh; R[4 * 7 + 4] = -0.5; R[4 * 7 + 6] = +byh;
  R[5 * 7 + 3] = +isroh; R[5 * 7 + 5] = +0.5; R[5 * 7 + 6] = -bzh;
  R[6 * 7 + 3] = +isroh; R[6 * 7 + 5] = -0.5; R[6 * 7 + 6] = +bzh;
}

/* Face values of one 5-point stencil for spatial order odr: the solvers'
 * calc_flr (2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:55-68). */
static void face_lr(int odr, const double *f, double *fl, double *fr)
{
  if (odr == 1) { *fl = f[2]; *fr = f[2]; }
  else if (odr == 2) oc_muscl_mm(f[1], f[2], f[3], fl, fr);
  else if (odr == 3) oc_wcns3(f[1], f[2], f[3], fl, fr);
  else oc_wcns4(f[0], f[1], f[2], f[3], f[4], fl, fr);
}

/* include/mhd_eigen.h:455-530 (kind 0) and :609-684 (kind 1); s = 5.
 * st[7*v+m]: variable v (ro,vn,vt,vu,bt,bu,pr) at stencil point m. */
void oc_reconst(int kind, int odr, const double *st, double bn, double gamma, double *ul, double *ur)
{
  double L[49], R[49], w[7][5], wl[7], wr[7], p[8];
  const double *ro = st, *vx = st + 5, *vy = st + 10, *vz = st + 15, *by = st + 20, *bz = st + 25,
               *pr = st + 30;
  int m, n;
  p[0] = ro[2]; p[1] = vx[2]; p[2] = vy[2]; p[3] = vz[2];
  p[4] = bn; p[5] = by[2]; p[6] = bz[2]; p[7] = pr[2];
  if (kind == 0) oc_eigen_st(p, gamma, 0, L, R);
  else oc_eigen_ae(p, gamma, L, R);
  for (m = 0; m < 5; m++) {
    for (n = 0; n < 7; n++) {
      const double *l = L + 7 * n;
      w[n][m] = (+l[0] * ro[m] + l[1] * vx[m] + l[2] * vy[m] + l[3] * vz[m] + l[4] * by[m] +
                 l[5] * bz[m] + l[6] * pr[m]);
    }
  }
  for (n = 0; n < 7; n++) face_lr(odr, w[n], &wl[n], &wr[n]);
  for (n = 0; n < 7; n++)
    ul[n] = R[0 + n] * wl[0] + R[7 + n] * wl[1] + R[14 + n] * wl[2] + R[21 + n] * wl[3] +
            R[28 + n] * wl[4] + R[35 + n] * wl[5] + R[42 + n] * wl[6];
  for (n = 0; n < 7; n++)
    ur[n] = R[0 + n] * wr[0] + R[7 + n] * wr[1] + R[14 + n] * wr[2] + R[21 + n] * wr[3] +
            R[28 + n] * wr[4] + R[35 + n] * wr[5] + R[42 + n] * wr[6];
  {
    int f0 = 1, f1 = 1, f2 = 1, f3 = 1, f4 = 1;
    for (m = 0; m < 5; m++) {
      if (vx[m] != 0.0) f0 = 0;
      if (vy[m] != 0.0) f1 = 0;
      if (vz[m] != 0.0) f2 = 0;
      if (by[m] != 0.0) f3 = 0;
      if (bz[m] != 0.0) f4 = 0;
    }
    if (f0) ul[1] = ur[1] = 0.0;
    if (f1) ul[2] = ur[2] = 0.0;
    if (f2) ul[3] = ur[3] = 0.0;
    if (f3) ul[4] = ur[4] = 0.0;
    if (f4) ul[5] = ur[5] = 0.0;
  }
}

/* include/mhd_flux.h:26-213 */
static void flux_mlau(const double *q, double bnc, double gamma, const double *dv, double *f)
{
  const double rol = q[0], vnl = q[1], vtl = q[2], vul = q[3], btl = q[4], bul = q[5], prl = q[6];
  const double ror = q[7], vnr = q[8], vtr = q[9], vur = q[10], btr = q[11], bur = q[12], prr = q[13];
  double gammam1i = 1.0 / (gamma - 1.0);
  double bnc2 = bnc * bnc;
  double sgn = (bnc > 0) ? (1.0) : (-1.0);
  double bnca = sgn * bnc;
  double roli = 1.0 / rol;
  double vl2 = vnl * vnl + vtl * vtl + vul * vul;
  double pml = 0.5 * (btl * btl + bul * bul);
  double ptl = prl + pml;
  double enl = gammam1i * prl + pml + 0.5 * rol * vl2;
  double hnl = (enl + ptl) * roli;
  double rori = 1.0 / ror;
  double vr2 = vnr * vnr + vtr * vtr + vur * vur;
  double pmr = 0.5 * (btr * btr + bur * bur);
  double ptr = prr + pmr;
  double enr = gammam1i * prr + pmr + 0.5 * ror * vr2;
  double hnr = (enr + ptr) * rori;
  double cl2 = gamma * prl * roli;
  double cr2 = gamma * prr * rori;
  double cal2 = bnc2 * roli;
  double car2 = bnc2 * rori;
  double cbl2 = cl2 + cal2 + 2.0 * pml * roli;
  double cbr2 = cr2 + car2 + 2.0 * pmr * rori;
  double cfl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * cl2 * cal2)));
  double cfr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * cr2 * car2)));
  cbl2 = vl2 + cal2 + 2.0 * pml * roli;
  cbr2 = vr2 + car2 + 2.0 * pmr * rori;
  double ccl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * vl2 * cal2)));
  double ccr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * vr2 * car2)));
  double cmax = sqrt(mx_(cfl2, cfr2));
  double cfn = cmax;
  double cfni = 1.0 / cfn;
  double sl = mn_(0.0, mn_(vnl, vnr) - cmax);
  double sr = mx_(0.0, mx_(vnl, vnr) + cmax);
  double mal = vnl * cfni;
  double mar = vnr * cfni;
  double snl = (mal > 0) ? 1.0 : (-1.0);
  double snr = (mar > 0) ? 1.0 : (-1.0);
  double theta = mn_(1.0, (cfn - mn_(dv[0], 0.0)) / (cfn - mn_(dv[1], 0.0)));
  theta *= theta;
  theta *= theta;
  double alpha = 0.1875, beta = 0.125;
  double malp1 = mal + 1.0;
  double marm1 = mar - 1.0;
  double mal2m1 = mal * mal - 1.0;
  double mar2m1 = mar * mar - 1.0;
  double bfl, ffl, bfr, ffr;
  if (snl * mal > 1.0) {
    bfl = 0.5 * (1.0 + snl);
    ffl = bfl * mal;
  } else {
    bfl = +0.25 * malp1 * malp1 * (2.0 - mal) + alpha * mal * mal2m1 * mal2m1;
    ffl = +0.25 * malp1 * malp1 + beta * mal2m1 * mal2m1;
  }
  if (snr * mar > 1.0) {
    bfr = 0.5 * (1.0 - snr);
    ffr = bfr * mar;
  } else {
    bfr = +0.25 * marm1 * marm1 * (2.0 + mar) - alpha * mar * mar2m1 * mar2m1;
    ffr = -0.25 * marm1 * marm1 - beta * mar2m1 * mar2m1;
  }
  double mnc = ffl + ffr;
  double mflux = mnc * cfn;
  mflux -= mx_(1.0 - fabs(mnc), 0) * theta * cfni * (ptr - ptl) / (rol + ror);
  double dir = (mflux > 0) ? 1.0 : (-1.0);
  double dirl = 0.5 * (1.0 + dir);
  double dirr = 1.0 - dirl;
  mflux *= (dirl * rol + dirr * ror);
  double mfl = mflux * dirl;
  double mfr = mflux - mfl;
  double ptot = bfl * ptl + bfr * ptr;
  double cfmod = sqrt(mx_(ccl2, ccr2));
  ptot -= 0.25 * bfl * bfr * (rol + ror) * cfmod * (vnr - vnl);
  ptot -= 0.5 * (1.0 - cfmod * cfni) * (bfl + bfr - 1.0) * (ptl + ptr);
  double slvl = sl - vnl;
  double srvr = sr - vnr;
  double snc = dirl * sl + dirr * sr;
  double vndnm = mflux + dirl * rol * slvl + dirr * ror * srvr;
  double vnc = ((fabs(vndnm) > EPSF) && (snc != 0)) ? (mflux * snc / vndnm) : (dirl * vnl + dirr * vnr);
  double slvc = sl - vnc;
  double srvc = sr - vnc;
  double slvci = 1.0 / slvc;
  double srvci = 1.0 / srvc;
  double rhdl = rol * slvl * slvc - bnc2;
  double rhdr = ror * srvr * srvc - bnc2;
  double dvtl = 0.0, dvul = 0.0, dbtl = 0.0, dbul = 0.0;
  double dvtr = 0.0, dvur = 0.0, dbtr = 0.0, dbur = 0.0;
  if (fabs(rhdl) > EPSF) {
    double dvrhdli = (vnc - vnl) / rhdl;
    dvtl = -bnc * btl * dvrhdli;
    dvul = -bnc * bul * dvrhdli;
    dbtl = bnc2 * btl * dvrhdli * slvci;
    dbul = bnc2 * bul * dvrhdli * slvci;
  }
  if (fabs(rhdr) > EPSF) {
    double dvrhdri = (vnc - vnr) / rhdr;
    dvtr = -bnc * btr * dvrhdri;
    dvur = -bnc * bur * dvrhdri;
    dbtr = bnc2 * btr * dvrhdri * srvci;
    dbur = bnc2 * bur * dvrhdri * srvci;
  }
  double cmpl = slvl * slvci;
  double cmpr = srvr * srvci;
  double ro2l = rol * cmpl;
  double srol = sqrt(ro2l);
  double vt2l = dvtl + vtl;
  double vu2l = dvul + vul;
  double bt2l = dbtl + btl * cmpl;
  double bu2l = dbul + bul * cmpl;
  double ro2r = ror * cmpr;
  double sror = sqrt(ro2r);
  double vt2r = dvtr + vtr;
  double vu2r = dvur + vur;
  double bt2r = dbtr + btr * cmpr;
  double bu2r = dbur + bur * cmpr;
  double sroi = 1.0 / (srol + sror);
  double srollri = srol * sroi;
  double srorlri = sror * sroi;
  double bnflb = sgn * mn_(mx_(0.0, srollri * (bnca + sror * vnc)), bnca);
  double bnfrb = sgn * mn_(mx_(0.0, srorlri * (bnca - srol * vnc)), bnca);
  double bnflv = sgn * mn_(mx_(0.0, srorlri * (bnca + mflux / sror)), bnca);
  double bnfrv = sgn * mn_(mx_(0.0, srollri * (bnca - mflux / srol)), bnca);
  double bnsrv = mx_(bnca - dir * vnc * (dirl * srol + dirr * sror), 0.0);
  double ccb = sroi * bnsrv;
  double ccv = srol * sror * ccb;
  double bnvt = bnflb * vt2l + bnfrb * vt2r + ccb * (bt2r - bt2l);
  double bnvu = bnflb * vu2l + bnfrb * vu2r + ccb * (bu2r - bu2l);
  double bnbt = bnflv * bt2l + bnfrv * bt2r + ccv * (vt2r - vt2l);
  double bnbu = bnflv * bu2l + bnfrv * bu2r + ccv * (vu2r - vu2l);
  bnvt -= vnc * (dirl * dbtl + dirr * dbtr);
  bnvu -= vnc * (dirl * dbul + dirr * dbur);
  bnbt -= mflux * (dirl * dvtl + dirr * dvtr);
  bnbu -= mflux * (dirl * dvul + dirr * dvur);
  double fro = mflux;
  double fmn = mfl * vnl + mfr * vnr + ptot - 0.5 * bnc2;
  double fmt = mfl * vtl + mfr * vtr - bnbt;
  double fmu = mfl * vul + mfr * vur - bnbu;
  double fbt = mfl * btl * roli + mfr * btr * rori - bnvt;
  double fbu = mfl * bul * roli + mfr * bur * rori - bnvu;
  double vb1 = dirl * (vtl * btl + vul * bul) + dirr * (vtr * btr + vur * bur);
  double vb2 = dirl * (vt2l * bt2l + vu2l * bu2l) + dirr * (vt2r * bt2r + vu2r * bu2r);
  double rv2b2 = (mflux * vnc - bnc2);
  rv2b2 *= rv2b2;
  double vb3 = 0.0;
  if (rv2b2 > EPSF * EPSF) {
    vb3 = (+(fmt * vnc + fbt * bnc) * (fmt * bnc + fbt * mflux) +
           (fmu * vnc + fbu * bnc) * (fmu * bnc + fbu * mflux)) / rv2b2;
  }
  double bnvb = sgn * (bnca * (snc * vb2 - vnc * vb1) * (dirl * slvci + dirr * srvci) + bnsrv * (vb3 - vb2));
  f[0] = fro; f[1] = fmn; f[2] = fmt; f[3] = fmu; f[4] = fbt; f[5] = fbu;
  f[6] = mfl * hnl + mfr * hnr - bnvb;
}

/* include/mhd_flux.h:215-454 */
static void flux_roe(const double *q, double bnc, double gamma, double *f)
{
  const double rol = q[0], vnl = q[1], vtl = q[2], vul = q[3], btl = q[4], bul = q[5], prl = q[6];
  const double ror = q[7], vnr = q[8], vtr = q[9], vur = q[10], btr = q[11], bur = q[12], prr = q[13];
  double gammam1 = gamma - 1.0;
  double gammam2 = gamma - 2.0;
  double gammam1i = 1.0 / gammam1;
  double roli = 1.0 / rol;
  double rori = 1.0 / ror;
  double rrol = sqrt(rol);
  double rror = sqrt(ror);
  double rroi = 1.0 / (rrol + rror);
  double robar = rrol * rror;
  double robari = 1.0 / robar;
  double rrobar = sqrt(robar);
  double rrobai = sqrt(robari);
  double vnbar = (rrol * vnl + rror * vnr) * rroi;
  double vtbar = (rrol * vtl + rror * vtr) * rroi;
  double vubar = (rrol * vul + rror * vur) * rroi;
  double bnbar = bnc;
  double btbar = (rrol * btr + rror * btl) * rroi;
  double bubar = (rrol * bur + rror * bul) * rroi;
  double hl = (+gamma * gammam1i * prl * roli + 0.5 * (vnl * vnl + vtl * vtl + vul * vul) +
               (bnbar * bnbar + btl * btl + bul * bul) * roli);
  double hr = (+gamma * gammam1i * prr * rori + 0.5 * (vnr * vnr + vtr * vtr + vur * vur) +
               (bnbar * bnbar + btr * btr + bur * bur) * rori);
  double hbar = (rrol * hl + rror * hr) * rroi;
  double btave = 0.5 * (btl + btr);
  double buave = 0.5 * (bul + bur);
  double sgn = (bnbar > 0) ? (1.0) : (-1.0);
  double db2 = 0.5 * gammam2 * gammam1i * ((btr - btl) * (btr - btl) + (bur - bul) * (bur - bul)) * rroi * rroi;
  double c2 = gammam1 * (+hbar - 0.5 * (vnbar * vnbar + vtbar * vtbar + vubar * vubar) - db2 -
                         (bnbar * bnbar + btbar * btbar + bubar * bubar) * robari);
  double a2 = (gammam1 * (+hbar - 0.5 * (vnbar * vnbar + vtbar * vtbar + vubar * vubar) - db2) -
               gammam2 * (bnbar * bnbar + btbar * btbar + bubar * bubar) * robari);
  double ca2 = bnbar * bnbar * robari;
  double cb2 = (btbar * btbar + bubar * bubar) * robari;
  double cf2 = 0.5 * (a2 + sqrt(cb2 * (a2 + c2 + ca2) + (c2 - ca2) * (c2 - ca2)));
  double cs2 = c2 * ca2 / cf2;
  double cf = sqrt(cf2);
  double cs = sqrt(cs2);
  double ca = sqrt(ca2);
  double c = sqrt(c2);
  double cfi = 1.0 / cf;
  double sgr1 = btbar * btbar + bubar * bubar - EPSF;
  double sf11 = (sgr1 > 0) ? 1.0 : 0.0;
  double sf21 = 1.0 - sf11;
  double beti = 1.0 / sqrt(btbar * btbar + bubar * bubar + sf21);
  double bett = sf11 * btbar * beti + sf21 * sqrt(0.5);
  double betu = sf11 * bubar * beti + sf21 * sqrt(0.5);
  double sgr2 = (btbar * btbar + bubar * bubar) * robari + fabs(ca2 - c2) - EPSF;
  double sf12 = (sgr2 > 0) ? 1.0 : 0.0;
  double sf22 = 1.0 - sf12;
  double cfca = mx_(0.0, cf2 - ca2);
  double cfcs = mx_(0.0, cf2 - cs2);
  double cfa = mx_(0.0, cf2 - c2);
  double alpi = 1.0 / sqrt(cfcs + sf22);
  double alpf = sf12 * sqrt(cfca) * alpi + sf22;
  double alps = sf12 * sqrt(cfa) * alpi;
  double eeps = 0.0;
  double evpf = mx_(fabs(vnbar + cf), eeps);
  double evmf = mx_(fabs(vnbar - cf), eeps);
  double evps = mx_(fabs(vnbar + cs), eeps);
  double evms = mx_(fabs(vnbar - cs), eeps);
  double evpa = mx_(fabs(vnbar + ca), eeps);
  double evma = mx_(fabs(vnbar - ca), eeps);
  double ev00 = mx_(fabs(vnbar), eeps);
  double dltro = ror - rol;
  double dltmn = robar * (vnr - vnl);
  double dltmt = robar * (vtr - vtl);
  double dltmu = robar * (vur - vul);
  double dltbt = btr - btl;
  double dltbu = bur - bul;
  double t1 = bett * dltbt + betu * dltbu;
  double t2 = (prr - prl + (btave * dltbt + buave * dltbu) + gammam2 * (btbar * dltbt + bubar * dltbu)) * gammam1i;
  double t3 = betu * dltbt - bett * dltbu;
  double s1 = dltmn;
  double s2 = bett * dltmt + betu * dltmu;
  double s3 = betu * dltmt - bett * dltmu;
  double p11 = alps * cf * rrobai;
  double p12 = -alpf * c2 * cfi * rrobai;
  double p21 = alpf * (cf2 - c2 * gammam2 * gammam1i);
  double p22 = alps * (cs2 - c2 * gammam2 * gammam1i);
  double q11 = alpf * cf;
  double q12 = alps * cs;
  double q21 = -alps * ca * sgn;
  double q22 = alpf * c * sgn;
  double detpi = 1.0 / (p11 * p22 - p12 * p21);
  double detqi = 1.0 / (q11 * q22 - q12 * q21);
  double wvpf = 0.5 * ((p22 * t1 - p12 * t2) * detpi + (q22 * s1 - q12 * s2) * detqi);
  double wvmf = 0.5 * ((p22 * t1 - p12 * t2) * detpi - (q22 * s1 - q12 * s2) * detqi);
  double wvps = 0.5 * ((-p21 * t1 + p11 * t2) * detpi + (-q21 * s1 + q11 * s2) * detqi);
  double wvms = 0.5 * ((-p21 * t1 + p11 * t2) * detpi - (-q21 * s1 + q11 * s2) * detqi);
  double wvpa = 0.5 * (rrobar * t3 - sgn * s3);
  double wvma = 0.5 * (rrobar * t3 + sgn * s3);
  double wv00 = dltro - alpf * (wvpf + wvmf) - alps * (wvps + wvms);
  double frol = rol * vnl;
  double fmnl = rol * vnl * vnl + prl + 0.5 * (-bnbar * bnbar + btl * btl + bul * bul);
  double fmtl = rol * vnl * vtl - bnbar * btl;
  double fmul = rol * vnl * vul - bnbar * bul;
  double fbtl = vnl * btl - vtl * bnbar;
  double fbul = vnl * bul - vul * bnbar;
  double fenl = rol * vnl * hl - bnbar * (bnbar * vnl + btl * vtl + bul * vul);
  double fror = ror * vnr;
  double fmnr = ror * vnr * vnr + prr + 0.5 * (-bnbar * bnbar + btr * btr + bur * bur);
  double fmtr = ror * vnr * vtr - bnbar * btr;
  double fmur = ror * vnr * vur - bnbar * bur;
  double fbtr = vnr * btr - vtr * bnbar;
  double fbur = vnr * bur - vur * bnbar;
  double fenr = ror * vnr * hr - bnbar * (bnbar * vnr + btr * vtr + bur * vur);
  double v2h = 0.5 * (vnbar * vnbar + vtbar * vtbar + vubar * vubar);
  double rvpfro = alpf;
  double rvpfmn = alpf * (vnbar + cf);
  double rvpfmt = alpf * vtbar - alps * bett * ca * sgn;
  double rvpfmu = alpf * vubar - alps * betu * ca * sgn;
  double rvpfbt = alps * bett * cf * rrobai;
  double rvpfbu = alps * betu * cf * rrobai;
  double rvpfen = alpf * (v2h + db2 + cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i) -
                  alps * ca * (bett * vtbar + betu * vubar) * sgn;
  double rvmfro = alpf;
  double rvmfmn = alpf * (vnbar - cf);
  double rvmfmt = alpf * vtbar + alps * bett * ca * sgn;
  double rvmfmu = alpf * vubar + alps * betu * ca * sgn;
  double rvmfbt = rvpfbt;
  double rvmfbu = rvpfbu;
  double rvmfen = alpf * (v2h + db2 - cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i) +
                  alps * ca * (bett * vtbar + betu * vubar) * sgn;
  double rvpsro = alps;
  double rvpsmn = alps * (vnbar + cs);
  double rvpsmt = alps * vtbar + c * sgn * alpf * bett;
  double rvpsmu = alps * vubar + c * sgn * alpf * betu;
  double rvpsbt = -rrobai * c2 * alpf * bett * cfi;
  double rvpsbu = -rrobai * c2 * alpf * betu * cfi;
  double rvpsen = alps * (v2h + db2 + cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i) +
                  alpf * c * (bett * vtbar + betu * vubar) * sgn;
  double rvmsro = alps;
  double rvmsmn = alps * (vnbar - cs);
  double rvmsmt = alps * vtbar - c * sgn * alpf * bett;
  double rvmsmu = alps * vubar - c * sgn * alpf * betu;
  double rvmsbt = rvpsbt;
  double rvmsbu = rvpsbu;
  double rvmsen = alps * (v2h + db2 - cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i) -
                  alpf * c * (bett * vtbar + betu * vubar) * sgn;
  double rvparo = 0.0, rvpamn = 0.0;
  double rvpamt = -sgn * betu;
  double rvpamu = sgn * bett;
  double rvpabt = rrobai * betu;
  double rvpabu = -rrobai * bett;
  double rvpaen = -(betu * vtbar - bett * vubar) * sgn;
  double rvmaro = 0.0, rvmamn = 0.0;
  double rvmamt = -rvpamt;
  double rvmamu = -rvpamu;
  double rvmabt = rvpabt;
  double rvmabu = rvpabu;
  double rvmaen = -rvpaen;
  double rv00ro = 1.0, rv00mn = vnbar, rv00mt = vtbar, rv00mu = vubar, rv00bt = 0.0, rv00bu = 0.0;
  double rv00en = v2h + db2;
#define ROE_SUM(FL, FR, S)                                                                        \
  0.5 * (+FL + FR - evpf * wvpf * rvpf##S - evmf * wvmf * rvmf##S - evps * wvps * rvps##S -         \
         evms * wvms * rvms##S - evpa * wvpa * rvpa##S - evma * wvma * rvma##S - ev00 * wv00 * rv00##S)
  f[0] = ROE_SUM(frol, fror, ro);
  f[1] = ROE_SUM(fmnl, fmnr, mn);
  f[2] = ROE_SUM(fmtl, fmtr, mt);
  f[3] = ROE_SUM(fmul, fmur, mu);
  f[4] = ROE_SUM(fbtl, fbtr, bt);
  f[5] = ROE_SUM(fbul, fbur, bu);
  f[6] = ROE_SUM(fenl, fenr, en);
#undef ROE_SUM
}

/* include/mhd_flux.h:456-582 (lowdiss = 0) and :584-719 (lowdiss = 1) */
static void flux_hlld(const double *q, double bnc, double gamma, const double *dv, double *f, int lowdiss)
{
  const double rol = q[0], vnl = q[1], vtl = q[2], vul = q[3], btl = q[4], bul = q[5], prl = q[6];
  const double ror = q[7], vnr = q[8], vtr = q[9], vur = q[10], btr = q[11], bur = q[12], prr = q[13];
  double gammam1i = 1.0 / (gamma - 1.0);
  double bnc2 = bnc * bnc;
  double sgn = (bnc > 0) ? (1.0) : (-1.0);
  double roli = 1.0 / rol;
  double vl2 = vnl * vnl + vtl * vtl + vul * vul;
  double pml = 0.5 * (btl * btl + bul * bul);
  double ptl = prl + pml;
  double enl = gammam1i * prl + pml + 0.5 * rol * vl2;
  double vbl = vtl * btl + vul * bul;
  double rori = 1.0 / ror;
  double vr2 = vnr * vnr + vtr * vtr + vur * vur;
  double pmr = 0.5 * (btr * btr + bur * bur);
  double ptr = prr + pmr;
  double enr = gammam1i * prr + pmr + 0.5 * ror * vr2;
  double vbr = vtr * btr + vur * bur;
  double cl2 = gamma * prl * roli;
  double cr2 = gamma * prr * rori;
  double cal2 = bnc2 * roli;
  double car2 = bnc2 * rori;
  double cbl2 = cl2 + cal2 + 2.0 * pml * roli;
  double cbr2 = cr2 + car2 + 2.0 * pmr * rori;
  double cfl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * cl2 * cal2)));
  double cfr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * cr2 * car2)));
  double cmax, vnc, ptc;
  double sl, sr, slvl, srvr, rslvl, rsrvr, drsvi;
  if (lowdiss) {
    cbl2 = vl2 + cal2 + 2.0 * pml * roli;
    cbr2 = vr2 + car2 + 2.0 * pmr * rori;
    double ccl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * vl2 * cal2)));
    double ccr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * vr2 * car2)));
    cmax = sqrt(mx_(cfl2, cfr2));
    sl = mn_(0.0, mn_(vnl, vnr) - cmax);
    sr = mx_(0.0, mx_(vnl, vnr) + cmax);
    slvl = sl - vnl;
    srvr = sr - vnr;
    rslvl = rol * slvl;
    rsrvr = ror * srvr;
    drsvi = 1.0 / (rsrvr - rslvl);
    double theta = mn_(1.0, (cmax - mn_(dv[0], 0.0)) / (cmax - mn_(dv[1], 0.0)));
    theta *= theta;
    theta *= theta;
    vnc = (rsrvr * vnr - rslvl * vnl - theta * (ptr - ptl)) * drsvi;
    double mcc = mn_(1.0, sqrt(mx_(ccl2, ccr2)) / cmax);
    ptc = (rsrvr * ptl - rslvl * ptr + (mcc * (2.0 - mcc)) * rsrvr * rslvl * (vnr - vnl)) * drsvi;
  } else {
    double cfl = sqrt(cfl2);
    double cfr = sqrt(cfr2);
    cmax = mx_(cfl, cfr);
    sl = mn_(0.0, mn_(vnl, vnr) - cmax);
    sr = mx_(0.0, mx_(vnl, vnr) + cmax);
    slvl = sl - vnl;
    srvr = sr - vnr;
    rslvl = rol * slvl;
    rsrvr = ror * srvr;
    drsvi = 1.0 / (rsrvr - rslvl);
    vnc = (rsrvr * vnr - rslvl * vnl - ptr + ptl) * drsvi;
    ptc = (rsrvr * ptl - rslvl * ptr + rsrvr * rslvl * (vnr - vnl)) * drsvi;
  }
  double slvc = sl - vnc;
  double srvc = sr - vnc;
  double ro2l = rslvl / slvc;
  double ro2r = rsrvr / srvc;
  double rhdl = rslvl * slvc - bnc2;
  double rhdr = rsrvr * srvc - bnc2;
  double vt2l, vu2l, bt2l, bu2l, vt2r, vu2r, bt2r, bu2r;
  if (fabs(rhdl) > EPSF) {
    double rhdli = 1.0 / rhdl;
    double rhnvl = (vnl - vnc) * bnc;
    double rhnbl = rslvl * slvl - bnc2;
    vt2l = vtl + rhnvl * rhdli * btl;
    vu2l = vul + rhnvl * rhdli * bul;
    bt2l = rhnbl * rhdli * btl;
    bu2l = rhnbl * rhdli * bul;
  } else {
    vt2l = vtl; vu2l = vul; bt2l = btl; bu2l = bul;
  }
  if (fabs(rhdr) > EPSF) {
    double rhdri = 1.0 / rhdr;
    double rhnvr = (vnr - vnc) * bnc;
    double rhnbr = rsrvr * srvr - bnc2;
    vt2r = vtr + rhnvr * rhdri * btr;
    vu2r = vur + rhnvr * rhdri * bur;
    bt2r = rhnbr * rhdri * btr;
    bu2r = rhnbr * rhdri * bur;
  } else {
    vt2r = vtr; vu2r = vur; bt2r = btr; bu2r = bur;
  }
  double vb2l = vt2l * bt2l + vu2l * bu2l;
  double vb2r = vt2r * bt2r + vu2r * bu2r;
  double en2l = (slvl * enl - ptl * vnl + ptc * vnc + bnc * (vbl - vb2l)) / slvc;
  double en2r = (srvr * enr - ptr * vnr + ptc * vnc + bnc * (vbr - vb2r)) / srvc;
  double rro2l = sqrt(ro2l);
  double rro2r = sqrt(ro2r);
  double rro2i = 1.0 / (rro2r + rro2l);
  double vt3m = (rro2r * vt2r + rro2l * vt2l + (bt2r - bt2l) * sgn) * rro2i;
  double vu3m = (rro2r * vu2r + rro2l * vu2l + (bu2r - bu2l) * sgn) * rro2i;
  double bt3m = (rro2l * bt2r + rro2r * bt2l + rro2r * rro2l * (vt2r - vt2l) * sgn) * rro2i;
  double bu3m = (rro2l * bu2r + rro2r * bu2l + rro2r * rro2l * (vu2r - vu2l) * sgn) * rro2i;
  double vb3m = vt3m * bt3m + vu3m * bu3m;
  double en3l = en2l - rro2l * (vb2l - vb3m) * sgn;
  double en3r = en2r + rro2r * (vb2r - vb3m) * sgn;
  double hl = (vnc > 0) ? (1.0) : (0.0);
  double hr = 1.0 - hl;
  double h2l = (vnc - fabs(bnc) / rro2l > 0) ? (1.0) : (0.0);
  double h3l = (1.0 - h2l) * hl;
  double h2r = (vnc + fabs(bnc) / rro2r > 0) ? (0.0) : (1.0);
  double h3r = (1.0 - h2r) * hr;
  double rou = ro2l * hl + ro2r * hr;
  double vtu = vt2l * h2l + vt3m * (h3l + h3r) + vt2r * h2r;
  double vuu = vu2l * h2l + vu3m * (h3l + h3r) + vu2r * h2r;
  double btu = bt2l * h2l + bt3m * (h3l + h3r) + bt2r * h2r;
  double buu = bu2l * h2l + bu3m * (h3l + h3r) + bu2r * h2r;
  double enu = en2l * h2l + en3l * h3l + en3r * h3r + en2r * h2r;
  f[0] = rou * vnc;
  f[1] = rou * vnc * vnc + ptc - bnc2 * 0.5;
  f[2] = rou * vtu * vnc - bnc * btu;
  f[3] = rou * vuu * vnc - bnc * buu;
  f[4] = btu * vnc - bnc * vtu;
  f[5] = buu * vnc - bnc * vuu;
  f[6] = (enu + ptc) * vnc - bnc * (vtu * btu + vuu * buu);
}

void oc_flux(int kind, const double *in, double bn, double gamma, const double *dv, double *out)
{
  switch (kind) {
  case OC_FLUX_MLAU: flux_mlau(in, bn, gamma, dv, out); break;
  case OC_FLUX_ROE: flux_roe(in, bn, gamma, out); break;
  case OC_FLUX_HLLD: flux_hlld(in, bn, gamma, dv, out, 0); break;
  default: flux_hlld(in, bn, gamma, dv, out, 1); break;
  }
}
