/*
 * kqed_pert_host.c -- host-side functions of the reference's pi_pert.c that do no table work: the
 * Dirac trace, the modified Bessel functions K_n and the analytic light-by-light results the
 * lepton-loop integrand is compared with.
 */
#include <math.h>

#include "kqed_assembly.h"

#define KQED_ZETA3 1.2020569031595942854
static const double kqed_alpha = 1.0 / 137.035999; /* AlfQED, reference KQED.h:42 */

/* 1/4 Tr(g_idx[0] ... g_idx[n-1]) (pi_pert.c:136-157) */
int diractrace(const int n, const int *idx) { return kqed_quarter_trace(n, idx); }

/* K_0 .. K_n at x, bv[n+1] (pi_pert.c:92-133): Abramowitz & Stegun 9.8.5-9.8.8 for K0 and K1 (the
 * I0, I1 series below x = 2, the large-x asymptotic forms above), then K_{j+1} = K_{j-1} + (2j/x) K_j */
void bessk(const int n, const double x, double *bv)
{
  static const double i0c[7] = { 1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.360768e-1, 0.45813e-2 };
  static const double i1c[7] = { 0.5, 0.87890594, 0.51498869, 0.15084934, 0.2658733e-1, 0.301532e-2, 0.32411e-3 };
  static const double k0s[7] = { -0.57721566, 0.42278420, 0.23069756, 0.3488590e-1, 0.262698e-2, 0.10750e-3, 0.74e-5 };
  static const double k1s[7] = { 1.0, 0.15443144, -0.67278579, -0.18156897, -0.1919402e-1, -0.110404e-2, -0.4686e-4 };
  static const double k0l[7] = { 1.25331414, -0.7832358e-1, 0.2189568e-1, -0.1062446e-1, 0.587872e-2, -0.251540e-2, 0.53208e-3 };
  static const double k1l[7] = { 1.25331414, 0.23498619, -0.3655620e-1, 0.1504268e-1, -0.780353e-2, 0.325614e-2, -0.68245e-3 };
  /* Horner form c0 + t (c1 + t (c2 + ...)), innermost first */
#define HORNER(c, t) (c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * (c[5] + t * c[6]))))))
  const double tox = 2.0 / x;
  double km, k;
  if (x <= 2.0) {
    const double y = (x / 3.75) * (x / 3.75);
    const double i0 = HORNER(i0c, y), i1 = x * HORNER(i1c, y);
    const double q = 0.25 * x * x, l = log(x / 2.0);
    km = (-l * i0) + HORNER(k0s, q);
    k = (l * i1) + (1.0 / x) * HORNER(k1s, q);
  } else {
    const double e = (exp(-x) / sqrt(x));
    km = e * HORNER(k0l, tox);
    k = e * HORNER(k1l, tox);
  }
#undef HORNER
  bv[0] = km;
  bv[1] = k;
  for (int j = 1; j < n; j++) {
    const double kp = km + j * tox * k;
    km = k;
    bv[j + 1] = k = kp;
  }
}

/* Light-by-light contribution of a heavy lepton loop to a_mu, mh = m_lepton / m_mu, including
 * (alpha/pi)^3: the large-mass expansion to 1/mh^10 of Remiddi and Laporta, PLB 301 (1993) 440,
 * Eq. (4) (pi_pert.c:50-68; the reference also prints the five terms) */
double amu_heavlept(const double mh)
{
  const double z2 = M_PI * M_PI / 6, m2 = mh * mh, m4 = m2 * m2, L = log(mh);
  const double c2 = 1.5 * KQED_ZETA3 - 19.0 / 16;
  const double c4 = -161.0 / 810 * L * L - 16189.0 / 48600 * L + 13.0 / 18 * KQED_ZETA3 - 161.0 / 9720 * M_PI * M_PI - 831931.0 / 972000;
  const double c6 = 17.0 / 36 * KQED_ZETA3 - 13.0 / 224 * z2 - 1840256147.0 / 3556224000 - 4381.0 / 120960 * 4.0 * L * L -
                    24761.0 / 317520 * 2.0 * L;
  const double c8 = 7.0 / 20 * KQED_ZETA3 - 2047.0 / 54000 * z2 - 453410778211.0 / 1200225600000 - 5207.0 / 189000 * 4.0 * L * L -
                    41940853.0 / 952560000 * 2.0 * L;
  const double c10 = 5.0 / 18 * KQED_ZETA3 - 1187.0 / 44550 * z2 - 86251554753071.0 / 287550049248000 -
                     328337.0 / 14968800 * 4.0 * L * L - 640572781.0 / 23051952000 * 2.0 * L;
  const double sum = c2 / m2 + c4 / (m2 * m2) + c6 / (m4 * m2) + c8 / (m4 * m4) + c10 / (m4 * m4 * m2);
  return kqed_alpha * kqed_alpha * kqed_alpha / (M_PI * M_PI * M_PI) * sum;
}

/* Light-by-light contribution of a light lepton loop, ml = m_lepton / m_mu, including (alpha/pi)^3:
 * the small-mass expansion to ml^4 of Remiddi and Laporta, PLB 301 (1993) 440, Eq. (2)
 * (pi_pert.c:70-89) */
double amu_lightlept(const double ml)
{
  const double p2 = M_PI * M_PI, p4 = p2 * p2, m2 = ml * ml, L = -log(ml), l2 = log(2.0);
  const double z3 = KQED_ZETA3;
  const double t1 = 2.0 / 3 * p2 * L;
  const double t2 = 59.0 / 270 * p4 - 3.0 * z3 - 10.0 / 3 * p2 + 2.0 / 3;
  const double t3 = ml * (4.0 / 3 * p2 * L - 196.0 / 3 * p2 * l2 + 424.0 / 9 * p2);
  const double t4 = m2 * (-2.0 / 3 * L * L * L + (p2 / 9 - 20.0 / 3) * L * L - (16.0 / 135 * p4 + 4.0 * z3 - 32.0 / 9 * p2 + 61.0 / 3) * L +
                          4.0 / 3 * z3 * p2 - 61.0 / 270 * p4 + 3.0 * z3 + 25.0 / 18 * p2 - 283.0 / 12);
  const double t5 = m2 * ml * (10.0 / 9 * p2 * L - 11.0 / 9 * p2);
  const double t6 = m2 * m2 * (7.0 / 9 * L * L * L + 41.0 / 18 * L * L + 13.0 / 9 * p2 * L + 517.0 / 108 * L + z3 / 2 +
                               191.0 / 216 * p2 + 13283.0 / 2592);
  return kqed_alpha * kqed_alpha * kqed_alpha / (M_PI * M_PI * M_PI) * (t1 + t2 + t3 + t4 + t5 + t6);
}

/* pi_pert.c:384-397, kept as the reference has it: a small-mass expansion with a4 = Li_4(1/2); the
 * reference notes that it does not reproduce the value it is meant to give */
double vac_pol(const double ml)
{
  const double a4 = 0.51747906167389938633;
  const double p2 = M_PI * M_PI, m2 = ml * ml, L = -log(ml), l2 = log(2.0), z3 = KQED_ZETA3;
  const double t1 = 2.0 / 9 * L * L;
  const double t2 = (z3 - 2.0 / 3 * p2 * l2 + p2 / 9 + 31.0 / 27) * L;
  const double t3 = 11.0 / 216 * p2 * p2 - 2.0 / 9 * p2 * l2 * l2 - 8.0 / 3 * a4 - 1.0 / 9 * l2 * l2 * l2 * l2 - 3.0 * z3 +
                    5.0 / 3 * p2 * l2 - 25.0 / 18 * p2 + 1075.0 / 216;
  const double t4 = ml * (-13.0 / 18 * p2 * M_PI - 16.0 / 9 * p2 * l2 + 3199.0 / 1080 * p2);
  const double t5 = m2 * (10.0 / 3 * L * L - 11.0 / 9 * L - 14.0 / 3 * p2 * l2 - 2.0 * z3 + 49.0 / 12 * p2 - 131.0 / 54);
  const double t6 = ml * m2 * (4.0 / 3 * p2 * L + 35.0 / 12 * p2 * M_PI - 16.0 / 3 * p2 * l2 - 5771.0 / 1080 * p2);
  return t1 + t2 + t3 + t4 + t5 + t6;
}
