// weighting.cpp -- data uncertainty, synthetic noise and the data misfit of
// the inversion (src/cal_uncertainty_noise.c:17-98, src/frannor.c:90-169,
// src/fg_fwi.c:207-225).
//
// The observed data of a gradient run carry noise drawn from frannor, so
// parity with the reference's obs_<shot>.txt files needs the same normal
// deviates bit for bit: the generator below is the RNOR method of Marsaglia
// and Tsang (SIAM J. Sci. Stat. Comput. 5(2), 1984) with the lagged
// Fibonacci uniform source and the constants published by Kahaner, Moler and
// Nash ("Numerical Methods and Software", 1988), evaluated with the same
// float/double steps as the reference so that every deviate rounds alike.
// The float/double mix of the uncertainty and misfit formulas is kept for
// the same reason (each line notes the precision the reference's expression
// evaluates in).
#include <cmath>
#include <cstring>

#include "emmi_cli.h"

namespace {

// RNOR: abscissae v[0..64] of the 64 equal-area strips under the normal
// density (the last three repeat the tail start) and the fast-path limits
constexpr double kAA = 12.37586, kB = 0.4878992, kC = 12.67706, kC1 = 0.9689279, kC2 = 1.301198;
constexpr double kPC = 0.01958303, kXN = 2.776994, kOXN = 0.3601016;
constexpr float kStrip[65] = {
    0.3409450f, 0.4573146f, 0.5397793f, 0.6062427f, 0.6631691f, 0.7136975f, 0.7596125f, 0.8020356f, 0.8417227f,
    0.8792102f, 0.9148948f, 0.9490791f, 0.9820005f, 1.0138492f, 1.0447810f, 1.0749254f, 1.1043917f, 1.1332738f,
    1.1616530f, 1.1896010f, 1.2171815f, 1.2444516f, 1.2714635f, 1.2982650f, 1.3249008f, 1.3514125f, 1.3778399f,
    1.4042211f, 1.4305929f, 1.4569915f, 1.4834526f, 1.5100121f, 1.5367061f, 1.5635712f, 1.5906454f, 1.6179680f,
    1.6455802f, 1.6735255f, 1.7018503f, 1.7306045f, 1.7598422f, 1.7896223f, 1.8200099f, 1.8510770f, 1.8829044f,
    1.9155830f, 1.9492166f, 1.9839239f, 2.0198430f, 2.0571356f, 2.0959930f, 2.1366450f, 2.1793713f, 2.2245175f,
    2.2725185f, 2.3239338f, 2.3795007f, 2.4402218f, 2.5075117f, 2.5834658f, 2.6713916f, 2.7769943f, 2.7769943f,
    2.7769943f, 2.7769943f};

// the lagged Fibonacci generator u[i] <- u[i] - u[j] mod 1 (lags 17 and 5)
// and the 17 words a fresh process starts from
constexpr float kFresh[17] = {0.8668672834288f,  0.3697986366357f,  0.8008968294805f, 0.4173889774680f,
                              0.8254561579836f,  0.9640965269077f,  0.4508667414265f, 0.6451309529668f,
                              0.1645456024730f,  0.2787901807898f,  0.06761531340295f, 0.9663226330820f,
                              0.01963343943798f, 0.02947398211399f, 0.1636231515294f, 0.3976343250467f,
                              0.2631008574685f};

struct Lagged {
  float u[17];
  int i = 16, j = 4;
  Lagged() { reset(); }
  void reset() {
    memcpy(u, kFresh, sizeof u);
    i = 16;
    j = 4;
  }
  float next() {
    float x = u[i] - u[j];
    if (x < 0.0) x = static_cast<float>(x + 1.0);
    u[i] = x;
    if (--i < 0) i = 16;
    if (--j < 0) j = 16;
    return x;
  }
};
Lagged g_uni;

inline float fabsf_(float x) { return x < 0 ? -x : x; }

// real-by-complex products of the reference's C99 arithmetic: each component
// scaled by a real (no cross terms), in the precision of the wider operand
inline emmi_c64 add_scaled(emmi_c64 acc, emmi_c64 z, double c, double sign) {
  return {static_cast<float>(acc.re + sign * (z.re * c)), static_cast<float>(acc.im + sign * (z.im * c))};
}
inline double abs_scaled(float a, emmi_c64 z) {  // cabs(a * z), a * z in float
  return std::hypot(static_cast<double>(a * z.re), static_cast<double>(a * z.im));
}
// sqrt of a float sum taken in double, rounded back (C's sqrt of a float)
inline float root(float x) { return static_cast<float>(std::sqrt(static_cast<double>(x))); }
inline double abs_c(emmi_c64 z) { return std::hypot(static_cast<double>(z.re), static_cast<double>(z.im)); }
}  // namespace

extern "C" {

float frannor(void) {
  const float uni = g_uni.next();
  const float vni = static_cast<float>((uni + uni) - 1.0);
  const int k = static_cast<int>(g_uni.u[g_uni.i] * 128) % 64;
  const float rnor = vni * kStrip[k + 1];
  if (fabsf_(rnor) <= kStrip[k]) return rnor;
  // the wedge and the tail
  float x = (fabsf_(rnor) - kStrip[k]) / (kStrip[k + 1] - kStrip[k]);
  float y = g_uni.next();
  const float s = x + y;
  if (s <= kC2) {
    if (s <= kC1) return rnor;
    const float bmbx = static_cast<float>(kB - kB * x);
    if (y <= kC - kAA * std::exp(-0.5 * bmbx * bmbx)) {
      const double vk = kStrip[k + 1];
      if (std::exp(-0.5 * vk * vk) + y * kPC / vk <= std::exp(-0.5 * rnor * rnor)) return rnor;
      do {
        y = g_uni.next();
        x = static_cast<float>(kOXN * std::log(static_cast<double>(y)));
        y = g_uni.next();
      } while (-2.0 * std::log(static_cast<double>(y)) <= x * x);
      const float xnmx = static_cast<float>(kXN - x);
      return rnor >= 0.0 ? fabsf_(xnmx) : -fabsf_(xnmx);
    }
  }
  const float bmbx = static_cast<float>(kB - kB * x);
  return rnor >= 0.0 ? fabsf_(bmbx) : -fabsf_(bmbx);
}

void srannor(int seed) {
  g_uni.i = 16;
  g_uni.j = 4;
  int ia = (seed < 0 ? -seed : seed) % 32707, ib = 1111, ic = 1947;
  for (int w = 0; w < 17; ++w) {
    float s = 0.0f, t = 0.5f;
    for (int b = 0; b < 64; ++b) {
      int id = ic - ia;
      if (id < 0) {
        id += 32707;
        s += t;
      }
      ia = ib;
      ib = ic;
      ic = id;
      t *= 0.5f;
    }
    g_uni.u[w] = s;
  }
}

void emmi_rannor_reset(void) { g_uni.reset(); }

float cal_uncertainty_noise(acq_t* acq, emf_t* emf) {
  const int nc = emf->nchrec, nf = emf->nfreq, nr = acq->nrec;
  memset(&emf->delta_emf[0][0][0], 0, sizeof(float) * nc * nf * nr);
  float mse = 0;
  for (int r = 0; r < nr; ++r) {
    // azimuth between the source towline and the receiver
    const float phi = acq->rec_azimuth[r] - acq->src_azimuth[0];
    const double cp = std::cos(static_cast<double>(phi)), sp = std::sin(static_cast<double>(phi));
    for (int f = 0; f < nf; ++f) {
      // the horizontal components as recorded and rotated into the towline
      emmi_c64 Ex{0, 0}, Ey{0, 0}, Hx{0, 0}, Hy{0, 0};
      emmi_c64 Exc{0, 0}, Eyc{0, 0}, Hxc{0, 0}, Hyc{0, 0};
      for (int c = 0; c < nc; ++c) {
        const char* ch = emf->chrec[c];
        const emmi_c64 d = emf->dobs_fd[c][f][r];
        if (!strcmp(ch, "Ex")) {
          Ex = d;
          Exc = add_scaled(Exc, d, cp, 1);
          Eyc = add_scaled(Eyc, d, sp, 1);
        } else if (!strcmp(ch, "Ey")) {
          Ey = d;
          Exc = add_scaled(Exc, d, sp, -1);
          Eyc = add_scaled(Eyc, d, cp, 1);
        } else if (!strcmp(ch, "Hx")) {
          Hx = d;
          Hxc = add_scaled(Hxc, d, cp, 1);
          Hyc = add_scaled(Hyc, d, sp, 1);
        } else if (!strcmp(ch, "Hy")) {
          Hy = d;
          Hxc = add_scaled(Hxc, d, sp, -1);
          Hyc = add_scaled(Hyc, d, cp, 1);
        }
      }
      const float amp = emf->amp_perc, dphi = emf->delta_phi;
      for (int c = 0; c < nc; ++c) {
        const char* ch = emf->chrec[c];
        float s1, s2, s3;
        float* delta = &emf->delta_emf[c][f][r];
        if (!strcmp(ch, "Ex")) {
          s1 = static_cast<float>(abs_scaled(amp, Ex) * cp);
          s2 = static_cast<float>(abs_scaled(amp, Ey) * sp);
          s3 = static_cast<float>(dphi * abs_c(Eyc));
          *delta = root(s1 * s1 + s2 * s2 + s3 * s3 + emf->noisefloorE * emf->noisefloorE);
        } else if (!strcmp(ch, "Ey")) {
          s1 = static_cast<float>(abs_scaled(amp, Ex) * sp);
          s2 = static_cast<float>(abs_scaled(amp, Ey) * cp);
          s3 = static_cast<float>(dphi * abs_c(Exc));
          *delta = root(s1 * s1 + s2 * s2 + s3 * s3 + emf->noisefloorE * emf->noisefloorE);
        } else if (!strcmp(ch, "Hx")) {
          s1 = static_cast<float>(abs_scaled(amp, Hx) * cp);
          s2 = static_cast<float>(abs_scaled(amp, Hy) * sp);
          s3 = static_cast<float>(dphi * abs_c(Hyc));
          *delta = root(s1 * s1 + s2 * s2 + s3 * s3 + emf->noisefloorH * emf->noisefloorH);
        } else if (!strcmp(ch, "Hy")) {
          s1 = static_cast<float>(abs_scaled(amp, Hx) * sp);
          s2 = static_cast<float>(abs_scaled(amp, Hy) * cp);
          s3 = static_cast<float>(abs_scaled(dphi, Hxc));  // the reference scales before the modulus here
          *delta = root(s1 * s1 + s2 * s2 + s3 * s3 + emf->noisefloorH * emf->noisefloorH);
        }
        if (emf->addnoise) {
          // complex N(0,1)/sqrt(2): the imaginary deviate is drawn first, the
          // order in which the reference's build evaluates the two calls of
          // (frannor() + I*frannor()) (its shipped obs files confirm it)
          const float im = frannor();
          const float re = frannor();
          const float gr = static_cast<float>(re / std::sqrt(2.)), gi = static_cast<float>(im / std::sqrt(2.));
          emmi_c64& d = emf->dobs_fd[c][f][r];
          d.re = d.re + *delta * gr;
          d.im = d.im + *delta * gi;
          mse = mse + (gr * gr + gi * gi);
        }
      }
    }
  }
  return mse;
}

float emmi_data_misfit(acq_t* acq, emf_t* emf) {
  constexpr double kPi = 3.14159265358979323846264, kMu0 = 4. * kPi * 1e-7;
  const int nc = emf->nchrec, nf = emf->nfreq, nr = acq->nrec;
  memset(&emf->rmse[0][0][0], 0, sizeof(float) * nc * nf * nr);
  float fcost = 0;
  for (int c = 0; c < nc; ++c)
    for (int f = 0; f < nf; ++f) {
      // only receivers beyond 5.5 skin depths in water count
      const float s = static_cast<float>(5.5 * std::sqrt(2. * emf->rho_water / (emf->omegas[f] * kMu0)));
      for (int r = 0; r < nr; ++r) {
        const float s1 = acq->rec_x1[r] - acq->src_x1[0];
        const float s2 = acq->rec_x2[r] - acq->src_x2[0];
        if (!(std::sqrt(static_cast<double>(s1 * s1 + s2 * s2)) > s)) continue;
        const float w = emf->delta_emf[c][f][r];
        const emmi_c64 o = emf->dobs_fd[c][f][r], k = emf->dcal_fd[c][f][r];
        emmi_c64 d{(o.re - k.re) / w, (o.im - k.im) / w};  // W (dobs - Ru)
        emf->rmse[c][f][r] = static_cast<float>(abs_c(d));
        fcost = fcost + (d.re * d.re + d.im * d.im);
        d = {d.re / w, d.im / w};                           // W^T W (dobs - Ru)
        emf->dres_fd[c][f][r] = {d.re, -d.im};
      }
    }
  return static_cast<float>(fcost * 0.5);
}
}
