// host_path.cpp -- the reference's host-side steps around the solve, kept
// bit-exact: grid construction (src/regrid_init_free.c, src/create_nugrid.c,
// src/homogenization.c, src/find_index.c), source injection and receiver /
// volume extraction (src/inject_extract.c) and the gradient product
// (src/build_gradient.c).
//
// These are a bit-exact contract, so the float/double evaluation order of the
// reference is kept operation by operation; complex arithmetic is spelled out
// on (re, im) pairs with the C99 formulas (no library complex type), and the
// file is compiled with -ffp-contract=off.  Math functions are always called
// on double arguments, as C calls them.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "emmi_mgfd.h"
#include "host_path.h"

namespace {

constexpr double kPI = 3.14159265358979323846264;
constexpr double kMU0 = 4. * kPI * 1e-7;

[[noreturn]] void fatal(const char* m) {
  fprintf(stderr, "emmi_mgfd: %s\n", m);
  exit(EXIT_FAILURE);
}

void* xmalloc(size_t n) {
  void* p = malloc(n ? n : 1);
  if (!p) fatal("out of host memory");
  return p;
}

// pointer-table 3D array over one contiguous block, freeable in three calls
float*** alloc3f(size_t n1, size_t n2, size_t n3) {
  float*** p = (float***)xmalloc(n3 * sizeof(float**));
  p[0] = (float**)xmalloc(n3 * n2 * sizeof(float*));
  p[0][0] = (float*)xmalloc(n3 * n2 * n1 * sizeof(float));
  for (size_t k = 0; k < n3; k++) {
    p[k] = p[0] + n2 * k;
    for (size_t j = 0; j < n2; j++) p[k][j] = p[0][0] + n1 * (j + n2 * k);
  }
  return p;
}
void free3f(float*** p) {
  if (!p) return;
  free(p[0][0]);
  free(p[0]);
  free(p);
}

struct C {
  double re, im;
};
inline C cadd(C a, C b) { return {a.re + b.re, a.im + b.im}; }
inline C csub(C a, C b) { return {a.re - b.re, a.im - b.im}; }
inline C cmulr(C a, double s) { return {a.re * s, a.im * s}; }
inline C rmulc(double s, C a) { return {s * a.re, s * a.im}; }
inline C cdivr(C a, double s) { return {a.re / s, a.im / s}; }
inline C cmul(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C ld(const emmi_c128* v, long i) { return {v[i].re, v[i].im}; }
inline void addto(emmi_c128* v, long i, C a) { v[i].re += a.re; v[i].im += a.im; }

// (I * omega) * mu0 as float complex promoted to double complex
inline C i_omega_mu0(float omega) {
  const float re_f = 0.0f * omega, im_f = 1.0f * omega;
  return {(double)re_f * kMU0, (double)im_f * kMU0};
}

}  // namespace

// ---------------------------------------------------------------------------
// grid construction
// ---------------------------------------------------------------------------
extern "C" int find_index(int n, float* x, float val) {
  int lo = 0, hi = n - 1, mid;
  while (lo < hi) {
    mid = (lo + hi) / 2;
    if (x[mid] <= val) lo = mid;
    if (x[mid] > val) hi = mid;
    if (lo == hi || lo == hi - 1) break;
  }
  return lo;
}

namespace emmi {
// the same bisection over an array of length len, refusing to read past it
int find_index_guarded(int n, const float* x, int len, float val) {
  int lo = 0, hi = n - 1, mid;
  while (lo < hi) {
    mid = (lo + hi) / 2;
    if (mid >= len) fatal("find_index would read past the coordinate array (n3 too small for the source depth)");
    if (x[mid] <= val) lo = mid;
    if (x[mid] > val) hi = mid;
    if (lo == hi || lo == hi - 1) break;
  }
  return lo;
}
}  // namespace emmi

extern "C" float create_nugrid(int n, float len, float dx, float* x) {
  float q = 1.1f, qq;
  const float eps = 1e-15f;
  if (n * dx >= len) {
    for (int i = 0; i <= n; i++) x[i] = i * dx;
    return 1;
  }
  for (;;) {
    qq = (float)pow((double)len * ((double)q - 1.) / (double)dx + 1., 1. / (double)n);
    if (fabs((double)(qq - q)) < (double)eps) break;
    q = qq;
  }
  x[0] = 0;
  for (int i = 1; i <= n; i++) x[i] = (float)((pow((double)q, (double)i) - 1.) * (double)dx / ((double)q - 1.));
  return q;
}

static void staggered(int n, const float* x, float* xs, float* d, float* ds) {
  for (int i = 0; i < n; i++) {
    ds[i] = x[i + 1] - x[i];
    xs[i] = (float)(0.5 * (double)(x[i] + x[i + 1]));
  }
  xs[n] = (float)((double)x[n] + 0.5 * (double)ds[n - 1]);
  ds[n] = xs[n] - x[n];
  for (int i = 1; i <= n; i++) d[i] = xs[i] - xs[i - 1];
  d[0] = xs[0] - x[0];
}

// one separable pass of the overlap average along a strided axis
static void overlap_pass(const float* in, float* out, int nin, const float* xin, int nout, const float* xout,
                         long step_in, long step_out) {
  int i = 0;
  for (int j = 0; j < nout; j++) {
    const float a = xout[j], b = xout[j + 1];
    float s = 0.0f;
    while (i < nin && xin[i + 1] <= a) i++;
    for (int k = i; k < nin && xin[k] < b; k++) {
      const float lo = (xin[k] > a) ? xin[k] : a;
      const float hi = (xin[k + 1] < b) ? xin[k + 1] : b;
      const float ov = hi - lo;
      if ((double)ov > 0.0) s += in[k * step_in] * ov;
    }
    out[j * step_out] = s / (b - a);
  }
}

extern "C" void homogenization(float*** in, float*** out, int nx, int ny, int nz, float* xx, float* yy, float* zz,
                               int n1, int n2, int n3, float* x1, float* x2, float* x3) {
  const float* src = &in[0][0][0];
  float* t1 = (float*)xmalloc(sizeof(float) * (size_t)n1 * ny * nz);
  float* t2 = (float*)xmalloc(sizeof(float) * (size_t)n1 * n2 * nz);
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      overlap_pass(src + (long)nx * (iy + (long)ny * iz), t1 + (long)n1 * (iy + (long)ny * iz), nx, xx, n1, x1, 1, 1);
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < n1; ix++)
      overlap_pass(t1 + ix + (long)n1 * ny * iz, t2 + ix + (long)n1 * n2 * iz, ny, yy, n2, x2, n1, n1);
  float* dst = &out[0][0][0];
  for (int iy = 0; iy < n2; iy++)
    for (int ix = 0; ix < n1; ix++)
      overlap_pass(t2 + ix + (long)n1 * iy, dst + ix + (long)n1 * iy, nz, zz, n3, x3, (long)n1 * n2, (long)n1 * n2);
  free(t1);
  free(t2);
}

extern "C" void regrid_init(acq_t* acq, emf_t* emf) {
  const int nb = emf->nb;
  const int nxp = emf->nx + 2 * nb, nyp = emf->ny + 2 * nb, nzp = emf->nz + 2 * nb;
  emf->xx = (float*)xmalloc(sizeof(float) * (nxp + 1));
  emf->yy = (float*)xmalloc(sizeof(float) * (nyp + 1));
  emf->zz = (float*)xmalloc(sizeof(float) * (nzp + 1));
  for (int i = 0; i <= emf->nx; i++) emf->xx[i + nb] = emf->ox + i * emf->dx;
  for (int i = 0; i <= emf->ny; i++) emf->yy[i + nb] = emf->oy + i * emf->dy;
  for (int i = 0; i <= emf->nz; i++) emf->zz[i + nb] = emf->oz + i * emf->dz;
  {
    float* x = (float*)xmalloc(sizeof(float) * (nb + 1));
    create_nugrid(nb, emf->lextend, emf->dx, x);
    for (int i = 0; i <= nb; i++) {
      emf->xx[i] = emf->ox - x[nb - i];
      emf->xx[i + emf->nx + nb] = emf->ox + emf->nx * emf->dx + x[i];
    }
    create_nugrid(nb, emf->lextend, emf->dy, x);
    for (int i = 0; i <= nb; i++) {
      emf->yy[i] = emf->oy - x[nb - i];
      emf->yy[i + emf->ny + nb] = emf->oy + emf->ny * emf->dy + x[i];
    }
    create_nugrid(nb, emf->lextend, emf->dz, x);
    for (int i = 0; i <= nb; i++) {
      emf->zz[i] = emf->oz - x[nb - i];
      emf->zz[i + emf->nz + nb] = emf->oz + emf->nz * emf->dz + x[i];
    }
    free(x);
  }
  emf->x1min = emf->xx[0];
  emf->x1max = emf->xx[nxp];
  emf->x2min = emf->yy[0];
  emf->x2max = emf->yy[nyp];
  emf->x3min = emf->zz[0];
  emf->x3max = emf->zz[nzp];

  // padded model: conductivities for the horizontal terms, resistivity for
  // the vertical one
  float*** s11 = alloc3f(nxp, nyp, nzp);
  float*** s22 = alloc3f(nxp, nyp, nzp);
  float*** r33 = alloc3f(nxp, nyp, nzp);
  for (int k = 0; k < emf->nz; k++)
    for (int j = 0; j < emf->ny; j++)
      for (int i = 0; i < emf->nx; i++) {
        s11[k + nb][j + nb][i + nb] = (float)(1. / (double)emf->rho11[k][j][i]);
        s22[k + nb][j + nb][i + nb] = (float)(1. / (double)emf->rho22[k][j][i]);
        r33[k + nb][j + nb][i + nb] = emf->rho33[k][j][i];
      }
  for (int k = 0; k < nzp; k++)
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < nb; i++) {
        s11[k][j][i] = s11[k][j][nb];
        s11[k][j][nxp - 1 - i] = s11[k][j][nxp - 1 - nb];
        s22[k][j][i] = s22[k][j][nb];
        s22[k][j][nxp - 1 - i] = s22[k][j][nxp - 1 - nb];
        r33[k][j][i] = r33[k][j][nb];
        r33[k][j][nxp - 1 - i] = r33[k][j][nxp - 1 - nb];
      }
  for (int k = 0; k < nzp; k++)
    for (int j = 0; j < nb; j++)
      for (int i = 0; i < nxp; i++) {
        s11[k][j][i] = s11[k][nb][i];
        s11[k][nyp - 1 - j][i] = s11[k][nyp - 1 - nb][i];
        s22[k][j][i] = s22[k][nb][i];
        s22[k][nyp - 1 - j][i] = s22[k][nyp - 1 - nb][i];
        r33[k][j][i] = r33[k][nb][i];
        r33[k][nyp - 1 - j][i] = r33[k][nyp - 1 - nb][i];
      }
  for (int k = 0; k < nb; k++)
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < nxp; i++) {
        s11[k][j][i] = (float)(1. / (double)emf->rho_air);
        s11[nzp - 1 - k][j][i] = s11[nzp - 1 - nb][j][i];
        s22[k][j][i] = (float)(1. / (double)emf->rho_air);
        s22[nzp - 1 - k][j][i] = s22[nzp - 1 - nb][j][i];
        r33[k][j][i] = emf->rho_air;
        r33[nzp - 1 - k][j][i] = r33[nzp - 1 - nb][j][i];
      }

  const int n1 = emf->n1, n2 = emf->n2, n3 = emf->n3;
  float** arr[12] = {&emf->x1, &emf->x2, &emf->x3, &emf->x1s, &emf->x2s, &emf->x3s,
                     &emf->d1, &emf->d2, &emf->d3, &emf->d1s, &emf->d2s, &emf->d3s};
  const int lens[3] = {n1 + 1, n2 + 1, n3 + 1};
  for (int q = 0; q < 12; q++) *arr[q] = (float*)xmalloc(sizeof(float) * lens[q % 3]);

  // x and y: geometric progressions from the source outwards
  float* x = (float*)xmalloc(sizeof(float) * (n1 + n2 + n3 + 3));
  {
    const int np1 = n1 / 2, np2 = n1 - np1;
    float dist = acq->src_x1[0] - emf->x1min;
    create_nugrid(np1, dist, emf->d1min, x);
    for (int i = 0; i <= np1; i++) emf->x1[np1 - i] = acq->src_x1[0] - x[i];
    dist = emf->x1max - acq->src_x1[0];
    create_nugrid(np2, dist, emf->d1min, x);
    for (int i = 0; i <= np2; i++) emf->x1[np1 + i] = acq->src_x1[0] + x[i];
  }
  {
    const int np1 = n2 / 2, np2 = n2 - np1;
    float dist = acq->src_x2[0] - emf->x2min;
    create_nugrid(np1, dist, emf->d2min, x);
    for (int i = 0; i <= np1; i++) emf->x2[np1 - i] = acq->src_x2[0] - x[i];
    dist = emf->x2max - acq->src_x2[0];
    create_nugrid(np2, dist, emf->d2min, x);
    for (int i = 0; i <= np2; i++) emf->x2[np1 + i] = acq->src_x2[0] + x[i];
  }
  {
    // input nodes down to five cells below the source cell; the reference
    // bisects zz with the count nx+2nb (src/regrid_init_free.c:194)
    const int np1 = emmi::find_index_guarded(emf->nx + 2 * nb, emf->zz, nzp + 1, acq->src_x3[0]) + 5;
    const int np2 = n3 - np1;
    if (np1 > nzp || np2 < 1) fatal("n3 too small for the source depth");
    for (int i = 0; i <= np1; i++) emf->x3[i] = emf->zz[i];
    const float dist = emf->x3max - emf->x3[np1];
    create_nugrid(np2, dist, emf->d3min, x);
    for (int i = 0; i <= np2; i++) emf->x3[np1 + i] = emf->x3[np1] + x[i];
  }
  free(x);
  staggered(n1, emf->x1, emf->x1s, emf->d1, emf->d1s);
  staggered(n2, emf->x2, emf->x2s, emf->d2, emf->d2s);
  staggered(n3, emf->x3, emf->x3s, emf->d3, emf->d3s);

  emf->sigma11 = alloc3f(n1, n2, n3);
  emf->sigma22 = alloc3f(n1, n2, n3);
  emf->sigma33 = alloc3f(n1, n2, n3);
  emf->invmur = alloc3f(n1, n2, n3);
  homogenization(s11, emf->sigma11, nxp, nyp, nzp, emf->xx, emf->yy, emf->zz, n1, n2, n3, emf->x1, emf->x2, emf->x3);
  homogenization(s22, emf->sigma22, nxp, nyp, nzp, emf->xx, emf->yy, emf->zz, n1, n2, n3, emf->x1, emf->x2, emf->x3);
  homogenization(r33, emf->sigma33, nxp, nyp, nzp, emf->xx, emf->yy, emf->zz, n1, n2, n3, emf->x1, emf->x2, emf->x3);
  float* s33 = &emf->sigma33[0][0][0];
  float* im = &emf->invmur[0][0][0];
  for (long c = 0; c < (long)n1 * n2 * n3; c++) {
    s33[c] = (float)(1. / (double)s33[c]);
    im[c] = 1.f;
  }
  free3f(s11);
  free3f(s22);
  free3f(r33);
}

extern "C" void regrid_free(emf_t* emf) {
  free(emf->xx); free(emf->yy); free(emf->zz);
  free(emf->x1); free(emf->x2); free(emf->x3);
  free(emf->x1s); free(emf->x2s); free(emf->x3s);
  free(emf->d1); free(emf->d2); free(emf->d3);
  free(emf->d1s); free(emf->d2s); free(emf->d3s);
  free3f(emf->sigma11); free3f(emf->sigma22); free3f(emf->sigma33); free3f(emf->invmur);
  emf->xx = emf->yy = emf->zz = nullptr;
  emf->sigma11 = emf->sigma22 = emf->sigma33 = emf->invmur = nullptr;
}

// ---------------------------------------------------------------------------
// sources and receivers
// ---------------------------------------------------------------------------
namespace emmi {

int channel_code(const char* ch) {
  static const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
  for (int i = 0; i < 6; i++)
    if (strcmp(ch, names[i]) == 0) return i;
  return -1;
}

// staggering per pattern Ex Ey Ez Hx Hy Hz: 1 = staggered nodes on that axis
static const int kStag[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

Tri locate(const emf_t* emf, int pat, float px, float py, float pz) {
  Tri t;
  const float* gx = kStag[pat][0] ? emf->x1s : emf->x1;
  const float* gy = kStag[pat][1] ? emf->x2s : emf->x2;
  const float* gz = kStag[pat][2] ? emf->x3s : emf->x3;
  t.i = find_index(emf->n1 + 1, (float*)gx, px);
  t.j = find_index(emf->n2 + 1, (float*)gy, py);
  t.k = find_index(emf->n3 + 1, (float*)gz, pz);
  t.w1 = (px - gx[t.i]) / (gx[t.i + 1] - gx[t.i]);
  t.w2 = (py - gy[t.j]) / (gy[t.j + 1] - gy[t.j]);
  t.w3 = (pz - gz[t.k]) / (gz[t.k + 1] - gz[t.k]);
  t.vol = (gx[t.i + 1] - gx[t.i]) * (gy[t.j + 1] - gy[t.j]) * (gz[t.k + 1] - gz[t.k]);
  return t;
}

}  // namespace emmi

using emmi::Tri;

namespace {
// destinations and sources of the trilinear stencils: a dense host vector, or
// a sparse list (device-resident vectors are reached through such lists)
struct DenseSink {
  emmi_c128* v;
  void add(long i, C a) { addto(v, i, a); }
};
struct DenseSource {
  const emmi_c128* v;
  C get(long i) const { return ld(v, i); }
};

// v[corner] += val * f1 * f2 * f3 in the reference's order
template <typename Sink>
void scatter8(Sink& v, long comp_off, const Tri& t, int n1, int n2, C val) {
  const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1);
  const long o = comp_off + t.i + s1 * t.j + s2 * t.k;
  const double w1 = t.w1, w2 = t.w2, w3 = t.w3;
  v.add(o, cmulr(cmulr(cmulr(val, 1. - w1), 1. - w2), 1. - w3));
  v.add(o + 1, cmulr(cmulr(cmulr(val, w1), 1. - w2), 1. - w3));
  v.add(o + s1, cmulr(cmulr(cmulr(val, 1. - w1), w2), 1. - w3));
  v.add(o + 1 + s1, cmulr(cmulr(cmulr(val, w1), w2), 1. - w3));
  v.add(o + s2, cmulr(cmulr(cmulr(val, 1. - w1), 1. - w2), w3));
  v.add(o + 1 + s2, cmulr(cmulr(cmulr(val, w1), 1. - w2), w3));
  v.add(o + s1 + s2, cmulr(cmulr(cmulr(val, 1. - w1), w2), w3));
  v.add(o + 1 + s1 + s2, cmulr(cmulr(cmulr(val, w1), w2), w3));
}

template <typename Source>
C gather8(const Source& v, long comp_off, int i, int j, int k, float w1f, float w2f, float w3f, int n1, int n2) {
  const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1);
  const long o = comp_off + i + s1 * j + s2 * k;
  const double w1 = w1f, w2 = w2f, w3 = w3f;
  C s = {0, 0};
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o), 1. - w1), 1. - w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o + 1), w1), 1. - w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o + s1), 1. - w1), w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o + 1 + s1), w1), w2), 1. - w3));
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o + s2), 1. - w1), 1. - w2), w3));
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o + 1 + s2), w1), 1. - w2), w3));
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o + s1 + s2), 1. - w1), w2), w3));
  s = cadd(s, cmulr(cmulr(cmulr(v.get(o + 1 + s1 + s2), w1), w2), w3));
  return s;
}

// dipole moment of the active channels, rotated by azimuth and dip
// (src/inject_extract.c:42-55); `first` is 0 for E channels, 3 for H
void moment(int nch, char** ch, int first, const C* amp, float phi, float theta, C* F) {
  F[0] = F[1] = F[2] = C{0, 0};
  const double cp = cos((double)phi), sp = sin((double)phi), ct = cos((double)theta), st = sin((double)theta);
  for (int ic = 0; ic < nch; ic++) {
    const int code = emmi::channel_code(ch[ic]);
    if (code == first + 0) {
      F[0] = cadd(F[0], cmulr(amp[ic], cp));
      F[1] = cadd(F[1], cmulr(amp[ic], sp));
    } else if (code == first + 1) {
      F[0] = csub(F[0], cmulr(cmulr(amp[ic], sp), ct));
      F[1] = cadd(F[1], cmulr(cmulr(amp[ic], cp), ct));
      F[2] = cadd(F[2], cmulr(amp[ic], st));
    } else if (code == first + 2) {
      F[0] = cadd(F[0], cmulr(cmulr(amp[ic], sp), st));
      F[1] = csub(F[1], cmulr(cmulr(amp[ic], cp), st));
      F[2] = cadd(F[2], cmulr(amp[ic], ct));
    }
  }
}

long n123_of(const emf_t* e) { return (long)(e->n1 + 1) * (e->n2 + 1) * (e->n3 + 1); }

}  // namespace

namespace emmi {
// source term of src/inject_extract.c:26-120 into any sink
template <typename Sink>
void source_terms(acq_t* acq, emf_t* emf, int ifreq, Sink& sink) {
  const long n123 = n123_of(emf);
  C* amp = (C*)xmalloc(sizeof(C) * (emf->nchsrc + 1));
  for (int i = 0; i < emf->nchsrc; i++) amp[i] = C{1, 0};
  const C iwm = i_omega_mu0(emf->omegas[ifreq]);
  for (int is = 0; is < acq->nsrc; is++) {
    C F[3];
    moment(emf->nchsrc, emf->chsrc, 0, amp, acq->src_azimuth[is], acq->src_dip[is], F);
    for (int comp = 0; comp < 3; comp++) {
      const Tri t = emmi::locate(emf, comp, acq->src_x1[is], acq->src_x2[is], acq->src_x3[is]);
      const C sv = cdivr(cmul(iwm, F[comp]), (double)t.vol);
      scatter8(sink, comp * n123, t, emf->n1, emf->n2, sv);
    }
  }
  free(amp);
}

struct SparseSink {
  std::map<long, C> m;
  void add(long i, C a) {
    auto it = m.find(i);
    if (it == m.end()) m.emplace(i, C{0.0 + a.re, 0.0 + a.im});
    else it->second = cadd(it->second, a);
  }
};

void source_entries(acq_t* acq, emf_t* emf, int ifreq, std::vector<long>& idx, std::vector<emmi_c128>& val) {
  SparseSink sk;
  source_terms(acq, emf, ifreq, sk);
  idx.clear();
  val.clear();
  for (auto& kv : sk.m) {
    idx.push_back(kv.first);
    val.push_back(emmi_c128{kv.second.re, kv.second.im});
  }
}
}  // namespace emmi

// src/inject_extract.c:26-120
extern "C" void inject_source(acq_t* acq, emf_t* emf, emmi_c128* b, int ifreq) {
  memset(b, 0, sizeof(emmi_c128) * 3 * n123_of(emf));
  DenseSink sk{b};
  emmi::source_terms(acq, emf, ifreq, sk);
}

namespace emmi {
// the scatters of src/inject_extract.c:133-302: E-channel residuals into b,
// H-channel residuals into the magnetic dipole M (x), in the reference's order
template <typename SinkB, typename SinkX>
void adj_terms(acq_t* acq, emf_t* emf, int ifreq, SinkB& sb, SinkX& sx) {
  const long n123 = n123_of(emf);
  const int n1 = emf->n1, n2 = emf->n2;
  C* amp = (C*)xmalloc(sizeof(C) * (emf->nchrec + 1));
  const C iwm = i_omega_mu0(emf->omegas[ifreq]);
  for (int ir = 0; ir < acq->nrec; ir++) {
    for (int ic = 0; ic < emf->nchrec; ic++) {
      const emmi_c64 d = emf->dres_fd[ic][ifreq][ir];
      amp[ic] = C{(double)d.re, (double)d.im};
    }
    C F[3];
    moment(emf->nchrec, emf->chrec, 0, amp, acq->rec_azimuth[ir], acq->rec_dip[ir], F);
    for (int comp = 0; comp < 3; comp++) {
      const Tri t = emmi::locate(emf, comp, acq->rec_x1[ir], acq->rec_x2[ir], acq->rec_x3[ir]);
      scatter8(sb, comp * n123, t, n1, n2, cdivr(cmul(iwm, F[comp]), (double)t.vol));
    }
    moment(emf->nchrec, emf->chrec, 3, amp, acq->rec_azimuth[ir], acq->rec_dip[ir], F);
    for (int comp = 0; comp < 3; comp++) {
      const Tri t = emmi::locate(emf, 3 + comp, acq->rec_x1[ir], acq->rec_x2[ir], acq->rec_x3[ir]);
      scatter8(sx, comp * n123, t, n1, n2, cdivr(F[comp], (double)t.vol));
    }
  }
  free(amp);
}

void adj_entries(acq_t* acq, emf_t* emf, int ifreq, std::vector<long>& bidx, std::vector<emmi_c128>& bval,
                 std::vector<long>& xidx, std::vector<emmi_c128>& xval) {
  SparseSink sb, sx;
  adj_terms(acq, emf, ifreq, sb, sx);
  bidx.clear(); bval.clear(); xidx.clear(); xval.clear();
  for (auto& kv : sb.m) { bidx.push_back(kv.first); bval.push_back(emmi_c128{kv.second.re, kv.second.im}); }
  for (auto& kv : sx.m) { xidx.push_back(kv.first); xval.push_back(emmi_c128{kv.second.re, kv.second.im}); }
}
}  // namespace emmi

// src/inject_extract.c:124-314
extern "C" void inject_adj_source(acq_t* acq, emf_t* emf, emmi_c128* x, emmi_c128* b, int ifreq) {
  const long n123 = n123_of(emf);
  const int n1 = emf->n1, n2 = emf->n2, n3 = emf->n3;
  memset(b, 0, sizeof(emmi_c128) * 3 * n123);
  memset(x, 0, sizeof(emmi_c128) * 3 * n123);
  DenseSink sb{b}, sx{x};
  emmi::adj_terms(acq, emf, ifreq, sb, sx);
  // b += curl M with the dual widths (mu_r = 1)
  const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1);
  for (int k = 1; k < n3 + 1; k++)
    for (int j = 1; j < n2 + 1; j++)
      for (int i = 1; i < n1 + 1; i++) {
        const long o = i + s1 * j + s2 * k;
        const double d1 = emf->d1[i], d2 = emf->d2[j], d3 = emf->d3[k];
        const C mx = ld(x, o), my = ld(x, n123 + o), mz = ld(x, 2 * n123 + o);
        addto(b, o, csub(cdivr(csub(mz, ld(x, 2 * n123 + o - s1)), d2), cdivr(csub(my, ld(x, n123 + o - s2)), d3)));
        addto(b, n123 + o, csub(cdivr(csub(mx, ld(x, o - s2)), d3), cdivr(csub(mz, ld(x, 2 * n123 + o - 1)), d1)));
        addto(b, 2 * n123 + o, csub(cdivr(csub(my, ld(x, n123 + o - 1)), d1), cdivr(csub(mx, ld(x, o - s1)), d2)));
      }
}

namespace {
emmi_c64 rotate(int which, float phi, float theta, C Fx, C Fy, C Fz) {
  const double cp = cos((double)phi), sp = sin((double)phi), ct = cos((double)theta), st = sin((double)theta);
  C r;
  if (which == 0) {
    r = cadd(rmulc(cp, Fx), rmulc(sp, Fy));
  } else if (which == 1) {
    r = csub(cadd(rmulc(-sp * ct, Fx), rmulc(cp * ct, Fy)), rmulc(st, Fz));
  } else {
    r = cadd(csub(rmulc(sp * st, Fx), rmulc(cp * st, Fy)), rmulc(ct, Fz));
  }
  return emmi_c64{(float)r.re, (float)r.im};
}
}  // namespace

namespace emmi {
// receiver data of src/inject_extract.c:317-474 from any source of E (x) and
// H (b) values
template <typename SrcE, typename SrcH>
void receiver_data(acq_t* acq, emf_t* emf, int ifreq, const SrcE& x, const SrcH& b) {
  const long n123 = n123_of(emf);
  const int n1 = emf->n1, n2 = emf->n2;
  const float* s33 = &emf->sigma33[0][0][0];
  for (int ir = 0; ir < acq->nrec; ir++) {
    const float phi = acq->rec_azimuth[ir], theta = acq->rec_dip[ir];
    const float px = acq->rec_x1[ir], py = acq->rec_x2[ir], pz = acq->rec_x3[ir];
    Tri t = emmi::locate(emf, 0, px, py, pz);
    const C Fx = gather8(x, 0, t.i, t.j, t.k, t.w1, t.w2, t.w3, n1, n2);
    t = emmi::locate(emf, 1, px, py, pz);
    const C Fy = gather8(x, n123, t.i, t.j, t.k, t.w1, t.w2, t.w3, n1, n2);
    t = emmi::locate(emf, 2, px, py, pz);
    C Fz = {0, 0};
    {
      // sigma33 * Ez interpolated with the unaveraged cell conductivity
      const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1);
      const double w1 = t.w1, w2 = t.w2, w3 = t.w3;
      const int i = t.i, j = t.j, k = t.k;
#define S33(ii, jj, kk) ((double)s33[(ii) + (long)n1 * ((jj) + (long)n2 * (kk))])
#define XZ(ii, jj, kk) x.get(2 * n123 + (ii) + s1 * (jj) + s2 * (kk))
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i, j, k), XZ(i, j, k)), 1. - w1), 1. - w2), 1. - w3));
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i + 1, j, k), XZ(i + 1, j, k)), w1), 1. - w2), 1. - w3));
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i, j + 1, k), XZ(i, j + 1, k)), 1. - w1), w2), 1. - w3));
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i + 1, j + 1, k), XZ(i + 1, j + 1, k)), w1), w2), 1. - w3));
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i, j, k + 1), XZ(i, j, k + 1)), 1. - w1), 1. - w2), w3));
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i + 1, j, k + 1), XZ(i + 1, j, k + 1)), w1), 1. - w2), w3));
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i, j + 1, k + 1), XZ(i, j + 1, k + 1)), 1. - w1), w2), w3));
      Fz = cadd(Fz, cmulr(cmulr(cmulr(rmulc(S33(i + 1, j + 1, k + 1), XZ(i + 1, j + 1, k + 1)), w1), w2), w3));
#undef S33
#undef XZ
      Fz = cmulr(Fz, (double)emf->rho_water);
    }
    for (int ic = 0; ic < emf->nchrec; ic++) {
      const int code = emmi::channel_code(emf->chrec[ic]);
      if (code >= 0 && code <= 2) emf->dcal_fd[ic][ifreq][ir] = rotate(code, phi, theta, Fx, Fy, Fz);
    }
    t = emmi::locate(emf, 3, px, py, pz);
    const C Hx = gather8(b, 0, t.i, t.j, t.k, t.w1, t.w2, t.w3, n1, n2);
    t = emmi::locate(emf, 4, px, py, pz);
    const C Hy = gather8(b, n123, t.i, t.j, t.k, t.w1, t.w2, t.w3, n1, n2);
    t = emmi::locate(emf, 5, px, py, pz);
    const C Hz = gather8(b, 2 * n123, t.i, t.j, t.k, t.w1, t.w2, t.w3, n1, n2);
    for (int ic = 0; ic < emf->nchrec; ic++) {
      const int code = emmi::channel_code(emf->chrec[ic]);
      if (code >= 3) emf->dcal_fd[ic][ifreq][ir] = rotate(code - 3, phi, theta, Hx, Hy, Hz);
    }
  }
}

// records the entries a receiver extraction reads
struct RecordingSource {
  std::vector<long>* idx;
  C get(long i) const { idx->push_back(i); return C{0, 0}; }
};
// serves previously gathered entries in the order they were recorded
struct ReplaySource {
  const std::vector<emmi_c128>* val;
  mutable size_t pos = 0;
  C get(long) const { const emmi_c128& v = (*val)[pos++]; return C{v.re, v.im}; }
};

void receiver_indices(acq_t* acq, emf_t* emf, std::vector<long>& ix, std::vector<long>& ib) {
  ix.clear();
  ib.clear();
  emmi_c64*** keep = emf->dcal_fd;
  // run against a scratch output table so the caller's data is untouched
  const int nch = emf->nchrec, nf = emf->nfreq, nr = acq->nrec;
  std::vector<emmi_c64> buf((size_t)nch * nf * (nr > 0 ? nr : 1));
  std::vector<emmi_c64*> p1((size_t)nch * nf);
  std::vector<emmi_c64**> p2(nch);
  for (int c = 0; c < nch; c++) {
    p2[c] = &p1[(size_t)c * nf];
    for (int f = 0; f < nf; f++) p1[(size_t)c * nf + f] = &buf[((size_t)c * nf + f) * nr];
  }
  emf->dcal_fd = p2.data();
  RecordingSource rx{&ix}, rb{&ib};
  receiver_data(acq, emf, 0, rx, rb);
  emf->dcal_fd = keep;
}

void receiver_data_from(acq_t* acq, emf_t* emf, int ifreq, const std::vector<emmi_c128>& vx,
                        const std::vector<emmi_c128>& vb) {
  ReplaySource sx{&vx}, sb{&vb};
  receiver_data(acq, emf, ifreq, sx, sb);
}
}  // namespace emmi

// src/inject_extract.c:317-474
extern "C" void extract_emf_data(acq_t* acq, emf_t* emf, emmi_c128* x, emmi_c128* b, int ifreq) {
  DenseSource sx{x}, sb{b};
  emmi::receiver_data(acq, emf, ifreq, sx, sb);
}

namespace emmi {
// per-axis tables of cell-centre interpolation (src/inject_extract.c:504-537)
void centre_table(int n, float o, float d, int ng, const float* xg, const float* xgs, int* in, float* wn, int* is,
                  float* ws) {
  for (int i = 0; i < n; i++) {
    const float c = (float)((double)o + ((double)i + 0.5) * (double)d);
    int ii = find_index(ng + 1, (float*)xg, c);
    wn[i] = (c - xg[ii]) / (xg[ii + 1] - xg[ii]);
    in[i] = ii;
    ii = find_index(ng + 1, (float*)xgs, c);
    ws[i] = (c - xgs[ii]) / (xgs[ii + 1] - xgs[ii]);
    is[i] = ii;
  }
}

CentreTables::CentreTables(const emf_t* emf) {
  nx = emf->nx; ny = emf->ny; nz = emf->nz;
  i1.resize(nx); i1s.resize(nx); w1.resize(nx); w1s.resize(nx);
  i2.resize(ny); i2s.resize(ny); w2.resize(ny); w2s.resize(ny);
  i3.resize(nz); i3s.resize(nz); w3.resize(nz); w3s.resize(nz);
  centre_table(nx, emf->ox, emf->dx, emf->n1, emf->x1, emf->x1s, i1.data(), w1.data(), i1s.data(), w1s.data());
  centre_table(ny, emf->oy, emf->dy, emf->n2, emf->x2, emf->x2s, i2.data(), w2.data(), i2s.data(), w2s.data());
  centre_table(nz, emf->oz, emf->dz, emf->n3, emf->x3, emf->x3s, i3.data(), w3.data(), i3s.data(), w3s.data());
}
}  // namespace emmi

// src/inject_extract.c:477-635
extern "C" void extract_electric_field(acq_t* acq, emf_t* emf, emmi_c128* x, int ifreq) {
  (void)acq;
  const emmi::CentreTables T(emf);
  const long n123 = n123_of(emf);
  const int n1 = emf->n1, n2 = emf->n2;
  const long s1 = n1 + 1, s2 = (long)(n1 + 1) * (n2 + 1);
  emmi_c64* E1 = &emf->E1[ifreq][0][0][0];
  emmi_c64* E2 = &emf->E2[ifreq][0][0][0];
  emmi_c64* E3 = &emf->E3[ifreq][0][0][0];
  const float* s33 = &emf->sigma33[0][0][0];
  const float* r33 = &emf->rho33[0][0][0];
  for (int k = 0; k < T.nz; k++)
    for (int j = 0; j < T.ny; j++)
      for (int i = 0; i < T.nx; i++) {
        const long o = i + (long)T.nx * (j + (long)T.ny * k);
        const DenseSource sx{x};
        C v = gather8(sx, 0, T.i1s[i], T.i2[j], T.i3[k], T.w1s[i], T.w2[j], T.w3[k], n1, n2);
        E1[o] = emmi_c64{(float)v.re, (float)v.im};
        v = gather8(sx, n123, T.i1[i], T.i2s[j], T.i3[k], T.w1[i], T.w2s[j], T.w3[k], n1, n2);
        E2[o] = emmi_c64{(float)v.re, (float)v.im};
      }
#define S33(a, b, c) s33[(a) + (long)n1 * ((b) + (long)n2 * (c))]
  for (int k = 0; k < T.nz; k++) {
    const int kk = T.i3s[k];
    const double wz = T.w3s[k];
    for (int j = 0; j < T.ny; j++) {
      const int jj = T.i2[j];
      const double wy = T.w2[j];
      for (int i = 0; i < T.nx; i++) {
        const int ii = T.i1[i];
        const double wx = T.w1[i];
        C s = {0, 0};
        auto term = [&](int a, int b2, int c, double f1, double f2, double f3) {
          // vertical current: four-cell mean of sigma33 around the Ez edge
          const float sg = (float)(0.25 * (double)(S33(a, b2, c) + S33(a, b2 - 1, c) + S33(a - 1, b2, c) + S33(a - 1, b2 - 1, c)));
          const C xv = ld(x, 2 * n123 + a + s1 * b2 + s2 * c);
          s = cadd(s, cmulr(cmulr(cmulr(rmulc((double)sg, xv), f1), f2), f3));
        };
        term(ii, jj, kk, 1. - wx, 1. - wy, 1. - wz);
        term(ii + 1, jj, kk, wx, 1. - wy, 1. - wz);
        term(ii, jj + 1, kk, 1. - wx, wy, 1. - wz);
        term(ii + 1, jj + 1, kk, wx, wy, 1. - wz);
        term(ii, jj, kk + 1, 1. - wx, 1. - wy, wz);
        term(ii + 1, jj, kk + 1, wx, 1. - wy, wz);
        term(ii, jj + 1, kk + 1, 1. - wx, wy, wz);
        term(ii + 1, jj + 1, kk + 1, wx, wy, wz);
        const long o = i + (long)T.nx * (j + (long)T.ny * k);
        const C e = cmulr(s, (double)r33[o]);
        E3[o] = emmi_c64{(float)e.re, (float)e.im};
      }
    }
  }
#undef S33
}

// src/inject_extract.c:638-785
extern "C" void extract_magnetic_field(acq_t* acq, emf_t* emf, emmi_c128* b, int ifreq) {
  (void)acq;
  const emmi::CentreTables T(emf);
  const long n123 = n123_of(emf);
  const int n1 = emf->n1, n2 = emf->n2;
  emmi_c64* H1 = &emf->H1[ifreq][0][0][0];
  emmi_c64* H2 = &emf->H2[ifreq][0][0][0];
  emmi_c64* H3 = &emf->H3[ifreq][0][0][0];
  for (int k = 0; k < T.nz; k++)
    for (int j = 0; j < T.ny; j++)
      for (int i = 0; i < T.nx; i++) {
        const long o = i + (long)T.nx * (j + (long)T.ny * k);
        const DenseSource sb{b};
        C v = gather8(sb, 0, T.i1[i], T.i2s[j], T.i3s[k], T.w1[i], T.w2s[j], T.w3s[k], n1, n2);
        H1[o] = emmi_c64{(float)v.re, (float)v.im};
        v = gather8(sb, n123, T.i1s[i], T.i2[j], T.i3s[k], T.w1s[i], T.w2[j], T.w3s[k], n1, n2);
        H2[o] = emmi_c64{(float)v.re, (float)v.im};
        v = gather8(sb, 2 * n123, T.i1s[i], T.i2s[j], T.i3[k], T.w1s[i], T.w2s[j], T.w3[k], n1, n2);
        H3[o] = emmi_c64{(float)v.re, (float)v.im};
      }
}

// ---------------------------------------------------------------------------
// gradient (src/build_gradient.c:13-111)
// ---------------------------------------------------------------------------
namespace {
inline float creal_mul_f(emmi_c64 a, emmi_c64 b) {
  // real part of a float complex product evaluated in float
  const float re = a.re * b.re - a.im * b.im;
  return re;
}
inline double cabs_d(emmi_c64 a) { return hypot((double)a.re, (double)a.im); }

void mute_box(const emf_t* emf, const fwi_t* fwi, float px, float py, float pz, float r) {
  int b[6];
  emmi::mute_bounds(emf, px, py, pz, r, b);
  for (int k = 0; k < fwi->npar; k++)
    for (int i3 = b[4]; i3 <= b[5]; i3++)
      for (int i2 = b[2]; i2 <= b[3]; i2++)
        for (int i1 = b[0]; i1 <= b[1]; i1++) fwi->grad[k][i3][i2][i1] = 0.f;
}
}  // namespace

namespace emmi {
// cells within r of (px, py, pz), clipped to the model (src/build_gradient.c:64-111)
void mute_bounds(const emf_t* emf, float px, float py, float pz, float r, int* b) {
  int i1min = (int)((px - r - emf->ox) / emf->dx), i1max = (int)((px + r - emf->ox) / emf->dx);
  int i2min = (int)((py - r - emf->oy) / emf->dy), i2max = (int)((py + r - emf->oy) / emf->dy);
  int i3min = (int)((pz - r - emf->oz) / emf->dz), i3max = (int)((pz + r - emf->oz) / emf->dz);
  auto clip = [](int v, int n) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); };
  b[0] = clip(i1min, emf->nx); b[1] = clip(i1max, emf->nx);
  b[2] = clip(i2min, emf->ny); b[3] = clip(i2max, emf->ny);
  b[4] = clip(i3min, emf->nz); b[5] = clip(i3max, emf->nz);
}
}  // namespace emmi

extern "C" void build_gradient(acq_t* acq, emf_t* emf, fwi_t* fwi, emmi_c64**** E1f, emmi_c64**** E2f,
                               emmi_c64**** E3f, emmi_c64**** E1a, emmi_c64**** E2a, emmi_c64**** E3a) {
  const float s = emf->dx * emf->dy * emf->dz;
  for (int k = 0; k < fwi->npar; k++)
    for (int i3 = 0; i3 < emf->nz; i3++)
      for (int i2 = 0; i2 < emf->ny; i2++)
        for (int i1 = 0; i1 < emf->nx; i1++) {
          float& g = fwi->grad[k][i3][i2][i1];
          g = 0;
          for (int f = 0; f < emf->nfreq; f++) {
            if (fwi->idxpar[k] == 1) {
              g = (float)((double)g + (double)creal_mul_f(E1f[f][i3][i2][i1], E1a[f][i3][i2][i1]));
              g = (float)((double)g + (double)creal_mul_f(E2f[f][i3][i2][i1], E2a[f][i3][i2][i1]));
            }
            if (fwi->idxpar[k] == 2) g = (float)((double)g + (double)creal_mul_f(E3f[f][i3][i2][i1], E3a[f][i3][i2][i1]));
          }
          if (fwi->idxpar[k] == 1) g *= s / emf->rho11[i3][i2][i1];
          if (fwi->idxpar[k] == 2) g *= s / emf->rho33[i3][i2][i1];
          if ((double)emf->oz + ((double)i3 + 0.5) * (double)emf->dz <= (double)(emf->bathy[i2][i1] + emf->dz)) g = 0.f;
          if (fwi->preco == 1) {
            float& h = fwi->phess[k][i3][i2][i1];
            h = 0.f;
            for (int f = 0; f < emf->nfreq; f++) {
              if (fwi->idxpar[k] == 1) {
                h = (float)((double)h + cabs_d(E1f[f][i3][i2][i1]) * cabs_d(E1f[f][i3][i2][i1]));
                h = (float)((double)h + cabs_d(E2f[f][i3][i2][i1]) * cabs_d(E2f[f][i3][i2][i1]));
              }
              if (fwi->idxpar[k] == 2) h = (float)((double)h + cabs_d(E3f[f][i3][i2][i1]) * cabs_d(E3f[f][i3][i2][i1]));
            }
            if (fwi->idxpar[k] == 1) h *= s / emf->rho11[i3][i2][i1];
            if (fwi->idxpar[k] == 2) h *= s / emf->rho33[i3][i2][i1];
          }
        }
  for (int is = 0; is < acq->nsrc; is++) mute_box(emf, fwi, acq->src_x1[is], acq->src_x2[is], acq->src_x3[is], emf->dsmute);
  for (int ir = 0; ir < acq->nrec; ir++) mute_box(emf, fwi, acq->rec_x1[ir], acq->rec_x2[ir], acq->rec_x3[ir], emf->drmute);
}
