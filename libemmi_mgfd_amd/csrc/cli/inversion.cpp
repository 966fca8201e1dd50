// inversion.cpp -- the model side of the inversion: Tikhonov and total-
// variation regularisation (src/regularization.c:11-130, src/fg_mod_reg.c:
// 19-94), the three gradient preconditioners (src/precondition.c:16-87),
// triangle smoothing (src/triangle_smoothing.c:9-107) and the l-BFGS method
// with the bisection Wolfe line search (src/opt.c:22-257).
//
// Results feed back into the next misfit evaluation, so every formula keeps
// the reference's float/double evaluation and its sequential float sums.
#include "inversion.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "arrays.h"

using namespace emmi_cli;

namespace {
constexpr double kPi = 3.14159265358979323846264, kMu0 = 4. * kPi * 1e-7;
// weights of horizontal and vertical roughness (src/regularization.c:19-20)
constexpr float kCh = 1, kCv = 0.03f;

// second difference -(a - 2 b + c), formed in double as the reference does
inline float second_diff(float a, float b, float c) { return static_cast<float>(-(a - 2.0 * b + c)); }

// a box filter of length nbox as causal running sums: y = x * box / nbox,
// ny = nx + nbox - 1 outputs (src/triangle_smoothing.c:33-46)
void box_conv(int nbox, int nx, const float* x, float* y, float* run) {
  const int ny = nx + nbox - 1;
  run[0] = x[0];
  for (int i = 1; i < nx; ++i) run[i] = run[i - 1] + x[i];
  for (int i = nx; i < ny; ++i) run[i] = run[i - 1];
  for (int i = 0; i < nbox; ++i) y[i] = run[i];
  for (int i = nbox; i < ny; ++i) y[i] = run[i] - run[i - nbox];
  for (int i = 0; i < ny; ++i) y[i] /= nbox;
}

// a triangle of half-width nbox applied to x[0..nd) in place: two boxes,
// the tails folded back at both ends (src/triangle_smoothing.c:48-60)
struct Triangle {
  int nbox, nd, np, nq;
  std::vector<float> p, q, b1, b2, y;
  Triangle(int nb, int n) : nbox(nb), nd(n), np(nb + n - 1), nq(nb + (nb + n - 1) - 1) {
    p.resize(np);
    q.resize(nq);
    b1.resize(np);
    b2.resize(nq);
    y.resize(nd);
  }
  void apply(float* x) {
    box_conv(nbox, nd, x, p.data(), b1.data());
    box_conv(nbox, np, p.data(), q.data(), b2.data());
    for (int i = 0; i < nd; ++i) y[i] = q[i + nbox - 1];
    for (int i = 0; i < nbox - 1; ++i) y[i] += q[nbox - 2 - i];
    for (int i = 0; i < nbox - 1; ++i) y[nd - i - 1] += q[nd + (nbox - 1) + i];
    memcpy(x, y.data(), sizeof(float) * nd);
  }
};

void write_floats(const char* name, const float* p, size_t n) {
  FILE* fp = fopen(name, "wb");
  if (!fp) emmi_cli_err("cannot open %s for writing", name);
  fwrite(p, sizeof(float) * n, 1, fp);
  fclose(fp);
}
}  // namespace

extern "C" {

float regularization_tikhonov(float* x, float* g, int n1, int n2, int n3, float, float, float) {
  // 0.5 |D x|^2 with first differences, horizontal weight kCh and vertical
  // kCv, and its gradient D^T D x as second differences; unit spacings
  memset(g, 0, sizeof(float) * n1 * n2 * n3);
  float s1 = 0, s2 = 0, s3 = 0;
  const int s[3] = {1, n1, n1 * n2};
  for (int i3 = 1; i3 < n3 - 1; ++i3)
    for (int i2 = 1; i2 < n2 - 1; ++i2)
      for (int i1 = 1; i1 < n1 - 1; ++i1) {
        const int k = i1 + n1 * (i2 + n2 * i3);
        const float t1 = second_diff(x[k - s[0]], x[k], x[k + s[0]]);
        float d = x[k + s[0]] - x[k];
        s1 += d * d;
        const float t2 = second_diff(x[k - s[1]], x[k], x[k + s[1]]);
        d = x[k + s[1]] - x[k];
        s2 += d * d;
        const float t3 = second_diff(x[k - s[2]], x[k], x[k + s[2]]);
        d = x[k + s[2]] - x[k];
        s3 += d * d;
        g[k] = kCh * (t1 + t2) + kCv * t3;
      }
  return kCh * (s1 + s2) + kCv * s3;
}

float regularization_tv(float* x, float* g, int n1, int n2, int n3, float, float, float) {
  // sum sqrt(|grad x|^2 + eps) with forward differences, and its gradient
  // -div(grad x / |grad x|) with backward differences
  const float eps = 1e-14f;
  const size_t n = static_cast<size_t>(n1) * n2 * n3;
  std::vector<float> g1(n, 0.0f), g2(n, 0.0f), g3(n, 0.0f);
  memset(g, 0, sizeof(float) * n);
  float f = 0;
  for (int i3 = 0; i3 < n3 - 1; ++i3)
    for (int i2 = 0; i2 < n2 - 1; ++i2)
      for (int i1 = 0; i1 < n1 - 1; ++i1) {
        const int k = i1 + n1 * (i2 + n2 * i3);
        const float s1 = x[k + 1] - x[k], s2 = x[k + n1] - x[k], s3 = x[k + n1 * n2] - x[k];
        const float t = static_cast<float>(std::sqrt(static_cast<double>(s1 * s1 + s2 * s2 + s3 * s3 + eps)));
        f += t;
        g1[k] = s1 / t;
        g2[k] = s2 / t;
        g3[k] = s3 / t;
      }
  for (int i3 = 1; i3 < n3; ++i3)
    for (int i2 = 1; i2 < n2; ++i2)
      for (int i1 = 1; i1 < n1; ++i1) {
        const int k = i1 + n1 * (i2 + n2 * i3);
        g[k] = -((g1[k] - g1[k - 1]) + (g2[k] - g2[k - n1]) + (g3[k] - g3[k - n1 * n2]));
      }
  return f;
}

void triangle_smoothing(float*** mod, int n1, int n2, int n3, int r1, int r2, int r3, int repeat) {
  std::vector<float> col(std::max(n2, n3));
  for (int rep = 0; rep < repeat; ++rep) {
    {
      Triangle t(r1, n1);
      for (int i3 = 0; i3 < n3; ++i3)
        for (int i2 = 0; i2 < n2; ++i2) t.apply(mod[i3][i2]);
    }
    {
      Triangle t(r2, n2);
      for (int i3 = 0; i3 < n3; ++i3)
        for (int i1 = 0; i1 < n1; ++i1) {
          for (int i2 = 0; i2 < n2; ++i2) col[i2] = mod[i3][i2][i1];
          t.apply(col.data());
          for (int i2 = 0; i2 < n2; ++i2) mod[i3][i2][i1] = col[i2];
        }
    }
    if (n3 > 1) {
      Triangle t(r3, n3);
      for (int i2 = 0; i2 < n2; ++i2)
        for (int i1 = 0; i1 < n1; ++i1) {
          for (int i3 = 0; i3 < n3; ++i3) col[i3] = mod[i3][i2][i1];
          t.apply(col.data());
          for (int i3 = 0; i3 < n3; ++i3) mod[i3][i2][i1] = col[i3];
        }
    }
  }
}

void precondition(emf_t* emf, fwi_t* fwi, float* g) {
  const int nx = emf->nx, ny = emf->ny, nz = emf->nz;
  const size_t nxyz = static_cast<size_t>(nx) * ny * nz;
  auto depth = [&](int i3) { return emf->oz + (i3 + 0.5) * emf->dz; };  // cell centre, double
  if (fwi->preco == 1) {
    // divide by the pseudo-Hessian, damped by 1e-5 of its largest value
    // below the seabed, then restore the gradient's norm
    float s1 = 0, s2 = 0;
    for (int k = 0; k < fwi->npar; ++k) {
      float z = 0;
      for (int i3 = 0; i3 < nz; ++i3)
        for (int i2 = 0; i2 < ny; ++i2)
          for (int i1 = 0; i1 < nx; ++i1)
            if (depth(i3) > emf->bathy[i2][i1]) z = std::max(z, fwi->phess[k][i3][i2][i1]);
      for (int i3 = 0; i3 < nz; ++i3)
        for (int i2 = 0; i2 < ny; ++i2)
          for (int i1 = 0; i1 < nx; ++i1) {
            float& v = g[i1 + nx * (i2 + ny * (i3 + static_cast<size_t>(nz) * k))];
            s1 += v * v;
            v = static_cast<float>(v / (fwi->phess[k][i3][i2][i1] + 1e-5 * z));
            s2 += v * v;
          }
    }
    const float r = static_cast<float>(std::sqrt(static_cast<double>(s1 / s2)));
    for (size_t i = 0; i < nxyz * fwi->npar; ++i) g[i] *= r;
  } else if (fwi->preco == 2) {
    // grow with depth below the seabed by the 1 ohm-m skin depth of the
    // first frequency
    const float skin = static_cast<float>(std::sqrt(2. / (emf->omegas[0] * kMu0)));
    for (int k = 0; k < fwi->npar; ++k)
      for (int i3 = 0; i3 < nz; ++i3)
        for (int i2 = 0; i2 < ny; ++i2)
          for (int i1 = 0; i1 < nx; ++i1) {
            float z = static_cast<float>(depth(i3));
            if (z > emf->bathy[i2][i1]) {
              z -= emf->bathy[i2][i1];
              float& v = g[i1 + nx * (i2 + ny * (i3 + static_cast<size_t>(nz) * k))];
              v = static_cast<float>(v * std::exp(static_cast<double>(z / skin)));
            }
          }
  } else if (fwi->preco == 3) {
    // smoothing as a model reparameterisation (Harlan 1995)
    memcpy(&fwi->grad[0][0][0][0], g, sizeof(float) * fwi->n);
    for (int k = 0; k < fwi->npar; ++k)
      triangle_smoothing(fwi->grad[k], nx, ny, nz, fwi->r1, fwi->r2, fwi->r3, fwi->repeat);
    memcpy(g, &fwi->grad[0][0][0][0], sizeof(float) * fwi->n);
  }
}

void fg_mod_reg(acq_t*, emf_t* emf, fwi_t* fwi, float* x, float* g) {
  const int nx = emf->nx, ny = emf->ny, nz = emf->nz;
  const size_t nxyz = static_cast<size_t>(nx) * ny * nz;
  const float cool = static_cast<float>(std::pow(0.8, fwi->iter));  // cooling with the iteration
  fwi->fcost_mod = 0;
  // the model change from the reference model below the seabed
  std::vector<float> xm(fwi->n), gm(fwi->n);
  for (int k = 0; k < fwi->npar; ++k)
    for (int i3 = 0; i3 < nz; ++i3)
      for (int i2 = 0; i2 < ny; ++i2)
        for (int i1 = 0; i1 < nx; ++i1) {
          const size_t i = i1 + nx * (i2 + ny * (i3 + static_cast<size_t>(nz) * k));
          xm[i] = (emf->oz + (i3 + 0.5) * emf->dz < emf->bathy[i2][i1] + emf->dz) ? 0 : (x[i] - fwi->xref[i]);
        }
  struct Term {
    float weight;
    float (*f)(float*, float*, int, int, int, float, float, float);
    const char *name_h, *name_v;
  } terms[2] = {{fwi->gamma1, regularization_tikhonov, "gradient_tikhonov_Rh", "gradient_tikhonov_Rv"},
                {fwi->gamma2, regularization_tv, "gradient_tv_Rh", "gradient_tv_Rv"}};
  for (const Term& t : terms) {
    if (!(t.weight > 0)) continue;
    std::fill(gm.begin(), gm.end(), 0.0f);
    const float gamma = t.weight * cool;
    for (int k = 0; k < fwi->npar; ++k) {
      const float f = t.f(&xm[k * nxyz], &gm[k * nxyz], nx, ny, nz, emf->dx, emf->dy, emf->dz);
      fwi->fcost_mod = static_cast<float>(fwi->fcost_mod + 0.5 * gamma * f);
      for (size_t i = k * nxyz; i < (k + 1) * nxyz; ++i) g[i] = static_cast<float>(g[i] + 0.5 * gamma * gm[i]);
    }
    if (emf->verb) {
      write_floats(t.name_h, gm.data(), nxyz);
      if (fwi->npar > 1) write_floats(t.name_v, gm.data() + nxyz, nxyz);
    }
  }
}
}  // extern "C"

namespace emmi_cli {

float l2norm(int n, const float* a) {
  float s = 0;
  for (int i = 0; i < n; ++i) s += a[i] * a[i];
  return std::sqrt(s);
}

float dotprod(int n, const float* a, const float* b) {
  float s = 0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

void Lbfgs::save(const float* x, const float* g) {
  // keep the last npair (x, g); the differences are formed by update()
  if (kpair < npair) {
    memcpy(sk[kpair].data(), x, sizeof(float) * n);
    memcpy(yk[kpair].data(), g, sizeof(float) * n);
    ++kpair;
  } else {
    for (int i = 0; i < npair - 1; ++i) {
      sk[i].swap(sk[i + 1]);
      yk[i].swap(yk[i + 1]);
    }
    memcpy(sk[npair - 1].data(), x, sizeof(float) * n);
    memcpy(yk[npair - 1].data(), g, sizeof(float) * n);
  }
}

void Lbfgs::update(const float* x, const float* g) {
  float *s = sk[kpair - 1].data(), *y = yk[kpair - 1].data();
  for (int i = 0; i < n; ++i) {
    s[i] = x[i] - s[i];
    y[i] = g[i] - y[i];
  }
}

void Lbfgs::first_loop(const float* g) {
  // q = g minus the projections on the stored pairs, newest first
  memcpy(q.data(), g, sizeof(float) * n);
  loop1 = false;
  if (!(l2norm(n, sk[kpair - 1].data()) > 0. && l2norm(n, yk[kpair - 1].data()) > 0.)) return;
  rho.assign(kpair, 0.0f);
  alp.assign(kpair, 0.0f);
  for (int i = kpair - 1; i >= 0; --i) {
    rho[i] = static_cast<float>(1. / dotprod(n, yk[i].data(), sk[i].data()));
    alp[i] = rho[i] * dotprod(n, sk[i].data(), q.data());
    for (int j = 0; j < n; ++j) q[j] -= alp[i] * yk[i][j];
  }
  loop1 = true;
}

void Lbfgs::second_loop(float* d) {
  // scale by gamma = s.y / y.y, add the pair corrections oldest first, and
  // return d = -q
  const float s_y = static_cast<float>(1. / rho[kpair - 1]);
  const float gamma = s_y / dotprod(n, yk[kpair - 1].data(), yk[kpair - 1].data());
  for (int j = 0; j < n; ++j) q[j] = gamma * q[j];
  for (int i = 0; i < kpair; ++i) {
    const float beta = rho[i] * dotprod(n, yk[i].data(), q.data());
    const float c = alp[i] - beta;
    for (int j = 0; j < n; ++j) q[j] += c * sk[i][j];
  }
  for (int j = 0; j < n; ++j) d[j] = -q[j];
}

void Lbfgs::direction(const float* g, float* d) {
  // the two loops without preconditioning; -g while no usable pair is stored
  const bool usable =
      kpair > 0 && l2norm(n, sk[kpair - 1].data()) > 0. && l2norm(n, yk[kpair - 1].data()) > 0.;
  if (!usable) {
    for (int j = 0; j < n; ++j) d[j] = -g[j];
    return;
  }
  first_loop(g);
  second_loop(d);
}

void line_search(LineSearch& ls, int n, float* x, float* g, const float* d, const std::function<float(float*, float*)>& fg,
                 const float* xmin, const float* xmax) {
  // bracketing, then bisection, on the Wolfe conditions (Nocedal and Wright,
  // eq. 3.6): alpha1 the largest step known too short, alpha2 the smallest
  // known too long
  const float inf = 0x3f3f3f3f;
  float a1 = 0, a2 = inf, fcost = 0;
  ls.alpha = 1.;
  std::vector<float> xk(x, x + n);
  const float gxd0 = dotprod(n, g, d);
  const float c1_gxd = ls.c1 * gxd0, c2_gxd = ls.c2 * gxd0;
  for (ls.ils = 0; ls.ils < ls.nls; ++ls.ils) {
    for (int j = 0; j < n; ++j) x[j] = xk[j] + ls.alpha * d[j];
    if (xmin)
      for (int j = 0; j < n; ++j) {
        if (x[j] < xmin[j]) x[j] = xmin[j];
        if (x[j] > xmax[j]) x[j] = xmax[j];
      }
    fcost = fg(x, g);
    ++ls.igrad;
    const float gxd = dotprod(n, g, d);
    const float fxx = ls.fk + ls.alpha * c1_gxd;
    if (fcost > fxx) {
      if (ls.verb) printf("Wolfe condition 1 fails: insufficient misfit decrease!\n");
      a2 = ls.alpha;
      ls.alpha = static_cast<float>(0.5 * (a1 + a2));
    } else if (gxd < c2_gxd) {
      if (ls.verb) printf("Wolfe condition 2 fails: stepsize is too small!\n");
      a1 = ls.alpha;
      if (a2 < inf) ls.alpha = static_cast<float>(0.5 * (a1 + a2));
      else ls.alpha = static_cast<float>(ls.alpha * 2.0);
    } else {
      break;
    }
    if (ls.verb) {
      printf("#line search %d, alp1=%f alp2=%f alp=%f\n", ls.ils, a1, a2, ls.alpha);
      printf("------------------------------------------------\n");
    }
  }
  if (fcost <= ls.fk) {
    ls.failed = false;
    ls.fk = fcost;
  } else {
    ls.failed = true;
  }
}

}  // namespace emmi_cli
