// inversion.h -- the l-BFGS method and its line search (src/opt.c:22-257)
// as used by the inversion loop of emmi_main; the regularisation and
// preconditioning entry points are declared in emmi_cli.h.
#pragma once

#include <functional>
#include <vector>

#include "emmi_cli.h"

namespace emmi_cli {

// sequential float sums, as src/opt.c:22-40 forms them
float l2norm(int n, const float* a);
float dotprod(int n, const float* a, const float* b);

// the stored pairs and the two-loop recursion (src/opt.c:52-186)
struct Lbfgs {
  int n, npair, kpair = 0;
  std::vector<std::vector<float>> sk, yk;
  std::vector<float> q, rho, alp;
  bool loop1 = false;
  Lbfgs(int n_, int npair_) : n(n_), npair(npair_), sk(npair_, std::vector<float>(n_)), yk(npair_, std::vector<float>(n_)), q(n_) {}
  void save(const float* x, const float* g);    // lbfgs_save
  void update(const float* x, const float* g);  // lbfgs_update: pair k <- (x - x_k, g - g_k)
  void first_loop(const float* g);              // lbfgs_descent1: q
  void second_loop(float* d);                   // lbfgs_descent2: d = -H q
  void direction(const float* g, float* d);     // lbfgs_descent
};

// bisection line search on the Wolfe conditions (src/opt.c:198-257)
struct LineSearch {
  int nls = 5, ils = 0, igrad = 0;
  float c1 = 1e-4f, c2 = 0.9f, alpha = 1, fk = 0;
  bool verb = false, failed = false;
};
// x <- the accepted point along d from x, g <- its gradient; xmin/xmax clip
// x when given (bound=1)
void line_search(LineSearch& ls, int n, float* x, float* g, const float* d,
                 const std::function<float(float*, float*)>& fg, const float* xmin, const float* xmax);

}  // namespace emmi_cli
