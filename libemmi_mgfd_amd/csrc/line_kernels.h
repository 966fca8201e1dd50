// line_kernels.h -- exact line block Gauss-Seidel (line_kernels.hip).
//
// A sweep along axis a visits the lines (a; pb, pc) in four colours
// (pb%2, pc%2); lines of one colour are independent.  Per colour the line
// systems are factored once (k_line_factor: records of FREC doubles per node,
// depending on grid, medium and frequency only) and every sweep then runs
// the residual and a forward/back substitution (k_line_solve, scratch of
// SREC doubles per node).  A launch takes a pack of up to LINE_PACK_MAX jobs
// (one per right-hand side or per level) that share the axis and colour.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include "emmi_dev.h"

namespace emmi {

constexpr int LINE_PACK_MAX = 8;

template <typename T>
struct LineJob {
  Level<T> L;
  typename cvec<T>::type* u;
  const typename cvec<T>::type* f;
  double* fac;   // factor records of this colour: [n_a][FREC][nlines]
  double* scr;   // substitution scratch: [n_a][SREC][nlines]
  int cb, cc;    // colour: parity of pb and pc (1 = odd)
  int nlo, nhi;  // lines of this colour along the low- and high-stride perpendicular axes
};

template <typename T>
struct LinePack {
  LineJob<T> job[LINE_PACK_MAX];
  int njobs;
};

// number of lines of a colour, split by perpendicular axis (low stride first)
void line_colour_counts(const int* n, int a, int colour, int* nlo, int* nhi);
// doubles of one colour's factor records / substitution scratch
size_t line_factor_doubles(const int* n, int a, int colour);
size_t line_scratch_doubles(const int* n, int a, int colour);
// fills the colour and counts of a job from its level shape
template <typename T>
inline void line_job_colour(LineJob<T>& J, int a, int colour) {
  J.cb = colour & 1;
  J.cc = (colour >> 1) & 1;
  line_colour_counts(J.L.n, a, colour, &J.nlo, &J.nhi);
}

template <typename T>
void launch_line_factor(const LinePack<T>& P, int axis, hipStream_t s);
template <typename T>
void launch_line_solve(const LinePack<T>& P, int axis, hipStream_t s);

}  // namespace emmi
