// line_kernels.h -- exact line block Gauss-Seidel (line_kernels.hip).
//
// A sweep along axis a visits the lines (a; pb, pc) in four colours
// (pb%2, pc%2); lines of one colour are independent.  Per colour the line
// systems are factored once (k_line_factor: FREC values per node, depending
// on grid, medium and frequency only) and every sweep then computes the
// right-hand sides (k_line_rhs, SREC values per node) and runs a forward/back
// substitution (k_line_solve).  Records hold values of type LineFac<T>::type
// (factors) or LineScr<T>::type (line records) in 16-byte chunks: value k of node p of line l is element k % V of chunk
// (p * nchunks + k / V) * nlines + l, V = 16 / sizeof(value), so a lane reads
// its node with 16-byte loads and a wave's load covers contiguous memory.  A launch takes a pack of up to
// LINE_PACK_MAX jobs (one per right-hand side or per level) that share the
// axis and colour.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include "emmi_dev.h"

namespace emmi {

constexpr int LINE_PACK_MAX = 16;

// Value types of the line factor records (LineFac) and of the right-hand
// side / substitution records (LineScr) for cycles of storage precision T.
// Factors are FP64 for both precisions: in air the line blocks are close to
// singular and entries of W^-1 exceed the FP32 range.  The line records
// (residuals sigma, rho and the forward intermediates z) take the cycle's
// precision.
template <typename T> struct LineFac { using type = double; };
template <typename T> struct LineScr { using type = T; };

template <typename T>
struct LineJob {
  Level<T> L;
  typename cvec<T>::type* u;
  const typename cvec<T>::type* f;
  typename LineFac<T>::type* fac;  // factor records of this colour, chunked: [n_a][FREC/V][nlines][V]
  typename LineScr<T>::type* scr;  // right-hand side / substitution records, chunked: [n_a][SREC/V][nlines][V]
  int cb, cc;    // colour: parity of pb and pc (1 = odd)
  int nlo, nhi;  // lines of this colour along the low- and high-stride perpendicular axes
};

template <typename T>
struct LinePack {
  LineJob<T> job[LINE_PACK_MAX];
  int njobs;
};
// a pack is a kernel argument: it must stay within the 4 KiB of kernel
// arguments a launch takes
static_assert(sizeof(LinePack<double>) <= 4096 && sizeof(LinePack<float>) <= 4096, "line pack exceeds 4 KiB");

// number of lines of a colour, split by perpendicular axis (low stride first)
void line_colour_counts(const int* n, int a, int colour, int* nlo, int* nhi);
// bytes of one colour's factor records / substitution records for values of
// esize bytes (4 or 8)
size_t line_factor_bytes(const int* n, int a, int colour, size_t esize);
size_t line_scratch_bytes(const int* n, int a, int colour, size_t esize);
// fills the colour and counts of a job from its level shape
template <typename T>
inline void line_job_colour(LineJob<T>& J, int a, int colour) {
  J.cb = colour & 1;
  J.cc = (colour >> 1) & 1;
  line_colour_counts(J.L.n, a, colour, &J.nlo, &J.nhi);
}

// substitution kernel of later launches: 0 by line count (or by the
// autotuner's table where it has the launch's shape), 1 one lane per line
// half, 2 four lanes per line half, 3 four lanes building their own
// right-hand sides, 4 four lanes after a right-hand-side launch; returns the
// previous mode or -1
int set_line_kernel(int mode);

// Autotuning: a launch shape (value size of the cycle precision, axis, line
// length, lines per job, jobs per pack) and the substitution choice measured
// fastest for it.  LINE_RULE: the built-in rule by line count.
// LINE_RULE_ONLY (forced launches only): the rule, ignoring the table.
enum LineChoice {
  LINE_RULE_ONLY = -2, LINE_RULE = -1, LINE_ONE_LANE = 0, LINE_QUAD_FUSED = 1, LINE_QUAD_SEPARATE = 2, LINE_NCHOICE = 3
};
struct LineShape {
  int esize, axis, n;
  long nl;
  int njobs;
  bool operator<(const LineShape& o) const {
    if (esize != o.esize) return esize < o.esize;
    if (axis != o.axis) return axis < o.axis;
    if (n != o.n) return n < o.n;
    if (nl != o.nl) return nl < o.nl;
    return njobs < o.njobs;
  }
};
template <typename T>
LineShape line_shape(const LinePack<T>& P, int axis);
// single-line launches (cyclic reduction) have no choice to tune
template <typename T>
bool line_pcr_shape(const LinePack<T>& P, int axis);
void set_line_choice(const LineShape& k, int choice, double us_choice, double us_rule);
bool has_line_choice(const LineShape& k);
void clear_line_choices();
// rows of (esize, axis, n, nl, njobs, choice) and (us of the choice, us of the
// rule); returns the number of entries
int line_choices(int* out, double* us, int maxn);

template <typename T>
void launch_line_factor(const LinePack<T>& P, int axis, hipStream_t s);
template <typename T>
void launch_line_solve(const LinePack<T>& P, int axis, hipStream_t s);
// the same with the substitution choice forced (LINE_RULE: the rule alone)
template <typename T>
void launch_line_solve_as(const LinePack<T>& P, int axis, int choice, hipStream_t s);

}  // namespace emmi
