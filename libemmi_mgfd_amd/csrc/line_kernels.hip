// line_kernels.hip -- exact line block Gauss-Seidel, factored once per level.
//
// Line (a; pb, pc) holds the vertex patches of the nodes (p, pb, pc),
// p = 1..N-1, N = n_a (src/gmg.c:340-1298): the along-line edges E_a(e),
// e = 0..N-1, and per node four perpendicular edges, slots
//   0: E_b(p, pb-1, pc)  1: E_b(p, pb, pc)  2: E_c(p, pb, pc-1)  3: E_c(p, pb, pc)
// with (a, b, c) cyclic.  Each line system is solved exactly for the
// correction of the current residual.  An along-line edge couples only to
// the perpendicular edges of its two end nodes, through the constant vector
// v = (Hc, -Hc, Hb, -Hb):
//     d_e dE_a(e) + v.da_e - v.da_{e+1} = rho_e,
// so it is eliminated first; the node unknowns then form a complex
// symmetric 4x4 block-tridiagonal system
//     D_p = K_p - v v^T (1/d_{p-1} + 1/d_p),   B_p = diag(l_{p-1}) + v v^T / d_{p-1}
//     W_1 = D_1,  W_p = D_p - B_p W_{p-1}^{-1} B_p
// whose factors depend on the grid, the medium and the frequency only.
// k_line_factor stores, per node, W_p^{-1} (symmetric, 10 entries) and
// 1/d_p once; k_line_rhs computes the residuals sigma_p, rho_p of the lines
// and k_line_solve then runs, per sweep,
//     y_p = g_p - B_p z_{p-1},  z_p = W_p^{-1} y_p          (forward)
//     da_p = z_p - W_p^{-1} B_{p+1} da_{p+1},
//     dE_a(p) = rho_p/d_p + v.(da_{p+1} - da_p) / d_p        (backward)
// with g_p = sigma_p + v (rho_{p-1}/d_{p-1} - rho_p/d_p) from the residual.
// Lines of one colour (pb%2, pc%2) are not coupled, so a launch sweeps one
// colour; every kernel computes in FP64 whatever the storage precision.
// Records hold LineFac<T> / LineScr<T> values in 16-byte chunks (layout in
// line_kernels.h);
// the substitution keeps a ring of the next D nodes' chunks in registers, so
// D nodes of loads are in flight while a node is computed.
#include "line_kernels.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>
#include "stencil.h"
#include "timing.h"

namespace emmi {

constexpr int FREC = 22;  // values per node in the factor records: W^-1 (10 complex), 1/d
constexpr int SREC = 10;  // values per node in the line records: sigma or z (4 complex), rho

// 16-byte chunk of record values
template <typename R>
struct alignas(16) Ch {
  static constexpr int V = 16 / (int)sizeof(R);
  R v[V];
};
template <typename R> constexpr int fchunks() { return (FREC + Ch<R>::V - 1) / Ch<R>::V; }
template <typename R> constexpr int schunks() { return (SREC + Ch<R>::V - 1) / Ch<R>::V; }
// minimum waves per SIMD the right-hand-side kernels are compiled for
// (bounds their registers; more waves keep more stencil loads in flight)
#ifndef EMMI_LINE_RHS_MIN_WAVES
#define EMMI_LINE_RHS_MIN_WAVES 1
#endif

template <typename T> using FacRec = typename LineFac<T>::type;
template <typename T> using ScrRec = typename LineScr<T>::type;

// nodes of records in flight in the substitution (build-time tunable)
// for FP32-storage and FP64 cycles
#ifndef EMMI_LINE_RING_F32
#define EMMI_LINE_RING_F32 2
#endif
#ifndef EMMI_LINE_RING_F64
#define EMMI_LINE_RING_F64 1
#endif
template <typename T> constexpr int ring_depth() { return sizeof(T) == 4 ? EMMI_LINE_RING_F32 : EMMI_LINE_RING_F64; }
// Launches of at most this many lines per job (FP32-storage / FP64 cycles)
// use the quad-lane substitution, whose ring holds EMMI_LINE_QUAD_RING nodes.
#ifndef EMMI_LINE_QUAD_MAX_LINES_F32
#define EMMI_LINE_QUAD_MAX_LINES_F32 4096
#endif
#ifndef EMMI_LINE_QUAD_MAX_LINES_F64
#define EMMI_LINE_QUAD_MAX_LINES_F64 4096
#endif
#ifndef EMMI_LINE_QUAD_RING
#define EMMI_LINE_QUAD_RING 3
#endif
// quad-lane launches of at most this many lines per job (FP32-storage /
// FP64 cycles) compute their right-hand sides themselves (no k_line_rhs
// launch)
#ifndef EMMI_LINE_FUSE_MAX_LINES_F32
#define EMMI_LINE_FUSE_MAX_LINES_F32 4096
#endif
#ifndef EMMI_LINE_FUSE_MAX_LINES_F64
#define EMMI_LINE_FUSE_MAX_LINES_F64 4096
#endif
// threads of a block of the fused launches (a multiple of 128; the waves
// beyond the first two only compute right-hand sides)
#ifndef EMMI_LINE_FUSE_THREADS
#define EMMI_LINE_FUSE_THREADS 512
#endif
// the same for lines of 64 nodes and more
#ifndef EMMI_LINE_FUSE_THREADS_LONG
#define EMMI_LINE_FUSE_THREADS_LONG 256
#endif
template <typename T> constexpr int fuse_max_lines() {
  return sizeof(T) == 4 ? EMMI_LINE_FUSE_MAX_LINES_F32 : EMMI_LINE_FUSE_MAX_LINES_F64;
}
template <typename T> constexpr int quad_max_lines() {
  return sizeof(T) == 4 ? EMMI_LINE_QUAD_MAX_LINES_F32 : EMMI_LINE_QUAD_MAX_LINES_F64;
}

// value k of a node's chunks, as double
template <typename R, int NC>
__device__ __forceinline__ double val(const Ch<R> (&c)[NC], int k) {
  return (double)c[k / Ch<R>::V].v[k % Ch<R>::V];
}
template <typename R, int NC>
__device__ __forceinline__ void put(Ch<R> (&c)[NC], int k, double x) {
  c[k / Ch<R>::V].v[k % Ch<R>::V] = (R)x;
}
// loads the NC chunks of one node; src points at chunk 0 of that node's line
template <typename R, int NC>
__device__ __forceinline__ void load_node(Ch<R> (&c)[NC], const Ch<R>* __restrict__ src, long stride) {
#pragma unroll
  for (int k = 0; k < NC; k++) c[k] = src[k * stride];
}

__device__ __forceinline__ int symidx(int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a * 4 - (a * (a - 1)) / 2 + (b - a);
}

__device__ __forceinline__ void inv4(cxd (&M)[4][4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const cxd piv = cinv(M[k][k]);
    M[k][k] = cxd(1.0, 0.0);
#pragma unroll
    for (int j = 0; j < 4; j++) M[k][j] = M[k][j] * piv;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (i == k) continue;
      const cxd fct = M[i][k];
      M[i][k] = cxd(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < 4; j++) M[i][j] -= fct * M[k][j];
    }
  }
}

// constants of one line
template <typename T, int A>
struct LineGeo {
  static constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  int N, pb, pc;
  long sa, base, oa, ob0, ob1, oc0, oc1;
  double Hb, Hc, Wb0, Wb1, Wc0, Wc1, sIb, sIc, dre_a;
  double v[4];

  __device__ __forceinline__ LineGeo() {}
  __device__ __forceinline__ LineGeo(const Level<T>& L, int pb_, int pc_) { init(L, pb_, pc_); }
  __device__ __forceinline__ void init(const Level<T>& L, int pb_, int pc_) {
    pb = pb_;
    pc = pc_;
    N = L.n[A];
    sa = L.es[A];
    base = (long)pb * L.es[B] + (long)pc * L.es[C];
    oa = A * L.nc + base;
    ob0 = B * L.nc + base - L.es[B];
    ob1 = B * L.nc + base;
    oc0 = C * L.nc + base - L.es[C];
    oc1 = C * L.nc + base;
    Hb = L.H[B][pb];
    Hc = L.H[C][pc];
    Wb0 = L.W[B][pb - 1];
    Wb1 = L.W[B][pb];
    Wc0 = L.W[C][pc - 1];
    Wc1 = L.W[C][pc];
    sIb = (double)L.IW[B][pb - 1] + (double)L.IW[B][pb];
    sIc = (double)L.IW[C][pc - 1] + (double)L.IW[C][pc];
    dre_a = Hc * sIb + Hb * sIc;
    v[0] = Hc; v[1] = -Hc; v[2] = Hb; v[3] = -Hb;
  }
  // diagonal of the along-line edge e
  __device__ __forceinline__ cxd d_edge(const Level<T>& L, int e) const {
    return cxd((double)L.W[A][e] * dre_a, -L.omu * (double)L.sbar[oa + e * sa]);
  }
  // along-line coupling of node p to node p+1, per slot
  __device__ __forceinline__ void lcoef(const Level<T>& L, int p, double (&l)[4]) const {
    const double iWa = L.IW[A][p];
    l[0] = -Wb0 * Hc * iWa; l[1] = -Wb1 * Hc * iWa; l[2] = -Wc0 * Hb * iWa; l[3] = -Wc1 * Hb * iWa;
  }
};

template <typename T, int A>
__device__ __forceinline__ bool line_of(const LineJob<T>& J, int tid, int& pb, int& pc) {
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  if (tid >= J.nlo * J.nhi) return false;
  const int sb = J.cb ? 1 : 2, sc = J.cc ? 1 : 2;
  // the perpendicular axis with the smaller stride varies fastest
  if (B < C) { pb = sb + 2 * (tid % J.nlo); pc = sc + 2 * (tid / J.nlo); }
  else { pc = sc + 2 * (tid % J.nlo); pb = sb + 2 * (tid / J.nlo); }
  return true;
}

// D_p: the diagonal block of node p after the along-line edges p-1 and p
// (with 1/d = rdm, rdp) are eliminated
template <typename T, int A>
__device__ __forceinline__ void diag_block(const Level<T>& L, const LineGeo<T, A>& G, int p, cxd rdm, cxd rdp,
                                           cxd (&D)[4][4]) {
  const double Ha = L.H[A][p];
  const double iWa0 = L.IW[A][p - 1], iWa1 = L.IW[A][p];
  const double Gb = Ha * G.sIc + G.Hc * (iWa1 + iWa0);
  const double Gc = G.Hb * (iWa1 + iWa0) + Ha * G.sIb;
  const long off = p * G.sa;
  D[0][0] = cxd(G.Wb0 * Gb, -L.omu * (double)L.sbar[G.ob0 + off]);
  D[1][1] = cxd(G.Wb1 * Gb, -L.omu * (double)L.sbar[G.ob1 + off]);
  D[2][2] = cxd(G.Wc0 * Gc, -L.omu * (double)L.sbar[G.oc0 + off]);
  D[3][3] = cxd(G.Wc1 * Gc, -L.omu * (double)L.sbar[G.oc1 + off]);
  D[0][1] = D[1][0] = cxd(0, 0);
  D[2][3] = D[3][2] = cxd(0, 0);
  D[0][2] = D[2][0] = cxd(-Ha, 0);
  D[0][3] = D[3][0] = cxd(Ha, 0);
  D[1][2] = D[2][1] = cxd(Ha, 0);
  D[1][3] = D[3][1] = cxd(-Ha, 0);
  const cxd s = rdm + rdp;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) D[i][j] -= s * (G.v[i] * G.v[j]);
}

// D -= B Wi B for the coupling B = diag(l) + v v^T rd of two neighbouring
// nodes (Wi: the inverse factor of the neighbour)
__device__ __forceinline__ void sub_bwb(cxd (&D)[4][4], const cxd (&Wi)[4][4], const double (&l)[4], cxd rd,
                                        const double (&v)[4]) {
  cxd wv[4];
#pragma unroll
  for (int i = 0; i < 4; i++) wv[i] = Wi[i][0] * v[0] + Wi[i][1] * v[1] + Wi[i][2] * v[2] + Wi[i][3] * v[3];
  cxd Pm[4][4];   // Wi B
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) Pm[i][j] = Wi[i][j] * l[j] + (wv[i] * rd) * v[j];
  cxd vP[4];
#pragma unroll
  for (int j = 0; j < 4; j++) vP[j] = Pm[0][j] * v[0] + Pm[1][j] * v[1] + Pm[2][j] * v[2] + Pm[3][j] * v[3];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) D[i][j] -= Pm[i][j] * l[i] + (vP[j] * rd) * v[i];
}

// inverts D in place and stores the symmetric part with 1/d of node p
template <typename R>
__device__ __forceinline__ void store_factor(Ch<R>* rec, long stride, int p, cxd (&D)[4][4], cxd rd) {
  constexpr int FC = fchunks<R>();
  inv4(D);
  Ch<R> c[FC] = {};
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = i; j < 4; j++) {
      const int sidx = symidx(i, j);
      const cxd w = (i == j) ? D[i][j] : (D[i][j] + D[j][i]) * 0.5;
      put(c, 2 * sidx, w.re);
      put(c, 2 * sidx + 1, w.im);
    }
  put(c, 20, rd.re);
  put(c, 21, rd.im);
  Ch<R>* r = rec + (long)p * FC * stride;
#pragma unroll
  for (int k = 0; k < FC; k++) r[k * stride] = c[k];
}

// Twisted factorization of each line, meeting at node m = N/2 of the nodes
// 1..M (M = N-1):
//   top     W_1 = D_1,  W_p = D_p - B_p W_{p-1}^-1 B_p        p = 2..m-1
//   bottom  U_M = D_M,  U_p = D_p - B_{p+1} U_{p+1}^-1 B_{p+1}  p = M-1..m+1
//   middle  S_m = D_m - B_m W_{m-1}^-1 B_m - B_{m+1} U_{m+1}^-1 B_{m+1}
// Node p stores the inverse of its factor (W, S or U) and 1/d_p; node 0
// stores 1/d_0.
template <typename T, int A>
__global__ void __launch_bounds__(64) k_line_factor(LinePack<T> P) {
  using R = FacRec<T>;
  constexpr int FC = fchunks<R>(), V = Ch<R>::V;
  const LineJob<T>& J = P.job[blockIdx.y];
  const Level<T>& L = J.L;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  int pb, pc;
  if (!line_of<T, A>(J, tid, pb, pc)) return;
  const LineGeo<T, A> G(L, pb, pc);
  const int N = G.N;
  if (N < 2) return;
  const int M = N - 1, m = N / 2;
  const long stride = (long)J.nlo * J.nhi;
  Ch<R>* rec = reinterpret_cast<Ch<R>*>(J.fac) + tid;
  const cxd rd0 = cinv(G.d_edge(L, 0));
  {
    // node 0 holds only 1/d_0 (values 20, 21: one chunk)
    Ch<R> c[FC] = {};
    put(c, 20, rd0.re);
    put(c, 21, rd0.im);
    rec[(20 / V) * stride] = c[20 / V];
  }
  cxd Wt[4][4], Wb[4][4], Dm[4][4];
  double l[4];
  cxd rd_prev = rd0;
  for (int p = 1; p <= m - 1; p++) {
    const cxd rd = cinv(G.d_edge(L, p));
    diag_block(L, G, p, rd_prev, rd, Dm);
    if (p >= 2) {
      G.lcoef(L, p - 1, l);
      sub_bwb(Dm, Wt, l, rd_prev, G.v);
    }
    store_factor(rec, stride, p, Dm, rd);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) Wt[i][j] = Dm[i][j];
    rd_prev = rd;
  }
  for (int p = M; p >= m + 1; p--) {
    const cxd rdm = cinv(G.d_edge(L, p - 1)), rdp = cinv(G.d_edge(L, p));
    diag_block(L, G, p, rdm, rdp, Dm);
    if (p <= M - 1) {
      G.lcoef(L, p, l);
      sub_bwb(Dm, Wb, l, rdp, G.v);
    }
    store_factor(rec, stride, p, Dm, rdp);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) Wb[i][j] = Dm[i][j];
  }
  {
    const cxd rdm = cinv(G.d_edge(L, m - 1)), rdp = cinv(G.d_edge(L, m));
    diag_block(L, G, m, rdm, rdp, Dm);
    if (m >= 2) {
      G.lcoef(L, m - 1, l);
      sub_bwb(Dm, Wt, l, rdm, G.v);
    }
    if (m <= M - 1) {
      G.lcoef(L, m, l);
      sub_bwb(Dm, Wb, l, rdp, G.v);
    }
    store_factor(rec, stride, m, Dm, rdp);
  }
}

// ---------------------------------------------------------------------------
// single-line launches: parallel cyclic reduction
// ---------------------------------------------------------------------------
// The coarsest level of a semi-coarsening pattern is one line along the
// kept axis (the full grid length), so the substitution would be one serial
// chain of N/2 nodes per half.  With one line per launch, the node system
//     B_p x_{p-1} + D_p x_p + B_{p+1} x_{p+1} = g_p,   p = 1..M, x_0 = x_N = 0
// is reduced instead: at stride s = 1, 2, 4, ... each equation eliminates its
// neighbours at distance s,
//     alpha_p = -A_p D_{p-s}^-1,  gamma_p = -C_p D_{p+s}^-1
//     D_p <- D_p + alpha_p C_{p-s} + gamma_p A_{p+s},  A_p <- alpha_p A_{p-s},
//     C_p <- gamma_p C_{p+s},  g_p <- g_p + alpha_p g_{p-s} + gamma_p g_{p+s}
// (A_p couples p to p-s, C_p to p+s, both start as the line's B), until no
// coupling is left and x_p = D_p^-1 g_p.  k_line_pcr_factor builds alpha,
// gamma and D^-1 once per frequency; k_line_pcr_solve then takes log2(M)
// block-wide steps.  Launches of one line of up to PCR_MAX_NODES nodes use it.
constexpr int PCR_MAX_NODES = 256;

__host__ __device__ constexpr int pcr_levels(int M) {
  int k = 0;
  for (int s = 1; s < M; s *= 2) k++;
  return k;
}
// complex entries of the factor buffer: alpha and gamma per level and node,
// D^-1 and 1/d per node, two work copies of (D, A, C) per node
__host__ __device__ constexpr long pcr_entries(int N) {
  return (long)N * (32L * pcr_levels(N - 1) + 16 + 1 + 2 * 48);
}

struct PcrLayout {
  cxd *alpha, *dinv, *rd, *work;
  int N, nlev;
  __device__ PcrLayout(void* base, int N_) : N(N_), nlev(pcr_levels(N_ - 1)) {
    cxd* b = reinterpret_cast<cxd*>(base);
    alpha = b;                                   // [lev][p][2][16]: alpha, gamma
    dinv = alpha + (long)nlev * N * 32;          // [p][16]
    rd = dinv + (long)N * 16;                    // [e]
    work = rd + N;                               // [2][p][48]: D, A, C
  }
  __device__ cxd* ag(int lev, int p, int which) const { return alpha + (((long)lev * N + p) * 2 + which) * 16; }
  __device__ cxd* w(int buf, int p) const { return work + ((long)buf * N + p) * 48; }
};

__device__ __forceinline__ void ld16(const cxd* s, cxd (&M)[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) M[i][j] = s[4 * i + j];
}
__device__ __forceinline__ void st16(cxd* d, const cxd (&M)[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) d[4 * i + j] = M[i][j];
}
// C = A B
__device__ __forceinline__ void mm4(const cxd (&A)[4][4], const cxd (&B)[4][4], cxd (&C)[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      cxd acc(0, 0);
#pragma unroll
      for (int k = 0; k < 4; k++) acc += A[i][k] * B[k][j];
      C[i][j] = acc;
    }
}
__device__ __forceinline__ void zero4(cxd (&M)[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) M[i][j] = cxd(0, 0);
}
// the coupling B = diag(l) + v v^T rd of two neighbouring nodes
template <typename T, int A>
__device__ __forceinline__ void coupling(const Level<T>& L, const LineGeo<T, A>& G, int e, cxd rd, cxd (&Bm)[4][4]) {
  double l[4];
  G.lcoef(L, e, l);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) Bm[i][j] = (G.v[i] * G.v[j]) * rd + cxd(i == j ? l[i] : 0.0, 0);
}

template <typename T, int A>
__global__ void __launch_bounds__(PCR_MAX_NODES) k_line_pcr_factor(LinePack<T> P) {
  const LineJob<T>& J = P.job[blockIdx.y];
  const Level<T>& L = J.L;
  int pb, pc;
  if (!line_of<T, A>(J, 0, pb, pc)) return;   // uniform over the block
  const LineGeo<T, A> G(L, pb, pc);
  const int N = G.N;
  if (N < 2 || N > PCR_MAX_NODES) return;     // uniform
  const int M = N - 1, p = threadIdx.x;
  const PcrLayout F(J.fac, N);
  for (int e = p; e <= M; e += blockDim.x) F.rd[e] = cinv(G.d_edge(L, e));
  __syncthreads();
  const bool own = p >= 1 && p <= M;
  if (own) {
    cxd Dm[4][4], Am[4][4], Cm[4][4];
    diag_block(L, G, p, F.rd[p - 1], F.rd[p], Dm);
    if (p >= 2) coupling(L, G, p - 1, F.rd[p - 1], Am);
    else zero4(Am);
    if (p <= M - 1) coupling(L, G, p, F.rd[p], Cm);
    else zero4(Cm);
    cxd* w = F.w(0, p);
    st16(w, Dm);
    st16(w + 16, Am);
    st16(w + 32, Cm);
  }
  __syncthreads();
  int cur = 0, lev = 0;
  for (int s = 1; s < M; s *= 2, lev++) {
    if (own) {
      cxd Dm[4][4], Am[4][4], Cm[4][4], al[4][4], ga[4][4], t[4][4], u[4][4];
      const cxd* w = F.w(cur, p);
      ld16(w, Dm);
      ld16(w + 16, Am);
      ld16(w + 32, Cm);
      cxd An[4][4], Cn[4][4];
      zero4(An);
      zero4(Cn);
      if (p - s >= 1) {
        const cxd* wm = F.w(cur, p - s);
        ld16(wm, t);
        inv4(t);
        mm4(Am, t, al);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) al[i][j] = -al[i][j];
        ld16(wm + 32, t);   // C_{p-s}
        mm4(al, t, u);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) Dm[i][j] += u[i][j];
        ld16(wm + 16, t);   // A_{p-s}
        mm4(al, t, An);
      } else {
        zero4(al);
      }
      if (p + s <= M) {
        const cxd* wp = F.w(cur, p + s);
        ld16(wp, t);
        inv4(t);
        mm4(Cm, t, ga);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) ga[i][j] = -ga[i][j];
        ld16(wp + 16, t);   // A_{p+s}
        mm4(ga, t, u);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) Dm[i][j] += u[i][j];
        ld16(wp + 32, t);   // C_{p+s}
        mm4(ga, t, Cn);
      } else {
        zero4(ga);
      }
      st16(F.ag(lev, p, 0), al);
      st16(F.ag(lev, p, 1), ga);
      cxd* wn = F.w(cur ^ 1, p);
      st16(wn, Dm);
      st16(wn + 16, An);
      st16(wn + 32, Cn);
    }
    __syncthreads();
    cur ^= 1;
  }
  if (own) {
    cxd Dm[4][4];
    ld16(F.w(cur, p), Dm);
    inv4(Dm);
    st16(F.dinv + (long)p * 16, Dm);
  }
}

template <typename T, int A>
__global__ void __launch_bounds__(4 * PCR_MAX_NODES) k_line_pcr_solve(LinePack<T> P) {
  using RS = ScrRec<T>;
  constexpr int SC = schunks<RS>();
  // LDS holds the complex values as double2 (cxd has a constructor)
  __shared__ double2 g_[2][PCR_MAX_NODES][4];
  __shared__ double2 xs_[PCR_MAX_NODES + 1][4];
  __shared__ double2 qs_[PCR_MAX_NODES];
  auto get = [](const double2& a) { return cxd(a.x, a.y); };
  auto set = [](double2& a, cxd v) { a.x = v.re; a.y = v.im; };
  const LineJob<T>& J = P.job[blockIdx.y];
  const Level<T>& L = J.L;
  int pb, pc;
  if (!line_of<T, A>(J, 0, pb, pc)) return;   // uniform over the block
  const LineGeo<T, A> G(L, pb, pc);
  const int N = G.N;
  if (N < 2 || N > PCR_MAX_NODES) return;     // uniform
  const int M = N - 1, tid = threadIdx.x;
  const PcrLayout F(J.fac, N);
  // residual records of the line's nodes (as k_line_rhs writes them)
  for (int p = tid; p <= M; p += blockDim.x) node_rhs<T, A>(J, 0, pb, pc, p, 1);
  __syncthreads();   // the records are complete (workgroup scope)
  const Ch<RS>* rec = reinterpret_cast<const Ch<RS>*>(J.scr);
  for (int e = tid; e <= M; e += blockDim.x) {
    Ch<RS> c[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) c[k] = rec[(long)e * SC + k];
    set(qs_[e], cval(c, 8) * F.rd[e]);
  }
  __syncthreads();
  const int p = tid / 4 + 1, r = tid % 4;
  const bool own = p <= M;
  if (own) {
    Ch<RS> c[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) c[k] = rec[(long)p * SC + k];
    set(g_[0][p][r], cval(c, 2 * r) + (get(qs_[p - 1]) - get(qs_[p])) * G.v[r]);
  }
  __syncthreads();
  int cur = 0, lev = 0;
  for (int s = 1; s < M; s *= 2, lev++) {
    if (own) {
      cxd acc = get(g_[cur][p][r]);
      if (p - s >= 1) {
        const cxd* a = F.ag(lev, p, 0) + 4 * r;
#pragma unroll
        for (int j = 0; j < 4; j++) acc += a[j] * get(g_[cur][p - s][j]);
      }
      if (p + s <= M) {
        const cxd* a = F.ag(lev, p, 1) + 4 * r;
#pragma unroll
        for (int j = 0; j < 4; j++) acc += a[j] * get(g_[cur][p + s][j]);
      }
      set(g_[cur ^ 1][p][r], acc);
    }
    __syncthreads();
    cur ^= 1;
  }
  if (own) {
    const cxd* di = F.dinv + (long)p * 16 + 4 * r;
    cxd acc(0, 0);
#pragma unroll
    for (int j = 0; j < 4; j++) acc += di[j] * get(g_[cur][p][j]);
    set(xs_[p][r], acc);
  }
  if (tid < 4) {
    set(xs_[0][tid], cxd(0, 0));
    set(xs_[N][tid], cxd(0, 0));
  }
  __syncthreads();
  typename cvec<T>::type* __restrict__ u = J.u;
  if (own) {
    const long o = (r == 0 ? G.ob0 : r == 1 ? G.ob1 : r == 2 ? G.oc0 : G.oc1) + (long)p * G.sa;
    st<T>(u, o, ld<T>(u, o) + get(xs_[p][r]));
  }
  for (int e = tid; e <= M; e += blockDim.x) {
    cxd dv(0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) dv += (get(xs_[e + 1][i]) - get(xs_[e][i])) * G.v[i];
    const long o = G.oa + (long)e * G.sa;
    st<T>(u, o, ld<T>(u, o) + (get(qs_[e]) + dv * F.rd[e]));
  }
}

// z = W^{-1} y with the packed symmetric inverse held in values 0..19
template <typename R, int NC>
__device__ __forceinline__ void symv(const Ch<R> (&w)[NC], const cxd (&y)[4], cxd (&z)[4]) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    cxd acc(0, 0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int s = symidx(i, j);
      acc += cxd(val(w, 2 * s), val(w, 2 * s + 1)) * y[j];
    }
    z[i] = acc;
  }
}

// Right-hand sides of the lines of one colour: one thread per (node, line),
// lines fastest, so the chunked records are written contiguously.
//   record p, slots 0..7: sigma_p, the residual of the four perpendicular
//                          edges of node p (p >= 1)
//   record p, slots 8..9: rho_p, the residual of the along-line edge p
// Lines of one colour are not coupled, so these are the residuals each line
// solve of the colour starts from.
// the records of node p of line `line` (at (pb, pc)) of job J
// the ten values of that record (node 0: rho only, the rest zero)
template <typename T, int A>
__device__ __forceinline__ void node_rhs_rec(const LineJob<T>& J, int pb, int pc, int p,
                                             Ch<ScrRec<T>> (&c)[schunks<ScrRec<T>>()]) {
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  const Level<T>& L = J.L;
  const typename cvec<T>::type* __restrict__ u = J.u;
  const typename cvec<T>::type* __restrict__ f = J.f;
  const long base = (long)pb * L.es[B] + (long)pc * L.es[C] + (long)p * L.es[A];
  {
    int q[3];
    q[A] = p; q[B] = pb; q[C] = pc;
    const cxd rho = ld<T>(f, A * L.nc + base) - apply_row<T, A>(L, u, q[A], q[B], q[C], base);
    put(c, 8, rho.re);
    put(c, 9, rho.im);
  }
  if (p >= 1) {
    int q[3];
    cxd g;
    const long ob0 = base - L.es[B], oc0 = base - L.es[C];
    q[A] = p; q[B] = pb - 1; q[C] = pc;
    g = ld<T>(f, B * L.nc + ob0) - apply_row<T, B>(L, u, q[B], q[C], q[A], ob0);
    put(c, 0, g.re); put(c, 1, g.im);
    q[B] = pb;
    g = ld<T>(f, B * L.nc + base) - apply_row<T, B>(L, u, q[B], q[C], q[A], base);
    put(c, 2, g.re); put(c, 3, g.im);
    q[C] = pc - 1;
    g = ld<T>(f, C * L.nc + oc0) - apply_row<T, C>(L, u, q[C], q[A], q[B], oc0);
    put(c, 4, g.re); put(c, 5, g.im);
    q[C] = pc;
    g = ld<T>(f, C * L.nc + base) - apply_row<T, C>(L, u, q[C], q[A], q[B], base);
    put(c, 6, g.re); put(c, 7, g.im);
  }
}

// the records of node p of a line into its chunked scratch records
template <typename T>
__device__ __forceinline__ void store_rhs_rec(const LineJob<T>& J, int line, int p, long nl,
                                              const Ch<ScrRec<T>> (&c)[schunks<ScrRec<T>>()]) {
  using R = ScrRec<T>;
  constexpr int SC = schunks<R>(), V = Ch<R>::V;
  Ch<R>* rec = reinterpret_cast<Ch<R>*>(J.scr) + line + (long)p * SC * nl;
  if (p >= 1) {
#pragma unroll
    for (int k = 0; k < SC; k++) rec[k * nl] = c[k];
  } else {
    rec[(8 / V) * nl] = c[8 / V];   // node 0: rho_0 only (values 8, 9: one chunk)
  }
}

template <typename T, int A>
__device__ __forceinline__ void node_rhs(const LineJob<T>& J, int line, int pb, int pc, int p, long nl) {
  using R = ScrRec<T>;
  constexpr int SC = schunks<R>(), V = Ch<R>::V;
  Ch<R> c[SC] = {};
  node_rhs_rec<T, A>(J, pb, pc, p, c);
  store_rhs_rec<T>(J, line, p, nl, c);
}

// Right-hand sides of consecutive nodes of one line, marching along it.
// The five rows of a node read 30 distinct values of u; 13 of them are read
// again by node p + 1 (the same (b, c) offsets one node further back), so the
// march carries those and loads 17 per node (with f and sigma-bar of the five
// rows), against about 40 loads per node when each node is built on its own.
// A column (k, db, dc) is component (A + k) % 3 at offsets db, dc along the
// perpendicular axes B, C; the window holds it at da = lo..hi along the line.
struct MarchCol {
  int k, db, dc, lo, hi;
};
constexpr MarchCol kMarch[] = {
    {0, -1, 0, -1, 0},  {0, 0, -1, -1, 0}, {0, 0, 0, -1, 0}, {0, 0, 1, -1, 0}, {0, 1, 0, -1, 0},
    {1, -1, -1, 0, 0},  {1, -1, 0, -1, 1}, {1, -1, 1, 0, 0}, {1, 0, -1, 0, 0}, {1, 0, 0, -1, 1},
    {1, 0, 1, 0, 0},    {2, -1, -1, 0, 0}, {2, -1, 0, 0, 0}, {2, 0, -1, -1, 1}, {2, 0, 0, -1, 1},
    {2, 1, -1, 0, 0},   {2, 1, 0, 0, 0}};
constexpr int kMarchCols = sizeof(kMarch) / sizeof(kMarch[0]);
constexpr int march_col(int k, int db, int dc) {
  for (int i = 0; i < kMarchCols; i++)
    if (kMarch[i].k == k && kMarch[i].db == db && kMarch[i].dc == dc) return i;
  return -1;
}
template <int I>
struct ColIdx {
  static constexpr int value = I;
};
template <typename F, int... I>
__device__ __forceinline__ void each_col(const F& f, std::integer_sequence<int, I...>) {
  (f(ColIdx<I>{}), ...);
}
template <typename F>
__device__ __forceinline__ void each_col(const F& f) {
  each_col(f, std::make_integer_sequence<int, kMarchCols>{});
}

struct MarchWin {
  cxd v[kMarchCols][3];   // [column][da + 1]
  // the value at (da, db, dc) from the node; a position the table lacks
  // does not compile
  template <int K, int DA, int DB, int DC>
  __device__ __forceinline__ cxd at() const {
    constexpr int c = march_col(K, DB, DC);
    static_assert(c >= 0 && DA >= kMarch[c].lo && DA <= kMarch[c].hi, "position outside the march window");
    return v[c][DA + 1];
  }
};

// value KK of row X (0: its own component, 1 and 2 the next ones
// cyclically) at offsets (dx, dy, dz) along the row's axes X, X+1, X+2, for
// the row of the edge at offsets (0, RB, RC) from the node; everything in
// the line frame (0 = A, 1 = B, 2 = C)
template <int X, int RB, int RC, int KK, int dx, int dy, int dz>
__device__ __forceinline__ cxd win_val(const MarchWin& w) {
  constexpr int Y = (X + 1) % 3, Z = (X + 2) % 3;
  constexpr int da = (X == 0 ? dx : 0) + (Y == 0 ? dy : 0) + (Z == 0 ? dz : 0);
  constexpr int db = RB + (X == 1 ? dx : 0) + (Y == 1 ? dy : 0) + (Z == 1 ? dz : 0);
  constexpr int dc = RC + (X == 2 ? dx : 0) + (Y == 2 ? dy : 0) + (Z == 2 ? dz : 0);
  return w.at<(X + KK) % 3, da, db, dc>();
}
// the thirteen values of that row in row_arith's order
template <int X, int RB, int RC>
__device__ __forceinline__ void win_row(const MarchWin& w, cxd (&v)[13]) {
  v[0] = win_val<X, RB, RC, 0, 0, 0, 0>(w);
  v[1] = win_val<X, RB, RC, 1, 1, 0, 0>(w);
  v[2] = win_val<X, RB, RC, 1, 0, 0, 0>(w);
  v[3] = win_val<X, RB, RC, 0, 0, 1, 0>(w);
  v[4] = win_val<X, RB, RC, 1, 1, -1, 0>(w);
  v[5] = win_val<X, RB, RC, 1, 0, -1, 0>(w);
  v[6] = win_val<X, RB, RC, 0, 0, -1, 0>(w);
  v[7] = win_val<X, RB, RC, 0, 0, 0, 1>(w);
  v[8] = win_val<X, RB, RC, 2, 1, 0, 0>(w);
  v[9] = win_val<X, RB, RC, 2, 0, 0, 0>(w);
  v[10] = win_val<X, RB, RC, 0, 0, 0, -1>(w);
  v[11] = win_val<X, RB, RC, 2, 1, 0, -1>(w);
  v[12] = win_val<X, RB, RC, 2, 0, 0, -1>(w);
}

// the records of nodes p0 .. p1-1 of the line at (pb, pc) of job J, handed
// to sink(p, c) in order; equal, value for value, to node_rhs_rec's (the
// rows' arithmetic is row_arith in both).  Every position read lies in the
// level: nodes -1 .. N along the line are read only where node_rhs_rec reads
// them (da = -1 from node 1 on).
template <typename T, int A, typename Sink>
__device__ __forceinline__ void line_rhs_march(const LineJob<T>& J, int pb, int pc, int p0, int p1,
                                               const Sink& sink) {
  using RS = ScrRec<T>;
  constexpr int B = (A + 1) % 3, C = (A + 2) % 3;
  const Level<T>& L = J.L;
  const typename cvec<T>::type* __restrict__ u = J.u;
  const typename cvec<T>::type* __restrict__ f = J.f;
  const long sa = L.es[A], sb = L.es[B], sc = L.es[C], nc = L.nc;
  // geometry of the perpendicular axes: fixed along the line
  const double iwb = L.IW[B][pb], iwbm = L.IW[B][pb - 1], hb = L.H[B][pb], wb = L.W[B][pb], wbm = L.W[B][pb - 1];
  const double iwc = L.IW[C][pc], iwcm = L.IW[C][pc - 1], hc = L.H[C][pc], wc = L.W[C][pc], wcm = L.W[C][pc - 1];
  // column ic (a compile-time index) at the node at flat index o, into the
  // window slot of da (compile-time too)
  auto load_col = [&](MarchWin& w, auto ic, auto dai, long o) {
    constexpr int ci = decltype(ic)::value, da = decltype(dai)::value;
    constexpr MarchCol m = kMarch[ci];
    w.v[ci][da + 1] = ld<T>(u, (long)((A + m.k) % 3) * nc + o + m.db * sb + m.dc * sc);
  };
  MarchWin w;
  long o = (long)pb * sb + (long)pc * sc + (long)p0 * sa;
  // the slots of da = -1 of a run from node 0 are never read (rows B and C
  // start at node 1, after a shift); they take node 0's values so that every
  // slot is written unconditionally and the window stays in registers
  const long om = p0 >= 1 ? o - sa : o;
  each_col([&](auto ic) {
    constexpr MarchCol m = kMarch[decltype(ic)::value];
    if constexpr (m.lo < 0) load_col(w, ic, ColIdx<-1>{}, om);
    load_col(w, ic, ColIdx<0>{}, o);
    if constexpr (m.hi > 0) load_col(w, ic, ColIdx<1>{}, o + sa);
  });
  double iwa_m = p0 >= 1 ? L.IW[A][p0 - 1] : 0.0;
  for (int p = p0; p < p1; p++, o += sa) {
    const double iwa = L.IW[A][p], wa = L.W[A][p], ha = L.H[A][p];
    Ch<RS> c[schunks<RS>()] = {};
    cxd v[13];
    {
      // rho_p: row A of the node, local indices (p, pb, pc)
      win_row<0, 0, 0>(w, v);
      const RowGeo g{iwa, iwb, iwbm, iwc, iwcm, hc, hb, wa, L.omu * (double)L.sbar[(long)A * nc + o]};
      const cxd rho = ld<T>(f, (long)A * nc + o) - row_arith(v, g);
      put(c, 8, rho.re);
      put(c, 9, rho.im);
    }
    if (p >= 1) {
      cxd g;
      // row B at pb - 1 and at pb: local indices (pb', pc, p)
      win_row<1, -1, 0>(w, v);
      g = ld<T>(f, (long)B * nc + o - sb) -
          row_arith(v, RowGeo{iwbm, iwc, iwcm, iwa, iwa_m, ha, hc, wbm, L.omu * (double)L.sbar[(long)B * nc + o - sb]});
      put(c, 0, g.re); put(c, 1, g.im);
      win_row<1, 0, 0>(w, v);
      g = ld<T>(f, (long)B * nc + o) -
          row_arith(v, RowGeo{iwb, iwc, iwcm, iwa, iwa_m, ha, hc, wb, L.omu * (double)L.sbar[(long)B * nc + o]});
      put(c, 2, g.re); put(c, 3, g.im);
      // row C at pc - 1 and at pc: local indices (pc', p, pb)
      win_row<2, 0, -1>(w, v);
      g = ld<T>(f, (long)C * nc + o - sc) -
          row_arith(v, RowGeo{iwcm, iwa, iwa_m, iwb, iwbm, hb, ha, wcm, L.omu * (double)L.sbar[(long)C * nc + o - sc]});
      put(c, 4, g.re); put(c, 5, g.im);
      win_row<2, 0, 0>(w, v);
      g = ld<T>(f, (long)C * nc + o) -
          row_arith(v, RowGeo{iwc, iwa, iwa_m, iwb, iwbm, hb, ha, wc, L.omu * (double)L.sbar[(long)C * nc + o]});
      put(c, 6, g.re); put(c, 7, g.im);
    }
    sink(p, c);
    if (p + 1 < p1) {
      // node p + 1: shift the window back one node and load its leading values
      each_col([&](auto ic) {
        constexpr int ci = decltype(ic)::value;
        constexpr MarchCol m = kMarch[ci];
        if constexpr (m.lo < 0) w.v[ci][0] = w.v[ci][1];
        if constexpr (m.hi > 0) {
          w.v[ci][1] = w.v[ci][2];
          load_col(w, ic, ColIdx<1>{}, o + 2 * sa);
        } else {
          load_col(w, ic, ColIdx<0>{}, o + sa);
        }
      });
    }
    iwa_m = iwa;
  }
}

template <typename T, int A>
__global__ void __launch_bounds__(256, EMMI_LINE_RHS_MIN_WAVES) k_line_rhs(LinePack<T> P) {
  const LineJob<T>& J = P.job[blockIdx.y];
  const long nl = (long)J.nlo * J.nhi;
  const int N = J.L.n[A];
  const long t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nl * N) return;
  const unsigned pu = (unsigned)t / (unsigned)nl;   // fewer than 2^32 nodes per launch
  const int p = (int)pu;
  const int line = (int)((unsigned)t - pu * (unsigned)nl);
  int pb, pc;
  line_of<T, A>(J, line, pb, pc);
  node_rhs<T, A>(J, line, pb, pc, p, nl);
}

// The same records built by marching (line_rhs_march): a thread per run of
// `run` consecutive nodes of a line, lines fastest (y and z lines, whose
// neighbours lie side by side along x)
template <typename T, int A>
__global__ void __launch_bounds__(256) k_line_rhs_run(LinePack<T> P, int run) {
  const LineJob<T>& J = P.job[blockIdx.y];
  const long nl = (long)J.nlo * J.nhi;
  const int N = J.L.n[A];
  const long nruns = (N + run - 1) / run;
  const long t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nl * nruns) return;
  const int r = (int)(t / nl), line = (int)(t - (long)r * nl);
  const int p0 = r * run, p1 = min(N, p0 + run);
  int pb, pc;
  line_of<T, A>(J, line, pb, pc);
  line_rhs_march<T, A>(J, pb, pc, p0, p1, [&](int p, const Ch<ScrRec<T>>(&c)[schunks<ScrRec<T>>()]) {
    store_rhs_rec<T>(J, line, p, nl, c);
  });
}

// Right-hand sides for lines along x (axis 0): consecutive lines lie in
// different x rows, so a thread per (node, line) with lines fastest would
// load one cache line per lane.  A block takes a tile of 16 nodes x 16 lines,
// computes with nodes fastest (loads contiguous along x) and transposes the
// records through LDS so the stores run along the lines.
constexpr int XT = 16;

template <typename T>
__global__ void __launch_bounds__(256, EMMI_LINE_RHS_MIN_WAVES) k_line_rhs_x(LinePack<T> P, int nlb) {
  constexpr int A = 0, B = 1, C = 2;
  __shared__ double tile[SREC][XT][XT + 1];
  const LineJob<T>& J = P.job[blockIdx.y];
  const Level<T>& L = J.L;
  const long nl = (long)J.nlo * J.nhi;
  const int N = L.n[A];
  // XCD-aware order: each XCD takes a contiguous range of tiles (a slab of
  // nodes over all lines), so the rows that neighbouring line tiles share
  // are fetched into one L2 (in dispatch order they land on different XCDs
  // and the level-0 launch fetched 2.4x its bytes)
  const long bl = xcd_block(blockIdx.x, gridDim.x);
  const int lb = (int)(bl % nlb), pbk = (int)(bl / nlb);
  const long l0 = (long)lb * XT;
  const int p0 = pbk * XT;
  if (l0 >= nl || p0 >= N) return;   // uniform over the block
  {
    const int pl = threadIdx.x % XT, ll = threadIdx.x / XT;
    const int p = p0 + pl;
    const long line = l0 + ll;
    double v[SREC];
#pragma unroll
    for (int k = 0; k < SREC; k++) v[k] = 0.0;
    if (p < N && line < nl) {
      int pb, pc;
      line_of<T, A>(J, (int)line, pb, pc);
      const typename cvec<T>::type* __restrict__ u = J.u;
      const typename cvec<T>::type* __restrict__ f = J.f;
      const long base = (long)pb * L.es[B] + (long)pc * L.es[C] + (long)p * L.es[A];
      int q[3];
      q[A] = p; q[B] = pb; q[C] = pc;
      const cxd rho = ld<T>(f, A * L.nc + base) - apply_row<T, A>(L, u, q[A], q[B], q[C], base);
      v[8] = rho.re;
      v[9] = rho.im;
      if (p >= 1) {
        cxd g;
        const long ob0 = base - L.es[B], oc0 = base - L.es[C];
        q[B] = pb - 1;
        g = ld<T>(f, B * L.nc + ob0) - apply_row<T, B>(L, u, q[B], q[C], q[A], ob0);
        v[0] = g.re; v[1] = g.im;
        q[B] = pb;
        g = ld<T>(f, B * L.nc + base) - apply_row<T, B>(L, u, q[B], q[C], q[A], base);
        v[2] = g.re; v[3] = g.im;
        q[C] = pc - 1;
        g = ld<T>(f, C * L.nc + oc0) - apply_row<T, C>(L, u, q[C], q[A], q[B], oc0);
        v[4] = g.re; v[5] = g.im;
        q[C] = pc;
        g = ld<T>(f, C * L.nc + base) - apply_row<T, C>(L, u, q[C], q[A], q[B], base);
        v[6] = g.re; v[7] = g.im;
      }
#pragma unroll
      for (int k = 0; k < SREC; k++) tile[k][pl][ll] = v[k];
    }
  }
  __syncthreads();
  {
    const int ll = threadIdx.x % XT, pl = threadIdx.x / XT;
    const int p = p0 + pl;
    const long line = l0 + ll;
    if (p < N && line < nl) {
      using R = ScrRec<T>;
      constexpr int SC = schunks<R>(), V = Ch<R>::V;
      Ch<R> c[SC] = {};
#pragma unroll
      for (int k = 0; k < SREC; k++) put(c, k, tile[k][pl][ll]);
      Ch<R>* rec = reinterpret_cast<Ch<R>*>(J.scr) + line + (long)p * SC * nl;
      if (p >= 1) {
#pragma unroll
        for (int k = 0; k < SC; k++) rec[k * nl] = c[k];
      } else {
        rec[(8 / V) * nl] = c[8 / V];
      }
    }
  }
}

// t = B z for the coupling B = diag(l) + v v^T rd of two neighbouring nodes
template <typename T, int A>
__device__ __forceinline__ void bmul(const LineGeo<T, A>& G, const cxd (&z)[4], double iWe, cxd rd, cxd (&t)[4]) {
  const double l[4] = {-G.Wb0 * G.Hc * iWe, -G.Wb1 * G.Hc * iWe, -G.Wc0 * G.Hb * iWe, -G.Wc1 * G.Hb * iWe};
  const cxd vz = z[0] * G.v[0] + z[1] * G.v[1] + z[2] * G.v[2] + z[3] * G.v[3];
#pragma unroll
  for (int i = 0; i < 4; i++) t[i] = z[i] * l[i] + (vz * rd) * G.v[i];
}

template <typename R, int NC>
__device__ __forceinline__ cxd cval(const Ch<R> (&c)[NC], int k) { return cxd(val(c, k), val(c, k + 1)); }

// Lines along x (axis 0) lie in different x rows, so a wave updating u node
// by node would write 64 scattered 8-byte values per store, and the L2
// evicts those partly written lines: the substitution along x wrote 2.7x the
// bytes of the other axes.  Instead it leaves each node's corrections in the
// node's line record (its z values are dead by then; values 0,1: the
// along-line edge, 2..9: the four node edges) and k_line_solve_x_apply adds
// them to u with the lanes running along x.
template <typename RS>
__device__ __forceinline__ void put_corr(Ch<RS>* __restrict__ rec, long stride, cxd ea, const cxd (&ap)[4]) {
  constexpr int SC = schunks<RS>();
  Ch<RS> c[SC] = {};
  put(c, 0, ea.re);
  put(c, 1, ea.im);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    put(c, 2 + 2 * i, ap[i].re);
    put(c, 3 + 2 * i, ap[i].im);
  }
#pragma unroll
  for (int k = 0; k < SC; k++) rec[k * stride] = c[k];
}

// u += the corrections left in the line records by k_line_solve<T, 0>: a
// block takes 16 nodes x 16 lines, reads the records with lines fastest and
// updates u through an LDS transpose with nodes fastest.  Node m = N/2 (its
// node edges are updated by the substitution itself) and node 0 hold no
// correction; the along-line edge of node p is p below m and p-1 above.
template <typename T>
__global__ void __launch_bounds__(XT * XT) k_line_solve_x_apply(LinePack<T> P, int nlb) {
  constexpr int A = 0;
  using RS = ScrRec<T>;
  constexpr int SC = schunks<RS>();
  __shared__ double tile[SREC][XT][XT + 1];
  const LineJob<T>& J = P.job[blockIdx.y];
  const Level<T>& L = J.L;
  const long nl = (long)J.nlo * J.nhi;
  const int N = L.n[A], m = N / 2;
  const int lb = blockIdx.x % nlb, pbk = blockIdx.x / nlb;
  const long l0 = (long)lb * XT;
  const int p0 = pbk * XT;
  if (l0 >= nl || p0 >= N) return;   // uniform over the block
  {
    const int ll = threadIdx.x % XT, pl = threadIdx.x / XT;
    const int p = p0 + pl;
    const long line = l0 + ll;
    if (p < N && line < nl && p >= 1 && p != m) {
      const Ch<RS>* rec = reinterpret_cast<const Ch<RS>*>(J.scr) + line + (long)p * SC * nl;
      Ch<RS> c[SC];
#pragma unroll
      for (int k = 0; k < SC; k++) c[k] = rec[k * nl];
#pragma unroll
      for (int k = 0; k < SREC; k++) tile[k][pl][ll] = val(c, k);
    }
  }
  __syncthreads();
  {
    const int pl = threadIdx.x % XT, ll = threadIdx.x / XT;
    const int p = p0 + pl;
    const long line = l0 + ll;
    if (p < N && line < nl && p >= 1 && p != m) {
      int pb, pc;
      line_of<T, A>(J, (int)line, pb, pc);
      const LineGeo<T, A> G(L, pb, pc);
      typename cvec<T>::type* __restrict__ u = J.u;
      const long o = (long)p * G.sa;
      const long ea = G.oa + (long)(p < m ? p : p - 1) * G.sa;
      st<T>(u, ea, ld<T>(u, ea) + cxd(tile[0][pl][ll], tile[1][pl][ll]));
      const long offs[4] = {G.ob0, G.ob1, G.oc0, G.oc1};
#pragma unroll
      for (int i = 0; i < 4; i++)
        st<T>(u, offs[i] + o, ld<T>(u, offs[i] + o) + cxd(tile[2 + 2 * i][pl][ll], tile[3 + 2 * i][pl][ll]));
    }
  }
}

// Block substitution of the lines of one colour on the twisted factors
// (k_line_factor).  A block of 128 threads takes 64 lines: wave 0 ("top")
// eliminates nodes 1..m-1 downwards and wave 1 ("bottom") nodes M..m+1
// upwards at the same time,
//   top     z_p = W_p^-1 (g_p - B_p z_{p-1})
//   bottom  z_p = U_p^-1 (g_p - B_{p+1} z_{p+1}),
// the waves exchange B_m z_{m-1} and B_{m+1} z_{m+1} through LDS, both solve
// the meeting node x_m = S_m^-1 (g_m - B_m z_{m-1} - B_{m+1} z_{m+1}), and
// each substitutes back over its half,
//   top     x_p = z_p - W_p^-1 B_{p+1} x_{p+1}     p = m-1..1
//   bottom  x_p = z_p - U_p^-1 B_p x_{p-1}         p = m+1..M,
// updating the along-line edges dE_a(e) = q_e + v.(x_{e+1} - x_e) / d_e
// (top: edges 0..m-1, bottom: m..N-1; q_e = rho_e / d_e) and the node
// edges.  g_p = sigma_p + v (q_{p-1} - q_p).  Each wave keeps the chunks of
// its next D nodes in flight in a register ring; ring loads past the end of
// a half are clamped into the line and never used.  The forward passes
// overwrite sigma_p with z_p.
template <typename T, int A>
__global__ void __launch_bounds__(128) k_line_solve(LinePack<T> P) {
  using E = typename cvec<T>::type;
  using RF = FacRec<T>;
  using RS = ScrRec<T>;
  constexpr int FC = fchunks<RF>(), SC = schunks<RS>(), VF = Ch<RF>::V, VS = Ch<RS>::V, D = ring_depth<T>();
  constexpr int ZC = 8 / VS;    // chunks holding values 0..7 (sigma, then z)
  constexpr int CD = 20 / VF;   // factor chunk holding 1/d (values 20, 21)
  constexpr int CQ = 8 / VS;    // record chunk holding rho (values 8, 9)
  __shared__ double xch[2][8][64];
  const LineJob<T>& J = P.job[blockIdx.y];
  const Level<T>& L = J.L;
  const int N = L.n[A];
  if (N < 2) return;   // uniform over the block: no interior node
  const int M = N - 1, m = N / 2;
  const int lane = threadIdx.x & 63, side = threadIdx.x >> 6;
  const int line = blockIdx.x * 64 + lane;
  int pb = 0, pc = 0;
  const bool live = line_of<T, A>(J, line, pb, pc);
  LineGeo<T, A> G;
  if (live) G.init(L, pb, pc);
  const long stride = (long)J.nlo * J.nhi;
  const Ch<RF>* __restrict__ fac = reinterpret_cast<const Ch<RF>*>(J.fac) + line;
  Ch<RS>* __restrict__ scr = reinterpret_cast<Ch<RS>*>(J.scr) + line;
  E* __restrict__ u = J.u;
  const double* __restrict__ iw = L.IW[A];
  const long fstep = (long)FC * stride, sstep = (long)SC * stride;
  const long off_u[5] = {G.oa, G.ob0, G.ob1, G.oc0, G.oc1};

  // D nodes ahead are in flight: the load for node p + D goes into the slot
  // node p - 1 used, so a slot is never copied before it is refilled
  constexpr int R = D + 1;
  Ch<RF> rf[R][FC];
  Ch<RS> rs[R][SC];
  double riw[R];
  cxd tm[4];   // this half's coupling term at the meeting node
#pragma unroll
  for (int i = 0; i < 4; i++) tm[i] = cxd(0, 0);

  // ---------------- forward elimination ----------------
  if (live && side == 0) {
    const int last = m - 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = max(min(1 + d, last), 1);
      load_node(rf[d], fac + pn * fstep, stride);
      load_node(rs[d], scr + pn * sstep, stride);
      riw[d] = iw[pn - 1];
    }
    const Ch<RF> f0 = fac[CD * stride];
    const Ch<RS> s0 = scr[CQ * stride];
    cxd rd_prev((double)f0.v[20 % VF], (double)f0.v[21 % VF]);
    cxd q_prev = cxd((double)s0.v[8 % VS], (double)s0.v[9 % VS]) * rd_prev;
    cxd z[4];
    auto step = [&](const int p, const int d) {
      const Ch<RF>(&cf)[FC] = rf[d];
      const Ch<RS>(&cr)[SC] = rs[d];
      const double iWe = (double)riw[d];
      {
        const int pn = min(p + D, last), dn = (d + D) % R;
        load_node(rf[dn], fac + pn * fstep, stride);
        load_node(rs[dn], scr + pn * sstep, stride);
        riw[dn] = iw[pn - 1];
      }
      const cxd rd = cval(cf, 20);
      const cxd q = cval(cr, 8) * rd;
      const cxd dq = q_prev - q;
      cxd g[4];
#pragma unroll
      for (int i = 0; i < 4; i++) g[i] = cval(cr, 2 * i) + dq * G.v[i];
      if (p >= 2) {
        cxd t[4];
        bmul(G, z, iWe, rd_prev, t);
#pragma unroll
        for (int i = 0; i < 4; i++) g[i] -= t[i];
      }
      symv(cf, g, z);
      Ch<RS> zc[SC] = {};
#pragma unroll
      for (int i = 0; i < 4; i++) { put(zc, 2 * i, z[i].re); put(zc, 2 * i + 1, z[i].im); }
      Ch<RS>* sp = scr + p * sstep;
#pragma unroll
      for (int k = 0; k < ZC; k++) sp[k * stride] = zc[k];
      rd_prev = rd;
      q_prev = q;
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = 1;
    for (; p0 + (R - 1) <= last; p0 += R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 + d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 + d <= last) step(p0 + d, d);
    if (m >= 2) bmul(G, z, (double)iw[m - 1], rd_prev, tm);   // B_m z_{m-1}
  } else if (live) {
    // ring entries also carry rho_{p-1} and 1/d_{p-1} of the node above
    Ch<RF> rdn[R];
    Ch<RS> rqn[R];
    const int last = m + 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = min(max(M - d, last), M);
      load_node(rf[d], fac + pn * fstep, stride);
      load_node(rs[d], scr + pn * sstep, stride);
      rdn[d] = fac[(pn - 1) * fstep + CD * stride];
      rqn[d] = scr[(pn - 1) * sstep + CQ * stride];
      riw[d] = iw[pn];
    }
    cxd zn[4], rd_last(0, 0);
    auto step = [&](const int p, const int d) {
      const Ch<RF>(&cf)[FC] = rf[d];
      const Ch<RS>(&cr)[SC] = rs[d];
      const Ch<RF>& cdm = rdn[d];
      const Ch<RS>& cqm = rqn[d];
      const double iWe = (double)riw[d];
      {
        const int pn = max(p - D, last), dn = (d + D) % R;
        load_node(rf[dn], fac + pn * fstep, stride);
        load_node(rs[dn], scr + pn * sstep, stride);
        rdn[dn] = fac[(pn - 1) * fstep + CD * stride];
        rqn[dn] = scr[(pn - 1) * sstep + CQ * stride];
        riw[dn] = iw[pn];
      }
      const cxd rd = cval(cf, 20);
      const cxd q = cval(cr, 8) * rd;
      const cxd rdm((double)cdm.v[20 % VF], (double)cdm.v[21 % VF]);
      const cxd qm = cxd((double)cqm.v[8 % VS], (double)cqm.v[9 % VS]) * rdm;
      const cxd dq = qm - q;
      cxd g[4];
#pragma unroll
      for (int i = 0; i < 4; i++) g[i] = cval(cr, 2 * i) + dq * G.v[i];
      if (p <= M - 1) {
        cxd t[4];
        bmul(G, zn, iWe, rd, t);
#pragma unroll
        for (int i = 0; i < 4; i++) g[i] -= t[i];
      }
      symv(cf, g, zn);
      Ch<RS> zc[SC] = {};
#pragma unroll
      for (int i = 0; i < 4; i++) { put(zc, 2 * i, zn[i].re); put(zc, 2 * i + 1, zn[i].im); }
      Ch<RS>* sp = scr + p * sstep;
#pragma unroll
      for (int k = 0; k < ZC; k++) sp[k * stride] = zc[k];
      rd_last = rdm;   // 1/d_{p-1}: after the loop, 1/d_m
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = M;
    for (; p0 - (R - 1) >= last; p0 -= R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 - d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 - d >= last) step(p0 - d, d);
    if (m + 1 <= M) bmul(G, zn, (double)iw[m], rd_last, tm);   // B_{m+1} z_{m+1}
  }

  // ---------------- meeting node ----------------
#pragma unroll
  for (int i = 0; i < 4; i++) {
    xch[side][2 * i][lane] = tm[i].re;
    xch[side][2 * i + 1][lane] = tm[i].im;
  }
  __syncthreads();
  if (!live) return;
  cxd xm[4], qmid, rdmid;
  {
    Ch<RF> cf[FC];
          Ch<RS> cr[SC];
    load_node(cf, fac + m * fstep, stride);
    load_node(cr, scr + m * sstep, stride);
    const Ch<RF> cdm = fac[(m - 1) * fstep + CD * stride];
    const Ch<RS> cqm = scr[(m - 1) * sstep + CQ * stride];
    rdmid = cval(cf, 20);
    qmid = cval(cr, 8) * rdmid;
    const cxd rdm((double)cdm.v[20 % VF], (double)cdm.v[21 % VF]);
    const cxd dq = cxd((double)cqm.v[8 % VS], (double)cqm.v[9 % VS]) * rdm - qmid;
    cxd g[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const cxd ttop(xch[0][2 * i][lane], xch[0][2 * i + 1][lane]);
      const cxd tbot(xch[1][2 * i][lane], xch[1][2 * i + 1][lane]);
      g[i] = ((cval(cr, 2 * i) + dq * G.v[i]) - ttop) - tbot;
    }
    symv(cf, g, xm);
  }

  // ---------------- back substitution ----------------
  E ru[R][5];
  if (side == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const long o = off_u[i + 1] + (long)m * G.sa;
      st<T>(u, o, ld<T>(u, o) + xm[i]);
    }
    const int last = m - 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = max(last - d, 1);
      load_node(rf[d], fac + pn * fstep, stride);
      load_node(rs[d], scr + pn * sstep, stride);
      riw[d] = iw[pn];
      if constexpr (A != 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) ru[d][k] = u[off_u[k] + (long)pn * G.sa];
      }
    }
    cxd an[4];   // x_{p+1}
#pragma unroll
    for (int i = 0; i < 4; i++) an[i] = xm[i];
    auto step = [&](const int p, const int d) {
      const Ch<RF>(&cf)[FC] = rf[d];
      const Ch<RS>(&cr)[SC] = rs[d];
      const E(&cu)[5] = ru[d];
      const double iWe = (double)riw[d];
      {
        const int pn = max(p - D, 1), dn = (d + D) % R;
        load_node(rf[dn], fac + pn * fstep, stride);
        load_node(rs[dn], scr + pn * sstep, stride);
        riw[dn] = iw[pn];
        if constexpr (A != 0) {
#pragma unroll
          for (int k = 0; k < 5; k++) ru[dn][k] = u[off_u[k] + (long)pn * G.sa];
        }
      }
      const cxd rd = cval(cf, 20);
      const cxd q = cval(cr, 8) * rd;
      cxd t[4], wt[4], ap[4];
      bmul(G, an, iWe, rd, t);
      symv(cf, t, wt);
#pragma unroll
      for (int i = 0; i < 4; i++) ap[i] = cval(cr, 2 * i) - wt[i];
      cxd dv(0, 0);
#pragma unroll
      for (int i = 0; i < 4; i++) dv += (an[i] - ap[i]) * G.v[i];
      if constexpr (A == 0) {
        put_corr(scr + p * sstep, stride, q + dv * rd, ap);
      } else {
        const long o = (long)p * G.sa;
        st<T>(u, G.oa + o, cxd((double)cu[0].x, (double)cu[0].y) + (q + dv * rd));
#pragma unroll
        for (int i = 0; i < 4; i++)
          st<T>(u, off_u[i + 1] + o, cxd((double)cu[i + 1].x, (double)cu[i + 1].y) + ap[i]);
      }
#pragma unroll
      for (int i = 0; i < 4; i++) an[i] = ap[i];
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = last;
    for (; p0 - (R - 1) >= 1; p0 -= R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 - d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 - d >= 1) step(p0 - d, d);
    // along-line edge 0, between the boundary node and node 1
    const Ch<RF> f0 = fac[CD * stride];
    const Ch<RS> s0 = scr[CQ * stride];
    const cxd rd0((double)f0.v[20 % VF], (double)f0.v[21 % VF]);
    const cxd q0 = cxd((double)s0.v[8 % VS], (double)s0.v[9 % VS]) * rd0;
    cxd dv(0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) dv += an[i] * G.v[i];
    st<T>(u, G.oa, ld<T>(u, G.oa) + (q0 + dv * rd0));
  } else {
    const int last = m + 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = min(last + d, M);
      load_node(rf[d], fac + pn * fstep, stride);
      load_node(rs[d], scr + pn * sstep, stride);
      riw[d] = iw[pn - 1];
      if constexpr (A != 0) {
        ru[d][0] = u[G.oa + (long)(pn - 1) * G.sa];
#pragma unroll
        for (int k = 1; k < 5; k++) ru[d][k] = u[off_u[k] + (long)pn * G.sa];
      }
    }
    cxd xp[4];   // x_{p-1}
#pragma unroll
    for (int i = 0; i < 4; i++) xp[i] = xm[i];
    cxd rdm = rdmid, qm = qmid;   // 1/d and q of edge p-1
    auto step = [&](const int p, const int d) {
      const Ch<RF>(&cf)[FC] = rf[d];
      const Ch<RS>(&cr)[SC] = rs[d];
      const E(&cu)[5] = ru[d];
      const double iWe = (double)riw[d];
      {
        const int pn = min(p + D, M), dn = (d + D) % R;
        load_node(rf[dn], fac + pn * fstep, stride);
        load_node(rs[dn], scr + pn * sstep, stride);
        riw[dn] = iw[pn - 1];
        if constexpr (A != 0) {
          ru[dn][0] = u[G.oa + (long)(pn - 1) * G.sa];
#pragma unroll
          for (int k = 1; k < 5; k++) ru[dn][k] = u[off_u[k] + (long)pn * G.sa];
        }
      }
      const cxd rd = cval(cf, 20);
      const cxd q = cval(cr, 8) * rd;
      cxd t[4], wt[4], ap[4];
      bmul(G, xp, iWe, rdm, t);
      symv(cf, t, wt);
#pragma unroll
      for (int i = 0; i < 4; i++) ap[i] = cval(cr, 2 * i) - wt[i];
      cxd dv(0, 0);
#pragma unroll
      for (int i = 0; i < 4; i++) dv += (ap[i] - xp[i]) * G.v[i];
      if constexpr (A == 0) {
        put_corr(scr + p * sstep, stride, qm + dv * rdm, ap);
      } else {
        const long o = (long)p * G.sa;
        st<T>(u, G.oa + o - G.sa, cxd((double)cu[0].x, (double)cu[0].y) + (qm + dv * rdm));
#pragma unroll
        for (int i = 0; i < 4; i++)
          st<T>(u, off_u[i + 1] + o, cxd((double)cu[i + 1].x, (double)cu[i + 1].y) + ap[i]);
      }
#pragma unroll
      for (int i = 0; i < 4; i++) xp[i] = ap[i];
      rdm = rd;
      qm = q;
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = last;
    for (; p0 + (R - 1) <= M; p0 += R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 + d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 + d <= M) step(p0 + d, d);
    // along-line edge M = N-1, between node M and the boundary node
    cxd dv(0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) dv += (cxd(0, 0) - xp[i]) * G.v[i];
    const long o = G.oa + (long)M * G.sa;
    st<T>(u, o, ld<T>(u, o) + (qm + dv * rdm));
  }
}

// ---------------------------------------------------------------------------
// quad-lane substitution for launches with few lines
// ---------------------------------------------------------------------------
// Lanes 4k..4k+3 of a wave hold the components 0..3 of one line's node
// vector and the matching rows of W^-1, so the serial chain of a node is a
// quarter of the arithmetic plus quad exchanges (DPP): on levels with few
// lines the launch is bound by that chain, not by memory.
template <int CTRL>
__device__ __forceinline__ double qdpp(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// sum over the quad (every lane gets the same value)
__device__ __forceinline__ cxd qsum(cxd a) {
  a += cxd(qdpp<0xB1>(a.re), qdpp<0xB1>(a.im));   // quad_perm [1,0,3,2]
  a += cxd(qdpp<0x4E>(a.re), qdpp<0x4E>(a.im));   // quad_perm [2,3,0,1]
  return a;
}
// component J of the quad's vector
template <int J>
__device__ __forceinline__ cxd qget(cxd a) {
  return cxd(qdpp<J * 0x55>(a.re), qdpp<J * 0x55>(a.im));
}
// row q of W^-1 (w) times the quad's vector y
__device__ __forceinline__ cxd qsymv(const cxd (&w)[4], cxd y) {
  const cxd y0 = qget<0>(y), y1 = qget<1>(y), y2 = qget<2>(y), y3 = qget<3>(y);
  cxd acc(0, 0);
  acc += w[0] * y0;
  acc += w[1] * y1;
  acc += w[2] * y2;
  acc += w[3] * y3;
  return acc;
}

// one node of one lane's records: row q of W^-1, 1/d, component q of the line
// record (sigma or z) and rho
template <typename RS>
struct QNode {
  double2 w[4];
  double2 rd;
  typename cvec<RS>::type s, rho;
  typename cvec<RS>::type pad;
};

template <typename RS>
__device__ __forceinline__ void qload_f(QNode<RS>& n, const Ch<double>* __restrict__ fnode, long stride, const int (&wch)[4]) {
#pragma unroll
  for (int j = 0; j < 4; j++) n.w[j] = *reinterpret_cast<const double2*>(&fnode[wch[j] * stride]);
  n.rd = *reinterpret_cast<const double2*>(&fnode[10 * stride]);
}

// fused launches with lines of NT nodes keep the block's line records in LDS
// when they take at most EMMI_QUAD_LDS_KB KiB (build-time tunable)
#ifndef EMMI_QUAD_LDS_KB
#define EMMI_QUAD_LDS_KB 80
#endif
// Block shape of the fused quad-lane launches (lines per block, threads;
// build-time tunables).  The FP32 128-node instance holds 314 registers per
// lane (one wave per SIMD) and, with 16 lines, 80 KiB of records plus the
// 2 KiB meeting exchange in LDS: one 256-thread block per CU, whose two
// further waves only help with the right-hand sides.  15 lines in
// 128-thread blocks (75 + 2 KiB) let two blocks and four substituting waves
// share a CU, and measured slower at 128^3 (three right-hand sides:
// 257 -> 305 us per y/z launch; one: 78 -> 128 us): the right-hand sides,
// built by two waves instead of four, cost more than the second block gains.
// Single-wave blocks of 8 lines from line length EMMI_LINE_WAVE_MIN_NT on
// (off by default) fit four to eight to a CU, but the one wave builds all
// its records and runs both halves' loops under divergent masks: measured at
// 128^3 they won on a few shapes (FP64 128-node lines, one right-hand side:
// 143 -> 125 us; FP32 64-node x lines, three: 77 -> 70 us) and lost on most
// (FP32 128-node y/z lines, three: 257 -> 357 us; FP32 64-node lines, one:
// 36 -> 62 us).
#ifndef EMMI_LINE_WAVE_MIN_NT
#define EMMI_LINE_WAVE_MIN_NT 1000000
#endif
#ifndef EMMI_LINE_LPB_F32_128
#define EMMI_LINE_LPB_F32_128 16
#endif
#ifndef EMMI_LINE_BT_F32_128
#define EMMI_LINE_BT_F32_128 EMMI_LINE_FUSE_THREADS_LONG
#endif
template <typename T>
__host__ __device__ constexpr int quad_lpb(int nt) {
  return nt >= EMMI_LINE_WAVE_MIN_NT ? 8 : (sizeof(T) == 4 && nt == 128 ? EMMI_LINE_LPB_F32_128 : 16);
}
template <typename T>
__host__ __device__ constexpr int quad_bt(int nt) {
  return nt >= EMMI_LINE_WAVE_MIN_NT                 ? 64
         : sizeof(T) == 4 && nt == 128               ? EMMI_LINE_BT_F32_128
         : nt >= 64                                  ? EMMI_LINE_FUSE_THREADS_LONG
                                                     : EMMI_LINE_FUSE_THREADS;
}

// nodes per thread whose right-hand sides a fused launch builds together
#ifndef EMMI_LINE_RHS_UNROLL
#define EMMI_LINE_RHS_UNROLL 4
#endif
// 1: a fused launch of lines of 64 nodes or more builds its right-hand sides
// by marching along the lines (line_rhs_march; shorter lines leave runs of a
// node or two, with nothing to carry); 0: node by node.  Lines along x keep
// the node-by-node loop: there a wave's lanes take runs 8 nodes apart in 16
// rows, twice the cache lines per load of four neighbouring nodes (and in
// FP64 the record stores scatter alike).  Measured at 128^3 per launch of a
// colour with one right-hand side: along y and z the march takes 143 -> 137
// and 145 -> 139 us (FP64, 128 nodes), along x 156 -> 296 us (FP64) and 100
// -> 107 us (FP32).
#ifndef EMMI_LINE_RHS_MARCH
#define EMMI_LINE_RHS_MARCH 1
#endif
// nodes per thread of the separate right-hand-side launch that marches
#ifndef EMMI_LINE_RHS_RUN
#define EMMI_LINE_RHS_RUN 8
#endif
// fused launches along x build their node-by-node records with a wave on
// this many neighbouring nodes of each of 64 / EMMI_LINE_RHS_XGROUP lines
// (0: 4 nodes of 16 lines, as along y and z)
#ifndef EMMI_LINE_RHS_XGROUP
#define EMMI_LINE_RHS_XGROUP 16
#endif
template <typename T, int A>
__host__ __device__ constexpr bool rhs_march(int nt) {
  return EMMI_LINE_RHS_MARCH != 0 && nt >= 64 && A != 0;
}

template <typename T>
__host__ __device__ constexpr bool quad_lds_records(int nt, int lpb = 16) {
  return nt > 0 && lpb * nt * SREC * (int)sizeof(ScrRec<T>) <= EMMI_QUAD_LDS_KB * 1024;
}

// NT > 0: every job of the launch has lines of NT nodes; up to 64 nodes the
// passes are unrolled completely (longer lines keep the loop: the code would
// not fit the instruction cache).  The rolled loop's guard-free groups keep
// the ring in place (see k_line_solve); the unrolled passes also drop the
// loop overhead of these short chains.
//
// With RHS the block first computes the records of its own LPB lines (the
// work of k_line_rhs, for launches too small to pay for a launch of their
// own) with all its BT threads, and then substitutes.  BT >= 128: a wave per
// line half side (top halves of the block's LPB <= 16 lines in wave 0,
// bottom halves in wave 1; four lanes per line half, lanes beyond 4 LPB
// idle), further waves only help with the right-hand sides, the halves meet
// through LDS.  BT = 64 (LPB <= 8): one wave holds both halves (lanes 0-31
// the top halves, 32-63 the bottom ones, each side under its own exec
// mask) and the halves meet through a cross-lane shuffle.
template <typename T, int A, int D, bool RHS, int BT, int NT, int LPB = 16>
__global__ void __launch_bounds__(BT) k_line_solve_q(LinePack<T> P) {
  constexpr bool ONEW = BT == 64;   // both sides in one wave
  static_assert(ONEW ? LPB <= 8 : (LPB <= 16 && BT >= 128), "lines per block");
  using E = typename cvec<T>::type;
  using RS = ScrRec<T>;
  using PS = typename cvec<RS>::type;
  constexpr int FC = fchunks<double>(), SC = schunks<RS>(), VS = Ch<RS>::V;
  __shared__ double xch[ONEW ? 1 : 2][2][ONEW ? 1 : 64];
  const LineJob<T>& J = P.job[blockIdx.y];
  const Level<T>& L = J.L;
  const int N = NT > 0 ? NT : L.n[A];
  constexpr int UNR = NT > 0 && NT <= 64 ? 64 : 1;   // the passes' unroll count
  if (N < 2) return;   // uniform over the block: no interior node
  const int M = N - 1, m = N / 2;
  // LR: the block's line records live in LDS (lrec[node][line][value]) and
  // never go to memory, which takes a store and a dependent load off each end
  // of the substitution
  constexpr bool LR = RHS && quad_lds_records<T>(NT, LPB);
  __shared__ RS lrec[LR ? LPB * NT * SREC : 1];
  if (RHS) {
    const long nl = (long)J.nlo * J.nhi;
    if constexpr (rhs_march<T, A>(NT)) {
      // each line's nodes in CH consecutive runs, a run per thread, lines
      // fastest (a wave's loads cover LPB neighbouring lines), each run
      // marching along its line
      constexpr int CH = BT / LPB;
      const int ll = (int)threadIdx.x % LPB, ch = (int)threadIdx.x / LPB;
      const int run = (N + CH - 1) / CH;
      const int p0 = ch * run, p1 = min(N, p0 + run);
      const int ln = blockIdx.x * LPB + ll;
      int b2, c2;
      if (ch < CH && p0 < p1 && line_of<T, A>(J, ln, b2, c2)) {
        line_rhs_march<T, A>(J, b2, c2, p0, p1, [&](int p, const Ch<RS>(&c)[SC]) {
          if constexpr (LR) {
#pragma unroll
            for (int v = 0; v < SREC; v++) lrec[(p * LPB + ll) * SREC + v] = (RS)val(c, v);
          } else {
            store_rhs_rec<T>(J, ln, p, nl, c);
          }
        });
      }
    } else {
      // unrolled: several nodes' stencil loads in flight per thread (the
      // substitution sets the kernel's register count, this loop fits in it)
#pragma unroll EMMI_LINE_RHS_UNROLL
      for (int idx = threadIdx.x; idx < LPB * N; idx += blockDim.x) {
        // along x (lines in separate rows) a wave takes XN neighbouring
        // nodes of 64 / XN lines instead of 4 nodes of 16 lines
        constexpr int XN = EMMI_LINE_RHS_XGROUP;
        constexpr bool XG = A == 0 && NT > 0 && XN > 0 && NT % XN == 0;
        const int ll = XG ? (idx / XN) % LPB : idx % LPB;
        const int p = XG ? idx / (XN * LPB) * XN + idx % XN : idx / LPB;
        const int ln = blockIdx.x * LPB + ll;
        int b2, c2;
        if (line_of<T, A>(J, ln, b2, c2)) {
          if constexpr (LR) {
            Ch<RS> c[SC] = {};
            node_rhs_rec<T, A>(J, b2, c2, p, c);
#pragma unroll
            for (int v = 0; v < SREC; v++) lrec[(p * LPB + ll) * SREC + v] = (RS)val(c, v);
          } else {
            node_rhs<T, A>(J, ln, b2, c2, p, nl);
          }
        }
      }
    }
    __syncthreads();   // the block's records are complete (workgroup scope)
  }
  const int lane = threadIdx.x & 63, side = ONEW ? lane >> 5 : (int)threadIdx.x >> 6;
  const int q = lane & 3;
  const int lloc = ONEW ? (lane & 31) >> 2 : lane >> 2;   // the lane's line within the block
  const int line = blockIdx.x * LPB + lloc;
  int pb = 0, pc = 0;
  const bool live = side < 2 && lloc < LPB && line_of<T, A>(J, line, pb, pc);
  LineGeo<T, A> G;
  if (live) G.init(L, pb, pc);
  const long stride = (long)J.nlo * J.nhi;
  const Ch<double>* __restrict__ fac = reinterpret_cast<const Ch<double>*>(J.fac) + line;
  Ch<RS>* __restrict__ scr = reinterpret_cast<Ch<RS>*>(J.scr) + line;
  E* __restrict__ u = J.u;
  const double* __restrict__ iw = L.IW[A];
  const long fstep = (long)FC * stride, sstep = (long)SC * stride;
  // this lane's component: v_q, the coupling coefficient (l_q = cq / W_a), the
  // chunks of row q of W^-1, and where component q of a line record sits
  const double vq = q == 0 ? G.Hc : q == 1 ? -G.Hc : q == 2 ? G.Hb : -G.Hb;
  const double cq = q == 0 ? -G.Wb0 * G.Hc : q == 1 ? -G.Wb1 * G.Hc : q == 2 ? -G.Wc0 * G.Hb : -G.Wc1 * G.Hb;
  const int wch[4] = {symidx(q, 0), symidx(q, 1), symidx(q, 2), symidx(q, 3)};
  const long soff = (long)((2 * q) / VS) * stride;
  const int sel = (2 * q) % VS;
  const long off_q = q == 0 ? G.ob0 : q == 1 ? G.ob1 : q == 2 ? G.oc0 : G.oc1;
  // component q of sigma (z once written) and rho of node p of this lane's line
  auto rec_s = [&](int p) -> PS {
    if constexpr (LR) return *reinterpret_cast<const PS*>(&lrec[(p * LPB + lloc) * SREC + 2 * q]);
    else return *reinterpret_cast<const PS*>(reinterpret_cast<const RS*>(scr + p * sstep + soff) + sel);
  };
  auto rec_rho = [&](int p) -> PS {
    if constexpr (LR) return *reinterpret_cast<const PS*>(&lrec[(p * LPB + lloc) * SREC + 8]);
    else return *reinterpret_cast<const PS*>(reinterpret_cast<const RS*>(scr + p * sstep + (8 / VS) * stride) + 8 % VS);
  };
  auto put_z = [&](int p, PS z) {
    if constexpr (LR) *reinterpret_cast<PS*>(&lrec[(p * LPB + lloc) * SREC + 2 * q]) = z;
    else *reinterpret_cast<PS*>(reinterpret_cast<RS*>(scr + p * sstep + soff) + sel) = z;
  };
  auto qnode = [&](QNode<RS>& n, int p) {
    qload_f(n, fac + p * fstep, stride, wch);
    n.s = rec_s(p);
    n.rho = rec_rho(p);
  };

  // D nodes ahead are in flight: the load for node p + D goes into the slot
  // node p - 1 used, so a slot is never copied before it is refilled
  constexpr int R = D + 1;
  QNode<RS> ring[R];
  double riw[R];
  cxd tm(0, 0);   // component q of this half's coupling term at the meeting node
  auto cw = [](const QNode<RS>& n, cxd (&w)[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++) w[j] = cxd(n.w[j].x, n.w[j].y);
  };

  // ---------------- forward elimination ----------------
  if (live && side == 0) {
    const int last = m - 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = max(min(1 + d, last), 1);
      qnode(ring[d], pn);
      riw[d] = iw[pn - 1];
    }
    const cxd rd0(fac[10 * stride].v[0], fac[10 * stride].v[1]);
    const PS r0 = rec_rho(0);
    cxd rd_prev = rd0;
    cxd q_prev = cxd((double)r0.x, (double)r0.y) * rd0;
    cxd z(0, 0);
    auto step = [&](const int p, const int d) {
      const QNode<RS>& nd = ring[d];
      const double iWe = (double)riw[d];
      {
        const int pn = min(p + D, last), dn = (d + D) % R;
        qnode(ring[dn], pn);
        riw[dn] = iw[pn - 1];
      }
      cxd w[4];
      cw(nd, w);
      const cxd rd(nd.rd.x, nd.rd.y);
      const cxd qp = cxd((double)nd.rho.x, (double)nd.rho.y) * rd;
      cxd g = cxd((double)nd.s.x, (double)nd.s.y) + (q_prev - qp) * vq;
      if (p >= 2) {
        const cxd vz = qsum(z * vq);
        g -= z * (cq * iWe) + (vz * rd_prev) * vq;
      }
      z = qsymv(w, g);
      PS zs;
      zs.x = (RS)z.re;
      zs.y = (RS)z.im;
      put_z(p, zs);
      rd_prev = rd;
      q_prev = qp;
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = 1;
#pragma unroll UNR
    for (; p0 + (R - 1) <= last; p0 += R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 + d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 + d <= last) step(p0 + d, d);
    if (m >= 2) {
      const cxd vz = qsum(z * vq);
      tm = z * (cq * (double)iw[m - 1]) + (vz * rd_prev) * vq;   // B_m z_{m-1}
    }
  } else if (live) {
    const int last = m + 1;
    // 1/d and rho of the node above each ring entry, kept as loaded: a
    // conversion at load time would wait for the load
    Ch<double> rdn[R];
    PS qn[R];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = min(max(M - d, last), M);
      qnode(ring[d], pn);
      riw[d] = iw[pn];
      rdn[d] = fac[(pn - 1) * fstep + 10 * stride];
      qn[d] = rec_rho(pn - 1);
    }
    cxd zn(0, 0), rd_last(0, 0);
    auto step = [&](const int p, const int d) {
      const QNode<RS>& nd = ring[d];
      const double iWe = (double)riw[d];
      const cxd rdm(rdn[d].v[0], rdn[d].v[1]), rhom((double)qn[d].x, (double)qn[d].y);
      {
        const int pn = max(p - D, last), dn = (d + D) % R;
        qnode(ring[dn], pn);
        riw[dn] = iw[pn];
        rdn[dn] = fac[(pn - 1) * fstep + 10 * stride];
        qn[dn] = rec_rho(pn - 1);
      }
      cxd w[4];
      cw(nd, w);
      const cxd rd(nd.rd.x, nd.rd.y);
      const cxd qp = cxd((double)nd.rho.x, (double)nd.rho.y) * rd;
      const cxd qm = rhom * rdm;
      cxd g = cxd((double)nd.s.x, (double)nd.s.y) + (qm - qp) * vq;
      if (p <= M - 1) {
        const cxd vz = qsum(zn * vq);
        g -= zn * (cq * iWe) + (vz * rd) * vq;
      }
      zn = qsymv(w, g);
      PS zs;
      zs.x = (RS)zn.re;
      zs.y = (RS)zn.im;
      put_z(p, zs);
      rd_last = rdm;
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = M;
#pragma unroll UNR
    for (; p0 - (R - 1) >= last; p0 -= R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 - d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 - d >= last) step(p0 - d, d);
    if (m + 1 <= M) {
      const cxd vz = qsum(zn * vq);
      tm = zn * (cq * (double)iw[m]) + (vz * rd_last) * vq;   // B_{m+1} z_{m+1}
    }
  }

  // ---------------- meeting node ----------------
  cxd tpeer(0, 0);   // one wave: the other side's coupling term, same line and component
  if constexpr (!ONEW) {
    if (side < 2) {
      xch[side][0][lane] = tm.re;
      xch[side][1][lane] = tm.im;
    }
    __syncthreads();
  } else {
    tpeer = cxd(__shfl_xor(tm.re, 32), __shfl_xor(tm.im, 32));
  }
  if (!live) return;
  cxd xm, qmid, rdmid;
  {
    QNode<RS> nd;
    qnode(nd, m);
    const Ch<double> c = fac[(m - 1) * fstep + 10 * stride];
    const PS r = rec_rho(m - 1);
    cxd w[4];
    cw(nd, w);
    rdmid = cxd(nd.rd.x, nd.rd.y);
    qmid = cxd((double)nd.rho.x, (double)nd.rho.y) * rdmid;
    const cxd rdm(c.v[0], c.v[1]);
    const cxd dq = cxd((double)r.x, (double)r.y) * rdm - qmid;
    cxd ttop, tbot;
    if constexpr (!ONEW) {
      ttop = cxd(xch[0][0][lane], xch[0][1][lane]);
      tbot = cxd(xch[1][0][lane], xch[1][1][lane]);
    } else {
      ttop = side == 0 ? tm : tpeer;
      tbot = side == 0 ? tpeer : tm;
    }
    const cxd g = ((cxd((double)nd.s.x, (double)nd.s.y) + dq * vq) - ttop) - tbot;
    xm = qsymv(w, g);
  }

  // ---------------- back substitution ----------------
  E ru[R], ra[R];
  if (side == 0) {
    {
      const long o = off_q + (long)m * G.sa;
      st<T>(u, o, ld<T>(u, o) + xm);
    }
    const int last = m - 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = max(last - d, 1);
      qnode(ring[d], pn);
      riw[d] = iw[pn];
      ru[d] = u[off_q + (long)pn * G.sa];
      ra[d] = u[G.oa + (long)pn * G.sa];
    }
    cxd an = xm;   // component q of x_{p+1}
    auto step = [&](const int p, const int d) {
      const QNode<RS>& nd = ring[d];
      const double iWe = (double)riw[d];
      const E cu = ru[d], ca = ra[d];
      {
        const int pn = max(p - D, 1), dn = (d + D) % R;
        qnode(ring[dn], pn);
        riw[dn] = iw[pn];
        ru[dn] = u[off_q + (long)pn * G.sa];
        ra[dn] = u[G.oa + (long)pn * G.sa];
      }
      cxd w[4];
      cw(nd, w);
      const cxd rd(nd.rd.x, nd.rd.y);
      const cxd qp = cxd((double)nd.rho.x, (double)nd.rho.y) * rd;
      const cxd va = qsum(an * vq);
      const cxd t = an * (cq * iWe) + (va * rd) * vq;
      const cxd ap = cxd((double)nd.s.x, (double)nd.s.y) - qsymv(w, t);
      const cxd dv = qsum((an - ap) * vq);
      const long o = (long)p * G.sa;
      if (q == 0) st<T>(u, G.oa + o, cxd((double)ca.x, (double)ca.y) + (qp + dv * rd));
      st<T>(u, off_q + o, cxd((double)cu.x, (double)cu.y) + ap);
      an = ap;
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = last;
#pragma unroll UNR
    for (; p0 - (R - 1) >= 1; p0 -= R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 - d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 - d >= 1) step(p0 - d, d);
    // along-line edge 0, between the boundary node and node 1
    const cxd rd0(fac[10 * stride].v[0], fac[10 * stride].v[1]);
    const PS r0 = rec_rho(0);
    const cxd q0 = cxd((double)r0.x, (double)r0.y) * rd0;
    const cxd dv = qsum(an * vq);
    if (q == 0) st<T>(u, G.oa, ld<T>(u, G.oa) + (q0 + dv * rd0));
  } else {
    const int last = m + 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int pn = min(last + d, M);
      qnode(ring[d], pn);
      riw[d] = iw[pn - 1];
      ru[d] = u[off_q + (long)pn * G.sa];
      ra[d] = u[G.oa + (long)(pn - 1) * G.sa];
    }
    cxd xp = xm;   // component q of x_{p-1}
    cxd rdm = rdmid, qm = qmid;
    auto step = [&](const int p, const int d) {
      const QNode<RS>& nd = ring[d];
      const double iWe = (double)riw[d];
      const E cu = ru[d], ca = ra[d];
      {
        const int pn = min(p + D, M), dn = (d + D) % R;
        qnode(ring[dn], pn);
        riw[dn] = iw[pn - 1];
        ru[dn] = u[off_q + (long)pn * G.sa];
        ra[dn] = u[G.oa + (long)(pn - 1) * G.sa];
      }
      cxd w[4];
      cw(nd, w);
      const cxd rd(nd.rd.x, nd.rd.y);
      const cxd qp = cxd((double)nd.rho.x, (double)nd.rho.y) * rd;
      const cxd vx = qsum(xp * vq);
      const cxd t = xp * (cq * iWe) + (vx * rdm) * vq;
      const cxd ap = cxd((double)nd.s.x, (double)nd.s.y) - qsymv(w, t);
      const cxd dv = qsum((ap - xp) * vq);
      const long o = (long)p * G.sa;
      if (q == 0) st<T>(u, G.oa + o - G.sa, cxd((double)ca.x, (double)ca.y) + (qm + dv * rdm));
      st<T>(u, off_q + o, cxd((double)cu.x, (double)cu.y) + ap);
      xp = ap;
      rdm = rd;
      qm = qp;
    };
    // whole groups of R nodes without a guard per node (guards inside the
    // group make the register allocator copy the ring), then the rest
    int p0 = last;
#pragma unroll UNR
    for (; p0 + (R - 1) <= M; p0 += R) {
#pragma unroll
      for (int d = 0; d < R; d++) step(p0 + d, d);
    }
#pragma unroll
    for (int d = 0; d < R - 1; d++)
      if (p0 + d <= M) step(p0 + d, d);
    // along-line edge M = N-1, between node M and the boundary node
    const cxd dv = qsum((cxd(0, 0) - xp) * vq);
    const long o = G.oa + (long)M * G.sa;
    if (q == 0) st<T>(u, o, ld<T>(u, o) + (qm + dv * rdm));
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void line_colour_counts(const int* n, int a, int colour, int* nlo, int* nhi) {
  const int b = (a + 1) % 3, c = (a + 2) % 3;
  const int cb = colour & 1, cc = (colour >> 1) & 1;
  auto cnt = [](int m, int par) { const int s = par ? 1 : 2; return m - 1 >= s ? (m - 1 - s) / 2 + 1 : 0; };
  const int nbv = cnt(n[b], cb), ncv = cnt(n[c], cc);
  if (b < c) { *nlo = nbv; *nhi = ncv; } else { *nlo = ncv; *nhi = nbv; }
}

static int g_line_kernel = 0;

int set_line_kernel(int mode) {
  if (mode < 0 || mode > 4) return -1;
  const int prev = g_line_kernel;
  g_line_kernel = mode;
  return prev;
}

// The autotuner's table: the substitution choice measured fastest for a
// launch shape.  Process-wide (every solver on the device shares it) and
// consulted only in the automatic mode (set_line_kernel(0)).
struct TunedLine {
  int choice;
  double us_choice, us_rule;
};
static std::map<LineShape, TunedLine> g_tuned;
static std::mutex g_tuned_mu;

template <typename T>
LineShape line_shape(const LinePack<T>& P, int axis) {
  LineShape k{(int)sizeof(T), axis, 0, 0, P.njobs};
  for (int j = 0; j < P.njobs; j++) {
    k.n = std::max(k.n, P.job[j].L.n[axis]);
    k.nl = std::max(k.nl, (long)P.job[j].nlo * P.job[j].nhi);
  }
  return k;
}

static int tuned_choice(const LineShape& k) {
  std::lock_guard<std::mutex> g(g_tuned_mu);
  const auto it = g_tuned.find(k);
  return it == g_tuned.end() ? LINE_RULE : it->second.choice;
}

void set_line_choice(const LineShape& k, int choice, double us_choice, double us_rule) {
  std::lock_guard<std::mutex> g(g_tuned_mu);
  g_tuned[k] = TunedLine{choice, us_choice, us_rule};
}

bool has_line_choice(const LineShape& k) {
  std::lock_guard<std::mutex> g(g_tuned_mu);
  return g_tuned.count(k) != 0;
}

void clear_line_choices() {
  std::lock_guard<std::mutex> g(g_tuned_mu);
  g_tuned.clear();
}

int line_choices(int* out, double* us, int maxn) {
  std::lock_guard<std::mutex> g(g_tuned_mu);
  int i = 0;
  for (const auto& kv : g_tuned) {
    if (i < maxn) {
      const LineShape& k = kv.first;
      const int row[7] = {k.esize, k.axis, k.n, (int)std::min<long>(k.nl, INT32_MAX), k.njobs, kv.second.choice, 0};
      for (int c = 0; c < 6; c++) out[6 * i + c] = row[c];
      us[2 * i] = kv.second.us_choice;
      us[2 * i + 1] = kv.second.us_rule;
    }
    i++;
  }
  return i;
}

static size_t chunks_of(int values, size_t esize) {
  const size_t v = 16 / esize;
  return (values + v - 1) / v;
}

size_t line_factor_bytes(const int* n, int a, int colour, size_t esize) {
  int nlo, nhi;
  line_colour_counts(n, a, colour, &nlo, &nhi);
  const size_t twisted = (size_t)nlo * nhi * (size_t)n[a] * chunks_of(FREC, esize) * 16;
  if ((long)nlo * nhi == 1 && n[a] >= 2 && n[a] <= PCR_MAX_NODES)
    return std::max(twisted, (size_t)pcr_entries(n[a]) * sizeof(cxd));
  return twisted;
}

size_t line_scratch_bytes(const int* n, int a, int colour, size_t esize) {
  int nlo, nhi;
  line_colour_counts(n, a, colour, &nlo, &nhi);
  return (size_t)nlo * nhi * (size_t)n[a] * chunks_of(SREC, esize) * 16;
}

// A launch; when timing is on, its events are recorded by the dispatch itself
// (hipExtLaunchKernelGGL), so their elapsed time is the kernel's alone.
// bytes = {compulsory, layout} of the launch
struct Bytes {
  double comp, layout;
};

template <typename K, typename... Args>
static void launch_timed(K kernel, dim3 g, dim3 b, hipStream_t s, int cls, Bytes bytes, Args... args) {
  hipEvent_t e0, e1;
  timing::kernel_events(cls, bytes.comp, &e0, &e1, bytes.layout);
  if (e0) hipExtLaunchKernelGGL(kernel, g, b, 0, s, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kernel, g, b, 0, s, args...);
}

template <typename T, int NT>
static void launch_quad_fused(const LinePack<T>& P, int axis, int maxl, hipStream_t s, Bytes by) {
  // the 128-node lines keep their loop (not unrolled): a deeper ring pays
  // there in FP32 cycles (in FP64 it spills)
  constexpr int QD = NT == 128 && sizeof(T) == 4 ? 5 : EMMI_LINE_QUAD_RING;
  // block shape: see quad_lpb / quad_bt
  constexpr int LPB = quad_lpb<T>(NT);
  constexpr int FT = quad_bt<T>(NT);
  const dim3 g((unsigned)((maxl + LPB - 1) / LPB), P.njobs), b(FT);
  if (axis == 0) launch_timed(k_line_solve_q<T, 0, QD, true, FT, NT, LPB>, g, b, s, timing::LINE, by, P);
  else if (axis == 1) launch_timed(k_line_solve_q<T, 1, QD, true, FT, NT, LPB>, g, b, s, timing::LINE, by, P);
  else launch_timed(k_line_solve_q<T, 2, QD, true, FT, NT, LPB>, g, b, s, timing::LINE, by, P);
}

template <typename T>
static void launch_pack(const LinePack<T>& P, int axis, bool factor, hipStream_t s, int choice = LINE_RULE) {
  int maxl = 0;
  long maxnodes = 0;
  double rhs_bytes = 0, solve_bytes = 0, rec_bytes = 0, comp = 0;
  const double es = sizeof(typename cvec<T>::type), cs = sizeof(T);
  for (int j = 0; j < P.njobs; j++) {
    // compulsory bytes of one colour: a quarter of a full sweep's
    // 3 Ne s + 3 Nc c (u read and written, f read, sbar read)
    const Level<T>& L = P.job[j].L;
    comp += 0.25 * (3.0 * 3.0 * (double)L.nc * es + 3.0 * (double)L.n[0] * L.n[1] * L.n[2] * cs);
    const long nl = (long)P.job[j].nlo * P.job[j].nhi;
    const int na = P.job[j].L.n[axis];
    if (nl > maxl) maxl = (int)nl;
    if (nl * na > maxnodes) maxnodes = nl * na;
    const double rows = (double)nl * (5.0 * na - 4.0);
    // right-hand sides: the colour's rows read f and sigma-bar once, u is
    // read once over the level, the records are written once
    rhs_bytes += rows * (es + cs) + 3.0 * (double)P.job[j].L.nc * es + (double)nl * na * SREC * sizeof(ScrRec<T>);
    // substitution: records and factors read once, the rows of u read and written
    solve_bytes += (double)nl * na * (SREC * sizeof(ScrRec<T>) + FREC * sizeof(FacRec<T>)) + rows * 2.0 * es;
    rec_bytes += (double)nl * na * SREC * sizeof(ScrRec<T>);
  }
  if (maxl == 0 || P.njobs == 0) return;
  int maxna = 0;
  for (int j = 0; j < P.njobs; j++) maxna = std::max(maxna, P.job[j].L.n[axis]);
  if (maxl == 1 && maxna <= PCR_MAX_NODES) {
    const dim3 g(1, P.njobs);
    if (factor) {
      timing::Scope ts(timing::OTHER, s, 0);
      if (axis == 0) hipLaunchKernelGGL((k_line_pcr_factor<T, 0>), g, dim3(PCR_MAX_NODES), 0, s, P);
      else if (axis == 1) hipLaunchKernelGGL((k_line_pcr_factor<T, 1>), g, dim3(PCR_MAX_NODES), 0, s, P);
      else hipLaunchKernelGGL((k_line_pcr_factor<T, 2>), g, dim3(PCR_MAX_NODES), 0, s, P);
      return;
    }
    const Bytes by{comp, rhs_bytes + solve_bytes};
    const dim3 b(4 * PCR_MAX_NODES);
    if (axis == 0) launch_timed(k_line_pcr_solve<T, 0>, g, b, s, timing::LINE, by, P);
    else if (axis == 1) launch_timed(k_line_pcr_solve<T, 1>, g, b, s, timing::LINE, by, P);
    else launch_timed(k_line_pcr_solve<T, 2>, g, b, s, timing::LINE, by, P);
    return;
  }
  if (factor) {
    const dim3 g((maxl + 63) / 64, P.njobs), b(64);
    timing::Scope ts(timing::OTHER, s, 0);
    if (axis == 0) hipLaunchKernelGGL((k_line_factor<T, 0>), g, b, 0, s, P);
    else if (axis == 1) hipLaunchKernelGGL((k_line_factor<T, 1>), g, b, 0, s, P);
    else hipLaunchKernelGGL((k_line_factor<T, 2>), g, b, 0, s, P);
    return;
  }
  // along x the quad-lane kernel's updates of u are scattered rows (see
  // k_line_solve_x_apply): the finest FP32 x sweeps of several right-hand
  // sides keep the one-lane kernel (3 slots at 128^3: 1% faster), a single
  // right-hand side takes the quad-lane one (430 against 567 us per sweep)
  if (choice == LINE_RULE && g_line_kernel == 0) choice = tuned_choice(line_shape(P, axis));
  if (choice == LINE_RULE && g_line_kernel == 3) choice = LINE_QUAD_FUSED;
  if (choice == LINE_RULE && g_line_kernel == 4) choice = LINE_QUAD_SEPARATE;
  bool quad, fuse;
  if (choice == LINE_ONE_LANE || choice == LINE_QUAD_FUSED || choice == LINE_QUAD_SEPARATE) {
    quad = choice != LINE_ONE_LANE;
    fuse = choice == LINE_QUAD_FUSED;
  } else {
    quad = g_line_kernel == 2 || (g_line_kernel == 0 && maxl <= quad_max_lines<T>() &&
                                  (axis != 0 || maxl <= 2048 || sizeof(T) == 8 || P.njobs == 1));
    fuse = quad && maxl <= fuse_max_lines<T>();
  }
  if (fuse) {
    // the substitution kernel computes its own right-hand sides
  } else if (axis == 0) {
    int maxn = 0;
    for (int j = 0; j < P.njobs; j++) maxn = std::max(maxn, P.job[j].L.n[0]);
    const int nlb = (maxl + XT - 1) / XT, npb = (maxn + XT - 1) / XT;
    const dim3 g((unsigned)(nlb * npb), P.njobs), b(XT * XT);
    launch_timed(k_line_rhs_x<T>, g, b, s, timing::LINE_RHS, Bytes{0.0, rhs_bytes}, P, nlb);
  } else if (EMMI_LINE_RHS_MARCH != 0 && maxna >= 64) {
    // lines of 64 nodes and more march (as the fused launches do)
    const int run = EMMI_LINE_RHS_RUN;
    const long nth = (long)maxl * ((maxna + run - 1) / run);
    const dim3 g((unsigned)((nth + 255) / 256), P.njobs), b(256);
    if (axis == 1) launch_timed(k_line_rhs_run<T, 1>, g, b, s, timing::LINE_RHS, Bytes{0.0, rhs_bytes}, P, run);
    else launch_timed(k_line_rhs_run<T, 2>, g, b, s, timing::LINE_RHS, Bytes{0.0, rhs_bytes}, P, run);
  } else {
    const dim3 g((unsigned)((maxnodes + 255) / 256), P.njobs), b(256);
    if (axis == 1) launch_timed(k_line_rhs<T, 1>, g, b, s, timing::LINE_RHS, Bytes{0.0, rhs_bytes}, P);
    else launch_timed(k_line_rhs<T, 2>, g, b, s, timing::LINE_RHS, Bytes{0.0, rhs_bytes}, P);
  }
  if (quad) {
    // few lines: four lanes per line half shorten the serial chain
    constexpr int QD = EMMI_LINE_QUAD_RING;
    const dim3 g((maxl + 15) / 16, P.njobs);
    if (fuse) {
      // a common line length with an instance of its own, else the generic one
      int nt = P.job[0].L.n[axis];
      for (int j = 1; j < P.njobs; j++)
        if (P.job[j].L.n[axis] != nt) nt = 0;
      if (nt != 4 && nt != 8 && nt != 16 && nt != 32 && nt != 64 && nt != 128) nt = 0;
      // records kept in LDS are neither written to nor read from memory
      const Bytes by{comp, rhs_bytes + solve_bytes - (quad_lds_records<T>(nt, quad_lpb<T>(nt)) ? 2.0 * rec_bytes : 0.0)};
      switch (nt) {
        case 4: launch_quad_fused<T, 4>(P, axis, maxl, s, by); break;
        case 8: launch_quad_fused<T, 8>(P, axis, maxl, s, by); break;
        case 16: launch_quad_fused<T, 16>(P, axis, maxl, s, by); break;
        case 32: launch_quad_fused<T, 32>(P, axis, maxl, s, by); break;
        case 64: launch_quad_fused<T, 64>(P, axis, maxl, s, by); break;
        case 128: launch_quad_fused<T, 128>(P, axis, maxl, s, by); break;
        default: launch_quad_fused<T, 0>(P, axis, maxl, s, by); break;
      }
    } else {
      const dim3 b(128);
      const Bytes by{comp, solve_bytes};
      if (axis == 0) launch_timed(k_line_solve_q<T, 0, QD, false, 128, 0>, g, b, s, timing::LINE, by, P);
      else if (axis == 1) launch_timed(k_line_solve_q<T, 1, QD, false, 128, 0>, g, b, s, timing::LINE, by, P);
      else launch_timed(k_line_solve_q<T, 2, QD, false, 128, 0>, g, b, s, timing::LINE, by, P);
    }
  } else {
    const dim3 g((maxl + 63) / 64, P.njobs), b(128);
    if (axis == 0) {
      // the substitution's algorithmic bytes include the update of u, which
      // the second kernel makes (counted once, with the substitution)
      int maxn = 0;
      for (int j = 0; j < P.njobs; j++) maxn = std::max(maxn, P.job[j].L.n[0]);
      const int nlb = (maxl + XT - 1) / XT, npb = (maxn + XT - 1) / XT;
      launch_timed(k_line_solve<T, 0>, g, b, s, timing::LINE, Bytes{comp, solve_bytes}, P);
      launch_timed(k_line_solve_x_apply<T>, dim3((unsigned)(nlb * npb), P.njobs), dim3(XT * XT), s, timing::LINE,
                   Bytes{0.0, 0.0}, P, nlb);
    } else if (axis == 1) launch_timed(k_line_solve<T, 1>, g, b, s, timing::LINE, Bytes{comp, solve_bytes}, P);
    else launch_timed(k_line_solve<T, 2>, g, b, s, timing::LINE, Bytes{comp, solve_bytes}, P);
  }
}

template <typename T>
void launch_line_factor(const LinePack<T>& P, int axis, hipStream_t s) { launch_pack<T>(P, axis, true, s); }
template <typename T>
void launch_line_solve(const LinePack<T>& P, int axis, hipStream_t s) { launch_pack<T>(P, axis, false, s); }
template <typename T>
void launch_line_solve_as(const LinePack<T>& P, int axis, int choice, hipStream_t s) {
  launch_pack<T>(P, axis, false, s, choice);
}
template <typename T>
bool line_pcr_shape(const LinePack<T>& P, int axis) {
  const LineShape k = line_shape(P, axis);
  return k.nl == 1 && k.n <= PCR_MAX_NODES;
}

template void launch_line_factor<double>(const LinePack<double>&, int, hipStream_t);
template void launch_line_factor<float>(const LinePack<float>&, int, hipStream_t);
template void launch_line_solve<double>(const LinePack<double>&, int, hipStream_t);
template void launch_line_solve<float>(const LinePack<float>&, int, hipStream_t);
template void launch_line_solve_as<double>(const LinePack<double>&, int, int, hipStream_t);
template void launch_line_solve_as<float>(const LinePack<float>&, int, int, hipStream_t);
template LineShape line_shape<double>(const LinePack<double>&, int);
template LineShape line_shape<float>(const LinePack<float>&, int);
template bool line_pcr_shape<double>(const LinePack<double>&, int);
template bool line_pcr_shape<float>(const LinePack<float>&, int);

}  // namespace emmi
