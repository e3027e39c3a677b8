// xf_kernels.inl -- gfx950 kernels of the FDTD time step (included by
// xf_engine.hip, one translation unit).
//
// All floating-point expressions keep the evaluation order of the reference
// kernels and the file is built with -ffp-contract=off, so the FP64 path
// rounds exactly like the CPU oracle built from the same order.
#include "xf_kernels.h"

namespace xf {

// Where a tile reads its coefficient table: TAB_ANY lets the compiler work
// it out (the single-grid kernels, whose table pointer it can trace to LDS or
// to a kernel argument); TAB_LDS and TAB_GLOBAL state it for the batched
// kernels, whose table is a wave's LDS slice or a global array chosen at run
// time (a pointer that may be either would be read with flat loads).
enum { TAB_ANY = 0, TAB_LDS = 1, TAB_GLOBAL = 2 };

template <int TAB, typename C>
__device__ __forceinline__ C tab_at(const C* tab, int off) {
  if constexpr (TAB == TAB_LDS)
    return ((const __attribute__((address_space(3))) C*)tab)[off];
  else if constexpr (TAB == TAB_GLOBAL)
    return gp(tab)[off];
  else
    return tab[off];
}

template <int TAB, typename C>
__device__ __forceinline__ void tab_row(const C* tab, int id, C (&t)[9]) {
#pragma unroll
  for (int n = 0; n < 9; ++n) t[n] = tab_at<TAB>(tab, 9 * id + n);
}

// ------------------------------------------------------------------ PML
// One face of correctPML (pml_corrector.cu:9-48) applied to the cell
// (i, j, k) of the current pass; v/have/dirty hold the three components of
// the pass being written (loaded lazily when not main-updated).
// USE_TAB: the caller holds the node's coefficient id and table, and a face
// whose c_psi values all equal a table coefficient (F.ta / F.tb >= 0,
// checked bit for bit at upload) reads them there instead of from its
// per-node array.
template <int AX, bool ISE, bool TWO_D, bool USE_TAB = false, int TAB = TAB_ANY, typename S, typename C>
__device__ __forceinline__ void pml_face(const PmlFace<C>& F, int i, int j, int k,
                                         int64_t idx, const Fields<S>& f, const Dims& d,
                                         const Scales<C>& sc, C (&v)[3], bool (&have)[3],
                                         bool (&dirty)[3], const C* tab = nullptr, int id = 0) {
  constexpr int A = (AX + 1) % 3, B = (AX + 2) % 3;
  const int a = AX == 0 ? j : AX == 1 ? k : i;
  const int b = AX == 0 ? k : AX == 1 ? i : j;
  const int c = AX == 0 ? i : AX == 1 ? j : k;
  if (c < F.c0 || c >= F.c1) return;
  const bool in_a = ISE ? (a >= F.a0 && a < F.a1 && b >= F.b0 && b <= F.b1)
                        : (a >= F.a0 && a <= F.a1 && b >= F.b0 && b < F.b1);
  const bool in_b = ISE ? (a >= F.a0 && a <= F.a1 && b >= F.b0 && b < F.b1)
                        : (a >= F.a0 && a < F.a1 && b >= F.b0 && b <= F.b1);
  if (!in_a && !in_b) return;
  const int64_t stc = AX == 0 ? d.sx : AX == 1 ? d.sy : 1;
  const int ci = c + F.off_c - F.gstart;
  const int layer = c - F.nstart;
  const C ca = gp(F.ca)[ci], cb = gp(F.cb)[ci];
  // psi index aBCToXYZ(a, b, layer)
  const int px = AX == 0 ? layer : AX == 1 ? b : a;
  const int py = AX == 0 ? a : AX == 1 ? layer : b;
  const int pz = AX == 0 ? b : AX == 1 ? a : layer;
  const C inv_self = ISE ? sc.inv_e : sc.inv_h;
  const C inv_dual = ISE ? sc.inv_h : sc.inv_e;
  S* const* self = ISE ? f.e : f.h;
  S* const* dual = ISE ? f.h : f.e;
  if (in_a && !(TWO_D && (self[A] == nullptr || dual[B] == nullptr))) {
    const C dp = Store<S>::ld(dual[B], idx, inv_dual);
    const C dq = Store<S>::ld(dual[B], ISE ? idx - stc : idx + stc, inv_dual);
    const uint64_t p = ((uint32_t)px * (uint32_t)F.sa1 + (uint32_t)py) * (uint32_t)F.sa2 + (uint32_t)pz;
    C psi = gp(F.psi_a)[p];
    if (ISE)
      psi = cb * psi + ca * (dp - dq);
    else
      psi = cb * psi + ca * (dq - dp);
    gp(F.psi_a)[p] = psi;
    if (!have[A]) {
      v[A] = Store<S>::ld(self[A], idx, inv_self);
      have[A] = true;
    }
    const C cp = (USE_TAB && F.ta >= 0) ? tab_at<TAB>(tab, 9 * id + F.ta) : gp(F.cpsi_a)[p];
    v[A] += cp * psi;
    dirty[A] = true;
  }
  if (in_b && !(TWO_D && (self[B] == nullptr || dual[A] == nullptr))) {
    const C dp = Store<S>::ld(dual[A], idx, inv_dual);
    const C dq = Store<S>::ld(dual[A], ISE ? idx - stc : idx + stc, inv_dual);
    const uint64_t p = ((uint32_t)px * (uint32_t)F.sb1 + (uint32_t)py) * (uint32_t)F.sb2 + (uint32_t)pz;
    C psi = gp(F.psi_b)[p];
    if (ISE)
      psi = cb * psi + ca * (dp - dq);
    else
      psi = cb * psi + ca * (dq - dp);
    gp(F.psi_b)[p] = psi;
    if (!have[B]) {
      v[B] = Store<S>::ld(self[B], idx, inv_self);
      have[B] = true;
    }
    const C cp = (USE_TAB && F.tb >= 0) ? tab_at<TAB>(tab, 9 * id + F.tb) : gp(F.cpsi_b)[p];
    v[B] += cp * psi;
    dirty[B] = true;
  }
}

// Every face of the pass in registration order (X, Y, Z faces); faces whose
// bit is clear in mask (pml_mask of the cell's column) cannot contain the
// cell and are skipped.
template <bool ISE, bool TWO_D, bool USE_TAB = false, int TAB = TAB_ANY, typename S, typename C>
__device__ __forceinline__ void pml_all(const PmlPass<C>& P, int i, int j, int k, int64_t idx,
                                        const Fields<S>& f, const Dims& d, const Scales<C>& sc,
                                        C (&v)[3], bool (&have)[3], bool (&dirty)[3],
                                        unsigned mask = ~0u, const C* tab = nullptr, int id = 0) {
  for (int q = 0; q < P.n; ++q) {
    if (!((mask >> q) & 1u)) continue;
    const PmlFace<C>& F = P.f[q];
    if (F.axis == 0)
      pml_face<0, ISE, TWO_D, USE_TAB, TAB>(F, i, j, k, idx, f, d, sc, v, have, dirty, tab, id);
    else if (F.axis == 1)
      pml_face<1, ISE, TWO_D, USE_TAB, TAB>(F, i, j, k, idx, f, d, sc, v, have, dirty, tab, id);
    else
      pml_face<2, ISE, TWO_D, USE_TAB, TAB>(F, i, j, k, idx, f, d, sc, v, have, dirty, tab, id);
  }
}

// One node of the z-face CPML corrections of a pass, run after the Yee
// kernel when that carries only the x and y faces (XF_TILE_PML_SPLIT).  The
// corrections of a node read only its own components, the dual field (not
// written in this pass) and their own psi, so per node the adds come in the
// order pml_all gives them: x and y faces in the Yee kernel, then z faces.
template <typename S, bool ISE>
__device__ __forceinline__ void pml_z_node(const Fields<S>& f, const PmlPass<typename Store<S>::C>& P,
                                           const PmlZBox& B, const Dims& d,
                                           const Scales<typename Store<S>::C>& sc,
                                           const RowRef<typename Store<S>::C>& R, int64_t t) {
  using C = typename Store<S>::C;
  if (t >= B.start[B.n]) return;
  const int q = (B.n > 1 && t >= B.start[1]) ? 1 : 0;
  const uint32_t local = (uint32_t)(t - B.start[q]);  // < 2^31 (checked by the engine)
  const uint32_t nk = (uint32_t)B.nk[q], nj = (uint32_t)B.nj[q];
  const int k = B.k0[q] + (int)(local % nk);
  const uint32_t r = local / nk;
  const int j = B.j0[q] + (int)(r % nj);
  const int i = B.i0[q] + (int)(r / nj);
  const int64_t idx = (int64_t)i * d.sx + (int64_t)j * d.sy + k;
  C v[3] = {0, 0, 0};
  bool have[3] = {false, false, false};
  bool dirty[3] = {false, false, false};
  // the node's row id where a face takes c_psi from the table (0 where the
  // node is not a cell, as stored)
  const bool rows = q == 0 ? (P.f[0].ta >= 0 || P.f[0].tb >= 0) : (P.f[1].ta >= 0 || P.f[1].tb >= 0);
  int id = 0;
  if (rows)
    id = R.wide ? (int)gp(static_cast<const uint16_t*>(R.ids))[idx] : (int)gp(static_cast<const uint8_t*>(R.ids))[idx];
  // static face indices keep a by-value pass out of scratch
  if (q == 0)
    pml_face<2, ISE, false, true, TAB_GLOBAL>(P.f[0], i, j, k, idx, f, d, sc, v, have, dirty, R.tab, id);
  else
    pml_face<2, ISE, false, true, TAB_GLOBAL>(P.f[1], i, j, k, idx, f, d, sc, v, have, dirty, R.tab, id);
  S* const* self = ISE ? f.e : f.h;
  const C s = ISE ? sc.s_e : sc.s_h;
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if (dirty[c]) Store<S>::st(self[c], idx, v[c], s);
}

template <typename S, bool ISE>
__global__ void __launch_bounds__(256)
    k_pml_z(Fields<S> f, PmlPass<typename Store<S>::C> P, PmlZBox B, Dims d,
            Scales<typename Store<S>::C> sc, RowRef<typename Store<S>::C> R) {
  pml_z_node<S, ISE>(f, P, B, d, sc, R, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Loads the pass's coefficient table into LDS (or points at global memory).
template <typename C, bool LDS>
__device__ __forceinline__ const C* load_table(const C* __restrict__ g, int nval, C* smem) {
  if (!LDS) return g;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  for (int n = tid; n < nval; n += nthr) smem[n] = g[n];
  __syncthreads();
  return smem;
}

// ------------------------------------------------------------------ TFSF
// exInc..hzInc (tfsf_corrector.cu:18-130): interpolated 1-D incident value.
// P is a TfsfPass or a TfsfZTail (the same field names).
template <template <typename> class Pass, typename C>
__device__ __forceinline__ C tfsf_inc(const Pass<C>& P, bool inc_is_e, int comp, int t,
                                      int i, int j, int k) {
  const int li = i - P.g[0], lj = j - P.g[1], lk = k - P.g[2];
  C proj;
  if (inc_is_e) {
    if (comp == 0)
      proj = gp(P.proj_half[0])[li + 1] + gp(P.proj_int[1])[lj] + gp(P.proj_int[2])[lk];
    else if (comp == 1)
      proj = gp(P.proj_int[0])[li] + gp(P.proj_half[1])[lj + 1] + gp(P.proj_int[2])[lk];
    else
      proj = gp(P.proj_int[0])[li] + gp(P.proj_int[1])[lj] + gp(P.proj_half[2])[lk + 1];
  } else {
    if (comp == 0)
      proj = gp(P.proj_int[0])[li] + gp(P.proj_half[1])[lj + 1] + gp(P.proj_half[2])[lk + 1] - (C)0.5;
    else if (comp == 1)
      proj = gp(P.proj_half[0])[li + 1] + gp(P.proj_int[1])[lj] + gp(P.proj_half[2])[lk + 1] - (C)0.5;
    else
      proj = gp(P.proj_half[0])[li + 1] + gp(P.proj_half[1])[lj + 1] + gp(P.proj_int[2])[lk] - (C)0.5;
  }
  const uint64_t index = (uint64_t)proj;
  const C weight = proj - (C)index;
  const __attribute__((address_space(1))) C* row = gp(P.inc) + (int64_t)t * P.n_aux;
  return P.tr[comp] * ((1 - weight) * row[index] + weight * row[index + 1]);
}

#include "xf_yee3d.inl"

// All TFSF face corrections of one pass in one launch.  Each thread owns a
// written cell of one job; cells shared by several faces of the same
// component are applied by the first of them, in the reference's launch
// order (TFSFCorrector3DAgency/2DAgency, tfsf_corrector.cu:155-271).
// (i, j, k) offsets of flat index r (k fastest) in a box of n1 x n2 rows;
// 32-bit division when r fits (every box here but the very largest), which
// is several times cheaper than the 64-bit routine.
__device__ __forceinline__ void box_coords(int64_t r, int n1, int n2, int& li, int& lj, int& lk) {
  if (r < (int64_t(1) << 32)) {
    uint32_t u = (uint32_t)r;
    lk = (int)(u % (uint32_t)n2);
    u /= (uint32_t)n2;
    lj = (int)(u % (uint32_t)n1);
    li = (int)(u / (uint32_t)n1);
  } else {
    lk = (int)(r % n2);
    r /= n2;
    lj = (int)(r % n1);
    li = (int)(r / n1);
  }
}

template <typename S>
__device__ __forceinline__ void tfsf_cell(const Fields<S>& f, const TfsfPass<typename Store<S>::C>& P,
                                          int t, const Scales<typename Store<S>::C>& sc, const Dims& d,
                                          int two_d, int64_t tid) {
  using C = typename Store<S>::C;
  int q = 0;
  while (q + 1 < P.njob && tid >= P.job[q + 1].start) ++q;
  const TfsfJob<C>& J = P.job[q];
  // a job whose inner nodes the Yee tiles correct enumerates only the two
  // end rows of its edge axis (offset 0 or n - 1 along it)
  const int ea = J.edge_axis;
  int li, lj, lk;
  box_coords(tid - J.start, ea == 1 ? 2 : J.n[1], ea == 2 ? 2 : J.n[2], li, lj, lk);
  if (ea == 0 && li) li = J.n[0] - 1;
  if (ea == 1 && lj) lj = J.n[1] - 1;
  if (ea == 2 && lk) lk = J.n[2] - 1;
  const int cell[3] = {J.o[0] + li, J.o[1] + lj, J.o[2] + lk};
  auto inside = [&](const TfsfJob<C>& X) {
    return cell[0] >= X.o[0] && cell[0] < X.o[0] + X.n[0] && cell[1] >= X.o[1] &&
           cell[1] < X.o[1] + X.n[1] && cell[2] >= X.o[2] && cell[2] < X.o[2] + X.n[2];
  };
  for (int p = 0; p < q; ++p)
    if (P.job[p].comp == J.comp && inside(P.job[p])) return;
  const int64_t idx = two_d ? (int64_t)cell[0] * d.sx + cell[1]
                            : (int64_t)cell[0] * d.sx + (int64_t)cell[1] * d.sy + cell[2];
  S* field = P.is_e ? f.e[J.comp] : f.h[J.comp];
  const C inv = P.is_e ? sc.inv_e : sc.inv_h;
  const C scl = P.is_e ? sc.s_e : sc.s_h;
  C v = Store<S>::ld(field, idx, inv);
  for (int p = q; p < P.njob; ++p) {
    const TfsfJob<C>& X = P.job[p];
    if (X.comp != J.comp || !inside(X)) continue;
    const int axis = X.face / 2;
    const bool negative = (X.face % 2) == 0;
    const int a_axis = (axis + 1) % 3, b_axis = (axis + 2) % 3;
    const int dual = (X.comp == a_axis) ? b_axis : a_axis;
    int pos[3] = {cell[0], cell[1], cell[2]};
    // E pass, negative face: incident H read at c-1; H pass, negative face:
    // H written at c-1 from incident E at c.
    if (negative) pos[axis] += P.is_e ? -1 : 1;
    const C inc = tfsf_inc(P, !P.is_e, dual, t, pos[0], pos[1], pos[2]);
    v += inc * X.coef;
  }
  Store<S>::st(field, idx, v, scl);
}

// The TF/SF corrections of a 2-D pass applied in the Yee thread that has
// just formed (i, j)'s values (FP64/FP32 storage, where a store and a load
// give back the same value, so the result equals k_tfsf's): every job whose
// written box holds the node, in job order, adds inc * coef to its
// component.  TF/SF faces lie outside the CPML slabs, so the node's CPML
// correction, applied before, touches other nodes.
// A 2-D pass has four jobs (one per face), so the loop is unrolled over
// that bound: the job descriptors are wave-uniform and their loads issue
// together at the start of the kernel, not one after another.  The time
// step is read only by a thread that applies a correction.
constexpr int kTfsf2dJobs = 4;
template <typename S>
__device__ __forceinline__ void tfsf_fused_2d(const Fields<S>& f, const TfsfPass<typename Store<S>::C>& P,
                                              int t_arg, const int* t_dev,
                                              const Scales<typename Store<S>::C>& sc, int i, int j,
                                              int64_t idx, typename Store<S>::C (&v)[3], bool (&have)[3],
                                              bool (&dirty)[3]) {
  using C = typename Store<S>::C;
#pragma unroll
  for (int p = 0; p < kTfsf2dJobs; ++p) {
    if (p >= P.njob) break;
    const TfsfJob<C>& X = P.job[p];
    if (i < X.o[0] || i >= X.o[0] + X.n[0] || j < X.o[1] || j >= X.o[1] + X.n[1] || 0 < X.o[2] ||
        0 >= X.o[2] + X.n[2])
      continue;
    const int c = X.comp;
    const int axis = X.face / 2;
    const bool negative = (X.face % 2) == 0;
    const int a_axis = (axis + 1) % 3, b_axis = (axis + 2) % 3;
    const int dual = (c == a_axis) ? b_axis : a_axis;
    int pos[3] = {i, j, 0};
    if (negative) pos[axis] += P.is_e ? -1 : 1;
    const int t = t_dev ? *t_dev : t_arg;
    const C inc = tfsf_inc(P, !P.is_e, dual, t, pos[0], pos[1], pos[2]);
    if (!have[c]) {
      v[c] = Store<S>::ld(P.is_e ? f.e[c] : f.h[c], idx, P.is_e ? sc.inv_e : sc.inv_h);
      have[c] = true;
    }
    v[c] += inc * X.coef;
    dirty[c] = true;
  }
}

template <typename S>
__global__ void __launch_bounds__(256)
    k_tfsf(Fields<S> f, TfsfPass<typename Store<S>::C> P, int t_arg, const int* t_dev,
           Scales<typename Store<S>::C> sc, Dims d, int two_d) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= P.total) return;
  tfsf_cell<S>(f, P, t_dev ? *t_dev : t_arg, sc, d, two_d, tid);
}

// Coefficients LO..HI-1 of a node's row: the dense per-cell arrays when
// they are set, else the table row of the node's id.  Filled element by
// element, with no pointer chosen between the two sources, so the row stays
// in registers (a pointer that may address a private array puts the array in
// scratch memory).
template <int LO, int HI, typename C, typename I>
__device__ __forceinline__ void coef_row(const Dense9<C>& dn, const C* __restrict__ tab,
                                         const I* __restrict__ ids, int64_t idx, C (&c)[9]) {
  if (dn.a[0]) {
#pragma unroll
    for (int n = LO; n < HI; ++n) c[n] = dn.a[n][idx];
  } else {
    const C* t = tab + 9 * (int)ids[idx];
#pragma unroll
    for (int n = LO; n < HI; ++n) c[n] = t[n];
  }
}

// ------------------------------------------------------- 2-D TMz, H and E
// The TE updator (basic_updator_te.cu:9-20, basic_updator.cuh:16-32) on a
// 2-D grid (nz = 1): Ex, Ey and Hz stay exactly zero there, so only Hx, Hy
// and Ez are stored and updated; the dropped terms are products with zero
// differences.  One thread per (i, j) cell, 64 lanes along j.
template <typename S, typename I, bool LDS, bool PML>
__global__ void __launch_bounds__(256)
    k_yee_h2(Fields<S> f, const I* __restrict__ hid,
             const typename Store<S>::C* __restrict__ htab, int ntab,
             PmlPass<typename Store<S>::C> pml, Dims d, Scales<typename Store<S>::C> sc,
             Dense9<typename Store<S>::C> dn, TfsfPass<typename Store<S>::C> tf, int t_arg,
             const int* t_dev) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = load_table<C, LDS>(htab, ntab * 9, reinterpret_cast<C*>(smem_raw));
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (j > d.ny || i > d.nx) return;
  const int64_t idx = (int64_t)i * d.sx + j;
  C v[3] = {0, 0, 0};
  bool have[3] = {false, false, false};
  bool dirty[3] = {false, false, false};
  if (i < d.nx && j < d.ny) {
    C t[9];
    coef_row<0, 6>(dn, tab, hid, idx, t);
    const C ez = Store<S>::ld(f.e[2], idx, sc.inv_e);
    const C ez_j = Store<S>::ld(f.e[2], idx + 1, sc.inv_e);
    const C ez_n = Store<S>::ld(f.e[2], idx + d.sx, sc.inv_e);
    const C zero = 0;
    // Hx: chxh*Hx + chxey*(Ey(k+1)-Ey(k)) + chxez*(Ez(j+1)-Ez(j)), Ey = 0
    v[0] = t[0] * Store<S>::ld(f.h[0], idx, sc.inv_h) + t[1] * zero + t[2] * (ez_j - ez);
    // Hy: chyh*Hy + chyez*(Ez(i+1)-Ez(i)) + chyex*(Ex(k+1)-Ex(k)), Ex = 0
    v[1] = t[3] * Store<S>::ld(f.h[1], idx, sc.inv_h) + t[4] * (ez_n - ez) + t[5] * zero;
    have[0] = have[1] = dirty[0] = dirty[1] = true;
  }
  if (PML) pml_all<false, true>(pml, i, j, 0, idx, f, d, sc, v, have, dirty);
  if (tf.njob > 0) tfsf_fused_2d<S>(f, tf, t_arg, t_dev, sc, i, j, idx, v, have, dirty);
  if (dirty[0]) Store<S>::st(f.h[0], idx, v[0], sc.s_h);
  if (dirty[1]) Store<S>::st(f.h[1], idx, v[1], sc.s_h);
}

template <typename S, typename I, bool LDS, bool PML>
__global__ void __launch_bounds__(256)
    k_yee_e2(Fields<S> f, const I* __restrict__ eid,
             const typename Store<S>::C* __restrict__ etab, int ntab,
             PmlPass<typename Store<S>::C> pml, Dims d, Scales<typename Store<S>::C> sc,
             Dense9<typename Store<S>::C> dn, TfsfPass<typename Store<S>::C> tf, int t_arg,
             const int* t_dev) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = load_table<C, LDS>(etab, ntab * 9, reinterpret_cast<C*>(smem_raw));
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (j > d.ny || i > d.nx) return;
  const int64_t idx = (int64_t)i * d.sx + j;
  C v[3] = {0, 0, 0};
  bool have[3] = {false, false, false};
  bool dirty[3] = {false, false, false};
  if (i < d.nx && j < d.ny && i > 0 && j > 0) {
    C t[9];
    coef_row<6, 9>(dn, tab, eid, idx, t);
    const C hx = Store<S>::ld(f.h[0], idx, sc.inv_h);
    const C hx_j = Store<S>::ld(f.h[0], idx - 1, sc.inv_h);
    const C hy = Store<S>::ld(f.h[1], idx, sc.inv_h);
    const C hy_i = Store<S>::ld(f.h[1], idx - d.sx, sc.inv_h);
    v[2] = t[6] * Store<S>::ld(f.e[2], idx, sc.inv_e) + t[7] * (hx - hx_j) + t[8] * (hy - hy_i);
    have[2] = dirty[2] = true;
  }
  if (PML) pml_all<true, true>(pml, i, j, 0, idx, f, d, sc, v, have, dirty);
  if (tf.njob > 0) tfsf_fused_2d<S>(f, tf, t_arg, t_dev, sc, i, j, idx, v, have, dirty);
  if (dirty[2]) Store<S>::st(f.e[2], idx, v[2], sc.s_e);
}

// ----------------------------------------------------------------- NF2FF
template <typename S>
__device__ __forceinline__ typename Store<S>::C fld(const S* p, int64_t idx,
                                                    typename Store<S>::C inv) {
  return p ? Store<S>::ld(p, idx, inv) : (typename Store<S>::C)0;
}

// interpolateSurfaceCenter<xyz, field> (interpolate_scheme.cuh:11-134).
template <typename S>
__device__ __forceinline__ typename Store<S>::C nf_interp(const Fields<S>& f, const Dims& d,
                                                          int two_d, int face_axis, bool is_e,
                                                          int comp, int i, int j, int k,
                                                          const Scales<typename Store<S>::C>& sc) {
  using C = typename Store<S>::C;
  const S* p = is_e ? f.e[comp] : f.h[comp];
  const C inv = is_e ? sc.inv_e : sc.inv_h;
  const int64_t sx = d.sx, sy = two_d ? 1 : d.sy, sz = two_d ? 0 : 1;
  const int64_t id = (int64_t)i * sx + (int64_t)j * sy + (int64_t)k * sz;
  if (p == nullptr) return 0;
  auto F = [&](int di, int dj, int dk) { return fld(p, id + di * sx + dj * sy + dk * sz, inv); };
  if (face_axis == 0) {
    if (!is_e && comp == 2) return (C)0.25 * (F(0, 0, 1) + F(0, 0, 0) + F(-1, 0, 1) + F(-1, 0, 0));
    if (!is_e && comp == 1) return (C)0.25 * (F(0, 1, 0) + F(0, 0, 0) + F(-1, 1, 0) + F(-1, 0, 0));
    if (is_e && comp == 2) return (C)0.5 * (F(0, 1, 0) + F(0, 0, 0));
    return (C)0.5 * (F(0, 0, 1) + F(0, 0, 0));
  } else if (face_axis == 1) {
    if (!is_e && comp == 2) return (C)0.25 * (F(0, 0, 1) + F(0, 0, 0) + F(0, -1, 1) + F(0, -1, 0));
    if (!is_e && comp == 0) return (C)0.25 * (F(1, 0, 0) + F(0, 0, 0) + F(1, -1, 0) + F(0, -1, 0));
    if (is_e && comp == 2) return (C)0.5 * (F(1, 0, 0) + F(0, 0, 0));
    return (C)0.5 * (F(0, 0, 1) + F(0, 0, 0));
  }
  if (!is_e && comp == 0) return (C)0.25 * (F(1, 0, 0) + F(0, 0, 0) + F(0, 0, -1) + F(1, 0, -1));
  if (!is_e && comp == 1) return (C)0.25 * (F(0, 0, 0) + F(0, 1, 0) + F(0, 0, -1) + F(0, 1, -1));
  if (is_e && comp == 0) return (C)0.5 * (F(0, 1, 0) + F(0, 0, 0));
  return (C)0.5 * (F(1, 0, 0) + F(0, 0, 0));
}

// a complex FP64 sum as a plain two-double vector (the accumulators and
// time tables are double2 arrays; this type is read and written through
// global-space pointers, which HIP's double2 class cannot bind to)
typedef double nf_d2 __attribute__((ext_vector_type(2)));

// calculateJ / calculateM of all six faces and all frequencies in one launch
// (nf2ff_frequency_domain.data.cu:80-232).  The four interpolated tangential
// fields are formed once per surface cell and accumulated into every
// frequency's complex FP64 sums.
template <typename S>
__device__ __forceinline__ void nf2ff_cell(const Fields<S>& f, const NfPass& P, int t, const Dims& d,
                                           int two_d, const Scales<typename Store<S>::C>& sc,
                                           int64_t tid) {
  int face = 0;
  while (face < 5 && tid >= P.face[face + 1].start) ++face;
  const NfFace& F = P.face[face];
  const int64_t cell = tid - F.start;
  int li, lj, lk;
  box_coords(cell, F.n[1], F.n[2], li, lj, lk);
  const int i = F.o[0] + li, j = F.o[1] + lj, k = F.o[2] + lk;
  const int axis = face / 2;
  const bool positive = face % 2;
  const int a_axis = (axis + 1) % 3, b_axis = (axis + 2) % 3;
  const double hb = (double)nf_interp(f, d, two_d, axis, false, b_axis, i, j, k, sc);
  const double ha = (double)nf_interp(f, d, two_d, axis, false, a_axis, i, j, k, sc);
  const double eb = (double)nf_interp(f, d, two_d, axis, true, b_axis, i, j, k, sc);
  const double ea = (double)nf_interp(f, d, two_d, axis, true, a_axis, i, j, k, sc);
  const double r_ja = (positive ? -1.0 : 1.0) * hb;
  const double r_jb = (positive ? 1.0 : -1.0) * ha;
  const double r_ma = (positive ? 1.0 : -1.0) * eb;
  const double r_mb = (positive ? -1.0 : 1.0) * ea;
  const int64_t nface = (int64_t)F.n[0] * F.n[1] * F.n[2];
  for (int q = 0; q < P.nfreq; ++q) {
    const nf_d2 th = gp(reinterpret_cast<const nf_d2*>(P.th))[(int64_t)q * P.nt + t];
    const nf_d2 te = gp(reinterpret_cast<const nf_d2*>(P.te))[(int64_t)q * P.nt + t];
    __attribute__((address_space(1))) nf_d2* base =
        gp(reinterpret_cast<nf_d2*>(P.cur)) + (int64_t)q * P.block * 4 + F.off * 4 + cell;
    nf_d2 v;
    v = base[0];
    v.x += r_ja * th.x;
    v.y += r_ja * th.y;
    base[0] = v;
    v = base[nface];
    v.x += r_jb * th.x;
    v.y += r_jb * th.y;
    base[nface] = v;
    v = base[2 * nface];
    v.x += r_ma * te.x;
    v.y += r_ma * te.y;
    base[2 * nface] = v;
    v = base[3 * nface];
    v.x += r_mb * te.x;
    v.y += r_mb * te.y;
    base[3 * nface] = v;
  }
}

template <typename S>
__global__ void __launch_bounds__(256)
    k_nf2ff(Fields<S> f, NfPass P, int t_arg, const int* t_dev, Dims d, int two_d,
            Scales<typename Store<S>::C> sc) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= P.total) return;
  nf2ff_cell<S>(f, P, t_dev ? *t_dev : t_arg, d, two_d, sc, tid);
}

// ------------------------------------------------------ NF2FF time domain
// One cell's record in a tile: its contributions (wa, wb from H, ua, ub from
// E; each as the (1 - f) part for bin nn and the f part for bin nn + 1) and
// its E and H bins relative to the tile's first (-2: no cell).
struct NtRec {
  double x[8];
  int re, rh;
  int pad[2];
};

// NF2FFTimeDomainData<D>::update<E> and update<H> (nf2ff_time_domain_data.cu:
// 103-188) for one tile of one face: prev fields updated in place, the
// contributions summed per time bin (each wave over its 64 cells in the
// reference's (i, j, k) order, then the four waves in order) into the tile's
// partial sums.  Block: 256 threads = 16 x 16 cells; dynamic LDS: 256 records
// and 4 waves x 4 potentials x maxspan doubles.
template <typename S>
__global__ void __launch_bounds__(256)
    k_nftd_tile(Fields<S> f, NtPass P, Dims d, Scales<typename Store<S>::C> sc, int t_arg,
                const int* t_dev) {
  extern __shared__ double nt_lds[];
  __shared__ int red[4][4];
  NtRec* rec = reinterpret_cast<NtRec*>(nt_lds);
  double* wpart = nt_lds + 256 * (sizeof(NtRec) / sizeof(double));
  const int t = t_dev ? *t_dev : t_arg;
  const int blk = blockIdx.x;
  int face = 0;
  while (face < 5 && blk >= P.face[face + 1].blk0) ++face;
  const NtFace& F = P.face[face];
  const int c = face >> 1, a = (c + 1) % 3, b = (c + 2) % 3;
  const int uax = a < b ? a : b, vax = a < b ? b : a;
  const int na = a == uax ? F.nu : F.nv, nb = b == uax ? F.nu : F.nv;
  const int tl = blk - F.blk0;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int u = (tl / F.tv) * NT_TILE + (tid >> 4);
  const int v = (tl % F.tv) * NT_TILE + (tid & (NT_TILE - 1));
  const bool valid = u < F.nu && v < F.nv;
  int nn_e = 0, nn_h = 0;
  double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (valid) {
    int idx[3];
    idx[c] = F.o[c];
    idx[uax] = F.o[uax] + u;
    idx[vax] = F.o[vax] + v;
    double r[3];
    for (int ax = 0; ax < 3; ++ax) r[ax] = ax == c ? P.e_node[ax][idx[ax]] : P.h_node[ax][idx[ax]];
    const double rd = r[0] * P.u[0] + r[1] * P.u[1] + r[2] * P.u[2];
    const double ds = P.e_size[a][idx[a]] * P.e_size[b][idx[b]];
    const bool positive = face & 1;
    const int64_t plane = (int64_t)na * nb;
    double* pv = P.prev + F.prev_off + (int64_t)(idx[a] - F.o[a]) * nb + (idx[b] - F.o[b]);
    {  // E: (ua, ub), offset 0.5
      double ca = 1.0, cb = -1.0;
      if (!positive) {
        ca *= -1.0;
        cb *= -1.0;
      }
      const double tsd = (P.dmax_e - rd) / P.c_dt;
      const double sn = tsd + 0.5 + (double)t;
      nn_e = (int)floor(sn);
      const double cf = sn - (double)nn_e;
      const double A = (double)nf_interp(f, d, 0, c, true, a, idx[0], idx[1], idx[2], sc);
      const double B = (double)nf_interp(f, d, 0, c, true, b, idx[0], idx[1], idx[2], sc);
      const double dA = A - pv[0];
      const double dB = B - pv[plane];
      pv[0] = A;
      pv[plane] = B;
      x[4] = ca * (1 - cf) * ds * dB;
      x[5] = ca * cf * ds * dB;
      x[6] = cb * (1 - cf) * ds * dA;
      x[7] = cb * cf * ds * dA;
    }
    {  // H: (wa, wb), offset 0
      double ca = -1.0, cb = 1.0;
      if (!positive) {
        ca *= -1.0;
        cb *= -1.0;
      }
      const double tsd = (P.dmax_h - rd) / P.c_dt;
      const double sn = tsd + 0.0 + (double)t;
      nn_h = (int)floor(sn);
      const double cf = sn - (double)nn_h;
      const double A = (double)nf_interp(f, d, 0, c, false, a, idx[0], idx[1], idx[2], sc);
      const double B = (double)nf_interp(f, d, 0, c, false, b, idx[0], idx[1], idx[2], sc);
      const double dA = A - pv[2 * plane];
      const double dB = B - pv[3 * plane];
      pv[2 * plane] = A;
      pv[3 * plane] = B;
      x[0] = ca * (1 - cf) * ds * dB;
      x[1] = ca * cf * ds * dB;
      x[2] = cb * (1 - cf) * ds * dA;
      x[3] = cb * cf * ds * dA;
    }
  }
  // the tile's first and last bins
  int mn_e = valid ? nn_e : INT_MAX, mx_e = valid ? nn_e : INT_MIN;
  int mn_h = valid ? nn_h : INT_MAX, mx_h = valid ? nn_h : INT_MIN;
  for (int o = 32; o > 0; o >>= 1) {
    mn_e = min(mn_e, __shfl_xor(mn_e, o));
    mx_e = max(mx_e, __shfl_xor(mx_e, o));
    mn_h = min(mn_h, __shfl_xor(mn_h, o));
    mx_h = max(mx_h, __shfl_xor(mx_h, o));
  }
  if (lane == 0) {
    red[wave][0] = mn_e;
    red[wave][1] = mx_e;
    red[wave][2] = mn_h;
    red[wave][3] = mx_h;
  }
  __syncthreads();
  int b0_e = red[0][0], b1_e = red[0][1], b0_h = red[0][2], b1_h = red[0][3];
  for (int w = 1; w < 4; ++w) {
    b0_e = min(b0_e, red[w][0]);
    b1_e = max(b1_e, red[w][1]);
    b0_h = min(b0_h, red[w][2]);
    b1_h = max(b1_h, red[w][3]);
  }
  const bool any = b0_e != INT_MAX;
  const int ms = P.maxspan;
  const int span_e = any ? min(b1_e - b0_e + 2, ms) : 0;
  const int span_h = any ? min(b1_h - b0_h + 2, ms) : 0;
  NtRec& R = rec[tid];
  for (int q = 0; q < 8; ++q) R.x[q] = x[q];
  R.re = valid ? nn_e - b0_e : -2;
  R.rh = valid ? nn_h - b0_h : -2;
  __syncthreads();
  // per-wave sums: lane m owns bins m, m + 64, ...
  const int span = max(span_e, span_h);
  const NtRec* W = rec + wave * 64;
  double* wp = wpart + wave * 4 * ms;
  for (int m = lane; m < span; m += 64) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < 64; ++q) {
      const NtRec& r = W[q];
      if (r.rh == m) {
        acc[0] += r.x[0];
        acc[1] += r.x[2];
      } else if (r.rh + 1 == m) {
        acc[0] += r.x[1];
        acc[1] += r.x[3];
      }
      if (r.re == m) {
        acc[2] += r.x[4];
        acc[3] += r.x[6];
      } else if (r.re + 1 == m) {
        acc[2] += r.x[5];
        acc[3] += r.x[7];
      }
    }
    for (int p = 0; p < 4; ++p) wp[p * ms + m] = acc[p];
  }
  __syncthreads();
  double* out = P.part + (int64_t)blk * 4 * ms;
  for (int m = tid; m < span; m += 256)
    for (int p = 0; p < 4; ++p)
      out[p * ms + m] = ((wpart[p * ms + m] + wpart[(4 + p) * ms + m]) + wpart[(8 + p) * ms + m]) +
                        wpart[(12 + p) * ms + m];
  if (tid == 0) {
    P.hdr[blk * 4 + 0] = b0_e;
    P.hdr[blk * 4 + 1] = span_e;
    P.hdr[blk * 4 + 2] = b0_h;
    P.hdr[blk * 4 + 3] = span_h;
  }
}

// One wave per time bin of a face: lane l adds the partial sums of tiles
// l, l + 64, ... that cover the bin, then a xor butterfly (a fixed pairing,
// so every run forms the same sum) combines the lanes and lane 0 adds the
// total to the potentials.  The face's first bin comes from the largest
// r . r_unit (the shortest delay).
__global__ void __launch_bounds__(256) k_nftd_reduce(NtPass P, int t_arg, const int* t_dev) {
  const int t = t_dev ? *t_dev : t_arg;
  const int blk = blockIdx.x;
  int face = 0;
  while (face < 5 && blk >= P.face[face + 1].rblk0) ++face;
  const NtFace& F = P.face[face];
  const int lane = (int)threadIdx.x & 63;
  const int slot = (blk - F.rblk0) * 4 + ((int)threadIdx.x >> 6);
  if (slot >= F.bins) return;
  const int n_e = (int)floor((P.dmax_e - F.rd_max) / P.c_dt + 0.5 + (double)t) + slot;
  const int n_h = (int)floor((P.dmax_h - F.rd_max) / P.c_dt + 0.0 + (double)t) + slot;
  const int ms = P.maxspan;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int cov_e = 0, cov_h = 0;
  const int ntile = F.tu * F.tv;
  for (int q = lane; q < ntile; q += 64) {
    const int4 h = reinterpret_cast<const int4*>(P.hdr)[F.blk0 + q];
    const double* pp = P.part + (int64_t)(F.blk0 + q) * 4 * ms;
    const int rh = n_h - h.z;
    if (rh >= 0 && rh < h.w) {
      acc[0] += pp[rh];
      acc[1] += pp[ms + rh];
      cov_h = 1;
    }
    const int re = n_e - h.x;
    if (re >= 0 && re < h.y) {
      acc[2] += pp[2 * ms + re];
      acc[3] += pp[3 * ms + re];
      cov_e = 1;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    for (int p = 0; p < 4; ++p) acc[p] += __shfl_xor(acc[p], o);
    cov_e |= __shfl_xor(cov_e, o);
    cov_h |= __shfl_xor(cov_h, o);
  }
  if (lane != 0) return;
  double* pot = P.pot + (int64_t)face * 4 * P.n_pot;
  if (cov_h && n_h >= 0 && n_h < P.n_pot) {
    pot[n_h] += acc[0];
    pot[P.n_pot + n_h] += acc[1];
  }
  if (cov_e && n_e >= 0 && n_e < P.n_pot) {
    pot[2 * P.n_pot + n_e] += acc[2];
    pot[3 * P.n_pot + n_e] += acc[3];
  }
}

// --------------------------------------------------------- point sources
// E_comp(node) += value[t], sources in registration order (one thread, so
// sources sharing a node add in a fixed order).
template <typename S>
__global__ void k_point_sources(Fields<S> f, PointDev P, Scales<typename Store<S>::C> sc,
                                int t_arg, const int* t_dev) {
  using C = typename Store<S>::C;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int t = t_dev ? *t_dev : t_arg;
  for (int q = 0; q < P.n; ++q) {
    S* fld = f.e[P.comp[q]];
    const int64_t i = P.idx[q];
    const C v = Store<S>::ld(fld, i, sc.inv_e) + (C)P.value[(int64_t)q * P.nt + t];
    Store<S>::st(fld, i, v, sc.s_e);
  }
}

// ------------------------------------------------------------- movies
// MovieMonitor<F>::update (movie_monitor.cu:17-45): every `interval`
// steps the component over the box goes to frame t / interval.  All movie
// monitors of a grid run in one launch: thread c of the concatenated boxes
// (start[q] = first thread of monitor q) copies its node when its monitor
// records this step.
template <typename S>
__device__ __forceinline__ void movie_node(const Fields<S>& f, const MovieDev* __restrict__ M,
                                           const int64_t* __restrict__ start, int n, const Dims& d, int two_d,
                                           const Scales<typename Store<S>::C>& sc, int t, int64_t c) {
  if (c >= start[n]) return;
  int q = 0;
  while (q + 1 < n && c >= start[q + 1]) ++q;
  const MovieDev& m = M[q];
  if (t % m.interval != 0) return;
  const int64_t frame = t / m.interval;
  if (frame >= m.frames) return;
  const int64_t r = c - start[q];
  int li, lj, lk;
  box_coords(r, m.n[1], m.n[2], li, lj, lk);
  const int i = m.o[0] + li, j = m.o[1] + lj, k = m.o[2] + lk;
  const S* p = m.field < 3 ? f.e[m.field] : f.h[m.field - 3];
  const int64_t idx = two_d ? (int64_t)i * d.sx + j : (int64_t)i * d.sx + (int64_t)j * d.sy + k;
  const double v = p ? (double)Store<S>::ld(p, idx, m.field < 3 ? sc.inv_e : sc.inv_h) : 0.0;
  m.data[frame * m.box + r] = v;
}

template <typename S>
__global__ void __launch_bounds__(256)
    k_movies(Fields<S> f, const MovieDev* __restrict__ M, const int64_t* __restrict__ start, int n, Dims d,
             int two_d, Scales<typename Store<S>::C> sc, int t_arg, const int* t_dev) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  movie_node<S>(f, M, start, n, d, two_d, sc, t_dev ? *t_dev : t_arg, c);
}

// An NF2FF box and the movie monitors in one launch: both only read the
// fields at the end of the step, so they are independent.  Threads
// [0, P.total) accumulate the box, the rest record the movies.
template <typename S>
__global__ void __launch_bounds__(256)
    k_nf2ff_movies(Fields<S> f, NfPass P, const MovieDev* __restrict__ M, const int64_t* __restrict__ start,
                   int n, Dims d, int two_d, Scales<typename Store<S>::C> sc, int t_arg, const int* t_dev) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = t_dev ? *t_dev : t_arg;
  if (tid < P.total)
    nf2ff_cell<S>(f, P, t, d, two_d, sc, tid);
  else
    movie_node<S>(f, M, start, n, d, two_d, sc, t, tid - P.total);
}

// --------------------------------------------------------- learned layer
// The column predictor of xf_learned_boundary runs on the matrix cores
// (k_lb_mfma): the Mur update plus s times the MLP of x / s, s the power of
// two above max |x_i| (exact scaling), with x / s and the layer-1
// activations in fp16, products v_mfma_f32_32x32x16_f16 (fp16 x fp16, fp32
// accumulation) and the output unit on the VALU.
typedef _Float16 lb_h8 __attribute__((ext_vector_type(8)));
typedef _Float16 lb_h4 __attribute__((ext_vector_type(4)));
typedef float lb_f16 __attribute__((ext_vector_type(16)));
constexpr int kLbTilesPerWave = 1;  // 64-column tiles per wave (1 and 2 measure alike)
constexpr int kLbMfmaCols = 4 * 64 * kLbTilesPerWave;  // columns per 256-thread block

// A column and the fields of its face it needs (copied, so no pointer into
// the kernel's argument escapes and the argument stays out of scratch).
struct LbCol {
  int f, c, a, b, positive;  // face, its normal axis and cyclic frame, side
  int tr;                    // LbFace::tr
  int cc;                    // component: 0 E_a, 1 E_b
  int p, q;                  // in-plane a- and b-index
  int p1, q1;                // last a- and b-index of the component
  int64_t hist_off, na1, nb1;
  int64_t abi_col;           // the column's index in xf_learned_boundary order
};

__device__ __forceinline__ LbCol lb_column(const LbDev& L, int64_t col) {
  int fi = 0;
  while (fi + 1 < L.nface && col >= L.face[fi + 1].start[0]) ++fi;
  LbCol r;
  r.f = L.face[fi].f;
  r.c = L.face[fi].c;
  r.a = L.face[fi].a;
  r.b = L.face[fi].b;
  r.positive = L.face[fi].positive;
  r.tr = L.face[fi].tr;
  r.hist_off = L.face[fi].hist_off;
  r.na1 = L.face[fi].na1;
  r.nb1 = L.face[fi].nb1;
  r.cc = col >= L.face[fi].start[1] ? 1 : 0;
  r.p1 = L.face[fi].p1[r.cc];
  r.q1 = L.face[fi].q1[r.cc];
  // a component of a face holds < 2^32 columns (checked at upload); thread
  // order is a-major (xf_learned_boundary order) or, on tr faces, b-major
  const int64_t start = L.face[fi].start[r.cc];
  const uint32_t local = (uint32_t)(col - start);
  const uint32_t width = (uint32_t)(r.q1 - L.inset + 1);
  if (r.tr) {
    const uint32_t height = (uint32_t)(r.p1 - L.inset + 1);
    r.q = L.inset + (int)(local / height);
    r.p = L.inset + (int)(local % height);
  } else {
    r.p = L.inset + (int)(local / width);
    r.q = L.inset + (int)(local % width);
  }
  r.abi_col = start + (int64_t)(r.p - L.inset) * width + (r.q - L.inset);
  return r;
}

__device__ __forceinline__ int64_t lb_node(const LbDev& L, const Dims& d, const LbCol& c,
                                           int layer) {
  int pos[3];
  pos[c.c] = c.positive ? L.n[c.c] - L.inset - layer : L.inset + layer;
  pos[c.a] = c.p;
  pos[c.b] = c.q;
  return (int64_t)pos[0] * d.sx + (int64_t)pos[1] * d.sy + pos[2];
}

__device__ __forceinline__ __attribute__((address_space(1))) float* lb_hist(const LbDev& L, const LbCol& c,
                                                                           int par, int layer, int p, int q) {
  const int64_t plane = ((int64_t)par * 2 + c.cc) * 4 + layer;
  return gp(L.hist) + c.hist_off + (c.tr ? (plane * c.nb1 + q) * c.na1 + p : (plane * c.na1 + p) * c.nb1 + q);
}

// The face a node of the column also lies on (its component tangential to
// both), or -1.
__device__ __forceinline__ int lb_other(const LbDev& L, const LbCol& c) {
  if (c.cc == 0) {
    if (c.q == 0 && ((L.faces >> (2 * c.b)) & 1)) return 2 * c.b;
    if (c.q == L.n[c.b] && ((L.faces >> (2 * c.b + 1)) & 1)) return 2 * c.b + 1;
  } else {
    if (c.p == 0 && ((L.faces >> (2 * c.a)) & 1)) return 2 * c.a;
    if (c.p == L.n[c.a] && ((L.faces >> (2 * c.a + 1)) & 1)) return 2 * c.a + 1;
  }
  return -1;
}

// The 12 inputs of a column at time t (x0..x2 u^{t+1} on layers 1..3, x3..x6
// u^t on layers 0..3, x7..x9 u^{t-1} on layers 0..2, x10/x11 the in-plane
// 5-point Laplacians of u^t on layers 1 and 0, a neighbour outside the
// column range counting as the centre value).
template <typename S>
__device__ __forceinline__ void lb_inputs(const Fields<S>& f, const LbDev& L, const Dims& d,
                                          const Scales<typename Store<S>::C>& sc, const LbCol& c,
                                          int par, float (&x)[XF_LB_IN]) {
  const S* fld = f.e[c.cc == 0 ? c.a : c.b];
#pragma unroll
  for (int l = 1; l <= 3; ++l) x[l - 1] = (float)Store<S>::ld(fld, lb_node(L, d, c, l), sc.inv_e);
#pragma unroll
  for (int l = 0; l < 4; ++l) x[3 + l] = *lb_hist(L, c, par, l, c.p, c.q);
#pragma unroll
  for (int l = 0; l < 3; ++l) x[7 + l] = *lb_hist(L, c, par ^ 1, l, c.p, c.q);
#pragma unroll
  for (int l = 1; l >= 0; --l) {
    const float u = x[3 + l];
    const float up = c.p < c.p1 ? *lb_hist(L, c, par, l, c.p + 1, c.q) : u;
    const float um = c.p > L.inset ? *lb_hist(L, c, par, l, c.p - 1, c.q) : u;
    const float vp = c.q < c.q1 ? *lb_hist(L, c, par, l, c.p, c.q + 1) : u;
    const float vm = c.q > L.inset ? *lb_hist(L, c, par, l, c.p, c.q - 1) : u;
    x[11 - l] = (((up + um) + (vp + vm)) - 4.0f * u);
  }
}

// One column of the layer at time t in the VALU phases: RECORD phase 0
// (training samples), APPLY phase 1 (edge nodes: the first-order Mur
// update, the network being trained on face interiors), phase 2 (history
// roll).  APPLY phase 0 is k_lb_mfma's.
template <typename S>
__device__ __forceinline__ void lb_column_step(const Fields<S>& f, const LbDev& L, const Dims& d,
                                               const Scales<typename Store<S>::C>& sc, int phase,
                                               int t, int64_t col) {
  using C = typename Store<S>::C;
  const int par = t & 1;
  const LbCol c = lb_column(L, col);
  S* fld = f.e[c.cc == 0 ? c.a : c.b];
  if (phase == 2) {
#pragma unroll
    for (int l = 0; l < 4; ++l)
      *lb_hist(L, c, par ^ 1, l, c.p, c.q) = (float)Store<S>::ld(fld, lb_node(L, d, c, l), sc.inv_e);
    return;
  }
  if (L.mode == XF_LB_RECORD) {
    if (phase != 0 || (t % L.rec_every) != 0 || t / L.rec_every >= L.rec_slots) return;
  } else {
    const int other = lb_other(L, c);
    if (phase != 1 || !(other >= 0 && c.f < other)) return;
  }
  float x[XF_LB_IN];
  lb_inputs<S>(f, L, d, sc, c, par, x);
  const int64_t node0 = lb_node(L, d, c, 0);
  if (L.mode == XF_LB_RECORD) {
    __attribute__((address_space(1))) float* r = gp(L.rec) + ((int64_t)(t / L.rec_every) * L.total + c.abi_col) * 13;
#pragma unroll
    for (int i = 0; i < XF_LB_IN; ++i) r[i] = x[i];
    r[12] = (float)Store<S>::ld(fld, node0, sc.inv_e);
    return;
  }
  const float u1 = x[4] + L.mur_k * (x[0] - x[3]);
  Store<S>::st(fld, node0, (C)u1, sc.s_e);
}

// The MLP of one 32-column product tile: bx holds the tile's scaled inputs
// in the B-operand layout (lane l: column l % 32, inputs 8g..8g+7); returns
// the output unit's r for column l % 32 in both lanes l and l + 32.
__device__ __forceinline__ float lb_mlp_tile(const lb_h8& bx, const lb_h8& a1, const lb_h8& a20,
                                             const lb_h8& a21, const lb_f16& c1, const lb_f16& c2,
                                             const float (&w3)[16], float b3) {
  const lb_f16 z1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, bx, c1, 0, 0, 0);
  lb_h8 h0, h1;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float u = z1[j] > 0.0f ? z1[j] : 0.0f;
    h0[j] = (_Float16)(u < 65504.0f ? u : 65504.0f);
    u = z1[8 + j] > 0.0f ? z1[8 + j] : 0.0f;
    h1[j] = (_Float16)(u < 65504.0f ? u : 65504.0f);
  }
  lb_f16 z2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a20, h0, c2, 0, 0, 0);
  z2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a21, h1, z2, 0, 0, 0);
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) s = s + w3[i] * (z2[i] > 0.0f ? z2[i] : 0.0f);
  return (s + __shfl_xor(s, 32)) + b3;
}

// APPLY phase 0 for the tiles [col0, col0 + 64 * ntiles) of one wave.  Lane
// l gathers and writes column col0 + l; the 64 columns are two product
// tiles, A (lanes 0..31's columns) and B (lanes 32..63's).  Lane half g
// feeds inputs 8g..8g+7 of its tile column, so lanes l and l ^ 32 swap the
// half of their scaled inputs the other one feeds; lane half g then holds
// accumulator rows 8(i / 4) + 4g + i % 4.  Every lane reaches every
// product (no lane exits).
template <typename S>
__device__ __forceinline__ void lb_mfma_tiles(const Fields<S>& f, const LbDev& L, const Dims& d,
                                              const Scales<typename Store<S>::C>& sc, int t,
                                              int64_t col0, int ntiles) {
  using C = typename Store<S>::C;
  const int lane = threadIdx.x & 63, g = lane >> 5;
  // the pack is read through a global-space pointer (it is found in a
  // descriptor, so the compiler cannot tell its address space)
  typedef unsigned lb_u4 __attribute__((ext_vector_type(4)));
  const __attribute__((address_space(1))) LbPack* P = gp(L.pack);
  const lb_h8 a1 = __builtin_bit_cast(lb_h8, *gp(reinterpret_cast<const lb_u4*>(L.pack->a1[lane])));
  const lb_h8 a20 = __builtin_bit_cast(lb_h8, *gp(reinterpret_cast<const lb_u4*>(L.pack->a2[0][lane])));
  const lb_h8 a21 = __builtin_bit_cast(lb_h8, *gp(reinterpret_cast<const lb_u4*>(L.pack->a2[1][lane])));
  lb_f16 c1, c2;
  float w3[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    c1[i] = P->c1[i][lane];
    c2[i] = P->c2[i][lane];
    w3[i] = P->w3[i][lane];
  }
  const float b3 = P->b3;
  const int par = t & 1;
  for (int q = 0; q < ntiles; ++q) {
    const int64_t col = col0 + 64 * q + lane;
    bool live = col < L.total;
    LbCol c{};
    if (live) {
      c = lb_column(L, col);
      live = lb_other(L, c) < 0;  // edge nodes are phase 1's
    }
    float x[XF_LB_IN];
#pragma unroll
    for (int i = 0; i < XF_LB_IN; ++i) x[i] = 0.0f;
    if (live) lb_inputs<S>(f, L, d, sc, c, par, x);
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < XF_LB_IN; ++i) {
      const float ax = fabsf(x[i]);
      if (ax > m) m = ax;
    }
    const float mur = x[4] + L.mur_k * (x[0] - x[3]);
    int e = 0;
    if (m > 0.0f) (void)frexpf(m, &e);
    lb_h8 lo;
    lb_h4 hi;
#pragma unroll
    for (int j = 0; j < 8; ++j) lo[j] = (_Float16)ldexpf(x[j], -e);
#pragma unroll
    for (int j = 0; j < 4; ++j) hi[j] = (_Float16)ldexpf(x[8 + j], -e);
    const uint4 lo_u = __builtin_bit_cast(uint4, lo);
    const uint2 hi_u = __builtin_bit_cast(uint2, hi);
    // half 0 sends inputs 8..11 (its partner feeds them for tile A), half 1
    // sends inputs 0..7 (its partner feeds them for tile B)
    // (selects dword by dword: a select of whole vectors goes through scratch)
    const bool h0 = g == 0;
    const unsigned r0 = __shfl_xor(h0 ? hi_u.x : lo_u.x, 32);
    const unsigned r1 = __shfl_xor(h0 ? hi_u.y : lo_u.y, 32);
    const unsigned r2 = __shfl_xor(h0 ? 0u : lo_u.z, 32);
    const unsigned r3 = __shfl_xor(h0 ? 0u : lo_u.w, 32);
    const lb_h8 bA = __builtin_bit_cast(lb_h8, make_uint4(h0 ? lo_u.x : r0, h0 ? lo_u.y : r1,
                                                          h0 ? lo_u.z : 0u, h0 ? lo_u.w : 0u));
    const lb_h8 bB = __builtin_bit_cast(lb_h8, make_uint4(h0 ? r0 : hi_u.x, h0 ? r1 : hi_u.y,
                                                          h0 ? r2 : 0u, h0 ? r3 : 0u));
    const float rA = lb_mlp_tile(bA, a1, a20, a21, c1, c2, w3, b3);
    const float rB = lb_mlp_tile(bB, a1, a20, a21, c1, c2, w3, b3);
    if (live) {
      const float r = g == 0 ? rA : rB;
      const float u1 = m == 0.0f ? mur : mur + ldexpf(r, e);
      S* fld = f.e[c.cc == 0 ? c.a : c.b];
      const int64_t node0 = lb_node(L, d, c, 0);
      Store<S>::st(fld, node0, (C)u1, sc.s_e);
      // the history roll of this column (phase 2's work): u^{t+1} of layers
      // 0..3 into the parity slot of t + 1.  Layer 0 is the value just
      // stored (read back, so a narrow storage format rounds it as it
      // would), layers 1..3 are x0..x2, interior nodes nothing else of the
      // step writes; the slot is u^{t-1}'s, which only this column reads.
      *lb_hist(L, c, par ^ 1, 0, c.p, c.q) = (float)Store<S>::ld(fld, node0, sc.inv_e);
#pragma unroll
      for (int l = 1; l < 4; ++l) *lb_hist(L, c, par ^ 1, l, c.p, c.q) = x[l - 1];
    }
  }
}

// APPLY phases 1 (edge nodes) and 2 (their history roll) over the edge
// columns only; the other columns roll their history in k_lb_mfma.
template <typename S>
__global__ void __launch_bounds__(256)
    k_lb_edges(Fields<S> f, LbDev L, Dims d, Scales<typename Store<S>::C> sc, int phase, int t_arg,
               const int* t_dev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n_edge) return;
  lb_column_step<S>(f, L, d, sc, phase, t_dev ? *t_dev : t_arg, (int64_t)L.edge[i]);
}

template <typename S>
__global__ void __launch_bounds__(256)
    k_lb_mfma(Fields<S> f, LbDev L, Dims d, Scales<typename Store<S>::C> sc, int t_arg,
              const int* t_dev) {
  const int64_t col0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 * kLbTilesPerWave);
  if (col0 >= L.total) return;  // whole waves only
  lb_mfma_tiles<S>(f, L, d, sc, t_dev ? *t_dev : t_arg, col0, kLbTilesPerWave);
}

template <typename S>
__global__ void __launch_bounds__(256)
    k_lb(Fields<S> f, LbDev L, Dims d, Scales<typename Store<S>::C> sc, int phase, int t_arg,
         const int* t_dev) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= L.total) return;
  lb_column_step<S>(f, L, d, sc, phase, t_dev ? *t_dev : t_arg, col);
}

// ----------------------------------------------------------- batched forms
// Item g owns flat indices [start[g], start[g + 1]); start[n] is the total.
__device__ __forceinline__ int owner_of(const int64_t* __restrict__ start, int n, int64_t x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// In the one-launch batch passes below, item q owns the 256-thread blocks
// covering [start[q], start[q + 1]) (ranges padded to whole blocks), so the
// item is found from the block's first thread and is the same for the whole
// block.
// The TFSF pass of every grid of a batch in one launch.
template <typename S>
__global__ void __launch_bounds__(256)
    k_tfsf_batch(const TfsfItem<S>* __restrict__ items, const int64_t* __restrict__ start, int n,
                 int t_arg, const int* t_dev) {
  const int64_t b0 = (int64_t)blockIdx.x * blockDim.x;
  if (b0 >= start[n]) return;
  const int g = owner_of(start, n, b0);
  const TfsfItem<S>& I = items[g];
  const int64_t local = b0 - start[g] + threadIdx.x;
  if (local >= I.P.total) return;
  tfsf_cell<S>(I.f, I.P, t_dev ? *t_dev : t_arg, I.sc, I.d, 0, local);
}

// The NF2FF accumulation of every box of a batch in one launch.
template <typename S>
__global__ void __launch_bounds__(256)
    k_nf2ff_batch(const NfItem<S>* __restrict__ items, const int64_t* __restrict__ start, int n,
                  int t_arg, const int* t_dev) {
  const int64_t b0 = (int64_t)blockIdx.x * blockDim.x;
  if (b0 >= start[n]) return;
  const int g = owner_of(start, n, b0);
  const NfItem<S>& I = items[g];
  const int64_t local = b0 - start[g] + threadIdx.x;
  if (local >= I.P.total) return;
  nf2ff_cell<S>(I.f, I.P, t_dev ? *t_dev : t_arg, I.d, 0, I.sc, local);
}

// The z-face CPML corrections of every grid of a batch that splits them
// off, in one launch per pass.
template <typename S, bool ISE>
__global__ void __launch_bounds__(256)
    k_pml_z_batch(const PmlZItem<S>* __restrict__ items, const int64_t* __restrict__ start, int n) {
  const int64_t b0 = (int64_t)blockIdx.x * blockDim.x;
  if (b0 >= start[n]) return;
  const int g = owner_of(start, n, b0);
  const PmlZItem<S>& I = items[g];
  const int64_t local = b0 - start[g] + threadIdx.x;
  if (local >= I.B.start[I.B.n]) return;
  pml_z_node<S, ISE>(I.f, I.P, I.B, I.d, I.sc, I.R, local);
}

// The point sources of every grid of a batch in one launch: thread g adds
// grid g's sources in registration order, as k_point_sources does.
template <typename S>
__global__ void __launch_bounds__(64)
    k_point_sources_batch(const PointItem<S>* __restrict__ items, int n, int t_arg, const int* t_dev) {
  using C = typename Store<S>::C;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int t = t_dev ? *t_dev : t_arg;
  const PointItem<S>& I = items[g];
  for (int q = 0; q < I.P.n; ++q) {
    S* fld = I.f.e[gp(I.P.comp)[q]];
    const int64_t i = gp(I.P.idx)[q];
    const C v = Store<S>::ld(fld, i, I.sc.inv_e) + (C)gp(I.P.value)[(int64_t)q * I.P.nt + t];
    Store<S>::st(fld, i, v, I.sc.s_e);
  }
}

// One VALU phase of the recording learned layers of a batch in one launch.
// Item g owns blocks [bstart[g], bstart[g + 1]) of 256 columns.
template <typename S>
__global__ void __launch_bounds__(256)
    k_lb_batch(const LbItem<S>* __restrict__ items, const int64_t* __restrict__ bstart, int n,
               int phase, int t_arg, const int* t_dev) {
  const int g = owner_of(bstart, n, (int64_t)blockIdx.x);
  const LbItem<S>& I = items[g];
  if (I.L.mode != XF_LB_RECORD) return;  // apply layers: k_lb_mfma_batch, k_lb_edges_batch
  const int64_t col = ((int64_t)blockIdx.x - bstart[g]) * blockDim.x + threadIdx.x;
  if (col >= I.L.total) return;
  lb_column_step<S>(I.f, I.L, I.d, I.sc, phase, t_dev ? *t_dev : t_arg, col);
}

// APPLY phases 1 and 2 of every learned layer of a batch over their edge
// columns; item g owns edge-list entries [estart[g], estart[g + 1]).
template <typename S>
__global__ void __launch_bounds__(256)
    k_lb_edges_batch(const LbItem<S>* __restrict__ items, const int64_t* __restrict__ estart, int n,
                     int phase, int t_arg, const int* t_dev) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= estart[n]) return;
  const int g = owner_of(estart, n, tid);
  const LbItem<S>& I = items[g];
  lb_column_step<S>(I.f, I.L, I.d, I.sc, phase, t_dev ? *t_dev : t_arg, (int64_t)I.L.edge[tid - estart[g]]);
}

// APPLY phase 0 of every learned layer of a batch in one launch.  Item g
// owns blocks [bstart[g], bstart[g + 1]) of kLbMfmaCols columns; RECORD
// layers own no blocks.
template <typename S>
__global__ void __launch_bounds__(256)
    k_lb_mfma_batch(const LbItem<S>* __restrict__ items, const int64_t* __restrict__ bstart, int n,
                    int t_arg, const int* t_dev) {
  const int g = owner_of(bstart, n, (int64_t)blockIdx.x);
  const LbItem<S>& I = items[g];
  if (I.L.mode != XF_LB_APPLY) return;
  const int64_t col0 = (((int64_t)blockIdx.x - bstart[g]) * 4 + (threadIdx.x >> 6)) * (64 * kLbTilesPerWave);
  if (col0 >= I.L.total) return;  // whole waves only
  lb_mfma_tiles<S>(I.f, I.L, I.d, I.sc, t_dev ? *t_dev : t_arg, col0, kLbTilesPerWave);
}

// ------------------------------------------------------------------ ADE
// Device state of the dispersive cells, compacted (the reference keeps
// full-grid 4-D arrays, ade_method.cuh:176-182).  Per-cell and per-pole
// coefficients and the J recursion stay in FP64.
struct AdeDev {
  int kind;
  int np;
  int64_t n;                // dispersive cells
  const int64_t* idx;       // padded-grid index of each cell
  // Coefficients: the cells of one material share them, so by default each
  // cell holds a row id into a table of the distinct rows (tw doubles each:
  // cejs, ceep, then per pole cjj, cje, cjsj, cjjp, cjen, cjep, cjsjp);
  // with more than 65535 distinct rows, or row == nullptr, the per-cell
  // arrays below are read instead.
  const uint16_t* row;      // [n]
  const double* tab;        // [rows][tw]
  int tw;
  const double* cejs;       // [n]
  const double* ceep;       // [n]
  const double* cjj;        // [np][n]
  const double* cjjp;
  const double* cjen;
  const double* cje;
  const double* cjep;
  const double* cjsj;
  const double* cjsjp;
  double* eprev;            // [3][n]
  double* jc;               // [3][np][n]
  double* jp;               // [3][np][n]
};

// TemplateADEUpdateScheme::updateE (template_ade_update_scheme.cuh:15-86)
// for the three components of each dispersive cell, with the model's
// calculateJSum / updateJ / recordEPrevious.
template <typename S, typename I>
__global__ void __launch_bounds__(256)
    k_ade_e3(Fields<S> f, const I* __restrict__ eid,
             const typename Store<S>::C* __restrict__ etab, AdeDev A, Dims d,
             Scales<typename Store<S>::C> sc, Dense9<typename Store<S>::C> dn) {
  using C = typename Store<S>::C;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.n) return;
  const int64_t idx = A.idx[c];
  // (i, j, k) from the padded index (32-bit division when the grid allows)
  int i, j, k;
  if ((uint64_t)d.sx * (uint64_t)(d.nx + 1) < (uint64_t(1) << 32)) {
    const uint32_t u = (uint32_t)idx, sx = (uint32_t)d.sx, sy = (uint32_t)d.sy;
    i = (int)(u / sx);
    const uint32_t r = u - (uint32_t)i * sx;
    j = (int)(r / sy);
    k = (int)(r - (uint32_t)j * sy);
  } else {
    i = (int)(idx / d.sx);
    const int64_t r = idx - (int64_t)i * d.sx;
    j = (int)(r / d.sy);
    k = (int)(r - (int64_t)j * d.sy);
  }
  // the cell's coefficient row, or the per-cell arrays
  const double* crow = A.row ? A.tab + (int64_t)A.row[c] * A.tw : nullptr;
  auto coef = [&](const double* per_cell, int64_t at, int off) -> double {
    return crow ? crow[off] : per_cell[at];
  };
  C t[9];
  coef_row<0, 9>(dn, etab, eid, idx, t);
  const int cell[3] = {i, j, k};
  const int64_t st[3] = {d.sx, d.sy, 1};
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const int a_axis = (comp + 1) % 3, b_axis = (comp + 2) % 3;
    const int a = cell[a_axis], b = cell[b_axis];
    if (a == 0 || b == 0) continue;
    // eNext with ha = H_a (difference along b), hb = H_b (along a)
    const double ha_p = (double)Store<S>::ld(f.h[a_axis], idx, sc.inv_h);
    const double ha_q = (double)Store<S>::ld(f.h[a_axis], idx - st[b_axis], sc.inv_h);
    const double hb_p = (double)Store<S>::ld(f.h[b_axis], idx, sc.inv_h);
    const double hb_q = (double)Store<S>::ld(f.h[b_axis], idx - st[a_axis], sc.inv_h);
    const double e_cur = (double)Store<S>::ld(f.e[comp], idx, sc.inv_e);
    double j_sum = 0;
    for (int p = 0; p < A.np; ++p) {
      const int64_t pj = ((int64_t)comp * A.np + p) * A.n + c;
      const int64_t pc = (int64_t)p * A.n + c;
      if (A.kind == 3)
        j_sum += coef(A.cjsj, pc, 2 + 7 * p + 2) * A.jc[pj] + coef(A.cjsjp, pc, 2 + 7 * p + 6) * A.jp[pj];
      else
        j_sum += coef(A.cjsj, pc, 2 + 7 * p + 2) * A.jc[pj];
    }
    double e_prev = 0.0, ceep = 0.0;
    if (A.kind == 3) {
      e_prev = A.eprev[(int64_t)comp * A.n + c];
      ceep = coef(A.ceep, c, 1);
    }
    const double tc = (double)t[3 * comp], ta = (double)t[3 * comp + 1],
                 tb = (double)t[3 * comp + 2];
    const double e_new = ceep * e_prev + (tc * e_cur + ta * (ha_p - ha_q) + tb * (hb_p - hb_q)) +
                         coef(A.cejs, c, 0) * j_sum;
    Store<S>::st(f.e[comp], idx, (C)e_new, sc.s_e);
    for (int p = 0; p < A.np; ++p) {
      const int64_t pj = ((int64_t)comp * A.np + p) * A.n + c;
      const int64_t pc = (int64_t)p * A.n + c;
      if (A.kind == 1) {
        A.jc[pj] = coef(A.cjj, pc, 2 + 7 * p) * A.jc[pj] + coef(A.cje, pc, 2 + 7 * p + 1) * (e_new + e_cur);
      } else if (A.kind == 2) {
        A.jc[pj] = coef(A.cjj, pc, 2 + 7 * p) * A.jc[pj] + coef(A.cje, pc, 2 + 7 * p + 1) * (e_new - e_cur);
      } else {
        const double jn = coef(A.cjen, pc, 2 + 7 * p + 4) * e_new + coef(A.cje, pc, 2 + 7 * p + 1) * e_cur +
                          coef(A.cjep, pc, 2 + 7 * p + 5) * e_prev + coef(A.cjj, pc, 2 + 7 * p) * A.jc[pj] +
                          coef(A.cjjp, pc, 2 + 7 * p + 3) * A.jp[pj];
        A.jp[pj] = A.jc[pj];
        A.jc[pj] = jn;
      }
    }
    if (A.kind == 3) A.eprev[(int64_t)comp * A.n + c] = e_cur;
  }
}

__global__ void k_next_step(int* t) { *t += 1; }

// The watchdog: every stored field value is read back and any NaN or
// infinity lowers *first to the step t being checked (*first starts at
// UINT_MAX, set by k_watch_clear; a fill kernel, not a memset, so a caller
// may capture it).
__global__ void k_watch_clear(unsigned* __restrict__ first) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *first = 0xffffffffu;
}

template <typename S>
__global__ void __launch_bounds__(256)
    k_watch(Fields<S> f, int64_t total, unsigned t, unsigned* __restrict__ first) {
  using C = typename Store<S>::C;
  const S* arr[6] = {f.e[0], f.e[1], f.e[2], f.h[0], f.h[1], f.h[2]};
  bool bad = false;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int c = 0; c < 6; ++c) {
    const S* p = arr[c];
    if (!p) continue;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
      const C v = Store<S>::ld(p, i, (C)1);
      bad |= !isfinite(v);
    }
  }
  if (bad) atomicMin(first, t);
}

// Plain streaming copy, 16 bytes per lane, grid-stride, U independent loads
// in flight per lane, optionally non-temporal.
typedef double xf_d2 __attribute__((ext_vector_type(2)));
template <int U, bool NT>
__global__ void __launch_bounds__(256)
    xf_stream_copy_kernel(const xf_d2* __restrict__ src, xf_d2* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    xf_d2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = NT ? __builtin_nontemporal_load(&src[i + u * stride]) : src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NT)
        __builtin_nontemporal_store(v[u], &dst[i + u * stride]);
      else
        dst[i + u * stride] = v[u];
    }
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

// One 16-byte element per lane, one pass (the BabelStream copy shape).
template <bool NT>
__global__ void __launch_bounds__(256)
    xf_stream_copy1_kernel(const xf_d2* __restrict__ src, xf_d2* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (NT)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
  else
    dst[i] = src[i];
}

}  // namespace xf
