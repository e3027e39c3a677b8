// xf_yee3d.inl -- the 3-D Yee passes (included by xf_kernels.inl).
//
// H pass: BasicUpdator::updateH (basic_updator.cuh:16-32) with hNext
// (update_scheme.cuh:58-63) for Hx, Hy, Hz fused, plus the H corrections of
// every PML face (pml_corrector.cu:180-270).
// E pass: BasicUpdator3D::updateE (basic_updator_3d.cu:10-34) with eNext
// (update_scheme.cuh:27-32, skipping a == 0 || b == 0) for Ex, Ey, Ez
// fused, plus the E corrections of every PML face (pml_corrector.cu:50-178).
//
// Work decomposition: a tile is 64 lanes along k x blockDim.y rows along j
// (one wave per row: blocks are 64 wide, so j is uniform across a wave);
// each wave marches i through the tile's x-chunk.
//  * Addressing: per plane, every array is read through a wave-uniform
//    pointer (SGPRs) at the lane index, and neighbours are uniform pointer
//    offsets (+-1 along k, +-sy along j, +-sx along i), so a load costs no
//    vector address arithmetic.
//  * Pipelining: two register sets alternate (plane i computed from one
//    while plane i+1 loads into the other), with no copy between them.
//  * The marching neighbour (E pass: Hy, Hz at i-1; H pass: Ey, Ez at i) is
//    carried from the previous plane.
//  * CPML: each lane works out once per tile which faces' corrected boxes
//    its column can meet in the chunk; cells of the interior skip the CPML
//    code entirely.
// Arithmetic and its order are the reference's; the file is built without
// FMA contraction.

template <typename S>
struct HLoads {
  typename Store<S>::C ex_i, ex_k, ex_j, ey_k, ey_n, ez_j, ez_n, hx, hy, hz;
  int id;
};

// Plane loads of the H pass at uniform offset `base` (lane 0's index); `o`
// is the lane's element offset from it.
template <typename S, typename I, bool NT>
__device__ __forceinline__ void h_load(HLoads<S>& L, const Fields<S>& f, const I* __restrict__ hid,
                                       int64_t base, uint32_t o, const Dims& d,
                                       typename Store<S>::C ie, typename Store<S>::C ih) {
  const S* e0 = f.e[0] + base;
  const S* e1 = f.e[1] + base;
  const S* e2 = f.e[2] + base;
  L.id = (int)(hid + base)[(uint64_t)o];
  L.ex_i = ld32(e0, o, ie);
  L.ex_k = ld32(e0 + 1, o, ie);
  L.ex_j = ld32(e0 + d.sy, o, ie);
  L.ey_k = ld32(e1 + 1, o, ie);
  L.ey_n = ld32(e1 + d.sx, o, ie);
  L.ez_j = ld32(e2 + d.sy, o, ie);
  L.ez_n = ld32(e2 + d.sx, o, ie);
  L.hx = NT ? ld32_nt(f.h[0] + base, o, ih) : ld32(f.h[0] + base, o, ih);
  L.hy = NT ? ld32_nt(f.h[1] + base, o, ih) : ld32(f.h[1] + base, o, ih);
  L.hz = NT ? ld32_nt(f.h[2] + base, o, ih) : ld32(f.h[2] + base, o, ih);
}

template <typename S>
struct ELoads {
  typename Store<S>::C ex, ey, ez, hx, hx_k, hx_j, hy, hy_k, hz, hz_j;
  int id;
};

template <typename S, typename I, bool NT>
__device__ __forceinline__ void e_load(ELoads<S>& L, const Fields<S>& f, const I* __restrict__ eid,
                                       int64_t base, uint32_t o, const Dims& d, bool j_in, bool k_in,
                                       typename Store<S>::C ie, typename Store<S>::C ih) {
  using C = typename Store<S>::C;
  const S* h0 = f.h[0] + base;
  const S* h1 = f.h[1] + base;
  const S* h2 = f.h[2] + base;
  L.id = (int)(eid + base)[(uint64_t)o];
  L.ex = NT ? ld32_nt(f.e[0] + base, o, ie) : ld32(f.e[0] + base, o, ie);
  L.ey = NT ? ld32_nt(f.e[1] + base, o, ie) : ld32(f.e[1] + base, o, ie);
  L.ez = NT ? ld32_nt(f.e[2] + base, o, ie) : ld32(f.e[2] + base, o, ie);
  L.hx = ld32(h0, o, ih);
  L.hy = ld32(h1, o, ih);
  L.hz = ld32(h2, o, ih);
  // k - 1 and j - 1 neighbours: read from an in-bounds address whatever the
  // lane and zeroed where they do not exist, so the loads need no branch.
  // Along k the pointer is one cell back for the whole wave and a lane
  // without a k - 1 neighbour reads its own cell (offset o + 1); along j
  // (uniform) row 0 reads itself.
  const uint32_t ok = k_in ? o : o + 1u;
  const int64_t sj = j_in ? d.sy : 0;
  const C hx_k = ld32(h0 - 1, ok, ih);
  const C hy_k = ld32(h1 - 1, ok, ih);
  const C hx_j = ld32(h0 - sj, o, ih);
  const C hz_j = ld32(h2 - sj, o, ih);
  L.hx_k = k_in ? hx_k : (C)0;
  L.hy_k = k_in ? hy_k : (C)0;
  L.hx_j = j_in ? hx_j : (C)0;
  L.hz_j = j_in ? hz_j : (C)0;
}

template <typename S, bool NT>
__device__ __forceinline__ void st3(S* const* dst, int64_t base, uint32_t o,
                                    const typename Store<S>::C (&v)[3], const bool (&dirty)[3],
                                    typename Store<S>::C scale) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!dirty[c]) continue;
    if (NT)
      st32_nt(dst[c] + base, o, v[c], scale);
    else
      st32(dst[c] + base, o, v[c], scale);
  }
}

// Faces of the pass whose corrected box the column (j, k) can meet for some
// i in [i0, i1): a superset of what pml_face accepts (lateral ranges taken
// with their one-node extension).
template <typename C>
__device__ __forceinline__ unsigned pml_mask(const PmlPass<C>& P, int i0, int i1, int j, int k) {
  unsigned bits = 0;
  for (int q = 0; q < P.n; ++q) {
    const PmlFace<C>& F = P.f[q];
    bool hit;
    if (F.axis == 0)  // c = i, a = j, b = k
      hit = F.c0 < i1 && F.c1 > i0 && j >= F.a0 && j <= F.a1 && k >= F.b0 && k <= F.b1;
    else if (F.axis == 1)  // c = j, a = k, b = i
      hit = j >= F.c0 && j < F.c1 && k >= F.a0 && k <= F.a1 && F.b0 < i1 && F.b1 >= i0;
    else  // c = k, a = i, b = j
      hit = k >= F.c0 && k < F.c1 && F.a0 < i1 && F.a1 >= i0 && j >= F.b0 && j <= F.b1;
    if (hit) bits |= 1u << q;
  }
  return bits;
}

template <bool ISE, typename S, typename C>
__device__ __forceinline__ void pml_some(const PmlPass<C>& P, unsigned bits, int i, int j, int k,
                                         int64_t idx, const Fields<S>& f, const Dims& d,
                                         const Scales<C>& sc, C (&v)[3], bool (&have)[3],
                                         bool (&dirty)[3]) {
  for (int q = 0; q < P.n; ++q) {
    if (!((bits >> q) & 1u)) continue;
    const PmlFace<C>& F = P.f[q];
    if (F.axis == 0)
      pml_face<0, ISE, false>(F, i, j, k, idx, f, d, sc, v, have, dirty);
    else if (F.axis == 1)
      pml_face<1, ISE, false>(F, i, j, k, idx, f, d, sc, v, have, dirty);
    else
      pml_face<2, ISE, false>(F, i, j, k, idx, f, d, sc, v, have, dirty);
  }
}

// Plane i of the H pass from loads L; ey_i/ez_i carry E at i and become E
// at i+1.
template <typename S, bool PML, bool NT>
__device__ __forceinline__ void h_plane(const YeePass<S>& D, const typename Store<S>::C* tab,
                                        const HLoads<S>& L, int i, int j, int k, int lane,
                                        int64_t base, bool jk_main, unsigned pml_bits,
                                        typename Store<S>::C& ey_i, typename Store<S>::C& ez_i) {
  using C = typename Store<S>::C;
  C v[3] = {0, 0, 0};
  bool have[3] = {false, false, false};
  bool dirty[3] = {false, false, false};
  if (jk_main && i < D.d.nx) {
    const C* t = tab + 9 * L.id;
    v[0] = t[0] * L.hx + t[1] * (L.ey_k - ey_i) + t[2] * (L.ez_j - ez_i);
    v[1] = t[3] * L.hy + t[4] * (L.ez_n - ez_i) + t[5] * (L.ex_k - L.ex_i);
    v[2] = t[6] * L.hz + t[7] * (L.ex_j - L.ex_i) + t[8] * (L.ey_n - ey_i);
    have[0] = have[1] = have[2] = true;
    dirty[0] = dirty[1] = dirty[2] = true;
    ey_i = L.ey_n;
    ez_i = L.ez_n;
  }
  if (PML && pml_bits)
    pml_some<false>(D.pml, pml_bits, i, j, k, base + lane, D.f, D.d, D.sc, v, have, dirty);
  st3<S, NT>(D.f.h, base, lane, v, dirty, D.sc.s_h);
}

// The march of one H-pass column through planes [i0, i1): two register
// sets alternate, plane i computed from one while plane i+1 loads into the
// other.
template <typename S, typename I, bool PML, bool NT>
__device__ __forceinline__ void h_march(const YeePass<S>& D, const typename Store<S>::C* tab,
                                        const I* __restrict__ hid, int64_t base, int i0, int i1,
                                        int j, int k, int lane, bool jk_main, unsigned pml_bits,
                                        typename Store<S>::C ie, typename Store<S>::C ih) {
  using C = typename Store<S>::C;
  const Dims& d = D.d;
  C ey_i = 0, ez_i = 0;
  HLoads<S> A{}, B{};
  // lanes off the updated cells load their row's first cell instead (in
  // bounds, ignored), so the loads of a plane need no per-lane branch
  const uint32_t ld_lane = jk_main ? (uint32_t)lane : 0u;
  // A plane that is not computed (i >= nx or past the chunk) is still
  // "loaded", from plane nx - 1 (whose i + 1 neighbour exists): the loads
  // then carry no branch and the wait counts stay exact across the loop.
  const int64_t last = (int64_t)(d.nx - 1) * d.sx + (base - (int64_t)i0 * d.sx);
  auto plane = [&](int ip, int64_t b) -> int64_t { return ip < d.nx && ip < i1 ? b : last; };
  ey_i = ld32(D.f.e[1] + plane(i0, base), ld_lane, ie);
  ez_i = ld32(D.f.e[2] + plane(i0, base), ld_lane, ie);
  h_load<S, I, NT>(A, D.f, hid, plane(i0, base), ld_lane, d, ie, ih);
  for (int i = i0; i < i1; i += 2) {
    h_load<S, I, NT>(B, D.f, hid, plane(i + 1, base + d.sx), ld_lane, d, ie, ih);
    h_plane<S, PML, NT>(D, tab, A, i, j, k, lane, base, jk_main, pml_bits, ey_i, ez_i);
    if (i + 1 >= i1) break;
    h_load<S, I, NT>(A, D.f, hid, plane(i + 2, base + 2 * d.sx), ld_lane, d, ie, ih);
    h_plane<S, PML, NT>(D, tab, B, i + 1, j, k, lane, base + d.sx, jk_main, pml_bits, ey_i, ez_i);
    base += 2 * d.sx;
  }
}

// One tile of the H pass.  Lanes beyond the grid return at once; the tile
// bodies hold no barrier, so the batched kernels may loop over tiles with
// some lanes idle.
template <typename S, typename I, bool PML, bool NT>
__device__ __forceinline__ void yee_h3_tile(const YeePass<S>& D, const typename Store<S>::C* tab,
                                            int kb, int jb, int ch) {
  using C = typename Store<S>::C;
  const Dims& d = D.d;
  const int lane = threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(jb * (int)blockDim.y + (int)threadIdx.y);
  const int k0 = kb * 64;
  const int k = k0 + lane;
  if (j > d.ny || k > d.nz) return;
  const I* __restrict__ hid = static_cast<const I*>(D.ids);
  const C ie = D.sc.inv_e, ih = D.sc.inv_h;
  const int i0 = ch * D.chunk;
  const int i1 = min(i0 + D.chunk, d.nx + 1);
  const bool jk_main = j < d.ny && k < d.nz;
  const unsigned pml_bits = PML ? pml_mask(D.pml, i0, i1, j, k) : 0u;
  const int64_t base = (int64_t)i0 * d.sx + (int64_t)j * d.sy + k0;
  // a wave whose columns meet no CPML box runs the march without the CPML
  // code (and its loads), which keeps the two-plane pipeline intact
  if (PML && __ballot(pml_bits != 0u) != 0)
    h_march<S, I, true, NT>(D, tab, hid, base, i0, i1, j, k, lane, jk_main, pml_bits, ie, ih);
  else
    h_march<S, I, false, NT>(D, tab, hid, base, i0, i1, j, k, lane, jk_main, 0u, ie, ih);
}

// Plane i of the E pass from loads L; hy_m/hz_m carry H at i-1 and become H
// at i.  Table rows flagged dispersive are left to the ADE kernel.
template <typename S, bool PML, bool ADE, bool NT>
__device__ __forceinline__ void e_plane(const YeePass<S>& D, const typename Store<S>::C* tab,
                                        const ELoads<S>& L, int i, int j, int k, int lane,
                                        int64_t base, bool jk_main, bool j_in, bool k_in,
                                        unsigned pml_bits, typename Store<S>::C& hy_m,
                                        typename Store<S>::C& hz_m) {
  using C = typename Store<S>::C;
  C v[3] = {0, 0, 0};
  bool have[3] = {false, false, false};
  bool dirty[3] = {false, false, false};
  if (jk_main && i < D.d.nx) {
    if (!(ADE && D.eflag[L.id])) {
      const C* t = tab + 9 * L.id;
      if (j_in && k_in) {
        v[0] = t[0] * L.ex + t[1] * (L.hy - L.hy_k) + t[2] * (L.hz - L.hz_j);
        have[0] = dirty[0] = true;
      }
      if (k_in && i > 0) {
        v[1] = t[3] * L.ey + t[4] * (L.hz - hz_m) + t[5] * (L.hx - L.hx_k);
        have[1] = dirty[1] = true;
      }
      if (i > 0 && j_in) {
        v[2] = t[6] * L.ez + t[7] * (L.hx - L.hx_j) + t[8] * (L.hy - hy_m);
        have[2] = dirty[2] = true;
      }
    }
    hy_m = L.hy;
    hz_m = L.hz;
  }
  if (PML && pml_bits)
    pml_some<true>(D.pml, pml_bits, i, j, k, base + lane, D.f, D.d, D.sc, v, have, dirty);
  st3<S, NT>(D.f.e, base, lane, v, dirty, D.sc.s_e);
}

template <typename S, typename I, bool PML, bool ADE, bool NT>
__device__ __forceinline__ void e_march(const YeePass<S>& D, const typename Store<S>::C* tab,
                                        const I* __restrict__ eid, int64_t base, int i0, int i1,
                                        int j, int k, int lane, bool jk_main, bool j_in, bool k_in,
                                        unsigned pml_bits, typename Store<S>::C ie,
                                        typename Store<S>::C ih) {
  using C = typename Store<S>::C;
  const Dims& d = D.d;
  C hy_m = 0, hz_m = 0;
  ELoads<S> A{}, B{};
  const uint32_t ld_lane = jk_main ? (uint32_t)lane : 0u;
  const bool k_ld = jk_main && k_in;  // k - 1 exists for the lane's load
  // planes not computed load plane nx - 1 instead (see h_march)
  const int64_t last = (int64_t)(d.nx - 1) * d.sx + (base - (int64_t)i0 * d.sx);
  auto plane = [&](int ip, int64_t b) -> int64_t { return ip < d.nx && ip < i1 ? b : last; };
  {
    const int64_t bm = i0 > 0 ? plane(i0 - 1, base - d.sx) : last;
    const C hy0 = ld32(D.f.h[1] + bm, ld_lane, ih);
    const C hz0 = ld32(D.f.h[2] + bm, ld_lane, ih);
    hy_m = i0 > 0 ? hy0 : (C)0;
    hz_m = i0 > 0 ? hz0 : (C)0;
  }
  e_load<S, I, NT>(A, D.f, eid, plane(i0, base), ld_lane, d, j_in, k_ld, ie, ih);
  for (int i = i0; i < i1; i += 2) {
    e_load<S, I, NT>(B, D.f, eid, plane(i + 1, base + d.sx), ld_lane, d, j_in, k_ld, ie, ih);
    e_plane<S, PML, ADE, NT>(D, tab, A, i, j, k, lane, base, jk_main, j_in, k_in, pml_bits, hy_m, hz_m);
    if (i + 1 >= i1) break;
    e_load<S, I, NT>(A, D.f, eid, plane(i + 2, base + 2 * d.sx), ld_lane, d, j_in, k_ld, ie, ih);
    e_plane<S, PML, ADE, NT>(D, tab, B, i + 1, j, k, lane, base + d.sx, jk_main, j_in, k_in, pml_bits,
                             hy_m, hz_m);
    base += 2 * d.sx;
  }
}

// One tile of the E pass.
template <typename S, typename I, bool PML, bool ADE, bool NT>
__device__ __forceinline__ void yee_e3_tile(const YeePass<S>& D, const typename Store<S>::C* tab,
                                            int kb, int jb, int ch) {
  using C = typename Store<S>::C;
  const Dims& d = D.d;
  const int lane = threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(jb * (int)blockDim.y + (int)threadIdx.y);
  const int k0 = kb * 64;
  const int k = k0 + lane;
  if (j > d.ny || k > d.nz) return;
  const I* __restrict__ eid = static_cast<const I*>(D.ids);
  const C ie = D.sc.inv_e, ih = D.sc.inv_h;
  const int i0 = ch * D.chunk;
  const int i1 = min(i0 + D.chunk, d.nx + 1);
  const bool jk_main = j < d.ny && k < d.nz;
  const bool j_in = j > 0, k_in = k > 0;
  const unsigned pml_bits = PML ? pml_mask(D.pml, i0, i1, j, k) : 0u;
  const int64_t base = (int64_t)i0 * d.sx + (int64_t)j * d.sy + k0;
  if (PML && __ballot(pml_bits != 0u) != 0)
    e_march<S, I, true, ADE, NT>(D, tab, eid, base, i0, i1, j, k, lane, jk_main, j_in, k_in, pml_bits, ie, ih);
  else
    e_march<S, I, false, ADE, NT>(D, tab, eid, base, i0, i1, j, k, lane, jk_main, j_in, k_in, 0u, ie, ih);
}

// One grid: the descriptor is a kernel argument, block -> tile by
// tile_coords (optionally XCD-aware).
template <typename S, typename I, bool LDS, bool PML, bool NT>
__global__ void __launch_bounds__(256) k_yee_h3(YeePass<S> D, int xcd) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = load_table<C, LDS>(D.tab, D.ntab * 9, reinterpret_cast<C*>(smem_raw));
  Tiling T{D.nkb, D.njb, D.nch, D.chunk, xcd};
  int kb, jb, ch;
  tile_coords(T, kb, jb, ch);
  yee_h3_tile<S, I, PML, NT>(D, tab, kb, jb, ch);
}

template <typename S, typename I, bool LDS, bool PML, bool ADE, bool NT>
__global__ void __launch_bounds__(256) k_yee_e3(YeePass<S> D, int xcd) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = load_table<C, LDS>(D.tab, D.ntab * 9, reinterpret_cast<C*>(smem_raw));
  Tiling T{D.nkb, D.njb, D.nch, D.chunk, xcd};
  int kb, jb, ch;
  tile_coords(T, kb, jb, ch);
  yee_e3_tile<S, I, PML, ADE, NT>(D, tab, kb, jb, ch);
}

// A batch of grids of different sizes in one persistent launch: the grid is
// sized to the chip's resident capacity and each block walks the
// concatenated tile list of all grids (tile t, t + gridDim.x, ...).
// tile_start[g] is the first tile of grid g (tile_start[n] = total).  All
// blocks run the same number of loop trips up to one, so the launch drains
// when the list is exhausted.
__device__ __forceinline__ int batch_instance(const int* __restrict__ tile_start, int n, int t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_start[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <typename S, typename I, bool PML, bool NT>
__global__ void __launch_bounds__(256)
    k_yee_h3_batch(const YeePass<S>* __restrict__ passes, const int* __restrict__ tile_start, int n) {
  const int total = tile_start[n];
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    const int g = batch_instance(tile_start, n, t);
    const YeePass<S>& D = passes[g];
    const int local = t - tile_start[g];
    const int kb = local % D.nkb;
    const int rest = local / D.nkb;
    yee_h3_tile<S, I, PML, NT>(D, D.tab, kb, rest % D.njb, rest / D.njb);
  }
}

template <typename S, typename I, bool PML, bool NT>
__global__ void __launch_bounds__(256)
    k_yee_e3_batch(const YeePass<S>* __restrict__ passes, const int* __restrict__ tile_start, int n) {
  const int total = tile_start[n];
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    const int g = batch_instance(tile_start, n, t);
    const YeePass<S>& D = passes[g];
    const int local = t - tile_start[g];
    const int kb = local % D.nkb;
    const int rest = local / D.nkb;
    yee_e3_tile<S, I, PML, false, NT>(D, D.tab, kb, rest % D.njb, rest / D.njb);
  }
}
