// xf_yee3d.inl -- the 3-D Yee passes (included by xf_kernels.inl).
//
// H pass: BasicUpdator::updateH (basic_updator.cuh:16-32) with hNext
// (update_scheme.cuh:58-63) for Hx, Hy, Hz fused, plus the H corrections of
// the PML faces the pass carries (pml_corrector.cu:180-270).
// E pass: BasicUpdator3D::updateE (basic_updator_3d.cu:10-34) with eNext
// (update_scheme.cuh:27-32, skipping a == 0 || b == 0) for Ex, Ey, Ez
// fused, plus the E corrections of the PML faces the pass carries
// (pml_corrector.cu:50-178).  The pass carries every face, or every face
// but the z faces, which k_pml_z then applies after it (they are the last
// corrections of a node either way).
//
// Work decomposition: a tile is blockDim.x (64, or 32) lanes along k x
// blockDim.y rows along j; each thread marches i through the tile's
// x-chunk, carrying the plane it
// has loaded as the next cell's i-1 (E pass: Hy, Hz) or i (H pass: Ey, Ez)
// neighbour, and issuing the loads of plane i+1 before computing plane i
// (two planes of loads in flight per wave).  Arithmetic and its order are
// the reference's; the file is built without FMA contraction.

// Whether a tile's in-march CPML corrections take c_psi from the table row
// (PmlFace::ta / tb): not with dense coefficients (no table), nor in the H
// pass of compensated bf16 storage, which it slowed from 1.70 to 2.21 ms at
// 536^3 (its E pass gained: 2.05 against 1.99 ms).
template <typename S, bool DN, bool ISE>
constexpr bool kCpsiRows = !DN && (ISE || !__is_same(S, Bf16c));

template <typename S>
struct HLoads {
  typename Store<S>::C ex_i, ex_k, ex_j, ey_k, ey_n, ez_j, ez_n, hx, hy, hz;
  int id;
};

template <typename S, typename I, bool NT>
__device__ __forceinline__ void h_load(HLoads<S>& L, const Fields<S>& f, const I* __restrict__ hid,
                                       int64_t idx, const Dims& d, typename Store<S>::C ie,
                                       typename Store<S>::C ih) {
  L.id = (int)gp(hid)[idx];
  L.ex_i = Store<S>::ld(f.e[0], idx, ie);
  L.ex_k = Store<S>::ld(f.e[0], idx + 1, ie);
  L.ex_j = Store<S>::ld(f.e[0], idx + d.sy, ie);
  L.ey_k = Store<S>::ld(f.e[1], idx + 1, ie);
  L.ey_n = Store<S>::ld(f.e[1], idx + d.sx, ie);
  L.ez_j = Store<S>::ld(f.e[2], idx + d.sy, ie);
  L.ez_n = Store<S>::ld(f.e[2], idx + d.sx, ie);
  L.hx = NT ? Store<S>::ld_nt(f.h[0], idx, ih) : Store<S>::ld(f.h[0], idx, ih);
  L.hy = NT ? Store<S>::ld_nt(f.h[1], idx, ih) : Store<S>::ld(f.h[1], idx, ih);
  L.hz = NT ? Store<S>::ld_nt(f.h[2], idx, ih) : Store<S>::ld(f.h[2], idx, ih);
}

template <typename S>
struct ELoads {
  typename Store<S>::C ex, ey, ez, hx, hx_k, hx_j, hy, hy_k, hz, hz_j;
  int id;
};

template <typename S, typename I, bool NT>
__device__ __forceinline__ void e_load(ELoads<S>& L, const Fields<S>& f, const I* __restrict__ eid,
                                       int64_t idx, const Dims& d, bool j_in, bool k_in,
                                       typename Store<S>::C ie, typename Store<S>::C ih) {
  using C = typename Store<S>::C;
  L.id = (int)gp(eid)[idx];
  L.ex = NT ? Store<S>::ld_nt(f.e[0], idx, ie) : Store<S>::ld(f.e[0], idx, ie);
  L.ey = NT ? Store<S>::ld_nt(f.e[1], idx, ie) : Store<S>::ld(f.e[1], idx, ie);
  L.ez = NT ? Store<S>::ld_nt(f.e[2], idx, ie) : Store<S>::ld(f.e[2], idx, ie);
  L.hx = Store<S>::ld(f.h[0], idx, ih);
  L.hy = Store<S>::ld(f.h[1], idx, ih);
  L.hz = Store<S>::ld(f.h[2], idx, ih);
  L.hx_k = k_in ? Store<S>::ld(f.h[0], idx - 1, ih) : (C)0;
  L.hy_k = k_in ? Store<S>::ld(f.h[1], idx - 1, ih) : (C)0;
  L.hx_j = j_in ? Store<S>::ld(f.h[0], idx - d.sy, ih) : (C)0;
  L.hz_j = j_in ? Store<S>::ld(f.h[2], idx - d.sy, ih) : (C)0;
}

template <typename S, bool NT>
__device__ __forceinline__ void st3(S* const* dst, int64_t idx, const typename Store<S>::C (&v)[3],
                                    const bool (&dirty)[3], typename Store<S>::C scale) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!dirty[c]) continue;
    if (NT)
      Store<S>::st_nt(dst[c], idx, v[c], scale);
    else
      Store<S>::st(dst[c], idx, v[c], scale);
  }
}

// Faces of the pass whose corrected box the column (j, k) can meet for some
// i in [i0, i1): a superset of what pml_face accepts (lateral ranges taken
// with their one-node extension).  Worked out once per tile, it lets the
// interior columns skip the CPML code of every plane.
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

// The z-face CPML corrections of the tile's column (j, k) over [i0, i1),
// applied after its march (XF_TILE_PML_ZTAIL): the lanes whose column lies
// in a z slab revisit their nodes, which this lane stored itself, so the
// lines are still in cache, instead of a separate launch fetching them
// again.  Per node the work and its order are k_pml_z's (pml_z_node): the
// x and y corrections came in the march, the z faces follow in
// registration order; the dual field is not written in this pass.
// USE_TAB: c_psi from the node's table row where the face allows it (the
// node's id read again here; 0 where the node is not a cell, as stored).
template <bool ISE, typename S, bool USE_TAB = false, int TAB = TAB_ANY, typename I = uint8_t>
__device__ __forceinline__ void pml_z_tail(const PmlZTail<typename Store<S>::C>& Z, int i0, int i1, int j,
                                           int k, int64_t idx0, const Fields<S>& f, const Dims& d,
                                           const Scales<typename Store<S>::C>& sc,
                                           const typename Store<S>::C* tab = nullptr,
                                           const I* __restrict__ ids = nullptr) {
  using C = typename Store<S>::C;
  auto meets = [&](const PmlFace<C>& F) {
    return k >= F.c0 && k < F.c1 && F.a0 < i1 && F.a1 >= i0 && j >= F.b0 && j <= F.b1;
  };
  const bool m0 = Z.n > 0 && meets(Z.f[0]);
  const bool m1 = Z.n > 1 && meets(Z.f[1]);
  if (!m0 && !m1) return;
  S* const* self = ISE ? f.e : f.h;
  const C s = ISE ? sc.s_e : sc.s_h;
  int64_t idx = idx0;
  for (int i = i0; i < i1; ++i, idx += d.sx) {
    C v[3] = {0, 0, 0};
    bool have[3] = {false, false, false};
    bool dirty[3] = {false, false, false};
    const int id = USE_TAB ? (int)gp(ids)[idx] : 0;
    // static face indices keep the descriptor out of scratch
    if (m0) pml_face<2, ISE, false, USE_TAB, TAB>(Z.f[0], i, j, k, idx, f, d, sc, v, have, dirty, tab, id);
    if (m1) pml_face<2, ISE, false, USE_TAB, TAB>(Z.f[1], i, j, k, idx, f, d, sc, v, have, dirty, tab, id);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (dirty[c]) Store<S>::st(self[c], idx, v[c], s);
  }
}

// The z = const TF/SF corrections of the tile's column (j, k) over
// [i0, i1), applied after its march (XF_TILE_TFSF_ZTAIL) instead of by
// k_tfsf, whose threads would each fetch a line for one node of such a
// face.  The host hands over only nodes that no other face of the same
// component writes, so each gets exactly the one correction k_tfsf would
// add (tfsf_cell), after the node's CPML corrections as there.
template <bool ISE, typename S>
__device__ __forceinline__ void tfsf_z_tail(const TfsfZTail<typename Store<S>::C>& Z, int i0, int i1, int j,
                                            int k, int64_t idx0, const Fields<S>& f, const Dims& d,
                                            const Scales<typename Store<S>::C>& sc) {
  using C = typename Store<S>::C;
  const C inv = ISE ? sc.inv_e : sc.inv_h;
  const C s = ISE ? sc.s_e : sc.s_h;
  const bool inc_e = !ISE;  // the H pass adds incident E, the E pass incident H
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= Z.n) break;
    const TfsfZJob<C>& J = Z.z[q];
    if (k != J.k || j < J.j0 || j >= J.j1) continue;
    const int ia = max(i0, J.i0), ib = min(i1, J.i1);
    if (ia >= ib) continue;
    const int t = Z.t_dev ? *Z.t_dev : Z.t_arg;
    S* field = ISE ? f.e[J.comp] : f.h[J.comp];
    // tfsf_inc for this job with its j and k terms read once: the x term is
    // on the half grid where tfsf_inc takes it so (incident E along x, or
    // incident H across x), the sums in tfsf_inc's order
    const bool xh = inc_e ? J.dual == 0 : J.dual != 0;
    const bool yh = inc_e ? J.dual == 1 : J.dual != 1;
    const __attribute__((address_space(1))) C* px = gp(xh ? Z.proj_half[0] + 1 : Z.proj_int[0]) - Z.g[0];
    const int lj = j - Z.g[1], lk = J.kinc - Z.g[2];
    const C py = yh ? gp(Z.proj_half[1])[lj + 1] : gp(Z.proj_int[1])[lj];
    const C pz = inc_e ? gp(Z.proj_int[2])[lk] : gp(Z.proj_half[2])[lk + 1];
    const __attribute__((address_space(1))) C* row = gp(Z.inc) + (int64_t)t * Z.n_aux;
    const C tr = Z.tr[J.dual];
    // four planes at a time, every load issued before the first use (the
    // last plane repeated where fewer are left; only real ones are stored)
    for (int i = ia; i < ib; i += 4) {
      C v[4], w0[4], w1[4], wt[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int iu = min(i + u, ib - 1);
        v[u] = Store<S>::ld(field, idx0 + (int64_t)(iu - i0) * d.sx, inv);
        C proj = px[iu] + py + pz;
        if (!inc_e) proj = proj - (C)0.5;
        const uint64_t index = (uint64_t)proj;
        wt[u] = proj - (C)index;
        w0[u] = row[index];
        w1[u] = row[index + 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (i + u >= ib) break;
        const C inc = tr * ((1 - wt[u]) * w0[u] + wt[u] * w1[u]);
        const C val = v[u] + inc * J.coef;
        Store<S>::st(field, idx0 + (int64_t)(i + u - i0) * d.sx, val, s);
      }
    }
  }
}

// One tile of the H pass.  Lanes beyond the grid return at once; the tile
// bodies hold no barrier, so the batched kernels may loop over tiles with
// some lanes idle.
template <typename S, typename I, int PML, bool NT, bool DN = false, int TAB = TAB_ANY>
__device__ __forceinline__ void yee_h3_tile(const YeePass<S>& D, const typename Store<S>::C* tab,
                                            int kb, int jb, int ch) {
  using C = typename Store<S>::C;
  const Dims& d = D.d;
  const int k = kb * blockDim.x + threadIdx.x;  // 64 (or 32) k-lanes per row
  const int j = jb * blockDim.y + threadIdx.y;
  if (k > d.nz || j > d.ny) return;
  const I* __restrict__ hid = static_cast<const I*>(D.ids);
  const Fields<S>& f = D.f;
  const Scales<C>& sc = D.sc;
  const int i0 = ch * D.chunk;
  const int i1 = min(i0 + D.chunk, d.nx + 1);
  const bool jk_main = j < d.ny && k < d.nz;
  const unsigned pmask = PML ? pml_mask(D.pml, i0, i1, j, k) : 0u;
  const bool any_pml = pmask != 0u;
  int64_t idx = (int64_t)i0 * d.sx + (int64_t)j * d.sy + k;
  const int64_t idx0 = idx;
  const C ie = sc.inv_e, ih = sc.inv_h;
  C ey_i = 0, ez_i = 0;
  HLoads<S> cur{}, nxt{};
  if (jk_main && i0 < d.nx) {
    ey_i = Store<S>::ld(f.e[1], idx, ie);
    ez_i = Store<S>::ld(f.e[2], idx, ie);
    h_load<S, I, NT>(cur, f, hid, idx, d, ie, ih);
  }
  for (int i = i0; i < i1; ++i, idx += d.sx) {
    if (jk_main && i + 1 < d.nx && i + 1 < i1) h_load<S, I, NT>(nxt, f, hid, idx + d.sx, d, ie, ih);
    C v[3] = {0, 0, 0};
    bool have[3] = {false, false, false};
    bool dirty[3] = {false, false, false};
    if (jk_main && i < d.nx) {
      if constexpr (DN) {
        C t[9];
#pragma unroll
        for (int n = 0; n < 9; ++n) t[n] = gp(D.dense[n])[idx];
        v[0] = t[0] * cur.hx + t[1] * (cur.ey_k - ey_i) + t[2] * (cur.ez_j - ez_i);
        v[1] = t[3] * cur.hy + t[4] * (cur.ez_n - ez_i) + t[5] * (cur.ex_k - cur.ex_i);
        v[2] = t[6] * cur.hz + t[7] * (cur.ex_j - cur.ex_i) + t[8] * (cur.ey_n - ey_i);
      } else {
        // the whole row at once (the three components all update here)
        C t[9];
        tab_row<TAB>(tab, cur.id, t);
        v[0] = t[0] * cur.hx + t[1] * (cur.ey_k - ey_i) + t[2] * (cur.ez_j - ez_i);
        v[1] = t[3] * cur.hy + t[4] * (cur.ez_n - ez_i) + t[5] * (cur.ex_k - cur.ex_i);
        v[2] = t[6] * cur.hz + t[7] * (cur.ex_j - cur.ex_i) + t[8] * (cur.ey_n - ey_i);
      }
      have[0] = have[1] = have[2] = true;
      dirty[0] = dirty[1] = dirty[2] = true;
      ey_i = cur.ey_n;
      ez_i = cur.ez_n;
    }
    // the node's row where it is a cell (id 0 elsewhere, as in the padded
    // id array): the row the c_psi match was made against
    if (PML && any_pml)
      pml_all<false, false, kCpsiRows<S, DN, false>, TAB>(D.pml, i, j, k, idx, f, d, sc, v, have, dirty, pmask, tab,
                                   (jk_main && i < d.nx) ? cur.id : 0);
    st3<S, NT>(f.h, idx, v, dirty, sc.s_h);
    cur = nxt;
  }
  if constexpr (PML == 2) {
    if (D.zt.n > 0) pml_z_tail<false, S, !DN, TAB>(D.zt, i0, i1, j, k, idx0, f, d, sc, tab, hid);
    if (D.tz.n > 0) tfsf_z_tail<false, S>(D.tz, i0, i1, j, k, idx0, f, d, sc);
  }
}

// One tile of the E pass; table rows flagged dispersive are left to the
// ADE kernel.
template <typename S, typename I, int PML, bool ADE, bool NT, bool DN = false, int TAB = TAB_ANY>
__device__ __forceinline__ void yee_e3_tile(const YeePass<S>& D, const typename Store<S>::C* tab,
                                            int kb, int jb, int ch) {
  using C = typename Store<S>::C;
  const Dims& d = D.d;
  const int k = kb * blockDim.x + threadIdx.x;  // 64 (or 32) k-lanes per row
  const int j = jb * blockDim.y + threadIdx.y;
  if (k > d.nz || j > d.ny) return;
  const I* __restrict__ eid = static_cast<const I*>(D.ids);
  const Fields<S>& f = D.f;
  const Scales<C>& sc = D.sc;
  const int i0 = ch * D.chunk;
  const int i1 = min(i0 + D.chunk, d.nx + 1);
  const bool jk_main = j < d.ny && k < d.nz;
  const bool j_in = j > 0, k_in = k > 0;
  const unsigned pmask = PML ? pml_mask(D.pml, i0, i1, j, k) : 0u;
  const bool any_pml = pmask != 0u;
  int64_t idx = (int64_t)i0 * d.sx + (int64_t)j * d.sy + k;
  const int64_t idx0 = idx;
  const C ie = sc.inv_e, ih = sc.inv_h;
  C hy_m = 0, hz_m = 0;  // H at i-1
  ELoads<S> cur{}, nxt{};
  if (jk_main && i0 < d.nx) {
    if (i0 > 0) {
      hy_m = Store<S>::ld(f.h[1], idx - d.sx, ih);
      hz_m = Store<S>::ld(f.h[2], idx - d.sx, ih);
    }
    e_load<S, I, NT>(cur, f, eid, idx, d, j_in, k_in, ie, ih);
  }
  for (int i = i0; i < i1; ++i, idx += d.sx) {
    if (jk_main && i + 1 < d.nx && i + 1 < i1)
      e_load<S, I, NT>(nxt, f, eid, idx + d.sx, d, j_in, k_in, ie, ih);
    C v[3] = {0, 0, 0};
    bool have[3] = {false, false, false};
    bool dirty[3] = {false, false, false};
    if (jk_main && i < d.nx) {
      if (!(ADE && D.eflag[cur.id])) {
        // coefficient n of the cell, read where its component is updated
        // (the E components have different boundary conditions)
        const int row = 9 * cur.id;
        auto t = [&](int n) -> C {
          if constexpr (DN)
            return gp(D.dense[n])[idx];
          else
            return tab_at<TAB>(tab, row + n);
        };
        if (j_in && k_in) {
          v[0] = t(0) * cur.ex + t(1) * (cur.hy - cur.hy_k) + t(2) * (cur.hz - cur.hz_j);
          have[0] = dirty[0] = true;
        }
        if (k_in && i > 0) {
          v[1] = t(3) * cur.ey + t(4) * (cur.hz - hz_m) + t(5) * (cur.hx - cur.hx_k);
          have[1] = dirty[1] = true;
        }
        if (i > 0 && j_in) {
          v[2] = t(6) * cur.ez + t(7) * (cur.hx - cur.hx_j) + t(8) * (cur.hy - hy_m);
          have[2] = dirty[2] = true;
        }
      }
      hy_m = cur.hy;
      hz_m = cur.hz;
    }
    // the node's row where it is a cell (id 0 elsewhere, as in the padded
    // id array): the row the c_psi match was made against
    if (PML && any_pml)
      pml_all<true, false, kCpsiRows<S, DN, true>, TAB>(D.pml, i, j, k, idx, f, d, sc, v, have, dirty, pmask, tab,
                                   (jk_main && i < d.nx) ? cur.id : 0);
    st3<S, NT>(f.e, idx, v, dirty, sc.s_e);
    cur = nxt;
  }
  if constexpr (PML == 2) {
    if (D.zt.n > 0) pml_z_tail<true, S, !DN, TAB>(D.zt, i0, i1, j, k, idx0, f, d, sc, tab, eid);
    if (D.tz.n > 0) tfsf_z_tail<true, S>(D.tz, i0, i1, j, k, idx0, f, d, sc);
  }
}

// One grid: the descriptor is a kernel argument, block -> tile by
// tile_coords (optionally XCD-aware).
// DN: coefficients from the dense per-cell arrays (the ids then only carry
// the E pass's dispersive flag) instead of the id table.
template <typename S, typename I, bool LDS, int PML, bool NT, bool DN = false>
__global__ void __launch_bounds__(256) k_yee_h3(YeePass<S> D, int xcd) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = DN ? nullptr : load_table<C, LDS>(D.tab, D.ntab * 9, reinterpret_cast<C*>(smem_raw));
  Tiling T{D.nkb, D.njb, D.nch, D.chunk, xcd};
  int kb, jb, ch;
  tile_coords(T, kb, jb, ch);
  yee_h3_tile<S, I, PML, NT, DN>(D, tab, kb, jb, ch);
}

template <typename S, typename I, bool LDS, int PML, bool ADE, bool NT, bool DN = false>
__global__ void __launch_bounds__(256) k_yee_e3(YeePass<S> D, int xcd) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = DN ? nullptr : load_table<C, LDS>(D.tab, D.ntab * 9, reinterpret_cast<C*>(smem_raw));
  Tiling T{D.nkb, D.njb, D.nch, D.chunk, xcd};
  int kb, jb, ch;
  tile_coords(T, kb, jb, ch);
  yee_e3_tile<S, I, PML, ADE, NT, DN>(D, tab, kb, jb, ch);
}

// Tiles taken from a list (tile id kb + nkb (jb + njb ch)): the engine
// launches a pass's CPML-free tiles and its CPML tiles apart, on two streams,
// so the CPML-free ones run in the leaner PML = false instantiation.
template <typename S, typename I, bool LDS, int PML, bool NT>
__global__ void __launch_bounds__(256) k_yee_h3_list(YeePass<S> D, const int* __restrict__ list) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = load_table<C, LDS>(D.tab, D.ntab * 9, reinterpret_cast<C*>(smem_raw));
  const int t = list[blockIdx.x];
  const int kb = t % D.nkb, rest = t / D.nkb;
  yee_h3_tile<S, I, PML, NT>(D, tab, kb, rest % D.njb, rest / D.njb);
}

template <typename S, typename I, bool LDS, int PML, bool ADE, bool NT>
__global__ void __launch_bounds__(256) k_yee_e3_list(YeePass<S> D, const int* __restrict__ list) {
  using C = typename Store<S>::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* tab = load_table<C, LDS>(D.tab, D.ntab * 9, reinterpret_cast<C*>(smem_raw));
  const int t = list[blockIdx.x];
  const int kb = t % D.nkb, rest = t / D.nkb;
  yee_e3_tile<S, I, PML, ADE, NT>(D, tab, kb, rest % D.njb, rest / D.njb);
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

// The tiles a persistent batch block takes.  xcd = 0: block b takes tiles
// b, b + G, b + 2G, ... (G blocks).  xcd = 1: blocks are dealt to the 8 XCDs
// round-robin by index, so the blocks b = x (mod 8) of XCD x take the
// contiguous eighth [x S, x S + S) of the tile list (S = ceil(T / 8)), with
// their count as stride: tiles next to each other, which read each other's
// halo rows, then go through one XCD's L2.  Every tile is taken once, and
// every block ends when its range does.
struct BatchWalk {
  int first, end, stride;
};
__device__ __forceinline__ BatchWalk batch_walk(int total, int xcd) {
  if (!xcd || gridDim.x < 8) return BatchWalk{(int)blockIdx.x, total, (int)gridDim.x};
  const int x = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int per = ((int)gridDim.x - x + 7) >> 3;
  const int span = (total + 7) >> 3;
  const int lo = min(x * span, total), hi = min(lo + span, total);
  return BatchWalk{lo + r, hi, per};
}

// Each wave keeps the table of the grid it is working on in its own slice
// of LDS (kBatchTab values, written by the wave when its grid changes; no
// block barrier), read like the single-grid kernels' staged table; grids
// with larger tables read theirs through L1.
constexpr int kBatchTab = 9 * 16;

template <typename C>
__device__ __forceinline__ const C* batch_table(const C* __restrict__ g_tab, int ntab, C* mine, int g,
                                                int& cur_g) {
  if (ntab * 9 > kBatchTab) return g_tab;
  if (g != cur_g) {
    const int lane = (threadIdx.x + blockDim.x * threadIdx.y) & 63;
    __builtin_amdgcn_wave_barrier();  // the wave's reads of the previous table are issued
    for (int n = lane; n < ntab * 9; n += 64) mine[n] = gp(g_tab)[n];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    cur_g = g;
  }
  return mine;
}

template <typename S, typename I, int PML, bool NT>
__global__ void __launch_bounds__(256)
    k_yee_h3_batch(const YeePass<S>* __restrict__ passes, const int* __restrict__ tile_start, int n,
                     int xcd) {
  using C = typename Store<S>::C;
  __shared__ C wtab[4][kBatchTab];
  C* mine = wtab[(threadIdx.x + blockDim.x * threadIdx.y) >> 6];
  int cur_g = -1;
  const BatchWalk w = batch_walk(tile_start[n], xcd);
  for (int t = w.first; t < w.end; t += w.stride) {
    const int g = batch_instance(tile_start, n, t);
    const YeePass<S>& D = passes[g];
    const C* tab = batch_table<C>(D.tab, D.ntab, mine, g, cur_g);
    const int local = t - tile_start[g];
    const int kb = local % D.nkb;
    const int rest = local / D.nkb;
    if (tab == mine)
      yee_h3_tile<S, I, PML, NT, false, TAB_LDS>(D, tab, kb, rest % D.njb, rest / D.njb);
    else
      yee_h3_tile<S, I, PML, NT, false, TAB_GLOBAL>(D, tab, kb, rest % D.njb, rest / D.njb);
  }
}

template <typename S, typename I, int PML, bool NT>
__global__ void __launch_bounds__(256)
    k_yee_e3_batch(const YeePass<S>* __restrict__ passes, const int* __restrict__ tile_start, int n,
                     int xcd) {
  using C = typename Store<S>::C;
  __shared__ C wtab[4][kBatchTab];
  C* mine = wtab[(threadIdx.x + blockDim.x * threadIdx.y) >> 6];
  int cur_g = -1;
  const BatchWalk w = batch_walk(tile_start[n], xcd);
  for (int t = w.first; t < w.end; t += w.stride) {
    const int g = batch_instance(tile_start, n, t);
    const YeePass<S>& D = passes[g];
    const C* tab = batch_table<C>(D.tab, D.ntab, mine, g, cur_g);
    const int local = t - tile_start[g];
    const int kb = local % D.nkb;
    const int rest = local / D.nkb;
    if (tab == mine)
      yee_e3_tile<S, I, PML, false, NT, false, TAB_LDS>(D, tab, kb, rest % D.njb, rest / D.njb);
    else
      yee_e3_tile<S, I, PML, false, NT, false, TAB_GLOBAL>(D, tab, kb, rest % D.njb, rest / D.njb);
  }
}
