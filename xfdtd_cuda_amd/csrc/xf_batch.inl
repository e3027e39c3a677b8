// xf_batch.inl -- many independent 3-D grids stepped together, and the
// tile-shape autotuner (single grid and batch).  Included at the end of
// xf_engine.hip: each grid of a batch is a full Engine (its own fields,
// tables, CPML, TFSF and NF2FF state, uploaded and downloaded as usual); the
// batch only replaces the per-grid launches of a step by one launch per
// kernel over all grids.
//
// Work list.  The H (and E) pass of grid g consists of nkb * njb * nch tiles
// (64 k x tile_y j x `planes` i).  The tiles of all grids are concatenated
// (tile_start[g] = first tile of grid g) and a persistent grid of blocks,
// sized to what the device holds resident, walks the list with a stride of
// gridDim.x.  Consecutive tiles are neighbours in k, then j, so the blocks in
// flight at any moment cover a compact slab of one or two grids.
#include <climits>
#include <memory>

namespace xf {

struct Batch {
  std::vector<std::unique_ptr<xf_handle>> inst;
  xf_plan plan{};
  int device{0};
  hipStream_t stream{};
  bool uploaded{false};
  int64_t cur_step{0};
  int64_t nt{INT64_MAX};  // end of the shortest time table
  int id_bytes{1};
  int pml_level{0};  // 0 no CPML, 1 CPML faces in the march, 2 and z faces at the tile end
  int64_t cells{0};
  xf_tuning tune{1, 4, 0, 1, 0, 1, 0, 0.0};
  // device descriptors (sizes fixed by the scenes, rewritten by configure)
  std::vector<void*> allocs;
  void* pass_h{};
  void* pass_e{};
  int* tile_start{};
  int total_tiles{0};
  int grid_blocks{0};
  int resident{0};
  void* tf_items[2]{};
  int64_t* tf_start[2]{};
  int n_tf[2]{0, 0};
  int64_t tf_total[2]{0, 0};
  void* nf_items{};
  int64_t* nf_start{};
  int n_nf{0};
  int64_t nf_total{0};
  void* pz_items[2]{};     // z-face CPML items of the H and E passes
  int64_t* pz_start[2]{};
  int n_pz[2]{0, 0};
  int64_t pz_total[2]{0, 0};
  void* pt_items{};
  int n_pt{0};
  void* lb_items{};
  int64_t* lb_bstart{};  // first 256-column block of each layer (VALU phases)
  int64_t* lb_mstart{};  // first kLbMfmaCols-column block of each layer (k_lb_mfma_batch)
  int64_t* lb_estart{};  // first edge-list entry of each layer (k_lb_edges_batch)
  int64_t lb_edges{0};
  int n_lb{0};
  int64_t lb_blocks{0}, lb_mblocks{0};
  bool lb_any_record{false};
  int* t_dev{};
  hipGraph_t graph{};
  hipGraphExec_t graph_exec{};
  bool graph_ready{false};
  xf_timing timing{};
  std::string desc;

  ~Batch() { release(); }

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_ready = false;
  }

  void release() {
    drop_graph();
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    pass_h = pass_e = nullptr;
    tile_start = nullptr;
    tf_items[0] = tf_items[1] = nf_items = pt_items = lb_items = nullptr;
    tf_start[0] = tf_start[1] = nf_start = lb_bstart = lb_mstart = lb_estart = nullptr;
    pz_items[0] = pz_items[1] = nullptr;
    pz_start[0] = pz_start[1] = nullptr;
    t_dev = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    uploaded = false;
  }

  int alloc(void** p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(XF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(*p);
    return XF_OK;
  }

  // copies host bytes into *p, allocating it on first use
  int put(void** p, const void* src, size_t bytes) {
    if (!*p) XF_TRY(alloc(p, bytes));
    if (bytes) XF_HIP(hipMemcpy(*p, src, bytes, hipMemcpyHostToDevice));
    return XF_OK;
  }

  Engine& eng(size_t g) { return inst[g]->eng; }

  int create(xf_scene* const* scenes, int n, const xf_plan* pl);
  int upload();
  int configure();
  template <typename S>
  int configure_t();
  template <typename S>
  int launch_step(int t, bool use_tdev, bool ev_on, hipEvent_t* ev);
  int run(int64_t n);
  int download();
  int snapshot();
  int restore();
  int autotune(int trial_steps, xf_tuning* best);
  // every grid's step counter (and learned-layer recording count) to cur_step
  void sync_steps() {
    for (auto& h : inst) {
      Engine& E = h->eng;
      E.cur_step = cur_step;
      if (E.has_lb && E.lbd.mode == XF_LB_RECORD) E.lb_rec_count = E.lb_recorded_after(cur_step);
    }
  }
};

int Batch::create(xf_scene* const* scenes, int n, const xf_plan* pl) {
  if (n < 1) return fail(XF_ERR_INVALID, "a batch needs at least one scene");
  if (pl) plan = *pl;
  if (plan.precision < XF_FP64 || plan.precision > XF_FP16_STORE) return fail(XF_ERR_INVALID, "unknown precision");
  device = plan.device;
  xf_plan p = plan;
  p.use_graph = 0;
  p.timing = 0;
  // the batched Yee kernels read coefficient tables only
  if (p.coeff_mode == XF_COEFF_DENSE)
    return fail(XF_ERR_UNSUPPORTED, "batches take the coefficient table mode only");
  p.coeff_mode = XF_COEFF_TABLE;
  for (int g = 0; g < n; ++g) {
    if (!scenes[g]) return fail(XF_ERR_INVALID, "null scene " + std::to_string(g));
    xf_handle* h = nullptr;
    const int st = xf_create(scenes[g], &p, &h);
    if (st != XF_OK) return fail(st, "scene " + std::to_string(g) + ": " + g_err);
    inst.emplace_back(h);
    Engine& E = h->eng;
    if (E.dim != 3) return fail(XF_ERR_UNSUPPORTED, "scene " + std::to_string(g) + ": batches take 3-D grids only");
    if (E.has_ade) return fail(XF_ERR_UNSUPPORTED, "scene " + std::to_string(g) + ": ADE media are not batched");
    // point sources, the learned layer, movies and NF2FF time-domain boxes
    // run per grid inside the batch's step (Engine::launch_additions / launch_td)
    if (g == 0) cur_step = E.cur_step;
    if (E.cur_step != cur_step) return fail(XF_ERR_INVALID, "scenes of a batch must share the current step");
    if (E.sc->tfsf || E.sc->n_nf2ff > 0 || E.sc->n_point > 0 || E.sc->n_nf2ff_td > 0) nt = std::min(nt, E.nt);
    cells += E.nx * E.ny * E.nz;
  }
  return XF_OK;
}

int Batch::upload() {
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return fail(XF_ERR_HIP, "no HIP device");
  if (device < 0 || device >= nd) return fail(XF_ERR_INVALID, "bad device ordinal");
  release();
  for (auto& h : inst) XF_TRY(h->eng.upload());
  // one id width for the whole batch (the kernels are compiled per width)
  bool wide = false;
  for (auto& h : inst) wide = wide || h->eng.id_bytes == 2;
  if (wide) {
    for (auto& h : inst) {
      if (h->eng.id_bytes == 2) continue;
      h->eng.wide_ids = true;
      XF_TRY(h->eng.upload());
    }
  }
  id_bytes = wide ? 2 : 1;
  XF_HIP(hipSetDevice(device));
  XF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  void* tp;
  XF_TRY(alloc(&tp, sizeof(int)));
  t_dev = static_cast<int*>(tp);
  XF_TRY(configure());
  uploaded = true;
  return XF_OK;
}

namespace {
template <typename S, typename I, int PM, bool NT>
int batch_occupancy(int threads) {
  int a = 0, b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_yee_h3_batch<S, I, PM, NT>, threads, 0) != hipSuccess) a = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_yee_e3_batch<S, I, PM, NT>, threads, 0) != hipSuccess) b = 1;
  return std::max(1, std::min(a, b));
}

template <typename S, typename I, int PM>
int batch_occupancy_nt(bool nt, int threads) {
  return nt ? batch_occupancy<S, I, PM, true>(threads) : batch_occupancy<S, I, PM, false>(threads);
}

template <typename S, typename I>
int batch_occupancy_lv(int level, bool nt, int threads) {
  if (level == 2) return batch_occupancy_nt<S, I, 2>(nt, threads);
  if (level == 1) return batch_occupancy_nt<S, I, 1>(nt, threads);
  return batch_occupancy_nt<S, I, 0>(nt, threads);
}

template <typename S>
int batch_occupancy_any(int id_bytes, int level, bool nt, int threads) {
  return id_bytes == 1 ? batch_occupancy_lv<S, uint8_t>(level, nt, threads)
                       : batch_occupancy_lv<S, uint16_t>(level, nt, threads);
}
}  // namespace

template <typename S>
int Batch::configure_t() {
  using C = typename Store<S>::C;
  const int n = static_cast<int>(inst.size());
  std::vector<YeePass<S>> ph(n), pe(n);
  std::vector<int> ts(n + 1, 0);
  int64_t tiles = 0;
  pml_level = 0;
  const int flags = XF_TILE_SET | (tune.nontemporal ? XF_TILE_NT : 0) | (tune.pml_split == 1 ? XF_TILE_PML_SPLIT : 0) |
                    (tune.pml_split == 2 ? XF_TILE_PML_ZTAIL : 0) | (tune.k32 ? XF_TILE_K32 : 0);
  for (int g = 0; g < n; ++g) {
    Engine& E = eng(g);
    const int64_t chunks = (E.nx + 1 + tune.planes - 1) / tune.planes;
    E.set_tiling(tune.tile_y, static_cast<int>(chunks), flags);
    ph[g] = E.yee_pass<S>(false);
    pe[g] = E.yee_pass<S>(true);
    if (ph[g].zt.n > 0 || pe[g].zt.n > 0) pml_level = 2;
    else if ((ph[g].pml.n > 0 || pe[g].pml.n > 0) && pml_level == 0) pml_level = 1;
    ts[g] = static_cast<int>(tiles);
    tiles += static_cast<int64_t>(ph[g].nkb) * ph[g].njb * ph[g].nch;
    if (tiles > INT_MAX / 2) return fail(XF_ERR_UNSUPPORTED, "batch too large for 32-bit tile ids");
  }
  ts[n] = static_cast<int>(tiles);
  total_tiles = static_cast<int>(tiles);
  XF_TRY(put(&pass_h, ph.data(), ph.size() * sizeof(YeePass<S>)));
  XF_TRY(put(&pass_e, pe.data(), pe.size() * sizeof(YeePass<S>)));
  void* p = tile_start;
  XF_TRY(put(&p, ts.data(), ts.size() * sizeof(int)));
  tile_start = static_cast<int*>(p);
  // Per-item thread ranges of the one-launch passes below are padded to
  // whole 256-thread blocks, so each block serves one item and finds it
  // from its first thread (a block-uniform index: the item's descriptor is
  // then read with scalar loads).
  auto whole_blocks = [](int64_t count) { return (count + 255) / 256 * 256; };
  // z-face CPML corrections of the grids that split them off
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<PmlZItem<S>> items;
    std::vector<int64_t> st(1, 0);
    for (int g = 0; g < n; ++g) {
      Engine& E = eng(g);
      if (!E.split_z()) continue;
      PmlZItem<S> it{};
      it.f = make_fields<S>(E);
      it.sc = make_scales<C>(E);
      it.d = make_dims(E);
      fill_pml_pass<C>(E, it.P, pass == 1, 2);
      it.B = pml_z_box<C>(E, it.P);
      it.R = row_ref<C>(E, pass == 1);
      if (it.B.start[it.B.n] == 0) continue;
      items.push_back(it);
      st.push_back(st.back() + whole_blocks(it.B.start[it.B.n]));
    }
    n_pz[pass] = static_cast<int>(items.size());
    pz_total[pass] = st.back();
    // sized for every grid, so a later configure with more splitting grids fits
    std::vector<PmlZItem<S>> full(n);
    std::copy(items.begin(), items.end(), full.begin());
    std::vector<int64_t> stf(n + 1, st.back());
    std::copy(st.begin(), st.end(), stf.begin());
    XF_TRY(put(&pz_items[pass], full.data(), full.size() * sizeof(PmlZItem<S>)));
    void* q = pz_start[pass];
    XF_TRY(put(&q, stf.data(), stf.size() * sizeof(int64_t)));
    pz_start[pass] = static_cast<int64_t*>(q);
  }
  // TFSF passes and NF2FF boxes
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<TfsfItem<S>> items;
    std::vector<int64_t> st(1, 0);
    for (int g = 0; g < n; ++g) {
      Engine& E = eng(g);
      if (!E.sc->tfsf) continue;
      TfsfItem<S> it{};
      it.f = make_fields<S>(E);
      it.sc = make_scales<C>(E);
      it.d = make_dims(E);
      fill_tfsf_pass<C>(E, it.P, pass);
      if (it.P.total == 0) continue;
      items.push_back(it);
      st.push_back(st.back() + whole_blocks(it.P.total));
    }
    n_tf[pass] = static_cast<int>(items.size());
    tf_total[pass] = st.back();
    XF_TRY(put(&tf_items[pass], items.data(), items.size() * sizeof(TfsfItem<S>)));
    void* q = tf_start[pass];
    XF_TRY(put(&q, st.data(), st.size() * sizeof(int64_t)));
    tf_start[pass] = static_cast<int64_t*>(q);
  }
  {
    std::vector<NfItem<S>> items;
    std::vector<int64_t> st(1, 0);
    for (int g = 0; g < n; ++g) {
      Engine& E = eng(g);
      for (auto& N : E.nf) {
        if (N.pass.total == 0) continue;
        NfItem<S> it{};
        it.f = make_fields<S>(E);
        it.sc = make_scales<C>(E);
        it.d = make_dims(E);
        it.P = N.pass;
        items.push_back(it);
        st.push_back(st.back() + whole_blocks(N.pass.total));
      }
    }
    n_nf = static_cast<int>(items.size());
    nf_total = st.back();
    XF_TRY(put(&nf_items, items.data(), items.size() * sizeof(NfItem<S>)));
    void* q = nf_start;
    XF_TRY(put(&q, st.data(), st.size() * sizeof(int64_t)));
    nf_start = static_cast<int64_t*>(q);
  }
  // point sources and learned layers
  {
    std::vector<PointItem<S>> items;
    for (int g = 0; g < n; ++g) {
      Engine& E = eng(g);
      if (E.pts.n == 0) continue;
      PointItem<S> it{};
      it.f = make_fields<S>(E);
      it.sc = make_scales<C>(E);
      it.P = E.pts;
      items.push_back(it);
    }
    n_pt = static_cast<int>(items.size());
    XF_TRY(put(&pt_items, items.data(), items.size() * sizeof(PointItem<S>)));
  }
  {
    std::vector<LbItem<S>> items;
    std::vector<int64_t> st(1, 0), sm(1, 0), se(1, 0);
    lb_any_record = false;
    for (int g = 0; g < n; ++g) {
      Engine& E = eng(g);
      if (!E.has_lb || E.lbd.total == 0) continue;
      LbItem<S> it{};
      it.f = make_fields<S>(E);
      it.sc = make_scales<C>(E);
      it.d = make_dims(E);
      it.L = E.lbd;
      items.push_back(it);
      st.push_back(st.back() + (E.lbd.total + 255) / 256);
      const bool apply = E.lbd.mode == XF_LB_APPLY;
      sm.push_back(sm.back() + (apply ? (E.lbd.total + kLbMfmaCols - 1) / kLbMfmaCols : 0));
      se.push_back(se.back() + (apply ? E.lbd.n_edge : 0));
      lb_any_record = lb_any_record || !apply;
    }
    lb_edges = se.back();
    n_lb = static_cast<int>(items.size());
    lb_blocks = st.back();
    lb_mblocks = sm.back();
    if (lb_blocks > INT_MAX) return fail(XF_ERR_UNSUPPORTED, "learned layers too large for one launch");
    XF_TRY(put(&lb_items, items.data(), items.size() * sizeof(LbItem<S>)));
    void* q = lb_bstart;
    XF_TRY(put(&q, st.data(), st.size() * sizeof(int64_t)));
    lb_bstart = static_cast<int64_t*>(q);
    q = lb_mstart;
    XF_TRY(put(&q, sm.data(), sm.size() * sizeof(int64_t)));
    lb_mstart = static_cast<int64_t*>(q);
    q = lb_estart;
    XF_TRY(put(&q, se.data(), se.size() * sizeof(int64_t)));
    lb_estart = static_cast<int64_t*>(q);
  }
  // persistent grid: what the device holds resident, unless set
  hipDeviceProp_t prop;
  XF_HIP(hipGetDeviceProperties(&prop, device));
  resident = batch_occupancy_any<S>(id_bytes, pml_level, tune.nontemporal != 0, 64 * tune.tile_y) *
             prop.multiProcessorCount;
  grid_blocks = tune.blocks > 0 ? tune.blocks : resident;
  grid_blocks = std::max(1, std::min(grid_blocks, total_tiles));
  return XF_OK;
}

int Batch::configure() {
  if (tune.tile_y < 1 || tune.tile_y > 4) return fail(XF_ERR_INVALID, "tile_y must be 1..4");
  if (tune.planes < 1) return fail(XF_ERR_INVALID, "planes must be >= 1");
  if (tune.blocks < 0) return fail(XF_ERR_INVALID, "blocks must be >= 0");
  XF_HIP(hipSetDevice(device));
  drop_graph();
  int st;
  switch (plan.precision) {
    case XF_FP64: st = configure_t<double>(); break;
    case XF_FP32: st = configure_t<float>(); break;
    case XF_BF16_COMP: st = configure_t<Bf16c>(); break;
    case XF_FP16_STORE: st = configure_t<__half>(); break;
    default: st = configure_t<__hip_bfloat16>(); break;
  }
  if (st != XF_OK) return st;
  // grids whose every c_psi value the tiles read from the coefficient table
  int cpsi_tab = 0;
  for (size_t g = 0; g < inst.size(); ++g)
    cpsi_tab += eng(static_cast<int>(g)).cpsi_matched() == 4 * static_cast<int>(eng(static_cast<int>(g)).pml.size()) ? 1 : 0;
  desc = "grids=" + std::to_string(inst.size()) + " cells=" + std::to_string(cells) +
         " tiles=" + std::to_string(total_tiles) + " blocks=" + std::to_string(grid_blocks) +
         " resident=" + std::to_string(resident) + " tile_y=" + std::to_string(tune.tile_y) +
         " planes=" + std::to_string(tune.planes) + " nt=" + std::to_string(tune.nontemporal) +
         " id_bytes=" + std::to_string(id_bytes) + " pml_split=" + std::to_string(n_pz[1]) +
         " pml_ztail=" + std::to_string(tune.pml_split == 2 ? 1 : 0) +
         " k32=" + std::to_string(tune.k32) + " xcd=" + std::to_string(tune.xcd_remap) +
         " tfsf=" + std::to_string(n_tf[0]) +
         " nf2ff=" + std::to_string(n_nf) + " cpsi_tab=" + std::to_string(cpsi_tab);
  return XF_OK;
}

template <typename S>
int Batch::launch_step(int t, bool use_tdev, bool ev_on, hipEvent_t* ev) {
  const int* tp = use_tdev ? t_dev : nullptr;
  const int n = static_cast<int>(inst.size());
  const dim3 block(tune.k32 ? 32 : 64, tune.k32 ? 2 * tune.tile_y : tune.tile_y), grid(grid_blocks);
  const bool nt_on = tune.nontemporal != 0;
#define XF_BY(KER, I, PM, NT, PASS) \
  hipLaunchKernelGGL((KER<S, I, PM, NT>), grid, block, 0, stream, static_cast<const YeePass<S>*>(PASS), tile_start, n, \
                     tune.xcd_remap)
#define XF_BYNT(KER, I, PM, PASS)                                                              \
  do {                                                                                         \
    if (nt_on) XF_BY(KER, I, PM, true, PASS); else XF_BY(KER, I, PM, false, PASS);             \
  } while (0)
#define XF_BYLV(KER, I, PASS)                                                                  \
  do {                                                                                         \
    if (pml_level == 2) XF_BYNT(KER, I, 2, PASS);                                              \
    else if (pml_level == 1) XF_BYNT(KER, I, 1, PASS);                                         \
    else XF_BYNT(KER, I, 0, PASS);                                                             \
  } while (0)
#define XF_BYALL(KER, PASS)                                                                    \
  do {                                                                                         \
    if (id_bytes == 1) XF_BYLV(KER, uint8_t, PASS); else XF_BYLV(KER, uint16_t, PASS);         \
  } while (0)
  if (ev_on) XF_HIP(hipEventRecord(ev[0], stream));
  XF_BYALL(k_yee_h3_batch, pass_h);
  if (pz_total[0] > 0)
    hipLaunchKernelGGL((k_pml_z_batch<S, false>), dim3((unsigned)((pz_total[0] + 255) / 256)), dim3(256), 0,
                       stream, static_cast<const PmlZItem<S>*>(pz_items[0]), pz_start[0], n_pz[0]);
  if (ev_on) XF_HIP(hipEventRecord(ev[1], stream));
  if (tf_total[0] > 0)
    hipLaunchKernelGGL((k_tfsf_batch<S>), dim3((unsigned)((tf_total[0] + 255) / 256)), dim3(256), 0, stream,
                       static_cast<const TfsfItem<S>*>(tf_items[0]), tf_start[0], n_tf[0], t, tp);
  if (ev_on) XF_HIP(hipEventRecord(ev[2], stream));
  XF_BYALL(k_yee_e3_batch, pass_e);
  if (pz_total[1] > 0)
    hipLaunchKernelGGL((k_pml_z_batch<S, true>), dim3((unsigned)((pz_total[1] + 255) / 256)), dim3(256), 0,
                       stream, static_cast<const PmlZItem<S>*>(pz_items[1]), pz_start[1], n_pz[1]);
  if (ev_on) XF_HIP(hipEventRecord(ev[3], stream));
  if (tf_total[1] > 0)
    hipLaunchKernelGGL((k_tfsf_batch<S>), dim3((unsigned)((tf_total[1] + 255) / 256)), dim3(256), 0, stream,
                       static_cast<const TfsfItem<S>*>(tf_items[1]), tf_start[1], n_tf[1], t, tp);
  // per grid: point sources, then the learned layer's phases, then movies,
  // as Engine::launch_additions orders them; the grids share no memory
  if (n_pt > 0)
    hipLaunchKernelGGL((k_point_sources_batch<S>), dim3((unsigned)((n_pt + 63) / 64)), dim3(64), 0, stream,
                       static_cast<const PointItem<S>*>(pt_items), n_pt, t, tp);
  if (n_lb > 0) {
    // phase 0 (RECORD layers on the VALU, APPLY layers on the matrix
    // cores), the edge nodes of APPLY layers, the history roll
    const LbItem<S>* li = static_cast<const LbItem<S>*>(lb_items);
    // recording layers: phase 0 (record) and 2 (roll) over all columns;
    // apply layers: the network with the roll of its columns, then the edge
    // columns' Mur update and roll
    if (lb_any_record)
      hipLaunchKernelGGL((k_lb_batch<S>), dim3((unsigned)lb_blocks), dim3(256), 0, stream, li, lb_bstart, n_lb, 0,
                         t, tp);
    if (lb_mblocks > 0)
      hipLaunchKernelGGL((k_lb_mfma_batch<S>), dim3((unsigned)lb_mblocks), dim3(256), 0, stream, li, lb_mstart,
                         n_lb, t, tp);
    if (lb_edges > 0)
      for (int ph = 1; ph < 3; ++ph)
        hipLaunchKernelGGL((k_lb_edges_batch<S>), dim3((unsigned)((lb_edges + 255) / 256)), dim3(256), 0, stream,
                           li, lb_estart, n_lb, ph, t, tp);
    if (lb_any_record)
      hipLaunchKernelGGL((k_lb_batch<S>), dim3((unsigned)lb_blocks), dim3(256), 0, stream, li, lb_bstart, n_lb, 2,
                         t, tp);
  }
  for (auto& h : inst) XF_TRY(h->eng.launch_movies<S>(t, tp, stream));
  if (nf_total > 0)
    hipLaunchKernelGGL((k_nf2ff_batch<S>), dim3((unsigned)((nf_total + 255) / 256)), dim3(256), 0, stream,
                       static_cast<const NfItem<S>*>(nf_items), nf_start, n_nf, t, tp);
  for (auto& h : inst) XF_TRY(h->eng.launch_td<S>(t, tp, stream));
  if (ev_on) XF_HIP(hipEventRecord(ev[4], stream));
  if (use_tdev) hipLaunchKernelGGL(k_next_step, dim3(1), dim3(1), 0, stream, t_dev);
#undef XF_BYALL
#undef XF_BYLV
#undef XF_BYNT
#undef XF_BY
  XF_HIP(hipGetLastError());
  return XF_OK;
}

int Batch::run(int64_t n) {
  if (!uploaded) return fail(XF_ERR_STATE, "run before upload");
  if (n < 0) return fail(XF_ERR_INVALID, "negative step count");
  if (n == 0) return XF_OK;
  if (nt != INT64_MAX && cur_step + n > nt)
    return fail(XF_ERR_INVALID, "run past the end of the time axis the TFSF/NF2FF tables were built for");
  XF_HIP(hipSetDevice(device));
  const bool timing_on = plan.timing != 0;
  const bool graph_on = plan.use_graph != 0 && !timing_on;
  auto step = [&](int t, bool tdev, bool ev_on, hipEvent_t* e) -> int {
    switch (plan.precision) {
      case XF_FP64: return launch_step<double>(t, tdev, ev_on, e);
      case XF_FP32: return launch_step<float>(t, tdev, ev_on, e);
      case XF_BF16_COMP: return launch_step<Bf16c>(t, tdev, ev_on, e);
      case XF_FP16_STORE: return launch_step<__half>(t, tdev, ev_on, e);
      default: return launch_step<__hip_bfloat16>(t, tdev, ev_on, e);
    }
  };
  HipEvents ev, runs;
  if (timing_on) XF_TRY(ev.create(5 * n));
  XF_TRY(runs.create(2));
  hipEvent_t run0 = runs[0], run1 = runs[1];
  int status = XF_OK;
  // the watchdog checks each grid's fields (Engine::launch_watch) on the
  // batch's stream
  const int64_t wd = plan.watchdog > 0 ? plan.watchdog : 0;
  auto watch = [&](int64_t done) -> int {
    for (auto& h : inst) {
      Engine& E = h->eng;
      int st;
      switch (plan.precision) {
        case XF_FP64: st = E.launch_watch<double>(cur_step + done, stream); break;
        case XF_FP32: st = E.launch_watch<float>(cur_step + done, stream); break;
        case XF_BF16_COMP: st = E.launch_watch<Bf16c>(cur_step + done, stream); break;
        case XF_FP16_STORE: st = E.launch_watch<__half>(cur_step + done, stream); break;
        default: st = E.launch_watch<__hip_bfloat16>(cur_step + done, stream);
      }
      if (st != XF_OK) return st;
    }
    return XF_OK;
  };
  if (wd)
    for (auto& h : inst) hipLaunchKernelGGL(k_watch_clear, dim3(1), dim3(64), 0, stream, h->eng.wd_dev);
  if (graph_on) {
    if (!graph_ready) {
      XF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      const int st = step(0, true, false, nullptr);
      XF_TRY(end_capture(stream, st, &graph));
      XF_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
      graph_ready = true;
    }
    const int t0 = static_cast<int>(cur_step);
    XF_HIP(hipMemcpyAsync(t_dev, &t0, sizeof(int), hipMemcpyHostToDevice, stream));
    XF_HIP(hipStreamSynchronize(stream));
    XF_HIP(hipEventRecord(run0, stream));
    for (int64_t s = 0; s < n; ++s) {
      XF_HIP(hipGraphLaunch(graph_exec, stream));
      if (wd && ((cur_step + s + 1) % wd == 0 || s + 1 == n)) XF_TRY(watch(s + 1));
    }
  } else {
    XF_HIP(hipEventRecord(run0, stream));
    for (int64_t s = 0; s < n && status == XF_OK; ++s) {
      status = step(static_cast<int>(cur_step + s), false, timing_on, timing_on ? &ev[5 * s] : nullptr);
      if (status == XF_OK && wd && ((cur_step + s + 1) % wd == 0 || s + 1 == n)) status = watch(s + 1);
    }
  }
  XF_HIP(hipEventRecord(run1, stream));
  XF_HIP(hipEventSynchronize(run1));
  if (status != XF_OK) return status;
  XF_HIP(hipGetLastError());
  int bad_grid = -1;
  unsigned first_bad = 0xffffffffu;
  if (wd)
    for (size_t g = 0; g < inst.size(); ++g) {
      unsigned f = 0xffffffffu;
      XF_HIP(hipMemcpy(&f, inst[g]->eng.wd_dev, sizeof(unsigned), hipMemcpyDeviceToHost));
      if (f < first_bad) {
        first_bad = f;
        bad_grid = static_cast<int>(g);
      }
    }
  float ms = 0;
  XF_HIP(hipEventElapsedTime(&ms, run0, run1));
  timing.run_ms = ms;
  timing.steps += n;
  timing.step_ms = ms / static_cast<double>(n);
  if (timing_on) {
    for (int64_t s = 0; s < n; ++s) {
      float a, b, c, e;
      XF_HIP(hipEventElapsedTime(&a, ev[5 * s], ev[5 * s + 1]));
      XF_HIP(hipEventElapsedTime(&b, ev[5 * s + 1], ev[5 * s + 2]));
      XF_HIP(hipEventElapsedTime(&c, ev[5 * s + 2], ev[5 * s + 3]));
      XF_HIP(hipEventElapsedTime(&e, ev[5 * s + 3], ev[5 * s + 4]));
      timing.yee_h_ms += a;
      timing.tfsf_ms += b;
      timing.yee_e_ms += c;
      timing.nf2ff_ms += e;
    }
    timing.yee_h_launches += n;
    timing.yee_e_launches += n;
  }
  // algorithmic bytes per cell-step, cell-weighted over the grids
  double bytes = 0;
  for (auto& h : inst) bytes += h->eng.bytes_per_cell_step() * static_cast<double>(h->eng.nx * h->eng.ny * h->eng.nz);
  timing.bytes_per_cell_step = bytes / static_cast<double>(cells);
  cur_step += n;
  sync_steps();
  if (bad_grid >= 0)
    return fail(XF_ERR_NUMERIC, "watchdog: non-finite field value in grid " + std::to_string(bad_grid) +
                                    " found after step " + std::to_string(first_bad));
  return XF_OK;
}

int Batch::download() {
  if (!uploaded) return fail(XF_ERR_STATE, "download before upload");
  XF_HIP(hipStreamSynchronize(stream));
  for (auto& h : inst) XF_TRY(h->eng.download());
  return XF_OK;
}

int Batch::snapshot() {
  for (auto& h : inst) XF_TRY(h->eng.snapshot());
  return XF_OK;
}

int Batch::restore() {
  XF_HIP(hipStreamSynchronize(stream));
  for (auto& h : inst) XF_TRY(h->eng.restore());
  return XF_OK;
}

// Candidate tilings: rows per tile x planes per march x persistent grid
// (resident, half of it) x non-temporal streams.  Each runs one warm-up
// step and trial_steps timed steps from the same state.
int Batch::autotune(int trial_steps, xf_tuning* best) {
  if (!uploaded) return fail(XF_ERR_STATE, "autotune before upload");
  if (trial_steps < 1) return fail(XF_ERR_INVALID, "trial_steps must be >= 1");
  if (nt != INT64_MAX && cur_step + trial_steps + 1 > nt)
    return fail(XF_ERR_INVALID, "not enough time-table steps left to tune");
  XF_TRY(snapshot());
  const int64_t step0 = cur_step;
  const xf_timing keep = timing;
  const int keep_timing = plan.timing;
  plan.timing = 0;
  xf_tuning win{};
  win.step_ms = 1e30;
  int status = XF_OK;
  // the z-face split and tail only where some grid can take them
  bool can_split = false, can_tail = false;
  for (auto& h : inst) {
    const int keep_split = h->eng.pml_split, keep_tail = h->eng.pml_ztail;
    h->eng.pml_split = 1;
    h->eng.pml_ztail = 0;
    can_split = can_split || h->eng.split_z();
    h->eng.pml_ztail = 1;
    can_tail = can_tail || h->eng.ztail();
    h->eng.pml_split = keep_split;
    h->eng.pml_ztail = keep_tail;
  }
  for (int ty : {1, 2, 4}) {
    // 6 planes: 30.29 against 30.08-30.11 Gcells/s for 4 on the FP32 64-grid
    // batch at the tuner's other choices (5: 30.36, 3: 29.66)
    for (int planes : {2, 4, 6, 8, 16}) {
      for (int nt_on : {1, 0}) {
        for (int cand = 0; cand < 32; ++cand) {
          // split: 0 z faces in the march, 1 split launch, 2 tile tail
          const int half = cand & 1, split = (cand >> 1) & 3, k32 = (cand >> 3) & 1, xcd = cand >> 4;
          if (split == 3 || (split == 1 && !can_split) || (split == 2 && !can_tail)) continue;
          if (xcd && half) continue;
          tune = xf_tuning{ty, planes, 0, nt_on, xcd, split, k32, 0.0};
          status = configure();
          if (status == XF_OK && half) {
            tune.blocks = std::max(1, resident / 2);
            status = configure();
          }
          if (status == XF_OK) status = run(1);
          if (status == XF_OK) status = run(trial_steps);
          if (status != XF_OK) break;
          const double ms = timing.run_ms / trial_steps;
          if (ms < win.step_ms) {
            win = tune;
            win.step_ms = ms;
          }
          cur_step = step0;
          sync_steps();
          status = restore();
          if (status != XF_OK) break;
        }
        if (status != XF_OK) break;
      }
      if (status != XF_OK) break;
    }
    if (status != XF_OK) break;
  }
  plan.timing = keep_timing;
  timing = keep;
  cur_step = step0;
  sync_steps();
  if (status != XF_OK) {
    (void)restore();
    return status;
  }
  tune = win;
  XF_TRY(configure());
  if (best) *best = win;
  return XF_OK;
}

// Single grid: rows per tile x planes per march x non-temporal streams.
int Engine::autotune(int trial_steps, xf_tuning* best) {
  if (!uploaded) return fail(XF_ERR_STATE, "autotune before upload");
  if (trial_steps < 1) return fail(XF_ERR_INVALID, "trial_steps must be >= 1");
  if ((sc->tfsf || !nf.empty() || !nftd.empty() || pts.n > 0) && cur_step + trial_steps + 1 > nt)
    return fail(XF_ERR_INVALID, "not enough time-table steps left to tune");
  XF_TRY(snapshot());
  const int64_t step0 = cur_step;
  const xf_timing keep = timing;
  const int keep_timing = plan.timing;
  plan.timing = 0;
  xf_tuning win{};
  win.step_ms = 1e30;
  int status = XF_OK;
  const std::vector<int> tys = dim == 3 ? std::vector<int>{1, 2, 4} : std::vector<int>{1};
  const std::vector<int> pls = dim == 3 ? std::vector<int>{2, 4, 8, 16} : std::vector<int>{1};
  // the z-face split and tail only where they change the launches
  const int keep_split = pml_split, keep_tail = pml_ztail;
  pml_split = 1;
  pml_ztail = 0;
  const bool can_split = split_z();
  pml_ztail = 1;
  const bool can_tail = ztail();
  pml_split = keep_split;
  pml_ztail = keep_tail;
  for (int ty : tys) {
    for (int planes : pls) {
      for (int nt_on : {1, 0}) {
        for (int cand = 0; cand < 6; ++cand) {
          // split: 0 z faces in the march, 1 split launch, 2 tile tail
          const int split = cand % 3, k32 = dim == 3 ? cand / 3 : 0;
          if ((split == 1 && !can_split) || (split == 2 && !can_tail) || (dim != 3 && cand >= 3)) continue;
          const int chunks = static_cast<int>((nx + 1 + planes - 1) / planes);
          set_tiling(ty, chunks, XF_TILE_SET | (nt_on ? XF_TILE_NT : 0) | (split == 1 ? XF_TILE_PML_SPLIT : 0) |
                                     (split == 2 ? XF_TILE_PML_ZTAIL : 0) | (k32 ? XF_TILE_K32 : 0));
          status = run(1);
          if (status == XF_OK) status = run(trial_steps);
          if (status != XF_OK) break;
          const double ms = timing.run_ms / trial_steps;
          if (ms < win.step_ms) win = xf_tuning{ty, planes, 0, nt_on, 0, split, k32, ms};
          cur_step = step0;
          status = restore();
          if (status != XF_OK) break;
        }
        if (status != XF_OK) break;
      }
      if (status != XF_OK) break;
    }
    if (status != XF_OK) break;
  }
  plan.timing = keep_timing;
  timing = keep;
  cur_step = step0;
  if (status != XF_OK) {
    (void)restore();
    return status;
  }
  set_tiling(win.tile_y, static_cast<int>((nx + 1 + win.planes - 1) / win.planes),
             XF_TILE_SET | (win.nontemporal ? XF_TILE_NT : 0) | (win.pml_split == 1 ? XF_TILE_PML_SPLIT : 0) |
                 (win.pml_split == 2 ? XF_TILE_PML_ZTAIL : 0) | (win.k32 ? XF_TILE_K32 : 0));
  if (best) *best = win;
  return XF_OK;
}

}  // namespace xf

struct xf_batch {
  xf::Batch b;
};

extern "C" {

int xf_autotune(xf_handle* h, int32_t trial_steps, xf_tuning* best) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null handle");
    return h->eng.autotune(trial_steps, best);
  });
}

int xf_batch_create(xf_scene* const* scenes, int32_t n, const xf_plan* plan, xf_batch** out) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!scenes || !out) return xf::fail(XF_ERR_INVALID, "null scenes or out");
    *out = nullptr;
    std::unique_ptr<xf_batch> b(new (std::nothrow) xf_batch);
    if (!b) return xf::fail(XF_ERR_NOMEM, "out of host memory");
    const int st = b->b.create(scenes, n, plan);
    if (st != XF_OK) return st;
    *out = b.release();
    return XF_OK;
  });
}

int xf_batch_upload(xf_batch* b) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!b) return xf::fail(XF_ERR_INVALID, "null batch");
    const int st = xf::guarded([&] { return b->b.upload(); });
    if (st != XF_OK) b->b.release();
    return st;
  });
}

int xf_batch_run(xf_batch* b, int64_t n_steps) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!b) return xf::fail(XF_ERR_INVALID, "null batch");
    return b->b.run(n_steps);
  });
}

int xf_batch_download(xf_batch* b) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!b) return xf::fail(XF_ERR_INVALID, "null batch");
    return b->b.download();
  });
}

int xf_batch_destroy(xf_batch* b) {
  return xf::guarded([&]() -> int {
    delete b;
    return XF_OK;
  });
}

int xf_batch_set_tuning(xf_batch* b, const xf_tuning* t) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!b || !t) return xf::fail(XF_ERR_INVALID, "null argument");
    const xf_tuning keep = b->b.tune;
    b->b.tune = *t;
    b->b.tune.step_ms = 0;
    if (t->tile_y < 1 || t->tile_y > 4 || t->planes < 1 || t->blocks < 0 || t->pml_split < 0 || t->pml_split > 2) {
      b->b.tune = keep;
      return xf::fail(XF_ERR_INVALID, "bad tuning (tile_y 1..4, planes >= 1, blocks >= 0, pml_split 0..2)");
    }
    if (!b->b.uploaded) return XF_OK;
    const int st = b->b.configure();
    if (st != XF_OK) {
      b->b.tune = keep;
      (void)b->b.configure();
    }
    return st;
  });
}

int xf_batch_get_tuning(const xf_batch* b, xf_tuning* t) {
  return xf::guarded([&]() -> int {
    if (!b || !t) return xf::fail(XF_ERR_INVALID, "null argument");
    *t = b->b.tune;
    return XF_OK;
  });
}

int xf_batch_autotune(xf_batch* b, int32_t trial_steps, xf_tuning* best) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!b) return xf::fail(XF_ERR_INVALID, "null batch");
    return b->b.autotune(trial_steps, best);
  });
}

int xf_batch_get_timing(const xf_batch* b, xf_timing* out) {
  return xf::guarded([&]() -> int {
    if (!b || !out) return xf::fail(XF_ERR_INVALID, "null argument");
    *out = b->b.timing;
    return XF_OK;
  });
}

int xf_batch_reset_timing(xf_batch* b) {
  return xf::guarded([&]() -> int {
    if (!b) return xf::fail(XF_ERR_INVALID, "null batch");
    b->b.timing = xf_timing{};
    return XF_OK;
  });
}

int xf_batch_set_timing(xf_batch* b, int32_t on) {
  return xf::guarded([&]() -> int {
    if (!b) return xf::fail(XF_ERR_INVALID, "null batch");
    b->b.plan.timing = on ? 1 : 0;
    return XF_OK;
  });
}

int xf_batch_describe(const xf_batch* b, char* buf, size_t len) {
  return xf::guarded([&]() -> int {
    if (!b || !buf || len == 0) return xf::fail(XF_ERR_INVALID, "null argument");
    std::snprintf(buf, len, "%s", b->b.desc.c_str());
    return XF_OK;
  });
}

}  // extern "C"
