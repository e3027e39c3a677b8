// xf_engine.hip -- host side of the gfx950 FDTD engine: the C ABI of
// include/xfdtd_hip.h.  Validates the scene, lays the state out in HBM,
// compresses the Yee coefficients, and runs the DomainHD step loop
// (src/domain/domain_hd.cu:35-87) as 2-6 launches per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/xfdtd_hip.h"
#include "xf_kernels.inl"

namespace xf {
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define XF_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(XF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define XF_TRY(call)              \
  do {                            \
    int s_ = (call);              \
    if (s_ != XF_OK) return s_;   \
  } while (0)

// Every extern "C" entry point runs its body through guarded(): no C++
// exception crosses the C ABI.  Host allocations (std::vector staging of a
// field, the hash tables of the coefficient compression) and thread start-up
// can throw; they come back as status codes.
template <typename F>
int guarded(F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(XF_ERR_NOMEM, "host allocation failed");
  } catch (const std::length_error& e) {
    return fail(XF_ERR_NOMEM, std::string("host allocation too large: ") + e.what());
  } catch (const std::system_error& e) {
    return fail(XF_ERR_NOMEM, std::string("host resources exhausted: ") + e.what());
  } catch (const std::exception& e) {
    return fail(XF_ERR_INVALID, std::string("internal error: ") + e.what());
  } catch (...) {
    return fail(XF_ERR_INVALID, "internal error: unknown exception");
  }
}

// HIP events owned for the length of one run: destroyed on every return.
struct HipEvents {
  std::vector<hipEvent_t> ev;
  HipEvents() = default;
  HipEvents(const HipEvents&) = delete;
  HipEvents& operator=(const HipEvents&) = delete;
  ~HipEvents() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  int create(size_t n) {
    ev.assign(n, nullptr);
    for (auto& e : ev) XF_HIP(hipEventCreate(&e));
    return XF_OK;
  }
  hipEvent_t& operator[](size_t q) { return ev[q]; }
};

// Ends a stream capture started by the caller; on a failed capture (status
// st of the captured launches, or the end itself) the partial graph is
// destroyed and the failure returned, else *out holds the graph.
int end_capture(hipStream_t stream, int st, hipGraph_t* out) {
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(stream, &g);
  if (st != XF_OK || ce != hipSuccess) {
    if (ce == hipSuccess && g) (void)hipGraphDestroy(g);
    if (st != XF_OK) return st;
    return fail(XF_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
  }
  *out = g;
  return XF_OK;
}

bool empty_task(const xf_task& t) {
  return !(t.x.start < t.x.end && t.y.start < t.y.end && t.z.start < t.z.end);
}

int64_t task_lo(const xf_task& t, int ax) { return ax == 0 ? t.x.start : ax == 1 ? t.y.start : t.z.start; }
int64_t task_hi(const xf_task& t, int ax) { return ax == 0 ? t.x.end : ax == 1 ? t.y.end : t.z.end; }

int check_array(const xf_array& a, std::vector<int64_t> shape, const std::string& name) {
  if (a.data == nullptr) return fail(XF_ERR_INVALID, name + ": null data");
  if (a.ndim != static_cast<int32_t>(shape.size()))
    return fail(XF_ERR_INVALID, name + ": expected " + std::to_string(shape.size()) + " dims, got " +
                                    std::to_string(a.ndim));
  for (size_t d = 0; d < shape.size(); ++d)
    if (a.shape[d] != shape[d])
      return fail(XF_ERR_INVALID, name + ": extent " + std::to_string(d) + " is " +
                                      std::to_string(a.shape[d]) + ", expected " +
                                      std::to_string(shape[d]));
  return XF_OK;
}

int64_t numel(const xf_array& a) {
  int64_t n = 1;
  for (int d = 0; d < a.ndim; ++d) n *= a.shape[d];
  return n;
}

// Field shape of component comp (0..2) of E (attr 0) or H (attr 1).
std::vector<int64_t> field_shape(int attr, int comp, int64_t nx, int64_t ny, int64_t nz) {
  const int64_t n[3] = {nx, ny, nz};
  std::vector<int64_t> s(3);
  for (int ax = 0; ax < 3; ++ax) s[ax] = (attr == 0) ? (ax == comp ? n[ax] : n[ax] + 1) : (ax == comp ? n[ax] + 1 : n[ax]);
  return s;
}

struct TfsfJobHost {
  int comp, face;
  int64_t o[3], n[3];
  double coef;
  // z = const job whose nodes other jobs of its component meet only at the
  // two ends of this axis (0 or 1): the inner ones may go to the Yee tiles
  int tail_axis{-1};
};

}  // namespace

// ---------------------------------------------------------------- engine
struct Engine {
  xf_scene* sc{};
  xf_plan plan{};
  int dim{3};
  int64_t nx{}, ny{}, nz{};
  int64_t sx{}, sy{}, total{};  // layout: 3-D idx = i*sx + j*sy + k; 2-D idx = i*sx + j
  size_t esize{8};              // bytes per stored value
  bool uploaded{false};
  int64_t cur_step{0};
  int64_t nt{0};
  int device{0};
  hipStream_t stream{};
  std::vector<void*> allocs;
  // device buffers a time step rewrites (fields, psi, NF2FF sums, ADE
  // state) and their backup for snapshot()/restore()
  std::vector<std::pair<void*, size_t>> state;
  std::vector<void*> backup;
  // fields
  void* fld[6]{};
  // coefficient ids and tables (compute type)
  void* eid{};
  void* hid{};
  int id_bytes{1};
  bool wide_ids{false};  // force 16-bit ids (a batch needs one id width)
  int ntab_e{0}, ntab_h{0};
  void* etab{};
  void* htab{};
  unsigned char* eflag{};
  std::vector<double> etab_h, htab_h;  // host copies (double)
  // per scene CPML face, per psi array (ea, eb, ha, hb): the coefficient of
  // the node's table row that equals each of its c_psi values, or -1
  std::vector<std::array<int, 4>> cpsi_row;
  std::vector<unsigned char> eflag_h;
  // dense mode (XF_COEFF_DENSE, or too many rows for a table): the 18
  // per-cell coefficient arrays in the padded layout (E pass 0..8, H pass
  // 9..17), compute type; the E ids then only flag dispersive cells
  bool dense{false};
  void* dcoef[18]{};
  int upload_dense_coeff();
  // PML
  std::vector<int> pml_order;  // scene index sorted by axis
  struct PmlDev {
    void *ca_e, *cb_e, *ca_h, *cb_h;
    void* cpsi[4];
    void* psi[4];
  };
  std::vector<PmlDev> pml;
  // TFSF
  std::vector<TfsfJobHost> tfsf_jobs[2];  // [0] H pass, [1] E pass
  void* proj_int[3]{};
  void* proj_half[3]{};
  void* e_inc{};
  void* h_inc{};
  // NF2FF
  struct NfDev {
    double2* cur;
    double2* te;
    double2* th;
    NfPass pass;
  };
  std::vector<NfDev> nf;
  // NF2FF time domain
  std::vector<NtPass> nftd;
  // ADE
  AdeDev ade{};
  int64_t ade_rows{-1};  // distinct ADE coefficient rows in the table, -1 per-cell arrays
  // point sources and the learned layer (additions)
  PointDev pts{};
  std::vector<MovieDev> movies;
  LbDev lbd{};
  bool has_lb{false};
  int64_t lb_rec_count{0}, lb_rec_count_saved{0};
  bool has_ade{false};
  std::vector<int64_t> ade_cells;  // host (i,j,k) linear index in (nx+1)(ny+1)(nz+1)
  // 2-D FP64/FP32: the TF/SF corrections run inside the Yee kernels
  bool tfsf_fused() const {
    return dim == 2 && sc && sc->tfsf && (plan.precision == XF_FP64 || plan.precision == XF_FP32) &&
           tfsf_jobs[0].size() <= static_cast<size_t>(kTfsf2dJobs) &&
           tfsf_jobs[1].size() <= static_cast<size_t>(kTfsf2dJobs);
  }
  // all movie monitors in one launch: their descriptors and first threads
  void* movie_items{};
  int64_t* movie_start{};
  // time counter for graph replay
  int* t_dev{};
  // watchdog: first step found with a non-finite field value (UINT_MAX: none)
  unsigned* wd_dev{};
  template <typename S>
  int launch_watch(int64_t step, hipStream_t st);
  hipGraph_t graph{};
  hipGraphExec_t graph_exec{};
  bool graph_ready{false};
  // timing
  xf_timing timing{};
  int tile_y{4};
  int chunk{1};
  int nkb{1}, njb{1}, nch_eff{1};
  int xcd{0};
  int nt_streams{1};
  int pml_split{0};  // XF_TILE_PML_SPLIT requested (see split_z())
  int pml_ztail{0};  // XF_TILE_PML_ZTAIL requested (see ztail())
  int tfsf_zt{0};    // XF_TILE_TFSF_ZTAIL requested (see tfsf_ztail())
  int klanes{64};    // k-lanes of a tile row: 64, or 32 (XF_TILE_K32)
  // XF_TILE_PML_STREAMS: per pass a tile list, CPML-free tiles first, then
  // the tiles that meet a face the Yee kernel carries; the CPML part runs on
  // stream2 beside the rest
  int pml_streams{0};
  hipStream_t stream2{};
  hipEvent_t ev_fork{}, ev_join{};
  void* tlist[2]{};
  int64_t tlist_cap[2]{0, 0};
  int n_int[2]{0, 0}, n_pml[2]{0, 0};
  int build_tile_lists();
  bool use_lists() const { return pml_streams && dim == 3 && !dense && tlist[0] && tlist[1]; }
  int rows() const { return tile_y * (64 / klanes); }  // j-rows of a block
  void set_tiling(int ty, int chunks_x, int flags);
  Tiling tiling() const;
  int snapshot();  // copy every state buffer aside (device to device)
  int restore();   // put the copies back; the step counter is the caller's
  int autotune(int trial_steps, xf_tuning* best);
  template <typename S>
  YeePass<S> yee_pass(bool is_e) const;
  // the dense coefficient arrays of a pass (all null in table mode)
  template <typename C>
  Dense9<C> dense9(bool is_e) const {
    Dense9<C> d{};
    if (dense)
      for (int n = 0; n < 9; ++n) d.a[n] = static_cast<const C*>(dcoef[(is_e ? 0 : 9) + n]);
    return d;
  }
  std::string layout_desc;

  ~Engine() { release(); }

  void release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_ready = false;
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    state.clear();
    backup.clear();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    if (stream2) (void)hipStreamDestroy(stream2);
    stream2 = nullptr;
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    ev_fork = ev_join = nullptr;
    tlist[0] = tlist[1] = nullptr;
    tlist_cap[0] = tlist_cap[1] = 0;
    uploaded = false;
  }

  int alloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(XF_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    allocs.push_back(*p);
    return XF_OK;
  }

  template <typename T>
  int upload_vec(void** p, const std::vector<T>& v) {
    XF_TRY(alloc(p, v.size() * sizeof(T)));
    if (!v.empty()) XF_HIP(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return XF_OK;
  }

  // compute-type copy of a host double array
  int upload_c(void** p, const double* src, int64_t n) {
    if (plan.precision == XF_FP64) {
      std::vector<double> v(src, src + n);
      return upload_vec(p, v);
    }
    std::vector<float> v(n);
    for (int64_t i = 0; i < n; ++i) v[i] = static_cast<float>(src[i]);
    return upload_vec(p, v);
  }

  int download_c(double* dst, const void* p, int64_t n) {
    if (plan.precision == XF_FP64) {
      XF_HIP(hipMemcpy(dst, p, n * sizeof(double), hipMemcpyDeviceToHost));
      return XF_OK;
    }
    std::vector<float> v(n);
    XF_HIP(hipMemcpy(v.data(), p, n * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) dst[i] = v[i];
    return XF_OK;
  }

  int validate();
  int build_tfsf_jobs();
  int check_overlaps();
  int check_ade_pml() const;
  bool dispersive(int64_t i, int64_t j, int64_t k) const;
  int upload();
  int pack_fields();
  int unpack_fields();
  int build_ids();
  int upload_pml();
  int upload_tfsf();
  int upload_nf2ff();
  int plan_nf2ff_td(const xf_nf2ff_td& T, NtPass& P, int64_t* prev_total) const;
  int upload_nf2ff_td();
  int upload_ade();
  int validate_additions();
  int upload_additions();
  int64_t lb_recorded_after(int64_t step) const;
  int download();
  int run(int64_t n);
  template <typename S>
  int launch_step(int t, bool use_tdev, bool timing_events, hipEvent_t* ev);
  template <typename S>
  int launch_h(int t, const int* tp);
  template <typename S>
  int launch_e(int t, const int* tp);
  template <typename S>
  int launch_post_h(int t, const int* tp);
  template <typename S>
  int launch_post_e(int t, const int* tp);
  // after the E corrections: point sources, the learned layer, movies (in
  // that order), and after the frequency-domain NF2FF the time-domain one;
  // on stream st (the engine's own, or a batch's)
  template <typename S>
  int launch_additions(int t, const int* tp, hipStream_t st, bool with_movies = true);
  template <typename S>
  int launch_movies(int t, const int* tp, hipStream_t st);
  // z-face CPML corrections in a launch of their own after each Yee pass
  bool split_z() const;
  bool ztail() const;
  void match_cpsi(const std::vector<uint32_t>& ide, const std::vector<uint32_t>& idh);
  int cpsi_matched() const {  // psi arrays whose c_psi the tiles take from the table
    int m = 0;
    for (const auto& r : cpsi_row)
      for (int v : r) m += (v >= 0 && !dense) ? 1 : 0;
    return m;
  }
  // the z = const TF/SF corrections applied by the Yee tiles: with the
  // z-face CPML tail (whose instantiation carries the code), one launch per
  // pass, and no ADE (whose kernel rewrites E after the Yee pass)
  bool tfsf_ztail() const { return tfsf_zt && dim == 3 && sc && sc->tfsf && !has_ade && !use_lists() && ztail(); }
  // device index of history entry i (xf_learned_boundary layout); y faces
  // are stored [b][a] on the device (LbFace::tr)
  int64_t lb_dev_hist_index(int64_t i) const;
  template <typename S>
  int launch_pml_z(bool is_e, hipStream_t st);
  template <typename S>
  int launch_td(int t, const int* tp, hipStream_t st);

  double bytes_per_cell_step() const;
};

// ------------------------------------------------------------ validation
int Engine::validate() {
  if (sc->abi_version != XF_ABI_VERSION) return fail(XF_ERR_INVALID, "scene ABI version mismatch");
  dim = sc->dimension;
  nx = sc->nx;
  ny = sc->ny;
  nz = sc->nz;
  if (dim != 2 && dim != 3) return fail(XF_ERR_INVALID, "dimension must be 2 or 3");
  if (nx < 1 || ny < 1 || nz < 1) return fail(XF_ERR_INVALID, "empty grid");
  if (dim == 2 && nz != 1) return fail(XF_ERR_INVALID, "a 2-D scene has nz = 1");
  if (nx > (1 << 20) || ny > (1 << 20) || nz > (1 << 20)) return fail(XF_ERR_INVALID, "grid too large");
  {
    // one field component on the padded node grid at 8 bytes per value
    // (rows padded to 128 bytes); (2^20 + 1)^3 * 8 does not fit in int64
    const unsigned __int128 per_field =
        static_cast<unsigned __int128>(nx + 1) * static_cast<unsigned __int128>(ny + 1) *
        static_cast<unsigned __int128>(nz + 1 + 16) * 8u;
    if (per_field > (static_cast<unsigned __int128>(1) << 50))
      return fail(XF_ERR_NOMEM, "packed grid too large: " + std::to_string(static_cast<double>(per_field)) +
                                    " bytes per field component (limit 2^50)");
  }
  if (!(sc->dt > 0)) return fail(XF_ERR_INVALID, "dt must be positive");
  nt = sc->start_step + sc->n_steps;
  cur_step = sc->current_step;
  const xf_array* E[3] = {&sc->ex, &sc->ey, &sc->ez};
  const xf_array* H[3] = {&sc->hx, &sc->hy, &sc->hz};
  const char* en[3] = {"ex", "ey", "ez"};
  const char* hn[3] = {"hx", "hy", "hz"};
  for (int c = 0; c < 3; ++c) {
    XF_TRY(check_array(*E[c], field_shape(0, c, nx, ny, nz), en[c]));
    XF_TRY(check_array(*H[c], field_shape(1, c, nx, ny, nz), hn[c]));
  }
  if (sc->coeff[0].data) {
    for (int n = 0; n < 18; ++n)
      XF_TRY(check_array(sc->coeff[n], field_shape(n / 9, (n % 9) / 3, nx, ny, nz),
                         "coeff." + std::to_string(n)));
  } else {
    if (!sc->cell_material || !sc->material_coeff || sc->n_material < 1)
      return fail(XF_ERR_INVALID, "no coefficients: dense arrays or a material table are required");
  }
  if (sc->n_pml < 0 || sc->n_pml > kMaxPml) return fail(XF_ERR_UNSUPPORTED, "at most 6 PML faces are supported");
  const int64_t n3[3] = {nx, ny, nz};
  for (int p = 0; p < sc->n_pml; ++p) {
    const xf_pml& P = sc->pml[p];
    const std::string nm = "pml." + std::to_string(p);
    if (P.main_axis < 0 || P.main_axis > 2) return fail(XF_ERR_INVALID, nm + ": bad axis");
    if (dim == 2 && P.main_axis == 2) return fail(XF_ERR_UNSUPPORTED, nm + ": z PML on a 2-D grid");
    if (empty_task(P.task)) continue;
    const int ax = P.main_axis, a = (ax + 1) % 3, b = (ax + 2) % 3;
    for (int d = 0; d < 3; ++d)
      if (task_lo(P.task, d) < 0 || task_hi(P.task, d) > n3[d])
        return fail(XF_ERR_INVALID, nm + ": task outside the grid");
    const int64_t cs = task_lo(P.task, ax), ce = task_hi(P.task, ax);
    const int64_t off = cs == 0 ? 1 : 0;
    const int64_t ce0 = cs + off, ce1 = ce + off;  // E nodes
    // E nodes along c must have an H neighbour at c-1 and lie inside the E
    // tangential extent (n+1); H at c+1 needs E extent n+1.
    if (ce0 < 1 || ce1 > n3[ax] + 1) return fail(XF_ERR_INVALID, nm + ": E range outside the grid");
    const int64_t L = ce - cs;
    const xf_array* c1[4] = {&P.coeff_a_e, &P.coeff_b_e, &P.coeff_a_h, &P.coeff_b_h};
    for (int q = 0; q < 4; ++q) {
      if (!c1[q]->data || c1[q]->ndim != 1) return fail(XF_ERR_INVALID, nm + ": 1-D coefficient missing");
    }
    // 1-D index c + off_c - gstart over the corrected ranges
    auto in1 = [&](const xf_array& arr, int64_t c0, int64_t c1_, int64_t gs) {
      return c0 + P.offset_c - gs >= 0 && c1_ - 1 + P.offset_c - gs < arr.shape[0];
    };
    if (!in1(P.coeff_a_e, ce0, ce1, P.global_e_start) || !in1(P.coeff_b_e, ce0, ce1, P.global_e_start) ||
        !in1(P.coeff_a_h, cs, ce, P.global_h_start) || !in1(P.coeff_b_h, cs, ce, P.global_h_start))
      return fail(XF_ERR_INVALID, nm + ": 1-D coefficients do not cover the slab");
    // psi arrays: index aBCToXYZ(a, b, layer) over the corrected ranges
    const int64_t a0 = task_lo(P.task, a), a1 = task_hi(P.task, a);
    const int64_t b0 = task_lo(P.task, b), b1 = task_hi(P.task, b);
    struct R {
      const xf_array* psi;
      const xf_array* cpsi;
      int64_t amax, bmax, l0, l1;
    } rs[4] = {
        {&P.ea_psi_hb, &P.c_ea_psi_hb, a1 - 1, b1, ce0 - P.node_e_start, ce1 - 1 - P.node_e_start},
        {&P.eb_psi_ha, &P.c_eb_psi_ha, a1, b1 - 1, ce0 - P.node_e_start, ce1 - 1 - P.node_e_start},
        {&P.ha_psi_eb, &P.c_ha_psi_eb, a1, b1 - 1, cs - P.node_h_start, ce - 1 - P.node_h_start},
        {&P.hb_psi_ea, &P.c_hb_psi_ea, a1 - 1, b1, cs - P.node_h_start, ce - 1 - P.node_h_start}};
    for (int q = 0; q < 4; ++q) {
      if (!rs[q].psi->data || !rs[q].cpsi->data || rs[q].psi->ndim != 3 || rs[q].cpsi->ndim != 3)
        return fail(XF_ERR_INVALID, nm + ": psi arrays missing");
      for (int d = 0; d < 3; ++d)
        if (rs[q].psi->shape[d] != rs[q].cpsi->shape[d]) return fail(XF_ERR_INVALID, nm + ": psi/c_psi shapes differ");
      if (rs[q].l0 < 0) return fail(XF_ERR_INVALID, nm + ": negative psi layer");
      int64_t mx[3];
      mx[a] = rs[q].amax;
      mx[b] = rs[q].bmax;
      mx[ax] = rs[q].l1;
      int64_t mn[3];
      mn[a] = a0;
      mn[b] = b0;
      mn[ax] = rs[q].l0;
      for (int d = 0; d < 3; ++d)
        if (mn[d] < 0 || mx[d] >= rs[q].psi->shape[d]) return fail(XF_ERR_INVALID, nm + ": psi array too small");
      // the kernels index psi with 32-bit arithmetic
      if (numel(*rs[q].psi) >= (int64_t(1) << 32)) return fail(XF_ERR_UNSUPPORTED, nm + ": psi array above 2^32 values");
    }
    (void)L;
  }
  if (sc->tfsf) XF_TRY(build_tfsf_jobs());
  for (int q = 0; q < sc->n_nf2ff; ++q) {
    const xf_nf2ff& N = sc->nf2ff[q];
    if (N.n_freq < 1 || !N.freq) return fail(XF_ERR_INVALID, "nf2ff: no frequency");
    for (int face = 0; face < 6; ++face) {
      const xf_task& t = N.face_task[face];
      if (empty_task(t)) continue;
      const int ax = face / 2;
      if (dim == 2 && ax == 2) return fail(XF_ERR_UNSUPPORTED, "nf2ff: z face on a 2-D grid");
      if (task_hi(t, ax) - task_lo(t, ax) != 1) return fail(XF_ERR_INVALID, "nf2ff: face task is not one cell thick");
      // interpolation reads index-1 along the face normal's tangential axes
      // and index+1; keep them inside [0, n]
      for (int d = 0; d < 3; ++d) {
        if (dim == 2 && d == 2) {
          if (t.z.start != 0 || t.z.end != 1) return fail(XF_ERR_INVALID, "nf2ff: 2-D face z range must be [0,1)");
          continue;
        }
        if (task_lo(t, d) < 1 || task_hi(t, d) > n3[d] - 1)
          return fail(XF_ERR_INVALID, "nf2ff: face task must stay one cell inside the grid");
      }
      for (int f = 0; f < N.n_freq; ++f)
        for (int c = 0; c < 4; ++c)
          XF_TRY(check_array(N.freq[f].cur[face][c],
                             {t.x.end - t.x.start, t.y.end - t.y.start, t.z.end - t.z.start, 2},
                             "nf2ff current"));
    }
    for (int f = 0; f < N.n_freq; ++f) {
      XF_TRY(check_array(N.freq[f].transform_e, {nt, 2}, "nf2ff transform_e"));
      XF_TRY(check_array(N.freq[f].transform_h, {nt, 2}, "nf2ff transform_h"));
    }
  }
  if (sc->ade && sc->ade->kind != XF_ADE_NONE) {
    const xf_ade& A = *sc->ade;
    if (dim != 3) return fail(XF_ERR_UNSUPPORTED, "ADE media need a 3-D grid");
    if (A.kind < 1 || A.kind > 3) return fail(XF_ERR_INVALID, "ade: unknown kind");
    if (A.num_pole < 1 || A.num_pole > 64) return fail(XF_ERR_INVALID, "ade: bad pole count");
    const std::vector<int64_t> s4 = {nx + 1, ny + 1, nz + 1, A.num_pole}, s3 = {nx + 1, ny + 1, nz + 1};
    XF_TRY(check_array(A.coeff_j_j, s4, "ade.coeff_j_j"));
    XF_TRY(check_array(A.coeff_j_e, s4, "ade.coeff_j_e"));
    XF_TRY(check_array(A.coeff_j_sum_j, s4, "ade.coeff_j_sum_j"));
    XF_TRY(check_array(A.coeff_e_j_sum, s3, "ade.coeff_e_j_sum"));
    XF_TRY(check_array(A.jx, s4, "ade.jx"));
    XF_TRY(check_array(A.jy, s4, "ade.jy"));
    XF_TRY(check_array(A.jz, s4, "ade.jz"));
    if (A.kind == XF_ADE_MLORENTZ) {
      XF_TRY(check_array(A.coeff_j_j_p, s4, "ade.coeff_j_j_p"));
      XF_TRY(check_array(A.coeff_j_e_n, s4, "ade.coeff_j_e_n"));
      XF_TRY(check_array(A.coeff_j_e_p, s4, "ade.coeff_j_e_p"));
      XF_TRY(check_array(A.coeff_j_sum_j_p, s4, "ade.coeff_j_sum_j_p"));
      XF_TRY(check_array(A.coeff_e_e_p, s3, "ade.coeff_e_e_p"));
      XF_TRY(check_array(A.ex_prev, s3, "ade.ex_prev"));
      XF_TRY(check_array(A.ey_prev, s3, "ade.ey_prev"));
      XF_TRY(check_array(A.ez_prev, s3, "ade.ez_prev"));
      XF_TRY(check_array(A.jx_prev, s4, "ade.jx_prev"));
      XF_TRY(check_array(A.jy_prev, s4, "ade.jy_prev"));
      XF_TRY(check_array(A.jz_prev, s4, "ade.jz_prev"));
    }
    has_ade = true;
  }
  if (sc->n_nf2ff_td < 0 || (sc->n_nf2ff_td > 0 && !sc->nf2ff_td))
    return fail(XF_ERR_INVALID, "nf2ff_td boxes missing");
  for (int32_t q = 0; q < sc->n_nf2ff_td; ++q) {
    NtPass P;
    int64_t prev_total = 0;
    XF_TRY(plan_nf2ff_td(sc->nf2ff_td[q], P, &prev_total));
  }
  return validate_additions();
}

// Geometry of an NF2FF time-domain box: faces, tiles, the bins a tile and a
// face can touch in one step, and a check that every bin of the remaining
// steps [cur_step, nt) lies inside the potential arrays.  The r . r_unit
// ranges use the kernel's own expression, so the kernels' bins are bounded
// by the ones computed here.
int Engine::plan_nf2ff_td(const xf_nf2ff_td& T, NtPass& P, int64_t* prev_total) const {
  P = NtPass{};
  if (dim != 3) return fail(XF_ERR_UNSUPPORTED, "nf2ff_td: needs a 3-D grid");
  const int64_t n3[3] = {nx, ny, nz};
  static const char* ax_name[3] = {"x", "y", "z"};
  for (int ax = 0; ax < 3; ++ax) {
    XF_TRY(check_array(T.e_node[ax], {n3[ax] + 1}, std::string("nf2ff_td e_node ") + ax_name[ax]));
    XF_TRY(check_array(T.h_node[ax], {n3[ax]}, std::string("nf2ff_td h_node ") + ax_name[ax]));
    XF_TRY(check_array(T.e_size[ax], {n3[ax]}, std::string("nf2ff_td e_size ") + ax_name[ax]));
  }
  const double c_dt = 299792458.0 * sc->dt;
  if (!(c_dt > 0)) return fail(XF_ERR_INVALID, "nf2ff_td: dt must be positive");
  for (int d = 0; d < 3; ++d) P.u[d] = T.r_unit[d];
  P.dmax_e = T.dmax_e;
  P.dmax_h = T.dmax_h;
  P.c_dt = c_dt;
  P.n_pot = -1;
  int blk = 0, rblk = 0, maxspan = 0;
  int64_t prev_off = 0;
  for (int face = 0; face < 6; ++face) {
    NtFace& F = P.face[face];
    F.blk0 = blk;
    F.rblk0 = rblk;
    F.prev_off = prev_off;
    const xf_task& t = T.face_task[face];
    if (empty_task(t)) continue;
    const int c = face / 2, a = (c + 1) % 3, b = (c + 2) % 3;
    const int uax = std::min(a, b), vax = std::max(a, b);
    if (task_hi(t, c) - task_lo(t, c) != 1) return fail(XF_ERR_INVALID, "nf2ff_td: face task is not one cell thick");
    for (int d = 0; d < 3; ++d)
      if (task_lo(t, d) < 1 || task_hi(t, d) > n3[d] - 1)
        return fail(XF_ERR_INVALID, "nf2ff_td: face task must stay one cell inside the grid");
    const int64_t ext[3] = {t.x.end - t.x.start, t.y.end - t.y.start, t.z.end - t.z.start};
    const std::string nm = "nf2ff_td face " + std::to_string(face);
    for (int q = 0; q < 4; ++q) {
      if (P.n_pot < 0) {
        if (!T.pot[face][q].data || T.pot[face][q].ndim != 1 || T.pot[face][q].shape[0] < 2)
          return fail(XF_ERR_INVALID, nm + ": potential arrays must be 1-D with at least 2 bins");
        P.n_pot = T.pot[face][q].shape[0];
      }
      XF_TRY(check_array(T.pot[face][q], {P.n_pot}, nm + " potential"));
      XF_TRY(check_array(T.prev[face][q], {ext[a], ext[b]}, nm + " prev"));
    }
    if (P.n_pot > INT_MAX / 2) return fail(XF_ERR_INVALID, nm + ": too many time bins");
    F.o[0] = static_cast<int>(t.x.start);
    F.o[1] = static_cast<int>(t.y.start);
    F.o[2] = static_cast<int>(t.z.start);
    F.nu = static_cast<int>(ext[uax]);
    F.nv = static_cast<int>(ext[vax]);
    F.tu = (F.nu + NT_TILE - 1) / NT_TILE;
    F.tv = (F.nv + NT_TILE - 1) / NT_TILE;
    F.nb = static_cast<int>(ext[b]);
    std::vector<double> tmin(static_cast<size_t>(F.tu) * F.tv, INFINITY), tmax(tmin.size(), -INFINITY);
    double fmin = INFINITY, fmax = -INFINITY;
    for (int u = 0; u < F.nu; ++u)
      for (int v = 0; v < F.nv; ++v) {
        int64_t idx[3];
        idx[c] = F.o[c];
        idx[uax] = F.o[uax] + u;
        idx[vax] = F.o[vax] + v;
        double r[3];
        for (int ax = 0; ax < 3; ++ax) r[ax] = ax == c ? T.e_node[ax].data[idx[ax]] : T.h_node[ax].data[idx[ax]];
        const double rd = r[0] * P.u[0] + r[1] * P.u[1] + r[2] * P.u[2];
        if (!std::isfinite(rd)) return fail(XF_ERR_INVALID, nm + ": non-finite surface position");
        const size_t tl = static_cast<size_t>(u / NT_TILE) * F.tv + v / NT_TILE;
        tmin[tl] = std::min(tmin[tl], rd);
        tmax[tl] = std::max(tmax[tl], rd);
        fmin = std::min(fmin, rd);
        fmax = std::max(fmax, rd);
      }
    for (size_t tl = 0; tl < tmin.size(); ++tl)
      maxspan = std::max(maxspan, static_cast<int>(std::ceil((tmax[tl] - tmin[tl]) / c_dt)) + 3);
    F.rd_min = fmin;
    F.rd_max = fmax;
    F.bins = static_cast<int>(std::ceil((fmax - fmin) / c_dt)) + 4;
    F.nrblk = (F.bins + 3) / 4;  // one wave per bin, four per block
    if (cur_step < nt) {
      const double offs[2] = {0.5, 0.0};
      const double dm[2] = {T.dmax_e, T.dmax_h};
      for (int e = 0; e < 2; ++e) {
        const double lo = std::floor((dm[e] - fmax) / c_dt + offs[e] + static_cast<double>(cur_step));
        const double hi = std::floor((dm[e] - fmin) / c_dt + offs[e] + static_cast<double>(nt - 1)) + 1;
        if (lo < 0 || hi >= static_cast<double>(P.n_pot))
          return fail(XF_ERR_INVALID, nm + ": time bins [" + std::to_string(lo) + ", " + std::to_string(hi) +
                                          "] of the run outside the " + std::to_string(P.n_pot) +
                                          " potential bins (dmax or the potential length)");
      }
    }
    blk += F.tu * F.tv;
    rblk += F.nrblk;
    prev_off += 4 * ext[a] * ext[b];
  }
  if (maxspan > NT_MAXSPAN)
    return fail(XF_ERR_UNSUPPORTED, "nf2ff_td: a 16x16 surface tile spans " + std::to_string(maxspan) +
                                        " time bins (limit " + std::to_string(NT_MAXSPAN) +
                                        "); cells are too long for c dt");
  P.nblk = blk;
  P.nrblk = rblk;
  P.maxspan = std::max(maxspan, 1);
  if (P.n_pot < 0) P.n_pot = 0;
  *prev_total = prev_off;
  return XF_OK;
}

// Point sources and the learned layer: every index the kernels will form.
int Engine::validate_additions() {
  const int64_t n3[3] = {nx, ny, nz};
  if (sc->n_point < 0 || (sc->n_point > 0 && !sc->point)) return fail(XF_ERR_INVALID, "point sources missing");
  for (int32_t q = 0; q < sc->n_point; ++q) {
    const xf_point_source& P = sc->point[q];
    if (P.comp < 0 || P.comp > 2) return fail(XF_ERR_INVALID, "point source: bad component");
    const auto sh = field_shape(0, P.comp, nx, ny, nz);
    if (P.i < 0 || P.j < 0 || P.k < 0 || P.i >= sh[0] || P.j >= sh[1] || P.k >= sh[2])
      return fail(XF_ERR_INVALID, "point source " + std::to_string(q) + " outside its component");
    if (!P.value.data || P.value.ndim != 1 || P.value.shape[0] < nt)
      return fail(XF_ERR_INVALID, "point source " + std::to_string(q) + ": value table shorter than the run");
  }
  if (sc->n_movie < 0 || (sc->n_movie > 0 && !sc->movie)) return fail(XF_ERR_INVALID, "movie monitors missing");
  for (int32_t q = 0; q < sc->n_movie; ++q) {
    const xf_movie& M = sc->movie[q];
    const std::string nm = "movie " + std::to_string(q);
    if (M.field < 0 || M.field > 5) return fail(XF_ERR_INVALID, nm + ": bad field");
    if (M.interval < 1) return fail(XF_ERR_INVALID, nm + ": interval must be >= 1");
    const auto sh = field_shape(M.field < 3 ? 0 : 1, M.field % 3, nx, ny, nz);
    const xf_range* r[3] = {&M.task.x, &M.task.y, &M.task.z};
    for (int d = 0; d < 3; ++d)
      if (r[d]->start < 0 || r[d]->end > sh[d] || r[d]->start >= r[d]->end)
        return fail(XF_ERR_INVALID, nm + ": box outside the component's array");
    if (!M.data.data || M.data.ndim != 4 || M.data.shape[0] < 1 ||
        M.data.shape[1] != M.task.x.end - M.task.x.start || M.data.shape[2] != M.task.y.end - M.task.y.start ||
        M.data.shape[3] != M.task.z.end - M.task.z.start)
      return fail(XF_ERR_INVALID, nm + ": data must be (frames, box extents)");
  }
  if (!sc->lb || sc->lb->mode == XF_LB_OFF) return XF_OK;
  const xf_learned_boundary& L = *sc->lb;
  if (dim != 3) return fail(XF_ERR_UNSUPPORTED, "learned layer: 3-D grids only");
  if (L.mode != XF_LB_APPLY && L.mode != XF_LB_RECORD) return fail(XF_ERR_INVALID, "learned layer: bad mode");
  if (L.faces <= 0 || L.faces > 63) return fail(XF_ERR_INVALID, "learned layer: bad face mask");
  if (L.mode == XF_LB_APPLY && (L.inset != 0 || !L.weights))
    return fail(XF_ERR_INVALID, "learned layer: APPLY needs inset 0 and weights");
  if (L.mode == XF_LB_RECORD && (L.inset < 1 || L.rec_every < 1 || L.rec_slots < 1 || !L.rec))
    return fail(XF_ERR_INVALID, "learned layer: bad RECORD schedule or buffer");
  int64_t hist = 0;
  for (int f = 0; f < 6; ++f) {
    const int c = f / 2, a = (c + 1) % 3, b = (c + 2) % 3;
    hist += 16 * (n3[a] + 1) * (n3[b] + 1);
    if (!((L.faces >> f) & 1)) continue;
    if (n3[c] < 2 * static_cast<int64_t>(L.inset) + 4) return fail(XF_ERR_INVALID, "learned layer: grid too thin");
    if (n3[a] - 1 - L.inset < L.inset || n3[b] - 1 - L.inset < L.inset)
      return fail(XF_ERR_INVALID, "learned layer: inset too large");
  }
  if (!L.history.data || L.history.ndim != 1 || L.history.shape[0] != hist)
    return fail(XF_ERR_INVALID, "learned layer: history must hold " + std::to_string(hist) + " values");
  has_lb = true;
  return XF_OK;
}

// TFSF jobs in the reference's launch order, with written boxes; checks every
// incident-table index the kernel will form (tfsf_corrector.cuh:20-170).
int Engine::build_tfsf_jobs() {
  const xf_tfsf& T = *sc->tfsf;
  if (T.dimension != dim) return fail(XF_ERR_INVALID, "tfsf: dimension does not match the grid");
  const xf_task& tt = T.total_task;
  if (empty_task(tt)) return fail(XF_ERR_INVALID, "tfsf: empty total task");
  const int64_t n3[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) {
    if (dim == 2 && d == 2) {
      if (tt.z.start != 0 || tt.z.end != 1) return fail(XF_ERR_INVALID, "tfsf: 2-D task z range must be [0,1)");
      continue;
    }
    if (task_lo(tt, d) < 1 || task_hi(tt, d) > n3[d] - 1) return fail(XF_ERR_INVALID, "tfsf: box must stay inside the grid");
  }
  const xf_array* pi[3] = {&T.proj_x_int, &T.proj_y_int, &T.proj_z_int};
  const xf_array* ph[3] = {&T.proj_x_half, &T.proj_y_half, &T.proj_z_half};
  for (int d = 0; d < 3; ++d)
    if (!pi[d]->data || !ph[d]->data || pi[d]->ndim != 1 || ph[d]->ndim != 1)
      return fail(XF_ERR_INVALID, "tfsf: projection arrays missing");
  if (!T.e_inc.data || !T.h_inc.data || T.e_inc.ndim != 2 || T.h_inc.ndim != 2)
    return fail(XF_ERR_INVALID, "tfsf: incident tables missing");
  if (T.e_inc.shape[0] < nt || T.h_inc.shape[0] < nt) return fail(XF_ERR_INVALID, "tfsf: incident tables shorter than the run");
  // Job order (tfsf_corrector.cu:155-271): (face, component)
  struct FC {
    int face, comp;
  };
  std::vector<FC> order_e, order_h;
  if (dim == 3) {
    const FC o[12] = {{XF_XN, 2}, {XF_XP, 2}, {XF_YN, 2}, {XF_YP, 2}, {XF_ZN, 0}, {XF_ZP, 0},
                      {XF_YN, 0}, {XF_YP, 0}, {XF_ZN, 1}, {XF_ZP, 1}, {XF_XN, 1}, {XF_XP, 1}};
    order_e.assign(o, o + 12);
    order_h.assign(o, o + 12);
  } else {
    order_e = {{XF_XN, 2}, {XF_XP, 2}, {XF_YN, 2}, {XF_YP, 2}};
    order_h = {{XF_XN, 1}, {XF_XP, 1}, {XF_YN, 0}, {XF_YP, 0}};
  }
  const double ca[3] = {T.cax, T.cay, T.caz}, cb[3] = {T.cbx, T.cby, T.cbz};
  for (int pass = 0; pass < 2; ++pass) {
    const bool is_e = pass == 1;
    tfsf_jobs[pass].clear();
    for (const FC& fc : (is_e ? order_e : order_h)) {
      const int ax = fc.face / 2, a = (ax + 1) % 3, b = (ax + 2) % 3;
      const bool negative = fc.face % 2 == 0;
      const int64_t cpos = negative ? task_lo(tt, ax) : task_hi(tt, ax);
      int64_t lo[3] = {tt.x.start, tt.y.start, tt.z.start}, hi[3] = {tt.x.end, tt.y.end, tt.z.end};
      lo[ax] = cpos;
      hi[ax] = cpos + 1;
      const int offset_a = ((is_e && fc.comp == b) || (!is_e && fc.comp == a)) ? 1 : 0;
      if (offset_a)
        hi[a] += 1;
      else
        hi[b] += 1;
      TfsfJobHost J;
      J.comp = fc.comp;
      J.face = fc.face;
      for (int d = 0; d < 3; ++d) {
        J.o[d] = lo[d];
        J.n[d] = hi[d] - lo[d];
      }
      if (negative && !is_e) J.o[ax] -= 1;  // H written at c-1
      const int equation_flag = is_e ? 1 : -1;
      const int different_flag = (fc.comp == a) ? -1 : 1;
      const int direction_flag = negative ? -1 : 1;
      J.coef = (is_e ? ca[ax] : cb[ax]) * (equation_flag * different_flag * direction_flag);
      // written cells inside the field's own shape
      const auto fs = field_shape(is_e ? 0 : 1, fc.comp, nx, ny, nz);
      for (int d = 0; d < 3; ++d)
        if (J.o[d] < 0 || J.o[d] + J.n[d] > fs[d]) return fail(XF_ERR_INVALID, "tfsf: correction outside the field");
      // incident-table indices at every cell
      const int dual = (fc.comp == a) ? b : a;
      const bool inc_e = !is_e;
      const xf_array& tab = inc_e ? T.e_inc : T.h_inc;
      const int64_t g[3] = {tt.x.start, tt.y.start, tt.z.start};
      for (int64_t i = J.o[0]; i < J.o[0] + J.n[0]; ++i)
        for (int64_t j = J.o[1]; j < J.o[1] + J.n[1]; ++j)
          for (int64_t k = J.o[2]; k < J.o[2] + J.n[2]; ++k) {
            int64_t p[3] = {i, j, k};
            if (negative) p[ax] += is_e ? -1 : 1;
            const int64_t l[3] = {p[0] - g[0], p[1] - g[1], p[2] - g[2]};
            // which axes are half-integer for this incident component
            bool half[3];
            for (int d = 0; d < 3; ++d) half[d] = inc_e ? (d == dual) : (d != dual);
            double proj = 0;
            for (int d = 0; d < 3; ++d) {
              const int64_t ix = half[d] ? l[d] + 1 : l[d];
              const xf_array& arr = half[d] ? *ph[d] : *pi[d];
              if (ix < 0 || ix >= arr.shape[0]) return fail(XF_ERR_INVALID, "tfsf: projection index out of range");
              proj += arr.data[ix];
            }
            if (!inc_e) proj -= 0.5;
            if (!(proj >= 0)) return fail(XF_ERR_INVALID, "tfsf: negative projection");
            const int64_t index = static_cast<int64_t>(proj);
            if (index + 1 >= tab.shape[1]) return fail(XF_ERR_INVALID, "tfsf: incident table too short");
          }
      tfsf_jobs[pass].push_back(J);
    }
    // z = const jobs whose inner nodes no other job of the same component
    // writes: XF_TILE_TFSF_ZTAIL hands those to the Yee tiles
    if (dim != 3) continue;
    auto& jobs = tfsf_jobs[pass];
    int tails = 0;
    for (size_t q = 0; q < jobs.size() && tails < 4; ++q) {
      TfsfJobHost& J = jobs[q];
      if (J.face / 2 != 2 || J.n[2] != 1) continue;
      for (int e = 0; e < 2 && J.tail_axis < 0; ++e) {
        if (J.n[e] < 3) continue;
        bool ends_only = true;
        for (size_t p = 0; p < jobs.size() && ends_only; ++p) {
          const TfsfJobHost& X = jobs[p];
          if (p == q || X.comp != J.comp) continue;
          int64_t lo[3], hi[3];
          bool empty = false;
          for (int d = 0; d < 3; ++d) {
            lo[d] = std::max(J.o[d], X.o[d]);
            hi[d] = std::min(J.o[d] + J.n[d], X.o[d] + X.n[d]);
            empty = empty || lo[d] >= hi[d];
          }
          if (empty) continue;
          ends_only = hi[e] == J.o[e] + 1 || lo[e] == J.o[e] + J.n[e] - 1;
        }
        if (ends_only) J.tail_axis = e;
      }
      if (J.tail_axis >= 0) ++tails;
    }
  }
  return XF_OK;
}

// Fusing the PML corrections into the Yee kernels applies them before the
// TFSF corrections; the reference applies TFSF first (simulation_hd.cu:
// 124-160).  Both only add to a field, so the result is the same unless one
// cell receives both, which is rejected here.
int Engine::check_overlaps() {
  XF_TRY(check_ade_pml());
  if (!sc->tfsf) return XF_OK;
  for (int pass = 0; pass < 2; ++pass) {
    const bool is_e = pass == 1;
    for (const auto& J : tfsf_jobs[pass]) {
      for (int p = 0; p < sc->n_pml; ++p) {
        const xf_pml& P = sc->pml[p];
        if (empty_task(P.task)) continue;
        const int ax = P.main_axis, a = (ax + 1) % 3, b = (ax + 2) % 3;
        const int64_t cs = task_lo(P.task, ax), ce = task_hi(P.task, ax);
        const int64_t off = (is_e && cs == 0) ? 1 : 0;
        int64_t lo[3], hi[3];
        lo[ax] = cs + off;
        hi[ax] = ce + off;
        lo[a] = task_lo(P.task, a);
        hi[a] = task_hi(P.task, a) + 1;
        lo[b] = task_lo(P.task, b);
        hi[b] = task_hi(P.task, b) + 1;
        // only the two tangential components are corrected
        if (J.comp == ax) continue;
        bool overlap = true;
        for (int d = 0; d < 3; ++d)
          if (J.o[d] >= hi[d] || J.o[d] + J.n[d] <= lo[d]) overlap = false;
        if (overlap) return fail(XF_ERR_UNSUPPORTED, "tfsf box overlaps a PML slab; move the TF/SF boundary inside the PML");
      }
    }
  }
  return XF_OK;
}

// A dispersive cell's E components take the ADE update (after the E pass)
// and then, in the reference, the CPML correction (domain_hd.cu:35-87:
// updateE, then correctE).  The engine applies the CPML E corrections inside
// the Yee kernels, before the ADE kernel runs, so a dispersive cell inside a
// CPML slab's corrected E box would be updated in the wrong order: rejected.
int Engine::check_ade_pml() const {
  if (!has_ade) return XF_OK;
  for (int p = 0; p < sc->n_pml; ++p) {
    const xf_pml& P = sc->pml[p];
    if (empty_task(P.task)) continue;
    const int ax = P.main_axis, a = (ax + 1) % 3, b = (ax + 2) % 3;
    const int64_t cs = task_lo(P.task, ax), ce = task_hi(P.task, ax);
    const int64_t off = cs == 0 ? 1 : 0;
    int64_t lo[3], hi[3];  // E nodes the face corrects, half-open
    lo[ax] = cs + off;
    hi[ax] = ce + off;
    lo[a] = task_lo(P.task, a);
    hi[a] = task_hi(P.task, a) + 1;
    lo[b] = task_lo(P.task, b);
    hi[b] = task_hi(P.task, b) + 1;
    const int64_t n3[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) hi[d] = std::min(hi[d], n3[d]);  // dispersive cells have i < nx, ...
    for (int64_t i = lo[0]; i < hi[0]; ++i)
      for (int64_t j = lo[1]; j < hi[1]; ++j)
        for (int64_t k = lo[2]; k < hi[2]; ++k)
          if (dispersive(i, j, k))
            return fail(XF_ERR_UNSUPPORTED, "dispersive (ADE) cell (" + std::to_string(i) + ", " + std::to_string(j) +
                                                ", " + std::to_string(k) +
                                                ") inside a CPML slab; keep dispersive media out of the PML");
  }
  return XF_OK;
}

double Engine::bytes_per_cell_step() const {
  // H pass: read H3 + E3, write H3; E pass: read E3 + H3, write E3, plus one
  // coefficient index per pass.
  const double fields = 2.0 * 9.0 * static_cast<double>(esize);
  return fields + 2.0 * id_bytes;
}

// ------------------------------------------------------------- upload
namespace {
// fn(lo, hi) over [0, n) split across the host's cores (the packing loops
// write disjoint elements)
template <typename F>
void host_parallel_for(int64_t n, F fn) {
  const int64_t hw = std::max<int64_t>(1, static_cast<int64_t>(std::thread::hardware_concurrency()));
  const int64_t nth = std::min<int64_t>(std::min<int64_t>(hw, 32), std::max<int64_t>(1, n / 4));
  if (nth <= 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nth);
  int64_t started = 0;
  try {
    for (; started < nth; ++started) {
      const int64_t t = started;
      pool.emplace_back([&, t] { fn(n * t / nth, n * (t + 1) / nth); });
    }
  } catch (const std::system_error&) {
    // a thread could not start: its range and the rest run on this thread
  }
  for (int64_t t = started; t < nth; ++t) fn(n * t / nth, n * (t + 1) / nth);
  for (auto& th : pool) th.join();
}
}  // namespace

int Engine::pack_fields() {
  const xf_array* src[6] = {&sc->ex, &sc->ey, &sc->ez, &sc->hx, &sc->hy, &sc->hz};
  const int prec = plan.precision;
  const double se = 1.0, sh = (prec == XF_BF16_STORE || prec == XF_BF16_COMP || prec == XF_FP16_STORE) ? 376.730313668 : 1.0;
  std::vector<unsigned char> buf(total * esize);
  for (int c = 0; c < 6; ++c) {
    const xf_array& a = *src[c];
    const bool live = dim == 3 || c == 2 || c == 3 || c == 4;
    if (!live) {
      // 2-D: Ex, Ey, Hz are identically zero in the TE updator
      for (int64_t n = 0; n < numel(a); ++n)
        if (a.data[n] != 0.0) return fail(XF_ERR_UNSUPPORTED, "2-D grid with nonzero Ex/Ey/Hz");
      fld[c] = nullptr;
      continue;
    }
    std::fill(buf.begin(), buf.end(), 0);
    const double scale = c < 3 ? se : sh;
    // compensated bf16 packs 64-value blocks: split on block-aligned planes
    host_parallel_for(a.shape[0], [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i)
      for (int64_t j = 0; j < a.shape[1]; ++j)
        for (int64_t k = 0; k < a.shape[2]; ++k) {
          const double v = a.data[(i * a.shape[1] + j) * a.shape[2] + k];
          const int64_t idx = dim == 3 ? i * sx + j * sy + k : i * sx + j;
          if (prec == XF_FP64)
            reinterpret_cast<double*>(buf.data())[idx] = v;
          else if (prec == XF_FP32)
            reinterpret_cast<float*>(buf.data())[idx] = static_cast<float>(v);
          else if (prec == XF_BF16_COMP) {
            // the kernels' blocked layout: 64 bf16 words, then 64 residuals
            uint16_t h;
            int8_t r;
            bf16c_encode(static_cast<float>(v * scale), h, r);
            unsigned char* b = buf.data() + (idx >> 6) * 192;
            reinterpret_cast<uint16_t*>(b)[idx & 63] = h;
            reinterpret_cast<int8_t*>(b + 128)[idx & 63] = r;
          } else if (prec == XF_FP16_STORE)
            reinterpret_cast<__half*>(buf.data())[idx] = __float2half(static_cast<float>(v * scale));
          else
            reinterpret_cast<__hip_bfloat16*>(buf.data())[idx] = __float2bfloat16(static_cast<float>(v * scale));
        }
    });
    XF_TRY(alloc(&fld[c], total * esize));
    XF_HIP(hipMemcpy(fld[c], buf.data(), total * esize, hipMemcpyHostToDevice));
    state.emplace_back(fld[c], total * esize);
  }
  return XF_OK;
}

int Engine::unpack_fields() {
  xf_array* dst[6] = {&sc->ex, &sc->ey, &sc->ez, &sc->hx, &sc->hy, &sc->hz};
  const int prec = plan.precision;
  const double sh = (prec == XF_BF16_STORE || prec == XF_BF16_COMP || prec == XF_FP16_STORE) ? 376.730313668 : 1.0;
  std::vector<unsigned char> buf(total * esize);
  for (int c = 0; c < 6; ++c) {
    if (!fld[c]) continue;
    XF_HIP(hipMemcpy(buf.data(), fld[c], total * esize, hipMemcpyDeviceToHost));
    xf_array& a = *dst[c];
    const double inv = c < 3 ? 1.0 : 1.0 / sh;
    host_parallel_for(a.shape[0], [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i)
      for (int64_t j = 0; j < a.shape[1]; ++j)
        for (int64_t k = 0; k < a.shape[2]; ++k) {
          const int64_t idx = dim == 3 ? i * sx + j * sy + k : i * sx + j;
          double v;
          if (prec == XF_FP64)
            v = reinterpret_cast<const double*>(buf.data())[idx];
          else if (prec == XF_FP32)
            v = reinterpret_cast<const float*>(buf.data())[idx];
          else if (prec == XF_BF16_COMP) {
            const unsigned char* b = buf.data() + (idx >> 6) * 192;
            v = static_cast<double>(bf16c_decode(reinterpret_cast<const uint16_t*>(b)[idx & 63],
                                                 reinterpret_cast<const int8_t*>(b + 128)[idx & 63])) * inv;
          } else if (prec == XF_FP16_STORE)
            v = static_cast<double>(__half2float(reinterpret_cast<const __half*>(buf.data())[idx])) * inv;
          else
            v = static_cast<double>(__bfloat162float(reinterpret_cast<const __hip_bfloat16*>(buf.data())[idx])) * inv;
          a.data[(i * a.shape[1] + j) * a.shape[2] + k] = v;
        }
    });
  }
  return XF_OK;
}

namespace {
struct KeyHash {
  size_t operator()(const std::vector<uint64_t>& k) const {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t v : k) {
      h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    }
    return static_cast<size_t>(h);
  }
};
uint64_t bits(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}
}  // namespace

// A cell is dispersive when any of its ADE coefficients is nonzero; the
// E pass leaves such cells to the ADE kernel.
bool Engine::dispersive(int64_t i, int64_t j, int64_t k) const {
  const xf_ade& A = *sc->ade;
  const int64_t np = A.num_pole;
  const int64_t l3 = (i * (ny + 1) + j) * (nz + 1) + k;
  if (A.coeff_e_j_sum.data[l3] != 0.0) return true;
  if (A.kind == XF_ADE_MLORENTZ && A.coeff_e_e_p.data[l3] != 0.0) return true;
  for (int64_t p = 0; p < np; ++p) {
    const int64_t l4 = l3 * np + p;
    if (A.coeff_j_j.data[l4] != 0.0 || A.coeff_j_e.data[l4] != 0.0 || A.coeff_j_sum_j.data[l4] != 0.0) return true;
    if (A.kind == XF_ADE_MLORENTZ &&
        (A.coeff_j_j_p.data[l4] != 0.0 || A.coeff_j_e_n.data[l4] != 0.0 || A.coeff_j_e_p.data[l4] != 0.0 ||
         A.coeff_j_sum_j_p.data[l4] != 0.0))
      return true;
  }
  return false;
}

// Per-cell coefficient indices.  Dense input: the 9 coefficients of each
// pass at (i, j, k) (each from its own field-shaped array) are deduplicated
// bit-exactly into a table.  Compact input: the material index is the id.
int Engine::build_ids() {
  std::vector<uint32_t> ide(nx * ny * nz), idh(nx * ny * nz);
  etab_h.clear();
  htab_h.clear();
  eflag_h.clear();
  // dispersive cells
  std::vector<unsigned char> disp;
  if (has_ade) {
    disp.assign(nx * ny * nz, 0);
    for (int64_t i = 0; i < nx; ++i)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k) disp[(i * ny + j) * nz + k] = dispersive(i, j, k) ? 1 : 0;
  }
  dense = false;
  if (sc->coeff[0].data == nullptr && plan.coeff_mode == XF_COEFF_DENSE)
    return fail(XF_ERR_UNSUPPORTED, "dense coefficient mode needs the dense coefficient input");
  if (plan.coeff_mode < XF_COEFF_AUTO || plan.coeff_mode > XF_COEFF_DENSE)
    return fail(XF_ERR_INVALID, "unknown coefficient mode");
  if (sc->coeff[0].data == nullptr) {
    const int nm = sc->n_material;
    if (has_ade) return fail(XF_ERR_UNSUPPORTED, "ADE media need the dense coefficient layout");
    for (int m = 0; m < nm; ++m) {
      for (int n = 0; n < 9; ++n) {
        etab_h.push_back(sc->material_coeff[m * 18 + n]);
        htab_h.push_back(sc->material_coeff[m * 18 + 9 + n]);
      }
      eflag_h.push_back(0);
    }
    for (int64_t c = 0; c < nx * ny * nz; ++c) {
      const uint16_t m = sc->cell_material[c];
      if (m >= nm) return fail(XF_ERR_INVALID, "cell material index out of range");
      ide[c] = idh[c] = m;
    }
    ntab_e = ntab_h = nm;
  } else if (plan.coeff_mode != XF_COEFF_DENSE) {
    std::unordered_map<std::vector<uint64_t>, uint32_t, KeyHash> me, mh;
    std::vector<uint64_t> ke(10), kh(9), last_e, last_h;
    uint32_t last_ide = 0, last_idh = 0;
    for (int64_t i = 0; i < nx && !dense; ++i)
      for (int64_t j = 0; j < ny && !dense; ++j)
        for (int64_t k = 0; k < nz; ++k) {
          // past the table's limit the per-cell arrays are streamed instead
          if (me.size() > 65536 || mh.size() > 65536) {
            if (plan.coeff_mode == XF_COEFF_TABLE) break;
            dense = true;
            break;
          }
          const int64_t c = (i * ny + j) * nz + k;
          for (int n = 0; n < 18; ++n) {
            const xf_array& a = sc->coeff[n];
            const double v = a.data[(i * a.shape[1] + j) * a.shape[2] + k];
            if (n < 9)
              ke[n] = bits(v);
            else
              kh[n - 9] = bits(v);
          }
          ke[9] = has_ade ? disp[c] : 0;
          if (ke == last_e) {
            ide[c] = last_ide;
          } else {
            auto it = me.find(ke);
            if (it == me.end()) {
              it = me.emplace(ke, static_cast<uint32_t>(me.size())).first;
              for (int n = 0; n < 9; ++n) {
                double v;
                std::memcpy(&v, &ke[n], 8);
                etab_h.push_back(v);
              }
              eflag_h.push_back(static_cast<unsigned char>(ke[9]));
            }
            ide[c] = it->second;
            last_e = ke;
            last_ide = it->second;
          }
          if (kh == last_h) {
            idh[c] = last_idh;
          } else {
            auto it = mh.find(kh);
            if (it == mh.end()) {
              it = mh.emplace(kh, static_cast<uint32_t>(mh.size())).first;
              for (int n = 0; n < 9; ++n) {
                double v;
                std::memcpy(&v, &kh[n], 8);
                htab_h.push_back(v);
              }
            }
            idh[c] = it->second;
            last_h = kh;
            last_idh = it->second;
          }
        }
    ntab_e = static_cast<int>(me.size());
    ntab_h = static_cast<int>(mh.size());
  } else {
    dense = true;
  }
  if (dense) {
    // the ids flag dispersive cells (E pass row 1), the tables are unused
    for (int64_t c = 0; c < nx * ny * nz; ++c) {
      ide[c] = has_ade ? disp[c] : 0;
      idh[c] = 0;
    }
    etab_h.assign(18, 0.0);
    htab_h.assign(9, 0.0);
    eflag_h = {0, 1};
    ntab_e = 2;
    ntab_h = 1;
    XF_TRY(upload_dense_coeff());
  }
  if (ntab_e > 65536 || ntab_h > 65536)
    return fail(XF_ERR_UNSUPPORTED, "more than 65536 distinct coefficient rows and the table-only coefficient mode");
  id_bytes = (ntab_e <= 256 && ntab_h <= 256 && !wide_ids) ? 1 : 2;
  // scatter ids into the padded layout
  auto put = [&](void** dptr, const std::vector<uint32_t>& id) -> int {
    if (id_bytes == 1) {
      std::vector<uint8_t> v(total, 0);
      for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
          for (int64_t k = 0; k < nz; ++k)
            v[dim == 3 ? i * sx + j * sy + k : i * sx + j] = static_cast<uint8_t>(id[(i * ny + j) * nz + k]);
      return upload_vec(dptr, v);
    }
    std::vector<uint16_t> v(total, 0);
    for (int64_t i = 0; i < nx; ++i)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k)
          v[dim == 3 ? i * sx + j * sy + k : i * sx + j] = static_cast<uint16_t>(id[(i * ny + j) * nz + k]);
    return upload_vec(dptr, v);
  };
  XF_TRY(put(&eid, ide));
  XF_TRY(put(&hid, idh));
  XF_TRY(upload_c(&etab, etab_h.data(), static_cast<int64_t>(etab_h.size())));
  XF_TRY(upload_c(&htab, htab_h.data(), static_cast<int64_t>(htab_h.size())));
  void* fl = nullptr;
  XF_TRY(upload_vec(&fl, eflag_h));
  eflag = static_cast<unsigned char*>(fl);
  if (has_ade) {
    ade_cells.clear();
    for (int64_t i = 0; i < nx; ++i)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k)
          if (disp[(i * ny + j) * nz + k]) ade_cells.push_back((i * (ny + 1) + j) * (nz + 1) + k);
  }
  match_cpsi(ide, idh);
  return XF_OK;
}

// The c_psi value of a corrected node is the curl coefficient, along the
// CPML axis, of the component it corrects, so for the usual scene it equals
// coefficient a 3 + 1 (psi_a) or b 3 + 2 (psi_b) of the node's own table
// row.  Where that holds bit for bit at every node the 3-D march corrects
// (pml_face's ranges), the tiles read it from the table they hold instead of
// the per-node array; anywhere else (kappa-stretched or graded layers, dense
// coefficients) the array stays in use.
void Engine::match_cpsi(const std::vector<uint32_t>& ide, const std::vector<uint32_t>& idh) {
  cpsi_row.assign(sc->n_pml, {-1, -1, -1, -1});
  if (dense || dim != 3) return;
  const int64_t n3[3] = {nx, ny, nz};
  for (int p = 0; p < sc->n_pml; ++p) {
    const xf_pml& S = sc->pml[p];
    if (empty_task(S.task)) continue;
    const int ax = S.main_axis, a = (ax + 1) % 3, b = (ax + 2) % 3;
    const xf_array* cp[4] = {&S.c_ea_psi_hb, &S.c_eb_psi_ha, &S.c_ha_psi_eb, &S.c_hb_psi_ea};
    for (int q = 0; q < 4; ++q) {
      const bool is_e = q < 2, comp_a = q % 2 == 0;
      const int n = comp_a ? a * 3 + 1 : b * 3 + 2;
      const std::vector<uint32_t>& id = is_e ? ide : idh;
      const std::vector<double>& tab = is_e ? etab_h : htab_h;
      const xf_array& C = *cp[q];
      const int64_t cs = task_lo(S.task, ax), ce = task_hi(S.task, ax);
      const int64_t off = (is_e && cs == 0) ? 1 : 0;
      const int64_t nstart = is_e ? S.node_e_start : S.node_h_start;
      // pml_face: a half-open and b closed for E psi_a and H psi_b, the
      // other way round for E psi_b and H psi_a; nodes up to n on each axis
      const bool a_open = is_e == comp_a;
      const int64_t c_lo = std::max<int64_t>(cs + off, 0), c_hi = std::min(ce + off, n3[ax] + 1);
      const int64_t a_lo = std::max<int64_t>(task_lo(S.task, a), 0);
      const int64_t a_hi = std::min(task_hi(S.task, a) + (a_open ? 0 : 1), n3[a] + 1);
      const int64_t b_lo = std::max<int64_t>(task_lo(S.task, b), 0);
      const int64_t b_hi = std::min(task_hi(S.task, b) + (a_open ? 1 : 0), n3[b] + 1);
      bool same = C.ndim == 3;
      for (int64_t c = c_lo; c < c_hi && same; ++c)
        for (int64_t u = a_lo; u < a_hi && same; ++u)
          for (int64_t w = b_lo; w < b_hi && same; ++w) {
            const int64_t layer = c - nstart;
            int64_t pxyz[3], node[3];
            pxyz[ax] = layer;
            pxyz[a] = u;
            pxyz[b] = w;
            node[ax] = c;
            node[a] = u;
            node[b] = w;
            if (layer < 0 || pxyz[0] >= C.shape[0] || pxyz[1] >= C.shape[1] || pxyz[2] >= C.shape[2]) {
              same = false;
              break;
            }
            const double cv = C.data[(pxyz[0] * C.shape[1] + pxyz[1]) * C.shape[2] + pxyz[2]];
            const bool cell = node[0] < nx && node[1] < ny && node[2] < nz;
            const uint32_t row = cell ? id[(node[0] * ny + node[1]) * nz + node[2]] : 0;
            same = bits(cv) == bits(tab[static_cast<size_t>(row) * 9 + n]);
          }
      if (same) cpsi_row[p][q] = n;
    }
  }
}

// The 18 per-cell coefficient arrays in the padded layout, in the compute
// type (converted exactly as the tables are).
int Engine::upload_dense_coeff() {
  const bool f64 = plan.precision == XF_FP64;
  const size_t cb = f64 ? sizeof(double) : sizeof(float);
  std::vector<unsigned char> buf(static_cast<size_t>(total) * cb);
  for (int n = 0; n < 18; ++n) {
    const xf_array& a = sc->coeff[n];
    std::fill(buf.begin(), buf.end(), 0);
    host_parallel_for(nx, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i)
        for (int64_t j = 0; j < ny; ++j)
          for (int64_t k = 0; k < nz; ++k) {
            const double v = a.data[(i * a.shape[1] + j) * a.shape[2] + k];
            const int64_t idx = dim == 3 ? i * sx + j * sy + k : i * sx + j;
            if (f64)
              reinterpret_cast<double*>(buf.data())[idx] = v;
            else
              reinterpret_cast<float*>(buf.data())[idx] = static_cast<float>(v);
          }
    });
    XF_TRY(upload_vec(&dcoef[n], buf));
  }
  return XF_OK;
}

int Engine::upload_pml() {
  pml_order.clear();
  for (int ax = 0; ax < 3; ++ax)
    for (int p = 0; p < sc->n_pml; ++p)
      if (sc->pml[p].main_axis == ax && !empty_task(sc->pml[p].task)) pml_order.push_back(p);
  pml.assign(pml_order.size(), PmlDev{});
  for (size_t q = 0; q < pml_order.size(); ++q) {
    const xf_pml& P = sc->pml[pml_order[q]];
    PmlDev& D = pml[q];
    XF_TRY(upload_c(&D.ca_e, P.coeff_a_e.data, numel(P.coeff_a_e)));
    XF_TRY(upload_c(&D.cb_e, P.coeff_b_e.data, numel(P.coeff_b_e)));
    XF_TRY(upload_c(&D.ca_h, P.coeff_a_h.data, numel(P.coeff_a_h)));
    XF_TRY(upload_c(&D.cb_h, P.coeff_b_h.data, numel(P.coeff_b_h)));
    const xf_array* cp[4] = {&P.c_ea_psi_hb, &P.c_eb_psi_ha, &P.c_ha_psi_eb, &P.c_hb_psi_ea};
    const xf_array* ps[4] = {&P.ea_psi_hb, &P.eb_psi_ha, &P.ha_psi_eb, &P.hb_psi_ea};
    for (int c = 0; c < 4; ++c) {
      XF_TRY(upload_c(&D.cpsi[c], cp[c]->data, numel(*cp[c])));
      XF_TRY(upload_c(&D.psi[c], ps[c]->data, numel(*ps[c])));
      state.emplace_back(D.psi[c], numel(*ps[c]) * (plan.precision == XF_FP64 ? 8 : 4));
    }
  }
  return XF_OK;
}

int Engine::upload_tfsf() {
  if (!sc->tfsf) return XF_OK;
  const xf_tfsf& T = *sc->tfsf;
  const xf_array* pi[3] = {&T.proj_x_int, &T.proj_y_int, &T.proj_z_int};
  const xf_array* ph[3] = {&T.proj_x_half, &T.proj_y_half, &T.proj_z_half};
  for (int d = 0; d < 3; ++d) {
    XF_TRY(upload_c(&proj_int[d], pi[d]->data, numel(*pi[d])));
    XF_TRY(upload_c(&proj_half[d], ph[d]->data, numel(*ph[d])));
  }
  XF_TRY(upload_c(&e_inc, T.e_inc.data, numel(T.e_inc)));
  XF_TRY(upload_c(&h_inc, T.h_inc.data, numel(T.h_inc)));
  return XF_OK;
}

int Engine::upload_nf2ff() {
  nf.clear();
  for (int q = 0; q < sc->n_nf2ff; ++q) {
    const xf_nf2ff& N = sc->nf2ff[q];
    NfDev D{};
    NfPass& P = D.pass;
    int64_t start = 0, off = 0;
    for (int face = 0; face < 6; ++face) {
      const xf_task& t = N.face_task[face];
      NfFace& F = P.face[face];
      F.start = start;
      F.off = off;
      if (empty_task(t)) {
        F.o[0] = F.o[1] = F.o[2] = 0;
        F.n[0] = F.n[1] = F.n[2] = 0;
        continue;
      }
      F.o[0] = static_cast<int>(t.x.start);
      F.o[1] = static_cast<int>(t.y.start);
      F.o[2] = static_cast<int>(t.z.start);
      F.n[0] = static_cast<int>(t.x.end - t.x.start);
      F.n[1] = static_cast<int>(t.y.end - t.y.start);
      F.n[2] = static_cast<int>(t.z.end - t.z.start);
      const int64_t cells = static_cast<int64_t>(F.n[0]) * F.n[1] * F.n[2];
      start += cells;
      off += cells;
    }
    P.total = start;
    P.block = off;
    P.nfreq = N.n_freq;
    P.nt = nt;
    // current buffer [freq][face][comp][cell]
    std::vector<double2> cur(static_cast<size_t>(N.n_freq) * off * 4);
    std::vector<double2> te(static_cast<size_t>(N.n_freq) * nt), th(te.size());
    for (int f = 0; f < N.n_freq; ++f) {
      for (int face = 0; face < 6; ++face) {
        const NfFace& F = P.face[face];
        const int64_t cells = static_cast<int64_t>(F.n[0]) * F.n[1] * F.n[2];
        for (int c = 0; c < 4; ++c) {
          const xf_array& a = N.freq[f].cur[face][c];
          for (int64_t n = 0; n < cells; ++n)
            cur[f * off * 4 + F.off * 4 + c * cells + n] = make_double2(a.data[2 * n], a.data[2 * n + 1]);
        }
      }
      for (int64_t t = 0; t < nt; ++t) {
        te[f * nt + t] = make_double2(N.freq[f].transform_e.data[2 * t], N.freq[f].transform_e.data[2 * t + 1]);
        th[f * nt + t] = make_double2(N.freq[f].transform_h.data[2 * t], N.freq[f].transform_h.data[2 * t + 1]);
      }
    }
    void *pc, *pe, *ph;
    XF_TRY(upload_vec(&pc, cur));
    XF_TRY(upload_vec(&pe, te));
    XF_TRY(upload_vec(&ph, th));
    D.cur = static_cast<double2*>(pc);
    state.emplace_back(pc, cur.size() * sizeof(double2));
    D.te = static_cast<double2*>(pe);
    D.th = static_cast<double2*>(ph);
    P.cur = D.cur;
    P.te = D.te;
    P.th = D.th;
    nf.push_back(D);
  }
  return XF_OK;
}

int Engine::upload_ade() {
  if (!has_ade) return XF_OK;
  const xf_ade& A = *sc->ade;
  const int64_t n = static_cast<int64_t>(ade_cells.size());
  const int64_t np = A.num_pole;
  ade = AdeDev{};
  ade.kind = A.kind;
  ade.np = static_cast<int>(np);
  ade.n = n;
  std::vector<int64_t> idx(n);
  std::vector<double> cejs(n), ceep(n), cjj(np * n), cjjp(np * n), cjen(np * n), cje(np * n),
      cjep(np * n), cjsj(np * n), cjsjp(np * n), eprev(3 * n), jc(3 * np * n), jp(3 * np * n);
  const xf_array* J[3] = {&A.jx, &A.jy, &A.jz};
  const xf_array* JP[3] = {&A.jx_prev, &A.jy_prev, &A.jz_prev};
  const xf_array* EP[3] = {&A.ex_prev, &A.ey_prev, &A.ez_prev};
  const bool ml = A.kind == XF_ADE_MLORENTZ;
  for (int64_t c = 0; c < n; ++c) {
    const int64_t l3 = ade_cells[c];
    const int64_t k = l3 % (nz + 1), j = (l3 / (nz + 1)) % (ny + 1), i = l3 / ((nz + 1) * (ny + 1));
    idx[c] = i * sx + j * sy + k;
    cejs[c] = A.coeff_e_j_sum.data[l3];
    ceep[c] = ml ? A.coeff_e_e_p.data[l3] : 0.0;
    for (int64_t p = 0; p < np; ++p) {
      const int64_t l4 = l3 * np + p, pc = p * n + c;
      cjj[pc] = A.coeff_j_j.data[l4];
      cje[pc] = A.coeff_j_e.data[l4];
      cjsj[pc] = A.coeff_j_sum_j.data[l4];
      if (ml) {
        cjjp[pc] = A.coeff_j_j_p.data[l4];
        cjen[pc] = A.coeff_j_e_n.data[l4];
        cjep[pc] = A.coeff_j_e_p.data[l4];
        cjsjp[pc] = A.coeff_j_sum_j_p.data[l4];
      }
      for (int comp = 0; comp < 3; ++comp) {
        jc[(comp * np + p) * n + c] = J[comp]->data[l4];
        if (ml) jp[(comp * np + p) * n + c] = JP[comp]->data[l4];
      }
    }
    if (ml)
      for (int comp = 0; comp < 3; ++comp) eprev[comp * n + c] = EP[comp]->data[l3];
  }
  void* p;
  XF_TRY(upload_vec(&p, idx));
  ade.idx = static_cast<const int64_t*>(p);
  // the distinct coefficient rows (compared by bit pattern, so the table
  // gives back exactly each cell's values) and each cell's row id; the
  // dense coefficient mode keeps the reference's per-cell arrays here too
  const int tw = static_cast<int>(2 + 7 * np);
  std::map<std::vector<uint64_t>, int> rows;
  std::vector<double> tab;
  std::vector<uint16_t> row_of(n);
  bool use_tab = !dense;
  std::vector<uint64_t> key(tw);
  for (int64_t c = 0; c < n && use_tab; ++c) {
    double r[2 + 7 * 64];
    r[0] = cejs[c];
    r[1] = ceep[c];
    for (int64_t q = 0; q < np; ++q) {
      const int64_t pc = q * n + c;
      double* o = r + 2 + 7 * q;
      o[0] = cjj[pc];
      o[1] = cje[pc];
      o[2] = cjsj[pc];
      o[3] = ml ? cjjp[pc] : 0.0;
      o[4] = ml ? cjen[pc] : 0.0;
      o[5] = ml ? cjep[pc] : 0.0;
      o[6] = ml ? cjsjp[pc] : 0.0;
    }
    for (int w = 0; w < tw; ++w) std::memcpy(&key[w], &r[w], sizeof(double));
    auto it = rows.find(key);
    if (it == rows.end()) {
      if (rows.size() >= 65535) {
        use_tab = false;
        break;
      }
      it = rows.emplace(key, static_cast<int>(rows.size())).first;
      tab.insert(tab.end(), r, r + tw);
    }
    row_of[c] = static_cast<uint16_t>(it->second);
  }
  ade.tw = tw;
  if (use_tab && n > 0) {
    XF_TRY(upload_vec(&p, row_of));
    ade.row = static_cast<const uint16_t*>(p);
    XF_TRY(upload_vec(&p, tab));
    ade.tab = static_cast<const double*>(p);
  } else {
    ade.row = nullptr;
    ade.tab = nullptr;
    XF_TRY(upload_vec(&p, cejs));
    ade.cejs = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, ceep));
    ade.ceep = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, cjj));
    ade.cjj = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, cjjp));
    ade.cjjp = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, cjen));
    ade.cjen = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, cje));
    ade.cje = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, cjep));
    ade.cjep = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, cjsj));
    ade.cjsj = static_cast<const double*>(p);
    XF_TRY(upload_vec(&p, cjsjp));
    ade.cjsjp = static_cast<const double*>(p);
  }
  ade_rows = use_tab ? static_cast<int64_t>(rows.size()) : -1;
  XF_TRY(upload_vec(&p, eprev));
  ade.eprev = static_cast<double*>(p);
  state.emplace_back(p, eprev.size() * sizeof(double));
  XF_TRY(upload_vec(&p, jc));
  ade.jc = static_cast<double*>(p);
  state.emplace_back(p, jc.size() * sizeof(double));
  XF_TRY(upload_vec(&p, jp));
  ade.jp = static_cast<double*>(p);
  state.emplace_back(p, jp.size() * sizeof(double));
  return XF_OK;
}

namespace {
// IEEE binary16 bits -> float (exact)
float half_bits_to_float(uint16_t h) {
  const int ex = (h >> 10) & 0x1f, man = h & 0x3ff;
  float v;
  if (ex == 0)
    v = std::ldexp(static_cast<float>(man), -24);
  else if (ex == 31)
    v = man ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
  else
    v = std::ldexp(static_cast<float>(man | 0x400), ex - 25);
  return (h >> 15) ? -v : v;
}
}  // namespace

int64_t Engine::lb_recorded_after(int64_t step) const {
  const xf_learned_boundary& L = *sc->lb;
  if (step <= 0) return lb_rec_count;
  const int64_t slots = std::min<int64_t>(L.rec_slots, (step + L.rec_every - 1) / L.rec_every);
  return std::max(lb_rec_count, slots);
}

int Engine::upload_additions() {
  pts = PointDev{};
  if (sc->n_point > 0) {
    const int n = sc->n_point;
    std::vector<int> comp(n);
    std::vector<int64_t> idx(n);
    std::vector<double> val(static_cast<size_t>(n) * nt);
    for (int q = 0; q < n; ++q) {
      const xf_point_source& P = sc->point[q];
      comp[q] = P.comp;
      idx[q] = P.i * sx + P.j * sy + P.k;
      std::copy(P.value.data, P.value.data + nt, val.begin() + static_cast<size_t>(q) * nt);
    }
    void *pc, *pi, *pv;
    XF_TRY(upload_vec(&pc, comp));
    XF_TRY(upload_vec(&pi, idx));
    XF_TRY(upload_vec(&pv, val));
    pts.n = n;
    pts.nt = nt;
    pts.comp = static_cast<const int*>(pc);
    pts.idx = static_cast<const int64_t*>(pi);
    pts.value = static_cast<const double*>(pv);
  }
  movies.clear();
  for (int32_t q = 0; q < sc->n_movie; ++q) {
    const xf_movie& M = sc->movie[q];
    MovieDev D{};
    D.field = M.field;
    const xf_range* r[3] = {&M.task.x, &M.task.y, &M.task.z};
    for (int d = 0; d < 3; ++d) {
      D.o[d] = static_cast<int>(r[d]->start);
      D.n[d] = static_cast<int>(r[d]->end - r[d]->start);
    }
    D.interval = M.interval;
    D.frames = M.data.shape[0];
    D.box = static_cast<int64_t>(D.n[0]) * D.n[1] * D.n[2];
    void* p;
    XF_TRY(alloc(&p, static_cast<size_t>(D.frames * D.box) * sizeof(double)));
    XF_HIP(hipMemcpy(p, M.data.data, static_cast<size_t>(D.frames * D.box) * sizeof(double), hipMemcpyHostToDevice));
    D.data = static_cast<double*>(p);
    state.emplace_back(p, static_cast<size_t>(D.frames * D.box) * sizeof(double));
    movies.push_back(D);
  }
  movie_items = nullptr;
  movie_start = nullptr;
  if (!movies.empty()) {
    std::vector<int64_t> st(1, 0);
    for (const MovieDev& D : movies) st.push_back(st.back() + D.box);
    XF_TRY(upload_vec(&movie_items, movies));
    void* sp = nullptr;
    XF_TRY(upload_vec(&sp, st));
    movie_start = static_cast<int64_t*>(sp);
  }
  lbd = LbDev{};
  if (!has_lb) return XF_OK;
  const xf_learned_boundary& L = *sc->lb;
  const int64_t n3[3] = {nx, ny, nz};
  lbd.mode = L.mode;
  lbd.faces = L.faces;
  lbd.inset = L.inset;
  lbd.mur_k = static_cast<float>(L.mur_k);
  lbd.rec_every = L.rec_every > 0 ? L.rec_every : 1;
  lbd.rec_slots = L.rec_slots;
  for (int d = 0; d < 3; ++d) lbd.n[d] = static_cast<int>(n3[d]);
  int64_t off = 0, col = 0;
  for (int f = 0; f < 6; ++f) {
    const int c = f / 2, a = (c + 1) % 3, b = (c + 2) % 3;
    const int64_t face_off = off;
    off += 16 * (n3[a] + 1) * (n3[b] + 1);
    if (!((L.faces >> f) & 1)) continue;
    LbFace& F = lbd.face[lbd.nface++];
    F.f = f;
    F.c = c;
    F.a = a;
    F.b = b;
    F.positive = f & 1;
    F.tr = c == 1 ? 1 : 0;
    F.hist_off = face_off;
    F.na1 = n3[a] + 1;
    F.nb1 = n3[b] + 1;
    for (int cc = 0; cc < 2; ++cc) {
      F.start[cc] = col;
      F.p1[cc] = static_cast<int>(n3[a] - L.inset - (cc == 0 ? 1 : 0));
      F.q1[cc] = static_cast<int>(n3[b] - L.inset - (cc == 1 ? 1 : 0));
      const int64_t cols = static_cast<int64_t>(F.p1[cc] - L.inset + 1) * (F.q1[cc] - L.inset + 1);
      if (cols >= (int64_t(1) << 32)) return fail(XF_ERR_UNSUPPORTED, "learned layer: a face holds 2^32 columns or more");
      col += cols;
    }
  }
  lbd.total = col;
  if (L.mode == XF_LB_APPLY && col >= (int64_t(1) << 32))
    return fail(XF_ERR_UNSUPPORTED, "learned layer: 2^32 columns or more");
  if (L.mode == XF_LB_APPLY) {
    // edge columns (their node also lies on another face of the layer),
    // found as lb_column / lb_other find them on the device
    std::vector<uint32_t> edges;
    for (int fi = 0; fi < lbd.nface; ++fi) {
      const LbFace& F = lbd.face[fi];
      for (int cc = 0; cc < 2; ++cc) {
        const int other_axis = cc == 0 ? F.b : F.a;
        const int lim = static_cast<int>(n3[other_axis]);
        for (int p = lbd.inset; p <= F.p1[cc]; ++p)
          for (int q = lbd.inset; q <= F.q1[cc]; ++q) {
            const int x = cc == 0 ? q : p;  // the in-plane index along other_axis
            const bool edge = (x == 0 && ((L.faces >> (2 * other_axis)) & 1)) ||
                              (x == lim && ((L.faces >> (2 * other_axis + 1)) & 1));
            if (!edge) continue;
            // thread-order index (lb_column's decode)
            const int64_t c = F.tr ? F.start[cc] + static_cast<int64_t>(q - lbd.inset) * (F.p1[cc] - lbd.inset + 1) +
                                         (p - lbd.inset)
                                   : F.start[cc] + static_cast<int64_t>(p - lbd.inset) * (F.q1[cc] - lbd.inset + 1) +
                                         (q - lbd.inset);
            edges.push_back(static_cast<uint32_t>(c));
          }
      }
    }
    lbd.n_edge = static_cast<int64_t>(edges.size());
    if (!edges.empty()) {
      void* pe;
      XF_TRY(upload_vec(&pe, edges));
      lbd.edge = static_cast<const uint32_t*>(pe);
    }
  }
  std::vector<float> hist(off);
  for (int64_t i = 0; i < off; ++i) hist[lb_dev_hist_index(i)] = static_cast<float>(L.history.data[i]);
  void* ph;
  XF_TRY(upload_vec(&ph, hist));
  lbd.hist = static_cast<float*>(ph);
  state.emplace_back(ph, hist.size() * sizeof(float));
  if (L.mode == XF_LB_APPLY) {
    // the weights as per-lane MFMA operands (LbPack)
    const uint16_t* W1 = L.weights;
    const uint16_t* b1 = W1 + XF_LB_HID * XF_LB_IN;
    const uint16_t* W2 = b1 + XF_LB_HID;
    const uint16_t* b2 = W2 + XF_LB_HID * XF_LB_HID;
    const uint16_t* w3 = b2 + XF_LB_HID;
    std::vector<LbPack> pk(1);
    LbPack& P = pk[0];
    for (int lane = 0; lane < 64; ++lane) {
      const int g = lane / 32, m = lane % 32;
      for (int j = 0; j < 8; ++j) {
        const int k = 8 * g + j;
        P.a1[lane][j] = k < XF_LB_IN ? W1[m * XF_LB_IN + k] : 0;
        for (int p = 0; p < 2; ++p) P.a2[p][lane][j] = W2[m * XF_LB_HID + 16 * p + 8 * (j / 4) + 4 * g + j % 4];
      }
      for (int i = 0; i < 16; ++i) {
        const int u = 8 * (i / 4) + 4 * g + i % 4;
        P.c1[i][lane] = half_bits_to_float(b1[u]);
        P.c2[i][lane] = half_bits_to_float(b2[u]);
        P.w3[i][lane] = half_bits_to_float(w3[u]);
      }
    }
    P.b3 = half_bits_to_float(w3[XF_LB_HID]);
    void* pw;
    XF_TRY(upload_vec(&pw, pk));
    lbd.pack = static_cast<const LbPack*>(pw);
  } else {
    const size_t nrec = static_cast<size_t>(L.rec_slots) * col * 13;
    std::vector<float> rec(L.rec, L.rec + nrec);
    void* pr;
    XF_TRY(upload_vec(&pr, rec));
    lbd.rec = static_cast<float*>(pr);
    state.emplace_back(pr, nrec * sizeof(float));
  }
  lb_rec_count = L.rec_count;
  return XF_OK;
}

int Engine::upload_nf2ff_td() {
  nftd.clear();
  for (int32_t q = 0; q < sc->n_nf2ff_td; ++q) {
    const xf_nf2ff_td& T = sc->nf2ff_td[q];
    NtPass P;
    int64_t prev_total = 0;
    XF_TRY(plan_nf2ff_td(T, P, &prev_total));
    for (int ax = 0; ax < 3; ++ax) {
      void* p;
      XF_TRY(upload_vec(&p, std::vector<double>(T.e_node[ax].data, T.e_node[ax].data + numel(T.e_node[ax]))));
      P.e_node[ax] = static_cast<const double*>(p);
      XF_TRY(upload_vec(&p, std::vector<double>(T.h_node[ax].data, T.h_node[ax].data + numel(T.h_node[ax]))));
      P.h_node[ax] = static_cast<const double*>(p);
      XF_TRY(upload_vec(&p, std::vector<double>(T.e_size[ax].data, T.e_size[ax].data + numel(T.e_size[ax]))));
      P.e_size[ax] = static_cast<const double*>(p);
    }
    std::vector<double> pot(static_cast<size_t>(24 * P.n_pot), 0.0), prev(static_cast<size_t>(prev_total));
    for (int face = 0; face < 6; ++face) {
      if (empty_task(T.face_task[face])) continue;
      const NtFace& F = P.face[face];
      const int64_t plane = numel(T.prev[face][0]);
      for (int c = 0; c < 4; ++c) {
        std::copy(T.pot[face][c].data, T.pot[face][c].data + P.n_pot, pot.begin() + (face * 4 + c) * P.n_pot);
        std::copy(T.prev[face][c].data, T.prev[face][c].data + plane, prev.begin() + F.prev_off + c * plane);
      }
    }
    void *pp, *pv, *pa, *ph;
    XF_TRY(upload_vec(&pp, pot));
    XF_TRY(upload_vec(&pv, prev));
    state.emplace_back(pp, pot.size() * sizeof(double));
    state.emplace_back(pv, prev.size() * sizeof(double));
    XF_TRY(alloc(&pa, static_cast<size_t>(P.nblk) * 4 * P.maxspan * sizeof(double)));
    XF_TRY(alloc(&ph, static_cast<size_t>(P.nblk) * 4 * sizeof(int)));
    P.pot = static_cast<double*>(pp);
    P.prev = static_cast<double*>(pv);
    P.part = static_cast<double*>(pa);
    P.hdr = static_cast<int*>(ph);
    nftd.push_back(P);
  }
  return XF_OK;
}

int Engine::upload() {
  release();
  XF_HIP(hipSetDevice(device));
  XF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  XF_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
  XF_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  XF_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  esize = plan.precision == XF_FP64 ? 8 : plan.precision == XF_FP32 ? 4 : plan.precision == XF_BF16_COMP ? 3 : 2;
  // rows start on a 128-byte boundary (a 64-value block for the 3-byte
  // compensated format, whose layout is blocked by 64)
  const int64_t align = plan.precision == XF_BF16_COMP ? 64 : 128 / static_cast<int64_t>(esize);
  if (dim == 3) {
    const int64_t pz = ((nz + 1 + align - 1) / align) * align;
    sy = pz;
    sx = (ny + 1) * pz;
    total = (nx + 1) * sx;
  } else {
    const int64_t py = ((ny + 1 + align - 1) / align) * align;
    sy = 1;
    sx = py;
    total = (nx + 1) * sx;
  }
  if (plan.precision == XF_BF16_COMP && static_cast<uint64_t>(total) * 3 >= (uint64_t(1) << 32))
    return fail(XF_ERR_UNSUPPORTED, "compensated bf16 storage: grid above 2^32 bytes per component");
  XF_TRY(build_ids());
  XF_TRY(pack_fields());
  XF_TRY(upload_pml());
  XF_TRY(upload_tfsf());
  XF_TRY(upload_nf2ff());
  XF_TRY(upload_nf2ff_td());
  XF_TRY(upload_ade());
  XF_TRY(upload_additions());
  void* tp;
  XF_TRY(alloc(&tp, sizeof(int)));
  t_dev = static_cast<int*>(tp);
  XF_TRY(alloc(&tp, 64));
  wd_dev = static_cast<unsigned*>(tp);
  set_tiling(plan.tile_y, plan.chunks_x, plan.tiling_flags);
  XF_HIP(hipDeviceSynchronize());
  uploaded = true;
  return XF_OK;
}

// Tiling of the 3-D marching kernels.  chunks_x = 0 gives short marches
// (3 or 4 planes FP64, 6 FP32, 16 for 16-bit and compensated storage):
// measured at 536^3 on MI355X, many short
// marches beat long ones (a long march leaves a partly empty tail round of
// blocks and keeps far-apart planes open at once), while the carried plane
// still saves a neighbour re-read every step of the march.
void Engine::set_tiling(int ty, int chunks_x, int flags) {
  // kept in the plan, so a later upload (which re-applies the plan) keeps it
  plan.tile_y = ty;
  plan.chunks_x = chunks_x;
  plan.tiling_flags = flags;
  // 64 x tile_y threads must fit the kernels' __launch_bounds__(256)
  tile_y = ty > 0 ? std::min(ty, 4) : 1;
  if (flags & XF_TILE_SET) {
    xcd = (flags & XF_TILE_XCD_REMAP) ? 1 : 0;
    nt_streams = (flags & XF_TILE_NT) ? 1 : 0;
    pml_split = (flags & XF_TILE_PML_SPLIT) ? 1 : 0;
    pml_ztail = (flags & XF_TILE_PML_ZTAIL) ? 1 : 0;
    tfsf_zt = (flags & XF_TILE_TFSF_ZTAIL) ? 1 : 0;
    klanes = (flags & XF_TILE_K32) ? 32 : 64;
    pml_streams = (flags & XF_TILE_PML_STREAMS) ? 1 : 0;
  } else {
    // defaults: non-temporal streams; z-face CPML corrections split off
    // (measured at 536^3: FP32 +3 %, FP64 even; 64 small grids: +13 %),
    // except FP64, where 32-lane tiles that apply them at the end of each
    // 3-plane march measured fastest at 536^3 (32.5 against 31.8 Gcells/s on
    // one box; FP32 keeps the split, 56.7 against 53.8)
    xcd = 0;
    nt_streams = 1;
    pml_split = 1;
    pml_ztail = esize == 8 ? 1 : 0;
    tfsf_zt = 0;
    pml_streams = 0;
    klanes = esize == 8 ? 32 : 64;
  }
  if (dim == 3) {
    nkb = static_cast<int>((nz + 1 + klanes - 1) / klanes);
    njb = static_cast<int>((ny + 1 + rows() - 1) / rows());
    // default march: 4 planes for 8-byte fields (3 with the z faces at the
    // tile end), 6 for 4-byte ones, 16 for narrower ones (the measured optima
    // on MI355X at 536^3; FP32 6 against 8 planes: 56.9 against 56.2 Gcells/s;
    // bf16 16 against 8: 63.5 against 61.4, fp16 61.6 against 59.8,
    // compensated bf16 37.0 against 35.8)
    const int64_t planes = esize == 8 ? (ztail() ? 3 : 4) : esize == 4 ? 6 : 16;
    int64_t nch = chunks_x;
    if (nch <= 0) nch = std::max<int64_t>(1, (nx + 1 + planes - 1) / planes);
    nch = std::min<int64_t>(nch, nx + 1);
    chunk = static_cast<int>((nx + 1 + nch - 1) / nch);
    nch_eff = static_cast<int>((nx + 1 + chunk - 1) / chunk);
  }
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  graph_exec = nullptr;
  graph = nullptr;
  graph_ready = false;
  if (pml_streams && dim == 3 && stream && build_tile_lists() != XF_OK) pml_streams = 0;
  layout_desc = std::string("dim=") + std::to_string(dim) + " pad_stride=" + std::to_string(dim == 3 ? sy : sx) +
                " id_bytes=" + std::to_string(id_bytes) + " coeff=" + (dense ? "dense" : "table") +
                " etab=" + std::to_string(ntab_e) +
                " htab=" + std::to_string(ntab_h) + " chunk=" + std::to_string(chunk) +
                " nchunk=" + std::to_string(nch_eff) + " tile_y=" + std::to_string(tile_y) +
                " xcd=" + std::to_string(xcd) + " nt=" + std::to_string(nt_streams) + " pml_split=" + std::to_string(split_z() ? 1 : 0) + " pml_ztail=" + std::to_string(ztail() ? 1 : 0) +
                " tfsf_ztail=" + std::to_string(tfsf_ztail() ? 1 : 0) +
                " klanes=" + std::to_string(klanes) + " pml_streams=" + std::to_string(pml_streams) +
                " pml_faces=" + std::to_string(pml.size()) + " cpsi_tab=" + std::to_string(cpsi_matched()) +
                " ade_cells=" + std::to_string(ade_cells.size()) +
                (has_ade ? " ade_rows=" + std::to_string(ade_rows) : std::string());
}

int Engine::snapshot() {
  if (!uploaded) return fail(XF_ERR_STATE, "snapshot before upload");
  XF_HIP(hipSetDevice(device));
  if (backup.size() != state.size()) {
    backup.assign(state.size(), nullptr);
    for (size_t q = 0; q < state.size(); ++q) XF_TRY(alloc(&backup[q], state[q].second));
  }
  for (size_t q = 0; q < state.size(); ++q)
    XF_HIP(hipMemcpyAsync(backup[q], state[q].first, state[q].second, hipMemcpyDeviceToDevice, stream));
  XF_HIP(hipStreamSynchronize(stream));
  lb_rec_count_saved = lb_rec_count;
  return XF_OK;
}

int Engine::restore() {
  if (backup.size() != state.size()) return fail(XF_ERR_STATE, "restore without snapshot");
  XF_HIP(hipSetDevice(device));
  for (size_t q = 0; q < state.size(); ++q)
    XF_HIP(hipMemcpyAsync(state[q].first, backup[q], state[q].second, hipMemcpyDeviceToDevice, stream));
  XF_HIP(hipStreamSynchronize(stream));
  lb_rec_count = lb_rec_count_saved;
  return XF_OK;
}

Tiling Engine::tiling() const {
  Tiling t;
  t.nkb = nkb;
  t.njb = njb;
  t.nch = nch_eff;
  t.chunk = chunk;
  t.xcd = xcd;
  return t;
}

// ------------------------------------------------------------- download
int Engine::download() {
  if (!uploaded) return fail(XF_ERR_STATE, "download before upload");
  XF_HIP(hipStreamSynchronize(stream));
  XF_TRY(unpack_fields());
  for (size_t q = 0; q < pml_order.size(); ++q) {
    xf_pml& P = sc->pml[pml_order[q]];
    xf_array* ps[4] = {&P.ea_psi_hb, &P.eb_psi_ha, &P.ha_psi_eb, &P.hb_psi_ea};
    for (int c = 0; c < 4; ++c) XF_TRY(download_c(ps[c]->data, pml[q].psi[c], numel(*ps[c])));
  }
  for (int q = 0; q < static_cast<int>(nf.size()); ++q) {
    const xf_nf2ff& N = sc->nf2ff[q];
    const NfPass& P = nf[q].pass;
    std::vector<double2> cur(static_cast<size_t>(N.n_freq) * P.block * 4);
    XF_HIP(hipMemcpy(cur.data(), nf[q].cur, cur.size() * sizeof(double2), hipMemcpyDeviceToHost));
    for (int f = 0; f < N.n_freq; ++f)
      for (int face = 0; face < 6; ++face) {
        const NfFace& F = P.face[face];
        const int64_t cells = static_cast<int64_t>(F.n[0]) * F.n[1] * F.n[2];
        for (int c = 0; c < 4; ++c) {
          xf_array& a = N.freq[f].cur[face][c];
          for (int64_t n = 0; n < cells; ++n) {
            const double2 v = cur[f * P.block * 4 + F.off * 4 + c * cells + n];
            a.data[2 * n] = v.x;
            a.data[2 * n + 1] = v.y;
          }
        }
      }
  }
  for (size_t q = 0; q < nftd.size(); ++q) {
    const xf_nf2ff_td& T = sc->nf2ff_td[q];
    const NtPass& P = nftd[q];
    for (int face = 0; face < 6; ++face) {
      if (empty_task(T.face_task[face])) continue;
      const int64_t plane = numel(T.prev[face][0]);
      for (int c = 0; c < 4; ++c) {
        XF_HIP(hipMemcpy(T.pot[face][c].data, P.pot + (face * 4 + c) * P.n_pot, P.n_pot * sizeof(double),
                         hipMemcpyDeviceToHost));
        XF_HIP(hipMemcpy(T.prev[face][c].data, P.prev + P.face[face].prev_off + c * plane,
                         plane * sizeof(double), hipMemcpyDeviceToHost));
      }
    }
  }
  if (has_ade) {
    const xf_ade& A = *sc->ade;
    const int64_t n = ade.n, np = ade.np;
    std::vector<double> eprev(3 * n), jc(3 * np * n), jp(3 * np * n);
    XF_HIP(hipMemcpy(jc.data(), ade.jc, jc.size() * 8, hipMemcpyDeviceToHost));
    XF_HIP(hipMemcpy(jp.data(), ade.jp, jp.size() * 8, hipMemcpyDeviceToHost));
    XF_HIP(hipMemcpy(eprev.data(), ade.eprev, eprev.size() * 8, hipMemcpyDeviceToHost));
    const xf_array* J[3] = {&A.jx, &A.jy, &A.jz};
    const xf_array* JP[3] = {&A.jx_prev, &A.jy_prev, &A.jz_prev};
    const xf_array* EP[3] = {&A.ex_prev, &A.ey_prev, &A.ez_prev};
    const bool ml = A.kind == XF_ADE_MLORENTZ;
    for (int64_t c = 0; c < n; ++c) {
      const int64_t l3 = ade_cells[c];
      for (int comp = 0; comp < 3; ++comp) {
        for (int64_t p = 0; p < np; ++p) {
          J[comp]->data[l3 * np + p] = jc[(comp * np + p) * n + c];
          if (ml) JP[comp]->data[l3 * np + p] = jp[(comp * np + p) * n + c];
        }
        if (ml) EP[comp]->data[l3] = eprev[comp * n + c];
      }
    }
  }
  for (size_t q = 0; q < movies.size(); ++q)
    XF_HIP(hipMemcpy(sc->movie[q].data.data, movies[q].data,
                     static_cast<size_t>(movies[q].frames * movies[q].box) * sizeof(double), hipMemcpyDeviceToHost));
  if (has_lb) {
    xf_learned_boundary& L = *sc->lb;
    std::vector<float> hist(L.history.shape[0]);
    XF_HIP(hipMemcpy(hist.data(), lbd.hist, hist.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < hist.size(); ++i) L.history.data[i] = hist[lb_dev_hist_index(static_cast<int64_t>(i))];
    if (L.mode == XF_LB_RECORD) {
      XF_HIP(hipMemcpy(L.rec, lbd.rec, static_cast<size_t>(L.rec_slots) * lbd.total * 13 * sizeof(float),
                       hipMemcpyDeviceToHost));
      L.rec_count = lb_rec_count;
    }
  }
  sc->current_step = cur_step;
  return XF_OK;
}

// ------------------------------------------------------------- launches
namespace {
// faces: 0 every face, 1 all but the z faces, 2 the z faces only (faces
// keep their registration order: X, Y, Z)
template <typename C>
void fill_pml_pass(const Engine& E, PmlPass<C>& P, bool is_e, int faces = 0) {
  P.n = 0;
  for (int q = 0; q < static_cast<int>(E.pml_order.size()); ++q) {
    const xf_pml& S = E.sc->pml[E.pml_order[q]];
    const auto& D = E.pml[q];
    const int ax = S.main_axis, a = (ax + 1) % 3, b = (ax + 2) % 3;
    if ((faces == 1 && ax == 2) || (faces == 2 && ax != 2)) continue;
    PmlFace<C>& F = P.f[P.n++];
    F.axis = ax;
    const int64_t cs = task_lo(S.task, ax), ce = task_hi(S.task, ax);
    const int64_t off = (is_e && cs == 0) ? 1 : 0;
    F.c0 = static_cast<int>(cs + off);
    F.c1 = static_cast<int>(ce + off);
    F.a0 = static_cast<int>(task_lo(S.task, a));
    F.a1 = static_cast<int>(task_hi(S.task, a));
    F.b0 = static_cast<int>(task_lo(S.task, b));
    F.b1 = static_cast<int>(task_hi(S.task, b));
    F.gstart = static_cast<int>(is_e ? S.global_e_start : S.global_h_start);
    F.nstart = static_cast<int>(is_e ? S.node_e_start : S.node_h_start);
    F.off_c = static_cast<int>(S.offset_c);
    F.ca = static_cast<const C*>(is_e ? D.ca_e : D.ca_h);
    F.cb = static_cast<const C*>(is_e ? D.cb_e : D.cb_h);
    const int ia = is_e ? 0 : 2, ib = is_e ? 1 : 3;
    const xf_array* ps[4] = {&S.ea_psi_hb, &S.eb_psi_ha, &S.ha_psi_eb, &S.hb_psi_ea};
    F.psi_a = static_cast<C*>(D.psi[ia]);
    F.cpsi_a = static_cast<const C*>(D.cpsi[ia]);
    F.psi_b = static_cast<C*>(D.psi[ib]);
    F.cpsi_b = static_cast<const C*>(D.cpsi[ib]);
    F.sa1 = ps[ia]->shape[1];
    F.sa2 = ps[ia]->shape[2];
    F.sb1 = ps[ib]->shape[1];
    F.sb2 = ps[ib]->shape[2];
    const bool rows = !E.dense && static_cast<size_t>(E.pml_order[q]) < E.cpsi_row.size();
    F.ta = rows ? E.cpsi_row[E.pml_order[q]][ia] : -1;
    F.tb = rows ? E.cpsi_row[E.pml_order[q]][ib] : -1;
  }
}

template <typename S>
Fields<S> make_fields(const Engine& E) {
  Fields<S> f;
  for (int c = 0; c < 3; ++c) {
    f.e[c] = static_cast<S*>(E.fld[c]);
    f.h[c] = static_cast<S*>(E.fld[3 + c]);
  }
  return f;
}

template <typename C>
Scales<C> make_scales(const Engine& E) {
  Scales<C> s;
  const double sh =
      (E.plan.precision == XF_BF16_STORE || E.plan.precision == XF_BF16_COMP ||
       E.plan.precision == XF_FP16_STORE) ? 376.730313668 : 1.0;
  s.inv_e = 1;
  s.s_e = 1;
  s.inv_h = static_cast<C>(1.0 / sh);
  s.s_h = static_cast<C>(sh);
  return s;
}

Dims make_dims(const Engine& E) {
  Dims d;
  d.nx = static_cast<int>(E.nx);
  d.ny = static_cast<int>(E.ny);
  d.nz = static_cast<int>(E.nz);
  d.sx = E.sx;
  d.sy = E.sy;
  return d;
}
}  // namespace

// The watchdog's check of every stored field value after step `step`.
template <typename S>
int Engine::launch_watch(int64_t step, hipStream_t st) {
  int cus = 0;
  XF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const int64_t want = (total + 255) / 256;
  const unsigned blocks = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(want, 8 * cus)));
  hipLaunchKernelGGL((k_watch<S>), dim3(blocks), dim3(256), 0, st, make_fields<S>(*this), total,
                     static_cast<unsigned>(std::min<int64_t>(step, 0xfffffffe)), wd_dev);
  XF_HIP(hipGetLastError());
  return XF_OK;
}

// The tile lists of XF_TILE_PML_STREAMS: a tile is a CPML tile when its
// (k, j) lanes and its x-chunk meet the corrected box of a face the pass's
// Yee kernel carries (the same test as pml_mask, per tile).
int Engine::build_tile_lists() {
  const int64_t ntile = static_cast<int64_t>(nkb) * njb * nch_eff;
  if (ntile > INT_MAX) return fail(XF_ERR_UNSUPPORTED, "too many tiles for the tile lists");
  const int64_t rws = rows();
  for (int pass = 0; pass < 2; ++pass) {
    PmlPass<double> P{};
    fill_pml_pass<double>(*this, P, pass == 1, split_z() ? 1 : 0);
    std::vector<unsigned char> mark(ntile, 0);
    auto lo_hi = [](int64_t a, int64_t b, int64_t w, int64_t n_tiles, int64_t& t0, int64_t& t1) {
      // tiles of width w meeting the closed range [a, b]
      t0 = std::max<int64_t>(0, a / w);
      t1 = std::min<int64_t>(n_tiles - 1, b / w);
    };
    for (int q = 0; q < P.n; ++q) {
      const PmlFace<double>& F = P.f[q];
      int64_t i_a, i_b, j_a, j_b, k_a, k_b;  // closed node ranges the face can touch
      if (F.axis == 0) {
        i_a = F.c0; i_b = F.c1 - 1; j_a = F.a0; j_b = F.a1; k_a = F.b0; k_b = F.b1;
      } else if (F.axis == 1) {
        j_a = F.c0; j_b = F.c1 - 1; k_a = F.a0; k_b = F.a1; i_a = F.b0; i_b = F.b1;
      } else {
        k_a = F.c0; k_b = F.c1 - 1; i_a = F.a0; i_b = F.a1; j_a = F.b0; j_b = F.b1;
      }
      if (i_b < i_a || j_b < j_a || k_b < k_a) continue;
      int64_t c0, c1, jb0, jb1, kb0, kb1;
      lo_hi(i_a, i_b, chunk, nch_eff, c0, c1);
      lo_hi(j_a, j_b, rws, njb, jb0, jb1);
      lo_hi(k_a, k_b, klanes, nkb, kb0, kb1);
      for (int64_t c = c0; c <= c1; ++c)
        for (int64_t jb = jb0; jb <= jb1; ++jb)
          for (int64_t kb = kb0; kb <= kb1; ++kb) mark[kb + nkb * (jb + njb * c)] = 1;
    }
    std::vector<int> list;
    list.reserve(ntile);
    for (int64_t t = 0; t < ntile; ++t)
      if (!mark[t]) list.push_back(static_cast<int>(t));
    n_int[pass] = static_cast<int>(list.size());
    for (int64_t t = 0; t < ntile; ++t)
      if (mark[t]) list.push_back(static_cast<int>(t));
    n_pml[pass] = static_cast<int>(list.size()) - n_int[pass];
    if (tlist_cap[pass] < ntile) {
      XF_TRY(alloc(&tlist[pass], static_cast<size_t>(ntile) * sizeof(int)));
      tlist_cap[pass] = ntile;
    }
    XF_HIP(hipMemcpy(tlist[pass], list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return XF_OK;
}

template <typename S>
YeePass<S> Engine::yee_pass(bool is_e) const {
  using C = typename Store<S>::C;
  YeePass<S> D{};
  D.f = make_fields<S>(*this);
  D.ids = is_e ? eid : hid;
  D.tab = static_cast<const C*>(is_e ? etab : htab);
  D.ntab = is_e ? ntab_e : ntab_h;
  D.eflag = is_e ? eflag : nullptr;
  const Dense9<C> dn = dense9<C>(is_e);
  for (int n = 0; n < 9; ++n) D.dense[n] = dn.a[n];
  D.d = make_dims(*this);
  D.sc = make_scales<C>(*this);
  D.nkb = nkb;
  D.njb = njb;
  D.chunk = chunk;
  D.nch = nch_eff;
  const bool tail = ztail();
  fill_pml_pass<C>(*this, D.pml, is_e, (split_z() || tail) ? 1 : 0);
  D.zt = PmlZTail<C>{};
  if (tail) {
    PmlPass<C> Z{};
    fill_pml_pass<C>(*this, Z, is_e, 2);
    D.zt.n = Z.n;
    for (int q = 0; q < Z.n; ++q) D.zt.f[q] = Z.f[q];
  }
  return D;
}

int64_t Engine::lb_dev_hist_index(int64_t i) const {
  // face blocks sit at 16 (n_a + 1)(n_b + 1) intervals in face order, with
  // faces outside the mask keeping their block (unused)
  const int64_t n3[3] = {nx, ny, nz};
  int64_t off = 0;
  for (int f = 0; f < 6; ++f) {
    const int c = f / 2, a = (c + 1) % 3, b = (c + 2) % 3;
    const int64_t na1 = n3[a] + 1, nb1 = n3[b] + 1, size = 16 * na1 * nb1;
    if (i < off + size) {
      if (c != 1) return i;
      const int64_t r = i - off, plane = r / (na1 * nb1), p = (r / nb1) % na1, q = r % nb1;
      return off + (plane * nb1 + q) * na1 + p;
    }
    off += size;
  }
  return i;
}

// The split applies to 3-D FP64/FP32 grids with one or two z faces whose
// corrected ranges do not overlap (each node then gets its z corrections
// from one thread) and fewer than 2^31 nodes per face box.  BF16 storage
// keeps them fused: the split would round the intermediate field.
// The tail applies to 3-D FP64/FP32 grids in table mode with one or two z
// faces in each pass.  BF16 storage keeps the z faces in the march: read back from storage,
// the intermediate field would be rounded once more.
bool Engine::ztail() const {
  if (!pml_ztail || dense || dim != 3 || (plan.precision != XF_FP64 && plan.precision != XF_FP32)) return false;
  PmlPass<double> P{};
  for (int pass = 0; pass < 2; ++pass) {
    fill_pml_pass<double>(*this, P, pass == 1, 2);
    if (P.n == 0 || P.n > 2) return false;
  }
  return true;
}

bool Engine::split_z() const {
  if (!pml_split || ztail() || dim != 3 || (plan.precision != XF_FP64 && plan.precision != XF_FP32))
    return false;
  PmlPass<double> P{};
  for (int pass = 0; pass < 2; ++pass) {
    fill_pml_pass<double>(*this, P, pass == 1, 2);
    if (P.n == 0 || P.n > 2) return false;
    if (P.n == 2 && P.f[0].c1 > P.f[1].c0 && P.f[1].c1 > P.f[0].c0) return false;
    for (int q = 0; q < P.n; ++q) {
      const int64_t ni = std::min<int64_t>(P.f[q].a1, nx) - P.f[q].a0 + 1;
      const int64_t nj = std::min<int64_t>(P.f[q].b1, ny) - P.f[q].b0 + 1;
      const int64_t nk = std::min<int64_t>(P.f[q].c1, nz + 1) - P.f[q].c0;
      if (ni > 0 && nj > 0 && nk > 0 && ni * nj * nk >= (int64_t(1) << 31)) return false;
    }
  }
  return true;
}

template <typename C>
PmlZBox pml_z_box(const Engine& E, const PmlPass<C>& P) {
  PmlZBox B{};
  B.n = P.n;
  B.start[0] = 0;
  for (int q = 0; q < P.n; ++q) {
    const PmlFace<C>& F = P.f[q];
    const int64_t ni = std::max<int64_t>(0, std::min<int64_t>(F.a1, E.nx) - F.a0 + 1);
    const int64_t nj = std::max<int64_t>(0, std::min<int64_t>(F.b1, E.ny) - F.b0 + 1);
    const int64_t nk = std::max<int64_t>(0, std::min<int64_t>(F.c1, E.nz + 1) - F.c0);
    B.i0[q] = F.a0;
    B.j0[q] = F.b0;
    B.k0[q] = F.c0;
    B.nj[q] = static_cast<int>(std::max<int64_t>(nj, 1));
    B.nk[q] = static_cast<int>(std::max<int64_t>(nk, 1));
    B.start[q + 1] = B.start[q] + ni * nj * nk;
  }
  return B;
}

// The table and ids a pass's CPML corrections read c_psi rows through.
template <typename C>
RowRef<C> row_ref(const Engine& E, bool is_e) {
  return RowRef<C>{static_cast<const C*>(is_e ? E.etab : E.htab), is_e ? E.eid : E.hid, E.id_bytes == 2 ? 1 : 0};
}

template <typename S>
int Engine::launch_pml_z(bool is_e, hipStream_t st) {
  using C = typename Store<S>::C;
  PmlPass<C> P{};
  fill_pml_pass<C>(*this, P, is_e, 2);
  const PmlZBox B = pml_z_box<C>(*this, P);
  const int64_t total = B.start[B.n];
  if (total == 0) return XF_OK;
  const dim3 grid((unsigned)((total + 255) / 256));
  if (is_e)
    hipLaunchKernelGGL((k_pml_z<S, true>), grid, dim3(256), 0, st, make_fields<S>(*this), P, B,
                       make_dims(*this), make_scales<C>(*this), row_ref<C>(*this, true));
  else
    hipLaunchKernelGGL((k_pml_z<S, false>), grid, dim3(256), 0, st, make_fields<S>(*this), P, B,
                       make_dims(*this), make_scales<C>(*this), row_ref<C>(*this, false));
  XF_HIP(hipGetLastError());
  return XF_OK;
}

namespace {
// What tfsf_inc reads, for a TfsfPass or a TfsfZTail.
template <template <typename> class Pass, typename C>
void fill_tfsf_inc(const Engine& E, Pass<C>& P, int pass) {
  for (int d = 0; d < 3; ++d) {
    P.proj_int[d] = static_cast<const C*>(E.proj_int[d]);
    P.proj_half[d] = static_cast<const C*>(E.proj_half[d]);
  }
  const xf_tfsf& T = *E.sc->tfsf;
  P.inc = static_cast<const C*>(pass == 1 ? E.h_inc : E.e_inc);
  P.n_aux = pass == 1 ? T.h_inc.shape[1] : T.e_inc.shape[1];
  P.g[0] = static_cast<int>(T.total_task.x.start);
  P.g[1] = static_cast<int>(T.total_task.y.start);
  P.g[2] = static_cast<int>(T.total_task.z.start);
  for (int d = 0; d < 3; ++d) P.tr[d] = static_cast<C>(pass == 1 ? T.transform_h[d] : T.transform_e[d]);
}

// tail: the inner nodes of the jobs with a tail axis go to the Yee tiles
// (fill_tfsf_ztail); their jobs then enumerate only the two end rows.
template <typename C>
void fill_tfsf_pass(const Engine& E, TfsfPass<C>& P, int pass, bool tail = false) {
  const auto& jobs = E.tfsf_jobs[pass];
  P.njob = static_cast<int>(jobs.size());
  P.is_e = pass == 1;
  int64_t start = 0;
  for (int q = 0; q < P.njob; ++q) {
    const TfsfJobHost& J = jobs[q];
    TfsfJob<C>& D = P.job[q];
    D.comp = J.comp;
    D.face = J.face;
    for (int d = 0; d < 3; ++d) {
      D.o[d] = static_cast<int>(J.o[d]);
      D.n[d] = static_cast<int>(J.n[d]);
    }
    D.start = start;
    D.coef = static_cast<C>(J.coef);
    D.edge_axis = tail ? J.tail_axis : -1;
    int64_t count = 1;
    for (int d = 0; d < 3; ++d) count *= d == D.edge_axis ? 2 : J.n[d];
    start += count;
  }
  P.total = start;
  fill_tfsf_inc(E, P, pass);
}

// The jobs fill_tfsf_pass(.., tail = true) leaves to the Yee tiles.
template <typename C>
void fill_tfsf_ztail(const Engine& E, TfsfZTail<C>& Z, int pass, int t, const int* tp) {
  Z = TfsfZTail<C>{};
  const bool is_e = pass == 1;
  for (const TfsfJobHost& J : E.tfsf_jobs[pass]) {
    if (J.tail_axis < 0) continue;
    TfsfZJob<C>& D = Z.z[Z.n++];
    const int e = J.tail_axis;
    D.comp = J.comp;
    D.dual = J.comp == 0 ? 1 : 0;  // face axis z: a = x, b = y
    D.k = static_cast<int>(J.o[2]);
    D.kinc = D.k + (J.face % 2 == 0 ? (is_e ? -1 : 1) : 0);
    D.i0 = static_cast<int>(J.o[0] + (e == 0 ? 1 : 0));
    D.i1 = static_cast<int>(J.o[0] + J.n[0] - (e == 0 ? 1 : 0));
    D.j0 = static_cast<int>(J.o[1] + (e == 1 ? 1 : 0));
    D.j1 = static_cast<int>(J.o[1] + J.n[1] - (e == 1 ? 1 : 0));
    D.coef = static_cast<C>(J.coef);
  }
  Z.t_arg = t;
  Z.t_dev = tp;
  fill_tfsf_inc(E, Z, pass);
}
}  // namespace

template <typename S>
int Engine::launch_h(int t, const int* tp) {
  using C = typename Store<S>::C;
  PmlPass<C> P{};
  fill_pml_pass<C>(*this, P, false);
  const auto f = make_fields<S>(*this);
  const auto scl = make_scales<C>(*this);
  const Dims d = make_dims(*this);
  const size_t tab_bytes = static_cast<size_t>(ntab_h) * 9 * sizeof(C);
  const bool lds = tab_bytes <= 32768;
  const size_t shm = lds ? tab_bytes : 0;
  const bool has_pml = P.n > 0;
  const C* tab = static_cast<const C*>(htab);
  if (dim == 3) {
    YeePass<S> D = yee_pass<S>(false);
    if (tfsf_ztail()) fill_tfsf_ztail<C>(*this, D.tz, 0, t, tp);
    const bool has_pml = D.pml.n > 0 || D.zt.n > 0;
    dim3 block(klanes, rows());
    dim3 grid((unsigned)(D.nkb * D.njb * D.nch));
#define XF_LH3(I, L, PM)                                                                     \
  do {                                                                                       \
    if (nt_streams)                                                                          \
      hipLaunchKernelGGL((k_yee_h3<S, I, L, PM, true>), grid, block, shm, stream, D, xcd);   \
    else                                                                                     \
      hipLaunchKernelGGL((k_yee_h3<S, I, L, PM, false>), grid, block, shm, stream, D, xcd);  \
  } while (0)
#define XF_LH3D(PM)                                                                                   \
  do {                                                                                                \
    if (nt_streams)                                                                                   \
      hipLaunchKernelGGL((k_yee_h3<S, uint8_t, false, PM, true, true>), grid, block, 0, stream, D, xcd);  \
    else                                                                                              \
      hipLaunchKernelGGL((k_yee_h3<S, uint8_t, false, PM, false, true>), grid, block, 0, stream, D, xcd); \
  } while (0)
    if (dense) {
      if (has_pml) XF_LH3D(true); else XF_LH3D(false);
    } else if (use_lists()) {
      // CPML tiles on stream2, the rest (no CPML code) on the engine's stream
      const int* li = static_cast<const int*>(tlist[0]);
      XF_HIP(hipEventRecord(ev_fork, stream));
      XF_HIP(hipStreamWaitEvent(stream2, ev_fork, 0));
#define XF_LHL(I, L, PM, ST, N, LIST)                                                                    \
  do {                                                                                                 \
    if (nt_streams)                                                                                    \
      hipLaunchKernelGGL((k_yee_h3_list<S, I, L, PM, true>), dim3(N), block, shm, ST, D, LIST);         \
    else                                                                                               \
      hipLaunchKernelGGL((k_yee_h3_list<S, I, L, PM, false>), dim3(N), block, shm, ST, D, LIST);        \
  } while (0)
#define XF_LHL2(I, L)                                                                                  \
  do {                                                                                                 \
    if (n_pml[0] > 0 && D.zt.n > 0) XF_LHL(I, L, 2, stream2, n_pml[0], li + n_int[0]);                 \
    else if (n_pml[0] > 0) XF_LHL(I, L, 1, stream2, n_pml[0], li + n_int[0]);                          \
    if (n_int[0] > 0) XF_LHL(I, L, false, stream, n_int[0], li);                                       \
  } while (0)
      if (id_bytes == 1) {
        if (lds) XF_LHL2(uint8_t, true); else XF_LHL2(uint8_t, false);
      } else {
        if (lds) XF_LHL2(uint16_t, true); else XF_LHL2(uint16_t, false);
      }
#undef XF_LHL2
#undef XF_LHL
      XF_HIP(hipEventRecord(ev_join, stream2));
      XF_HIP(hipStreamWaitEvent(stream, ev_join, 0));
    } else {
      // CPML level: 0 none, 1 faces in the march, 2 and z faces at the tile end
#define XF_LH3P(I, L)                       \
  do {                                      \
    if (!has_pml) XF_LH3(I, L, 0);          \
    else if (D.zt.n > 0) XF_LH3(I, L, 2);   \
    else XF_LH3(I, L, 1);                   \
  } while (0)
      if (id_bytes == 1) {
        if (lds) XF_LH3P(uint8_t, true); else XF_LH3P(uint8_t, false);
      } else {
        if (lds) XF_LH3P(uint16_t, true); else XF_LH3P(uint16_t, false);
      }
#undef XF_LH3P
    }
#undef XF_LH3
#undef XF_LH3D
    if (split_z()) XF_TRY(launch_pml_z<S>(false, stream));
  } else {
    dim3 block(64, 4);
    dim3 grid((unsigned)((ny + 1 + 63) / 64), (unsigned)((nx + 1 + 3) / 4));
    TfsfPass<C> tf{};
    if (tfsf_fused()) fill_tfsf_pass<C>(*this, tf, 0);
#define XF_LH2(I, L, PM)                                                                              \
  hipLaunchKernelGGL((k_yee_h2<S, I, L, PM>), grid, block, shm, stream, f, static_cast<const I*>(hid), \
                     tab, ntab_h, P, d, scl, dense9<C>(false), tf, t, tp)
    if (id_bytes == 1) {
      if (lds) { if (has_pml) XF_LH2(uint8_t, true, true); else XF_LH2(uint8_t, true, false); }
      else { if (has_pml) XF_LH2(uint8_t, false, true); else XF_LH2(uint8_t, false, false); }
    } else {
      if (lds) { if (has_pml) XF_LH2(uint16_t, true, true); else XF_LH2(uint16_t, true, false); }
      else { if (has_pml) XF_LH2(uint16_t, false, true); else XF_LH2(uint16_t, false, false); }
    }
#undef XF_LH2
  }
  XF_HIP(hipGetLastError());
  return XF_OK;
}

template <typename S>
int Engine::launch_e(int t, const int* tp) {
  using C = typename Store<S>::C;
  PmlPass<C> P{};
  fill_pml_pass<C>(*this, P, true);
  const auto f = make_fields<S>(*this);
  const auto scl = make_scales<C>(*this);
  const Dims d = make_dims(*this);
  const size_t tab_bytes = static_cast<size_t>(ntab_e) * 9 * sizeof(C);
  const bool lds = tab_bytes <= 32768;
  const size_t shm = lds ? tab_bytes : 0;
  const bool has_pml = P.n > 0;
  const C* tab = static_cast<const C*>(etab);
  if (dim == 3) {
    YeePass<S> D = yee_pass<S>(true);
    if (tfsf_ztail()) fill_tfsf_ztail<C>(*this, D.tz, 1, t, tp);
    const bool has_pml = D.pml.n > 0 || D.zt.n > 0;
    dim3 block(klanes, rows());
    dim3 grid((unsigned)(D.nkb * D.njb * D.nch));
#define XF_LE3(I, L, PM, AD)                                                                     \
  do {                                                                                           \
    if (nt_streams)                                                                              \
      hipLaunchKernelGGL((k_yee_e3<S, I, L, PM, AD, true>), grid, block, shm, stream, D, xcd);   \
    else                                                                                         \
      hipLaunchKernelGGL((k_yee_e3<S, I, L, PM, AD, false>), grid, block, shm, stream, D, xcd);  \
  } while (0)
#define XF_LE3A(I, L, PM) \
  do { if (has_ade) XF_LE3(I, L, PM, true); else XF_LE3(I, L, PM, false); } while (0)
#define XF_LE3D(PM, AD)                                                                                   \
  do {                                                                                                    \
    if (nt_streams)                                                                                       \
      hipLaunchKernelGGL((k_yee_e3<S, uint8_t, false, PM, AD, true, true>), grid, block, 0, stream, D, xcd);  \
    else                                                                                                  \
      hipLaunchKernelGGL((k_yee_e3<S, uint8_t, false, PM, AD, false, true>), grid, block, 0, stream, D, xcd); \
  } while (0)
    if (dense) {
      if (has_pml) { if (has_ade) XF_LE3D(true, true); else XF_LE3D(true, false); }
      else { if (has_ade) XF_LE3D(false, true); else XF_LE3D(false, false); }
    } else if (use_lists()) {
      const int* li = static_cast<const int*>(tlist[1]);
      XF_HIP(hipEventRecord(ev_fork, stream));
      XF_HIP(hipStreamWaitEvent(stream2, ev_fork, 0));
#define XF_LEL(I, L, PM, AD, ST, N, LIST)                                                                   \
  do {                                                                                                    \
    if (nt_streams)                                                                                       \
      hipLaunchKernelGGL((k_yee_e3_list<S, I, L, PM, AD, true>), dim3(N), block, shm, ST, D, LIST);        \
    else                                                                                                  \
      hipLaunchKernelGGL((k_yee_e3_list<S, I, L, PM, AD, false>), dim3(N), block, shm, ST, D, LIST);       \
  } while (0)
#define XF_LEL2(I, L, AD)                                                                                 \
  do {                                                                                                    \
    if (n_pml[1] > 0 && D.zt.n > 0) XF_LEL(I, L, 2, AD, stream2, n_pml[1], li + n_int[1]);                \
    else if (n_pml[1] > 0) XF_LEL(I, L, 1, AD, stream2, n_pml[1], li + n_int[1]);                         \
    if (n_int[1] > 0) XF_LEL(I, L, false, AD, stream, n_int[1], li);                                      \
  } while (0)
#define XF_LEL3(I, L) do { if (has_ade) XF_LEL2(I, L, true); else XF_LEL2(I, L, false); } while (0)
      if (id_bytes == 1) {
        if (lds) XF_LEL3(uint8_t, true); else XF_LEL3(uint8_t, false);
      } else {
        if (lds) XF_LEL3(uint16_t, true); else XF_LEL3(uint16_t, false);
      }
#undef XF_LEL3
#undef XF_LEL2
#undef XF_LEL
      XF_HIP(hipEventRecord(ev_join, stream2));
      XF_HIP(hipStreamWaitEvent(stream, ev_join, 0));
    } else {
#define XF_LE3P(I, L)                        \
  do {                                       \
    if (!has_pml) XF_LE3A(I, L, 0);          \
    else if (D.zt.n > 0) XF_LE3A(I, L, 2);   \
    else XF_LE3A(I, L, 1);                   \
  } while (0)
      if (id_bytes == 1) {
        if (lds) XF_LE3P(uint8_t, true); else XF_LE3P(uint8_t, false);
      } else {
        if (lds) XF_LE3P(uint16_t, true); else XF_LE3P(uint16_t, false);
      }
#undef XF_LE3P
    }
#undef XF_LE3A
#undef XF_LE3
#undef XF_LE3D
    XF_HIP(hipGetLastError());
    if (split_z()) XF_TRY(launch_pml_z<S>(true, stream));
    if (has_ade && ade.n > 0) {
      dim3 b1(256), g1((unsigned)((ade.n + 255) / 256));
      if (id_bytes == 1)
        hipLaunchKernelGGL((k_ade_e3<S, uint8_t>), g1, b1, 0, stream, f, static_cast<const uint8_t*>(eid), tab, ade, d, scl,
                           dense9<C>(true));
      else
        hipLaunchKernelGGL((k_ade_e3<S, uint16_t>), g1, b1, 0, stream, f, static_cast<const uint16_t*>(eid), tab, ade, d, scl,
                           dense9<C>(true));
    }
  } else {
    dim3 block(64, 4);
    dim3 grid((unsigned)((ny + 1 + 63) / 64), (unsigned)((nx + 1 + 3) / 4));
    TfsfPass<C> tf{};
    if (tfsf_fused()) fill_tfsf_pass<C>(*this, tf, 1);
#define XF_LE2(I, L, PM)                                                                              \
  hipLaunchKernelGGL((k_yee_e2<S, I, L, PM>), grid, block, shm, stream, f, static_cast<const I*>(eid), \
                     tab, ntab_e, P, d, scl, dense9<C>(true), tf, t, tp)
    if (id_bytes == 1) {
      if (lds) { if (has_pml) XF_LE2(uint8_t, true, true); else XF_LE2(uint8_t, true, false); }
      else { if (has_pml) XF_LE2(uint8_t, false, true); else XF_LE2(uint8_t, false, false); }
    } else {
      if (lds) { if (has_pml) XF_LE2(uint16_t, true, true); else XF_LE2(uint16_t, true, false); }
      else { if (has_pml) XF_LE2(uint16_t, false, true); else XF_LE2(uint16_t, false, false); }
    }
#undef XF_LE2
  }
  XF_HIP(hipGetLastError());
  return XF_OK;
}


template <typename S>
int Engine::launch_post_h(int t, const int* tp) {
  using C = typename Store<S>::C;
  if (sc->tfsf && !tfsf_fused()) {
    TfsfPass<C> P{};
    fill_tfsf_pass<C>(*this, P, 0, tfsf_ztail());
    if (P.total > 0)
      hipLaunchKernelGGL((k_tfsf<S>), dim3((unsigned)((P.total + 255) / 256)), dim3(256), 0, stream,
                         make_fields<S>(*this), P, t, tp, make_scales<C>(*this), make_dims(*this), dim == 2 ? 1 : 0);
  }
  XF_HIP(hipGetLastError());
  return XF_OK;
}

template <typename S>
int Engine::launch_post_e(int t, const int* tp) {
  using C = typename Store<S>::C;
  if (sc->tfsf && !tfsf_fused()) {
    TfsfPass<C> P{};
    fill_tfsf_pass<C>(*this, P, 1, tfsf_ztail());
    if (P.total > 0)
      hipLaunchKernelGGL((k_tfsf<S>), dim3((unsigned)((P.total + 255) / 256)), dim3(256), 0, stream,
                         make_fields<S>(*this), P, t, tp, make_scales<C>(*this), make_dims(*this), dim == 2 ? 1 : 0);
  }
  // the movies share the first NF2FF box's launch (both only read the
  // fields as the step leaves them)
  const bool nf_movies = !movies.empty() && !nf.empty() && nf[0].pass.total > 0;
  XF_TRY(launch_additions<S>(t, tp, stream, !nf_movies));
  for (size_t q = 0; q < nf.size(); ++q) {
    const NfPass& P = nf[q].pass;
    if (P.total == 0) continue;
    if (q == 0 && nf_movies) {
      int64_t total_box = 0;
      for (const MovieDev& M : movies) total_box += M.box;
      hipLaunchKernelGGL((k_nf2ff_movies<S>), dim3((unsigned)((P.total + total_box + 255) / 256)), dim3(256), 0,
                         stream, make_fields<S>(*this), P, static_cast<const MovieDev*>(movie_items), movie_start,
                         static_cast<int>(movies.size()), make_dims(*this), dim == 2 ? 1 : 0,
                         make_scales<C>(*this), t, tp);
    } else {
      hipLaunchKernelGGL((k_nf2ff<S>), dim3((unsigned)((P.total + 255) / 256)), dim3(256), 0, stream,
                         make_fields<S>(*this), P, t, tp, make_dims(*this), dim == 2 ? 1 : 0,
                         make_scales<C>(*this));
    }
  }
  XF_TRY(launch_td<S>(t, tp, stream));
  XF_HIP(hipGetLastError());
  return XF_OK;
}

template <typename S>
int Engine::launch_additions(int t, const int* tp, hipStream_t st, bool with_movies) {
  using C = typename Store<S>::C;
  if (pts.n > 0)
    hipLaunchKernelGGL((k_point_sources<S>), dim3(1), dim3(64), 0, st, make_fields<S>(*this), pts,
                       make_scales<C>(*this), t, tp);
  if (has_lb && lbd.total > 0) {
    const dim3 b(256);
    if (lbd.mode == XF_LB_APPLY) {
      // the network on the matrix cores (with the history roll of its
      // columns), then the edge columns: Mur update, then their roll
      hipLaunchKernelGGL((k_lb_mfma<S>), dim3((unsigned)((lbd.total + kLbMfmaCols - 1) / kLbMfmaCols)), b, 0,
                         st, make_fields<S>(*this), lbd, make_dims(*this), make_scales<C>(*this), t, tp);
      if (lbd.n_edge > 0)
        for (int ph = 1; ph < 3; ++ph)
          hipLaunchKernelGGL((k_lb_edges<S>), dim3((unsigned)((lbd.n_edge + 255) / 256)), b, 0, st,
                             make_fields<S>(*this), lbd, make_dims(*this), make_scales<C>(*this), ph, t, tp);
    } else {
      // RECORD: phase 0 records, phase 2 rolls the history
      const dim3 g((unsigned)((lbd.total + 255) / 256));
      for (int ph : {0, 2})
        hipLaunchKernelGGL((k_lb<S>), g, b, 0, st, make_fields<S>(*this), lbd, make_dims(*this),
                           make_scales<C>(*this), ph, t, tp);
    }
  }
  return with_movies ? launch_movies<S>(t, tp, st) : XF_OK;
}

template <typename S>
int Engine::launch_movies(int t, const int* tp, hipStream_t st) {
  using C = typename Store<S>::C;
  if (movies.empty()) return XF_OK;
  int64_t total_box = 0;
  for (const MovieDev& M : movies) total_box += M.box;
  hipLaunchKernelGGL((k_movies<S>), dim3((unsigned)((total_box + 255) / 256)), dim3(256), 0, st,
                     make_fields<S>(*this), static_cast<const MovieDev*>(movie_items), movie_start,
                     static_cast<int>(movies.size()), make_dims(*this), dim == 2 ? 1 : 0, make_scales<C>(*this), t,
                     tp);
  XF_HIP(hipGetLastError());
  return XF_OK;
}

template <typename S>
int Engine::launch_td(int t, const int* tp, hipStream_t st) {
  using C = typename Store<S>::C;
  for (const NtPass& P : nftd) {
    if (P.nblk == 0) continue;
    const size_t lds = 256 * sizeof(NtRec) + 16 * static_cast<size_t>(P.maxspan) * sizeof(double);
    hipLaunchKernelGGL((k_nftd_tile<S>), dim3(P.nblk), dim3(256), lds, st, make_fields<S>(*this), P,
                       make_dims(*this), make_scales<C>(*this), t, tp);
    hipLaunchKernelGGL(k_nftd_reduce, dim3(P.nrblk), dim3(256), 0, st, P, t, tp);
  }
  XF_HIP(hipGetLastError());
  return XF_OK;
}

template <typename S>
int Engine::launch_step(int t, bool use_tdev, bool timing_events, hipEvent_t* ev) {
  const int* tp = use_tdev ? t_dev : nullptr;
  if (timing_events) XF_HIP(hipEventRecord(ev[0], stream));
  XF_TRY(launch_h<S>(t, tp));
  if (timing_events) XF_HIP(hipEventRecord(ev[1], stream));
  XF_TRY(launch_post_h<S>(t, tp));
  if (timing_events) XF_HIP(hipEventRecord(ev[2], stream));
  XF_TRY(launch_e<S>(t, tp));
  if (timing_events) XF_HIP(hipEventRecord(ev[3], stream));
  XF_TRY(launch_post_e<S>(t, tp));
  if (timing_events) XF_HIP(hipEventRecord(ev[4], stream));
  if (use_tdev) {
    hipLaunchKernelGGL(k_next_step, dim3(1), dim3(1), 0, stream, t_dev);
    XF_HIP(hipGetLastError());
  }
  return XF_OK;
}

int Engine::run(int64_t n) {
  if (!uploaded) return fail(XF_ERR_STATE, "run before upload");
  if (n < 0) return fail(XF_ERR_INVALID, "negative step count");
  if (n == 0) return XF_OK;
  if ((sc->tfsf || !nf.empty() || !nftd.empty() || pts.n > 0) && cur_step + n > nt)
    return fail(XF_ERR_INVALID, "run past the end of the time axis the TFSF/NF2FF/source tables were built for");
  XF_HIP(hipSetDevice(device));
  const bool timing_on = plan.timing != 0;
  const bool graph_on = plan.use_graph != 0;
  HipEvents ev, runs;
  XF_TRY(runs.create(2));
  hipEvent_t run0 = runs[0], run1 = runs[1];
  if (timing_on) XF_TRY(ev.create(5 * n));
  const int64_t wd = plan.watchdog > 0 ? plan.watchdog : 0;
  auto watch = [&](int64_t done) -> int {  // after `done` steps of this run
    switch (plan.precision) {
      case XF_FP64: return launch_watch<double>(cur_step + done, stream);
      case XF_FP32: return launch_watch<float>(cur_step + done, stream);
      case XF_BF16_COMP: return launch_watch<Bf16c>(cur_step + done, stream);
      case XF_FP16_STORE: return launch_watch<__half>(cur_step + done, stream);
      default: return launch_watch<__hip_bfloat16>(cur_step + done, stream);
    }
  };
  if (wd) hipLaunchKernelGGL(k_watch_clear, dim3(1), dim3(64), 0, stream, wd_dev);
  XF_HIP(hipEventRecord(run0, stream));
  int status = XF_OK;
  auto step = [&](int64_t s) -> int {
    const int t = static_cast<int>(cur_step + s);
    hipEvent_t* e = timing_on ? &ev[5 * s] : nullptr;
    switch (plan.precision) {
      case XF_FP64:
        return launch_step<double>(t, false, timing_on, e);
      case XF_FP32:
        return launch_step<float>(t, false, timing_on, e);
      case XF_BF16_COMP:
        return launch_step<Bf16c>(t, false, timing_on, e);
      case XF_FP16_STORE:
        return launch_step<__half>(t, false, timing_on, e);
      default:
        return launch_step<__hip_bfloat16>(t, false, timing_on, e);
    }
  };
  if (graph_on && !timing_on) {
    if (!graph_ready) {
      XF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      int st;
      switch (plan.precision) {
        case XF_FP64:
          st = launch_step<double>(0, true, false, nullptr);
          break;
        case XF_FP32:
          st = launch_step<float>(0, true, false, nullptr);
          break;
        case XF_BF16_COMP:
          st = launch_step<Bf16c>(0, true, false, nullptr);
          break;
        case XF_FP16_STORE:
          st = launch_step<__half>(0, true, false, nullptr);
          break;
        default:
          st = launch_step<__hip_bfloat16>(0, true, false, nullptr);
      }
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
    for (int64_t s = 0; s < n && status == XF_OK; ++s) {
      status = step(s);
      if (status == XF_OK && wd && ((cur_step + s + 1) % wd == 0 || s + 1 == n)) status = watch(s + 1);
    }
  }
  XF_HIP(hipEventRecord(run1, stream));
  XF_HIP(hipEventSynchronize(run1));
  if (status != XF_OK) return status;
  XF_HIP(hipGetLastError());
  unsigned first_bad = 0xffffffffu;
  if (wd) XF_HIP(hipMemcpy(&first_bad, wd_dev, sizeof(unsigned), hipMemcpyDeviceToHost));
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
  timing.bytes_per_cell_step = bytes_per_cell_step();
  cur_step += n;
  if (has_lb && lbd.mode == XF_LB_RECORD) lb_rec_count = lb_recorded_after(cur_step);
  if (first_bad != 0xffffffffu)
    return fail(XF_ERR_NUMERIC, "watchdog: non-finite field value found after step " + std::to_string(first_bad) +
                                    " (the run went on to step " + std::to_string(cur_step) + ")");
  return XF_OK;
}

}  // namespace xf

struct xf_handle {
  xf::Engine eng;
};

extern "C" {

const char* xf_last_error(void) { return xf::g_err.c_str(); }

int xf_device_count(int* count) {
  return xf::guarded([&]() -> int {
    if (!count) return xf::fail(XF_ERR_INVALID, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return XF_OK;
  });
}

int xf_create(xf_scene* scene, const xf_plan* plan, xf_handle** out) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!scene || !out) return xf::fail(XF_ERR_INVALID, "null scene or out");
    *out = nullptr;
    std::unique_ptr<xf_handle> h(new (std::nothrow) xf_handle);
    if (!h) return xf::fail(XF_ERR_NOMEM, "out of host memory");
    h->eng.sc = scene;
    if (plan) h->eng.plan = *plan;
    if (h->eng.plan.precision < XF_FP64 || h->eng.plan.precision > XF_FP16_STORE)
      return xf::fail(XF_ERR_INVALID, "unknown precision");
    h->eng.device = h->eng.plan.device;
    int st = h->eng.validate();
    if (st == XF_OK) st = h->eng.check_overlaps();
    if (st != XF_OK) return st;
    *out = h.release();
    return XF_OK;
  });
}

int xf_upload(xf_handle* h) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null handle");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return xf::fail(XF_ERR_HIP, "no HIP device");
    if (h->eng.device < 0 || h->eng.device >= n) return xf::fail(XF_ERR_INVALID, "bad device ordinal");
    // a failed upload (status or exception) leaves nothing on the device
    const int st = xf::guarded([&] { return h->eng.upload(); });
    if (st != XF_OK) h->eng.release();
    return st;
  });
}

int xf_run(xf_handle* h, int64_t n_steps) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null handle");
    return h->eng.run(n_steps);
  });
}

int xf_download(xf_handle* h) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null handle");
    return h->eng.download();
  });
}

int xf_destroy(xf_handle* h) {
  return xf::guarded([&]() -> int {
    delete h;
    return XF_OK;
  });
}

int xf_current_step(const xf_handle* h, int64_t* step) {
  return xf::guarded([&]() -> int {
    if (!h || !step) return xf::fail(XF_ERR_INVALID, "null argument");
    *step = h->eng.cur_step;
    return XF_OK;
  });
}

int xf_get_timing(const xf_handle* h, xf_timing* out) {
  return xf::guarded([&]() -> int {
    if (!h || !out) return xf::fail(XF_ERR_INVALID, "null argument");
    *out = h->eng.timing;
    out->bytes_per_cell_step = h->eng.bytes_per_cell_step();
    return XF_OK;
  });
}

int xf_reset_timing(xf_handle* h) {
  return xf::guarded([&]() -> int {
    if (!h) return xf::fail(XF_ERR_INVALID, "null handle");
    h->eng.timing = xf_timing{};
    return XF_OK;
  });
}

int xf_stream_bench(int64_t bytes, int32_t reps, int32_t device, int32_t variant, int32_t blocks,
                    double* gbps) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!gbps || bytes < 16 || reps < 1) return xf::fail(XF_ERR_INVALID, "bad stream bench arguments");
    if (variant < 0 || variant > 6) return xf::fail(XF_ERR_INVALID, "bad stream bench variant");
    if (hipSetDevice(device) != hipSuccess) return xf::fail(XF_ERR_HIP, "hipSetDevice failed");
    const int64_t n = bytes / 16;
    xf::xf_d2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return xf::fail(XF_ERR_NOMEM, "stream bench allocation failed");
    if (hipMalloc(&b, n * 16) != hipSuccess) {
      (void)hipFree(a);
      return xf::fail(XF_ERR_NOMEM, "stream bench allocation failed");
    }
    hipEvent_t e0, e1;
    (void)hipMemset(a, 0, n * 16);
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const dim3 block(256), grid(blocks > 0 ? blocks : 256 * 8 * 4);
    auto launch = [&]() {
      switch (variant) {
        case 0: hipLaunchKernelGGL((xf::xf_stream_copy_kernel<1, false>), grid, block, 0, 0, a, b, n); break;
        case 1: hipLaunchKernelGGL((xf::xf_stream_copy_kernel<4, false>), grid, block, 0, 0, a, b, n); break;
        case 2: hipLaunchKernelGGL((xf::xf_stream_copy_kernel<4, true>), grid, block, 0, 0, a, b, n); break;
        case 3: hipLaunchKernelGGL((xf::xf_stream_copy_kernel<8, false>), grid, block, 0, 0, a, b, n); break;
        case 4: hipLaunchKernelGGL((xf::xf_stream_copy_kernel<8, true>), grid, block, 0, 0, a, b, n); break;
        case 5: hipLaunchKernelGGL((xf::xf_stream_copy1_kernel<false>), dim3((unsigned)((n + 255) / 256)), block, 0, 0, a, b, n); break;
        default: hipLaunchKernelGGL((xf::xf_stream_copy1_kernel<true>), dim3((unsigned)((n + 255) / 256)), block, 0, 0, a, b, n); break;
      }
    };
    launch();  // warm-up
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return xf::fail(XF_ERR_HIP, hipGetErrorString(err));
    *gbps = 2.0 * static_cast<double>(n) * 16.0 * reps / (ms * 1e-3) / 1e9;
    return XF_OK;
  });
}

int xf_set_tiling(xf_handle* h, int32_t tile_y, int32_t chunks_x, int32_t flags) {
  return xf::guarded([&]() -> int {
    if (!h) return xf::fail(XF_ERR_INVALID, "null handle");
    // the Yee kernels are compiled for at most 256 threads per block (64 x 4)
    if (tile_y < 0 || tile_y > 4 || chunks_x < 0) return xf::fail(XF_ERR_INVALID, "bad tiling (tile_y 0..4)");
    h->eng.set_tiling(tile_y, chunks_x, flags);
    return XF_OK;
  });
}

int xf_set_timing(xf_handle* h, int32_t on) {
  return xf::guarded([&]() -> int {
    if (!h) return xf::fail(XF_ERR_INVALID, "null handle");
    h->eng.plan.timing = on ? 1 : 0;
    return XF_OK;
  });
}

int xf_describe(const xf_handle* h, char* buf, size_t len) {
  return xf::guarded([&]() -> int {
    if (!h || !buf || len == 0) return xf::fail(XF_ERR_INVALID, "null argument");
    std::snprintf(buf, len, "%s", h->eng.layout_desc.c_str());
    return XF_OK;
  });
}

}  // extern "C"

#include "xf_batch.inl"
#include "xf_lb_train.inl"
