// capi.cpp -- C ABI over the host scene model and the SimulationHD drop-in
// (include/xfdtd_host.h).
#include <cmath>
#include <cstring>
#include <filesystem>
#include <memory>
#include <string>

#include "../../../include/xfdtd_cuda_amd/simulation_hd.h"
#include "../../../include/xfdtd_cuda_amd/xfdtd_core.h"
#include "../../../include/xfdtd_host.h"

struct xfh_sim {
  std::unique_ptr<xfdtd::Simulation> sim;
};

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return XF_ERR_INVALID;
}

template <class F>
int guard(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return XF_ERR_INVALID;
  } catch (...) {
    g_err = "unknown exception";
    return XF_ERR_INVALID;
  }
}

std::unique_ptr<xfdtd::Material> makeMaterial(const xfh_material* m) {
  using xfdtd::Material;
  if (m == nullptr) return Material::createAir();
  std::vector<double> p0, p1, p2;
  const double* p = m->p;
  if (m->kind >= 2 && m->kind <= 5) {
    // eps_inf and the poles are read from p: it must hold them
    if (p == nullptr) throw xfdtd::XFDTDException("dispersive material: null parameter array");
    if (m->n_pole < 1 || m->n_pole > 64) throw xfdtd::XFDTDException("dispersive material: n_pole must be 1..64");
  }
  switch (m->kind) {
    case 0:
      return std::make_unique<Material>(
          "material", xfdtd::ElectroMagneticProperty{m->eps_r, m->mu_r, m->sigma_e, m->sigma_m});
    case 1:
      return Material::createPec();
    case 2:
      for (int i = 0; i < m->n_pole; ++i) {
        p0.push_back(p[1 + 2 * i]);
        p1.push_back(p[2 + 2 * i]);
      }
      return Material::makeDrudeMedium("drude", p[0], p0, p1);
    case 3:
      for (int i = 0; i < m->n_pole; ++i) {
        p0.push_back(p[1 + 2 * i]);
        p1.push_back(p[2 + 2 * i]);
      }
      return Material::makeDebyeMedium("debye", p[0], p0, p1);
    case 4:
      for (int i = 0; i < m->n_pole; ++i) {
        p0.push_back(p[1 + 3 * i]);
        p1.push_back(p[2 + 3 * i]);
        p2.push_back(p[3 + 3 * i]);
      }
      return Material::makeLorentzMedium("lorentz", p[0], p0, p1, p2);
    case 5: {
      std::vector<xfdtd::MLorentzPole> poles;
      for (int i = 0; i < m->n_pole; ++i)
        poles.push_back({p[1 + 5 * i], p[2 + 5 * i], p[3 + 5 * i], p[4 + 5 * i], p[5 + 5 * i]});
      return Material::makeMLorentz("mlorentz", p[0], poles);
    }
    default:
      throw xfdtd::XFDTDException("unknown material kind");
  }
}

std::unique_ptr<xfdtd::Waveform> makeWave(int kind, double tau, double t0, double f) {
  switch (kind) {
    case 0:
      return xfdtd::Waveform::gaussian(tau, t0);
    case 1:
      return xfdtd::Waveform::cosineModulatedGaussian(tau, t0, f);
    case 2:
      return xfdtd::Waveform::sine(f);
    default:
      throw xfdtd::XFDTDException("unknown waveform kind");
  }
}

const xf_array* pmlMember(const xf_pml& p, const std::string& m) {
  if (m == "coeff_a_e") return &p.coeff_a_e;
  if (m == "coeff_b_e") return &p.coeff_b_e;
  if (m == "coeff_a_h") return &p.coeff_a_h;
  if (m == "coeff_b_h") return &p.coeff_b_h;
  if (m == "c_ea_psi_hb") return &p.c_ea_psi_hb;
  if (m == "c_eb_psi_ha") return &p.c_eb_psi_ha;
  if (m == "c_ha_psi_eb") return &p.c_ha_psi_eb;
  if (m == "c_hb_psi_ea") return &p.c_hb_psi_ea;
  if (m == "ea_psi_hb") return &p.ea_psi_hb;
  if (m == "eb_psi_ha") return &p.eb_psi_ha;
  if (m == "ha_psi_eb") return &p.ha_psi_eb;
  if (m == "hb_psi_ea") return &p.hb_psi_ea;
  return nullptr;
}

const xf_array* tfsfMember(const xf_tfsf& t, const std::string& m) {
  if (m == "proj_x_int") return &t.proj_x_int;
  if (m == "proj_y_int") return &t.proj_y_int;
  if (m == "proj_z_int") return &t.proj_z_int;
  if (m == "proj_x_half") return &t.proj_x_half;
  if (m == "proj_y_half") return &t.proj_y_half;
  if (m == "proj_z_half") return &t.proj_z_half;
  if (m == "e_inc") return &t.e_inc;
  if (m == "h_inc") return &t.h_inc;
  return nullptr;
}

const xf_array* adeMember(const xf_ade& a, const std::string& m) {
  if (m == "coeff_j_j") return &a.coeff_j_j;
  if (m == "coeff_j_j_p") return &a.coeff_j_j_p;
  if (m == "coeff_j_e_n") return &a.coeff_j_e_n;
  if (m == "coeff_j_e") return &a.coeff_j_e;
  if (m == "coeff_j_e_p") return &a.coeff_j_e_p;
  if (m == "coeff_j_sum_j") return &a.coeff_j_sum_j;
  if (m == "coeff_j_sum_j_p") return &a.coeff_j_sum_j_p;
  if (m == "coeff_e_j_sum") return &a.coeff_e_j_sum;
  if (m == "coeff_e_e_p") return &a.coeff_e_e_p;
  if (m == "ex_prev") return &a.ex_prev;
  if (m == "ey_prev") return &a.ey_prev;
  if (m == "ez_prev") return &a.ez_prev;
  if (m == "jx") return &a.jx;
  if (m == "jy") return &a.jy;
  if (m == "jz") return &a.jz;
  if (m == "jx_prev") return &a.jx_prev;
  if (m == "jy_prev") return &a.jy_prev;
  if (m == "jz_prev") return &a.jz_prev;
  return nullptr;
}

std::vector<std::string> split(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == '.') {
      out.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  out.push_back(cur);
  return out;
}

int faceIndex(const std::string& f) {
  static const char* names[6] = {"XN", "XP", "YN", "YP", "ZN", "ZP"};
  for (int i = 0; i < 6; ++i)
    if (f == names[i]) return i;
  return -1;
}

}  // namespace

extern "C" {

const char* xfh_last_error(void) { return g_err.c_str(); }

int xfh_sim_create(double dx, double dy, double dz, double cfl, xfh_sim** out) {
  return guard([&] {
    if (!out) return fail("out is null");
    auto h = new xfh_sim;
    h->sim = std::make_unique<xfdtd::Simulation>(dx, dy, dz, cfl);
    *out = h;
    return 0;
  });
}

int xfh_sim_destroy(xfh_sim* s) {
  delete s;
  return 0;
}

int xfh_sim_add_cube(xfh_sim* s, const char* name, const double o[3], const double sz[3],
                     const xfh_material* m) {
  return guard([&] {
    if (!s) return fail("null sim");
    s->sim->addObject(std::make_shared<xfdtd::Object>(
        name ? name : "cube",
        std::make_unique<xfdtd::Cube>(xfdtd::Vector{o[0], o[1], o[2]},
                                      xfdtd::Vector{sz[0], sz[1], sz[2]}),
        makeMaterial(m)));
    return 0;
  });
}

int xfh_sim_add_sphere(xfh_sim* s, const char* name, const double c[3], double r,
                       const xfh_material* m) {
  return guard([&] {
    if (!s) return fail("null sim");
    s->sim->addObject(std::make_shared<xfdtd::Object>(
        name ? name : "sphere",
        std::make_unique<xfdtd::Sphere>(xfdtd::Vector{c[0], c[1], c[2]}, r), makeMaterial(m)));
    return 0;
  });
}

int xfh_sim_add_cylinder(xfh_sim* s, const char* name, const double c[3], double r,
                         double height, const xfh_material* m) {
  return guard([&] {
    if (!s) return fail("null sim");
    s->sim->addObject(std::make_shared<xfdtd::Object>(
        name ? name : "cylinder",
        std::make_unique<xfdtd::Cylinder>(xfdtd::Vector{c[0], c[1], c[2]}, r, height),
        makeMaterial(m)));
    return 0;
  });
}

int xfh_sim_add_pml(xfh_sim* s, int32_t direction, int64_t thickness, int64_t order,
                    double sigma_ratio, double alpha_min, double alpha_max, double kappa_max) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (direction < 0 || direction > 5) return fail("bad PML direction");
    if (thickness <= 0) return fail("bad PML thickness");
    s->sim->addBoundary(std::make_shared<xfdtd::PML>(
        static_cast<xfdtd::Index>(thickness), static_cast<xfdtd::Axis::Direction>(direction),
        static_cast<xfdtd::Index>(order), sigma_ratio, alpha_min, alpha_max, kappa_max));
    return 0;
  });
}

int xfh_sim_add_tfsf3d(xfh_sim* s, const int64_t st[3], double theta, double phi, double psi,
                       int32_t wave_kind, double tau, double t0, double freq) {
  return guard([&] {
    if (!s) return fail("null sim");
    s->sim->addWaveformSource(std::make_shared<xfdtd::TFSF3D>(
        st[0], st[1], st[2], theta, phi, psi, makeWave(wave_kind, tau, t0, freq)));
    return 0;
  });
}

int xfh_sim_add_tfsf2d(xfh_sim* s, const int64_t st[2], double phi, int32_t wave_kind,
                       double tau, double t0, double freq) {
  return guard([&] {
    if (!s) return fail("null sim");
    s->sim->addWaveformSource(std::make_shared<xfdtd::TFSF2D>(
        st[0], st[1], phi, makeWave(wave_kind, tau, t0, freq)));
    return 0;
  });
}

int xfh_sim_add_nf2ff_fd(xfh_sim* s, const int64_t st[3], const double* freqs, int32_t n) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (!freqs || n <= 0) return fail("no frequency");
    s->sim->addNF2FF(std::make_shared<xfdtd::NFFFTFrequencyDomain>(
        st[0], st[1], st[2], std::vector<double>(freqs, freqs + n)));
    return 0;
  });
}

int xfh_sim_add_nf2ff_td(xfh_sim* s, const int64_t st[3], double theta, double phi) {
  return guard([&] {
    if (!s || !st) return fail("null argument");
    if (!std::isfinite(theta) || !std::isfinite(phi)) return fail("bad observation direction");
    s->sim->addNF2FF(std::make_shared<xfdtd::NFFFTTimeDomain>(st[0], st[1], st[2], theta, phi));
    return 0;
  });
}

int xfh_sim_add_point_source(xfh_sim* s, int32_t comp, const int64_t idx[3],
                             int32_t wave_kind, double tau, double t0, double freq,
                             double amplitude) {
  return guard([&] {
    if (!s || !idx) return fail("null argument");
    if (comp < 0 || comp > 2) return fail("bad point source component");
    if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0) return fail("negative point source index");
    s->sim->addWaveformSource(std::make_shared<xfdtd::PointSource>(
        static_cast<xfdtd::Axis::XYZ>(comp), idx[0], idx[1], idx[2],
        makeWave(wave_kind, tau, t0, freq), amplitude));
    return 0;
  });
}

int xfh_sim_add_learned_boundary(xfh_sim* s, int32_t mode, int32_t faces, int64_t size,
                                 const uint16_t* weights, int32_t n_weights, int64_t every,
                                 int64_t slots) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (size < 0 || every < 0 || slots < 0) return fail("negative learned-layer size");
    if (mode == XF_LB_APPLY) {
      if (!weights || n_weights != XF_LB_NPARAM) return fail("learned layer: bad weights");
      s->sim->addBoundary(xfdtd::LearnedBoundary::makeApply(
          faces, static_cast<xfdtd::Index>(size),
          std::vector<uint16_t>(weights, weights + n_weights)));
    } else if (mode == XF_LB_RECORD) {
      s->sim->addBoundary(xfdtd::LearnedBoundary::makeRecord(
          faces, static_cast<xfdtd::Index>(size), static_cast<xfdtd::Index>(every),
          static_cast<xfdtd::Index>(slots)));
    } else {
      return fail("learned layer: bad mode");
    }
    return 0;
  });
}

int xfh_sim_lb_samples(xfh_sim* s, const float** data, int64_t* columns, int64_t* recorded) {
  return guard([&] {
    if (!s || !data || !columns || !recorded) return fail("null argument");
    for (auto& b : s->sim->boundaries())
      if (auto l = dynamic_cast<xfdtd::LearnedBoundary*>(b.get())) {
        *data = l->samples().data();
        *columns = static_cast<int64_t>(l->columns());
        *recorded = static_cast<int64_t>(l->recordedSteps());
        return 0;
      }
    return fail("no learned layer");
  });
}

int xfh_sim_add_movie_monitor(xfh_sim* s, int32_t field, const double lo[3], const double size[3],
                              int64_t interval, const char* output_dir) {
  return guard([&] {
    if (!s || !lo || !size) return fail("null argument");
    if (field < 0 || field > 5) return fail("bad movie field");
    if (interval < 1) return fail("movie interval must be >= 1");
    auto cube = std::make_unique<xfdtd::Cube>(xfdtd::Vector{lo[0], lo[1], lo[2]},
                                              xfdtd::Vector{size[0], size[1], size[2]});
    auto fm = std::make_unique<xfdtd::FieldMonitor>(std::move(cube), static_cast<xfdtd::EMF::Field>(field));
    s->sim->addMonitor(std::make_shared<xfdtd::MovieMonitor>(std::move(fm), static_cast<xfdtd::Index>(interval),
                                                             "movie", output_dir ? output_dir : ""));
    return 0;
  });
}

int xfh_sim_output_movies(xfh_sim* s, int64_t* frames) {
  return guard([&] {
    if (!s) return fail("null sim");
    int64_t n = 0;
    for (auto& m : s->sim->monitors())
      if (auto mm = dynamic_cast<xfdtd::MovieMonitor*>(m.get())) n += static_cast<int64_t>(mm->output());
    if (frames) *frames = n;
    return 0;
  });
}

int xfh_sim_set_coefficient_layout(xfh_sim* s, int32_t layout) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (layout != 0 && layout != 1) return fail("bad layout");
    s->sim->setCoefficientLayout(static_cast<xfdtd::Simulation::CoefficientLayout>(layout));
    return 0;
  });
}

int xfh_sim_set_cell_sizes(xfh_sim* s, int32_t axis, const double* sizes, int64_t n) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (axis < 0 || axis > 2) return fail("bad axis");
    if (!sizes || n < 1) return fail("no cell sizes");
    s->sim->setCellSizes(static_cast<xfdtd::Axis::XYZ>(axis), std::vector<double>(sizes, sizes + n));
    return 0;
  });
}

int xfh_sim_init(xfh_sim* s, int64_t n) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (n <= 0) return fail("time_step must be positive");
    s->sim->init(static_cast<xfdtd::Index>(n));
    return 0;
  });
}

int xfh_sim_scene(xfh_sim* s, xf_scene** out) {
  return guard([&] {
    if (!s || !out) return fail("null argument");
    if (!s->sim->initialized()) return fail("simulation not initialised");
    *out = s->sim->scene();
    return 0;
  });
}

int xfh_sim_info(xfh_sim* s, int64_t dims[3], double* dt, int32_t* dimension) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (!s->sim->initialized()) return fail("simulation not initialised");
    const auto& g = *s->sim->gridSpace();
    if (dims) {
      dims[0] = g.sizeX();
      dims[1] = g.sizeY();
      dims[2] = g.sizeZ();
    }
    if (dt) *dt = s->sim->timeParam().dt();
    if (dimension) *dimension = static_cast<int32_t>(g.dimension());
    return 0;
  });
}

int xfh_scene_array(const xf_scene* sc, const char* name, double** data, int64_t shape[5],
                    int32_t* ndim) {
  return guard([&] {
    if (!sc || !name || !data || !shape || !ndim) return fail("null argument");
    const auto parts = split(name);
    const xf_array* a = nullptr;
    const std::string& h = parts[0];
    if (parts.size() == 1) {
      if (h == "ex") a = &sc->ex;
      else if (h == "ey") a = &sc->ey;
      else if (h == "ez") a = &sc->ez;
      else if (h == "hx") a = &sc->hx;
      else if (h == "hy") a = &sc->hy;
      else if (h == "hz") a = &sc->hz;
    } else if (h == "coeff" && parts.size() == 2) {
      const int n = std::stoi(parts[1]);
      if (n >= 0 && n < 18) a = &sc->coeff[n];
    } else if (h == "pml" && parts.size() == 3) {
      const int p = std::stoi(parts[1]);
      if (p >= 0 && p < sc->n_pml) a = pmlMember(sc->pml[p], parts[2]);
    } else if (h == "tfsf" && parts.size() == 2) {
      if (sc->tfsf) a = tfsfMember(*sc->tfsf, parts[1]);
    } else if (h == "ade" && parts.size() == 2) {
      if (sc->ade) a = adeMember(*sc->ade, parts[1]);
    } else if (h == "movie" && parts.size() == 3 && parts[2] == "data") {
      const int q = std::stoi(parts[1]);
      if (q >= 0 && q < sc->n_movie) a = &sc->movie[q].data;
    } else if (h == "lb" && parts.size() == 2 && parts[1] == "history") {
      if (sc->lb) a = &sc->lb->history;
    } else if (h == "point" && parts.size() == 3 && parts[2] == "value") {
      const int p = std::stoi(parts[1]);
      if (p >= 0 && p < sc->n_point) a = &sc->point[p].value;
    } else if (h == "nf2ff_td" && parts.size() >= 4) {
      const int q = std::stoi(parts[1]);
      const int face = faceIndex(parts[2]);
      if (q >= 0 && q < sc->n_nf2ff_td && face >= 0) {
        const xf_nf2ff_td& td = sc->nf2ff_td[q];
        static const char* pn[4] = {"wa", "wb", "ua", "ub"};
        static const char* qn[4] = {"ea", "eb", "ha", "hb"};
        for (int c = 0; c < 4; ++c) {
          if (parts.size() == 4 && parts[3] == pn[c]) a = &td.pot[face][c];
          if (parts.size() == 5 && parts[3] == "prev" && parts[4] == qn[c]) a = &td.prev[face][c];
        }
      }
    } else if (h == "nf2ff" && parts.size() >= 4) {
      const int q = std::stoi(parts[1]);
      const int f = std::stoi(parts[2]);
      if (q >= 0 && q < sc->n_nf2ff && f >= 0 && f < sc->nf2ff[q].n_freq) {
        const xf_nf2ff_freq& fr = sc->nf2ff[q].freq[f];
        if (parts.size() == 4 && parts[3] == "transform_e") a = &fr.transform_e;
        else if (parts.size() == 4 && parts[3] == "transform_h") a = &fr.transform_h;
        else if (parts.size() == 5) {
          const int face = faceIndex(parts[3]);
          static const char* cn[4] = {"ja", "jb", "ma", "mb"};
          for (int c = 0; c < 4 && face >= 0; ++c)
            if (parts[4] == cn[c]) a = &fr.cur[face][c];
        }
      }
    }
    if (a == nullptr) return fail(std::string("no array named ") + name);
    *data = a->data;
    *ndim = a->ndim;
    for (int d = 0; d < 5; ++d) shape[d] = a->shape[d];
    return 0;
  });
}

int xfh_sim_run_hd(xfh_sim* s, int64_t n, const int32_t grid[3], const int32_t block[3],
                   int32_t precision, int32_t use_graph) {
  return guard([&] {
    if (!s) return fail("null sim");
    if (n <= 0) return fail("time_step must be positive");
    xfdtd::cuda::SimulationHD hd(s->sim.get());
    if (grid) hd.setGridDim(dim3(grid[0], grid[1], grid[2]));
    if (block) hd.setBlockDim(dim3(block[0], block[1], block[2]));
    hd.setPrecision(static_cast<xf_precision>(precision));
    hd.setUseGraph(use_graph != 0);
    hd.run(static_cast<xfdtd::Index>(n));
    return 0;
  });
}

int xfh_sim_far_field(xfh_sim* s, int32_t q, int32_t f, const double* theta, int64_t n,
                      double phi, double inc_re, double inc_im, double* e_theta,
                      double* e_phi, double* rcs) {
  return guard([&] {
    if (!s) return fail("null sim");
    std::vector<xfdtd::NFFFTFrequencyDomain*> boxes;
    for (auto& b : s->sim->nf2ffs())
      if (auto p = dynamic_cast<xfdtd::NFFFTFrequencyDomain*>(b.get())) boxes.push_back(p);
    if (q < 0 || q >= static_cast<int32_t>(boxes.size())) return fail("bad NF2FF index");
    std::vector<double> th(theta, theta + n);
    auto ff = boxes[q]->processFarField(static_cast<xfdtd::Index>(f), th, phi,
                                        std::complex<double>(inc_re, inc_im));
    for (int64_t i = 0; i < n; ++i) {
      if (e_theta) {
        e_theta[2 * i] = ff.e_theta[i].real();
        e_theta[2 * i + 1] = ff.e_theta[i].imag();
      }
      if (e_phi) {
        e_phi[2 * i] = ff.e_phi[i].real();
        e_phi[2 * i + 1] = ff.e_phi[i].imag();
      }
      if (rcs) rcs[i] = ff.rcs[i];
    }
    return 0;
  });
}

int xfh_sim_far_field_td(xfh_sim* s, int32_t q, int64_t* n, double* time, double* e_theta,
                         double* e_phi) {
  return guard([&] {
    if (!s || !n) return fail("null argument");
    std::vector<xfdtd::NFFFTTimeDomain*> boxes;
    for (auto& b : s->sim->nf2ffs())
      if (auto p = dynamic_cast<xfdtd::NFFFTTimeDomain*>(b.get())) boxes.push_back(p);
    if (q < 0 || q >= static_cast<int32_t>(boxes.size())) return fail("bad NF2FF-TD index");
    *n = static_cast<int64_t>(boxes[q]->potentialSize());
    if (!time && !e_theta && !e_phi) return 0;
    auto ff = boxes[q]->farField();
    for (int64_t i = 0; i < *n; ++i) {
      if (time) time[i] = ff.time[i];
      if (e_theta) e_theta[i] = ff.e_theta[i];
      if (e_phi) e_phi[i] = ff.e_phi[i];
    }
    return 0;
  });
}

int xfh_sim_output_far_field(xfh_sim* s, int32_t q, const double* theta, int64_t n_theta,
                             const double* phi, int64_t n_phi, const char* sub_dir,
                             const char* output_dir) {
  return guard([&] {
    if (!s || !theta || !phi || !sub_dir) return fail("null argument");
    if (n_theta < 1 || n_phi < 1 || (n_theta > 1 && n_phi > 1))
      return fail("give a theta cut at one phi or a phi cut at one theta");
    std::vector<xfdtd::NFFFTFrequencyDomain*> boxes;
    for (auto& b : s->sim->nf2ffs())
      if (auto p = dynamic_cast<xfdtd::NFFFTFrequencyDomain*>(b.get())) boxes.push_back(p);
    if (q < 0 || q >= static_cast<int32_t>(boxes.size())) return fail("bad NF2FF index");
    if (output_dir) boxes[q]->setOutputDir(output_dir);
    if (n_phi == 1)
      boxes[q]->processFarField(std::vector<double>(theta, theta + n_theta), phi[0], sub_dir);
    else
      boxes[q]->processFarField(theta[0], std::vector<double>(phi, phi + n_phi), sub_dir);
    return 0;
  });
}

int xfh_sim_output_far_field_td(xfh_sim* s, int32_t q, const char* output_dir) {
  return guard([&] {
    if (!s) return fail("null sim");
    std::vector<xfdtd::NFFFTTimeDomain*> boxes;
    for (auto& b : s->sim->nf2ffs())
      if (auto p = dynamic_cast<xfdtd::NFFFTTimeDomain*>(b.get())) boxes.push_back(p);
    if (q < 0 || q >= static_cast<int32_t>(boxes.size())) return fail("bad NF2FF-TD index");
    if (output_dir) boxes[q]->setOutputDir(output_dir);
    boxes[q]->processFarField();
    return 0;
  });
}

int xfh_sim_output_waveform(xfh_sim* s, const char* output_dir) {
  return guard([&] {
    if (!s || !output_dir) return fail("null argument");
    for (auto& w : s->sim->waveformSources()) {
      auto t = dynamic_cast<xfdtd::TFSF*>(w.get());
      if (!t) continue;
      const auto dir = std::filesystem::path(output_dir);
      xfdtd::dumpNpy((dir / "time.npy").string(), t->waveform()->time());
      xfdtd::dumpNpy((dir / "incident_wave.npy").string(), t->waveform()->value());
      return 0;
    }
    return fail("no TFSF source");
  });
}

int xfh_sim_waveform(xfh_sim* s, double* time, double* value, int64_t n) {
  return guard([&] {
    if (!s) return fail("null sim");
    for (auto& w : s->sim->waveformSources()) {
      if (auto t = dynamic_cast<xfdtd::TFSF*>(w.get())) {
        const auto& tv = t->waveform()->time();
        const auto& vv = t->waveform()->value();
        if (static_cast<int64_t>(tv.size()) < n) return fail("waveform shorter than requested");
        for (int64_t i = 0; i < n; ++i) {
          if (time) time[i] = tv[i];
          if (value) value[i] = vv[i];
        }
        return 0;
      }
    }
    return fail("no TFSF source");
  });
}

int xfh_material_permittivity(const xfh_material* m, double w, double* re, double* im) {
  return guard([&] {
    auto mat = makeMaterial(m);
    const auto e = mat->relativePermittivity(w);
    if (re) *re = e.real();
    if (im) *im = e.imag();
    return 0;
  });
}

}  // extern "C"
