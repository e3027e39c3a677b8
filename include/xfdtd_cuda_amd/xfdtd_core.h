// xfdtd_core.h -- host-side scene model and set-up of the FDTD engine.
//
// XFDTD_CUDA only time-steps state that its host library XFDTD_CORE builds
// (SimulationHD::init calls host()->init(n), src/simulation/simulation_hd.cu:
// 259-267).  XFDTD_CORE is a separate project that this repository does not
// have, so this header restates the subset of it that SimulationHD consumes
// (the accessors listed at simulation_hd.cu:275-503 and in the HD mirrors):
// Simulation, GridSpace, EMF, FDTD coefficients, PML, TFSF2D/3D, Waveform,
// NFFFTFrequencyDomain and the Drude / Debye / M-Lorentz ADE storages.
// Class and method names follow XFDTD_CORE's as used by the reference
// drivers (test/dielectric_sphere_scatter.cu, test/benchmark.cu,
// test/cylinder_scatter_2d.cu, test/dispersive_sphere_scatter/*).
//
// The numerical set-up (voxelisation, CPML profiles, incident-field table,
// NF2FF phasors, ADE coefficients) is this project's own choice wherever
// XFDTD_CORE's is not visible in the reference; each choice is documented
// at its definition.
#ifndef XFDTD_CUDA_AMD_XFDTD_CORE_H_
#define XFDTD_CUDA_AMD_XFDTD_CORE_H_

#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../xfdtd_hip.h"

namespace xfdtd {

using Real = double;
using Index = std::size_t;

namespace constant {
inline constexpr Real PI = 3.14159265358979323846;
inline constexpr Real EPSILON_0 = 8.854187817e-12;
inline constexpr Real MU_0 = 4.0 * PI * 1e-7;
inline const Real C_0 = 1.0 / __builtin_sqrt(EPSILON_0 * MU_0);
inline const Real ETA_0 = __builtin_sqrt(MU_0 / EPSILON_0);
inline constexpr Real INF = std::numeric_limits<Real>::infinity();
}  // namespace constant

struct XFDTDException : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Row-major N-d array of doubles (last index fastest) that owns its memory.
class Array {
 public:
  Array() = default;
  explicit Array(std::vector<int64_t> shape, Real fill = 0.0);
  Real* data() { return _data.data(); }
  const Real* data() const { return _data.data(); }
  std::size_t size() const { return _data.size(); }
  const std::vector<int64_t>& shape() const { return _shape; }
  Real& operator[](std::size_t i) { return _data[i]; }
  Real operator[](std::size_t i) const { return _data[i]; }
  Real& at(int64_t i, int64_t j, int64_t k) {
    return _data[(i * _shape[1] + j) * _shape[2] + k];
  }
  Real at(int64_t i, int64_t j, int64_t k) const {
    return _data[(i * _shape[1] + j) * _shape[2] + k];
  }
  xf_array view();
  void fill(Real v);
  bool empty() const { return _data.empty(); }

 private:
  std::vector<Real> _data;
  std::vector<int64_t> _shape;
};

struct Vector {
  Real x{}, y{}, z{};
};

// .npy output (the drivers' xt::dump_npy): C-order doubles, or complex
// doubles (interleaved re, im) written as '<c16'.
void dumpNpy(const std::string& path, const double* data, const std::vector<int64_t>& shape,
             bool is_complex = false);
void dumpNpy(const std::string& path, const std::vector<Real>& v);

namespace Axis {
enum class XYZ { X = 0, Y = 1, Z = 2 };
enum class Direction { XN = 0, XP = 1, YN = 2, YP = 3, ZN = 4, ZP = 5 };
}  // namespace Axis

// ------------------------------------------------------------------ shapes
class Shape {
 public:
  virtual ~Shape() = default;
  virtual bool isInside(const Vector& p) const = 0;
  virtual Vector lower() const = 0;  // bounding box
  virtual Vector upper() const = 0;
};

class Cube : public Shape {
 public:
  Cube(Vector origin, Vector size) : _origin(origin), _size(size) {}
  bool isInside(const Vector& p) const override;
  Vector lower() const override { return _origin; }
  Vector upper() const override;

 private:
  Vector _origin, _size;
};

class Sphere : public Shape {
 public:
  Sphere(Vector center, Real radius) : _center(center), _radius(radius) {}
  bool isInside(const Vector& p) const override;
  Vector lower() const override;
  Vector upper() const override;
  Vector center() const { return _center; }
  Real radius() const { return _radius; }

 private:
  Vector _center;
  Real _radius;
};

// Circular cylinder along z (the 2-D scatterer of test/cylinder_scatter_2d.cu).
class Cylinder : public Shape {
 public:
  Cylinder(Vector center, Real radius, Real height)
      : _center(center), _radius(radius), _height(height) {}
  bool isInside(const Vector& p) const override;
  Vector lower() const override;
  Vector upper() const override;

 private:
  Vector _center;
  Real _radius, _height;
};

// --------------------------------------------------------------- materials
struct ElectroMagneticProperty {
  Real epsilon_r{1}, mu_r{1}, sigma_e{0}, sigma_m{0};
};

// One second-order ADE pole in the M-Lorentz form
//   chi_p(w) = (a0 + a1 jw) / (b0 + b1 jw + b2 (jw)^2).
// Lorentz: a0 = de*w0^2, b0 = w0^2, b1 = 2*delta, b2 = 1.
// Drude:   a0 = wp^2, b0 = 0, b1 = gamma, b2 = 1.
struct MLorentzPole {
  Real a0{}, a1{}, b0{}, b1{}, b2{};
};

class Material {
 public:
  enum class Dispersion { None = 0, Drude = 1, Debye = 2, MLorentz = 3 };
  Material(std::string name, ElectroMagneticProperty emp)
      : _name(std::move(name)), _emp(emp) {}
  static std::unique_ptr<Material> createAir();
  static std::unique_ptr<Material> createPec();
  const std::string& name() const { return _name; }
  const ElectroMagneticProperty& emp() const { return _emp; }
  Dispersion dispersion() const { return _dispersion; }
  // Drude: omega_p, gamma per pole.  Debye: delta_eps, tau per pole.
  const std::vector<Real>& p0() const { return _p0; }
  const std::vector<Real>& p1() const { return _p1; }
  const std::vector<MLorentzPole>& mlorPoles() const { return _mlor; }

  // Drude medium: eps(w) = eps_inf - sum wp^2 / (w^2 - j w gamma).
  static std::unique_ptr<Material> makeDrudeMedium(
      std::string name, Real eps_inf, std::vector<Real> omega_p,
      std::vector<Real> gamma);
  // Debye medium: eps(w) = eps_inf + sum de / (1 + j w tau).
  static std::unique_ptr<Material> makeDebyeMedium(
      std::string name, Real eps_inf, std::vector<Real> delta_eps,
      std::vector<Real> tau);
  // Lorentz medium: eps = eps_inf + sum de w0^2 / (w0^2 + 2 j w delta - w^2).
  static std::unique_ptr<Material> makeLorentzMedium(
      std::string name, Real eps_inf, std::vector<Real> delta_eps,
      std::vector<Real> omega_0, std::vector<Real> delta);
  // General M-Lorentz medium (XFDTD_CORE MLorentzMaterial::makeMLorentz).
  static std::unique_ptr<Material> makeMLorentz(std::string name,
                                                Real eps_inf,
                                                std::vector<MLorentzPole> p);
  // Relative permittivity at angular frequency w (analytic).
  std::complex<Real> relativePermittivity(Real w) const;
  // The same medium as M-Lorentz poles (Drude and Lorentz map exactly).
  std::vector<MLorentzPole> asMLorentz() const;

 private:
  std::string _name;
  ElectroMagneticProperty _emp;
  Dispersion _dispersion{Dispersion::None};
  std::vector<Real> _p0, _p1;
  std::vector<MLorentzPole> _mlor;
};

class Object {
 public:
  Object(std::string name, std::unique_ptr<Shape> shape,
         std::unique_ptr<Material> material)
      : _name(std::move(name)),
        _shape(std::move(shape)),
        _material(std::move(material)) {}
  const std::string& name() const { return _name; }
  const Shape& shape() const { return *_shape; }
  const Material& material() const { return *_material; }

 private:
  std::string _name;
  std::unique_ptr<Shape> _shape;
  std::unique_ptr<Material> _material;
};

// --------------------------------------------------------------- waveforms
class Waveform {
 public:
  enum class Kind { Gaussian = 0, CosineModulatedGaussian = 1, Sine = 2 };
  // exp(-((t - t0)/tau)^2)
  static std::unique_ptr<Waveform> gaussian(Real tau, Real t0);
  // exp(-((t - t0)/tau)^2) * cos(2 pi f (t - t0))
  static std::unique_ptr<Waveform> cosineModulatedGaussian(Real tau, Real t0,
                                                           Real f);
  static std::unique_ptr<Waveform> sine(Real f);
  Real operator()(Real t) const;
  // Sampled at t = n dt, n = 0..nt-1 (filled by init).
  void init(Real dt, Index nt);
  const std::vector<Real>& time() const { return _time; }
  const std::vector<Real>& value() const { return _value; }
  Kind kind() const { return _kind; }

 private:
  Kind _kind{Kind::Gaussian};
  Real _tau{}, _t0{}, _f{};
  std::vector<Real> _time, _value;
};

class Simulation;

// ------------------------------------------------------------------- grid
// The grid (grid_space.cuh:18-120): per axis the E nodes (cell corners,
// n + 1), the H nodes (cell centres, n), the E sizes (cell edges, n) and
// the H sizes (dual edges around each E node, n + 1; the two end values
// are the end cells' sizes).  An axis is uniform (based size) or graded
// (Simulation::setCellSizes).
class GridSpace {
 public:
  enum class Dimension { TWO = 2, THREE = 3 };
  Dimension dimension() const { return _dim; }
  Index sizeX() const { return _nx; }
  Index sizeY() const { return _ny; }
  Index sizeZ() const { return _nz; }
  Real basedDx() const { return _dx; }
  Real basedDy() const { return _dy; }
  Real basedDz() const { return _dz; }
  Real minDx() const { return minSize(0); }
  Real minDy() const { return minSize(1); }
  Real minDz() const { return minSize(2); }
  Vector origin() const { return _origin; }
  bool graded(int axis) const { return !_e_node[axis].empty(); }
  bool graded() const { return graded(0) || graded(1) || graded(2); }
  // node coordinate
  Real nodeX(Index i) const { return node(0, i); }
  Real nodeY(Index j) const { return node(1, j); }
  Real nodeZ(Index k) const { return node(2, k); }
  Real node(int axis, Index i) const {
    if (graded(axis)) return _e_node[axis][i];
    const Real o = axis == 0 ? _origin.x : axis == 1 ? _origin.y : _origin.z;
    return o + based(axis) * static_cast<Real>(i);
  }
  Real based(int axis) const { return axis == 0 ? _dx : axis == 1 ? _dy : _dz; }
  // size of cell i (E edge) and of the dual edge around E node i
  Real cellSize(int axis, Index i) const { return graded(axis) ? _e_size[axis][i] : based(axis); }
  Real dualSize(int axis, Index i) const {
    if (!graded(axis)) return based(axis);
    const Index n = size(axis);
    if (i == 0) return cellSize(axis, 0);
    if (i >= n) return cellSize(axis, n - 1);
    return (_e_size[axis][i - 1] + _e_size[axis][i]) / 2;
  }
  Index size(int axis) const { return axis == 0 ? _nx : axis == 1 ? _ny : _nz; }
  std::vector<Real> eNode(int axis) const;
  std::vector<Real> hNode(int axis) const;
  std::vector<Real> eSize(int axis) const;
  std::vector<Real> hSize(int axis) const;
  std::vector<Real> eNodeX() const { return eNode(0); }
  std::vector<Real> eNodeY() const { return eNode(1); }
  std::vector<Real> eNodeZ() const { return eNode(2); }
  std::vector<Real> hNodeX() const { return hNode(0); }
  std::vector<Real> hNodeY() const { return hNode(1); }
  std::vector<Real> hNodeZ() const { return hNode(2); }
  std::vector<Real> eSizeX() const { return eSize(0); }
  std::vector<Real> eSizeY() const { return eSize(1); }
  std::vector<Real> eSizeZ() const { return eSize(2); }
  std::vector<Real> hSizeX() const { return hSize(0); }
  std::vector<Real> hSizeY() const { return hSize(1); }
  std::vector<Real> hSizeZ() const { return hSize(2); }
  // the cell whose [node(i), node(i + 1)) holds x, clamped to [0, n - 1]
  Index cellOf(int axis, Real x) const;

 private:
  friend class Simulation;
  Real minSize(int axis) const;
  Dimension _dim{Dimension::THREE};
  Index _nx{}, _ny{}, _nz{};
  Real _dx{}, _dy{}, _dz{};
  Vector _origin{};
  std::vector<Real> _e_node[3];  // graded axes only: node coordinates
  std::vector<Real> _e_size[3];  // and cell sizes (as given)
};

class TimeParam {
 public:
  Real dt() const { return _dt; }
  Index startTimeStep() const { return _start; }
  Index size() const { return _size; }
  Index currentTimeStep() const { return _current; }
  Index endTimeStep() const { return _start + _size; }
  void nextStep() { ++_current; }

 private:
  friend class Simulation;
  Real _dt{};
  Index _start{}, _size{}, _current{};
};

// --------------------------------------------------------------- boundary
class Boundary {
 public:
  virtual ~Boundary() = default;
};

// Convolutional PML (Roden & Gedney 2000, CFS-CPML) on one face.
// Profiles (this project's choice; XFDTD_CORE's defaults are not visible):
//   sigma(r) = sigma_max (r/L)^m, sigma_max = sigma_ratio * 0.8 (m+1)/(eta0 dl)
//   kappa(r) = 1 + (kappa_max - 1)(r/L)^m
//   alpha(r) = alpha_min + (alpha_max - alpha_min)(1 - r/L)
//   b = exp(-(sigma/kappa + alpha) dt/eps0), a = sigma/(sigma kappa +
//   kappa^2 alpha) (b - 1), with r the depth from the inner interface.
class PML : public Boundary {
 public:
  PML(Index thickness, Axis::Direction direction, Index order = 3,
      Real sigma_ratio = 1.0, Real alpha_min = 0.0, Real alpha_max = 0.05,
      Real kappa_max = 1.0);
  Index thickness() const { return _thickness; }
  Axis::Direction direction() const { return _direction; }
  Axis::XYZ mainAxis() const;
  xf_pml& desc() { return _desc; }

 private:
  friend class Simulation;
  Index _thickness;
  Axis::Direction _direction;
  Index _order;
  Real _sigma_ratio, _alpha_min, _alpha_max, _kappa_max;
  Array _ca_e, _cb_e, _ca_h, _cb_h;
  Array _c_ea, _c_eb, _c_ha, _c_hb;
  Array _psi_ea, _psi_eb, _psi_ha, _psi_hb;
  std::vector<Real> _kappa_e, _kappa_h;
  xf_pml _desc{};
};

// ------------------------------------------------------------------- TFSF
class WaveformSource {
 public:
  virtual ~WaveformSource() = default;
};

// Total-field / scattered-field plane wave (3-D, or 2-D TMz with
// dimension 2).  The incident field comes from a 1-D auxiliary FDTD grid
// with the dispersion-matching cell ratio of test/2d.cu:793-843, an
// incident point two aux cells before the first TF corner, and a
// first-order Mur end condition (test/2d.cu:850-890); the aux E is driven
// by the waveform sampled at t = n dt.
class TFSF : public WaveformSource {
 public:
  TFSF(Index start_x, Index start_y, Index start_z, Real theta, Real phi,
       Real psi, std::unique_ptr<Waveform> waveform, int dimension);
  Waveform* waveform() { return _waveform.get(); }
  xf_tfsf& desc() { return _desc; }
  int dimension() const { return _dimension; }
  Real theta() const { return _theta; }
  Real phi() const { return _phi; }
  Real psi() const { return _psi; }
  const std::array<Real, 3>& kVector() const { return _k; }

 private:
  friend class Simulation;
  Index _sx, _sy, _sz;
  Real _theta, _phi, _psi;
  std::unique_ptr<Waveform> _waveform;
  int _dimension;
  std::array<Real, 3> _k{};
  Array _pxi, _pyi, _pzi, _pxh, _pyh, _pzh, _e_inc, _h_inc;
  xf_tfsf _desc{};
};

class TFSF3D : public TFSF {
 public:
  TFSF3D(Index start_x, Index start_y, Index start_z, Real theta, Real phi,
         Real psi, std::unique_ptr<Waveform> waveform)
      : TFSF(start_x, start_y, start_z, theta, phi, psi, std::move(waveform),
             3) {}
};

class TFSF2D : public TFSF {
 public:
  // TMz plane wave in the xy plane at azimuth phi.
  TFSF2D(Index start_x, Index start_y, Real phi,
         std::unique_ptr<Waveform> waveform)
      : TFSF(start_x, start_y, 0, constant::PI / 2, phi, 0,
             std::move(waveform), 2) {}
};

// ---------------------------------------------------------------- monitors
struct EMF {
  enum class Field { EX = 0, EY = 1, EZ = 2, HX = 3, HY = 4, HZ = 5 };
};

class Monitor {
 public:
  virtual ~Monitor() = default;
};

// A field component over the cells a shape covers (its bounding box,
// snapped to the grid; an infinite extent takes the whole axis).
class FieldMonitor {
 public:
  FieldMonitor(std::unique_ptr<Shape> shape, EMF::Field field, std::string name = "",
               std::string output_dir = "")
      : _shape(std::move(shape)), _field(field), _name(std::move(name)), _dir(std::move(output_dir)) {}
  EMF::Field field() const { return _field; }
  const Shape& shape() const { return *_shape; }

 private:
  std::unique_ptr<Shape> _shape;
  EMF::Field _field;
  std::string _name, _dir;
};

// Reference MovieMonitor (src/monitor/movie_monitor.cu): a frame of the
// field monitor's box every `frame_interval` steps; output() writes frame n
// as <output_dir>/%05d.npy (movie_monitor.cu:64-104).
class MovieMonitor : public Monitor {
 public:
  MovieMonitor(std::unique_ptr<FieldMonitor> frame, Index frame_interval, std::string name = "movie",
               std::string output_dir = "");
  FieldMonitor& frame() { return *_frame; }
  Index frameInterval() const { return _interval; }
  const std::string& outputDir() const { return _dir; }
  void setOutputDir(std::string d) { _dir = std::move(d); }
  Array& data() { return _data; }  // (frames, box extents)
  xf_movie& desc() { return _desc; }
  // writes every frame as <output_dir>/%05d.npy; returns the frame count
  Index output() const;

 private:
  friend class Simulation;
  std::unique_ptr<FieldMonitor> _frame;
  Index _interval;
  std::string _name, _dir;
  Array _data;
  xf_movie _desc{};
};

// ---------------------------------------------------------- point source
// Soft point source (no counterpart in the reference's GPU path): after the
// E pass of step n, amplitude * g(n dt) is added to E_component at index
// (i, j, k) of that component's array (grid indices, boundary cells
// included).
class PointSource : public WaveformSource {
 public:
  PointSource(Axis::XYZ component, Index i, Index j, Index k,
              std::unique_ptr<Waveform> waveform, Real amplitude = 1.0);
  Waveform* waveform() { return _waveform.get(); }
  xf_point_source& desc() { return _desc; }

 private:
  friend class Simulation;
  Axis::XYZ _comp;
  Index _i, _j, _k;
  std::unique_ptr<Waveform> _waveform;
  Real _amp;
  Array _value;
  xf_point_source _desc{};
};

// ------------------------------------------------------ learned boundary
// Learned absorbing layer (no reference counterpart; the column model is
// documented at xf_learned_boundary in xfdtd_hip.h).
//  apply:  `thickness` cells of the domain's material are added outside the
//          domain on every face in `faces` (bit f = face f, XN..ZP); the
//          outer node of each is driven by the column MLP (fp16 weights).
//  record: nothing is added; every `every` steps the inputs and the actual
//          next value of every column of the plane `inset` nodes inside
//          each face are kept (training data), up to `slots` steps.
// All three cell sizes must be equal.
class LearnedBoundary : public Boundary {
 public:
  static std::shared_ptr<LearnedBoundary> makeApply(int faces, Index thickness,
                                                    std::vector<uint16_t> weights);
  static std::shared_ptr<LearnedBoundary> makeRecord(int faces, Index inset, Index every,
                                                     Index slots);
  int mode() const { return _desc.mode; }
  int faces() const { return _desc.faces; }
  Index thickness() const { return _thickness; }
  Index columns() const { return _columns; }  // samples per recorded step
  Index recordedSteps() const { return static_cast<Index>(_desc.rec_count); }
  const std::vector<float>& samples() const { return _rec; }
  std::vector<uint16_t>& weights() { return _weights; }
  xf_learned_boundary& desc() { return _desc; }

 private:
  friend class Simulation;
  Index _thickness{0};
  Index _columns{0};
  std::vector<uint16_t> _weights;
  Array _history;
  std::vector<float> _rec;
  xf_learned_boundary _desc{};
};

// ---------------------------------------------------------------- NF2FF
class NF2FF {
 public:
  virtual ~NF2FF() = default;
};

// Frequency-domain Huygens box at an inset of (sx, sy, sz) cells from the
// outer boundary.  Phasors (this project's choice, matching the reference's
// time axes eTime = (n+1) dt, hTime = (n+1/2) dt, time_param.cu:13-31):
//   transform_e[n] = dt exp(-j w (n+1) dt),  transform_h[n] = dt exp(-j w (n+1/2) dt).
class NFFFTFrequencyDomain : public NF2FF {
 public:
  NFFFTFrequencyDomain(Index sx, Index sy, Index sz,
                       std::vector<Real> frequencies, std::string output_dir = "");
  Index freqCount() const { return _freqs.size(); }
  const std::vector<Real>& frequencies() const { return _freqs; }
  xf_nf2ff& desc() { return _desc; }
  void setOutputDir(std::string dir) { _output_dir = std::move(dir); }
  const std::string& outputDir() const { return _output_dir; }
  // Far field along directions (theta_i, phi_i): complex E_theta, E_phi
  // (divided by the incident spectrum when it is nonzero) and bistatic RCS
  // 4 pi (|E_theta|^2 + |E_phi|^2).
  struct FarField {
    std::vector<std::complex<Real>> e_theta, e_phi;
    std::vector<Real> rcs;
  };
  FarField farField(Index freq_index, const std::vector<Real>& theta,
                    const std::vector<Real>& phi, std::complex<Real> incident) const;
  FarField processFarField(Index freq_index, const std::vector<Real>& theta,
                           Real phi, std::complex<Real> incident) const;
  // The drivers' calls (dielectric_sphere_scatter.cu:86-91,
  // dispersive_sphere_scatter.cuh:113-124): a theta cut at one phi or a phi
  // cut at one theta, every frequency, written unnormalised to
  // <outputDir>/<sub_dir>/: theta.npy and phi.npy (one value per direction),
  // freq.npy, e_theta.npy and e_phi.npy (complex, frequencies x directions).
  void processFarField(const std::vector<Real>& theta, Real phi, const std::string& sub_dir) const;
  void processFarField(Real theta, const std::vector<Real>& phi, const std::string& sub_dir) const;

 private:
  friend class Simulation;
  Index _sx, _sy, _sz;
  std::vector<Real> _freqs;
  std::string _output_dir;
  void writeFarField(const std::vector<Real>& theta, const std::vector<Real>& phi,
                     const std::string& sub_dir) const;
  const GridSpace* _grid{};
  std::vector<Array> _cur;           // [freq][face][4]
  std::vector<Array> _te, _th;       // [freq]
  std::vector<xf_nf2ff_freq> _fdesc;
  xf_nf2ff _desc{};
};

// Time-domain Huygens box at an inset of (sx, sy, sz) cells for one
// observation direction (theta, phi): the host side the reference's
// NF2FFTimeDomainHD consumes (nf2ff_time_domain_hd.cu:35-59):
// observationDirection, distanceRange<E|H>, the per-face potentials
// (equivalentSurfaceCurrent: wa, wb from H, ua, ub from E) and the previous
// face fields.  The potentials hold n_pot = nt + ceil((dmax - dmin) / (c dt))
// + 2 time bins; bin m is the observer time m dt measured from the arrival
// of a wave leaving the surface point of largest r . rhat at time 0.
class NFFFTTimeDomain : public NF2FF {
 public:
  NFFFTTimeDomain(Index sx, Index sy, Index sz, Real theta, Real phi);
  // the driver form (dielectric_sphere_scatter.cu:72-74)
  NFFFTTimeDomain(Index sx, Index sy, Index sz, std::string output_dir, Real theta, Real phi);
  void setOutputDir(std::string dir) { _output_dir = std::move(dir); }
  const std::string& outputDir() const { return _output_dir; }
  Vector observationDirection() const { return _r_unit; }
  Real theta() const { return _theta; }
  Real phi() const { return _phi; }
  // min / max of r . rhat over the surface cells (metres); E and H use the
  // same surface points (rVector, nf2ff_time_domain_data.cu:54-68).
  std::pair<Real, Real> distanceRange() const { return {_dmin, _dmax}; }
  Index potentialSize() const { return _n_pot; }
  xf_nf2ff_td& desc() { return _desc; }
  const Array& potential(int face, int q) const { return _pot[face * 4 + q]; }
  // Far field R E_theta(t_m), R E_phi(t_m), t_m = m dt (Luebbers et al.,
  // IEEE TAP 39(4), 1991): W = w / (4 pi c dt), U = u / (4 pi c dt) summed
  // over the faces as vectors, R E_theta = -(eta W_theta + U_phi),
  // R E_phi = -eta W_phi + U_theta.
  struct FarField {
    std::vector<Real> time, e_theta, e_phi;
  };
  FarField farField() const;
  // Writes time.npy, e_theta.npy and e_phi.npy to outputDir().
  void processFarField() const;

 private:
  friend class Simulation;
  Index _sx, _sy, _sz;
  Real _theta, _phi;
  std::string _output_dir;
  Vector _r_unit{};
  Real _dmin{}, _dmax{}, _dt{};
  Index _n_pot{};
  std::vector<Array> _pot;   // [face][wa, wb, ua, ub]
  std::vector<Array> _prev;  // [face][ea, eb, ha, hb]
  std::array<Array, 3> _e_node, _h_node, _e_size;
  xf_nf2ff_td _desc{};
};

// --------------------------------------------------------------- ADE
// ADE storage in the reference's layout (ade_method.cuh:14-183).
// Coefficients (this project's derivation, second-order accurate):
//  Drude  dJ/dt + g J = e0 wp^2 E   at n+1/2:
//    cJJ = (1 - g dt/2)/(1 + g dt/2), cJE = e0 wp^2 dt/2 / (1 + g dt/2).
//  Debye  tau dJ/dt + J = e0 de dE/dt at n+1/2:
//    cJJ = (1 - dt/2tau)/(1 + dt/2tau), cJE = (e0 de / tau)/(1 + dt/2tau).
//  M-Lorentz  b0 J + b1 J' + b2 J'' = e0 (a0 E' + a1 E'') central at n:
//    K = b1/2dt + b2/dt^2, cJJ = (2 b2/dt^2 - b0)/K,
//    cJJp = (b1/2dt - b2/dt^2)/K, cJEn = e0 (a0/2dt + a1/dt^2)/K,
//    cJE = -2 e0 a1/(dt^2 K), cJEp = e0 (a1/dt^2 - a0/2dt)/K.
//  Ampere at n+1/2 with J^{n+1/2} = (J^{n+1} + J^n)/2 gives the Yee
//  coefficients and cEJsum = -1/(2 D), cJsumJ = 1 + cJJ, cJsumJp = cJJp,
//  cEEp = -sum cJEp/(2 D), D = e0 eps_inf/dt + sigma/2 + sum(cJE or cJEn)/2.
class ADEMethodStorage {
 public:
  int kind() const { return _desc.kind; }
  Index numPole() const { return _np; }
  xf_ade& desc() { return _desc; }

 private:
  friend class Simulation;
  Index _np{};
  Array _cjj, _cjjp, _cjen, _cje, _cjep, _cjsj, _cjsjp, _cejs, _ceep;
  Array _exp, _eyp, _ezp, _jx, _jy, _jz, _jxp, _jyp, _jzp;
  xf_ade _desc{};
};

// -------------------------------------------------------------- Simulation
// Scene + set-up.  Simulation(dx, dy, dz, cfl): the grid covers the first
// object ("domain"); each PML adds its thickness outside it.  A domain
// infinite along z gives a 2-D grid (nz = 1).
class Simulation {
 public:
  Simulation(Real dx, Real dy, Real dz, Real cfl);
  void addObject(std::shared_ptr<Object> o) { _objects.push_back(std::move(o)); }
  void addBoundary(std::shared_ptr<Boundary> b) { _boundaries.push_back(std::move(b)); }
  void addWaveformSource(std::shared_ptr<WaveformSource> w) { _sources.push_back(std::move(w)); }
  void addNF2FF(std::shared_ptr<NF2FF> n) { _nf2ffs.push_back(std::move(n)); }
  void addMonitor(std::shared_ptr<Monitor> m) { _monitors.push_back(std::move(m)); }

  // Coefficient layout handed to the engine: dense per-field arrays (the
  // reference's layout) or a per-cell material index + material table.
  enum class CoefficientLayout { Dense = 0, Compact = 1 };
  void setCoefficientLayout(CoefficientLayout l) { _layout = l; }

  // A graded axis: the sizes of the domain's cells along it, from its lower
  // end (they must add up to the domain's extent); PML cells outside the
  // domain keep the based size.  The Yee coefficients then use each node's
  // own cell and dual-edge sizes (dense layout only), dt the smallest cells;
  // TF/SF boxes and NF2FF time-domain boxes need uniform cells where they
  // lie, learned layers a uniform grid.
  void setCellSizes(Axis::XYZ axis, std::vector<Real> sizes);

  // Builds grid, materials, coefficients, PML, TFSF tables, NF2FF phasors,
  // ADE storage and zeroed fields for time_step steps.
  void init(Index time_step);
  bool initialized() const { return _initialized; }

  const std::shared_ptr<GridSpace>& gridSpace() const { return _grid; }
  TimeParam& timeParam() { return _time; }
  const std::vector<std::shared_ptr<Object>>& objects() const { return _objects; }
  const std::vector<std::shared_ptr<Boundary>>& boundaries() const { return _boundaries; }
  const std::vector<std::shared_ptr<WaveformSource>>& waveformSources() const { return _sources; }
  const std::vector<std::shared_ptr<NF2FF>>& nf2ffs() const { return _nf2ffs; }
  const std::vector<std::shared_ptr<Monitor>>& monitors() const { return _monitors; }
  ADEMethodStorage* aDEMethodStorage() { return _ade.get(); }
  Real cfl() const { return _cfl; }

  // The mirrored state as POD (pointers into this object's arrays).
  xf_scene* scene() { return &_scene; }
  Array& field(int attr_e0_h1, int axis) { return _fields[attr_e0_h1 * 3 + axis]; }
  Array& coefficient(int n) { return _coeff[n]; }
  const std::vector<uint16_t>& cellMaterial() const { return _cell_mat; }
  Index materialCount() const { return _mat_table.size(); }
  const Material& material(Index m) const { return *_mat_table[m]; }

 private:
  Real _dx, _dy, _dz, _cfl;
  CoefficientLayout _layout{CoefficientLayout::Dense};
  std::vector<std::shared_ptr<Object>> _objects;
  std::vector<std::shared_ptr<Boundary>> _boundaries;
  std::vector<std::shared_ptr<WaveformSource>> _sources;
  std::vector<std::shared_ptr<NF2FF>> _nf2ffs;
  std::vector<std::shared_ptr<Monitor>> _monitors;
  std::vector<xf_movie> _movie_desc;
  std::shared_ptr<GridSpace> _grid;
  TimeParam _time;
  std::array<Array, 6> _fields;
  std::array<Array, 18> _coeff;
  std::vector<uint16_t> _cell_mat;
  std::vector<const Material*> _mat_table;
  std::vector<Real> _mat_coeff;  // (n_mat, 18)
  std::unique_ptr<ADEMethodStorage> _ade;
  std::vector<xf_pml> _pml_desc;
  std::vector<xf_nf2ff> _nf_desc;
  std::vector<xf_nf2ff_td> _nftd_desc;
  xf_scene _scene{};
  bool _initialized{false};

  void buildGrid();
  void voxelize();
  void buildCoefficients();
  void buildPML(PML& p);
  void buildTFSF(TFSF& t);
  void buildNF2FF(NFFFTFrequencyDomain& n);
  void buildNF2FFTD(NFFFTTimeDomain& n);
  void buildADE();
  void buildPointSource(PointSource& p);
  void buildLearnedBoundary(LearnedBoundary& l);
  void buildMovie(MovieMonitor& m);
  void fillScene();
  std::vector<xf_point_source> _pt_desc;
  std::vector<Real> _graded_sizes[3];
  // per material: the E and H denominators, and for dispersive media the
  // ADE curl numerator (the curl coefficients divide these by a size)
  std::vector<Real> _mat_den_e, _mat_den_h, _mat_ccurl;
  std::vector<unsigned char> _mat_disp;
  // Yee coefficient n (0..17) of the cell (i, j, k) from its material and
  // the local sizes (graded grids)
  Real gradedCoeff(int n, uint16_t m, int64_t i, int64_t j, int64_t k) const;
};

}  // namespace xfdtd

#endif  // XFDTD_CUDA_AMD_XFDTD_CORE_H_
