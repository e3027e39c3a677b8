/*
 * xfdtd_host.h -- C ABI of the host-side scene model (libxfdtd_host.so).
 *
 * Lets non-C++ callers (the Python package, tests, bench) build the scenes
 * the reference drivers build with XFDTD_CORE (test/dielectric_sphere_
 * scatter.cu:20-107, test/benchmark.cu:105-163, test/cylinder_scatter_2d.cu:
 * 28-183, test/dispersive_sphere_scatter/), run them through the
 * xfdtd::cuda::SimulationHD drop-in, and read the arrays back.  The scene
 * is the C++ xfdtd::Simulation of include/xfdtd_cuda_amd/xfdtd_core.h.
 *
 * All functions return 0 on success and a nonzero status otherwise;
 * xfh_last_error() holds the message.
 */
#ifndef XFDTD_HOST_H_
#define XFDTD_HOST_H_

#include <stddef.h>
#include <stdint.h>

#include "xfdtd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xfh_sim xfh_sim;

/* xfdtd::Simulation{dx, dy, dz, cfl} */
int xfh_sim_create(double dx, double dy, double dz, double cfl, xfh_sim** out);
int xfh_sim_destroy(xfh_sim* s);

/* Material kinds for xfh_sim_add_*:
 *  0 plain (eps_r, mu_r, sigma_e, sigma_m)
 *  1 PEC (Material::createPec)
 *  2 Drude   (eps_inf = p[0]; poles p[1+2i] = omega_p, p[2+2i] = gamma)
 *  3 Debye   (eps_inf = p[0]; poles p[1+2i] = delta_eps, p[2+2i] = tau)
 *  4 Lorentz (eps_inf = p[0]; poles p[1+3i] = delta_eps, omega_0, delta)
 *  5 M-Lorentz (eps_inf = p[0]; poles p[1+5i..5+5i] = a0, a1, b0, b1, b2)
 * n_pole counts poles; p holds the numbers above. */
typedef struct xfh_material {
  int32_t kind;
  double eps_r, mu_r, sigma_e, sigma_m;
  int32_t n_pole;
  const double* p;
} xfh_material;

/* Object shapes: Cube(origin, size), Sphere(center, radius),
 * Cylinder(center, radius, height) along z (height may be infinite). */
int xfh_sim_add_cube(xfh_sim* s, const char* name, const double origin[3],
                     const double size[3], const xfh_material* m);
int xfh_sim_add_sphere(xfh_sim* s, const char* name, const double center[3],
                       double radius, const xfh_material* m);
int xfh_sim_add_cylinder(xfh_sim* s, const char* name, const double center[3],
                         double radius, double height, const xfh_material* m);

/* PML(thickness, direction, order, sigma_ratio, alpha_min, alpha_max,
 * kappa_max); direction 0..5 = XN..ZP. */
int xfh_sim_add_pml(xfh_sim* s, int32_t direction, int64_t thickness,
                    int64_t order, double sigma_ratio, double alpha_min,
                    double alpha_max, double kappa_max);

/* Waveform kinds: 0 gaussian(tau, t0), 1 cosine-modulated gaussian(tau, t0,
 * f), 2 sine(f). */
int xfh_sim_add_tfsf3d(xfh_sim* s, const int64_t start[3], double theta,
                       double phi, double psi, int32_t wave_kind, double tau,
                       double t0, double freq);
int xfh_sim_add_tfsf2d(xfh_sim* s, const int64_t start[2], double phi,
                       int32_t wave_kind, double tau, double t0, double freq);

int xfh_sim_add_nf2ff_fd(xfh_sim* s, const int64_t start[3],
                         const double* freqs, int32_t n_freq);

/* NF2FF time-domain box (NFFFTTimeDomain, the host side of the reference's
 * NF2FFTimeDomainHD, nf2ff_time_domain_hd.cu:35-59) at the same inset
 * convention, observed along (theta, phi). */
int xfh_sim_add_nf2ff_td(xfh_sim* s, const int64_t start[3], double theta, double phi);

/* Additions without a reference counterpart.  A soft point source on
 * E_comp at index idx of that component's array (grid indices), driving
 * amplitude * waveform(n dt) after each E pass. */
int xfh_sim_add_point_source(xfh_sim* s, int32_t comp, const int64_t idx[3],
                             int32_t wave_kind, double tau, double t0, double freq,
                             double amplitude);

/* The learned absorbing layer (xf_learned_boundary in xfdtd_hip.h).
 * mode XF_LB_APPLY: `size` cells outside the domain on every face of the
 * mask, weights = XF_LB_NPARAM fp16 values.  mode XF_LB_RECORD: samples of
 * the plane `size` nodes inside each face every `every` steps, `slots`
 * recorded steps at most (weights unused). */
int xfh_sim_add_learned_boundary(xfh_sim* s, int32_t mode, int32_t faces, int64_t size,
                                 const uint16_t* weights, int32_t n_weights, int64_t every,
                                 int64_t slots);

/* Recorded samples of the learned layer after a run: a pointer to
 * recorded x columns x 13 floats (12 inputs, then the target). */
int xfh_sim_lb_samples(xfh_sim* s, const float** data, int64_t* columns, int64_t* recorded);

/* A movie monitor (reference MovieMonitor + FieldMonitor over a cube):
 * field 0..5 = Ex..Hz, the cube [lo, lo + size) in metres (an infinite
 * extent takes the whole axis), a frame every `interval` steps. */
int xfh_sim_add_movie_monitor(xfh_sim* s, int32_t field, const double lo[3], const double size[3],
                              int64_t interval, const char* output_dir);

/* Writes every movie monitor's frames as <output_dir>/%05d.npy
 * (MovieMonitor::output, movie_monitor.cu:64-94); *frames = frames written. */
int xfh_sim_output_movies(xfh_sim* s, int64_t* frames);

/* 0 dense (the reference's 18 coefficient arrays), 1 compact. */
int xfh_sim_set_coefficient_layout(xfh_sim* s, int32_t layout);

/* Simulation::setCellSizes: a graded axis (0 x, 1 y, 2 z) -- the sizes of
 * the domain's n cells along it from its lower end (the reference's
 * GridSpace e/h node and size arrays, grid_space.cuh:36-90). */
int xfh_sim_set_cell_sizes(xfh_sim* s, int32_t axis, const double* sizes, int64_t n);

/* Simulation::init(n): builds everything for n steps. */
int xfh_sim_init(xfh_sim* s, int64_t time_step);

/* The mirrored state (valid until the next init / destroy). */
int xfh_sim_scene(xfh_sim* s, xf_scene** out);
int xfh_sim_info(xfh_sim* s, int64_t dims[3], double* dt, int32_t* dimension);

/* Named array of any scene: "ex".."hz", "coeff.N", "pml.P.<member>",
 * "tfsf.<member>", "nf2ff.Q.F.<face>.<ja|jb|ma|mb>", "nf2ff.Q.F.transform_e",
 * "nf2ff_td.Q.<face>.<wa|wb|ua|ub>" (potentials),
 * "nf2ff_td.Q.<face>.prev.<ea|eb|ha|hb>", "ade.<member>", "movie.N.data",
 * "lb.history", "point.N.value".  Returns a pointer into the scene's storage. */
int xfh_scene_array(const xf_scene* sc, const char* name, double** data,
                    int64_t shape[5], int32_t* ndim);

/* Runs n steps through xfdtd::cuda::SimulationHD (init + copy to device +
 * DomainHD loop + copy back), the drop-in path of the reference's drivers.
 * grid/block are the setGridDim/setBlockDim values. */
int xfh_sim_run_hd(xfh_sim* s, int64_t time_step, const int32_t grid[3],
                   const int32_t block[3], int32_t precision,
                   int32_t use_graph);

/* Far field of NF2FF box q at frequency f along (theta[i], phi), normalised
 * by the incident spectrum (re, im) when nonzero: E_theta, E_phi as
 * (re, im) pairs and RCS. */
int xfh_sim_far_field(xfh_sim* s, int32_t q, int32_t f, const double* theta,
                      int64_t n, double phi, double inc_re, double inc_im,
                      double* e_theta, double* e_phi, double* rcs);

/* Far field of NF2FF time-domain box q: *n = potential bins; with non-NULL
 * arrays of *n doubles, the time axis and R E_theta(t), R E_phi(t)
 * (NFFFTTimeDomain::processFarField). */
int xfh_sim_far_field_td(xfh_sim* s, int32_t q, int64_t* n, double* time,
                         double* e_theta, double* e_phi);

/* The drivers' on-disk outputs.  Far field of frequency-domain box q for
 * every frequency along a theta cut at one phi (n_phi = 1) or a phi cut at
 * one theta (n_theta = 1), unnormalised, to <output_dir>/<sub_dir>/
 * {theta,phi,freq,e_theta,e_phi}.npy (NFFFTFrequencyDomain::processFarField
 * as called in dielectric_sphere_scatter.cu:86-91); output_dir NULL keeps
 * the box's own. */
int xfh_sim_output_far_field(xfh_sim* s, int32_t q, const double* theta, int64_t n_theta,
                             const double* phi, int64_t n_phi, const char* sub_dir,
                             const char* output_dir);

/* Time-domain box q: <output_dir>/{time,e_theta,e_phi}.npy. */
int xfh_sim_output_far_field_td(xfh_sim* s, int32_t q, const char* output_dir);

/* The TFSF waveform as the drivers dump it: <output_dir>/time.npy and
 * incident_wave.npy (dielectric_sphere_scatter.cu:96-99). */
int xfh_sim_output_waveform(xfh_sim* s, const char* output_dir);

/* Waveform samples of the TFSF source (time(), value()), length n_steps. */
int xfh_sim_waveform(xfh_sim* s, double* time, double* value, int64_t n);

/* Relative permittivity of material kind m at angular frequency w. */
int xfh_material_permittivity(const xfh_material* m, double w, double* re,
                              double* im);

const char* xfh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* XFDTD_HOST_H_ */
