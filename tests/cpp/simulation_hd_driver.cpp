// simulation_hd_driver.cpp -- the reference driver's calling sequence
// (test/dielectric_sphere_scatter.cu:20-107) against this repository's
// xfdtd::cuda::SimulationHD, as a user switching libraries would write it.
//
// usage: simulation_hd_driver <n_cells> <steps> <out.bin> [data_dir]
// Builds an n^3 vacuum cube at dl = 5e-3 with a dielectric sphere
// (r = 0.1 m scaled to the cube, eps_r 3, mu_r 2), 6 CPML faces of 6 cells,
// a Gaussian TFSF plane wave at a 9-cell inset and a 1 GHz NF2FF box at a
// 7-cell inset; runs it through SimulationHD; writes the grid size, the
// final Ex and the xz far-field RCS cut to out.bin.  With data_dir it also
// writes the driver's files (dielectric_sphere_scatter.cu:85-99): the xz
// and yz far-field cuts under data_dir/fd/, time.npy and incident_wave.npy.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/xfdtd_cuda_amd/simulation_hd.h"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: " << argv[0] << " <n_cells> <steps> <out.bin>\n";
    return 2;
  }
  const int n = std::atoi(argv[1]);
  const int steps = std::atoi(argv[2]);
  constexpr double dl{5e-3};
  const double edge = n * dl, half = edge / 2;
  try {
    auto domain{std::make_shared<xfdtd::Object>(
        "domain",
        std::make_unique<xfdtd::Cube>(xfdtd::Vector{-half, -half, -half},
                                      xfdtd::Vector{edge, edge, edge}),
        xfdtd::Material::createAir())};
    auto sphere{std::make_shared<xfdtd::Object>(
        "dielectric_sphere",
        std::make_unique<xfdtd::Sphere>(xfdtd::Vector{0, 0, 0}, edge * 0.1 / 0.35),
        std::make_unique<xfdtd::Material>("a", xfdtd::ElectroMagneticProperty{3, 2, 0, 0}))};
    const double tau{dl * 20 / 6e8};
    auto tfsf{std::make_shared<xfdtd::TFSF3D>(9, 9, 9, 0, 0, 0,
                                               xfdtd::Waveform::gaussian(tau, 4.5 * tau))};
    auto nffft{std::make_shared<xfdtd::NFFFTFrequencyDomain>(7, 7, 7, std::vector<double>{1e9})};

    xfdtd::Simulation s{dl, dl, dl, 0.9};
    s.addObject(domain);
    s.addObject(sphere);
    s.addWaveformSource(tfsf);
    for (auto d : {xfdtd::Axis::Direction::XN, xfdtd::Axis::Direction::XP,
                   xfdtd::Axis::Direction::YN, xfdtd::Axis::Direction::YP,
                   xfdtd::Axis::Direction::ZN, xfdtd::Axis::Direction::ZP})
      s.addBoundary(std::make_shared<xfdtd::PML>(6, d));
    s.addNF2FF(nffft);

    auto s_hd = xfdtd::cuda::SimulationHD{&s};
    s_hd.setGridDim(dim3{4, 4, 4});
    s_hd.setBlockDim(dim3{8, 8, 8});
    s_hd.run(steps);

    // far field, normalised by the incident spectrum (the drivers follow
    // run() with processFarField on the same host objects)
    const auto& wt = tfsf->waveform()->time();
    const auto& wv = tfsf->waveform()->value();
    std::complex<double> inc{0, 0};
    const double dt = s.timeParam().dt();
    for (std::size_t i = 0; i < wt.size(); ++i)
      inc += wv[i] * dt * std::exp(std::complex<double>(0, -2 * xfdtd::constant::PI * 1e9 * wt[i]));
    std::vector<double> theta(360);
    for (int i = 0; i < 360; ++i) theta[i] = -xfdtd::constant::PI + 2 * xfdtd::constant::PI * i / 359.0;
    auto ff = nffft->processFarField(0, theta, 0.0, inc);

    if (argc > 4) {
      const std::string data_path = argv[4];
      std::vector<double> cut(360);
      for (int i = 0; i < 360; ++i) cut[i] = -xfdtd::constant::PI + 2 * xfdtd::constant::PI * i / 359.0;
      nffft->setOutputDir(data_path + "/fd");
      nffft->processFarField(cut, 0, "xz");
      nffft->processFarField(cut, xfdtd::constant::PI * 0.5, "yz");
      xfdtd::dumpNpy(data_path + "/time.npy", tfsf->waveform()->time());
      xfdtd::dumpNpy(data_path + "/incident_wave.npy", tfsf->waveform()->value());
    }

    std::FILE* f = std::fopen(argv[3], "wb");
    if (!f) throw std::runtime_error("cannot open output");
    const auto& g = *s.gridSpace();
    const int64_t dims[4] = {static_cast<int64_t>(g.sizeX()), static_cast<int64_t>(g.sizeY()),
                             static_cast<int64_t>(g.sizeZ()),
                             static_cast<int64_t>(s.timeParam().currentTimeStep())};
    std::fwrite(dims, sizeof dims, 1, f);
    auto& ex = s.field(0, 0);
    std::fwrite(ex.data(), sizeof(double), ex.size(), f);
    std::fwrite(ff.rcs.data(), sizeof(double), ff.rcs.size(), f);
    std::fclose(f);
    std::cout << "ok " << dims[0] << "x" << dims[1] << "x" << dims[2] << " steps " << dims[3]
              << " rcs[180] " << ff.rcs[180] << "\n";
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
