// simulation_hd.cpp -- xfdtd::cuda::SimulationHD over the gfx950 engine.
// Mirrors src/simulation/simulation_hd.cu:44-273 of XFDTD_CUDA: run(n) =
// init(n), copy to device, DomainHD loop, copy back; errors become
// std::runtime_error.
#include "../../../include/xfdtd_cuda_amd/simulation_hd.h"

#include <algorithm>
#include <string>

namespace xfdtd::cuda {

SimulationHD::SimulationHD(xfdtd::Simulation* host) : _host(host) {
  if (host == nullptr) throw std::runtime_error("SimulationHD: host is nullptr");
}

SimulationHD::~SimulationHD() {
  if (_handle) xf_destroy(_handle);
  _handle = nullptr;
}

auto SimulationHD::check(int status, const char* what) -> void {
  if (status != XF_OK) {
    throw std::runtime_error(std::string("SimulationHD::") + what + ": " + xf_last_error());
  }
}

// simulation_hd.cu:259-267
auto SimulationHD::init(Index time_step) -> void { _host->init(time_step); }

// simulation_hd.cu:48-60 (plus the agency assembly of :119-224)
auto SimulationHD::copyHostToDevice() -> void {
  if (!_host->initialized()) throw std::runtime_error("SimulationHD::copyHostToDevice: host not initialised");
  releaseDevice();
  xf_plan plan{};
  plan.precision = _precision;
  plan.grid_dim[0] = static_cast<int32_t>(_grid_dim.x);
  plan.grid_dim[1] = static_cast<int32_t>(_grid_dim.y);
  plan.grid_dim[2] = static_cast<int32_t>(_grid_dim.z);
  plan.block_dim[0] = static_cast<int32_t>(_block_dim.x);
  plan.block_dim[1] = static_cast<int32_t>(_block_dim.y);
  plan.block_dim[2] = static_cast<int32_t>(_block_dim.z);
  plan.use_graph = _use_graph ? 1 : 0;
  plan.watchdog = static_cast<int32_t>(std::min<Index>(_watchdog, 0x7fffffff));
  check(xf_create(_host->scene(), &plan, &_handle), "copyHostToDevice");
  check(xf_upload(_handle), "copyHostToDevice");
}

// simulation_hd.cu:62-86
auto SimulationHD::copyDeviceToHost() -> void {
  if (_handle == nullptr) throw std::runtime_error("SimulationHD::copyDeviceToHost: nothing on the device");
  check(xf_download(_handle), "copyDeviceToHost");
  // keep the host TimeParam in step with the device (time_param_hd.cu:27-37)
  int64_t step = 0;
  check(xf_current_step(_handle, &step), "copyDeviceToHost");
  while (static_cast<int64_t>(_host->timeParam().currentTimeStep()) < step) _host->timeParam().nextStep();
}

// simulation_hd.cu:88-104
auto SimulationHD::releaseDevice() -> void {
  if (_handle) xf_destroy(_handle);
  _handle = nullptr;
}

// simulation_hd.cu:106-257
auto SimulationHD::run(Index time_step) -> void {
  init(time_step);
  copyHostToDevice();
  check(xf_run(_handle, static_cast<int64_t>(time_step)), "run");
  copyDeviceToHost();
}

auto SimulationHD::setGridDim(dim3 grid_dim) -> void { _grid_dim = grid_dim; }

auto SimulationHD::setBlockDim(dim3 block_dim) -> void { _block_dim = block_dim; }

}  // namespace xfdtd::cuda
