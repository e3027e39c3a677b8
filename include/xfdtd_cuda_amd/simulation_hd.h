// simulation_hd.h -- drop-in for xfdtd::cuda::SimulationHD
// (include/xfdtd_cuda/simulation/simulation_hd.cuh:37-88 of XFDTD_CUDA).
//
// Same constructor, setGridDim/setBlockDim, init, run and copy/release
// methods, with C++ exceptions on error like the reference
// (common.cuh:16-26, simulation_hd.cu:251-256).  The time stepping runs in
// the gfx950 HIP engine behind the C ABI of include/xfdtd_hip.h; grid and
// block dimensions are accepted as hints (the engine tiles for wave64).
#ifndef XFDTD_CUDA_AMD_SIMULATION_HD_H_
#define XFDTD_CUDA_AMD_SIMULATION_HD_H_

#include <hip/hip_runtime_api.h>

#include "xfdtd_core.h"

namespace xfdtd::cuda {

class SimulationHD {
 public:
  explicit SimulationHD(xfdtd::Simulation* host);
  ~SimulationHD();
  SimulationHD(const SimulationHD&) = delete;
  SimulationHD& operator=(const SimulationHD&) = delete;

  auto copyHostToDevice() -> void;
  auto copyDeviceToHost() -> void;
  auto releaseDevice() -> void;
  auto run(Index time_step) -> void;
  auto init(Index time_step) -> void;
  auto setGridDim(dim3 grid_dim) -> void;
  auto setBlockDim(dim3 block_dim) -> void;

  // Additions (no reference counterpart): arithmetic mode and step capture.
  auto setPrecision(xf_precision p) -> void { _precision = p; }
  auto setUseGraph(bool g) -> void { _use_graph = g; }
  // check every stored field value every n steps (0: off); run() then throws
  // std::runtime_error when one is a NaN or an infinity (xf_plan.watchdog)
  auto setFieldCheckInterval(Index n) -> void { _watchdog = n; }
  auto host() -> xfdtd::Simulation* { return _host; }

 private:
  xfdtd::Simulation* _host;
  dim3 _grid_dim{1, 1, 1};
  dim3 _block_dim{1, 1, 1};
  xf_precision _precision{XF_FP64};
  bool _use_graph{false};
  Index _watchdog{0};
  xf_handle* _handle{nullptr};
  auto check(int status, const char* what) -> void;
};

}  // namespace xfdtd::cuda

#endif  // XFDTD_CUDA_AMD_SIMULATION_HD_H_
