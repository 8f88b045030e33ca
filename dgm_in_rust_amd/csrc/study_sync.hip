// Launch-boundary study (not part of libdgm): the cost of a grid-wide barrier inside one launch against
// the cost of a boundary between two launches of a captured HIP graph, for the same light passes.
//
// A pass adds 1 to one float per thread (n = blocks * threads floats), so its own work is negligible and
// the time of `phases` passes is dominated by how the passes are separated:
//   graph: `phases` launches of k_pass captured into one graph, replayed `reps` times;
//   grid:  one cooperative launch of k_passes, `phases` passes separated by cooperative_groups grid.sync().
// The difference between phases = P and phases = 1, over P - 1, is the cost of one separation.
// Built into libdgm_study.so by build.py; driven by `python -m dgm_in_rust_amd.studies --sync`.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <vector>

namespace cg = cooperative_groups;

namespace {

__global__ void k_pass(float* buf) { buf[blockIdx.x * blockDim.x + threadIdx.x] += 1.0f; }

__global__ void k_passes(float* buf, int phases) {
  cg::grid_group g = cg::this_grid();
  float* p = buf + blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = 0; i < phases; ++i) {
    *p += 1.0f;
    if (i + 1 < phases) g.sync();
  }
}

}  // namespace

extern "C" {

// The largest cooperative grid of k_passes with `threads` per block on device 0 (blocks per CU that fit
// times the CUs), or -1 when the device does not support cooperative launches.
int dgm_study_max_grid(int threads) {
  int coop = 0, cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, 0) != hipSuccess || !coop) return -1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_passes), threads, 0) !=
      hipSuccess)
    return -1;
  return per_cu * cus;
}

// Mean milliseconds of `phases` passes of blocks x threads, separated by launch boundaries inside a graph
// (*graph_ms) and by grid barriers inside one cooperative launch (*grid_ms), over `reps` repetitions after
// one untimed warm-up. Returns 0, 1 for bad arguments or a grid too large to be co-resident (the
// cooperative launch is then not attempted), 2 on a HIP error; the final buffer is checked for
// blocks * threads floats of (2 + 2 reps) * phases.
int dgm_study_sync(int blocks, int threads, int phases, int reps, double* graph_ms, double* grid_ms) {
  if (blocks < 1 || threads < 64 || threads > 1024 || phases < 1 || reps < 1 || !graph_ms || !grid_ms) return 1;
  const int maxg = dgm_study_max_grid(threads);
  if (maxg < blocks) return 1;
  const size_t n = (size_t)blocks * threads;
  float* buf = nullptr;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int status = 0;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && !status) status = 2;
    return status == 0;
  };
  float ms = 0.f;
  void* args[] = {&buf, &phases};
  std::vector<float> host;
  if (!ok(hipMalloc(&buf, n * sizeof(float))) || !ok(hipMemset(buf, 0, n * sizeof(float)))) goto done;
  if (!ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) || !ok(hipEventCreate(&e0)) || !ok(hipEventCreate(&e1)))
    goto done;
  // launch boundaries: the passes captured into one graph
  if (!ok(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal))) goto done;
  for (int i = 0; i < phases; ++i) hipLaunchKernelGGL(k_pass, dim3(blocks), dim3(threads), 0, st, buf);
  if (!ok(hipStreamEndCapture(st, &graph)) || !ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0))) goto done;
  if (!ok(hipGraphLaunch(exec, st))) goto done;
  if (!ok(hipEventRecord(e0, st))) goto done;
  for (int r = 0; r < reps; ++r)
    if (!ok(hipGraphLaunch(exec, st))) goto done;
  if (!ok(hipEventRecord(e1, st)) || !ok(hipEventSynchronize(e1)) || !ok(hipEventElapsedTime(&ms, e0, e1))) goto done;
  *graph_ms = ms / reps;
  // grid barriers: one cooperative launch per repetition
  if (!ok(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_passes), dim3(blocks), dim3(threads), args, 0, st)))
    goto done;
  if (!ok(hipEventRecord(e0, st))) goto done;
  for (int r = 0; r < reps; ++r)
    if (!ok(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_passes), dim3(blocks), dim3(threads), args, 0,
                                       st)))
      goto done;
  if (!ok(hipEventRecord(e1, st)) || !ok(hipEventSynchronize(e1)) || !ok(hipEventElapsedTime(&ms, e0, e1))) goto done;
  *grid_ms = ms / reps;
  host.resize(n);
  if (!ok(hipMemcpy(host.data(), buf, n * sizeof(float), hipMemcpyDeviceToHost))) goto done;
  for (size_t i = 0; i < n; ++i)
    if (host[i] != (float)((2 + 2 * reps) * phases)) {
      status = 2;
      break;
    }
done:
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (buf) (void)hipFree(buf);
  return status;
}

}  // extern "C"
