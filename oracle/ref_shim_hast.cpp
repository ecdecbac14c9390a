// Test harness: the reference's CC driver flow (CC/main.cu:43-203) restated
// over the reference's own headers and CPU path (CC/host_spanning_tree.cu,
// built alongside) with libphem.so providing SpanningTree and
// gpu_spanning_tree in place of the CUDA sources.  Test infrastructure only.
#include <climits>
#include <cstdint>
#include <cstring>
#include <hip/hip_runtime.h>

#include "include/hbcg_utils.cuh"
#include "include/ExternalSpanningTree.cuh"
#include "include/GPUSpanningtree.cuh"
#include "include/host_spanning_tree.cuh"

extern int Batch_Size;
extern float GPU_share;
extern "C" void phem_set_host_spanning_tree(int (*fn)(struct graph_data_host *));

extern "C" __attribute__((visibility("default"))) int ref_hast_run(
    int V, long E, const uint64_t *edges, float gpu_share, int batch,
    uint32_t *labels_out, uint64_t *t1_out, uint64_t *t2_out, float *ms_out) {
  Batch_Size = batch;
  GPU_share = gpu_share;
  phem_set_host_spanning_tree(host_spanning_tree);
  long egpu = gpu_share >= 1.0f ? E : (long)((double)gpu_share * (double)E);
  if (egpu > E) egpu = E;
  const long ecpu = E - egpu;
  int hV = V;
  long hE = E, hes = ecpu, hels = egpu;
  graph_data_host h_input;
  graph_data d_input, h_input_gpu;
  std::memset(&d_input, 0, sizeof(d_input));
  std::memset(&h_input_gpu, 0, sizeof(h_input_gpu));
  h_input.V = &hV;
  h_input.E = &hE;
  h_input.edges_size = &hes;
  h_input.edge_list_size = &hels;
  h_input_gpu.V = &hV;
  h_input_gpu.E = &hE;
  uint64_t *pinned = nullptr;
  if (hipHostMalloc((void **)&pinned, sizeof(uint64_t) * (egpu ? egpu : 1)) != hipSuccess)
    return -1;
  if (egpu) std::memcpy(pinned, edges, sizeof(uint64_t) * egpu);
  h_input_gpu.edges = pinned;
  const size_t b = batch > 0 ? batch : 1;
  if (hipMalloc((void **)&d_input.edges, 8 * b) || hipMalloc((void **)&d_input.edges2, 8 * b) ||
      hipMalloc((void **)&d_input.label, 4 * (size_t)V) ||
      hipMalloc((void **)&d_input.temp_label, 4 * (size_t)V) ||
      hipMalloc((void **)&d_input.T1edges, 8 * (size_t)V) ||
      hipMalloc((void **)&d_input.T2edges, 8 * (size_t)V))
    return -1;
  h_input.label = parlay::sequence<int>::from_function(V, [](size_t i) { return (int)i; });
  h_input.temp_label = parlay::sequence<int>::from_function(V, [](size_t i) { return (int)i; });
  h_input.sptree = parlay::sequence<uint64_t>(V, (uint64_t)INT_MAX);
  h_input.edges = parlay::sequence<uint64_t>(edges + egpu, edges + E);

  float ms = SpanningTree(&h_input_gpu, &d_input, &h_input);
  uint64_t *temp_array = nullptr;
  if (hipMalloc((void **)&temp_array, 8 * (size_t)V) ||
      hipMemcpy(temp_array, h_input.sptree.begin(), 8 * (size_t)V, hipMemcpyHostToDevice))
    return -1;
  ms += gpu_spanning_tree(&d_input, temp_array, V);
  if (hipMemcpy(labels_out, d_input.label, 4 * (size_t)V, hipMemcpyDeviceToHost) ||
      hipMemcpy(t1_out, d_input.T1edges, 8 * (size_t)V, hipMemcpyDeviceToHost) ||
      hipMemcpy(t2_out, d_input.T2edges, 8 * (size_t)V, hipMemcpyDeviceToHost))
    return -1;
  *ms_out = ms;
  (void)hipFree(temp_array);
  (void)hipFree(d_input.edges); (void)hipFree(d_input.edges2); (void)hipFree(d_input.label);
  (void)hipFree(d_input.temp_label); (void)hipFree(d_input.T1edges); (void)hipFree(d_input.T2edges);
  (void)hipHostFree(pinned);
  return 0;
}
