// kernels_host.hip - the device half of a generation whose fitness the
// caller computes on the host (configs with a host fitness_function and no
// engine decoder; the reference evaluates every individual that way,
// core/population.hpp:626-653): one workgroup per child slot builds the
// child's keys (elite / migrant copy with its fitness, mutant, or biased
// crossover) exactly as the fused step does, and stops before the decode.
// The caller reads the rows, evaluates them and hands the fitness back; the
// partition then ranks the islands (kernels_brkga.hip).
#include "launch.hpp"
#include "step_device.hpp"

namespace brkga {

constexpr int kHostChildBlock = 256;

__global__ __launch_bounds__(kHostChildBlock) void k_host_children(InstDesc d, int cur, int init)
{
    extern __shared__ __align__(16) char lds[];
    const int gs = blockIdx.x;
    const int island = gs / d.Pi;
    const int s = gs - island * d.Pi;
    const int nxt = init ? 0 : cur ^ 1;
    const int t = init ? 0 : *d.gen;
    float4* out = reinterpret_cast<float4*>(d.keys[nxt] + (size_t)gs * d.ld);
    float* fit_out = d.fit[nxt] + (size_t)gs * d.prob.num_obj;
    const int* order = d.perm[cur] + (size_t)island * d.Pi;
    (void)build_child<kHostChildBlock>(d, island, s, cur, init, t, order, nullptr, out,
                                       reinterpret_cast<float4*>(lds), fit_out);
}

hipError_t launch_host_children(const InstDesc& d, int cur, int init, hipStream_t s)
{
    const int lds = round_up_int(d.L, 32) * 4;
    hipLaunchKernelGGL(k_host_children, dim3(d.pop), dim3(kHostChildBlock), lds, s, d, cur, init);
    return hipGetLastError();
}

}  // namespace brkga
