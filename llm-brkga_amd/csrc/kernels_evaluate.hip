// kernels_evaluate.hip - decode + fitness of caller-owned key rows, one
// workgroup per row: the evaluate_population_gpu hook (core/config.hpp:50-53).
#include "launch.hpp"
#include "launch_util.hpp"
#include "step_device.hpp"

namespace brkga {

template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_evaluate(ProblemDesc p, const float* __restrict__ keys, long long ld,
                                                    int count, float* __restrict__ fit, int* __restrict__ decoded)
{
    extern __shared__ __align__(16) char lds[];
    const int row = blockIdx.x;
    if (row >= count) return;
    const int L = p.chrom_len;
    const int nb = sort_buckets(p.n);
    const DecodeLayout lay = decode_layout(KIND, p.n, L, nb, p.aux);
    float* k = reinterpret_cast<float*>(lds + lay.keys_off);
    const float* src = keys + (size_t)row * ld;
    for (int j = threadIdx.x; j < L; j += BLOCK) k[j] = src[j];
    __syncthreads();
    decode_block<KIND, BLOCK>(p, lds, lay, nb, fit ? fit + (size_t)row * p.num_obj : nullptr,
                              decoded ? decoded + (size_t)row * L : nullptr);
}

namespace {

template <int KIND, int BLOCK>
hipError_t eval_one(const ProblemDesc& p, const float* keys, long long ld, int count, float* fit, int* decoded,
                    int lds, hipStream_t s)
{
    ensure_lds(k_evaluate<KIND, BLOCK>, lds);
    hipLaunchKernelGGL((k_evaluate<KIND, BLOCK>), dim3(count), dim3(BLOCK), lds, s, p, keys, ld, count, fit,
                       decoded);
    return hipGetLastError();
}

template <int KIND>
hipError_t eval_kind(int block, const ProblemDesc& p, const float* keys, long long ld, int count, float* fit,
                     int* decoded, int lds, hipStream_t s)
{
    switch (block) {
    case 64: return eval_one<KIND, 64>(p, keys, ld, count, fit, decoded, lds, s);
    case 128: return eval_one<KIND, 128>(p, keys, ld, count, fit, decoded, lds, s);
    case 256: return eval_one<KIND, 256>(p, keys, ld, count, fit, decoded, lds, s);
    case 512: return eval_one<KIND, 512>(p, keys, ld, count, fit, decoded, lds, s);
    case 1024: return eval_one<KIND, 1024>(p, keys, ld, count, fit, decoded, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_evaluate(const ProblemDesc& p, const float* keys, long long ld, int count, float* fit,
                           int* decoded, int block, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int lds = decode_layout(p.kind, p.n, p.chrom_len, sort_buckets(p.n), p.aux).bytes;
    switch (p.kind) {
    case KIND_TSP_COORDS: return eval_kind<KIND_TSP_COORDS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_TSP_MATRIX: return eval_kind<KIND_TSP_MATRIX>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_KNAP_THRESHOLD:
        return eval_kind<KIND_KNAP_THRESHOLD>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_KNAP_GREEDY: return eval_kind<KIND_KNAP_GREEDY>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_TSPJ: return eval_kind<KIND_TSPJ>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_BITSP_COORDS: return eval_kind<KIND_BITSP_COORDS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_BITSP_MATRIX: return eval_kind<KIND_BITSP_MATRIX>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_CONTINUOUS: return eval_kind<KIND_CONTINUOUS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_JOBSHOP: return eval_kind<KIND_JOBSHOP>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_ZDT1: return eval_kind<KIND_ZDT1>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_VRP_DAYS: return eval_kind<KIND_VRP_DAYS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_MOVRP: return eval_kind<KIND_MOVRP>(block, p, keys, ld, count, fit, decoded, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace brkga
