// Entry points whose kernels are not built yet report MQ_ERR_UNSUPPORTED.
#include "mq_internal.hpp"

hipError_t launch_init_csr(mq_state*, const Group&, Key2) { return hipErrorNotSupported; }
hipError_t launch_sweep_csr(mq_state*, const Group&, int, uint32_t, Key2) { return hipErrorNotSupported; }
hipError_t launch_measure_csr(mq_state*, const Group&) { return hipErrorNotSupported; }
hipError_t launch_wolff_csr(mq_state*, const Group&, int, uint32_t, Key2) { return hipErrorNotSupported; }

extern "C" int mq_lattice_csr(mq_ctx* ctx, int64_t, const int32_t*, const int32_t*, mq_lattice** out)
{
    if (out) *out = nullptr;
    return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "CSR lattices not built yet");
}
extern "C" int mq_renormalize(mq_state* st) { return mq_fail(st ? st->ctx : nullptr, MQ_ERR_UNSUPPORTED, "not built yet"); }
extern "C" int mq_pt_swap(mq_state* st, int, int, uint64_t, uint64_t, int32_t*) { return mq_fail(st ? st->ctx : nullptr, MQ_ERR_UNSUPPORTED, "not built yet"); }
extern "C" int mq_pt_get_temps(mq_state* st, int32_t*) { return mq_fail(st ? st->ctx : nullptr, MQ_ERR_UNSUPPORTED, "not built yet"); }
extern "C" int mq_run_graph(mq_state* st, int, int, uint64_t, uint64_t, uint64_t, int, int) { return mq_fail(st ? st->ctx : nullptr, MQ_ERR_UNSUPPORTED, "not built yet"); }
