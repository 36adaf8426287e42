// pq_learned.hip -- learned index over the suffix arrays (placeholder until the
// model kernels land): training is refused, so learned search is unavailable.
#include "pq_ctx.h"

struct Seed;

void pq_learned_free(pq_ctx_t *ctx)
{
    ctx->learned = nullptr;
}

int pq_learned_seeds(pq_ctx_t *ctx, uint64_t, uint64_t, Seed *, uint32_t *)
{
    return ctx->fail(PQ_ERR_STATE, "learned index not available");
}

extern "C" void pq_learned_default(pq_learned_cfg_t *cfg)
{
    if (!cfg)
        return;
    cfg->k = 21;
    cfg->hidden = 64;
    cfg->leaves = 4096;
    cfg->epochs = 4;
    cfg->batch = 4096;
    cfg->samples = 1 << 20;
    cfg->lr = 1e-3f;
    cfg->seed = 12345;
}

extern "C" int pq_learned_train(pq_ctx_t *ctx, const pq_learned_cfg_t *)
{
    if (!ctx)
        return PQ_ERR_ARG;
    return ctx->fail(PQ_ERR_STATE, "learned index not available");
}

extern "C" int pq_learned_stats(pq_ctx_t *ctx, pq_learned_stats_t *)
{
    if (!ctx)
        return PQ_ERR_ARG;
    return ctx->fail(PQ_ERR_STATE, "learned index not trained");
}
