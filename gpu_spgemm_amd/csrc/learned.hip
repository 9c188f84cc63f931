// learned.hip -- learned scheduler (placeholder until the MLP path lands).
#include "tsg_internal.h"

namespace tsg {
int spgemm_plain(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                 const tsg_options *opts, tsg_stats *st);

int spgemm_learned(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts, tsg_stats *st) {
    (void)ctx; (void)A; (void)B; (void)C; (void)opts; (void)st;
    return fail(TSG_ERR_UNSUPPORTED, "learned scheduler not available in this build");
}
void learned_release(Context *ctx) { (void)ctx; }
}  // namespace tsg

extern "C" int tsg_spgemm_batched(tsg_context *c, int32_t count, const tsg_tile_matrix *A,
                                  const tsg_tile_matrix *B, tsg_tile_matrix *C, const tsg_options *opts,
                                  tsg_stats *stats) {
    (void)c; (void)count; (void)A; (void)B; (void)C; (void)opts; (void)stats;
    return tsg::fail(TSG_ERR_UNSUPPORTED, "batched launch not available in this build");
}
