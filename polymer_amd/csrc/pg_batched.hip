// pg_batched.hip -- batched forms (filled in below).
#include "pg_engine.h"
#include "pg_internal.h"

extern "C" int pg_msbfs(pg_graph *, const int64_t *, int, int32_t *, uint32_t, pg_stats *) {
    return PG_ERR_UNSUPPORTED;
}
extern "C" int pg_mssssp(pg_graph *, const int64_t *, int, int32_t *, uint32_t, pg_stats *) {
    return PG_ERR_UNSUPPORTED;
}
extern "C" int pg_ppr_batched(pg_graph *, const int64_t *, int, int, double, float *, uint32_t,
                              pg_stats *) {
    return PG_ERR_UNSUPPORTED;
}
extern "C" int pg_ppr(pg_graph *, int64_t, int, double, float *, uint32_t, pg_stats *) {
    return PG_ERR_UNSUPPORTED;
}
extern "C" int pg_graph_multi(int, const int64_t *, int64_t, const int64_t *, const int32_t *,
                              pg_graph **) {
    return PG_ERR_UNSUPPORTED;
}
extern "C" int pg_pagerank_multigraph(pg_graph *, int, double, float *, uint32_t, pg_stats *) {
    return PG_ERR_UNSUPPORTED;
}
extern "C" int pg_graph_autotune(pg_graph *, int32_t *, int64_t *) { return PG_ERR_UNSUPPORTED; }
