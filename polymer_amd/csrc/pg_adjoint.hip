// pg_adjoint.hip -- weighted PageRank forward and adjoint (filled in below).
#include "pg_engine.h"
#include "pg_internal.h"

extern "C" int pg_wpr_forward(pg_graph *, const float *, const float *, double, int, double,
                              float *, uint32_t, pg_stats *) {
    return PG_ERR_UNSUPPORTED;
}
extern "C" int pg_wpr_adjoint(pg_graph *, const float *, int, double, float *, float *, float *,
                              uint32_t, pg_stats *) {
    return PG_ERR_UNSUPPORTED;
}
