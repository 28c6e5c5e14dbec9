// temporary: 3PCF/4PCF entry points until their kernels land
#include "gs_api_internal.h"
using namespace gs;
namespace gs {
void extra_release(GsExtra& x) { x.params.release(); x.out.release(); x.queue.release(); x.dbl.release(); x.i32.release(); }
}
extern "C" int gs_count_3pcf(gs_ctx*, const gs_params*, uint64_t*, double*, double*, double*, gs_stats*)
{ return fail(GS_ELIMIT, "3PCF not built yet"); }
extern "C" int gs_count_4pcf(gs_ctx*, const gs_params*, uint64_t*, double*, double*, double*, double*, gs_stats*)
{ return fail(GS_ELIMIT, "4PCF not built yet"); }
