// Placeholder until the tiled gridder lands.
#include "internal.hpp"

namespace migrid {
int launch_tiled(migrid_ctx*, const GridArgs&, const migrid_opts& o, hipStream_t)
{
    return fail(MIGRID_ENOTSUP, "strategy %d not built yet", o.strategy);
}
}  // namespace migrid
