// AWProject gridder (filled in below the WProject path).
#include "gridders.h"

namespace ygrid {
std::unique_ptr<Gridder> make_awproject(std::shared_ptr<Ctx>, const Params&) {
  throw Error(5, "GPUAWProject is not available in this build");
}
}  // namespace ygrid
