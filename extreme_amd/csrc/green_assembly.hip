// Device assembly of the AtoA Green tensor (placeholder until the kernels land).
#include "xtr_green.h"
#include "xtr_common.h"

namespace xtr {

void build_green_atoa(const GreenInput &, double2 *, const double2 *, const double2 *, hipStream_t)
{
    throw hip_error("device Green-tensor assembly is not available in this build");
}

}  // namespace xtr
