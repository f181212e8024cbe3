// Cartesian factorised signal: complex GEMM on the matrix cores.
// (sampling analysis only in this revision; execution falls back to direct)
#include "signal_plan.h"

namespace fm {

bool cartesian_analyse(fm_signal_plan *, const double *, const double *, const double *, const double *)
{
    return false;
}
int cartesian_upload(fm_signal_plan *) { return FM_OK; }
void cartesian_free(fm_signal_plan *) {}
size_t cartesian_workspace(const fm_signal_plan *, int, int64_t, int) { return 0; }
int cartesian_execute(const fm_signal_plan *, int, int64_t, int, const double *, void *, size_t, double *,
                      hipStream_t)
{
    set_error("Cartesian GEMM path not available");
    return FM_ERR_UNSUPPORTED;
}

}  // namespace fm
