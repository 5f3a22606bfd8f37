// Batched driven solves (many grids x frequencies x sources).
#include "runtime.h"
#include "../../include/fdfd.h"

struct fdfd_batch { fdfd_handle* h; };
extern "C" {
int fdfd_batch_create(fdfd_handle* h, fdfd_batch** out) {
    return fdfd::fail(h ? &h->impl : nullptr, FDFD_ERR_INVALID, "batch: not built");
}
int fdfd_batch_add(fdfd_batch*, int, int, double, double, int, int, double, int,
                   const fdfd_complex*, int, const double*, int, const fdfd_complex*, int*) {
    return FDFD_ERR_INVALID;
}
int fdfd_batch_run(fdfd_batch*, int, double, fdfd_stats*) { return FDFD_ERR_INVALID; }
int fdfd_batch_result(fdfd_batch*, int, int, int, fdfd_complex*) { return FDFD_ERR_INVALID; }
int fdfd_batch_destroy(fdfd_batch*) { return FDFD_ERR_INVALID; }
}
