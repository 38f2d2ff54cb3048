// kst_stage.h -- the driver's direct path into the backend's pinned staging
// for spawn commands: reserve room for `n`, build the commands there, commit
// how many were built. kst_spawn is reserve + copy + commit; the driver's own
// block loop skips the copy. Both return KST_OK or an error code with the
// message in kst_last_error().
#pragma once
#include <cstddef>

#include "kestrel_hip.h"

namespace kst {
int spawn_reserve(kst_backend *b, size_t n, kst_spawn_cmd **out);
int spawn_commit(kst_backend *b, size_t n);
}  // namespace kst
