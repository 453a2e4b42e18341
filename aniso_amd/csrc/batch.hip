// Batched multi-view forward: many cameras of one scene per pass of batched launches.
#include "aniso_amd.h"
#include "engine.hpp"

namespace ans {

struct ViewBatch {};

void ViewBatchDeleter::operator()(ViewBatch* b) const { delete b; }

}  // namespace ans
