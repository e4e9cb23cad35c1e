// NexDynAdapter.h:17 of the reference: the standalone name of the oracle.
#pragma once

#include "SdfOracle/SdfOracle.h"

namespace sdf {

using StandaloneSdfOracle = SdfOracle;

} // namespace sdf
