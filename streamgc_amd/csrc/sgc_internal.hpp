// sgc_internal.hpp -- host-side helpers shared by the engine translation units.
#pragma once
#include <stdint.h>

#include <string>

namespace sgc {

// Tier-2 host repair over packed words; returns vertices recoloured.
uint32_t tier2_host(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, uint32_t* word, uint32_t passes);

}  // namespace sgc
