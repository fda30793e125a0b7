// paint_common.h -- host-side declarations shared by the painter's sources.
#pragma once

#include <cstddef>

namespace mb3d {

// The 8 diffuse look-up tables of 128 entries (cos_tabs,
// src/lighting.rs:464-499): the 4 light functions, then their rough
// variants.
void paint_cos_tables(float tab[8][128]);

// A device buffer of at least `bytes` owned by the library context (null
// and an error status on failure); valid until the next call that asks.
void *lib_work(size_t bytes, int *status);

}  // namespace mb3d
