// Host helpers shared by the C-ABI translation units (implemented in mp_api.hip).
#pragma once

#include <string>

#include "../../include/micropolis.h"
#include "mp_types.h"

namespace mp {

// Projection + framebuffer record of one view (src/base/Projection.cpp:23-50,
// glm::perspective with far 10000; clear value of src/Reyes/Framebuffer.cpp:65-68).
void make_view(const mp_projection& pr, const float clear[4], int subpixel_bits, DevView& v);

// 0 with `why` set when the options are outside what the library supports.
int validate_config(const mp_config* c, std::string& why);

// Display thresholds: tab[k - 1] = smallest linear value whose display byte
// is at least k (k = 1..255); tab[255] = NaN.
void display_thresholds(float tab[256]);

// Reason of the last failure of a context-free entry point (mp_create_error()).
void set_create_error(const std::string& msg);

// DevParams of a configuration with the library's default scheduling choices.
DevParams device_params(const mp_config* cfg);

}  // namespace mp
