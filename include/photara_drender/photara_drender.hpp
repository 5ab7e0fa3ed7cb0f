// Umbrella header of the drop-in C++ API (raster / interpolate / texture / refine path).
#pragma once

#include "photara_drender/context.hpp"
#include "photara_drender/rasterizer.hpp"
#include "photara_drender/texture_baker.hpp"
#include "photara_drender/texture_padding.hpp"
#include "photara_drender/texture_refiner.hpp"
#include "photara_drender/types.hpp"
