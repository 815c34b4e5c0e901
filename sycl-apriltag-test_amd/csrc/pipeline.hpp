// pipeline.hpp - the whole-frame detector (threshold -> quads).
#pragma once
#include "common.hpp"
