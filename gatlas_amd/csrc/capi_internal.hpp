// Definitions shared by the extern "C" translation units.
#pragma once

#include "context.hpp"

struct gat_ctx {
    gat::Context impl;
};
